"""Benchmark of the STM B+tree batch path on one MI355X.

Workload (the 2^26-key, 95 % lookup / 5 % insert configuration the headline metric is
quoted on): bulk-load 2^26 unique keys drawn from Gamma(k=2, theta=2^56), values
splitmix64(key); then batches of 2^22 ops, 95 % point lookups (90 % of them hits sampled
uniformly from the loaded keys, 10 % fresh misses) and 5 % inserts of fresh keys of the
same distribution.  A step is one batch.  All batches are generated on the host and copied
to HBM before the timed region.  Each route (plain descent, learned routing, the static leaf
index; at the library defaults and at the launch parameters of the tuning record) runs them
back to back through stm_run_batch_device, one synchronous call per batch; the learned and
descent routes also run them as one device-resident stream (stm_run_batches_device, graph
replay).

Prints one JSON line.  `value` is Mops/s over the K timed batches of the learned route at
the tuned parameters, as a stream (wall clock from the first launch to the last batch's
completion), the mean of two such timed streams (the stream routes alternate).  Per-kernel
device times come from HIP events the library records on its own stream around every phase
of the synchronous runs (stm_config.timing).

A second section ("uneven") runs the uneven-batching workload: 2^26 Normal(2^62, 2^58)
keys and a stream of batches with sizes log-uniform in [2^10, 2^22], replayed from captured
hipGraphs with launch parameters chosen by eirene_amd.autotune, beside the same stream
launched batch by batch from the host with the library defaults.  A third ("serving") runs the
continuous-batching request loop open-loop: 2^25 uniform keys, Poisson arrivals of requests
of 1..2^16 ops (90% lookups, 8% inserts, 2% range scans), request latency percentiles.

    python bench.py [--steps K] [--warmup W] [--route descent|learned|index] [--no-cpu]
                    [--uneven-batches N] [--tuning FILE]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mops/s (95% lookup/5% insert, 2^26 keys): learned vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per op with a 64-slot u64 accounting node: a lookup reads its key (8),
# the leaf's keys (512) and the value (8) and writes the result (8); an insert reads its key
# and value (16) and the leaf's keys (512, routed by the lookup kernel) and its metadata line
# (64), appends one key and value (16), writes the count and version (16) and the result (8)
# (leaf entries are unordered, so nothing else of the leaf is read or written).
BYTES_LOOKUP = 536
BYTES_INSERT = 632


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--load-log2", type=int, default=26)
    ap.add_argument("--batch-log2", type=int, default=22)
    ap.add_argument("--fanout", type=int, default=64)
    ap.add_argument("--lanes", type=int, default=16)
    ap.add_argument("--group", type=int, default=16)
    ap.add_argument("--route", choices=["descent", "learned", "index", "both"], default="both",
                    help="both = descent, learned and the static index")
    ap.add_argument("--tuning-headline", default=str(ROOT / "eirene_amd" / "tuning" /
                                                     "gamma_2p26_batch_2p22.json"),
                    help="tuning record (eirene_amd.autotune) for the tuned routes; '' = none")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU-baseline budget (batches run until it is spent)")
    ap.add_argument("--no-stream", dest="stream", action="store_false",
                    help="time only one synchronous call per batch (no stream runs)")
    ap.add_argument("--stream-rounds", type=int, default=2,
                    help="timed streams per route, the routes alternating; a route's figure "
                         "(and `value`) is their mean")
    ap.add_argument("--no-headline", action="store_true",
                    help="skip the learned-vs-plain section (profiling the other sections alone)")
    ap.add_argument("--serve-requests", type=int, default=8000,
                    help="serving loop: requests offered (0 = skip)")
    ap.add_argument("--serve-warmup", type=int, default=2000,
                    help="serving requests run before the measured ones (device clocks and host "
                         "threads reach steady state)")
    ap.add_argument("--serve-mops", type=float, default=200.0,
                    help="serving loop: offered load in Mops/s (Poisson arrivals)")
    ap.add_argument("--uneven-batches", type=int, default=256,
                    help="uneven stream: batches log-uniform in [2^10, 2^22] (0 = skip)")
    ap.add_argument("--tuning", default=None,
                    help="tuning record to use for the uneven stream instead of autotuning")
    ap.add_argument("--tune-batches", type=int, default=40,
                    help="batches of the stream the autotuner measures")
    ap.add_argument("--no-small-batches", dest="small_batches", action="store_false",
                    help="skip the per-route streams of 2^12..2^16-op batches")
    ap.add_argument("--plain-trials", type=int, default=3,
                    help="plain-uniform section: fresh-tree trials (0 = skip the section)")
    ap.add_argument("--contended-batches", type=int, default=8,
                    help="contended-Zipf section: timed batches (0 = skip the section)")
    ap.add_argument("--only", choices=["plain", "contended", "uneven", "serving"], default=None,
                    help="run one section alone and print its JSON (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's copy kernel 4 times over MIB MiB (the known-size "
                         "kernel HBM traffic counters are scaled by) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    import eirene_amd as ea
    n = mib * (1 << 20) // 8
    print(json.dumps({"kernel": "k_copy16", "bytes_read": n * 8, "bytes_written": n * 8,
                      "dispatches": 5, "gbs_read_plus_write": round(copy_ceiling_gbs(ea, mib), 1)}))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """The host CPU: model, logical CPUs, the CPUs this process may run on, and the worker
    threads the CPU baseline uses (every CPU granted to the process, at most 16: a GPU job on
    these machines is given 16 host threads)."""
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        granted = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        granted = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "granted": granted,
            "threads_used": max(1, min(granted, 16))}


def record_name(path):
    """A tuning record's path relative to the repository when it lies inside it."""
    p = Path(path).resolve()
    try:
        return str(p.relative_to(ROOT.resolve()))
    except ValueError:
        return str(p)


def copy_ceiling_gbs(ea, mib=4096, reps=4):
    """Read + write GB/s of the library's streaming copy kernel over mib MiB: the achievable
    HBM rate beside the nominal peak."""
    n = mib * (1 << 20) // 8
    src = ea.DeviceBuffer(n, np.uint64)
    dst = ea.DeviceBuffer(n, np.uint64)
    try:
        dst.copy_from(src)  # first touch
        t0 = time.perf_counter()
        for _ in range(reps):
            dst.copy_from(src)
        dt = (time.perf_counter() - t0) / reps
    finally:
        src.free()
        dst.free()
    return 2 * n * 8 / dt / 1e9


def run_uneven(args, ea, w):
    """Uneven batching: 2^load_log2 Normal(2^62, 2^58) keys, a stream of batches with sizes
    log-uniform in [2^10, 2^22] at 95/5, replayed from captured graphs with the autotuned
    launch parameters, against the same stream launched batch by batch from the host with the
    library defaults."""
    from eirene_amd import autotune as at
    n_load = 1 << args.load_log2
    t0 = time.time()
    load = w.normal_keys(n_load, args.seed + 3)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    keys, vals, ops, offsets = at.uneven_sample(srt, args.uneven_batches, args.seed + 3, 10, 22,
                                                w.DIST_NORMAL, 2.0 ** 62, 2.0 ** 58)
    n_ops = int(offsets[-1])
    log(f"[uneven] generated {n_load} keys and {args.uneven_batches} batches ({n_ops} ops) "
        f"in {time.time() - t0:.1f}s")
    sample = at.Sample(load, lv, keys, vals, ops, offsets)
    base = ea.StmConfig(max_keys=n_load + n_ops + (1 << 20), max_batch=1 << 22)
    t0 = time.time()
    if args.tuning:
        rec = at.load_tuning(args.tuning)
        tuned_from = args.tuning
    else:
        tb = max(1, min(args.tune_batches, args.uneven_batches))
        tune_sample = at.Sample(load, lv, keys[: int(offsets[tb])], vals[: int(offsets[tb])],
                                ops[: int(offsets[tb])], offsets[: tb + 1])
        rec = at.tune(tune_sample, base, log=lambda m: log(f"[tune] {m}"))
        tune_sample.free()
        tuned_from = f"autotuned in this run on the first {tb} batches"
    log(f"[uneven] tuning {rec['best']} ({time.time() - t0:.1f}s)")

    def timed(cfg, graphs):
        cfg = ea.StmConfig(**{**cfg.__dict__, "use_graphs": graphs})
        with ea.StmTree(cfg) as t:
            t.bulk_load_device(sample.load_k, sample.load_v, sample.n_load)
            if graphs:
                t.capture_graphs()
            t.sync()
            t0 = time.perf_counter()
            t.run_batches_device(sample.k, sample.v, sample.o, sample.out_v, sample.out_f, offsets)
            dt = time.perf_counter() - t0
            st = t.stats()
            out = (sample.out_v.download(), sample.out_f.download(), t.dump())
        return n_ops / dt / 1e6, dt, st, out

    tuned = at.apply_tuning(base, rec["best"])
    m_tg, dt_tg, st_tg, out_tg = timed(tuned, True)
    m_dg, dt_dg, _, _ = timed(base, True)
    m_dh, dt_dh, _, out_dh = timed(base, False)
    same = bool(all(np.array_equal(a, b) for a, b in zip(out_tg[:2], out_dh[:2]))
                and all(np.array_equal(a, b) for a, b in zip(out_tg[2], out_dh[2])))
    log(f"[uneven] graphs+tuned {m_tg:.1f}, graphs+default {m_dg:.1f}, host+default {m_dh:.1f} "
        f"Mops/s; identical={same}")
    sample.free()
    return {
        "workload": f"bulk-load 2^{args.load_log2} Normal(2^62, 2^58) keys; {args.uneven_batches} "
                    f"batches log-uniform in [2^10, 2^22] ({n_ops} ops), 95% lookup / 5% insert",
        "mops_graphs_tuned": round(m_tg, 2), "mops_graphs_default": round(m_dg, 2),
        "mops_host_default": round(m_dh, 2),
        "ms_graphs_tuned": round(dt_tg * 1e3, 2), "ms_host_default": round(dt_dh * 1e3, 2),
        "tuned_config": rec["best"], "tuned_from": tuned_from,
        "autotune_trials": len(rec.get("trials", [])),
        "results_identical_tuned_vs_default": same,
        "stm": {k: st_tg[k] for k in ("upserts", "aborts", "leaf_splits", "deferred", "height")},
    }


def phase_summary(st, batches):
    return {k: round(v / batches, 4) for k, v in st["phase_ms"].items() if st["phase_launches"][k]}


def run_plain(args, ea, w):
    """Plain method, small: 2^23 uniform keys (values splitmix64(key)), one batch of 2^22 ops,
    95% lookups (90% of them hits) / 5% inserts of fresh keys.  Timed on a fresh tree per trial
    after one untimed batch of the same shape: the device-resident call (stm_run_batch_device),
    then the host-buffer call (stm_run_batch: pageable host arrays in, PCIe both ways, the
    reference's transferData / transferAns, bpt-gpustm.cpp:313-335).  Median of the trials."""
    n_load, n = 1 << 23, 1 << 22
    t0 = time.time()
    load = w.uniform_keys(n_load, args.seed + 11)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    keys, vals, ops = w.mixed_batch(srt, n, args.seed + 12, 0.95, 0.9)
    wk, wv, wo = w.mixed_batch(srt, n, args.seed + 13, 0.95, 0.9)
    log(f"[plain] generated in {time.time() - t0:.1f}s")
    cfg = ea.StmConfig(max_keys=n_load + 2 * n + (1 << 20), max_batch=n, timing=True)
    dk, dv = ea.DeviceBuffer.from_array(load), ea.DeviceBuffer.from_array(lv)
    bk, bv, bo = (ea.DeviceBuffer.from_array(a) for a in (keys, vals, ops))
    xk, xv, xo = (ea.DeviceBuffer.from_array(a) for a in (wk, wv, wo))
    ov, of = ea.DeviceBuffer(n, np.uint64), ea.DeviceBuffer(n, np.uint8)
    dev, host, phases = [], [], []
    same = True
    for _ in range(args.plain_trials):
        with ea.StmTree(cfg) as t:
            t.bulk_load_device(dk, dv)
            t.run_batch_device(xk, xv, xo, ov, of)
            t.reset_stats()
            t.sync()
            t1 = time.perf_counter()
            t.run_batch_device(bk, bv, bo, ov, of)
            dev.append(time.perf_counter() - t1)
            phases.append(phase_summary(t.stats(), 1))
            dres = (ov.download(), of.download())
        with ea.StmTree(cfg) as t:
            t.bulk_load_device(dk, dv)
            t.run_batch(wk, wv, wo)
            t.sync()
            t1 = time.perf_counter()
            hres = t.run_batch(keys, vals, ops)
            host.append(time.perf_counter() - t1)
        same = same and bool(np.array_equal(dres[0], hres[0]) and np.array_equal(dres[1], hres[1]))
    for b in (dk, dv, bk, bv, bo, xk, xv, xo, ov, of):
        b.free()
    md, mh = float(np.median(dev)), float(np.median(host))
    mid = phases[int(np.argsort(dev)[len(dev) // 2])]
    log(f"[plain] device {n / md / 1e6:.1f} Mops/s, host buffers {n / mh / 1e6:.1f} Mops/s")
    return {
        "workload": "bulk-load 2^23 uniform keys; one batch of 2^22 ops, 95% lookup (90% hits) / "
                    "5% insert of fresh keys",
        "mops_device": round(n / md / 1e6, 2), "ms_device": round(md * 1e3, 3),
        "mops_host_buffers": round(n / mh / 1e6, 2), "ms_host_buffers": round(mh * 1e3, 3),
        "host_buffers_note": "stm_run_batch from pageable host arrays: 17 B/op in and 9 B/op out "
                             "over PCIe inside the timed call",
        "phase_ms": mid, "trials": args.plain_trials, "identical_device_vs_host_api": same,
    }


def run_contended(args, ea, w):
    """Contended STM: 2^24 uniform keys; 8 batches of 2^22 ops, 50% lookups / 50% inserts, op
    keys at Zipf(0.99) ranks of the loaded keys, inserts next to the chosen key so they hammer
    the hot leaves.  One synchronous stm_run_batch_device call per batch after one untimed
    batch of the same shape; abort / deferral counts from the library's counters."""
    n_load, n, nb = 1 << 24, 1 << 22, args.contended_batches
    t0 = time.time()
    load = w.uniform_keys(n_load, args.seed + 21)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    keys = np.empty((nb + 1) * n, dtype=np.uint64)
    vals = np.empty_like(keys)
    ops = np.empty((nb + 1) * n, dtype=np.uint8)
    for b in range(nb + 1):
        s_ = slice(b * n, (b + 1) * n)
        keys[s_], vals[s_], ops[s_] = w.zipf_batch(srt, n, args.seed * 100 + 30 + b, 0.99, 0.5)
    log(f"[contended] generated in {time.time() - t0:.1f}s")
    cfg = ea.StmConfig(max_keys=n_load + (nb + 1) * n + (1 << 20), max_batch=n, timing=True)
    bk, bv, bo = (ea.DeviceBuffer.from_array(a) for a in (keys, vals, ops))
    ov, of = ea.DeviceBuffer(keys.size, np.uint64), ea.DeviceBuffer(keys.size, np.uint8)
    with ea.StmTree(cfg) as t:
        dk, dv = ea.DeviceBuffer.from_array(load), ea.DeviceBuffer.from_array(lv)
        t.bulk_load_device(dk, dv)
        dk.free()
        dv.free()
        t.run_batch_device(bk, bv, bo, ov, of, n=n, offset=0)  # untimed
        t.sync()
        t.reset_stats()
        lat = []
        t1 = time.perf_counter()
        for b in range(1, nb + 1):
            tb = time.perf_counter()
            t.run_batch_device(bk, bv, bo, ov, of, n=n, offset=b * n)
            lat.append(time.perf_counter() - tb)
        elapsed = time.perf_counter() - t1
        st = t.stats()
        bad = t.check()
    for b in (bk, bv, bo, ov, of):
        b.free()
    mops = nb * n / elapsed / 1e6
    mix = 0.5 * BYTES_LOOKUP + 0.5 * BYTES_INSERT
    log(f"[contended] {mops:.1f} Mops/s, aborts {st['aborts']}, deferred {st['deferred']}, "
        f"phases {phase_summary(st, nb)}")
    return {
        "workload": "bulk-load 2^24 uniform keys; 8 batches of 2^22 ops, 50% lookup / 50% insert, "
                    "keys at Zipf(0.99) ranks, inserts next to the chosen key",
        "mops": round(mops, 2), "ms_per_batch": round(elapsed / nb * 1e3, 3),
        "p50_batch_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_batch_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "phase_ms_per_batch": phase_summary(st, nb),
        "unique_ops": st["unique_ops"], "upserts": st["upserts"], "commits": st["commits"],
        "aborts": st["aborts"], "aborts_per_upsert": round(st["aborts"] / max(1, st["upserts"]), 4),
        "deferred": st["deferred"], "deferred_per_upsert": round(st["deferred"] / max(1, st["upserts"]), 5),
        "sort_fallbacks": st["sort_fallbacks"],
        "leaf_splits": st["leaf_splits"], "inner_splits": st["inner_splits"],
        "whole_batch_gbs": round(nb * n * mix / elapsed / 1e9, 1), "bytes_per_op": mix,
        "tree_check_violations": bad,
    }


def run_serving(args, ea, w):
    """Serving loop: 2^25 uniform keys resident in HBM; open-loop Poisson arrivals at
    serve_mops; requests of floor(2^(16u)) ops (1 .. 2^16), 90% lookups, 8% inserts, 2% range
    scans of length 16..256; FIFO admission by HBM-budget slots; request latency percentiles."""
    from eirene_amd.serving import open_loop
    n_load = 1 << 25
    t0 = time.time()
    load = w.uniform_keys(n_load, args.seed + 5)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    cfg = ea.StmConfig(max_keys=n_load + (1 << 24), max_batch=1 << 22)
    with ea.StmTree(cfg) as t:
        dk, dv = ea.DeviceBuffer.from_array(load), ea.DeviceBuffer.from_array(lv)
        t.bulk_load_device(dk, dv)
        dk.free()
        dv.free()
        log(f"[serve] loaded 2^25 keys in {time.time() - t0:.1f}s")
        rep = open_loop(t, srt, requests=args.serve_requests, offered_mops=args.serve_mops,
                        ops_log2=(0, 16), lookup_frac=0.90, insert_frac=0.08, scan_len=(16, 256),
                        seed=args.seed + 7, warmup=args.serve_warmup)
        bad = t.check()
    log(f"[serve] {rep['requests']} requests, {rep['mops']:.1f} Mops/s, p50 {rep['p50_ms']:.2f} ms, "
        f"p99 {rep['p99_ms']:.2f} ms, {rep['batches']} batches")
    return {
        "workload": "2^25 uniform keys; open-loop Poisson arrivals; requests of 1..2^16 ops "
                    "(log-uniform), 90% lookup / 8% insert / 2% scan (16..256)",
        "offered_mops_target": args.serve_mops,
        "offered_mops": round(rep["offered_mops"], 2), "mops": round(rep["mops"], 2),
        "requests": rep["requests"], "ops": rep["ops"], "failed": rep["failed"],
        "latency_ms": {k: round(rep[k + "_ms"], 3) for k in ("p50", "p90", "p99", "p999", "max", "mean")},
        "batches": rep["batches"], "mean_batch_ops": round(rep["mean_batch_ops"], 1),
        "max_pending": rep["max_pending"], "busy_frac": round(rep["busy_frac"], 3),
        "scans": rep["scans"], "scan_pairs": rep["scan_pairs"],
        "hbm_budget_bytes": rep["hbm_budget"], "slot_ops": rep["slot_ops"],
        # requests run first on the same arrival process, outside the figures above
        "warmup_requests": rep["warmup_requests"], "warmup_p99_ms": round(rep["warmup_p99_ms"], 3),
        "tree_check_violations": bad,
    }


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the batch path has no "
                          "multi-GPU mode (key-range sharding is a design note)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import eirene_amd as ea
    from eirene_amd import workloads as w

    if args.only:
        fn = {"plain": run_plain, "contended": run_contended, "uneven": run_uneven,
              "serving": run_serving}[args.only]
        print(json.dumps({args.only: fn(args, ea, w)}), flush=True)
        return
    if args.no_headline:
        print(json.dumps({"uneven": run_uneven(args, ea, w) if args.uneven_batches > 0 else None,
                          "serving": run_serving(args, ea, w) if args.serve_requests > 0 else None}),
              flush=True)
        return

    n_load = 1 << args.load_log2
    n_batch = 1 << args.batch_log2
    n_batches = args.warmup + args.steps
    t0 = time.time()
    load = w.gamma_keys(n_load, args.seed, 2.0, 2.0 ** 56)
    load_vals = w.values_for(load)
    load_sorted = w.sorted_copy(load)
    log(f"generated {n_load} keys in {time.time() - t0:.1f}s")
    t0 = time.time()
    keys = np.empty(n_batches * n_batch, dtype=np.uint64)
    vals = np.empty_like(keys)
    ops = np.empty(n_batches * n_batch, dtype=np.uint8)
    for b in range(n_batches):
        s = slice(b * n_batch, (b + 1) * n_batch)
        keys[s], vals[s], ops[s] = w.mixed_batch(load_sorted, n_batch, args.seed * 1000 + b, 0.95,
                                                 0.9, w.DIST_GAMMA, 2.0, 2.0 ** 56)
    log(f"generated {n_batches} batches in {time.time() - t0:.1f}s")

    tuned = None
    if args.tuning_headline and Path(args.tuning_headline).exists():
        from eirene_amd import autotune as at
        tuned = at.load_tuning(args.tuning_headline)["best"]

    def route_config(route):
        cfg = ea.StmConfig(fanout=args.fanout, max_keys=n_load + n_batches * n_batch + (1 << 20),
                           max_batch=n_batch, lanes_per_op=args.lanes, ops_per_group=args.group,
                           route={"learned": ea.ROUTE_LEARNED, "index": ea.ROUTE_INDEX}.get(
                               route.replace("_tuned", ""), ea.ROUTE_DESCENT),
                           retrain_every=8, timing=True)
        if route.endswith("_tuned"):
            from eirene_amd import autotune as at
            cfg = at.apply_tuning(cfg, tuned)
        return cfg

    tail_batches = min(8, args.steps)

    def tail():
        """Answers of the last timed batches, to check other runs against."""
        o = (n_batches - tail_batches) * n_batch
        return (out_v.download(tail_batches * n_batch, offset=o),
                out_f.download(tail_batches * n_batch, offset=o))

    def run_route(route):
        cfg = route_config(route)
        tree = ea.StmTree(cfg)
        t0 = time.time()
        dk, dv = ea.DeviceBuffer.from_array(load), ea.DeviceBuffer.from_array(load_vals)
        tree.bulk_load_device(dk, dv)
        dk.free()
        dv.free()
        log(f"[{route}] bulk load {time.time() - t0:.2f}s, {tree.stats()['height']} levels")
        for b in range(args.warmup):
            tree.run_batch_device(bk, bv, bo, out_v, out_f, n=n_batch, offset=b * n_batch)
        tree.sync()
        first = (out_v.download(n_batch), out_f.download(n_batch)) if args.warmup else None
        tree.reset_stats()
        lat = []
        t_start = time.perf_counter()
        for b in range(args.warmup, n_batches):
            tb = time.perf_counter()
            tree.run_batch_device(bk, bv, bo, out_v, out_f, n=n_batch, offset=b * n_batch)
            lat.append(time.perf_counter() - tb)
        tree.sync()
        elapsed = time.perf_counter() - t_start
        st = tree.stats()
        info = tree.route_info() if route.startswith("learned") else None
        tree.close()
        mops = args.steps * n_batch / elapsed / 1e6
        log(f"[{route}] {args.steps} batches in {elapsed:.3f}s -> {mops:.1f} Mops/s; "
            f"phases {st['phase_ms']}")
        return {"mops": mops, "elapsed": elapsed, "lat": lat, "st": st, "first": first, "info": info,
                "tail": tail() if args.stream else None}

    def run_stream(route):
        """The same batches through stm_run_batches_device: the warm-up batches as one stream,
        then the timed batches as one device-resident stream (one captured graph per size class,
        replayed; retrains between replays), on a fresh tree from the same bulk load."""
        cfg = route_config(route)
        cfg.timing = False
        tree = ea.StmTree(cfg)
        dk, dv = ea.DeviceBuffer.from_array(load), ea.DeviceBuffer.from_array(load_vals)
        tree.bulk_load_device(dk, dv)
        dk.free()
        dv.free()
        tree.capture_graphs()
        offs = np.arange(n_batches + 1, dtype=np.uint64) * np.uint64(n_batch)
        if args.warmup:
            tree.run_batches_device(bk, bv, bo, out_v, out_f, offs[:args.warmup + 1])
        tree.sync()
        t0 = time.perf_counter()
        tree.run_batches_device(bk, bv, bo, out_v, out_f, offs[args.warmup:])
        elapsed = time.perf_counter() - t0
        tree.close()
        mops = args.steps * n_batch / elapsed / 1e6
        log(f"[{route} stream] {args.steps} batches in {elapsed:.3f}s -> {mops:.1f} Mops/s")
        return {"mops": mops, "elapsed": elapsed, "tail": tail()}

    def small_batch_routes():
        """Routing at small batch sizes (the sizes of the uneven stream and of serving
        requests): per route, on a fresh tree from the same bulk load, streams of batches of
        2^12, 2^14 and 2^16 ops (2^21 ops each, 95/5, Gamma keys) replayed from graphs.  The
        router is trained once before the streams (no periodic retrain inside them), so the
        comparison is the routing itself."""
        sizes = [(12, 512), (14, 128), (16, 32)]
        sets = []
        for lg, cnt in sizes:
            nb_ = 1 << lg
            sk = np.empty(cnt * nb_, dtype=np.uint64)
            sv = np.empty_like(sk)
            so = np.empty(cnt * nb_, dtype=np.uint8)
            for b in range(cnt):
                s_ = slice(b * nb_, (b + 1) * nb_)
                sk[s_], sv[s_], so[s_] = w.mixed_batch(load_sorted, nb_, args.seed * 7000 + lg * 1000 + b,
                                                       0.95, 0.9, w.DIST_GAMMA, 2.0, 2.0 ** 56)
            offs = np.arange(cnt + 1, dtype=np.uint64) * np.uint64(nb_)
            sets.append((lg, cnt, [ea.DeviceBuffer.from_array(a) for a in (sk, sv, so)], offs))
        n_small = max(c * (1 << lg) for lg, c in sizes)
        sov, sof = ea.DeviceBuffer(n_small, np.uint64), ea.DeviceBuffer(n_small, np.uint8)
        out = {}
        for route in ("descent", "learned", "index"):
            cfg = ea.StmConfig(fanout=args.fanout, max_keys=n_load + (1 << 24), max_batch=n_batch,
                               lanes_per_op=args.lanes, ops_per_group=args.group,
                               route={"learned": ea.ROUTE_LEARNED, "index": ea.ROUTE_INDEX}.get(
                                   route, ea.ROUTE_DESCENT), retrain_every=1 << 20)
            with ea.StmTree(cfg) as t:
                dk, dv = ea.DeviceBuffer.from_array(load), ea.DeviceBuffer.from_array(load_vals)
                t.bulk_load_device(dk, dv)
                dk.free()
                dv.free()
                t.capture_graphs()
                r = {}
                for lg, cnt, (sk, sv, so), offs in sets:
                    t.sync()
                    t0 = time.perf_counter()
                    t.run_batches_device(sk, sv, so, sov, sof, offs)
                    r[f"2^{lg}"] = round(cnt * (1 << lg) / (time.perf_counter() - t0) / 1e6, 1)
                out[route] = r
            log(f"[small batches] {route}: {out[route]} Mops/s")
        for _, _, bufs, _ in sets:
            for b_ in bufs:
                b_.free()
        sov.free()
        sof.free()
        out["learned_over_descent"] = {k: round(out["learned"][k] / out["descent"][k], 3)
                                       for k in out["descent"]}
        out["learned_over_index"] = {k: round(out["learned"][k] / out["index"][k], 3)
                                     for k in out["descent"]}
        return out

    bk, bv, bo = (ea.DeviceBuffer.from_array(a) for a in (keys, vals, ops))
    out_v = ea.DeviceBuffer(keys.size, np.uint64)
    out_f = ea.DeviceBuffer(keys.size, np.uint8)
    routes = ["descent", "learned", "index"] if args.route == "both" else [args.route]
    # The tuning record's launch parameters under both routings.
    if tuned and "descent" in routes:
        routes.append("descent_tuned")
    if tuned and "learned" in routes:
        routes.append("learned_tuned")
    res = {r: run_route(r) for r in routes}
    # The workload is the learned-routing configuration: `value` is its learned route, at the
    # tuned launch parameters when a tuning record is present, with the timed batches submitted
    # as one device-resident stream.  Plain descent runs the same way beside it.  Per-call
    # latency and the lookup kernel's roofline come from the synchronous run of the same route.
    head_name = next((r for r in ("learned_tuned", "learned") if r in res), routes[0])
    head = res[head_name]
    st, lat = head["st"], head["lat"]
    streams = {}
    if args.stream and head_name.startswith("learned"):
        plain = head_name.replace("learned", "descent")
        # the learned route at the defaults too, so the tuning's gain is measured on the stream;
        # every route twice, the routes alternating (a route's figure is the mean of its two
        # timed streams, so clock drift over the run does not favour the route timed last)
        order = [r for r in dict.fromkeys((head_name, plain, "learned")) if r in res]
        runs = {r: [] for r in order}
        for _ in range(args.stream_rounds):
            for r in order:
                x = run_stream(r)
                x["identical_to_sync"] = bool(np.array_equal(x["tail"][0], res[r]["tail"][0]) and
                                              np.array_equal(x["tail"][1], res[r]["tail"][1]))
                runs[r].append(x)
        for r in order:
            xs = runs[r]
            streams[r] = {"mops": float(np.mean([x["mops"] for x in xs])),
                          "elapsed": float(np.mean([x["elapsed"] for x in xs])),
                          "runs_mops": [round(x["mops"], 2) for x in xs],
                          "identical_to_sync": all(x["identical_to_sync"] for x in xs)}
    if head_name in streams:
        mops, elapsed = streams[head_name]["mops"], streams[head_name]["elapsed"]
    else:
        mops, elapsed = head["mops"], head["elapsed"]
    total_ops = args.steps * n_batch
    same = None
    firsts = [x["first"] for x in res.values() if x["first"] is not None]
    if len(firsts) >= 2:
        same = bool(all(np.array_equal(firsts[0][0], f[0]) and np.array_equal(firsts[0][1], f[1])
                        for f in firsts[1:]))

    # Roofline of the read-only lookup kernel (the tree kernel every lookup takes).
    lookups = st["unique_ops"] - st["upserts"]
    lk_launch = max(st["phase_launches"]["lookup"], 1)
    lk_ms = st["phase_ms"]["lookup"] / lk_launch
    lk_bytes = lookups / lk_launch * BYTES_LOOKUP
    achieved = lk_bytes / (lk_ms * 1e-3) / 1e9 if lk_ms > 0 else 0.0
    mix_bytes = 0.95 * BYTES_LOOKUP + 0.05 * BYTES_INSERT
    head_fanout = tuned["fanout"] if head_name.endswith("_tuned") else args.fanout
    head_lanes = tuned["lanes_per_op"] if head_name.endswith("_tuned") else args.lanes
    kname = "k_lookup<%d,%d,%s>" % (head_fanout, head_lanes,
                                    "true" if head_name.startswith("learned") else "false")
    # Measured L2<->memory bytes per launch of this kernel for this workload (rocprofv3 PMC
    # FETCH_SIZE / WRITE_SIZE, scaled by the copy-kernel calibration), committed with the profile.
    traffic = None
    tf = ROOT / "profiles" / "hbm_traffic_gamma_2p26_batch_2p22.json"
    if tf.exists() and args.load_log2 == 26 and args.batch_log2 == 22:
        kt = json.loads(tf.read_text())["routes"].get(head_name)
        kk = next((k for k in kt if k.startswith(kname)), None) if kt else None
        if kk:
            traffic = kt[kk]["read_bytes"] + kt[kk]["written_bytes"]
    roofline = {
        "kernel": kname, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "traffic_source": "profiles/hbm_traffic_gamma_2p26_batch_2p22.json" if traffic else None,
        "avg_launch_ms": round(lk_ms, 4), "lookups_per_launch": round(lookups / lk_launch),
        "bytes_per_lookup": BYTES_LOOKUP,
        "whole_batch_gbs": round(total_ops * mix_bytes / elapsed / 1e9, 1),
    }
    # The measured streaming copy (read + write) is the achievable HBM rate. The lookup's
    # algorithmic bytes include inner nodes the caches serve, so the rate of its counted HBM
    # traffic is the figure to hold against that ceiling.
    ceiling = copy_ceiling_gbs(ea)
    roofline["copy_ceiling_gbs"] = round(ceiling, 1)
    if traffic:
        traffic_gbs = traffic / (lk_ms * 1e-3) / 1e9
        roofline["traffic_gbs"] = round(traffic_gbs, 1)
        roofline["traffic_frac_of_copy"] = round(traffic_gbs / ceiling, 4)
    first_gpu = head["first"]

    cpu = None
    agree = None
    if not args.no_cpu:
        import oracle
        oracle.build()
        host = host_cpu()
        threads = host["threads_used"]
        t0 = time.time()
        ex = oracle.CpuExecutor(load_sorted, w.values_for(load_sorted), threads)
        log(f"cpu executor built in {time.time() - t0:.1f}s with {threads} threads")
        cpu_sec = 0.0
        cpu_ops = 0
        b = 0
        while b < n_batches and (cpu_sec < args.cpu_seconds or b < 2):
            s = slice(b * n_batch, (b + 1) * n_batch)
            cv, cf, sec = ex.run_batch(keys[s], vals[s], ops[s])
            if b == 0 and first_gpu is not None:  # first batch: same starting tree
                agree = bool(np.array_equal(first_gpu[0], cv) and np.array_equal(first_gpu[1], cf))
            cpu_sec += sec
            cpu_ops += n_batch
            b += 1
        ex.close()
        cpu = {"value": round(cpu_ops / cpu_sec / 1e6, 2), "unit": "Mops/s", "cores": threads,
               "kind": "port",
               "sample": f"first {b} batches of the same workload ({cpu_ops} ops) on a "
                         f"{threads}-thread key-range-partitioned CPU B+tree (persistent "
                         f"worker threads)",
               "what": "this repository's CPU restatement of the batch semantics "
                       "(oracle/cpu_baseline.cpp); the reference has no CPU path",
               "host": host}

    def summary(r):
        x = res[r]
        s_ = x["st"]
        return {"mops": round(x["mops"], 2), "ms_per_batch": round(x["elapsed"] / args.steps * 1e3, 3),
                "p50_batch_ms": round(float(np.percentile(x["lat"], 50)) * 1e3, 3),
                "p99_batch_ms": round(float(np.percentile(x["lat"], 99)) * 1e3, 3),
                "phase_ms_per_batch": {k: round(v / args.steps, 4) for k, v in s_["phase_ms"].items()
                                       if s_["phase_launches"][k]},
                "model_window_hits": s_["route_hits"], "model_window_misses": s_["route_misses"],
                "fence_window_hit_rate": (round(x["info"]["window_hits"] / max(1, x["info"]["leaves"]), 4)
                                          if x["info"] else None)}

    uneven = None
    serving = None
    bk.free(); bv.free(); bo.free(); out_v.free(); out_f.free()
    small = small_batch_routes() if args.small_batches and args.route == "both" else None
    plain = run_plain(args, ea, w) if args.plain_trials > 0 else None
    contended = run_contended(args, ea, w) if args.contended_batches > 0 else None
    if args.uneven_batches > 0:
        uneven = run_uneven(args, ea, w)
    if args.serve_requests > 0:
        serving = run_serving(args, ea, w)

    out = {
        "metric": METRIC, "value": round(mops, 2), "unit": "Mops/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"learned_gamma: bulk-load 2^{args.load_log2} Gamma(k=2, theta=2^56) "
                               f"keys, {args.steps} batches of 2^{args.batch_log2} ops, 95% lookup "
                               f"(90% hits) / 5% insert, router retrained every 8 batches",
                   "global_batch": n_batch, "value_route": head_name,
                   "value_api": ("stm_run_batches_device: the timed batches as one device-resident "
                                 "stream (graph replay)" if head_name in streams else
                                 "stm_run_batch_device: one synchronous call per batch"),
                   "defaults": {"fanout": args.fanout, "lanes_per_op": args.lanes,
                                "ops_per_group": args.group},
                   "tuned": tuned,  # the launch parameters of descent_tuned and learned_tuned
                   "tuning_record": record_name(args.tuning_headline) if tuned else None,
                   "parallelism": "single GPU"},
        # per synchronous call of the value route (stm_run_batch_device, one batch per call)
        "latency_ms": {"p50_batch": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99_batch": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "routes": {r: summary(r) for r in res},  # one synchronous call per batch
        "streams": {r: {"mops": round(x["mops"], 2), "runs_mops": x["runs_mops"],
                        "ms_per_batch": round(x["elapsed"] / args.steps * 1e3, 3),
                        "last_batches_identical_to_sync": x["identical_to_sync"]}
                    for r, x in streams.items()},
        "learned_over_plain_stream": (round(streams[head_name]["mops"] /
                                            streams[head_name.replace("learned", "descent")]["mops"], 4)
                                      if head_name.replace("learned", "descent") in streams else None),
        # the tuning record's launch parameters over the defaults, both on the learned stream
        "tuned_over_default_stream": (round(streams["learned_tuned"]["mops"] / streams["learned"]["mops"], 4)
                                      if "learned_tuned" in streams and "learned" in streams else None),
        "small_batches": small,
        # learned over plain descent at the same launch parameters (defaults, tuned)
        "learned_over_plain": {("descent" + k): round(res["learned" + k]["mops"] / res["descent" + k]["mops"], 4)
                               for k in ("", "_tuned") if "learned" + k in res and "descent" + k in res},
        # the same leaf directory searched without the model (a rebuilt static index)
        "learned_over_index": (round(res["learned"]["mops"] / res["index"]["mops"], 4)
                               if "learned" in res and "index" in res else None),
        "learned_and_plain_results_identical": same,
        "stm": {k: st[k] for k in ("upserts", "commits", "aborts", "leaf_splits", "inner_splits",
                                   "root_splits", "deferred", "n_keys", "height", "node_reads")},
        "roofline": roofline, "cpu_baseline": cpu, "cpu_gpu_first_batch_agree": agree,
        "plain": plain,
        "contended": contended,
        "uneven": uneven,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
