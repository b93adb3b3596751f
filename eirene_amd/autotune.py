"""Autotuner of the batch path's launch parameters, measured on the device it runs on.

Search space (every point is a valid StmConfig):
    fanout               node slots                      32, 64, 128
    lanes_per_op         lanes that search one node      8, 16, 32 (at most fanout)
    ops_per_group        sorted ops walked per lane group 8, 16, 32
    sort_tile_items      radix-sort elements per thread  8, 16  (tiles of 2048 / 4096 keys)
    upsert_ops_per_wave  batch slots per upsert wavefront 256, 512, 1024
    upsert_lanes         lanes per insert on the upsert fast path 8, 16, 32 (at most fanout)

Method: for each fanout (the tree must be rebuilt for it), coordinate descent over the other
parameters from the library defaults, two sweeps, keeping a change only when it is faster by
more than `min_gain`.  A trial bulk-loads a fresh tree from device-resident keys and replays the
sample stream once untimed (first touch), then times the stream on fresh trees loaded from the
same keys, so every trial starts from the same tree.  Each timed tree captures its graphs and
(with learned routing) trains its router before the clock starts; throughput is ops / wall
time of stm_run_batches_device, which returns after the stream has completed, and a trial's
figure is the median of three such timed streams.

After the search the defaults and the best configuration are timed alternately (`final`:
three replays each, one at a time), since the search's own figures come from trials minutes
apart.  The result is a JSON tuning file:
    {"device": ..., "workload": ..., "best": {<StmConfig fields>}, "best_mops": ...,
     "default_mops": ..., "final": {"default_median": ..., "best_median": ..., ...},
     "trials": [{"config": {...}, "mops": ...}, ...]}
read back with `load_tuning()`.

    python -m eirene_amd.autotune --out tuning.json [--keys-log2 22] [--batches 48]
        [--route learned --retrain-every 8]
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import asdict
from pathlib import Path
from typing import Callable, Dict, List, Optional

import numpy as np

from . import DeviceBuffer, ROUTE_DESCENT, ROUTE_LEARNED, StmConfig, StmTree

SPACE = {
    "lanes_per_op": [8, 16, 32],
    "ops_per_group": [8, 16, 32],
    "sort_tile_items": [8, 16],
    "upsert_ops_per_wave": [256, 512, 1024],
    "upsert_lanes": [8, 16, 32],
}
FANOUTS = [32, 64, 128]
TUNED_FIELDS = ["fanout"] + list(SPACE)


class Sample:
    """A device-resident workload sample: bulk-load pairs plus a stream of batches."""

    def __init__(self, load_keys, load_vals, keys, vals, ops, offsets):
        self.n_load = int(load_keys.size)
        self.load_k = DeviceBuffer.from_array(np.ascontiguousarray(load_keys, dtype=np.uint64))
        self.load_v = DeviceBuffer.from_array(np.ascontiguousarray(load_vals, dtype=np.uint64))
        self.k = DeviceBuffer.from_array(np.ascontiguousarray(keys, dtype=np.uint64))
        self.v = DeviceBuffer.from_array(np.ascontiguousarray(vals, dtype=np.uint64))
        self.o = DeviceBuffer.from_array(np.ascontiguousarray(ops, dtype=np.uint8))
        self.out_v = DeviceBuffer(int(keys.size), np.uint64)
        self.out_f = DeviceBuffer(int(keys.size), np.uint8)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.n_ops = int(self.offsets[-1] - self.offsets[0])

    def free(self) -> None:
        for b in (self.load_k, self.load_v, self.k, self.v, self.o, self.out_v, self.out_f):
            b.free()


def _trial(sample: Sample, cfg: StmConfig, repeats: int = 3) -> float:
    """Mops/s of the sample stream under cfg: the median of `repeats` timed replays, each on a
    fresh tree bulk-loaded from the same keys (so every trial starts from the same tree)."""
    runs: List[float] = []
    routed = cfg.route != ROUTE_DESCENT
    with StmTree(cfg) as t:
        t.bulk_load_device(sample.load_k, sample.load_v, sample.n_load)
        # untimed pass: first-touch costs
        t.run_batches_device(sample.k, sample.v, sample.o, sample.out_v, sample.out_f,
                             sample.offsets)
    for _ in range(repeats):
        with StmTree(cfg) as t:
            t.bulk_load_device(sample.load_k, sample.load_v, sample.n_load)
            if routed:  # the first router build trains from scratch: outside the timed stream
                t.retrain()
            t.capture_graphs()  # graphs belong to the tree: capture them before the clock starts
            t.sync()
            t0 = time.perf_counter()
            t.run_batches_device(sample.k, sample.v, sample.o, sample.out_v, sample.out_f,
                                 sample.offsets)
            dt = time.perf_counter() - t0
        runs.append(sample.n_ops / dt / 1e6)
    return float(np.median(runs))


def tune(sample: Sample, base: StmConfig, fanouts=FANOUTS, sweeps: int = 2,
         min_gain: float = 0.01, log: Optional[Callable[[str], None]] = None,
         final_rounds: int = 3) -> Dict:
    """Coordinate descent per fanout; returns the tuning record (see module doc)."""
    log = log or (lambda s: None)
    trials: List[Dict] = []
    seen: Dict[tuple, float] = {}

    def measure(c: StmConfig) -> float:
        key = tuple(getattr(c, f) for f in TUNED_FIELDS)
        if key not in seen:
            seen[key] = _trial(sample, c)
            trials.append({"config": {f: getattr(c, f) for f in TUNED_FIELDS},
                           "mops": round(seen[key], 2)})
            log(f"trial {dict(zip(TUNED_FIELDS, key))}: {seen[key]:.1f} Mops/s")
        return seen[key]

    default_mops = measure(base)
    best_cfg, best_mops = base, default_mops
    for F in fanouts:
        cur = StmConfig(**{**asdict(base), "fanout": F,
                           "lanes_per_op": min(base.lanes_per_op, F),
                           "upsert_lanes": min(base.upsert_lanes, F)})
        cur_mops = measure(cur)
        for _ in range(sweeps):
            changed = False
            for field, values in SPACE.items():
                for v in values:
                    if v == getattr(cur, field):
                        continue
                    if field in ("lanes_per_op", "upsert_lanes") and v > F:
                        continue
                    cand = StmConfig(**{**asdict(cur), field: v})
                    m = measure(cand)
                    if m > cur_mops * (1.0 + min_gain):
                        cur, cur_mops, changed = cand, m, True
            if not changed:
                break
        if cur_mops > best_mops:
            best_cfg, best_mops = cur, cur_mops
    # The search's figures come from trials minutes apart (the defaults' trial is the process's
    # first): the final comparison times the defaults and the best alternately, one replay each,
    # `final_rounds` times, and records both medians.
    final = {"default": [], "best": []}
    for _ in range(final_rounds):
        final["default"].append(_trial(sample, base, repeats=1))
        final["best"].append(_trial(sample, best_cfg, repeats=1))
    log(f"final: defaults {final['default']} best {final['best']}")
    return {
        "best": {f: getattr(best_cfg, f) for f in TUNED_FIELDS},
        "best_mops": round(best_mops, 2),
        "default": {f: getattr(base, f) for f in TUNED_FIELDS},
        "default_mops": round(default_mops, 2),
        "final": {"rounds": final_rounds,
                  "default_mops": [round(x, 2) for x in final["default"]],
                  "best_mops": [round(x, 2) for x in final["best"]],
                  "default_median": round(float(np.median(final["default"])), 2),
                  "best_median": round(float(np.median(final["best"])), 2)},
        "trials": trials,
    }


def apply_tuning(cfg: StmConfig, best: Dict) -> StmConfig:
    """cfg with the tuned fields of a tuning record's "best" entry."""
    return StmConfig(**{**asdict(cfg), **{k: int(v) for k, v in best.items() if k in TUNED_FIELDS}})


def load_tuning(path) -> Dict:
    with open(path) as f:
        rec = json.load(f)
    if "best" not in rec:
        raise ValueError(f"{path}: not a tuning record")
    return rec


def uneven_sample(load_sorted, n_batches: int, seed: int, lo_exp: float, hi_exp: float,
                  dist: int, p0: float, p1: float, lookup_frac: float = 0.95):
    """A stream of n_batches mixed batches with sizes log-uniform in [2^lo_exp, 2^hi_exp]."""
    from . import workloads as w
    sizes = w.batch_sizes(n_batches, seed, lo_exp, hi_exp)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(offsets[-1])
    keys = np.empty(n, dtype=np.uint64)
    vals = np.empty(n, dtype=np.uint64)
    ops = np.empty(n, dtype=np.uint8)
    for b in range(n_batches):
        a, z = int(offsets[b]), int(offsets[b + 1])
        keys[a:z], vals[a:z], ops[a:z] = w.mixed_batch(load_sorted, z - a, seed * 7919 + b,
                                                       lookup_frac, 0.9, dist, p0, p1)
    return keys, vals, ops, offsets


def main() -> None:
    from . import workloads as w
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--keys-log2", type=int, default=24)
    ap.add_argument("--dist", choices=["normal", "gamma"], default="normal")
    ap.add_argument("--batches", type=int, default=48)
    ap.add_argument("--lo", type=float, default=10)
    ap.add_argument("--hi", type=float, default=22)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--route", choices=["descent", "learned"], default="descent",
                    help="leaf routing the trials run with")
    ap.add_argument("--retrain-every", type=int, default=8,
                    help="batches between router retrains (learned route)")
    args = ap.parse_args()
    n = 1 << args.keys_log2
    if args.dist == "gamma":
        dist, p0, p1, dname = w.DIST_GAMMA, 2.0, 2.0 ** 56, "Gamma(k=2, theta=2^56)"
        load = w.gamma_keys(n, args.seed, p0, p1)
    else:
        dist, p0, p1, dname = w.DIST_NORMAL, 2.0 ** 62, 2.0 ** 58, "Normal(2^62, 2^58)"
        load = w.normal_keys(n, args.seed, p0, p1)
    srt = w.sorted_copy(load)
    keys, vals, ops, offsets = uneven_sample(srt, args.batches, args.seed, args.lo, args.hi,
                                             dist, p0, p1)
    sample = Sample(load, w.values_for(load), keys, vals, ops, offsets)
    base = StmConfig(max_keys=n + int(offsets[-1]) + (1 << 20), max_batch=1 << int(args.hi),
                     route=ROUTE_LEARNED if args.route == "learned" else ROUTE_DESCENT,
                     retrain_every=args.retrain_every)
    rec = tune(sample, base, log=print)
    sample.free()
    size = (f"of 2^{args.hi:g} ops" if args.lo == args.hi
            else f"log-uniform in [2^{args.lo:g}, 2^{args.hi:g}]")
    rec["workload"] = f"2^{args.keys_log2} {dname} keys, {args.batches} batches {size}, 95/5 lookup/insert"
    rec["route"] = args.route
    if args.route == "learned":
        rec["retrain_every"] = args.retrain_every
    Path(args.out).write_text(json.dumps(rec, indent=1) + "\n")
    print(json.dumps({k: rec[k] for k in ("best", "best_mops", "default_mops")}))


if __name__ == "__main__":
    main()
