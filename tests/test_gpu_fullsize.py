"""GPU parity at the full sizes of the five benchmark workloads (bit-exact).

Each test runs one workload at the size, key distribution, op mix and batch shape the
benchmark quotes it at, through the C ABI, and compares every op's (found, value) and the
final in-order dump with a CPU checker fed the same seeded inputs:

  plain uniform     2^23 uniform keys, one 2^22-op 95/5 batch, host-buffer API
                    (stm_bulk_load / stm_run_batch): checker Oracle (std::map replay)
  contended Zipf    2^24 uniform keys, 8 batches of 2^22 ops, 50/50, Zipf(0.99) ranks with
                    inserts next to hot keys, device API: checker CpuExecutor; aborts > 0
  uneven stream     2^26 Normal(2^62, 2^58) keys, 256 batches log-uniform in [2^10, 2^22]
                    as one stream replayed from captured graphs (every size class up to 2^22)
  learned Gamma     2^26 Gamma(2, 2^56) keys, 64 batches of 2^22 ops with learned routing
                    retrained every 8 batches: 8 synchronous calls, then one stream
  serving loop      2^25 uniform keys, open-loop Poisson arrivals at 200 Mops/s, requests of
                    1..2^16 ops, 90/8/2 with range scans: every request's answers and scan
                    pairs against Oracle's replay of the requests in admission order

CpuExecutor (the key-range-partitioned CPU B+tree) is pinned to Oracle answer by answer and
dump by dump in tests/test_oracle.py.  Generated insert values are splitmix64(key ^ salt of
the batch's seed), so they differ from the stored splitmix64(key) and an update that returned
the new value instead of the old one would show.  Reference sizes: evalutaion/random_testall_v38.sh:9
loops over trees of 2^23..2^26 keys.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def log(msg):
    print(f"[fullsize] {msg}", flush=True)


def assert_same(gv, gf, wv, wf, what):
    bad = np.nonzero((gf != wf) | (gv != wv))[0]
    assert bad.size == 0, (f"{what}: {bad.size} mismatches, first at {bad[:5].tolist()}: got "
                           f"{[(int(gf[i]), int(gv[i])) for i in bad[:3]]} want "
                           f"{[(int(wf[i]), int(wv[i])) for i in bad[:3]]}")


def assert_dump(t, want_k, want_v):
    k, v = t.dump()
    assert k.size == want_k.size, (k.size, want_k.size)
    assert np.array_equal(k, want_k), "dump keys differ"
    assert np.array_equal(v, want_v), "dump values differ"
    assert t.check() == 0


def gen_batches(srt, n_batches, n, seed, dist=0, p0=0.0, p1=0.0):
    from eirene_amd import workloads as w
    keys = np.empty(n_batches * n, dtype=np.uint64)
    vals = np.empty_like(keys)
    ops = np.empty(n_batches * n, dtype=np.uint8)
    for b in range(n_batches):
        s = slice(b * n, (b + 1) * n)
        keys[s], vals[s], ops[s] = w.mixed_batch(srt, n, seed + b, 0.95, 0.9, dist, p0, p1)
    return keys, vals, ops


@pytest.mark.timeout(900)
def test_plain_uniform_full_size(oracle_mod):
    from eirene_amd import StmConfig, StmTree
    from eirene_amd import workloads as w
    t0 = time.time()
    load = w.uniform_keys(1 << 23, 101)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    keys, vals, ops = w.mixed_batch(srt, 1 << 22, 102, 0.95, 0.9)
    o = oracle_mod.Oracle()
    o.bulk_load(srt, w.values_for(srt))
    wv, wf = o.run_batch(keys, vals, ops)
    log(f"plain: oracle ready in {time.time() - t0:.1f}s")
    with StmTree(StmConfig(max_keys=(1 << 23) + (1 << 22), max_batch=1 << 22)) as t:
        t.bulk_load(load, lv)
        gv, gf = t.run_batch(keys, vals, ops)
        assert_same(gv, gf, wv, wf, "plain uniform batch")
        ek, ev = o.dump()
        assert_dump(t, ek, ev)
        st = t.stats()
        assert st["upserts"] > 0 and st["n_keys"] == ek.size
    log(f"plain: done in {time.time() - t0:.1f}s")


@pytest.mark.timeout(900)
def test_contended_zipf_full_size(oracle_mod):
    from eirene_amd import DeviceBuffer, StmConfig, StmTree
    from eirene_amd import workloads as w
    t0 = time.time()
    n_load, n = 1 << 24, 1 << 22
    load = w.uniform_keys(n_load, 201)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    ex = oracle_mod.CpuExecutor(srt, w.values_for(srt), THREADS)
    with StmTree(StmConfig(max_keys=n_load + 8 * n, max_batch=n)) as t:
        dk, dv = DeviceBuffer.from_array(load), DeviceBuffer.from_array(lv)
        t.bulk_load_device(dk, dv)
        dk.free()
        dv.free()
        ov, of = DeviceBuffer(n, np.uint64), DeviceBuffer(n, np.uint8)
        for b in range(8):
            keys, vals, ops = w.zipf_batch(srt, n, 210 + b, 0.99, 0.5)
            wv, wf, _ = ex.run_batch(keys, vals, ops)
            bk, bv, bo = (DeviceBuffer.from_array(a) for a in (keys, vals, ops))
            t.run_batch_device(bk, bv, bo, ov, of)
            for x in (bk, bv, bo):
                x.free()
            assert_same(ov.download(), of.download(), wv, wf, f"zipf batch {b}")
            assert t.check() == 0
            log(f"zipf: batch {b} ok ({time.time() - t0:.1f}s)")
        st = t.stats()
        log(f"zipf: upserts {st['upserts']} aborts {st['aborts']} deferred {st['deferred']} "
            f"leaf splits {st['leaf_splits']}")
        assert st["upserts"] > 0 and st["leaf_splits"] > 0
        assert st["aborts"] > 0  # hot leaves make transactions collide
        ek, ev = ex.dump()
        assert_dump(t, ek, ev)
    ex.close()


@pytest.mark.timeout(900)
def test_uneven_normal_stream_full_size(oracle_mod):
    from eirene_amd import DeviceBuffer, StmConfig, StmTree
    from eirene_amd import autotune as at
    from eirene_amd import workloads as w
    t0 = time.time()
    n_load = 1 << 26
    load = w.normal_keys(n_load, 301)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    keys, vals, ops, offsets = at.uneven_sample(srt, 256, 302, 10, 22, w.DIST_NORMAL,
                                                2.0 ** 62, 2.0 ** 58)
    n_ops = int(offsets[-1])
    ex = oracle_mod.CpuExecutor(srt, w.values_for(srt), THREADS)
    wv, wf, _ = ex.run_batches(keys, vals, ops, offsets)
    log(f"uneven: {n_ops} ops, checker done in {time.time() - t0:.1f}s")
    # The launch shape the autotuner picks for this stream (fanout 32, 16 lanes per op, 8 ops
    # per group, 2048-key sort tiles), all graph size classes up to 2^22 captured first.
    cfg = StmConfig(fanout=32, lanes_per_op=16, ops_per_group=8, sort_tile_items=8,
                    max_keys=n_load + n_ops + (1 << 20), max_batch=1 << 22, use_graphs=True)
    with StmTree(cfg) as t:
        dk, dv = DeviceBuffer.from_array(load), DeviceBuffer.from_array(lv)
        t.bulk_load_device(dk, dv)
        dk.free()
        dv.free()
        t.capture_graphs()
        bk, bv, bo = (DeviceBuffer.from_array(a) for a in (keys, vals, ops))
        ov, of = DeviceBuffer(n_ops, np.uint64), DeviceBuffer(n_ops, np.uint8)
        t.run_batches_device(bk, bv, bo, ov, of, offsets)
        assert_same(ov.download(), of.download(), wv, wf, "uneven stream")
        for x in (bk, bv, bo, ov, of):
            x.free()
        ek, ev = ex.dump()
        assert_dump(t, ek, ev)
    ex.close()
    log(f"uneven: done in {time.time() - t0:.1f}s")


@pytest.mark.timeout(900)
def test_learned_gamma_full_size(oracle_mod):
    from eirene_amd import ROUTE_LEARNED, DeviceBuffer, StmConfig, StmTree
    from eirene_amd import workloads as w
    t0 = time.time()
    n_load, n, nb = 1 << 26, 1 << 22, 64
    load = w.gamma_keys(n_load, 401, 2.0, 2.0 ** 56)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    keys, vals, ops = gen_batches(srt, nb, n, 402 * 1000, w.DIST_GAMMA, 2.0, 2.0 ** 56)
    offsets = np.arange(nb + 1, dtype=np.uint64) * np.uint64(n)
    ex = oracle_mod.CpuExecutor(srt, w.values_for(srt), THREADS)
    wv, wf, _ = ex.run_batches(keys, vals, ops, offsets)
    log(f"learned: checker done in {time.time() - t0:.1f}s")
    # The benchmark's value route: learned routing at the headline tuning record's launch
    # parameters (fanout 64, 8 lanes per op, 16 ops per group), retrained every 8 batches.
    cfg = StmConfig(fanout=64, lanes_per_op=8, ops_per_group=16, route=ROUTE_LEARNED,
                    retrain_every=8, max_keys=n_load + nb * n + (1 << 20), max_batch=n)
    with StmTree(cfg) as t:
        dk, dv = DeviceBuffer.from_array(load), DeviceBuffer.from_array(lv)
        t.bulk_load_device(dk, dv)
        dk.free()
        dv.free()
        bk, bv, bo = (DeviceBuffer.from_array(a) for a in (keys, vals, ops))
        ov, of = DeviceBuffer(keys.size, np.uint64), DeviceBuffer(keys.size, np.uint8)
        for b in range(8):  # one synchronous call per batch
            t.run_batch_device(bk, bv, bo, ov, of, n=n, offset=b * n)
        t.capture_graphs()
        t.run_batches_device(bk, bv, bo, ov, of, offsets[8:])  # the rest as one stream
        assert_same(ov.download(), of.download(), wv, wf, "learned batches")
        for x in (bk, bv, bo, ov, of):
            x.free()
        st = t.stats()
        log(f"learned: model window hits {st['route_hits']} misses {st['route_misses']}, "
            f"leaf splits {st['leaf_splits']}")
        assert st["route_hits"] > 0 and st["leaf_splits"] > 0
        ek, ev = ex.dump()
        assert_dump(t, ek, ev)
    ex.close()
    log(f"learned: done in {time.time() - t0:.1f}s")


def scan_pair_mask(vals, ops, written):
    """Positions of a request's scan area that its scans wrote: scan op i owns the next L_i
    slots (L_i = vals[i]) and writes the first `written[i]` of them."""
    sc = ops == 3
    L = vals[sc].astype(np.int64)
    got = written[sc].astype(np.int64)
    total = int(L.sum())
    mask = np.zeros(total, dtype=bool)
    if total:
        starts = np.concatenate([[0], np.cumsum(L)[:-1]])
        within = np.arange(total) - np.repeat(starts, L)
        mask = within < np.repeat(got, L)
    return mask


@pytest.mark.timeout(900)
def test_serving_open_loop_full_size(oracle_mod):
    from eirene_amd import DeviceBuffer, StmConfig, StmTree
    from eirene_amd import workloads as w
    from eirene_amd.serving import open_loop
    t0 = time.time()
    n_load = 1 << 25
    load = w.uniform_keys(n_load, 501)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    with StmTree(StmConfig(max_keys=n_load + (1 << 24), max_batch=1 << 22)) as t:
        dk, dv = DeviceBuffer.from_array(load), DeviceBuffer.from_array(lv)
        t.bulk_load_device(dk, dv)
        dk.free()
        dv.free()
        rep = open_loop(t, srt, requests=800, offered_mops=200.0, ops_log2=(0, 16),
                        lookup_frac=0.90, insert_frac=0.08, scan_len=(16, 256), seed=502,
                        warmup=200, record=True)
        rec = rep["record"]
        log(f"serving: {rec['tickets'].size} requests, "
            f"{rep['batches']} batches, p99 {rep['p99_ms']:.2f} ms ({time.time() - t0:.1f}s)")
        assert rep["failed"] == 0
        o = oracle_mod.Oracle()
        o.bulk_load(srt, w.values_for(srt))
        ro, po = rec["req_offsets"], rec["pair_offsets"]
        tickets = rec["tickets"]
        assert np.unique(tickets).size == tickets.size
        n_scans = 0
        for r in np.argsort(tickets, kind="stable"):  # admission order
            a, z = int(ro[r]), int(ro[r + 1])
            k, v, op = rec["keys"][a:z], rec["vals"][a:z], rec["ops"][a:z]
            wv, wf, sk, sv = o.run_batch_scan(k, v, op)
            assert_same(rec["out_vals"][a:z], rec["out_found"][a:z], wv, wf, f"request {r}")
            m = scan_pair_mask(v, op, wv)
            pa, pz = int(po[r]), int(po[r + 1])
            assert pz - pa == m.size
            assert np.array_equal(rec["scan_keys"][pa:pz][m], sk[m]), f"request {r} scan keys"
            assert np.array_equal(rec["scan_vals"][pa:pz][m], sv[m]), f"request {r} scan values"
            n_scans += int((op == 3).sum())
        assert n_scans > 0
        ek, ev = o.dump()
        assert_dump(t, ek, ev)
    log(f"serving: {n_scans} scans checked, done in {time.time() - t0:.1f}s")
