"""The continuous-batching serving loop: each request's result is the same alone and in the mix.

Requests are submitted while earlier ones run, so batches mix many requests; every request's
answers and scan pairs must equal running the requests one at a time, in submission order, on
a second tree (stm_run_batch_scan), and the CPU oracle's sequential replay.  Exact equality is
the stated tolerance (integer keys and values).
"""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_requests(seed, load, count, max_log2=12, scan_frac=0.05):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        n = int(np.exp(rng.uniform(0, np.log(1 << max_log2))))
        n = max(n, 1)
        u = rng.random(n)
        ops = np.where(u < 0.7, 1, np.where(u < 1 - scan_frac, 0, 3)).astype(np.uint8)
        keys = np.where(rng.random(n) < 0.5, rng.choice(load, n),
                        rng.integers(0, 1 << 40, size=n, dtype=np.uint64)).astype(np.uint64)
        vals = rng.integers(0, 1 << 62, size=n, dtype=np.uint64)
        s = ops == 3
        vals[s] = rng.integers(16, 257, size=int(s.sum()))
        reqs.append((keys, vals, ops))
    return reqs


def new_tree(load, **kw):
    from eirene_amd import StmConfig, StmTree
    cfg = dict(max_keys=1 << 20, max_batch=1 << 16)
    cfg.update(kw)
    t = StmTree(StmConfig(**cfg))
    t.bulk_load(load, load * np.uint64(3))
    return t


def serve_all(t, reqs, pace_s=0.0, **srv_kw):
    """Submits from a client thread (optionally paced) while the caller waits in order."""
    from eirene_amd.serving import Server
    results = [None] * len(reqs)
    with Server(t, **srv_kw) as srv:
        tickets = [None] * len(reqs)
        ready = [threading.Event() for _ in reqs]

        def client():
            for i, (k, v, o) in enumerate(reqs):
                tickets[i] = srv.submit(k, v, o)
                ready[i].set()
                if pace_s:
                    time.sleep(pace_s)

        th = threading.Thread(target=client)
        th.start()
        for i in range(len(reqs)):
            ready[i].wait()
            results[i] = srv.wait(tickets[i])
        th.join()
        stats = srv.stats()
    return results, stats


def check_alone(results, reqs, load, oracle_mod, **tree_kw):
    o = oracle_mod.Oracle()
    o.bulk_load(load, load * np.uint64(3))
    with new_tree(load, **tree_kw) as alone:
        for i, ((k, v, op), got) in enumerate(zip(reqs, results)):
            a = alone.run_batch_scan(k, v, op)
            e = o.run_batch_scan(k, v, op)
            for x, y, z, what in zip(got[:2], a[:2], e[:2], ("values", "found")):
                np.testing.assert_array_equal(x, y, err_msg=f"request {i} {what}: mix vs alone")
                np.testing.assert_array_equal(x, z, err_msg=f"request {i} {what}: mix vs oracle")
            # scan pairs: the written part of every scan (unwritten slots are unspecified)
            s = op == 3
            off = np.concatenate([[0], np.cumsum(v[s].astype(np.int64))])
            for j, pos in enumerate(np.nonzero(s)[0]):
                c, b = int(got[0][pos]), int(off[j])
                for q in (2, 3):
                    np.testing.assert_array_equal(got[q][b:b + c], a[q][b:b + c])
                    np.testing.assert_array_equal(got[q][b:b + c], e[q][b:b + c])
        return alone.dump(), o.dump()


def test_same_alone_and_in_mix(oracle_mod):
    rng = np.random.default_rng(1)
    load = np.unique(rng.integers(0, 1 << 40, size=40000, dtype=np.uint64))
    reqs = make_requests(2, load, 300)
    with new_tree(load) as t:
        results, stats = serve_all(t, reqs, hbm_budget=64 << 20)
        mixed_dump = t.dump()
    assert stats["requests"] == len(reqs)
    assert stats["batches"] < len(reqs), "requests were never batched together"
    (ak, av), (ok, ov) = check_alone(results, reqs, load, oracle_mod)
    np.testing.assert_array_equal(mixed_dump[0], ak)
    np.testing.assert_array_equal(mixed_dump[1], av)
    np.testing.assert_array_equal(ak, ok)
    np.testing.assert_array_equal(av, ov)


def test_admission_waits_for_memory(oracle_mod):
    """Tiny slots: requests queue for admission, batches stay within a slot, results exact."""
    rng = np.random.default_rng(3)
    load = np.unique(rng.integers(0, 1 << 40, size=20000, dtype=np.uint64))
    reqs = make_requests(4, load, 200, max_log2=11)
    with new_tree(load) as t:
        results, stats = serve_all(t, reqs, hbm_budget=12 << 20, max_batch_ops=2048, slots=2)
    assert stats["slot_ops"] == 2048
    assert stats["batches"] >= sum(k.size for k, _, _ in reqs) // 2048
    assert stats["max_pending"] > 1
    check_alone(results, reqs, load, oracle_mod)


def test_request_larger_than_slot_rejected():
    from eirene_amd import StmError
    from eirene_amd.serving import Server
    load = np.arange(1, 1000, dtype=np.uint64)
    with new_tree(load) as t:
        with Server(t, hbm_budget=12 << 20, max_batch_ops=1024, slots=2) as srv:
            n = 2000
            with pytest.raises(StmError) as e:
                srv.submit(np.arange(n, dtype=np.uint64), np.zeros(n, dtype=np.uint64),
                           np.ones(n, dtype=np.uint8))
            assert e.value.code == 2
            tk = srv.submit(np.arange(10, dtype=np.uint64), np.zeros(10, dtype=np.uint64),
                            np.ones(10, dtype=np.uint8))
            v, f, _, _, lat = srv.wait(tk)
            assert f.tolist() == [0] + [1] * 9 and lat >= 0


def test_open_loop_report():
    from eirene_amd import workloads as w
    from eirene_amd.serving import open_loop
    load = w.uniform_keys(1 << 18, 5)
    srt = w.sorted_copy(load)
    with new_tree(load, max_keys=1 << 21, max_batch=1 << 20) as t:
        rep = open_loop(t, srt, requests=400, offered_mops=20.0, ops_log2=(0, 14), seed=3,
                        hbm_budget=256 << 20, return_latencies=True)
        assert t.check() == 0
    assert rep["failed"] == 0 and rep["requests"] == 400
    assert 0 < rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
    assert rep["batches"] >= 1 and rep["scans"] > 0
    lat = rep["latencies_ms"]
    assert np.all(lat > 0)
