"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the GPU checker.

Checks, each chosen so that a plausible mistake fails one of them:
  * hand-written known answers (old value returned by INSERT, value seen by a later GET,
    op codes 2/3 acting as GET, last-writer-wins bulk load);
  * agreement of the two restatements (sequential replay and the reference's
    sort/RLE/scan/mergeMission/writeback pipeline) on random and skewed batches;
  * an independent Python dict replay on small random batches (brute force);
  * the committed golden fixtures (written by tests/golden/make_golden.py from the same
    definition, independently of the C++ code);
  * the stable-sort building block against numpy's stable argsort;
  * the multi-threaded CPU baseline executor against the oracle for several T.
"""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def py_replay(state, keys, vals, ops):
    out_v = np.zeros(len(keys), dtype=np.uint64)
    out_f = np.zeros(len(keys), dtype=np.uint8)
    for i, (k, v, o) in enumerate(zip(keys.tolist(), vals.tolist(), ops.tolist())):
        if k in state:
            out_f[i], out_v[i] = 1, state[k]
        if o == 0:
            state[k] = v
    return out_v, out_f


def test_known_answers(oracle_mod):
    for runner in ("run_batch", "pipeline_batch"):
        o = oracle_mod.Oracle()
        keys = np.array([5, 5, 5, 5, 5, 7, 7], dtype=np.uint64)
        vals = np.array([0, 10, 0, 11, 0, 70, 0], dtype=np.uint64)
        ops = np.array([1, 0, 1, 0, 3, 0, 2], dtype=np.uint8)
        v, f = getattr(o, runner)(keys, vals, ops)
        # GET miss, INSERT (no old), GET -> 10, INSERT returns 10, RANGE-as-GET -> 11,
        # INSERT 7 (no old), DELETE-as-GET -> 70
        assert f.tolist() == [0, 0, 1, 1, 1, 0, 1], runner
        assert v.tolist() == [0, 0, 10, 10, 11, 0, 70], runner
        k, d = o.dump()
        assert k.tolist() == [5, 7] and d.tolist() == [11, 70]


def test_bulk_load_last_writer_wins(oracle_mod):
    o = oracle_mod.Oracle()
    o.bulk_load(np.array([3, 1, 3, 2], dtype=np.uint64), np.array([30, 10, 31, 20], dtype=np.uint64))
    k, v = o.dump()
    assert k.tolist() == [1, 2, 3] and v.tolist() == [10, 20, 31]
    assert o.size() == 3


@pytest.mark.parametrize("seed", range(6))
def test_restatements_agree_with_python(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    domain = rng.integers(0, 1 << 63, size=int(rng.integers(5, 300)), dtype=np.uint64)
    load = rng.choice(domain, size=len(domain) // 2)
    load_vals = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    a, b = oracle_mod.Oracle(), oracle_mod.Oracle()
    a.bulk_load(load, load_vals)
    b.bulk_load(load, load_vals)
    state = {}
    for k, v in zip(load.tolist(), load_vals.tolist()):
        state[k] = v
    for _ in range(3):
        n = int(rng.integers(1, 2000))
        keys = rng.choice(domain, size=n)
        vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
        ops = rng.integers(0, 4, size=n).astype(np.uint8)
        ev, ef = py_replay(state, keys, vals, ops)
        v1, f1 = a.run_batch(keys, vals, ops)
        v2, f2 = b.pipeline_batch(keys, vals, ops)
        np.testing.assert_array_equal(f1, ef)
        np.testing.assert_array_equal(v1, ev)
        np.testing.assert_array_equal(f2, ef)
        np.testing.assert_array_equal(v2, ev)
    ka, va = a.dump()
    kb, vb = b.dump()
    np.testing.assert_array_equal(ka, kb)
    np.testing.assert_array_equal(va, vb)
    assert np.all(np.diff(ka.astype(np.float64)) >= 0) and len(np.unique(ka)) == len(ka)
    assert ka.tolist() == sorted(state)


@pytest.mark.parametrize("name", sorted(p.stem for p in GOLDEN.glob("*.npz")))
def test_golden_fixtures(oracle_mod, name):
    z = np.load(GOLDEN / f"{name}.npz")
    for runner in ("run_batch", "pipeline_batch"):
        o = oracle_mod.Oracle()
        o.bulk_load(z["load_keys"], z["load_vals"])
        for b in range(int(z["n_batches"][0])):
            v, f = getattr(o, runner)(z[f"b{b}_keys"], z[f"b{b}_vals"], z[f"b{b}_ops"])
            np.testing.assert_array_equal(f, z[f"b{b}_out_found"])
            np.testing.assert_array_equal(v, z[f"b{b}_out_vals"])
        k, v = o.dump()
        np.testing.assert_array_equal(k, z["dump_keys"])
        np.testing.assert_array_equal(v, z["dump_vals"])


def test_stable_sort_perm(oracle_mod):
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 50, size=5000).astype(np.uint64)  # many ties
    perm = oracle_mod.stable_sort_perm(keys)
    np.testing.assert_array_equal(perm, np.argsort(keys, kind="stable"))
    big = rng.integers(0, 1 << 64, size=5000, dtype=np.uint64, endpoint=False)
    np.testing.assert_array_equal(oracle_mod.stable_sort_perm(big), np.argsort(big, kind="stable"))


@pytest.mark.parametrize("threads", [1, 3, 4])
def test_cpu_executor_matches_oracle(oracle_mod, threads):
    rng = np.random.default_rng(threads)
    load = np.unique(rng.integers(0, 1 << 63, size=20000, dtype=np.uint64))
    load_vals = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    ex = oracle_mod.CpuExecutor(load, load_vals, threads)
    o = oracle_mod.Oracle()
    o.bulk_load(load, load_vals)
    for b in range(4):
        n = 30000
        keys = np.concatenate([rng.choice(load, n // 2),
                               rng.integers(0, 1 << 63, size=n - n // 2, dtype=np.uint64)])
        if b == 3:  # duplicate-heavy batch and keys below every loaded key
            keys = np.concatenate([rng.choice(load[:20], n // 2), rng.integers(0, 1000, n - n // 2).astype(np.uint64)])
        rng.shuffle(keys)
        vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
        ops = (rng.random(n) < 0.5).astype(np.uint8)
        ev, ef = o.run_batch(keys, vals, ops)
        v, f, sec = ex.run_batch(keys, vals, ops)
        np.testing.assert_array_equal(f, ef)
        np.testing.assert_array_equal(v, ev)
        assert sec > 0
    # the in-order dump of the partitions equals the sequential replay's
    k, v = ex.dump()
    ek, ev = o.dump()
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)
    ex.close()


def py_replay_scan(state, keys, vals, ops):
    """Brute force: the sorted key list is rebuilt for every scan."""
    out_v = np.zeros(len(keys), dtype=np.uint64)
    out_f = np.zeros(len(keys), dtype=np.uint8)
    sk, sv = [], []
    for i, (k, v, o) in enumerate(zip(keys.tolist(), vals.tolist(), ops.tolist())):
        if o == 3:
            got = [x for x in sorted(state) if x >= k][:v]
            sk += got + [0] * (v - len(got))
            sv += [state[x] for x in got] + [0] * (v - len(got))
            out_v[i], out_f[i] = len(got), 1 if got else 0
            continue
        if k in state:
            out_f[i], out_v[i] = 1, state[k]
        if o == 0:
            state[k] = v
    return out_v, out_f, np.array(sk, dtype=np.uint64), np.array(sv, dtype=np.uint64)


def test_scan_known_answers(oracle_mod):
    o = oracle_mod.Oracle()
    o.bulk_load(np.array([10, 20, 30], dtype=np.uint64), np.array([1, 2, 3], dtype=np.uint64))
    keys = np.array([15, 25, 15, 40, 0, 31], dtype=np.uint64)
    vals = np.array([2, 9, 3, 1, 5, 2], dtype=np.uint64)
    ops = np.array([3, 0, 3, 3, 3, 3], dtype=np.uint8)
    v, f, sk, sv = o.run_batch_scan(keys, vals, ops)
    # scan(15,2) -> 20,30; insert 25=9; scan(15,3) -> 20,25,30; scan(40,1) -> none;
    # scan(0,5) -> all four; scan(31,2) -> none
    assert v.tolist() == [2, 0, 3, 0, 4, 0] and f.tolist() == [1, 0, 1, 0, 1, 0]
    assert sk.size == 2 + 3 + 1 + 5 + 2
    assert sk[:2].tolist() == [20, 30] and sv[:2].tolist() == [2, 3]
    assert sk[2:5].tolist() == [20, 25, 30] and sv[2:5].tolist() == [2, 9, 3]
    assert sk[6:10].tolist() == [10, 20, 25, 30] and sv[6:10].tolist() == [1, 2, 9, 3]


@pytest.mark.parametrize("seed", range(4))
def test_scan_replay_matches_python(oracle_mod, seed):
    rng = np.random.default_rng(100 + seed)
    domain = rng.integers(0, 1 << 20, size=400, dtype=np.uint64)
    load = rng.choice(domain, 150)
    o = oracle_mod.Oracle()
    o.bulk_load(load, load + 1)
    state = {int(k): int(k) + 1 for k in load}
    for _ in range(3):
        n = 300
        keys = rng.choice(domain, n)
        ops = rng.choice(np.array([0, 1, 3], dtype=np.uint8), n, p=[0.3, 0.5, 0.2])
        vals = rng.integers(0, 1 << 40, size=n, dtype=np.uint64)
        vals[ops == 3] = rng.integers(1, 40, size=int((ops == 3).sum()))
        ev, ef, esk, esv = py_replay_scan(state, keys, vals, ops)
        v, f, sk, sv = o.run_batch_scan(keys, vals, ops)
        np.testing.assert_array_equal(v, ev)
        np.testing.assert_array_equal(f, ef)
        # compare only the written part of every scan's slot range
        off = np.concatenate([[0], np.cumsum(vals[ops == 3])])[:-1].astype(np.int64)
        for j, i in enumerate(np.nonzero(ops == 3)[0]):
            c, a = int(v[i]), int(off[j])
            np.testing.assert_array_equal(sk[a:a + c], esk[a:a + c])
            np.testing.assert_array_equal(sv[a:a + c], esv[a:a + c])
