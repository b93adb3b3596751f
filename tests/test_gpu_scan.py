"""GPU parity of range scans (op 3 under stm_run_batch_scan) against the CPU oracle.

SCAN(k, L) at input position i must return exactly the first L pairs with key >= k of the
state before op i: the tree before the batch plus the batch's earlier inserts (the oracle's
sequential replay, oracle_run_batch_scan).  Compared: every op's (value, found), each scan's
written pairs, and the final dump.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tree(**kw):
    from eirene_amd import StmConfig, StmTree
    base = dict(max_keys=1 << 20, max_batch=1 << 16)
    base.update(kw)
    return StmTree(StmConfig(**base))


def check_scan_batch(t, o, keys, vals, ops, what=""):
    v, f, sk, sv = t.run_batch_scan(keys, vals, ops)
    ev, ef, esk, esv = o.run_batch_scan(keys, vals, ops)
    bad = np.nonzero((v != ev) | (f != ef))[0]
    assert bad.size == 0, f"{what}: {bad.size} op mismatches, first {bad[:5].tolist()}"
    scans = np.nonzero(ops == 3)[0]
    off = np.concatenate([[0], np.cumsum(vals[scans].astype(np.int64))])
    for j, i in enumerate(scans):
        a, c = int(off[j]), int(v[i])
        np.testing.assert_array_equal(sk[a:a + c], esk[a:a + c], err_msg=f"{what} scan {i} keys")
        np.testing.assert_array_equal(sv[a:a + c], esv[a:a + c], err_msg=f"{what} scan {i} vals")


def test_scan_known_answers(oracle_mod):
    with tree() as t:
        t.bulk_load(np.array([10, 20, 30], dtype=np.uint64), np.array([1, 2, 3], dtype=np.uint64))
        keys = np.array([15, 25, 15, 40, 0, 31], dtype=np.uint64)
        vals = np.array([2, 9, 3, 1, 5, 2], dtype=np.uint64)
        ops = np.array([3, 0, 3, 3, 3, 3], dtype=np.uint8)
        v, f, sk, sv = t.run_batch_scan(keys, vals, ops)
        assert v.tolist() == [2, 0, 3, 0, 4, 0] and f.tolist() == [1, 0, 1, 0, 1, 0]
        assert sk[:2].tolist() == [20, 30] and sv[:2].tolist() == [2, 3]
        assert sk[2:5].tolist() == [20, 25, 30] and sv[2:5].tolist() == [2, 9, 3]
        assert sk[6:10].tolist() == [10, 20, 25, 30] and sv[6:10].tolist() == [1, 2, 9, 3]
        k, d = t.dump()
        assert k.tolist() == [10, 20, 25, 30] and d.tolist() == [1, 2, 9, 3]


def test_scan_on_empty_tree(oracle_mod):
    with tree() as t:
        o = oracle_mod.Oracle()
        keys = np.array([5, 7, 5, 1, 6], dtype=np.uint64)
        vals = np.array([3, 70, 4, 50, 2], dtype=np.uint64)
        ops = np.array([3, 0, 3, 0, 3], dtype=np.uint8)
        check_scan_batch(t, o, keys, vals, ops, "empty")


@pytest.mark.parametrize("fanout", [32, 64, 128])
def test_scan_random_batches(oracle_mod, fanout):
    rng = np.random.default_rng(fanout)
    domain = rng.integers(0, 1 << 62, size=60000, dtype=np.uint64)
    load = rng.choice(domain, 20000)
    o = oracle_mod.Oracle()
    o.bulk_load(load, load + 1)
    with tree(fanout=fanout) as t:
        t.bulk_load(load, load + 1)
        for b in range(4):
            n = 20000
            keys = rng.choice(domain, n)
            ops = rng.choice(np.array([0, 1, 3], dtype=np.uint8), n, p=[0.3, 0.6, 0.1])
            vals = rng.integers(0, 1 << 40, size=n, dtype=np.uint64)
            s = ops == 3
            vals[s] = rng.integers(1, 300, size=int(s.sum()))
            check_scan_batch(t, o, keys, vals, ops, f"F={fanout} batch {b}")
            assert t.check() == 0
        k, v = t.dump()
        ek, ev = o.dump()
        np.testing.assert_array_equal(k, ek)
        np.testing.assert_array_equal(v, ev)


def test_scan_dense_batch_over_sparse_tree(oracle_mod):
    """Most scanned pairs come from the batch's own inserts (B stream dominates)."""
    rng = np.random.default_rng(7)
    o = oracle_mod.Oracle()
    load = np.arange(0, 1 << 30, 1 << 20, dtype=np.uint64)
    o.bulk_load(load, load)
    with tree(fanout=64) as t:
        t.bulk_load(load, load)
        n = 30000
        keys = rng.integers(0, 1 << 30, size=n, dtype=np.uint64)
        keys[rng.random(n) < 0.2] = 12345  # a hot key inserted and scanned over and over
        ops = rng.choice(np.array([0, 1, 3], dtype=np.uint8), n, p=[0.6, 0.2, 0.2])
        vals = rng.integers(0, 1 << 40, size=n, dtype=np.uint64)
        s = ops == 3
        vals[s] = rng.integers(1, 2000, size=int(s.sum()))
        check_scan_batch(t, o, keys, vals, ops, "dense")
        assert t.check() == 0


def test_long_scans_cross_many_leaves(oracle_mod):
    rng = np.random.default_rng(11)
    load = rng.integers(0, 1 << 63, size=200000, dtype=np.uint64)
    o = oracle_mod.Oracle()
    o.bulk_load(load, load ^ 5)
    with tree(fanout=32, max_keys=1 << 21) as t:
        t.bulk_load(load, load ^ 5)
        keys = np.concatenate([rng.integers(0, 1 << 63, size=6, dtype=np.uint64),
                               np.array([0, (1 << 63) - 10, 0xFFFFFFFFFFFFFFFE], dtype=np.uint64)])
        vals = np.array([65536, 1, 40000, 7, 65536, 3, 65536, 100, 9], dtype=np.uint64)
        ops = np.full(keys.size, 3, dtype=np.uint8)
        ops[1] = 0
        check_scan_batch(t, o, keys, vals, ops, "long")


def test_scan_device_api_and_chunking(oracle_mod):
    from eirene_amd import DeviceBuffer
    rng = np.random.default_rng(21)
    load = rng.integers(0, 1 << 40, size=30000, dtype=np.uint64)
    n = 25000
    keys = rng.integers(0, 1 << 40, size=n, dtype=np.uint64)
    ops = rng.choice(np.array([0, 1, 3], dtype=np.uint8), n, p=[0.2, 0.7, 0.1])
    vals = rng.integers(0, 1 << 40, size=n, dtype=np.uint64)
    s = ops == 3
    vals[s] = rng.integers(1, 64, size=int(s.sum()))
    o = oracle_mod.Oracle()
    o.bulk_load(load, load)
    ev, ef, esk, esv = o.run_batch_scan(keys, vals, ops)
    with tree(max_batch=4093) as t:  # chunked: scan offsets continue across chunks
        t.bulk_load(load, load)
        bufs = [DeviceBuffer.from_array(a) for a in (keys, vals, ops)]
        ov, of = DeviceBuffer(n, np.uint64), DeviceBuffer(n, np.uint8)
        cap = int(vals[s].sum())
        dsk, dsv = DeviceBuffer(cap, np.uint64), DeviceBuffer(cap, np.uint64)
        t.run_batch_scan_device(*bufs, ov, of, dsk, dsv)
        v, f, sk, sv = ov.download(), of.download(), dsk.download(), dsv.download()
    np.testing.assert_array_equal(v, ev)
    np.testing.assert_array_equal(f, ef)
    off = np.concatenate([[0], np.cumsum(vals[s].astype(np.int64))])
    for j, i in enumerate(np.nonzero(s)[0]):
        a, c = int(off[j]), int(v[i])
        np.testing.assert_array_equal(sk[a:a + c], esk[a:a + c])
        np.testing.assert_array_equal(sv[a:a + c], esv[a:a + c])


def test_scan_bad_length_rejected():
    from eirene_amd import StmError
    with tree() as t:
        for L in (0, 65537):
            with pytest.raises(StmError) as e:
                t.run_batch_scan(np.array([1], dtype=np.uint64), np.array([L], dtype=np.uint64),
                                 np.array([3], dtype=np.uint8))
            assert e.value.code == 1
        assert t.check() == 0
