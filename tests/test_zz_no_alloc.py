"""The entry point that allocates nothing (stm_run_batch_device_noalloc) against
stm_run_batch_device on a twin tree: identical answers, dumps and statistics, bit-exact and
equal to the oracle, on two calls in a row (the second smaller, so stale scratch contents of
the first call must not leak into it), plus its size limit and its routing precondition."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def batch(rng, load, n):
    keys = np.concatenate([rng.choice(load, n // 2),
                           rng.integers(0, 1 << 63, size=n - n // 2, dtype=np.uint64)])
    rng.shuffle(keys)
    rep = rng.random(n) < 0.1  # repeated keys give the per-key fold work
    keys[1:][rep[1:]] = keys[:-1][rep[1:]]
    vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    ops = (rng.random(n) < 0.7).astype(np.uint8)  # 30 % inserts
    return keys, vals, ops


@pytest.mark.parametrize("route", [0, 1])
def test_noalloc_matches_device_entry_point(oracle_mod, route):
    from eirene_amd import NOALLOC_MAX_OPS, DeviceBuffer, StmConfig, StmError, StmTree
    rng = np.random.default_rng(77 + route)
    load = rng.integers(0, 1 << 63, size=200000, dtype=np.uint64)
    lv = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    cfg = StmConfig(max_keys=1 << 20, max_batch=1 << 16, route=route, train_steps=40)
    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    with StmTree(cfg) as a, StmTree(cfg) as b:
        a.bulk_load(load, lv)
        b.bulk_load(load, lv)
        first = True
        for n in (NOALLOC_MAX_OPS, 20000):
            keys, vals, ops = batch(rng, load, n)
            want_v, want_f = o.run_batch(keys, vals, ops)
            dk, dv, do = (DeviceBuffer.from_array(x) for x in (keys, vals, ops))
            ova, ofa = DeviceBuffer(n, np.uint64), DeviceBuffer(n, np.uint8)
            ovb, ofb = DeviceBuffer(n, np.uint64), DeviceBuffer(n, np.uint8)
            if route and first:
                # learned routing: the router's buffers come from the first training
                with pytest.raises(StmError) as e:
                    b.run_batch_device_noalloc(dk, dv, do, ovb, ofb)
                assert e.value.code == 4
                b.retrain()
            first = False
            a.run_batch_device(dk, dv, do, ova, ofa)
            b.run_batch_device_noalloc(dk, dv, do, ovb, ofb)
            ga, fa = ova.download(), ofa.download()
            gb, fb = ovb.download(), ofb.download()
            assert np.array_equal(fa, fb) and np.array_equal(ga, gb), f"n={n}: entry points differ"
            assert np.array_equal(fb, want_f) and np.array_equal(gb, want_v), f"n={n}: oracle"
            for x in (dk, dv, do, ova, ofa, ovb, ofb):
                x.free()
        ka, va = a.dump()
        kb, vb = b.dump()
        ek, ev = o.dump()
        assert np.array_equal(ka, kb) and np.array_equal(va, vb)
        assert np.array_equal(kb, ek) and np.array_equal(vb, ev)
        assert a.check() == 0 and b.check() == 0
        assert b.stats()["upserts"] == a.stats()["upserts"]
        # the size limit is an error, not a fault
        n = NOALLOC_MAX_OPS + 1
        keys, vals, ops = batch(rng, load, n)
        dk, dv, do = (DeviceBuffer.from_array(x) for x in (keys, vals, ops))
        ov, of = DeviceBuffer(n, np.uint64), DeviceBuffer(n, np.uint8)
        with pytest.raises(StmError) as e:
            b.run_batch_device_noalloc(dk, dv, do, ov, of)
        assert e.value.code == 1
        for x in (dk, dv, do, ov, of):
            x.free()
