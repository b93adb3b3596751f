"""GPU parity of the HIP batch path against the CPU oracle (bit-exact).

Every test calls the product through the C ABI (libstmtree.so via ctypes) and compares
each op's (found, value) and the final in-order dump with the oracle's sequential replay
on the same inputs.  The structural checker (stm_check) must report a well-formed tree
with every lock released after each batch.
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def tree(**kw):
    from eirene_amd import StmConfig, StmTree
    base = dict(max_keys=1 << 20, max_batch=1 << 16)
    base.update(kw)
    return StmTree(StmConfig(**base))


def assert_same(got, want, what=""):
    gv, gf = got
    wv, wf = want
    bad = np.nonzero((gf != wf) | (gv != wv))[0]
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5].tolist()}: " \
        f"got {[(int(gf[i]), int(gv[i])) for i in bad[:3]]} want {[(int(wf[i]), int(wv[i])) for i in bad[:3]]}"


def assert_dump(t, o):
    k, v = t.dump()
    ek, ev = o.dump()
    assert k.size == ek.size, (k.size, ek.size)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)
    assert t.size() == ek.size
    assert t.check() == 0


def test_known_answers():
    with tree(fanout=64) as t:
        keys = np.array([5, 5, 5, 5, 5, 7, 7], dtype=np.uint64)
        vals = np.array([0, 10, 0, 11, 0, 70, 0], dtype=np.uint64)
        ops = np.array([1, 0, 1, 0, 3, 0, 2], dtype=np.uint8)
        v, f = t.run_batch(keys, vals, ops)
        assert f.tolist() == [0, 0, 1, 1, 1, 0, 1]
        assert v.tolist() == [0, 0, 10, 10, 11, 0, 70]
        k, d = t.dump()
        assert k.tolist() == [5, 7] and d.tolist() == [11, 70]
        assert t.check() == 0


@pytest.mark.parametrize("fanout", [32, 64, 128])
@pytest.mark.parametrize("name", sorted(p.stem for p in GOLDEN.glob("*.npz")))
def test_golden_fixtures(oracle_mod, name, fanout):
    z = np.load(GOLDEN / f"{name}.npz")
    with tree(fanout=fanout, lanes_per_op=16) as t:
        t.bulk_load(z["load_keys"], z["load_vals"])
        assert t.check() == 0
        for b in range(int(z["n_batches"][0])):
            got = t.run_batch(z[f"b{b}_keys"], z[f"b{b}_vals"], z[f"b{b}_ops"])
            assert_same(got, (z[f"b{b}_out_vals"], z[f"b{b}_out_found"]), f"{name} batch {b}")
            assert t.check() == 0
        k, v = t.dump()
        np.testing.assert_array_equal(k, z["dump_keys"])
        np.testing.assert_array_equal(v, z["dump_vals"])


@pytest.mark.parametrize("fanout,lanes", [(32, 8), (32, 32), (64, 8), (64, 16), (64, 64),
                                          (128, 16), (128, 64)])
def test_random_batches(oracle_mod, fanout, lanes):
    rng = np.random.default_rng(fanout * 100 + lanes)
    domain = rng.integers(0, 1 << 63, size=30000, dtype=np.uint64)
    load = rng.choice(domain, 8000)
    load_vals = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    o = oracle_mod.Oracle()
    o.bulk_load(load, load_vals)
    with tree(fanout=fanout, lanes_per_op=lanes, ops_per_group=int(rng.integers(1, 40))) as t:
        t.bulk_load(load, load_vals)
        for b in range(4):
            n = int(rng.integers(1, 40000))
            keys = rng.choice(domain, n)
            vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
            ops = rng.integers(0, 4, size=n).astype(np.uint8)
            assert_same(t.run_batch(keys, vals, ops), o.run_batch(keys, vals, ops), f"batch {b}")
        assert_dump(t, o)


@pytest.mark.parametrize("fanout", [32, 64, 128])
def test_empty_tree_growth_with_root_splits(oracle_mod, fanout):
    rng = np.random.default_rng(fanout)
    o = oracle_mod.Oracle()
    with tree(fanout=fanout) as t:
        for b in range(3):
            n = 60000
            keys = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
            if b == 2:
                keys[: n // 2] = rng.integers(0, 64, size=n // 2).astype(np.uint64)
            vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
            ops = (rng.random(n) < 0.1).astype(np.uint8)  # 90 % inserts
            assert_same(t.run_batch(keys, vals, ops), o.run_batch(keys, vals, ops), f"batch {b}")
            assert t.check() == 0
        st = t.stats()
        assert st["leaf_splits"] > 0 and st["root_splits"] > 0 and st["height"] >= 3
        # Node reads (the reference's traversal steps): every merged op reads at least its leaf
        # once between the lookup and upsert kernels, and at most a root-to-leaf path per
        # attempt (lookup routing, transactions and their retries).
        assert st["node_reads"] >= st["unique_ops"]
        assert st["node_reads"] <= (st["height"] + 2) * (2 * st["unique_ops"] + st["aborts"])
        assert_dump(t, o)


def test_chunking_does_not_change_results(oracle_mod):
    rng = np.random.default_rng(5)
    load = rng.integers(0, 1 << 20, size=20000).astype(np.uint64)
    vals0 = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    n = 50000
    keys = rng.integers(0, 1 << 20, size=n).astype(np.uint64)
    vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    ops = (rng.random(n) < 0.6).astype(np.uint8)
    outs = []
    for mb in (997, 1 << 16):
        with tree(max_batch=mb) as t:
            t.bulk_load(load, vals0)
            outs.append(t.run_batch(keys, vals, ops) + t.dump())
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)
    o = oracle_mod.Oracle()
    o.bulk_load(load, vals0)
    assert_same(outs[0][:2], o.run_batch(keys, vals, ops))


def test_contended_zipf(oracle_mod):
    from eirene_amd import workloads as w
    load = w.uniform_keys(1 << 18, 21)
    lv = w.values_for(load)
    srt = w.sorted_copy(load)
    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    with tree(max_keys=1 << 21, max_batch=1 << 18) as t:
        t.bulk_load(load, lv)
        for b in range(4):
            keys, vals, ops = w.zipf_batch(srt, 1 << 18, 100 + b, 0.99, 0.5)
            assert_same(t.run_batch(keys, vals, ops), o.run_batch(keys, vals, ops), f"zipf {b}")
            assert t.check() == 0
        st = t.stats()
        # a transaction may commit several upserts of one leaf (merged), so commits <= upserts
        assert st["upserts"] > 0 and st["commits"] > 0
        assert st["aborts"] > 0  # Zipf-hot leaves make transactions collide
        assert_dump(t, o)


@pytest.mark.parametrize("fanout", [32, 64])
def test_forced_aborts_run_the_serial_pass(oracle_mod, fanout):
    """Thousands of distinct new keys packed into the range of a few leaves, one upsert per
    wavefront (upsert_ops_per_wave = 1) and a single attempt per transaction: nearly every
    handed-back insert collides on a leaf lock or a parent split, gives up after its first
    abort and runs in the serial pass (k_upsert_serial, the only forward-progress guarantee).
    Every answer, the dump and the tree's invariants must still equal the sequential replay
    (reference retry loop: src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:107-237)."""
    rng = np.random.default_rng(fanout + 7)
    load = (np.arange(4096, dtype=np.uint64) << np.uint64(40)) + np.uint64(12345)
    lv = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    with tree(fanout=fanout, max_keys=1 << 16, max_batch=1 << 14, max_attempts=1,
              upsert_ops_per_wave=1) as t:
        t.bulk_load(load, lv)
        for b in range(3):
            n = 6000 if b < 2 else 3000
            lo, hi = int(load[100 + 40 * b]), int(load[103 + 40 * b])
            fresh = rng.integers(lo, hi, size=n, dtype=np.uint64)
            keys = np.concatenate([fresh, rng.choice(fresh, n // 4), rng.choice(load, n // 4)])
            rng.shuffle(keys)
            vals = rng.integers(0, 1 << 63, size=keys.size, dtype=np.uint64)
            ops = (rng.random(keys.size) < 0.2).astype(np.uint8)  # 80 % inserts
            assert_same(t.run_batch(keys, vals, ops), o.run_batch(keys, vals, ops), f"batch {b}")
            assert t.check() == 0
        st = t.stats()
        assert st["aborts"] > 0, st
        assert st["deferred"] > 0, st
        assert st["leaf_splits"] > 0
        assert_dump(t, o)


@pytest.mark.parametrize("fanout,per_wave", [(32, 64), (64, 256), (128, 64)])
def test_merged_leaf_transactions(oracle_mod, fanout, per_wave):
    """Thousands of new keys packed into a few leaves' range: the wave-wide upsert kernel merges
    the ops of a round that share a leaf into one transaction, in place or with a split of up
    to 2F entries (at most F ops per transaction when F < 64), while the fast path commits one
    op per leaf.  Updates of present keys ride in the same transactions.  Every answer, the
    dump and the tree's invariants equal the sequential replay."""
    rng = np.random.default_rng(fanout + per_wave)
    load = (np.arange(2048, dtype=np.uint64) << np.uint64(40)) + np.uint64(777)
    lv = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    with tree(fanout=fanout, max_keys=1 << 17, max_batch=1 << 15,
              upsert_ops_per_wave=per_wave) as t:
        t.bulk_load(load, lv)
        for b in range(4):
            lo, hi = int(load[200 + 50 * b]), int(load[202 + 50 * b])
            fresh = rng.integers(lo, hi, size=5000 + 2000 * b, dtype=np.uint64)
            old = load[190 + 50 * b: 215 + 50 * b]  # present keys around the range: updates
            keys = np.concatenate([fresh, rng.choice(fresh, 1000), np.repeat(old, 3)])
            rng.shuffle(keys)
            vals = rng.integers(0, 1 << 63, size=keys.size, dtype=np.uint64)
            ops = (rng.random(keys.size) < 0.1).astype(np.uint8)  # 90 % inserts
            assert_same(t.run_batch(keys, vals, ops), o.run_batch(keys, vals, ops), f"batch {b}")
            assert t.check() == 0
        st = t.stats()
        assert st["commits"] < st["upserts"], st  # some transactions carried several upserts
        assert st["leaf_splits"] > 0 and st["inner_splits"] > 0, st
        assert_dump(t, o)


@pytest.mark.parametrize("graphs", [False, True])
def test_hot_get_groups_fan_out(oracle_mod, graphs):
    """Keys read by many ops of one batch (more than 32 waiting ops: the lookup kernel queues
    their answers in chunks of 1024 that the whole grid writes): group sizes on both sides of
    both limits, present and absent keys, a hot key whose GETs precede an INSERT of it, batches
    through the synchronous call and replayed graphs."""
    from eirene_amd import DeviceBuffer
    rng = np.random.default_rng(61)
    load = rng.integers(0, 1 << 63, size=50000, dtype=np.uint64)
    lv = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    parts, opp = [], []
    absent = rng.integers(0, 1 << 63, size=4, dtype=np.uint64)
    for i, m in enumerate([1, 31, 32, 33, 1023, 1024, 1025, 2049, 7000]):
        parts.append(np.full(m, load[100 * i], dtype=np.uint64))
        opp.append(np.ones(m, dtype=np.uint8))
    for i, m in enumerate([33, 3000]):
        parts.append(np.full(m, absent[i], dtype=np.uint64))
        opp.append(np.ones(m, dtype=np.uint8))
    # hot GETs then one INSERT of the same key (present and absent), then more GETs
    for k in (load[5], absent[2]):
        parts.append(np.full(2500, k, dtype=np.uint64))
        op = np.ones(2500, dtype=np.uint8)
        op[1500] = 0
        opp.append(op)
    parts.append(rng.choice(load, 20000))
    opp.append((rng.random(20000) < 0.8).astype(np.uint8))
    keys = np.concatenate(parts)
    ops = np.concatenate(opp)
    perm = rng.permutation(keys.size)
    keys, ops = keys[perm], ops[perm]
    vals = rng.integers(0, 1 << 63, size=keys.size, dtype=np.uint64)
    offsets = np.array([0, keys.size // 3, keys.size], dtype=np.uint64)
    want_v = np.zeros(keys.size, dtype=np.uint64)
    want_f = np.zeros(keys.size, dtype=np.uint8)
    for b in range(2):
        a, z = int(offsets[b]), int(offsets[b + 1])
        want_v[a:z], want_f[a:z] = o.run_batch(keys[a:z], vals[a:z], ops[a:z])
    with tree(use_graphs=graphs) as t:
        t.bulk_load(load, lv)
        dk, dv, do = (DeviceBuffer.from_array(x) for x in (keys, vals, ops))
        ov, of = DeviceBuffer(keys.size, np.uint64), DeviceBuffer(keys.size, np.uint8)
        t.run_batches_device(dk, dv, do, ov, of, offsets)
        assert_same((ov.download(), of.download()), (want_v, want_f), f"graphs={graphs}")
        assert_dump(t, o)


def test_sort_fallback_then_exact_batches(oracle_mod):
    """Batches above 2^18 ops sort normalized keys; a hot key with many close, distinct
    neighbours leaves a long bucket of distinct keys, so that batch re-sorts with the exact
    passes (sort_fallbacks = 1) and the following batches sort exactly from the start (no new
    fallback), through the synchronous call and through replayed graphs (recaptured in the
    exact form).  Every answer and the dump against the oracle."""
    from eirene_amd import DeviceBuffer
    rng = np.random.default_rng(83)
    load = rng.integers(0, 1 << 63, size=100000, dtype=np.uint64)
    lv = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    hot = load[7]

    def hot_batch(n):
        k = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
        m = n // 4
        k[:m] = hot
        k[m: 2 * m] = hot + rng.integers(1, 1 << 30, m).astype(np.uint64)
        rng.shuffle(k)
        v = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
        o = (rng.random(n) < 0.6).astype(np.uint8)
        return k, v, o

    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    with tree(max_keys=1 << 22, max_batch=1 << 19) as t:
        t.bulk_load(load, lv)
        for b in range(2):
            keys, vals, ops = hot_batch(300000)
            assert_same(t.run_batch(keys, vals, ops), o.run_batch(keys, vals, ops), f"batch {b}")
            assert t.stats()["sort_fallbacks"] == 1  # the second batch sorted exactly at once
        parts = [hot_batch(m) for m in (300000, 270000)]
        keys = np.concatenate([p_[0] for p_ in parts])
        vals = np.concatenate([p_[1] for p_ in parts])
        ops = np.concatenate([p_[2] for p_ in parts])
        offsets = np.array([0, 300000, 570000], dtype=np.uint64)
        want_v, want_f = [], []
        for b in range(2):
            a, z = int(offsets[b]), int(offsets[b + 1])
            wv, wf = o.run_batch(keys[a:z], vals[a:z], ops[a:z])
            want_v.append(wv)
            want_f.append(wf)
        dk, dv, do = (DeviceBuffer.from_array(x) for x in (keys, vals, ops))
        ov, of = DeviceBuffer(keys.size, np.uint64), DeviceBuffer(keys.size, np.uint8)
        t.run_batches_device(dk, dv, do, ov, of, offsets)
        assert_same((ov.download(), of.download()), (np.concatenate(want_v), np.concatenate(want_f)),
                    "stream")
        assert t.stats()["sort_fallbacks"] == 1
        assert_dump(t, o)


def test_plain_uniform_reduced(oracle_mod):
    from eirene_amd import workloads as w
    load = w.uniform_keys(1 << 20, 1)
    lv = w.values_for(load)
    keys, vals, ops = w.mixed_batch(w.sorted_copy(load), 1 << 19, 2, 0.95, 0.9)
    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    with tree(max_keys=1 << 21, max_batch=1 << 19) as t:
        t.bulk_load(load, lv)
        assert_same(t.run_batch(keys, vals, ops), o.run_batch(keys, vals, ops))
        assert_dump(t, o)


def test_reserved_key_rejected():
    from eirene_amd import StmError
    with tree() as t:
        keys = np.array([1, 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
        with pytest.raises(StmError) as e:
            t.run_batch(keys, np.zeros(2, dtype=np.uint64), np.zeros(2, dtype=np.uint8))
        assert e.value.code == 1
        with pytest.raises(StmError):
            t.run_batch(keys[:1], np.zeros(1, dtype=np.uint64), np.array([9], dtype=np.uint8))
        assert t.size() == 0 and t.check() == 0


def test_pool_exhaustion_reports_nomem():
    from eirene_amd import StmError
    rng = np.random.default_rng(1)
    with tree(max_keys=1000, fanout=32) as t:
        keys = rng.integers(0, 1 << 63, size=200000, dtype=np.uint64)
        with pytest.raises(StmError) as e:
            t.run_batch(keys, keys, np.zeros(keys.size, dtype=np.uint8))
        assert e.value.code == 2
        assert t.check() == 0  # the tree stays well formed after refused inserts


def test_device_api_matches_host_api(oracle_mod):
    from eirene_amd import DeviceBuffer
    rng = np.random.default_rng(9)
    load = rng.integers(0, 1 << 63, size=10000, dtype=np.uint64)
    n = 30000
    keys = np.concatenate([rng.choice(load, n // 2), rng.integers(0, 1 << 63, size=n // 2, dtype=np.uint64)])
    vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    ops = (rng.random(n) < 0.8).astype(np.uint8)
    o = oracle_mod.Oracle()
    o.bulk_load(load, load)
    want = o.run_batch(keys, vals, ops)
    with tree() as t:
        t.bulk_load_device(DeviceBuffer.from_array(load), DeviceBuffer.from_array(load))
        dk, dv, do = (DeviceBuffer.from_array(a) for a in (keys, vals, ops))
        ov, of = DeviceBuffer(n, np.uint64), DeviceBuffer(n, np.uint8)
        half = n // 2  # two device batches back to back over one set of buffers
        t.run_batch_device(dk, dv, do, ov, of, n=half)
        t.run_batch_device(dk, dv, do, ov, of, n=n - half, offset=half)
        assert_same((ov.download(), of.download()), want)
        assert_dump(t, o)


@pytest.mark.parametrize("n16", [1, 255, 1024, 4 * 1024 * 3 + 5, (1 << 20) + 17])
def test_device_copy_ragged(n16):
    """The streaming copy kernel moves every 16-byte word of ragged sizes, including sizes
    that leave a partial unrolled step and a partial block."""
    from eirene_amd import DeviceBuffer
    rng = np.random.default_rng(n16)
    a = rng.integers(0, 2**63, size=2 * n16, dtype=np.uint64)
    src = DeviceBuffer.from_array(a)
    dst = DeviceBuffer.from_array(np.zeros_like(a))
    try:
        dst.copy_from(src)
        assert np.array_equal(dst.download(), a)
    finally:
        src.free()
        dst.free()


def test_device_buffer_ranges():
    """download(n, offset) returns exactly elements [offset, offset + n) and rejects ranges
    outside the buffer."""
    from eirene_amd import DeviceBuffer
    a = np.arange(1000, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    b = DeviceBuffer.from_array(a)
    try:
        assert np.array_equal(b.download(), a)
        assert np.array_equal(b.download(10, offset=990), a[990:])
        assert np.array_equal(b.download(offset=500), a[500:])
        assert b.download(0, offset=1000).size == 0
        with pytest.raises(ValueError):
            b.download(11, offset=990)
        with pytest.raises(ValueError):
            b.download(offset=1001)
    finally:
        b.free()


@pytest.mark.parametrize("n", [1 << 18, (1 << 18) + 1])
def test_sort_forms_at_their_boundary(oracle_mod, n):
    """Batches of at most 2^18 ops sort in a single sweep, larger ones in three kernels per pass:
    both sides of the boundary, with heavy key duplication (stability decides the answers)."""
    rng = np.random.default_rng(n)
    load = rng.integers(0, 1 << 16, size=30000).astype(np.uint64)
    lv = rng.integers(0, 1 << 62, size=load.size).astype(np.uint64)
    keys = rng.integers(0, 1 << 16, size=n).astype(np.uint64)
    vals = rng.integers(0, 1 << 62, size=n).astype(np.uint64)
    ops = (rng.random(n) < 0.3).astype(np.uint8) ^ 1  # 30 % inserts
    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    with tree(max_keys=1 << 20, max_batch=(1 << 18) + 1) as t:
        t.bulk_load(load, lv)
        assert_same(t.run_batch(keys, vals, ops), o.run_batch(keys, vals, ops), f"n={n}")
        assert_dump(t, o)
