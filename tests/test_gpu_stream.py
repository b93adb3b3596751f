"""GPU parity of batch streams (stm_run_batches_device) against the CPU oracle (bit-exact).

A stream is many batches of uneven size stored back to back in device memory.  With
use_graphs the library replays one captured hipGraph per size class, reading each batch
from a device-side descriptor; the results must equal the oracle's batch-by-batch replay
(the reference's gpu_work_main loop, src/gpu-stm-tree-v38/bpt-gpustm.cpp:77-135), for every
op and for the final in-order dump.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RESERVED = np.uint64(0xFFFFFFFFFFFFFFFF)


def tree(**kw):
    from eirene_amd import StmConfig, StmTree
    base = dict(max_keys=1 << 20, max_batch=1 << 14)
    base.update(kw)
    return StmTree(StmConfig(**base))


def make_stream(seed, load, n_batches, max_batch):
    """Batch sizes log-uniform in [1, 2^14] plus the size-class edges, one batch longer than
    max_batch (chunked) and empty batches; 70% GET / 30% INSERT with stray op codes."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(0, np.log(1 << 14), n_batches)).astype(np.int64)
    edges = [0, 1, 1024, 1025, 4096, 4097, max_batch, 3 * max_batch + 5, 0]
    sizes = np.concatenate([sizes[: n_batches // 2], edges, sizes[n_batches // 2:]])
    n = int(sizes.sum())
    hit = rng.random(n) < 0.5
    keys = np.where(hit, rng.choice(load, n), rng.integers(0, 1 << 63, size=n, dtype=np.uint64))
    keys = keys.astype(np.uint64)
    # repeat some keys inside a batch so the per-key fold has work
    rep = rng.random(n) < 0.1
    keys[1:][rep[1:]] = keys[:-1][rep[1:]]
    vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    ops = np.where(rng.random(n) < 0.3, 0, 1).astype(np.uint8)
    ops[rng.random(n) < 0.02] = 2
    ops[rng.random(n) < 0.02] = 3
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return keys, vals, ops, offsets


def oracle_replay(oracle_mod, load, lv, keys, vals, ops, offsets):
    o = oracle_mod.Oracle()
    o.bulk_load(load, lv)
    out_v = np.zeros(keys.size, dtype=np.uint64)
    out_f = np.zeros(keys.size, dtype=np.uint8)
    for b in range(offsets.size - 1):
        a, z = int(offsets[b]), int(offsets[b + 1])
        if z > a:
            out_v[a:z], out_f[a:z] = o.run_batch(keys[a:z], vals[a:z], ops[a:z])
    return o, out_v, out_f


def run_stream(t, keys, vals, ops, offsets, calls=1):
    """Runs the stream through the C ABI, split into `calls` consecutive calls."""
    from eirene_amd import DeviceBuffer
    n = keys.size
    dk, dv, do = (DeviceBuffer.from_array(a) for a in (keys, vals, ops))
    ov, of = DeviceBuffer(n, np.uint64), DeviceBuffer(n, np.uint8)
    nb = offsets.size - 1
    cuts = np.linspace(0, nb, calls + 1).astype(int)
    for c in range(calls):
        t.run_batches_device(dk, dv, do, ov, of, offsets[cuts[c]: cuts[c + 1] + 1])
    return ov.download(), of.download()


def check_same(got_v, got_f, want_v, want_f, mask=None):
    bad = (got_f != want_f) | (got_v != want_v)
    if mask is not None:
        bad &= mask
    idx = np.nonzero(bad)[0]
    assert idx.size == 0, f"{idx.size} mismatches, first at {idx[:5].tolist()}"


@pytest.mark.parametrize("fanout,tile,graphs,lanes,group,per_wave,ulanes", [
    (32, 8, True, 8, 8, 256, 8),
    (64, 16, True, 16, 16, 512, 16),
    (64, 8, True, 32, 32, 1024, 32),
    (128, 16, True, 16, 16, 512, 0),  # automatic: 16 lanes per insert
    (64, 16, False, 16, 16, 512, 0),  # automatic: 8
])
def test_stream_matches_oracle(oracle_mod, fanout, tile, graphs, lanes, group, per_wave, ulanes):
    rng = np.random.default_rng(fanout + tile)
    load = rng.integers(0, 1 << 63, size=40000, dtype=np.uint64)
    lv = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    keys, vals, ops, offsets = make_stream(fanout * 7 + tile, load, 40, 1 << 14)
    o, want_v, want_f = oracle_replay(oracle_mod, load, lv, keys, vals, ops, offsets)
    with tree(fanout=fanout, sort_tile_items=tile, use_graphs=graphs, lanes_per_op=lanes,
              ops_per_group=group, upsert_ops_per_wave=per_wave, upsert_lanes=ulanes) as t:
        t.bulk_load(load, lv)
        got_v, got_f = run_stream(t, keys, vals, ops, offsets, calls=2)
        check_same(got_v, got_f, want_v, want_f)
        k, v = t.dump()
        ek, ev = o.dump()
        np.testing.assert_array_equal(k, ek)
        np.testing.assert_array_equal(v, ev)
        assert t.check() == 0
        st = t.stats()
        assert st["ops"] == keys.size
        assert st["upserts"] > 0


def test_stream_large_size_classes(oracle_mod):
    """max_batch = 2^20: graph size classes 2^16..2^20, batches on both sides of the 2^18-key
    single-sweep boundary (so the three-kernel sort reads its length from the device-side
    descriptor inside a replayed graph), a batch longer than max_batch (chunked) and small
    batches between them; every op and the dump against the oracle."""
    rng = np.random.default_rng(41)
    load = rng.integers(0, 1 << 63, size=200000, dtype=np.uint64)
    lv = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    sizes = np.array([(1 << 18) - 3, (1 << 18) + 1, 3000, (1 << 20), (1 << 19) + 77, 70000,
                      (1 << 18), 1, (1 << 20) + 4097, (1 << 16) + 1], dtype=np.int64)
    n = int(sizes.sum())
    hit = rng.random(n) < 0.6
    keys = np.where(hit, rng.choice(load, n), rng.integers(0, 1 << 63, size=n, dtype=np.uint64))
    keys = keys.astype(np.uint64)
    rep = rng.random(n) < 0.1
    keys[1:][rep[1:]] = keys[:-1][rep[1:]]
    vals = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    ops = np.where(rng.random(n) < 0.2, 0, 1).astype(np.uint8)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    o, want_v, want_f = oracle_replay(oracle_mod, load, lv, keys, vals, ops, offsets)
    with tree(max_keys=1 << 22, max_batch=1 << 20) as t:
        t.bulk_load(load, lv)
        t.capture_graphs()
        got_v, got_f = run_stream(t, keys, vals, ops, offsets)
        check_same(got_v, got_f, want_v, want_f)
        k, v = t.dump()
        ek, ev = o.dump()
        np.testing.assert_array_equal(k, ek)
        np.testing.assert_array_equal(v, ev)
        assert t.check() == 0


def test_stream_equals_single_batches(oracle_mod):
    """The graph replay and per-batch host launches leave identical trees and outputs."""
    rng = np.random.default_rng(3)
    load = rng.integers(0, 1 << 40, size=30000, dtype=np.uint64)
    keys, vals, ops, offsets = make_stream(11, load, 30, 1 << 14)
    res = []
    for graphs in (True, False):
        with tree(use_graphs=graphs) as t:
            t.bulk_load(load, load)
            res.append(run_stream(t, keys, vals, ops, offsets) + t.dump())
    for a, b in zip(res[0], res[1]):
        np.testing.assert_array_equal(a, b)


def test_stream_learned_routing(oracle_mod):
    """Learned routing inside replayed graphs, retrained every 3 batches between replays."""
    from eirene_amd import ROUTE_LEARNED
    rng = np.random.default_rng(17)
    load = np.unique(rng.gamma(2.0, 2.0 ** 40, size=60000).astype(np.uint64))
    lv = rng.integers(0, 1 << 63, size=load.size, dtype=np.uint64)
    keys, vals, ops, offsets = make_stream(23, load, 24, 1 << 14)
    o, want_v, want_f = oracle_replay(oracle_mod, load, lv, keys, vals, ops, offsets)
    with tree(route=ROUTE_LEARNED, retrain_every=3, train_steps=50, timing=True) as t:
        t.bulk_load(load, lv)
        got_v, got_f = run_stream(t, keys, vals, ops, offsets)
        check_same(got_v, got_f, want_v, want_f)
        k, v = t.dump()
        ek, ev = o.dump()
        np.testing.assert_array_equal(k, ek)
        np.testing.assert_array_equal(v, ev)
        assert t.check() == 0
        st = t.stats()
        assert st["route_hits"] > 0
        assert st["phase_launches"]["train"] >= 3


def test_stream_stops_after_reserved_key(oracle_mod):
    """A batch holding the reserved key is reported once the stream ends; the tree then holds
    the batches before it and that batch's other ops, and no later batch."""
    from eirene_amd import StmError
    rng = np.random.default_rng(29)
    load = rng.integers(0, 1 << 63, size=20000, dtype=np.uint64)
    keys, vals, ops, offsets = make_stream(31, load, 12, 1 << 14)
    keys[keys == RESERVED] = 1
    bad_batch = 8  # the 1024-op batch of make_stream's size-class edges
    a, z = int(offsets[bad_batch]), int(offsets[bad_batch + 1])
    assert z - a >= 2
    keys[a + (z - a) // 2] = RESERVED
    o = oracle_mod.Oracle()
    o.bulk_load(load, load)
    want_v = np.zeros(keys.size, dtype=np.uint64)
    want_f = np.zeros(keys.size, dtype=np.uint8)
    for b in range(bad_batch + 1):
        lo, hi = int(offsets[b]), int(offsets[b + 1])
        m = keys[lo:hi] != RESERVED
        if hi > lo:
            v, f = o.run_batch(keys[lo:hi][m], vals[lo:hi][m], ops[lo:hi][m])
            want_v[lo:hi][m], want_f[lo:hi][m] = v, f
    mask = np.zeros(keys.size, dtype=bool)
    mask[:z] = True
    mask &= keys != RESERVED
    from eirene_amd import DeviceBuffer
    with tree() as t:
        t.bulk_load(load, load)
        dk, dv, do = (DeviceBuffer.from_array(x) for x in (keys, vals, ops))
        ov, of = DeviceBuffer(keys.size, np.uint64), DeviceBuffer(keys.size, np.uint8)
        with pytest.raises(StmError) as e:
            t.run_batches_device(dk, dv, do, ov, of, offsets)
        assert e.value.code == 1
        assert "later batches skipped" in str(e.value)
        check_same(ov.download(), of.download(), want_v, want_f, mask)
        k, v = t.dump()
        ek, ev = o.dump()
        np.testing.assert_array_equal(k, ek)
        np.testing.assert_array_equal(v, ev)
        assert t.check() == 0
