"""GPU tests of the learned leaf router: device gradients against the CPU restatement,
bit-exact batch results with learned routing and with the static index over the same
directory, the warm directory rewrite, and what the trained network gains over the untrained
router (the linear root with identity segment models)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mlp_gradients_match_cpu():
    import eirene_amd as ea
    from oracle import mlp
    rng = np.random.default_rng(3)
    for n in (1, 63, 64, 65, 1000, 4096):
        p = rng.normal(0, 1, size=49).astype(np.float32)
        x = rng.uniform(0, 1, size=n).astype(np.float32)
        t = np.sort(rng.uniform(0, 1, size=n)).astype(np.float32)
        loss, g = ea.mlp_gradients(x, t, p)
        want_loss, want_g = mlp.loss_and_grad(p.astype(np.float64), x.astype(np.float64),
                                              t.astype(np.float64))
        scale = max(np.abs(want_g).max(), 1e-3)
        # float32 accumulation in a different order: relative tolerance on the gradient scale
        np.testing.assert_allclose(g, want_g, rtol=0, atol=2e-5 * scale + 1e-7)
        assert abs(loss - want_loss) <= 2e-5 * max(want_loss, 1e-6) + 1e-7


def gamma_tree(ea, w, n, seed, **kw):
    load = w.gamma_keys(n, seed, 2.0, 2.0 ** 56)
    vals = w.values_for(load)
    cfg = dict(max_keys=4 * n, max_batch=1 << 18, route=ea.ROUTE_LEARNED)
    cfg.update(kw)
    t = ea.StmTree(ea.StmConfig(**cfg))
    t.bulk_load(load, vals)
    return t, load, vals


def test_learned_routing_results_match_oracle(oracle_mod):
    import eirene_amd as ea
    from eirene_amd import workloads as w
    t, load, vals = gamma_tree(ea, w, 1 << 18, 5, retrain_every=2)
    o = oracle_mod.Oracle()
    o.bulk_load(load, vals)
    srt = w.sorted_copy(load)
    for b in range(5):  # retrains after batches 2 and 4, stale directory in between
        lookup = 0.95 if b % 2 == 0 else 0.5  # insert-heavy batches split many leaves
        keys, bv, ops = w.mixed_batch(srt, 1 << 18, 40 + b, lookup, 0.9, w.DIST_GAMMA, 2.0, 2.0 ** 56)
        gv, gf = t.run_batch(keys, bv, ops)
        ev, ef = o.run_batch(keys, bv, ops)
        np.testing.assert_array_equal(gf, ef)
        np.testing.assert_array_equal(gv, ev)
        assert t.check() == 0
    st = t.stats()
    assert st["route_hits"] > 0
    assert st["route_hits"] > 4 * st["route_misses"], st
    k, v = t.dump()
    ek, evv = o.dump()
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, evv)
    t.close()


def test_index_routing_results_match_oracle(oracle_mod):
    """The static index (the directory searched without the model) routes every op exactly:
    results and the final dump equal the oracle across directory rewrites."""
    import eirene_amd as ea
    from eirene_amd import workloads as w
    t, load, vals = gamma_tree(ea, w, 1 << 18, 11, retrain_every=2, route=ea.ROUTE_INDEX)
    o = oracle_mod.Oracle()
    o.bulk_load(load, vals)
    srt = w.sorted_copy(load)
    for b in range(5):
        lookup = 0.95 if b % 2 == 0 else 0.5
        keys, bv, ops = w.mixed_batch(srt, 1 << 18, 110 + b, lookup, 0.9, w.DIST_GAMMA, 2.0, 2.0 ** 56)
        gv, gf = t.run_batch(keys, bv, ops)
        ev, ef = o.run_batch(keys, bv, ops)
        np.testing.assert_array_equal(gf, ef)
        np.testing.assert_array_equal(gv, ev)
        assert t.check() == 0
    st = t.stats()
    assert st["route_hits"] > 0
    assert st["route_misses"] == 0, st  # the search always lands on the key's entry
    k, v = t.dump()
    ek, evv = o.dump()
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, evv)
    t.close()


def test_warm_directory_rewrite_matches_cold_build():
    """A periodic retrain rewrites the directory from the split chains; after insert-heavy
    batches it must list exactly the leaves a rebuild from every node lists."""
    import eirene_amd as ea
    from eirene_amd import workloads as w
    t, load, vals = gamma_tree(ea, w, 1 << 18, 7, retrain_every=1, train_steps=32)
    t.retrain()  # the directory of the loaded tree; later retrains rewrite it
    f0, _ = t.route_directory()
    srt = w.sorted_copy(load)
    probe = np.ascontiguousarray(srt[:1024])
    for b in range(3):
        keys, bv, ops = w.mixed_batch(srt, 1 << 18, 70 + b, 0.2, 0.5, w.DIST_GAMMA, 2.0, 2.0 ** 56)
        t.run_batch(keys, bv, ops)
        # the retrain runs before each batch: a lookup-only batch rewrites the directory
        # and splits nothing
        t.run_batch(probe, np.zeros_like(probe), np.full(len(probe), ea.OP_GET, dtype=np.uint8))
        fw, iw = t.route_directory()
        assert np.all(fw[1:] > fw[:-1])
    assert len(fw) > len(f0) + 1000, (len(f0), len(fw))
    assert t.check() == 0
    t.retrain(0)  # rebuild from every node
    fc, ic = t.route_directory()
    np.testing.assert_array_equal(fw, fc)
    np.testing.assert_array_equal(iw, ic)
    t.close()


def test_trained_router_beats_untrained():
    import eirene_amd as ea
    from eirene_amd import workloads as w
    # 2^24 keys: ~85 leaves per segment, enough that the identity-like segment models miss
    t, load, vals = gamma_tree(ea, w, 1 << 24, 6)
    t.retrain(0)
    untrained = t.route_info()
    t.retrain(300)
    trained = t.route_info()
    assert trained["leaves"] == untrained["leaves"] > 100000
    rate_u = untrained["window_hits"] / trained["leaves"]
    rate_t = trained["window_hits"] / trained["leaves"]
    print(f"window hit rate untrained {rate_u:.4f} trained {rate_t:.4f}")
    assert trained["window_misses"] * 2 < untrained["window_misses"]
    assert rate_t > 0.9
    t.close()


@pytest.mark.parametrize("per_group", [1, 16])
def test_trained_router_reads_fewer_nodes_than_descent(oracle_mod, per_group):
    """What the network buys on the lookup path: on the same small batch (2^14 lookups over
    2^22 Gamma keys, so lookups rarely share a leaf), root-to-leaf descent reads inner nodes to
    route (every lookup when a lane group walks one op; the first of a span and every parent
    change when it walks 16 sorted ops), while the trained router routes nearly every key from a
    directory window straight to its leaf, and the untrained router (linear root, identity-like
    segment models) misses its window more often.  Tree nodes read per merged lookup, with
    identical answers (checked against the oracle).  A window the untrained model misses costs
    one more directory window load rather than a node read, so between the two routers the
    difference shows in the model's window hit rate."""
    import eirene_amd as ea
    from eirene_amd import workloads as w
    n = 1 << 22
    load = w.gamma_keys(n, 31, 2.0, 2.0 ** 56)
    vals = w.values_for(load)
    srt = w.sorted_copy(load)
    rng = np.random.default_rng(5)
    keys = rng.choice(load, 1 << 14)
    ops = np.full(keys.size, ea.OP_GET, dtype=np.uint8)
    o = oracle_mod.Oracle()
    o.bulk_load(srt, w.values_for(srt))
    want = o.run_batch(keys, np.zeros_like(keys), ops)
    reads, hit_rate = {}, {}
    for name, route, steps in (("descent", ea.ROUTE_DESCENT, None), ("untrained", ea.ROUTE_LEARNED, 0),
                               ("trained", ea.ROUTE_LEARNED, 300)):
        with ea.StmTree(ea.StmConfig(max_keys=2 * n, max_batch=1 << 18, route=route,
                                     ops_per_group=per_group)) as t:
            t.bulk_load(load, vals)
            if steps is not None:
                t.retrain(steps)
            t.reset_stats()
            got = t.run_batch(keys, np.zeros_like(keys), ops)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), name
            st = t.stats()
            reads[name] = st["node_reads"] / st["unique_ops"]
            hit_rate[name] = st["route_hits"] / max(1, st["route_hits"] + st["route_misses"])
    print(f"tree nodes read per merged lookup, {per_group} ops per lane group: {reads}; "
          f"model window hit rate {hit_rate}")
    assert reads["trained"] < reads["descent"] and reads["trained"] <= reads["untrained"]
    assert reads["trained"] < 1.2  # the leaf, and rarely more
    if per_group == 1:
        assert reads["descent"] >= 3.0  # root, inner levels and the leaf of every lookup
    assert hit_rate["trained"] >= hit_rate["untrained"] and hit_rate["trained"] > 0.99


def test_split_heavy_batches_trigger_a_retrain(oracle_mod):
    """With periodic retraining out of the way (retrain_every = 1000), insert-heavy batches that
    split more than a 32nd of the directory's leaves make the next batch rebuild the directory
    and retrain: its leaf count then covers the splits.  Answers against the oracle throughout."""
    import eirene_amd as ea
    from eirene_amd import workloads as w
    n = 1 << 18
    load = w.gamma_keys(n, 12, 2.0, 2.0 ** 56)
    vals = w.values_for(load)
    srt = w.sorted_copy(load)
    o = oracle_mod.Oracle()
    o.bulk_load(srt, w.values_for(srt))
    cfg = ea.StmConfig(max_keys=4 * n, max_batch=1 << 18, route=ea.ROUTE_LEARNED,
                       retrain_every=1000, train_steps=50)
    with ea.StmTree(cfg) as t:
        t.bulk_load(load, vals)
        probe = srt[::97].copy()
        gets = np.full(probe.size, ea.OP_GET, dtype=np.uint8)
        got = t.run_batch(probe, np.zeros_like(probe), gets)  # first batch: directory + training
        want = o.run_batch(probe, np.zeros_like(probe), gets)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        leaves0 = len(t.route_directory()[0])
        splits = 0
        b = 0
        while splits < leaves0 // 32:
            # all inserts, small batches: leaves fill over many batches, so later batches
            # route through a directory whose leaves have already split
            keys, bv, ops = w.mixed_batch(srt, 1 << 12, 90 + b, 0.0, 0.0, w.DIST_GAMMA, 2.0,
                                          2.0 ** 56)
            got = t.run_batch(keys, bv, ops)
            want = o.run_batch(keys, bv, ops)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), b
            splits = t.stats()["leaf_splits"]
            assert len(t.route_directory()[0]) == leaves0  # no periodic retrain yet
            b += 1
        # routing through the stale directory walked right siblings of split leaves
        assert b >= 2 and t.stats()["leaf_steps"] > 0, (b, t.stats()["leaf_steps"])
        got = t.run_batch(probe, np.zeros_like(probe), gets)
        want = o.run_batch(probe, np.zeros_like(probe), gets)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        leaves1 = len(t.route_directory()[0])
        assert leaves1 >= leaves0 + splits, (leaves0, splits, leaves1)
        assert t.check() == 0
