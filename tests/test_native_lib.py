"""CPU-side checks of the native library: it loads, exports every symbol declared in
include/stmtree.h, rejects bad configurations without a GPU call reaching a kernel, and
its host-side workload generators are deterministic and shaped as specified."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "stmtree.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(stm_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("stm_create", "stm_destroy", "stm_bulk_load", "stm_run_batch", "stm_dump",
              "stm_get_stats", "stm_last_error", "stm_run_batch_device", "stm_check"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import ctypes
    import eirene_amd
    lib = eirene_amd.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert isinstance(lib.stm_last_error(), bytes)
    assert isinstance(lib, ctypes.CDLL)


def test_config_validation_without_device():
    import eirene_amd
    lib = eirene_amd.load_library()
    cfg = eirene_amd.StmConfig(fanout=48).to_c()
    import ctypes
    h = ctypes.c_void_p()
    rc = lib.stm_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == 1 and b"fanout" in lib.stm_last_error()
    cfg = eirene_amd.StmConfig(lanes_per_op=12).to_c()
    assert lib.stm_create(ctypes.byref(cfg), ctypes.byref(h)) == 1


def py_splitmix64(x):
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def test_splitmix64_known_values():
    from eirene_amd import workloads as w
    # First outputs of Vigna's splitmix64 seeded with 0 are splitmix64(0), splitmix64(g), ...
    assert w.splitmix64(0) == 0xE220A8397B1DCDAF
    for x in (1, 2, 12345, (1 << 63) + 7, (1 << 64) - 1):
        assert w.splitmix64(x) == py_splitmix64(x)
    keys = np.arange(1000, dtype=np.uint64) * np.uint64(0x12345)
    np.testing.assert_array_equal(w.values_for(keys),
                                  np.array([py_splitmix64(int(k)) for k in keys], dtype=np.uint64))


@pytest.mark.parametrize("dist", ["uniform", "normal", "gamma"])
def test_key_generators(dist):
    from eirene_amd import workloads as w
    fn = {"uniform": w.uniform_keys, "normal": w.normal_keys, "gamma": w.gamma_keys}[dist]
    a = fn(200000, 42)
    b = fn(200000, 42)
    c = fn(200000, 43)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    assert len(np.unique(a)) == a.size
    assert a.max() < (1 << 63)
    f = a.astype(np.float64)
    if dist == "uniform":
        assert abs(f.mean() / 2.0 ** 62 - 1) < 0.01
    elif dist == "normal":
        assert abs(f.mean() / 2.0 ** 62 - 1) < 0.001 and abs(f.std() / 2.0 ** 58 - 1) < 0.02
    else:
        theta = 2.0 ** 56
        assert abs(f.mean() / (2 * theta) - 1) < 0.02 and abs(f.var() / (2 * theta ** 2) - 1) < 0.05


def test_mixed_batch_shape():
    from eirene_amd import workloads as w
    load = w.sorted_copy(w.uniform_keys(100000, 1))
    keys, vals, ops = w.mixed_batch(load, 1 << 16, 9, 0.95, 0.9)
    n_look = int(np.floor((1 << 16) * 0.95))
    assert (ops == 1).sum() == n_look and (ops == 0).sum() == (1 << 16) - n_look
    hit = np.isin(keys, load)
    assert hit[ops == 0].sum() == 0  # inserts are fresh keys
    assert abs(hit[ops == 1].mean() - 0.9) < 0.001
    np.testing.assert_array_equal(vals, w.batch_values(keys, 9))
    # insert values differ from the stored values splitmix64(key)
    assert not np.any(vals == w.values_for(keys))
    k2, _, o2 = w.mixed_batch(load, 1 << 16, 9, 0.95, 0.9)
    np.testing.assert_array_equal(keys, k2)
    np.testing.assert_array_equal(ops, o2)


def test_zipf_batch_is_skewed():
    from eirene_amd import workloads as w
    load = w.sorted_copy(w.uniform_keys(1 << 16, 3))
    keys, vals, ops = w.zipf_batch(load, 1 << 16, 5, 0.99, 0.5)
    look = keys[ops == 1]
    ranks = np.searchsorted(load, look)
    assert np.all(load[ranks] == look)
    # rank 0 is the most frequent and far above the uniform share
    counts = np.bincount(ranks, minlength=load.size)
    assert counts.argmax() == 0 and counts[0] > 50 * (look.size / load.size)
    assert abs((ops == 1).mean() - 0.5) < 0.02


def test_batch_sizes_log_uniform():
    from eirene_amd import workloads as w
    s = w.batch_sizes(4096, 11)
    assert s.min() >= 1 << 10 and s.max() < 1 << 22
    lg = np.log2(s.astype(np.float64))
    assert abs(lg.mean() - 16) < 0.3
