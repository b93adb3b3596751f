"""Seeded synthetic workloads of the five benchmark configurations (host side).

Generators run in libstmtree.so's host code (counter-based splitmix64 streams, so the
output does not depend on thread count).  Bulk-load values are splitmix64(key); the insert
values of a generated batch are splitmix64(key ^ salt(seed)), so they differ from the stored
ones (an update that answered the new value instead of the old would show).

    uniform_keys(n, seed)                     keys uniform in [0, 2^63)
    normal_keys(n, seed, mean, sd)            N(mean, sd) clamped to [0, 2^63), unique
    gamma_keys(n, seed, k, theta)             Gamma(k, theta) clamped, unique
    mixed_batch(loaded_sorted, n_ops, seed, lookup_frac, hit_frac, dist, p0, p1)
    zipf_batch(loaded_sorted, n_ops, seed, theta, lookup_frac)
    batch_sizes(count, seed, lo_exp, hi_exp)
"""
from __future__ import annotations

import numpy as np

from . import load_library

DIST_UNIFORM, DIST_NORMAL, DIST_GAMMA = 0, 1, 2


def splitmix64(x: int) -> int:
    return load_library().stm_splitmix64(int(x) & 0xFFFFFFFFFFFFFFFF)


def values_for(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty_like(k)
    load_library().stm_gen_values(k.ctypes.data, k.size, out.ctypes.data)
    return out


def batch_values(keys: np.ndarray, seed: int) -> np.ndarray:
    """Insert values a batch generated with `seed` carries: splitmix64(key ^ salt(seed))."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty_like(k)
    load_library().stm_gen_batch_values(k.ctypes.data, k.size, seed & 0xFFFFFFFFFFFFFFFF,
                                        out.ctypes.data)
    return out


def _keys(dist: int, n: int, seed: int, p0: float = 0.0, p1: float = 0.0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    rc = load_library().stm_gen_keys(dist, n, seed & 0xFFFFFFFFFFFFFFFF, p0, p1, out.ctypes.data)
    if rc != 0:
        raise RuntimeError("key generator could not draw enough unique keys")
    return out


def uniform_keys(n: int, seed: int) -> np.ndarray:
    return _keys(DIST_UNIFORM, n, seed)


def normal_keys(n: int, seed: int, mean: float = 2.0 ** 62, sd: float = 2.0 ** 58) -> np.ndarray:
    return _keys(DIST_NORMAL, n, seed, mean, sd)


def gamma_keys(n: int, seed: int, k: float = 2.0, theta: float = 2.0 ** 56) -> np.ndarray:
    return _keys(DIST_GAMMA, n, seed, k, theta)


def sorted_copy(keys: np.ndarray) -> np.ndarray:
    s = np.array(keys, dtype=np.uint64, copy=True)
    load_library().stm_sort_u64(s.ctypes.data, s.size)
    return s


def mixed_batch(loaded_sorted: np.ndarray, n_ops: int, seed: int, lookup_frac: float = 0.95,
                hit_frac: float = 0.9, dist: int = DIST_UNIFORM, p0: float = 0.0,
                p1: float = 0.0):
    """(keys, vals, ops) of a mixed batch; see stm_gen_mixed_batch."""
    ls = np.ascontiguousarray(loaded_sorted, dtype=np.uint64)
    keys = np.empty(n_ops, dtype=np.uint64)
    vals = np.empty(n_ops, dtype=np.uint64)
    ops = np.empty(n_ops, dtype=np.uint8)
    rc = load_library().stm_gen_mixed_batch(ls.ctypes.data, ls.size, n_ops, seed, lookup_frac,
                                            hit_frac, dist, p0, p1, keys.ctypes.data,
                                            vals.ctypes.data, ops.ctypes.data)
    if rc != 0:
        raise RuntimeError("mixed batch generator failed")
    return keys, vals, ops


def zipf_batch(loaded_sorted: np.ndarray, n_ops: int, seed: int, theta: float = 0.99,
               lookup_frac: float = 0.5):
    ls = np.ascontiguousarray(loaded_sorted, dtype=np.uint64)
    keys = np.empty(n_ops, dtype=np.uint64)
    vals = np.empty(n_ops, dtype=np.uint64)
    ops = np.empty(n_ops, dtype=np.uint8)
    rc = load_library().stm_gen_zipf_batch(ls.ctypes.data, ls.size, n_ops, seed, theta,
                                           lookup_frac, keys.ctypes.data, vals.ctypes.data,
                                           ops.ctypes.data)
    if rc != 0:
        raise RuntimeError("zipf batch generator failed")
    return keys, vals, ops


def batch_sizes(count: int, seed: int, lo_exp: float = 10.0, hi_exp: float = 22.0) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    load_library().stm_gen_batch_sizes(count, seed, lo_exp, hi_exp, out.ctypes.data)
    return out


# Named workload shapes (sizes at full scale).
CONFIGS = {
    "plain_uniform": dict(load=1 << 23, dist=DIST_UNIFORM, batches=1, batch=1 << 22,
                          lookup=0.95, hit=0.9),
    "contended_zipf": dict(load=1 << 24, dist=DIST_UNIFORM, batches=8, batch=1 << 22,
                           lookup=0.5, zipf=0.99),
    "uneven_normal": dict(load=1 << 26, dist=DIST_NORMAL, p0=2.0 ** 62, p1=2.0 ** 58,
                          batches=256, lookup=0.95, hit=0.9),
    "learned_gamma": dict(load=1 << 26, dist=DIST_GAMMA, p0=2.0, p1=2.0 ** 56, batches=64,
                          batch=1 << 22, lookup=0.95, hit=0.9, retrain_every=8),
    "serving_uniform": dict(load=1 << 25, dist=DIST_UNIFORM, rate_mops=200.0),
}
