"""Writes the committed golden fixtures under tests/golden/.

The reference ships no tests, fixtures or data (its op files are git-ignored), so the
fixtures are computed here from the definition of the batch semantics, independently of
the C++ oracle: a plain Python dict replays every op in input order
(src/input-helper.h:27-38 of the reference: INSERT returns the old value and assigns,
GET returns the value; codes other than INSERT act as GET as in
src/gpu-stm-tree-v38/bpt-gpustm-kernel-aux.cuh:34).  Keys and op mixes follow the
benchmark shapes at small sizes; values are splitmix64(key) as in the workloads, except in
the "random_values" fixture, whose values are arbitrary so that returning a new value
instead of the old one cannot go unnoticed.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
M64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    # Vigna's splitmix64 finaliser (public domain reference implementation).
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def replay(state: dict, keys, vals, ops):
    out_v = np.zeros(len(keys), dtype=np.uint64)
    out_f = np.zeros(len(keys), dtype=np.uint8)
    for i, (k, v, o) in enumerate(zip(keys.tolist(), vals.tolist(), ops.tolist())):
        if k in state:
            out_f[i] = 1
            out_v[i] = state[k]
        if o == 0:
            state[k] = v
    return out_v, out_f


def fixture(name, rng, n_load, batches):
    """batches: list of (n_ops, key_source, lookup_frac)."""
    load = np.unique(rng.integers(0, 1 << 63, size=n_load, dtype=np.uint64))
    rng.shuffle(load)
    load_vals = np.array([splitmix64(int(k)) for k in load], dtype=np.uint64)
    state = {int(k): int(v) for k, v in zip(load, load_vals)}
    arrays = {"load_keys": load, "load_vals": load_vals}
    for b, (n_ops, source, lookup_frac, rand_vals) in enumerate(batches):
        if source == "mixed":
            hits = rng.choice(load, size=n_ops // 2)
            fresh = rng.integers(0, 1 << 63, size=n_ops - n_ops // 2, dtype=np.uint64)
            keys = np.concatenate([hits, fresh])
            rng.shuffle(keys)
        elif source == "hot":  # few distinct keys: long per-key runs inside the batch
            pool = np.concatenate([load[:8], rng.integers(0, 1 << 63, size=8, dtype=np.uint64)])
            keys = rng.choice(pool, size=n_ops)
        else:  # "fresh": all new keys, forces leaf and inner splits
            keys = rng.integers(0, 1 << 63, size=n_ops, dtype=np.uint64)
        ops = np.where(rng.random(n_ops) < lookup_frac, 1, 0).astype(np.uint8)
        # a sprinkle of DELETE / RANGE codes, which execute as GET
        odd = rng.random(n_ops) < 0.02
        ops[odd & (ops == 1)] = rng.choice([2, 3], size=int((odd & (ops == 1)).sum()))
        if rand_vals:
            vals = rng.integers(0, 1 << 64, size=n_ops, dtype=np.uint64, endpoint=False)
        else:
            vals = np.array([splitmix64(int(k)) for k in keys], dtype=np.uint64)
        out_v, out_f = replay(state, keys, vals, ops)
        arrays[f"b{b}_keys"] = keys.astype(np.uint64)
        arrays[f"b{b}_vals"] = vals
        arrays[f"b{b}_ops"] = ops
        arrays[f"b{b}_out_vals"] = out_v
        arrays[f"b{b}_out_found"] = out_f
    dk = np.array(sorted(state), dtype=np.uint64)
    arrays["dump_keys"] = dk
    arrays["dump_vals"] = np.array([state[int(k)] for k in dk], dtype=np.uint64)
    arrays["n_batches"] = np.array([len(batches)])
    np.savez_compressed(HERE / f"{name}.npz", **arrays)
    print(name, {k: v.shape for k, v in arrays.items() if k.endswith("keys")})


def main():
    fixture("mixed_95_5", np.random.default_rng(1), 1000,
            [(4096, "mixed", 0.95, False), (4096, "mixed", 0.95, False)])
    fixture("hot_keys_50_50", np.random.default_rng(2), 1000,
            [(4096, "hot", 0.5, True), (4096, "hot", 0.5, True)])
    fixture("random_values", np.random.default_rng(3), 500,
            [(2048, "mixed", 0.5, True), (2048, "hot", 0.3, True)])
    fixture("split_heavy", np.random.default_rng(4), 64,
            [(4096, "fresh", 0.0, False), (4096, "fresh", 0.1, False), (4096, "mixed", 0.5, True)])
    fixture("empty_tree", np.random.default_rng(5), 0,
            [(1024, "fresh", 0.2, True), (1024, "hot", 0.5, True)])


if __name__ == "__main__":
    main()
