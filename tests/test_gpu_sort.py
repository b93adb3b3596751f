"""The batch path's device sort (stm_sort_pairs_device) against numpy's stable argsort,
bit-exact (reference: the cub::DeviceRadixSort::SortPairs call it replaces,
src/gpu-stm-tree-v38/bpt-gpustm-kernel.cu:293).

Every form is covered: the single sweep (at most 2^18 keys); above it, the passes over
normalized 24-bit (up to 2^23 keys) or 32-bit keys, with k_fix ordering each bucket of equal
normalized keys and, when a long bucket holds distinct keys, the exact eight-pass fallback in
one persistent launch.  The distributions are chosen to drive each: uniform 63-bit keys and
Gamma keys (short buckets only), the full 64-bit range with both extremes, a narrow range (exact
normalization, long buckets of equal keys), all keys equal, a hot key with distinct neighbours
and a dense cluster inside a wide range (long buckets of distinct keys: the fallback)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def keys_of(dist, n, rng):
    if dist == "uniform63":
        return rng.integers(0, 1 << 63, n, dtype=np.uint64)
    if dist == "full64":
        k = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
        k[: min(n, 8)] = np.array([0, 2**64 - 1, 0, 2**64 - 2, 1, 2**64 - 1, 7, 0],
                                  dtype=np.uint64)[: min(n, 8)]
        return k
    if dist == "gamma":
        return rng.gamma(2.0, 2.0 ** 56, n).astype(np.uint64)
    if dist == "narrow":
        return rng.integers(10 ** 12, 10 ** 12 + 5000, n).astype(np.uint64)
    if dist == "equal":
        return np.full(n, 0x123456789ABCDEF, dtype=np.uint64)
    if dist == "hot":  # one key a third of the time, distinct keys just above it, the rest wide
        m = n // 3
        k = rng.integers(0, 1 << 63, n, dtype=np.uint64)
        k[:m] = np.uint64(1 << 55)
        k[m: 2 * m] = np.uint64(1 << 55) + rng.integers(1, 1 << 30, m).astype(np.uint64)
        return k
    if dist == "medium":  # groups of 1..64 keys (some equal), each alone in a 2^39-wide bucket
        sizes = rng.integers(1, 65, max(2, n // 16))
        slots = rng.choice(1 << 23, sizes.size, replace=False).astype(np.uint64) * np.uint64(2)
        slots[:2] = [0, (1 << 24) - 2]  # the range spans 2^63: normalized buckets 2^39 wide
        base = (slots << np.uint64(39)) + rng.integers(0, 1 << 38, sizes.size).astype(np.uint64)
        k = np.repeat(base, sizes) + rng.integers(0, 1 << 20, int(sizes.sum())).astype(np.uint64)
        return k[:n]
    if dist == "cluster":  # a dense cluster of distinct keys inside a 2^63-wide range
        k = rng.integers(0, 1 << 63, n, dtype=np.uint64)
        m = n // 2
        k[:m] = np.uint64(5 << 40) + rng.integers(0, 1 << 20, m).astype(np.uint64)
        return k
    raise ValueError(dist)


def check(keys):
    from eirene_amd import sort_pairs
    got_k, got_i = sort_pairs(keys)
    want_i = np.argsort(keys, kind="stable").astype(np.uint32)
    bad = np.nonzero(got_i != want_i)[0]
    assert bad.size == 0, f"{bad.size} positions differ, first {bad[:5].tolist()}"
    assert np.array_equal(got_k, keys[want_i])


DISTS = ["uniform63", "full64", "gamma", "narrow", "equal", "hot", "medium", "cluster"]


@pytest.mark.parametrize("dist", DISTS)
@pytest.mark.parametrize("n", [1, 1000, (1 << 18) + 1, (1 << 20) + 3])
def test_sort_matches_stable_argsort(dist, n):
    rng = np.random.default_rng(n * 7 + DISTS.index(dist))
    keys = keys_of(dist, n, rng)
    rng.shuffle(keys)
    check(keys)


@pytest.mark.parametrize("dist", ["uniform63", "gamma", "hot", "medium"])
@pytest.mark.parametrize("n", [1 << 22, (1 << 23) + 5])
def test_sort_large(dist, n):
    """2^22 keys (three 8-bit normalized passes) and just above 2^23 (four)."""
    rng = np.random.default_rng(n + DISTS.index(dist))
    keys = keys_of(dist, n, rng)
    rng.shuffle(keys)
    check(keys)
