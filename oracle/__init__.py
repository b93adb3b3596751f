"""CPU oracle of the STM B+tree batch path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker (or the reported CPU baseline); the product path in
eirene_amd never touches it.

    Oracle          sequential replay on std::map (oracle.cpp) and the restated v38
                    pipeline (stable sort, RLE, scan, mergeMission, writeback)
    CpuExecutor     T-thread partitioned CPU B+tree (cpu_baseline.cpp), the timed baseline
                    and the checker of the full-size GPU tests (pinned to Oracle by
                    tests/test_oracle.py: every answer and the in-order dump)
    build()         compiles liboracle.so with g++ (no GPU needed)

Parity status of each function is recorded in the header of oracle.cpp and in the tests
that pin it (tests/test_oracle.py).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
SOURCES = [HERE / "oracle.cpp", HERE / "cpu_baseline.cpp"]

_lib = None


def build(force: bool = False) -> Path:
    newest = max(s.stat().st_mtime for s in SOURCES)
    if not force and LIB.exists() and LIB.stat().st_mtime >= newest:
        return LIB
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", str(LIB),
           *map(str, SOURCES)]
    subprocess.run(cmd, check=True)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        P = C.c_void_p
        for name, res, args in [
            ("oracle_create", P, []),
            ("oracle_destroy", None, [P]),
            ("oracle_size", C.c_uint64, [P]),
            ("oracle_bulk_load", None, [P, P, P, C.c_size_t]),
            ("oracle_run_batch", None, [P, P, P, P, C.c_size_t, P, P]),
            ("oracle_pipeline_batch", None, [P, P, P, P, C.c_size_t, P, P]),
            ("oracle_run_batch_scan", C.c_uint64, [P, P, P, P, C.c_size_t, P, P, P, P,
                                                   C.c_uint64]),
            ("oracle_dump", None, [P, P, P]),
            ("oracle_stable_sort_perm", None, [P, C.c_size_t, P]),
            ("cpu_exec_create", P, [P, P, C.c_size_t, C.c_int]),
            ("cpu_exec_destroy", None, [P]),
            ("cpu_exec_run_batch", C.c_double, [P, P, P, P, C.c_size_t, P, P]),
            ("cpu_exec_size", C.c_uint64, [P]),
            ("cpu_exec_dump", None, [P, P, P]),
        ]:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _a(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


class Oracle:
    def __init__(self):
        self._lib = _load()
        self._h = self._lib.oracle_create()

    def close(self):
        if self._h:
            self._lib.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bulk_load(self, keys, vals):
        k, v = _a(keys, np.uint64), _a(vals, np.uint64)
        self._lib.oracle_bulk_load(self._h, k.ctypes.data, v.ctypes.data, k.size)

    def _batch(self, fn, keys, vals, ops):
        k, o = _a(keys, np.uint64), _a(ops, np.uint8)
        v = _a(vals if vals is not None else np.zeros(k.size), np.uint64)
        out_v = np.zeros(k.size, dtype=np.uint64)
        out_f = np.zeros(k.size, dtype=np.uint8)
        fn(self._h, k.ctypes.data, v.ctypes.data, o.ctypes.data, k.size, out_v.ctypes.data,
           out_f.ctypes.data)
        return out_v, out_f

    def run_batch(self, keys, vals, ops):
        """Sequential replay in input order."""
        return self._batch(self._lib.oracle_run_batch, keys, vals, ops)

    def run_batch_scan(self, keys, vals, ops):
        """Sequential replay with op 3 = SCAN(key, L = vals[i]).  Returns (values, found,
        scan_keys, scan_vals); scan op i's pairs start at the sum of the L of earlier scans."""
        k, o = _a(keys, np.uint64), _a(ops, np.uint8)
        v = _a(vals, np.uint64)
        cap = int(v[o == 3].sum()) if k.size else 0
        out_v = np.zeros(k.size, dtype=np.uint64)
        out_f = np.zeros(k.size, dtype=np.uint8)
        sk = np.zeros(max(cap, 1), dtype=np.uint64)
        sv = np.zeros(max(cap, 1), dtype=np.uint64)
        used = self._lib.oracle_run_batch_scan(self._h, k.ctypes.data, v.ctypes.data,
                                               o.ctypes.data, k.size, out_v.ctypes.data,
                                               out_f.ctypes.data, sk.ctypes.data, sv.ctypes.data,
                                               cap)
        assert used == cap
        return out_v, out_f, sk[:cap], sv[:cap]

    def pipeline_batch(self, keys, vals, ops):
        """The reference's sort / RLE / scan / mergeMission / writeback pipeline."""
        return self._batch(self._lib.oracle_pipeline_batch, keys, vals, ops)

    def size(self) -> int:
        return int(self._lib.oracle_size(self._h))

    def dump(self):
        n = self.size()
        k = np.empty(n, dtype=np.uint64)
        v = np.empty(n, dtype=np.uint64)
        self._lib.oracle_dump(self._h, k.ctypes.data, v.ctypes.data)
        return k, v


def stable_sort_perm(keys) -> np.ndarray:
    k = _a(keys, np.uint64)
    perm = np.empty(k.size, dtype=np.uint32)
    _load().oracle_stable_sort_perm(k.ctypes.data, k.size, perm.ctypes.data)
    return perm


class CpuExecutor:
    """Partitioned multi-threaded CPU B+tree executor (the timed CPU baseline)."""

    def __init__(self, sorted_keys, vals, threads: int):
        self._lib = _load()
        self._k = _a(sorted_keys, np.uint64)
        self._v = _a(vals, np.uint64)
        self.threads = threads
        self._h = self._lib.cpu_exec_create(self._k.ctypes.data, self._v.ctypes.data,
                                            self._k.size, threads)

    def run_batch(self, keys, vals, ops):
        """Returns (values, found, seconds)."""
        k, v, o = _a(keys, np.uint64), _a(vals, np.uint64), _a(ops, np.uint8)
        out_v = np.zeros(k.size, dtype=np.uint64)
        out_f = np.zeros(k.size, dtype=np.uint8)
        sec = self._lib.cpu_exec_run_batch(self._h, k.ctypes.data, v.ctypes.data, o.ctypes.data,
                                           k.size, out_v.ctypes.data, out_f.ctypes.data)
        return out_v, out_f, sec

    def run_batches(self, keys, vals, ops, offsets):
        """Batches [offsets[b], offsets[b+1]) one after another; returns (values, found,
        seconds summed over the batches)."""
        k, v, o = _a(keys, np.uint64), _a(vals, np.uint64), _a(ops, np.uint8)
        out_v = np.zeros(k.size, dtype=np.uint64)
        out_f = np.zeros(k.size, dtype=np.uint8)
        sec = 0.0
        for b in range(len(offsets) - 1):
            a, z = int(offsets[b]), int(offsets[b + 1])
            if z > a:
                sec += self._lib.cpu_exec_run_batch(self._h, k[a:].ctypes.data, v[a:].ctypes.data,
                                                    o[a:].ctypes.data, z - a,
                                                    out_v[a:].ctypes.data, out_f[a:].ctypes.data)
        return out_v, out_f, sec

    def size(self) -> int:
        return int(self._lib.cpu_exec_size(self._h))

    def dump(self):
        n = self.size()
        k = np.empty(n, dtype=np.uint64)
        v = np.empty(n, dtype=np.uint64)
        self._lib.cpu_exec_dump(self._h, k.ctypes.data, v.ctypes.data)
        return k, v

    def close(self):
        if self._h:
            self._lib.cpu_exec_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
