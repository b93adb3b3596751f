"""eirene_amd — MI355X-native transactional B+tree batch path (Python host layer).

The product is the C ABI of ``include/stmtree.h`` implemented by ``libstmtree.so``
(hand-written HIP kernels for gfx950).  This module binds it with ctypes and mirrors the
reference's batch interface (``GPU_BTree`` in the reference's src/tree.h:28-43):

    prepareGPU()            -> StmTree(...)
    prepareInput + gpu_work -> StmTree.run_batch(keys, vals, ops)
    first "-i" insert file  -> StmTree.bulk_load(keys, vals)
    test()                  -> the (values, found) arrays returned by run_batch

Op codes are the reference's (src/input-helper.h:14-17): INSERT=0, GET=1, DELETE=2,
RANGE=3; codes other than INSERT run with GET semantics, as in the reference's merge.
There is no CPU fallback: importing works without a GPU, but creating a tree needs the
HIP library and a device, and fails with StmError otherwise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

__all__ = [
    "StmTree", "StmError", "StmConfig", "DeviceBuffer", "load_library", "LIB_PATH",
    "OP_INSERT", "OP_GET", "OP_DELETE", "OP_RANGE", "KEY_RESERVED",
    "ROUTE_DESCENT", "ROUTE_LEARNED", "ROUTE_INDEX", "PHASES",
]

LIB_PATH = Path(__file__).resolve().parent / "libstmtree.so"

OP_INSERT, OP_GET, OP_DELETE, OP_RANGE = 0, 1, 2, 3
KEY_RESERVED = 0xFFFFFFFFFFFFFFFF
ROUTE_DESCENT, ROUTE_LEARNED, ROUTE_INDEX = 0, 1, 2
NOALLOC_MAX_OPS = 65536  # largest batch of stm_run_batch_device_noalloc
PHASES = ("sort", "merge", "lookup", "upsert", "deferred", "route", "train", "scan")

STATUS = {0: "OK", 1: "INVALID", 2: "NOMEM", 3: "DEVICE", 4: "STATE"}


class StmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code


class _Config(C.Structure):
    _fields_ = [
        ("fanout", C.c_uint32), ("leaf_fill", C.c_uint32), ("max_keys", C.c_uint64),
        ("max_batch", C.c_uint64), ("device", C.c_int32), ("lanes_per_op", C.c_uint32),
        ("ops_per_group", C.c_uint32), ("upsert_ops_per_wave", C.c_uint32),
        ("max_attempts", C.c_uint32), ("route", C.c_uint32), ("timing", C.c_uint32),
        ("retrain_every", C.c_uint32), ("train_steps", C.c_uint32),
        ("sort_tile_items", C.c_uint32), ("use_graphs", C.c_uint32), ("upsert_lanes", C.c_uint32),
        ("retrain_splits", C.c_uint32),
    ]


class _Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "n_keys", "n_nodes", "capacity_nodes", "height", "batches", "ops", "unique_ops",
        "upserts", "commits", "aborts", "leaf_splits", "inner_splits", "root_splits",
        "deferred", "route_hits", "route_misses")] + [
        ("phase_ms", C.c_double * 8), ("phase_launches", C.c_uint64 * 8),
        ("node_reads", C.c_uint64), ("leaf_steps", C.c_uint64), ("sort_fallbacks", C.c_uint64)]


_lib: Optional[C.CDLL] = None
_P = C.c_void_p
_U64P = C.POINTER(C.c_uint64)
_U8P = C.POINTER(C.c_uint8)


def load_library(path: Optional[os.PathLike] = None) -> C.CDLL:
    """Loads libstmtree.so (built by ``python -m eirene_amd.build``); raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise StmError(3, f"native library {p} is missing; run `python -m eirene_amd.build`")
    lib = C.CDLL(str(p))
    sig = {
        "stm_config_init": (C.c_int, [C.POINTER(_Config)]),
        "stm_create": (C.c_int, [C.POINTER(_Config), C.POINTER(_P)]),
        "stm_destroy": (C.c_int, [_P]),
        "stm_bulk_load": (C.c_int, [_P, _P, _P, C.c_size_t]),
        "stm_bulk_load_device": (C.c_int, [_P, _P, _P, C.c_size_t]),
        "stm_run_batch": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, _P]),
        "stm_run_batch_device": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, _P]),
        "stm_run_batch_device_noalloc": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, _P]),
        "stm_run_batches_device": (C.c_int, [_P, _P, _P, _P, _P, C.c_size_t, _P, _P]),
        "stm_run_batch_scan": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, _P, _P, _P, C.c_size_t]),
        "stm_run_batch_scan_device": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, _P, _P, _P,
                                                C.c_size_t]),
        "stm_capture_graphs": (C.c_int, [_P]),
        "stm_sort_pairs_device": (C.c_int, [C.c_int, _P, C.c_size_t, _P, _P]),
        "stm_size": (C.c_int, [_P, _U64P]),
        "stm_dump": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "stm_check": (C.c_int, [_P, _U64P]),
        "stm_get_stats": (C.c_int, [_P, C.POINTER(_Stats)]),
        "stm_reset_stats": (C.c_int, [_P]),
        "stm_get_stream": (C.c_int, [_P, C.POINTER(_P)]),
        "stm_last_error": (C.c_char_p, []),
        "stm_device_alloc": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(_P)]),
        "stm_device_free": (C.c_int, [_P]),
        "stm_memcpy_h2d": (C.c_int, [_P, _P, C.c_size_t]),
        "stm_memcpy_d2h": (C.c_int, [_P, _P, C.c_size_t]),
        "stm_device_copy": (C.c_int, [_P, _P, C.c_size_t]),
        "stm_sync": (C.c_int, [_P]),
        "stm_retrain": (C.c_int, [_P, C.c_uint32]),
        "stm_route_info": (C.c_int, [_P, _U64P, _U64P, _U64P]),
        "stm_route_directory": (C.c_int, [_P, _P, _P, C.c_uint64, _U64P]),
        "stm_mlp_gradients": (C.c_int, [_P, _P, C.c_size_t, _P, _P, C.POINTER(C.c_float)]),
        # workload generators (host only)
        "stm_splitmix64": (C.c_uint64, [C.c_uint64]),
        "stm_gen_values": (None, [_P, C.c_size_t, _P]),
        "stm_gen_batch_values": (None, [_P, C.c_size_t, C.c_uint64, _P]),
        "stm_gen_keys": (C.c_int, [C.c_int, C.c_size_t, C.c_uint64, C.c_double, C.c_double, _P]),
        "stm_gen_mixed_batch": (C.c_int, [_P, C.c_size_t, C.c_size_t, C.c_uint64, C.c_double,
                                          C.c_double, C.c_int, C.c_double, C.c_double, _P, _P, _P]),
        "stm_gen_zipf_batch": (C.c_int, [_P, C.c_size_t, C.c_size_t, C.c_uint64, C.c_double,
                                         C.c_double, _P, _P, _P]),
        "stm_gen_batch_sizes": (None, [C.c_size_t, C.c_uint64, C.c_double, C.c_double, _P]),
        "stm_sort_u64": (None, [_P, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise StmError(rc, load_library().stm_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint8)


class DeviceBuffer:
    """A device allocation made through the library's own HIP runtime.

    Keeps batches resident in HBM without bringing a second GPU runtime (such as the one
    bundled with PyTorch) into the process."""

    def __init__(self, n: int, dtype, device: int = 0):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        _check(load_library().stm_device_alloc(device, self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.size, a.dtype, device)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size > self.n:
            raise ValueError("array larger than the buffer")
        _check(load_library().stm_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, n: Optional[int] = None, offset: int = 0) -> np.ndarray:
        """Elements [offset, offset + n) (default: to the end) copied to the host."""
        offset = int(offset)
        n = self.n - offset if n is None else int(n)
        if offset < 0 or n < 0 or offset + n > self.n:
            raise ValueError("range outside the buffer")
        out = np.empty(n, dtype=self.dtype)
        _check(load_library().stm_memcpy_d2h(out.ctypes.data, self.ptr + offset * self.dtype.itemsize,
                                             out.nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer") -> None:
        """Device-to-device copy of other's bytes (same size) by the library's copy kernel."""
        nbytes = self.n * np.dtype(self.dtype).itemsize
        if other.n * np.dtype(other.dtype).itemsize != nbytes:
            raise ValueError("buffers differ in size")
        _check(load_library().stm_device_copy(self.ptr, other.ptr, nbytes))

    def free(self) -> None:
        if getattr(self, "ptr", None):
            load_library().stm_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sort_pairs(keys, device: int = 0):
    """The library's stable device sort of u64 keys: returns (sorted keys, input positions)."""
    k = _u64(keys)
    n = k.size
    if n == 0:
        return k.copy(), np.zeros(0, dtype=np.uint32)
    src = DeviceBuffer.from_array(k)
    ok, oi = DeviceBuffer(n, np.uint64), DeviceBuffer(n, np.uint32)
    try:
        _check(load_library().stm_sort_pairs_device(device, src.ptr, n, ok.ptr, oi.ptr))
        return ok.download(), oi.download()
    finally:
        for b in (src, ok, oi):
            b.free()


@dataclass
class StmConfig:
    fanout: int = 64
    leaf_fill: int = 0
    max_keys: int = 1 << 24
    max_batch: int = 1 << 22
    device: int = 0
    lanes_per_op: int = 16
    ops_per_group: int = 16
    upsert_ops_per_wave: int = 512
    max_attempts: int = 64
    route: int = ROUTE_DESCENT
    timing: bool = False
    retrain_every: int = 8
    train_steps: int = 300
    sort_tile_items: int = 16
    use_graphs: bool = True
    upsert_lanes: int = 0  # 0 = automatic
    retrain_splits: int = 0  # 0 = automatic: a 32nd of the directory's leaves

    def to_c(self) -> _Config:
        c = _Config()
        _check(load_library().stm_config_init(C.byref(c)))
        for f in ("fanout", "leaf_fill", "max_keys", "max_batch", "device", "lanes_per_op",
                  "ops_per_group", "upsert_ops_per_wave", "max_attempts", "route",
                  "retrain_every", "train_steps", "sort_tile_items", "use_graphs", "upsert_lanes",
                  "retrain_splits"):
            setattr(c, f, int(getattr(self, f)))
        c.timing = 1 if self.timing else 0
        return c


class StmTree:
    """One device-resident tree.  Not thread-safe: drive it from one thread."""

    def __init__(self, config: Optional[StmConfig] = None, **kw):
        cfg = config or StmConfig(**kw)
        self.config = cfg
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.stm_create(C.byref(cfg.to_c()), C.byref(h)))
        self._h = h

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.stm_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- operations
    def bulk_load(self, keys, vals) -> None:
        k, v = _u64(keys), _u64(vals)
        if k.shape != v.shape:
            raise ValueError("keys and vals differ in length")
        _check(self._lib.stm_bulk_load(self._h, _ptr(k), _ptr(v), k.size))

    def bulk_load_device(self, keys: "DeviceBuffer", vals: "DeviceBuffer", n: Optional[int] = None) -> None:
        n = keys.n if n is None else int(n)
        _check(self._lib.stm_bulk_load_device(self._h, keys.ptr, vals.ptr, n))

    def run_batch(self, keys, vals, ops):
        """Runs one batch; returns (values u64[n], found u8[n])."""
        k, o = _u64(keys), _u8(ops)
        v = _u64(vals) if vals is not None else None
        n = k.size
        if o.size != n or (v is not None and v.size != n):
            raise ValueError("keys, vals and ops differ in length")
        out_v = np.zeros(n, dtype=np.uint64)
        out_f = np.zeros(n, dtype=np.uint8)
        _check(self._lib.stm_run_batch(self._h, _ptr(k), _ptr(v) if v is not None else None,
                                       _ptr(o), n, _ptr(out_v), _ptr(out_f)))
        return out_v, out_f

    def run_batch_scan(self, keys, vals, ops):
        """One batch in which op 3 is SCAN(key, L = vals[i]).  Returns (values, found,
        scan_keys, scan_vals): scan op i's pairs start at the sum of the L of the scans before
        it, and values[i] is the number of pairs it found."""
        k, v, o = _u64(keys), _u64(vals), _u8(ops)
        n = k.size
        if o.size != n or v.size != n:
            raise ValueError("keys, vals and ops differ in length")
        cap = int(v[o == 3].sum()) if n else 0
        out_v = np.zeros(n, dtype=np.uint64)
        out_f = np.zeros(n, dtype=np.uint8)
        sk = np.zeros(max(cap, 1), dtype=np.uint64)
        sv = np.zeros(max(cap, 1), dtype=np.uint64)
        _check(self._lib.stm_run_batch_scan(self._h, _ptr(k), _ptr(v), _ptr(o), n, _ptr(out_v),
                                            _ptr(out_f), _ptr(sk), _ptr(sv), cap))
        return out_v, out_f, sk[:cap], sv[:cap]

    def run_batch_scan_device(self, keys: "DeviceBuffer", vals: "DeviceBuffer",
                              ops: "DeviceBuffer", out_vals: "DeviceBuffer",
                              out_found: "DeviceBuffer", scan_keys: "DeviceBuffer",
                              scan_vals: "DeviceBuffer", n: Optional[int] = None,
                              offset: int = 0, scan_offset: int = 0) -> None:
        """Device-resident scan batch of n ops from element `offset`; scan pairs go to
        scan_keys / scan_vals from element `scan_offset`."""
        n = keys.n - offset if n is None else int(n)
        cap = scan_keys.n - scan_offset
        _check(self._lib.stm_run_batch_scan_device(
            self._h, keys.ptr + 8 * offset, vals.ptr + 8 * offset, ops.ptr + offset, n,
            out_vals.ptr + 8 * offset, out_found.ptr + offset, scan_keys.ptr + 8 * scan_offset,
            scan_vals.ptr + 8 * scan_offset, cap))

    def run_batch_device(self, keys: "DeviceBuffer", vals: "DeviceBuffer", ops: "DeviceBuffer",
                         out_vals: "DeviceBuffer", out_found: "DeviceBuffer",
                         n: Optional[int] = None, offset: int = 0) -> None:
        """Device-resident batch of n ops starting at element `offset` of each buffer."""
        n = keys.n - offset if n is None else int(n)
        _check(self._lib.stm_run_batch_device(
            self._h, keys.ptr + 8 * offset, vals.ptr + 8 * offset, ops.ptr + offset, n,
            out_vals.ptr + 8 * offset, out_found.ptr + offset))

    def run_batch_device_noalloc(self, keys: "DeviceBuffer", vals: "DeviceBuffer",
                                 ops: "DeviceBuffer", out_vals: "DeviceBuffer",
                                 out_found: "DeviceBuffer", n: Optional[int] = None,
                                 offset: int = 0) -> None:
        """run_batch_device through the entry point that allocates nothing (its scratch is a
        fixed device array per GPU); n <= NOALLOC_MAX_OPS."""
        n = keys.n - offset if n is None else int(n)
        _check(self._lib.stm_run_batch_device_noalloc(
            self._h, keys.ptr + 8 * offset, vals.ptr + 8 * offset, ops.ptr + offset, n,
            out_vals.ptr + 8 * offset, out_found.ptr + offset))

    def run_batches_device(self, keys: "DeviceBuffer", vals: "DeviceBuffer", ops: "DeviceBuffer",
                           out_vals: "DeviceBuffer", out_found: "DeviceBuffer", offsets) -> None:
        """Stream of device-resident batches: batch b is elements [offsets[b], offsets[b+1])."""
        off = _u64(offsets)
        if off.size < 1:
            raise ValueError("offsets needs n_batches + 1 entries")
        if off.size > 1 and int(off[-1]) > keys.n:
            raise ValueError("offsets run past the buffers")
        _check(self._lib.stm_run_batches_device(
            self._h, keys.ptr, vals.ptr, ops.ptr, _ptr(off), off.size - 1, out_vals.ptr,
            out_found.ptr))

    def capture_graphs(self) -> None:
        """Captures every batch size class's graph now (otherwise done on first use)."""
        _check(self._lib.stm_capture_graphs(self._h))

    def sync(self) -> None:
        _check(self._lib.stm_sync(self._h))

    def size(self) -> int:
        n = C.c_uint64()
        _check(self._lib.stm_size(self._h, C.byref(n)))
        return n.value

    def dump(self):
        """In-order (keys, values) of the whole tree."""
        n = C.c_size_t()
        _check(self._lib.stm_dump(self._h, None, None, 0, C.byref(n)))
        keys = np.empty(n.value, dtype=np.uint64)
        vals = np.empty(n.value, dtype=np.uint64)
        _check(self._lib.stm_dump(self._h, _ptr(keys), _ptr(vals), n.value, C.byref(n)))
        return keys[: n.value], vals[: n.value]

    def check(self) -> int:
        """Number of violated structural invariants (0 = well formed)."""
        bad = C.c_uint64()
        _check(self._lib.stm_check(self._h, C.byref(bad)))
        return bad.value

    def stats(self) -> dict:
        s = _Stats()
        _check(self._lib.stm_get_stats(self._h, C.byref(s)))
        d = {name: getattr(s, name) for name, _ in _Stats._fields_[:16]}
        d["phase_ms"] = {p: s.phase_ms[i] for i, p in enumerate(PHASES)}
        d["phase_launches"] = {p: s.phase_launches[i] for i, p in enumerate(PHASES)}
        d["node_reads"] = s.node_reads
        d["leaf_steps"] = s.leaf_steps
        d["sort_fallbacks"] = s.sort_fallbacks
        return d

    def reset_stats(self) -> None:
        _check(self._lib.stm_reset_stats(self._h))

    def retrain(self, train_steps: Optional[int] = None) -> None:
        """Rebuilds the leaf directory and trains the learned router now."""
        steps = self.config.train_steps if train_steps is None else int(train_steps)
        _check(self._lib.stm_retrain(self._h, steps))

    def route_info(self) -> dict:
        """Directory size and window hits / misses of the router on its own training pairs."""
        n, h, m = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(self._lib.stm_route_info(self._h, C.byref(n), C.byref(h), C.byref(m)))
        return {"leaves": n.value, "window_hits": h.value, "window_misses": m.value}

    def route_directory(self):
        """The leaf directory: (low fences uint64[L], leaf ids uint32[L]) in key order."""
        n = C.c_uint64()
        _check(self._lib.stm_route_directory(self._h, None, None, 0, C.byref(n)))
        f = np.empty(n.value, dtype=np.uint64)
        ids = np.empty(n.value, dtype=np.uint32)
        _check(self._lib.stm_route_directory(self._h, _ptr(f), _ptr(ids), n.value, C.byref(n)))
        return f, ids

    def stream(self) -> int:
        p = C.c_void_p()
        _check(self._lib.stm_get_stream(self._h, C.byref(p)))
        return p.value or 0


def mlp_gradients(x, t, params):
    """Loss and gradient of the router MLP computed by the device kernels (float32)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.size != 49 or x.size != t.size:
        raise ValueError("need 49 parameters and matching x / t")
    g = np.zeros(49, dtype=np.float32)
    loss = C.c_float()
    _check(load_library().stm_mlp_gradients(x.ctypes.data, t.ctypes.data, x.size, p.ctypes.data,
                                            g.ctypes.data, C.byref(loss)))
    return float(loss.value), g
