"""Continuous-batching serving loop (stm_server_* of libstmtree.so) from Python.

    srv = Server(tree, hbm_budget=1 << 30)
    t = srv.submit(keys, vals, ops)          # returns at once
    vals, found, scan_keys, scan_vals, latency_s = srv.wait(t)

Requests are admitted FIFO into the batch being formed while its slot of the HBM budget
has room; a request's results equal running the requests one by one in submission order.
The server owns the tree until close().  `open_loop` drives the native open-loop harness
(Poisson arrivals, latency percentiles).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import StmError, StmTree, load_library


class _SrvConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("slots", C.c_uint32), ("max_batch_ops", C.c_uint32),
                ("scan_fraction", C.c_double), ("reserved", C.c_uint32 * 4)]


class _SrvStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("requests", "ops", "batches", "scans", "scan_pairs",
                                          "max_pending")] + [
        ("busy_s", C.c_double), ("hbm_budget", C.c_uint64), ("slot_ops", C.c_uint64),
        ("slot_pairs", C.c_uint64)]


class _Workload(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("offered_mops", C.c_double),
                ("ops_lo_log2", C.c_uint32), ("ops_hi_log2", C.c_uint32),
                ("lookup_frac", C.c_double), ("insert_frac", C.c_double),
                ("scan_lo", C.c_uint32), ("scan_hi", C.c_uint32), ("seed", C.c_uint64),
                ("warmup_requests", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class _Report(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("requests", "ops", "batches", "scans", "scan_pairs",
                                          "max_pending", "failed")] + [
        (n, C.c_double) for n in ("wall_s", "mops", "offered_mops", "p50_ms", "p90_ms", "p99_ms",
                                  "p999_ms", "max_ms", "mean_ms", "busy_frac", "mean_batch_ops")] + [
        ("hbm_budget", C.c_uint64), ("slot_ops", C.c_uint64), ("warmup_requests", C.c_uint64),
        ("warmup_p99_ms", C.c_double)]


class _Record(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("ops", C.c_uint64), ("pairs", C.c_uint64),
                ("req_offsets", C.POINTER(C.c_uint64)), ("keys", C.POINTER(C.c_uint64)),
                ("vals", C.POINTER(C.c_uint64)), ("ops_codes", C.POINTER(C.c_uint8)),
                ("out_vals", C.POINTER(C.c_uint64)), ("out_found", C.POINTER(C.c_uint8)),
                ("pair_offsets", C.POINTER(C.c_uint64)), ("scan_keys", C.POINTER(C.c_uint64)),
                ("scan_vals", C.POINTER(C.c_uint64)), ("tickets", C.POINTER(C.c_uint64)),
                ("opaque", C.c_void_p)]


_P = C.c_void_p
_bound = False


def _lib():
    global _bound
    lib = load_library()
    if not _bound:
        sig = {
            "stm_server_config_init": (C.c_int, [C.POINTER(_SrvConfig)]),
            "stm_server_create": (C.c_int, [_P, C.POINTER(_SrvConfig), C.POINTER(_P)]),
            "stm_server_submit": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, _P, _P, _P, C.c_size_t,
                                            C.c_double, C.POINTER(C.c_uint64)]),
            "stm_server_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_double)]),
            "stm_server_drain": (C.c_int, [_P]),
            "stm_server_get_stats": (C.c_int, [_P, C.POINTER(_SrvStats)]),
            "stm_server_destroy": (C.c_int, [_P]),
            "stm_server_now": (C.c_double, []),
            "stm_server_last_error": (C.c_char_p, []),
            "stm_serve_open_loop": (C.c_int, [_P, C.POINTER(_SrvConfig), C.POINTER(_Workload), _P,
                                              C.c_size_t, C.POINTER(_Report), _P]),
            "stm_serve_open_loop_record": (C.c_int, [_P, C.POINTER(_SrvConfig), C.POINTER(_Workload),
                                                     _P, C.c_size_t, C.POINTER(_Report), _P,
                                                     C.POINTER(C.POINTER(_Record))]),
            "stm_serve_record_free": (C.c_int, [C.POINTER(_Record)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = (_lib().stm_server_last_error() or b"").decode()
        if not msg:
            msg = (load_library().stm_last_error() or b"").decode()
        raise StmError(rc, msg)


def _config(hbm_budget: int, slots: int, max_batch_ops: int, scan_fraction: float) -> _SrvConfig:
    c = _SrvConfig()
    _check(_lib().stm_server_config_init(C.byref(c)))
    c.hbm_budget, c.slots, c.max_batch_ops = int(hbm_budget), int(slots), int(max_batch_ops)
    c.scan_fraction = float(scan_fraction)
    return c


@dataclass
class _Ticket:
    id: int
    arrays: tuple  # inputs and outputs kept alive until wait()


class Server:
    """Not thread-safe from Python's side: submit and wait from one thread."""

    def __init__(self, tree: StmTree, hbm_budget: int = 0, slots: int = 3, max_batch_ops: int = 0,
                 scan_fraction: float = 0.5):
        self._tree = tree
        self._c = _config(hbm_budget, slots, max_batch_ops, scan_fraction)
        h = C.c_void_p()
        _check(_lib().stm_server_create(tree._h, C.byref(self._c), C.byref(h)))
        self._h = h

    def submit(self, keys, vals, ops, arrival: Optional[float] = None) -> _Ticket:
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        v = np.ascontiguousarray(vals, dtype=np.uint64)
        o = np.ascontiguousarray(ops, dtype=np.uint8)
        if not (k.size == v.size == o.size):
            raise ValueError("keys, vals and ops differ in length")
        pairs = int(v[o == 3].sum()) if k.size else 0
        ov = np.zeros(k.size, dtype=np.uint64)
        of = np.zeros(k.size, dtype=np.uint8)
        sk = np.zeros(max(pairs, 1), dtype=np.uint64)
        sv = np.zeros(max(pairs, 1), dtype=np.uint64)
        t = C.c_uint64()
        _check(_lib().stm_server_submit(self._h, k.ctypes.data, v.ctypes.data, o.ctypes.data,
                                        k.size, ov.ctypes.data, of.ctypes.data, sk.ctypes.data,
                                        sv.ctypes.data, pairs,
                                        -1.0 if arrival is None else float(arrival), C.byref(t)))
        return _Ticket(t.value, (k, v, o, ov, of, sk[:pairs], sv[:pairs]))

    def wait(self, ticket: _Ticket):
        """(values, found, scan_keys, scan_vals, latency_s) of a submitted request."""
        lat = C.c_double()
        _check(_lib().stm_server_wait(self._h, ticket.id, C.byref(lat)))
        _, _, _, ov, of, sk, sv = ticket.arrays
        return ov, of, sk, sv, lat.value

    def stats(self) -> Dict:
        s = _SrvStats()
        _check(_lib().stm_server_get_stats(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in _SrvStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().stm_server_destroy(self._h)
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


def now() -> float:
    return _lib().stm_server_now()


def open_loop(tree: StmTree, loaded_sorted, requests: int, offered_mops: float,
              ops_log2=(0, 16), lookup_frac=0.90, insert_frac=0.08, scan_len=(16, 256),
              seed: int = 1, hbm_budget: int = 0, slots: int = 3, max_batch_ops: int = 0,
              scan_fraction: float = 0.5, return_latencies: bool = False,
              warmup: int = 0, record: bool = False) -> Dict:
    """Native open-loop run of the serving loop; returns its report (latencies in ms).  The
    first `warmup` requests run on the same arrival process before the measured ones and are
    left out of the report's figures (their p99 is `warmup_p99_ms`).  With `record` the report
    also holds "record": every request (warm-up ones first) with its answers and admission
    ticket, as numpy arrays (keys, vals, ops, req_offsets, out_vals, out_found, pair_offsets,
    scan_keys, scan_vals, tickets)."""
    ls = np.ascontiguousarray(loaded_sorted, dtype=np.uint64)
    w = _Workload(requests=int(requests), offered_mops=float(offered_mops),
                  ops_lo_log2=int(ops_log2[0]), ops_hi_log2=int(ops_log2[1]),
                  lookup_frac=float(lookup_frac), insert_frac=float(insert_frac),
                  scan_lo=int(scan_len[0]), scan_hi=int(scan_len[1]), seed=int(seed),
                  warmup_requests=int(warmup))
    c = _config(hbm_budget, slots, max_batch_ops, scan_fraction)
    rep = _Report()
    lat = np.zeros(int(requests), dtype=np.float64)
    rec = None
    if record:
        rp = C.POINTER(_Record)()
        _check(_lib().stm_serve_open_loop_record(tree._h, C.byref(c), C.byref(w), ls.ctypes.data,
                                                 ls.size, C.byref(rep), lat.ctypes.data, C.byref(rp)))
        r = rp.contents
        n_req, n_ops, n_pairs = int(r.requests), int(r.ops), int(r.pairs)

        def copy(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].astype(dt, copy=True)

        rec = {"req_offsets": copy(r.req_offsets, n_req + 1, np.uint64),
               "keys": copy(r.keys, n_ops, np.uint64), "vals": copy(r.vals, n_ops, np.uint64),
               "ops": copy(r.ops_codes, n_ops, np.uint8),
               "out_vals": copy(r.out_vals, n_ops, np.uint64),
               "out_found": copy(r.out_found, n_ops, np.uint8),
               "pair_offsets": copy(r.pair_offsets, n_req + 1, np.uint64),
               "scan_keys": copy(r.scan_keys, n_pairs, np.uint64),
               "scan_vals": copy(r.scan_vals, n_pairs, np.uint64),
               "tickets": copy(r.tickets, n_req, np.uint64)}
        _lib().stm_serve_record_free(rp)
    else:
        _check(_lib().stm_serve_open_loop(tree._h, C.byref(c), C.byref(w), ls.ctypes.data,
                                          ls.size, C.byref(rep), lat.ctypes.data))
    out = {n: getattr(rep, n) for n, _ in _Report._fields_}
    if rec is not None:
        out["record"] = rec
    if return_latencies:
        out["latencies_ms"] = lat
    return out
