"""heterosync_amd -- MI355X-native HeteroSync synchronization hot path.

The product is the C-ABI library libheterosync_amd.so (include/heterosync.h)
plus the drop-in command lines under heterosync_amd/bin/.  This module is a
thin ctypes binding to that library for tests, the benchmark and
Python callers.  There is no fallback: if the library is missing or cannot
be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libheterosync_amd.so")
BIN_DIR = os.path.join(HERE, "bin")

HS_OK = 0
ERRORS = {
    -1: "HS_EINVAL", -2: "HS_EHIP", -3: "HS_ENOMEM", -4: "HS_EOCCUPANCY",
    -5: "HS_ETIMEOUT", -6: "HS_ENODEV", -7: "HS_EUNSUPPORTED", -8: "HS_ESCRATCH",
}

# Every symbol include/heterosync.h declares.
EXPORTED = [
    "hs_strerror", "hs_open", "hs_close", "hs_device_info", "hs_prim_id",
    "hs_prim_name", "hs_prim_supported", "hs_storage_words", "hs_syncprim_expected",
    "hs_syncprim_run", "hs_syncprim_run_noalloc",
    "hs_seqlock_run", "hs_flags_run", "hs_refcounter_run", "hs_splitcounter_run",
    "hs_uts_ref_run", "hs_uts_geo_run", "hs_uts_ref_run_cpu", "hs_uts_geo_run_cpu",
    "hs_batch_run", "hs_batch_autotune",
    "hs_uts_serve",
]


class HSError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, code)} ({_lib().hs_strerror(code).decode()})")


class SyncPrimCfg(ctypes.Structure):
    _fields_ = [
        ("prim", ctypes.c_int), ("num_ldst", ctypes.c_int), ("num_wgs", ctypes.c_int),
        ("num_cs_iters", ctypes.c_int), ("wg_size", ctypes.c_int), ("reps", ctypes.c_int),
        ("mode", ctypes.c_int), ("backoff_min", ctypes.c_int), ("backoff_max", ctypes.c_int),
        ("poll_sleep", ctypes.c_int), ("timeout_s", ctypes.c_double),
        ("sync_scope", ctypes.c_int),
    ]


class SyncPrimResult(ctypes.Structure):
    _fields_ = [
        ("roi_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
        ("cs_entries", ctypes.c_int64), ("rd_entries", ctypes.c_int64),
        ("episodes", ctypes.c_int64), ("state_ok", ctypes.c_int),
        ("lock_words", ctypes.c_uint32 * 8), ("sync_scope_used", ctypes.c_int),
    ]


class RelaxedResult(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("atomic_ops", ctypes.c_uint64),
                ("aborted", ctypes.c_int)]


class UtsResult(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("nodes", ctypes.c_uint64),
                ("leaves", ctypes.c_uint64), ("leaf_sum", ctypes.c_uint64),
                ("max_depth", ctypes.c_int), ("aborted", ctypes.c_int)]


class UtsCpuResult(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint64), ("deliveries", ctypes.c_uint64),
                ("busy_s", ctypes.c_double), ("threads", ctypes.c_int)]


class BatchInstance(ctypes.Structure):
    _fields_ = [("prim", ctypes.c_int), ("num_wgs", ctypes.c_int),
                ("num_cs_iters", ctypes.c_int), ("num_ldst", ctypes.c_int)]


class BatchCfg(ctypes.Structure):
    _fields_ = [("wg_size", ctypes.c_int), ("reps", ctypes.c_int), ("mode", ctypes.c_int),
                ("backoff_min", ctypes.c_int), ("backoff_max", ctypes.c_int),
                ("poll_sleep", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("persistent_wgs", ctypes.c_int), ("timeout_s", ctypes.c_double),
                ("probe", ctypes.c_int), ("cohort", ctypes.c_int)]


class BatchResult(ctypes.Structure):
    _fields_ = [("roi_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("acquires", ctypes.c_int64), ("all_ok", ctypes.c_int)]


class UtsRequest(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int), ("gen_mx", ctypes.c_int), ("b0", ctypes.c_double),
                ("arrival_s", ctypes.c_double)]


class UtsResponse(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint64), ("leaves", ctypes.c_uint64),
                ("max_depth", ctypes.c_int), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double)]


class ServeCfg(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("num_blocks", ctypes.c_int),
                ("deque_nodes", ctypes.c_int), ("timeout_s", ctypes.c_double)]


class ServeSummary(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("total_nodes", ctypes.c_uint64),
                ("p50_s", ctypes.c_double), ("p99_s", ctypes.c_double),
                ("cut_batches", ctypes.c_uint32), ("deque_high_water", ctypes.c_uint32),
                ("deque_soft", ctypes.c_uint32)]


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m heterosync_amd.build` "
                "(there is no non-HIP fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        lib.hs_strerror.restype = ctypes.c_char_p
        lib.hs_strerror.argtypes = [ctypes.c_int]
        lib.hs_open.argtypes = [ctypes.c_int, P(ctypes.c_void_p)]
        lib.hs_close.argtypes = [ctypes.c_void_p]
        lib.hs_close.restype = None
        lib.hs_device_info.argtypes = [ctypes.c_void_p, P(ctypes.c_int), P(ctypes.c_int),
                                       P(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
        lib.hs_prim_id.argtypes = [ctypes.c_char_p]
        lib.hs_prim_name.restype = ctypes.c_char_p
        lib.hs_prim_name.argtypes = [ctypes.c_int]
        lib.hs_prim_supported.argtypes = [ctypes.c_int]
        lib.hs_storage_words.restype = ctypes.c_long
        lib.hs_storage_words.argtypes = [ctypes.c_int] * 5
        lib.hs_syncprim_expected.argtypes = [ctypes.c_int] * 8 + [P(ctypes.c_int32),
                                                                  ctypes.c_long]
        lib.hs_syncprim_run.argtypes = [ctypes.c_void_p, P(SyncPrimCfg), P(ctypes.c_int32),
                                        ctypes.c_long, P(SyncPrimResult)]
        lib.hs_syncprim_run_noalloc.argtypes = [ctypes.c_void_p, P(SyncPrimCfg),
                                                ctypes.c_void_p, ctypes.c_long,
                                                P(SyncPrimResult)]
        lib.hs_seqlock_run.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [
            P(ctypes.c_uint32), ctypes.c_int, P(ctypes.c_int32), P(ctypes.c_uint64),
            P(RelaxedResult)]
        lib.hs_flags_run.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4 + [
            ctypes.c_float, ctypes.c_int, P(ctypes.c_float), P(ctypes.c_uint32),
            P(ctypes.c_uint32), P(RelaxedResult)]
        lib.hs_refcounter_run.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [
            P(ctypes.c_uint32)] * 4 + [P(RelaxedResult)]
        lib.hs_splitcounter_run.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4 + [
            P(ctypes.c_uint32), P(ctypes.c_uint32), P(RelaxedResult)]
        lib.hs_uts_ref_run.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_uint32),
                                       P(ctypes.c_int32), P(ctypes.c_int32), ctypes.c_int,
                                       P(UtsResult)]
        lib.hs_uts_geo_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, P(UtsResult)]
        lib.hs_uts_ref_run_cpu.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_uint32),
                                           P(ctypes.c_int32), P(ctypes.c_int32), ctypes.c_int,
                                           ctypes.c_int, P(UtsResult), P(UtsCpuResult)]
        lib.hs_uts_geo_run_cpu.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           P(UtsResult), P(UtsCpuResult)]
        lib.hs_batch_run.argtypes = [ctypes.c_void_p, P(BatchInstance), ctypes.c_int,
                                     P(BatchCfg), P(P(ctypes.c_int32)), P(ctypes.c_int64),
                                     P(BatchResult)]
        lib.hs_batch_autotune.argtypes = [ctypes.c_void_p, P(BatchInstance), ctypes.c_int,
                                          P(BatchCfg), P(ctypes.c_double)]
        lib.hs_uts_serve.argtypes = [ctypes.c_void_p, P(UtsRequest), ctypes.c_int,
                                     P(ServeCfg), P(UtsResponse), P(ServeSummary)]
        _LIB = lib
    return _LIB


def lib():
    """The loaded ctypes library (raises if it is not built)."""
    return _lib()


def _check(rc: int, what: str) -> None:
    if rc != HS_OK:
        raise HSError(rc, what)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def prim_id(name: str) -> int:
    return _lib().hs_prim_id(name.encode())


def prim_name(pid: int):
    r = _lib().hs_prim_name(pid)
    return r.decode() if r else None


def storage_words(prim: int, num_wgs: int, num_cu: int, wg_size: int, num_ldst: int) -> int:
    return _lib().hs_storage_words(prim, num_wgs, num_cu, wg_size, num_ldst)


def expected_storage(prim, num_ldst, num_wgs, num_cs_iters, reps, wg_size, num_cu, mode):
    """The storage a correct run leaves (hs_syncprim_expected: the reference
    binary's own closed-form self-check).  Host code; needs no device."""
    pid = prim if isinstance(prim, int) else prim_id(prim)
    n = storage_words(pid, num_wgs, num_cu, wg_size, num_ldst)
    if n < 0:
        raise HSError(n, "hs_storage_words")
    out = np.empty(n, np.int32)
    _check(_lib().hs_syncprim_expected(pid, num_ldst, num_wgs, num_cs_iters, reps, wg_size,
                                       num_cu, mode, _ptr(out, ctypes.c_int32), n),
           "hs_syncprim_expected")
    return out


@dataclass
class DeviceInfo:
    num_cu: int
    max_threads: int
    max_wgs: int
    arch: str


class Device:
    """An open device context (hs_open / hs_close)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        _check(_lib().hs_open(device, ctypes.byref(self._h)), "hs_open")
        ncu, mt, mw = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        arch = ctypes.create_string_buffer(64)
        _check(_lib().hs_device_info(self._h, ctypes.byref(ncu), ctypes.byref(mt),
                                     ctypes.byref(mw), arch, 64), "hs_device_info")
        self.info = DeviceInfo(ncu.value, mt.value, mw.value, arch.value.decode())

    def close(self):
        if self._h:
            _lib().hs_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- syncPrims -------------------------------------------------------
    def syncprim(self, prim, num_ldst, num_wgs, num_cs_iters, wg_size=64, reps=10,
                 mode=0, backoff=(0, 0), poll_sleep=0, timeout_s=10.0, storage=None,
                 sync_scope=0):
        """Runs one microbenchmark; returns (storage, SyncPrimResult)."""
        pid = prim if isinstance(prim, int) else prim_id(prim)
        n = storage_words(pid, num_wgs, self.info.num_cu, wg_size, num_ldst)
        if n < 0:
            raise HSError(n, "hs_storage_words")
        if storage is None:
            storage = np.arange(n, dtype=np.int32)
        cfg = SyncPrimCfg(pid, num_ldst, num_wgs, num_cs_iters, wg_size, reps, mode,
                          backoff[0], backoff[1], poll_sleep, timeout_s, sync_scope)
        res = SyncPrimResult()
        _check(_lib().hs_syncprim_run(self._h, ctypes.byref(cfg), _ptr(storage, ctypes.c_int32),
                                      n, ctypes.byref(res)), f"hs_syncprim_run({pid})")
        return storage, res

    # ---- relaxed atomics -------------------------------------------------
    def seqlock(self, num_wgs, wg_size, group_size, writer_period=16, iters=10, variant=0):
        n_seq = max(group_size, num_wgs // max(group_size, 1))
        seq = np.zeros(n_seq, np.uint32)
        out = np.full(num_wgs * wg_size, -1, np.int32)
        torn = ctypes.c_uint64()
        res = RelaxedResult()
        _check(_lib().hs_seqlock_run(self._h, num_wgs, wg_size, group_size, writer_period,
                                     iters, variant, _ptr(seq, ctypes.c_uint32), n_seq,
                                     _ptr(out, ctypes.c_int32), ctypes.byref(torn),
                                     ctypes.byref(res)), "hs_seqlock_run")
        return seq, out, torn.value, res

    def flags(self, num_tbs, tb_size, num_accs_per_thr, num_repeats, cutoff, num_iters,
              rand_arr):
        out = np.zeros(num_tbs * tb_size, np.uint32)
        dirty = ctypes.c_uint32()
        res = RelaxedResult()
        r = np.ascontiguousarray(rand_arr, np.float32)
        _check(_lib().hs_flags_run(self._h, num_tbs, tb_size, num_accs_per_thr, num_repeats,
                                   ctypes.c_float(cutoff), num_iters, _ptr(r, ctypes.c_float),
                                   _ptr(out, ctypes.c_uint32), ctypes.byref(dirty),
                                   ctypes.byref(res)), "hs_flags_run")
        return out, dirty.value, res

    def refcounter(self, sharers, num_tbs, tb_size, num_repeats, spacing=-1, strict=False):
        nthr = num_tbs * tb_size
        c0 = np.zeros(nthr // sharers, np.uint32)
        c1 = np.zeros_like(c0)
        d0 = np.zeros(nthr, np.uint32)
        d1 = np.zeros_like(d0)
        res = RelaxedResult()
        _check(_lib().hs_refcounter_run(self._h, sharers, num_tbs, tb_size, num_repeats,
                                        spacing, int(strict), _ptr(c0, ctypes.c_uint32),
                                        _ptr(c1, ctypes.c_uint32), _ptr(d0, ctypes.c_uint32),
                                        _ptr(d1, ctypes.c_uint32), ctypes.byref(res)),
               "hs_refcounter_run")
        return c0, c1, d0, d1, res

    def splitcounter(self, num_tbs, tb_size, iters, read_every=0):
        nthr = num_tbs * tb_size
        counters = np.zeros(nthr // 2, np.uint32)
        out = np.zeros(nthr, np.uint32)
        res = RelaxedResult()
        _check(_lib().hs_splitcounter_run(self._h, num_tbs, tb_size, iters, read_every,
                                          _ptr(counters, ctypes.c_uint32),
                                          _ptr(out, ctypes.c_uint32), ctypes.byref(res)),
               "hs_splitcounter_run")
        return counters, out, res

    # ---- UTS ---------------------------------------------------------------
    def uts_ref(self, value, child_a, child_b, num_blocks=0):
        res = UtsResult()
        v = np.ascontiguousarray(value, np.uint32)
        a = np.ascontiguousarray(child_a, np.int32)
        b = np.ascontiguousarray(child_b, np.int32)
        _check(_lib().hs_uts_ref_run(self._h, len(v), _ptr(v, ctypes.c_uint32),
                                     _ptr(a, ctypes.c_int32), _ptr(b, ctypes.c_int32),
                                     num_blocks, ctypes.byref(res)), "hs_uts_ref_run")
        return res

    def uts_geo(self, seed, b0, gen_mx, num_blocks=0):
        res = UtsResult()
        _check(_lib().hs_uts_geo_run(self._h, seed, b0, gen_mx, num_blocks,
                                     ctypes.byref(res)), "hs_uts_geo_run")
        return res

    def uts_geo_cpu(self, seed, b0, gen_mx, cpu_threads, num_blocks=0):
        """A geometric tree with host worker threads taking part; returns
        (UtsResult, UtsCpuResult)."""
        res, cpu = UtsResult(), UtsCpuResult()
        _check(_lib().hs_uts_geo_run_cpu(self._h, seed, b0, gen_mx, num_blocks, cpu_threads,
                                         ctypes.byref(res), ctypes.byref(cpu)),
               "hs_uts_geo_run_cpu")
        return res, cpu

    def uts_ref_cpu(self, value, child_a, child_b, cpu_threads, num_blocks=0):
        res, cpu = UtsResult(), UtsCpuResult()
        v = np.ascontiguousarray(value, np.uint32)
        a = np.ascontiguousarray(child_a, np.int32)
        b = np.ascontiguousarray(child_b, np.int32)
        _check(_lib().hs_uts_ref_run_cpu(self._h, len(v), _ptr(v, ctypes.c_uint32),
                                         _ptr(a, ctypes.c_int32), _ptr(b, ctypes.c_int32),
                                         num_blocks, cpu_threads, ctypes.byref(res),
                                         ctypes.byref(cpu)), "hs_uts_ref_run_cpu")
        return res, cpu

    def uts_serve(self, requests, slots=0, num_blocks=0, deque_nodes=0, timeout_s=60.0):
        n = len(requests)
        req = (UtsRequest * n)(*[UtsRequest(*r) for r in requests])
        resp = (UtsResponse * n)()
        cfg = ServeCfg(slots, num_blocks, deque_nodes, timeout_s)
        summ = ServeSummary()
        _check(_lib().hs_uts_serve(self._h, req, n, ctypes.byref(cfg), resp,
                                   ctypes.byref(summ)), "hs_uts_serve")
        return list(resp), summ

    # ---- batched ---------------------------------------------------------
    def batch_autotune(self, instances, reps=10, mode=0, use_graph=1, timeout_s=20.0):
        """Returns the tuned knobs as a dict: wg_size, persistent_wgs,
        poll_sleep, backoff_min, backoff_max, probe, cohort, and best_ms (one
        launch of the whole batch in the chosen configuration, from the
        cohort stage).  batch(**{k: v for k in TUNED}) runs them."""
        n = len(instances)
        inst = (BatchInstance * n)(*[BatchInstance(*i) for i in instances])
        cfg = BatchCfg(64, reps, mode, 0, 0, 0, use_graph, 0, timeout_s, 0, 0)
        best = ctypes.c_double()
        _check(_lib().hs_batch_autotune(self._h, inst, n, ctypes.byref(cfg), ctypes.byref(best)),
               "hs_batch_autotune")
        return {"wg_size": cfg.wg_size, "persistent_wgs": cfg.persistent_wgs,
                "poll_sleep": cfg.poll_sleep, "backoff_min": cfg.backoff_min,
                "backoff_max": cfg.backoff_max, "probe": cfg.probe, "cohort": cfg.cohort,
                "best_ms": best.value}

    @staticmethod
    def tuned_kwargs(t):
        """batch() keyword arguments for a batch_autotune() result."""
        return dict(wg_size=t["wg_size"], persistent_wgs=t["persistent_wgs"],
                    poll_sleep=t["poll_sleep"], backoff=(t["backoff_min"], t["backoff_max"]),
                    probe=t["probe"], cohort=t["cohort"])

    def batch(self, instances, storages, wg_size=64, reps=10, mode=0, backoff=(0, 0),
              poll_sleep=0, use_graph=1, persistent_wgs=0, timeout_s=20.0, probe=0,
              cohort=0):
        n = len(instances)
        inst = (BatchInstance * n)(*[BatchInstance(*i) for i in instances])
        ptrs = (ctypes.POINTER(ctypes.c_int32) * n)(*[_ptr(s, ctypes.c_int32) for s in storages])
        entries = np.zeros(n, np.int64)
        cfg = BatchCfg(wg_size, reps, mode, backoff[0], backoff[1], poll_sleep, use_graph,
                       persistent_wgs, timeout_s, probe, cohort)
        res = BatchResult()
        _check(_lib().hs_batch_run(self._h, inst, n, ctypes.byref(cfg), ptrs,
                                   _ptr(entries, ctypes.c_int64), ctypes.byref(res)),
               "hs_batch_run")
        return entries, res
