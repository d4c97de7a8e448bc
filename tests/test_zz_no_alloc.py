"""hs_syncprim_run_noalloc (the entry point that allocates nothing: its
state lives in a static per-device scratch array) against hs_syncprim_run
on the same workloads: storage bit-exact, the same counts and idle state,
and the oracle's closed form, on two calls in a row with the second on a
smaller input (the scratch keeps the first call's contents).  The caller's
device storage is allocated here through the HIP runtime."""
import ctypes

import numpy as np
import pytest

import heterosync_amd as hs

pytestmark = pytest.mark.gpu

HS_ESCRATCH = -8


@pytest.fixture(scope="module")
def hip():
    rt = ctypes.CDLL("libamdhip64.so")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return rt


def noalloc(dev, hip, prim, L, nw, iters, wg=64, reps=2, mode=1, scope=0):
    pid = hs.prim_id(prim)
    n = hs.storage_words(pid, nw, dev.info.num_cu, wg, L)
    host = np.arange(n, dtype=np.int32)
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), n * 4) == 0
    try:
        assert hip.hipMemcpy(d, host.ctypes.data, n * 4, 1) == 0  # host to device
        cfg = hs.SyncPrimCfg(pid, L, nw, iters, wg, reps, mode, 0, 0, 0, 20.0, scope)
        res = hs.SyncPrimResult()
        rc = hs.lib().hs_syncprim_run_noalloc(dev._h, ctypes.byref(cfg), d, n, ctypes.byref(res))
        if rc == 0:
            assert hip.hipMemcpy(host.ctypes.data, d, n * 4, 2) == 0  # device to host
    finally:
        hip.hipFree(d)
    return rc, host, res


# (prim, first call (L, WGs per CU or WGs, iters, wg), second call, smaller)
CASES = [
    ("atomicTreeBarrUniq", (10, "4cu", 200, 256), (4, "1cu", 50, 64)),
    ("atomicTreeBarrUniqLocalExch", (4, "2cu", 100, 256), (2, "1cu", 20, 64)),
    ("lfTreeBarrUniq", (10, "4cu", 200, 256), (3, "2cu", 30, 64)),
    ("atomicTreeBarrUniqNaiveSRB", (4, "2cu", 50, 64), (1, 32, 10, 64)),
    ("faMutex", (10, 32, 100, 64), (3, 8, 10, 64)),
    ("spinMutexEBO", (10, 64, 30, 64), (1, 4, 5, 64)),
    ("sleepMutex", (4, 128, 100, 64), (2, 16, 10, 64)),
    ("spinSem10", (10, "2cu", 20, 64), (4, 64, 10, 64)),
    ("spinSemEBOUniq2", (4, "2cu", 10, 64), (2, "1cu", 5, 64)),
    ("PRspinSem10", (4, "2cu", 10, 64), (2, 64, 5, 64)),
]


def nwgs(dev, x):
    return x if isinstance(x, int) else int(x[:-2]) * dev.info.num_cu


@pytest.mark.parametrize("prim,first,second", CASES)
def test_noalloc_equals_allocating_entry_point(dev, hip, ohelp, prim, first, second):
    ncu = dev.info.num_cu
    pid = hs.prim_id(prim)
    for L, nw, iters, wg in (first, second):
        nw = nwgs(dev, nw)
        rc, got, res = noalloc(dev, hip, prim, L, nw, iters, wg=wg)
        assert rc == 0, (prim, rc)
        want, ref = dev.syncprim(pid, L, nw, iters, wg_size=wg, reps=2, mode=1, timeout_s=20.0)
        assert np.array_equal(got, want), (prim, L, nw, iters)
        assert np.array_equal(got, ohelp.golden(pid, L, nw, iters, 2, wg, ncu, 1))
        assert (res.cs_entries, res.rd_entries, res.episodes, res.state_ok) == \
            (ref.cs_entries, ref.rd_entries, ref.episodes, 1)
        assert res.sync_scope_used == ref.sync_scope_used


def test_noalloc_rejects_state_beyond_the_scratch(dev, hip):
    # a queue lock of 512 WGs x 1000 iterations x 2 launches: a ring of
    # 1,024,001 one-word slots exceeds the 2^20-word scratch (refused before
    # any launch)
    rc, _, _ = noalloc(dev, hip, "sleepMutex", 1, 512, 1000)
    assert rc == HS_ESCRATCH
    assert hs.lib().hs_strerror(HS_ESCRATCH).decode().startswith("workload needs more")
    # and the next call still runs
    rc, got, res = noalloc(dev, hip, "faMutex", 2, 8, 5)
    assert rc == 0 and res.state_ok == 1
