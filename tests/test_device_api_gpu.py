"""The public device API (include/heterosync_device.h) used from a user
kernel: tests/device_api/user_kernels.hip re-implements the reference's
microbenchmark loops with the API's mutexes, semaphores and barriers and a
plain-load/store in-place-increment critical section; the final storage
must equal the oracle's closed form (inc mode) bit for bit, and the
barriers' episode counts must be exact."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "device_api",
                   "libhs_device_api_test.so")


@pytest.fixture(scope="module")
def user(dev):
    assert os.path.exists(LIB), f"{LIB} is not built (heterosync_amd.build)"
    lib = ctypes.CDLL(LIB)
    lib.hsu_run.argtypes = [ctypes.c_int] * 8 + [ctypes.POINTER(ctypes.c_int32), ctypes.c_long,
                                                 ctypes.c_double, ctypes.POINTER(ctypes.c_uint32)]
    return lib


def run(user, ohelp, dev, test, kind, prim, nw, L, iters, sem_count=0, wg=64):
    ncu = dev.info.num_cu
    n = ohelp.storage_words(prim, nw, ncu, wg, L)
    s = np.arange(n, dtype=np.int32)
    ep = ctypes.c_uint32()
    rc = user.hsu_run(test, kind, sem_count, nw, ncu, L, iters, wg,
                      s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n, 20.0, ctypes.byref(ep))
    assert rc == 0, rc
    want = ohelp.golden(prim, L, nw, iters, 1, wg, ncu, 1)
    assert np.array_equal(s, want), (test, kind, prim, nw, L, iters, int((s != want).sum()))
    return ep.value


# wg 256: four waves share each critical section and each barrier call
@pytest.mark.parametrize("wg", [64, 256])
@pytest.mark.parametrize("kind,prim", [(0, 4), (1, 5), (3, 7)])  # spin, EBO, ticket
@pytest.mark.parametrize("nw,L,iters", [(32, 10, 50), (300, 3, 20)])
def test_user_kernel_mutex(user, ohelp, dev, kind, prim, nw, L, iters, wg):
    run(user, ohelp, dev, 0, kind, prim, nw, L, iters, wg=wg)


@pytest.mark.parametrize("wg", [64, 256])
@pytest.mark.parametrize("kind,prim,count", [(0, 9, 2), (0, 11, 120), (1, 14, 10)])
@pytest.mark.parametrize("nw,L,iters", [(64, 10, 20), (512, 4, 10)])
def test_user_kernel_semaphore(user, ohelp, dev, kind, prim, count, nw, L, iters, wg):
    run(user, ohelp, dev, 1, kind, prim, nw, L, iters, sem_count=count, wg=wg)


@pytest.mark.parametrize("test,prim", [(2, 0), (3, 2)])  # atomic tree, lock-free tree
@pytest.mark.parametrize("per_cu,L,iters,wg", [(1, 10, 100, 64), (4, 10, 200, 64),
                                               (8, 2, 50, 64), (1, 10, 100, 256),
                                               (4, 10, 200, 256)])
def test_user_kernel_barrier(user, ohelp, dev, test, prim, per_cu, L, iters, wg):
    ep = run(user, ohelp, dev, test, 0, prim, per_cu * dev.info.num_cu, L, iters, wg=wg)
    assert ep == iters
