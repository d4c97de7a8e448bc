"""CPU tests of the C ABI: the library loads, exports every declared symbol,
and its host-only functions (primitive table, storage sizes, error paths)
behave; the command lines print usage and fail on bad arguments.  No
compute call is made here (this host has no GPU)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import heterosync_amd as hs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "heterosync.h")).read()
    return sorted(set(re.findall(r"\b(hs_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = hs.lib()
    decl = declared_symbols()
    assert len(decl) >= 17
    missing = [s for s in decl if not hasattr(lib, s)]
    assert not missing
    assert sorted(hs.EXPORTED) == decl


def test_prim_table_and_sizes(oracle):
    from tests.test_oracle import REF_NAMES
    for name, pid in REF_NAMES.items():
        assert hs.prim_id(name) == pid
        assert hs.prim_name(pid) == name
    assert hs.prim_id("bogus") == -1
    lib = hs.lib()
    supported = list(range(36)) + list(range(37, 43))
    assert [p for p in range(47) if lib.hs_prim_supported(p)] == supported
    for pid in supported:
        for nw, ncu, wg, L in [(32, 256, 64, 10), (1024, 256, 256, 10), (3, 4, 64, 0),
                               (512, 256, 64, 100)]:
            assert hs.storage_words(pid, nw, ncu, wg, L) == \
                oracle.hso_num_storage_locs(pid, nw, ncu, wg, L)
    assert hs.storage_words(7, 32, 256, 64, 10) == 704  # C1: 64 * 11


def test_open_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert hs.lib().hs_open(0, ctypes.byref(h)) == -6  # HS_ENODEV
    assert hs.lib().hs_strerror(-6) == b"no HIP device"
    assert hs.lib().hs_syncprim_run(None, None, None, 0, None) == -1  # HS_EINVAL


@pytest.mark.parametrize("cli,args", [("allSyncPrims-1kernel", ["faMutex", "10"]),
                                      ("seqlocks", ["1"]), ("flags", ["1", "2"]),
                                      ("refCounter", []), ("splitCounter", ["1"])])
def test_cli_usage(cli, args):
    path = os.path.join(hs.BIN_DIR, cli)
    assert os.path.exists(path), path
    p = subprocess.run([path, *args], capture_output=True, text=True, timeout=60)
    assert p.returncode == 255  # exit(-1), as the reference
    assert "./" in p.stderr


@pytest.mark.parametrize("mode", [0, 1])
def test_library_self_check_equals_oracle(ohelp, mode):
    """hs_syncprim_expected (the library's closed form of the reference's
    host golden, used by the command line's PASSED! check and by the
    measurement harnesses) equals the oracle's closed form at every
    primitive id, and the oracle's literal golden where that applies."""
    from tests.test_oracle import PRIO_EBO, SUPPORTED
    for prim in SUPPORTED:
        for nw, ncu, wg, L, iters, reps in [(4, 4, 64, 2, 3, 2), (12, 4, 64, 0, 2, 1),
                                            (3, 4, 128, 1, 5, 2), (64, 16, 64, 3, 2, 3),
                                            (1024, 256, 256, 10, 1000, 10)]:
            if hs.storage_words(prim, nw, ncu, wg, L) < 0:
                continue
            got = hs.expected_storage(prim, L, nw, iters, reps, wg, ncu, mode)
            assert np.array_equal(got, ohelp.golden(prim, L, nw, iters, reps, wg, ncu, mode)), \
                (prim, nw, ncu, wg, L, mode)
            if nw <= 64 and prim not in PRIO_EBO:
                lit = ohelp.golden(prim, L, nw, iters, reps, wg, ncu, mode, literal=True)
                assert np.array_equal(got, lit), (prim, nw, L, mode)


def test_library_self_check_rejects_bad_arguments():
    lib = hs.lib()
    out = np.zeros(704, np.int32)
    p = out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    assert lib.hs_syncprim_expected(7, 10, 32, 100, 10, 64, 256, 0, p, 704) == 0
    assert lib.hs_syncprim_expected(7, 10, 32, 100, 10, 64, 256, 0, p, 703) == -1  # size
    assert lib.hs_syncprim_expected(36, 10, 32, 100, 10, 64, 256, 0, p, 704) == -1  # reserved
    assert lib.hs_syncprim_expected(7, 10, 32, 100, 10, 64, 256, 2, p, 704) == -1  # mode
