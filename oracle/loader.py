"""ctypes loader of the C oracle (liboracle.so) and numpy-facing helpers.

Test infrastructure: the oracle is the checker of the HIP library, never
part of it.  Imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg only.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")


def load_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = ctypes.CDLL(ORACLE_SO)
    P = ctypes.POINTER
    lib.hso_prim_id.argtypes = [ctypes.c_char_p]
    lib.hso_num_locs_mult.argtypes = [ctypes.c_int] * 3
    lib.hso_num_storage_locs.restype = ctypes.c_long
    lib.hso_num_storage_locs.argtypes = [ctypes.c_int] * 5
    for fn in (lib.hso_golden_literal, lib.hso_golden_closed):
        fn.argtypes = [ctypes.c_int] * 8 + [P(ctypes.c_int32), ctypes.c_long]
    lib.hso_region_passes.argtypes = [ctypes.c_int] * 5 + [P(ctypes.c_int64), ctypes.c_int]
    lib.hso_expected_counts.argtypes = [ctypes.c_int] * 5 + [P(ctypes.c_int64)] * 3
    lib.hso_seqlock_expected.argtypes = [ctypes.c_int] * 4 + [P(ctypes.c_uint32), ctypes.c_int]
    lib.hso_flags_rand.argtypes = [P(ctypes.c_float), ctypes.c_long, ctypes.c_uint]
    lib.hso_flags_dirty.restype = ctypes.c_uint64
    lib.hso_flags_dirty.argtypes = [P(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_float]
    lib.hso_sha1.argtypes = [ctypes.c_char_p, ctypes.c_size_t, P(ctypes.c_uint8)]
    lib.hso_uts_root.argtypes = [ctypes.c_int, P(ctypes.c_uint8)]
    lib.hso_uts_spawn.argtypes = [P(ctypes.c_uint8), ctypes.c_int, P(ctypes.c_uint8)]
    lib.hso_uts_rand.restype = ctypes.c_int32
    lib.hso_uts_rand.argtypes = [P(ctypes.c_uint8)]
    lib.hso_uts_geo_children.argtypes = [P(ctypes.c_uint8), ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int]
    lib.hso_uts_geo_thresholds.argtypes = [ctypes.c_double, P(ctypes.c_uint32), ctypes.c_int]
    lib.hso_uts_geo_count.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                      P(ctypes.c_uint64), P(ctypes.c_uint64), P(ctypes.c_int)]
    lib.hso_uts_ref_tree.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_uint32),
                                     P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_uint32),
                                     P(ctypes.c_uint32)]
    lib.hso_cpu_run.restype = ctypes.c_double
    lib.hso_cpu_run.argtypes = [ctypes.c_int] * 8 + [P(ctypes.c_int32), ctypes.c_long,
                                                     P(ctypes.c_int64)]
    lib.hso_cpu_barrier_run.restype = ctypes.c_double
    lib.hso_cpu_barrier_run.argtypes = [ctypes.c_int] * 9 + [P(ctypes.c_int32), ctypes.c_long]
    return lib


class OracleHelpers:
    """numpy-facing wrappers around the C oracle."""

    def __init__(self, lib):
        self.lib = lib

    def storage_words(self, prim, nw, ncu, wg, L):
        return self.lib.hso_num_storage_locs(prim, nw, ncu, wg, L)

    def golden(self, prim, L, nw, iters, reps, wg, ncu, mode, literal=False):
        n = self.storage_words(prim, nw, ncu, wg, L)
        out = np.zeros(n, np.int32)
        fn = self.lib.hso_golden_literal if literal else self.lib.hso_golden_closed
        rc = fn(prim, L, nw, iters, reps, wg, ncu, mode,
                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n)
        assert rc == 0, (prim, L, nw, iters, reps, wg, ncu, mode)
        return out

    def counts(self, prim, nw, iters, reps, ncu):
        w, r, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        assert self.lib.hso_expected_counts(prim, nw, iters, reps, ncu, ctypes.byref(w),
                                            ctypes.byref(r), ctypes.byref(e)) == 0
        return w.value, r.value, e.value
