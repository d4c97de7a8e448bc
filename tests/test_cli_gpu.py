"""The drop-in command lines on the GPU.  Each binary runs its own
self-check (the closed-form golden of allSyncPrims-1kernel, the relaxed-
atomics expected counts, the UTS host golden) and must report success the
way the reference's binaries do; the invocations are the reference's run
lines (hip/syncPrims/uvm/syncPrims_results.sh, cuda/relaxedAtomics/uvm/*/run,
cuda/uts/uvm/run) with work-group sizes raised to the 64-lane wavefront."""
import os
import subprocess

import pytest

import heterosync_amd as hs
from tests.test_oracle import REF_NAMES

pytestmark = pytest.mark.gpu


def run(cli, *args, timeout=120):
    return subprocess.run([os.path.join(hs.BIN_DIR, cli), *map(str, args)],
                          capture_output=True, text=True, timeout=timeout)


# main.hip.cpp:2177-2340, including the reference's "2" for ids 11, 19, 27
BANNER = {0: "atomic_tree_barrier_003", 7: "fetchadd_mutex_003", 11: "spin_lock_sem_002_003",
          15: "ebo_sem_120_003", 19: "prior_spin_lock_sem_002_003",
          23: "fetchadd_mutex_uniq_003", 27: "spin_lock_sem_uniq_002_003",
          35: "prior_ebo_sem_120_003", 42: "atomic_tree_barrier_localExch_NaiveAllSRB_003"}


@pytest.mark.parametrize("name,pid", sorted(REF_NAMES.items(), key=lambda kv: kv[1]))
def test_all_sync_prims_cli(dev, name, pid):
    """./allSyncPrims-1kernel <syncPrim> <numLdSt> <numWGs> <numCSIters>"""
    nw = 2 * dev.info.num_cu
    p = run("allSyncPrims-1kernel", name, 4, nw, 3, "--reps=2", "--mode=inc")
    assert p.returncode == 0, (name, p.stderr[-2000:])
    lines = p.stdout.splitlines()
    assert "PASSED!" in lines, name
    avg = [l for l in lines if l.startswith("elapsed time:") and "average:" in l]
    assert len(avg) == 1  # the line syncPrims_results.sh parses
    float(avg[0].split(":")[2].split("m")[0])
    if pid in BANNER:
        assert BANNER[pid] in lines, (pid, lines[:6])


def test_results_sweep_reference_sets():
    """syncPrims_results.sh's own sweep (its 16 primitives, numWGs 2 4 8 16,
    numLdSt 10, numCSIters 10) through the sweep harness: every point passes
    its self-check and yields the duration the script parses."""
    import io
    from heterosync_amd import sweep
    out = io.StringIO()
    rows = sweep.sweep(out=out)
    assert len(rows) == len(sweep.SYNCPRIMS) * len(sweep.NUM_WGS)
    bad = [r for r in rows if not r["passed"] or not r["average_ms"] > 0]
    assert not bad, (bad[:3], out.getvalue()[-2000:])
    for nw in (2, 4, 8, 16):  # and a sense-reversing barrier at the same shapes
        p = run("allSyncPrims-1kernel", "atomicTreeBarrUniqSRB", 10, nw, 10)
        assert p.returncode == 0 and "PASSED!" in p.stdout, (nw, p.stderr[-2000:])


def test_all_sync_prims_cli_rejects_reserved_id():
    p = run("allSyncPrims-1kernel", "noSuchPrim", 10, 8, 1)
    assert p.returncode == 255 and "Unknown synchronization primitive" in p.stderr


@pytest.mark.parametrize("args,ok_line,stream", [
    (("seqlocks", 512, 128, 4, 0), "PASSED", "stdout"),
    (("seqlocks", 512, 128, 4, 1), "PASSED", "stdout"),
    (("seqlocks", 512, 128, 4, 0, "--strict"), "PASSED", "stdout"),
    (("flags", 160, 64, 16, 4, 0.33, 1, 1), "Test PASSED!", "stdout"),
    (("refCounter", 32, 64, 64, 4), "TEST PASSED!", "stderr"),
    (("refCounter", 32, 64, 64, 4, "--strict"), "TEST PASSED!", "stderr"),
    (("splitCounter", 112, 128), "PASSED", "stdout"),  # writer and reader waves
    (("uts", 1, 8192, 90, 32, 128), "SUCCESS!", "stdout"),
])
def test_relaxed_and_uts_clis(args, ok_line, stream):
    p = run(*args)
    assert p.returncode == 0, (args, p.stderr[-2000:])
    assert ok_line in getattr(p, stream).splitlines(), (args, p.stdout[-1000:])


def test_uts_geo_cli_sample_tree():
    """./uts --geo 19 4 10: the published UTS sample tree T1."""
    p = run("uts", "--geo", 19, 4, 10)
    assert p.returncode == 0, p.stderr
    assert "Tree size = 4130071, tree depth = 10, num leaves = 3305118 (80.03%)" in p.stdout
