"""bench.py's command-line contract, on the CPU: one GPU only."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU" in (p.stdout + p.stderr)
    assert not p.stdout.strip().startswith("{")  # no result line


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def test_lf_form_choice_follows_the_library_rule():
    """The XCD form of the lock-free barrier runs when an XCD's leader can
    gather the XCD's other WGs' flags at up to 8 per lane (256 CUs, 8
    XCDs)."""
    b = _bench()
    assert b.lf_xcd_form(1024, 256)       # 128 WGs per XCD, 256 lanes
    assert b.lf_xcd_form(256, 64)         # 32 per XCD
    assert b.lf_xcd_form(512, 512)
    assert b.lf_xcd_form(5, 64)           # one WG on each of 5 XCDs
    assert b.lf_xcd_form(1024, 64)        # 128 per XCD: two flags per lane
    assert b.lf_xcd_form(4096, 64)        # 512 per XCD: eight per lane
    assert not b.lf_xcd_form(8192, 64)    # 1024 per XCD: sixteen per lane


def test_barrier_floor_terms():
    """Each barrier's floor is its data pass plus its dependent hops."""
    b = _bench()
    t = {"wgs": 1024, "lanes": 256, "l2_pass_us": 1.5, "rmw_rt_us": 0.3,
         "hop_same_xcd_us": 0.4, "hop_same_xcd_l2_us": 0.2, "hop_other_xcd_us": 0.45,
         "hop_far_xcd_us": 0.5}
    assert abs(b.barrier_floor_us("atomicTreeBarrUniq", t) - (1.5 + 0.6 + 0.5)) < 1e-9
    assert abs(b.barrier_floor_us("atomicTreeBarrUniqLocalExch", t) - (3.0 + 0.6 + 0.5 + 0.4)) < 1e-9
    assert abs(b.barrier_floor_us("lfTreeBarrUniq", t) - (1.5 + 0.4 + 0.5)) < 1e-9
    assert abs(b.barrier_floor_us("lfTreeBarrUniqLocalExch", t) - (3.0 + 0.6 + 0.5)) < 1e-9
    t64 = dict(t, wgs=8192, lanes=64)  # device form: member hop + two cross-XCD hops
    assert abs(b.barrier_floor_us("lfTreeBarrUniq", t64) - (1.5 + 0.2 + 1.0)) < 1e-9
