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


class _R:
    def __init__(self, nodes, leaves, depth, ms=0.0):
        self.nodes, self.leaves, self.max_depth, self.kernel_ms = nodes, leaves, depth, ms


class _FakeDev:
    """Stands in for the device: fixed counts per call site."""

    def __init__(self, alt=None, serve=None, runs_ms=(130.0, 128.0, 129.0)):
        self.calls = []
        self.ms = list(runs_ms)
        self.alt, self.serve = alt, serve

    def uts_geo(self, seed, b0, depth, blocks=0):
        self.calls.append((seed, b0, depth, blocks))
        if depth == 4:
            return _R(100, 80, 4)
        if blocks:
            return self.alt or _R(1000, 800, 15)
        return _R(1000, 800, 15, self.ms.pop(0))

    def uts_serve(self, reqs, slots=0, timeout_s=0.0):
        return [self.serve or _R(1000, 800, 15)], None


def test_uts_tree_extra_parity_and_rate():
    """The single-tree line agrees only when every run, the second grid and
    the serving kernel give the same counts; its rate is the median run's."""
    b = _bench()
    out = b.extra_uts_tree(_FakeDev())
    assert out["parity_ok"] is True
    assert out["kernel_ms"] == [128.0, 129.0, 130.0]
    assert abs(out["nodes_per_s"] - 1000 / 0.129) <= 0.05  # rounded to 0.1
    peak = out["roofline"]["peak_nodes_per_s"]
    assert abs(out["roofline"]["frac"] - round(out["nodes_per_s"] / peak, 4)) < 1e-9
    assert not b.extra_uts_tree(_FakeDev(alt=_R(1000, 799, 15)))["parity_ok"]
    assert not b.extra_uts_tree(_FakeDev(serve=_R(1001, 800, 15)))["parity_ok"]
    assert not b.extra_uts_tree(_FakeDev(serve=_R(1000, 800, 14)))["parity_ok"]
