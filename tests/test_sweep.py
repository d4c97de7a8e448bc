"""The sweep harness's parsing and report lines, on the CPU (the GPU runs
are in test_cli_gpu.py)."""
import io

from heterosync_amd import sweep


def test_parse_average_matches_the_scripts_cut():
    # the CLI's line; syncPrims_results.sh:34 keeps field 3 of ':' up to 'm'
    out = "atomic_tree_barrier_003\nelapsed time: 12.500000 ms, average: 1.250000 ms\nPASSED!\n"
    assert sweep.parse_average(out) == 1.25


def test_reference_sets():
    assert sweep.NUM_WGS == [2, 4, 8, 16] and sweep.NUM_LDST == [10]
    assert sweep.NUM_CS_ITERS == 10 and len(sweep.SYNCPRIMS) == 16


def test_report_lines(monkeypatch):
    calls = []

    def fake(prim, ld, nw, iters, extra=(), timeout=120):
        calls.append((prim, ld, nw, iters))
        return True, 0.5, None

    monkeypatch.setattr(sweep, "run_one", fake)
    out = io.StringIO()
    rows = sweep.sweep(["faMutex"], [10], [2, 4], 10, runs=2, out=out)
    assert calls == [("faMutex", 10, 2, 10)] * 2 + [("faMutex", 10, 4, 10)] * 2
    text = out.getvalue()
    assert "\t\t\t\t./allSyncPrims-1kernel faMutex 10 2 10" in text
    assert "All durations for allSyncPrims-1kernel: 0.5, 0.5, " in text
    assert [r["average_ms"] for r in rows] == [0.5, 0.5]
