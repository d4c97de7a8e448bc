"""The syncPrims results sweep (hip/syncPrims/uvm/syncPrims_results.sh:3-60)
over the drop-in command line, as a regression harness.

For every primitive x numLdSt x numWGs it runs

    allSyncPrims-1kernel <syncPrim> <numLdSt> <numWGs> <numCSIters>

NUM_RUNS times, takes the duration from the line the script parses
(`grep average | cut -f3 -d: | cut -f1 -dm`, syncPrims_results.sh:34) and
prints the script's own report lines.  Unlike the script it also requires
`PASSED!` from every run (the CLI's closed-form self-check), so a wrong
result fails the sweep instead of printing a time.

    python -m heterosync_amd.sweep                  # the script's own sets
    python -m heterosync_amd.sweep --per-cu         # 1, 2, 4, 8 WGs per CU
    python -m heterosync_amd.sweep --json out.json  # rows as JSON too
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "bin", "allSyncPrims-1kernel")

# syncPrims_results.sh:5-10
SYNCPRIMS = ["atomicTreeBarrUniq", "atomicTreeBarrUniqLocalExch", "lfTreeBarrUniq",
             "lfTreeBarrUniqLocalExch", "spinMutex", "spinMutexEBO", "faMutex", "sleepMutex",
             "spinSem10", "spinSemEBO10", "spinMutexUniq", "spinMutexEBOUniq", "faMutexUniq",
             "sleepMutexUniq", "spinSemUniq10", "spinSemEBOUniq10"]
NUM_LDST = [10]
NUM_WGS = [2, 4, 8, 16]  # the script's 2-CU GPU: 1, 2, 4, 8 WGs per CU
NUM_CS_ITERS = 10
NUM_RUNS = 1


def parse_average(stdout: str) -> float:
    """The script's `grep "average" | cut -f3 -d: | cut -f1 -dm`."""
    for line in stdout.splitlines():
        if "average" in line:
            return float(line.split(":")[2].split("m")[0])
    raise ValueError("no 'average' line in the output")


def run_one(prim: str, num_ldst: int, num_wgs: int, iters: int, extra=(), timeout=120):
    p = subprocess.run([CLI, prim, str(num_ldst), str(num_wgs), str(iters), *extra],
                       capture_output=True, text=True, timeout=timeout)
    passed = p.returncode == 0 and "PASSED!" in p.stdout.splitlines()
    avg = parse_average(p.stdout) if passed else None
    return passed, avg, p


def sweep(prims=SYNCPRIMS, num_ldst=NUM_LDST, num_wgs=NUM_WGS, iters=NUM_CS_ITERS,
          runs=NUM_RUNS, extra=(), out=sys.stdout):
    """Runs the sweep, printing the script's lines; returns one row per point."""
    rows = []
    print("Beginning allSyncPrims-1kernel's tests", file=out)
    for prim in prims:
        print(f"\tsyncPrim = {prim}", file=out)
        for ld in num_ldst:
            print(f"\t\tnumLdSt = {ld}", file=out)
            for nw in num_wgs:
                print(f"\t\t\tnumWGs = {nw}", file=out)
                print(f"\t\t\t\t./allSyncPrims-1kernel {prim} {ld} {nw} {iters}", file=out)
                durations, passed = [], True
                for _ in range(runs):
                    ok, avg, p = run_one(prim, ld, nw, iters, extra)
                    passed = passed and ok
                    if not ok:
                        print(f"\t\t\t\t\tFAILED (exit {p.returncode}): {p.stderr.strip()[-300:]}",
                              file=out)
                        break
                    durations.append(avg)
                print("\t\t\t\t\tAll durations for allSyncPrims-1kernel: " +
                      "".join(f"{d}, " for d in durations), file=out)
                mean = sum(durations) / len(durations) if durations else None
                print(f"\t\t\t\tAverage = {mean}", file=out)
                rows.append({"prim": prim, "num_ldst": ld, "num_wgs": nw, "iters": iters,
                             "durations_ms": durations, "average_ms": mean, "passed": passed})
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--prims", nargs="*", default=SYNCPRIMS)
    ap.add_argument("--ldst", nargs="*", type=int, default=NUM_LDST)
    ap.add_argument("--wgs", nargs="*", type=int, default=None)
    ap.add_argument("--per-cu", action="store_true",
                    help="numWGs = 1, 2, 4, 8 per CU of this GPU (the script's intent)")
    ap.add_argument("--iters", type=int, default=NUM_CS_ITERS)
    ap.add_argument("--runs", type=int, default=NUM_RUNS)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    wgs = a.wgs
    if wgs is None and a.per_cu:
        sys.path.insert(0, os.path.dirname(HERE))
        import heterosync_amd as hs
        with hs.Device(0) as dev:
            ncu = dev.info.num_cu
        wgs = [k * ncu for k in (1, 2, 4, 8)]
    rows = sweep(a.prims, a.ldst, wgs or NUM_WGS, a.iters, a.runs)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rows, f, indent=1)
    return 0 if all(r["passed"] for r in rows) else 1


if __name__ == "__main__":
    sys.exit(main())
