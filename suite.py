"""Measurement suite for the synchronization hot path on one GPU.

  C1  plain faMutex 32 WGs x 64 lanes, numLdSt=10, numCSIters=100 (10 launches)
  C2  grid barriers 1024 WGs x 256 lanes, numLdSt=10, numCSIters=1000
  C3  64-instance mutex/semaphore mix: batched persistent launch vs plain
  C4  relaxed atomics 1024 WGs x 256 lanes x 10^4 iterations
  C5  UTS serving loop vs one launch per tree: saturation throughput and
      latency under Poisson arrivals, geometric SHA-1 trees

Every measurement is parity-checked (oracle golden / closed form / exact
counts) and reported next to its unit rate.  Usage:
  python suite.py [c1 c2 c3 c4 c5]
"""
from __future__ import annotations

import ctypes
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import heterosync_amd as hs  # noqa: E402
from heterosync_amd.mix import make_mix  # noqa: E402


def _oracle():
    from oracle.loader import OracleHelpers, load_oracle
    o = load_oracle()
    return o, OracleHelpers(o)


def c1(dev, oh):
    s, r = dev.syncprim("faMutex", 10, 32, 100, reps=10, timeout_s=30.0)
    ok = bool(np.array_equal(s, oh.golden(7, 10, 32, 100, 10, 64, dev.info.num_cu, 0)))
    return {"acq_rel_per_s": r.cs_entries / (r.roi_ms / 1e3), "roi_ms": r.roi_ms,
            "acquires": r.cs_entries, "parity_ok": ok and r.state_ok == 1}


SRB_NAMES = ("atomicTreeBarrUniqSRB", "atomicTreeBarrUniqLocalExchSRB",
             "atomicTreeBarrUniqNaiveSRB", "atomicTreeBarrUniqLocalExchNaiveSRB",
             "atomicTreeBarrUniqNaiveAllSRB", "atomicTreeBarrUniqLocalExchNaiveAllSRB")


def c2(dev, oh, scopes=(0, 1)):
    out = {}
    ncu = dev.info.num_cu
    for name in ("atomicTreeBarrUniq", "atomicTreeBarrUniqLocalExch", "lfTreeBarrUniq",
                 "lfTreeBarrUniqLocalExch") + SRB_NAMES:
        pid = hs.prim_id(name)
        for sc in scopes:
            s, r = dev.syncprim(pid, 10, 4 * ncu, 1000, wg_size=256, reps=10, timeout_s=30.0,
                                sync_scope=sc)
            ok = bool(np.array_equal(s, oh.golden(pid, 10, 4 * ncu, 1000, 10, 256, ncu, 0)))
            # named for the form that ran (a failed XCD placement check falls
            # back to the strict form)
            form = "strict" if r.sync_scope_used else "xcd"
            out[f"{name}_{form}{'' if r.sync_scope_used == sc else '_fallback'}"] = {
                "episodes_per_s": r.episodes / (r.roi_ms / 1e3),
                "parity_ok": ok and r.state_ok == 1}
    return out


def c3(dev, oh, seed=3, reps=1, tune=True):
    """One launch (reps=1) of the seeded 64-instance mix, batched vs plain."""
    mix = make_mix(seed=seed, n=64)
    ncu = dev.info.num_cu
    kw = {}
    tuned = None
    if tune:
        tuned = dev.batch_autotune(mix, reps=reps)
        kw = dev.tuned_kwargs(tuned)
    st = [np.arange(hs.storage_words(p, nw, ncu, 64, L), dtype=np.int32) for p, nw, it, L in mix]
    entries, rb = dev.batch(mix, st, reps=reps, use_graph=1, timeout_s=60.0, **kw)
    ok = rb.all_ok == 1
    for (p, nw, it, L), s in zip(mix, st):
        ok = ok and bool(np.array_equal(s, oh.golden(p, L, nw, it, reps, 64, ncu, 0)))
    plain_ms = 0.0
    plain_acq = 0
    plain_ok = True
    for (p, nw, it, L), s in zip(mix, st):
        ps, r = dev.syncprim(p, L, nw, it, reps=reps, timeout_s=60.0)
        plain_ms += r.roi_ms
        plain_acq += r.cs_entries + r.rd_entries
        plain_ok = plain_ok and bool(np.array_equal(ps, s)) and r.state_ok == 1
    batched = rb.acquires / (rb.roi_ms / 1e3)
    plain = plain_acq / (plain_ms / 1e3)
    return {"instances": len(mix), "acquires": int(rb.acquires),
            "batched_acq_rel_per_s": batched, "batched_roi_ms": rb.roi_ms,
            "plain_acq_rel_per_s": plain, "plain_roi_ms_sum": plain_ms,
            "speedup": batched / plain, "tuned": tuned,
            "parity_ok": ok and plain_ok and plain_acq == rb.acquires}


def c4(dev, oh, iters=10000):
    o = oh.lib
    ncu = dev.info.num_cu
    nw = 4 * ncu
    out = {}
    # seqlock: 1 writer WG per 64 readers (writer period 65), 16 locks; the
    # default (drained agent-scope accesses) and the strict (fenced) form
    for key, variant in (("seqlock", 0), ("seqlock_strict", 2)):
        seq, outa, torn, r = dev.seqlock(nw, 256, 16, 65, iters, variant)
        want = np.zeros(len(seq), np.uint32)
        o.hso_seqlock_expected(nw, 16, 65, iters,
                               want.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(seq))
        out[key] = {"ms": r.kernel_ms, "atomic_ops": r.atomic_ops,
                    "atomic_ops_per_s": r.atomic_ops / (r.kernel_ms / 1e3),
                    "passes_per_s": nw * iters / (r.kernel_ms / 1e3),
                    "parity_ok": bool(np.array_equal(seq, want)) and torn == 0}
    # flags
    rnd = np.zeros(nw * iters, np.float32)
    o.hso_flags_rand(rnd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), rnd.size, 2018)
    wd = o.hso_flags_dirty(rnd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), nw, iters,
                           ctypes.c_float(0.33))
    fo, dirty, r = dev.flags(nw, 256, 16, 4, 0.33, iters, rnd)
    out["flags"] = {"ms": r.kernel_ms, "dirty": dirty,
                    "worker_iters_per_s": (nw - 1) * iters / (r.kernel_ms / 1e3),
                    "parity_ok": dirty == wd}
    # refCounter: 32 sharers, 10^4 inc/dec pairs, spacing 8; default and
    # strict (fenced) forms
    for key, strict in (("refCounter", False), ("refCounter_strict", True)):
        c0, c1_, d0, d1, r = dev.refcounter(32, nw, 256, iters, 8, strict=strict)
        groups = nw // 32
        okr = not c0.any() and not c1_.any() and bool(
            np.all(d0.reshape(32, groups, 256).max(axis=0) == 1)) and bool(
            np.all(d1.reshape(32, groups, 256).max(axis=0) == 1))
        out[key] = {"ms": r.kernel_ms, "atomic_ops": r.atomic_ops,
                    "atomic_ops_per_s": r.atomic_ops / (r.kernel_ms / 1e3),
                    "parity_ok": okr}
    # split counter
    cnt, so, r = dev.splitcounter(nw, 256, iters)
    out["splitCounter"] = {"ms": r.kernel_ms, "atomic_ops": r.atomic_ops,
                           "adds_per_s": cnt.size * iters / (r.kernel_ms / 1e3),
                           "parity_ok": bool(np.all(cnt == iters)) and
                           int(so.max()) <= cnt.size * iters}
    return out


def c5(dev, oh, n=256, seed=9, load=0.8, slots=16, depth=(6, 11), more_loads=(0.5, 0.95)):
    """UTS serving vs one launch per tree, geometric trees (b0 = 4) with
    depth limits drawn from `depth` (about 4K to 5M nodes each).

    saturation: every request present at t = 0; serving nodes/s (wall clock
      of the whole session) against the plain calls back to back, both by
      wall clock (engine kept between calls) and by kernel time alone.
    poisson: arrivals at `load` x the plain kernel capacity; latency =
      completion - arrival.  Plain latencies come from a FIFO queue served by
      the measured per-call wall times (and, as the plain side's best case,
      by the kernel times alone).  `more_loads` repeats it at other offered
      loads (poisson_by_load).
    parity: every tree's (nodes, leaves, max depth) is the same served in the
      mix as launched alone."""
    rng = np.random.default_rng(seed)
    trees = [(int(rng.integers(0, 1 << 30)), int(rng.integers(depth[0], depth[1] + 1)))
             for _ in range(n)]
    dev.uts_geo(1, 4.0, 4)  # builds the context's engine once
    plain = []
    t0 = time.perf_counter()
    for sd, d in trees:
        t1 = time.perf_counter()
        r = dev.uts_geo(sd, 4.0, d)
        plain.append((r.nodes, r.leaves, r.max_depth, r.kernel_ms / 1e3,
                      time.perf_counter() - t1))
    plain_wall = time.perf_counter() - t0
    total_nodes = sum(p[0] for p in plain)
    plain_kernel_s = sum(p[3] for p in plain)
    capacity = total_nodes / plain_kernel_s

    def counts_ok(resp):
        return all((r.nodes, r.leaves, r.max_depth) == p[:3] for r, p in zip(resp, plain))

    def pct(v, q):
        v = np.sort(np.asarray(v))
        return float(v[int(q * (len(v) - 1) + 0.5)])

    def split(resp, arr):
        wait = [r.admit_s - a for r, a in zip(resp, arr)]
        svc = [r.done_s - r.admit_s for r in resp]
        return {"admit_wait_p50_s": pct(wait, 0.5), "admit_wait_p99_s": pct(wait, 0.99),
                "service_p50_s": pct(svc, 0.5), "service_p99_s": pct(svc, 0.99)}

    # saturation
    more = {}
    for sl in sorted({slots, 64}):
        resp, summ = dev.uts_serve([(sd, d, 4.0, 0.0) for sd, d in trees], slots=sl,
                                   timeout_s=120.0)
        more[f"slots_{sl}_nodes_per_s"] = summ.total_nodes / summ.wall_s
    resp, summ = dev.uts_serve([(sd, d, 4.0, 0.0) for sd, d in trees], slots=slots,
                               timeout_s=120.0)
    sat = {"serving_nodes_per_s": summ.total_nodes / summ.wall_s, "by_slots": more,
           "serving_wall_s": summ.wall_s,
           "plain_nodes_per_s_kernel": capacity, "plain_kernel_s": plain_kernel_s,
           "plain_nodes_per_s_wall": total_nodes / plain_wall, "plain_wall_s": plain_wall,
           "speedup_vs_kernel": (summ.total_nodes / summ.wall_s) / capacity,
           "speedup_vs_wall": (summ.total_nodes / summ.wall_s) / (total_nodes / plain_wall)}
    ok = counts_ok(resp)

    def poisson(rho):
        lam = rho * capacity / (total_nodes / n)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, n))
        arrivals -= arrivals[0]
        resp, summ = dev.uts_serve(
            [(sd, d, 4.0, float(a)) for (sd, d), a in zip(trees, arrivals)], slots=slots,
            timeout_s=120.0)

        def fifo(col):
            t, lat = 0.0, []
            for a, p in zip(arrivals, plain):
                t = max(t, a) + p[col]
                lat.append(t - a)
            return {"p50_s": pct(lat, 0.5), "p99_s": pct(lat, 0.99)}

        return counts_ok(resp), {
            "offered_load": rho, "arrival_rate_per_s": lam,
            "serving": {"p50_s": summ.p50_s, "p99_s": summ.p99_s,
                        "nodes_per_s": summ.total_nodes / summ.wall_s,
                        **split(resp, arrivals)},
            "plain_fifo_kernel": fifo(3), "plain_fifo_wall": fifo(4)}

    good, pois = poisson(load)
    ok = ok and good
    by_load = {}
    for rho in more_loads:
        good, by_load[str(rho)] = poisson(rho)
        ok = ok and good
    return {"trees": n, "depths": list(depth), "total_nodes": int(total_nodes), "slots": slots,
            "saturation": sat, "poisson": pois, "poisson_by_load": by_load, "parity_ok": ok}


def main(argv):
    which = argv or ["c1", "c2", "c3", "c4", "c5"]
    o, oh = _oracle()
    dev = hs.Device(0)
    out = {}
    for w in which:
        out[w] = globals()[w](dev, oh)
        print(json.dumps({w: out[w]}), flush=True)
    dev.close()
    return out


if __name__ == "__main__":
    main(sys.argv[1:])
