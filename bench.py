#!/usr/bin/env python3
"""Benchmark of the synchronization hot path on one MI355X.

Headline line: barrier episodes/s of the full-co-residency grid barrier
(atomicTreeBarrUniq, 1024 workgroups = 4 per CU x 256 CUs, 256 threads,
numLdSt=10, numCSIters=1000, 10 launches per step) through the C ABI.  One
step = one hs_syncprim_run call: 10 launches x 1000 episodes, each launch
followed by a device synchronize (the reference's plain ROI,
hip/syncPrims/uvm/main.hip.cpp:2354-2511).  The timed runs use the
in-place-increment critical section, so the parity check (the library's own
closed-form self-check, the reference binary's golden) sees any lost,
doubled or stale pass.

Roofline: the episode floor measured live by `hs_probe floor` -- the data
pass of one episode through the L2s at this footprint, plus the dependent
synchronization hops of the barrier (two RMW round trips and one flag
hand-off) -- against the measured episode time of the kernel.  The HBM view
(algorithmic bytes over 8 TB/s, with the rocprof-measured HBM traffic) is
kept beside it.

`extras` carry the other barriers, the plain faMutex (C1), the batched
mutex/semaphore mix against its plain per-instance launches (C3) and the
UTS serving loop against one launch per tree (C5), each with its parity
check.  `cpu_baseline` times the same workloads on the host's cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, public spec
# HBM bytes per launch of k_barrier (atomicTreeBarrUniq, XCD-aware, 1024 WGs
# x 256 lanes, numLdSt=10, numCSIters=1000), from rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE per dispatch, each scaled by what the same counter reads for
# a 256 MiB copy kernel (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000); the command
# lines and raw figures are in profiles/k_barrier_atomicTreeBarrUniq_1024x256.txt.
# Key: (prim, sync scope, WG size, numLdSt, numCSIters, num WGs, mode); mode 0
# is the reference's copy chain, mode 1 the in-place increment the headline
# runs (FETCH 89432.4 KiB x 1.99982 + WRITE 226699.0 KiB x 1.000).
MEASURED_TRAFFIC = {("atomicTreeBarrUniq", 0, 256, 10, 1000, 1024, 0): 405270138,
                    ("atomicTreeBarrUniq", 0, 256, 10, 1000, 1024, 1): 415280494}
METRIC = ("acquire+release/s & barrier episodes/s at 1 GPU, batched vs plain; "
          "UTS nodes/s, p99")
BARRIERS = ("atomicTreeBarrUniq", "atomicTreeBarrUniqLocalExch", "lfTreeBarrUniq",
            "lfTreeBarrUniqLocalExch")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--prim", default="atomicTreeBarrUniq")
    ap.add_argument("--num-ldst", type=int, default=10)
    ap.add_argument("--iters", type=int, default=1000)
    ap.add_argument("--wg", type=int, default=256)
    ap.add_argument("--strict", action="store_true",
                    help="agent-scope release/acquire in every WG")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--only", default="",
                    help="comma list of extras to run (barriers,c1,c3,c5,uts,uts_cpu); "
                         "default all")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the floor probe (a child process); the roofline then has no "
                         "floor (for runs under a profiler)")
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="length of each CPU-baseline sample (about this much wall time)")
    return ap.parse_args()


# ---- measured floors --------------------------------------------------------

def probe_floor(nw, wg, L):
    """Terms of the synchronization floors from `hs_probe floor` (a separate
    process, run before the device is opened here)."""
    exe = os.path.join(ROOT, "heterosync_amd", "bin", "hs_probe")
    p = subprocess.run([exe, "floor", str(nw), str(wg), str(L)], capture_output=True, text=True,
                       timeout=120)
    if p.returncode != 0:
        raise RuntimeError(f"hs_probe floor failed: {p.stdout}{p.stderr}")
    return json.loads(p.stdout.strip().splitlines()[-1])


def lf_xcd_form(nw, wg, ncu=256, nxcd=8):
    """Whether the XCD-aware lock-free barrier runs its XCD form (one
    sub-group per XCD, the XCD leaders all-to-all, each releasing its own
    XCD): the library takes it when the XCD leader's lanes can gather its
    XCD's other WGs' flags at up to 8 flags per lane (lf_xcd_form in
    hs_syncprims.hip)."""
    nab = min(nw, ncu)
    per = -(-nw // nab)
    return (-(-nab // nxcd)) * per - 1 <= 8 * wg


def barrier_floor_us(name, t, ncu=256):
    """Lower bound on one episode of a grid barrier at the probed shape
    (t["wgs"] x t["lanes"]): the data pass (the local-exchange forms move
    two regions per episode) plus the dependent steps no episode can overlap
    with its data.  The XCD-aware atomic tree: two RMW round trips
    (sub-counter, device counter), then a release that must reach every XCD
    (the slowest cross-XCD hop).  The flag tree: the member flags reach
    their leader inside the XCD (plain stores kept in the XCD's L2:
    hop_same_xcd_l2), then, in the XCD form, the XCD leaders' flags reach
    every other leader (the slowest cross-XCD hop) and each leader releases
    its own XCD (same-XCD hop again); in the device form the leaders' flags
    reach the device master and its release reaches every XCD (two
    cross-XCD hops).  The local exchange adds one same-XCD hop (the group's
    all-to-all flags, or its counter and generation)."""
    lx = "LocalExch" in name
    data = t["l2_pass_us"] * (2 if lx else 1)
    far = t.get("hop_far_xcd_us", t["hop_other_xcd_us"])
    same_l2 = t.get("hop_same_xcd_l2_us", t["hop_same_xcd_us"])
    if name.startswith("atomic"):
        sync = 2 * t["rmw_rt_us"] + far
        if lx:
            sync += t["hop_same_xcd_us"]
    else:
        if lf_xcd_form(t["wgs"], t["lanes"], ncu):
            sync = 2 * same_l2 + far
        else:
            sync = same_l2 + 2 * far
        if lx:
            sync += same_l2
    return data + sync


def floor_text(name, nw, wg, ncu):
    """What barrier_floor_us charges for `name`, in words."""
    data = ("one episode's data pass through the L2s at this footprint (hs_probe floor, the "
            "kernel's own walk with no barrier)" + (", twice (two regions)" if "LocalExch" in name
                                                   else ""))
    if name.startswith("atomic"):
        sync = ("2 dependent RMW round trips (mean over the 8 issuing XCDs) + the release "
                "hand-off to the farthest XCD (slowest of the 7 cross-XCD hops)")
        if "LocalExch" in name:
            sync += " + one same-XCD hop (the group's counter and generation)"
    elif lf_xcd_form(nw, wg, ncu):
        sync = ("the member flags' hop to their XCD leader (plain stores kept in the XCD's L2) "
                "+ the XCD leaders' all-to-all (slowest cross-XCD hop) + each leader's "
                "release of its own XCD (same-XCD hop)")
    else:
        sync = ("the member flags' hop to their leader (same XCD) + the leaders' flags to the "
                "device master and its release to every XCD (two slowest cross-XCD hops)")
    if not name.startswith("atomic") and "LocalExch" in name:
        sync += " + the group's all-to-all flags (same-XCD hop)"
    return data + " + " + sync


def handoff_floor_us(t, L, lanes=64):
    """Lower bound on one mutex hand-off with an L-line critical section: the
    release is seen by the next holder (one hand-off), which loads the lines
    its predecessor wrote through (a 16-line chunk per round trip, chunks
    pipelined, so at least one load round trip) and drains its stores
    before releasing (one store round trip).  The load is charged at the
    probe's round trip for lines just written through by agent-scope
    stores from every XCD (the next holder runs on another XCD 7 times in
    8, where the lines are not in its L2)."""
    return t["hop_other_xcd_us"] + t.get("remote_load_rt_us", t["load_rt_us"]) + t["store_rt_us"]


# ---- the configurations -----------------------------------------------------

def run_c2(dev, hs, a, nw, prim, scope):
    reps = 10

    def step():
        return dev.syncprim(prim, a.num_ldst, nw, a.iters, wg_size=a.wg, reps=reps, mode=1,
                            timeout_s=30.0, sync_scope=scope)

    for _ in range(a.warmup):
        step()
    roi_ms = kern_ms = 0.0
    ok = True
    used = set()  # the scope each timed run actually used (placement fallback)
    want = hs.expected_storage(prim, a.num_ldst, nw, a.iters, reps, a.wg, dev.info.num_cu, 1)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s, r = step()
        roi_ms += r.roi_ms
        kern_ms += r.kernel_ms
        ok = ok and r.state_ok == 1 and r.episodes == a.iters * reps and bool(
            np.array_equal(s, want))
        used.add(r.sync_scope_used)
    wall = time.perf_counter() - t0
    return roi_ms, kern_ms, ok, used, wall


def extra_barriers(dev, hs, a, nw, floor_terms):
    out = {}
    for name in BARRIERS:
        for sc in (0, 1):
            s, r = dev.syncprim(name, a.num_ldst, nw, a.iters, wg_size=a.wg, reps=10, mode=1,
                                timeout_s=30.0, sync_scope=sc)
            want = hs.expected_storage(name, a.num_ldst, nw, a.iters, 10, a.wg,
                                       dev.info.num_cu, 1)
            form = "strict" if r.sync_scope_used else "xcd"  # the form that ran
            key = f"{name}_{form}{'' if r.sync_scope_used == sc else '_fallback'}"
            ep_us = r.kernel_ms * 1e3 / r.episodes
            line = {"episodes_per_s": round(r.episodes / (r.roi_ms / 1e3), 1),
                    "episode_us": round(ep_us, 3),
                    "parity_ok": bool(np.array_equal(s, want)) and r.state_ok == 1}
            if floor_terms and not r.sync_scope_used:
                fl = barrier_floor_us(name, floor_terms, dev.info.num_cu)
                line.update(floor_us=round(fl, 3), frac_of_floor=round(fl / ep_us, 4))
            out[key] = line
    return out


def extra_c1(dev, hs, floor_terms):
    s, r = dev.syncprim("faMutex", 10, 32, 100, reps=10, mode=1, timeout_s=30.0)
    want = hs.expected_storage("faMutex", 10, 32, 100, 10, 64, dev.info.num_cu, 1)
    per_us = r.roi_ms * 1e3 / r.cs_entries
    out = {"workload": "faMutex 32 WGs x 64, numLdSt=10, numCSIters=100, 10 launches, inc mode",
           "acq_rel_per_s": round(r.cs_entries / (r.roi_ms / 1e3), 1),
           "us_per_acquire": round(per_us, 4),
           "parity_ok": bool(np.array_equal(s, want)) and r.state_ok == 1
           and r.cs_entries == 32000}
    if floor_terms:
        fl = handoff_floor_us(floor_terms, 10)
        out.update(floor_us=round(fl, 4), frac_of_floor=round(fl / per_us, 4))
    return out


def handoff_us(dev, L):
    """Measured time per acquire+release of the plain faMutex at numLdSt L
    (32 WGs contending, every hand-off serialised)."""
    _, r = dev.syncprim("faMutex", L, 32, 50, reps=1, mode=1, timeout_s=30.0)
    return r.roi_ms * 1e3 / r.cs_entries


def serialized_acquires(p, nw, iters, ncu):
    """Acquires of one instance that exclude each other: every mutex
    acquire, or the writers' entries of a semaphore (one writer WG per
    group of min(nw, NUM_CU) WGs, main.hip.cpp:883)."""
    if p <= 7:
        return nw * iters
    nab = min(nw, ncu)
    per = -(-nw // nab)
    writers = sum(1 for b in range(nw) if b // nab == (b % nab) % per)
    return writers * iters


def batched_acquire_floor_us(p, nw, L, ncu, t, cohort=1):
    """Lower bound on one exclusive acquire+release of a batched instance:
    the release is seen by the next holder (a hand-off across XCDs), which
    makes its CAS where the lock needs one (spin, EBO, semaphores), loads
    its lines (at least one round trip per 64 lines: the widest worker
    splits a chain over four waves, 16 lines per load batch) and drains its
    stores (one store round trip) before releasing.  A semaphore writer
    walks its group's perCU regions.  The next holder's loads are charged
    as in handoff_floor_us.  With the cohort a spin / EBO lock or a
    semaphore writer's hold serves up to `cohort` critical sections on one
    worker, back to back: only the first pays the hand-off and the CAS, and
    each section at least one load round trip through the L2 (no drain
    between them); the bound takes every hold at its fullest."""
    regions = 1 if p <= 7 else -(-nw // min(nw, ncu))
    cas = t["rmw_rt_us"] if p in (4, 5) or p >= 8 else 0.0
    rounds = max(1, -(-L // 64)) * regions
    # the next holder's loads as in C1's floor; a cohort's critical
    # sections stay on one worker (L2 hits)
    cold = t.get("remote_load_rt_us", t["load_rt_us"])
    cs = rounds * t["load_rt_us"]
    handoff = t["hop_other_xcd_us"] + cas + rounds * cold + t["store_rt_us"]
    c = max(1, cohort) if p in (4, 5) or p >= 8 else 1
    return (handoff + (c - 1) * cs) / c


def extra_c3(dev, hs, floor_terms):
    """The batched configuration: the seed-3 mix of 64 instances in one
    persistent hipGraph-captured launch (autotuned, inc mode) against the 64
    plain launches of the same instances."""
    from heterosync_amd.mix import make_mix
    mix = make_mix(seed=3, n=64)
    ncu = dev.info.num_cu
    t0 = time.perf_counter()
    tuned = dev.batch_autotune(mix, reps=1, mode=1)
    tune_s = time.perf_counter() - t0
    st = [np.arange(hs.storage_words(p, nw, ncu, 64, L), dtype=np.int32) for p, nw, it, L in mix]
    entries, rb = dev.batch(mix, st, reps=1, mode=1, use_graph=1, timeout_s=60.0,
                            **dev.tuned_kwargs(tuned))
    ok = rb.all_ok == 1
    for (p, nw, it, L), s in zip(mix, st):
        ok = ok and bool(np.array_equal(s, hs.expected_storage(p, L, nw, it, 1, 64, ncu, 1)))
    plain_ms, plain_acq, plain_ok = 0.0, 0, True
    for (p, nw, it, L), s in zip(mix, st):
        ps, r = dev.syncprim(p, L, nw, it, reps=1, mode=1, timeout_s=60.0)
        plain_ms += r.roi_ms
        plain_acq += r.cs_entries + r.rd_entries
        plain_ok = plain_ok and bool(np.array_equal(ps, s)) and r.state_ok == 1
    batched = rb.acquires / (rb.roi_ms / 1e3)
    plain = plain_acq / (plain_ms / 1e3)
    out = {"workload": "64 seeded instances (seed 3) of {spinMutex, spinMutexEBO, faMutex, "
                        "sleepMutex, spinSem1/2/10/120}, numTBs 8-512 x 64 threads, numCSIters "
                        "log-uniform 10-1000, numLdSt in {1,10,100}; one launch, inc mode",
            "acquires": int(rb.acquires),
            "batched_acq_rel_per_s": round(batched, 1), "batched_roi_ms": round(rb.roi_ms, 3),
            "plain_acq_rel_per_s": round(plain, 1), "plain_roi_ms_sum": round(plain_ms, 3),
            "speedup": round(batched / plain, 3),
            "tuned": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tuned.items()},
            "autotune_s": round(tune_s, 2),
            "parity_ok": ok and plain_ok and plain_acq == rb.acquires}
    if not floor_terms:
        return out
    # critical-path bound: the instances run concurrently, so the launch
    # takes at least the longest chain of mutually exclusive acquires, each
    # at least the floor of one hand-off (measured round trips), or of one
    # cohort critical section where the tuned run uses the cohort
    coh = tuned.get("cohort", 0)
    chains = [(serialized_acquires(p, nw, it, ncu) *
               batched_acquire_floor_us(p, nw, L, ncu, floor_terms, coh), i)
              for i, (p, nw, it, L) in enumerate(mix)]
    bound_us, crit = max(chains)
    hand = {L: handoff_us(dev, L) for L in sorted({L for *_, L in mix})}
    # what one-at-a-time hand-offs at the plain form's measured rate would
    # need for the longest chain (the cohort and the local successor let
    # the batched run go below it)
    plain_ref_us, plain_crit = max(
        (serialized_acquires(p, nw, it, ncu) * hand[L], i) for i, (p, nw, it, L) in enumerate(mix))
    cp, cnw, cit, cL = mix[crit]
    out["critical_path"] = {
        "instance": crit, "prim": hs.prim_name(cp), "num_wgs": cnw, "num_cs_iters": cit,
        "num_ldst": cL, "serialized_acquires": serialized_acquires(cp, cnw, cit, ncu),
        "acquire_floor_us": round(
            batched_acquire_floor_us(cp, cnw, cL, ncu, floor_terms, coh), 4),
        "cohort": coh,
        "achieved_us_per_acquire": round(
            rb.roi_ms * 1e3 / serialized_acquires(cp, cnw, cit, ncu), 4),
        "bound_ms": round(bound_us / 1e3, 3),
        "frac_of_bound": round(bound_us / 1e3 / rb.roi_ms, 4),
        "plain_handoff_chain_ms": round(plain_ref_us / 1e3, 3),
        "plain_handoff_chain_instance": plain_crit,
        "plain_famutex_us_per_acquire_by_numldst": {
            str(k): round(v, 4) for k, v in hand.items()}}
    return out


def c5_trees(n=256, seed=9, depth=(6, 14)):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(0, 1 << 30)), int(rng.integers(depth[0], depth[1] + 1)))
            for _ in range(n)]


def pct(v, q):
    v = np.sort(np.asarray(v, dtype=float))
    return float(v[int(q * (len(v) - 1) + 0.5)])


def extra_c5(dev, slots=64, loads=(0.8,), n=256, seed=9, depth=(6, 14)):
    """The serving configuration: geometric SHA-1 trees (b0 = 4, depth limit
    uniform in 6-14, seeded roots), one persistent serving launch against one
    launch per tree, both measured in real time.

    saturation: every request present at t = 0.
    poisson: arrivals at `load` x the plain form's measured capacity (wall
      clock); the plain side is a real FIFO run (each tree launched when it
      has arrived and the previous one is done), so both latencies are
      measured.  The first of `loads` is reported as `poisson`, any others
      under `poisson_other_loads`.
    parity: every tree's (nodes, leaves, max depth) is the same served in the
      mix as launched alone."""
    trees = c5_trees(n, seed, depth)
    # the context's two engines (the grids of small and of large trees),
    # built once, before anything is timed
    dev.uts_geo(1, 4.0, 4)
    dev.uts_geo(1, 4.0, 14)
    plain = []
    t0 = time.perf_counter()
    for sd, d in trees:
        t1 = time.perf_counter()
        r = dev.uts_geo(sd, 4.0, d)
        plain.append((r.nodes, r.leaves, r.max_depth, r.kernel_ms / 1e3,
                      time.perf_counter() - t1))
    plain_wall = time.perf_counter() - t0
    total = sum(p[0] for p in plain)
    cap_wall = total / plain_wall

    def same(resp):
        return all((r.nodes, r.leaves, r.max_depth) == p[:3] for r, p in zip(resp, plain))

    resp, summ = dev.uts_serve([(sd, d, 4.0, 0.0) for sd, d in trees], slots=slots,
                               timeout_s=300.0)
    sat = {"serving_nodes_per_s": round(summ.total_nodes / summ.wall_s, 1),
           "serving_wall_s": round(summ.wall_s, 4),
           "plain_nodes_per_s": round(cap_wall, 1), "plain_wall_s": round(plain_wall, 4),
           "plain_nodes_per_s_kernel_only": round(total / sum(p[3] for p in plain), 1),
           "speedup": round((summ.total_nodes / summ.wall_s) / cap_wall, 3)}
    ok = same(resp)

    def poisson(load):
        """Poisson arrivals at `load` x plain capacity (by wall clock)."""
        rng = np.random.default_rng(seed + 1)
        lam = load * cap_wall / (total / n)
        arr = np.cumsum(rng.exponential(1.0 / lam, n))
        arr -= arr[0]
        resp, summ = dev.uts_serve([(sd, d, 4.0, float(t)) for (sd, d), t in zip(trees, arr)],
                                   slots=slots, timeout_s=300.0)
        good = same(resp)
        serve_lat = [r.done_s - t for r, t in zip(resp, arr)]
        # the plain form serving the same arrivals: FIFO, one launch at a time
        lat = []
        t0 = time.perf_counter()
        for (sd, d), t in zip(trees, arr):
            now = time.perf_counter() - t0
            if now < t:
                time.sleep(t - now)
            r = dev.uts_geo(sd, 4.0, d)
            lat.append(time.perf_counter() - t0 - t)
            good = good and (r.nodes, r.leaves, r.max_depth) == plain[len(lat) - 1][:3]
        return good, {"offered_load": load, "arrival_rate_per_s": round(lam, 2),
                      "serving_p50_ms": round(pct(serve_lat, 0.5) * 1e3, 3),
                      "serving_p99_ms": round(pct(serve_lat, 0.99) * 1e3, 3),
                      "serving_nodes_per_s": round(summ.total_nodes / summ.wall_s, 1),
                      "plain_p50_ms": round(pct(lat, 0.5) * 1e3, 3),
                      "plain_p99_ms": round(pct(lat, 0.99) * 1e3, 3)}

    pois = []
    for load in loads:
        good, line = poisson(load)
        ok = ok and good
        pois.append(line)
    out = {"workload": f"{n} geometric SHA-1 trees, b0=4, depth limit U[{depth[0]},{depth[1]}], "
                       f"root seeds from seed {seed}; {slots} slots",
           "trees": n, "total_nodes": int(total), "slots": slots,
           "largest_tree_nodes": int(max(p[0] for p in plain)),
           "saturation": sat, "poisson": pois[0]}
    if len(pois) > 1:
        out["poisson_other_loads"] = pois[1:]
    out["parity_ok"] = bool(ok)
    return out


# VALU issue peak of the GPU: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per
# wave64 instruction, in wave-instructions per second.
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4
# The geometric kernel's VALU lane-slots per node on the depth-15 tree below
# (SQ_INSTS_VALU x 64 / nodes; profiles/k_uts_geometric_seed7_depth15.txt).
UTS_LANE_SLOTS_PER_NODE = 625


def extra_uts_tree(dev, seed=7, depth=15, runs=3):
    """One large geometric tree (b0 = 4, depth limit 15: 7.85e9 nodes) by the
    single-tree kernel at its default grid, against the VALU issue bound.
    Parity: every run's counts, the single-tree kernel's at a smaller grid
    (2048 WGs) and the serving kernel's (the tree as its only request) all
    agree."""
    dev.uts_geo(1, 4.0, 4)
    rs = [dev.uts_geo(seed, 4.0, depth) for _ in range(runs)]
    alt = dev.uts_geo(seed, 4.0, depth, 2048)
    resp, _ = dev.uts_serve([(seed, depth, 4.0, 0.0)], slots=1, timeout_s=120.0)
    key = (rs[0].nodes, rs[0].leaves, rs[0].max_depth)
    ok = all((r.nodes, r.leaves, r.max_depth) == key for r in rs + [alt] + list(resp))
    ms = sorted(r.kernel_ms for r in rs)
    med = ms[len(ms) // 2]
    bound = VALU_ISSUE_PER_S * 64 / UTS_LANE_SLOTS_PER_NODE
    rate = rs[0].nodes / (med / 1e3)
    return {"workload": f"geometric tree seed {seed}, b0=4, depth {depth}: {rs[0].nodes} nodes, "
                        f"{rs[0].leaves} leaves",
            "kernel_ms": [round(m, 3) for m in ms], "nodes_per_s": round(rate, 1),
            "roofline": {"bound": "valu-issue", "peak_nodes_per_s": round(bound, 1),
                         "frac": round(rate / bound, 4),
                         "lane_slots_per_node": UTS_LANE_SLOTS_PER_NODE},
            "parity_ok": bool(ok)}


def extra_uts_host_workers(dev, seed=7, depth=13, threads=(0, 4, 16)):
    """One large geometric tree (b0 = 4) searched by the GPU alone and with
    host worker threads taking part (each served by one GPU WG through a
    pinned mailbox): what the host's cores add to the GPU's rate."""
    dev.uts_geo(1, 4.0, 4)
    base = dev.uts_geo(seed, 4.0, depth)
    out = {"workload": f"geometric tree seed {seed}, b0=4, depth {depth}: {base.nodes} nodes",
           "runs": {}}
    ok = True
    for t in threads:
        if t == 0:
            r, cpu = dev.uts_geo(seed, 4.0, depth), None
        else:
            r, cpu = dev.uts_geo_cpu(seed, 4.0, depth, t)
        ok = ok and (r.nodes, r.leaves, r.max_depth) == (base.nodes, base.leaves, base.max_depth)
        line = {"nodes_per_s": round(r.nodes / (r.kernel_ms / 1e3), 1),
                "kernel_ms": round(r.kernel_ms, 3)}
        if cpu is not None:
            line.update(host_nodes=int(cpu.nodes), host_share=round(cpu.nodes / r.nodes, 8),
                        host_batches=int(cpu.deliveries), host_busy_s=round(cpu.busy_s, 4))
        out["runs"][f"host_threads_{t}"] = line
    out["parity_ok"] = bool(ok)
    return out


# ---- CPU baselines ------------------------------------------------------------

def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def cpu_baselines(prim, L, nw, wg, ncu, seconds):
    """The oracle's std::thread + std::atomic ports of the same workloads, on
    bounded samples (about `seconds` each), checked against the closed form."""
    from oracle.loader import OracleHelpers, load_oracle
    o = load_oracle()
    oh = OracleHelpers(o)
    P32 = ctypes.POINTER(ctypes.c_int32)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))

    # C2: the barrier workload on a pool of `threads` threads
    n = o.hso_num_storage_locs(prim, nw, ncu, wg, L)
    s = np.zeros(n, np.int32)
    probe = 1500
    secs = o.hso_cpu_barrier_run(prim, L, nw, probe, 1, wg, ncu, 1, threads,
                                 s.ctypes.data_as(P32), n)
    episodes = probe
    if secs > 0 and seconds > 0:
        episodes = int(min(300000, max(probe, seconds * probe / secs)))
        secs = o.hso_cpu_barrier_run(prim, L, nw, episodes, 1, wg, ncu, 1, threads,
                                     s.ctypes.data_as(P32), n)
    ok2 = bool(np.array_equal(s, oh.golden(prim, L, nw, episodes, 1, wg, ncu, 1)))
    c2 = {"value": round(episodes / secs, 2) if secs > 0 else None,
          "unit": "barrier episodes/s", "cores": threads, "kind": "port",
          "sample": (f"{episodes} episodes ({secs:.1f} s) of the same {nw}-WG x {wg}-lane, "
                     f"numLdSt={L} workload (inc mode) on {threads} std::threads (each owns "
                     f"~{nw // threads} emulated WGs, one std::atomic barrier per episode); "
                     f"storage matches the closed form: {ok2}")}

    def syncprim_rate(p, Lc, nwc, iters):
        m = o.hso_num_storage_locs(p, nwc, ncu, 64, Lc)
        buf = np.zeros(m, np.int32)
        cs = ctypes.c_int64()
        t = o.hso_cpu_run(p, Lc, nwc, iters, 1, 64, ncu, 1, buf.ctypes.data_as(P32), m,
                          ctypes.byref(cs))
        w, r, _ = oh.counts(p, nwc, iters, 1, ncu)
        good = bool(np.array_equal(buf, oh.golden(p, Lc, nwc, iters, 1, 64, ncu, 1)))
        return t, w + r, good

    # C1: faMutex, 32 emulated WGs (one thread each), numLdSt 10
    t, acq, good = syncprim_rate(7, 10, 32, 100)
    iters = int(min(20000, max(100, 100 * seconds / max(t, 1e-6))))
    t, acq, good = syncprim_rate(7, 10, 32, iters)
    c1 = {"value": round(acq / t, 1), "unit": "acquire+release/s", "cores": 32, "kind": "port",
          "sample": f"faMutex 32 emulated WGs (32 std::threads) x {iters} iterations, numLdSt=10, "
                    f"{acq} acquires in {t:.2f} s; storage matches the closed form: {good}"}

    # C3: the 64 instances one after another (the plain form), iterations
    # scaled down by one factor so the sample takes about `seconds`
    from heterosync_amd.mix import make_mix
    mix = make_mix(seed=3, n=64)
    scale, tot_t, tot_acq, good3, threads_max = 0.02, 0.0, 0, True, 0
    for attempt in range(2):
        tot_t, tot_acq, good3, threads_max = 0.0, 0, True, 0
        for p, nwc, it, Lc in mix:
            itc = max(1, int(round(it * scale)))
            t, acq, good = syncprim_rate(p, Lc, nwc, itc)
            tot_t += t
            tot_acq += acq
            good3 = good3 and good
            threads_max = max(threads_max, nwc)
        if attempt == 0 and tot_t > 0:
            scale = min(1.0, scale * seconds / tot_t)
    c3 = {"value": round(tot_acq / tot_t, 1), "unit": "acquire+release/s",
          "cores": threads_max, "kind": "port",
          "sample": f"the seed-3 64-instance mix, one instance at a time (one std::thread per "
                    f"emulated WG, up to {threads_max}), numCSIters x {scale:.4f}: {tot_acq} "
                    f"acquires in {tot_t:.2f} s; every storage matches the closed form: {good3}"}

    # C5: the serving mix's trees counted by `threads` host threads (the
    # oracle's SHA-1 depth-first search; ctypes releases the GIL), on the
    # trees of the mix in order until about `seconds` of work
    from concurrent.futures import ThreadPoolExecutor
    trees = [tr for tr in c5_trees() if tr[1] <= 10]

    def count(tr):
        nn, lv, md = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        o.hso_uts_geo_count(tr[0], 4.0, tr[1], ctypes.byref(nn), ctypes.byref(lv),
                            ctypes.byref(md))
        return nn.value
    t0 = time.perf_counter()
    done_nodes, done_trees = 0, 0
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for k in range(0, len(trees), threads):
            done_nodes += sum(ex.map(count, trees[k:k + threads]))
            done_trees += len(trees[k:k + threads])
            if time.perf_counter() - t0 > seconds:
                break
    t5 = time.perf_counter() - t0
    c5 = {"value": round(done_nodes / t5, 1), "unit": "UTS nodes/s", "cores": threads,
          "kind": "port",
          "sample": f"{done_trees} trees of the serving mix with depth limit <= 10 "
                    f"({done_nodes} nodes, {t5:.2f} s) searched by {threads} std::threads "
                    f"(one tree each at a time, SHA-1 depth-first)"}
    out = dict(c2)
    out["host"] = host_info()
    out["by_config"] = {"c1": c1, "c2": c2, "c3": c3, "c5": c5}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {a.gpus} requested); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import heterosync_amd as hs

    only = set(filter(None, a.only.split(","))) or {"barriers", "c1", "c3", "c5", "uts",
                                                     "uts_cpu"}
    # the probe runs first, in its own process, before this one opens the device
    ncu_guess = 256
    floor_terms = None if a.no_probe else probe_floor(4 * ncu_guess, a.wg, a.num_ldst)
    dev = hs.Device(0)
    ncu = dev.info.num_cu
    nw = 4 * ncu  # full co-residency: 4 WGs per CU
    if floor_terms and nw != floor_terms["wgs"]:
        dev.close()
        floor_terms = probe_floor(nw, a.wg, a.num_ldst)
        dev = hs.Device(0)
    prim = hs.prim_id(a.prim)
    reps = 10
    scope = 1 if a.strict else 0
    roi_ms, kern_ms, ok, used, wall = run_c2(dev, hs, a, nw, prim, scope)
    used_scope = used.pop() if len(used) == 1 else None

    episodes = a.iters * reps * a.steps
    value = episodes / (roi_ms / 1e3)
    avg_launch_ms = kern_ms / (reps * a.steps)
    ep_us = avg_launch_ms * 1e3 / a.iters
    floor_us = barrier_floor_us(a.prim, floor_terms, ncu) if floor_terms else None
    # HBM view: algorithmic bytes per launch = episodes/launch x WGs x lanes x
    # numLdSt x (4 B load + 4 B store), over the average launch duration
    bytes_per_episode = nw * a.wg * a.num_ldst * 8
    achieved_gbs = bytes_per_episode * a.iters / (avg_launch_ms / 1e3) / 1e9
    traffic = MEASURED_TRAFFIC.get((a.prim, used_scope, a.wg, a.num_ldst, a.iters, nw, 1))
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "barrier episodes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(roi_ms / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (storage[i] = i, the reference's initialisation)",
        "config": {
            "workload": f"{a.prim}: {nw} WGs x {a.wg} threads, numLdSt={a.num_ldst}, "
                        f"numCSIters={a.iters}, {reps} launches per step (plain form), "
                        f"in-place-increment critical section",
            "prim": a.prim, "num_wgs": nw, "wg_size": a.wg, "num_ldst": a.num_ldst,
            "num_cs_iters": a.iters, "launches_per_step": reps, "mode": "inc",
            "sync_scope": {0: "xcd-aware", 1: "strict-agent"}.get(
                used_scope, "mixed (the XCD placement check fell back to strict)"),
            "parity_ok": ok, "wall_s": round(wall, 3),
        },
        "roofline": {
            "bound": "sync-latency",
            "achieved": round(1e6 / ep_us, 1),
            "peak": round(1e6 / floor_us, 1) if floor_us else None,
            "unit": "episodes/s",
            "frac": round(floor_us / ep_us, 4) if floor_us else None,
            "traffic": traffic,
            "kernel": f"k_barrier ({a.prim})",
            "avg_launch_ms": round(avg_launch_ms, 4),
            "episode_us": round(ep_us, 4),
            "floor_us": round(floor_us, 4) if floor_us else None,
            "floor": floor_text(a.prim, nw, a.wg, ncu),
            "floor_terms": floor_terms,
            "hbm": {"achieved_gbs": round(achieved_gbs, 1), "peak_gbs": HBM_PEAK_GBS,
                    "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
                    "bytes_per_episode": bytes_per_episode,
                    "note": "algorithmic bytes over HBM peak; the walk keeps regions in L2, "
                            "so measured HBM traffic (traffic, per launch) is ~2% of them"},
        },
    }
    if not a.no_extras:
        extras = {}

        def run_extra(key, fn, *args):
            # an extra that fails is reported in its own entry; the headline
            # line above is printed either way
            try:
                extras[key] = fn(*args)
            except Exception as e:  # noqa: BLE001 -- any failure of one extra
                extras[key] = {"error": f"{type(e).__name__}: {e}", "parity_ok": False}
                print(f"bench.py: extra {key} failed: {e}", file=sys.stderr)

        if "barriers" in only:
            run_extra("barriers", extra_barriers, dev, hs, a, nw, floor_terms)
        if "c1" in only:
            run_extra("c1_faMutex", extra_c1, dev, hs, floor_terms)
        if "c3" in only:
            run_extra("c3_batched", extra_c3, dev, hs, floor_terms)
        if "c5" in only:
            run_extra("c5_serving", extra_c5, dev)
        if "uts" in only:
            run_extra("uts_tree", extra_uts_tree, dev)
        if "uts_cpu" in only:
            run_extra("uts_host_workers", extra_uts_host_workers, dev)
        out["extras"] = extras
    dev.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baselines(prim, a.num_ldst, nw, a.wg, ncu, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
