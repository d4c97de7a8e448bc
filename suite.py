#!/usr/bin/env python3
"""Measurement suite for the synchronization hot path on one GPU.

  C1  plain faMutex 32 WGs x 64 lanes, numLdSt=10, numCSIters=100 (10 launches)
  C2  grid barriers 1024 WGs x 256 lanes, numLdSt=10, numCSIters=1000, every
      barrier form (tree, local exchange, lock-free, sense-reversing) in both
      sync scopes
  C3  64-instance mutex/semaphore mix: batched persistent launch vs plain
  C4  relaxed atomics 1024 WGs x 256 lanes x 10^4 iterations
  C5  UTS serving loop vs one launch per tree (geometric SHA-1 trees, depth
      limit U[6,14]), at 64 and 16 slots; Poisson arrivals at 0.5, 0.8 and
      0.95 x the plain form's capacity (64 slots)

C1, C3 and C5 are bench.py's own measurements.  Every line carries its
parity check: the library's closed-form self-check (the reference binary's
golden), exact counts, or the same tree alone and in the mix.  Usage:
  python suite.py [--no-probe] [c1 c2 c3 c4 c5]
--no-probe skips the floor probe (a child process), for runs under a
profiler; the floor fields are then left out.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
import heterosync_amd as hs  # noqa: E402

SRB_NAMES = ("atomicTreeBarrUniqSRB", "atomicTreeBarrUniqLocalExchSRB",
             "atomicTreeBarrUniqNaiveSRB", "atomicTreeBarrUniqLocalExchNaiveSRB",
             "atomicTreeBarrUniqNaiveAllSRB", "atomicTreeBarrUniqLocalExchNaiveAllSRB")


def c1(dev, terms):
    return bench.extra_c1(dev, hs, terms)


def c2(dev, terms):
    out = {}
    ncu = dev.info.num_cu
    for name in bench.BARRIERS + SRB_NAMES:
        for sc in (0, 1):
            s, r = dev.syncprim(name, 10, 4 * ncu, 1000, wg_size=256, reps=10, mode=1,
                                timeout_s=30.0, sync_scope=sc)
            want = hs.expected_storage(name, 10, 4 * ncu, 1000, 10, 256, ncu, 1)
            form = "strict" if r.sync_scope_used else "xcd"  # the form that ran
            ep_us = r.kernel_ms * 1e3 / r.episodes
            line = {"episodes_per_s": r.episodes / (r.roi_ms / 1e3), "episode_us": ep_us,
                    "parity_ok": bool(np.array_equal(s, want)) and r.state_ok == 1}
            if terms and name in bench.BARRIERS and not r.sync_scope_used:
                fl = bench.barrier_floor_us(name, terms, ncu)
                line.update(floor_us=fl, frac_of_floor=fl / ep_us)
            out[f"{name}_{form}{'' if r.sync_scope_used == sc else '_fallback'}"] = line
    return out


def c3(dev, terms):
    return bench.extra_c3(dev, hs, terms)


def c4(dev, terms, iters=10000):
    ncu = dev.info.num_cu
    nw = 4 * ncu
    out = {}
    # seqlock: 1 writer WG per 64 readers (writer period 65), 16 locks; the
    # default (drained agent-scope accesses) and the strict (fenced) form.
    # Lock l ends at 2 x iters x (writer WGs b with b % 16 == l).
    writers = np.bincount([b % 16 for b in range(0, nw, 65)], minlength=16)
    for key, variant in (("seqlock", 0), ("seqlock_strict", 2)):
        seq, _, torn, r = dev.seqlock(nw, 256, 16, 65, iters, variant)
        out[key] = {"ms": r.kernel_ms, "atomic_ops": r.atomic_ops,
                    "atomic_ops_per_s": r.atomic_ops / (r.kernel_ms / 1e3),
                    "passes_per_s": nw * iters / (r.kernel_ms / 1e3),
                    "parity_ok": bool(np.array_equal(seq[:16], 2 * iters * writers))
                    and torn == 0}
    # flags: worker WG b's iteration it reads rand[b + it * nw]; the input is
    # uniform on [0, 0.5] like the reference's rand() / UINT32_MAX
    # (flags.cu:92-101), from numpy's generator; dirty counts the worker
    # reads at or under 0.33
    rnd = (np.random.default_rng(2018).random(nw * iters) * 0.5).astype(np.float32)
    want_dirty = int(np.count_nonzero(rnd.reshape(iters, nw)[:, :nw - 1] <= np.float32(0.33)))
    fo, dirty, r = dev.flags(nw, 256, 16, 4, 0.33, iters, rnd)
    out["flags"] = {"ms": r.kernel_ms, "dirty": dirty,
                    "worker_iters_per_s": (nw - 1) * iters / (r.kernel_ms / 1e3),
                    "parity_ok": dirty == want_dirty and bool(np.all(
                        fo.reshape(nw, 256)[-1] == (4 * 16 * 16 if dirty else 0)))}
    # refCounter: 32 sharers, 10^4 inc/dec pairs, spacing 8; default and
    # strict (fenced) forms.  Every counter ends at 0 and saw its last release.
    for key, strict in (("refCounter", False), ("refCounter_strict", True)):
        c0, c1_, d0, d1, r = dev.refcounter(32, nw, 256, iters, 8, strict=strict)
        groups = nw // 32
        okr = not c0.any() and not c1_.any() and bool(
            np.all(d0.reshape(32, groups, 256).max(axis=0) == 1)) and bool(
            np.all(d1.reshape(32, groups, 256).max(axis=0) == 1))
        out[key] = {"ms": r.kernel_ms, "atomic_ops": r.atomic_ops,
                    "atomic_ops_per_s": r.atomic_ops / (r.kernel_ms / 1e3), "parity_ok": okr}
    # split counter: every counter ends at iters; reads lie in [0, total]
    cnt, so, r = dev.splitcounter(nw, 256, iters)
    out["splitCounter"] = {"ms": r.kernel_ms, "atomic_ops": r.atomic_ops,
                           "adds_per_s": cnt.size * iters / (r.kernel_ms / 1e3),
                           "parity_ok": bool(np.all(cnt == iters)) and
                           int(so.max()) <= cnt.size * iters}
    return out


def c5(dev, terms):
    return {"slots_64": bench.extra_c5(dev, slots=64, loads=(0.8, 0.5, 0.95)),
            "slots_16": bench.extra_c5(dev, slots=16)}


def main(argv):
    probe = "--no-probe" not in argv
    argv = [x for x in argv if x != "--no-probe"]
    which = argv or ["c1", "c2", "c3", "c4", "c5"]
    terms = bench.probe_floor(1024, 256, 10) if probe else None
    dev = hs.Device(0)
    out = {"floor_terms": terms}
    print(json.dumps(out), flush=True)
    for w in which:
        out[w] = globals()[w](dev, terms)
        print(json.dumps({w: out[w]}), flush=True)
    dev.close()
    return out


if __name__ == "__main__":
    main(sys.argv[1:])
