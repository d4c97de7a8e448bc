"""Seeded generator of the batched mutex/semaphore mix: 64 independent
instances with types uniform over {spinMutex, spinMutexEBO, faMutex,
sleepMutex, spinSem1, spinSem2, spinSem10, spinSem120}, numTBs uniform in
[8, 512] (64 threads each), numCSIters log-uniform in [10, 1000] and numLdSt
in {1, 10, 100}."""
import numpy as np

TYPES = ["spinMutex", "spinMutexEBO", "faMutex", "sleepMutex",
         "spinSem1", "spinSem2", "spinSem10", "spinSem120"]
TYPE_IDS = [4, 5, 7, 6, 8, 9, 10, 11]


def make_mix(seed=3, n=64, iters_cap=None):
    """List of (prim_id, num_wgs, num_cs_iters, num_ldst) tuples."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        t = int(rng.integers(0, len(TYPES)))
        nw = int(rng.integers(8, 513))
        iters = int(round(10 ** rng.uniform(1.0, 3.0)))
        if iters_cap is not None:
            iters = min(iters, iters_cap)
        L = int(rng.choice([1, 10, 100]))
        out.append((TYPE_IDS[t], nw, iters, L))
    return out
