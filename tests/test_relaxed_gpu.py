"""GPU parity of the relaxed-atomics kernels against the oracle's exact
invariants: seqlock sequence numbers (closed form) and no torn snapshot;
flags' dirty count and output rows; reference counters back to zero with a
release recorded for every counter; split counters exact, reader totals in
range."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def seq_expected(oracle, nw, gs, period, iters, n):
    out = np.zeros(n, np.uint32)
    oracle.hso_seqlock_expected(nw, gs, period, iters,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n)
    return out


# bit 0: the reference variant (0 hipAtomics, 1 useTFs); bit 1: strict fences
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("nw,wg,gs,period,iters", [
    (512, 128, 4, 16, 10),     # the reference's run file: ./seqlocks 512 128 4 {0,1}
    (256, 64, 16, 65, 200),
    (1024, 256, 16, 65, 50),   # full-occupancy shape, 1 writer WG per 64 readers
    (1024, 256, 16, 65, 10000),  # relaxed-atomics configuration: 10^4 iterations
    (2048, 64, 8, 33, 100),    # 8 WGs per CU, more writers
    (128, 1024, 2, 16, 20),    # 1024-thread WGs
    (1, 64, 1, 1, 1),          # one WG: a lone writer, no readers
    (65, 64, 1, 65, 1),        # one writer, 64 readers, one pass each
])
def test_seqlock(dev, oracle, variant, nw, wg, gs, period, iters):
    seq, out, torn, r = dev.seqlock(nw, wg, gs, period, iters, variant)
    want = seq_expected(oracle, nw, gs, period, iters, len(seq))
    assert np.array_equal(seq[:gs], want[:gs])
    assert all(int(x) % 2 == 0 for x in seq)
    if variant % 2 == 0:  # readers load (do not exchange): snapshots are checkable
        assert torn == 0
        # every reader snapshot is r1 + r2 with r1 == r2 == a writer's even value
        readers = np.array([b for b in range(nw) if b % period], np.int64)
        o = out.reshape(nw, wg)[readers]
        assert np.all(o % 4 == 0) and np.all(o >= 0)
    assert r.aborted == 0


@pytest.mark.parametrize("ntb,tb,accs,reps,cut,iters", [
    (160, 64, 16, 4, 0.33, 1),      # the reference's run file (tbSize 32 -> wave64)
    (256, 64, 16, 4, 0.2, 100),
    (1024, 256, 16, 4, 0.33, 200),
    (64, 64, 16, 4, 0.45, 2),       # the next value is read one iteration ahead
    (1024, 256, 16, 4, 0.33, 10000),  # C4
    (2, 64, 1, 1, 0.33, 1),         # the main WG and a single worker
])
def test_flags(dev, oracle, ntb, tb, accs, reps, cut, iters):
    rnd = np.zeros(ntb * iters, np.float32)
    oracle.hso_flags_rand(rnd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), rnd.size, 2018)
    want = oracle.hso_flags_dirty(rnd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ntb, iters,
                                  ctypes.c_float(cut))
    out, dirty, r = dev.flags(ntb, tb, accs, reps, cut, iters, rnd)
    assert dirty == want
    o = out.reshape(ntb, tb)
    assert np.all(o[:-1] == 0)
    assert np.all(o[-1] == (reps * accs * accs if dirty else 0))


def test_flags_never_dirty(dev):
    rnd = np.ones(64 * 10, np.float32)  # all > cutoff
    out, dirty, _ = dev.flags(64, 64, 8, 2, 0.5, 10, rnd)
    assert dirty == 0 and not out.any()


@pytest.mark.parametrize("strict", [False, True])
@pytest.mark.parametrize("sh,ntb,tb,reps,spacing", [
    (32, 64, 32, 4, -1),            # the reference's run file
    (32, 1024, 256, 100, 8),
    (32, 1024, 256, 10000, 8),   # relaxed-atomics configuration: 10^4 inc/dec pairs
    (4, 512, 64, 50, 0),
    (1, 1, 64, 1, 0),               # one WG sharing its counters with nobody
])
def test_refcounter(dev, sh, ntb, tb, reps, spacing, strict):
    c0, c1, d0, d1, r = dev.refcounter(sh, ntb, tb, reps, spacing, strict=strict)
    assert not c0.any() and not c1.any()
    groups = ntb // sh
    for d in (d0, d1):
        per_counter = d.reshape(ntb, tb).reshape(sh, groups, tb).max(axis=0)
        assert np.all(per_counter == 1)  # every counter saw its last release
    assert r.atomic_ops == 4 * ntb * tb * reps


@pytest.mark.parametrize("ntb,tb,iters,read_every", [
    (112, 128, 1, 0), (256, 256, 1000, 0), (1024, 256, 200, 0), (64, 1024, 300, 0),
    (1024, 256, 10000, 0),    # relaxed-atomics configuration: 10^4 iterations
    (1024, 256, 10000, 1),    # a WG read after every writer iteration
    (1024, 256, 10000, 2500),  # four reads per WG
    (256, 256, 100, 1000),    # read period past the run: one read at the end
    (1, 128, 1, 0),           # one WG: one writer wave, one reader wave
])
def test_splitcounter(dev, ntb, tb, iters, read_every):
    counters, out, r = dev.splitcounter(ntb, tb, iters, read_every)
    assert np.all(counters == iters)
    total = iters * counters.size
    assert out.max() <= total
    w = out.reshape(ntb, tb // 64, 64)
    assert np.all(w[:, 0::2, :] == 0)  # writer waves write no output
    # the read is WG-collective (CSK:33-76): every reader thread of a WG
    # outputs the same sum
    readers = w[:, 1::2, :].reshape(ntb, -1)
    assert np.all(readers == readers[:, :1])
    # operation count: every writer add, plus each WG's reads (its own
    # counters' lines, then every WG's partial, plus the partial's store)
    period = min(read_every if read_every > 0 else 64, max(iters, 1))
    reads = iters // period
    per_read = (tb // 2 + 63) // 64 + (ntb + 63) // 64 + 1
    assert r.atomic_ops == counters.size * iters + ntb * reads * per_read * 64
