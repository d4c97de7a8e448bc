"""GPU parity of the syncPrims hot path (hs_syncprim_run through the C ABI)
against the oracle: final storage bit-exact with the closed form (and with
the literal restatement of the reference golden at small sizes), CS entry /
episode counts exact, lock and barrier words back in their idle state."""
import numpy as np
import pytest

import heterosync_amd as hs

pytestmark = pytest.mark.gpu

MUTEX = [4, 5, 6, 7]
UMUTEX = [20, 21, 22, 23]
SEM = list(range(8, 16))
PRIO = [16, 17, 18, 19, 32, 33, 34, 35]
USEM = list(range(24, 32))
BARRIER = [0, 1, 2, 3]
SRB = [37, 38, 39, 40, 41, 42]


def srb_scope(prim, nw, ncu, scope):
    """Every SRB form honours the requested scope: the tree runs XCD-aware,
    the flat barriers fence-free over agent-scope data."""
    return scope


def check(dev, ohelp, prim, L, nw, iters, reps=10, wg=64, mode=0, literal=False, **kw):
    ncu = dev.info.num_cu
    s, r = dev.syncprim(prim, L, nw, iters, wg_size=wg, reps=reps, mode=mode,
                        timeout_s=10.0, **kw)
    want = ohelp.golden(prim, L, nw, iters, reps, wg, ncu, mode, literal=literal)
    assert np.array_equal(s, want), (prim, L, nw, iters, mode, int((s != want).sum()))
    w, rd, ep = ohelp.counts(prim, nw, iters, reps, ncu)
    assert (r.cs_entries, r.rd_entries, r.episodes) == (w, rd, ep)
    assert r.state_ok == 1
    return r


# sync_scope 0: fence-free hand-off over agent-scope data; 1: strict
# release/acquire with plain data accesses
@pytest.mark.parametrize("prim", MUTEX)
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("scope", [0, 1])
def test_mutex(dev, ohelp, prim, mode, scope):
    r = check(dev, ohelp, prim, 10, 32, 20, mode=mode, sync_scope=scope)
    assert r.sync_scope_used == scope
    check(dev, ohelp, prim, 3, 4, 5, mode=mode, literal=True, sync_scope=scope)


@pytest.mark.parametrize("prim", UMUTEX)
@pytest.mark.parametrize("scope", [0, 1])
def test_uniq_mutex(dev, ohelp, prim, scope):
    ncu = dev.info.num_cu
    check(dev, ohelp, prim, 4, 2 * ncu, 10, mode=1, sync_scope=scope)
    check(dev, ohelp, prim, 4, ncu + 3, 5, mode=1, sync_scope=scope)   # uneven WGs per lock
    check(dev, ohelp, prim, 2, 8, 5, mode=0, literal=True, sync_scope=scope)


@pytest.mark.parametrize("prim", SEM)
@pytest.mark.parametrize("scope", [0, 1])
def test_semaphore(dev, ohelp, prim, scope):
    ncu = dev.info.num_cu
    r = check(dev, ohelp, prim, 4, 16, 10, mode=1, sync_scope=scope)
    assert r.sync_scope_used == scope
    check(dev, ohelp, prim, 4, 2 * ncu, 2, reps=2, mode=1, sync_scope=scope)  # 1 writer + readers per CU


@pytest.mark.parametrize("prim", PRIO)
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("scope", [0, 1])
def test_priority_semaphore(dev, ohelp, prim, mode, scope):
    """Entries yield to posts in flight; 32-35 run the global kernel over
    num_wgs regions (the reference's dispatch), so only regions < perCU move."""
    ncu = dev.info.num_cu
    check(dev, ohelp, prim, 4, 16, 10, mode=mode, sync_scope=scope)
    check(dev, ohelp, prim, 4, 2 * ncu, 2, reps=2, mode=mode, sync_scope=scope)
    r = check(dev, ohelp, prim, 2, 4 * ncu, 3, reps=2, mode=mode, sync_scope=scope)
    assert r.state_ok == 1 and r.sync_scope_used == scope


@pytest.mark.parametrize("prim", USEM)
@pytest.mark.parametrize("scope", [0, 1])
def test_uniq_semaphore(dev, ohelp, prim, scope):
    ncu = dev.info.num_cu
    check(dev, ohelp, prim, 4, 2 * ncu, 5, mode=1, sync_scope=scope)
    check(dev, ohelp, prim, 2, 4 * ncu, 3, reps=3, mode=0, sync_scope=scope)


@pytest.mark.parametrize("prim", BARRIER)
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("scope", [0, 1])
def test_barrier_small(dev, ohelp, prim, mode, scope):
    ncu = dev.info.num_cu
    # up to 16 WGs per CU: the XCD-aware sub-groups take 16, 32 and 64 WGs
    for nw in (1, 5, 8, 64, ncu, 4 * ncu, 8 * ncu, 16 * ncu):
        r = check(dev, ohelp, prim, 4, nw, 20, mode=mode, sync_scope=scope)
        assert r.sync_scope_used == scope
    check(dev, ohelp, prim, 2, 2 * ncu, 3, reps=2, mode=mode, wg=128, sync_scope=scope)


@pytest.mark.parametrize("prim", BARRIER)
@pytest.mark.parametrize("scope", [0, 1])
def test_barrier_uneven_xcds(dev, ohelp, prim, scope):
    """Grids that do not spread evenly over the XCDs (up to NUM_CU WGs; a
    larger grid is a multiple of it): the XCDs hold different numbers of
    WGs and sub-groups, so the XCD leaders gather different numbers of
    flags.  Several flags per lane: test_barrier_small's 64-lane grids of
    4 to 16 WGs per CU."""
    for nw in (3, 13, 100, 203):
        for wg in (64, 256):
            r = check(dev, ohelp, prim, 3, nw, 40, reps=2, wg=wg, mode=1, sync_scope=scope)
            assert r.sync_scope_used == scope


@pytest.mark.parametrize("prim", BARRIER + SRB)
@pytest.mark.parametrize("scope", [0, 1])
def test_barrier_regions_return(dev, ohelp, prim, scope):
    """More episodes per launch than WGs: every region passes through every
    XCD's WGs and comes back (the XCD-aware walks keep a region in one L2
    between hand-overs, so its return must not meet a stale copy)."""
    ncu = dev.info.num_cu
    for nw, iters in ((64, 200), (ncu, 3 * ncu + 5), (4 * ncu, 2 * 4 * ncu + 3)):
        check(dev, ohelp, prim, 3, nw, iters, reps=2, mode=1, sync_scope=scope)


@pytest.mark.parametrize("prim", MUTEX + UMUTEX + SEM + USEM)
@pytest.mark.parametrize("scope", [0, 1])
def test_locks_large_wgs(dev, ohelp, prim, scope):
    """256- and 1024-thread WGs: the critical section's band walk spans 16
    or 64 chains per WG."""
    ncu = dev.info.num_cu
    check(dev, ohelp, prim, 10, 16, 4, reps=2, wg=1024, mode=1, sync_scope=scope)
    check(dev, ohelp, prim, 5, 2 * ncu, 3, reps=2, wg=256, mode=0, sync_scope=scope)


@pytest.mark.parametrize("prim", BARRIER + SRB)
@pytest.mark.parametrize("scope", [0, 1])
def test_barrier_large_wgs(dev, ohelp, prim, scope):
    """512- and 1024-thread WGs: a WG's region spans 32 or 64 chains, so the
    band walk's lanes cover many chains each."""
    ncu = dev.info.num_cu
    check(dev, ohelp, prim, 10, ncu, 30, reps=2, wg=1024, mode=1, sync_scope=scope)
    check(dev, ohelp, prim, 7, 2 * ncu, 30, reps=2, wg=512, mode=0, sync_scope=scope)


@pytest.mark.parametrize("prim", BARRIER)
@pytest.mark.parametrize("scope", [0, 1])
def test_barrier_full_config(dev, ohelp, prim, scope):
    """1024 WGs x 256 threads, L=10, 1000 iterations x 10 launches."""
    ncu = dev.info.num_cu
    r = check(dev, ohelp, prim, 10, 4 * ncu, 1000, wg=256, mode=1, sync_scope=scope)
    assert r.episodes == 10000 and r.sync_scope_used == scope


@pytest.mark.parametrize("prim", SRB)
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("scope", [0, 1])
def test_srb_small(dev, ohelp, prim, mode, scope):
    """Sense-reversing family: flat below 5 WGs per group, tree above."""
    ncu = dev.info.num_cu
    for nw in (1, 5, 64, ncu, 2 * ncu, 4 * ncu, 8 * ncu, 16 * ncu):
        r = check(dev, ohelp, prim, 3, nw, 12, reps=3, mode=mode, sync_scope=scope)
        assert r.sync_scope_used == srb_scope(prim, nw, ncu, scope)
    check(dev, ohelp, prim, 2, 2 * ncu, 3, reps=2, mode=mode, wg=128, sync_scope=scope)
    check(dev, ohelp, prim, 2, 3, 4, reps=2, mode=mode, literal=True, sync_scope=scope)


@pytest.mark.parametrize("prim", SRB)
@pytest.mark.parametrize("scope", [0, 1])
def test_srb_full_config(dev, ohelp, prim, scope):
    """1024 WGs x 256 threads, L=10, 1000 iterations x 2 launches."""
    ncu = dev.info.num_cu
    r = check(dev, ohelp, prim, 10, 4 * ncu, 1000, reps=2, wg=256, mode=1, sync_scope=scope)
    assert r.episodes == 2000 and r.sync_scope_used == scope


def test_fa_mutex_c1(dev, ohelp):
    """faMutex 32 WG x 64, numLdSt=10, 100 iterations: 32000 acquires."""
    for mode in (0, 1):
        for scope in (0, 1):
            r = check(dev, ohelp, 7, 10, 32, 100, mode=mode, sync_scope=scope)
            assert r.lock_words[0] == r.lock_words[4] == 32000  # ticket == turn


@pytest.mark.parametrize("prim", [i for i in range(43) if i != 36])
def test_edge_shapes(dev, ohelp, prim):
    """Every primitive id at the edges of its arguments: numLdSt 0 (an empty
    critical section), numCSIters 0 (no critical section at all, storage
    untouched), a single WG, and one WG per CU; both modes."""
    ncu = dev.info.num_cu
    for L, nw, it in ((0, 1, 1), (1, 1, 0), (0, ncu, 3), (3, 1, 4), (2, 8, 0)):
        for mode in (0, 1):
            check(dev, ohelp, prim, L, nw, it, reps=2, mode=mode)


@pytest.mark.parametrize("prim", [4, 7, 11, 23, 31])
def test_long_critical_section(dev, ohelp, prim):
    """A long critical section (numLdSt 500: each thread's chain is 501
    lines), in inc mode, where every line of every pass is counted."""
    check(dev, ohelp, prim, 500, 8, 6, reps=2, mode=1)


def test_error_paths(dev):
    ncu = dev.info.num_cu
    with pytest.raises(hs.HSError) as e:
        dev.syncprim(36, 10, 8, 1)           # id 36 is reserved in the reference
    assert e.value.code == -1
    with pytest.raises(hs.HSError) as e:
        dev.syncprim(0, 10, ncu + 1, 1)      # barrier grid not a multiple of NUM_CU
    assert e.value.code == -1
    with pytest.raises(hs.HSError) as e:
        dev.syncprim(0, 10, 8 * ncu, 1, wg_size=1024)  # cannot all be co-resident
    assert e.value.code == -4
