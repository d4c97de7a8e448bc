"""GPU parity of UTS work stealing against the oracle: the reference binary
tree (uts.cu:290-339) gives exact leaves / leaf-sum / node counts, and
geometric SHA-1 trees give the exact node, leaf and depth counts of the
oracle's exhaustive search, including the published sample tree T1."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P32 = ctypes.POINTER(ctypes.c_int32)


def ref_tree(oracle, n, fert):
    val = np.zeros(n, np.uint32)
    ca = np.zeros(n, np.int32)
    cb = np.zeros(n, np.int32)
    lg, ls = ctypes.c_uint32(), ctypes.c_uint32()
    oracle.hso_uts_ref_tree(n, fert, val.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                            ca.ctypes.data_as(P32), cb.ctypes.data_as(P32),
                            ctypes.byref(lg), ctypes.byref(ls))
    return val, ca, cb, lg.value, ls.value


@pytest.mark.parametrize("n,fert,blocks", [(8193, 90, 0), (8193, 90, 128), (1, 50, 0),
                                           (200001, 60, 0), (1000001, 95, 512)])
def test_uts_reference_tree(dev, oracle, n, fert, blocks):
    val, ca, cb, lg, ls = ref_tree(oracle, n, fert)
    r = dev.uts_ref(val, ca, cb, blocks)
    assert (r.nodes, r.leaves, r.leaf_sum) == (n, lg, ls)
    assert r.aborted == 0


def geo_count(oracle, seed, b0, d):
    n, lv, md = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
    oracle.hso_uts_geo_count(seed, b0, d, ctypes.byref(n), ctypes.byref(lv), ctypes.byref(md))
    return n.value, lv.value, md.value


@pytest.mark.parametrize("seed,d,blocks", [(0, 1, 0), (1, 4, 0), (19, 6, 0), (7, 8, 64),
                                           (123, 9, 0), (5, 0, 0), (42, 10, 0)])
def test_uts_geo_trees(dev, oracle, seed, d, blocks):
    want = geo_count(oracle, seed, 4.0, d)
    r = dev.uts_geo(seed, 4.0, d, blocks)
    assert (r.nodes, r.leaves, r.max_depth) == want


@pytest.mark.parametrize("seed,b0,d", [(11, 2.0, 16), (12, 8.0, 6), (13, 3.5, 9),
                                       (24, 20.0, 5), (22, 40.0, 4)])
def test_uts_geo_other_branching(dev, oracle, seed, b0, d):
    """Other expected branching factors: the child-count threshold table is
    built per b0.  At b0 = 20 and 40 a parent has up to 100 children, so a
    batch's records with children overflow the 64 kept in LDS and spill to
    the deque in most generate rounds."""
    want = geo_count(oracle, seed, b0, d)
    r = dev.uts_geo(seed, b0, d, 0)
    assert (r.nodes, r.leaves, r.max_depth) == want


def test_uts_geo_T1(dev):
    """UTS sample tree T1 (-t 1 -a 3 -d 10 -b 4 -r 19): 4130071 nodes, 3305118 leaves."""
    r = dev.uts_geo(19, 4.0, 10)
    assert (r.nodes, r.leaves, r.max_depth) == (4130071, 3305118, 10)


def test_uts_serve_alone_equals_mixed(dev, oracle):
    """Continuous batching: every tree's counts are the same alone and in the mix."""
    reqs = []
    rng = np.random.default_rng(5)
    t = 0.0
    for i in range(24):
        t += rng.exponential(0.002)
        reqs.append((int(rng.integers(0, 1 << 30)), int(rng.integers(3, 9)), 4.0, t))
    resp, summ = dev.uts_serve(reqs, slots=8)
    for (seed, d, b0, _), r in zip(reqs, resp):
        assert (r.nodes, r.leaves, r.max_depth) == geo_count(oracle, seed, b0, d)
        alone = dev.uts_geo(seed, b0, d)
        assert (alone.nodes, alone.leaves, alone.max_depth) == (r.nodes, r.leaves, r.max_depth)
    assert summ.total_nodes == sum(r.nodes for r in resp)
    assert 0 < summ.p50_s <= summ.p99_s


def test_uts_serve_rejects_grid_that_cannot_be_resident(dev):
    """The serving loop is persistent (its dispatcher is the grid's last WG),
    so a grid larger than what can be co-resident is refused before launch.
    The default grid is every WG that can be resident (at least 8 per CU at
    the kernel's register use): it runs, and so do 8 per CU given
    explicitly."""
    import heterosync_amd as hs
    with pytest.raises(hs.HSError) as ei:
        dev.uts_serve([(1, 3, 4.0, 0.0)], slots=2, num_blocks=64 * dev.info.num_cu)
    assert ei.value.code == -4  # HS_EOCCUPANCY
    for nb in (0, 8 * dev.info.num_cu):
        resp, _ = dev.uts_serve([(1, 3, 4.0, 0.0)], slots=2, num_blocks=nb)
        assert resp[0].nodes > 0


def poisson_trees(seed, n, dlo, dhi, mean_gap_s):
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for _ in range(n):
        out.append((int(rng.integers(0, 1 << 30)), int(rng.integers(dlo, dhi + 1)), 4.0, t))
        t += rng.exponential(mean_gap_s)
    return out


def test_uts_serve_stated_depths(dev, oracle):
    """The serving configuration's tree shapes: geometric b0 = 4, depth
    limit uniform in 6-14, Poisson arrivals.  Every tree's counts equal its
    own alone run, and the oracle's exhaustive search where the tree is
    small enough for the CPU (depth <= 11, at most 5e6 nodes)."""
    reqs = poisson_trees(77, 48, 6, 14, 0.002)
    assert {d for _, d, _, _ in reqs} >= {6, 14}
    resp, summ = dev.uts_serve(reqs, slots=16, timeout_s=120.0)
    checked = 0
    for (seed, d, b0, _), r in zip(reqs, resp):
        alone = dev.uts_geo(seed, b0, d)
        assert (alone.nodes, alone.leaves, alone.max_depth) == (r.nodes, r.leaves, r.max_depth)
        if d <= 11 and alone.nodes <= 5_000_000:
            assert (r.nodes, r.leaves, r.max_depth) == geo_count(oracle, seed, b0, d)
            checked += 1
    assert checked >= 10
    assert summ.total_nodes == sum(r.nodes for r in resp)
    assert summ.cut_batches == 0  # the default deques never reach the bound here
    assert 0 < summ.p50_s <= summ.p99_s


def test_uts_serve_bounded_deques(dev, oracle):
    """Admission and work by deque memory: with deques too small for the
    mix's unbounded high water, batches are cut and deliveries wait instead
    of a deque overflowing, and every tree still equals its alone run."""
    import heterosync_amd as hs
    reqs = [(s, d, 4.0, 0.0) for s, d, _, _ in poisson_trees(91, 32, 8, 10, 0.0)]
    most = most_children(oracle, 4.0)  # most children a node can have
    assert most == 96
    reserve = max(d for _, d, _, _ in reqs) * (most - 1)  # deepest tree x (most - 1)
    _, free = dev.uts_serve(reqs, slots=16, num_blocks=64)
    assert free.cut_batches == 0
    assert free.deque_high_water > reserve + 256, free.deque_high_water
    cap = max(reserve + 128, free.deque_high_water // 2)
    resp, summ = dev.uts_serve(reqs, slots=16, num_blocks=64, deque_nodes=cap)
    assert summ.deque_soft == cap - reserve
    assert summ.cut_batches > 0
    assert summ.deque_high_water <= cap < free.deque_high_water
    for (seed, d, b0, _), r in zip(reqs, resp):
        alone = dev.uts_geo(seed, b0, d)
        assert (alone.nodes, alone.leaves, alone.max_depth) == (r.nodes, r.leaves, r.max_depth)
    # a deque that cannot hold the reserve is refused before launch
    with pytest.raises(hs.HSError) as ei:
        dev.uts_serve(reqs, slots=4, deque_nodes=reserve)
    assert ei.value.code == -1  # HS_EINVAL


def most_children(oracle, b0):
    thr = np.zeros(128, np.uint32)
    return oracle.hso_uts_geo_thresholds(b0, thr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                         128)


def test_uts_geo_depth_past_the_deque_bound_is_refused(dev, oracle):
    """A depth limit whose reserve (depth x (most children - 1)) exceeds the
    65536-record deques is refused; one at the limit runs (a subcritical
    b0, so the tree stays small) and matches the oracle."""
    import heterosync_amd as hs
    with pytest.raises(hs.HSError) as ei:
        dev.uts_geo(1, 4.0, 690)  # 690 x 95 records of reserve > 65536
    assert ei.value.code == -1
    b0 = 0.9  # near-critical: seeds 1612 and 2003 reach depths 45 and 41
    d = (65536 - 1) // (most_children(oracle, b0) - 1)
    with pytest.raises(hs.HSError):
        dev.uts_geo(3, b0, d + 1)
    for seed in (3, 1612, 2003):
        r = dev.uts_geo(seed, b0, d)
        assert (r.nodes, r.leaves, r.max_depth) == geo_count(oracle, seed, b0, d)


@pytest.mark.parametrize("seed,d,threads", [(19, 10, 4), (42, 10, 16), (7, 12, 8), (3, 13, 16)])
def test_uts_geo_with_host_workers(dev, oracle, seed, d, threads):
    """Host worker threads take part in the search (each served by one GPU
    WG through a pinned mailbox): the counts are the GPU-only run's and the
    oracle's, and the host workers expanded part of the tree."""
    r, cpu = dev.uts_geo_cpu(seed, 4.0, d, threads)
    alone = dev.uts_geo(seed, 4.0, d)
    assert (r.nodes, r.leaves, r.max_depth) == (alone.nodes, alone.leaves, alone.max_depth)
    if d <= 10:
        assert (r.nodes, r.leaves, r.max_depth) == geo_count(oracle, seed, 4.0, d)
    assert cpu.threads == threads
    assert cpu.nodes < r.nodes
    if alone.nodes > 10_000_000:  # a run long enough for the host to be handed work
        assert cpu.deliveries > 0 and cpu.nodes > 0


@pytest.mark.parametrize("n,fert,threads", [(200001, 60, 4), (1000001, 95, 16)])
def test_uts_reference_tree_with_host_workers(dev, oracle, n, fert, threads):
    val, ca, cb, lg, ls = ref_tree(oracle, n, fert)
    r, cpu = dev.uts_ref_cpu(val, ca, cb, threads)
    assert (r.nodes, r.leaves, r.leaf_sum) == (n, lg, ls)
    assert cpu.nodes < n


def test_uts_host_workers_refused_beyond_the_grid(dev):
    import heterosync_amd as hs
    with pytest.raises(hs.HSError) as ei:
        dev.uts_geo_cpu(1, 4.0, 4, 200, num_blocks=64)
    assert ei.value.code == -1
