"""GPU parity of the batched persistent form: every instance of a seeded
mix ends bit-exact with the oracle (and with its own plain run), with exact
acquire counts, whether launched plainly or as a captured hipGraph.

Inc mode (in-place increment of every line) makes the storage count-exact:
with the reference's copy chain every pass past the L-th leaves the same
array, so only inc mode can see a lost, doubled or overlapping critical
section.  The full seeded mix of the batched configuration (64 instances,
numCSIters log-uniform in 10-1000) runs in inc mode, in the default and in
the autotuned configuration."""
import numpy as np
import pytest

import heterosync_amd as hs
from heterosync_amd.mix import make_mix

pytestmark = pytest.mark.gpu


def storages_for(dev, mix):
    return [np.arange(hs.storage_words(p, nw, dev.info.num_cu, 64, L), dtype=np.int32)
            for p, nw, it, L in mix]


def check_mix(dev, ohelp, mix, st, entries, reps, mode):
    ncu = dev.info.num_cu
    for (p, nw, it, L), s, e in zip(mix, st, entries):
        assert np.array_equal(s, ohelp.golden(p, L, nw, it, reps, 64, ncu, mode)), \
            (p, nw, it, L, mode)
        w, rd, _ = ohelp.counts(p, nw, it, reps, ncu)
        assert e == w + rd, (p, nw, it, L)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("use_graph", [0, 1])
def test_batch_mix_parity(dev, ohelp, use_graph, mode):
    mix = make_mix(seed=3, n=64, iters_cap=20)
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=3, mode=mode, use_graph=use_graph)
    assert r.all_ok == 1
    check_mix(dev, ohelp, mix, st, entries, 3, mode)
    assert r.acquires == int(entries.sum())


def test_batch_full_mix_inc(dev, ohelp):
    """The batched configuration at its stated size: the seed-3 mix of 64
    instances, numCSIters uncapped (10-1000), one persistent launch, inc
    mode; every instance's storage count-exact against the oracle."""
    mix = make_mix(seed=3, n=64)
    assert max(it for _, _, it, _ in mix) > 500
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=1, mode=1, use_graph=1, timeout_s=60.0)
    assert r.all_ok == 1
    check_mix(dev, ohelp, mix, st, entries, 1, 1)


def test_batch_full_mix_inc_autotuned(dev, ohelp):
    """The same mix in the configuration the autotuner picks (tuned in inc
    mode), run twice in a row: the tuned knobs change scheduling only."""
    mix = make_mix(seed=3, n=64)
    t = dev.batch_autotune(mix, reps=1, mode=1)
    assert t["wg_size"] in (64, 128, 256) and t["backoff_min"] in (1, 4, 16)
    assert t["probe"] in (0, 1) and t["cohort"] in (0, 8, 64)
    assert t["persistent_wgs"] > 0 and t["best_ms"] > 0
    for _ in range(2):
        st = storages_for(dev, mix)
        entries, r = dev.batch(mix, st, reps=1, mode=1, use_graph=1, timeout_s=60.0,
                               **dev.tuned_kwargs(t))
        assert r.all_ok == 1
        check_mix(dev, ohelp, mix, st, entries, 1, 1)


@pytest.mark.parametrize("probe", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("wg_size", [64, 128, 256])
def test_batch_worker_width(dev, ohelp, wg_size, mode, probe):
    """Workers of 2 and 4 waves split each critical section's lines over
    their waves (numLdSt 1, 10 and 100 all occur in the mix); both attempt
    forms (lane 0 in turn, or one probe per lane)."""
    mix = make_mix(seed=23, n=48, iters_cap=30)
    assert {L for _, _, _, L in mix} == {1, 10, 100}
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=2, mode=mode, wg_size=wg_size, use_graph=1,
                           probe=probe)
    assert r.all_ok == 1
    check_mix(dev, ohelp, mix, st, entries, 2, mode)


@pytest.mark.parametrize("probe", [0, 1])
@pytest.mark.parametrize("seed", [17, 29, 41])
@pytest.mark.parametrize("workers", [256, 2048])
def test_batch_mix_seeds_and_workers(dev, ohelp, seed, workers, probe):
    """Other mixes, with the virtual WGs spread over few workers (many per
    worker) or many (a few each)."""
    mix = make_mix(seed=seed, n=32, iters_cap=15)
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=2, mode=1, use_graph=1, persistent_wgs=workers,
                           probe=probe)
    assert r.all_ok == 1
    check_mix(dev, ohelp, mix, st, entries, 2, 1)


@pytest.mark.parametrize("probe", [0, 1])
@pytest.mark.parametrize("cohort", [8, 64])
def test_batch_full_mix_inc_cohort(dev, ohelp, cohort, probe):
    """The full mix with the spin / EBO mutex cohort: a lock holder passes
    the lock to its worker's other waiting virtual WGs of the instance (up
    to `cohort` critical sections per hold).  Inc mode, so a critical
    section that overlapped another or ran twice changes the storage; every
    instance still makes exactly numWGs x numCSIters entries."""
    mix = make_mix(seed=3, n=64)
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=1, mode=1, use_graph=1, timeout_s=60.0,
                           wg_size=256, probe=probe, cohort=cohort)
    assert r.all_ok == 1
    check_mix(dev, ohelp, mix, st, entries, 1, 1)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("wg_size", [64, 128, 256])
def test_batch_cohort_widths_and_modes(dev, ohelp, wg_size, mode):
    """The cohort with every worker width, in both modes, over a mix with
    many virtual WGs per worker (blocks of one instance share a worker)."""
    mix = make_mix(seed=31, n=40, iters_cap=25)
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=2, mode=mode, wg_size=wg_size, use_graph=1,
                           cohort=16, persistent_wgs=128)
    assert r.all_ok == 1
    check_mix(dev, ohelp, mix, st, entries, 2, mode)


@pytest.mark.parametrize("cohort", [0, 8])
def test_batch_edge_instances(dev, ohelp, cohort):
    """Instances at the edges of their arguments in one batch: a single
    virtual WG, no critical sections at all (storage untouched, no entries),
    empty critical sections (numLdSt 0), beside ordinary ones."""
    mix = [(4, 1, 5, 1), (5, 3, 0, 10), (6, 2, 4, 0), (7, 1, 3, 100), (8, 1, 2, 1),
           (11, 5, 0, 1), (10, 9, 3, 0), (4, 40, 6, 10), (9, 17, 4, 10)]
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=2, mode=1, use_graph=1, cohort=cohort)
    assert r.all_ok == 1
    check_mix(dev, ohelp, mix, st, entries, 2, 1)
    assert entries[1] == 0 and entries[5] == 0


def test_batch_alone_equals_plain(dev):
    mix = make_mix(seed=11, n=8, iters_cap=10)
    for mode in (0, 1):
        st = storages_for(dev, mix)
        dev.batch(mix, st, reps=2, mode=mode)
        for (p, nw, it, L), s in zip(mix, st):
            plain, r = dev.syncprim(p, L, nw, it, reps=2, mode=mode)
            assert np.array_equal(plain, s), (p, nw, it, L, mode)


def test_batch_rejects_bad_width_and_mode(dev):
    mix = make_mix(seed=11, n=2, iters_cap=2)
    for kw in (dict(wg_size=96), dict(wg_size=512), dict(mode=2), dict(probe=2),
               dict(cohort=-1)):
        with pytest.raises(hs.HSError) as ei:
            dev.batch(mix, storages_for(dev, mix), reps=1, **kw)
        assert ei.value.code == -1  # HS_EINVAL
