"""GPU parity of the batched persistent form: every instance of a seeded
mix ends bit-exact with the oracle (and with its own plain run), with exact
acquire counts, whether launched plainly or as a captured hipGraph."""
import numpy as np
import pytest

import heterosync_amd as hs
from heterosync_amd.mix import make_mix

pytestmark = pytest.mark.gpu


def storages_for(dev, mix):
    return [np.arange(hs.storage_words(p, nw, dev.info.num_cu, 64, L), dtype=np.int32)
            for p, nw, it, L in mix]


@pytest.mark.parametrize("use_graph", [0, 1])
def test_batch_mix_parity(dev, ohelp, use_graph):
    mix = make_mix(seed=3, n=64, iters_cap=20)
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=3, use_graph=use_graph)
    assert r.all_ok == 1
    ncu = dev.info.num_cu
    for (p, nw, it, L), s, e in zip(mix, st, entries):
        assert np.array_equal(s, ohelp.golden(p, L, nw, it, 3, 64, ncu, 0)), (p, nw, it, L)
        w, rd, _ = ohelp.counts(p, nw, it, 3, ncu)
        assert e == w + rd
    assert r.acquires == int(entries.sum())


@pytest.mark.parametrize("seed", [17, 29, 41])
@pytest.mark.parametrize("workers", [256, 2048])
def test_batch_mix_seeds_and_workers(dev, ohelp, seed, workers):
    """Other mixes, with the virtual WGs spread over few workers (many per
    worker) or many (a few each)."""
    mix = make_mix(seed=seed, n=32, iters_cap=15)
    st = storages_for(dev, mix)
    entries, r = dev.batch(mix, st, reps=2, use_graph=1, persistent_wgs=workers)
    assert r.all_ok == 1
    ncu = dev.info.num_cu
    for (p, nw, it, L), s, e in zip(mix, st, entries):
        assert np.array_equal(s, ohelp.golden(p, L, nw, it, 2, 64, ncu, 0)), (p, nw, it, L)
        w, rd, _ = ohelp.counts(p, nw, it, 2, ncu)
        assert e == w + rd


def test_batch_alone_equals_plain(dev):
    mix = make_mix(seed=11, n=8, iters_cap=10)
    st = storages_for(dev, mix)
    dev.batch(mix, st, reps=2)
    for (p, nw, it, L), s in zip(mix, st):
        plain, r = dev.syncprim(p, L, nw, it, reps=2)
        assert np.array_equal(plain, s)


def test_batch_autotune(dev):
    mix = make_mix(seed=5, n=16, iters_cap=50)
    wgs, poll, bmax, best = dev.batch_autotune(mix)
    assert wgs > 0 and best > 0
    st = storages_for(dev, mix)
    _, r = dev.batch(mix, st, reps=2, persistent_wgs=wgs, poll_sleep=poll,
                     backoff=(0, bmax))
    assert r.all_ok == 1
