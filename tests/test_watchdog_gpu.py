"""The device watchdog, forced from the host: with a per-launch timeout far
below the run's length, every bounded wait gives up (the shared abort word
drains the grid), the call returns HS_ETIMEOUT, and the next call on the
same context runs normally and bit-exact.  No kernel hangs or faults: the
condition is a tiny limit on each wait's own deadline."""
import numpy as np
import pytest

import heterosync_amd as hs

pytestmark = pytest.mark.gpu

HS_ETIMEOUT = -5


@pytest.mark.parametrize("prim,nw,iters", [
    ("spinMutex", 512, 1000),        # CAS polls on one word
    ("faMutex", 512, 1000),          # ticket holders behind one that gave up
    ("sleepMutex", 256, 1000),       # queue slots behind one that gave up
    ("spinSemEBO10", 512, 500),      # reader / writer waits with backoff
    ("atomicTreeBarrUniq", 1024, 20000),  # release polls of a grid barrier
    ("lfTreeBarrUniq", 1024, 20000),      # flag gathers of a grid barrier
])
def test_wait_gives_up_and_next_call_runs(dev, ohelp, prim, nw, iters):
    ncu = dev.info.num_cu
    nw = min(nw, 4 * ncu)
    with pytest.raises(hs.HSError) as ei:
        dev.syncprim(prim, 10, nw, iters, reps=1, timeout_s=2e-4)
    assert ei.value.code == HS_ETIMEOUT
    # the state is per call: the same context runs the next call bit-exact
    pid = hs.prim_id(prim)
    s, r = dev.syncprim(pid, 10, 32 if pid >= 4 else 2 * ncu, 5, reps=2, mode=1,
                        timeout_s=10.0)
    want = ohelp.golden(pid, 10, 32 if pid >= 4 else 2 * ncu, 5, 2, 64, ncu, 1)
    assert np.array_equal(s, want) and r.state_ok == 1
