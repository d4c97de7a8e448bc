"""CPU tests that pin the oracle (oracle/hs_oracle.c, oracle/cpu_emul.cpp).

Each function of the oracle is checked against something it does not
restate: published known-answer vectors and the system glibc's own rand()
output (fixtures in tests/golden, written by make_fixtures.py), an
independent implementation (hashlib, a Python glibc-rand restatement),
brute force over the kernels' access pattern, or a real concurrent run
(std::thread emulation).  Planted mistakes (a wrong pass count, a dropped
line) must be caught.
"""
import ctypes
import hashlib
import json
import math
import os
import random

import numpy as np
import pytest

P8 = ctypes.POINTER(ctypes.c_uint8)
P32 = ctypes.POINTER(ctypes.c_int32)
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _fips_message(spec):
    return spec.encode() if isinstance(spec, str) else (spec["repeat"] * spec["times"]).encode()

SUPPORTED = list(range(36)) + list(range(37, 43))
# priority EBO semaphores: the reference dispatches them to the global kernel
# (main.hip.cpp:2455-2470) but checks them with the uniq golden
# (main.hip.cpp:2667-2680), so its own check disagrees with its kernel
PRIO_EBO = [32, 33, 34, 35]

REF_NAMES = {  # hip/syncPrims/uvm/main.hip.cpp:2036-2089
    "atomicTreeBarrUniq": 0, "atomicTreeBarrUniqLocalExch": 1, "lfTreeBarrUniq": 2,
    "lfTreeBarrUniqLocalExch": 3, "spinMutex": 4, "spinMutexEBO": 5, "sleepMutex": 6,
    "faMutex": 7, "spinSem1": 8, "spinSem2": 9, "spinSem10": 10, "spinSem120": 11,
    "spinSemEBO1": 12, "spinSemEBO2": 13, "spinSemEBO10": 14, "spinSemEBO120": 15,
    "PRspinSem1": 16, "PRspinSem2": 17, "PRspinSem10": 18, "PRspinSem120": 19,
    "spinMutexUniq": 20, "spinMutexEBOUniq": 21, "sleepMutexUniq": 22, "faMutexUniq": 23,
    "spinSemUniq1": 24, "spinSemUniq2": 25, "spinSemUniq10": 26, "spinSemUniq120": 27,
    "spinSemEBOUniq1": 28, "spinSemEBOUniq2": 29, "spinSemEBOUniq10": 30,
    "spinSemEBOUniq120": 31, "PRspinSemEBO1": 32, "PRspinSemEBO2": 33,
    "PRspinSemEBO10": 34, "PRspinSemEBO120": 35, "atomicTreeBarrUniqSRB": 37,
    "atomicTreeBarrUniqLocalExchSRB": 38, "atomicTreeBarrUniqNaiveSRB": 39,
    "atomicTreeBarrUniqLocalExchNaiveSRB": 40, "atomicTreeBarrUniqNaiveAllSRB": 41,
    "atomicTreeBarrUniqLocalExchNaiveAllSRB": 42,
}


def sha1(oracle, msg: bytes) -> bytes:
    out = (ctypes.c_uint8 * 20)()
    oracle.hso_sha1(msg, len(msg), out)
    return bytes(out)


# ---- SHA-1 (FIPS 180-4 known answers) --------------------------------------

@pytest.mark.parametrize("vec", fixture("sha1_fips180.json")["vectors"],
                         ids=lambda v: str(v["message"])[:12])
def test_sha1_known_answers(oracle, vec):
    assert sha1(oracle, _fips_message(vec["message"])).hex() == vec["sha1"]


def test_sha1_matches_hashlib_all_lengths(oracle):
    rng = random.Random(7)
    for n in range(0, 200):
        m = bytes(rng.randrange(256) for _ in range(n))
        assert sha1(oracle, m) == hashlib.sha1(m).digest(), n


# ---- UTS geometric trees ---------------------------------------------------

def test_uts_rng_definition(oracle):
    """Root = SHA1(16 zero bytes || seed BE32); child i = SHA1(parent || i BE32)."""
    root = (ctypes.c_uint8 * 20)()
    oracle.hso_uts_root(19, root)
    assert bytes(root) == hashlib.sha1(bytes(16) + (19).to_bytes(4, "big")).digest()
    child = (ctypes.c_uint8 * 20)()
    oracle.hso_uts_spawn(root, 3, child)
    assert bytes(child) == hashlib.sha1(bytes(root) + (3).to_bytes(4, "big")).digest()
    want = int.from_bytes(bytes(child)[16:20], "big") & 0x7FFFFFFF
    assert oracle.hso_uts_rand(child) == want


def test_uts_T1_published_sample_tree(oracle):
    """UTS sample tree T1 (-t 1 -a 3 -d 10 -b 4 -r 19): 4130071 nodes,
    3305118 leaves, depth 10 -- the public UTS distribution's figures."""
    t1 = fixture("uts_t1.json")
    pr = t1["params"]
    n, lv, md = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
    oracle.hso_uts_geo_count(pr["root_seed"], pr["b0"], pr["gen_mx"], ctypes.byref(n),
                             ctypes.byref(lv), ctypes.byref(md))
    assert (n.value, lv.value, md.value) == (t1["nodes"], t1["leaves"], t1["max_depth"])


def _py_children(h, b0, depth, gen_mx):
    b = b0 if (depth < gen_mx or depth == 0) else 0.0
    p = 1.0 / (1.0 + b)
    u = h / 2147483648.0
    if p >= 1.0:
        return 0
    return min(100, int(math.floor(math.log(1.0 - u) / math.log(1.0 - p))))


def test_uts_geo_bruteforce_small_trees(oracle):
    """Exhaustive count against an independent hashlib-based recursion."""
    for seed, depth in [(0, 3), (1, 4), (19, 4), (42, 5), (7, 2)]:
        def walk(st, d):
            h = int.from_bytes(st[16:20], "big") & 0x7FFFFFFF
            nc = _py_children(h, 4.0, d, depth)
            if nc == 0:
                return 1, 1, d
            n, l, m = 1, 0, d
            for i in range(nc):
                c = hashlib.sha1(st + i.to_bytes(4, "big")).digest()
                a, b, e = walk(c, d + 1)
                n, l, m = n + a, l + b, max(m, e)
            return n, l, m
        root = hashlib.sha1(bytes(16) + seed.to_bytes(4, "big")).digest()
        want = walk(root, 0)
        n, lv, md = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        oracle.hso_uts_geo_count(seed, 4.0, depth, ctypes.byref(n), ctypes.byref(lv),
                                 ctypes.byref(md))
        assert (n.value, lv.value, md.value) == want


def test_uts_integer_thresholds_equal_float_formula(oracle):
    thr = (ctypes.c_uint32 * 128)()
    finite = oracle.hso_uts_geo_thresholds(4.0, thr, 128)
    assert 90 <= finite <= 100
    t = list(thr)[:finite]
    assert t == sorted(t) and t[0] > 0
    rng = random.Random(3)
    samples = [rng.randrange(1 << 31) for _ in range(20000)]
    for x in t:  # every boundary and its neighbours
        samples += [x - 1, x, x + 1]
    for h in samples:
        if not 0 <= h < (1 << 31):
            continue
        by_thr = sum(1 for x in t if h >= x)
        assert by_thr == _py_children(h, 4.0, 0, 10), h
    # closed form of the first threshold: 2^31 * (1 - 0.8)
    assert abs(t[0] - (1 << 31) * 0.2) <= 1


# ---- glibc rand (flags input) ----------------------------------------------

def _glibc_rand(seed, n):
    """The published glibc TYPE_3 additive-feedback generator (random_r.c)."""
    r = [0] * 34
    r[0] = seed if seed else 1
    for i in range(1, 31):
        hi, lo = divmod(r[i - 1], 127773)
        word = 16807 * lo - 2836 * hi
        if word < 0:
            word += 2147483647
        r[i] = word
    for i in range(31, 34):
        r[i] = r[i - 31]
    state = r[:]
    out = []
    k = 34
    while len(out) < n + 310:
        v = (state[k - 31] + state[k - 3]) & 0xFFFFFFFF
        state.append(v)
        k += 1
        if k - 1 >= 344:
            out.append(v >> 1)
    return out[:n]


@pytest.mark.parametrize("seed", [2018, 200])
def test_glibc_rand_matches_the_system_glibc(oracle, seed):
    """The oracle's generator and the Python restatement against the system
    glibc's own rand() (fixture): the flags input (srand(2018), flags.cu:92-101)
    and the UTS reference tree (srand(200), uts.cu:290)."""
    real = fixture("glibc_rand.json")[f"seed_{seed}"]
    assert _glibc_rand(seed, len(real)) == real
    arr = np.zeros(len(real), np.float32)
    oracle.hso_flags_rand(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), arr.size, seed)
    want = np.array(real, np.float32) / np.float32(4294967295.0)  # flags.cu:100
    assert np.array_equal(arr, want)


def test_flags_rand_is_glibc(oracle):
    n = 1000
    arr = np.zeros(n, np.float32)
    oracle.hso_flags_rand(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, 2018)
    want = np.array([np.float32(v) / np.float32(4294967295.0) for v in _glibc_rand(2018, n)],
                    np.float32)
    assert np.array_equal(arr, want)
    assert arr.max() <= 0.5  # rand() <= 2^31-1 -> values in [0, 0.5]


def test_flags_dirty_count_bruteforce(oracle):
    ntb, iters = 7, 13
    arr = np.zeros(ntb * iters, np.float32)
    oracle.hso_flags_rand(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), arr.size, 2018)
    want = sum(1 for b in range(ntb - 1) for it in range(iters) if arr[b + it * ntb] <= 0.33)
    got = oracle.hso_flags_dirty(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ntb,
                                 iters, ctypes.c_float(0.33))
    assert got == want


# ---- seqlock ------------------------------------------------------------------

@pytest.mark.parametrize("nw,gs,period,iters", [(512, 4, 16, 10), (1024, 16, 65, 10000),
                                                (33, 5, 3, 7)])
def test_seqlock_expected_bruteforce(oracle, nw, gs, period, iters):
    seq = np.zeros(gs, np.uint32)
    oracle.hso_seqlock_expected(nw, gs, period, iters,
                                seq.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), gs)
    want = np.zeros(gs, np.uint64)
    for b in range(nw):
        if b % 16 == 0 if period == 16 else b % period == 0:
            for _ in range(iters):
                want[b % gs] += 2
    assert np.array_equal(seq, want.astype(np.uint32))
    assert all(int(s) % 2 == 0 for s in seq)


# ---- syncPrims goldens -----------------------------------------------------

def test_prim_table_matches_reference(oracle):
    for name, pid in REF_NAMES.items():
        assert oracle.hso_prim_id(name.encode()) == pid
    assert oracle.hso_prim_id(b"noSuchPrim") == -1


def _simulate(prim, L, nw, iters, reps, wg, ncu, mode):
    """Brute force: replay the kernels' per-WG region walk, one pass at a time."""
    barrier = prim < 4 or 37 <= prim <= 42
    gsem = 8 <= prim <= 19 or 32 <= prim <= 35   # global semaphore kernels
    if barrier or 24 <= prim <= 35:
        mult = nw
    elif prim <= 7:
        mult = 1
    elif prim <= 19:
        mult = -(-nw // ncu)
    else:
        mult = min(nw, ncu)
    n = mult * wg * (L + 1)
    s = np.arange(n, dtype=np.int64)
    nab = min(nw, ncu)
    per = -(-nw // nab)
    lx = prim in (1, 3, 38, 40, 42) and per > 1

    def one_pass(region):
        for t in range(wg):
            tid = region * wg + t
            base = (tid // 16) * (L + 1)
            off = t % 16
            for k in range(L - 1, -1, -1):
                dst = (base + k + 1) * 16 + off
                if mode == 0:
                    s[dst] = s[(base + k) * 16 + off]
                else:
                    s[dst] += 1

    for _ in range(reps):
        for i in range(iters):
            if barrier:
                # exchange stride: the group count, or the grid for NaiveSRB
                xs = 0 if prim == 40 else nab
                for b in range(nw):  # episode i: WG b at region (b + shift)
                    cur = (b + i * ((xs + 1) if lx else 1)) % nw
                    one_pass(cur)
                    if lx:
                        one_pass((cur + xs) % nw)
            elif prim <= 7:
                for b in range(nw):
                    one_pass(0)
            elif gsem:
                for b in range(nw):
                    if b // nab == (b % nab) % per:
                        for j in range(per):
                            one_pass(j)
            elif prim <= 23:
                for b in range(nw):
                    one_pass(b % ncu)
            else:
                for b in range(nw):
                    if b // nab == (b % nab) % per:
                        cur = b
                        for j in range(per):
                            one_pass(cur)
                            cur = (cur + nab) % nw
    return s.astype(np.int32)


SHAPES = [(4, 4), (8, 4), (2, 4), (12, 4), (16, 8)]


@pytest.mark.parametrize("prim", [p for p in SUPPORTED if p not in PRIO_EBO])
def test_literal_golden_equals_closed_form(ohelp, prim):
    for nw, ncu in SHAPES:
        for L in (0, 1, 3):
            for mode in (0, 1):
                try:
                    closed = ohelp.golden(prim, L, nw, 3, 2, 64, ncu, mode)
                except AssertionError:
                    continue  # shape the primitive does not accept
                lit = ohelp.golden(prim, L, nw, 3, 2, 64, ncu, mode, literal=True)
                assert np.array_equal(lit, closed), (prim, nw, ncu, L, mode)


@pytest.mark.parametrize("prim", SUPPORTED)
def test_closed_form_equals_bruteforce_walk(ohelp, prim):
    for nw, ncu in [(4, 4), (8, 4), (12, 4), (3, 4)]:
        for mode in (0, 1):
            try:
                closed = ohelp.golden(prim, 2, nw, 2, 2, 64, ncu, mode)
            except AssertionError:
                continue
            sim = _simulate(prim, 2, nw, 2, 2, 64, ncu, mode)
            assert np.array_equal(sim, closed), (prim, nw, ncu, mode)


@pytest.mark.parametrize("prim", PRIO_EBO)
def test_prio_ebo_reference_golden_disagrees_with_its_kernel(ohelp, prim):
    """The reference's golden for 32-35 is the uniq one (every region once per
    iteration); its kernel is the global one (each group's writer walks
    regions 0..perCU-1).  The oracle follows the kernel: regions from perCU
    on keep their initial values."""
    for nw, ncu in [(4, 4), (8, 4), (16, 8)]:
        closed = ohelp.golden(prim, 2, nw, 3, 2, 64, ncu, 1)
        lit = ohelp.golden(prim, 2, nw, 3, 2, 64, ncu, 1, literal=True)
        assert not np.array_equal(lit, closed)
        per = -(-nw // min(nw, ncu))
        region = 64 * 3
        assert np.array_equal(closed[per * region:], np.arange(per * region, closed.size))
        assert np.array_equal(closed, _simulate(prim, 2, nw, 3, 2, 64, ncu, 1))


def test_planted_mistakes_are_caught(ohelp):
    """A wrong pass count or a dropped line changes the expected state."""
    good = ohelp.golden(7, 3, 4, 1, 1, 64, 4, 1)  # inc mode, k = 4
    lit = ohelp.golden(7, 3, 4, 1, 1, 64, 4, 1, literal=True)
    assert np.array_equal(good, lit)
    off_by_one = ohelp.golden(7, 3, 3, 1, 1, 64, 4, 1)  # one WG fewer
    assert not np.array_equal(good, off_by_one)
    dropped = good.copy()
    dropped[16 * 2 + 5] -= 1
    assert not np.array_equal(dropped, lit)
    # copy mode with k < L is sensitive to k as well
    a = ohelp.golden(0, 10, 4, 3, 1, 64, 4, 0)
    b = ohelp.golden(0, 10, 4, 4, 1, 64, 4, 0)
    assert not np.array_equal(a, b)


@pytest.mark.parametrize("prim", SUPPORTED)
def test_cpu_emulation_matches_closed_form(oracle, ohelp, prim):
    """A real concurrent run (std::thread + std::atomic) lands on the closed form."""
    for nw, ncu, L, mode in [(8, 4, 3, 1), (4, 4, 2, 0), (16, 8, 1, 1)]:
        try:
            want = ohelp.golden(prim, L, nw, 3, 2, 64, ncu, mode)
        except AssertionError:
            continue
        got = np.zeros_like(want)
        cs = ctypes.c_int64()
        t = oracle.hso_cpu_run(prim, L, nw, 3, 2, 64, ncu, mode, got.ctypes.data_as(P32),
                               got.size, ctypes.byref(cs))
        assert t >= 0
        assert np.array_equal(got, want), (prim, nw, ncu, L, mode)
        w, r, e = ohelp.counts(prim, nw, 3, 2, ncu)
        if e == 0:
            assert cs.value == w, (prim, cs.value, w)


def test_expected_counts(ohelp):
    assert ohelp.counts(7, 32, 100, 10, 256) == (32000, 0, 0)  # C1
    assert ohelp.counts(0, 1024, 1000, 10, 256) == (0, 0, 10000)  # C2
    w, r, _ = ohelp.counts(8, 512, 5, 10, 256)
    assert w == 256 * 50 and r == 256 * 50


# ---- reference UTS tree (cuda/uts/uvm/uts.cu:290-339) ----------------------

def test_uts_reference_tree_invariants(oracle):
    n = 8193  # `./uts 1 8192 90 32 128` -> num_nodes = N | 1
    val = np.zeros(n, np.uint32)
    ca = np.zeros(n, np.int32)
    cb = np.zeros(n, np.int32)
    lg, ls = ctypes.c_uint32(), ctypes.c_uint32()
    oracle.hso_uts_ref_tree(n, 90, val.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                            ca.ctypes.data_as(P32), cb.ctypes.data_as(P32),
                            ctypes.byref(lg), ctypes.byref(ls))
    internal = ca != 0
    assert np.array_equal(internal, cb != 0)
    assert np.all(cb[internal] == ca[internal] + 1)
    assert lg.value == (n + 1) // 2 == int((~internal).sum())
    assert ls.value == int(val[~internal].sum())
    # every node except the root is the child of exactly one node
    kids = np.concatenate([ca[internal], cb[internal]])
    assert sorted(kids.tolist()) == list(range(1, n))
    # deterministic (glibc srand(200))
    val2 = np.zeros(n, np.uint32)
    oracle.hso_uts_ref_tree(n, 90, val2.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                            ca.ctypes.data_as(P32), cb.ctypes.data_as(P32),
                            ctypes.byref(lg), ctypes.byref(ls))
    assert np.array_equal(val, val2)


@pytest.mark.parametrize("prim", [0, 1, 2, 3])
def test_cpu_barrier_pool_matches_closed_form(oracle, ohelp, prim):
    """The threaded CPU baseline of the barrier workload lands on the closed form."""
    for nw, ncu, threads, mode in [(16, 4, 3, 1), (8, 8, 4, 0), (32, 8, 5, 1)]:
        want = ohelp.golden(prim, 3, nw, 4, 2, 64, ncu, mode)
        got = np.zeros_like(want)
        t = oracle.hso_cpu_barrier_run(prim, 3, nw, 4, 2, 64, ncu, mode, threads,
                                       got.ctypes.data_as(P32), got.size)
        assert t >= 0 and np.array_equal(got, want), (prim, nw, ncu, threads, mode)


def test_cpu_emulation_under_thread_sanitizer():
    """The std::thread + std::atomic emulation of every primitive, and the
    threaded barrier pool, built with -fsanitize=thread: no data race is
    reported and every run lands on the closed form."""
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    subprocess.run(["make", "-s", "-C", here, "tsan"], check=True, timeout=300)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    p = subprocess.run([os.path.join(here, "tsan_emul")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert "ThreadSanitizer" not in p.stderr, p.stderr[-2000:]
    assert p.returncode == 0, p.stdout + p.stderr[-2000:]
    assert "0 mismatches" in p.stdout
