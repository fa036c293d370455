"""Pins the CPU oracle before anything is compared against it.

Each test names the oracle function it pins and what fixes the answer:
reference-held known-answer vectors, the reference's own Random123 code built
in oracle/_ref, closed forms, or exact enumeration.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
PROBE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "r123_probe")


def _probe(lines):
    out = subprocess.run([PROBE], input="\n".join(lines) + "\n", capture_output=True,
                         text=True, check=True).stdout.split("\n")
    return [ln for ln in out if ln]


needs_probe = pytest.mark.skipif(not os.path.exists(PROBE),
                                 reason="oracle/_ref not built (reference absent)")


# ---- R1 Philox4x32-R -------------------------------------------------------
def test_philox_known_answers(orc):
    kat = json.load(open(os.path.join(GOLD, "philox4x32_kat.json")))["vectors"]
    assert len(kat) == 110
    for v in kat:
        assert orc.philox4x32(v["rounds"], v["ctr"], v["key"]) == v["expected"], v


@needs_probe
def test_philox_matches_reference_code_on_engine_keys(orc):
    """The exact key/counter shapes the reference uses, through its own Philox."""
    rng = np.random.default_rng(7)
    qs = []
    for _ in range(300):
        w = int(rng.integers(0, 1 << 20))
        seed = int(rng.integers(0, 1 << 32))
        it = int(rng.integers(0, 1 << 16))
        c0 = int(rng.integers(0, 1 << 32))
        for key in ([2 * w, 0xdecafbad], [(2 * w + 50000) & 0xFFFFFFFF, 0x11cafead],
                    [w, 0xdecafbad], [w + 40000, 0xdecafead]):
            for R in (7, 10):
                qs.append((R, [c0, seed, it, 0xBADC0DED], key))
    lines = ["P %d %x %x %x %x %x %x" % (R, *c, *k) for R, c, k in qs]
    ref = _probe(lines)
    for (R, c, k), line in zip(qs, ref):
        assert orc.philox4x32(R, c, k) == [int(t, 16) for t in line.split()]


# ---- R3 u01fixedpt<float> ----------------------------------------------------
def test_u01fixedpt_known_answers(orc):
    kat = json.load(open(os.path.join(GOLD, "u01fixedpt_f32_kat.json")))["vectors"]
    assert len(kat) == 144
    for v in kat:
        assert orc.u01fixedpt_f32(v["r"]) == np.float32(float(v["expected"])), v


@needs_probe
def test_u01fixedpt_matches_reference_code(orc):
    rng = np.random.default_rng(3)
    rs = [int(x) for x in rng.integers(0, 1 << 32, size=2000)] + [0, 1, 255, 256,
                                                                  0x7FFFFFFF, 0x80000000,
                                                                  0xFFFFFFFF]
    ref = _probe(["U %x" % r for r in rs])
    for r, line in zip(rs, ref):
        got = np.float32(orc.u01fixedpt_f32(r)).view(np.uint32)
        assert int(got) == int(line, 16)


def test_half_comparison_is_sign_bit():
    """u01fixedpt<float>(r) < 0.5  <=>  r < 2^31, over every value of r >> 8.

    The engine proposes the spin step from the top bit of r3; the reference
    writes the comparison (muca_delta_multigpu.cu:800)."""
    x = np.arange(1 << 24, dtype=np.uint64)
    u = ((x | 1).astype(np.float64) * 2.0 ** -24).astype(np.float32)
    lhs = u.astype(np.float64) < 0.5
    rhs = (x << 8) < (1 << 31)
    assert np.array_equal(lhs, rhs)


# ---- R4 acceptance log -------------------------------------------------------
def test_acceptance_log_is_correctly_rounded_everywhere(orc):
    """All 2^23 distinct u01fixedpt<float> values against 113-bit logq."""
    assert orc.logf_u01_exhaustive(os.cpu_count() or 4) == 0


def test_acceptance_log_special_values(orc):
    assert orc.logf_u01(0) == np.float32(-24 * np.log(2.0))
    assert orc.logf_u01(0xFFFFFFFF) == np.float32(np.log1p(-2.0 ** -24))
    assert orc.logf_u01(0x80000000) == np.float32(np.log(0.5 + 2.0 ** -24))


# ---- R6 tables -----------------------------------------------------------
@pytest.mark.parametrize("L", [2, 4, 6, 16, 72])
def test_tables_closed_form(orc, L):
    naver, order = orc.tables(L)
    N = L * L
    i = np.arange(N)
    x, y = i % L, i // L
    nv = naver.reshape(N, 4)
    assert np.array_equal(nv[:, 0], y * L + (x + 1) % L)
    assert np.array_equal(nv[:, 1], y * L + (x - 1) % L)
    assert np.array_equal(nv[:, 2], ((y + 1) % L) * L + x)
    assert np.array_equal(nv[:, 3], ((y - 1) % L) * L + x)
    # even L: all even linear indices in increasing order, then all odd ones
    assert np.array_equal(order, np.concatenate([np.arange(0, N, 2), np.arange(1, N, 2)]))


def test_tables_odd_L_quirk(orc):
    """Odd L: the reference formula revisits sites (muca_delta_multigpu.cu:173)."""
    _, order = orc.tables(3)
    assert list(order) == [0, 2, 4, 6, 0, 2, 4, 6, 7]


# ---- R5 lattice init -------------------------------------------------------
@needs_probe
@pytest.mark.parametrize("keying", [0, 1])
def test_init_lattice_keying(orc, keying):
    L, seed, mm, gpu = 8, 1234, 3, 1
    lat = orc.init_lattice(keying, L, 5, mm=mm, gpu=gpu, seed=seed, w0=17)
    lines, want = [], []
    for j in range(5):
        w = 17 + j
        for blk in range(L * L // 4):
            if keying == 0:
                c = [gpu * 2 + blk + 1, seed, 0xBADCAB1E, 0xBADC0DED]
                k = [10 * w + 2 * mm, 0xdecafbad]
            else:
                c = [gpu * 1000, seed, blk + 1, 0xBADCAFE3]
                k = [10 * w, 0xdecafbad]
            lines.append("P 7 %x %x %x %x %x %x" % (*c, *k))
    words = [int(t, 16) for ln in _probe(lines) for t in ln.split()]
    u = (((np.array(words, np.uint64) >> 8) | 1).astype(np.float64) * 2.0 ** -24)
    spins = (u < 1 / 3.0).astype(int) + (u < 2 / 3.0).astype(int) - 1
    assert np.array_equal(lat.reshape(-1), spins)


def test_init_lattice_is_uniform(orc):
    lat = orc.init_lattice(0, 16, 64, seed=100)
    frac = np.bincount(lat.reshape(-1) + 1, minlength=3) / lat.size
    assert np.all(np.abs(frac - 1 / 3) < 0.02)


def test_stratified_start(orc):
    """strat_W = 0 is the reference start; otherwise walker w's spins follow
    the closed form of the same uniforms, so D/N tracks (w + 1/2)/W and the
    signs are balanced."""
    L, W = 16, 32
    a = orc.init_lattice(0, L, 4, seed=9, w0=3)
    b = orc.init_lattice(0, L, 4, seed=9, w0=3, strat_W=0)
    assert np.array_equal(a, b)
    lat = orc.init_lattice(0, L, W, seed=9, strat_W=W)
    ref = orc.init_lattice(0, L, W, seed=9)          # same stream: recover u's bins
    N = L * L
    D = (lat != 0).sum(1) / N
    p = (np.arange(W) + 0.5) / W
    assert np.all(np.abs(D - p) < 4 * np.sqrt(p * (1 - p) / N) + 1e-9)
    # the reference start has -1 exactly where u >= 2/3, and +1 where u < 1/3:
    # there the stratified spin is 0 when p < 2/3, and -1 when p/2 > 1/3
    for w in range(W):
        if p[w] < 2 / 3:
            assert np.all(lat[w][ref[w] == -1] == 0)
        if p[w] / 2 > 1 / 3:
            assert np.all(lat[w][ref[w] == 1] == -1)
    assert abs((lat == 1).sum() - (lat == -1).sum()) < 0.05 * (lat != 0).sum()


# ---- R7 state ----------------------------------------------------------------
def test_state_brute_force(orc):
    L = 12
    lat = orc.init_lattice(0, L, 8, seed=5)
    E, D, M = orc.state(L, lat)
    s = lat.reshape(8, L, L).astype(np.int64)
    Ej = -(s * np.roll(s, -1, 2)).sum((1, 2)) - (s * np.roll(s, -1, 1)).sum((1, 2))
    assert np.array_equal(E, Ej)
    assert np.array_equal(D, (s * s).sum((1, 2)))
    assert np.array_equal(M, s.sum((1, 2)))


# ---- R8 / R18 chain ----------------------------------------------------------
@pytest.mark.parametrize("fp64,keying", [(False, 0), (True, 1)])
def test_chain_invariants(orc, fp64, keying):
    """E, D, M carried incrementally equal the recomputed ones; histogram
    totals equal the attempt count; results do not depend on threading."""
    L, W = 8, 12
    N = L * L
    lat = orc.init_lattice(keying, L, W, seed=11)
    E, D, M = orc.state(L, lat)
    lnw = -0.3 * np.arange(N + 1) + np.sin(np.arange(N + 1))
    kw = dict(keying=keying, fp64=fp64, lnw32=lnw.astype(np.float32), lnw64=lnw, iter=2,
              seed=77, therm=N, nupdates=3 * N + 1)
    lat2, E2, D2, M2 = lat.copy(), E.copy(), D.copy(), M.copy()
    h1 = orc.mcs(L, lat, E, D, M, nthreads=1, **kw)
    h2 = orc.mcs(L, lat2, E2, D2, M2, nthreads=3, **kw)
    assert np.array_equal(lat, lat2) and np.array_equal(h1, h2)
    assert np.array_equal(E, E2) and np.array_equal(D, D2) and np.array_equal(M, M2)
    Er, Dr, Mr = orc.state(L, lat)
    assert np.array_equal(E, Er) and np.array_equal(D, Dr) and np.array_equal(M, Mr)
    assert int(h1.sum()) == W * (3 * N + 1)
    assert set(np.unique(lat)) <= {-1, 0, 1}


@pytest.mark.slow
@pytest.mark.parametrize("fp64,keying,beta", [(False, 0, 2.0), (True, 1, 2.0),
                                              (False, 0, 0.7)])
def test_chain_stationary_law_exact_enumeration(orc, fp64, keying, beta):
    """At L=4 the chain's long-run D (and D,M) histogram must follow
    exp(lnW(D)) * g_beta(D, M) from exact enumeration of all 3^16 states.
    lnW = -ln g_beta(D) plus a known tilt, so a wrong sign or index in the
    acceptance (weights, dE, dDel) shows up as an order-one deviation."""
    import exact
    L, W = 4, 64
    N = L * L
    g, _ = exact.density_of_states(L, beta)
    gD = g.sum(1)
    tilt = 0.15 * np.cos(np.arange(N + 1))
    lnw = -np.log(gD) + tilt
    lnw -= lnw[0]
    lat = orc.init_lattice(keying, L, W, seed=3)
    E, D, M = orc.state(L, lat)
    joint = np.zeros((N + 1) * (N + 2) // 2, np.uint64) if keying == 1 else None
    kw = dict(keying=keying, fp64=fp64, lnw32=lnw.astype(np.float32), lnw64=lnw, beta=beta,
              seed=21)
    orc.mcs(L, lat, E, D, M, iter=0, therm=200 * N, nupdates=0, nthreads=4, **kw)
    nthreads = 1 if joint is not None else 8
    h = orc.mcs(L, lat, E, D, M, iter=1, therm=0, nupdates=40000 * N, nthreads=nthreads,
                joint=joint, **kw)
    p_emp = h / h.sum()
    p_ex = np.exp(lnw) * gD
    p_ex /= p_ex.sum()
    assert np.max(np.abs(p_emp - p_ex) / p_ex) < 0.05, (p_emp, p_ex)
    if joint is not None:
        # joint law at fixed D: proportional to g(D, M)
        for d in (6, 10, 14):
            ms = np.arange(-d, d + 1, 2)
            hj = np.array([joint[(m + d * d + 2 * d) // 2] for m in ms], np.float64)
            gj = g[d, ms + N]
            keep = gj / gj.sum() > 0.02
            emp = hj / hj.sum()
            ex = gj / gj.sum()
            assert np.max(np.abs(emp[keep] - ex[keep]) / ex[keep]) < 0.1


# ---- R10-R14 host algorithms --------------------------------------------------
def test_hist_range_closed_form(orc):
    assert orc.hist_range(np.array([0, 1, 2, 0, 5, 1], np.uint64)) == (2, 4)
    assert orc.hist_range(np.array([1, 1, 0, 1], np.uint64)) == (0, 0)
    assert orc.hist_range(np.array([7, 0, 0, 9], np.uint64)) == (0, 3)


def test_kull_leibler_closed_form(orc):
    assert orc.kull_leibler(np.array([5, 5, 0, 5], np.uint64)) == 0.0
    v = orc.kull_leibler(np.array([1, 3, 0], np.uint64))
    assert v == pytest.approx(0.25 * np.log(0.5) + 0.75 * np.log(1.5), rel=1e-15)


def test_trivial_update_closed_form(orc):
    h = np.array([4, 1, 0, 9, 2], np.uint64)
    lnw = np.array([0.5, 1.0, 2.0, 3.0, 4.0])
    orc.update_trivial(h, lnw)
    assert np.allclose(lnw, [0.5 - np.log(4), 1.0, 2.0, 3.0 - np.log(9), 4.0 - np.log(2)],
                       rtol=0, atol=1e-15)
    f = orc.normalise(lnw)
    assert lnw[0] == 0.0 and f.dtype == np.float32


def test_recursion_two_bins_by_hand(orc):
    h = np.array([6, 3], np.uint64)
    lnw = np.array([1.0, -2.0])
    p_acc = np.array([0.5, 0.0])
    orc.weight_recursion(h, p_acc, 0, 1, lnw)
    p = 18 / 9
    k = p / (p + 0.5)
    assert lnw[1] == pytest.approx(1.0 + (-1.0 - 2.0 + k * np.log(2.0)), rel=1e-15)
    assert p_acc[0] == pytest.approx(2.5)


def test_recursion_first_pass_equals_trivial_update(orc):
    """With p_acc = 0 and all bins of [start, end] visited, kappa = 1 and the
    recursion reproduces the trivial update up to the normalisation."""
    rng = np.random.default_rng(1)
    N = 40
    h = rng.integers(2, 1000, size=N + 1).astype(np.uint64)
    lnw0 = rng.normal(size=N + 1)
    a, b = lnw0.copy(), lnw0.copy()
    orc.weight_recursion(h, np.zeros(N + 1), 0, N, a)
    orc.update_trivial(h, b)
    orc.normalise(a)
    orc.normalise(b)
    assert np.allclose(a, b, rtol=0, atol=1e-11)


def test_jackknife_of_mean_is_standard_error(orc):
    x = np.array([1.0, 4.0, 2.5, 7.0, 3.0])
    avg, err = orc.jackknife(x)
    assert avg == pytest.approx(x.mean())
    assert err == pytest.approx(x.std(ddof=1) / np.sqrt(len(x)), rel=1e-12)
    a1, e1 = orc.jackknife(np.array([2.0]))   # reference divides by J-1 = 0
    assert a1 == 2.0 and np.isnan(e1)


def test_weight_shift_and_boltz(orc):
    N = 30
    lnw = np.linspace(0, -45, N + 1) + np.cos(np.arange(N + 1))
    beta, delta = 2.0, 1.98
    orc.weight_shift(beta, delta, lnw)
    v = -lnw - beta * delta * np.arange(N + 1)
    assert abs(v.max()) < 1e-12
    b = orc.boltz(beta, delta, lnw)
    assert b.max() == pytest.approx(1.0, rel=1e-12)
    assert np.allclose(b, np.exp(v), rtol=1e-12, atol=0)


def test_observables_closed_form(orc):
    """Single D shell with counts at chosen m: <|M|> and chi by hand."""
    N = 16
    joint = np.zeros((N + 1) * (N + 2) // 2, np.uint64)
    d = 4
    for m, c in ((-4, 3), (-2, 1), (0, 5), (2, 2), (4, 1)):
        joint[(m + d * d + 2 * d) // 2] = c
    boltz = np.zeros(N + 1)
    boltz[d] = 0.25
    absM, chi = orc.observables(2.0, N, joint, boltz)
    ms = np.array([-4, -2, 0, 2, 4])
    cs = np.array([3, 1, 5, 2, 1])
    m1 = (np.abs(ms) * cs).sum() / cs.sum()
    m2 = (ms * ms * cs).sum() / cs.sum()
    assert absM == pytest.approx(m1, rel=1e-15)
    assert chi == pytest.approx(2.0 * (m2 - m1 * m1) / N, rel=1e-14)
    S0, S1, S2, S4 = orc.joint_moments(N, joint)
    assert S0[d] == cs.sum() and S1[d] == (np.abs(ms) * cs).sum()
    assert S2[d] == (ms ** 2 * cs).sum() and S4[d] == (ms ** 4 * cs).sum()


def test_moments_exact_closed_form(orc):
    """joint_moments_exact on the hand-made shell of the test above."""
    N, d = 16, 4
    joint = np.zeros((N + 1) * (N + 2) // 2, np.uint64)
    for m, c in ((-4, 3), (-2, 1), (0, 5), (2, 2), (4, 1)):
        joint[(m + d * d + 2 * d) // 2] = c
    S0, S1, S2, S4 = orc.joint_moments_exact(N, joint)
    assert (S0[d], S1[d], S2[d], S4[d]) == (12, 22, 76, 1072)
    assert sum(S0) == 12 and sum(S4) == 1072


def test_moment_add_128_bit_words(orc):
    """orc_moment_add at |M| = 2^16 (M^4 = 2^64 exactly: high word only) and
    at 2^16 - 1 three times (the low word of sum M^4 wraps twice)."""
    nb = 3
    mom = np.zeros(6 * nb, np.uint64)
    p = mom.ctypes.data_as(orc._u64p)
    orc.lib().orc_moment_add(p, nb, 2, -65536)
    S0, S1, S2, S4 = orc.moments_as_int(mom, nb - 1)
    assert (S0[2], S1[2], S2[2], S4[2]) == (1, 65536, 2 ** 32, 2 ** 64)
    assert mom[4 * nb + 2] == 0 and mom[5 * nb + 2] == 1
    for _ in range(3):
        orc.lib().orc_moment_add(p, nb, 1, 65535)
    S0, S1, S2, S4 = orc.moments_as_int(mom, nb - 1)
    assert (S0[1], S1[1], S2[1], S4[1]) == (3, 3 * 65535, 3 * 65535 ** 2, 3 * 65535 ** 4)
    assert int(mom[5 * nb + 1]) == (3 * 65535 ** 4) >> 64 == 2


def test_oracle_moments_equal_joint_moments(orc):
    """The chain's per-D moments equal the moments of its own joint
    histogram, exactly; the long-double observables agree both ways."""
    L, W = 6, 12
    N = L * L
    lat = orc.init_lattice(orc.KEYING_PRODUCE, L, W, seed=7)
    E, D, Mg = orc.state(L, lat)
    lnw = -np.log(np.arange(1, N + 2, dtype=np.float64)) * 2.0
    joint = np.zeros((N + 1) * (N + 2) // 2, np.uint64)
    mom = np.zeros(6 * (N + 1), np.uint64)
    orc.mcs(L, lat, E, D, Mg, keying=orc.KEYING_PRODUCE, fp64=True, lnw64=lnw, iter=1,
            seed=3, therm=N, nupdates=40 * N, joint=joint, moments=mom)
    S = orc.moments_as_int(mom, N)
    X = orc.joint_moments_exact(N, joint)
    for k in range(4):
        assert list(S[k]) == list(X[k]), k
    assert sum(S[0]) == W * 40 * N
    w = lnw.copy()
    orc.weight_shift(2.0, 1.95, w)
    bz = orc.boltz(2.0, 1.95, w)
    a0, c0 = orc.observables(2.0, N, joint, bz)
    a1, c1 = orc.observables_moments(2.0, N, S[0], S[1], S[2], bz)
    assert a1 == pytest.approx(a0, rel=1e-14) and c1 == pytest.approx(c0, rel=1e-11)


@pytest.mark.slow
def test_weight_loop_converges_to_exact_lnw(orc):
    """L=4: the multigpu_weight loop (trivial update) converges and its lnW
    matches -ln g_2(D) from exact enumeration within the statistical noise."""
    import exact
    L, N = 4, 16
    g, _ = exact.density_of_states(L, 2.0)
    exact_lnw = -np.log(g.sum(1))
    exact_lnw -= exact_lnw[0]
    recs = orc.weight_iteration_reference(L, 256, 2000, seed=1234, nthreads=8)
    last = recs[-1]
    assert last["converged"], (len(recs), last["width"], last["dkl"])
    assert np.max(np.abs(last["lnw"] - exact_lnw)) < 0.15
