"""GPU parity: libmuca's HIP path against the CPU oracle, bit for bit.

Lattices, E, D, M and every integer histogram must be identical for the same
seed, walker ids, lattice and weights.  Oracle lattices are walker-major
[walker][site]; the engine's are site-major [site][walker] (the reference's
interleaving), hence the transposes.
"""
import os

import numpy as np
import pytest

import muca_amd as M

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 4)


def oracle_init(orc, keying, L, W, mm=0, gpu=0, seed=M.SEED, rounds=7):
    lat = orc.init_lattice(keying, L, W, mm=mm, gpu=gpu, seed=seed, rounds=rounds)
    E, D, Mg = orc.state(L, lat)
    return lat, E, D, Mg


def test_acclog_device_exhaustive(orc):
    n, step = 1 << 23, 1 << 22
    for k0 in range(0, n, step):
        a = M.acclog_device(k0, step)
        b = orc.logf_u01_range(k0, step)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k0


def test_acclog_fast_band_exhaustive(orc):
    """The update kernels screen each acceptance with la = v_log_f32(u) * ln 2
    and fall back to the exact log only when diff lies within
    band = |la| * 2^-20 + 2^-36 of la.  That is exact iff |la - log(u)| <
    band for every input; require half the band on all 2^23 inputs."""
    n, step = 1 << 23, 1 << 22
    worst = 0.0
    for k0 in range(0, n, step):
        la = M.acclog_fast_device(k0, step)
        ex = orc.logf_u01_range(k0, step)
        band = (np.abs(la) * np.float32(M.ACCLOG_FAST_REL) +
                np.float32(M.ACCLOG_FAST_ABS)).astype(np.float64)
        err = np.abs(la.astype(np.float64) - ex.astype(np.float64))
        worst = max(worst, float(np.max(err / band)))
    print(f"worst |la - log u| / band = {worst:.4g}")
    assert worst <= 0.5


@pytest.mark.parametrize("keying", [M.KEYING_WEIGHT, M.KEYING_PRODUCE])
@pytest.mark.parametrize("L,W,mm,gpu,seed,rounds", [(16, 16, 0, 0, 1234, 7),
                                                     (10, 300, 2, 1, 100, 7),
                                                     (5, 7, 1, 0, 9, 10)])
def test_init_lattice_and_state(orc, keying, L, W, mm, gpu, seed, rounds):
    with M.Context(L, W, keying=keying, rounds=rounds) as ctx:
        ctx.init_lattice(mm=mm, gpu=gpu, seed=seed)
        ctx.compute_state()
        lat = ctx.get_lattice()
        E, D, Mg = ctx.get_state()
    ref, rE, rD, rM = oracle_init(orc, keying, L, W, mm, gpu, seed, rounds)
    assert np.array_equal(lat, ref.T)
    assert np.array_equal(E, rE) and np.array_equal(D, rD) and np.array_equal(Mg, rM)


@pytest.mark.parametrize("keying", [M.KEYING_WEIGHT, M.KEYING_PRODUCE])
def test_stratified_start_matches_oracle(orc, keying):
    L, W = 12, 77
    with M.Context(L, W, keying=keying) as ctx:
        ctx.init_lattice(mm=1, gpu=0, seed=31, stratified=True)
        lat = ctx.get_lattice()
    ref = orc.init_lattice(keying, L, W, mm=1, gpu=0, seed=31, strat_W=W)
    assert np.array_equal(lat, ref.T)


CASES = [
    # L, W, precision, keying, beta, rounds, therm, nup
    (16, 16, M.REF32, M.KEYING_WEIGHT, 2.0, 7, 256, 1),
    (16, 200, M.REF32, M.KEYING_WEIGHT, 2.0, 7, 256, 1027),
    (12, 64, M.FP64, M.KEYING_PRODUCE, 2.0, 7, 144 * 3, 144 * 5 + 3),
    (8, 33, M.REF32, M.KEYING_WEIGHT, 0.7, 10, 5, 299),
    (8, 65, M.FP64, M.KEYING_WEIGHT, 1.3, 7, 0, 250),
    (5, 20, M.REF32, M.KEYING_WEIGHT, 2.0, 7, 25, 61),       # odd L (reference order quirk)
    (2, 9, M.FP64, M.KEYING_PRODUCE, 2.0, 7, 4, 33),
    (24, 513, M.REF32, M.KEYING_PRODUCE, 2.0, 7, 576, 2),
]


KERNELS = pytest.mark.parametrize("kernel", [M.KERNEL_GENERIC, M.KERNEL_TEAM, M.KERNEL_LANE],
                                  ids=["generic", "team", "lane"])


@KERNELS
@pytest.mark.parametrize("case", CASES, ids=lambda c: "L%d_W%d_p%d_k%d_b%g_r%d_t%d_n%d" % c)
def test_single_launch_parity(orc, case, kernel):
    L, W, prec, keying, beta, rounds, therm, nup = case
    N = L * L
    rng = np.random.default_rng(L * 1000 + W)
    lnw = np.cumsum(rng.normal(scale=1.5, size=N + 1))
    lnw -= lnw[0]
    lat, E, D, Mg = oracle_init(orc, keying, L, W, seed=77, rounds=rounds)
    with M.Context(L, W, keying=keying, precision=prec, beta=beta, rounds=rounds,
                   kernel=kernel) as ctx:
        ctx.set_lattice(lat.T)
        ctx.set_state(E, D, Mg)
        ctx.set_log_weight(lnw)
        hist, _ = ctx.run(3, 1234, therm, nup)
        glat = ctx.get_lattice()
        gE, gD, _ = ctx.get_state()
    h = orc.mcs(L, lat, E, D, Mg, keying=keying, fp64=prec == M.FP64,
                lnw32=lnw.astype(np.float32), lnw64=lnw, iter=3, seed=1234, therm=therm,
                nupdates=nup, beta=beta, rounds=rounds, nthreads=NTHREADS)
    assert np.array_equal(glat, lat.T)
    assert np.array_equal(gE, E) and np.array_equal(gD, D)
    assert np.array_equal(hist, h)
    assert int(hist.sum()) == W * nup


def test_autotune_keeps_state_and_results():
    """muca_autotune times the candidates on a scratch copy: the walkers are
    untouched and later launches equal those of a fixed-kernel context."""
    L, W = 16, 2048
    N = L * L
    lnw = -0.4 * np.arange(N + 1) + np.sin(np.arange(N + 1))
    with M.Context(L, W) as a, M.Context(L, W, kernel=M.KERNEL_GENERIC) as b:
        for c in (a, b):
            c.init_lattice(seed=3)
            c.compute_state()
            c.set_log_weight(lnw)
        lat0, st0 = a.get_lattice(), a.get_state()
        k, blk, ms = a.autotune(2 * N)
        assert k in (M.KERNEL_GENERIC, M.KERNEL_TEAM, M.KERNEL_LANE) and ms > 0
        assert np.array_equal(a.get_lattice(), lat0)
        assert all(np.array_equal(x, y) for x, y in zip(a.get_state(), st0))
        ha, _ = a.run(5, 77, N, 3 * N + 1)
        hb, _ = b.run(5, 77, N, 3 * N + 1)
        assert np.array_equal(ha, hb)
        assert np.array_equal(a.get_lattice(), b.get_lattice())


def test_zero_attempts_leaves_state(orc):
    L, W = 8, 40
    with M.Context(L, W) as ctx:
        ctx.init_lattice(seed=5)
        ctx.compute_state()
        ctx.set_log_weight(np.zeros(L * L + 1))
        before = ctx.get_lattice()
        hist, _ = ctx.run(0, 100, 0, 0)
        assert np.array_equal(before, ctx.get_lattice())
        assert int(hist.sum()) == 0


def _weight_loop_parity(orc, L, W, iters, seed, recursion=False, kernel=M.KERNEL_AUTO):
    recs = orc.weight_iteration_reference(L, W, iters, seed=seed, recursion=recursion,
                                          nthreads=NTHREADS)
    with M.Context(L, W, kernel=kernel) as ctx:
        ctx.init_lattice(mm=0, gpu=0, seed=seed)
        ctx.compute_state()
        got = []
        M.weight_iteration(ctx, iters, seed=seed, recursion=recursion,
                           callback=lambda r: got.append(
                               dict(r, lat=ctx.get_lattice(), ED=ctx.get_state())))
    assert len(got) == len(recs)
    for g, r in zip(got, recs):
        assert np.array_equal(g["hist"], r["hist"]), g["k"]
        assert np.array_equal(g["lat"], r["lat"].T), g["k"]
        assert np.array_equal(g["ED"][0], r["E"]) and np.array_equal(g["ED"][1], r["D"])
        assert g["width"] == r["width"] and g["nup"] == r["nup"] and g["dkl"] == r["dkl"]
        assert np.array_equal(g["lnw"], r["lnw"]), g["k"]


@KERNELS
def test_weight_loop_L16_16walkers_10iters(orc, kernel):
    """Plain method, small: L=16, 16 walkers, seed 1234, 10 trivial iterations."""
    _weight_loop_parity(orc, 16, 16, 10, 1234, kernel=kernel)


@KERNELS
def test_weight_loop_recursive_L16(orc, kernel):
    _weight_loop_parity(orc, 16, 64, 12, 1234, recursion=True, kernel=kernel)


@KERNELS
def test_weight_loop_L72_4096walkers_20iters(orc, kernel):
    """Plain method at reference scale: L=72, 4096 walkers, 20 iterations."""
    _weight_loop_parity(orc, 72, 4096, 20, 1234, kernel=kernel)


@KERNELS
def test_produce_joint_histograms_per_block(orc, kernel):
    L, W, B = 8, 96, 3
    N = L * L
    rng = np.random.default_rng(5)
    lnw = np.cumsum(rng.normal(size=N + 1))
    lat, E, D, Mg = oracle_init(orc, M.KEYING_PRODUCE, L, W, seed=100)
    with M.Context(L, W, precision=M.FP64, keying=M.KEYING_PRODUCE, joint_blocks=B,
                   kernel=kernel) as ctx:
        ctx.set_lattice(lat.T)
        ctx.set_state(E, D, Mg)
        ctx.set_log_weight(lnw)
        ctx.run(0, 101, 2 * N, 0)          # therm-only launch: no joint accumulation
        ctx.run(1, 200, 0, 7 * N + 1)
        joints = [ctx.get_joint(b) for b in range(B)]
        _, _, gM = ctx.get_state()
    kw = dict(keying=M.KEYING_PRODUCE, fp64=True, lnw64=lnw, nthreads=1)
    orc.mcs(L, lat, E, D, Mg, iter=0, seed=101, therm=2 * N, nupdates=0, **kw)
    bw = W // B
    for b in range(B):
        j = np.zeros((N + 1) * (N + 2) // 2, np.uint64)
        sub = np.ascontiguousarray(lat[b * bw:(b + 1) * bw])
        e, d, m = E[b * bw:(b + 1) * bw].copy(), D[b * bw:(b + 1) * bw].copy(), \
            Mg[b * bw:(b + 1) * bw].copy()
        orc.mcs(L, sub, e, d, m, iter=1, seed=200, therm=0, nupdates=7 * N + 1, w0=b * bw,
                joint=j, **kw)
        assert np.array_equal(joints[b], j), b
        assert np.array_equal(gM[b * bw:(b + 1) * bw], m)


@pytest.mark.parametrize("kernel", [M.KERNEL_GENERIC, M.KERNEL_TEAM], ids=["generic", "team"])
@pytest.mark.parametrize("L,W,B,start,nup_mult", [
    (8, 96, 3, "random", 7),    # many D changes
    (72, 48, 3, "random", 3),   # three rounds per half-row, the last one partial
    (64, 128, 2, "ones", 10),   # sum M^4 per D passes 2^64 (high words)
    (256, 64, 1, "ones", 0),    # |M| = 2^16 at the start: M^4 = 2^64 exactly
    (256, 32, 2, "random", 0),  # rough ln W read from HBM (team)
    (64, 32, 2, "zeros", 2),    # D climbs one bin per step from 0: 16 bins per round
], ids=["L8-random", "L72-random", "L64-ordered", "L256-ordered", "L256-random", "L64-climb"])
def test_produce_moments_per_block(orc, L, W, B, start, nup_mult, kernel):
    """Per-D moments accumulated by the update kernel (OBS_MOMENTS) in place
    of the joint histogram: identical to the oracle chain's exact 128-bit
    sums, block by block, with the final M.  The generic kernel keeps one
    walker's run per lane; the team kernel adds each run of attempts at one
    (D, M) to its LDS window of D slots, or straight to HBM when D moved
    beyond the window within 16 steps (the climb case)."""
    N = L * L
    nup = nup_mult * N + 1 if nup_mult else 2000
    rng = np.random.default_rng(L)
    if start == "random":
        lnw = np.cumsum(rng.normal(size=N + 1))
        lat, E, D, Mg = oracle_init(orc, M.KEYING_PRODUCE, L, W, seed=100)
    else:
        # "zeros": every proposal on a zero spin raises D, and a steep ln W
        # accepts it
        lnw = np.zeros(N + 1) if start == "ones" else 40.0 * np.arange(N + 1)
        lat = np.full((W, N), 1 if start == "ones" else 0, np.int8)
        E, D, Mg = orc.state(L, lat)
    with M.Context(L, W, precision=M.FP64, keying=M.KEYING_PRODUCE, joint_blocks=B,
                   observables=M.OBS_MOMENTS, kernel=kernel) as ctx:
        assert ctx.info.kernel == kernel and ctx.info.joint_bins == 0
        ctx.set_lattice(lat.T)
        ctx.set_state(E, D, Mg)
        ctx.set_log_weight(lnw)
        ghist, _ = ctx.run(1, 200, 0, nup)
        moms = [ctx.get_moments(b) for b in range(B)]
        _, _, gM = ctx.get_state()
    kw = dict(keying=M.KEYING_PRODUCE, fp64=True, lnw64=lnw, nthreads=1)
    bw = W // B
    hi = 0
    ohist = np.zeros(N + 1, np.uint64)
    for b in range(B):
        mom = np.zeros(6 * (N + 1), np.uint64)
        sub = np.ascontiguousarray(lat[b * bw:(b + 1) * bw])
        e, d, m = E[b * bw:(b + 1) * bw].copy(), D[b * bw:(b + 1) * bw].copy(), \
            Mg[b * bw:(b + 1) * bw].copy()
        ohist += orc.mcs(L, sub, e, d, m, iter=1, seed=200, therm=0, nupdates=nup, w0=b * bw,
                         moments=mom, **kw)
        want = orc.moments_as_int(mom, N)
        for k in range(4):
            assert list(moms[b][k]) == list(want[k]), (b, k)
        assert sum(moms[b][0]) == bw * nup
        assert np.array_equal(gM[b * bw:(b + 1) * bw], m)
        hi = max(hi, max(int(x) >> 64 for x in want[3]))
    # the launch's D histogram (the blocks' S0 summed on the device)
    assert np.array_equal(ghist, ohist)
    if start == "ones":
        assert hi > 0   # the 128-bit carry path was taken


def test_reweighting_from_moments_matches_joint(orc):
    """The same production run accumulated both ways: device reweighting of
    the per-D moments equals that of the joint histograms, and the oracle's
    long-double sums over the exact moments."""
    L, W, B = 8, 128, 4
    N = L * L
    lnw = -np.log(np.arange(1, N + 2, dtype=np.float64)) * 3.0
    lnw -= lnw[0]
    deltas = np.array([1.9, 1.95, 1.98, 2.0])
    out = {}
    for obs in (M.OBS_JOINT, M.OBS_MOMENTS):
        with M.Context(L, W, precision=M.FP64, keying=M.KEYING_PRODUCE, joint_blocks=B,
                       kernel=M.KERNEL_GENERIC, observables=obs) as ctx:
            ctx.init_lattice(seed=100)
            ctx.compute_state()
            ctx.set_log_weight(lnw)
            ctx.run(0, 100, 50 * N, 0)
            ctx.run(1, 2100, 0, 100 * N)
            out[obs] = ctx.reweight(deltas)
            if obs == M.OBS_MOMENTS:
                moms = [ctx.get_moments(b) for b in range(B)]
    for k in range(3):
        np.testing.assert_allclose(out[M.OBS_MOMENTS][k], out[M.OBS_JOINT][k], rtol=1e-12)
    absM, chi, _ = out[M.OBS_MOMENTS]
    for i, dl in enumerate(deltas):
        w = lnw.copy()
        orc.weight_shift(2.0, dl, w)
        bz = orc.boltz(2.0, dl, w)
        for b in range(B):
            a_ref, c_ref = orc.observables_moments(2.0, N, moms[b][0], moms[b][1], moms[b][2],
                                                   bz)
            assert absM[i, b] == pytest.approx(a_ref, rel=1e-12)
            assert chi[i, b] == pytest.approx(c_ref, rel=1e-9)


def test_reweighting_matches_oracle_observables(orc):
    """Device FP64 reweighting of the joint histograms vs the reference's
    long-double observables() (produce2_multigpu.cu:498-518)."""
    L, W, B = 8, 128, 4
    N = L * L
    lnw = -np.log(np.arange(1, N + 2, dtype=np.float64)) * 3.0
    lnw -= lnw[0]
    deltas = np.array([1.9, 1.95, 1.98, 2.0])
    with M.Context(L, W, precision=M.FP64, keying=M.KEYING_PRODUCE, joint_blocks=B) as ctx:
        ctx.init_lattice(seed=100)
        ctx.compute_state()
        ctx.set_log_weight(lnw)
        ctx.run(0, 100, 50 * N, 0)
        ctx.run(1, 2100, 0, 100 * N)
        absM, chi, binder = ctx.reweight(deltas)
        joints = [ctx.get_joint(b) for b in range(B)]
    for i, dl in enumerate(deltas):
        w = lnw.copy()
        orc.weight_shift(2.0, dl, w)
        bz = orc.boltz(2.0, dl, w)
        for b in range(B):
            a_ref, c_ref = orc.observables(2.0, N, joints[b], bz)
            assert absM[i, b] == pytest.approx(a_ref, rel=1e-12)
            assert chi[i, b] == pytest.approx(c_ref, rel=1e-9)
            S0, S1, S2, S4 = orc.joint_moments(N, joints[b])
            z = (bz * S0).sum()
            m2, m4 = (bz * S2).sum() / z, (bz * S4).sum() / z
            assert binder[i, b] == pytest.approx(1 - m4 / (3 * m2 * m2), rel=1e-10)


@pytest.mark.parametrize("kernel,L,W,prec,nup", [
    (M.KERNEL_TEAM, 16, 40, M.REF32, 600),
    (M.KERNEL_TEAM, 72, 48, M.REF32, 700),
    (M.KERNEL_TEAM, 24, 20, M.FP64, 500),
    (M.KERNEL_LANE, 16, 300, M.REF32, 400),
    (M.KERNEL_LANE, 12, 70, M.FP64, 300),
], ids=["team-L16", "team-L72", "team-L24-f64", "lane-L16", "lane-L12-f64"])
def test_lds_histogram_wrap_carry(orc, kernel, L, W, prec, nup):
    """The u32 LDS histogram bins start 16 short of 2^32 (test hook), so every
    bin a workgroup fills 16 or more times wraps and carries 2^32 into its
    uint64 HBM bin; the returned histogram must still be the oracle's exact
    count."""
    N = L * L
    rng = np.random.default_rng(N + W)
    lnw = -1.2 * np.arange(N + 1) + np.cumsum(rng.normal(scale=0.8, size=N + 1))
    lnw -= lnw[0]
    lat, E, D, Mg = oracle_init(orc, M.KEYING_WEIGHT, L, W, seed=21)
    with M.Context(L, W, precision=prec, kernel=kernel) as ctx:
        ctx.set_test_hist_preload(0xFFFFFFF0)
        ctx.set_lattice(lat.T)
        ctx.set_state(E, D, Mg)
        ctx.set_log_weight(lnw)
        hist, _ = ctx.run(1, 99, N, nup)
        glat = ctx.get_lattice()
    h = orc.mcs(L, lat, E, D, Mg, keying=M.KEYING_WEIGHT, fp64=prec == M.FP64,
                lnw32=lnw.astype(np.float32), lnw64=lnw, iter=1, seed=99, therm=N,
                nupdates=nup, beta=2.0, nthreads=NTHREADS)
    assert np.array_equal(glat, lat.T)
    assert np.array_equal(hist, h)
    # bins filled 16 or more times by one workgroup took the carry
    assert int(h.max()) >= 16 * 4
