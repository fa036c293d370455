"""Full-size lattices on the GPU: L = 128 (the team kernel with lattices in
HBM instead of LDS, the lane kernel with 64 KB of ln W in LDS) and L = 256
(65 536 spins, 65 537 D-bins: ln W fits no LDS; the generic kernel, and
the team kernel reading ln W from HBM one window per team).

Every walker is an independent chain, so a sample of walkers pins the
launch: the oracle replays walkers 0..7 (and the last walker) bit for bit;
the rest of the launch is checked through properties that hold at any size
(E, D, M recomputed from the final lattices equal the incremental ones; the
histogram counts exactly walkers x NUPDATES entries)."""
import numpy as np
import pytest

import muca_amd as M

pytestmark = pytest.mark.gpu

SAMPLE = [0, 1, 2, 3, 4, 5, 6, 7]


def _weights(N, beta):
    # a tilted, rough ln W so that acceptances are far from trivial
    rng = np.random.default_rng(N)
    lnw = -1.6 * beta * np.arange(N + 1) + np.cumsum(rng.normal(scale=0.7, size=N + 1))
    return lnw - lnw[0]


@pytest.mark.parametrize("L,W,kernel,prec", [
    (128, 96, M.KERNEL_TEAM, M.REF32),
    (128, 96, M.KERNEL_TEAM, M.FP64),
    (128, 256, M.KERNEL_LANE, M.REF32),
    (128, 96, M.KERNEL_GENERIC, M.REF32),
    (256, 64, M.KERNEL_GENERIC, M.REF32),
    (256, 64, M.KERNEL_GENERIC, M.FP64),
    (256, 48, M.KERNEL_TEAM, M.REF32),
    (256, 48, M.KERNEL_TEAM, M.FP64),
    (200, 40, M.KERNEL_TEAM, M.FP64),
], ids=["L128-team-f32", "L128-team-f64", "L128-lane-f32", "L128-generic-f32",
        "L256-generic-f32", "L256-generic-f64", "L256-team-f32", "L256-team-f64",
        "L200-team-f64"])
def test_large_lattice_walker_sample(orc, L, W, kernel, prec):
    N = L * L
    beta = 2.0
    lnw = _weights(N, beta)
    therm, nup = N // 2, N + 3
    with M.Context(L, W, kernel=kernel, precision=prec, beta=beta) as ctx:
        ctx.init_lattice(mm=0, gpu=0, seed=41)
        ctx.compute_state()
        E0, D0, M0 = ctx.get_state()
        ctx.set_log_weight(lnw)
        hist, _ = ctx.run(2, 555, therm, nup)
        glat = ctx.get_lattice()
        gE, gD, gM = ctx.get_state()
        ctx.compute_state()
        rE, rD, rM = ctx.get_state()
    assert int(hist.sum()) == W * nup
    assert np.array_equal(gE, rE) and np.array_equal(gD, rD) and np.array_equal(gM, rM)
    assert np.any(gD != D0)   # the walk moved
    for w in SAMPLE + [W - 1]:
        lat = orc.init_lattice(M.KEYING_WEIGHT, L, 1, seed=41, w0=w)
        E, D, Mg = orc.state(L, lat)
        assert (E[0], D[0], Mg[0]) == (E0[w], D0[w], M0[w])
        orc.mcs(L, lat, E, D, Mg, keying=M.KEYING_WEIGHT, fp64=prec == M.FP64,
                lnw32=lnw.astype(np.float32), lnw64=lnw, iter=2, seed=555, therm=therm,
                nupdates=nup, beta=beta, w0=w)
        assert np.array_equal(glat[:, w], lat[0]), w
        assert (gE[w], gD[w]) == (E[0], D[0]), w


@pytest.mark.parametrize("rounds,keying,prec", [
    (10, M.KEYING_WEIGHT, M.REF32),
    (7, M.KEYING_PRODUCE, M.FP64),
    (10, M.KEYING_PRODUCE, M.FP64),
], ids=["r10-weight-f32", "r7-produce-f64", "r10-produce-f64"])
def test_team_kernel_hbm_weights_rounds_and_keyings(orc, rounds, keying, prec):
    """The team kernel at L = 256 (ln W read from HBM) with Philox4x32-10 and
    with the produce keying (whose acceptance window is 6 D values wide
    against 8 on the weight path): sampled walkers bit for bit."""
    L, W = 256, 48
    N = L * L
    beta = 2.0
    lnw = _weights(N, beta)
    therm, nup = N // 2, N + 3
    with M.Context(L, W, kernel=M.KERNEL_TEAM, precision=prec, beta=beta, rounds=rounds,
                   keying=keying) as ctx:
        assert ctx.info.kernel == M.KERNEL_TEAM
        ctx.init_lattice(mm=0, gpu=0, seed=41)
        ctx.compute_state()
        E0, D0, M0 = ctx.get_state()
        ctx.set_log_weight(lnw)
        hist, _ = ctx.run(2, 555, therm, nup)
        glat = ctx.get_lattice()
        gE, gD, gM = ctx.get_state()
    assert int(hist.sum()) == W * nup
    assert np.any(gD != D0)
    for w in [0, 1, W - 1]:
        lat = orc.init_lattice(keying, L, 1, seed=41, w0=w, rounds=rounds)
        E, D, Mg = orc.state(L, lat)
        assert (E[0], D[0], Mg[0]) == (E0[w], D0[w], M0[w])
        orc.mcs(L, lat, E, D, Mg, keying=keying, fp64=prec == M.FP64,
                lnw32=lnw.astype(np.float32), lnw64=lnw, iter=2, seed=555, therm=therm,
                nupdates=nup, beta=beta, rounds=rounds, w0=w)
        assert np.array_equal(glat[:, w], lat[0]), w
        assert (gE[w], gD[w], gM[w]) == (E[0], D[0], Mg[0]), w


@pytest.mark.parametrize("L,prec,therm,nup", [
    (96, M.FP64, 1, 2 * 96 * 96 + 7),
    (160, M.REF32, 37, 160 * 160 + 5),
    (128, M.REF32, 128 * 128 // 2 + 70, 3),
    (136, M.REF32, 68 * 3 + 1, 136 * 68 + 64),
], ids=["L96-f64-therm1", "L160-f32-therm37", "L128-f32-midrow", "L136-f32-rows"])
def test_team_kernel_hbm_lattices_chunk_boundaries(orc, L, prec, therm, nup):
    """The team kernel with lattices in HBM and ln W in LDS, where each
    chunk's spins are loaded while the chunk before it runs: phases that end
    inside a half-row or a chunk, a thermalisation of one attempt (the next
    chunk reads the sites the one before wrote), and half-rows of two chunks
    (L = 136, 160).  Sampled walkers bit for bit against the oracle chain."""
    W = 32
    N = L * L
    beta = 2.0
    lnw = _weights(N, beta)
    with M.Context(L, W, kernel=M.KERNEL_TEAM, precision=prec, beta=beta) as ctx:
        assert ctx.info.kernel == M.KERNEL_TEAM
        ctx.init_lattice(mm=0, gpu=0, seed=43)
        ctx.compute_state()
        E0, D0, M0 = ctx.get_state()
        ctx.set_log_weight(lnw)
        hist, _ = ctx.run(3, 777, therm, nup)
        glat = ctx.get_lattice()
        gE, gD, gM = ctx.get_state()
        ctx.compute_state()
        rE, rD, rM = ctx.get_state()
    assert int(hist.sum()) == W * nup
    assert np.array_equal(gE, rE) and np.array_equal(gD, rD) and np.array_equal(gM, rM)
    for w in [0, 5, 17, W - 1]:
        lat = orc.init_lattice(M.KEYING_WEIGHT, L, 1, seed=43, w0=w)
        E, D, Mg = orc.state(L, lat)
        assert (E[0], D[0], Mg[0]) == (E0[w], D0[w], M0[w])
        orc.mcs(L, lat, E, D, Mg, keying=M.KEYING_WEIGHT, fp64=prec == M.FP64,
                lnw32=lnw.astype(np.float32), lnw64=lnw, iter=3, seed=777, therm=therm,
                nupdates=nup, beta=beta, w0=w)
        assert np.array_equal(glat[:, w], lat[0]), w
        assert (gE[w], gD[w], gM[w]) == (E[0], D[0], Mg[0]), w
