"""The production step at the C2 size: L = 72, 4096 walkers, ten jackknife
blocks (410 walkers each, the last 406), produce keying, FP64 acceptance,
the team kernel (produce2_multigpu.cu:401-496 for the update, :489 for the
joint histogram, :293-358 and :498-518 for the reweighting).

One production run (therm N, then NUPDATES 20 N per walker, stratified
start, near-flat T = 0.5 weights) is made twice from the same state, once
with the reference's joint H(D,M) per block and once with exact per-D
moments.  Checked:
  - block 0's joint histogram bit for bit against the oracle chain of its
    410 walkers;
  - for every block, the team kernel's moments equal, as exact integers,
    the moments of that block's joint histogram;
  - for every block and four Delta values, the device reweighting of the
    joint histograms against the oracle's long-double observables()
    (1e-12 relative for <|M|>, 1e-9 for chi), and the moments' reweighting
    against the joint histograms' (1e-12).
"""
import numpy as np
import pytest

import muca_amd as M
from test_gpu_timed_path import scaled_weights

pytestmark = pytest.mark.gpu

L, W, B = 72, 4096, 10
N = L * L
THERM, NUP = N, 20 * N
DELTAS = np.array([1.90, 1.95, 1.98, 2.00])


def _run(obs, lnw):
    with M.Context(L, W, precision=M.FP64, keying=M.KEYING_PRODUCE, joint_blocks=B,
                   observables=obs, kernel=M.KERNEL_TEAM) as ctx:
        assert ctx.info.kernel == M.KERNEL_TEAM
        ctx.init_lattice(mm=0, gpu=0, seed=100, stratified=True)
        ctx.compute_state()
        ctx.set_log_weight(lnw)
        ctx.run(0, 100, THERM, 0)
        hist, _ = ctx.run(1, 2100, 0, NUP)
        rw = ctx.reweight(DELTAS)
        if obs == M.OBS_JOINT:
            blocks = [ctx.get_joint(b) for b in range(B)]
        else:
            blocks = [ctx.get_moments(b) for b in range(B)]
        state = ctx.get_state()
    return hist, rw, blocks, state


def test_c2_production_joint_and_moments(orc):
    lnw = scaled_weights(L)
    hj, rwj, joints, sj = _run(M.OBS_JOINT, lnw)
    print("joint run done", flush=True)
    hm, rwm, moms, sm = _run(M.OBS_MOMENTS, lnw)
    print("moments run done", flush=True)
    # the same chains: the same D histogram and final state
    assert np.array_equal(hj, hm)
    for a, b in zip(sj, sm):
        assert np.array_equal(a, b)
    assert int(hj.sum()) == W * NUP

    # block 0 replayed by the oracle
    bw = (W + B - 1) // B
    lat = orc.init_lattice(M.KEYING_PRODUCE, L, bw, seed=100, strat_W=W)
    E, D, Mg = orc.state(L, lat)
    kw = dict(keying=M.KEYING_PRODUCE, fp64=True, lnw64=lnw)
    orc.mcs(L, lat, E, D, Mg, iter=0, seed=100, therm=THERM, nupdates=0, nthreads=8, **kw)
    j0 = np.zeros((N + 1) * (N + 2) // 2, np.uint64)
    orc.mcs(L, lat, E, D, Mg, iter=1, seed=2100, therm=0, nupdates=NUP, joint=j0, nthreads=1,
            **kw)
    assert np.array_equal(joints[0], j0)
    print("block 0 joint histogram equals the oracle's", flush=True)

    # per-D S0, S1, S2 of every block's joint histogram (exact in float64 at
    # these counts: every partial sum stays below 2^53), S4 as exact
    # integers for block 0
    Dj = np.repeat(np.arange(N + 1), np.arange(1, N + 2))
    base = (np.arange(N + 1) * np.arange(1, N + 2)) // 2
    am = np.abs(2 * (np.arange(Dj.size) - base[Dj]) - Dj).astype(np.float64)
    for b in range(B):
        c = joints[b].astype(np.float64)
        want = [np.bincount(Dj, weights=c, minlength=N + 1),
                np.bincount(Dj, weights=am * c, minlength=N + 1),
                np.bincount(Dj, weights=am * am * c, minlength=N + 1)]
        for k in range(3):
            got = np.array([float(x) for x in moms[b][k]])
            assert np.array_equal(got, want[k]), (b, k)
    exact = orc.joint_moments_exact(N, joints[0])
    assert list(moms[0][3]) == list(exact[3])
    print("moments of all blocks equal the joint histograms' moments", flush=True)

    absj, chij, binj = rwj
    Dg = np.arange(N + 1, dtype=np.float64)
    for b in range(B):
        S0 = np.bincount(Dj, weights=joints[b].astype(np.float64), minlength=N + 1)
        for i, dl in enumerate(DELTAS):
            # Boltzmann factors shifted over the block's visited bins, as the
            # device does (the reference's all-bin shift can underflow there)
            expo = -2.0 * dl * Dg - lnw
            bz = np.where(S0 > 0, np.exp(np.minimum(expo - np.max(expo[S0 > 0]), 0.0)), 0.0)
            a_ref, c_ref = orc.observables(2.0, N, joints[b], bz)
            assert absj[i, b] == pytest.approx(a_ref, rel=1e-12), (dl, b)
            assert chij[i, b] == pytest.approx(c_ref, rel=1e-9), (dl, b)
    for x, y in zip(rwm, rwj):
        np.testing.assert_allclose(x, y, rtol=1e-12)
