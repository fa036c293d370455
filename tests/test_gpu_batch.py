"""GPU parity of the batched engine (muca_batch_*): many independent weight
problems, the update on persistent team or lane tiles and the host loop body
(muca_delta_multigpu.cu:379-476) in a device kernel, optionally replayed from a
captured hipGraph.

Each instance is followed by its own oracle chain, iteration by iteration:
the histogram and lattice-derived state must be identical given the weights the
device used; the device weight update (FP64, device log) must agree with the
host formula applied to the same histogram within a few ulps; width, range,
Kullback-Leibler divergence and the NUPDATES schedule must match the host rule.
"""
import numpy as np
import pytest

import muca_amd as M

pytestmark = pytest.mark.gpu

# (L, walkers, seed, init_nup, beta, recursion)
TEAM_SET = [(8, 16, 11, 40, 2.0, False), (16, 40, 12, 20, 2.0, True),
            (24, 20, 13, 10, 1.6, False), (8, 5, 14, 64, 2.5, True)]
LANE_SET = [(6, 70, 21, 30, 2.0, False), (10, 261, 22, 12, 2.0, True),
            (16, 33, 23, 8, 1.7, False), (4, 64, 24, 50, 2.0, False)]
# team tiles whose lattices do not fit a workgroup's LDS (lattices in HBM)
HBM_SET = [(96, 32, 31, 4, 2.0, True), (128, 16, 32, 2, 2.0, False)]


def _instances(spec):
    return [dict(L=L, walkers=W, seed=s, init_nup=n, beta=b, recursion=r)
            for (L, W, s, n, b, r) in spec]


class OracleInstance:
    """One instance's chain on the oracle, fed the device's weights."""

    def __init__(self, orc, q, stratified=False):
        self.orc, self.q = orc, q
        self.L, self.N = q["L"], q["L"] * q["L"]
        self.lat = orc.init_lattice(M.KEYING_WEIGHT, self.L, q["walkers"], mm=0, gpu=0,
                                    seed=q["seed"], rounds=7,
                                    strat_W=q["walkers"] if stratified else 0)
        self.E, self.D, self.Mg = orc.state(self.L, self.lat)
        self.width, self.nup = 10, q["init_nup"]
        self.p_acc = np.zeros(self.N + 1)
        self.converged = False

    def step(self, k, lnw_dev):
        """Iteration k with the device's ln W; returns (hist, expected ln W,
        start, end, dkl, nup used, converged)."""
        orc, N = self.orc, self.N
        if self.width == N:
            self.nup = int((self.nup + 1) * 1.01)
        nup = self.nup
        hist = orc.mcs(self.L, self.lat, self.E, self.D, self.Mg, keying=M.KEYING_WEIGHT,
                       fp64=False, lnw32=lnw_dev.astype(np.float32), iter=k,
                       seed=self.q["seed"], therm=N, nupdates=nup, beta=self.q["beta"],
                       rounds=7, nthreads=8)
        s, e = orc.hist_range(hist)
        self.width = max(self.width, e - s)
        dkl = orc.kull_leibler(hist)
        conv = self.width == N and dkl < 1e-4
        lnw = lnw_dev.copy()
        if not conv:
            if self.q["recursion"]:
                orc.weight_recursion(hist, self.p_acc, s, e, lnw)
            else:
                orc.update_trivial(hist, lnw)
            orc.normalise(lnw)
        self.converged = conv
        return hist, lnw, s, e, dkl, nup, conv


def _run_against_oracle(orc, spec, tiles, use_graph, iters, stratified=False):
    ins = _instances(spec)
    with M.Batch(ins, tiles=tiles) as b:
        assert b.tiles == tiles
        b.init(0, stratified=stratified)
        ref = [OracleInstance(orc, q, stratified) for q in ins]
        for k in range(iters):
            before = [b.get_log_weight(i) for i in range(b.n)]
            b.iterate(1, use_graph=use_graph)
            st = b.status()
            for i, o in enumerate(ref):
                if o.converged:
                    assert st[i]["converged"] == 1
                    assert np.array_equal(b.get_log_weight(i), before[i])
                    continue
                hist, lnw, s, e, dkl, nup, conv = o.step(k, before[i])
                tag = f"instance {i} {spec[i]} iteration {k}"
                assert np.array_equal(b.get_hist(i), hist), tag
                got = b.get_log_weight(i)
                assert got[0] == 0.0, tag
                np.testing.assert_allclose(got, lnw, rtol=1e-13, atol=1e-11, err_msg=tag)
                assert (st[i]["start"], st[i]["end"]) == (s, e), tag
                assert st[i]["width"] == o.width, tag
                assert abs(st[i]["dkl"] - dkl) <= 1e-12 * max(1.0, abs(dkl)), tag
                assert bool(st[i]["converged"]) == conv, tag
                assert st[i]["iter"] == k + 1, tag
                if not conv:
                    nxt = int((nup + 1) * 1.01) if o.width == o.N else nup
                    assert st[i]["nup"] == nxt, tag
                W = spec[i][1]
                assert st[i]["attempts"] >= W * (o.N + nup), tag


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_team_tiles_match_oracle(orc, use_graph):
    _run_against_oracle(orc, TEAM_SET, M.KERNEL_TEAM, use_graph, iters=4)


def test_batch_team_tiles_lattices_in_hbm_match_oracle(orc):
    """L = 96 and 128: the batch engine launches its build of the team kernel
    for lattices in HBM (the next chunk's spins loaded during the current
    chunk's chain), replayed from a graph."""
    _run_against_oracle(orc, HBM_SET, M.KERNEL_TEAM, True, iters=2)


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_lane_tiles_match_oracle(orc, use_graph):
    _run_against_oracle(orc, LANE_SET, M.KERNEL_LANE, use_graph, iters=4)


def test_batch_stratified_start_matches_oracle(orc):
    _run_against_oracle(orc, TEAM_SET, M.KERNEL_TEAM, True, iters=3, stratified=True)


def test_batch_fp64_lane_first_iteration(orc):
    """FP64 acceptance: the first iteration's histograms match the FP64 oracle."""
    ins = _instances(LANE_SET[:3])
    with M.Batch(ins, precision=M.FP64, tiles=M.KERNEL_LANE) as b:
        b.init(0)
        b.iterate(1, use_graph=True)
        for i, q in enumerate(ins):
            L, N = q["L"], q["L"] ** 2
            lat = orc.init_lattice(M.KEYING_WEIGHT, L, q["walkers"], seed=q["seed"])
            E, D, Mg = orc.state(L, lat)
            h = orc.mcs(L, lat, E, D, Mg, keying=M.KEYING_WEIGHT, fp64=True,
                        lnw64=np.zeros(N + 1), iter=0, seed=q["seed"], therm=N,
                        nupdates=q["init_nup"], beta=q["beta"], nthreads=8)
            assert np.array_equal(b.get_hist(i), h), i


def test_batch_graph_equals_eager():
    """Replaying captured iterations gives the same results as launching
    them: a graph of 3 iterations, then one of 8 replayed three times with
    eager remainders and host reads in between (3 + 11 + 16 iterations)."""
    ins = _instances(TEAM_SET)
    out = []
    for g in (True, False):
        with M.Batch(ins, tiles=M.KERNEL_TEAM) as b:
            b.init(3)
            snaps = []
            for n in (3, 11, 16):
                b.iterate(n, use_graph=g)
                snaps.append(([b.get_log_weight(i) for i in range(b.n)],
                              [b.get_hist(i) for i in range(b.n)], b.status()))
            out.append(snaps)
    for k, (sg, se) in enumerate(zip(*out)):
        for i in range(len(ins)):
            assert np.array_equal(sg[0][i], se[0][i]), (k, i)
            assert np.array_equal(sg[1][i], se[1][i]), (k, i)
        assert sg[2] == se[2], k


def test_batch_converged_instance_stops():
    """An instance that meets width == N and D_KL < 1e-4 is left alone while
    the others keep iterating."""
    ins = [dict(L=4, walkers=4096, seed=5, init_nup=2000),
           dict(L=16, walkers=16, seed=6, init_nup=4)]
    with M.Batch(ins, tiles=M.KERNEL_LANE) as b:
        b.init(0)
        done = None
        for k in range(60):
            b.iterate(1)
            st = b.status()
            if st[0]["converged"]:
                done = k
                break
        assert done is not None, "L=4 did not converge in 60 iterations"
        h = b.get_hist(0)
        assert st[0]["width"] == 16
        assert M.kull_leibler(h) < 1e-4
        lnw, it1 = b.get_log_weight(0), st[1]["iter"]
        b.iterate(3)
        st2 = b.status()
        assert st2[0]["iter"] == st[0]["iter"] and st2[0]["done_iter"] == done
        assert np.array_equal(b.get_log_weight(0), lnw)
        assert np.array_equal(b.get_hist(0), h)
        assert st2[1]["iter"] == it1 + 3


def test_batch_auto_tiles():
    """Team tiles while every L is a multiple of 8 and walkers are few, lane
    tiles otherwise."""
    with M.Batch(_instances(TEAM_SET)) as b:
        assert b.tiles == M.KERNEL_TEAM
        assert b.ntiles == sum((w + 15) // 16 for (_, w, *_r) in TEAM_SET)
    with M.Batch(_instances(LANE_SET)) as b:
        assert b.tiles == M.KERNEL_LANE
        assert b.walkers == sum(w for (_, w, *_r) in LANE_SET)


def test_batch_autotune_leaves_fresh_state():
    """After the autotuner the batch is as after init(mm): its iterations
    match a batch created with the chosen tiles and initialised directly."""
    ins = _instances(TEAM_SET)
    with M.Batch(ins) as b:
        kind, block, ms = b.autotune(2)
        assert kind in (M.KERNEL_TEAM, M.KERNEL_LANE) and ms > 0
        assert b.tiles == kind and b.block == block
        b.iterate(3)
        got = [b.get_log_weight(i) for i in range(b.n)], [b.get_hist(i) for i in range(b.n)]
    with M.Batch(ins, tiles=kind) as b:
        b.init(2)
        u, w = b.iterate_profiled(3)
        assert u > 0 and w > 0
        for i in range(b.n):
            assert np.array_equal(b.get_log_weight(i), got[0][i])
            assert np.array_equal(b.get_hist(i), got[1][i])


def test_batch_rejects_bad_instances():
    with pytest.raises(M.MucaError):
        M.Batch([dict(L=7, walkers=4, seed=1)])
    with pytest.raises(M.MucaError):
        M.Batch([dict(L=12, walkers=4, seed=1)], tiles=M.KERNEL_TEAM)


def test_batch_autotune_keeps_weights_set_before():
    """Weights set with set_log_weight (a surrogate seed) survive the
    autotuner, which re-initialises the walkers for every candidate."""
    ins = _instances(TEAM_SET)
    seeds = []
    with M.Batch(ins) as b:
        b.init(1)
        for i, q in enumerate(ins):
            N = q["L"] ** 2
            lnw = -0.7 * np.arange(N + 1) + np.sin(0.3 * np.arange(N + 1))
            seeds.append(lnw - lnw[0])
            b.set_log_weight(i, seeds[-1])
        b.autotune(1)
        for i in range(b.n):
            assert np.array_equal(b.get_log_weight(i), seeds[i])
        b.iterate(1)
        kind = b.tiles
        got = [b.get_hist(i) for i in range(b.n)]
    with M.Batch(ins, tiles=kind) as b:
        b.init(1)
        for i in range(b.n):
            b.set_log_weight(i, seeds[i])
        b.iterate(1)
        for i in range(b.n):
            assert np.array_equal(b.get_hist(i), got[i])


def test_batch_nupdates_growth_and_cap():
    """A seeded-protocol schedule: once width == N, NUPDATES = (NUPDATES + 1)
    * growth each iteration, stopping at nup_max; the reference's 1.01 stays
    the default.  Each iteration's histogram holds walkers x NUPDATES entries."""
    from math import comb, log
    # 16 walkers: too few samples for D_KL < 1e-4, so neither converges
    ins = [dict(L=4, walkers=16, seed=3, init_nup=10, nup_growth=2.0, nup_max=1000, beta=0.01),
           dict(L=4, walkers=16, seed=4, init_nup=10, beta=0.01)]
    # near-flat weights at small beta (minus the log of the number of states
    # with each D): the histogram spans all 17 bins from the first iteration
    flat = np.array([-(log(comb(16, D)) + D * log(2)) for D in range(17)])
    with M.Batch(ins, tiles=M.KERNEL_LANE) as b:
        b.init(0, stratified=True)
        for i in range(2):
            b.set_log_weight(i, flat - flat[0])
        expect = [10, 10]
        for k in range(12):
            b.iterate(1)
            st = b.status()
            for i in range(2):
                assert int(b.get_hist(i).sum()) == 16 * expect[i], (k, i)
                if st[i]["converged"]:
                    continue
                if st[i]["width"] == 16:
                    g = 2.0 if i == 0 else 1.01
                    expect[i] = min(int((expect[i] + 1) * g), 1000 if i == 0 else 1 << 48)
                assert st[i]["nup"] == expect[i], (k, i, st[i])
            assert st[0]["nup_capped"] == (1 if expect[0] == 1000 else 0)
        assert expect[0] == 1000 and not any(x["converged"] for x in b.status())
