"""The 96-instance batch (L in {32,48,64,72,96,128} x Delta in {1.90,1.95,
1.98,2.00} x T in {0.5,0.6,0.7,0.8}, seed = instance id, recursive update) on
the GPU, with 256 walkers per instance in place of 4096.

Every walker is an independent chain, so per instance the oracle replays
walkers 0, 1 and the last one bit for bit through two weight iterations (the
second with the device's updated ln W); for all walkers, E, D and M
recomputed from the final lattices equal the engine's incremental ones, and
each histogram holds exactly walkers x NUPDATES entries.  Both tile kinds
run the same batch.
"""
import numpy as np
import pytest

import muca_amd as M

pytestmark = pytest.mark.gpu

W = 256


def batch96():
    ins, meta = [], []
    for L in (32, 48, 64, 72, 96, 128):
        for d in (1.90, 1.95, 1.98, 2.00):
            for T in (0.5, 0.6, 0.7, 0.8):
                ins.append(dict(L=L, walkers=W, seed=len(ins), beta=1.0 / T, recursion=True))
                meta.append((L, d, T))
    return ins, meta


def state_of(L, lat_site_major):
    """E, D, M of every walker from a site-major lattice, by brute force."""
    s = lat_site_major.astype(np.int32).reshape(L, L, -1)
    E = -(s * (np.roll(s, -1, axis=1) + np.roll(s, -1, axis=0))).sum(axis=(0, 1))
    return E, (s * s).sum(axis=(0, 1)), s.sum(axis=(0, 1))


@pytest.mark.parametrize("tiles", [M.KERNEL_TEAM, M.KERNEL_LANE])
def test_batch96_sampled_walkers_and_invariants(orc, tiles):
    ins, meta = batch96()
    sample = [0, 1, W - 1]
    with M.Batch(ins, tiles=tiles) as b:
        assert b.n == 96 and b.tiles == tiles
        b.init(0)
        # per instance and sampled walker: the oracle's lattice and state
        rep = {}
        for i, q in enumerate(ins):
            for w in sample:
                lat = orc.init_lattice(M.KEYING_WEIGHT, q["L"], 1, seed=q["seed"], w0=w)
                rep[i, w] = (lat, *orc.state(q["L"], lat))
        for k in range(2):
            before = [b.get_log_weight(i) for i in range(b.n)]
            nups = [s["nup"] for s in b.status()]
            b.iterate(1, use_graph=True)
            for i, q in enumerate(ins):
                L, N = q["L"], q["L"] ** 2
                tag = f"instance {i} {meta[i]} iteration {k}"
                print(tag, flush=True)
                h = b.get_hist(i)
                assert int(h.sum()) == W * nups[i], tag
                glat = b.get_lattice(i)
                gE, gD, gM = b.get_state(i)
                rE, rD, rM = state_of(L, glat)
                assert np.array_equal(gE, rE) and np.array_equal(gD, rD), tag
                assert np.array_equal(gM, rM), tag
                for w in sample:
                    lat, E, D, Mg = rep[i, w]
                    orc.mcs(L, lat, E, D, Mg, keying=M.KEYING_WEIGHT, fp64=False,
                            lnw32=before[i].astype(np.float32), iter=k, seed=q["seed"],
                            therm=N, nupdates=nups[i], beta=q["beta"], w0=w)
                    assert np.array_equal(glat[:, w], lat[0]), (tag, w)
                    assert (gE[w], gD[w], gM[w]) == (E[0], D[0], Mg[0]), (tag, w)
        st = b.status()
        assert all(s["iter"] == 2 for s in st)
