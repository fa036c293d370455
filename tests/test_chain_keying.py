"""Pins the Philox keying of the oracle's update chain (R2) on the draws of
the reference's own Random123 code.

The keys and counters below are written out from the reference kernels, not
taken from the oracle:

    weight path   muca_delta_multigpu.cu:740-742
        k1 = {2*WORKER, 0xdecafbad}, k3 = {2*WORKER+50000, 0x11cafead}
        c  = {0, seed, iter, 0xBADC0DED}
    produce path  produce2_multigpu.cu:406-409
        k1 = {WORKER, 0xdecafbad},   k3 = {WORKER+40000, 0xdecafead}

and both kernels do `if (i % 4 == 0) { ++c[0]; r1 = rng(c, k1); r3 = rng(c, k3); }`
with i restarting at 0 for the production loop while c runs on
(muca_delta_multigpu.cu:756-759,792-797; produce2_multigpu.cu:425-428,458-462).
r3 picks the proposal through u01fixedpt<float>(r3[i%4]) < 0.5, r1 the
acceptance through log(u01fixedpt<float>(r1[i%4])) < diff.  Every block of
draws comes from oracle/_ref/r123_probe (the reference's philox.h and
uniform.hpp compiled from their sources); the chain is restated here step by
step, and after every single attempt the oracle's lattice, E, D, M and
histogram must equal the restatement's.  A swapped key, a wrong key offset,
a counter that restarts with the phase, or r1/r3 used the other way round
each change some attempt's outcome within the first few dozen attempts.
"""
import math
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "r123_probe")

pytestmark = pytest.mark.skipif(not os.path.exists(PROBE),
                                reason="oracle/_ref not built (reference absent)")

M32 = 0xFFFFFFFF


def ref_keys(keying, w):
    if keying == 0:   # muca_delta_multigpu.cu:740-741
        return [(2 * w) & M32, 0xdecafbad], [(2 * w + 50000) & M32, 0x11cafead]
    return [w & M32, 0xdecafbad], [(w + 40000) & M32, 0xdecafead]   # produce2:406-409


def probe(lines):
    out = subprocess.run([PROBE], input="\n".join(lines) + "\n", capture_output=True,
                         text=True, check=True).stdout.split()
    return out


def ref_draws(keying, w, seed, it, nblocks, rounds=7):
    """r1, r3 blocks for counter c[0] = 1 .. nblocks, and the u01fixedpt<float>
    of every word, all from the reference's code."""
    k1, k3 = ref_keys(keying, w)
    lines = []
    for b in range(1, nblocks + 1):
        c = [b, seed, it, 0xBADC0DED]
        lines.append("P %d %x %x %x %x %x %x" % (rounds, *c, *k1))
        lines.append("P %d %x %x %x %x %x %x" % (rounds, *c, *k3))
    words = [int(t, 16) for t in probe(lines)]
    r = np.array(words, np.uint32).reshape(nblocks, 2, 4)
    u = probe(["U %x" % int(x) for x in r.reshape(-1)])
    uf = np.array([int(t, 16) for t in u], np.uint32).view(np.float32).reshape(nblocks, 2, 4)
    return r, uf


def site_order(L):
    """muca_delta_multigpu.cu:173 for even L: even linear indices, then odd."""
    N = L * L
    return list(range(0, N, 2)) + list(range(1, N, 2))


def neighbours(L, s):
    x, y = s % L, s // L
    return [y * L + (x + 1) % L, y * L + (x - 1) % L, ((y + 1) % L) * L + x,
            ((y - 1) % L) * L + x]


def restated_chain(L, spins, keying, fp64, lnw, seed, it, therm, nup, w, beta2=True):
    """The chain of one walker, one state per attempt: yields (lattice, E, D,
    M, hist) after every attempt."""
    N = L * L
    order = site_order(L)
    s = [int(v) for v in spins]
    E = sum(-s[i] * (s[nb[0]] + s[nb[2]]) for i in range(N)
            for nb in [neighbours(L, i)])
    D = sum(v * v for v in s)
    Mg = sum(s)
    hist = np.zeros(N + 1, np.uint64)
    nblk = (therm + 3) // 4 + (nup + 3) // 4
    r, u = ref_draws(keying, w, seed, it, nblk)
    blk = -1
    for phase, n in ((0, therm), (1, nup)):
        for i in range(n):
            if i % 4 == 0:
                blk += 1
            j = i % 4
            idx = order[i % N]
            old = s[idx]
            step = 1 if u[blk, 1, j] < np.float32(0.5) else -1   # r3: the proposal
            new = old + step
            new = -1 if new == 2 else (1 if new == -2 else new)
            nb = sum(s[k] for k in neighbours(L, idx))
            dE = -(new - old) * nb
            dDel = new * new - old * old
            if fp64:
                diff = float(-dE * 2) + float(lnw[D + dDel]) - float(lnw[D])
            else:
                w32 = np.asarray(lnw, np.float32)
                diff = (np.float32(-dE * 2) + w32[D + dDel]) - w32[D]
            lg = np.float32(math.log(float(u[blk, 0, j])))             # r1: the acceptance
            if diff >= 0 or lg < diff:
                s[idx] = new
                E += dE
                D += dDel
                Mg += new - old
            if phase:
                hist[D] += 1
            yield np.array(s, np.int8), E, D, Mg, hist.copy()


@pytest.mark.parametrize("keying,fp64,w,seed,it,therm,nup", [
    (0, False, 0, 1234, 0, 6, 30),
    (0, False, 7, 100, 3, 64, 21),
    (0, False, 40000, 0xdeadbeef, 65535, 9, 14),
    (1, True, 0, 100, 0, 6, 30),
    (1, True, 5, 2100, 4, 11, 23),
    (1, True, 123457, 7, 1, 3, 17),
])
def test_oracle_chain_follows_reference_draws(orc, keying, fp64, w, seed, it, therm, nup):
    L = 8
    N = L * L
    rng = np.random.default_rng(w + 17 * seed)
    spins = rng.integers(-1, 2, size=N).astype(np.int8)
    # tilted rough weights: acceptances are neither all nor none
    lnw = -1.4 * np.arange(N + 1) + np.cumsum(rng.normal(scale=1.5, size=N + 1))
    lnw -= lnw[0]
    states = list(restated_chain(L, spins, keying, fp64, lnw, seed, it, therm, nup, w))
    assert len(states) == therm + nup
    moves = 0
    prev = spins
    for k, (ls, E, D, Mg, hist) in enumerate(states, start=1):
        lat = spins.reshape(1, N).copy()
        oE, oD, oM = orc.state(L, lat)
        t, n = min(k, therm), max(0, k - therm)
        oh = orc.mcs(L, lat, oE, oD, oM, keying=keying, fp64=fp64,
                     lnw32=lnw.astype(np.float32), lnw64=lnw, iter=it, seed=seed, therm=t,
                     nupdates=n, beta=2.0, w0=w)
        tag = f"attempt {k} (therm {t}, nupdates {n})"
        assert np.array_equal(lat[0], ls), tag
        assert (int(oE[0]), int(oD[0]), int(oM[0])) == (E, D, Mg), tag
        assert np.array_equal(oh, hist), tag
        moves += int(np.any(ls != prev))
        prev = ls
    # the draws decided something: some attempts were accepted, some not
    assert 0 < moves < therm + nup


def test_keying_mistakes_are_seen(orc):
    """The restated chain is sensitive to what this file pins: with k1 and k3
    swapped the restatement and the oracle part within the first attempts."""
    L, N, w, seed, it = 8, 64, 3, 1234, 0
    rng = np.random.default_rng(5)
    spins = rng.integers(-1, 2, size=N).astype(np.int8)
    lnw = -1.4 * np.arange(N + 1) + np.cumsum(rng.normal(scale=1.5, size=N + 1))
    lnw -= lnw[0]
    global ref_keys
    good = ref_keys
    try:
        ref_keys = lambda keying, w_: tuple(reversed(good(keying, w_)))   # noqa: E731
        states = list(restated_chain(L, spins, 0, False, lnw, seed, it, 0, 40, w))
    finally:
        ref_keys = good
    lat = spins.reshape(1, N).copy()
    oE, oD, oM = orc.state(L, lat)
    orc.mcs(L, lat, oE, oD, oM, keying=0, fp64=False, lnw32=lnw.astype(np.float32),
            iter=it, seed=seed, therm=0, nupdates=40, beta=2.0, w0=w)
    assert not np.array_equal(lat[0], states[-1][0])
