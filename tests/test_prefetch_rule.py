"""The rule by which the team kernel (lattices in HBM, ln W in LDS) loads a
chunk's spins while the chunk before it runs its D chain
(muca_amd/csrc/team_run.h, the PF block after phase A), restated here and
enumerated over every chunk transition of whole launches: a prefetched site
must never be one the current chunk writes, and where the rule forwards the
lower neighbours from registers, the forwarded site must be the one the
next chunk reads.

Sweep order of a phase (the update kernels' order for even L): attempt rs
is at position rs mod N, first all sites with even x row by row (half h = 0),
then all with odd x (h = 1); a half-row of L/2 sites is cut into chunks of
up to 16 NR attempts; both phases (therm, NUPDATES) start at position 0.
"""
import pytest


def _chunks(L, NR, therm, nup):
    N, Rw = L * L, L // 2
    for phase, n in ((0, therm), (1, nup)):
        rs, h, y, jb = 0, 0, 0, 0
        while rs < n:
            cnt = min(Rw - jb, 16 * NR, n - rs)
            yield phase, rs, n, h, y, jb, cnt
            rs += cnt
            jb += cnt
            if jb == Rw:
                jb, y = 0, y + 1
                if y == L:
                    y, h = 0, h ^ 1


def _decide(L, NR, cur, nup):
    """(have, fwd, next geometry) as the kernel decides them."""
    Rw = L // 2
    phase, rs, n, h, y, jb, cnt = cur
    rs_n, n_n, nx = rs + cnt, n, rs + cnt < n
    h_n, y_n, jb_n = h, y, jb + cnt
    if jb_n == Rw:
        jb_n, y_n = 0, y_n + 1
        if y_n == L:
            y_n, h_n = 0, h_n ^ 1
    if not nx and phase == 0 and nup > 0:
        rs_n, n_n, nx = 0, nup, True
        h_n = y_n = jb_n = 0
    cnt_n = min(Rw - jb_n, 16 * NR, n_n - rs_n) if nx else 0
    yu_n = 0 if y_n + 1 == L else y_n + 1
    yd_n = L - 1 if y_n == 0 else y_n - 1
    ovl = jb_n < jb + cnt and jb < jb_n + cnt_n
    have, fwd = nx, False
    if y_n == y and h_n != h:
        have = False
    if h_n == h and (y_n == y or yu_n == y) and ovl:
        have = False
    if h_n == h and yd_n == y and ovl:
        if jb_n == jb:
            fwd = True
        else:
            have = False
    return have, fwd, (h_n, y_n, jb_n, cnt_n)


@pytest.mark.parametrize("L", [8, 16, 40, 72, 96, 128, 136, 160, 200, 256])
def test_prefetched_sites_are_not_written_by_the_chunk_before(L):
    N, Rw = L * L, L // 2
    NR = min(4, (Rw + 15) // 16)
    pairs = [(N, 1), (N, N), (N, 3 * N + 7), (1, 5), (0, N), (N, 0), (Rw - 3, N),
             (37, 2 * N + 5), (NR * 16 + 1, Rw * 3 + 2), (N // 2 + 5, N // 2 - 3)]
    seen = {True: 0, False: 0}
    for therm, nup in pairs:
        seq = list(_chunks(L, NR, therm, nup))
        for cur, nxt in zip(seq, seq[1:]):
            have, fwd, geo = _decide(L, NR, cur, nup)
            assert geo == nxt[3:], (L, cur, nxt)
            seen[fwd] += have
            if not have:
                continue
            _, _, _, h, y, jb, cnt = cur
            h_n, y_n, jb_n, cnt_n = geo
            written = {(y, 2 * (jb + j) + h) for j in range(cnt)}
            yu_n = 0 if y_n + 1 == L else y_n + 1
            yd_n = L - 1 if y_n == 0 else y_n - 1
            for lane in range(min(16 * NR, cnt_n)):   # lanes past cnt_n keep nothing
                x = 2 * (jb_n + lane) + h_n
                xr = 0 if x + 1 == L else x + 1
                xl = L - 1 if x == 0 else x - 1
                for site in ((y_n, x), (y_n, xr), (y_n, xl), (yu_n, x)):
                    assert site not in written, (L, therm, nup, cur, lane, site)
                if fwd and lane < cnt:
                    # forwarded: this lane's own site of the chunk before
                    assert (yd_n, x) == (y, 2 * (jb + lane) + h), (L, cur, lane)
                else:
                    assert (yd_n, x) not in written, (L, therm, nup, cur, lane)
    assert seen[False] > 0
    # one chunk per half-row (L <= 128): the next row's chunk starts at the
    # same place, and its lower neighbours are forwarded
    assert (seen[True] > 0) == (Rw <= 16 * NR)
