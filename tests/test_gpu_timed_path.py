"""The timed benchmark path at the size it is timed, against the oracle.

bench.py's default workload drives the batched engine: L = 72 (5185 D-bins),
4096 walkers, seed 1234, recursive weight update, NUPDATES = 1 from ln W = 0,
team tiles, each iteration a replayed hipGraph of the update kernel and the
device weight update (k_batch_update).  Here the same batch runs the warm-up
and timed iterations (2 + 20) and every iteration is checked against the
oracle chain fed the device's weights: the histogram bit for bit, the device
weight_recursion against the host formula (muca_delta_multigpu.cu:833-859)
within a few ulps, range, width, D_KL and the NUPDATES schedule.

A second case starts the walkers spread over D (stratified start) under
near-flat weights, so the histogram ranges exceed 2048 bins and
k_batch_update walks the recursion in several LDS-staged chunks, carrying the
running ln W from one chunk to the next.
"""
import numpy as np
import pytest

import muca_amd as M
from test_gpu_batch import OracleInstance

pytestmark = pytest.mark.gpu


def _check_iterations(orc, q, iters, stratified, min_span=None, lnw0=None):
    spans = []
    with M.Batch([q]) as b:
        assert b.tiles == M.KERNEL_TEAM
        b.init(0, stratified=stratified)
        if lnw0 is not None:
            b.set_log_weight(0, lnw0)
        ref = OracleInstance(orc, q, stratified)
        for k in range(iters):
            before = b.get_log_weight(0)
            b.iterate(1, use_graph=True)
            st = b.status()[0]
            hist, lnw, s, e, dkl, nup, conv = ref.step(k, before)
            tag = f"L={q['L']} W={q['walkers']} iteration {k}"
            print(tag, "range", s, e, flush=True)
            assert np.array_equal(b.get_hist(0), hist), tag
            got = b.get_log_weight(0)
            np.testing.assert_allclose(got, lnw, rtol=1e-13, atol=1e-11, err_msg=tag)
            assert (st["start"], st["end"]) == (s, e), tag
            assert st["width"] == ref.width, tag
            assert abs(st["dkl"] - dkl) <= 1e-12 * max(1.0, abs(dkl)), tag
            assert st["iter"] == k + 1 and not st["converged"], tag
            nxt = int((nup + 1) * 1.01) if ref.width == ref.N else nup
            assert st["nup"] == nxt, tag
            spans.append(e - s)
        # the walkers' E, D, M after the last iteration are the oracle's
        gE, gD, gM = b.get_state(0)
        assert np.array_equal(gE, ref.E) and np.array_equal(gD, ref.D)
        assert np.array_equal(gM, ref.Mg)
    if min_span is not None:
        assert max(spans) > min_span, spans
    return spans


def test_bench_c2_workload_every_iteration(orc):
    q = dict(L=72, walkers=4096, seed=1234, beta=2.0, recursion=True, init_nup=1)
    _check_iterations(orc, q, iters=22, stratified=False)


def scaled_weights(L):
    """An approximately flat-histogram ln W at T = 0.5 for side L: the
    committed converged L = 64 curve, extensive in N, interpolated in D/N."""
    from muca_amd import surrogate as S
    runs = [r for r in S.load_corpus() if r["L"] == 64 and abs(r["beta"] - 2.0) < 1e-9]
    src = runs[0]["lnw"]
    n0, N = len(src) - 1, L * L
    lnw = np.interp(np.arange(N + 1) / N, np.arange(n0 + 1) / n0, src / n0) * N
    return lnw - lnw[0]


def test_recursion_over_several_lds_chunks(orc):
    """Walkers spread over D (stratified start) under near-flat weights give
    histogram ranges of ~3000 bins, more than the 2048 recursion steps the
    device update stages in LDS at a time: the recursion runs in two chunks
    and carries ln W across the boundary."""
    q = dict(L=72, walkers=4096, seed=99, beta=2.0, recursion=True, init_nup=1)
    spans = _check_iterations(orc, q, iters=20, stratified=True, min_span=2048,
                              lnw0=scaled_weights(72))
    print("ranges", spans)
