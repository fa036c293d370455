"""muca_reweight_noalloc against muca_reweight, through the C ABI.

The no-alloc variant shares the reweighting kernels of muca_reweight but keeps
its scratch (FP64 per-D moments, crystal fields, output planes, range flag)
in a fixed device array instead of buffers the context allocates; the array
keeps the previous call's contents, so two calls in a row, the second with a
smaller input, must both equal muca_reweight bit for bit.  Sizes beyond the
array are refused with MUCA_E_RANGE before any device work.
"""
import ctypes

import numpy as np
import pytest

import muca_amd as M

pytestmark = pytest.mark.gpu

_f64p = ctypes.POINTER(ctypes.c_double)


def _noalloc(ctx, deltas):
    fn = M.lib().muca_reweight_noalloc
    d = np.ascontiguousarray(deltas, np.float64)
    out = [np.full((d.size, ctx.joint_blocks), np.nan) for _ in range(3)]
    rc = fn(ctx._h, d.ctypes.data_as(_f64p), d.size, *[o.ctypes.data_as(_f64p) for o in out])
    return rc, out


def _production(L, W, blocks, obs, seed):
    ctx = M.Context(L, W, keying=M.KEYING_PRODUCE, precision=M.FP64, joint_blocks=blocks,
                    observables=obs)
    N = L * L
    rng = np.random.default_rng(seed)
    lnw = -1.9 * np.arange(N + 1) + np.cumsum(rng.normal(scale=0.5, size=N + 1))
    ctx.init_lattice(mm=0, gpu=0, seed=seed)
    ctx.compute_state()
    ctx.set_log_weight(lnw - lnw[0])
    ctx.run(0, seed, N, 4 * N)
    return ctx


@pytest.mark.parametrize("obs", [M.OBS_MOMENTS, M.OBS_JOINT])
def test_noalloc_equals_reweight_twice_then_smaller(obs):
    big = _production(32, 640, 10, obs, 11)
    small = _production(16, 96, 3, obs, 12)
    try:
        for ctx, deltas in ((big, np.linspace(1.90, 2.00, 21)), (small, [1.95, 1.98, 2.01])):
            print(f"noalloc: L={ctx.L} blocks={ctx.joint_blocks} ndelta={len(deltas)}",
                  flush=True)
            want = ctx.reweight(deltas)
            rc, got = _noalloc(ctx, deltas)
            assert rc == 0, M.lib().muca_last_error()
            for g, w in zip(got, want):
                assert np.array_equal(g, w)
                assert np.all(np.isfinite(g))
    finally:
        big.close()
        small.close()


def test_noalloc_refuses_beyond_its_scratch():
    """4 * blocks * (N+1) doubles at L = 128 with 64 blocks are 33.6 MB, more
    than the 32 MiB array: MUCA_E_RANGE, and the context still works."""
    ctx = _production(128, 64, 64, M.OBS_MOMENTS, 13)
    try:
        rc, _ = _noalloc(ctx, [1.98])
        assert rc == -6
        assert b"maximum" in M.lib().muca_last_error()
        a, c, b = ctx.reweight([1.98])
        assert np.all(np.isfinite(a))
    finally:
        ctx.close()
