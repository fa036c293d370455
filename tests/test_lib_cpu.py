"""libmuca without a GPU: the library loads, exports every symbol that
include/muca.h declares, and its host-side code agrees with the oracle."""
import os
import re

import numpy as np
import pytest

import muca_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "muca.h")).read()
    return sorted(set(re.findall(r"\b(muca_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    lib = M.lib()
    syms = header_symbols()
    assert len(syms) >= 24
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(M._SIGS), set(syms) ^ set(M._SIGS)


def test_acclog_host_instantiation_exhaustive(orc):
    """The acceptance log compiled for the host equals the oracle's correctly
    rounded log on every one of the 2^23 inputs."""
    n = 1 << 23
    step = 1 << 21
    for k0 in range(0, n, step):
        a = M.acclog_host_range(k0, step)
        b = orc.logf_u01_range(k0, step)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k0


def _random_hists(rng, N, count=20):
    for t in range(count):
        h = rng.integers(0, 50, size=N + 1).astype(np.uint64)
        if t % 3 == 0:
            h[: rng.integers(0, N)] = 0
        if t % 4 == 1:
            h[rng.integers(0, N):] = rng.integers(0, 2)
        if t % 5 == 2:
            h[:] = 1
        yield h


def test_host_weight_pieces_match_oracle(orc):
    rng = np.random.default_rng(42)
    N = 64
    for h in _random_hists(rng, N):
        assert M.hist_range(h) == orc.hist_range(h)
        a, b = M.kull_leibler(h), orc.kull_leibler(h)
        assert (np.isnan(a) and np.isnan(b)) or a == b
        lnw0 = rng.normal(size=N + 1) * 10
        x, y = lnw0.copy(), lnw0.copy()
        M.update_trivial(h, x)
        orc.update_trivial(h, y)
        assert np.array_equal(x, y)
        s, e = M.hist_range(h)
        pa, pb = rng.random(N + 1), None
        pb = pa.copy()
        x, y = lnw0.copy(), lnw0.copy()
        M.update_recursive(h, s, e, pa, x)
        orc.weight_recursion(h, pb, s, e, y)
        assert np.array_equal(x, y) and np.array_equal(pa, pb)
        M.normalise(x)
        orc.normalise(y)
        assert np.array_equal(x, y)


def test_jackknife_matches_oracle(orc):
    rng = np.random.default_rng(0)
    for J in (2, 3, 10):
        x = rng.normal(size=J)
        assert M.jackknife(x) == orc.jackknife(x)


def test_bad_arguments_raise():
    with pytest.raises(M.MucaError):
        M.update_recursive(np.ones(5, np.uint64), 3, 1, np.zeros(5), np.zeros(5))
    with pytest.raises(M.MucaError):
        M.jackknife(np.zeros(0))


@pytest.mark.parametrize("kw,msg", [
    (dict(observables=7), "observables must be"),
    (dict(precision=M.REF32, observables=M.OBS_MOMENTS), "MUCA_FP64 mode only"),
    (dict(precision=M.FP64, observables=M.OBS_MOMENTS, kernel=M.KERNEL_LANE),
     "generic and team kernels only"),
])
def test_moments_mode_arguments_checked_before_any_device_call(kw, msg):
    """muca_create validates the produce-path options before it touches a
    device, so these fail the same way with or without a GPU."""
    with pytest.raises(M.MucaError, match=msg):
        M.Context(16, 64, keying=M.KEYING_PRODUCE, joint_blocks=2, **kw)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it says so and exits 2
    before touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr and p.stdout == ""


def test_batch_nupdates_is_64_bit():
    """The seeded protocol's NUPDATES at L = 128 (24 N^2 = 6.4e9, 16 N^2 =
    2^32) reach the library unchanged instead of wrapping to 32 bits."""
    from muca_amd import surrogate as S
    for factor in (16, 24):
        nup = S.equilibrium_nupdates(128, factor)
        arr = M.pack_instances([dict(L=128, walkers=4096, seed=1, init_nup=nup)])
        assert arr[0].init_nup == nup and nup >= 1 << 32


@pytest.mark.parametrize("field,value", [("init_nup", M.MAX_NUPDATES + 1), ("init_nup", -1),
                                         ("walkers", 1 << 32), ("seed", -5), ("L", 2.5)])
def test_batch_rejects_out_of_range_fields_before_packing(field, value):
    q = dict(L=16, walkers=16, seed=1, init_nup=1)
    q[field] = value
    with pytest.raises(M.MucaError):
        M.pack_instances([q])


def test_batch_create_rejects_nupdates_beyond_max_before_any_device_call():
    """muca_batch_create checks every instance before it touches a device, so
    this runs (and fails with MUCA_E_RANGE) without a GPU."""
    import ctypes
    arr = (M.Instance * 1)()
    arr[0] = M.Instance(16, 16, 1, 0, M.MAX_NUPDATES + 1, 2.0, 0.0, 0)
    h = ctypes.c_void_p()
    rc = M.lib().muca_batch_create(0, arr, 1, M.REF32, M.KERNEL_AUTO, ctypes.byref(h))
    assert rc == -6 and not h.value
    assert b"MUCA_MAX_NUPDATES" in M.lib().muca_batch_last_error()
    arr[0] = M.Instance(16, 16, 1, 0, 1, 2.0, 0.0, M.MAX_NUPDATES + 1)
    assert M.lib().muca_batch_create(0, arr, 1, M.REF32, M.KERNEL_AUTO, ctypes.byref(h)) == -6
    arr[0] = M.Instance(16, 16, 1, 0, 1, 2.0, 0.5, 0)
    assert M.lib().muca_batch_create(0, arr, 1, M.REF32, M.KERNEL_AUTO, ctypes.byref(h)) == -1
    arr[0] = M.Instance(7, 16, 1, 0, 1, 2.0, 0.0, 0)
    assert M.lib().muca_batch_create(0, arr, 1, M.REF32, M.KERNEL_AUTO, ctypes.byref(h)) == -1
