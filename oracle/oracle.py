"""ctypes view of the CPU oracle (oracle/libmuca_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py -- never by the
product package.  Each wrapper mirrors one function of muca_oracle.h, which
cites the reference file:line it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmuca_oracle.so")

KEYING_WEIGHT = 0
KEYING_PRODUCE = 1
SEED_DEFAULT = 100          # muca_delta_multigpu.cu:19

_u8p = ctypes.POINTER(ctypes.c_int8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


class RunParams(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint32), ("keying", ctypes.c_int), ("fp64", ctypes.c_int),
                ("rounds", ctypes.c_uint), ("beta", ctypes.c_double),
                ("iter", ctypes.c_uint32), ("seed", ctypes.c_uint32),
                ("therm", ctypes.c_uint64), ("nupdates", ctypes.c_uint64)]


def build():
    """Compile the oracle (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_philox4x32.argtypes = [ctypes.c_uint, _u32p, _u32p, _u32p]
        L.orc_u01fixedpt_f32.restype = ctypes.c_float
        L.orc_u01fixedpt_f32.argtypes = [ctypes.c_uint32]
        L.orc_logf_u01.restype = ctypes.c_float
        L.orc_logf_u01.argtypes = [ctypes.c_uint32]
        L.orc_logf_u01_quad.restype = ctypes.c_float
        L.orc_logf_u01_quad.argtypes = [ctypes.c_uint32]
        L.orc_logf_u01_range.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f32p]
        L.orc_logf_u01_exhaustive.restype = ctypes.c_long
        L.orc_logf_u01_exhaustive.argtypes = [ctypes.c_int]
        L.orc_tables.argtypes = [ctypes.c_uint32, _u32p, _u32p]
        L.orc_init_walker.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint, _u8p]
        L.orc_init_walker_strat.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint, ctypes.c_uint32, _u8p]
        L.orc_state.argtypes = [ctypes.c_uint32, _u32p, _u8p, _i32p, _i32p, _i32p]
        L.orc_mcs_batch.argtypes = [ctypes.POINTER(RunParams), _f32p, _f64p, ctypes.c_uint32,
                                    ctypes.c_uint32, _u8p, _i32p, _i32p, _i32p, _u64p, _u64p,
                                    _u64p, ctypes.c_int]
        L.orc_moment_add.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32]
        L.orc_hist_range.argtypes = [_u64p, ctypes.c_int, _i32p, _i32p]
        L.orc_kull_leibler.restype = ctypes.c_double
        L.orc_kull_leibler.argtypes = [_u64p, ctypes.c_int]
        L.orc_update_trivial.argtypes = [ctypes.c_int, _u64p, _f64p]
        L.orc_weight_recursion.argtypes = [ctypes.c_int, _u64p, _f64p, ctypes.c_int,
                                           ctypes.c_int, _f64p]
        L.orc_normalise.argtypes = [ctypes.c_int, _f64p, _f32p]
        L.orc_jackknife.argtypes = [_f64p, ctypes.c_uint, _f64p, _f64p]
        L.orc_weight_shift.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _f64p]
        L.orc_boltz.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _f64p, _f64p]
        L.orc_observables.argtypes = [ctypes.c_double, ctypes.c_uint, _u64p, _f64p, _f64p, _f64p]
        L.orc_joint_moments.argtypes = [ctypes.c_uint, _u64p, _f64p, _f64p, _f64p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def philox4x32(rounds, ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32(rounds, _p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return [int(x) for x in o]


def u01fixedpt_f32(r):
    return lib().orc_u01fixedpt_f32(r)


def logf_u01(r):
    return lib().orc_logf_u01(r)


def logf_u01_quad(r):
    return lib().orc_logf_u01_quad(r)


def logf_u01_range(k0, n):
    out = np.zeros(n, np.float32)
    lib().orc_logf_u01_range(k0, n, _p(out, _f32p))
    return out


def logf_u01_exhaustive(nthreads):
    return lib().orc_logf_u01_exhaustive(nthreads)


def tables(L):
    N = L * L
    naver = np.zeros(4 * N, dtype=np.uint32)
    order = np.zeros(N, dtype=np.uint32)
    lib().orc_tables(L, _p(naver, _u32p), _p(order, _u32p))
    return naver, order


def init_lattice(keying, L, walkers, mm=0, gpu=0, seed=SEED_DEFAULT, rounds=7, w0=0,
                 strat_W=0):
    """Walker-major lattices [walker][site] for walkers w0..w0+walkers-1
    (strat_W > 0: the stratified start over strat_W walkers)."""
    N = L * L
    lat = np.zeros((walkers, N), dtype=np.int8)
    for j in range(walkers):
        row = lat[j]
        lib().orc_init_walker_strat(keying, L, w0 + j, mm, gpu, seed, rounds, strat_W,
                                    _p(row, _u8p))
    return lat


def state(L, lat):
    """(E, D, M) arrays for walker-major lattices."""
    naver, _ = tables(L)
    W = lat.shape[0]
    E = np.zeros(W, np.int32)
    D = np.zeros(W, np.int32)
    M = np.zeros(W, np.int32)
    e, d, m = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    for j in range(W):
        row = np.ascontiguousarray(lat[j])
        lib().orc_state(L, _p(naver, _u32p), _p(row, _u8p), ctypes.byref(e), ctypes.byref(d),
                        ctypes.byref(m))
        E[j], D[j], M[j] = e.value, d.value, m.value
    return E, D, M


def mcs(L, lat, E, D, M, *, keying, fp64, lnw32=None, lnw64=None, iter=0,
        seed=SEED_DEFAULT, therm=0, nupdates=0, beta=2.0, rounds=7, w0=0, hist=None,
        joint=None, moments=None, nthreads=1):
    """Run the per-walker chain for every walker of lat (in place).

    Returns the D histogram (uint64, N+1) summed over walkers; the joint
    histogram and the per-D moments (6 (N+1) uint64, see orc_moment_add) are
    accumulated into the arrays passed in."""
    N = L * L
    assert lat.dtype == np.int8 and lat.flags.c_contiguous and lat.shape[1] == N
    if nthreads > 1 and (joint is not None or moments is not None):
        raise ValueError("joint histograms and moments are accumulated with nthreads=1 only")
    W = lat.shape[0]
    if hist is None:
        hist = np.zeros(N + 1, dtype=np.uint64)
    p = RunParams(L, keying, int(fp64), rounds, beta, iter & 0xFFFFFFFF, seed, therm, nupdates)
    if lnw32 is not None:
        lnw32 = np.ascontiguousarray(lnw32, dtype=np.float32)
    if lnw64 is not None:
        lnw64 = np.ascontiguousarray(lnw64, dtype=np.float64)
    if M is None:
        M = np.zeros(W, np.int32)
    lib().orc_mcs_batch(ctypes.byref(p), _p(lnw32, _f32p), _p(lnw64, _f64p), w0, W,
                        _p(lat, _u8p), _p(E, _i32p), _p(D, _i32p), _p(M, _i32p),
                        _p(hist, _u64p), _p(joint, _u64p), _p(moments, _u64p), nthreads)
    return hist


def hist_range(hist):
    N = len(hist) - 1
    h = np.ascontiguousarray(hist, dtype=np.uint64)
    s, e = ctypes.c_int32(), ctypes.c_int32()
    lib().orc_hist_range(_p(h, _u64p), N, ctypes.byref(s), ctypes.byref(e))
    return s.value, e.value


def kull_leibler(hist):
    h = np.ascontiguousarray(hist, dtype=np.uint64)
    return lib().orc_kull_leibler(_p(h, _u64p), len(h) - 1)


def update_trivial(hist, lnw):
    h = np.ascontiguousarray(hist, dtype=np.uint64)
    lib().orc_update_trivial(len(h) - 1, _p(h, _u64p), _p(lnw, _f64p))


def weight_recursion(hist, p_acc, start, end, lnw):
    h = np.ascontiguousarray(hist, dtype=np.uint64)
    lib().orc_weight_recursion(len(h) - 1, _p(h, _u64p), _p(p_acc, _f64p), start, end,
                               _p(lnw, _f64p))


def normalise(lnw):
    f = np.zeros(len(lnw), np.float32)
    lib().orc_normalise(len(lnw) - 1, _p(lnw, _f64p), _p(f, _f32p))
    return f


def jackknife(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    a, e = ctypes.c_double(), ctypes.c_double()
    lib().orc_jackknife(_p(x, _f64p), len(x), ctypes.byref(a), ctypes.byref(e))
    return a.value, e.value


def weight_shift(beta, delta, lnw):
    lib().orc_weight_shift(beta, delta, len(lnw) - 1, _p(lnw, _f64p))


def boltz(beta, delta, lnw):
    b = np.zeros(len(lnw), np.float64)
    lib().orc_boltz(beta, delta, len(lnw) - 1, _p(lnw, _f64p), _p(b, _f64p))
    return b


def observables(beta, N, joint, boltz_):
    a, c = ctypes.c_double(), ctypes.c_double()
    lib().orc_observables(beta, N, _p(joint, _u64p), _p(boltz_, _f64p), ctypes.byref(a),
                          ctypes.byref(c))
    return a.value, c.value


def joint_moments(N, joint):
    S = [np.zeros(N + 1, np.float64) for _ in range(4)]
    lib().orc_joint_moments(N, _p(joint, _u64p), *[_p(s, _f64p) for s in S])
    return S


def joint_moments_exact(N, joint):
    """Per-D moments of a joint H(D,M) as exact Python integers: S0 = sum_M H,
    S1 = sum |M| H, S2 = sum M^2 H, S4 = sum M^4 H, over the bins
    (M + D^2 + 2D)/2 of produce2_multigpu.cu:489."""
    S = [[0] * (N + 1) for _ in range(4)]
    for D in range(N + 1):
        base = (D * D + D) // 2          # M = -D
        h = joint[base:base + D + 1]
        for j in np.nonzero(h)[0]:
            c = int(h[j])
            am = abs(2 * int(j) - D)
            S[0][D] += c
            S[1][D] += am * c
            S[2][D] += am * am * c
            S[3][D] += am ** 4 * c
    return S


def moments_as_int(mom, N):
    """6 (N+1) uint64 words (orc_moment_add's layout) -> S0, S1, S2, S4 as
    arrays of Python integers."""
    w = np.asarray(mom, np.uint64).reshape(6, N + 1).astype(object)
    return w[0], w[1], w[2] + (w[3] << 64), w[4] + (w[5] << 64)


def observables_moments(beta, N, S0, S1, S2, boltz_):
    """observables() (produce2_multigpu.cu:498-518) summed per D from exact
    integer moments, in long double: <|M|> = sum S1 b / sum S0 b and
    chi = beta (sum S2 b / sum S0 b - <|M|>^2) / N."""
    def ld(x):
        x = int(x)
        return np.longdouble(np.uint64(x >> 64)) * np.longdouble(2.0) ** 64 + \
            np.longdouble(np.uint64(x & 0xFFFFFFFFFFFFFFFF))
    z = m1 = m2 = np.longdouble(0)
    for D in range(N + 1):
        if S0[D]:
            b = np.longdouble(boltz_[D])
            z += ld(S0[D]) * b
            m1 += ld(S1[D]) * b
            m2 += ld(S2[D]) * b
    a = m1 / z
    return float(a), float((m2 / z - a * a) * np.longdouble(beta) / np.longdouble(N))


def weight_iteration_reference(L, walkers, iters, *, seed=SEED_DEFAULT, init_nup=1,
                               recursion=False, rounds=7, beta=2.0, nthreads=1, mm=0,
                               lnw_init=None):
    """The host loop of multigpu_weight (muca_delta_multigpu.cu:329-477) for one
    GPU and `iters` iterations, driving the oracle chain.  Returns per-iteration
    records (hist, lnw64 after update, lattice, E, D, width, dkl, nup)."""
    N = L * L
    lat = init_lattice(KEYING_WEIGHT, L, walkers, mm=mm, gpu=0, seed=seed, rounds=rounds)
    E, D, _ = state(L, lat)
    lnw = np.zeros(N + 1, np.float64) if lnw_init is None else np.array(lnw_init, np.float64)
    lnw32 = lnw.astype(np.float32)
    p_acc = np.zeros(N + 1, np.float64)
    width, nup, recs = 10, init_nup, []
    for k in range(iters):
        if width == N:
            nup = int((nup + 1) * 1.01)
        hist = mcs(L, lat, E, D, None, keying=KEYING_WEIGHT, fp64=False, lnw32=lnw32, iter=k,
                   seed=seed, therm=N, nupdates=nup, beta=beta, rounds=rounds,
                   nthreads=nthreads)
        s, e = hist_range(hist)
        wt = (e - s) & 0xFFFFFFFF
        if wt > width:
            width = wt
        dkl = kull_leibler(hist)
        converged = width == N and dkl < 1e-4
        rec = dict(k=k, hist=hist.copy(), width_temp=wt, width=width, dkl=dkl, nup=nup,
                   converged=converged)
        if not converged:
            if recursion:
                weight_recursion(hist, p_acc, s, e, lnw)
            else:
                update_trivial(hist, lnw)
            lnw32 = normalise(lnw)
        rec.update(lnw=lnw.copy(), lnw32=lnw32.copy(), lat=lat.copy(), E=E.copy(), D=D.copy())
        recs.append(rec)
        if converged:
            break
    return recs
