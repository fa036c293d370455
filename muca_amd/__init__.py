"""muca_amd -- MI355X-native multicanonical Blume-Capel engine.

Python view of libmuca.so (include/muca.h) through ctypes.  The library
must be built in-tree first (``python -m muca_amd.build`` or
``__graft_entry__.build()``); importing any engine entry point without it
raises -- there is no CPU fallback.

The reference's own boundary is its two CLIs (muca_delta_multigpu.cu,
produce2_multigpu.cu); muca_amd.bin holds drop-in builds of them.  This module
exposes the same steps to Python: Context (one device, one walker set),
run() for the `mcs` kernel, and the O(N) host pieces of the weight loop.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmuca.so")
BIN_DIR = os.path.join(PKG, "bin")

REF32, FP64 = 0, 1
KEYING_WEIGHT, KEYING_PRODUCE = 0, 1
KERNEL_AUTO, KERNEL_GENERIC, KERNEL_TEAM, KERNEL_LANE = 0, 1, 2, 3
OBS_JOINT, OBS_MOMENTS = 0, 1
SEED = 100                       # muca_delta_multigpu.cu:19

_u8p = ctypes.POINTER(ctypes.c_int8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


class MucaError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint32), ("walkers", ctypes.c_uint32),
                ("device", ctypes.c_int32), ("rounds", ctypes.c_uint32),
                ("beta", ctypes.c_double), ("precision", ctypes.c_int32),
                ("keying", ctypes.c_int32), ("joint_blocks", ctypes.c_uint32),
                ("kernel", ctypes.c_int32), ("block", ctypes.c_uint32),
                ("observables", ctypes.c_int32)]


class Info(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint32), ("N", ctypes.c_uint32), ("walkers", ctypes.c_uint32),
                ("nbins", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("kernel", ctypes.c_int32), ("cus", ctypes.c_uint32),
                ("block", ctypes.c_uint32), ("lattice_bytes", ctypes.c_uint64),
                ("joint_bins", ctypes.c_uint64), ("lds_per_block", ctypes.c_uint32),
                ("lds_lattice", ctypes.c_uint32)]


MAX_NUPDATES = 1 << 48           # include/muca.h MUCA_MAX_NUPDATES


class Instance(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint32), ("walkers", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("recursion", ctypes.c_int32),
                ("init_nup", ctypes.c_uint64), ("beta", ctypes.c_double),
                ("nup_growth", ctypes.c_double), ("nup_max", ctypes.c_uint64)]


class InstanceStatus(ctypes.Structure):
    _fields_ = [("iter", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("converged", ctypes.c_uint32), ("reach_iter", ctypes.c_uint32),
                ("done_iter", ctypes.c_uint32), ("start", ctypes.c_int32),
                ("end", ctypes.c_int32), ("nup_capped", ctypes.c_uint32),
                ("nup", ctypes.c_uint64), ("dkl", ctypes.c_double),
                ("attempts", ctypes.c_uint64)]


def pack_instances(instances):
    """ctypes array of muca_instance; every field is range-checked first, since
    ctypes would wrap an out-of-range integer silently."""
    ins = list(instances)
    arr = (Instance * len(ins))()
    for i, q in enumerate(ins):
        vals = {"L": q["L"], "walkers": q["walkers"], "seed": q["seed"],
                "init_nup": q.get("init_nup", 1), "nup_max": q.get("nup_max", 0)}
        for k, v in vals.items():
            hi = MAX_NUPDATES if k in ("init_nup", "nup_max") else 0xFFFFFFFF
            if int(v) != v or not 0 <= v <= hi:
                raise MucaError(f"instance {i}: {k} = {v} is outside [0, {hi}]")
        arr[i] = Instance(int(q["L"]), int(q["walkers"]), int(q["seed"]),
                          1 if q.get("recursion") else 0, int(vals["init_nup"]),
                          float(q.get("beta", 2.0)), float(q.get("nup_growth", 0.0)),
                          int(vals["nup_max"]))
    return arr


_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32

# every symbol declared in include/muca.h, with its ctypes signature
_SIGS = {
    "muca_batch_last_error": (ctypes.c_char_p, []),
    "muca_batch_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(Instance), _u32,
                                         ctypes.c_int32, ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "muca_batch_destroy": (ctypes.c_int, [_vp]),
    "muca_batch_query": (ctypes.c_int, [_vp, _i32p, ctypes.POINTER(_u32), ctypes.POINTER(_u32),
                                        ctypes.POINTER(_u32), _u64p]),
    "muca_batch_iterate_profiled": (ctypes.c_int, [_vp, _u32, _f32p, _f32p]),
    "muca_batch_autotune": (ctypes.c_int, [_vp, _u32, _i32p, ctypes.POINTER(_u32), _f32p]),
    "muca_batch_init": (ctypes.c_int, [_vp, _u32]),
    "muca_batch_init_stratified": (ctypes.c_int, [_vp, _u32]),
    "muca_init_lattice_stratified": (ctypes.c_int, [_vp, _u32, _u32, _u32]),
    "muca_batch_set_log_weight": (ctypes.c_int, [_vp, _u32, _f64p, _u32]),
    "muca_batch_get_log_weight": (ctypes.c_int, [_vp, _u32, _f64p, _u32]),
    "muca_batch_get_hist": (ctypes.c_int, [_vp, _u32, _u64p, _u32]),
    "muca_batch_get_lattice": (ctypes.c_int, [_vp, _u32, _u8p, ctypes.c_uint64]),
    "muca_batch_get_state": (ctypes.c_int, [_vp, _u32, _i32p, _i32p, _i32p]),
    "muca_batch_iterate": (ctypes.c_int, [_vp, _u32, ctypes.c_int32, _f32p]),
    "muca_batch_status": (ctypes.c_int, [_vp, ctypes.POINTER(InstanceStatus), _u32]),
    "muca_mlp_last_error": (ctypes.c_char_p, []),
    "muca_mlp_param_count": (ctypes.c_uint32, []),
    "muca_mlp_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "muca_mlp_destroy": (ctypes.c_int, [_vp]),
    "muca_mlp_set_params": (ctypes.c_int, [_vp, _f32p, _u32]),
    "muca_mlp_get_params": (ctypes.c_int, [_vp, _f32p, _u32]),
    "muca_mlp_predict": (ctypes.c_int, [_vp, _f32p, _u32, _f32p]),
    "muca_mlp_gradient": (ctypes.c_int, [_vp, _f32p, _f32p, _u32, _f32p, _f32p]),
    "muca_mlp_set_data": (ctypes.c_int, [_vp, _f32p, _f32p, _f32p, _u32]),
    "muca_mlp_train": (ctypes.c_int, [_vp, _u32, _u32, ctypes.c_float, ctypes.c_float, _f32p,
                                      _f32p]),
    "muca_last_error": (ctypes.c_char_p, []),
    "muca_default_walkers": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32)]),
    "muca_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p)]),
    "muca_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "muca_query": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Info)]),
    "muca_init_lattice": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32]),
    "muca_set_lattice": (ctypes.c_int, [ctypes.c_void_p, _u8p, ctypes.c_uint64]),
    "muca_get_lattice": (ctypes.c_int, [ctypes.c_void_p, _u8p, ctypes.c_uint64]),
    "muca_compute_state": (ctypes.c_int, [ctypes.c_void_p]),
    "muca_get_state": (ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p, _i32p]),
    "muca_set_state": (ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p, _i32p]),
    "muca_set_log_weight": (ctypes.c_int, [ctypes.c_void_p, _f64p, ctypes.c_uint32]),
    "muca_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint64, ctypes.c_uint64, _u64p, _f32p]),
    "muca_get_hist": (ctypes.c_int, [ctypes.c_void_p, _u64p, ctypes.c_uint32, _f32p]),
    "muca_set_test_hist_preload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "muca_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _i32p,
                                     ctypes.POINTER(ctypes.c_uint32), _f32p]),
    "muca_get_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _u64p,
                                      ctypes.c_uint64]),
    "muca_get_moments": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _u64p,
                                        ctypes.c_uint64]),
    "muca_hist_range": (ctypes.c_int, [_u64p, ctypes.c_uint32, _i32p, _i32p]),
    "muca_kull_leibler": (ctypes.c_double, [_u64p, ctypes.c_uint32]),
    "muca_update_trivial": (ctypes.c_int, [_u64p, ctypes.c_uint32, _f64p]),
    "muca_update_recursive": (ctypes.c_int, [_u64p, ctypes.c_uint32, ctypes.c_int32,
                                             ctypes.c_int32, _f64p, _f64p]),
    "muca_normalise": (ctypes.c_int, [ctypes.c_uint32, _f64p]),
    "muca_jackknife": (ctypes.c_int, [_f64p, ctypes.c_uint32, _f64p, _f64p]),
    "muca_reweight": (ctypes.c_int, [ctypes.c_void_p, _f64p, ctypes.c_uint32, _f64p, _f64p,
                                     _f64p]),
    "muca_reweight_noalloc": (ctypes.c_int, [ctypes.c_void_p, _f64p, ctypes.c_uint32, _f64p,
                                             _f64p, _f64p]),
    "muca_acclog_host": (ctypes.c_float, [ctypes.c_uint32]),
    "muca_acclog_host_range": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, _f32p]),
    "muca_acclog_device": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32,
                                          _f32p]),
    "muca_acclog_fast_device": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint32,
                                               ctypes.c_uint32, _f32p]),
}

ACCLOG_FAST_REL = 2.0 ** -20     # screening band of the fast log (acclog.h)
ACCLOG_FAST_ABS = 2.0 ** -36

_lib = None


def lib():
    """Load libmuca.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MucaError(f"{LIB_PATH} is not built; run `python -m muca_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MucaError(f"libmuca error {rc}: {lib().muca_last_error().decode()}")


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def default_walkers(device=0):
    w = ctypes.c_uint32()
    _check(lib().muca_default_walkers(device, ctypes.byref(w)))
    return w.value


class Context:
    """One device and one set of walkers (the per-GPU state of the reference's
    main(), muca_delta_multigpu.cu:177-205 / produce2_multigpu.cu:247-275)."""

    def __init__(self, L, walkers=0, device=0, *, rounds=7, beta=2.0, precision=REF32,
                 keying=KEYING_WEIGHT, joint_blocks=0, kernel=KERNEL_AUTO, block=0,
                 observables=OBS_JOINT):
        self._h = ctypes.c_void_p()
        p = Params(L, walkers, device, rounds, beta, precision, keying, joint_blocks, kernel,
                   block, observables)
        _check(lib().muca_create(ctypes.byref(p), ctypes.byref(self._h)))
        info = Info()
        _check(lib().muca_query(self._h, ctypes.byref(info)))
        self.info = info
        self.L, self.N, self.W = info.L, info.N, info.walkers
        self.nbins = info.nbins
        self.joint_blocks = joint_blocks

    def close(self):
        if self._h:
            lib().muca_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init_lattice(self, mm=0, gpu=0, seed=SEED, stratified=False):
        fn = lib().muca_init_lattice_stratified if stratified else lib().muca_init_lattice
        _check(fn(self._h, mm, gpu, seed))

    def set_lattice(self, site_major):
        a = np.ascontiguousarray(site_major, dtype=np.int8)
        _check(lib().muca_set_lattice(self._h, _p(a, _u8p), a.size))

    def get_lattice(self):
        """Site-major lattice, shape (N, walkers)."""
        a = np.empty((self.N, self.W), np.int8)
        _check(lib().muca_get_lattice(self._h, _p(a, _u8p), a.size))
        return a

    def compute_state(self):
        _check(lib().muca_compute_state(self._h))

    def get_state(self):
        E, D, M = (np.empty(self.W, np.int32) for _ in range(3))
        _check(lib().muca_get_state(self._h, _p(E, _i32p), _p(D, _i32p), _p(M, _i32p)))
        return E, D, M

    def set_state(self, E=None, D=None, M=None):
        a = [None if x is None else np.ascontiguousarray(x, np.int32) for x in (E, D, M)]
        _check(lib().muca_set_state(self._h, *[_p(x, _i32p) for x in a]))

    def set_log_weight(self, lnw):
        a = np.ascontiguousarray(lnw, np.float64)
        _check(lib().muca_set_log_weight(self._h, _p(a, _f64p), a.size))

    def run(self, iter, seed, therm, nupdates, want_hist=True):
        """One update launch; returns (hist uint64[N+1] or None, kernel ms)."""
        h = np.empty(self.nbins, np.uint64) if want_hist else None
        ms = ctypes.c_float()
        _check(lib().muca_run(self._h, iter & 0xFFFFFFFF, seed, therm, nupdates, _p(h, _u64p),
                              ctypes.byref(ms)))
        return h, ms.value

    def set_test_hist_preload(self, preload):
        """Test hook: LDS histogram bins of the team and lane kernels start at
        `preload` (so the u32 wrap carry runs with few counts)."""
        _check(lib().muca_set_test_hist_preload(self._h, preload))

    def autotune(self, attempts):
        """Time every eligible (kernel, workgroup) on a scratch copy of the
        walkers and keep the fastest; returns (kernel, block, ms)."""
        k, b, ms = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_float()
        _check(lib().muca_autotune(self._h, attempts, ctypes.byref(k), ctypes.byref(b),
                                   ctypes.byref(ms)))
        self.kernel, self.block = k.value, b.value
        return k.value, b.value, ms.value

    def get_joint(self, block=0):
        n = (self.N + 1) * (self.N + 2) // 2
        a = np.empty(n, np.uint64)
        _check(lib().muca_get_joint(self._h, block, _p(a, _u64p), n))
        return a

    def get_moments(self, block=0):
        """Exact per-D moments of one walker block (OBS_MOMENTS) as Python
        integers: arrays S0, S1, S2, S4 over D (object dtype, 128-bit sums)."""
        n = 6 * (self.N + 1)
        a = np.empty(n, np.uint64)
        _check(lib().muca_get_moments(self._h, block, _p(a, _u64p), n))
        w = a.reshape(6, self.N + 1).astype(object)
        return w[0], w[1], w[2] + (w[3] << 64), w[4] + (w[5] << 64)

    def reweight(self, deltas):
        """(<|M|>, chi, Binder) arrays of shape (len(deltas), joint_blocks)."""
        d = np.ascontiguousarray(deltas, np.float64)
        shp = (d.size, self.joint_blocks)
        out = [np.empty(shp, np.float64) for _ in range(3)]
        _check(lib().muca_reweight(self._h, _p(d, _f64p), d.size, *[_p(o, _f64p) for o in out]))
        return out


def _bcheck(rc):
    if rc != 0:
        raise MucaError(f"libmuca batch error {rc}: {lib().muca_batch_last_error().decode()}")


class Batch:
    """Many independent weight problems advanced together on one device (each
    the weight loop of multigpu_weight for one jack,
    muca_delta_multigpu.cu:314-476); the loop body runs on the device.

    instances: iterable of dicts with keys L, walkers, seed, and optionally
    beta (2.0), init_nup (1, at most MAX_NUPDATES), recursion (False),
    nup_growth (0: the reference's 1.01) and nup_max (0: MAX_NUPDATES) of the
    NUPDATES schedule once the histogram spans all bins."""

    def __init__(self, instances, device=0, *, precision=REF32, tiles=KERNEL_AUTO):
        ins = list(instances)
        arr = pack_instances(ins)
        self._h = ctypes.c_void_p()
        _bcheck(lib().muca_batch_create(device, arr, len(ins), precision, tiles,
                                        ctypes.byref(self._h)))
        self.n = len(ins)
        self.N = [q["L"] * q["L"] for q in ins]
        self.W = [q["walkers"] for q in ins]
        self._query()

    def _query(self):
        k, bl, nt, g, w = ctypes.c_int32(), _u32(), _u32(), _u32(), ctypes.c_uint64()
        _bcheck(lib().muca_batch_query(self._h, ctypes.byref(k), ctypes.byref(bl),
                                       ctypes.byref(nt), ctypes.byref(g), ctypes.byref(w)))
        self.tiles, self.block, self.ntiles = k.value, bl.value, nt.value
        self.grid, self.walkers = g.value, w.value

    def close(self):
        if self._h:
            lib().muca_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init(self, mm=0, stratified=False):
        fn = lib().muca_batch_init_stratified if stratified else lib().muca_batch_init
        _bcheck(fn(self._h, mm))

    def set_log_weight(self, i, lnw):
        a = np.ascontiguousarray(lnw, np.float64)
        _bcheck(lib().muca_batch_set_log_weight(self._h, i, _p(a, _f64p), a.size))

    def get_log_weight(self, i):
        a = np.empty(self.N[i] + 1, np.float64)
        _bcheck(lib().muca_batch_get_log_weight(self._h, i, _p(a, _f64p), a.size))
        return a

    def get_hist(self, i):
        a = np.empty(self.N[i] + 1, np.uint64)
        _bcheck(lib().muca_batch_get_hist(self._h, i, _p(a, _u64p), a.size))
        return a

    def get_lattice(self, i):
        """Site-major lattice of instance i, shape (N, walkers)."""
        a = np.empty((self.N[i], self.W[i]), np.int8)
        _bcheck(lib().muca_batch_get_lattice(self._h, i, _p(a, _u8p), a.size))
        return a

    def get_state(self, i):
        E, D, Mg = (np.empty(self.W[i], np.int32) for _ in range(3))
        _bcheck(lib().muca_batch_get_state(self._h, i, _p(E, _i32p), _p(D, _i32p),
                                           _p(Mg, _i32p)))
        return E, D, Mg

    def iterate(self, iterations=1, use_graph=True):
        """Returns the elapsed device milliseconds."""
        ms = ctypes.c_float()
        _bcheck(lib().muca_batch_iterate(self._h, iterations, 1 if use_graph else 0,
                                         ctypes.byref(ms)))
        return ms.value

    def iterate_profiled(self, iterations=1):
        """Eager launches with HIP events around each: returns (summed update
        kernel ms, summed device weight-update ms)."""
        u, w = ctypes.c_float(), ctypes.c_float()
        _bcheck(lib().muca_batch_iterate_profiled(self._h, iterations, ctypes.byref(u),
                                                  ctypes.byref(w)))
        return u.value, w.value

    def autotune(self, mm=0):
        """Pick the fastest tile kind / workgroup size; leaves the batch
        initialised with mm, keeping each instance's ln W as it was before the
        call.  Returns (kind, block, ms of one iteration)."""
        k, bl, ms = ctypes.c_int32(), _u32(), ctypes.c_float()
        _bcheck(lib().muca_batch_autotune(self._h, mm, ctypes.byref(k), ctypes.byref(bl),
                                          ctypes.byref(ms)))
        self._query()
        return k.value, bl.value, ms.value

    def status(self):
        arr = (InstanceStatus * self.n)()
        _bcheck(lib().muca_batch_status(self._h, arr, self.n))
        return [{f: getattr(s, f) for f, _ in InstanceStatus._fields_} for s in arr]


def _mcheck(rc):
    if rc != 0:
        raise MucaError(f"libmuca mlp error {rc}: {lib().muca_mlp_last_error().decode()}")


class Mlp:
    """The weight surrogate network on one device (include/muca.h, muca_mlp_*)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _mcheck(lib().muca_mlp_create(device, ctypes.byref(self._h)))
        self.nparams = lib().muca_mlp_param_count()

    def close(self):
        if self._h:
            lib().muca_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, p):
        a = np.ascontiguousarray(p, np.float32)
        _mcheck(lib().muca_mlp_set_params(self._h, _p(a, _f32p), a.size))

    def get_params(self):
        a = np.empty(self.nparams, np.float32)
        _mcheck(lib().muca_mlp_get_params(self._h, _p(a, _f32p), a.size))
        return a

    def predict(self, x):
        xa = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        y = np.empty(len(xa), np.float32)
        _mcheck(lib().muca_mlp_predict(self._h, _p(xa, _f32p), len(xa), _p(y, _f32p)))
        return y

    def gradient(self, x, t):
        xa = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        ta = np.ascontiguousarray(t, np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = ctypes.c_float()
        _mcheck(lib().muca_mlp_gradient(self._h, _p(xa, _f32p), _p(ta, _f32p), len(xa),
                                        _p(g, _f32p), ctypes.byref(loss)))
        return g, loss.value

    def set_data(self, x, t, w=None):
        xa = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        ta = np.ascontiguousarray(t, np.float32)
        wa = None if w is None else np.ascontiguousarray(w, np.float32)
        _mcheck(lib().muca_mlp_set_data(self._h, _p(xa, _f32p), _p(ta, _f32p), _p(wa, _f32p),
                                        len(xa)))

    def train(self, steps, batch, lr, lr_final):
        """Returns (last minibatch loss, device ms)."""
        loss, ms = ctypes.c_float(), ctypes.c_float()
        _mcheck(lib().muca_mlp_train(self._h, steps, batch, lr, lr_final, ctypes.byref(loss),
                                     ctypes.byref(ms)))
        return loss.value, ms.value


# ---- host pieces of the weight iteration (muca_delta_multigpu.cu:431-476) ----
def hist_range(hist):
    h = np.ascontiguousarray(hist, np.uint64)
    s, e = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().muca_hist_range(_p(h, _u64p), h.size - 1, ctypes.byref(s), ctypes.byref(e)))
    return s.value, e.value


def kull_leibler(hist):
    h = np.ascontiguousarray(hist, np.uint64)
    return lib().muca_kull_leibler(_p(h, _u64p), h.size - 1)


def update_trivial(hist, lnw):
    h = np.ascontiguousarray(hist, np.uint64)
    _check(lib().muca_update_trivial(_p(h, _u64p), h.size - 1, _p(lnw, _f64p)))


def update_recursive(hist, start, end, p_acc, lnw):
    h = np.ascontiguousarray(hist, np.uint64)
    _check(lib().muca_update_recursive(_p(h, _u64p), h.size - 1, start, end, _p(p_acc, _f64p),
                                       _p(lnw, _f64p)))


def normalise(lnw):
    _check(lib().muca_normalise(lnw.size - 1, _p(lnw, _f64p)))


def jackknife(x):
    a = np.ascontiguousarray(x, np.float64)
    avg, err = ctypes.c_double(), ctypes.c_double()
    _check(lib().muca_jackknife(_p(a, _f64p), a.size, ctypes.byref(avg), ctypes.byref(err)))
    return avg.value, err.value


def acclog_host(r):
    return lib().muca_acclog_host(r)


def acclog_host_range(k0, n):
    out = np.empty(n, np.float32)
    _check(lib().muca_acclog_host_range(k0, n, _p(out, _f32p)))
    return out


def acclog_device(k0, n, device=0):
    out = np.empty(n, np.float32)
    _check(lib().muca_acclog_device(device, k0, n, _p(out, _f32p)))
    return out


def acclog_fast_device(k0, n, device=0):
    out = np.empty(n, np.float32)
    _check(lib().muca_acclog_fast_device(device, k0, n, _p(out, _f32p)))
    return out


def weight_iteration(ctx, iters, *, seed=SEED, init_nup=1, recursion=False, lnw_init=None,
                     callback=None, max_converged_stop=True):
    """The weight loop of multigpu_weight (muca_delta_multigpu.cu:374-477) on one
    device: update launch with iter=k, width/D_KL bookkeeping, NUPDATES
    schedule, trivial or recursive update, float copy.  Returns the final
    double ln W and a list of per-iteration dicts."""
    N = ctx.N
    lnw = np.zeros(N + 1) if lnw_init is None else np.array(lnw_init, np.float64)
    ctx.set_log_weight(lnw)
    p_acc = np.zeros(N + 1)
    width, nup, recs = 10, init_nup, []
    for k in range(iters):
        if width == N:
            nup = int((nup + 1) * 1.01)
        hist, ms = ctx.run(k, seed, N, nup)
        s, e = hist_range(hist)
        wt = (e - s) & 0xFFFFFFFF
        width = max(width, wt)
        dkl = kull_leibler(hist)
        conv = width == N and dkl < 1e-4
        rec = dict(k=k, hist=hist, width_temp=wt, width=width, dkl=dkl, nup=nup,
                   kernel_ms=ms, converged=conv)
        if not conv:
            if recursion:
                update_recursive(hist, s, e, p_acc, lnw)
            else:
                update_trivial(hist, lnw)
            normalise(lnw)
            ctx.set_log_weight(lnw)
        rec["lnw"] = lnw.copy()
        recs.append(rec)
        if callback:
            callback(rec)
        if conv and max_converged_stop:
            break
    return lnw, recs
