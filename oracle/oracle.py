"""ctypes wrapper of the FP64 CPU oracle (liboracle_sdpd.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it.

Arrays are numpy, particle-major in Python (shape (N, dim)); they are
handed to C as the reference's column-major (dim, N) layout, which is the
same memory.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_sdpd.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("num_dim", C.c_int32),
        ("symmetry", C.c_int32),
        ("kernel_type", C.c_int32),
        ("rhs_density_type", C.c_int32),
        ("rhs_force_type", C.c_int32),
        ("eos_type", C.c_int32),
        ("newtonian", C.c_int32),
        ("brownian", C.c_int32),
        ("body_force_type", C.c_int32),
        ("pad0", C.c_int32),
        ("min_phys", C.c_double * 3),
        ("max_phys", C.c_double * 3),
        ("cut_off", C.c_double),
        ("c", C.c_double),
        ("rho", C.c_double),
        ("rho_ref", C.c_double),
        ("gamma", C.c_double),
        ("eta", C.c_double),
        ("eta_coef", C.c_double),
        ("kt", C.c_double),
        ("dt", C.c_double),
        ("tau", C.c_double),
        ("n_p", C.c_double),
        ("kt_p", C.c_double),
        ("body_force", C.c_double * 3),
        ("seed", C.c_uint64),
        ("le_dim", C.c_int32),
        ("pad1", C.c_int32),
        ("le_vlo", C.c_double * 3),
        ("le_vhi", C.c_double * 3),
        ("le_time", C.c_double),
        ("coll", C.c_void_p),
    ]


class OrcColloids(C.Structure):
    _fields_ = [
        ("num", C.c_int32),
        ("noslip", C.c_int32),
        ("rho_type", C.c_int32),
        ("translate", C.c_int32),
        ("rotate", C.c_int32),
        ("ab", C.c_int32),
        ("sub_steps", C.c_int32),
        ("body_force_type", C.c_int32),
        ("body_force", C.c_double * 3),
        ("dout", C.c_double),
        ("le_time", C.c_double),
        ("step_start", C.c_int64),
    ] + [(name, C.c_void_p) for name in (
        "radius", "m", "mmi", "x", "v", "omega", "f", "alpha", "rot_matrix", "acc_matrix",
        "theta", "drag", "torque")]


class Colloids:
    """Rigid colloids for the oracle: numpy arrays that orc_vv_step updates
    in place (vectors stored with 3 components per colloid, histories of
    `ab` levels newest first)."""

    ARRAYS = ("radius", "m", "mmi", "x", "v", "omega", "f", "alpha", "rot_matrix", "acc_matrix",
              "theta", "drag", "torque")

    def __init__(self, radius, m, mmi, x, v=None, omega=None, noslip=2, rho_type=0,
                 translate=False, rotate=False, ab=1, sub_steps=1, body_force_type=0,
                 body_force=(0.0, 0.0, 0.0), dout=0.05, acc_matrix=None, step_start=0):
        k = len(radius)
        self.num = k
        self.noslip, self.rho_type = int(noslip), int(rho_type)
        self.translate, self.rotate = bool(translate), bool(rotate)
        self.ab, self.sub_steps = int(ab), int(sub_steps)
        self.body_force_type = int(body_force_type)
        self.body_force = [float(b) for b in list(body_force) + [0.0] * (3 - len(body_force))][:3]
        self.dout = float(dout)
        self.step_start = int(step_start)
        self.le_time = 0.0

        def vec(a):
            out = np.zeros((k, 3))
            if a is not None:
                a = np.asarray(a, np.float64).reshape(k, -1)
                out[:, : a.shape[1]] = a
            return out
        self.radius = np.ascontiguousarray(radius, np.float64)
        self.m = np.ascontiguousarray(m, np.float64)
        self.mmi = vec(mmi)
        self.x = vec(x)
        self.v = np.zeros((self.ab, k, 3))
        self.v[0] = vec(v)
        self.omega = np.zeros((self.ab, k, 3))
        self.omega[0] = vec(omega)
        self.f = np.zeros((self.ab, k, 3))
        self.alpha = np.zeros((self.ab, k, 3))
        self.rot_matrix = np.zeros((k, 3, 3))
        self.acc_matrix = np.tile(np.eye(3), (k, 1, 1)) if acc_matrix is None else \
            np.array(acc_matrix, np.float64).reshape(k, 3, 3).copy()
        self.theta = np.zeros((k, 3))
        self.drag = np.zeros((k, 3))
        self.torque = np.zeros((k, 3))

    def struct(self) -> OrcColloids:
        c = OrcColloids()
        c.num = self.num
        c.noslip, c.rho_type = self.noslip, self.rho_type
        c.translate, c.rotate = int(self.translate), int(self.rotate)
        c.ab, c.sub_steps, c.body_force_type = self.ab, self.sub_steps, self.body_force_type
        for d in range(3):
            c.body_force[d] = self.body_force[d]
        c.dout, c.le_time, c.step_start = self.dout, float(self.le_time), self.step_start
        for name in self.ARRAYS:
            a = getattr(self, name)
            assert a.flags["C_CONTIGUOUS"] and a.dtype == np.float64
            setattr(c, name, a.ctypes.data)
        # acc_matrix / rot_matrix are (k, 3, 3) row-major = column-major
        # (3, 3, k) of the transposes: stored as the C code reads them,
        # element (r, c) at [r + 3 c]
        return c


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU needed), also when
    the library is up to date."""
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        P = C.POINTER(OrcParams)
        L.orc_kernel.argtypes = [P, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_rank.argtypes = [C.c_int32, dp, C.c_int32, dp, dp, ip, ip, ip, ip, C.POINTER(C.c_int32)]
        L.orc_cell_counts.argtypes = [P, ip]
        L.orc_ghost_get.argtypes = [P, dp, C.c_int32, dp, ip, C.c_int32, C.POINTER(C.c_int32)]
        L.orc_neighbor_pairs.argtypes = [P, dp, C.c_int32, ip, C.c_int64, C.POINTER(C.c_int64)]
        L.orc_brute_pairs.argtypes = [P, dp, C.c_int32, ip, C.c_int64, C.POINTER(C.c_int64)]
        L.orc_density.argtypes = [P, dp, dp, ip, C.c_int32, dp]
        L.orc_pressure.argtypes = [P, dp, dp, C.c_int32, dp]
        L.orc_interaction.argtypes = [P, C.c_int64, dp, dp, dp, dp, dp, ip, ip, C.c_void_p, C.c_int32, dp, C.c_void_p]
        L.orc_compute_act.argtypes = [P, dp, dp, C.c_int32, dp]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_pair_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.POINTER(C.c_float)]
        L.orc_pair_normals.restype = None
        fp = C.POINTER(C.c_float)
        L.orc_noise_log_n.argtypes = [fp, C.c_int64, fp]
        L.orc_noise_log_n.restype = None
        L.orc_noise_sincos_n.argtypes = [C.POINTER(C.c_uint32), C.c_int64, fp, fp]
        L.orc_noise_sincos_n.restype = None
        up = C.POINTER(C.c_uint32)
        L.orc_pair_normals_n.argtypes = [C.c_uint64, up, up, C.c_int64, C.c_int64, fp]
        L.orc_pair_normals_n.restype = None
        L.orc_philox4x32_10.restype = None
        state = [P, C.c_int64, C.c_int32, dp, dp, dp, ip, ip, dp, dp, dp, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_prologue.argtypes = state
        L.orc_vv_step.argtypes = state
        L.orc_statistics.argtypes = [P, dp, dp, ip, C.c_int32, C.POINTER(C.c_double), dp]
        L.orc_colloid_nearest.argtypes = [P, C.c_double, dp, C.c_int32, dp, dp, dp]
        L.orc_colloid_noslip.argtypes = [P, dp, dp, dp, C.c_int32, dp]
        L.orc_colloid_collect.argtypes = [P, dp, dp, dp, ip, C.c_int32]
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_set_noise_table.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
        L.orc_vv_begin.argtypes = state
        L.orc_vv_finish.argtypes = state
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# injected noise: a check of a generator other than the oracle's own hands
# the oracle that generator's normals for the pairs of each evaluation
# --------------------------------------------------------------------------
_noise_source = None


def set_noise_source(fn) -> None:
    """fn(seed, lo, hi, step) -> (n, 4) float32 normals of the pairs
    (lo[k], hi[k]), lo < hi pids; None: the oracle's own counter-based
    stream.  While set, every force evaluation with thermal noise takes its
    normals from fn for exactly the pairs inside the cut-off."""
    global _noise_source
    _noise_source = fn


class _Injected:
    """Install the source's normals for the pairs of positions x at a step
    for the duration of one evaluation."""

    def __init__(self, prm, x, pid, step, le_time=0.0):
        self.active = (_noise_source is not None and bool(prm.Brownian)
                       and float(prm.kt) > 1e-14)
        if not self.active:
            return
        pairs = neighbor_pairs(prm, x, le_time=le_time)
        pid = np.asarray(pid, np.int64)
        a, b = pid[pairs[:, 0]], pid[pairs[:, 1]]
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        keys = np.unique((lo.astype(np.uint64) << np.uint64(32)) | hi.astype(np.uint64))
        if keys.size == 0:  # no pair inside the cut-off: no noise to draw
            self.active = False
            return
        self.keys = np.ascontiguousarray(keys, np.uint64)
        lo_k = (self.keys >> np.uint64(32)).astype(np.uint32)
        hi_k = (self.keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        g = _noise_source(int(prm.random_seed), lo_k, hi_k, int(step))
        self.normals = np.ascontiguousarray(g, np.float32).reshape(-1, 4)
        assert self.normals.shape[0] == self.keys.shape[0]
        self.step = int(step)

    def __enter__(self):
        if self.active:
            _check(lib().orc_set_noise_table(self.step, self.keys.shape[0],
                                             self.keys.ctypes.data, self.normals.ctypes.data),
                   "set_noise_table")
        return self

    def __exit__(self, *exc):
        if self.active:
            lib().orc_set_noise_table(0, 0, None, None)


def use_library(path: str) -> None:
    """Load the oracle from another build of the same source (the
    benchmark's natively tuned CPU baseline); before the first call."""
    global _lib, LIB_PATH
    _lib = None
    LIB_PATH = path


def set_threads(n: int) -> None:
    """Threads of the density and force traversals (1: the reference's
    serial order, the default; more: a colouring of the home cells whose
    sums accumulate in another order -- for timing only)."""
    _check(lib().orc_set_threads(int(n)), "set_threads")


_keep = []  # colloid records referenced by the last OrcParams built


def to_orc(prm, le_time: float = 0.0, coll: "Colloids | None" = None) -> OrcParams:
    """Convert an SDPDParams-like object into the oracle's record; le_time
    is the time the Lees-Edwards shear length is evaluated at; coll the
    colloids of the box, if any."""
    o = OrcParams()
    o.num_dim = int(prm.num_dim)
    o.symmetry = 1 if prm.symmetry else 0
    o.kernel_type = int(prm.kernel_type)
    o.rhs_density_type = int(prm.rhs_density_type)
    o.rhs_force_type = int(prm.rhs_force_type)
    o.eos_type = int(prm.stateEquation_type)
    o.newtonian = 1 if prm.Newtonian else 0
    o.brownian = 1 if prm.Brownian else 0
    o.body_force_type = int(prm.body_force_type)
    for d in range(3):
        o.min_phys[d] = float(prm.min_phys[d])
        o.max_phys[d] = float(prm.max_phys[d])
        o.body_force[d] = float(prm.body_force[d])
    o.cut_off = float(prm.cut_off)
    o.c = float(prm.c)
    o.rho = float(prm.rho)
    o.rho_ref = float(prm.rho_ref)
    o.gamma = float(prm.gamma)
    o.eta = float(prm.eta)
    o.eta_coef = float(prm.eta_coef)
    o.kt = float(prm.kt) if prm.Brownian else 0.0
    o.dt = float(prm.dt)
    o.tau = float(prm.tau)
    o.n_p = float(prm.n_p)
    o.kt_p = float(prm.kt_p)
    o.seed = int(prm.random_seed) & 0xFFFFFFFFFFFFFFFF
    o.le_dim = int(getattr(prm, "le_dim", 0))
    for d in range(3):
        o.le_vlo[d] = float(getattr(prm, "le_v_lower", [0.0] * 3)[d])
        o.le_vhi[d] = float(getattr(prm, "le_v_upper", [0.0] * 3)[d])
    o.le_time = float(le_time)
    if coll is not None:
        rec = coll.struct()
        _keep[:] = [rec]
        o.coll = C.addressof(rec)
    return o


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed (stat_info={rc})")


def _f64(a, dim=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def kernel(prm, r: float):
    w = C.c_double()
    g = C.c_double()
    _check(lib().orc_kernel(C.byref(to_orc(prm)), float(r), C.byref(w), C.byref(g)), "kernel")
    return w.value, g.value


def cell_counts(prm):
    nm = np.zeros(3, np.int32)
    lib().orc_cell_counts(C.byref(to_orc(prm)), nm)
    return nm


def rank(dim, x, xmin, xmax, nm, ngl):
    x = _f64(x)
    n = x.shape[0]
    nmt = [int(nm[d]) + int(ngl[d]) + int(ngl[dim + d]) for d in range(dim)]
    nbox = int(np.prod(nmt))
    lhbx = np.zeros(nbox + 1, np.int32)
    lpdx = np.zeros(max(n, 1), np.int32)
    cnt = C.c_int32()
    _check(lib().orc_rank(dim, x, n, _f64(xmin), _f64(xmax), _i32(nm), _i32(ngl), lhbx, lpdx,
                          C.byref(cnt)), "rank")
    return lhbx, lpdx[: cnt.value]


def ghosts(prm, x, le_time=0.0):
    x = _f64(x)
    n, dim = x.shape
    cap = 8 * n + 16
    xg = np.zeros((cap, dim))
    owner = np.zeros(cap, np.int32)
    ng = C.c_int32()
    _check(lib().orc_ghost_get(C.byref(to_orc(prm, le_time)), x, n, xg, owner, cap, C.byref(ng)),
           "ghost_get")
    return xg[: ng.value].copy(), owner[: ng.value].copy()


def neighbor_pairs(prm, x, brute=False, le_time=0.0):
    x = _f64(x)
    n = x.shape[0]
    cap = max(1024, n * 200)
    pairs = np.zeros(2 * cap, np.int32)
    np_ = C.c_int64()
    fn = lib().orc_brute_pairs if brute else lib().orc_neighbor_pairs
    _check(fn(C.byref(to_orc(prm, le_time)), x, n, pairs, cap, C.byref(np_)), "pairs")
    return pairs[: 2 * np_.value].reshape(-1, 2).copy()


def density(prm, x, m, sid=None, le_time=0.0, coll=None):
    x = _f64(x)
    n = x.shape[0]
    sid = np.zeros(n, np.int32) if sid is None else _i32(sid)
    rho = np.zeros(n)
    _check(lib().orc_density(C.byref(to_orc(prm, le_time, coll)), x, _f64(m), sid, n, rho), "density")
    return rho


def pressure(prm, rho, m):
    rho = _f64(rho)
    p = np.zeros_like(rho)
    _check(lib().orc_pressure(C.byref(to_orc(prm)), rho, _f64(m), rho.shape[0], p), "pressure")
    return p


def interaction(prm, step, x, v, rho, press, m, pid, sid=None, ct=None, le_time=0.0, coll=None):
    x = _f64(x)
    n, dim = x.shape
    sid = np.zeros(n, np.int32) if sid is None else _i32(sid)
    f = np.zeros((n, dim))
    vgt = np.zeros((n, dim * dim))
    ctp = None
    if ct is not None:
        ct = _f64(ct)
        ctp = ct.ctypes.data_as(C.c_void_p)
    with _Injected(prm, x, pid, step, le_time):
        _check(lib().orc_interaction(C.byref(to_orc(prm, le_time, coll)), int(step), x, _f64(v),
                                     _f64(rho), _f64(press), _f64(m), _i32(pid), sid, ctp, n, f,
                                     vgt.ctypes.data_as(C.c_void_p)), "interaction")
    return f, vgt


def colloid_nearest(prm, coll, t, xp, sid):
    """Nearest image of colloid sid's centre to xp: (xim, rx, vim)."""
    xim, rx, vim = np.zeros(3), np.zeros(3), np.zeros(3)
    xp = np.concatenate([np.asarray(xp, np.float64), np.zeros(3)])[:3]
    _check(lib().orc_colloid_nearest(C.byref(to_orc(prm, 0.0, coll)), float(t), xp, int(sid), xim, rx,
                                     vim), "colloid_nearest")
    return xim, rx, vim


def colloid_noslip(prm, coll, xf, xc, vf, sid, le_time=0.0):
    """Velocity of boundary particle xc (colloid sid) seen by fluid
    particle xf moving at vf."""
    vc = np.zeros(3)
    pad = lambda a: np.concatenate([np.asarray(a, np.float64), np.zeros(3)])[:3]  # noqa: E731
    _check(lib().orc_colloid_noslip(C.byref(to_orc(prm, le_time, coll)), pad(xf), pad(xc), pad(vf),
                                    int(sid), vc), "colloid_noslip")
    return vc[: prm.num_dim]


def colloid_collect(prm, coll, x, f, m, sid):
    """Drag and torque of every colloid into coll.drag / coll.torque."""
    x = _f64(x)
    _check(lib().orc_colloid_collect(C.byref(to_orc(prm, 0.0, coll)), x, _f64(f), _f64(m), _i32(sid),
                                     x.shape[0]), "colloid_collect")
    return coll.drag.copy(), coll.torque.copy()


def compute_act(prm, ct, vgt):
    ct = _f64(ct)
    act = np.zeros_like(ct)
    _check(lib().orc_compute_act(C.byref(to_orc(prm)), ct, _f64(vgt), ct.shape[0], act), "act")
    return act


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def pair_normals(seed, pid_lo, pid_hi, step):
    g = (C.c_float * 4)()
    lib().orc_pair_normals(int(seed) & 0xFFFFFFFFFFFFFFFF, int(pid_lo), int(pid_hi), int(step), g)
    return np.array([g[i] for i in range(4)], np.float32)


def pair_normals_n(seed, lo, hi, step):
    """pair_normals over arrays of keys: (n, 4) float32."""
    lo = np.ascontiguousarray(lo, np.uint32)
    hi = np.ascontiguousarray(hi, np.uint32)
    g = np.empty((lo.size, 4), np.float32)
    up = C.POINTER(C.c_uint32)
    lib().orc_pair_normals_n(int(seed) & 0xFFFFFFFFFFFFFFFF, lo.ctypes.data_as(up),
                             hi.ctypes.data_as(up), int(step), lo.size,
                             g.ctypes.data_as(C.POINTER(C.c_float)))
    return g


def noise_log(u):
    """The noise stream's ln u over a float32 array (u in (0, 1])."""
    u = np.ascontiguousarray(u, np.float32)
    out = np.empty_like(u)
    fp = C.POINTER(C.c_float)
    lib().orc_noise_log_n(u.ctypes.data_as(fp), u.size, out.ctypes.data_as(fp))
    return out


def noise_sincos(k):
    """The noise stream's (cos, sin) of 2 pi k / 2^24 over a uint32 array."""
    k = np.ascontiguousarray(k, np.uint32)
    cs = np.empty(k.shape, np.float32)
    sn = np.empty(k.shape, np.float32)
    fp = C.POINTER(C.c_float)
    lib().orc_noise_sincos_n(k.ctypes.data_as(C.POINTER(C.c_uint32)), k.size,
                             cs.ctypes.data_as(fp), sn.ctypes.data_as(fp))
    return cs, sn


class State:
    """Particle state for multi-step oracle runs (particle-major arrays)."""

    def __init__(self, prm, x, v, m, pid, sid=None, ct=None, coll=None):
        self.prm = prm
        self.coll = coll
        self.x = np.array(x, np.float64, copy=True)
        n, dim = self.x.shape
        self.v = np.array(v, np.float64, copy=True)
        self.m = np.ascontiguousarray(m, np.float64)
        self.pid = _i32(pid)
        self.sid = np.zeros(n, np.int32) if sid is None else _i32(sid)
        self.rho = np.zeros(n)
        self.p = np.zeros(n)
        self.f = np.zeros((n, dim))
        d2 = dim * dim
        if ct is None:
            ct = np.tile(np.eye(dim).reshape(-1, order="F"), (n, 1))
        self.ct = np.array(ct, np.float64, copy=True).reshape(n, d2)
        self.act = np.zeros((n, d2))
        self.vgt = np.zeros((n, d2))
        self.time = 0.0  # time_current (accumulated by repeated addition of dt)
        self.le_last = 0.0  # time of the last shear-length update

    def _args(self):
        vp = C.c_void_p
        return (self.rho, self.p, self.f, self.ct.ctypes.data_as(vp), self.act.ctypes.data_as(vp),
                self.vgt.ctypes.data_as(vp))

    def prologue(self, step=0):
        # marching_marching.F90:437: the boundary is updated at the current
        # time before the initial evaluation
        n = self.x.shape[0]
        self.le_last = self.time
        with _Injected(self.prm, self.x, self.pid, step, self.time):
            _check(lib().orc_prologue(C.byref(to_orc(self.prm, self.time, self.coll)), int(step), n,
                                      self.x, self.v, self.m, self.pid, self.sid, *self._args()),
                   "prologue")

    def vv_step(self, step):
        # marching_marching.F90:1392-1404 advances time_current before the
        # step; marching_integrate_VV.F90:784-786 evaluates the shear length
        # at time_current + dt
        # the colloid block runs before the step's boundary update, with the
        # shear length of the previous one (marching_integrate_VV.F90:455-797)
        n = self.x.shape[0]
        self.time = self.time + self.prm.dt
        if self.coll is not None:
            self.coll.le_time = self.le_last
        self.le_last = self.time + self.prm.dt
        rec = to_orc(self.prm, self.le_last, self.coll)
        p = C.byref(rec)
        args = (int(step), n, self.x, self.v, self.m, self.pid, self.sid) + self._args()
        if _noise_source is None:
            _check(lib().orc_vv_step(p, *args), "vv_step")
            return
        # the step's noise for the pairs at the positions after the drift
        _check(lib().orc_vv_begin(p, *args), "vv_begin")
        with _Injected(self.prm, self.x, self.pid, step, self.le_last):
            _check(lib().orc_vv_finish(p, *args), "vv_finish")

    def statistics(self):
        n = self.x.shape[0]
        e = C.c_double()
        mom = np.zeros(3)
        _check(lib().orc_statistics(C.byref(to_orc(self.prm, 0.0, self.coll)), self.v, self.m, self.sid, n,
                                    C.byref(e), mom), "statistics")
        return e.value, mom[: self.x.shape[1]]
