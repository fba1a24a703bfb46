"""Host-side mirror of the reference's per-step interface, over the C-ABI.

The reference's boundary is the set of module procedures that
marching_integrate_VV calls on TYPE(Particles) and TYPE(Technique)
(mcf/src/Class_Particles.F90:43-311, Class_Technique.F90:40-97).  The
class below keeps those names and their argument meaning:

    technique_build_list            -> build_list()
    particles_compute_density       -> compute_density()
    particles_compute_pressure      -> compute_pressure()
    particles_compute_interaction   -> compute_interaction()
    particles_integrate_velocity    -> integrate_velocity(lambda_)
    particles_integrate_position +
    particles_adjust_particles      -> integrate_position_and_adjust()
    particles_integrate_ct          -> integrate_ct(lambda_)
    statistic_compute_statistic     -> compute_statistic()
    marching_integrate_VV           -> integrate_VV(nsteps)

Errors: the reference returns stat_info = -1 and prints; here a non-zero
return code raises SDPDError carrying the library's message.  Everything
runs in libsdpd.so on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .params import SDPDParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdpd.so")


class SDPDError(RuntimeError):
    pass


class CParams(C.Structure):
    """Mirror of sdpd_params in include/sdpd.h."""

    _fields_ = [
        ("num_dim", C.c_int32),
        ("symmetry", C.c_int32),
        ("kernel_type", C.c_int32),
        ("rhs_density_type", C.c_int32),
        ("rhs_force_type", C.c_int32),
        ("state_equation_type", C.c_int32),
        ("newtonian", C.c_int32),
        ("brownian", C.c_int32),
        ("body_force_type", C.c_int32),
        ("precision", C.c_int32),
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
    ]


class CTuning(C.Structure):
    """struct sdpd_tuning of include/sdpd.h."""
    _fields_ = [("density_shape", C.c_int32), ("chunk_order", C.c_int32)]


class CColloidParams(C.Structure):
    """struct sdpd_colloid_params of include/sdpd.h."""
    _fields_ = [
        ("noslip", C.c_int32),
        ("rho_type", C.c_int32),
        ("translate", C.c_int32),
        ("rotate", C.c_int32),
        ("integrate_ab", C.c_int32),
        ("sub_steps", C.c_int32),
        ("body_force_type", C.c_int32),
        ("pad", C.c_int32),
        ("body_force", C.c_double * 3),
        ("dout", C.c_double),
    ]


# (name, restype, argtypes) of every symbol include/sdpd.h declares
_P = C.c_void_p
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_UP = C.POINTER(C.c_uint32)
_LP = C.POINTER(C.c_int64)
_FP = C.POINTER(C.c_float)
SYMBOLS = [
    ("sdpd_create", C.c_int, [C.POINTER(CParams), C.POINTER(C.c_void_p)]),
    ("sdpd_upload", C.c_int, [_P, _DP, _DP, _DP, _IP, C.c_int32]),
    ("sdpd_set_replicas", C.c_int, [_P, C.c_int32, _DP, _DP]),
    ("sdpd_upload_replicas", C.c_int, [_P, _DP, _DP, _DP, _IP, _IP, C.c_int32]),
    ("sdpd_upload_ct", C.c_int, [_P, _DP]),
    ("sdpd_build_cells", C.c_int, [_P]),
    ("sdpd_density", C.c_int, [_P]),
    ("sdpd_eos", C.c_int, [_P]),
    ("sdpd_forces", C.c_int, [_P]),
    ("sdpd_kick", C.c_int, [_P, C.c_double]),
    ("sdpd_drift_wrap", C.c_int, [_P]),
    ("sdpd_conformation", C.c_int, [_P, C.c_double]),
    ("sdpd_stats", C.c_int, [_P, _DP, _DP]),
    ("sdpd_stats_replicas", C.c_int, [_P, _DP, _DP]),
    ("sdpd_download", C.c_int, [_P, _DP, _DP, _DP, _DP, _DP]),
    ("sdpd_download_tensors", C.c_int, [_P, _DP, _DP]),
    ("sdpd_step", C.c_int, [_P, C.c_int32]),
    ("sdpd_set_lees_edwards", C.c_int, [_P, C.c_int32, _DP, _DP]),
    ("sdpd_set_colloids", C.c_int, [_P, C.POINTER(CColloidParams), C.c_int32, _DP, _DP, _DP, _DP,
                                    _DP, _DP]),
    ("sdpd_get_colloids", C.c_int, [_P, _DP, _DP, _DP, _DP, _DP]),
    ("sdpd_set_graphs", C.c_int, [_P, C.c_int32]),
    ("sdpd_set_list_capacity", C.c_int, [_P, C.c_int32]),
    ("sdpd_get_list_capacity", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("sdpd_set_skin", C.c_int, [_P, C.c_double]),
    ("sdpd_get_skin", C.c_int, [_P, _DP]),
    ("sdpd_get_list_builds", C.c_int, [_P, _LP]),
    ("sdpd_set_tuning", C.c_int, [_P, C.POINTER(CTuning)]),
    ("sdpd_get_tuning", C.c_int, [_P, C.POINTER(CTuning)]),
    ("sdpd_autotune", C.c_int, [_P, C.c_int32, C.POINTER(CTuning), _DP]),
    ("sdpd_set_step", C.c_int, [_P, C.c_int64]),
    ("sdpd_get_step", C.c_int, [_P, _LP]),
    ("sdpd_get_cells", C.c_int, [_P, _IP, _IP, _IP]),
    ("sdpd_get_counters", C.c_int, [_P, _LP, _LP]),
    ("sdpd_neighbor_signature", C.c_int, [_P, C.c_int32, _IP, _UP]),
    ("sdpd_pair_normals", C.c_int, [_UP, _UP, C.c_int32, C.c_uint64, C.c_int64, _FP]),
    ("sdpd_set_timing", C.c_int, [_P, C.c_int32]),
    ("sdpd_get_phase_times", C.c_int, [_P, _DP, _LP]),
    ("sdpd_sync", C.c_int, [_P]),
    ("sdpd_evaluate_noalloc", C.c_int, [C.POINTER(CParams), C.c_int64, _P, _P, _P, _P, C.c_int32,
                                        _P, _P, _P, _P]),
    ("sdpd_last_error", C.c_char_p, [_P]),
    ("sdpd_destroy", C.c_int, [_P]),
]

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libsdpd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SDPDError(
            f"{path} is missing: build the HIP library first (python -m mcf_amd.build)")
    lib = C.CDLL(path)
    for name, res, args in SYMBOLS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def to_c(prm: SDPDParams) -> CParams:
    c = CParams()
    c.num_dim = int(prm.num_dim)
    c.symmetry = 1 if prm.symmetry else 0
    c.kernel_type = int(prm.kernel_type)
    c.rhs_density_type = int(prm.rhs_density_type)
    c.rhs_force_type = int(prm.rhs_force_type)
    c.state_equation_type = int(prm.stateEquation_type)
    c.newtonian = 1 if prm.Newtonian else 0
    c.brownian = 1 if prm.Brownian else 0
    c.body_force_type = int(prm.body_force_type)
    c.precision = int(prm.precision)
    for d in range(3):
        c.min_phys[d] = float(prm.min_phys[d])
        c.max_phys[d] = float(prm.max_phys[d])
        c.body_force[d] = float(prm.body_force[d])
    c.cut_off = float(prm.cut_off)
    c.c = float(prm.c)
    c.rho = float(prm.rho)
    c.rho_ref = float(prm.rho_ref)
    c.gamma = float(prm.gamma)
    c.eta = float(prm.eta)
    c.eta_coef = float(prm.eta_coef)
    c.kt = float(prm.kt)
    c.dt = float(prm.dt)
    c.tau = float(prm.tau)
    c.n_p = float(prm.n_p)
    c.kt_p = float(prm.kt_p)
    c.seed = int(prm.random_seed) & 0xFFFFFFFFFFFFFFFF
    return c


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


class SDPD:
    """One box of particles on the GPU (one sdpd_ctx)."""

    def __init__(self, prm: SDPDParams):
        self.lib = load_library()
        self.prm = prm
        self._cp = to_c(prm)
        h = C.c_void_p()
        rc = self.lib.sdpd_create(C.byref(self._cp), C.byref(h))
        if rc != 0 or not h.value:
            raise SDPDError(f"sdpd_create failed (stat_info={rc}); see stderr")
        self.h = h
        self.n = 0
        self.ncoll = 0
        if getattr(prm, "le_dim", 0):
            self.set_lees_edwards(prm.le_dim, prm.le_v_lower, prm.le_v_upper)

    # -- lifetime -------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.sdpd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self.lib.sdpd_last_error(self.h)
            raise SDPDError(f"{what}: stat_info={rc}: {msg.decode() if msg else ''}")

    # -- data -----------------------------------------------------------
    def upload(self, x, v, m, pid, sid=None) -> None:
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        m = np.ascontiguousarray(m, np.float64)
        n = x.shape[0]
        ids = np.zeros((n, 2), np.int32)
        ids[:, 0] = pid
        if sid is not None:
            ids[:, 1] = sid
        self._check(self.lib.sdpd_upload(self.h, _dp(x), _dp(v), _dp(m), _ip(ids), n), "sdpd_upload")
        self.n = n

    def set_replicas(self, box_lo, box_hi) -> None:
        """Batched ensemble of independent periodic boxes; box_lo, box_hi:
        (nrep, 3) lower and upper corners."""
        lo = np.ascontiguousarray(box_lo, np.float64).reshape(-1, 3)
        hi = np.ascontiguousarray(box_hi, np.float64).reshape(-1, 3)
        if lo.shape != hi.shape:
            raise ValueError("box_lo and box_hi differ in shape")
        self._check(self.lib.sdpd_set_replicas(self.h, lo.shape[0], _dp(lo), _dp(hi)),
                    "sdpd_set_replicas")
        self.nrep = lo.shape[0]
        self.n = 0

    def upload_replicas(self, x, v, m, pid, rep) -> None:
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        m = np.ascontiguousarray(m, np.float64)
        rep = np.ascontiguousarray(rep, np.int32)
        n = x.shape[0]
        ids = np.zeros((n, 2), np.int32)
        ids[:, 0] = pid
        self._check(self.lib.sdpd_upload_replicas(self.h, _dp(x), _dp(v), _dp(m), _ip(ids), _ip(rep), n),
                    "sdpd_upload_replicas")
        self.n = n

    def compute_statistic_replicas(self):
        """(ekin[nrep], mom[nrep, dim]) of every replica."""
        nrep = getattr(self, "nrep", 1)
        mom = np.zeros((nrep, 3))
        ek = np.zeros(nrep)
        self._check(self.lib.sdpd_stats_replicas(self.h, _dp(mom), _dp(ek)), "sdpd_stats_replicas")
        return ek, mom[:, : self.prm.num_dim]

    def upload_ct(self, ct) -> None:
        ct = np.ascontiguousarray(ct, np.float64)
        self._check(self.lib.sdpd_upload_ct(self.h, _dp(ct)), "sdpd_upload_ct")

    def download(self):
        dim = self.prm.num_dim
        x = np.zeros((self.n, dim))
        v = np.zeros((self.n, dim))
        self._check(self.lib.sdpd_download(self.h, _dp(x), _dp(v), None, None, None), "sdpd_download")
        return x, v

    def download_density(self):
        rho = np.zeros(self.n)
        p = np.zeros(self.n)
        self._check(self.lib.sdpd_download(self.h, None, None, _dp(rho), _dp(p), None), "sdpd_download")
        return rho, p

    def download_force(self):
        f = np.zeros((self.n, self.prm.num_dim))
        self._check(self.lib.sdpd_download(self.h, None, None, None, None, _dp(f)), "sdpd_download")
        return f

    def download_tensors(self, vgt=True):
        d2 = self.prm.num_dim ** 2
        ct = np.zeros((self.n, d2))
        g = np.zeros((self.n, d2)) if vgt else None
        self._check(self.lib.sdpd_download_tensors(self.h, _dp(ct), _dp(g) if vgt else None),
                    "sdpd_download_tensors")
        return ct, g

    # -- the reference's per-step procedures ------------------------------
    def build_list(self) -> None:
        self._check(self.lib.sdpd_build_cells(self.h), "technique_build_list")

    def compute_density(self) -> None:
        self._check(self.lib.sdpd_density(self.h), "particles_compute_density")

    def compute_pressure(self) -> None:
        self._check(self.lib.sdpd_eos(self.h), "particles_compute_pressure")

    def compute_interaction(self) -> None:
        self._check(self.lib.sdpd_forces(self.h), "particles_compute_interaction")

    def integrate_velocity(self, lambda_: float) -> None:
        self._check(self.lib.sdpd_kick(self.h, float(lambda_)), "particles_integrate_velocity")

    def integrate_position_and_adjust(self) -> None:
        self._check(self.lib.sdpd_drift_wrap(self.h), "particles_integrate_position")

    def integrate_ct(self, lambda_: float) -> None:
        self._check(self.lib.sdpd_conformation(self.h, float(lambda_)), "particles_integrate_ct")

    def compute_statistic(self):
        mom = np.zeros(3)
        e = C.c_double()
        self._check(self.lib.sdpd_stats(self.h, _dp(mom), C.byref(e)), "statistic_compute_statistic")
        return e.value, mom[: self.prm.num_dim]

    def set_lees_edwards(self, le_dim: int, v_lower, v_upper) -> None:
        """Lees-Edwards faces on dimension le_dim (2 = y, 3 = z; 0 = off)
        moving with v_lower / v_upper (3-vectors, x components only)."""
        lo = np.ascontiguousarray(list(v_lower) + [0.0] * (3 - len(v_lower)), np.float64)
        hi = np.ascontiguousarray(list(v_upper) + [0.0] * (3 - len(v_upper)), np.float64)
        self._check(self.lib.sdpd_set_lees_edwards(self.h, int(le_dim), _dp(lo), _dp(hi)),
                    "sdpd_set_lees_edwards")

    def set_colloids(self, cs) -> None:
        """Rigid colloids (a mcf_amd.workloads.ColloidSet: radius, m, mmi,
        x, v, omega and the coll_* parameters); call before upload, whose
        sid then names each boundary particle's colloid (1..num)."""
        k = cs.num
        cp = CColloidParams()
        cp.noslip, cp.rho_type = int(cs.noslip), int(cs.rho_type)
        cp.translate, cp.rotate = int(bool(cs.translate)), int(bool(cs.rotate))
        cp.integrate_ab, cp.sub_steps = int(cs.ab), int(cs.sub_steps)
        cp.body_force_type = int(cs.body_force_type)
        for d in range(3):
            cp.body_force[d] = float(cs.body_force[d])
        cp.dout = float(cs.dout)
        arr = lambda a, w: np.ascontiguousarray(np.asarray(a, np.float64).reshape(k, w))  # noqa: E731
        radius, m = arr(cs.radius, 1), arr(cs.m, 1)
        mmi, x, v, w = arr(cs.mmi, 3), arr(cs.x, 3), arr(cs.v, 3), arr(cs.omega, 3)
        self._check(self.lib.sdpd_set_colloids(self.h, C.byref(cp), k, _dp(radius), _dp(m), _dp(mmi),
                                               _dp(x), _dp(v), _dp(w)), "sdpd_set_colloids")
        self.ncoll = k

    def colloids(self):
        """(x, v, omega, drag, torque) of the colloids, each (num, 3)."""
        k = self.ncoll
        out = [np.zeros((k, 3)) for _ in range(5)]
        self._check(self.lib.sdpd_get_colloids(self.h, *[_dp(a) for a in out]), "sdpd_get_colloids")
        return tuple(out)

    def set_graphs(self, enable: bool) -> None:
        """Run sdpd_step's steps through captured HIP graphs (default) or
        launch every kernel directly."""
        self._check(self.lib.sdpd_set_graphs(self.h, 1 if enable else 0), "sdpd_set_graphs")

    def tuning(self):
        """(density_shape, chunk_order) in use."""
        t = CTuning()
        self._check(self.lib.sdpd_get_tuning(self.h, C.byref(t)), "sdpd_get_tuning")
        return int(t.density_shape), int(t.chunk_order)

    def set_list_capacity(self, cap: int) -> None:
        """Stored neighbour entries per particle (0: automatic); before upload."""
        self._check(self.lib.sdpd_set_list_capacity(self.h, int(cap)), "sdpd_set_list_capacity")

    def list_capacity(self) -> int:
        cap = C.c_int32()
        self._check(self.lib.sdpd_get_list_capacity(self.h, C.byref(cap)), "sdpd_get_list_capacity")
        return int(cap.value)

    def set_skin(self, skin: float) -> None:
        """Verlet skin (< 0: automatic, 0: lists rebuilt every step); before
        upload."""
        self._check(self.lib.sdpd_set_skin(self.h, float(skin)), "sdpd_set_skin")

    def skin(self) -> float:
        s = C.c_double()
        self._check(self.lib.sdpd_get_skin(self.h, C.byref(s)), "sdpd_get_skin")
        return s.value

    def list_builds(self) -> int:
        """Verlet list builds since upload (-1: lists rebuilt every step)."""
        b = C.c_int64()
        self._check(self.lib.sdpd_get_list_builds(self.h, C.byref(b)), "sdpd_get_list_builds")
        return b.value

    def set_tuning(self, density_shape: int, chunk_order: int) -> None:
        t = CTuning(int(density_shape), int(chunk_order))
        self._check(self.lib.sdpd_set_tuning(self.h, C.byref(t)), "sdpd_set_tuning")

    def autotune(self, reps: int = 3):
        """Time every launch shape on the current state, apply the fastest;
        returns ((density_shape, chunk_order), ms per density + force pass)."""
        t = CTuning()
        ms = C.c_double()
        self._check(self.lib.sdpd_autotune(self.h, int(reps), C.byref(t), C.byref(ms)), "sdpd_autotune")
        return (int(t.density_shape), int(t.chunk_order)), ms.value

    def integrate_VV(self, nsteps: int = 1) -> None:
        self._check(self.lib.sdpd_step(self.h, int(nsteps)), "marching_integrate_VV")

    def evaluate(self) -> None:
        """Initial evaluation of marching_marching (cells, density, forces)."""
        self.build_list()
        self.compute_density()
        self.compute_interaction()

    # -- helpers ----------------------------------------------------------
    @property
    def step(self) -> int:
        s = C.c_int64()
        self._check(self.lib.sdpd_get_step(self.h, C.byref(s)), "sdpd_get_step")
        return s.value

    @step.setter
    def step(self, value: int) -> None:
        self._check(self.lib.sdpd_set_step(self.h, int(value)), "sdpd_set_step")

    def cells(self):
        nm = np.zeros(3, np.int32)
        self._check(self.lib.sdpd_get_cells(self.h, _ip(nm), None, None), "sdpd_get_cells")
        ncell = int(np.prod(nm))
        lhbx = np.zeros(ncell + 1, np.int32)
        lpdx = np.zeros(self.n, np.int32)
        self._check(self.lib.sdpd_get_cells(self.h, _ip(nm), _ip(lhbx), _ip(lpdx)), "sdpd_get_cells")
        return nm, lhbx, lpdx

    def counters(self):
        a = C.c_int64()
        b = C.c_int64()
        self._check(self.lib.sdpd_get_counters(self.h, C.byref(a), C.byref(b)), "sdpd_get_counters")
        return a.value, b.value

    def neighbor_signature(self, mixed_decision: bool = False):
        cnt = np.zeros(self.n, np.int32)
        hs = np.zeros(self.n, np.uint32)
        self._check(self.lib.sdpd_neighbor_signature(self.h, 1 if mixed_decision else 0, _ip(cnt),
                                                     hs.ctypes.data_as(_UP)),
                    "sdpd_neighbor_signature")
        return cnt, hs

    PHASES = ("boundary", "sort", "density", "forces", "stats")

    def set_timing(self, enable: bool) -> None:
        self._check(self.lib.sdpd_set_timing(self.h, 1 if enable else 0), "sdpd_set_timing")

    def phase_times(self):
        ms = np.zeros(len(self.PHASES))
        cnt = np.zeros(len(self.PHASES), np.int64)
        self._check(self.lib.sdpd_get_phase_times(self.h, _dp(ms), cnt.ctypes.data_as(_LP)),
                    "sdpd_get_phase_times")
        return {k: (float(ms[q]), int(cnt[q])) for q, k in enumerate(self.PHASES)}

    def sync(self) -> None:
        self._check(self.lib.sdpd_sync(self.h), "sdpd_sync")


def pair_normals(lo, hi, seed: int, step: int):
    """Device evaluation of the counter-based noise stream (4 normals each)."""
    lib = load_library()
    lo = np.ascontiguousarray(lo, np.uint32)
    hi = np.ascontiguousarray(hi, np.uint32)
    g = np.zeros((lo.shape[0], 4), np.float32)
    rc = lib.sdpd_pair_normals(lo.ctypes.data_as(_UP), hi.ctypes.data_as(_UP), lo.shape[0],
                               int(seed) & 0xFFFFFFFFFFFFFFFF, int(step), g.ctypes.data_as(_FP))
    if rc != 0:
        raise SDPDError(f"sdpd_pair_normals failed (stat_info={rc})")
    return g


def evaluate_noalloc(prm: SDPDParams, step: int, x, v, m, ids, n: int, rho, p, f,
                     stream=None) -> None:
    """sdpd_evaluate_noalloc on device buffers given as addresses (ints):
    x, v (n, dim) float64, m (n) float64, ids (n, 2) int32, outputs rho,
    p (n) float64 or None, f (n, dim) float64; queued on `stream` (a
    hipStream_t address, None for the null stream), not waited for."""
    lib = load_library()
    cp = to_c(prm)
    rc = lib.sdpd_evaluate_noalloc(C.byref(cp), int(step), x, v, m, ids, int(n), rho, p, f, stream)
    if rc != 0:
        raise SDPDError(f"sdpd_evaluate_noalloc failed (stat_info={rc}); see stderr")


def pair_hash(a, b):
    """The 32-bit pair hash of the neighbour-set signature (host side)."""
    a = np.asarray(a, np.uint32)
    b = np.asarray(b, np.uint32)
    with np.errstate(over="ignore"):
        h = (a * np.uint32(0x9E3779B1)) ^ ((b + np.uint32(0x7F4A7C15)) * np.uint32(0x85EBCA77))
        h ^= h >> np.uint32(15)
        h = h * np.uint32(0x2C1B3C6D)
        h ^= h >> np.uint32(12)
    return h.astype(np.uint32)
