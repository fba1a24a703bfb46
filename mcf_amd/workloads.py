"""Synthetic initial conditions for the SDPD benchmark and test boxes.

The lattice rule follows the reference's square / simple-cubic lattice
(mcf/src/particles_init_global_inter_square.F90:133-154,
particles_init_global_inter_cubic.F90:133-163): site centres at
(i - 1/2) dx.  Each component gets a seeded uniform jitter of +-0.05 dx.
Masses follow particles_compute_mass.F90:426-450 for a lattice:
m = rho * dx^dim.  Particle ids are 1..N with species id 0 (fluid), as
assign_id gives them.

Workloads
---------
shear_wave_2d  2D periodic box [0,1]^2, 64^2 particles, quintic kernel,
               cut_off 3 dx, p = c^2 rho (Morris EOS, gamma 1, rho_ref 0),
               c = 10, eta = 0.1, u_x = 0.1 sin(2 pi y), dt = 1e-4.
tait_box_3d    3D periodic box of 128^3 particles, dx = 1 (m = 1),
               cut_off 3, Tait EOS gamma 7, c = 10, eta = 0.1,
               Maxwellian velocities at kT = 1e-3, dt = 5e-5, Brownian.
oldroyd_box_3d Oldroyd-B variant of tait_box_3d (tau 0.1, eta_p 0.05).
oldroyd_ensemble
               batch of independent periodic Oldroyd-B boxes: particle
               counts drawn log-uniformly in [2000, 60000], each a cubic
               lattice of edge round(N^(1/3)) dx (N is then that cube),
               tau 0.1, eta_p 0.05, kT 1e-3.
"""
from __future__ import annotations

import math

import numpy as np

from .params import PRECISION_FP64, PRECISION_MIXED, SDPDParams


def lattice(num_dim: int, n_per_dim, dx: float, jitter: float, seed: int,
            origin=None) -> np.ndarray:
    """Jittered square/cubic lattice, particle-major (N, dim)."""
    if np.isscalar(n_per_dim):
        n_per_dim = [int(n_per_dim)] * num_dim
    origin = [0.0] * num_dim if origin is None else origin
    axes = [origin[d] + (np.arange(n_per_dim[d]) + 0.5) * dx for d in range(num_dim)]
    # x runs fastest, like the reference's lattice loops
    grids = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([g.transpose(tuple(reversed(range(num_dim)))).ravel() for g in grids], axis=1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        pts = pts + rng.uniform(-jitter, jitter, size=pts.shape) * dx
    return np.ascontiguousarray(pts)


def wrap(x: np.ndarray, prm: SDPDParams) -> np.ndarray:
    x = x.copy()
    for d in range(prm.num_dim):
        L = prm.max_phys[d] - prm.min_phys[d]
        x[:, d] = np.where(x[:, d] < prm.min_phys[d], x[:, d] + L, x[:, d])
        x[:, d] = np.where(x[:, d] >= prm.max_phys[d], x[:, d] - L, x[:, d])
    return x


class Box:
    """A workload: parameters plus initial particle arrays."""

    def __init__(self, name, prm, x, v, m, pid, sid=None):
        self.name = name
        self.prm = prm
        self.x = np.ascontiguousarray(x, np.float64)
        self.v = np.ascontiguousarray(v, np.float64)
        self.m = np.ascontiguousarray(m, np.float64)
        self.pid = np.ascontiguousarray(pid, np.int32)
        n = self.x.shape[0]
        self.sid = np.zeros(n, np.int32) if sid is None else np.ascontiguousarray(sid, np.int32)

    @property
    def n(self) -> int:
        return self.x.shape[0]


def shear_wave_2d(n_side: int = 64, brownian: bool = False, kt: float = 1e-4,
                  seed: int = 1, symmetry: bool = True) -> Box:
    dx = 1.0 / n_side
    prm = SDPDParams(num_dim=2, min_phys=[0, 0, 0], max_phys=[1, 1, 0], cut_off=3 * dx,
                     symmetry=symmetry, kernel_type=1, rhs_density_type=2, rhs_force_type=3,
                     stateEquation_type=1, c=10.0, rho=1.0, rho_ref=0.0, gamma=1.0, eta=0.1,
                     Newtonian=True, Brownian=brownian, kt=kt if brownian else 0.0, dt=1e-4,
                     random_seed=seed)
    x = wrap(lattice(2, n_side, dx, 0.05, seed), prm)
    v = np.zeros_like(x)
    v[:, 0] = 0.1 * np.sin(2 * math.pi * x[:, 1])
    m = np.full(x.shape[0], prm.rho * dx * dx)
    pid = np.arange(1, x.shape[0] + 1, dtype=np.int32)
    return Box("shear_wave_2d", prm, x, v, m, pid)


def tait_box_3d(n_side: int = 128, precision: int = PRECISION_FP64, kt: float = 1e-3,
                brownian: bool = True, seed: int = 1, eta: float = 0.1,
                symmetry: bool = True) -> Box:
    dx = 1.0
    L = n_side * dx
    prm = SDPDParams(num_dim=3, min_phys=[0, 0, 0], max_phys=[L, L, L], cut_off=3 * dx,
                     symmetry=symmetry, kernel_type=1, rhs_density_type=2, rhs_force_type=3,
                     stateEquation_type=2, c=10.0, rho=1.0, rho_ref=1.0, gamma=7.0, eta=eta,
                     Newtonian=True, Brownian=brownian, kt=kt if brownian else 0.0, dt=5e-5,
                     random_seed=seed, precision=precision)
    x = wrap(lattice(3, n_side, dx, 0.05, seed), prm)
    m = np.full(x.shape[0], prm.rho * dx ** 3)
    rng = np.random.default_rng(seed + 7919)
    sigma = math.sqrt(max(kt, 0.0) / m[0]) if kt > 0 else 0.0
    v = rng.normal(0.0, 1.0, size=x.shape) * sigma
    v -= v.mean(axis=0, keepdims=True)
    pid = np.arange(1, x.shape[0] + 1, dtype=np.int32)
    return Box("tait_box_3d", prm, x, v, m, pid)


def oldroyd_box_3d(n_side: int = 24, precision: int = PRECISION_FP64, kt: float = 1e-3,
                   brownian: bool = True, seed: int = 1) -> Box:
    box = tait_box_3d(n_side, precision, kt, brownian, seed)
    # eta_p = G tau with G = n_p kt_p (particles_compute_pressure_tensor.F90:94)
    tau, eta_p = 0.1, 0.05
    box.prm = box.prm.replace(Newtonian=False, tau=tau, n_p=1.0, kt_p=eta_p / tau)
    box.name = "oldroyd_box_3d"
    return box


class Ensemble:
    """A batched workload: parameters shared by all replicas, one periodic
    box per replica ((nrep, 3) corners), particles concatenated with their
    replica index.  Particle ids are unique across the batch, so replicas
    draw independent thermal noise."""

    def __init__(self, name, prm, box_lo, box_hi, x, v, m, pid, rep):
        self.name = name
        self.prm = prm
        self.box_lo = np.ascontiguousarray(box_lo, np.float64)
        self.box_hi = np.ascontiguousarray(box_hi, np.float64)
        self.x = np.ascontiguousarray(x, np.float64)
        self.v = np.ascontiguousarray(v, np.float64)
        self.m = np.ascontiguousarray(m, np.float64)
        self.pid = np.ascontiguousarray(pid, np.int32)
        self.rep = np.ascontiguousarray(rep, np.int32)

    @property
    def n(self) -> int:
        return self.x.shape[0]

    @property
    def nrep(self) -> int:
        return self.box_lo.shape[0]

    def replica(self, q: int) -> Box:
        """Replica q as a single-box workload (same ids and state)."""
        sel = self.rep == q
        lo = list(self.box_lo[q]) + [0.0] * (3 - self.box_lo.shape[1])
        hi = list(self.box_hi[q]) + [0.0] * (3 - self.box_hi.shape[1])
        prm = self.prm.replace(min_phys=lo, max_phys=hi)
        return Box(f"{self.name}[{q}]", prm, self.x[sel], self.v[sel], self.m[sel], self.pid[sel])


def ensemble_of_cubes(sides, prm: SDPDParams, kt: float, seed: int = 1, name="ensemble") -> Ensemble:
    """Replicas that are jittered cubic lattices of the given edges (dx 1,
    m = rho) at the origin, Maxwellian velocities at kT with each replica's
    mean momentum removed."""
    xs, vs, reps = [], [], []
    lo = np.zeros((len(sides), 3))
    hi = np.zeros((len(sides), 3))
    rng = np.random.default_rng(seed + 104729)
    for q, side in enumerate(sides):
        side = int(side)
        hi[q] = side
        pr = prm.replace(min_phys=[0.0, 0.0, 0.0], max_phys=[float(side)] * 3)
        x = wrap(lattice(3, side, 1.0, 0.05, seed + q), pr)
        sigma = math.sqrt(kt / prm.rho) if kt > 0 else 0.0
        v = rng.normal(0.0, 1.0, size=x.shape) * sigma
        v -= v.mean(axis=0, keepdims=True)
        xs.append(x)
        vs.append(v)
        reps.append(np.full(x.shape[0], q, np.int32))
    x = np.concatenate(xs)
    v = np.concatenate(vs)
    rep = np.concatenate(reps)
    m = np.full(x.shape[0], prm.rho)
    pid = np.arange(1, x.shape[0] + 1, dtype=np.int32)
    return Ensemble(name, prm, lo, hi, x, v, m, pid, rep)


def oldroyd_ensemble(nrep: int = 256, nmin: int = 2000, nmax: int = 60000, seed: int = 1,
                     precision: int = PRECISION_FP64, kt: float = 1e-3) -> Ensemble:
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=nrep)))
    # edge round(N^(1/3)), at least 10 dx (3 cells of 1.0001 * cut_off)
    sides = np.maximum(np.round(np.cbrt(counts)), 10).astype(int)
    prm = oldroyd_box_3d(10, precision, kt).prm
    return ensemble_of_cubes(sides, prm, kt, seed, name="oldroyd_ensemble")


# --------------------------------------------------------------------------
# colloids
# --------------------------------------------------------------------------

# mcf_header.F90:131,141: inner / outer layer coefficients
COLLOID_IN_LAYER_COEFF = 0.05
COLLOID_OUT_LAYER_COEFF = 0.05


class ColloidSet:
    """Rigid colloids of a box: disks (2D) or spheres (3D), with the
    parameters of the reference's colloid block (physics_config.mcf keys
    coll_*).  Vectors carry 3 components per colloid."""

    def __init__(self, radius, x, m, mmi, v=None, omega=None, noslip=2, rho_type=0,
                 translate=False, rotate=False, ab=1, sub_steps=1, body_force_type=0,
                 body_force=(0.0, 0.0, 0.0), dout=0.05):
        k = len(radius)
        self.radius = np.ascontiguousarray(radius, np.float64)
        self.x = np.zeros((k, 3))
        self.x[:, : np.asarray(x).shape[1]] = x
        self.m = np.ascontiguousarray(m, np.float64)
        self.mmi = np.zeros((k, 3))
        self.mmi[:, : np.asarray(mmi).reshape(k, -1).shape[1]] = np.asarray(mmi).reshape(k, -1)
        self.v = np.zeros((k, 3)) if v is None else np.array(v, np.float64).reshape(k, 3)
        self.omega = np.zeros((k, 3)) if omega is None else np.array(omega, np.float64).reshape(k, 3)
        self.noslip, self.rho_type = int(noslip), int(rho_type)
        self.translate, self.rotate = bool(translate), bool(rotate)
        self.ab, self.sub_steps = int(ab), int(sub_steps)
        self.body_force_type = int(body_force_type)
        self.body_force = [float(b) for b in list(body_force) + [0.0] * 3][:3]
        self.dout = float(dout)

    @property
    def num(self) -> int:
        return int(self.radius.shape[0])


def sphere_shell(radius: float, dx: float, din: float) -> np.ndarray:
    """Boundary particles of a sphere centred at the origin, placed parallel
    to the surface in layers dx apart (coll_place 5,
    colloid_create_boundary_particle_3D_sphere.F90:160-230)."""
    pts = []
    num_layer = math.ceil(din / dx - 0.5)
    for m in range(1, num_layer + 1):
        rad = radius - m * dx + dx / 2.0
        pts.append((rad, 0.0, 0.0))
        d_phi = dx / rad
        num_phi = int(round(math.pi / d_phi))
        d_phi = math.pi / num_phi
        phi = 0.0
        for _ in range(1, num_phi):
            phi = phi + d_phi
            d_theta = dx / (rad * math.sin(phi))
            num_theta = int(round(2.0 * math.pi / d_theta))
            d_theta = 2.0 * math.pi / num_theta
            theta = 0.0
            for _k in range(num_theta):
                pts.append((rad * math.cos(phi), rad * math.sin(phi) * math.sin(theta),
                            rad * math.sin(phi) * math.cos(theta)))
                theta = theta + d_theta
        pts.append((-rad, 0.0, 0.0))
    return np.array(pts)


def disk_shell(radius: float, dx: float, din: float) -> np.ndarray:
    """Boundary particles of a disk centred at the origin, rings dx apart,
    every other ring turned by half a step (coll_place 5,
    colloid_create_boundary_particle_2D.F90:148-173)."""
    pts = []
    shift = 0
    for i in range(math.ceil(din / dx)):
        rad = radius - i * dx - dx / 2.0
        if rad >= dx / 2.0:
            angle = math.acos(1.0 - (dx / rad) ** 2 / 2.0)
            num_sur = math.ceil(2.0 * math.pi / angle)
            angle = 2.0 * math.pi / num_sur
            for k in range(num_sur):
                a = k * angle + 0.5 * shift * angle
                pts.append((rad * math.cos(a), rad * math.sin(a)))
            shift = (shift + 1) % 2
        else:
            pts.append((0.0, 0.0))
    return np.array(pts)


def nearest_dist(x: np.ndarray, c: np.ndarray, L: np.ndarray) -> np.ndarray:
    """Periodic minimum-image distance of points x (n, d) to centre c (d)."""
    r = x - c[None, :]
    r = r - L[None, :] * np.round(r / L[None, :])
    return np.sqrt((r * r).sum(1))


def with_colloids(box: Box, centres, radius: float, dx: float = None, rho_type: int = 0,
                  **kw):
    """Put rigid colloids of the given radius at `centres` (k, dim) into a
    lattice box: solvent sites within radius + dout of a centre are dropped
    (particles_init_global_assign_id.F90, colloid_check_boundary_particle.F90:
    189-205), each colloid gets its boundary-particle shell (sid = k + 1) with
    the solvent particle mass (particles_compute_mass.F90:360-470), moving
    rigidly with the colloid.  Colloid mass and moment of inertia follow
    particles_compute_mass.F90:153-357 (disk: m = rho pi R^2, I = m R^2 / 2;
    sphere: m = rho 4/3 pi R^3, I = 2/5 m R^2).  Extra keywords go to
    ColloidSet.  Returns (box, colloids)."""
    prm = box.prm
    d = prm.num_dim
    if dx is None:
        dx = (prm.max_phys[0] - prm.min_phys[0]) / round((box.n) ** (1.0 / d))
    centres = np.asarray(centres, np.float64).reshape(-1, d)
    k = centres.shape[0]
    L = np.array(prm.length[:d])
    dout = dx * COLLOID_OUT_LAYER_COEFF
    din = (1.0 + COLLOID_IN_LAYER_COEFF) * prm.cut_off * (2.0 if rho_type == 1 else 1.0)
    keep = np.ones(box.n, bool)
    reach = radius + dout
    lo = np.array(prm.min_phys[:d])
    # bin the solvent into cells at least `reach` wide and test each centre
    # against the 3^d cells around it only
    nc = np.maximum((L // reach).astype(int), 1)
    cell = np.minimum(((box.x - lo[None, :]) / L[None, :] * nc[None, :]).astype(int), nc[None, :] - 1)
    flat = np.ravel_multi_index(tuple(cell.T), tuple(nc))
    order = np.argsort(flat, kind="stable")
    bounds = np.searchsorted(flat[order], np.arange(np.prod(nc) + 1))
    offs = np.array(np.meshgrid(*[[-1, 0, 1]] * d, indexing="ij")).reshape(d, -1).T
    for c in centres:
        cc = np.minimum(((c - lo) / L * nc).astype(int), nc - 1)
        cand = []
        for o in offs:
            q = np.ravel_multi_index(tuple(np.mod(cc + o, nc)), tuple(nc))
            cand.append(order[bounds[q]:bounds[q + 1]])
        cand = np.unique(np.concatenate(cand))
        keep[cand[nearest_dist(box.x[cand], c, L) <= reach]] = False
    shell = sphere_shell(radius, dx, din) if d == 3 else disk_shell(radius, dx, din)
    xs, vs, sids = [box.x[keep]], [box.v[keep]], [np.zeros(int(keep.sum()), np.int32)]
    vcol = np.zeros((k, 3)) if kw.get("v") is None else np.array(kw["v"], np.float64).reshape(k, 3)
    wcol = np.zeros((k, 3)) if kw.get("omega") is None else np.array(kw["omega"], np.float64).reshape(k, 3)
    for q, c in enumerate(centres):
        xq = shell + c[None, :]
        xq = xq - L[None, :] * np.floor((xq - np.array(prm.min_phys[:d])[None, :]) / L[None, :])
        r3 = np.zeros((shell.shape[0], 3))
        r3[:, :d] = shell
        vq = vcol[q][None, :] + np.cross(wcol[q][None, :], r3)
        xs.append(xq)
        vs.append(vq[:, :d])
        sids.append(np.full(shell.shape[0], q + 1, np.int32))
    x = np.concatenate(xs)
    v = np.concatenate(vs)
    sid = np.concatenate(sids)
    m = np.full(x.shape[0], box.m[0])
    pid = np.arange(1, x.shape[0] + 1, dtype=np.int32)
    if d == 3:
        mass = prm.rho * (4.0 * math.pi * radius ** 3.0 / 3.0)
        mmi = np.full((k, 3), 0.4 * mass * radius ** 2)
    else:
        mass = prm.rho * (math.pi * radius ** 2.0)
        mmi = np.zeros((k, 3))
        mmi[:, 2] = 0.5 * mass * radius ** 2
    coll = ColloidSet(np.full(k, float(radius)), centres, np.full(k, mass), mmi, dout=dout,
                      rho_type=rho_type, **kw)
    out = Box(box.name + "_colloids", prm, x, v, m, pid, sid)
    return out, coll


def colloid_suspension_3d(n_side: int = 128, ncoll: int = 512, radius: float = 5.0,
                          shear_rate: float = 0.5, kt: float = 1e-3, seed: int = 1,
                          precision: int = PRECISION_FP64, translate: bool = True,
                          rotate: bool = True, symmetry: bool = True):
    """Spheres of radius 5 dx in the 128^3 Tait box, Lees-Edwards faces on y
    sheared at `shear_rate` (face velocities -+ shear_rate L / 2) with the
    solvent started on the linear profile.  The centres sit on a cubic
    lattice of spacing L / ncoll^(1/3) (16 dx for 512 spheres: a surface gap
    of 6 dx = 2 cut-offs), shifted by a seeded random offset: random
    sequential addition cannot place 512 such spheres with that gap (their
    excluded volume fraction would be 0.52, above its 0.38 jamming limit)."""
    box = tait_box_3d(n_side, precision=precision, kt=kt, seed=seed, symmetry=symmetry)
    L = float(n_side)
    per = round(ncoll ** (1.0 / 3.0))
    if per ** 3 != ncoll:
        raise ValueError("ncoll must be a cube")
    a = L / per
    rng = np.random.default_rng(seed + 4242)
    off = rng.uniform(0.0, a, size=3)
    g = (np.arange(per) + 0.0) * a
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    centres = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1) + off[None, :]
    centres = np.mod(centres, L)
    u = shear_rate * L / 2.0
    box.v[:, 0] += shear_rate * (box.x[:, 1] - L / 2.0)
    box.prm = box.prm.replace(le_dim=2, le_v_lower=[-u, 0.0, 0.0], le_v_upper=[u, 0.0, 0.0])
    vcol = np.zeros((ncoll, 3))
    vcol[:, 0] = shear_rate * (centres[:, 1] - L / 2.0)
    out, coll = with_colloids(box, centres, radius, dx=1.0, translate=translate, rotate=rotate,
                              v=vcol)
    out.name = "colloid_suspension_3d"
    return out, coll


def mixed(box: Box) -> Box:
    b = Box(box.name + "_mixed", box.prm.replace(precision=PRECISION_MIXED), box.x, box.v, box.m,
            box.pid, box.sid)
    return b


def write_external(path: str, box: Box) -> None:
    """Write the reference's read_external particle file
    (mcf/src/io_read_particles.F90:162-204): N, then x v rho m pid sid."""
    dim = box.prm.num_dim
    with open(path, "w") as fh:
        fh.write(f"{box.n}\n")
        rho = np.full(box.n, box.prm.rho)
        for i in range(box.n):
            vals = list(box.x[i, :dim]) + list(box.v[i, :dim]) + [rho[i], box.m[i]]
            fh.write(" ".join(f"{v:.17e}" for v in vals) + f" {int(box.pid[i])} {int(box.sid[i])}\n")


def read_external(path: str, num_dim: int):
    """Read a read_external particle file; returns x, v, rho, m, pid, sid.
    Errors mirror io_read_particles.F90:121-240 (missing file, short file)."""
    with open(path) as fh:
        first = fh.readline()
        if not first:
            raise ValueError("reading particles has problem at line 1")
        n = int(float(first.split()[0]))
        data = np.loadtxt(fh, ndmin=2) if n > 0 else np.zeros((0, 2 * num_dim + 4))
    if data.shape[0] != n:
        raise ValueError("actual number of lines is not equal to number given !")
    d = num_dim
    x = data[:, 0:d]
    v = data[:, d:2 * d]
    rho = data[:, 2 * d]
    m = data[:, 2 * d + 1]
    pid = data[:, 2 * d + 2].astype(np.int32)
    sid = data[:, 2 * d + 3].astype(np.int32)
    return x, v, rho, m, pid, sid
