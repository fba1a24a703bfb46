"""Verlet lists in boxes with sheared (Lees-Edwards) faces and with colloids,
on the GPU.  The lists keep every pair within cut_off + skin for several
steps; each step decides in / out on the current positions, the image across
a sheared face taken from them, and a build happens once the particles'
displacements plus the drift of the sheared copies since the build could
have closed a pair in by the skin.

The bars are those of the per-step lists: trajectories against the oracle
within 1e-10 with particles crossing the sheared faces (3D and 2D boxes),
neighbour sets equal to the oracle's bit for bit at the end, and the same
pairs as the per-step lists at every checked step.

Reference: marching_integrate_VV.F90:968-971 rebuilds the list every step;
the sheared copies are particles_set_boundary_ghost_*'s Lees-Edwards ghosts
(shear length of boundary_lees_edwards), and ppm_neighlist_vlist.F90:40-50
keeps pairs within cutoff + skin.
"""
import numpy as np
import pytest

from mcf_amd import workloads as W
from mcf_amd.sdpd import SDPD, pair_hash
from tests.colloid_helpers import marching_time, to_oracle, two_spheres
from tests.gpu_helpers import vec_err

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, flush=True)


def oracle_signature_le(O, box, t):
    """The oracle's neighbour counts and pair-hash sums with the sheared
    faces at shear time t."""
    pairs = O.neighbor_pairs(box.prm, box.x, le_time=t)
    if box.prm.symmetry:
        a, b = pairs[:, 0], pairs[:, 1]
        aa, bb = np.concatenate([a, b]), np.concatenate([b, a])
    else:
        aa, bb = pairs[:, 0], pairs[:, 1]
    cnt = np.bincount(aa, minlength=box.n).astype(np.int32)
    h = pair_hash(box.pid[aa], box.pid[bb]).astype(np.uint64)
    hs = np.zeros(box.n, np.uint64)
    np.add.at(hs, aa, h)
    return cnt, (hs & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def sheared_box(kind="y3d", sym=True, n_side=13):
    """A box with sheared faces (shear rate 0.5), the solvent on the linear
    profile and a time step large enough that a thin skin (box.skin, 0.05 dx)
    lasts a few steps: the sheared copies drift by 0.003 to 0.0065 dx per
    step.  The lattice is shifted along the faces' normal so that its top
    plane sits 0.05 dx below the upper face; that plane moves towards it by
    0.004 dx per step (0.003 in 2D) and crosses it.  A crossing counts the
    tangential shift it took as displacement, so it rebuilds the lists unless
    that shift is smaller than what is left of the skin.
    y3d / z3d: a 13^3 Tait box sheared across y or z.  y2d: the 24^2
    shear-wave box (unit square) sheared across y, where the whole fluid
    moves along the normal (its light particles would stop a single plane
    within a few steps)."""
    if kind == "y2d":
        box = W.shear_wave_2d(24, brownian=True, kt=1e-6, symmetry=sym)
        L, dx, le_dim, dt_factor, w = 1.0, 1.0 / 24, 2, 2.5, 0.003
    else:
        box = W.tait_box_3d(n_side, kt=1e-3, symmetry=sym)
        L, dx, le_dim, dt_factor, w = float(n_side), 1.0, 2 if kind == "y3d" else 3, 20, 0.004
    g = le_dim - 1
    u = 0.5 * L / 2
    dt = box.prm.dt * dt_factor
    box.prm = box.prm.replace(le_dim=le_dim, le_v_lower=[-u, 0, 0], le_v_upper=[u, 0, 0], dt=dt)
    x = box.x.copy()
    x[:, g] = np.mod(x[:, g] + 0.45 * dx, L)
    v = box.v.copy()
    v[:, 0] += 0.5 * (x[:, g] - L / 2)
    moving = x[:, g] > (L - 0.3 * dx if kind != "y2d" else -1.0)
    v[moving, g] += w * dx / dt
    box.x = np.ascontiguousarray(x)
    box.v = np.ascontiguousarray(v)
    box.skin = 0.05 * dx
    return box


def gpu_run(box, steps, skin, sid=None, colloids=None):
    with SDPD(box.prm) as g:
        g.set_skin(skin)
        if colloids is not None:
            g.set_colloids(colloids)
        g.upload(box.x, box.v, box.m, box.pid, sid)
        g.integrate_VV(steps)
        x, v = g.download()
        f = g.download_force()
        rho, _ = g.download_density()
        pairs, _ = g.counters()
        return dict(x=x, v=v, f=f, rho=rho, pairs=pairs, builds=g.list_builds(),
                    skin=g.skin(), sig=g.neighbor_signature())


def periodic_err(a, b, L):
    d = np.abs(a - b)
    return np.minimum(d, L - d).max()


@pytest.mark.parametrize("kind,sym",
                         [("y3d", True), ("y3d", False), ("z3d", True), ("y2d", True), ("y2d", False)])
def test_sheared_verlet_lists_against_oracle(oracle, kind, sym):
    box = sheared_box(kind, sym)
    steps = 30
    announce(f"sheared Verlet lists: {kind} sym={sym}, {steps} steps")
    out = gpu_run(box, steps, skin=box.skin)
    assert out["skin"] == pytest.approx(box.skin)
    assert out["builds"] >= 3  # rebuilt within the run, not every step
    assert out["builds"] < steps
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, steps + 1):
        st.vv_step(s)
    L = box.prm.max_phys[0]
    g = box.prm.le_dim - 1
    # the crossers went through the sheared faces
    assert (np.abs(out["x"][:, g] - box.x[:, g]) > 0.5 * L).sum() >= 4
    assert periodic_err(out["x"], st.x, L) < 1e-10 * L
    assert vec_err(out["v"], st.v) < 1e-10
    assert vec_err(out["f"], st.f) < 1e-10
    b = W.Box("end", box.prm, st.x, st.v, box.m, box.pid)
    cnt, hs = oracle_signature_le(oracle, b, marching_time(steps, box.prm.dt))
    assert np.array_equal(out["sig"][0], cnt) and np.array_equal(out["sig"][1], hs)


@pytest.mark.parametrize("kind", ["y3d", "z3d", "y2d"])
def test_sheared_verlet_agrees_with_per_step_lists(kind):
    box = sheared_box(kind, True)
    steps = 20
    announce(f"sheared Verlet vs per-step lists: {kind}")
    a = gpu_run(box, steps, skin=box.skin)
    b = gpu_run(box, steps, skin=0.0)
    assert a["builds"] >= 2 and b["builds"] == -1
    assert a["pairs"] == b["pairs"]
    assert np.array_equal(a["sig"][0], b["sig"][0]) and np.array_equal(a["sig"][1], b["sig"][1])
    L = box.prm.max_phys[0]
    assert periodic_err(a["x"], b["x"], L) < 1e-10 * L
    assert vec_err(a["v"], b["v"]) < 1e-10


def test_colloid_sheared_verlet_against_oracle(oracle):
    # two spheres, one straddling the periodic faces, in a sheared box; the
    # colloids translate and rotate while the lists live over several steps
    box, cs = two_spheres(16, radius=4.0, kt=1e-3, translate=True, rotate=True)
    L = float(16)
    u = 0.3
    box.prm = box.prm.replace(le_dim=2, le_v_lower=[-u, 0, 0], le_v_upper=[u, 0, 0],
                              dt=box.prm.dt * 20)
    steps = 12
    announce(f"colloids + sheared Verlet lists, {steps} steps")
    # the sheared copies drift by 6e-4 dx per step: a skin of 0.003 lasts
    # about five steps
    out = gpu_run(box, steps, skin=0.003, sid=box.sid, colloids=cs)
    assert out["builds"] >= 2 and out["builds"] < steps
    ref = gpu_run(box, steps, skin=0.0, sid=box.sid, colloids=cs)
    assert out["pairs"] == ref["pairs"]
    c = to_oracle(oracle, cs)
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid, box.sid, coll=c)
    st.prologue(0)
    for s in range(1, steps + 1):
        st.vv_step(s)
    assert periodic_err(out["x"], st.x, L) < 1e-10 * L
    assert vec_err(out["v"], st.v) < 1e-10
    assert vec_err(out["f"], st.f) < 1e-10
