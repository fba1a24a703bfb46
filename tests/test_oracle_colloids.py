"""Colloids in the oracle: boundary-particle density, the Morris no-slip
velocity (colloid_noslip_Morris_sphere.F90), pair rules of
pp_interaction_fc/cf.inc, drag and torque collection
(particles_collect_colloid_interaction.F90) and the rigid-body block of
marching_integrate_VV.F90:455-765.

The reference holds no test vectors for any of these, so each is pinned by
closed forms worked out by hand, brute-force sums, or conservation laws."""
import math

import numpy as np
import pytest

from mcf_amd import workloads as W
from mcf_amd.params import SDPDParams
from tests.colloid_helpers import to_oracle, two_disks, two_spheres


def unit_box(L=10.0, dim=3, cut_off=1.0, **kw):
    return SDPDParams(num_dim=dim, min_phys=[0, 0, 0], max_phys=[L, L, L if dim == 3 else 1.0],
                      cut_off=cut_off, **kw)


def one_colloid(O, x, radius=2.0, v=(0, 0, 0), omega=(0, 0, 0), noslip=2, rotate=False, **kw):
    return O.Colloids([radius], [1.0], [[0.4, 0.4, 0.4]], [x], [v], [omega], noslip=noslip,
                      rotate=rotate, dout=0.05, **kw)


# ----------------------------------------------------------------------------
# nearest image (colloid_in_nearest_image.F90)
# ----------------------------------------------------------------------------

def test_nearest_image_periodic(oracle):
    prm = unit_box()
    c = one_colloid(oracle, [1.0, 1.0, 1.0])
    xim, rx, vim = oracle.colloid_nearest(prm, c, 0.0, [9.5, 1.0, 1.0], 1)
    assert np.allclose(xim, [11.0, 1.0, 1.0]) and np.allclose(rx, [-1.5, 0, 0])
    xim, rx, _ = oracle.colloid_nearest(prm, c, 0.0, [0.5, 9.4, 9.8], 1)
    assert np.allclose(rx, [-0.5, -1.6, -1.2])
    # within d_max = radius + dout no image is taken, even past L / 2
    c = one_colloid(oracle, [1.0, 1.0, 1.0], radius=6.0)
    _, rx, _ = oracle.colloid_nearest(prm, c, 0.0, [7.0, 1.0, 1.0], 1)
    assert np.allclose(rx, [6.0, 0, 0])


def test_nearest_image_lees_edwards(oracle):
    # y faces sheared at -1 / +1: at t = 0.3 the upper images move by +0.6 in x
    prm = unit_box(le_dim=2, le_v_lower=[-1, 0, 0], le_v_upper=[1, 0, 0])
    c = one_colloid(oracle, [5.0, 0.5, 5.0], v=[0.25, 0, 0])
    xim, rx, vim = oracle.colloid_nearest(prm, c, 0.3, [5.2, 9.8, 5.0], 1)
    assert np.allclose(xim, [5.6, 10.5, 5.0])
    assert np.allclose(rx, [-0.4, -0.7, 0.0])
    assert np.allclose(vim, [2.25, 0, 0])
    # the shifted image wraps in x by MODULO
    c = one_colloid(oracle, [9.9, 0.5, 5.0])
    xim, rx, _ = oracle.colloid_nearest(prm, c, 0.3, [0.3, 9.8, 5.0], 1)
    assert np.allclose(xim[:2], [0.5, 10.5]) and np.allclose(rx[:2], [-0.2, -0.7])
    # a centre near the top seen from the bottom: lower image, shear reversed
    c = one_colloid(oracle, [5.0, 9.5, 5.0])
    xim, rx, vim = oracle.colloid_nearest(prm, c, 0.3, [4.0, 0.2, 5.0], 1)
    assert np.allclose(xim[:2], [4.4, -0.5]) and np.allclose(vim[0], -2.0)


# ----------------------------------------------------------------------------
# no-slip velocity (colloid_noslip_Morris_sphere.F90, _frozen.F90)
# ----------------------------------------------------------------------------

@pytest.mark.parametrize("xc,xf,corr", [
    ([1.5, 0, 0], [3.0, 0, 0], 0.5),       # d_cs 0.5, d_fs 1
    ([1.5, 0, 0], [4.0, 0, 0], 0.25),      # d_cs 0.5, d_fs 2
    ([1.0, 0, 0], [3.0, 0, 0], 0.5),       # d_cs 1 / d_fs 1 clamps to 0.5
    ([2.04, 0, 0], [3.0, 0, 0], -0.04),    # boundary particle in the outer layer
    ([0.0, 1.5, 0], [3.0, 0, 0], 0.5),     # d_cn 0: d_cs 2 / d_fs 1 clamps
    ([1.2, 0.9, 0], [0.0, 8.0, 0], 1.1 / 6.0),  # d_cn 0.9: d_cs 1.1, d_fs 6
    ([0.3, 1.6, 0], [6.0, 8.0, 0], 0.0675),  # n = (0.6, 0.8): d_cn 1.46, d_cs 0.54, d_fs 8
])
def test_morris_extrapolation(oracle, xc, xf, corr):
    # colloid of radius 2 at rest in the middle of a large box; the
    # boundary particle takes -corr times the fluid velocity, corr =
    # d_cs / d_fs clamped to +-0.5 (mcf_colloid_dist_ratio_max)
    prm = unit_box(L=30.0)
    o = np.array([15.0, 15.0, 15.0])
    c = one_colloid(oracle, o)
    vf = np.array([0.3, -0.2, 0.1])
    vc = oracle.colloid_noslip(prm, c, np.array(xf) + o, np.array(xc) + o, vf, 1)
    assert np.allclose(vc, -corr * vf, atol=1e-15, rtol=1e-13)


def test_morris_moving_rotating_colloid(oracle):
    prm = unit_box(L=20.0)
    o = np.array([10.0, 10.0, 10.0])
    V = np.array([0.1, 0.2, -0.3])
    w = 0.7
    c = one_colloid(oracle, o, v=V, omega=[0, 0, w], rotate=True)
    xc, xf = o + [1.5, 0, 0], o + [4.0, 0, 0]
    vf = np.array([1.0, 0.5, 0.25])
    # r_s = R n = (2, 0, 0): surface velocity V + w e_z x r_s = V + (0, 2w, 0)
    vs = V + np.array([0.0, 2.0 * w, 0.0])
    vc = oracle.colloid_noslip(prm, c, xf, xc, vf, 1)
    assert np.allclose(vc, -0.25 * (vf - vs) + vs, atol=1e-15)
    # a fluid particle moving with the surface sees it at rest: vc = v_s
    vc = oracle.colloid_noslip(prm, c, xf, xc, vs, 1)
    assert np.allclose(vc, vs, atol=1e-15)


def test_penetration_and_frozen(oracle):
    prm = unit_box(L=20.0)
    o = np.array([10.0, 10.0, 10.0])
    V = np.array([0.1, 0.2, -0.3])
    c = one_colloid(oracle, o, v=V, omega=[0, 0, 0.5], rotate=True)
    xc = o + [0.0, 1.5, 0.0]
    rigid = V + np.cross([0, 0, 0.5], [0.0, 1.5, 0.0])
    # fluid particle inside the sphere: the boundary particle moves rigidly
    vc = oracle.colloid_noslip(prm, c, o + [1.9, 0, 0], xc, np.array([5.0, 5, 5]), 1)
    assert np.allclose(vc, rigid)
    cf = one_colloid(oracle, o, v=V, omega=[0, 0, 0.5], rotate=True, noslip=1)
    vc = oracle.colloid_noslip(prm, cf, o + [4.0, 0, 0], xc, np.array([5.0, 5, 5]), 1)
    assert np.allclose(vc, rigid)


def test_morris_across_lees_edwards_face(oracle):
    # the boundary particle sits above the top face of a colloid centred near
    # the bottom: the centre's image carries the velocity jump +2
    prm = unit_box(le_dim=2, le_v_lower=[-1, 0, 0], le_v_upper=[1, 0, 0])
    c = one_colloid(oracle, [5.0, 0.5, 5.0])
    xc = np.array([5.6, 9.0, 5.0])   # 1.5 below the image centre (5.6, 10.5)
    xf = np.array([5.6, 6.5, 5.0])   # 4 below it: d_cs 0.5, d_fs 2
    vf = np.array([0.5, 0.0, 0.0])
    vc = oracle.colloid_noslip(prm, c, xf, xc, vf, 1, le_time=0.3)
    vimg = np.array([2.0, 0, 0])
    assert np.allclose(vc, -0.25 * (vf - vimg) + vimg, atol=1e-15)


# ----------------------------------------------------------------------------
# drag and torque
# ----------------------------------------------------------------------------

def test_collect_matches_brute_force(oracle):
    box, cs = two_spheres()
    c = to_oracle(oracle, cs)
    rng = np.random.default_rng(3)
    f = rng.normal(size=box.x.shape)
    drag, torque = oracle.colloid_collect(box.prm, c, box.x, f, box.m, box.sid)
    L = np.array(box.prm.length)
    for k in range(cs.num):
        sel = box.sid == k + 1
        r = box.x[sel] - cs.x[k][None, :]
        r -= L[None, :] * np.round(r / L[None, :])
        F = box.m[sel, None] * f[sel]
        assert np.allclose(drag[k], F.sum(0), rtol=1e-12, atol=1e-12)
        assert np.allclose(torque[k], np.cross(r, F).sum(0), rtol=1e-12, atol=1e-10)
    # fluid particles do not count
    f2 = f.copy()
    f2[box.sid == 0] += 100.0
    d2, t2 = oracle.colloid_collect(box.prm, c, box.x, f2, box.m, box.sid)
    assert np.array_equal(d2, drag) and np.array_equal(t2, torque)


# ----------------------------------------------------------------------------
# density
# ----------------------------------------------------------------------------

@pytest.mark.parametrize("sym", [True, False])
@pytest.mark.parametrize("rho_type", [0, 1])
def test_density_brute_force(oracle, sym, rho_type):
    box, cs = two_spheres(sym=sym, rho_type=rho_type)
    c = to_oracle(oracle, cs)
    rho = oracle.density(box.prm, box.x, box.m, box.sid, coll=c)
    pairs = oracle.neighbor_pairs(box.prm, box.x, brute=True)
    L = np.array(box.prm.length)
    d = box.x[pairs[:, 0]] - box.x[pairs[:, 1]]
    d -= L[None, :] * np.round(d / L[None, :])
    w = np.array([oracle.kernel(box.prm, r)[0] for r in np.sqrt((d * d).sum(1))])
    w0 = oracle.kernel(box.prm, 0.0)[0]
    si, sj = box.sid[pairs[:, 0]], box.sid[pairs[:, 1]]
    use = ~((si > 0) & (sj > 0)) if rho_type == 0 else np.ones(len(pairs), bool)
    ref = np.full(box.n, w0)
    np.add.at(ref, pairs[use, 0], w[use])
    np.add.at(ref, pairs[use, 1], w[use])
    b = box.sid > 0
    if rho_type == 0:
        ref[b] = box.prm.rho / box.m[b]
    assert np.allclose(rho, ref, rtol=1e-12)
    assert b.any() and (~b).any()


# ----------------------------------------------------------------------------
# forces
# ----------------------------------------------------------------------------

def eval_forces(O, box, c, step=3, v=None):
    rho = O.density(box.prm, box.x, box.m, box.sid, coll=c)
    p = O.pressure(box.prm, rho, box.m)
    f, _ = O.interaction(box.prm, step, box.x, box.v if v is None else v, rho, p, box.m, box.pid,
                         box.sid, coll=c)
    return f


@pytest.mark.parametrize("maker,sym", [(two_spheres, True), (two_spheres, False), (two_disks, True)])
@pytest.mark.parametrize("noslip", [1, 2])
def test_pair_forces_conserve_momentum(oracle, maker, sym, noslip):
    box, cs = maker(sym=sym, noslip=noslip, rotate=True, omega=[[0, 0, 0.02], [0.01, 0, 0.03]])
    box.prm = box.prm.replace(body_force_type=0)
    c = to_oracle(oracle, cs)
    f = eval_forces(oracle, box, c)
    mf = box.m[:, None] * f
    # fluid-boundary pairs take the same inputs from both sides: the pair
    # forces cancel, thermal noise included
    assert np.abs(mf.sum(0)).max() < 1e-11 * np.abs(mf).sum()
    assert np.abs(f[box.sid > 0]).max() > 0


def test_forces_are_galilean_invariant(oracle):
    box, cs = two_spheres(kt=1e-3)
    V = np.array([0.05, -0.02, 0.03])
    c0 = to_oracle(oracle, cs)
    f0 = eval_forces(oracle, box, c0)
    cs.v = cs.v + V[None, :]
    c1 = to_oracle(oracle, cs)
    f1 = eval_forces(oracle, box, c1, v=box.v + V[None, :])
    scale = np.sqrt((f0 ** 2).sum(1).mean())
    assert np.abs(f1 - f0).max() < 1e-11 * scale


def test_boundary_particles_alone_feel_nothing(oracle):
    box, cs = two_spheres(kt=1e-3)
    keep = box.sid > 0
    b = W.Box("shells", box.prm.replace(body_force_type=1, body_force=[0.5, 0, 0]), box.x[keep],
              box.v[keep], box.m[keep], box.pid[keep], box.sid[keep])
    f = eval_forces(oracle, b, to_oracle(oracle, cs))
    assert np.array_equal(f, np.zeros_like(f))


def test_fluid_boundary_pair_force(oracle):
    # one fluid particle next to one boundary particle, kT = 0: the force
    # is the Hu-Adams pair force with the boundary particle moving at the
    # Morris velocity and carrying the fluid particle's pressure, and the
    # boundary particle's constant number density rho / m
    prm = unit_box(L=20.0, cut_off=3.0, stateEquation_type=2, c=10.0, rho=1.0, rho_ref=1.0,
                   gamma=7.0, eta=0.1)
    o = np.array([10.0, 10.0, 10.0])
    c = one_colloid(oracle, o, radius=2.0)
    x = np.array([o + [3.0, 0.2, 0.0], o + [1.5, 0.0, 0.0]])
    v = np.array([[0.3, -0.2, 0.1], [7.0, 7.0, 7.0]])
    m = np.array([1.0, 1.0])
    pid = np.array([1, 2], np.int32)
    sid = np.array([0, 1], np.int32)
    rho = oracle.density(prm, x, m, sid, coll=c)
    d = x[0] - x[1]
    r = np.linalg.norm(d)
    w, gw = oracle.kernel(prm, r)
    w0, _ = oracle.kernel(prm, 0.0)
    assert rho[0] == pytest.approx(w0 + w) and rho[1] == 1.0
    p = oracle.pressure(prm, rho, m)
    f, _ = oracle.interaction(prm, 0, x, v, rho, p, m, pid, sid, coll=c)
    vc = oracle.colloid_noslip(prm, c, x[0], x[1], v[0], 1)
    ni, nj = rho
    e = d / r
    fi = -(p[0] / ni ** 2 + p[0] / nj ** 2) * gw * e + prm.eta * (1 / ni ** 2 + 1 / nj ** 2) * gw / r * (v[0] - vc)
    assert np.allclose(f[0], fi, rtol=1e-13)
    assert np.allclose(f[1], -fi, rtol=1e-13)


@pytest.mark.parametrize("sym", [True, False])
def test_fluid_boundary_pair_oldroyd(oracle, sym):
    # Oldroyd-B (pp_interaction_fc.inc:80-113, pp_interaction_cf.inc:80-110):
    # both sides of a fluid-boundary pair take the fluid particle's number
    # density and pressure tensor Pi = (p + G) I - G c, the boundary particle
    # moving at the Morris velocity (kT = 0).  The boundary particle's own
    # number density (rho / m = 1), pressure and conformation differ from
    # the fluid particle's, so taking any of them would move the result.
    G = 0.5
    prm = unit_box(L=20.0, cut_off=3.0, stateEquation_type=2, c=10.0, rho=1.0, rho_ref=1.0,
                   gamma=7.0, eta=0.1, Newtonian=False, tau=0.1, n_p=1.0, kt_p=G, symmetry=sym)
    o = np.array([10.0, 10.0, 10.0])
    c = one_colloid(oracle, o, radius=2.0)
    x = np.array([o + [3.0, 0.2, 0.1], o + [1.5, 0.0, 0.0]])
    v = np.array([[0.3, -0.2, 0.1], [7.0, 7.0, 7.0]])
    m = np.array([1.0, 1.0])
    pid = np.array([1, 2], np.int32)
    sid = np.array([0, 1], np.int32)
    cf = np.array([[1.2, 0.1, -0.05], [0.1, 0.9, 0.2], [-0.05, 0.2, 1.1]])
    cb = np.array([[0.7, -0.3, 0.0], [-0.3, 1.4, 0.1], [0.0, 0.1, 0.8]])
    ct = np.stack([cf.reshape(-1, order="F"), cb.reshape(-1, order="F")])
    rho = oracle.density(prm, x, m, sid, coll=c)
    p = oracle.pressure(prm, rho, m)
    f, vgt = oracle.interaction(prm, 0, x, v, rho, p, m, pid, sid, ct=ct, coll=c)
    vc = oracle.colloid_noslip(prm, c, x[0], x[1], v[0], 1)
    nf = rho[0]
    assert rho[1] == 1.0 and abs(nf - 1.0) > 0.01
    d = x[0] - x[1]
    r = np.linalg.norm(d)
    e = d / r
    _, gw = oracle.kernel(prm, r)
    pi_f = (p[0] + G) * np.eye(3) - G * cf
    # rhs_force_ff_non_Newtonian_HuAdams.F90:105-131 with n_j = n_i, Pi_j = Pi_i
    fi = -(2.0 * pi_f @ e / nf ** 2) * gw + prm.eta * (2.0 / nf ** 2) * gw / r * (v[0] - vc)
    assert np.allclose(f[0], fi, rtol=1e-13, atol=0)
    assert np.allclose(f[1], -fi, rtol=1e-13, atol=0)
    # pp_vgt_ip.inc / pp_vgt_jp.inc: L(a, b) += (v_j - v_i)_a (x_i - x_j)_b W' / (n r)
    li = np.outer(vc - v[0], d) * gw / (nf * r)
    lb = np.outer(v[0] - vc, -d) * gw / (nf * r)
    assert np.allclose(vgt[0].reshape(3, 3, order="F"), li, rtol=1e-13, atol=1e-16)
    assert np.allclose(vgt[1].reshape(3, 3, order="F"), lb, rtol=1e-13, atol=1e-16)


# ----------------------------------------------------------------------------
# rigid-body motion
# ----------------------------------------------------------------------------

def shell_only(box, cs, q=1):
    keep = box.sid == q
    # a large time step keeps the displacements well above the rounding of
    # the positions
    return W.Box("shell", box.prm.replace(Brownian=False, kt=0.0, dt=0.05), box.x[keep], box.v[keep],
                 box.m[keep], box.pid[keep], np.full(int(keep.sum()), 1, np.int32))


def run_state(O, box, c, steps):
    st = O.State(box.prm, box.x, box.v, box.m, box.pid, box.sid, coll=c)
    st.prologue(0)
    out = [st.x.copy()]
    for s in range(1, steps + 1):
        st.vv_step(s)
        out.append(st.x.copy())
    return st, out


def unwrap(d, L):
    return d - L[None, :] * np.round(d / L[None, :])


@pytest.mark.parametrize("ab,sub", [(1, 1), (2, 1), (1, 2), (3, 1)])
def test_translation_under_constant_force(oracle, ab, sub):
    # shells alone feel no pair force, so the colloid body force is the
    # only drag: a = body_force, and the prologue starts the acceleration
    # history at a.  AB1 over N sub-steps of dt / sub: x_N - x_0 =
    # a (dt / sub)^2 N (N - 1) / 2.  AB2 and AB3 integrate the linear
    # velocity history exactly once their order has ramped up from 1:
    # x_n - x_0 = a dt^2 (n^2 - 1) / 2.  Either way v_n = n a dt.
    box, cs = two_spheres(kt=0.0)
    b = shell_only(box, cs)
    a = np.array([0.3, -0.1, 0.2])
    c = oracle.Colloids(cs.radius[:1], cs.m[:1], cs.mmi[:1], cs.x[:1], translate=True, ab=ab,
                        sub_steps=sub, body_force_type=1, body_force=a, dout=cs.dout)
    n = 6
    dt = b.prm.dt
    st, xs = run_state(oracle, b, c, n)
    L = np.array(b.prm.length)
    if ab == 1:
        N = n * sub
        expect = a * (dt / sub) ** 2 * N * (N - 1) / 2
    else:
        expect = a * dt ** 2 * (n * n - 1) / 2
    disp = unwrap(xs[-1] - xs[0], L)
    assert np.allclose(disp, expect[None, :], rtol=1e-9, atol=1e-15)
    assert np.allclose(unwrap(c.x[:1] - cs.x[:1], L), expect[None, :], rtol=1e-9, atol=1e-15)
    assert np.allclose(c.v[0, 0], a * dt * n, rtol=1e-12)


@pytest.mark.parametrize("ab", [1, 2])
def test_rotation_at_constant_spin(oracle, ab):
    box, cs = two_spheres(kt=0.0)
    b = shell_only(box, cs)
    w = np.array([0.0, 0.0, 40.0])
    c = oracle.Colloids(cs.radius[:1], cs.m[:1], cs.mmi[:1], cs.x[:1], omega=[w], rotate=True,
                        ab=ab, dout=cs.dout)
    n = 5
    st, xs = run_state(oracle, b, c, n)
    L = np.array(b.prm.length)
    r0 = unwrap(xs[0] - cs.x[0][None, :], L)
    r1 = unwrap(xs[-1] - cs.x[0][None, :], L)
    ang = n * w[2] * b.prm.dt
    rot = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
    assert np.allclose(r1, r0 @ rot.T, atol=1e-12)
    assert c.theta[0, 2] == pytest.approx(ang, rel=1e-12)
    # accumulated rotation, element (r, c) stored at [c][r]
    assert np.allclose(c.acc_matrix[0].T, rot, atol=1e-12)


def test_colloid_crosses_lees_edwards_face(oracle):
    box, cs = two_spheres(kt=0.0)
    b = shell_only(box, cs, q=2)
    L = np.array(b.prm.length)
    u = 2.0
    b.prm = b.prm.replace(le_dim=2, le_v_lower=[-u, 0, 0], le_v_upper=[u, 0, 0], dt=0.05)
    X0 = cs.x[1].copy()
    vy = (L[1] - X0[1] + 0.1) / b.prm.dt  # crosses the top face in the first step
    c = oracle.Colloids(cs.radius[1:], cs.m[1:], cs.mmi[1:], [X0], v=[[0.0, vy, 0.0]],
                        translate=True, dout=cs.dout)
    st = oracle.State(b.prm, b.x, b.v, b.m, b.pid, b.sid, coll=c)
    st.prologue(0)
    st.vv_step(1)
    # the step's colloid block uses the shear length of the prologue (t = 0):
    # the centre moves up by vy dt, leaves through the top and comes back at
    # the bottom with x unchanged (shear length 0) and v_x - 2u
    assert c.x[0, 1] == pytest.approx(X0[1] + vy * b.prm.dt - L[1])
    assert c.x[0, 0] == pytest.approx(X0[0])
    assert c.v[0, 0, 0] == pytest.approx(-2 * u)
    assert np.all((st.x[:, 1] >= 0) & (st.x[:, 1] < L[1]))


def test_statistics_add_colloid_momentum(oracle):
    box, cs = two_spheres(kt=1e-3)
    cs.v[:] = [[0.1, 0.0, 0.0], [0.0, -0.2, 0.05]]
    c = to_oracle(oracle, cs)
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid, box.sid, coll=c)
    ek, mom = st.statistics()
    fl = box.sid == 0
    ek_ref = 0.5 * (box.m[fl, None] * box.v[fl] ** 2).sum() + 0.5 * (cs.m[:, None] * cs.v ** 2).sum()
    mom_ref = (box.m[fl, None] * box.v[fl]).sum(0) + (cs.m[:, None] * cs.v).sum(0)
    assert ek == pytest.approx(ek_ref, rel=1e-12)
    assert np.allclose(mom, mom_ref, rtol=1e-10, atol=1e-12)
