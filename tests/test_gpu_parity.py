"""GPU parity of the HIP path against the FP64 oracle (run with -m gpu).

Bars: bit-exact for the cell list and the neighbour sets; 1e-12 relative
for density / pressure and for the forces (max over particles of |df_i|
normalised by the rms force) in FP64 mode on identical inputs;
trajectories after several steps are reported with a looser bar (1e-10)
because the summation order differs from the reference's.  The oracle
takes the thermal noise's normals from the device's stream (injected per
evaluation, tests/conftest.py); the stream itself is checked against the
oracle's exactly rounded one and for its statistics.
"""
import sys

import numpy as np
import pytest

from mcf_amd import workloads as W
from mcf_amd.sdpd import SDPD, SDPDError, pair_normals

from gpu_helpers import (gpu_evaluate, oracle_evaluate, oracle_signature, rel_err,
                         vec_err)

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


BOXES = {
    "shear2d_sym": lambda: W.shear_wave_2d(20),
    "shear2d_full": lambda: W.shear_wave_2d(17, symmetry=False),
    "tait3d_sym_kt0": lambda: W.tait_box_3d(12, brownian=False),
    "tait3d_sym_noise": lambda: W.tait_box_3d(13, kt=1e-3),
    "tait3d_full_noise": lambda: W.tait_box_3d(11, kt=1e-3, symmetry=False),
}


def test_noise_stream_close_to_exact(oracle):
    # the device draws Box-Muller with the hardware's log2, sqrt and
    # sin / cos; the oracle's stream does the same arithmetic in exactly
    # rounded float operations: the two differ by a few ulp of the radius
    # and of the angle's sine / cosine
    announce("noise stream: device vs exactly rounded")
    rng = np.random.default_rng(4)
    lo = rng.integers(0, 2 ** 31, 200000).astype(np.uint32)
    hi = (lo + rng.integers(1, 1000, 200000)).astype(np.uint32)
    lo[:3] = [0, 0xFFFFFFFF, 7]
    hi[:3] = [0, 0xFFFFFFFF, 7]
    worst = 0.0
    for seed, step in [(1, 0), (0xDEADBEEFCAFE, 12345), (7, (1 << 33) + 5)]:
        g = pair_normals(lo, hi, seed, step)
        ref = oracle.pair_normals_n(seed, lo, hi, step)
        assert np.isfinite(g).all()
        err = np.abs(g.astype(np.float64) - ref) / np.maximum(1.0, np.abs(ref))
        worst = max(worst, float(err.max()))
    announce(f"noise stream: max |dg| / max(1, |g|) = {worst:.2e}")
    assert worst < 1e-5


def test_noise_stream_statistics():
    # eight million normals of the device stream: N(0, 1) moments, the
    # Kolmogorov-Smirnov distance, no correlation between the four of a
    # pair, the radius bounded by sqrt(2 ln 2^24) (24-bit uniforms)
    from scipy import stats

    announce("noise stream: statistics")
    lo = np.arange(2000000, dtype=np.uint32) * np.uint32(2654435761)
    hi = lo ^ np.uint32(0x5bd1e995)
    g = pair_normals(lo, hi, 0x1234567890ABCDEF, 77).astype(np.float64)
    n = g.size
    x = g.ravel()
    assert abs(x.mean()) < 5.0 / np.sqrt(n)
    assert abs(x.var() - 1.0) < 5.0 * np.sqrt(2.0 / n)
    assert abs(stats.kurtosis(x)) < 5.0 * np.sqrt(24.0 / n)
    assert stats.kstest(x, "norm").pvalue > 1e-4
    c = np.corrcoef(g.T)
    assert np.abs(c - np.eye(4)).max() < 5.0 / np.sqrt(g.shape[0])
    assert np.abs(x).max() <= np.sqrt(2.0 * 24.0 * np.log(2.0)) * (1.0 + 1e-6)


@pytest.mark.parametrize("name", ["shear2d_sym", "tait3d_sym_kt0", "tait3d_full_noise"])
def test_cell_list_bitexact(oracle, name):
    box = BOXES[name]()
    announce(f"cell list: {name} n={box.n}")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        nm, lhbx, lpdx = g.cells()
    dim = box.prm.num_dim
    ref_lhbx, ref_lpdx = oracle.rank(dim, box.x, np.array(box.prm.min_phys[:dim]),
                                     np.array(box.prm.max_phys[:dim]), nm[:dim].astype(np.int32),
                                     np.zeros(2 * dim, np.int32))
    assert list(nm[:dim]) == box.prm.cell_counts()[:dim]
    assert np.array_equal(lhbx, ref_lhbx)
    assert np.array_equal(lpdx, ref_lpdx)


def test_cell_list_faces_and_ties(oracle):
    # particles exactly on cell faces and on the lower box face
    box = W.tait_box_3d(12, brownian=False)
    x = box.x.copy()
    cs = 12.0 / 3
    x[:40, 0] = cs * (np.arange(40) % 3)
    x[40:60, 1] = 0.0
    x[60:80, 2] = np.nextafter(12.0, 0.0)
    box.x = x
    announce("cell list: faces and ties")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        nm, lhbx, lpdx = g.cells()
    ref_lhbx, ref_lpdx = oracle.rank(3, x, np.zeros(3), np.full(3, 12.0), nm.astype(np.int32),
                                     np.zeros(6, np.int32))
    assert np.array_equal(lhbx, ref_lhbx) and np.array_equal(lpdx, ref_lpdx)


@pytest.mark.parametrize("name", list(BOXES))
def test_neighbor_sets_bitexact(oracle, name):
    box = BOXES[name]()
    announce(f"neighbour sets: {name} n={box.n}")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        cnt, hs = g.neighbor_signature()
    rc, rh = oracle_signature(oracle, box)
    assert np.array_equal(cnt, rc)
    assert np.array_equal(hs, rh)


def test_neighbor_sets_perfect_lattice(oracle):
    # 30 lattice neighbours sit exactly on r = 3 dx: strict cut-off
    box = W.tait_box_3d(12, brownian=False)
    box.x = W.lattice(3, 12, 1.0, 0.0, 0)
    announce("neighbour sets: perfect lattice")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        cnt, hs = g.neighbor_signature()
    rc, rh = oracle_signature(oracle, box)
    assert np.array_equal(cnt, rc) and np.array_equal(hs, rh)
    assert (cnt == cnt[0]).all()


@pytest.mark.parametrize("name", list(BOXES))
def test_density_pressure_parity(oracle, name):
    box = BOXES[name]()
    announce(f"density: {name}")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        g.compute_density()
        rho, p = g.download_density()
    ro, po = oracle.density(box.prm, box.x, box.m), None
    po = oracle.pressure(box.prm, ro, box.m)
    assert rel_err(rho, ro) < 1e-13
    assert rel_err(p, po) < 1e-12


@pytest.mark.parametrize("name", list(BOXES))
@pytest.mark.parametrize("force_type", [3, 4])
def test_force_parity(oracle, name, force_type):
    box = BOXES[name]()
    box.prm = box.prm.replace(rhs_force_type=force_type)
    announce(f"forces: {name} type {force_type}")
    g = gpu_evaluate(box, step=17)
    f = g.download_force()
    pairs, _ = g.counters()
    g.close()
    rho, p, fo, _ = oracle_evaluate(oracle, box, step=17)
    assert vec_err(f, fo) < 1e-12
    on = oracle.neighbor_pairs(box.prm, box.x)
    assert pairs == (2 if box.prm.symmetry else 1) * len(on)


@pytest.mark.parametrize("kind,ft,dens,brown,sym", [
    ("3d", 1, 1, False, True), ("3d", 1, 2, False, False),
    ("3d", 2, 2, True, True), ("3d", 2, 2, True, False), ("2d", 2, 2, True, True)])
def test_morris_espanol_parity(oracle, kind, ft, dens, brown, sym):
    # rhs_force_ff_Newtonian_Morris.F90 (type 1, no noise) and
    # rhs_force_ff_Newtonian_Espanol.F90 (type 2, its trace noise from the
    # fourth normal of the pair block), FP64
    if kind == "3d":
        box = W.tait_box_3d(10, kt=1e-3, brownian=brown, symmetry=sym)
    else:
        box = W.shear_wave_2d(16, brownian=brown, kt=1e-4, symmetry=sym)
    box.prm = box.prm.replace(rhs_force_type=ft, rhs_density_type=dens)
    announce(f"forces: rhs_force_type {ft}, density type {dens}, {kind} brownian={brown} sym={sym}")
    g = gpu_evaluate(box, step=4)
    rho, _ = g.download_density()
    f = g.download_force()
    g.close()
    ro, _, fo, _ = oracle_evaluate(oracle, box, step=4)
    assert rel_err(rho, ro) < 1e-13
    assert vec_err(f, fo) < 1e-12
    # mixed precision: FP32 pair arithmetic on the same records and normals
    gm = gpu_evaluate(W.mixed(box), step=4)
    fm = gm.download_force()
    gm.close()
    assert vec_err(fm, fo) < (5e-5 if kind == "2d" else 1e-5)


@pytest.mark.parametrize("ft", [1, 2])
def test_morris_espanol_trajectory(oracle, ft):
    box = W.tait_box_3d(10, kt=1e-3, brownian=ft == 2)
    box.prm = box.prm.replace(rhs_force_type=ft, dt=box.prm.dt * 10)
    announce(f"rhs_force_type {ft}: 6-step trajectory")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(6)
        x, v = g.download()
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, 7):
        st.vv_step(s)
    assert vec_err(v, st.v) < 1e-10
    L = 10.0
    dx = np.abs(x - st.x)
    assert np.minimum(dx, L - dx).max() < 1e-10 * L


@pytest.mark.parametrize("kind,btype", [("2d", 1), ("2d", 2), ("2d", 3), ("3d", 2), ("3d", 3)])
def test_body_force_parity(oracle, kind, btype):
    # particles_apply_body_force.F90: 1 constant, 2 +-g_x on the halves of
    # the last dimension, 3 g_x sin(k y); mixed precision shares the epilogue
    box = W.shear_wave_2d(16) if kind == "2d" else W.tait_box_3d(10, kt=1e-3)
    box.prm = box.prm.replace(body_force_type=btype, body_force=[0.5, 0.2, -0.1])
    announce(f"forces: body force type {btype}, {kind}")
    g = gpu_evaluate(box)
    f = g.download_force()
    g.close()
    _, _, fo, _ = oracle_evaluate(oracle, box)
    assert vec_err(f, fo) < 1e-12
    gm = gpu_evaluate(W.mixed(box))
    fm = gm.download_force()
    gm.close()
    assert vec_err(fm, fo) < (5e-5 if kind == "2d" else 1e-5)


@pytest.mark.parametrize("btype", [2, 3])
def test_body_force_trajectory(oracle, btype):
    box = W.tait_box_3d(10, kt=1e-3)
    box.prm = box.prm.replace(body_force_type=btype, body_force=[2.0, 0.0, 0.0], dt=box.prm.dt * 20)
    announce(f"body force type {btype}: 6-step trajectory")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(6)
        x, v = g.download()
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, 7):
        st.vv_step(s)
    assert vec_err(v, st.v) < 1e-10
    L = 10.0
    dx = np.abs(x - st.x)
    assert np.minimum(dx, L - dx).max() < 1e-10 * L


def test_body_force_types_2_3_are_for_one_box():
    prm = W.tait_box_3d(10).prm.replace(body_force_type=2, body_force=[0.1, 0.0, 0.0])
    ens = W.ensemble_of_cubes((10, 11), prm, kt=0.0)
    with SDPD(ens.prm) as g:
        with pytest.raises(SDPDError):
            g.set_replicas(ens.box_lo, ens.box_hi)


@pytest.mark.parametrize("name", ["shear2d_sym", "tait3d_sym_noise", "tait3d_full_noise"])
def test_trajectory_parity(oracle, name):
    box = BOXES[name]()
    nsteps = 6
    announce(f"trajectory: {name} {nsteps} steps")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(nsteps)
        x, v = g.download()
        e, mom = g.compute_statistic()
        rho, p = g.download_density()
        f = g.download_force()
        assert g.step == nsteps
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, nsteps + 1):
        st.vv_step(s)
    L = np.array(box.prm.length[: box.prm.num_dim])
    dx = x - st.x
    dx -= L * np.round(dx / L)
    assert np.abs(dx).max() < 1e-10 * L.max()
    assert vec_err(v, st.v) < 1e-10
    assert rel_err(rho, st.rho) < 1e-12
    assert vec_err(f, st.f) < 1e-10
    eo, mo = st.statistics()
    assert e == pytest.approx(eo, rel=1e-10)
    assert np.abs(mom - mo).max() < 1e-10 * (np.abs(box.m[:, None] * box.v).sum() + 1e-30)


@pytest.mark.parametrize("kt", [0.0, 1e-4])
def test_shear_wave_1000_steps(oracle, kt):
    # the 2D reference case: 64^2 particles, quintic kernel, c 10, eta 0.1,
    # sinusoidal shear, dt 1e-4, 1000 velocity-Verlet steps, deterministic
    # and with thermal noise (the same counter-based stream on both sides)
    box = W.shear_wave_2d(64, brownian=kt > 0, kt=kt if kt > 0 else 1e-4)
    nsteps = 1000
    announce(f"shear wave 64^2, kT {kt}: {nsteps} steps")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(nsteps)
        x, v = g.download()
        e, mom = g.compute_statistic()
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, nsteps + 1):
        st.vv_step(s)
    L = np.array(box.prm.length[:2])
    dx = x - st.x
    dx -= L * np.round(dx / L)
    ex, ev = np.abs(dx).max() / L.max(), vec_err(v, st.v)
    eo, mo = st.statistics()
    announce(f"  max |dx| / L = {ex:.2e}, velocity {ev:.2e}, E_k rel {abs(e / eo - 1):.2e}")
    # measured on MI355X: |dx| / L below 3e-15, velocities within 1.2e-12
    assert ex < 1e-13
    assert ev < 1e-10
    assert e == pytest.approx(eo, rel=1e-11)


def test_tait_box_200_steps(oracle):
    # the 3D Tait case (gamma 7, c 10, kT 1e-3, dt 5e-5) over its 200 steps
    # on a 16^3 box
    box = W.tait_box_3d(16, kt=1e-3)
    nsteps = 200
    announce(f"Tait 16^3: {nsteps} steps")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(nsteps)
        x, v = g.download()
        rho, _ = g.download_density()
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, nsteps + 1):
        st.vv_step(s)
    L = np.array(box.prm.length[:3])
    dx = x - st.x
    dx -= L * np.round(dx / L)
    ex, ev = np.abs(dx).max() / L.max(), vec_err(v, st.v)
    announce(f"  max |dx| / L = {ex:.2e}, velocity {ev:.2e}, density {rel_err(rho, st.rho):.2e}")
    # measured on MI355X: 1e-16 of the box, velocities 5e-15, density 3e-15
    assert ex < 1e-13
    assert ev < 1e-11
    assert rel_err(rho, st.rho) < 1e-12


def test_step_api_matches_individual_calls(oracle):
    # sdpd_step == kick, drift+wrap, build, density, forces, kick
    box = BOXES["tait3d_sym_noise"]()
    announce("step API vs individual calls")
    with SDPD(box.prm) as a, SDPD(box.prm) as b:
        for g in (a, b):
            g.upload(box.x, box.v, box.m, box.pid)
        a.integrate_VV(3)
        b.evaluate()
        for s in range(1, 4):
            b.step = s
            b.integrate_velocity(0.5)
            b.integrate_position_and_adjust()
            b.build_list()
            b.compute_density()
            b.compute_interaction()
            b.integrate_velocity(0.5)
        xa, va = a.download()
        xb, vb = b.download()
    assert np.array_equal(xa, xb)
    assert np.array_equal(va, vb)


def test_oldroyd_parity(oracle):
    box = W.oldroyd_box_3d(12)
    announce("Oldroyd-B forces, vgt, act")
    rng = np.random.default_rng(9)
    ct = np.tile(np.eye(3).reshape(-1), (box.n, 1))
    a = rng.normal(0, 0.05, size=(box.n, 3, 3))
    ct = ct + (a + a.transpose(0, 2, 1)).reshape(box.n, 9)
    g = gpu_evaluate(box, step=2, ct=ct)
    f = g.download_force()
    ctg, vgt = g.download_tensors()
    g.close()
    rho, p, fo, vgto = oracle_evaluate(oracle, box, step=2, ct=ct)
    assert vec_err(f, fo) < 1e-12
    assert vec_err(vgt, vgto) < 1e-12
    assert np.array_equal(ctg, ct)


def _oldroyd_2d():
    box = W.shear_wave_2d(24, brownian=True, kt=1e-4)
    box.prm = box.prm.replace(Newtonian=False, tau=0.1, n_p=1.0, kt_p=0.5)
    return box


def test_oldroyd_2d_parity(oracle):
    # the 2D pressure-tensor record (xx, yy, 0, xy) and the 2x2 tensors
    box = _oldroyd_2d()
    announce("Oldroyd-B 2D forces, vgt")
    rng = np.random.default_rng(11)
    a = rng.normal(0, 0.05, size=(box.n, 2, 2))
    ct = (np.tile(np.eye(2), (box.n, 1, 1)) + a + a.transpose(0, 2, 1)).reshape(box.n, 4)
    g = gpu_evaluate(box, step=3, ct=ct)
    f = g.download_force()
    _, vgt = g.download_tensors()
    g.close()
    rho, p, fo, vgto = oracle_evaluate(oracle, box, step=3, ct=ct)
    assert vec_err(f, fo) < 1e-12
    assert vec_err(vgt, vgto) < 1e-12
    gm = gpu_evaluate(W.mixed(box), step=3, ct=ct)
    fm = gm.download_force()
    gm.close()
    assert vec_err(fm, fo) < 5e-5  # the 2D mixed bar (FORCE_BAR)


@pytest.mark.parametrize("kind,ft,brown", [("3d", 2, True), ("3d", 4, False), ("2d", 2, True),
                                           ("2d", 4, False)])
def test_oldroyd_force_laws_parity(oracle, kind, ft, brown):
    # rhs_force_ff_non_Newtonian.F90: Espanol (2: the Hu-Adams terms and
    # Espanol's thermal force) and the angular form (4: the diagonal of the
    # pressure tensors averaged, dissipation along e, no thermal force)
    box = W.oldroyd_box_3d(12) if kind == "3d" else _oldroyd_2d()
    box.prm = box.prm.replace(rhs_force_type=ft, Brownian=brown)
    announce(f"Oldroyd-B rhs_force_type {ft}, {kind}, brownian={brown}")
    d = box.prm.num_dim
    rng = np.random.default_rng(13)
    a = rng.normal(0, 0.05, size=(box.n, d, d))
    ct = (np.tile(np.eye(d), (box.n, 1, 1)) + a + a.transpose(0, 2, 1)).reshape(box.n, d * d)
    g = gpu_evaluate(box, step=2, ct=ct)
    f = g.download_force()
    _, vgt = g.download_tensors()
    g.close()
    _, _, fo, vgto = oracle_evaluate(oracle, box, step=2, ct=ct)
    assert vec_err(f, fo) < 1e-12
    assert vec_err(vgt, vgto) < 1e-12
    # mixed precision: FP32 pair arithmetic on the Pi / n^2 records
    gm = gpu_evaluate(W.mixed(box), step=2, ct=ct)
    fm = gm.download_force()
    _, vgtm = gm.download_tensors()
    gm.close()
    assert vec_err(fm, fo) < (5e-5 if kind == "2d" else 1e-5)
    assert vec_err(vgtm, vgto) < 1e-5


@pytest.mark.parametrize("ft", [2, 4])
def test_oldroyd_force_laws_trajectory(oracle, ft):
    box = W.oldroyd_box_3d(12)
    box.prm = box.prm.replace(rhs_force_type=ft, Brownian=ft == 2)
    announce(f"Oldroyd-B rhs_force_type {ft}: trajectory")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(4)
        x, v = g.download()
        ct, vgt = g.download_tensors()
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, 5):
        st.vv_step(s)
    assert vec_err(v, st.v) < 1e-10
    assert vec_err(ct, st.ct) < 1e-10
    assert vec_err(vgt, st.vgt) < 1e-9


def test_oldroyd_trajectory(oracle):
    box = W.oldroyd_box_3d(12)
    announce("Oldroyd-B trajectory")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(4)
        x, v = g.download()
        ct, vgt = g.download_tensors()
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, 5):
        st.vv_step(s)
    assert vec_err(v, st.v) < 1e-10
    assert vec_err(ct, st.ct) < 1e-10
    assert vec_err(vgt, st.vgt) < 1e-9


def test_run_to_run_bitwise():
    box = BOXES["tait3d_sym_noise"]()
    announce("determinism")
    outs = []
    for _ in range(2):
        with SDPD(box.prm) as g:
            g.upload(box.x, box.v, box.m, box.pid)
            g.integrate_VV(3)
            outs.append(g.download() + (g.download_force(),))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_error_paths():
    box = BOXES["shear2d_sym"]()
    announce("error paths")
    with SDPD(box.prm) as g:
        with pytest.raises(SDPDError):
            g.build_list()  # no particles yet
        g.upload(box.x, box.v, box.m, box.pid)
        with pytest.raises(SDPDError):
            g.compute_density()  # no cell list yet
        g.build_list()
        with pytest.raises(SDPDError):
            g.download_force()  # forces not current
        bad = box.x.copy()
        bad[0, 0] = 2.0
        with pytest.raises(SDPDError):
            g.upload(bad, box.v, box.m, box.pid)


def test_full_size_tait_box_properties():
    # 128^3 = 2,097,152 particles: size-independent properties
    box = W.tait_box_3d(128)
    announce(f"full size: n={box.n}")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(2)
        f = g.download_force()
        rho, p = g.download_density()
        pairs, _ = g.counters()
        e, mom = g.compute_statistic()
    mf = box.m[:, None] * f
    assert np.abs(mf.sum(0)).max() < 1e-10 * np.abs(mf).sum()
    nb = pairs / box.n
    assert 100 < nb < 115
    assert abs(rho.mean() - 360.0 / 359.0) < 0.01
    assert np.isfinite(f).all()


# ---------------------------------------------------------------------
# mixed precision: FP32 pair arithmetic on cell-relative positions, FP64
# state; same noise stream.  Bars: bit-exact neighbour sets, density to
# 1e-6, forces to 1e-5 of the rms force against the FP64 path.
# ---------------------------------------------------------------------
def _tait2d():
    box = W.shear_wave_2d(20, brownian=True, kt=1e-4)
    box.prm = box.prm.replace(stateEquation_type=2, gamma=7.0, rho_ref=1.0)
    return box


MIXED_BOXES = {
    "tait3d_mixed_noise": lambda: W.tait_box_3d(13, kt=1e-3),
    "tait3d_mixed_full": lambda: W.tait_box_3d(11, kt=1e-3, symmetry=False),
    "tait2d_mixed": _tait2d,
    # p = c^2 rho carries a background pressure of 100: the pair terms
    # cancel to a few per cent of their size, which costs FP32 about one
    # decimal digit; checked at 5e-5
    "morris2d_mixed": lambda: W.shear_wave_2d(20),
}
FORCE_BAR = {"morris2d_mixed": 5e-5}


@pytest.mark.parametrize("name", list(MIXED_BOXES))
def test_mixed_neighbor_sets_bitexact(oracle, name):
    box = W.mixed(MIXED_BOXES[name]())
    announce(f"mixed neighbour sets: {name}")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        cnt, hs = g.neighbor_signature(mixed_decision=True)
    rc, rh = oracle_signature(oracle, box)
    assert np.array_equal(cnt, rc) and np.array_equal(hs, rh)


def test_mixed_neighbor_sets_perfect_lattice_ties(oracle):
    # every on-sphere lattice neighbour falls in the FP64 re-decision band
    box = W.mixed(W.tait_box_3d(12, brownian=False))
    box.x = W.lattice(3, 12, 1.0, 0.0, 0)
    announce("mixed neighbour sets: perfect lattice")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        g.compute_density()
        # the density pass's FP32 decisions sent the ties to FP64
        pairs, rechecks = g.counters()
        cnt, hs = g.neighbor_signature(mixed_decision=True)
    rc, rh = oracle_signature(oracle, box)
    assert np.array_equal(cnt, rc) and np.array_equal(hs, rh)
    assert rechecks > 0 and pairs == rc.sum()


@pytest.mark.parametrize("name", list(MIXED_BOXES))
@pytest.mark.parametrize("force_type", [3, 4])
def test_mixed_vs_fp64(name, force_type):
    box = MIXED_BOXES[name]()
    box.prm = box.prm.replace(rhs_force_type=force_type)
    mbox = W.mixed(box)
    announce(f"mixed vs fp64: {name} type {force_type}")
    g64 = gpu_evaluate(box, step=9)
    gmx = gpu_evaluate(mbox, step=9)
    rho64, p64 = g64.download_density()
    rhomx, pmx = gmx.download_density()
    f64 = g64.download_force()
    fmx = gmx.download_force()
    pairs64, _ = g64.counters()
    pairsmx, _ = gmx.counters()
    g64.close()
    gmx.close()
    assert pairs64 == pairsmx
    assert rel_err(rhomx, rho64) < 1e-6
    assert vec_err(fmx, f64) < FORCE_BAR.get(name, 1e-5)
    mf = box.m[:, None] * fmx
    assert np.abs(mf.sum(0)).max() < 1e-5 * np.abs(mf).sum()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("force_type", [3, 4])
def test_record_refresh_equals_density_pass(precision, force_type):
    # after a kick (or a separate EOS call) the pair records are rebuilt
    # before the force pass; they must give the forces that the records of a
    # fresh density pass at the same state give, bit for bit
    box = W.tait_box_3d(12, kt=1e-3)
    if precision == "mixed":
        box = W.mixed(box)
    box.prm = box.prm.replace(rhs_force_type=force_type)
    announce(f"record refresh: {precision} type {force_type}")
    with SDPD(box.prm) as a, SDPD(box.prm) as b:
        for g in (a, b):
            g.upload(box.x, box.v, box.m, box.pid)
            g.step = 3
            g.build_list()
            g.compute_density()
            g.compute_interaction()
        f0 = a.download_force()
        for g in (a, b):
            g.integrate_velocity(0.5)
        a.compute_interaction()  # records refreshed from the kicked velocities
        b.compute_density()
        b.compute_interaction()  # records written by the density epilogue
        fa, fb = a.download_force(), b.download_force()
        assert np.array_equal(fa, fb)
        assert not np.array_equal(fa, f0)
        a.compute_pressure()     # the EOS again: records refreshed, same state
        a.compute_interaction()
        assert np.array_equal(a.download_force(), fa)


def test_mixed_trajectory_close_to_fp64():
    box = W.tait_box_3d(13, kt=1e-3)
    announce("mixed trajectory vs fp64")
    out = []
    for b in (box, W.mixed(box)):
        with SDPD(b.prm) as g:
            g.upload(b.x, b.v, b.m, b.pid)
            g.integrate_VV(5)
            out.append(g.download() + (g.download_force(),))
    (x64, v64, f64), (xmx, vmx, fmx) = out
    assert vec_err(vmx, v64) < 1e-5
    assert vec_err(fmx, f64) < 1e-4
    assert np.abs(xmx - x64).max() < 1e-8


def test_mixed_full_size_properties():
    box = W.mixed(W.tait_box_3d(128))
    announce(f"mixed full size: n={box.n}")
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(2)
        f = g.download_force()
        pairs, rechecks = g.counters()
    mf = box.m[:, None] * f
    assert np.abs(mf.sum(0)).max() < 1e-6 * np.abs(mf).sum()
    assert 100 < pairs / box.n < 115
    # the jittered lattice piles ~30 neighbours near r = 3 dx: < 1 FP64 re-decision per particle
    assert rechecks < 0.01 * pairs
    assert np.isfinite(f).all()


def _perturbed_ct(n, seed=9):
    rng = np.random.default_rng(seed)
    ct = np.tile(np.eye(3).reshape(-1), (n, 1))
    a = rng.normal(0, 0.05, size=(n, 3, 3))
    return ct + (a + a.transpose(0, 2, 1)).reshape(n, 9)


def test_mixed_oldroyd_vs_fp64():
    # Pi / n^2 records and the velocity gradient in FP32 (Kahan sums)
    box = W.oldroyd_box_3d(12)
    announce("mixed Oldroyd-B vs fp64")
    ct = _perturbed_ct(box.n)
    g = gpu_evaluate(box, step=2, ct=ct)
    f64 = g.download_force()
    _, vgt64 = g.download_tensors()
    g.close()
    g = gpu_evaluate(W.mixed(box), step=2, ct=ct)
    fmx = g.download_force()
    _, vgtmx = g.download_tensors()
    g.close()
    assert vec_err(fmx, f64) < 1e-5
    assert vec_err(vgtmx, vgt64) < 1e-5


def test_mixed_oldroyd_trajectory_close_to_fp64():
    box = W.oldroyd_box_3d(12)
    announce("mixed Oldroyd-B trajectory")
    out = []
    for b in (box, W.mixed(box)):
        with SDPD(b.prm) as g:
            g.upload(b.x, b.v, b.m, b.pid)
            g.integrate_VV(4)
            x, v = g.download()
            ct, vgt = g.download_tensors()
        out.append((v, ct))
    assert vec_err(out[1][0], out[0][0]) < 1e-5
    assert vec_err(out[1][1] - np.tile(np.eye(3).reshape(-1), (box.n, 1)),
                   out[0][1] - np.tile(np.eye(3).reshape(-1), (box.n, 1))) < 1e-4


@pytest.mark.parametrize("variant", ["tait_fp64", "oldroyd_mixed"])
def test_graph_steps_match_direct_launches(variant):
    # sdpd_step replays two-step graphs; every replay must equal direct
    # launches, also after eager work in between (downloads, a re-sort that
    # flips the ping-pong parity, separate calls)
    box = BOXES["tait3d_sym_noise"]() if variant == "tait_fp64" else W.mixed(W.oldroyd_box_3d(12))
    announce(f"graph replays vs direct launches: {variant}")
    with SDPD(box.prm) as a, SDPD(box.prm) as b:
        b.set_graphs(False)
        for g in (a, b):
            g.upload(box.x, box.v, box.m, box.pid)
        for chunk in (5, 4, 7, 3):
            a.integrate_VV(chunk)
            b.integrate_VV(chunk)
            xa, va = a.download()
            xb, vb = b.download()
            assert np.array_equal(xa, xb) and np.array_equal(va, vb), chunk
            assert a.step == b.step
            if chunk == 4:
                for g in (a, b):
                    g.build_list()
                    g.compute_density()
                    g.compute_interaction()


def _list_capacity_boxes():
    from tests.colloid_helpers import two_spheres
    le = W.tait_box_3d(12, kt=1e-3)
    le.prm = le.prm.replace(le_dim=2, le_v_lower=[-0.3, 0, 0], le_v_upper=[0.3, 0, 0])
    sph, cs = two_spheres(sym=True, noslip=2, rotate=True, omega=[[0, 0, 0.02], [0.01, 0, 0]])
    return {
        "tait3d_fp64": (BOXES["tait3d_sym_noise"](), None),
        "tait3d_mixed": (W.mixed(BOXES["tait3d_sym_noise"]()), None),
        "shear2d": (BOXES["shear2d_full"](), None),
        "oldroyd3d": (W.oldroyd_box_3d(10), None),
        "oldroyd3d_mixed": (W.mixed(W.oldroyd_box_3d(10)), None),
        "lees_edwards": (le, None),
        "colloids": (sph, cs),
    }


@pytest.mark.parametrize("name", ["tait3d_fp64", "tait3d_mixed", "shear2d", "oldroyd3d",
                                  "oldroyd3d_mixed", "lees_edwards", "colloids"])
def test_list_capacity_gives_identical_results(name):
    # capacities below the neighbour counts send every chunk (8) or many
    # (48) to the force pass's stencil rescan; the results must not move
    box, cs = _list_capacity_boxes()[name]
    announce(f"list capacity: {name} n={box.n}")
    outs = []
    for cap in (0, 48, 8):
        with SDPD(box.prm) as g:
            g.set_list_capacity(cap)
            if cs is not None:
                g.set_colloids(cs)
            g.upload(box.x, box.v, box.m, box.pid, box.sid if cs is not None else None)
            if cap:
                assert g.list_capacity() == cap
            g.step = 5
            g.build_list()
            g.compute_density()
            g.compute_interaction()
            outs.append((g.download_density()[0], g.download_force()))
    for rho, f in outs[1:]:
        assert np.array_equal(rho, outs[0][0])
        assert np.array_equal(f, outs[0][1])


def test_list_capacity_arguments():
    box = BOXES["tait3d_sym_kt0"]()
    with SDPD(box.prm) as g:
        for bad in (-8, 4, 12, 8192):
            with pytest.raises(SDPDError):
                g.set_list_capacity(bad)
        g.upload(box.x, box.v, box.m, box.pid)
        assert g.list_capacity() >= 16
        with pytest.raises(SDPDError):
            g.set_list_capacity(64)  # after upload


def test_launch_shapes_give_identical_results():
    box = BOXES["tait3d_sym_noise"]()
    announce("launch shapes and autotuner")
    outs = {}
    for prec_box in (box, W.mixed(box)):
        with SDPD(prec_box.prm) as g:
            g.upload(prec_box.x, prec_box.v, prec_box.m, prec_box.pid)
            g.step = 3
            g.build_list()
            ref = None
            for shape in range(4):
                for order in (1, 0):
                    g.set_tuning(shape, order)
                    g.compute_density()
                    g.compute_interaction()
                    out = (g.download_density()[0], g.download_force())
                    if ref is None:
                        ref = out
                    assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])
            (shape, order), ms = g.autotune(2)
            assert g.tuning() == (shape, order) and ms > 0
            assert np.array_equal(g.download_force(), ref[1])
            outs[prec_box.prm.precision] = ref


@pytest.mark.parametrize("variant", ["tait_fp64", "oldroyd_mixed"])
def test_autotune_between_steps_keeps_trajectory(variant):
    # the autotuner times density + force passes at the current state; after
    # sdpd_step those would see v(t + dt), so it must put back the forces
    # (and act, vgt) of the last step: step(3) + autotune + step(3) equals
    # step(6), bit for bit
    box = BOXES["tait3d_sym_noise"]() if variant == "tait_fp64" else W.mixed(W.oldroyd_box_3d(11))
    announce(f"autotune between steps: {variant}")
    with SDPD(box.prm) as a, SDPD(box.prm) as b:
        for g in (a, b):
            g.upload(box.x, box.v, box.m, box.pid)
        a.integrate_VV(3)
        a.autotune(1)
        a.integrate_VV(3)
        b.set_tuning(*a.tuning())
        b.integrate_VV(6)
        xa, va = a.download()
        xb, vb = b.download()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        assert np.array_equal(a.download_force(), b.download_force())
        if variant == "oldroyd_mixed":
            cta, ga = a.download_tensors()
            ctb, gb = b.download_tensors()
            assert np.array_equal(cta, ctb) and np.array_equal(ga, gb)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_forces_after_step_see_kicked_velocities(precision):
    # sdpd_step ends with the closing kick; a force pass straight after it
    # must rebuild the pair records from the kicked velocities, i.e. equal a
    # fresh density + force pass on the same state
    box = BOXES["tait3d_sym_noise"]()
    if precision == "mixed":
        box = W.mixed(box)
    announce(f"forces after sdpd_step: {precision}")
    with SDPD(box.prm) as a:
        a.upload(box.x, box.v, box.m, box.pid)
        a.integrate_VV(2)
        a.compute_interaction()
        fa = a.download_force()
        a.compute_density()
        a.compute_interaction()
        fb = a.download_force()
    assert np.array_equal(fa, fb)
