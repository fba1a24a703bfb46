"""The Fortran host (mcf_amd/fortran/mcf_sdpd): MCF-style config files in,
marching_integrate_VV call order through the C-ABI, read_external particle
files out.  The GPU tests check it step for step against the Python
binding of the same library."""
import os
import subprocess

import numpy as np
import pytest

from mcf_amd import build as B
from mcf_amd import workloads as W
from mcf_amd.sdpd import SDPD

HOST = B.FORTRAN_HOST

needs_host = pytest.mark.skipif(not os.path.exists(HOST), reason="Fortran host not built (no amdflang)")


def announce(msg):
    print(msg, flush=True)


def write_configs(tmp, box, nsteps, fused=False, precision="fp64", stat_freq=2, extra=""):
    prm = box.prm
    d = prm.num_dim
    W.write_external(os.path.join(tmp, "particles.dat"), box)
    with open(os.path.join(tmp, "ctrl.mcf"), "w") as fh:
        fh.write(f"""# control file
mcf_physics_config_file = physics_config.mcf
mcf_io_config_file      = io_config.mcf
read_external = T
kernel_type = {prm.kernel_type}
symmetry = {'T' if prm.symmetry else 'F'}
rhs_density_type = {prm.rhs_density_type}
stateEquation_type = {prm.stateEquation_type}
Newtonian = {'T' if prm.Newtonian else 'F'}
Brownian = {'T' if prm.Brownian else 'F'}
random_seed = {prm.random_seed}
rhs_force_type = {prm.rhs_force_type}
integrate_type = 2
precision = {precision}
fused_step = {'T' if fused else 'F'}
""")
    fmt = lambda v: ", ".join(repr(float(a)) for a in v[:d])  # noqa: E731
    with open(os.path.join(tmp, "physics_config.mcf"), "w") as fh:
        fh.write(f"""num_dim = {d}
min_phys = {fmt(prm.min_phys)}
max_phys = {fmt(prm.max_phys)}
cut_off = {prm.cut_off!r}
dt = {prm.dt!r}
step_end = {nsteps}
rho = {prm.rho!r}
rho_ref = {prm.rho_ref!r}
c = {prm.c!r}
gamma = {prm.gamma!r}
eta = {prm.eta!r}
eta_coef = {prm.eta_coef!r}
kt = {prm.kt!r}
{extra}""")
    with open(os.path.join(tmp, "io_config.mcf"), "w") as fh:
        fh.write(f"""read_particles_file = particles.dat
statistic_file = stat.dat
statistic_freq_step = {stat_freq}
output_particles_file = parts
""")
    return os.path.join(tmp, "ctrl.mcf")


def run_host(ctrl, cwd):
    return subprocess.run([HOST, ctrl], cwd=cwd, capture_output=True, text=True, timeout=120)


@needs_host
def test_host_links_the_library():
    out = subprocess.run(["readelf", "-d", HOST], capture_output=True, text=True).stdout
    assert "libsdpd.so" in out


@needs_host
def test_host_reports_missing_config(tmp_path):
    r = run_host(str(tmp_path / "nope.mcf"), str(tmp_path))
    assert r.returncode != 0
    assert "cannot open control file" in r.stdout


@needs_host
def test_host_rejects_other_integrators(tmp_path):
    box = W.tait_box_3d(6, kt=1e-3)
    ctrl = write_configs(str(tmp_path), box, 2)
    with open(ctrl, "a") as fh:
        fh.write("integrate_type = 1\n")
    r = run_host(ctrl, str(tmp_path))
    assert r.returncode != 0 and "integrate_type" in r.stdout


def read_parts(path, d):
    with open(path) as fh:
        n = int(fh.readline())
        data = np.loadtxt(fh, ndmin=2)
    assert data.shape[0] == n
    return data[:, :d], data[:, d:2 * d]


@pytest.mark.gpu
@needs_host
@pytest.mark.parametrize("fused", [False, True])
def test_host_matches_python_binding(tmp_path, fused):
    box = W.tait_box_3d(12, kt=1e-3)
    nsteps = 4
    announce(f"Fortran host, fused={fused}, {nsteps} steps")
    ctrl = write_configs(str(tmp_path), box, nsteps, fused=fused)
    r = run_host(ctrl, str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    x, v = read_parts(str(tmp_path / f"parts{nsteps:08d}.dat"), 3)
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(nsteps)
        xr, vr = g.download()
        ek, mom = g.compute_statistic()
    # the particle file carries 18 significant digits: doubles round-trip
    assert np.array_equal(x, xr) and np.array_equal(v, vr)
    stat = np.loadtxt(str(tmp_path / "stat.dat"), ndmin=2)
    assert list(stat[:, 0]) == [0, 2, 4]
    assert stat[-1, 2] == pytest.approx(ek, rel=1e-7)


def shear_v_line(d, le_dim, v_lo, v_hi):
    """The reference's shear_v key (io_read_physics.F90:1378-1402): per face,
    its velocity components along the face, faces x1, x2, y1, y2[, z1, z2]."""
    vals = []
    for face in range(2 * d):
        along = [c for c in range(d) if c != face // 2]
        vel = [0.0] * 3
        if face // 2 == le_dim - 1:
            vel = v_lo if face % 2 == 0 else v_hi
        vals += [vel[c] for c in along]
    return "shear_v = " + ", ".join(repr(float(a)) for a in vals)


def le_extra(d, le_dim, v_lo, v_hi):
    bc = [1] * (2 * d)
    bc[2 * le_dim - 2] = bc[2 * le_dim - 1] = 10
    return "bcdef = " + ", ".join(map(str, bc)) + "\n" + shear_v_line(d, le_dim, v_lo, v_hi) + "\n"


@needs_host
@pytest.mark.parametrize("bc,msg", [("10, 1, 1, 1, 1, 1", "faces of x"),
                                    ("1, 1, 10, 1, 1, 1", "both faces"),
                                    ("1, 1, 10, 10, 10, 10", "one dimension"),
                                    ("1, 1, 7, 7, 1, 1", "only periodic")])
def test_host_rejects_bad_faces(tmp_path, bc, msg):
    box = W.tait_box_3d(10, kt=1e-3)
    ctrl = write_configs(str(tmp_path), box, 2, extra=f"bcdef = {bc}\n")
    r = run_host(ctrl, str(tmp_path))
    assert r.returncode != 0 and msg in r.stdout, r.stdout


@pytest.mark.gpu
@needs_host
@pytest.mark.parametrize("le_dim,fused", [(2, False), (3, True)])
def test_host_lees_edwards_matches_python_binding(tmp_path, le_dim, fused):
    box = W.tait_box_3d(12, kt=1e-3)
    u = 3.0
    v_lo, v_hi = [-u, 0.0, 0.0], [u, 0.0, 0.0]
    box.prm = box.prm.replace(dt=box.prm.dt * 100)
    v = box.v.copy()
    v[:, 0] += 2 * u / 12.0 * (box.x[:, le_dim - 1] - 6.0)
    box.v = v
    nsteps = 4
    announce(f"Fortran host with Lees-Edwards faces on dim {le_dim}, fused={fused}")
    ctrl = write_configs(str(tmp_path), box, nsteps, fused=fused, extra=le_extra(3, le_dim, v_lo, v_hi))
    r = run_host(ctrl, str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    x, vh = read_parts(str(tmp_path / f"parts{nsteps:08d}.dat"), 3)
    with SDPD(box.prm.replace(le_dim=le_dim, le_v_lower=v_lo, le_v_upper=v_hi)) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(nsteps)
        xr, vr = g.download()
    assert np.array_equal(x, xr) and np.array_equal(vh, vr)


def colloid_extra(cs, d):
    """num_colloid and the coll_* keys, once per colloid as the reference
    reads them (io_read_physics.F90:1242-1350)."""
    lines = [f"num_colloid = {cs.num}", f"coll_noslip = {cs.noslip}", f"coll_rho_type = {cs.rho_type}",
             f"coll_translate = {'T' if cs.translate else 'F'}",
             f"coll_rotate = {'T' if cs.rotate else 'F'}", f"coll_sub_time_step = {cs.sub_steps}",
             f"integrate_colloid_ab = {cs.ab}", "integrate_colloid_type = 2"]
    for k in range(cs.num):
        f = lambda a: ", ".join(repr(float(u)) for u in a)  # noqa: E731
        lines += ["coll_shape = 2", f"coll_radius = {f([cs.radius[k]] + [0.0] * (d - 1))}",
                  f"coll_x = {f(cs.x[k][:d])}", f"coll_v = {f(cs.v[k][:d])}",
                  f"coll_omega = {f(cs.omega[k])}"]
    return "\n".join(lines) + "\n"


@pytest.mark.gpu
@needs_host
@pytest.mark.parametrize("fused", [False, True])
def test_host_colloids_match_python_binding(tmp_path, fused):
    from tests.colloid_helpers import two_spheres
    box, cs = two_spheres(n_side=16, radius=4.0, kt=1e-3, translate=True, rotate=True,
                          omega=[[0.0, 0.0, 0.02], [0.01, 0.0, 0.0]], v=[[0.01, 0, 0], [0, 0.02, 0]])
    # the host derives dx = cut_off / 3 when num_part is absent: dout 0.05
    assert abs(cs.dout - 0.05) < 1e-15
    box.prm = box.prm.replace(dt=box.prm.dt * 20)
    nsteps = 4
    announce(f"Fortran host with colloids, fused={fused}")
    ctrl = write_configs(str(tmp_path), box, nsteps, fused=fused, extra=colloid_extra(cs, 3))
    r = run_host(ctrl, str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    x, vh = read_parts(str(tmp_path / f"parts{nsteps:08d}.dat"), 3)
    with SDPD(box.prm) as g:
        g.set_colloids(cs)
        g.upload(box.x, box.v, box.m, box.pid, box.sid)
        g.integrate_VV(nsteps)
        xr, vr = g.download()
    assert np.array_equal(x, xr) and np.array_equal(vh, vr)
