"""The Fortran host (mcf_amd/fortran/mcf_sdpd): MCF-style config files in,
marching_integrate_VV call order through the C-ABI, read_external particle
files out.  The GPU tests check it step for step against the Python
binding of the same library."""
import os
import subprocess

import numpy as np
import pytest

from mcf_amd import build as B
from mcf_amd import io as IO
from mcf_amd import workloads as W
from mcf_amd.sdpd import SDPD

HOST = B.FORTRAN_HOST

needs_host = pytest.mark.skipif(not os.path.exists(HOST), reason="Fortran host not built (no amdflang)")


def announce(msg):
    print(msg, flush=True)


def write_configs(tmp, box, nsteps, fused=False, precision="fp64", stat_freq=2, extra="",
                  out_fmt="UNFORMATTED", io_extra=""):
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
output_particles_fmt = {out_fmt}
restart_particles_file = restart
{io_extra}""")
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
    """Binary particle output of the host (full precision): x, v."""
    _, data = IO.read_binary_records(path, 2 * d + 4)
    return data[:, :d], data[:, d:2 * d]


def parts_name(nsteps):
    # read_external runs name their outputs <file>_r<step>.out
    return f"parts_r{nsteps:08d}.out"


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
    x, v = read_parts(str(tmp_path / parts_name(nsteps)), 3)
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(nsteps)
        xr, vr = g.download()
        ek, mom = g.compute_statistic()
    # binary particle output: doubles round-trip
    assert np.array_equal(x, xr) and np.array_equal(v, vr)
    # statistic records (I9, k E16.8) of step, time, kinetic energy, momentum
    lines = open(tmp_path / "stat.dat").read().splitlines()
    t = 0.0
    for _ in range(nsteps):
        t = t + box.prm.dt
    assert lines[-1] == IO.statistic_record(nsteps, t, ek, mom)
    stat = IO.read_statistic(str(tmp_path / "stat.dat"))
    assert list(stat[:, 0]) == [0, 2, 4]
    # the restart file of the last step reads back as a read_external file
    xs, vs, _, _, pid, sid = IO.read_particles(str(tmp_path / f"restart{nsteps:08d}.dat"), 3)
    assert np.allclose(xs, xr, rtol=5.1e-8) and np.array_equal(pid, box.pid)


@needs_host
@pytest.mark.gpu
@pytest.mark.parametrize("ft,bft", [(2, 3), (1, 2)])
def test_host_force_laws_and_body_forces(tmp_path, ft, bft):
    # the keys rhs_force_type (1 Morris, 2 Espanol) and body_force_type
    # (2 two-way, 3 sine) reach the library through the host unchanged
    box = W.tait_box_3d(12, kt=1e-3, brownian=ft == 2)
    box.prm = box.prm.replace(rhs_force_type=ft, body_force_type=bft, body_force=[0.5, 0.0, 0.0])
    nsteps = 4
    announce(f"Fortran host, rhs_force_type {ft}, body_force_type {bft}")
    extra = f"body_force_type = {bft}\nbody_force = 0.5, 0.0, 0.0\n"
    ctrl = write_configs(str(tmp_path), box, nsteps, extra=extra)
    r = run_host(ctrl, str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    x, v = read_parts(str(tmp_path / parts_name(nsteps)), 3)
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(nsteps)
        xr, vr = g.download()
    assert np.array_equal(x, xr) and np.array_equal(v, vr)


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
    x, vh = read_parts(str(tmp_path / parts_name(nsteps)), 3)
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
    x, vh = read_parts(str(tmp_path / parts_name(nsteps)), 3)
    with SDPD(box.prm) as g:
        g.set_colloids(cs)
        g.upload(box.x, box.v, box.m, box.pid, box.sid)
        g.integrate_VV(nsteps)
        xr, vr = g.download()
    assert np.array_equal(x, xr) and np.array_equal(vh, vr)


@pytest.mark.gpu
@needs_host
def test_host_formatted_outputs(tmp_path):
    # FORMATTED particle output: (k E16.8) records identical to the Python
    # writer's on the same (bitwise equal) state; Oldroyd-B conformation
    # read from a file and written back
    box = W.oldroyd_box_3d(10, kt=1e-3)
    nsteps = 3
    rng = np.random.default_rng(2)
    ct = np.tile(np.eye(3).reshape(-1), (box.n, 1)) + 0.01 * rng.normal(size=(box.n, 9))
    IO.write_conformation(str(tmp_path / "conf_in.dat"), ct, count_line=True)
    ct_in = IO.read_conformation(str(tmp_path / "conf_in.dat"), 3)
    announce("Fortran host formatted outputs (Oldroyd-B)")
    extra = f"tau = {box.prm.tau!r}\nn_p = {box.prm.n_p!r}\nkt_p = {box.prm.kt_p!r}\n"
    ctrl = write_configs(str(tmp_path), box, nsteps, extra=extra, out_fmt="FORMATTED",
                         io_extra="read_conformation_file = conf_in.dat\noutput_particles_freq_step = 1\n"
                                  "output_conformation_file = conf\noutput_conformation_freq_step = 3\n"
                                  "restart_conformation_file = rconf\n")
    with open(ctrl, "a") as fh:
        fh.write("Newtonian = F\n")
    r = run_host(ctrl, str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.upload_ct(ct_in)
        g.integrate_VV(nsteps)
        xr, vr = g.download()
        rho, _ = g.download_density()
        ctr, _ = g.download_tensors()
    ref = tmp_path / "ref.out"
    IO.write_particles(str(ref), xr, vr, rho, box.m, box.pid, box.sid)
    assert open(tmp_path / f"parts_r{nsteps:08d}.out").read() == open(ref).read()
    assert (tmp_path / "parts_r00000001.out").exists()
    refc = tmp_path / "refc.out"
    IO.write_conformation(str(refc), ctr)
    assert open(tmp_path / f"conf_r{nsteps:08d}.out").read() == open(refc).read()
    assert np.allclose(IO.read_conformation(str(tmp_path / f"rconf{nsteps:08d}.dat"), 3), ctr,
                       rtol=5.1e-8, atol=1e-300)


def _resume_dir(tmp, box, src, step, nsteps, extra=""):
    """A resumed run's files: the first run's control keys, its restart
    physics file as the physics config and its restart particles as the
    read_external file (marching_marching.F90:325-331 takes step_start and
    time_start from the physics config)."""
    import shutil

    ctrl = write_configs(str(tmp), box, nsteps, fused=True, extra=extra, stat_freq=1)
    for name in (IO.restart_physics_name("mcf_restart_physics", step), f"restart{step:08d}.dat"):
        shutil.copy(os.path.join(src, name), os.path.join(tmp, name))
    text = open(ctrl).read().replace("physics_config.mcf",
                                     IO.restart_physics_name("mcf_restart_physics", step))
    open(ctrl, "w").write(text)
    io = os.path.join(tmp, "io_config.mcf")
    text = open(io).read().replace("particles.dat", f"restart{step:08d}.dat")
    open(io, "w").write(text)
    return ctrl


@pytest.mark.gpu
@needs_host
@pytest.mark.parametrize("newtonian", [True, False])
def test_host_resumes_from_its_restart_files(tmp_path, newtonian):
    # the first run writes restart particles and restart physics at step 3;
    # a second run started from those files alone continues to step 6
    box = W.tait_box_3d(12, kt=1e-3) if newtonian else W.oldroyd_box_3d(12, kt=1e-3)
    n1, n2 = 3, 6
    extra = "" if newtonian else f"tau = {box.prm.tau!r}\nn_p = {box.prm.n_p!r}\nkt_p = {box.prm.kt_p!r}\n"
    first = tmp_path / "first"
    first.mkdir()
    announce(f"Fortran host restart at step {n1}, newtonian={newtonian}")
    ctrl = write_configs(str(first), box, n2, fused=True, extra=extra,
                         io_extra=f"restart_freq_step = {n1}\n")
    if not newtonian:
        with open(ctrl, "a") as fh:
            fh.write("Newtonian = F\n")
    r = run_host(ctrl, str(first))
    assert r.returncode == 0, r.stdout + r.stderr
    t = 0.0
    for _ in range(n1):
        t = t + box.prm.dt
    # the file the reference's writer makes for this state, record for record
    phys = first / IO.restart_physics_name("mcf_restart_physics", n1)
    assert phys.read_text().splitlines() == IO.restart_physics_lines(box.prm, n1, t, n2)
    cfg = IO.read_config(str(phys))
    assert int(cfg["step_start"]) == n1 and int(cfg["step_end"]) == n2

    second = tmp_path / "second"
    second.mkdir()
    ctrl2 = _resume_dir(str(second), box, str(first), n1, n2, extra)
    if not newtonian:
        # the conformation of the restart step comes with it
        import shutil

        shutil.copy(first / f"mcf_restart_conformation{n1:08d}.dat", second / "rconf.dat")
        with open(ctrl2, "a") as fh:
            fh.write("Newtonian = F\n")
        with open(second / "io_config.mcf", "a") as fh:
            fh.write("read_conformation_file = rconf.dat\n")
    r = run_host(ctrl2, str(second))
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{n2 - n1} steps" in r.stdout
    x, v = read_parts(str(second / parts_name(n2)), 3)

    # the library resumed from the same restart data at step n1: the host
    # resumes at the right step (noise keys, times), bit for bit
    xs, vs, _, ms, pid, _ = IO.read_particles(str(first / f"restart{n1:08d}.dat"), 3)
    with SDPD(box.prm) as g:
        g.upload(xs, vs, ms, pid)
        if not newtonian:
            g.upload_ct(IO.read_conformation(str(first / f"mcf_restart_conformation{n1:08d}.dat"), 3))
        g.step = n1
        g.integrate_VV(n2 - n1)
        xr, vr = g.download()
        ek, mom = g.compute_statistic()
    assert np.array_equal(x, xr) and np.array_equal(v, vr)
    # the statistic file continues from step_start at time_start
    tr = float(cfg["time_start"])
    for _ in range(n2 - n1):
        tr = tr + box.prm.dt
    lines = open(second / "stat.dat").read().splitlines()
    assert lines[0].split()[0] == str(n1)
    assert lines[-1] == IO.statistic_record(n2, tr, ek, mom)
    # and it stays with the uninterrupted run: the restart records keep 8
    # digits, and the resumed prologue evaluates the forces at v(t_n1)
    # rather than at the half-step velocity (marching_marching.F90 does the
    # initial evaluation on every start)
    xf, vf = read_parts(str(first / parts_name(n2)), 3)
    dx = x - xf
    dx -= 12.0 * np.round(dx / 12.0)
    assert np.abs(dx).max() < 1e-6
    assert np.abs(v - vf).max() < 1e-6 * max(np.abs(vf).max(), 1e-3) + 1e-8
