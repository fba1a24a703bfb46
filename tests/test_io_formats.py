"""The reference's file formats (mcf_amd/io.py) against a Fortran compiler's
own edit-descriptor output, and round trips through the readers."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mcf_amd import io as IO

FC = shutil.which("amdflang") or ("/opt/rocm/llvm/bin/amdflang"
                                 if os.path.exists("/opt/rocm/llvm/bin/amdflang") else None)
needs_fc = pytest.mark.skipif(FC is None, reason="no Fortran compiler")


def fortran_lines(tmp_path, values, fmt):
    """Write each value with a Fortran format through the compiler."""
    src = tmp_path / "w.f90"
    data = tmp_path / "in.txt"
    np.savetxt(data, values, fmt="%.17e")
    src.write_text(f"""program w
  implicit none
  double precision :: v
  integer :: ios
  open(10, file='{data}', status='old', action='read')
  do
    read(10, *, iostat=ios) v
    if (ios /= 0) exit
    write(*, '({fmt})') v
  end do
end program
""")
    exe = tmp_path / "w"
    subprocess.run([FC, str(src), "-o", str(exe)], check=True, capture_output=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()


@needs_fc
def test_e16_8_matches_fortran(tmp_path):
    rng = np.random.default_rng(7)
    vals = np.concatenate([
        rng.normal(size=200) * 10.0 ** rng.integers(-30, 30, size=200),
        [0.0, -0.0, 1.0, -1.0, 0.5, 9.999999996, 9.999999994, 1e-99, 1e-100, 1e99, 1e100, -2.5e150,
         123456789.123, 5e-324, 1.7976931348623157e308, 0.0128, 1.0 / 3.0],
    ])
    ref = fortran_lines(tmp_path, vals, "E16.8")
    got = [IO.fortran_e(v) for v in vals]
    assert got == ref


@needs_fc
def test_statistic_record_matches_fortran(tmp_path):
    src = tmp_path / "s.f90"
    src.write_text("""program s
  write(*, '(I9,5E16.8)') 832, 10.6496d0, 1.25d-3, -2.0d-5, 3.0d0, 0.0d0
end program
""")
    exe = tmp_path / "s"
    subprocess.run([FC, str(src), "-o", str(exe)], check=True, capture_output=True)
    ref = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.rstrip("\n")
    assert IO.statistic_record(832, 10.6496, 1.25e-3, [-2.0e-5, 3.0, 0.0]) == ref


def test_fortran_i_and_overflow():
    assert IO.fortran_i(832, 9) == "      832"
    assert IO.fortran_i(1234567890, 9) == "*********"
    assert IO.fortran_e(1e300, 8, 3) == "********" or len(IO.fortran_e(1e300, 8, 3)) == 8


def test_restart_particles_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n, d = 57, 3
    x = rng.uniform(0, 10, size=(n, d))
    v = rng.normal(size=(n, d))
    rho = rng.uniform(0.9, 1.1, size=n)
    m = np.full(n, 1.0)
    pid = np.arange(1, n + 1)
    sid = (np.arange(n) % 3).astype(np.int32)
    path = IO.write_restart_particles(str(tmp_path / "mcf_restart_particles"), 1200, x, v, rho, m, pid, sid)
    assert os.path.basename(path) == "mcf_restart_particles00001200.dat"
    xr, vr, rr, mr, pr, sr = IO.read_particles(path, d)
    # E16.8 keeps 8 significant digits: relative rounding up to 5e-8
    assert np.allclose(xr, x, rtol=5.1e-8, atol=0) and np.allclose(vr, v, rtol=5.1e-8, atol=1e-300)
    assert np.array_equal(pr, pid) and np.array_equal(sr, sid)
    lines = open(path).read().splitlines()
    assert lines[0].split() == [str(n)] and len(lines) == n + 1
    assert all(len(l) == 16 * (2 * d + 4) for l in lines[1:])


def test_particle_output_name_and_layout(tmp_path):
    assert IO.output_name("mcf_particles", 200) == "mcf_particles00000200.out"
    assert IO.output_name("mcf_particles", 200, read_external=True) == "mcf_particles_r00000200.out"
    p = tmp_path / "o.out"
    IO.write_particles(str(p), [[1.0, 2.0]], [[0.0, -1.0]], [1.0], [0.5], [7], [0])
    assert open(p).read() == "".join(IO.fortran_e(v) for v in [1, 2, 0, -1, 1, 0.5, 7, 0]) + "\n"


def test_conformation_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(5)
    ct = rng.normal(size=(11, 9))
    p = tmp_path / "c.dat"
    IO.write_conformation(str(p), ct, count_line=True)
    back = IO.read_conformation(str(p), 3)
    assert np.allclose(back, ct, rtol=5.1e-8, atol=1e-300)
    short = tmp_path / "short.dat"
    short.write_text(" 12\n" + "\n".join(open(p).read().splitlines()[1:]) + "\n")
    with pytest.raises(ValueError):
        IO.read_conformation(str(short), 3)


def test_statistic_file_round_trip(tmp_path):
    p = tmp_path / "stat.dat"
    IO.write_statistic(str(p), [IO.statistic_record(s, s * 0.5, 1.0 / (s + 1), [0.1 * s, -0.2, 0.0])
                                for s in range(5)])
    st = IO.read_statistic(str(p))
    assert st.shape == (5, 6)
    assert list(st[:, 0]) == [0, 1, 2, 3, 4]


@needs_fc
def test_restart_physics_records_match_fortran(tmp_path):
    # the records of io_write_restart_physics.F90: (A,E16.8), (3(A,E16.8)),
    # (A,I10), (3(A,I4)), (A,L) with gfortran's default width 2
    src = tmp_path / "r.f90"
    src.write_text("""program r
  character(len=512) :: cbuf
  write(cbuf, '(A,E16.8)') 'dt = ', 5.0d-5
  write(*, '(A)') trim(cbuf)
  write(cbuf, '(3(A,E16.8))') 'min_phys = ', 0.0d0, ',', -1.5d0, ',', 128.0d0
  write(*, '(A)') trim(cbuf)
  write(cbuf, '(A,I10)') 'step_start = ', 1234
  write(*, '(A)') trim(cbuf)
  write(cbuf, '(3(A,I4))') 'num_part = ', 64, ',', 64, ',', 1
  write(*, '(A)') trim(cbuf)
  write(cbuf, '(A,L2)') 'eigen_dynamics = ', .false.
  write(*, '(A)') trim(cbuf)
end program
""")
    exe = tmp_path / "r"
    subprocess.run([FC, str(src), "-o", str(exe)], check=True, capture_output=True)
    ref = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    got = [IO.config_line("dt", 5e-5), IO.config_line("min_phys", [0.0, -1.5, 128.0]),
           IO.config_line("step_start", 1234, 10), IO.config_line("num_part", [64, 64, 1], 4),
           IO.config_line("eigen_dynamics", False)]
    assert got == ref


def test_restart_physics_reads_back(tmp_path):
    from mcf_amd import workloads as W

    box = W.oldroyd_box_3d(10)
    prm = box.prm.replace(body_force_type=1, body_force=[1e-3, 0.0, -2e-3])
    t = 0.0
    for _ in range(7):
        t = t + prm.dt
    path = tmp_path / IO.restart_physics_name("mcf_restart_physics", 7)
    assert path.name == "mcf_restart_physics00000007.mcf"
    path.write_text("\n".join(IO.restart_physics_lines(prm, 7, t, 20, {"num_part": [10, 10, 10]})) + "\n")
    cfg = IO.read_config(str(path))
    assert int(cfg["step_start"]) == 7 and int(cfg["step_end"]) == 20
    # E16.8 keeps eight significant digits
    assert float(cfg["time_start"]) == float(f"{t:.7e}")
    assert IO.config_reals(cfg["max_phys"]) == [10.0, 10.0, 10.0]
    assert IO.config_reals(cfg["body_force"]) == [1e-3, 0.0, -2e-3]
    for key in ("cut_off", "dt", "rho", "eta", "eta_coef", "c", "rho_ref", "gamma", "tau", "n_p", "kt_p"):
        assert float(cfg[key]) == float(f"{getattr(prm, key):.7e}"), key
    assert float(cfg["kt"]) == float(f"{prm.kt_effective:.7e}")
    assert cfg["eigen_dynamics"] == "F" and cfg["num_colloid"] == "0"
    assert [int(v) for v in IO.config_reals(cfg["bcdef"])] == [1] * 6
    assert IO.config_reals(cfg["num_part"]) == [10.0, 10.0, 10.0]
