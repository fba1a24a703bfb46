"""The reference's file formats on the host side (Class_IO.F90, io_*.F90).

Files are written with Fortran edit descriptors, reproduced here:

* statistic file (io_write_statistic.F90:224-254): one record per output
  step, format (I9, k E16.8): step, time, kinetic energy, momentum(1:dim);
* particle output (io_write_particles.F90:220-431): file
  <output_particles_file>[_r]<step, 8 digits>.out ("_r" when the run read
  its particles from a file), one record per particle with format
  (k E16.8): x(1:dim), v(1:dim), rho, m, pid, sid;
* restart particles (io_write_restart_particles.F90:105-215): file
  <restart_particles_file><step, 8 digits>.dat, the particle count in a
  list-directed record, then the records of the particle output; this is
  the layout io_read_particles.F90:162-240 reads back (read_external);
* conformation output (io_write_conformation.F90:175-245): file
  <output_conformation_file>[_r]<step>.out, records (dim*dim E16.8);
  restart conformation (io_write_restart_conformation.F90): the count,
  then those records; io_read_conformation.F90 reads the count and then
  dim*dim values per particle;
* restart physics (io_write_restart_physics.F90:236-953): file
  <restart_physics_file><step, 8 digits>.mcf of key = value lines in the
  layout of physics_config.mcf (reals E16.8, vectors joined by commas),
  step_start and time_start being the step written; read_config reads it
  and the other key = value files.

Readers take the values list-directed, as the reference's READ(unit, *)
does.  Host <-> device copies happen only when a file is written: the
marching loop between output steps stays on the device (sdpd_step).
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

import numpy as np


def fortran_e(value: float, w: int = 16, d: int = 8) -> str:
    """Fortran Ew.d of one value: sign, 0., d digits, exponent E+xx (three
    exponent digits drop the E), right-justified in w columns."""
    v = float(value)
    if math.isnan(v):
        return "NaN".rjust(w)
    if math.isinf(v):
        return ("-Infinity" if v < 0 else "Infinity").rjust(w)
    neg = math.copysign(1.0, v) < 0
    a = abs(v)
    if a == 0.0:
        digits, exp = "0" * d, 0
    else:
        # d significant digits: Python's e-format has one digit before the
        # point, Fortran's E has none, so the exponent moves up by one
        s = f"{a:.{d - 1}e}"
        mant, ex = s.split("e")
        digits = mant.replace(".", "")
        exp = int(ex) + 1
    if -99 <= exp <= 99:
        es = f"E{'-' if exp < 0 else '+'}{abs(exp):02d}"
    else:
        es = f"{'-' if exp < 0 else '+'}{abs(exp):03d}"
    out = ("-" if neg else "") + "0." + digits + es
    if len(out) > w:
        return "*" * w
    return out.rjust(w)


def fortran_i(value: int, w: int) -> str:
    """Fortran Iw: right-justified integer, asterisks when it does not fit."""
    s = str(int(value))
    return s.rjust(w) if len(s) <= w else "*" * w


def list_directed_int(value: int) -> str:
    """An integer written list-directed (the particle count of restart
    files, ppm_io_write_ascii.F90): a blank, then the value, as this
    image's Fortran compiler writes it.  The width of list-directed output
    depends on the compiler; readers take it format-free."""
    return " " + str(int(value))


# ---------------------------------------------------------------------------
# statistic file
# ---------------------------------------------------------------------------

def statistic_record(step: int, time: float, ekin: float, mom: Sequence[float],
                     extra: Iterable[float] = ()) -> str:
    """One record of the statistic file: (I9, k E16.8) of step, time,
    kinetic energy, momentum and any further columns."""
    vals = [time, ekin] + [float(m) for m in mom] + [float(e) for e in extra]
    return fortran_i(step, 9) + "".join(fortran_e(v) for v in vals)


def read_statistic(path: str) -> np.ndarray:
    """The statistic file as an array of records (step, time, ekin, mom...)."""
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.strip():
                rows.append([float(t) for t in line.split()])
    return np.array(rows)


# ---------------------------------------------------------------------------
# particles and conformation
# ---------------------------------------------------------------------------

def output_name(base: str, step: int, read_external: bool = False, ext: str = ".out") -> str:
    """<base>[_r]<step:08d><ext> (io_write_particles.F90:224-232)."""
    return f"{base}{'_r' if read_external else ''}{int(step):08d}{ext}"


def _records(columns: Sequence[np.ndarray]) -> Iterable[str]:
    data = np.concatenate([np.asarray(c, np.float64).reshape(len(c), -1) for c in columns], axis=1)
    for row in data:
        yield "".join(fortran_e(v) for v in row)


def write_particles(path: str, x, v, rho, m, pid, sid, count_line: bool = False) -> None:
    """Particle records (k E16.8): x, v, rho, m, pid, sid, one per line;
    count_line adds the list-directed particle count first (restart
    files)."""
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    with open(path, "w") as fh:
        if count_line:
            fh.write(list_directed_int(n) + "\n")
        for rec in _records([x, v, rho, m, np.asarray(pid, np.float64), np.asarray(sid, np.float64)]):
            fh.write(rec + "\n")


def write_restart_particles(base: str, step: int, x, v, rho, m, pid, sid) -> str:
    """<restart_particles_file><step:08d>.dat: count, then particle records."""
    path = output_name(base, step, False, ".dat")
    write_particles(path, x, v, rho, m, pid, sid, count_line=True)
    return path


def read_particles(path: str, num_dim: int):
    """A read_external particle file (io_read_particles.F90:162-240): the
    count, then x v rho m pid sid per particle.  Returns x, v, rho, m, pid,
    sid; raises ValueError with the reference's messages on short files."""
    with open(path) as fh:
        first = fh.readline()
        if not first.strip():
            raise ValueError("reading particles has problem at line 1")
        n = int(float(first.split()[0]))
        data = np.loadtxt(fh, ndmin=2) if n > 0 else np.zeros((0, 2 * num_dim + 4))
    if data.shape[0] != n:
        raise ValueError("actual number of lines is not equal to number given !")
    d = num_dim
    return (data[:, 0:d], data[:, d:2 * d], data[:, 2 * d], data[:, 2 * d + 1],
            data[:, 2 * d + 2].astype(np.int32), data[:, 2 * d + 3].astype(np.int32))


def write_conformation(path: str, ct, count_line: bool = False) -> None:
    """Conformation records (dim*dim E16.8) in the reference's column-major
    order per particle; count_line for restart files."""
    ct = np.asarray(ct, np.float64)
    with open(path, "w") as fh:
        if count_line:
            fh.write(list_directed_int(ct.shape[0]) + "\n")
        for rec in _records([ct]):
            fh.write(rec + "\n")


def read_conformation(path: str, num_dim: int) -> np.ndarray:
    """io_read_conformation.F90: the count, then dim*dim values per
    particle; (n, dim*dim)."""
    with open(path) as fh:
        first = fh.readline()
        if not first.strip():
            raise ValueError("reading conformation tensor has problem at line 1")
        n = int(float(first.split()[0]))
        vals = np.array(fh.read().split(), dtype=np.float64)
    d2 = num_dim * num_dim
    if vals.size != n * d2:
        raise ValueError("actual number of lines is not equal to number given !")
    return vals.reshape(n, d2)


def config_line(key: str, values, width: int = 0) -> str:
    """One key = value record of a restart physics file: reals as E16.8,
    integers as I<width> (width > 0), logicals as L2, several values joined
    by commas (the (n(A,E16.8)) records of io_write_restart_physics.F90)."""
    vals = list(values) if isinstance(values, (list, tuple, np.ndarray)) else [values]
    if isinstance(vals[0], (bool, np.bool_)):
        return f"{key} =  {'T' if vals[0] else 'F'}"
    if width > 0:
        return f"{key} = " + ",".join(fortran_i(int(v), width) for v in vals)
    return f"{key} = " + ",".join(fortran_e(float(v)) for v in vals)


def restart_physics_lines(prm, step: int, time: float, step_end: int, keys=None) -> list:
    """The restart physics file of a run without colloids at `step`
    (io_write_restart_physics.F90:236-953): the physics keys in the
    reference's order with step_start = step and time_start = time, each
    record followed by a blank line.  keys overrides the values of keys the
    parameter record does not hold (num_part, time_end, ksai, relax keys,
    eval / evec, body_force_d, flow keys, bcdef, shear_type, shear_v,
    shear_freq, wall keys); absent ones get the reference's defaults."""
    keys = dict(keys or {})
    d = prm.num_dim
    k = lambda name, default: keys.get(name, default)  # noqa: E731
    recs = [("num_species", 1, 2), ("num_dim", d, 2),
            ("min_phys", prm.min_phys[:d], 0), ("max_phys", prm.max_phys[:d], 0),
            ("num_part", k("num_part", [0] * d), 4), ("cut_off", prm.cut_off, 0),
            ("dt", prm.dt, 0), ("step_start", step, 10), ("step_end", step_end, 10),
            ("time_start", time, 0), ("time_end", k("time_end", 0.0), 0),
            ("rho", prm.rho, 0), ("eta", prm.eta, 0), ("eta_coef", prm.eta_coef, 0),
            ("ksai", k("ksai", 0.0), 0), ("kt", prm.kt_effective, 0), ("c", prm.c, 0),
            ("rho_ref", prm.rho_ref, 0), ("gamma", prm.gamma, 0),
            ("relax_type", k("relax_type", 0), 10), ("dt_relax", k("dt_relax", 0.0), 0),
            ("step_relax", k("step_relax", 0), 10), ("time_relax", k("time_relax", 0.0), 0),
            ("disorder_level", k("disorder_level", 0.0), 0), ("kt_relax", k("kt_relax", 0.0), 0),
            ("c_relax", k("c_relax", 0.0), 0)]
    if not prm.Newtonian:
        recs += [("tau", prm.tau, 0), ("n_p", prm.n_p, 0), ("kt_p", prm.kt_p, 0),
                 ("eigen_dynamics", bool(k("eigen_dynamics", False)), 0),
                 ("eval", k("eval", [0.0] * d), 0), ("eval_tolerance", k("eval_tolerance", 0.0), 0),
                 ("evec", k("evec", [0.0] * d * d), 0),
                 ("evec_normalize", bool(k("evec_normalize", False)), 0),
                 ("evec_tolerance", k("evec_tolerance", 0.0), 0)]
    recs += [("body_force_type", prm.body_force_type, 3), ("body_force", prm.body_force[:d], 0),
             ("body_force_d", k("body_force_d", [0.0] * d), 0),
             ("flow_direction", k("flow_direction", 1), 5), ("flow_width", k("flow_width", 0.0), 0),
             ("flow_v", k("flow_v", 0.0), 0), ("flow_adjust_freq", k("flow_adjust_freq", 0), 5),
             ("num_colloid", 0, 3), ("bcdef", k("bcdef", [1] * 2 * d), 2),
             ("shear_type", k("shear_type", [0] * 2 * d), 2),
             ("shear_v", k("shear_v", [0.0] * 2 * d * (d - 1)), 0),
             ("shear_freq", k("shear_freq", [0.0] * 2 * d), 0),
             ("wall_rho_type", k("wall_rho_type", 0), 3), ("wall_noslip", k("wall_noslip", 0), 2)]
    bar = "#---------------------------------------------"
    lines = [bar, "# physics keys of the run at this step, read as",
             "# the keys of physics_config.mcf", bar, " "]
    for name, val, width in recs:
        lines += [config_line(name, val, width), " "]
    lines += [bar, "# end of the restart physics keys", bar]
    return lines


def restart_physics_name(base: str, step: int) -> str:
    """<restart_physics_file><step:08d>.mcf"""
    return f"{base}{step:08d}.mcf"


def read_config(path: str) -> dict:
    """key = value lines of a configuration file (ctrl.mcf,
    physics_config.mcf, io_config.mcf, a restart physics file): '#' starts
    a comment, keys are case-insensitive (lower-cased here), values stay
    strings without surrounding blanks, and a key given twice keeps its
    last value (io_read_ctrl.F90:157-357 reads line by line the same way)."""
    out = {}
    with open(path) as fh:
        for line in fh:
            line = line.split("#", 1)[0]
            if "=" not in line:
                continue
            key, val = line.split("=", 1)
            key = key.strip().lower()
            if key:
                out[key] = val.strip()
    return out


def config_reals(value: str) -> list:
    """The numbers of a comma- or blank-separated value (list-directed)."""
    return [float(t) for t in value.replace(",", " ").split()]


def write_statistic(path: str, records: Iterable[str]) -> None:
    with open(path, "w") as fh:
        for r in records:
            fh.write(r + "\n")



def read_binary_records(path: str, ncol: int, count_first: bool = False):
    """UNFORMATTED files (ppm_io_write_binary.F90 through Fortran sequential
    records with 4-byte length markers): [count pair,] the value count, then
    the values; returns (count or None, values as (n, ncol))."""
    raw = open(path, "rb").read()
    pos = 0
    recs = []
    while pos < len(raw):
        ln = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        recs.append(raw[pos + 4: pos + 4 + ln])
        end = int(np.frombuffer(raw, np.int32, 1, pos + 4 + ln)[0])
        if end != ln:
            raise ValueError("corrupt record markers")
        pos += ln + 8
    count = None
    if count_first:
        count = int(np.frombuffer(recs[1], np.int32)[0])
        recs = recs[2:]
    total = int(np.frombuffer(recs[0], np.int32)[0])
    vals = np.frombuffer(recs[1], np.float64)
    if vals.size != total or total % ncol:
        raise ValueError("record length does not match the value count")
    return count, vals.reshape(-1, ncol)
