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
  dim*dim values per particle.

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
