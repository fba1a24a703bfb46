#!/usr/bin/env python3
"""Benchmark of the SDPD step on one MI355X.

Headline (default): config C2, the 3D periodic Tait box of 128^3 =
2,097,152 particles, dx = 1, quintic kernel with cut-off 3 dx, gamma 7,
c = 10, eta = 0.1, Maxwellian velocities and thermal noise at kT = 1e-3,
dt = 5e-5, in plain FP64, over its 200 steps.  One step = one full
velocity-Verlet step of every particle: step boundary (kicks, drift, wrap,
cell keys), cell sort, density + EOS, pair forces with noise.

In the same run (unless --no-extras) the other single-GPU configurations
are timed over their own step counts and reported under "configs":
  C3  the same box in mixed precision (FP32 pair arithmetic, FP64 state),
      200 steps, with its particle-steps/s ratio to C2;
  C4  the box with 512 rigid spheres of radius 5 dx (boundary-particle
      shells, Morris no-slip, translating and rotating) under Lees-Edwards
      shear at rate 0.5 on the y faces, FP64, 200 steps;
  C5  256 independent Oldroyd-B boxes of 2k-60k particles (4.3M in total)
      in one context, mixed precision, 500 steps, with its ratio to C2.
--workload ensemble / colloids makes C5 / C4 the headline instead.

Prints one JSON line.  value = particle-steps/s of the whole job with the
state resident in HBM.  roofline = the pair-force kernel's stencil-model
bytes per launch over its average launch time (HIP events on the context
stream), against the 8 TB/s HBM peak; the unit that binds it and its VALU
issue fraction come from the rocprofv3 counters committed under profiles/
(named with their source).  copy_gbs = a plain 1 GiB device copy, measured
in the run.  cpu_baseline = the FP64 oracle (a C restatement of the
reference path), built -O3 -march=native on this host, timed on one VV step
of C2 itself with all the host threads OpenMP gives it and on one core.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X: 256 CUs; one wave64 VALU instruction per CU per clock at the
# 2.4 GHz peak engine clock (the published FP64 vector rate, 78.6 TFLOP/s,
# is that rate of FMAs)
VALU_PEAK_INSTS = 256 * 2.4e9
COUNTERS_FILE = os.path.join(ROOT, "profiles", "counters_tait128.json")
# Stencil-model algorithmic bytes per particle and launch (27 neighbour-cell
# records per particle): FP64 force record 68 B -> 27*68 + f 24 = 1860 B;
# mixed 36 B -> 27*36 + 24 = 996 B.
FORCE_BYTES = {"fp64": 1860.0, "mixed": 996.0}
# Oldroyd-B adds Pi to the force record: 6 components, FP64 48 B / FP32 24 B
# (FP64 record 68 + 48 = 116 B, mixed 36 + 24 = 60 B), plus f 24 B.
OB_FORCE_BYTES = {"fp64": 27 * 116.0 + 24.0, "mixed": 27 * 60.0 + 24.0}
METRIC = "particle-steps/s and pair-interactions/s at 2M 3D particles; % HBM peak"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


# --------------------------------------------------------------------------
# GPU workloads
# --------------------------------------------------------------------------

def make_workload(kind: str, precision: str, n_side: int, nrep: int, ncoll: int):
    """(box, colloids, nrep, volume, description, force bytes per particle)"""
    from mcf_amd import PRECISION_FP64, PRECISION_MIXED
    from mcf_amd import workloads as W

    prec = PRECISION_FP64 if precision == "fp64" else PRECISION_MIXED
    if kind == "colloids":
        box, coll = W.colloid_suspension_3d(n_side, ncoll=ncoll)
        desc = (f"3D Tait box {n_side}^3 with {ncoll} rigid spheres of radius 5 dx "
                f"(boundary-particle shells, Morris no-slip, translating and rotating), "
                f"Lees-Edwards shear rate 0.5 on y, {box.n} particles "
                f"({int((box.sid > 0).sum())} boundary), kT 1e-3, dt 5e-5, fp64")
        return box, coll, 1, float(n_side) ** 3, desc, FORCE_BYTES["fp64"]
    if kind == "tait128":
        box = W.tait_box_3d(n_side, precision=prec, kt=1e-3)
        desc = (f"3D periodic Tait box {n_side}^3 = {box.n} particles, dx 1, cut-off 3, "
                f"gamma 7, c 10, eta 0.1, kT 1e-3, dt 5e-5, {precision}")
        return box, None, 1, float(n_side) ** 3, desc, FORCE_BYTES[precision]
    box = W.oldroyd_ensemble(nrep, precision=prec)
    volume = float(np.prod(box.box_hi - box.box_lo, axis=1).sum())
    desc = (f"batch of {box.nrep} periodic Oldroyd-B boxes (N log-uniform in [2000, 60000], "
            f"edge round(N^(1/3)) dx), {box.n} particles in total, tau 0.1, eta_p 0.05, "
            f"eta 0.1, kT 1e-3, dt 5e-5, {precision}")
    return box, None, box.nrep, volume, desc, OB_FORCE_BYTES[precision]


def run_workload(kind: str, precision: str, steps: int, warmup: int, n_side: int = 128,
                 nrep: int = 256, ncoll: int = 512, autotune: bool = True) -> dict:
    from mcf_amd import autotune as A
    from mcf_amd.sdpd import SDPD

    box, coll, nrep, volume, desc, force_bytes = make_workload(kind, precision, n_side, nrep,
                                                               ncoll)
    log(f"bench: {kind} n={box.n} nrep={nrep} precision={precision} warmup={warmup} "
        f"steps={steps}")
    g = SDPD(box.prm)
    if kind == "ensemble":
        g.set_replicas(box.box_lo, box.box_hi)
        g.upload_replicas(box.x, box.v, box.m, box.pid, box.rep)
    elif coll is not None:
        g.set_colloids(coll)
        g.upload(box.x, box.v, box.m, box.pid, box.sid)
    else:
        g.upload(box.x, box.v, box.m, box.pid)
    g.integrate_VV(max(warmup, 1))
    tuned = None
    if autotune:
        tuned = A.tune(g, box.prm, box.n, volume, nrep,
                       tag=f"c{coll.num}" if coll is not None else "")
        log(f"bench: launch shape {tuned['shape']} ({'cached' if tuned['cached'] else 'tuned'})")
    g.sync()
    t0 = time.perf_counter()
    g.integrate_VV(steps)
    g.sync()
    wall = time.perf_counter() - t0
    pairs_ordered, _ = g.counters()
    log(f"bench: timed {steps} steps in {wall:.3f} s")
    # per-phase kernel times from HIP events on the context stream, over a
    # few more steps launched kernel by kernel (30: single passes vary by a
    # few per cent from run to run)
    k_prof = max(3, min(steps, 30))
    g.set_timing(True)
    g.integrate_VV(k_prof)
    ph = g.phase_times()
    g.set_timing(False)
    g.close()
    per_launch = {k: v[0] / max(v[1], 1) for k, v in ph.items()}
    per_step = {k: v[0] / k_prof for k, v in ph.items()}
    log("bench: ms per step " + json.dumps({k: round(v, 4) for k, v in per_step.items()}))
    return {
        "kind": kind,
        "precision": precision,
        "workload": desc,
        "particles": int(box.n),
        "replicas": int(nrep),
        "colloids": 0 if coll is None else int(coll.num),
        "steps": steps,
        "warmup": warmup,
        "wall_s": wall,
        "ms_per_step": wall / steps * 1e3,
        "value": box.n * steps / wall,
        "pairs_per_step": pairs_ordered / 2.0,
        "pair_interactions_per_s": pairs_ordered / 2.0 * steps / wall,
        "phase_ms_per_step": per_step,
        "phase_ms_per_launch": per_launch,
        "force_ms": per_launch["forces"],
        "force_bytes_per_particle": force_bytes,
        "launch_shape": None if tuned is None else tuned["shape"],
    }


def _hip_runtime():
    """The HIP runtime libsdpd.so links (torch bundles a runtime of its own,
    which cannot share the process with it)."""
    import ctypes as C

    from mcf_amd.sdpd import load_library

    load_library()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    h.hipEventSynchronize.argtypes = [C.c_void_p]
    h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    h.hipEventDestroy.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _device_copies(nbytes: int, reps: int) -> float:
    """reps device-to-device copies of nbytes on the null stream; returns the
    milliseconds per copy from HIP events around them."""
    import ctypes as C

    h = _hip_runtime()
    src, dst = C.c_void_p(), C.c_void_p()
    a, b = C.c_void_p(), C.c_void_p()
    if h.hipMalloc(C.byref(src), nbytes) or h.hipMalloc(C.byref(dst), nbytes):
        raise RuntimeError("copy probe: hipMalloc failed")
    try:
        h.hipMemset(src, 1, nbytes)
        h.hipMemcpyAsync(dst, src, nbytes, 3, None)  # hipMemcpyDeviceToDevice
        h.hipEventCreate(C.byref(a))
        h.hipEventCreate(C.byref(b))
        h.hipEventRecord(a, None)
        for _ in range(reps):
            h.hipMemcpyAsync(dst, src, nbytes, 3, None)
        h.hipEventRecord(b, None)
        h.hipEventSynchronize(b)
        ms = C.c_float()
        h.hipEventElapsedTime(C.byref(ms), a, b)
        return ms.value / reps
    finally:
        h.hipDeviceSynchronize()
        for e in (a, b):
            if e.value:
                h.hipEventDestroy(e)
        h.hipFree(src)
        h.hipFree(dst)


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """GB/s of a plain device copy (nbytes read + nbytes written)."""
    return 2.0 * nbytes / (_device_copies(nbytes, reps) * 1e-3) / 1e9


def copy_probe(nbytes: int = 1 << 30, reps: int = 5) -> int:
    """Plain device copies of nbytes: what FETCH_SIZE / WRITE_SIZE read for
    them scales the counters of the pair kernels to bytes."""
    ms = _device_copies(nbytes, reps)
    log(f"copy probe: {reps} copies of {nbytes} B, {ms:.3f} ms each")
    return 0


def profiled_counters(precision: str):
    """Per-launch counters of the pair-force kernel on the 128^3 box from
    the committed rocprofv3 summary: HBM bytes (FETCH_SIZE / WRITE_SIZE
    scaled by a plain copy's reading) and VALU instructions, with the
    source the figures come from; None when absent."""
    if not os.path.exists(COUNTERS_FILE):
        return None
    with open(COUNTERS_FILE) as fh:
        data = json.load(fh)
    entry = data.get("kernels", {}).get(f"{precision}_forces")
    if entry is None:
        return None
    return dict(entry, source=data.get("source"), commit=data.get("commit"))


# --------------------------------------------------------------------------
# CPU baseline
# --------------------------------------------------------------------------

def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def native_oracle() -> str:
    """Build the oracle -O3 -march=native on this host (into a temporary
    directory); returns the flags used, or the shipped build's on failure."""
    from oracle import oracle as O

    flags = ["-O3", "-march=native", "-std=c11", "-fPIC", "-ffp-contract=off", "-fopenmp"]
    out = os.path.join(tempfile.mkdtemp(prefix="sdpd_cpu_"), "liboracle_native.so")
    try:
        subprocess.run(["gcc", *flags, "-shared", "-o", out,
                        os.path.join(ROOT, "oracle", "sdpd_oracle.c"), "-lm"],
                       check=True, capture_output=True, timeout=180)
    except (OSError, subprocess.SubprocessError):
        return "-O2 -fopenmp (the shipped build)"
    O.use_library(out)
    return " ".join(flags)


def cpu_baseline(n_side: int = 128) -> dict:
    """One velocity-Verlet step of C2 itself on the host: with all the
    threads OpenMP is given, then on one core (the reference runs one MPI
    rank per core)."""
    from mcf_amd import workloads as W
    from oracle import oracle as O

    flags = native_oracle()
    threads = host_threads()
    box = W.tait_box_3d(n_side, kt=1e-3)
    st = O.State(box.prm, box.x, box.v, box.m, box.pid)
    O.set_threads(threads)
    st.prologue(0)
    t0 = time.perf_counter()
    st.vv_step(1)
    t_all = time.perf_counter() - t0
    O.set_threads(1)
    t0 = time.perf_counter()
    st.vv_step(2)
    t_one = time.perf_counter() - t0
    model = cpu_model()
    log(f"bench: cpu baseline {threads} threads {t_all:.2f} s, 1 thread {t_one:.2f} s per step")
    return {
        "value": box.n / t_all,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP64 oracle (C restatement of the reference path: PPM cell list with the "
                   f"symmetric half stencil and ghost fold-back, Hu-Adams forces, VV), one VV "
                   f"step of the {n_side}^3 = {box.n} particle Tait box in {t_all:.2f} s on "
                   f"{threads} OpenMP threads of '{model}' (nproc {os.cpu_count()}); gcc {flags}"),
        "single_core": {"value": box.n / t_one, "cores": 1, "seconds_per_step": t_one},
        "cpu_model": model,
        "nproc": os.cpu_count(),
    }


# --------------------------------------------------------------------------

def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the configuration's own: 200, C5 500)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["tait128", "ensemble", "colloids"], default="tait128",
                    help="tait128: the 3D Tait box (size from --n-side); ensemble: a batch of "
                         "independent Oldroyd-B boxes (count from --nrep); colloids: the Tait "
                         "box with --ncoll rigid spheres under Lees-Edwards shear")
    ap.add_argument("--ncoll", type=int, default=512)
    ap.add_argument("--n-side", type=int, default=128)
    ap.add_argument("--nrep", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="default: fp64 for tait128 and colloids, mixed for ensemble")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="time the headline workload only (no C3 / C4 / C5 legs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 1 GiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "how the path would shard across GPUs is described, not built", file=sys.stderr)
        return 2
    if args.precision is None:
        args.precision = "mixed" if args.workload == "ensemble" else "fp64"
    if args.workload == "colloids" and args.precision != "fp64":
        print("bench.py: colloids run in FP64 only", file=sys.stderr)
        return 2
    default_steps = {"tait128": 200, "colloids": 200, "ensemble": 500}
    steps = args.steps if args.steps is not None else default_steps[args.workload]
    auto = not args.no_autotune

    head = run_workload(args.workload, args.precision, steps, args.warmup, args.n_side, args.nrep,
                        args.ncoll, auto)
    configs = {}
    if not args.no_extras and args.workload == "tait128" and args.precision == "fp64" \
            and args.n_side == 128:
        k = args.steps
        configs["C3"] = run_workload("tait128", "mixed", k or 200, args.warmup, 128, autotune=auto)
        configs["C4"] = run_workload("colloids", "fp64", k or 200, args.warmup, 128, ncoll=512,
                                     autotune=auto)
        configs["C5"] = run_workload("ensemble", "mixed", k or 500, args.warmup, nrep=256,
                                     autotune=auto)
        for name in ("C3", "C5"):
            configs[name]["ratio_vs_C2_fp64"] = configs[name]["value"] / head["value"]
    copy_gbs = copy_bandwidth()
    log(f"bench: device copy {copy_gbs:.0f} GB/s")

    t_force = head["force_ms"] / 1e3
    force_gbs = head["force_bytes_per_particle"] * head["particles"] / t_force / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": "pair forces (k_forces list walk + overflow rescan)",
        "achieved": force_gbs,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": force_gbs / HBM_PEAK_GBS,
        "traffic": None,
        "bytes_per_particle": head["force_bytes_per_particle"],
        "launch_ms": head["force_ms"],
        "copy_gbs": copy_gbs,
        "frac_of_copy": force_gbs / copy_gbs,
    }
    prof = profiled_counters(args.precision) if (args.workload == "tait128"
                                                 and args.n_side == 128) else None
    if prof is not None:
        roofline["traffic"] = prof["bytes"]
        roofline["traffic_source"] = f"{prof['source']} (profiled at commit {prof['commit']})"
        insts = prof["valu_insts_per_pair"] * head["pairs_per_step"] * 2.0
        roofline["limiter"] = prof.get("limiter")
        roofline["valu"] = {
            "insts_per_launch": insts,
            "insts_per_ordered_pair": prof["valu_insts_per_pair"],
            "frac": insts / (VALU_PEAK_INSTS * t_force),
            "peak_insts_per_s": VALU_PEAK_INSTS,
        }

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 pairs / f64 state",
        "data": "synthetic (seeded jittered lattices, Maxwellian velocities)",
        "config": {
            "workload": head["workload"],
            "particles": head["particles"],
            "replicas": head["replicas"],
            "colloids": head["colloids"],
            "precision": args.precision,
            "launch_shape": head["launch_shape"],
        },
        "pair_interactions_per_s": head["pair_interactions_per_s"],
        "pairs_per_step": head["pairs_per_step"],
        "phase_ms_per_step": head["phase_ms_per_step"],
        "phase_ms_per_launch": head["phase_ms_per_launch"],
        "roofline": roofline,
    }
    if configs:
        out["configs"] = {
            name: {k: c[k] for k in ("workload", "precision", "particles", "steps", "value",
                                     "ms_per_step", "pair_interactions_per_s", "pairs_per_step",
                                     "phase_ms_per_step", "launch_shape")
                   if k in c} | ({"ratio_vs_C2_fp64": c["ratio_vs_C2_fp64"]}
                                 if "ratio_vs_C2_fp64" in c else {})
            for name, c in configs.items()
        }
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(args.n_side if args.workload == "tait128" else 128)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
