#!/usr/bin/env python3
"""Benchmark of the SDPD step on one MI355X.

Workload (default): the 3D periodic Tait box of 128^3 = 2,097,152 particles,
dx = 1, quintic kernel with cut-off 3 dx, gamma 7, c = 10, eta = 0.1,
Maxwellian velocities and thermal noise at kT = 1e-3, dt = 5e-5, in plain
FP64 (--precision fp64).  One step = one full velocity-Verlet step of every
particle: step boundary (kicks, drift, wrap, cell keys), cell sort,
density + EOS, pair forces with noise.  --workload ensemble runs a batch of
256 independent Oldroyd-B boxes (mixed precision) in one context instead;
--workload colloids the same Tait box with 512 rigid spheres of radius 5 dx
(boundary-particle shells, Morris no-slip, translating and rotating) under
Lees-Edwards shear at rate 0.5 on the y faces, in FP64.
The launch shapes come from the autotuner (cached per workload class).

Prints one JSON line.  value = particle-steps/s of the whole job with the
state resident in HBM.  roofline = the pair-force kernel's stencil-model
bytes per launch over its average launch time (HIP events on the context
stream), against the 8 TB/s HBM peak.  cpu_baseline = the FP64 oracle (a
C restatement of the reference path) on one host core, on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
# Stencil-model algorithmic bytes per particle and launch (27 neighbour-cell
# records per particle): FP64 force record 68 B -> 27*68 + f 24 = 1860 B;
# density record 24 B -> 27*24 + n, p 16 = 664 B + self 8 = 672 B.
FORCE_BYTES = {"fp64": 1860.0, "mixed": 996.0}
# Oldroyd-B adds Pi to the force record: 6 components, FP64 48 B / FP32 24 B
# (FP64 record 68 + 48 = 116 B, mixed 36 + 24 = 60 B), plus f 24 B.
OB_FORCE_BYTES = {"fp64": 27 * 116.0 + 24.0, "mixed": 27 * 60.0 + 24.0}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seconds_hint: float = 12.0) -> dict:
    from mcf_amd import workloads as W
    from oracle import oracle as O

    box = W.tait_box_3d(32, kt=1e-3)
    st = O.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    steps = 0
    t0 = time.perf_counter()
    while True:
        steps += 1
        st.vv_step(steps)
        if time.perf_counter() - t0 > seconds_hint or steps >= 40:
            break
    dt = time.perf_counter() - t0
    return {
        "value": box.n * steps / dt,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"FP64 oracle, same physics on a 32^3 = {box.n} particle box, {steps} VV steps "
                  f"in {dt:.1f} s on one core of {os.uname().nodename}",
    }


def measured_traffic(n_side: int, precision: str):
    """HBM bytes per pair-force launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the copy calibration), for
    the workload it was measured on; None otherwise."""
    if n_side != 128 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        data = json.load(fh)
    entry = data.get(f"tait128_{precision}_forces")
    return None if entry is None else float(entry["bytes"])


def copy_probe(nbytes: int = 1 << 30, reps: int = 5) -> int:
    """Plain elementwise copy of nbytes (reads nbytes, writes nbytes) on the
    GPU: what FETCH_SIZE / WRITE_SIZE read for it scales the counters of the
    pair kernels to bytes."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy probe: {reps} copies of {nbytes} B")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["tait128", "ensemble", "colloids"], default="tait128",
                    help="tait128: the 3D Tait box (size from --n-side); ensemble: a batch of "
                         "independent Oldroyd-B boxes (count from --nrep); colloids: the Tait "
                         "box with --ncoll rigid spheres under Lees-Edwards shear")
    ap.add_argument("--ncoll", type=int, default=512)
    ap.add_argument("--n-side", type=int, default=128)
    ap.add_argument("--nrep", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="default: fp64 for tait128, mixed for ensemble")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 1 GiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU design is described but not built", file=sys.stderr)
        return 2
    if args.precision is None:
        args.precision = "mixed" if args.workload == "ensemble" else "fp64"

    from mcf_amd import PRECISION_FP64, PRECISION_MIXED
    from mcf_amd import autotune as A
    from mcf_amd import workloads as W
    from mcf_amd.sdpd import SDPD

    prec = PRECISION_FP64 if args.precision == "fp64" else PRECISION_MIXED
    colloids = None
    if args.workload == "colloids":
        if args.precision != "fp64":
            print("bench.py: colloids run in FP64 only", file=sys.stderr)
            return 2
        box, colloids = W.colloid_suspension_3d(args.n_side, ncoll=args.ncoll)
        nrep, volume = 1, float(args.n_side) ** 3
        workload = (f"3D Tait box {args.n_side}^3 with {args.ncoll} rigid spheres of radius 5 dx "
                    f"(boundary-particle shells, Morris no-slip, translating and rotating), "
                    f"Lees-Edwards shear rate 0.5 on y, {box.n} particles "
                    f"({int((box.sid > 0).sum())} boundary), kT 1e-3, dt 5e-5, fp64")
        force_bytes = FORCE_BYTES["fp64"]
    elif args.workload == "tait128":
        box = W.tait_box_3d(args.n_side, precision=prec, kt=1e-3)
        nrep, volume = 1, float(args.n_side) ** 3
        workload = (f"3D periodic Tait box {args.n_side}^3 = {box.n} particles, dx 1, cut-off 3, "
                    f"gamma 7, c 10, eta 0.1, kT 1e-3, dt 5e-5, {args.precision}")
        force_bytes = FORCE_BYTES[args.precision]
    else:
        box = W.oldroyd_ensemble(args.nrep, precision=prec)
        nrep = box.nrep
        volume = float(np.prod(box.box_hi - box.box_lo, axis=1).sum())
        workload = (f"batch of {nrep} periodic Oldroyd-B boxes (N log-uniform in [2000, 60000], "
                    f"edge round(N^(1/3)) dx), {box.n} particles in total, tau 0.1, eta_p 0.05, "
                    f"eta 0.1, kT 1e-3, dt 5e-5, {args.precision}")
        force_bytes = OB_FORCE_BYTES[args.precision]
    log(f"bench: {box.name} n={box.n} nrep={nrep} precision={args.precision} "
        f"warmup={args.warmup} steps={args.steps}")
    g = SDPD(box.prm)
    if nrep > 1 or args.workload == "ensemble":
        g.set_replicas(box.box_lo, box.box_hi)
        g.upload_replicas(box.x, box.v, box.m, box.pid, box.rep)
    elif colloids is not None:
        g.set_colloids(colloids)
        g.upload(box.x, box.v, box.m, box.pid, box.sid)
    else:
        g.upload(box.x, box.v, box.m, box.pid)
    g.integrate_VV(max(args.warmup, 1))
    tuned = None
    if not args.no_autotune:
        tuned = A.tune(g, box.prm, box.n, volume, nrep,
                       tag=f"c{colloids.num}" if colloids is not None else "")
        log(f"bench: launch shape {tuned['shape']} ({'cached' if tuned['cached'] else 'tuned'})")
    g.sync()
    log("bench: warmup done")
    t0 = time.perf_counter()
    g.integrate_VV(args.steps)
    g.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    pairs_ordered, _ = g.counters()
    unordered = pairs_ordered / 2.0
    log(f"bench: timed {args.steps} steps in {wall:.3f} s")

    # per-phase kernel times from HIP events on the context stream
    k_prof = max(3, min(args.steps, 10))
    g.set_timing(True)
    g.integrate_VV(k_prof)
    ph = g.phase_times()
    g.set_timing(False)
    g.close()
    t_force = ph["forces"][0] / max(ph["forces"][1], 1) / 1e3
    force_gbs = force_bytes * box.n / t_force / 1e9
    log("bench: phases " + json.dumps({k: round(v[0] / max(v[1], 1), 4) for k, v in ph.items()}))

    out = {
        "metric": "particle-steps/s and pair-interactions/s at 2M 3D particles; % HBM peak",
        "value": box.n * args.steps / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 pairs / f64 state",
        "data": "synthetic (seeded jittered lattices, Maxwellian velocities)",
        "config": {
            "workload": workload,
            "particles": box.n,
            "replicas": nrep,
            "colloids": 0 if colloids is None else colloids.num,
            "precision": args.precision,
            "launch_shape": None if tuned is None else tuned["shape"],
        },
        "pair_interactions_per_s": unordered * args.steps / wall,
        "pairs_per_step": unordered,
        "phase_ms_per_launch": {k: v[0] / max(v[1], 1) for k, v in ph.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": "pair forces",
            "achieved": force_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": force_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.n_side, args.precision) if args.workload == "tait128" else None,
            "bytes_per_particle": force_bytes,
            "launch_ms": t_force * 1e3,
        },
    }
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle, one core)")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
