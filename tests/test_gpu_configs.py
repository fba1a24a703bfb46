"""The benchmark configurations at their full sizes on the GPU.

C3: the 128^3 Tait box (2,097,152 particles) in mixed precision against the
    FP64 path on identical states of the FP64 trajectory: per step, max and
    RMS relative errors of the number density and the forces, and the total
    momentum of the two trajectories, all within the 1e-5 mixed bar.
C4: 512 rigid spheres of radius 5 dx in the 128^3 box with Lees-Edwards
    shear at rate 0.5 on the y faces: pair forces conserve momentum, the
    drag and torque equal the boundary particles' sums, the state stays
    finite over steps; and a scaled copy (8 spheres in 32^3, same radius and
    shear) against the oracle.
C5: 256 Oldroyd-B replicas of 2k-60k particles (about 4.3M) in mixed
    precision in one context: per-replica momentum and finite state, and the
    smallest, median and largest replicas equal to single-box contexts bit
    for bit.

Reference: marching_integrate_VV.F90:303-1619 (the step), colloid_create_
boundary_particle_3D_sphere.F90:160-230 and marching_integrate_VV.F90:
455-765 (colloids), particles_compute_pressure_tensor.F90:114-133 (Pi).
"""
import sys

import numpy as np
import pytest

from mcf_amd import workloads as W
from mcf_amd.sdpd import SDPD

from colloid_helpers import to_oracle
from gpu_helpers import rel_err, vec_err

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _evaluate(prm, x, v, m, pid, step):
    with SDPD(prm) as g:
        g.upload(x, v, m, pid)
        g.step = step
        g.evaluate()
        rho, _ = g.download_density()
        f = g.download_force()
    return rho, f


def test_c3_mixed_against_fp64_per_step():
    box = W.tait_box_3d(128, kt=1e-3)
    mbox = W.mixed(box)
    announce(f"C3: n={box.n}")
    scale_p = np.abs(box.m[:, None] * box.v).sum()
    with SDPD(box.prm) as a, SDPD(mbox.prm) as b:
        a.upload(box.x, box.v, box.m, box.pid)
        b.upload(mbox.x, mbox.v, mbox.m, mbox.pid)
        for step in range(1, 4):
            a.integrate_VV(1)
            b.integrate_VV(1)
            x, v = a.download()
            # identical inputs: the FP64 trajectory's state, both precisions
            r64, f64 = _evaluate(box.prm, x, v, box.m, box.pid, step)
            rmx, fmx = _evaluate(mbox.prm, x, v, box.m, box.pid, step)
            dn = np.abs(rmx - r64) / r64
            fscale = np.sqrt((f64 ** 2).sum(1).mean())
            df = np.sqrt(((fmx - f64) ** 2).sum(1)) / fscale
            # the two trajectories' total momentum (noise streams identical)
            _, pa = a.compute_statistic()
            _, pb = b.compute_statistic()
            dp = np.abs(pb - pa).max() / scale_p
            announce(f"C3 step {step}: density max {dn.max():.2e} rms {np.sqrt((dn ** 2).mean()):.2e}; "
                     f"force max {df.max():.2e} rms {np.sqrt((df ** 2).mean()):.2e}; "
                     f"momentum {dp:.2e}")
            assert dn.max() < 1e-5 and np.sqrt((dn ** 2).mean()) < 1e-6
            assert df.max() < 1e-5 and np.sqrt((df ** 2).mean()) < 1e-6
            assert dp < 1e-5
            assert np.isfinite(fmx).all()


def _drag_torque(box, coll, x, f):
    """Sum of m f and of r x m f over each colloid's boundary particles,
    r from the centre's periodic nearest image (no shear at time 0), and
    the sums of their magnitudes (the scale of the rounding)."""
    L = np.array(box.prm.length[:3])
    out = np.zeros((4, coll.num, 3))
    for k in range(coll.num):
        sel = box.sid == k + 1
        mf = box.m[sel, None] * f[sel]
        r = x[sel] - np.asarray(coll.x)[k][None, :]
        r -= L[None, :] * np.round(r / L[None, :])
        rmf = np.cross(r, mf)
        out[0, k], out[1, k] = mf.sum(0), rmf.sum(0)
        out[2, k], out[3, k] = np.abs(mf).sum(0), np.abs(rmf).sum(0)
    return out


def test_c4_colloid_suspension_full_size():
    box, coll = W.colloid_suspension_3d(128, ncoll=512)
    nb = int((box.sid > 0).sum())
    announce(f"C4: n={box.n} boundary={nb} colloids={coll.num}")
    with SDPD(box.prm) as g:
        g.set_colloids(coll)
        g.upload(box.x, box.v, box.m, box.pid, box.sid)
        g.evaluate()  # the initial evaluation collects drag and torque
        f = g.download_force()
        x, _ = g.download()
        pairs, rechecks = g.counters()
        _, _, _, drag, torque = g.colloids()
        mf = box.m[:, None] * f
        assert np.isfinite(f).all()
        assert np.abs(mf.sum(0)).max() < 1e-10 * np.abs(mf).sum()
        assert 80 < pairs / box.n < 140 and rechecks < 0.01 * pairs
        d_ref, t_ref, d_abs, t_abs = _drag_torque(box, coll, x, f)
        assert (np.abs(drag - d_ref) <= 1e-12 * d_abs + 1e-300).all()
        assert (np.abs(torque - t_ref) <= 1e-12 * t_abs + 1e-300).all()
        g.integrate_VV(3)
        x, v = g.download()
        f = g.download_force()
        cx, cv, cw, drag, torque = g.colloids()
        for a in (x, v, f, cx, cv, cw, drag, torque):
            assert np.isfinite(a).all()
        assert (x >= 0).all() and (x < 128.0).all()
        mf = box.m[:, None] * f
        assert np.abs(mf.sum(0)).max() < 1e-10 * np.abs(mf).sum()
        announce(f"C4: {pairs / 2:.3e} pairs, |sum m f| / sum |m f| = "
                 f"{np.abs(mf.sum(0)).max() / np.abs(mf).sum():.1e}")


def test_c4_scaled_against_oracle(oracle):
    # 8 spheres of radius 5 dx in a 32^3 box (lattice spacing 16 dx), the
    # same shear rate 0.5 on the y faces, 3 steps
    box, coll = W.colloid_suspension_3d(32, ncoll=8)
    announce(f"C4 scaled: n={box.n} boundary={int((box.sid > 0).sum())}")
    nsteps = 3
    with SDPD(box.prm) as g:
        g.set_colloids(coll)
        g.upload(box.x, box.v, box.m, box.pid, box.sid)
        g.evaluate()
        rho0, _ = g.download_density()
        f0 = g.download_force()
        g.integrate_VV(nsteps)
        x, v = g.download()
        f = g.download_force()
        cx, cv, cw, drag, torque = g.colloids()
    oc = to_oracle(oracle, coll)
    ro = oracle.density(box.prm, box.x, box.m, box.sid, 0.0, oc)
    po = oracle.pressure(box.prm, ro, box.m)
    fo, _ = oracle.interaction(box.prm, 0, box.x, box.v, ro, po, box.m, box.pid, box.sid,
                               le_time=0.0, coll=oc)
    assert rel_err(rho0, ro) < 1e-13
    assert vec_err(f0, fo) < 1e-12
    oc = to_oracle(oracle, coll)
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid, box.sid, coll=oc)
    st.prologue(0)
    for s in range(1, nsteps + 1):
        st.vv_step(s)
    L = 32.0
    dx = x - st.x
    dx -= L * np.round(dx / L)
    assert np.abs(dx).max() < 1e-10 * L
    assert vec_err(v, st.v) < 1e-10
    assert vec_err(f, st.f) < 1e-10
    dc = cx - oc.x
    dc -= L * np.round(dc / L)
    assert np.abs(dc).max() < 1e-10 * L
    assert np.abs(cv - oc.v[0]).max() < 1e-9 * (np.abs(oc.v[0]).max() + 1e-12)
    assert np.abs(cw - oc.omega[0]).max() < 1e-9 * (np.abs(oc.omega[0]).max() + 1e-12)
    assert np.abs(drag - oc.drag).max() < 1e-9 * (np.abs(oc.drag).max() + 1e-300)
    assert np.abs(torque - oc.torque).max() < 1e-9 * (np.abs(oc.torque).max() + 1e-300)


def test_c5_oldroyd_ensemble_full_size():
    ens = W.oldroyd_ensemble(256, precision=1)
    counts = np.bincount(ens.rep, minlength=ens.nrep)
    announce(f"C5: n={ens.n} replicas={ens.nrep} sizes {counts.min()}..{counts.max()}")
    steps = 2
    with SDPD(ens.prm) as g:
        g.set_replicas(ens.box_lo, ens.box_hi)
        g.upload_replicas(ens.x, ens.v, ens.m, ens.pid, ens.rep)
        ek0, mom0 = g.compute_statistic_replicas()
        g.integrate_VV(steps)
        x, v = g.download()
        f = g.download_force()
        ct, vgt = g.download_tensors()
        ek, mom = g.compute_statistic_replicas()
    for a in (x, v, f, ct, vgt, ek, mom):
        assert np.isfinite(a).all()
    for q in range(ens.nrep):
        sel = ens.rep == q
        scale = np.abs(ens.m[sel, None] * ens.v[sel]).sum()
        # each replica's momentum is conserved by its own antisymmetric pairs
        assert np.abs(mom[q] - mom0[q]).max() < 1e-6 * scale, q
    order = np.argsort(counts, kind="stable")
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    for q in picks:
        sel = ens.rep == q
        box = ens.replica(q)
        announce(f"C5: replica {q} alone (n={box.n})")
        with SDPD(box.prm) as g:
            g.upload(box.x, box.v, box.m, box.pid)
            g.integrate_VV(steps)
            xs, vs = g.download()
            fs = g.download_force()
            cts, vgs = g.download_tensors()
        assert np.array_equal(x[sel], xs) and np.array_equal(v[sel], vs), q
        assert np.array_equal(f[sel], fs), q
        assert np.array_equal(ct[sel], cts) and np.array_equal(vgt[sel], vgs), q
