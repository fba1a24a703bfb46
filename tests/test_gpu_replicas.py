"""Batched ensembles of independent periodic boxes in one context.

A batch must advance every replica exactly as a context holding that
replica alone would (same kernels, same per-particle summation order, same
noise keys), so the checks are bitwise against single-box runs, which the
parity tests pin to the oracle."""
import numpy as np
import pytest

from mcf_amd import workloads as W
from mcf_amd.sdpd import SDPD, SDPDError

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, flush=True)


def small_ensemble(kind="tait", sides=(10, 12, 11), precision=0, ft=3):
    if kind == "tait":
        prm = W.tait_box_3d(10, precision=precision, kt=1e-3).prm
    else:
        prm = W.oldroyd_box_3d(10, precision=precision, kt=1e-3).prm
    # Morris and the Oldroyd-B angular form have no thermal force
    noisy = not (ft == 1 or (kind == "oldroyd" and ft == 4))
    prm = prm.replace(rhs_force_type=ft, Brownian=noisy)
    return W.ensemble_of_cubes(sides, prm, kt=1e-3, seed=3, name=f"{kind}_ensemble")


def run_batch(ens, steps):
    with SDPD(ens.prm) as g:
        g.set_replicas(ens.box_lo, ens.box_hi)
        g.upload_replicas(ens.x, ens.v, ens.m, ens.pid, ens.rep)
        g.integrate_VV(steps)
        x, v = g.download()
        f = g.download_force()
        rho, _ = g.download_density()
        ek, mom = g.compute_statistic_replicas()
    return x, v, f, rho, ek, mom


def run_single(box, steps):
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.integrate_VV(steps)
        x, v = g.download()
        f = g.download_force()
        rho, _ = g.download_density()
    return x, v, f, rho


@pytest.mark.parametrize("kind,precision,ft", [("tait", 0, 3), ("tait", 1, 3), ("oldroyd", 0, 3),
                                               ("oldroyd", 1, 3), ("tait", 0, 2), ("tait", 1, 1),
                                               ("oldroyd", 0, 4)])
def test_batch_equals_single_boxes(kind, precision, ft):
    ens = small_ensemble(kind, precision=precision, ft=ft)
    announce(f"batch vs single boxes: {kind} precision={precision} force type {ft} n={ens.n}")
    steps = 3
    xb, vb, fb, rb, _, _ = run_batch(ens, steps)
    for q in range(ens.nrep):
        sel = ens.rep == q
        xs, vs, fs, rs = run_single(ens.replica(q), steps)
        assert np.array_equal(xb[sel], xs), q
        assert np.array_equal(vb[sel], vs), q
        assert np.array_equal(fb[sel], fs), q
        assert np.array_equal(rb[sel], rs), q


def test_replica_statistics_and_momentum():
    ens = small_ensemble("tait", sides=(10, 13, 10, 11))
    announce(f"replica statistics n={ens.n}")
    x, v, f, rho, ek, mom = run_batch(ens, 2)
    for q in range(ens.nrep):
        sel = ens.rep == q
        m = ens.m[sel]
        ek_ref = 0.5 * (m[:, None] * v[sel] ** 2).sum()
        mom_ref = (m[:, None] * v[sel]).sum(0)
        assert ek[q] == pytest.approx(ek_ref, rel=1e-12)
        assert np.abs(mom[q] - mom_ref).max() < 1e-12 * (np.abs(m[:, None] * v[sel]).sum() + 1e-30)
        # pair forces are antisymmetric inside each replica (noise included)
        mf = m[:, None] * f[sel]
        assert np.abs(mf.sum(0)).max() < 1e-10 * np.abs(mf).sum()


def test_replicas_are_isolated():
    # two copies of the same box with disjoint ids: identical deterministic
    # dynamics (kT = 0), so no particle of one sees the other
    prm = W.tait_box_3d(10, kt=0.0, brownian=False).prm
    ens = W.ensemble_of_cubes((11, 11), prm, kt=0.0, seed=5)
    # make replica 1 an exact copy of replica 0 (ids stay distinct)
    n0 = int((ens.rep == 0).sum())
    ens.x[n0:] = ens.x[:n0]
    ens.v[n0:] = ens.v[:n0]
    announce("replica isolation")
    x, v, f, rho, _, _ = run_batch(ens, 2)
    assert np.array_equal(x[:n0], x[n0:])
    assert np.array_equal(f[:n0], f[n0:])


def test_batch_error_paths():
    ens = small_ensemble("tait", sides=(10, 11))
    announce("batch error paths")
    with SDPD(ens.prm) as g:
        g.set_replicas(ens.box_lo, ens.box_hi)
        with pytest.raises(SDPDError):
            g.upload(ens.x, ens.v, ens.m, ens.pid)
        bad = ens.rep.copy()
        bad[0] = 7
        with pytest.raises(SDPDError):
            g.upload_replicas(ens.x, ens.v, ens.m, ens.pid, bad)
        wrong = ens.rep.copy()
        wrong[ens.rep == 1] = 0  # replica 1 particles outside replica 0's box
        with pytest.raises(SDPDError):
            g.upload_replicas(ens.x, ens.v, ens.m, ens.pid, wrong)
        g.upload_replicas(ens.x, ens.v, ens.m, ens.pid, ens.rep)
        with pytest.raises(SDPDError):
            g.cells()
    with SDPD(ens.prm) as g:
        with pytest.raises(SDPDError):
            g.set_replicas([[0, 0, 0]], [[8.0, 10, 10]])  # below 3 cells of 1.0001 rc
