"""Pins of the oracle's noise stream, time integration and statistics.

Reference: random_random.F90:48-73 and random_random_Gaussian.F90:79-175
(N(0,1) truncated at |x| < 6), the symmetric traceless dW of
rhs_force_ff_Newtonian_HuAdams.F90:186-219, marching_integrate_VV.F90:303-1619,
particles_compute_act.F90:140-155, statistic_compute_statistic.F90:230-283.

Pins: Philox-4x32-10 known-answer vectors (Random123 kat_vectors),
moments / Kolmogorov-Smirnov of the normals, the covariance of the force
noise against the reference's 9-normal recipe, the fluctuation-dissipation
temperature, VV closed forms and time reversibility, the Oldroyd-B
homogeneous-shear steady state, closed statistic sums.
"""
import math

import numpy as np
import pytest
from scipy import stats

from mcf_amd import workloads as W
from mcf_amd.params import SDPDParams


@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(oracle, ctr, key, out):
    assert oracle.philox4x32_10(ctr, key) == out


def _normals(oracle, count, seed=11, step=5):
    g = np.empty((count, 4), np.float32)
    for k in range(count):
        g[k] = oracle.pair_normals(seed, k, k + 1 + (k % 7), step)
    return g.ravel().astype(np.float64)


def test_normals_distribution(oracle):
    g = _normals(oracle, 60000)
    assert abs(g.mean()) < 4 * 1 / math.sqrt(g.size)
    assert g.var() == pytest.approx(1.0, abs=0.015)
    assert abs(stats.skew(g)) < 0.03
    assert stats.kurtosis(g, fisher=False) == pytest.approx(3.0, abs=0.06)
    assert stats.kstest(g, "norm").pvalue > 1e-3
    # tails: Box-Muller on 24-bit uniforms cannot exceed sqrt(48 ln 2)
    assert np.abs(g).max() <= math.sqrt(-2 * math.log(2.0 ** -24)) + 1e-5


def test_normals_depend_on_every_key_field(oracle):
    base = oracle.pair_normals(3, 10, 20, 7)
    for args in [(4, 10, 20, 7), (3, 11, 20, 7), (3, 10, 21, 7), (3, 10, 20, 8),
                 (3 + (1 << 40), 10, 20, 7), (3, 10, 20, 7 + (1 << 33))]:
        assert not np.array_equal(oracle.pair_normals(*args), base)
    assert np.array_equal(oracle.pair_normals(3, 10, 20, 7), base)


def test_noise_log_within_two_ulp_of_libm(oracle):
    # every uniform the stream can draw: u = k 2^-24, k = 1 .. 2^24
    k = np.arange(1, 2 ** 24 + 1, dtype=np.uint64)
    u = (k.astype(np.float32) * np.float32(2.0 ** -24)).astype(np.float32)
    got = oracle.noise_log(u).astype(np.float64)
    ref = np.log(u.astype(np.float64))
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    ulp[ref == 0] = np.spacing(np.float32(0))
    assert (np.abs(got - ref) / ulp).max() <= 2.0
    assert got[-1] == 0.0  # ln 1


def test_noise_sincos_against_libm(oracle):
    # every angle index: cos and sin of 2 pi k / 2^24 to about an ulp of 1
    k = np.arange(2 ** 24, dtype=np.uint32)
    cs, sn = oracle.noise_sincos(k)
    th = 2 * np.pi * k.astype(np.float64) / 2 ** 24
    assert np.abs(cs - np.cos(th)).max() < 1.2e-7
    assert np.abs(sn - np.sin(th)).max() < 1.2e-7
    # quadrant starts are exact
    q = np.array([0, 1 << 22, 2 << 22, 3 << 22], np.uint32)
    c4, s4 = oracle.noise_sincos(q)
    assert np.array_equal(c4, [1, 0, -1, 0]) and np.array_equal(s4, [0, 1, 0, -1])


def _reference_we(rng, e, count):
    """The reference's recipe: 3x3 iid normals, symmetrised, made
    traceless, times e (rhs_force_ff_Newtonian_HuAdams.F90:186-219)."""
    dW = rng.normal(size=(count, 3, 3))
    tr = np.trace(dW, axis1=1, axis2=2) / 3.0
    S = 0.5 * (dW + dW.transpose(0, 2, 1))
    S -= tr[:, None, None] * np.eye(3)[None]
    return S @ e


def test_noise_covariance_matches_reference_recipe(oracle):
    # force noise of one pair through the oracle, over many steps
    prm = SDPDParams(num_dim=3, max_phys=[12.0] * 3, cut_off=3.0, stateEquation_type=2, gamma=7.0,
                     c=10.0, eta=0.1, Brownian=True, kt=1e-3, dt=1e-4)
    x = np.array([[5.0, 6.0, 6.0], [6.3, 6.4, 5.8]])
    v = np.zeros_like(x)
    m = np.ones(2)
    rho = oracle.density(prm, x, m)
    p = oracle.pressure(prm, rho, m)
    f_det, _ = oracle.interaction(prm.replace(Brownian=False), 0, x, v, rho, p, m, np.array([1, 2]))
    e = (x[0] - x[1]) / np.linalg.norm(x[0] - x[1])
    r = np.linalg.norm(x[0] - x[1])
    _, gw = oracle.kernel(prm, r)
    f_d = prm.eta * (2.0 / rho[0] ** 2) * gw / r
    f_r = math.sqrt(-4 * prm.kt * f_d) / math.sqrt(prm.dt)
    samples = []
    for step in range(4000):
        f, _ = oracle.interaction(prm, step, x, v, rho, p, m, np.array([1, 2]))
        samples.append((f[0] - f_det[0]) * m[0] / f_r)
        assert np.allclose(f[0] - f_det[0], -(f[1] - f_det[1]), rtol=1e-9, atol=1e-12)
    samples = np.array(samples)
    cov = samples.T @ samples / len(samples)
    ref = _reference_we(np.random.default_rng(0), e, 200000)
    cov_ref = ref.T @ ref / len(ref)
    analytic = 0.5 * np.eye(3) + np.outer(e, e) / 6.0
    assert np.allclose(cov_ref, analytic, atol=0.01)
    assert np.allclose(cov, analytic, atol=0.04)


@pytest.mark.parametrize("force_type", [3, 4])
def test_fluctuation_dissipation_temperature(oracle, force_type):
    # equilibrium <m v_a^2> = kT per component for the noise amplitude
    # sqrt(-4 kT f_d / dt) (Hu-Adams) or its angular counterpart
    kt = 1e-4
    box = W.shear_wave_2d(16, brownian=True, kt=kt)
    prm = box.prm.replace(rhs_force_type=force_type, stateEquation_type=2, gamma=7.0, rho_ref=1.0,
                          eta=0.1 if force_type == 3 else 0.4)
    st = oracle.State(prm, box.x, np.zeros_like(box.v), box.m, box.pid)
    st.prologue(0)
    temps = []
    for s in range(1, 1201):
        st.vv_step(s)
        if s > 400:
            e, mom = st.statistics()
            temps.append(2.0 * e / (2 * (box.n - 1)))
    assert np.mean(temps) == pytest.approx(kt, rel=0.12)


def test_vv_constant_force_closed_form(oracle):
    # a lattice at rest under a body force g: v = g t, x = x0 + g t^2 / 2
    prm = SDPDParams(num_dim=2, max_phys=[1.0, 1.0, 0], cut_off=3.0 / 16, stateEquation_type=1, c=10.0,
                     gamma=1.0, rho_ref=0.0, eta=0.1, body_force_type=1, body_force=[0.3, -0.2, 0], dt=1e-3)
    x0 = W.lattice(2, 16, 1.0 / 16, 0.0, 0)
    st = oracle.State(prm, x0, np.zeros_like(x0), np.full(len(x0), 1 / 256), np.arange(1, 257))
    st.prologue(0)
    n = 40
    for s in range(1, n + 1):
        st.vv_step(s)
    t = n * prm.dt
    g = np.array([0.3, -0.2])
    assert np.allclose(st.v, g * t, rtol=1e-9, atol=1e-12)
    d = W.wrap(x0 + 0.5 * g * t * t, prm) - st.x
    assert np.abs(d).max() < 1e-12


def test_vv_time_reversible_without_dissipation(oracle):
    box = W.shear_wave_2d(16)
    prm = box.prm.replace(eta=0.0)
    st = oracle.State(prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, 31):
        st.vv_step(s)
    st.v = -st.v
    st.prologue(0)
    for s in range(1, 31):
        st.vv_step(s)
    d = st.x - box.x
    d -= np.round(d)
    assert np.abs(d).max() < 1e-11
    assert np.abs(st.v + box.v).max() < 1e-9


def test_oldroyd_homogeneous_shear_steady_state(oracle):
    # dc/dt = L c + c L^T - (c - I)/tau -> c_xy = tau g, c_xx = 1 + 2 (tau g)^2, c_yy = 1
    tau, gdot, dt = 0.1, 2.0, 1e-3
    prm = SDPDParams(num_dim=3, max_phys=[10] * 3, Newtonian=False, tau=tau, dt=dt)
    L = np.zeros((3, 3))
    L[0, 1] = gdot  # dv_x / dy
    vgt = L.reshape(1, -1, order="F")
    ct = np.eye(3).reshape(1, -1)
    for _ in range(4000):
        act = oracle.compute_act(prm, ct, vgt)
        ct = ct + act * dt * 0.5
        act = oracle.compute_act(prm, ct, vgt)
        ct = ct + act * dt * 0.5
    c = ct.reshape(3, 3, order="F")
    tg = tau * gdot
    assert c[0, 1] == pytest.approx(tg, rel=1e-6)
    assert c[1, 0] == pytest.approx(tg, rel=1e-6)
    assert c[0, 0] == pytest.approx(1 + 2 * tg * tg, rel=1e-6)
    assert c[1, 1] == pytest.approx(1.0, abs=1e-9)
    assert c[2, 2] == pytest.approx(1.0, abs=1e-9)


def test_statistics_closed_sums(oracle):
    prm = SDPDParams(num_dim=3, max_phys=[10] * 3)
    rng = np.random.default_rng(2)
    v = rng.normal(size=(50, 3))
    m = rng.uniform(0.5, 2.0, 50)
    st = oracle.State(prm, np.zeros((50, 3)) + 1.0, v, m, np.arange(1, 51))
    e, mom = st.statistics()
    assert e == pytest.approx(0.5 * (m * (v ** 2).sum(1)).sum(), rel=1e-14)
    assert np.allclose(mom, (m[:, None] * v).sum(0), rtol=1e-14)
