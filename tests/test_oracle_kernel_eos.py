"""Pins of the oracle's smoothing kernels and equations of state.

Reference: mcf/src/kernel_new.F90:47-93, kernel_kernel_quintic_spline.F90,
kernel_kernel_Lucy.F90, stateEquation_compute_pressure.F90:158,176,
stateEquation_new.F90:102.  Pins: normalisation (fixes the coefficient),
finite differences of W (fixes W' and coef_grad), compact support and
smoothness at the knots (fixes the piecewise structure), closed-form EOS.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from mcf_amd.params import SDPDParams


def prm(dim, kernel_type=1, cut_off=3.0, **kw):
    return SDPDParams(num_dim=dim, cut_off=cut_off, kernel_type=kernel_type,
                      max_phys=[20.0, 20.0, 20.0], **kw)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel_type,cut", [(1, 3.0), (1, 0.6), (2, 1.3)])
def test_normalisation(oracle, dim, kernel_type, cut):
    p = prm(dim, kernel_type, cut)
    w = lambda r: oracle.kernel(p, r)[0]
    shell = (lambda r: 2 * math.pi * r) if dim == 2 else (lambda r: 4 * math.pi * r * r)
    h = cut / 3.0 if kernel_type == 1 else cut
    knots = [h, 2 * h] if kernel_type == 1 else []
    val, _ = integrate.quad(lambda r: w(r) * shell(r), 0.0, cut, points=knots, epsabs=1e-13,
                            epsrel=1e-12, limit=200)
    # The reference's 3D quintic coefficient 3/(359 pi h^3) (kernel_new.F90:59)
    # is not the exact normaliser 1/(120 pi h^3): the integral is 360/359.
    # That quirk is kept; the exact ratio pins the coefficient.
    expect = 360.0 / 359.0 if (kernel_type == 1 and dim == 3) else 1.0
    assert abs(val - expect) < 1e-10


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel_type", [1, 2])
def test_gradient_is_derivative(oracle, dim, kernel_type):
    cut = 3.0 if kernel_type == 1 else 1.0
    p = prm(dim, kernel_type, cut)
    for r in np.linspace(0.05, cut - 0.05, 37):
        eps = 1e-5
        wp = oracle.kernel(p, r + eps)[0]
        wm = oracle.kernel(p, r - eps)[0]
        _, g = oracle.kernel(p, r)
        fd = (wp - wm) / (2 * eps)
        assert abs(g - fd) <= 1e-7 * max(1.0, abs(fd)), (r, g, fd)


def test_quintic_support_and_knots(oracle):
    p = prm(3)
    w0, g0 = oracle.kernel(p, 0.0)
    coef = 3.0 / (359.0 * math.pi)
    assert w0 == pytest.approx(66.0 * coef, rel=1e-15)
    assert g0 == 0.0
    for r in [3.0, 3.5, 10.0]:
        assert oracle.kernel(p, r) == (0.0, 0.0)
    # fourth-order contact at the cut-off: W ~ (3-s)^5
    w, g = oracle.kernel(p, 3.0 - 1e-3)
    assert w == pytest.approx(coef * 1e-15, rel=1e-6)
    # continuity of W and W' across the knots s = 1 and s = 2
    for knot in (1.0, 2.0):
        a = oracle.kernel(p, knot - 1e-9)
        b = oracle.kernel(p, knot + 1e-9)
        assert a[0] == pytest.approx(b[0], rel=1e-7)
        assert a[1] == pytest.approx(b[1], rel=1e-7)


def test_negative_distance_is_an_error(oracle):
    with pytest.raises(RuntimeError):
        oracle.kernel(prm(3), -1.0)


def test_eos_morris_vector_form(oracle):
    # p = c^2 (rho^gamma - rho_ref) (stateEquation_compute_pressure.F90:158)
    p = prm(2, stateEquation_type=1, c=10.0, gamma=1.0, rho_ref=0.0, rhs_density_type=1)
    rho = np.array([0.5, 1.0, 1.25])
    out = oracle.pressure(p, rho, np.ones(3))
    assert np.allclose(out, 100.0 * rho, rtol=1e-15)
    p2 = p.replace(gamma=2.0, rho_ref=0.5)
    assert np.allclose(oracle.pressure(p2, rho, np.ones(3)), 100.0 * (rho ** 2 - 0.5), rtol=1e-14)


def test_eos_tait_and_number_density(oracle):
    # p = p0 ((rho/rho_ref)^gamma - 1), p0 = c^2 rho (rho_ref/rho)^gamma / gamma,
    # and number density is turned into mass density (particles_compute_pressure.F90)
    p = prm(3, stateEquation_type=2, c=10.0, gamma=7.0, rho=1.0, rho_ref=1.0, rhs_density_type=2)
    n = np.array([0.9, 1.0, 1.1]) * 2.0
    m = np.full(3, 0.5)
    out = oracle.pressure(p, n, m)
    assert out[1] == 0.0
    assert np.allclose(out, (100.0 / 7.0) * ((n * m) ** 7 - 1.0), rtol=1e-13)
    with pytest.raises(RuntimeError):
        oracle.pressure(p.replace(rho_ref=0.0), n, m)
