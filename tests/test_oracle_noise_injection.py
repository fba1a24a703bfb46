"""The oracle's injected noise stream (the route by which GPU checks hand
the oracle the device's normals): injecting the oracle's own stream
changes nothing, bit for bit; another stream changes the forces; a table
that lacks an in-cut-off pair fails the evaluation instead of falling back.
"""
import numpy as np
import pytest

from mcf_amd import workloads as W


@pytest.fixture
def inject(oracle):
    yield oracle.set_noise_source
    oracle.set_noise_source(None)


def _run(oracle, box, steps):
    st = oracle.State(box.prm, box.x, box.v, box.m, box.pid)
    st.prologue(0)
    for s in range(1, steps + 1):
        st.vv_step(s)
    return st


@pytest.mark.parametrize("make", [lambda: W.tait_box_3d(10, kt=1e-3),
                                  lambda: W.shear_wave_2d(16, brownian=True, symmetry=False)])
def test_injecting_the_own_stream_is_bitwise_neutral(oracle, inject, make):
    box = make()
    a = _run(oracle, box, 3)
    inject(lambda seed, lo, hi, step: oracle.pair_normals_n(seed, lo, hi, step))
    b = _run(oracle, box, 3)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.v, b.v) and np.array_equal(a.f, b.f)


def test_injected_stream_is_used(oracle, inject):
    box = W.tait_box_3d(10, kt=1e-3)
    rho = oracle.density(box.prm, box.x, box.m)
    p = oracle.pressure(box.prm, rho, box.m)
    f0, _ = oracle.interaction(box.prm, 4, box.x, box.v, rho, p, box.m, box.pid)
    inject(lambda seed, lo, hi, step: -oracle.pair_normals_n(seed, lo, hi, step))
    f1, _ = oracle.interaction(box.prm, 4, box.x, box.v, rho, p, box.m, box.pid)
    assert not np.array_equal(f0, f1)
    # the noise enters linearly: flipping every normal flips the noise force
    box0 = W.tait_box_3d(10, brownian=False)
    fd, _ = oracle.interaction(box0.prm.replace(eta=box.prm.eta), 4, box.x, box.v, rho, p, box.m,
                               box.pid)
    assert np.allclose(f0 - fd, -(f1 - fd), rtol=1e-9, atol=1e-9 * np.abs(f0).max())


def test_missing_pair_fails(oracle, inject):
    import oracle.oracle as O

    box = W.tait_box_3d(10, kt=1e-3)
    rho = oracle.density(box.prm, box.x, box.m)
    p = oracle.pressure(box.prm, rho, box.m)
    inject(lambda seed, lo, hi, step: oracle.pair_normals_n(seed, lo, hi, step))
    real = O._Injected.__init__

    def drop_first(self, *a, **k):
        # the table without its first pair
        real(self, *a, **k)
        if self.active:
            self.keys = np.ascontiguousarray(self.keys[1:])
            self.normals = np.ascontiguousarray(self.normals[1:])

    O._Injected.__init__ = drop_first
    try:
        with pytest.raises(RuntimeError):
            oracle.interaction(box.prm, 4, box.x, box.v, rho, p, box.m, box.pid)
    finally:
        O._Injected.__init__ = real
