"""sdpd_evaluate_noalloc (the allocation-free evaluation) against the
context path it mirrors: upload, sdpd_build_cells, sdpd_density,
sdpd_forces, download.  Same kernels and launch shapes, so density,
pressure and accelerations must agree bit for bit; two calls in a row (the
second on fewer particles, so leftovers of the first sit in the static
array) must both agree."""
import sys

import numpy as np
import pytest

from mcf_amd import workloads as W
from mcf_amd.params import SDPDParams
from mcf_amd.sdpd import SDPDError, evaluate_noalloc

from gpu_helpers import gpu_evaluate
from hip_buffers import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc(box, step):
    n, dim = box.x.shape
    ids = np.zeros((n, 2), np.int32)
    ids[:, 0] = box.pid
    # (n, dim) row-major is the reference's (dim, n) column-major
    x, v, m, idd = (DeviceArray(a) for a in (box.x, box.v, box.m, ids))
    rho = DeviceArray(np.full(n, -1.0))
    p = DeviceArray(np.full(n, -1.0))
    f = DeviceArray(np.full((n, dim), -1.0))
    evaluate_noalloc(box.prm, step, x.addr, v.addr, m.addr, idd.addr, n, rho.addr, p.addr, f.addr)
    synchronize()
    return rho.get(), p.get(), f.get()


def context_path(box, step):
    g = gpu_evaluate(box, step=step)
    rho, p = g.download_density()
    f = g.download_force()
    g.close()
    return rho, p, f


SEQUENCES = {
    "fp64": [lambda: W.tait_box_3d(12, kt=1e-3), lambda: W.shear_wave_2d(16)],
    "mixed": [lambda: W.mixed(W.tait_box_3d(13, kt=1e-3)),
              lambda: W.mixed(W.tait_box_3d(10, kt=1e-3, symmetry=False))],
    "angular": [lambda: _angular(W.tait_box_3d(11, kt=1e-3)),
                lambda: _angular(W.shear_wave_2d(14, brownian=True))],
}


def _angular(box):
    box.prm = box.prm.replace(rhs_force_type=4)
    return box


@pytest.mark.parametrize("name", list(SEQUENCES))
def test_noalloc_equals_context_path(name):
    for k, make in enumerate(SEQUENCES[name]):
        box = make()
        step = 7 + k
        announce(f"noalloc: {name} call {k} n={box.n}")
        got = noalloc(box, step)
        ref = context_path(box, step)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_noalloc_refuses_what_it_cannot_hold():
    prm = SDPDParams(num_dim=3, max_phys=[30.0] * 3, cut_off=3.0)
    with pytest.raises(SDPDError):  # too many particles (checked before any launch)
        evaluate_noalloc(prm, 0, 1, 1, 1, 1, 70000, 0, 0, 1)
    with pytest.raises(SDPDError):  # conformation-tensor flows are not served
        evaluate_noalloc(prm.replace(Newtonian=False, tau=0.1), 0, 1, 1, 1, 1, 10, 0, 0, 1)
    big = SDPDParams(num_dim=3, max_phys=[600.0] * 3, cut_off=3.0)
    with pytest.raises(SDPDError):  # 398^3 sort-grid cells exceed the array
        evaluate_noalloc(big, 0, 1, 1, 1, 1, 1000, 0, 0, 1)
