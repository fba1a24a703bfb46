"""Edge cases of the GPU path against the oracle: sparse boxes whose cells
and rows are mostly empty, a crowded cluster inside them (long lists,
ragged chunks), a box holding one particle, and the particle-count limits
of the upload."""
import sys

import numpy as np
import pytest

from mcf_amd import workloads as W
from mcf_amd.sdpd import SDPD, SDPDError, _dp, _ip

from gpu_helpers import gpu_evaluate, oracle_evaluate, oracle_signature, rel_err, vec_err

pytestmark = pytest.mark.gpu


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sparse_box(dim, L, n_spread, n_cluster, seed, symmetry=True):
    """n_spread particles spread uniformly over an L^dim box plus a cluster
    of n_cluster inside a 3-wide square or cube (two to three times the
    mean spacing density), in a random id order."""
    base = W.tait_box_3d(8, kt=1e-3, symmetry=symmetry) if dim == 3 else \
        W.shear_wave_2d(8, brownian=True, kt=1e-4, symmetry=symmetry)
    # unit spacing scale: cut-off 3 in both dimensions
    prm = base.prm.replace(min_phys=[0.0] * 3, max_phys=[L] * dim + [0.0] * (3 - dim), cut_off=3.0)
    rng = np.random.default_rng(seed)
    spread = rng.uniform(0.0, L, size=(n_spread, dim))
    centre = rng.uniform(0.2 * L, 0.8 * L, size=dim)
    cluster = centre + rng.uniform(-1.5, 1.5, size=(n_cluster, dim))
    x = np.ascontiguousarray(np.concatenate([spread, cluster]))
    n = x.shape[0]
    v = rng.normal(0.0, 0.03, size=(n, dim))
    m = np.full(n, 1.0)
    pid = (rng.permutation(n) + 1).astype(np.int32)
    return W.Box(f"sparse{dim}d", prm, x, v, m, pid)


SPARSE = {
    "sparse3d_sym": lambda: sparse_box(3, 24.0, 400, 60, 11),
    "sparse3d_full": lambda: sparse_box(3, 21.0, 300, 40, 12, symmetry=False),
    "sparse2d_sym": lambda: sparse_box(2, 30.0, 200, 30, 13),
}


@pytest.mark.parametrize("name", list(SPARSE))
def test_sparse_boxes(oracle, name):
    box = SPARSE[name]()
    announce(f"sparse box: {name} n={box.n} L={box.prm.max_phys[0]}")
    step = 5
    with SDPD(box.prm) as g:
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        cnt, hs = g.neighbor_signature()
    rc, rh = oracle_signature(oracle, box)
    assert np.array_equal(cnt, rc) and np.array_equal(hs, rh)
    # isolated particles beside crowded ones: ragged lists
    assert rc.min() <= 2 and rc.max() >= 30
    with gpu_evaluate(box, step=step) as g:
        rho, p = g.download_density()
        f = g.download_force()
    ro, po, fo, _ = oracle_evaluate(oracle, box, step=step)
    assert rel_err(rho, ro) < 1e-13
    assert rel_err(p, po) < 1e-12
    assert vec_err(f, fo) < 1e-12


def test_sparse_box_mixed(oracle):
    box = sparse_box(3, 24.0, 400, 60, 11)
    announce(f"sparse box, mixed: n={box.n}")
    with gpu_evaluate(box, step=2) as g:
        f64 = g.download_force()
        r64, _ = g.download_density()
    mx = W.mixed(box)
    with SDPD(mx.prm) as g:
        g.upload(mx.x, mx.v, mx.m, mx.pid)
        g.build_list()
        cnt, hs = g.neighbor_signature()
    rc, rh = oracle_signature(oracle, box)
    assert np.array_equal(cnt, rc) and np.array_equal(hs, rh)
    with gpu_evaluate(mx, step=2) as g:
        fmx = g.download_force()
        rmx, _ = g.download_density()
    assert rel_err(rmx, r64) < 1e-6
    assert vec_err(fmx, f64) < 1e-5


@pytest.mark.parametrize("dim", [2, 3])
def test_one_particle(oracle, dim):
    # no neighbours: the density is the particle's own kernel value and the
    # pair force is zero
    base = sparse_box(dim, 12.0, 1, 0, 21)
    announce(f"one particle: dim={dim}")
    with gpu_evaluate(base, step=1) as g:
        rho, p = g.download_density()
        f = g.download_force()
        pairs, _ = g.counters()
    ro, po, fo, _ = oracle_evaluate(oracle, base, step=1)
    assert pairs == 0
    assert rel_err(rho, ro) < 1e-15 and rel_err(p, po) < 1e-12
    assert np.array_equal(f, fo) and not f.any()


def test_upload_particle_count_limits():
    box = sparse_box(3, 12.0, 8, 0, 31)
    announce("upload limits: n = 0 and n = 2^25")
    with SDPD(box.prm) as g:
        with pytest.raises(SDPDError):
            g.upload(box.x[:0], box.v[:0], box.m[:0], box.pid[:0])
        # the count is checked before any array is read, so small buffers
        # stand in for the 2^25 particles
        ids = np.zeros((box.n, 2), np.int32)
        rc = g.lib.sdpd_upload(g.h, _dp(box.x), _dp(box.v), _dp(box.m), _ip(ids), 1 << 25)
        assert rc != 0
        assert b"2^25" in g.lib.sdpd_last_error(g.h)
        # the context still takes a valid upload afterwards
        g.upload(box.x, box.v, box.m, box.pid)
        g.build_list()
        g.compute_density()
        rho, _ = g.download_density()
    assert np.isfinite(rho).all()
