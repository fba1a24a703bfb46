"""Shared helpers of the GPU parity tests: run the HIP path and the oracle
on the same seeded inputs and measure the differences."""
import numpy as np

from mcf_amd.sdpd import SDPD, pair_hash


def rel_err(a, b):
    """max |a - b| / max |b|  (for scalar fields)."""
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def vec_err(a, b):
    """max_i |a_i - b_i| / rms_i |b_i|  (for vector fields per particle)."""
    d = np.sqrt(((a - b) ** 2).sum(-1)).max()
    scale = np.sqrt(((b ** 2).sum(-1)).mean())
    return float(d / max(scale, 1e-300))


def gpu_evaluate(box, step=0, ct=None):
    """Upload, cell list, density, forces on the GPU; returns the context."""
    g = SDPD(box.prm)
    g.upload(box.x, box.v, box.m, box.pid)
    if ct is not None:
        g.upload_ct(ct)
    g.step = step
    g.build_list()
    g.compute_density()
    g.compute_interaction()
    return g


def oracle_evaluate(O, box, step=0, ct=None):
    rho = O.density(box.prm, box.x, box.m)
    p = O.pressure(box.prm, rho, box.m)
    f, vgt = O.interaction(box.prm, step, box.x, box.v, rho, p, box.m, box.pid, ct=ct)
    return rho, p, f, vgt


def oracle_signature(O, box):
    pairs = O.neighbor_pairs(box.prm, box.x)
    if box.prm.symmetry:
        a, b = pairs[:, 0], pairs[:, 1]
        aa = np.concatenate([a, b])
        bb = np.concatenate([b, a])
    else:
        aa, bb = pairs[:, 0], pairs[:, 1]
    cnt = np.bincount(aa, minlength=box.n).astype(np.int32)
    h = pair_hash(box.pid[aa], box.pid[bb]).astype(np.uint64)
    hs = np.zeros(box.n, np.uint64)
    np.add.at(hs, aa, h)
    return cnt, (hs & np.uint64(0xFFFFFFFF)).astype(np.uint32)
