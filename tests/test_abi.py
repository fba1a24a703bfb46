"""The C-ABI library loads on the CPU host and exports every symbol that
include/sdpd.h declares; parameter validation needs no GPU."""
import ctypes as C
import os
import re

import pytest

from mcf_amd import sdpd as S
from mcf_amd.params import SDPDParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sdpd.h")).read()
    return sorted(set(re.findall(r"\b(sdpd_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    names = declared_symbols()
    assert len(names) >= 20
    assert sorted(n for n, _, _ in S.SYMBOLS) == names


def test_library_exports_every_declared_symbol():
    lib = S.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(S.SDPDError):
        S._lib_saved = S._lib
        try:
            S._lib = None
            S.load_library(str(tmp_path / "nope.so"))
        finally:
            S._lib = S._lib_saved


@pytest.mark.parametrize("bad", [
    dict(num_dim=4),
    dict(cut_off=-1.0),
    dict(stateEquation_type=2, rho_ref=0.0),
    dict(stateEquation_type=1, c=0.0),
    dict(Brownian=True, rhs_density_type=1),
    dict(Newtonian=False, tau=0.0),
    dict(max_phys=[5.0, 5.0, 5.0], cut_off=3.0),  # fewer than 3 cells
    dict(rhs_force_type=5),
    dict(rhs_force_type=1, Brownian=True),  # Morris has no thermal noise
    dict(Newtonian=False, tau=0.1, rhs_force_type=1),  # no Oldroyd-B Morris
    dict(Newtonian=False, tau=0.1, rhs_force_type=4, Brownian=True),  # no noise for it
    dict(body_force_type=4, body_force=[0.1, 0.0, 0.0]),  # the stochastic model
    dict(body_force_type=-1),
])
def test_create_rejects_bad_parameters(bad):
    lib = S.load_library()
    base = dict(num_dim=3, max_phys=[30.0] * 3, cut_off=3.0)
    base.update(bad)
    prm = SDPDParams(**base)
    h = C.c_void_p()
    rc = lib.sdpd_create(C.byref(S.to_c(prm)), C.byref(h))
    assert rc == -1 and not h.value


def test_null_context_calls_fail():
    lib = S.load_library()
    assert lib.sdpd_build_cells(None) == -1
    assert lib.sdpd_step(None, 1) == -1
    assert lib.sdpd_destroy(None) == 0
