"""CPU-side checks of the C ABI: the library loads and exports every symbol
declared in include/*.h; host-only entry points (basis tables, key parsing,
error reporting) behave without a GPU."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(nmf_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def nmf():
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


def test_header_declares_symbols():
    assert len(_declared_symbols()) >= 15


def test_library_exports_every_declared_symbol(nmf):
    L = nmf.lib()
    missing = [s for s in _declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


# bases and points per direction: hex, tet, prism, pyramid
_BTYPE = {0: (O.MOD_A,) * 3, 1: (O.MOD_A, O.MOD_B, O.MOD_C), 2: (O.MOD_A, O.MOD_A, O.MOD_B),
          3: (O.MOD_A, O.MOD_A, O.MOD_PYR_C)}
_PTYPE = {0: (O.GLL,) * 3, 1: (O.GLL, O.RADAU_A1, O.RADAU_A2), 2: (O.GLL, O.GLL, O.RADAU_A1),
          3: (O.GLL, O.GLL, O.RADAU_A2)}


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("nm", [2, 5, 8, 10])
def test_product_basis_tables_match_oracle(nmf, shape, nm):
    """The product's closed-form bases (nektar_amd/csrc/basis.cpp) against
    the oracle's restatement of Basis.cpp (block copies for Modified_C and
    ModifiedPyr_C)."""
    nq0 = nm + 1
    for d in range(3):
        t = nmf.basis_tables(shape, d, nm, nq0)
        btype = _BTYPE[shape][d]
        ptype = _PTYPE[shape][d]
        b, db = O.basis(btype, nm, t["nq"])
        z, w, D = O.points(ptype, t["nq"])
        assert np.max(np.abs(t["z"] - z)) < 2e-16
        assert np.max(np.abs(t["w"] - w)) < 1e-15
        assert np.max(np.abs(t["D"] - D)) < 1e-13 * t["nq"] ** 2
        assert np.max(np.abs(t["bdata"] - b.ravel())) < 1e-14
        assert np.max(np.abs(t["dbdata"] - db.ravel())) < 1e-12 * t["nq"] ** 2


def test_bad_keys_report_errors(nmf):
    h = ctypes.c_void_p()
    L = nmf.lib()
    rc = L.nmf_operator_create(b"Helmholtz_Quad_Regular", 4, 5, 1, 0, ctypes.byref(h))
    assert rc == 2 and b"Quad" in L.nmf_last_error()
    rc = L.nmf_operator_create(b"Foo_Hex_Regular", 4, 5, 1, 0, ctypes.byref(h))
    assert rc == 2
    rc = L.nmf_operator_create(b"Helmholtz_Hex", 4, 5, 1, 0, ctypes.byref(h))
    assert rc == 1
    # orders without a compile-time kernel run the runtime-size kernels;
    # beyond those kernels' limits the order is refused before any device work
    rc = L.nmf_operator_create(b"Helmholtz_Hex_Regular", 33, 34, 1, 0, ctypes.byref(h))
    assert rc == 2 and b"modes per direction" in L.nmf_last_error()
    m, q = (ctypes.c_int * 3)(4, 3, 5), (ctypes.c_int * 3)(5, 5, 5)
    rc = L.nmf_operator_create_basis(b"Helmholtz_Tet_Regular", m, q, 1, 0, ctypes.byref(h))
    assert rc == 2 and b"nm0 <= nm1 <= nm2" in L.nmf_last_error()
    rc = L.nmf_operator_create(b"Helmholtz_Hex_Regular", 4, 5, 1, 7, ctypes.byref(h))
    assert rc == 1
    assert L.nmf_operator_set_lambda(None, 1.0) == 1
