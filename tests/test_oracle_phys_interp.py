"""Pins of the oracle's PhysInterp1DScaled (PhysInterp1DScaled.h:360-430,
PhysInterp1DScaledKernels.hpp:151-217; interpolation matrices from
GaussPoints.cpp:283-380).

1. The 1-D interpolation matrices agree with the reference's own Polylib
   Imglj / Imgrjm (oracle/_ref, compiled unmodified from the reference
   tree; skipped where that tree was not available) to 1e-13 -- the oracle
   evaluates the same Lagrange polynomials by the product formula.
2. Mathematics: the tensor interpolation reproduces, at the output points,
   any polynomial of degree < nq_in in each direction (closed form); factor
   1 is the identity; the point counts follow Operator.hpp:568-575."""
import ctypes
import os

import numpy as np
import pytest

import _oracle as O

_D = ctypes.POINTER(ctypes.c_double)
PT = {O.GLL: ("_ZN7Polylib5ImgljEPdPKdS2_iidd", 0.0), O.RADAU_A1: ("_ZN7Polylib6ImgrjmEPdPKdS2_iidd", 1.0),
      O.RADAU_A2: ("_ZN7Polylib6ImgrjmEPdPKdS2_iidd", 2.0)}


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(O.REF_POLYLIB):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return ctypes.CDLL(O.REF_POLYLIB)


@pytest.mark.parametrize("ptype", [O.GLL, O.RADAU_A1, O.RADAU_A2], ids=["gll", "radau1", "radau2"])
@pytest.mark.parametrize("nin,nout", [(3, 4), (5, 7), (8, 12), (9, 13), (12, 18), (7, 7)])
def test_matrix_matches_reference_polylib(ref, ptype, nin, nout):
    zi = O.points(ptype, nin)[0]
    zo = O.points(ptype, nout)[0]
    name, alpha = PT[ptype]
    im = np.zeros(nin * nout)
    getattr(ref, name)(im.ctypes.data_as(_D), zi.ctypes.data_as(_D), zo.ctypes.data_as(_D), ctypes.c_int(nin),
                       ctypes.c_int(nout), ctypes.c_double(alpha), ctypes.c_double(0.0))
    mine = O.interp_matrix(ptype, nin, nout)
    assert np.max(np.abs(mine - im)) < 1e-13


def test_point_counts():
    assert O.interp_nq(O.HEX, 8, 1.5) == [12, 12, 12]
    assert O.interp_nq(O.TET, 8, 1.5) == [12, 11, 11]
    assert O.interp_nq(O.HEX, 5, 2.0) == [10, 10, 10]
    assert O.interp_nq(O.TET, 5, 2.0) == [10, 9, 9]
    assert O.interp_nq(O.TET, 7, 1.0) == [7, 6, 6]


def _grid(ptypes, n):
    z = [O.points(pt, k)[0] for pt, k in zip(ptypes, n)]
    # i (direction 0) fastest
    Z2, Z1, Z0 = np.meshgrid(z[2], z[1], z[0], indexing="ij")
    return Z0.ravel(), Z1.ravel(), Z2.ravel()


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
@pytest.mark.parametrize("nq0", [3, 5, 8, 11])
@pytest.mark.parametrize("factor", [1.5, 2.0])
def test_reproduces_polynomials(shape, nq0, factor):
    ni = O.default_nq(shape, nq0)
    no = O.interp_nq(shape, nq0, factor)
    pts = [O.GLL, O.RADAU_A1, O.RADAU_A2] if shape == O.TET else [O.GLL] * 3
    rng = np.random.default_rng(nq0)
    ca = [rng.uniform(-1, 1, k) for k in ni]   # degree ni[d]-1 in direction d
    cb = [rng.uniform(-1, 1, k) for k in ni]

    def f(x, y, z):
        pa = np.polynomial.polynomial
        return pa.polyval(x, ca[0]) * pa.polyval(y, ca[1]) * pa.polyval(z, ca[2]) + \
            pa.polyval(x, cb[0]) * pa.polyval(y, cb[1]) * pa.polyval(z, cb[2])

    nel = 3
    u = np.concatenate([(e + 1.0) * f(*_grid(pts, ni)) for e in range(nel)])
    out = O.phys_interp1d_scaled(shape, nq0, nel, factor, u)
    exact = np.concatenate([(e + 1.0) * f(*_grid(pts, no)) for e in range(nel)])
    assert np.max(np.abs(out - exact)) <= 1e-11 * np.max(np.abs(exact))


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
def test_factor_one_is_identity(shape):
    nq0 = 7
    n = O.nqtot(shape, nq0)
    u = np.random.default_rng(3).uniform(-1, 1, 4 * n)
    assert np.array_equal(O.phys_interp1d_scaled(shape, nq0, 4, 1.0, u), u)
