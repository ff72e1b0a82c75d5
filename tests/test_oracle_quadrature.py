"""Pins of the oracle's quadrature (Polylib restatement).

1. bit-identical to the reference's own Polylib.cpp compiled unmodified
   into oracle/_ref (skipped where the reference tree was not available to
   build it);
2. mathematics: monomial exactness of GLL / Gauss-Radau-Jacobi rules,
   exact differentiation of polynomials by D, closed-form GLL corner entries
   (TestPolylib.cpp:79-320 style checks)."""
import ctypes
import os

import numpy as np
import pytest
from numpy.polynomial import legendre as L

import _oracle as O

_D = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(O.REF_POLYLIB):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return ctypes.CDLL(O.REF_POLYLIB)


def _zw(fn, np_, a, b):
    z, w = np.zeros(np_), np.zeros(np_)
    fn(z.ctypes.data_as(_D), w.ctypes.data_as(_D), ctypes.c_int(np_), ctypes.c_double(a), ctypes.c_double(b))
    return z, w


def _dm(fn, z, a, b):
    D = np.zeros(len(z) ** 2)
    fn(D.ctypes.data_as(_D), z.ctypes.data_as(_D), ctypes.c_int(len(z)), ctypes.c_double(a), ctypes.c_double(b))
    return D


ZW = [("or_zwglj", "_ZN7Polylib5zwgljEPdS0_idd", 0.0), ("or_zwgrjm", "_ZN7Polylib6zwgrjmEPdS0_idd", 1.0),
      ("or_zwgrjm", "_ZN7Polylib6zwgrjmEPdS0_idd", 2.0)]
DM = [("or_dglj", "_ZN7Polylib4DgljEPdPKdidd", "or_zwglj", 0.0),
      ("or_dgrjm", "_ZN7Polylib5DgrjmEPdPKdidd", "or_zwgrjm", 1.0),
      ("or_dgrjm", "_ZN7Polylib5DgrjmEPdPKdidd", "or_zwgrjm", 2.0)]


@pytest.mark.parametrize("fo,fr,a", ZW)
@pytest.mark.parametrize("np_", range(2, 17))
def test_zw_bit_identical_to_reference(ref, fo, fr, a, np_):
    z1, w1 = _zw(getattr(O.lib(), fo), np_, a, 0.0)
    z2, w2 = _zw(getattr(ref, fr), np_, a, 0.0)
    assert np.array_equal(z1, z2) and np.array_equal(w1, w2)


@pytest.mark.parametrize("fo,fr,zf,a", DM)
@pytest.mark.parametrize("np_", range(2, 17))
def test_D_bit_identical_to_reference(ref, fo, fr, zf, a, np_):
    z, _ = _zw(getattr(O.lib(), zf), np_, a, 0.0)
    assert np.array_equal(_dm(getattr(O.lib(), fo), z, a, 0.0), _dm(getattr(ref, fr), z, a, 0.0))


@pytest.mark.parametrize("n", [0, 1, 2, 5, 9])
@pytest.mark.parametrize("ab", [(0.0, 0.0), (1.0, 1.0), (2.0, 1.0), (3.0, 1.0)])
def test_jacobfd_bit_identical_to_reference(ref, n, ab):
    z = np.linspace(-0.9, 0.9, 7)
    p1, d1 = O.jacobfd(z, n, *ab)
    p2, d2 = np.zeros(7), np.zeros(7)
    ref._ZN7Polylib7jacobfdEiPKdPdS2_idd(ctypes.c_int(7), z.ctypes.data_as(_D), p2.ctypes.data_as(_D),
                                         d2.ctypes.data_as(_D), ctypes.c_int(n), ctypes.c_double(ab[0]),
                                         ctypes.c_double(ab[1]))
    assert np.array_equal(p1, p2) and np.array_equal(d1, d2)


def _weighted_moment(alpha, k):
    x, w = L.leggauss(80)
    return float(np.sum(w * (1 - x) ** alpha * x ** k))


@pytest.mark.parametrize("nq", range(2, 16))
def test_gll_exactness(nq):
    z, w, D = O.points(O.GLL, nq)
    for k in range(0, 2 * nq - 2):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, z ** k) - exact) < 1e-13
    # one degree beyond must fail (the rule is not exact at 2nq-2)
    k = 2 * nq - 2
    assert abs(np.dot(w, z ** k) - 2.0 / (k + 1)) > 1e-8


@pytest.mark.parametrize("ptype,alpha", [(O.RADAU_A1, 1), (O.RADAU_A2, 2)])
@pytest.mark.parametrize("nq", range(2, 15))
def test_radau_exactness(ptype, alpha, nq):
    z, w, _ = O.points(ptype, nq)
    assert z[0] == -1.0
    for k in range(0, 2 * nq - 1):
        assert abs(np.dot(w, z ** k) - _weighted_moment(alpha, k)) < 1e-12 * max(1, abs(_weighted_moment(alpha, k)))


@pytest.mark.parametrize("ptype", [O.GLL, O.RADAU_A1, O.RADAU_A2])
@pytest.mark.parametrize("nq", range(2, 15))
def test_D_differentiates_polynomials(ptype, nq):
    z, _, D = O.points(ptype, nq)
    D = D.reshape(nq, nq)  # D[k, i] = l_k'(z_i)
    for m in range(nq):
        du = D.T @ z ** m
        exact = m * z ** (m - 1) if m else np.zeros(nq)
        assert np.max(np.abs(du - exact)) < 1e-10 * max(1.0, nq ** 2)


@pytest.mark.parametrize("nq", range(3, 15))
def test_gll_D_corners(nq):
    _, _, D = O.points(O.GLL, nq)
    assert abs(D[0] + nq * (nq - 1) / 4.0) < 1e-12 * nq * nq
    assert abs(D[nq * nq - 1] - nq * (nq - 1) / 4.0) < 1e-12 * nq * nq
