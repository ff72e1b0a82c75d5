"""Pins of the oracle's Modified_A/B/C tables (Basis.cpp:466-645) against
independent closed forms built from scipy Jacobi polynomials, and of the
derivative tables against exact polynomial derivatives."""
import numpy as np
import pytest

import _oracle as O
import brute as B


@pytest.mark.parametrize("nm", range(2, 11))
def test_modified_a(nm):
    nq = nm + 1
    z, _, _ = O.points(O.GLL, nq)
    b, db = O.basis(O.MOD_A, nm, nq)
    for p in range(nm):
        f = B.modified_a(p)
        assert np.max(np.abs(b[p] - f(z))) < 1e-13
        assert np.max(np.abs(db[p] - f.deriv()(z))) < 1e-11 * nq


@pytest.mark.parametrize("nm", range(2, 11))
def test_modified_b(nm):
    nq = nm
    z, _, _ = O.points(O.RADAU_A1, nq)
    b, db = O.basis(O.MOD_B, nm, nq)
    m = 0
    for p in range(nm):
        for q in range(nm - p):
            f = B.modified_b(p, q)
            assert np.max(np.abs(b[m] - f(z))) < 1e-13
            assert np.max(np.abs(db[m] - f.deriv()(z))) < 1e-10 * nq
            m += 1
    assert m == b.shape[0]


@pytest.mark.parametrize("nm", range(2, 11))
def test_modified_c(nm):
    nq = nm
    z, _, _ = O.points(O.RADAU_A2, nq)
    b, db = O.basis(O.MOD_C, nm, nq)
    for m, (p, q, r) in enumerate(B.tet_modes(nm)):
        f = B.modified_b(p + q, r)
        assert np.max(np.abs(b[m] - f(z))) < 1e-13
        assert np.max(np.abs(db[m] - f.deriv()(z))) < 1e-10 * nq


def test_modified_a_symmetry():
    # phi_p(-z) = (-1)^p phi_p(z) for bubbles; vertex modes swap
    nm, nq = 9, 10
    z, _, _ = O.points(O.GLL, nq)
    b, _ = O.basis(O.MOD_A, nm, nq)
    assert np.allclose(b[0][::-1], b[1], atol=1e-15)
    for p in range(2, nm):
        assert np.allclose(b[p][::-1], (-1) ** p * b[p], atol=1e-14)
