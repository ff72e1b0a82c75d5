"""Pins of the oracle's LinearAdvectionDiffusionReaction (the MatrixFree
operator: out = lambda IProduct(u + V . grad_x u) + L u with u = BwdTrans(x),
LinearAdvectionDiffusionReaction.h:600-658, advection step
LinearAdvectionDiffusionReactionKernels.hpp:193-386).

* The advection step is checked against an independent code path: the
  oracle's PhysDeriv (pinned by polynomial exactness and brute force) gives
  grad_x u, the pinned IProduct integrates V . grad_x u, and the pinned
  Helmholtz operator gives the rest.  The identity holds for any
  coefficients, velocities, J and df, so it fixes the df index convention,
  the collapsed-coordinate factors of the tet, the per-point velocity
  layout and the lambda scaling of the advection term (the reference
  scales it with lambda; the Collections wrapper passes V / (-lambda)).
* V = 0 reduces exactly to the Helmholtz operator.
* Linearity in V."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G


def _geometry(shape, nq0, deformed, nel):
    if shape == O.HEX:
        n = round(nel ** (1 / 3))
        _, jac, df = G.hex_geometry(n, nq0, deformed)
        return n ** 3, jac, df
    _, jac, df = G.tet_geometry(nel, nq0, deformed)
    return nel, jac, df


def _advect_reference(shape, nm, nq0, nel, deformed, jac, df, vel, lam, x, diff6=None):
    u = O.bwdtrans(shape, nm, nq0, nel, x)
    g = O.physderiv(shape, nm, nq0, nel, deformed, df, u)
    v = vel.reshape(-1, 3)
    adv = v[:, 0] * g[0] + v[:, 1] * g[1] + v[:, 2] * g[2]
    return O.helmholtz(shape, nm, nq0, nel, deformed, jac, df, lam, x, diff6) + \
        lam * O.iproduct(shape, nm, nq0, nel, deformed, jac, adv)


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
@pytest.mark.parametrize("deformed", [False, True], ids=["regular", "deformed"])
@pytest.mark.parametrize("nm", [2, 3, 5, 7])
def test_advection_equals_physderiv_then_iproduct(shape, deformed, nm):
    nq0 = nm + 1
    nel, jac, df = _geometry(shape, nq0, deformed, 8)
    nqt = O.nqtot(shape, nq0)
    x = G.coeffs(nel * O.ncoeffs(shape, nm), 21)
    vel = G.coeffs(nel * nqt * 3, 22)
    lam = 0.75
    out = O.adv_diff_react(shape, nm, nq0, nel, deformed, jac, df, vel, lam, x)
    ref = _advect_reference(shape, nm, nq0, nel, deformed, jac, df, vel, lam, x)
    assert G.rel_err(out, ref) < 1e-12


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
def test_advection_with_diffusion_tensor_uses_the_plain_metric(shape):
    # the constant diffusion tensor changes the Laplacian part only; the
    # advection term keeps grad_x u
    nm, nq0 = 4, 5
    nel, jac, df = _geometry(shape, nq0, True, 8)
    nqt = O.nqtot(shape, nq0)
    x = G.coeffs(nel * O.ncoeffs(shape, nm), 31)
    vel = G.coeffs(nel * nqt * 3, 32)
    d6 = np.array([2.0, 0.3, 1.5, -0.2, 0.1, 1.2])
    out = O.adv_diff_react(shape, nm, nq0, nel, True, jac, df, vel, 1.3, x, d6)
    ref = _advect_reference(shape, nm, nq0, nel, True, jac, df, vel, 1.3, x, d6)
    assert G.rel_err(out, ref) < 1e-12


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
@pytest.mark.parametrize("deformed", [False, True], ids=["regular", "deformed"])
def test_zero_velocity_is_helmholtz(shape, deformed):
    nm, nq0 = 5, 6
    nel, jac, df = _geometry(shape, nq0, deformed, 8)
    x = G.coeffs(nel * O.ncoeffs(shape, nm), 41)
    vel = np.zeros(nel * O.nqtot(shape, nq0) * 3)
    out = O.adv_diff_react(shape, nm, nq0, nel, deformed, jac, df, vel, 2.0, x)
    ref = O.helmholtz(shape, nm, nq0, nel, deformed, jac, df, 2.0, x)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("shape", [O.HEX, O.TET], ids=["hex", "tet"])
def test_linear_in_velocity(shape):
    nm, nq0 = 4, 5
    nel, jac, df = _geometry(shape, nq0, False, 8)
    n = nel * O.nqtot(shape, nq0) * 3
    x = G.coeffs(nel * O.ncoeffs(shape, nm), 51)
    v1, v2 = G.coeffs(n, 52), G.coeffs(n, 53)
    z = np.zeros(n)
    a = O.adv_diff_react(shape, nm, nq0, nel, False, jac, df, v1, 1.0, x)
    b = O.adv_diff_react(shape, nm, nq0, nel, False, jac, df, v2, 1.0, x)
    c = O.adv_diff_react(shape, nm, nq0, nel, False, jac, df, v1 + v2, 1.0, x)
    h = O.adv_diff_react(shape, nm, nq0, nel, False, jac, df, z, 1.0, x)
    assert G.rel_err(c - h, (a - h) + (b - h)) < 1e-12
