"""Pins of the oracle's per-direction (runtime-size) operators against
dense evaluations that share nothing with it (tests/brute.py):

* hex: Kronecker products of closed-form Modified_A tables on GLL points
  with nm[3] modes and nq[3] points per direction (any of them unequal);
* tet: the basis (closed forms, singular-vertex treatment) evaluated at the
  operator's own collapsed quadrature points, with the Duffy-collapsed
  weights, standard-coordinate gradients from exact polynomial derivatives
  and the collapsed chain rule, modes p < nm0, q < nm1-p, r < nm2-p-q.

Geometry is per point (the deformed path): random J > 0 and random df at
every quadrature point, so that an index mistake in the per-point access
(point q of element e, factor c) changes the result.  The MODIFIED default
sizes are included, which pins the oracle's deformed tet path on its own."""
import numpy as np
import pytest

import _oracle as O
import brute as B
import geometry_cases as G

TOL = 1e-13
NE = 3  # elements per case


def _geom(nqt, seed):
    rng = np.random.default_rng(seed)
    jac = rng.uniform(0.5, 1.5, NE * nqt)
    df = rng.uniform(-1.0, 1.0, NE * nqt * 9)
    return jac, df


def _elem(a, e, n):
    return a[e * n:(e + 1) * n]


HEX_CASES = [((3, 4, 5), (4, 5, 6)), ((4, 4, 4), (6, 5, 7)), ((2, 5, 3), (3, 5, 4)), ((5, 5, 5), (5, 5, 5)),
             ((6, 3, 4), (9, 4, 6)), ((4, 4, 4), (5, 5, 5))]
TET_CASES = [((3, 3, 3), (4, 3, 3)), ((4, 4, 4), (5, 4, 4)), ((5, 5, 5), (7, 6, 6)), ((3, 4, 5), (5, 5, 5)),
             ((2, 3, 4), (4, 6, 5)), ((4, 4, 4), (6, 4, 7))]


@pytest.mark.parametrize("nm,nq", HEX_CASES)
def test_hex_aniso_against_dense(nm, nq):
    Phi, Dd, W = B.hex_dense_aniso(nm, nq)
    nqt, nc = int(np.prod(nq)), int(np.prod(nm))
    assert O.ncoeffs_aniso(O.HEX, nm) == nc
    jac, df = _geom(nqt, sum(nm) + 10 * sum(nq))
    x = G.coeffs(NE * nc, 1)
    u = O.bwdtrans_aniso(O.HEX, nm, nq, NE, x)
    assert G.rel_err(u, np.concatenate([Phi @ _elem(x, e, nc) for e in range(NE)])) < TOL
    f = G.coeffs(NE * nqt, 2)
    y = O.iproduct_aniso(O.HEX, nm, nq, NE, True, jac, f)
    yb = np.concatenate([Phi.T @ (W * _elem(jac, e, nqt) * _elem(f, e, nqt)) for e in range(NE)])
    assert G.rel_err(y, yb) < TOL
    d = O.physderiv_aniso(O.HEX, nm, nq, NE, True, df, f)
    # PhysDeriv of point values: collocation derivative matrices of the
    # point sets (exact for the interpolating polynomial), then df
    Dc = [np.kron(np.eye(nq[2] * nq[1]), B.lagrange_deriv_matrix(B.gll(nq[0])[0]).reshape(nq[0], nq[0]).T),
          np.kron(np.eye(nq[2]), np.kron(B.lagrange_deriv_matrix(B.gll(nq[1])[0]).reshape(nq[1], nq[1]).T,
                                         np.eye(nq[0]))),
          np.kron(B.lagrange_deriv_matrix(B.gll(nq[2])[0]).reshape(nq[2], nq[2]).T, np.eye(nq[1] * nq[0]))]
    for e in range(NE):
        fe = _elem(f, e, nqt)
        dxi = [M @ fe for M in Dc]
        dfe = _elem(df, e, 9 * nqt).reshape(nqt, 9)
        for j in range(3):
            want = sum(dfe[:, j * 3 + k] * dxi[k] for k in range(3))
            assert G.rel_err(_elem(d[j], e, nqt), want) < 1e-12
    d6 = np.array([1.7, 0.2, 1.1, -0.3, 0.15, 0.9])
    for diff in (None, d6):
        h = O.helmholtz_aniso(O.HEX, nm, nq, NE, True, jac, df, 0.7, x, diff)
        hb = np.concatenate([B.dense_helmholtz(Phi, Dd, W, _elem(jac, e, nqt), _elem(df, e, 9 * nqt).reshape(nqt, 9),
                                               0.7, diff) @ _elem(x, e, nc) for e in range(NE)])
        assert G.rel_err(h, hb) < TOL


@pytest.mark.parametrize("nm,nq", TET_CASES)
def test_tet_aniso_own_points_against_dense(nm, nq):
    vals, grads, W = B.tet_dense_own_points(nm, nq)
    nqt, nc = int(np.prod(nq)), len(B.tet_modes_aniso(nm))
    assert O.ncoeffs_aniso(O.TET, nm) == nc
    jac, df = _geom(nqt, 3 * sum(nm) + sum(nq))
    x = G.coeffs(NE * nc, 3)
    u = O.bwdtrans_aniso(O.TET, nm, nq, NE, x)
    assert G.rel_err(u, np.concatenate([vals @ _elem(x, e, nc) for e in range(NE)])) < TOL
    f = G.coeffs(NE * nqt, 4)
    y = O.iproduct_aniso(O.TET, nm, nq, NE, True, jac, f)
    yb = np.concatenate([vals.T @ (W * _elem(jac, e, nqt) * _elem(f, e, nqt)) for e in range(NE)])
    assert G.rel_err(y, yb) < TOL
    # PhysDeriv of u = BwdTrans(x): the exact gradient of the expansion at the points
    d = O.physderiv_aniso(O.TET, nm, nq, NE, True, df, u)
    for e in range(NE):
        xe = _elem(x, e, nc)
        dfe = _elem(df, e, 9 * nqt).reshape(nqt, 9)
        for j in range(3):
            want = sum(dfe[:, j * 3 + k] * (grads[k] @ xe) for k in range(3))
            assert G.rel_err(_elem(d[j], e, nqt), want) < 1e-12
    # IProductWRTDerivBase: sum_j int (dphi/dx_j) f_j
    f1, f2 = G.coeffs(NE * nqt, 5), G.coeffs(NE * nqt, 6)
    yd = O.iproduct_deriv_aniso(O.TET, nm, nq, NE, True, jac, df, f, f1, f2)
    for e in range(NE):
        dfe = _elem(df, e, 9 * nqt).reshape(nqt, 9)
        wj = W * _elem(jac, e, nqt)
        want = sum((sum(dfe[:, j * 3 + k][:, None] * grads[k] for k in range(3))).T @ (wj * _elem(ff, e, nqt))
                   for j, ff in enumerate((f, f1, f2)))
        assert G.rel_err(_elem(yd, e, nc), want) < TOL
    d6 = np.array([1.3, -0.1, 0.8, 0.2, 0.05, 1.4])
    for diff in (None, d6):
        h = O.helmholtz_aniso(O.TET, nm, nq, NE, True, jac, df, 1.9, x, diff)
        hb = np.concatenate([B.dense_helmholtz(vals, grads, W, _elem(jac, e, nqt), _elem(df, e, 9 * nqt).reshape(nqt, 9),
                                               1.9, diff) @ _elem(x, e, nc) for e in range(NE)])
        assert G.rel_err(h, hb) < TOL


def test_default_sizes_match_isotropic_entry_points():
    """The per-direction entry points at the MODIFIED defaults are the
    isotropic ones, bit for bit."""
    for shape, nm, nq0 in ((O.HEX, 5, 6), (O.TET, 5, 6), (O.TET, 4, 6)):
        nq = O.default_nq(shape, nq0)
        nqt = int(np.prod(nq))
        jac, df = _geom(nqt, 9)
        x = G.coeffs(NE * O.ncoeffs(shape, nm), 7)
        a = O.helmholtz(shape, nm, nq0, NE, True, jac, df, 1.1, x)
        b = O.helmholtz_aniso(shape, (nm, nm, nm), nq, NE, True, jac, df, 1.1, x)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape_case", ["tet", "hex"])
def test_planted_per_point_index_mistake_fails(shape_case):
    """The dense pin is sensitive to the per-point geometry index: the
    oracle given factors with two quadrature points of one element swapped
    (what a point-index mistake inside it would read) no longer matches the
    dense evaluation built from the true factors."""
    if shape_case == "tet":
        nm, nq = (4, 4, 4), (5, 4, 4)
        vals, grads, W = B.tet_dense_own_points(nm, nq)
        shape = O.TET
    else:
        nm, nq = (3, 4, 5), (4, 5, 6)
        Phi, Dd, W = B.hex_dense_aniso(nm, nq)
        shape = O.HEX
    nqt = int(np.prod(nq))
    nc = O.ncoeffs_aniso(shape, nm)
    jac, df = _geom(nqt, 31)
    x = G.coeffs(NE * nc, 8)
    good = O.helmholtz_aniso(shape, nm, nq, NE, True, jac, df, 1.9, x)
    for planted in ("jac", "df"):
        j2, d2 = jac.copy(), df.copy()
        e, q0, q1 = 1, 2, nqt - 3
        if planted == "jac":
            j2[e * nqt + q0], j2[e * nqt + q1] = jac[e * nqt + q1], jac[e * nqt + q0]
        else:
            a, b = (e * nqt + q0) * 9, (e * nqt + q1) * 9
            d2[a:a + 9], d2[b:b + 9] = df[b:b + 9], df[a:a + 9]
        bad = O.helmholtz_aniso(shape, nm, nq, NE, True, j2, d2, 1.9, x)
        assert G.rel_err(bad, good) > 1e-6, planted
    if shape_case == "tet":
        hb = np.concatenate([B.dense_helmholtz(vals, grads, W, _elem(jac, e, nqt), _elem(df, e, 9 * nqt).reshape(nqt, 9),
                                               1.9) @ _elem(x, e, nc) for e in range(NE)])
        assert G.rel_err(good, hb) < TOL


PRISM_CASES = [((3, 3, 3), (4, 4, 3)), ((4, 4, 4), (5, 5, 4)), ((3, 4, 5), (5, 4, 5)), ((2, 3, 3), (4, 3, 3)),
               ((5, 5, 5), (7, 7, 6)), ((4, 2, 4), (5, 6, 5))]


def _collapsed_against_dense(shape, dense, nm, nq):
    vals, grads, W = dense
    nqt, nc = int(np.prod(nq)), vals.shape[1]
    assert O.ncoeffs_aniso(shape, nm) == nc
    jac, df = _geom(nqt, 5 * sum(nm) + sum(nq))
    x = G.coeffs(NE * nc, 13)
    u = O.bwdtrans_aniso(shape, nm, nq, NE, x)
    assert G.rel_err(u, np.concatenate([vals @ _elem(x, e, nc) for e in range(NE)])) < TOL
    f = G.coeffs(NE * nqt, 14)
    y = O.iproduct_aniso(shape, nm, nq, NE, True, jac, f)
    yb = np.concatenate([vals.T @ (W * _elem(jac, e, nqt) * _elem(f, e, nqt)) for e in range(NE)])
    assert G.rel_err(y, yb) < TOL
    d = O.physderiv_aniso(shape, nm, nq, NE, True, df, u)
    for e in range(NE):
        xe = _elem(x, e, nc)
        dfe = _elem(df, e, 9 * nqt).reshape(nqt, 9)
        for j in range(3):
            want = sum(dfe[:, j * 3 + k] * (grads[k] @ xe) for k in range(3))
            assert G.rel_err(_elem(d[j], e, nqt), want) < 1e-12
    f1, f2 = G.coeffs(NE * nqt, 15), G.coeffs(NE * nqt, 16)
    yd = O.iproduct_deriv_aniso(shape, nm, nq, NE, True, jac, df, f, f1, f2)
    for e in range(NE):
        dfe = _elem(df, e, 9 * nqt).reshape(nqt, 9)
        wj = W * _elem(jac, e, nqt)
        want = sum((sum(dfe[:, j * 3 + k][:, None] * grads[k] for k in range(3))).T @ (wj * _elem(ff, e, nqt))
                   for j, ff in enumerate((f, f1, f2)))
        assert G.rel_err(_elem(yd, e, nc), want) < TOL
    d6 = np.array([1.3, -0.1, 0.8, 0.2, 0.05, 1.4])
    for diff in (None, d6):
        h = O.helmholtz_aniso(shape, nm, nq, NE, True, jac, df, 1.9, x, diff)
        hb = np.concatenate([B.dense_helmholtz(vals, grads, W, _elem(jac, e, nqt), _elem(df, e, 9 * nqt).reshape(nqt, 9),
                                               1.9, diff) @ _elem(x, e, nc) for e in range(NE)])
        assert G.rel_err(h, hb) < TOL
    # LinearAdvectionDiffusionReaction: lambda M + L plus lambda (phi_i, V . grad u)
    vel = G.coeffs(NE * nqt * 3, 17)
    a = O.adv_diff_react_aniso(shape, nm, nq, NE, True, jac, df, vel, 1.9, x)
    for e in range(NE):
        xe = _elem(x, e, nc)
        dfe = _elem(df, e, 9 * nqt).reshape(nqt, 9)
        ve = _elem(vel, e, 3 * nqt).reshape(nqt, 3)
        gx = [sum(dfe[:, j * 3 + k] * (grads[k] @ xe) for k in range(3)) for j in range(3)]
        adv = sum(ve[:, j] * gx[j] for j in range(3))
        want = B.dense_helmholtz(vals, grads, W, _elem(jac, e, nqt), dfe, 1.9) @ xe + 1.9 * vals.T @ (
            W * _elem(jac, e, nqt) * adv)
        assert G.rel_err(_elem(a, e, nc), want) < TOL


@pytest.mark.parametrize("nm,nq", PRISM_CASES)
def test_prism_own_points_against_dense(nm, nq):
    """Prism operators (BwdTransKernels.hpp:334-405, IProductKernels.hpp:
    519-655, the prism chain rule and metric, HelmholtzKernels.hpp:918-1119,
    IProductWRTDerivBaseKernels.hpp:338-411) against the dense prism
    expansion at the operator's own points, per-point geometry."""
    assert len(B.prism_modes_aniso(nm)) == O.ncoeffs_aniso(O.PRISM, nm)
    _collapsed_against_dense(O.PRISM, B.prism_dense_own_points(nm, nq), nm, nq)


PYR_CASES = [((2, 2, 2), (3, 3, 2)), ((3, 3, 3), (4, 4, 3)), ((4, 4, 4), (5, 5, 4)), ((4, 4, 4), (6, 5, 5)),
             ((5, 5, 5), (7, 7, 6)), ((6, 6, 6), (7, 7, 6))]


@pytest.mark.parametrize("nm,nq", PYR_CASES)
def test_pyramid_own_points_against_dense(nm, nq):
    """Pyramid operators (BwdTransKernels.hpp:407-489, IProductKernels.hpp:
    660-815, the pyramid chain rule PhysDerivKernels.hpp:301-311, the metric
    HelmholtzKernels.hpp:1119-1326, IProductWRTDerivBaseKernels.hpp:417-499,
    LinearAdvectionDiffusionReactionKernels.hpp:475-560) against the dense
    pyramid expansion at the operator's own points, per-point geometry.  The
    direction-2 factors of the dense expansion come from their closed forms,
    not from the oracle's block copies of Basis.cpp."""
    assert len(B.pyr_modes(nm[0])) == O.ncoeffs_aniso(O.PYR, nm)
    _collapsed_against_dense(O.PYR, B.pyr_dense_own_points(nm[0], nq), nm, nq)


@pytest.mark.parametrize("nm,nq", PYR_CASES[1:4])
def test_pyramid_without_top_vertex_term_fails_the_pin(nm, nq):
    """The dense pin sees the top-vertex term (BwdTransKernels.hpp:474-485,
    IProductKernels.hpp:780-812): the expansion without it differs from the
    oracle by far more than the pin's tolerance."""
    bad, bgrads, W = B.pyr_dense_own_points(nm[0], nq, correct=False)
    nqt, nc = int(np.prod(nq)), bad.shape[1]
    x = G.coeffs(NE * nc, 13)
    u = O.bwdtrans_aniso(O.PYR, nm, nq, NE, x)
    assert G.rel_err(u, np.concatenate([bad @ _elem(x, e, nc) for e in range(NE)])) > 1e-3
    jac, df = _geom(nqt, 5 * sum(nm) + sum(nq))
    h = O.helmholtz_aniso(O.PYR, nm, nq, NE, True, jac, df, 1.9, x)
    hb = np.concatenate([B.dense_helmholtz(bad, bgrads, W, _elem(jac, e, nqt), _elem(df, e, 9 * nqt).reshape(nqt, 9),
                                           1.9) @ _elem(x, e, nc) for e in range(NE)])
    assert G.rel_err(h, hb) > 1e-3


def test_pyramid_regular_geometry_and_defaults():
    """Per-element geometry gives the per-point result with constant
    factors, and the isotropic entry points use the pyramid defaults
    (Q, Q, Q-1) and nm(nm+1)(2nm+1)/6 coefficients."""
    nm, nq0 = 4, 5
    nq = O.default_nq(O.PYR, nq0)
    assert nq == (5, 5, 4) and O.ncoeffs(O.PYR, nm) == 30
    nqt = int(np.prod(nq))
    rng = np.random.default_rng(4)
    jac_e = rng.uniform(0.5, 1.5, NE)
    df_e = rng.uniform(-1, 1, NE * 9)
    jac_q = np.repeat(jac_e, nqt)
    df_q = np.concatenate([np.tile(df_e[e * 9:(e + 1) * 9], nqt) for e in range(NE)])
    x = G.coeffs(NE * O.ncoeffs(O.PYR, nm), 19)
    a = O.helmholtz(O.PYR, nm, nq0, NE, False, jac_e, df_e, 0.7, x)
    b = O.helmholtz_aniso(O.PYR, (nm, nm, nm), nq, NE, True, jac_q, df_q, 0.7, x)
    assert G.rel_err(a, b) < 1e-13


@pytest.mark.parametrize("nm,nq", PRISM_CASES[:3])
def test_prism_without_edge_term_fails_the_pin(nm, nq):
    """The dense pin sees the collapsed-edge term (BwdTransKernels.hpp:
    387-397): the expansion without it differs from the oracle by far more
    than the pin's tolerance, in BwdTrans and in the Helmholtz matrix."""
    good, grads, W = B.prism_dense_own_points(nm, nq)
    bad, bgrads, _ = B.prism_dense_own_points(nm, nq, correct=False)
    nqt, nc = int(np.prod(nq)), len(B.prism_modes_aniso(nm))
    x = G.coeffs(NE * nc, 13)
    u = O.bwdtrans_aniso(O.PRISM, nm, nq, NE, x)
    ub = np.concatenate([bad @ _elem(x, e, nc) for e in range(NE)])
    assert G.rel_err(u, ub) > 1e-3
    jac, df = _geom(nqt, 5 * sum(nm) + sum(nq))
    h = O.helmholtz_aniso(O.PRISM, nm, nq, NE, True, jac, df, 1.9, x)
    hb = np.concatenate([B.dense_helmholtz(bad, bgrads, W, _elem(jac, e, nqt), _elem(df, e, 9 * nqt).reshape(nqt, 9),
                                           1.9) @ _elem(x, e, nc) for e in range(NE)])
    assert G.rel_err(h, hb) > 1e-3


def test_prism_regular_geometry_and_defaults():
    """Per-element geometry gives the per-point result with constant
    factors, and the isotropic entry points use the prism defaults
    (Q, Q, Q-1)."""
    nm, nq0 = 4, 5
    nq = O.default_nq(O.PRISM, nq0)
    assert nq == (5, 5, 4)
    nqt = int(np.prod(nq))
    rng = np.random.default_rng(3)
    jac_e = rng.uniform(0.5, 1.5, NE)
    df_e = rng.uniform(-1, 1, NE * 9)
    jac_q = np.repeat(jac_e, nqt)
    df_q = np.concatenate([np.tile(df_e[e * 9:(e + 1) * 9], nqt) for e in range(NE)])
    x = G.coeffs(NE * O.ncoeffs(O.PRISM, nm), 18)
    a = O.helmholtz(O.PRISM, nm, nq0, NE, False, jac_e, df_e, 0.7, x)
    b = O.helmholtz_aniso(O.PRISM, (nm, nm, nm), nq, NE, True, jac_q, df_q, 0.7, x)
    assert G.rel_err(a, b) < 1e-13
