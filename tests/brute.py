"""Independent dense ('brute force') evaluation of the MODIFIED hex/tet
expansions with numpy/scipy, used to pin the oracle.

Nothing here calls the oracle: points come from numpy Legendre/Jacobi root
finders, basis functions from scipy's Jacobi polynomials, derivatives from
exact polynomial differentiation (numpy.polynomial), and tet integrals from a
Duffy-transformed Gauss-Legendre rule of much higher order."""
from __future__ import annotations

import numpy as np
from numpy.polynomial import Polynomial as Poly
from numpy.polynomial import legendre as L
from scipy.special import eval_jacobi, roots_jacobi


# ---------------------------------------------------------------- points
def gll(n):
    """Gauss-Lobatto-Legendre points/weights (n points)."""
    c = np.zeros(n)
    c[-1] = 1.0
    inner = np.sort(np.real(L.legroots(L.legder(c)))) if n > 2 else np.array([])
    z = np.concatenate([[-1.0], inner, [1.0]])
    Pn1 = L.legval(z, c)
    w = 2.0 / (n * (n - 1) * Pn1 ** 2)
    return z, w


def radau_m(n, alpha):
    """Gauss-Radau-Jacobi points with z=-1 for weight (1-z)^alpha, and the
    weights of int_{-1}^{1} (1-z)^alpha f(z) dz (the un-prescaled Polylib
    weights): exact for degree 2n-2."""
    inner = np.sort(roots_jacobi(n - 1, alpha, 1.0)[0]) if n > 1 else np.array([])
    z = np.concatenate([[-1.0], inner])
    # weights from exactness on monomials (small n, well conditioned enough)
    V = np.vander(z, n, increasing=True).T
    mom = np.array([_moment(alpha, k) for k in range(n)])
    w = np.linalg.solve(V, mom)
    return z, w


def _moment(alpha, k):
    # int_{-1}^{1} (1-z)^alpha z^k dz, exact by high-order Gauss-Legendre
    x, wx = L.leggauss(64)
    return float(np.sum(wx * (1 - x) ** alpha * x ** k))


def lagrange_deriv_matrix(z):
    """D[k*n + i] = l_k'(z_i) by exact polynomial differentiation."""
    n = len(z)
    D = np.zeros((n, n))
    for k in range(n):
        others = np.delete(z, k)
        p = Poly.fromroots(others)
        p = p / p(z[k])
        dp = p.deriv()
        D[k] = dp(z)
    return D.reshape(-1)


# ---------------------------------------------------------------- bases
def _poly_jacobi(n, a, b):
    """P_n^{a,b} as a numpy Polynomial (exact coefficients via fitting)."""
    if n < 0:
        return Poly([0.0])
    x = np.cos(np.linspace(0, np.pi, n + 1)) if n > 0 else np.array([0.0])
    return Poly.fit(x, eval_jacobi(n, a, b, x), n, domain=[-1, 1], window=[-1, 1]) if n > 0 else Poly([1.0])


HALF_M = Poly([0.5, -0.5])  # (1-z)/2
HALF_P = Poly([0.5, 0.5])   # (1+z)/2


def modified_a(p):
    if p == 0:
        return HALF_M
    if p == 1:
        return HALF_P
    return HALF_M * HALF_P * _poly_jacobi(p - 2, 1.0, 1.0)


def modified_b(p, q):
    """Row p, column q of Modified_B (Basis.cpp:499-589 packing)."""
    if p == 0:
        return modified_a(q)
    if p == 1:
        return HALF_M if q == 0 else HALF_M * HALF_P * _poly_jacobi(q - 1, 1.0, 1.0)
    mp = HALF_M ** p
    return mp if q == 0 else mp * HALF_P * _poly_jacobi(q - 1, 2.0 * p - 1.0, 1.0)


def hex_modes(nm):
    return [(p, q, r) for r in range(nm) for q in range(nm) for p in range(nm)]


def tet_modes(nm):
    return [(p, q, r) for p in range(nm) for q in range(nm - p) for r in range(nm - p - q)]


# ------------------------------------------------------- hex dense operator
def hex_dense(nm, nq):
    """Phi[c, m] and its reference-direction derivatives on GLL^3."""
    z, w = gll(nq)
    B = np.array([[modified_a(p)(zi) for p in range(nm)] for zi in z])       # [i, p]
    Bd = np.array([[modified_a(p).deriv()(zi) for p in range(nm)] for zi in z])
    Phi = np.kron(B, np.kron(B, B))
    D0 = np.kron(B, np.kron(B, Bd))
    D1 = np.kron(B, np.kron(Bd, B))
    D2 = np.kron(Bd, np.kron(B, B))
    W = np.kron(w, np.kron(w, w))
    return Phi, (D0, D1, D2), W


def hex_helmholtz_dense(nm, nq, jac, df, lam, diff6=None):
    """Element matrix lam*M + L for one hex (jac scalar or per point, df 9 or
    [nq^3, 9]); returns the dense matrix."""
    Phi, Dd, W = hex_dense(nm, nq)
    nqt = nq ** 3
    jac = np.broadcast_to(np.asarray(jac, dtype=float).reshape(-1), (nqt,)) if np.size(jac) in (1, nqt) else None
    df = np.asarray(df, dtype=float).reshape(-1, 9)
    if df.shape[0] == 1:
        df = np.repeat(df, nqt, axis=0)
    K = np.eye(3)
    if diff6 is not None:
        d = diff6
        K = np.array([[d[0], d[1], d[3]], [d[1], d[2], d[4]], [d[3], d[4], d[5]]])
    A = lam * (Phi.T * (W * jac)) @ Phi
    # physical gradient: grad_x u = sum_k df[j*3+k] d_k u  -> G_j = sum_k df[:, j*3+k] * Dk
    G = [sum(df[:, j * 3 + k][:, None] * Dd[k] for k in range(3)) for j in range(3)]
    for a in range(3):
        for b in range(3):
            if K[a, b] != 0.0:
                A += K[a, b] * (G[a].T * (W * jac)) @ G[b]
    return A


# ------------------------------------------------------- tet dense operator
def tet_basis_fn(nm, mode_index):
    """Full tet basis function (with Nektar's singular-vertex treatment) and
    its exact partial derivatives in collapsed coordinates (a, b, c).
    Returns f(a, b, c) -> (value, d/da, d/db, d/dc)."""
    modes = tet_modes(nm)
    p, q, r = modes[mode_index]
    fa, fb, fc = modified_a(p), modified_b(p, q), modified_b(p + q, r)
    one = Poly([1.0])
    if mode_index == 1:  # top vertex: (psi_a0 + psi_a1)(psi_b00 + psi_b01) psi_c001 = (1+c)/2
        terms = [(one, one, HALF_P)]
    elif p == 0 and q == 1:  # bottom vertex (r = 0) / singular edge (r >= 1)
        terms = [(fa, fb, fc), (HALF_P, HALF_P, fc)]
    else:
        terms = [(fa, fb, fc)]

    def f(a, b, c):
        v = sum(ta(a) * tb(b) * tc(c) for ta, tb, tc in terms)
        da = sum(ta.deriv()(a) * tb(b) * tc(c) for ta, tb, tc in terms)
        db = sum(ta(a) * tb.deriv()(b) * tc(c) for ta, tb, tc in terms)
        dc = sum(ta(a) * tb(b) * tc.deriv()(c) for ta, tb, tc in terms)
        return v, da, db, dc
    return f


def duffy_points(n):
    """Gauss-Legendre rule on the collapsed cube with the Duffy Jacobian
    ((1-b)/2) ((1-c)/2)^2, integrating over the reference tet."""
    x, w = L.leggauss(n)
    A, Bb, C = np.meshgrid(x, x, x, indexing="ij")
    Wt = (w[:, None, None] * w[None, :, None] * w[None, None, :])
    Wt = Wt * (0.5 * (1 - Bb)) * (0.5 * (1 - C)) ** 2
    return A.ravel(), Bb.ravel(), C.ravel(), Wt.ravel()


def tet_dense_mass_stiff(nm, jac, df, lam, nquad=None):
    """Exact (high-order Duffy) lam*M + L for one affine tet; basis gradients
    w.r.t. xi via the collapsed chain rule from exact polynomial derivatives."""
    nc = len(tet_modes(nm))
    nquad = nquad or (nm + 4)
    a, b, c, w = duffy_points(nquad)
    vals = np.zeros((len(a), nc))
    grads = np.zeros((3, len(a), nc))
    for m in range(nc):
        v, da, db, dc = tet_basis_fn(nm, m)(a, b, c)
        vals[:, m] = v
        # chain rule collapsed -> standard (PhysDerivKernels.hpp:198-259)
        d0 = da * 4.0 / ((1 - b) * (1 - c))
        d1 = 0.5 * (1 + a) * d0 + db * 2.0 / (1 - c)
        d2 = 0.5 * (1 + a) * d0 + 0.5 * (1 + b) * db * 2.0 / (1 - c) + dc
        grads[0, :, m], grads[1, :, m], grads[2, :, m] = d0, d1, d2
    df = np.asarray(df).reshape(9)
    G = [sum(df[j * 3 + k] * grads[k] for k in range(3)) for j in range(3)]
    A = lam * (vals.T * (w * jac)) @ vals
    for j in range(3):
        A += (G[j].T * (w * jac)) @ G[j]
    return A, vals, w


# ------------------------------------- per-direction sizes (runtime-size ops)
def hex_dense_aniso(nm, nq):
    """Phi[c, m] and its reference-direction derivatives for modes nm[3] and
    GLL points nq[3] per direction; rows (k, j, i) with i fastest, columns
    (r, q, p) with p fastest."""
    Bs, Bds, Ws = [], [], []
    for d in range(3):
        z, w = gll(nq[d])
        Bs.append(np.array([[modified_a(p)(zi) for p in range(nm[d])] for zi in z]))
        Bds.append(np.array([[modified_a(p).deriv()(zi) for p in range(nm[d])] for zi in z]))
        Ws.append(w)
    Phi = np.kron(Bs[2], np.kron(Bs[1], Bs[0]))
    D0 = np.kron(Bs[2], np.kron(Bs[1], Bds[0]))
    D1 = np.kron(Bs[2], np.kron(Bds[1], Bs[0]))
    D2 = np.kron(Bds[2], np.kron(Bs[1], Bs[0]))
    W = np.kron(Ws[2], np.kron(Ws[1], Ws[0]))
    return Phi, (D0, D1, D2), W


def tet_modes_aniso(nm):
    return [(p, q, r) for p in range(nm[0]) for q in range(nm[1] - p) for r in range(nm[2] - p - q)]


def tet_dense_own_points(nm, nq):
    """Tet basis values and standard-coordinate derivatives at the operator's
    own collapsed quadrature points (GLL x Gauss-Radau(1,0) x Gauss-Radau(2,0),
    nq[3] points, i fastest), and the weights W = w_a (w_b / 2) (w_c / 4) of
    the Duffy-collapsed rule, for modes p < nm0, q < nm1-p, r < nm2-p-q.
    Values come from the closed-form modified bases (with the reference's
    singular-vertex treatment), derivatives from exact polynomial
    differentiation and the collapsed chain rule."""
    za, wa = gll(nq[0])
    zb, wb = radau_m(nq[1], 1.0)
    zc, wc = radau_m(nq[2], 2.0)
    C, Bb, A = np.meshgrid(zc, zb, za, indexing="ij")
    a, b, c = A.ravel(), Bb.ravel(), C.ravel()
    W = (wc[:, None, None] * 0.25 * (wb[None, :, None] * 0.5) * wa[None, None, :]).ravel()
    modes = tet_modes_aniso(nm)
    one = Poly([1.0])
    vals = np.zeros((len(a), len(modes)))
    grads = np.zeros((3, len(a), len(modes)))
    for m, (p, q, r) in enumerate(modes):
        fa, fb, fc = modified_a(p), modified_b(p, q), modified_b(p + q, r)
        if (p, q, r) == (0, 0, 1):  # top vertex: (1+c)/2
            terms = [(one, one, HALF_P)]
        elif p == 0 and q == 1:  # bottom vertex (r = 0) / singular edge (r >= 1)
            terms = [(fa, fb, fc), (HALF_P, HALF_P, fc)]
        else:
            terms = [(fa, fb, fc)]
        v = sum(ta(a) * tb(b) * tc(c) for ta, tb, tc in terms)
        da = sum(ta.deriv()(a) * tb(b) * tc(c) for ta, tb, tc in terms)
        db = sum(ta(a) * tb.deriv()(b) * tc(c) for ta, tb, tc in terms)
        dc = sum(ta(a) * tb(b) * tc.deriv()(c) for ta, tb, tc in terms)
        d0 = da * 4.0 / ((1 - b) * (1 - c))
        d1 = 0.5 * (1 + a) * d0 + db * 2.0 / (1 - c)
        d2 = 0.5 * (1 + a) * d0 + 0.5 * (1 + b) * db * 2.0 / (1 - c) + dc
        vals[:, m] = v
        grads[0, :, m], grads[1, :, m], grads[2, :, m] = d0, d1, d2
    return vals, grads, W


def prism_modes_aniso(nm):
    """Prism modes in the reference's order (StdPrismExp::GetMode): p < nm0,
    q < nm1, r < nm2 - p, r fastest."""
    return [(p, q, r) for p in range(nm[0]) for q in range(nm[1]) for r in range(nm[2] - p)]


def prism_dense_own_points(nm, nq, correct=True):
    """Prism basis values and standard-coordinate derivatives at the
    operator's own collapsed points (GLL x GLL x Gauss-Radau(1,0), nq[3]
    points, i fastest) and the weights W = w_a w_b (w_c / 2) of the collapsed
    rule (xi0 = (1 + a)(1 - c)/2 - 1, xi1 = b, xi2 = c).  Mode (p, q, r) is
    psi^A_p(a) psi^A_q(b) psi^B_pr(c); the collapsed-edge modes (0, q, 1)
    are psi^A_q(b) (1 + c)/2, independent of a.  Derivatives from exact
    polynomial differentiation and the collapsed chain rule
    d/dxi0 = 2/(1 - c) d/da, d/dxi2 = (1 + a)/(1 - c) d/da + d/dc.
    correct=False leaves the collapsed-edge modes as the plain products
    psi^A_0(a) psi^A_q(b) psi^B_01(c), the expansion an operator without the
    collapsed-edge term computes."""
    za, wa = gll(nq[0])
    zb, wb = gll(nq[1])
    zc, wc = radau_m(nq[2], 1.0)
    C, Bb, A = np.meshgrid(zc, zb, za, indexing="ij")
    a, b, c = A.ravel(), Bb.ravel(), C.ravel()
    W = (wc[:, None, None] * 0.5 * wb[None, :, None] * wa[None, None, :]).ravel()
    modes = prism_modes_aniso(nm)
    one = Poly([1.0])
    vals = np.zeros((len(a), len(modes)))
    grads = np.zeros((3, len(a), len(modes)))
    for m, (p, q, r) in enumerate(modes):
        if correct and p == 0 and r == 1:
            ta, tb, tc = one, modified_a(q), HALF_P
        else:
            ta, tb, tc = modified_a(p), modified_a(q), modified_b(p, r)
        v = ta(a) * tb(b) * tc(c)
        da = ta.deriv()(a) * tb(b) * tc(c)
        db = ta(a) * tb.deriv()(b) * tc(c)
        dc = ta(a) * tb(b) * tc.deriv()(c)
        vals[:, m] = v
        grads[0, :, m] = da * 2.0 / (1 - c)
        grads[1, :, m] = db
        grads[2, :, m] = da * (1 + a) / (1 - c) + dc
    return vals, grads, W


def pyr_modes(nm):
    """Pyramid modes in the reference's order (BwdTransKernels.hpp:419-446):
    p < nm, q < nm, r < nm - max(p, q), r fastest."""
    return [(p, q, r) for p in range(nm) for q in range(nm) for r in range(nm - max(p, q))]


def pyr_c(p, q, r):
    """Direction-2 factor of pyramid mode (p, q, r), from its definition in
    Basis.cpp:646-741: Modified_B row max(p, q) when p or q is a vertex
    index (<= 1), else ((1-z)/2)^(p+q-2) for r = 0 and
    ((1-z)/2)^(p+q-2) (1+z)/2 P_{r-1}^{(2p+2q-3, 1)} for r >= 1."""
    if p <= 1 or q <= 1:
        return modified_b(max(p, q), r)
    s = p + q - 2
    if r == 0:
        return HALF_M ** s
    return HALF_M ** s * HALF_P * _poly_jacobi(r - 1, 2.0 * p + 2.0 * q - 3.0, 1.0)


def pyr_dense_own_points(nm, nq, correct=True):
    """Pyramid basis values and standard-coordinate derivatives at the
    operator's own collapsed points (GLL x GLL x Gauss-Radau(2,0), i
    fastest) and the weights W = w_a w_b (w_c / 4) of the collapsed rule
    (xi0 = (1 + a)(1 - c)/2 - 1, xi1 = (1 + b)(1 - c)/2 - 1, xi2 = c).
    Mode (p, q, r) is psi^A_p(a) psi^A_q(b) psi^C_pqr(c); the top vertex
    (0, 0, 1) is (1 + c)/2 on its own (the reference adds
    (psi_0 psi_1 + psi_1 psi_0 + psi_1 psi_1)(1+c)/2 to psi_0 psi_0 (1+c)/2).
    correct=False leaves it as that single product.  The a and b
    derivatives are exact; the c derivative is that of the factor's
    interpolant through the c points (the reference differentiates by
    collocation, dbdata = D bdata, Basis.cpp:736-739, and the factors
    ((1-c)/2)^(p+q-2) reach degree 2nm-4, above what nq2 points interpolate
    exactly).  Chain rule d/dxi0 = 2/(1-c) d/da, d/dxi1 = 2/(1-c) d/db,
    d/dxi2 = (1+a)/(1-c) d/da + (1+b)/(1-c) d/db + d/dc."""
    za, wa = gll(nq[0])
    zb, wb = gll(nq[1])
    zc, wc = radau_m(nq[2], 2.0)
    C, Bb, A = np.meshgrid(zc, zb, za, indexing="ij")
    a, b, c = A.ravel(), Bb.ravel(), C.ravel()
    W = (wc[:, None, None] * 0.25 * wb[None, :, None] * wa[None, None, :]).ravel()
    Dc = lagrange_deriv_matrix(zc).reshape(len(zc), len(zc))  # [k, i] = l_k'(z_i)
    kc = np.repeat(np.arange(len(zc)), len(za) * len(zb))      # c index of each point
    modes = pyr_modes(nm)
    one = Poly([1.0])
    vals = np.zeros((len(a), len(modes)))
    grads = np.zeros((3, len(a), len(modes)))
    for m, (p, q, r) in enumerate(modes):
        if correct and (p, q, r) == (0, 0, 1):
            ta, tb, tc = one, one, HALF_P
        else:
            ta, tb, tc = modified_a(p), modified_a(q), pyr_c(p, q, r)
        vals[:, m] = ta(a) * tb(b) * tc(c)
        da = ta.deriv()(a) * tb(b) * tc(c)
        db = ta(a) * tb.deriv()(b) * tc(c)
        dtc = Dc.T @ tc(zc)
        dc = ta(a) * tb(b) * dtc[kc]
        grads[0, :, m] = da * 2.0 / (1 - c)
        grads[1, :, m] = db * 2.0 / (1 - c)
        grads[2, :, m] = (da * (1 + a) + db * (1 + b)) / (1 - c) + dc
    return vals, grads, W


def dense_helmholtz(vals, grads, W, jac, df, lam, diff6=None):
    """lam M + L from point values / reference-direction gradients (per
    point), per-point jac [nqt] and df [nqt, 9] (df[j*3+k] = dxi_k/dx_j)."""
    K = np.eye(3)
    if diff6 is not None:
        d = diff6
        K = np.array([[d[0], d[1], d[3]], [d[1], d[2], d[4]], [d[3], d[4], d[5]]])
    G = [sum(df[:, j * 3 + k][:, None] * grads[k] for k in range(3)) for j in range(3)]
    A = lam * (vals.T * (W * jac)) @ vals
    for a in range(3):
        for b in range(3):
            if K[a, b] != 0.0:
                A += K[a, b] * (G[a].T * (W * jac)) @ G[b]
    return A
