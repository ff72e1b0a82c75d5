"""Manufactured-solution helpers on the oracle (CPU): -lap(u) + lam u = f on
[0,1]^3 with u = prod sin(k pi x_i), homogeneous Dirichlet data."""
import numpy as np

import _oracle as O


def oracle_solve(n, nm, nq, lam, k=1, amp=0.0, seed=1, tol=1e-10):
    deformed = amp > 0
    verts = O.hex_mesh(n, amp, seed)
    jac, df = O.hex_geom(verts, n ** 3, nq, deformed)
    xyz = O.hex_coords(verts, n ** 3, nq)
    f = (3 * (k * np.pi) ** 2 + lam) * np.prod(np.sin(k * np.pi * xyz), axis=2).ravel()
    bl = O.iproduct(O.HEX, nm, nq, n ** 3, deformed, jac, f)
    m, dirichlet = O.hex_l2g(n, nm)
    b = np.zeros(len(dirichlet))
    np.add.at(b, m, bl)
    dl = O.helmholtz_diag(O.HEX, nm, nq, n ** 3, deformed, jac, df, lam)
    dg = np.zeros(len(dirichlet))
    np.add.at(dg, m, dl)
    invd = np.where(dirichlet, 0.0, 1.0 / dg)
    x, info = O.cg_helmholtz_hex(n, nm, nq, deformed, jac, df, lam, b, invd, tol, 5000)
    u = O.bwdtrans(O.HEX, nm, nq, n ** 3, x[m])
    exact = np.prod(np.sin(k * np.pi * xyz), axis=2).ravel()
    _, w, _ = O.points(O.GLL, nq)
    W = np.kron(w, np.kron(w, w))
    jq = jac if deformed else np.repeat(jac, nq ** 3)
    l2 = float(np.sqrt(np.sum(np.tile(W, n ** 3) * jq * (u - exact) ** 2)))
    return x, info, l2
