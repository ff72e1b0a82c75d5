"""The statically condensed boundary system on the CPU (test
infrastructure), as GlobalLinSysStaticCond builds it
(GlobalLinSysStaticCond.cpp:98-230): element Schur complements assembled
over the global boundary DOFs (Dirichlet DOFs first, then the other
boundary DOFs in increasing global id), the condensed right-hand side
F_bnd - A_bi A_ii^-1 F_int, the rhs magnitude of v_PreSolve
(GlobalLinSysIterativeStaticCond.cpp:436-492: |AssembleBnd(F_bnd)|^2 of the
uncondensed forcing over all boundary rows) and the back-substitution of the
interior."""
import numpy as np


def condensed(blocks, m, sg, fl, bl, ng, ndir):
    rows = np.unique(m[~fl])
    rows = np.concatenate([np.arange(ndir), rows[rows >= ndir]])
    bid = -np.ones(ng, dtype=np.int64)
    bid[rows] = np.arange(len(rows))
    nbg = len(rows)
    S = np.zeros((nbg, nbg))
    b = np.zeros(nbg)
    braw = np.zeros(nbg)
    elem = []
    off = 0
    for Ae in blocks:
        k = len(Ae)
        idx, s, f = m[off:off + k], sg[off:off + k], fl[off:off + k]
        ii, bb = np.flatnonzero(f), np.flatnonzero(~f)
        A = Ae  # the element's own basis; signs enter through the assembly
        fe = bl[off:off + k]
        Aii, Aib, Abb = A[np.ix_(ii, ii)], A[np.ix_(ii, bb)], A[np.ix_(bb, bb)]
        Se = Abb - Aib.T @ np.linalg.solve(Aii, Aib) if len(ii) else Abb
        fc = fe[bb] - (Aib.T @ np.linalg.solve(Aii, fe[ii]) if len(ii) else 0.0)
        gb = bid[idx[bb]]
        sb = s[bb]
        S[np.ix_(gb, gb)] += sb[:, None] * Se * sb[None, :]
        np.add.at(b, gb, sb * fc)
        np.add.at(braw, gb, sb * fe[bb])
        elem.append((off, ii, bb, A, fe, gb, sb))
        off += k
    mag = float(braw @ braw)
    mag = mag if mag > 1e-6 else 1.0
    return dict(rows=rows, S=S, b=b, rhs_mag=mag, elem=elem, nbg=nbg)


def back_substitute(sys, xb, nloc):
    """Local solution from the boundary values xb (boundary-system order)."""
    x = np.zeros(nloc)
    for off, ii, bb, A, fe, gb, sb in sys["elem"]:
        vb = sb * xb[gb]
        x[off + bb] = vb
        if len(ii):
            x[off + ii] = np.linalg.solve(A[np.ix_(ii, ii)], fe[ii] - A[np.ix_(ii, bb)] @ vb)
    return x
