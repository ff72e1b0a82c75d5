"""ctypes bindings of oracle/liboracle.so (the CPU FP64 restatement of the
reference path).  Test infrastructure: used only by tests/, smoke() and
bench.py's CPU-baseline leg, always as the checker."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_POLYLIB = os.path.join(ROOT, "oracle", "_ref", "libpolylib_ref.so")

HEX, TET, PRISM, PYR = 0, 1, 2, 3
GLL, RADAU_A1, RADAU_A2 = 0, 1, 2
MOD_A, MOD_B, MOD_C, MOD_PYR_C = 0, 1, 2, 3

_D = ctypes.POINTER(ctypes.c_double)
_lib = None


class CGInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("final_error", ctypes.c_double),
                ("rhs_magnitude", ctypes.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run `make -C oracle` (or __graft_entry__.build())")
        _lib = ctypes.CDLL(ORACLE_PATH)
    return _lib


def p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def set_threads(n: int):
    lib().or_set_threads(ctypes.c_int(n))


def ncoeffs(shape, nm):
    if shape == PRISM:
        return nm * nm * (nm + 1) // 2
    if shape == PYR:
        return nm * (nm + 1) * (2 * nm + 1) // 6
    return nm ** 3 if shape == HEX else nm * (nm + 1) * (nm + 2) // 6


def default_nq(shape, nq0):
    if shape in (PRISM, PYR):
        return (nq0, nq0, nq0 - 1)
    return (nq0, nq0, nq0) if shape == HEX else (nq0, nq0 - 1, nq0 - 1)


def nqtot(shape, nq0):
    a, b, c = default_nq(shape, nq0)
    return a * b * c


def points(ptype, nq):
    z, w, D = np.zeros(nq), np.zeros(nq), np.zeros(nq * nq)
    lib().or_points(ptype, nq, p(z), p(w), p(D))
    return z, w, D


def basis(btype, nm, nq):
    n = lib().or_basis_nmodes(btype, nm)
    b, db = np.zeros(n * nq), np.zeros(n * nq)
    lib().or_basis(btype, nm, nq, p(b), p(db))
    return b.reshape(n, nq), db.reshape(n, nq)


def jacobfd(z, n, a, b):
    z = np.ascontiguousarray(z, dtype=np.float64)
    poly, polyd = np.zeros_like(z), np.zeros_like(z)
    lib().or_jacobfd(len(z), p(z), p(poly), p(polyd), n, ctypes.c_double(a), ctypes.c_double(b))
    return poly, polyd


def bwdtrans(shape, nm, nq0, nelmt, x):
    out = np.zeros(nelmt * nqtot(shape, nq0))
    lib().or_bwdtrans(shape, nm, nq0, nelmt, p(np.ascontiguousarray(x)), p(out))
    return out


def iproduct(shape, nm, nq0, nelmt, deformed, jac, f):
    out = np.zeros(nelmt * ncoeffs(shape, nm))
    lib().or_iproduct(shape, nm, nq0, nelmt, int(deformed), p(jac), p(np.ascontiguousarray(f)), p(out))
    return out


def physderiv(shape, nm, nq0, nelmt, deformed, df, u):
    n = nelmt * nqtot(shape, nq0)
    o = [np.zeros(n) for _ in range(3)]
    lib().or_physderiv(shape, nm, nq0, nelmt, int(deformed), p(df), p(np.ascontiguousarray(u)),
                       p(o[0]), p(o[1]), p(o[2]))
    return o


def iproduct_deriv(shape, nm, nq0, nelmt, deformed, jac, df, f0, f1, f2):
    out = np.zeros(nelmt * ncoeffs(shape, nm))
    c = [np.ascontiguousarray(f, dtype=np.float64) for f in (f0, f1, f2)]
    lib().or_iproduct_deriv(shape, nm, nq0, nelmt, int(deformed), p(jac), p(df), p(c[0]), p(c[1]), p(c[2]),
                            p(out))
    return out


def helmholtz(shape, nm, nq0, nelmt, deformed, jac, df, lam, x, diff6=None):
    out = np.zeros(nelmt * ncoeffs(shape, nm))
    d = None if diff6 is None else np.ascontiguousarray(diff6, dtype=np.float64)
    lib().or_helmholtz(shape, nm, nq0, nelmt, int(deformed), p(jac), p(df), ctypes.c_double(lam),
                       p(d), p(np.ascontiguousarray(x)), p(out))
    return out


def adv_diff_react(shape, nm, nq0, nelmt, deformed, jac, df, vel, lam, x, diff6=None):
    """LinearAdvectionDiffusionReaction: vel[(e*nqtot + q)*3 + j] per point."""
    out = np.zeros(nelmt * ncoeffs(shape, nm))
    d = None if diff6 is None else np.ascontiguousarray(diff6, dtype=np.float64)
    lib().or_adv_diff_react(shape, nm, nq0, nelmt, int(deformed), p(jac), p(df),
                            p(np.ascontiguousarray(vel, dtype=np.float64)), ctypes.c_double(lam), p(d),
                            p(np.ascontiguousarray(x)), p(out))
    return out


def interp_nq(shape, nq0, factor):
    out = (ctypes.c_int * 3)()
    lib().or_interp_nq(shape, nq0, ctypes.c_double(factor), out)
    return list(out)


def interp_matrix(ptype, nq_in, nq_out):
    """I[p*nq_out + i] = h_p(zo_i)."""
    out = np.zeros(nq_in * nq_out)
    lib().or_interp_matrix(ptype, nq_in, nq_out, p(out))
    return out


def phys_interp1d_scaled(shape, nq0, nelmt, factor, u):
    no = interp_nq(shape, nq0, factor)
    out = np.zeros(nelmt * no[0] * no[1] * no[2])
    lib().or_phys_interp1d_scaled(shape, nq0, nelmt, ctypes.c_double(factor),
                                  p(np.ascontiguousarray(u, dtype=np.float64)), p(out))
    return out


def helmholtz_diag(shape, nm, nq0, nelmt, deformed, jac, df, lam):
    out = np.zeros(nelmt * ncoeffs(shape, nm))
    lib().or_helmholtz_diag_bruteforce(shape, nm, nq0, nelmt, int(deformed), p(jac), p(df),
                                       ctypes.c_double(lam), p(out))
    return out


def hex_mesh(n, amp, seed):
    v = np.zeros(n ** 3 * 24)
    lib().or_hex_mesh(n, ctypes.c_double(amp), ctypes.c_uint64(seed), p(v))
    return v


def hex_is_regular(verts8x3):
    v = np.ascontiguousarray(verts8x3, dtype=np.float64)
    return bool(lib().or_hex_is_regular(p(v)))


def hex_geom(verts, nelmt, nq, deformed):
    gp = nelmt * (nq ** 3 if deformed else 1)
    jac, df = np.zeros(gp), np.zeros(9 * gp)
    lib().or_hex_geom(p(np.ascontiguousarray(verts)), nelmt, nq, int(deformed), p(jac), p(df))
    return jac, df


def hex_coords(verts, nelmt, nq):
    xyz = np.zeros(nelmt * nq ** 3 * 3)
    lib().or_hex_coords(p(np.ascontiguousarray(verts)), nelmt, nq, p(xyz))
    return xyz.reshape(nelmt, nq ** 3, 3)


def tet_geom(verts, nelmt):
    jac, df = np.zeros(nelmt), np.zeros(9 * nelmt)
    lib().or_tet_geom(p(np.ascontiguousarray(verts)), nelmt, p(jac), p(df))
    return jac, df


def hex_l2g(n, nm):
    nloc = n ** 3 * nm ** 3
    nglob = lib().or_hex_nglobal(n, nm)
    m = np.zeros(nloc, dtype=np.int32)
    d = np.zeros(nglob, dtype=np.uint8)
    lib().or_hex_l2g(n, nm, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return m, d.astype(bool)


def cg_helmholtz_hex(n, nm, nq, deformed, jac, df, lam, b, inv_diag, tol, maxit):
    x = np.zeros_like(b)
    info = CGInfo()
    lib().or_cg_helmholtz_hex(n, nm, nq, int(deformed), p(jac), p(df), ctypes.c_double(lam), p(b),
                              p(inv_diag), ctypes.c_double(tol), maxit, p(x), ctypes.byref(info))
    return x, info


def cg_dense(A, b, ndir=0, inv_diag=None, tol=1e-10, maxit=5000):
    A = np.ascontiguousarray(A, dtype=np.float64)
    x = np.zeros(len(b))
    info = CGInfo()
    lib().or_cg_dense(len(b), ndir, p(A), p(np.ascontiguousarray(b, dtype=np.float64)),
                      p(inv_diag), ctypes.c_double(tol), maxit, p(x), ctypes.byref(info))
    return x, info


def cg_dense_mag(A, b, ndir, inv_diag, rhs_mag, tol=1e-10, maxit=5000):
    A = np.ascontiguousarray(A, dtype=np.float64)
    x = np.zeros(len(b))
    info = CGInfo()
    lib().or_cg_dense_mag(len(b), ndir, p(A), p(np.ascontiguousarray(b, dtype=np.float64)), p(inv_diag),
                          ctypes.c_double(tol), maxit, ctypes.c_double(rhs_mag), p(x), ctypes.byref(info))
    return x, info


def cg_loc(blocks, mp, sign, nglob, ndir, b_loc, inv_diag=None, tol=1e-10, maxit=5000):
    """Element-local CG (or_cg_loc): blocks = list of dense element
    matrices, whose local vectors concatenated in order are mp's layout."""
    boff = np.zeros(len(blocks) + 1, dtype=np.int32)
    boff[1:] = np.cumsum([len(a) for a in blocks])
    A = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(a, dtype=np.float64).ravel() for a in blocks]))
    mp = np.ascontiguousarray(mp, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    sign = None if sign is None else np.ascontiguousarray(sign, dtype=np.float64)
    x = np.zeros(int(boff[-1]))
    info = CGInfo()
    lib().or_cg_loc(len(blocks), boff.ctypes.data_as(ip), p(A), mp.ctypes.data_as(ip), p(sign), nglob, ndir,
                    p(inv_diag), p(np.ascontiguousarray(b_loc, dtype=np.float64)), ctypes.c_double(tol), maxit,
                    p(x), ctypes.byref(info))
    return x, info


def interior_modes(shape, nm):
    """Flags of the interior modes in the coefficient order: hex p, q, r >= 2
    (p fastest); tet p >= 2, q >= 1, r >= 1 (r fastest,
    StdTetExp.cpp:1156-1190)."""
    if shape == HEX:
        r, q, p = np.meshgrid(np.arange(nm), np.arange(nm), np.arange(nm), indexing="ij")
        return ((p >= 2) & (q >= 2) & (r >= 2)).ravel()
    return np.array([p >= 2 and q >= 1 and r >= 1 for p in range(nm) for q in range(nm - p) for r in range(nm - p - q)])


def sc_diagonal(blocks, interior, mp, sign, nglob, ndir):
    """Static-condensation diagonal (or_sc_diagonal); interior: flags over
    the concatenated local entries."""
    boff = np.zeros(len(blocks) + 1, dtype=np.int32)
    boff[1:] = np.cumsum([len(a) for a in blocks])
    A = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(a, dtype=np.float64).ravel() for a in blocks]))
    fl = np.ascontiguousarray(interior, dtype=np.uint8)
    mp = np.ascontiguousarray(mp, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    sign = None if sign is None else np.ascontiguousarray(sign, dtype=np.float64)
    d = np.zeros(nglob)
    rc = lib().or_sc_diagonal(len(blocks), boff.ctypes.data_as(ip), p(A), fl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                              mp.ctypes.data_as(ip), p(sign), nglob, ndir, p(d))
    if rc:
        raise ValueError(f"interior block of element {rc - 1} is not positive definite")
    return d


def splitmix_uniform(seed: int, n: int) -> np.ndarray:
    """SplitMix64 stream mapped to [-1, 1) as 2*(x>>11)*2^-53 - 1."""
    out = np.empty(n)
    s = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out[i] = 2.0 * (z >> 11) * 2.0 ** -53 - 1.0
    return out


# ---- per-direction modes and points (runtime-size operators) ------------
def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def ncoeffs_aniso(shape, nm):
    return int(lib().or_ncoeffs_aniso(shape, _i3(nm)))


def bwdtrans_aniso(shape, nm, nq, nelmt, x):
    out = np.zeros(nelmt * int(np.prod(nq)))
    lib().or_bwdtrans_aniso(shape, _i3(nm), _i3(nq), nelmt, p(np.ascontiguousarray(x)), p(out))
    return out


def iproduct_aniso(shape, nm, nq, nelmt, deformed, jac, f):
    out = np.zeros(nelmt * ncoeffs_aniso(shape, nm))
    lib().or_iproduct_aniso(shape, _i3(nm), _i3(nq), nelmt, int(deformed), p(jac), p(np.ascontiguousarray(f)),
                            p(out))
    return out


def physderiv_aniso(shape, nm, nq, nelmt, deformed, df, u):
    n = nelmt * int(np.prod(nq))
    o = [np.zeros(n) for _ in range(3)]
    lib().or_physderiv_aniso(shape, _i3(nm), _i3(nq), nelmt, int(deformed), p(df), p(np.ascontiguousarray(u)),
                             p(o[0]), p(o[1]), p(o[2]))
    return o


def iproduct_deriv_aniso(shape, nm, nq, nelmt, deformed, jac, df, f0, f1, f2):
    out = np.zeros(nelmt * ncoeffs_aniso(shape, nm))
    c = [np.ascontiguousarray(f, dtype=np.float64) for f in (f0, f1, f2)]
    lib().or_iproduct_deriv_aniso(shape, _i3(nm), _i3(nq), nelmt, int(deformed), p(jac), p(df), p(c[0]), p(c[1]),
                                  p(c[2]), p(out))
    return out


def helmholtz_aniso(shape, nm, nq, nelmt, deformed, jac, df, lam, x, diff6=None):
    out = np.zeros(nelmt * ncoeffs_aniso(shape, nm))
    d = None if diff6 is None else np.ascontiguousarray(diff6, dtype=np.float64)
    lib().or_helmholtz_aniso(shape, _i3(nm), _i3(nq), nelmt, int(deformed), p(jac), p(df), ctypes.c_double(lam),
                             p(d), p(np.ascontiguousarray(x)), p(out))
    return out


def adv_diff_react_aniso(shape, nm, nq, nelmt, deformed, jac, df, vel, lam, x, diff6=None):
    out = np.zeros(nelmt * ncoeffs_aniso(shape, nm))
    d = None if diff6 is None else np.ascontiguousarray(diff6, dtype=np.float64)
    lib().or_adv_diff_react_aniso(shape, _i3(nm), _i3(nq), nelmt, int(deformed), p(jac), p(df),
                                  p(np.ascontiguousarray(vel, dtype=np.float64)), ctypes.c_double(lam), p(d),
                                  p(np.ascontiguousarray(x)), p(out))
    return out
