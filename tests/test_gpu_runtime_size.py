"""Runtime-size operators (any modes and points per direction) against the
oracle's per-direction restatement, which tests/test_oracle_aniso.py pins
on dense evaluations.  The reference runs these shapes through its
runtime-size operator3D (MatrixFreeOps/SwitchNodesPoints.h:302-343,
Helmholtz.h:374-464).

Cases: anisotropic modes and points on hexes, tets and prisms, equal modes
with any points on pyramids (prisms and pyramids run the runtime-size
kernels at every size), isotropic orders
without a compile-time kernel (hex P=11 / Q=14, Q = 2 nm; tet P=10), and
sizes whose buffers exceed one workgroup's LDS, which run from the global
workspace (including more elements than workgroups of that grid).  Every
operator of the factory (BwdTrans, IProduct, PhysDeriv,
IProductWRTDerivBase, Helmholtz with and without a constant diffusion
tensor, LinearAdvectionDiffusionReaction), regular and deformed geometry:
FP64 relative max error <= 1e-12 up to 10 modes per direction, FP32
<= 5e-5.  Above 10 modes the FP64 bound is 1e-11: the library's quadrature
tables (basis.cpp: Golub-Welsch points, barycentric derivative matrices)
and the oracle's (a restatement of the reference's Polylib) agree to the
last few bits only, and at 11 modes the collapsed-coordinate factors
2/(1-z) of the tet's PhysDeriv amplify that difference to 2-3e-12 on a
smooth field (the same element code compiled for the host shows it too,
tests/cpu_generic).  The tet PhysDeriv multiplies the collapsed
derivatives by up to max(2/(1-z1)) max(2/(1-z2)) (the chain rule of
PhysDerivKernels.hpp:198-259), which grows with the Radau point counts
(10 at 3 points, 800 at 11, 2000 at 14); its bounds (FP64 and FP32) are
multiplied by a tenth of that factor when that is larger than 1."""
import numpy as np
import pytest

import _oracle as O
import geometry_cases as G

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HEX = [((3, 4, 5), (4, 5, 6), 37), ((4, 4, 4), (6, 5, 7), 20), ((6, 3, 4), (9, 4, 6), 15),
       ((4, 4, 4), (8, 8, 8), 24), ((2, 5, 3), (3, 5, 4), 33), ((12, 12, 12), (14, 14, 14), 9),
       ((11, 9, 10), (13, 13, 13), 1100)]
TET = [((3, 4, 5), (5, 5, 5), 37), ((2, 3, 4), (4, 6, 5), 21), ((4, 4, 4), (6, 4, 7), 19),
       ((5, 5, 5), (7, 7, 7), 16), ((11, 11, 11), (12, 11, 11), 10), ((8, 10, 12), (16, 14, 14), 1030)]
# prisms always run the runtime-size kernels: the MODIFIED defaults (P=3, 6,
# 9), anisotropic sizes, and a size from the global workspace with more
# elements than workgroups
PRISM = [((4, 4, 4), (5, 5, 4), 37), ((7, 7, 7), (8, 8, 7), 64), ((3, 4, 5), (5, 4, 5), 15),
         ((4, 2, 4), (5, 6, 5), 33), ((10, 10, 10), (11, 11, 10), 12), ((12, 12, 12), (14, 14, 13), 1030)]


# pyramids: equal modes in the three directions (the reference's
# precondition), the MODIFIED defaults, other points, a global-workspace size
PYR = [((4, 4, 4), (5, 5, 4), 37), ((7, 7, 7), (8, 8, 7), 64), ((3, 3, 3), (5, 4, 5), 15),
       ((2, 2, 2), (3, 3, 2), 33), ((10, 10, 10), (11, 11, 10), 12), ((12, 12, 12), (14, 14, 13), 1030)]


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


def _geom(E, nqt, deformed, seed):
    rng = np.random.default_rng(seed)
    n = E * (nqt if deformed else 1)
    jac = rng.uniform(0.5, 1.5, n)
    df = rng.uniform(-1.0, 1.0, 9 * n)
    return jac, df


def _run(nmf, key, nm, nq, E, prec, jac, df, ins, outs, lam=None, diff=None, vel=None):
    op = nmf.GetOperatorFactory().CreateInstance(key, list(nm), list(nq), E, prec)
    assert op.runtime_size, key
    if op.op in ("IProduct", "Helmholtz", "IProductWRTDerivBase", "LinearAdvectionDiffusionReaction"):
        op.SetJac(jac)
    if op.op in ("PhysDeriv", "Helmholtz", "IProductWRTDerivBase", "LinearAdvectionDiffusionReaction"):
        op.SetDF(df)
    if lam is not None:
        op.SetLambda(lam)
    if diff is not None:
        op.SetConstVarDiffusion(diff)
    if vel is not None:
        op.SetAdvectionVelocities(vel)
    dt = torch.float32 if prec == nmf.FP32 else torch.float64
    tin = [torch.from_numpy(a).to(dt).cuda() for a in ins]
    tout = [torch.full((n,), float("nan"), dtype=dt, device="cuda") for n in outs]
    if op.op == "IProductWRTDerivBase":
        op(tin, tout[0])
    else:
        op(tin[0], *tout)
    torch.cuda.synchronize()
    return [t.double().cpu().numpy() for t in tout]


def _collapse_factor(nmf, nq, shape="Tet"):
    """max 2/(1-z1) * max 2/(1-z2) over the tet's Radau points; max
    2/(1-z2) over the prism's and the pyramid's"""
    if shape in ("Prism", "Pyr"):
        sh = nmf.SHAPE_PRISM if shape == "Prism" else nmf.SHAPE_PYR
        z2 = nmf.basis_tables(sh, 2, 2, nq[2] + 1)["z"]
        return float(np.max(2 / (1 - z2)))
    z1 = nmf.basis_tables(nmf.SHAPE_TET, 1, 2, nq[1] + 1)["z"]
    z2 = nmf.basis_tables(nmf.SHAPE_TET, 2, 2, nq[2] + 1)["z"]
    return float(np.max(2 / (1 - z1)) * np.max(2 / (1 - z2)))


def _check(got, want, prec, nmf, what):
    nm = what[1]
    tol = 5e-5 if prec == nmf.FP32 else (1e-12 if max(nm) <= 10 else 1e-11)
    if what[0] in ("Tet", "Prism", "Pyr") and what[-1] == "PhysDeriv":
        tol *= max(1.0, _collapse_factor(nmf, what[2], what[0]) / 10)
    for g, w in zip(got, want):
        err = G.rel_err(g, w)
        assert err <= tol, (what[-1], what, err)


@pytest.mark.parametrize("shape", ["Hex", "Tet", "Prism", "Pyr"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("deformed", [False, True], ids=["regular", "deformed"])
def test_runtime_size_operators_against_oracle(nmf, shape, prec, deformed):
    P = nmf.FP64 if prec == "fp64" else nmf.FP32
    S = {"Hex": O.HEX, "Tet": O.TET, "Prism": O.PRISM, "Pyr": O.PYR}[shape]
    cases = {"Hex": HEX, "Tet": TET, "Prism": PRISM, "Pyr": PYR}[shape]
    geo = "Deformed" if deformed else "Regular"
    for ci, (nm, nq, E) in enumerate(cases):
        if prec == "fp32" and E > 100:
            continue
        nqt, nc = int(np.prod(nq)), O.ncoeffs_aniso(S, nm)
        jac, df = _geom(E, nqt, deformed, 100 + ci)
        x = G.coeffs(E * nc, 200 + ci)
        f = [G.coeffs(E * nqt, 300 + ci + k) for k in range(3)]
        what = (shape, nm, nq, E, prec, geo)
        # BwdTrans
        u = O.bwdtrans_aniso(S, nm, nq, E, x)
        _check(_run(nmf, f"BwdTrans_{shape}_Regular", nm, nq, E, P, None, None, [x], [E * nqt]), [u], P, nmf,
               what + ("BwdTrans",))
        # IProductWRTBase
        y = O.iproduct_aniso(S, nm, nq, E, deformed, jac, f[0])
        _check(_run(nmf, f"IProduct_{shape}_{geo}", nm, nq, E, P, jac, df, [f[0]], [E * nc]), [y], P, nmf,
               what + ("IProduct",))
        # PhysDeriv
        d = O.physderiv_aniso(S, nm, nq, E, deformed, df, u)
        _check(_run(nmf, f"PhysDeriv_{shape}_{geo}", nm, nq, E, P, jac, df, [u], [E * nqt] * 3), list(d), P, nmf,
               what + ("PhysDeriv",))
        # IProductWRTDerivBase
        yd = O.iproduct_deriv_aniso(S, nm, nq, E, deformed, jac, df, f[0], f[1], f[2])
        _check(_run(nmf, f"IProductWRTDerivBase_{shape}_{geo}", nm, nq, E, P, jac, df, f, [E * nc]), [yd], P, nmf,
               what + ("IProductWRTDerivBase",))
        # Helmholtz, identity and constant diffusion tensor
        for diff in (None, np.array([1.3, -0.1, 0.8, 0.2, 0.05, 1.4])):
            h = O.helmholtz_aniso(S, nm, nq, E, deformed, jac, df, 0.7, x, diff)
            _check(_run(nmf, f"Helmholtz_{shape}_{geo}", nm, nq, E, P, jac, df, [x], [E * nc], lam=0.7, diff=diff),
                   [h], P, nmf, what + ("Helmholtz", diff is not None))
        # LinearAdvectionDiffusionReaction
        vel = np.random.default_rng(400 + ci).uniform(-1.0, 1.0, E * nqt * 3)
        a = O.adv_diff_react_aniso(S, nm, nq, E, deformed, jac, df, vel, 1.3, x)
        _check(_run(nmf, f"LinearAdvectionDiffusionReaction_{shape}_{geo}", nm, nq, E, P, jac, df, [x], [E * nc],
                    lam=1.3, vel=vel), [a], P, nmf, what + ("LADR",))
        print(f"{shape} {geo} {prec} nm={nm} nq={nq} E={E}: all operators within tolerance", flush=True)


def test_compiled_and_runtime_choice(nmf):
    """Equal modes with the default points of an instantiated pair take the
    compile-time kernels; anything else takes the runtime-size ones.
    PhysInterp1DScaled is refused off the default points."""
    F = nmf.GetOperatorFactory()
    assert not F.CreateInstance("Helmholtz_Hex_Regular", [5, 5, 5], [6, 6, 6], 4).runtime_size
    assert not F.CreateInstance("Helmholtz_Tet_Regular", [5, 5, 5], [6, 5, 5], 4).runtime_size
    assert F.CreateInstance("Helmholtz_Hex_Regular", [5, 5, 6], [6, 6, 7], 4).runtime_size
    assert F.CreateInstance("Helmholtz_Tet_Regular", [5, 5, 5], [6, 6, 6], 4).runtime_size
    assert F.CreateInstance("Helmholtz_Hex_Regular", 12, 14, 4).runtime_size
    with pytest.raises(nmf.NektarError):
        F.CreateInstance("PhysInterp1DScaled_Hex_Regular", [5, 5, 6], [6, 6, 7], 4)
    with pytest.raises(nmf.NektarError):
        F.CreateInstance("Helmholtz_Tet_Regular", [5, 4, 5], [6, 5, 5], 4)  # nm0 <= nm1 <= nm2
    with pytest.raises(nmf.NektarError):
        F.CreateInstance("Helmholtz_Hex_Regular", [33, 4, 4], [34, 6, 6], 4)
    # prisms: runtime-size at the defaults too, nm0 <= nm2, no PhysInterp1DScaled
    assert F.CreateInstance("Helmholtz_Prism_Regular", 7, 8, 4).runtime_size
    with pytest.raises(nmf.NektarError):
        F.CreateInstance("Helmholtz_Prism_Regular", [5, 4, 4], [6, 6, 5], 4)
    with pytest.raises(nmf.NektarError):
        F.CreateInstance("PhysInterp1DScaled_Prism_Regular", 5, 6, 4)
    # pyramids: runtime-size, equal modes only, no PhysInterp1DScaled
    assert F.CreateInstance("Helmholtz_Pyr_Regular", 7, 8, 4).runtime_size
    with pytest.raises(nmf.NektarError):
        F.CreateInstance("Helmholtz_Pyr_Regular", [5, 5, 6], [6, 6, 6], 4)
    with pytest.raises(nmf.NektarError):
        F.CreateInstance("PhysInterp1DScaled_Pyr_Regular", 5, 6, 4)
