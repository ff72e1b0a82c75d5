"""The reference's own regression test of the C0 Helmholtz solve,
library/Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex.tst, on the GPU path:
the session file (8 hexahedra of [0,1]^3, NUMMODES 6, lambda 1,
u = sin(pi x/2) sin(pi y/2) sin(pi z/2) imposed on the whole boundary) read
by nektar_amd.contfield, the Dirichlet data projected as ContField does,
rhs -IProductWRTBase(f) from the GPU operator, the lifted CG solve on the
GPU with each of the reference's iterative solver choices, BwdTrans on the
GPU.  The metrics must match the .tst's values within its tolerance
(L2 1.11818e-07, Linf 5.66374e-07, 1e-12); likewise Helmholtz3D_CG_Tet.tst
(48 tetrahedra, NUMMODES 5) and Helmholtz3D_CG_Hex_AllBCs.tst (Neumann and
Robin faces)."""
import os

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
HEX = os.path.join(HERE, "golden", "helmholtz3d", "Helmholtz3D_Hex_P6.xml")
TET = os.path.join(HERE, "golden", "helmholtz3d", "Helmholtz3D_Tet.xml")


@pytest.fixture(scope="module")
def nmf():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import nektar_amd
    nektar_amd.lib()
    return nektar_amd


@pytest.mark.parametrize("solver", ["ConjugateGradient", "ConjugateGradientLoc", "IterativeStaticCond"])
def test_helmholtz3d_cg_hex(nmf, solver):
    from nektar_amd.contfield import helmholtz_solve

    l2, linf, info, _ = helmholtz_solve(HEX, solver=solver, tol=1e-14)
    print(f"Helmholtz3D_CG_Hex ({solver}): L2 {l2:.10e} Linf {linf:.10e}, {info.iterations} iterations", flush=True)
    assert info.converged
    assert abs(l2 - 1.11818e-07) <= 1e-12, l2
    assert abs(linf - 5.66374e-07) <= 1e-12, linf


@pytest.mark.parametrize("solver", ["ConjugateGradient", "ConjugateGradientLoc", "IterativeStaticCond"])
def test_helmholtz3d_cg_tet(nmf, solver):
    """Helmholtz3D_CG_Tet.tst: 48 tets, NUMMODES 5, the same conditions;
    L2 8.458e-05 (tolerance 1e-8), Linf 0.000862348 (1e-9)."""
    from nektar_amd.contfield import helmholtz_solve

    l2, linf, info, _ = helmholtz_solve(TET, solver=solver, tol=1e-14)
    print(f"Helmholtz3D_CG_Tet ({solver}): L2 {l2:.10e} Linf {linf:.10e}, {info.iterations} iterations", flush=True)
    assert info.converged
    assert abs(l2 - 8.458e-05) <= 1e-8, l2
    assert abs(linf - 0.000862348) <= 1e-9, linf


ALLBCS = os.path.join(HERE, "golden", "helmholtz3d", "Helmholtz3D_Hex_AllBCs_P6.xml")


@pytest.mark.parametrize("solver", ["ConjugateGradient", "ConjugateGradientLoc"])
def test_helmholtz3d_cg_hex_allbcs(nmf, solver):
    """Helmholtz3D_CG_Hex_AllBCs.tst: 27 hexahedra, NUMMODES 4, Dirichlet,
    Neumann and Robin (PRIMCOEFF 2, 3) faces; the reference prints
    L2 0.000416575 and Linf 0.000871589 (tolerance 1e-12 on six digits).
    The Robin face mass runs in the GPU matrix multiply, the lifting and
    the Jacobi diagonal."""
    from nektar_amd.contfield import helmholtz_solve

    l2, linf, info, _ = helmholtz_solve(ALLBCS, solver=solver, tol=1e-14)
    print(f"Helmholtz3D_CG_Hex_AllBCs ({solver}): L2 {l2:.10e} Linf {linf:.10e}, {info.iterations} iterations",
          flush=True)
    assert info.converged
    assert f"{l2:.6g}" == "0.000416575", l2
    assert f"{linf:.6g}" == "0.000871589", linf


def test_allbcs_matches_dense_and_rejects_static_condensation(nmf):
    """The GPU solution equals the oracle's dense solve with the Robin mass;
    static condensation refuses Robin rows."""
    import numpy as np

    import test_contfield_cpu as T
    from nektar_amd import contfield as cf

    s = cf.read_session(ALLBCS)
    xd, _ = T.dense_solve(s, cf.dirichlet_problem(s))
    _, _, info, xl = cf.helmholtz_solve(s, tol=1e-14)
    assert info.converged
    assert np.max(np.abs(xl - xd)) <= 1e-10 * np.max(np.abs(xd))
    with pytest.raises(nmf.NektarError):
        cf.helmholtz_solve(s, solver="IterativeStaticCond", tol=1e-14)


def test_trilinear_hexahedra(nmf):
    """The same session with its inner vertex coordinates perturbed
    (trilinear elements, per-point geometry, boundary faces that are not
    parallelograms): the GPU solve equals the oracle's dense solve."""
    import numpy as np

    import test_contfield_cpu as T
    from nektar_amd.contfield import dirichlet_problem, helmholtz_solve

    s = T.perturbed_hex_session()
    xd, (l2d, linfd) = T.dense_solve(s, dirichlet_problem(s))
    l2, linf, info, xl = helmholtz_solve(s, tol=1e-14)
    print(f"trilinear hexes: L2 {l2:.10e} (dense {l2d:.10e}), Linf {linf:.10e} (dense {linfd:.10e})", flush=True)
    assert info.converged
    # the error norms move by at most the pointwise difference of the two
    # solutions (a few 1e-15 here, the solve's rounding level), so they are
    # compared absolutely, 1000x finer than the reference's own 1e-9
    assert abs(l2 - l2d) <= 1e-12 and abs(linf - linfd) <= 1e-12
    assert np.max(np.abs(xl - xd)) <= 1e-10 * np.max(np.abs(xd))


def test_repeatable(nmf):
    """Two solves of the same session give the same solution bit for bit."""
    import numpy as np

    from nektar_amd import contfield as cf

    l2a, _, _, xa = cf.helmholtz_solve(HEX, tol=1e-14)
    l2b, _, _, xb = cf.helmholtz_solve(HEX, tol=1e-14)
    assert l2a == l2b and np.array_equal(xa, xb)
