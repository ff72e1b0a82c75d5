// Quadrature points, weights, differentiation matrices and the MODIFIED
// (hierarchical) bases Modified_A/B/C on the host, in FP64.
//
// Conventions (the drop-in contract of MatrixFreeOps' Helper tables,
// library/MatrixFreeOps/Operator.hpp:529-641):
//   bdata[m*nq + i]  = phi_m(z_i)
//   dbdata[m*nq + i] = phi_m'(z_i)
//   D[k*nq + i]      = l_k'(z_i)  (column-major derivative matrix)
//   w[i]             = quadrature weight; Radau(1,0) x0.5, Radau(2,0) x0.25
//                      (Operator.hpp:600-616) when `prescaled` is requested.
#pragma once

#include <vector>

namespace nektar_amd {

enum class PointsType { GLL = 0, RadauM_A1B0 = 1, RadauM_A2B0 = 2 };
enum class BasisType { ModifiedA = 0, ModifiedB = 1, ModifiedC = 2, ModifiedPyrC = 3 };
enum class Shape { Hex = 0, Tet = 1, Prism = 2, Pyr = 3 };

struct Points1D {
    int nq = 0;
    PointsType type = PointsType::GLL;
    std::vector<double> z, w, D;
};

struct Basis1D {
    BasisType type = BasisType::ModifiedA;
    int nm = 0;      // number of modes per direction (P+1)
    int nmodes = 0;  // entries in the table (nm, nm(nm+1)/2, ...)
    Points1D pts;
    std::vector<double> bdata, dbdata;
};

// Jacobi polynomial P_n^{a,b} at x by the three-term recurrence.
double jacobi(int n, double a, double b, double x);
// Its derivative, dP_n^{a,b}/dx = (n+a+b+1)/2 P_{n-1}^{a+1,b+1}.
double jacobi_deriv(int n, double a, double b, double x);

Points1D make_points(PointsType t, int nq);
Basis1D make_basis(BasisType t, int nm, int nq);

int num_coeffs(Shape s, int nm);
// MODIFIED expansion defaults (MeshGraph.cpp:1688-1752).
void default_points(Shape s, int nq0, int nq[3]);
BasisType default_basis(Shape s, int dir);
PointsType default_points_type(Shape s, int dir);
// Radau weight pre-scaling used by the matrix-free operators.
double weight_prescale(PointsType t);

// Interpolation from nin points of family t to nout points of the same
// family: I[p*nout + i] = h_p(zo_i), the Lagrange polynomial of input point
// p at output point i (GaussPoints::CalculateInterpMatrix,
// GaussPoints.cpp:283-380, via Polylib Imglj / Imgrjm).
std::vector<double> interp_matrix(PointsType t, int nin, int nout);
// Points per direction of PhysInterp1DScaled (Operator.hpp:568-575).
void interp_points(Shape s, int nq0, double factor, int out[3]);

}  // namespace nektar_amd
