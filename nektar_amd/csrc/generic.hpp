// Runtime-size operators (k_generic.hip): any modes and points per
// direction.  Reference: the runtime-size operator3D that the compile-time
// switch falls back to (MatrixFreeOps/SwitchNodesPoints.h:302-343,
// Helmholtz.h:374-464).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "basis.hpp"
#include "launch.hpp"

namespace nektar_amd {

struct GenArgs {
    OpType op;
    Shape shape;
    bool fp32, deformed;
    int nelmt;
    int nm[3], nq[3], nqtot, ncoeff;
    int npq;   // tets: (p, q) pairs, sum_p (nm1 - p)
    int buf;   // values per intermediate buffer (four per element)
    // offsets (in values) of the tables in `tab`: B_d [mode*nq_d + i],
    // D_d [l*nq_d + i] (= l_l'(z_i)), W_d (Radau pre-scaled), and the tet
    // half-space factors h0 (nq0), h1, h2 (nq1), h3 (nq2)
    int oB[3], oD[3], oW[3], oH[4];
    const void *tab;
    const int *itab;  // tets: index tables (see GenLayout)
    const void *geo, *adv;
    double lambda;
    const void *in0, *in1, *in2;
    void *out0, *out1, *out2;
    void *ws;    // global workspace (groups * 4 * buf values) when the buffers exceed LDS
    int groups;  // workgroups of the global-workspace form
    hipStream_t stream;
};

constexpr size_t kGenMaxLds = 65536;
constexpr int kGenGroups = 1024;
namespace gen {
constexpr size_t kMaxLds = kGenMaxLds;
}
inline size_t lds_bytes(const GenArgs &a) { return (size_t)4 * a.buf * (a.fp32 ? sizeof(float) : sizeof(double)); }

// Host-side description of one (shape, modes, points) kind: FP64 tables,
// tet index tables and the sizes the kernels need.
struct GenLayout {
    std::vector<double> tab;
    std::vector<int> itab;  // tets, prisms: cP | cPair | cM2 | pC0 | pM20 | pNR | pP
    int oB[3], oD[3], oW[3], oH[4];
    int npq = 0, buf = 0, nqtot = 0, ncoeff = 0;
};

// Coefficients of the expansion: hex nm0*nm1*nm2; tet with nm0 <= nm1 <= nm2:
// p < nm0, q < nm1 - p, r < nm2 - p - q (BwdTransKernels.hpp:244-266);
// prism with nm0 <= nm2: p < nm0, q < nm1, r < nm2 - p
// (StdPrismExp.cpp:1947-1956, BwdTransKernels.hpp:345-360).
int generic_ncoeffs(Shape s, const int nm[3]);
// Throws NMF_ERR_UNSUPPORTED for shapes the runtime-size kernels do not take.
void generic_check(Shape s, const int nm[3], const int nq[3]);
GenLayout generic_layout(Shape s, const int nm[3], const int nq[3]);
hipError_t launch_generic(const GenArgs &a);

}  // namespace nektar_amd
