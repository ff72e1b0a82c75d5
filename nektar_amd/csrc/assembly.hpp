// General C0 assembly and the iterative Helmholtz solve over any conforming
// mesh: the GPU counterpart of AssemblyMapCG (GlobalToLocal / Assemble,
// AssemblyMapCG.cpp:2804-2897) and of GlobalLinSysIterativeFull with
// NekLinSysIterCG and the Diagonal preconditioner
// (GlobalLinSysIterativeFull.cpp:96-383, NekLinSysIterCG.cpp:118-372,
// PreconditionerDiagonal.cpp:93-189), for collections of hex and tet
// Helmholtz operators that share one local-to-global map.
//
// The caller supplies the map as AssemblyMapCG holds it: for every entry i
// of the concatenated element-local coefficient vector, its global DOF
// map[i] and sign[i] (+-1, the orientation sign of edge and face modes; all
// +1 when NULL), with the Dirichlet DOFs numbered first (global ids
// 0..ndir-1, m_numGlobalDirBndCoeffs).  Homogeneous Dirichlet data.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "mf_operator.hpp"
#include "solver.hpp"

namespace nektar_amd {

class AssemblyMap {
public:
    AssemblyMap(const int *map, const double *sign, int64_t nloc, int nglobal, int ndir);

    // loc[i] = sign[i] glob[map[i]]
    void global_to_local(const double *glob, double *loc, hipStream_t s) const;
    // glob[g] = sum_{map[i] = g} sign[i] loc[i], in increasing i; Dirichlet
    // rows zeroed when zero_dir; partial sums of glob.wdot into partials
    // (reduction_blocks(nglobal) blocks) when wdot is given; sq: sum of
    // sign^2 loc (for diagonals).
    void assemble(const double *loc, double *glob, hipStream_t s, bool zero_dir = false, const double *wdot = nullptr,
                  double *partials = nullptr, bool sq = false) const;

    int64_t nloc() const { return nloc_; }
    int nglobal() const { return nglobal_; }
    int ndir() const { return ndir_; }

private:
    int64_t nloc_;
    int nglobal_, ndir_;
    bool has_sign_;
    DeviceBuffer map_, sign_, row_, col_, csign_;
};

class IterativeSolver {
public:
    // ops: Helmholtz operators (FP64) whose element-local vectors,
    // concatenated in this order, are the map's local vector.  The operators
    // must outlive the solver and keep their geometry and lambda while it
    // is used (the diagonal is built here).
    IterativeSolver(const AssemblyMap &map, const std::vector<MFOperator *> &ops, hipStream_t stream);
    ~IterativeSolver();
    IterativeSolver(const IterativeSolver &) = delete;
    IterativeSolver &operator=(const IterativeSolver &) = delete;

    // b = Assemble(bloc) (bloc: device, local layout, e.g. IProductWRTBase
    // of the forcing), ContField::v_HelmSolve's right-hand side
    void set_rhs_local(const double *bloc);
    SolveInfo solve(double tol, int maxit);
    // y = A x on global device vectors, Dirichlet rows zeroed
    void apply_global(const double *x, double *y);
    const double *solution() const { return cg_[0].as<double>(); }
    const double *inv_diag() const { return invd_.as<double>(); }
    int nglobal() const { return map_.nglobal(); }

private:
    void apply_local(const double *xloc, double *yloc);
    void build_diagonal();

    const AssemblyMap &map_;
    std::vector<MFOperator *> ops_;
    std::vector<int64_t> off_;
    hipStream_t stream_;
    int nb_ = 1;
    DeviceBuffer xloc_, yloc_, b_, invd_, state_, pa_, pmu_;
    DeviceBuffer cg_[6];  // x r p q w s
    solver::CGState *host_state_ = nullptr;
    bool have_rhs_ = false;
};

}  // namespace nektar_amd
