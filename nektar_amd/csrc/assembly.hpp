// General C0 assembly and the iterative Helmholtz solve over any conforming
// mesh: the GPU counterpart of AssemblyMapCG (GlobalToLocal / Assemble,
// AssemblyMapCG.cpp:2804-2897) and of GlobalLinSysIterativeFull with
// NekLinSysIterCG and the Diagonal preconditioner
// (GlobalLinSysIterativeFull.cpp:96-383, NekLinSysIterCG.cpp:118-372 or
// NekLinSysIterCGLoc.cpp:114-308, PreconditionerDiagonal.cpp:93-189), for collections of hex and tet
// Helmholtz operators that share one local-to-global map.
//
// The caller supplies the map as AssemblyMapCG holds it: for every entry i
// of the concatenated element-local coefficient vector, its global DOF
// map[i] and sign[i] (+-1, the orientation sign of edge and face modes; all
// +1 when NULL), with the Dirichlet DOFs numbered first (global ids
// 0..ndir-1, m_numGlobalDirBndCoeffs).  Dirichlet data are zero unless the
// caller gives their values (IterativeSolver::set_dirichlet_local).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
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
    // the same for the local entries [off, off + n)
    void global_to_local_range(int64_t off, int64_t n, const double *glob, double *loc, hipStream_t s) const;
    // the signed map on the device (entry m >= 0: global m, sign +1; m < 0: ~m, sign -1)
    const int *signed_map() const { return map_.as<int>(); }
    // glob[g] = sum_{map[i] = g} sign[i] loc[i], in increasing i; Dirichlet
    // rows zeroed when zero_dir; partial sums of glob.wdot into partials
    // (reduction_blocks(nglobal) blocks) when wdot is given; sq: sum of
    // sign^2 loc (for diagonals).
    void assemble(const double *loc, double *glob, hipStream_t s, bool zero_dir = false, const double *wdot = nullptr,
                  double *partials = nullptr, bool sq = false) const;

    // the element-local CG's preconditioner step (NekLinSysIterCGLoc with
    // PreconditionerDiagonal in local form): wg = inv_diag Assemble(r) off
    // the Dirichlet rows, with the partial sums of its dots
    // (solver::launch_cgloc_precon); scatter wg back with global_to_local
    void precon_local(bool init, const double *r, const double *inv_diag, double *wg, double *partials,
                      hipStream_t s) const;
    // glob[g] = sign loc of g's first local entry (loc continuous)
    void global_values(const double *loc, double *glob, hipStream_t s) const;
    // the map and signs back on the host (synchronous copies; signs empty
    // when the map has none)
    std::vector<int> host_map() const;
    std::vector<double> host_sign() const;

    int64_t nloc() const { return nloc_; }
    int nglobal() const { return nglobal_; }
    int ndir() const { return ndir_; }

private:
    int64_t nloc_;
    int nglobal_, ndir_;
    bool has_sign_;
    // signed indices (entry m >= 0: index m, sign +1; m < 0: index ~m, sign -1):
    // map_ local -> global, col_ the CSR columns of the transpose
    DeviceBuffer map_, row_, col_;
};

// The reference's solver choices for this system: GlobalSysSoln
// IterativeFull with LinSysIterSolver NekLinSysIterCG on the assembled
// vectors or NekLinSysIterCGLoc on the element-local ones
// (GlobalLinSysIterative.cpp:86-88), or GlobalSysSoln IterativeStaticCond
// (CG on the assembled Schur complement of the interior modes).
enum class CGSpace { Global, Local, StaticCond };

struct StaticCondData;  // static_cond.cpp

class IterativeSolver {
public:
    // ops: Helmholtz operators (FP64) whose element-local vectors,
    // concatenated in this order, are the map's local vector.  The operators
    // must outlive the solver.  The diagonal (and, for StaticCond, the
    // interior inverses) is built here from the operators' geometry and
    // lambda; a later change to an operator (its version counter) makes the
    // next call rebuild them, as the reference builds a new GlobalLinSys for
    // a new lambda key.
    IterativeSolver(const AssemblyMap &map, const std::vector<MFOperator *> &ops, hipStream_t stream,
                    CGSpace space = CGSpace::Global);
    ~IterativeSolver();
    IterativeSolver(const IterativeSolver &) = delete;
    IterativeSolver &operator=(const IterativeSolver &) = delete;

    // the right-hand side in the local layout (bloc: device, e.g.
    // IProductWRTBase of the forcing), ContField::v_HelmSolve's; the global
    // CG solves with b = Assemble(bloc), the local CG starts from r = bloc
    void set_rhs_local(const double *bloc);
    // Dirichlet values in the local layout (device; zero off the Dirichlet
    // DOFs; NULL: homogeneous), as ContField imposes them on the local
    // coefficients before the solve (ContField.cpp:557-608).  The solve then
    // lifts them as GlobalLinSysIterativeFull::v_Solve does (:118-169):
    // rhs = bloc - A dloc element by element, CG for the perturbation with
    // zero Dirichlet rows, solution = perturbation + dloc.
    void set_dirichlet_local(const double *dloc);
    // The Robin boundary mass of Helmholtz solves with Robin conditions
    // du/dn + alpha u = g: what GlobalLinSysIterativeFull adds to every
    // element operator apply (AddRobinTraceContribution,
    // GlobalLinSysIterativeFull.cpp:134-150 in the Dirichlet lifting and
    // :219-233 in the matrix multiply), the face mass matrices
    // int_face alpha phi_i phi_j of the elements with Robin faces, given as
    // rows of the local vector: for k < nrows, y[rows[k]] +=
    // sum_{m in [rowptr[k], rowptr[k+1])} vals[m] x[cols[m]] (host arrays,
    // local indices, rows distinct).  nrows = 0 removes it.  The Jacobi
    // diagonal takes the entries with cols[m] = rows[k].  Not available with
    // static condensation.
    void set_robin_local(int64_t nrows, const int64_t *rows, const int64_t *rowptr, const int64_t *cols,
                         const double *vals);
    // rebuild what depends on the operators if any of them changed since
    // the last build (called by every public entry that uses them)
    void refresh();
    SolveInfo solve(double tol, int maxit);
    // y = A x on global device vectors, Dirichlet rows zeroed
    void apply_global(const double *x, double *y);
    // the solution as global values (nglobal) / element-local (nloc), on
    // the device, enqueued on the solver's stream
    const double *solution();
    void solution_local(double *xloc);
    // the preconditioner's inverse diagonal in the caller's numbering: the
    // Jacobi one, or (StaticCond) the inverse Schur diagonal on the
    // boundary DOFs, 0 on interior DOFs
    const double *inv_diag() const;
    int nglobal() const { return map_.nglobal(); }
    CGSpace space() const { return space_; }
    // the static-condensation diagonal (PreconditionerDiagonal.cpp:132-154
    // before its inversion; Preconditioner.cpp:185-236): each element's
    // Schur complement A_bb - A_bi A_ii^-1 A_ib of its interior modes,
    // diagonal assembled over the boundary DOFs into dglob (device,
    // nglobal; Dirichlet rows and interior DOFs 0).  Needs interior modes
    // that no other local entry shares and distinct global DOFs for the
    // boundary modes of one element.
    void sc_diagonal(double *dglob);
    // StaticCond: CG iterations of the last solve ran on this many boundary
    // unknowns (global boundary DOFs, Dirichlet included)
    int nboundary() const;

private:
    void apply_local(const double *xloc, double *yloc);
    // yloc = A GlobalToLocal(xg): hex collections gather inside the operator
    // kernel, the others through the signed gather into xloc_
    void apply_from_global(const double *xg, double *yloc);
    void build_diagonal();
    void setup_static_cond();
    SolveInfo solve_static_cond(double tol, int maxit);

    const AssemblyMap &map_;
    std::vector<MFOperator *> ops_;
    std::vector<int64_t> off_;
    std::vector<uint64_t> versions_;  // the operators' versions the diagonal was built from
    hipStream_t stream_;
    CGSpace space_;
    int nb_ = 1;
    DeviceBuffer xloc_, yloc_, b_, invd_, state_, pa_, pmu_, xg_;
    DeviceBuffer cg_[6];  // x r p q w s: global (Global) or local (Local) vectors
    solver::CGState *host_state_ = nullptr;
    bool have_rhs_ = false;
    std::unique_ptr<StaticCondData> sc_;
    // the caller's local right-hand side and Dirichlet values; the
    // right-hand side the CG sees is built from them at every solve
    DeviceBuffer bloc_, dloc_, dglob_;
    bool have_dir_ = false;
    // Robin rows (set_robin_local): row index, row start, column, value
    DeviceBuffer rrow_, rptr_, rcol_, rval_;
    int64_t nrobin_ = 0;
    void add_robin(const double *x, bool global, double *yloc);
    void load_rhs();
};

}  // namespace nektar_amd
