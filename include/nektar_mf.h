/*
 * nektar_mf.h -- C ABI of the MI355X matrix-free spectral/hp operator
 * library (libnektar_amd.so).
 *
 * The surface mirrors the two plugin interfaces the reference path sits
 * behind (paths relative to the Nektar++ 5.10.0 tree):
 *   - MatrixFree::Operator and its string-keyed factory
 *     (library/MatrixFreeOps/Operator.hpp:59-366, Operator.cpp:40-90);
 *   - Collections::Helmholtz_MatrixFree / *_MatrixFree wrappers
 *     (library/Collections/Helmholtz.cpp:586-783, BwdTrans.cpp:160-244,
 *      IProductWRTBase.cpp:182-269, PhysDeriv.cpp:272-402),
 * plus the iterative C0 Helmholtz solve they serve
 * (MultiRegions/GlobalLinSysIterativeFull.cpp:96-383 driving
 *  LibUtilities/LinearAlgebra/NekLinSysIterCG.cpp:118-372 with
 *  MultiRegions/PreconditionerDiagonal.cpp:69-189).
 *
 * Conventions
 *   - Every function returns NMF_OK (0) or a non-zero NMF_ERR_* code;
 *     nmf_last_error() returns a thread-local description of the last
 *     failure (the C++ wrappers turn it into an exception, as NEKERROR does).
 *   - Array arguments of *_apply are DEVICE pointers of the operator's
 *     precision, element-major and unpadded exactly as the Collections
 *     interface passes them: coefficients [e][ncoeff], physical values
 *     [e][nqtot] with direction 0 fastest.  *_host variants take host FP64
 *     arrays and synchronise.
 *   - `stream` is a hipStream_t (NULL = default stream).  A handle is not
 *     thread-safe; concurrent use needs external synchronisation (as the
 *     reference operators, which hold mutable lambda state).
 *   - Geometry is given per element, not interleaved:
 *       regular  jac[e],           df[e*9 + d]
 *       deformed jac[e*nqtot + q], df[(e*nqtot + q)*9 + d]
 *     with df[j*3 + k] = d xi_k / d x_j (GeomFactors.cpp:401-477).
 */
#ifndef NEKTAR_MF_H
#define NEKTAR_MF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NMF_OK 0
#define NMF_ERR_INVALID 1     /* bad argument                        */
#define NMF_ERR_UNSUPPORTED 2 /* shape / order / key not provided    */
#define NMF_ERR_STATE 3       /* e.g. geometry not set before apply  */
#define NMF_ERR_DEVICE 4      /* HIP runtime or kernel launch error  */
#define NMF_ERR_MEMORY 5      /* device allocation failed / budget   */

#define NMF_FP64 0
#define NMF_FP32 1

#define NMF_SHAPE_HEX 0
#define NMF_SHAPE_TET 1
#define NMF_SHAPE_PRISM 2
#define NMF_SHAPE_PYR 3

typedef struct nmf_operator_s *nmf_operator;
typedef struct nmf_solver_s *nmf_solver;
typedef struct nmf_batch_s *nmf_batch;
typedef struct nmf_server_s *nmf_server;
typedef struct nmf_assembly_s *nmf_assembly;
typedef struct nmf_linsys_s *nmf_linsys;

/* Last error message of the calling thread ("" if none). */
const char *nmf_last_error(void);
/* Library version as major*10000 + minor*100 + patch. */
int nmf_version(void);

/* ---------------------------------------------------------------------
 * Device memory helpers (for callers without their own allocator).
 * ------------------------------------------------------------------- */
int nmf_device_malloc(void **ptr, size_t bytes);
int nmf_device_free(void *ptr);
int nmf_memcpy_h2d(void *dst, const void *src, size_t bytes);
int nmf_memcpy_d2h(void *dst, const void *src, size_t bytes);
int nmf_device_synchronize(void);

/* ---------------------------------------------------------------------
 * Quadrature / basis tables of the MODIFIED expansion (host, FP64).
 * Replaces LibUtilities::Basis::GetBdata/GetDbdata/GetZ/GetW and
 * Points::GetD (Foundations/Basis.cpp:276-645, GaussPoints.cpp:117-281).
 * shape NMF_SHAPE_*, dir 0..2, numpoints0 = points in direction 0.
 * bdata/dbdata sizes nmodes_dir*nq_dir, D nq_dir^2 (D[k*nq+i]=l_k'(z_i)),
 * z, w nq_dir (w without the Radau pre-scaling).  Any pointer may be NULL.
 * ------------------------------------------------------------------- */
int nmf_basis_tables(int shape, int dir, int nummodes, int numpoints0,
                     int *nq_dir, int *nmodes_dir, double *bdata,
                     double *dbdata, double *z, double *w, double *D);

/* ---------------------------------------------------------------------
 * MatrixFree::Operator
 * ------------------------------------------------------------------- */
/* MatrixFree::GetOperatorFactory().CreateInstance(key, basis, nElmt)
 * (Operator.cpp:40-90).  key = "<Op>_<Shape>_<Regular|Deformed>" as built
 * by GetOpstring, Op in {BwdTrans, IProduct, IProductWRTDerivBase,
 * PhysDeriv, Helmholtz, LinearAdvectionDiffusionReaction,
 * PhysInterp1DScaled},
 * Shape in {Hex, Tet, Prism, Pyr}.  nummodes = P+1 per direction,
 * numpoints0 = Q in direction 0 (tets use Q-1 Gauss-Radau points in
 * directions 1, 2, prisms and pyramids in direction 2).  Prisms
 * (Modified_A, Modified_A, Modified_B; BwdTransKernels.hpp:334-405,
 * IProductKernels.hpp:519-655, HelmholtzKernels.hpp:918-1119) and pyramids
 * (Modified_A, Modified_A, ModifiedPyr_C on Gauss-Radau(2,0) points, equal
 * modes in the three directions; BwdTransKernels.hpp:407-489,
 * IProductKernels.hpp:660-815, HelmholtzKernels.hpp:1119-1326) always run
 * the runtime-size kernels; PhysInterp1DScaled takes hexes and tets only. */
int nmf_operator_create(const char *key, int nummodes, int numpoints0,
                        int nelmt, int precision, nmf_operator *out);
/* The same with modes and points per direction, as the reference's
 * factory takes one BasisKey per direction (Operator.hpp:59-63,
 * std::vector<LibUtilities::BasisSharedPtr>): nummodes[3], numpoints[3].
 * Bases are the MODIFIED ones (hex Modified_A on GLL points; tet
 * Modified_A/B/C on GLL, Gauss-Radau(1,0), Gauss-Radau(2,0) points).
 * Equal modes with the default points of an instantiated (modes, points)
 * pair run the compile-time kernels; every other combination -- any
 * anisotropic order, point counts of 2..64 and modes of 2..32 per
 * direction (tets: nm0 <= nm1 <= nm2) -- runs the runtime-size kernels, as
 * the reference falls back to its runtime-size operators
 * (SwitchNodesPoints.h:302-343).  Anything else: NMF_ERR_UNSUPPORTED.
 * PhysInterp1DScaled takes the default points only. */
int nmf_operator_create_basis(const char *key, const int *nummodes,
                              const int *numpoints, int nelmt, int precision,
                              nmf_operator *out);
/* *generic = 1 when op runs the runtime-size kernels, 0 for a compile-time
 * kernel.  No counterpart in the reference (whose choice is internal to
 * SwitchNodesPoints.h:220-348); for callers that report which path ran. */
int nmf_operator_runtime_size(nmf_operator op, int *generic);
int nmf_operator_destroy(nmf_operator op);
/* GetInputSize/GetOutputSize per element (Collections/Operator.h:178-230) */
int nmf_operator_info(nmf_operator op, int *ncoeff, int *nqtot, int *nelmt,
                      int *deformed, int *precision);
/* Operator::SetJac (Operator.hpp:104-105) -- host FP64, see layout above */
int nmf_operator_set_jac(nmf_operator op, const double *jac);
/* Operator::SetDF (Operator.hpp:99-101) */
int nmf_operator_set_df(nmf_operator op, const double *df);
/* Helmholtz::SetLambda (Operator.hpp:298-301) */
int nmf_operator_set_lambda(nmf_operator op, double lambda);
/* Helmholtz::SetConstVarDiffusion (Operator.hpp:303-313): d6 = {d00, d01,
 * d11, d02, d12, d22}, symmetric positive definite; NULL resets to I. */
int nmf_operator_set_const_var_diffusion(nmf_operator op, const double *d6);
/* LinearAdvectionDiffusionReaction::SetAdvectionVelocities
 * (Operator.hpp:476-481): vel[(e*nqtot + q)*3 + j], velocity component j
 * at point q of element e (the de-interleaved form of the reference's
 * [block][q][j][lane] array, LinearAdvectionDiffusionReaction.cpp:690-714).
 * The operator computes lambda IProduct(u + V . grad u) + L u, as the
 * MatrixFree operator does; the Collections wrapper passes -lambda and
 * V / (-lambda) (LinearAdvectionDiffusionReaction.cpp:578-589, 660-669).
 * SetLambda and SetConstVarDiffusion apply to this operator as well. */
int nmf_operator_set_advection(nmf_operator op, const double *vel);
/* PhysInterp1DScaled::SetScalingFactor + SetUpInterp1D (Operator.hpp
 * :151-155, 561-589; default 1.5): point values at the quadrature points
 * -> values at (int)(nq*factor) points of the same family per direction
 * ((int)(nq0*factor - 1) in the tet's collapsed directions 1 and 2).
 * Output sizes beyond one workgroup's 64 KiB of staging are rejected with
 * NMF_ERR_UNSUPPORTED. */
int nmf_operator_set_scaling_factor(nmf_operator op, double factor);
/* Collections::Operator::GetInputSize / GetOutputSize (Operator.h:178-230):
 * values in the whole input and output arrays of operator(). */
int nmf_operator_io_size(nmf_operator op, int64_t *in_size, int64_t *out_size);
/* operator()(input, output) for BwdTrans / IProduct / Helmholtz
 * (Operator.hpp:119-121, 181-183, 294-296) */
int nmf_operator_apply(nmf_operator op, const void *in, void *out,
                       void *stream);
/* PhysDeriv operator()(input, output[3]) (Operator.hpp:238-240) */
int nmf_operator_apply_physderiv(nmf_operator op, const void *in, void *out0,
                                 void *out1, void *out2, void *stream);
/* Collections::Operator::GetWspSize (Collections/Operator.h:178-230):
 * workspace values (of the operator's precision) that the single-direction
 * PhysDeriv form needs, 2*nelmt*nqtot; 0 for the other operators. */
int nmf_operator_wsp_size(nmf_operator op, int64_t *wsp_size);
/* PhysDeriv operator()(dir, input, output, wsp) (Collections/PhysDeriv.cpp
 * :312-317): the derivative d/dx_dir (dir 0..2) into out; the other two
 * directions are computed into wsp (device, nmf_operator_wsp_size values),
 * as the reference computes all three and copies one.  wsp NULL: a
 * stream-ordered workspace is allocated and freed for this call (the
 * reference's Collection::ApplyOperator allocates one per call,
 * Collection.h:148-155). */
int nmf_operator_apply_physderiv_dir(nmf_operator op, int dir, const void *in,
                                     void *out, void *wsp, void *stream);
/* The single-direction PhysDeriv with no device allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  The two directions not asked for go to a 64 MiB static
 * device array of the library (one per GPU, held until the process ends)
 * instead of a workspace, so 2 * nelmt * nqtot values of the operator's
 * precision must fit 64 MiB (NMF_ERR_UNSUPPORTED otherwise), and the
 * operator's geometry must already be on the device: after SetDF, one
 * ordinary apply uploads it (NMF_ERR_STATE until then).  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (different
 * streams or threads) race; nothing in it needs clearing between calls.
 * Results are those of nmf_operator_apply_physderiv_dir, bit for bit. */
int nmf_operator_apply_physderiv_dir_noalloc(nmf_operator op, int dir,
                                             const void *in, void *out,
                                             void *stream);
/* IProductWRTDerivBase operator()(input[3], output) (Operator.hpp:192-217,
 * IProductWRTDerivBase.h:592-825): out = sum_j IProduct w.r.t. dphi/dx_j of
 * in_j; inputs [e][nqtot] at the quadrature points, output [e][ncoeff].
 * Key "IProductWRTDerivBase_<Hex|Tet>_<Regular|Deformed>"; needs SetJac and
 * SetDF. */
int nmf_operator_apply_iproduct_deriv(nmf_operator op, const void *in0,
                                      const void *in1, const void *in2,
                                      void *out, void *stream);
int nmf_operator_apply_iproduct_deriv_host(nmf_operator op, const double *in0,
                                           const double *in1,
                                           const double *in2, double *out);
/* Host FP64 convenience forms (copy in, apply, copy out, synchronise). */
int nmf_operator_apply_host(nmf_operator op, const double *in, double *out);
int nmf_operator_apply_physderiv_host(nmf_operator op, const double *in,
                                      double *out0, double *out1,
                                      double *out2);

/* ---------------------------------------------------------------------
 * Workgroup tiles and the autotuner.  No counterpart in the reference,
 * whose kernels have one shape per vector width (VecDataType in
 * MatrixFreeOps/Operator.hpp:56-57); here the hex Helmholtz kernels are
 * instantiated for several tiles (elements per workgroup) and every tet
 * kernel for two (work-items per element: nq0*(nq0-1), elements sharing
 * waves, or that rounded up to a multiple of 64, one element per wave at
 * nq0 <= 8), and the fastest is measured per (key, nummodes, numpoints,
 * precision) on the caller's data.  The choice is remembered for the rest of the process:
 * operators and solvers created later for the same kind start from it.
 * tile 0 = the library default.  Results do not depend on the tile,
 * except that hex Helmholtz tiles NMF_TILE_MFMA + E (E = 1, 2, 4 elements
 * per workgroup, NUMMODES >= 5) run the per-direction products on the
 * matrix cores (MFMA), which sum in fours: their results agree with the
 * vector tiles to rounding and are bit-identical among themselves.
 * ------------------------------------------------------------------- */
#define NMF_TILE_MFMA 1024
/* Tiles available to op (count 0 when its kernel has a single shape). */
int nmf_operator_tile_candidates(nmf_operator op, int *tiles, int max,
                                 int *count);
int nmf_operator_set_tile(nmf_operator op, int tile);
int nmf_operator_get_tile(nmf_operator op, int *tile);
/* Time every candidate on op's geometry (allocates scratch arrays of the
 * operator's size), keep the fastest; times_us[i] (may be NULL, else
 * `max` entries) receives the mean launch time of candidate i. */
int nmf_operator_autotune(nmf_operator op, void *stream, int *tile,
                          double *times_us, int max);

/* ---------------------------------------------------------------------
 * Batches of operators of uneven kinds and sizes.  Replaces the loop over
 * collections in ExpList::GeneralMatrixOp (MultiRegions/ExpList.cpp:
 * 2666-2757: one Collection::ApplyOperator per shape/order collection)
 * with one submission: Helmholtz entries of one kind (shape, nummodes,
 * numpoints, precision, geometry type, tile) run as a single grouped
 * launch, kinds are spread over `streams` streams, and the whole batch is
 * captured once into a hipGraph that every later run replays
 * (NMF_BATCH_GRAPH, the default).  NMF_BATCH_GROUPED launches the same
 * grouped kernels without a graph; NMF_BATCH_PER_ITEM applies entry by
 * entry on the given stream (the reference's loop).  Results are identical
 * in every mode.  Operators must outlive the batch; in/out arrays are
 * fixed when added.  A change to an operator (geometry, lambda, diffusion,
 * tile) is picked up by the next run, which then rebuilds the item tables
 * and graph (synchronising the device).
 * ------------------------------------------------------------------- */
#define NMF_BATCH_GRAPH 0
#define NMF_BATCH_GROUPED 1
#define NMF_BATCH_PER_ITEM 2
int nmf_batch_create(nmf_batch *out);
int nmf_batch_destroy(nmf_batch b);
/* operator()(in, out) of op, as nmf_operator_apply */
int nmf_batch_add(nmf_batch b, nmf_operator op, const void *in, void *out);
int nmf_batch_add_physderiv(nmf_batch b, nmf_operator op, const void *in,
                            void *out0, void *out1, void *out2);
int nmf_batch_set_mode(nmf_batch b, int mode);
/* streams 1..8 (default 4) */
int nmf_batch_set_streams(nmf_batch b, int streams);
/* Autotune each Helmholtz kind (hex or tet) on its largest member; every
 * member of the kind takes the chosen tile. */
int nmf_batch_autotune(nmf_batch b, void *stream);
int nmf_batch_run(nmf_batch b, void *stream);
/* entries, grouped launches, and total element-local input DOFs */
int nmf_batch_info(nmf_batch b, int *items, int *groups, int64_t *dofs);

/* ---------------------------------------------------------------------
 * C0 Helmholtz solve on a structured hex mesh of [0,1]^3
 * (ContField::HelmSolve -> GlobalLinSysIterativeFull -> NekLinSysIterCG with
 *  the Diagonal preconditioner; ContField.cpp:797-925,
 *  GlobalLinSysIterativeFull.cpp:96-383, NekLinSysIterCG.cpp:118-372).
 * Solves -lap(u) + lambda u = f with u = 0 on the boundary.  Global
 * vectors hold all N^3 C0 modes (N = n*(nummodes-1)+1) in the order
 * ((gz*N + gy)*N + gx); Dirichlet rows are zero in the solution.
 * ------------------------------------------------------------------- */
typedef struct {
    int iterations;       /* CG iterations (total inner ones when mixed)  */
    int outer_iterations; /* refinement steps (1 for FP64)                */
    int converged;
    double final_error;   /* sqrt(||r||^2/||b||^2), FP64 residual          */
    double rhs_magnitude; /* ||b||^2 over all global rows                  */
    double solve_ms;      /* device time of the solve                     */
} nmf_solve_info;

/* Structured n^3 hex mesh; interior vertices perturbed by U[-perturb*h,
 * perturb*h] per coordinate from SplitMix64(seed) (perturb = 0: regular,
 * else deformed geometry).  stream: hipStream_t to run on (NULL: own). */
int nmf_solver_create_hex_box(int n, int nummodes, int numpoints0,
                              double perturb, uint64_t seed, double lambda,
                              void *stream, nmf_solver *out);
int nmf_solver_destroy(nmf_solver s);
int nmf_solver_sizes(nmf_solver s, int64_t *nglobal, int64_t *nlocal,
                     int64_t *ndirichlet, int64_t *device_bytes);
/* Device memory a solver of this shape needs (admission control). */
int64_t nmf_solver_estimate_bytes(int n, int nummodes, int numpoints0,
                                  int deformed, int mixed);
/* Forcing f = (3 k^2 pi^2 + lambda) sin(k pi x) sin(k pi y) sin(k pi z)
 * (manufactured solution sin sin sin). */
int nmf_solver_set_forcing_manufactured(nmf_solver s, int k);
/* Forcing at the quadrature points (device FP64, element-major, element
 * e = ex + n*(ey + n*ez)). */
int nmf_solver_set_forcing_phys(nmf_solver s, const double *f_dev);
/* mixed = 0: FP64 CG; mixed = 1: FP32-inner iterative-refinement CG with
 * FP64 residual, dots and diagonal.  Stops at ||r|| < tol ||b||. */
int nmf_solver_solve(nmf_solver s, int mixed, double tol, int maxit,
                     nmf_solve_info *info);
/* The same solve without blocking, for callers that interleave many solves
 * in their own loop: begin enqueues the set-up; each advance enqueues the
 * next 8 iterations (or a reliable update) once the previous chunk's state
 * has reached the host, and sets *done when the solve has finished
 * (wait != 0: block for the chunk in flight instead of returning at once);
 * finish returns the result.  While a solve is in progress the solver's
 * other calls (forcing, solve, apply_global, L2 error, autotune) return
 * NMF_ERR_STATE. */
int nmf_solver_begin(nmf_solver s, int mixed, double tol, int maxit);
int nmf_solver_advance(nmf_solver s, int wait, int *done);
int nmf_solver_finish(nmf_solver s, nmf_solve_info *info);
/* L2 norm of u_h - sin(k pi x) sin(k pi y) sin(k pi z). */
int nmf_solver_l2_error_manufactured(nmf_solver s, int k, double *l2);
/* Host copies of the global solution / assembled rhs / inverse diagonal. */
int nmf_solver_get_solution(nmf_solver s, double *x_host);
int nmf_solver_get_rhs(nmf_solver s, double *b_host);
int nmf_solver_get_inv_diag(nmf_solver s, double *d_host);
/* y = A x for global device vectors (FP64), Dirichlet rows set to 0. */
int nmf_solver_apply_global(nmf_solver s, const double *x_dev, double *y_dev);
/* Autotune the solver's gather-apply kernel (FP64 and FP32 iterations);
 * iteration graphs are re-captured on the next solve. */
int nmf_solver_autotune(nmf_solver s, int *tile64, int *tile32);
/* Forget every tile the autotuner recorded in this process. */
int nmf_clear_tuning(void);
/* ---------------------------------------------------------------------
 * Serving loop for Helmholtz solves.  No counterpart in the reference
 * (which solves one system per process, ContField::HelmSolve); it admits
 * many solves of different meshes and orders as they arrive.  Requests
 * are structured hex boxes as nmf_solver_create_hex_box (perturb 0) with
 * the manufactured forcing of frequency k.  One host loop queues
 * arrivals (FIFO), admits while fewer than max_active solves run and the
 * request's device memory fits the HBM budget (solvers -- geometry,
 * diagonal, Krylov vectors, captured iterations -- stay resident, pooled
 * by mesh shape and reused with a new lambda; idle ones are evicted
 * least-recently-used), and advances every admitted solve chunk by chunk
 * (8 iterations) without blocking, on `streams` streams, so the work of
 * different requests overlaps on the device.  Results do not depend on
 * what runs beside a solve.
 * ------------------------------------------------------------------- */
typedef struct {
    int n, nummodes, numpoints; /* mesh n^3, P+1, Q                         */
    int k;                      /* forcing frequency                        */
    int mixed;                  /* 0: FP64 CG, 1: mixed-precision PCG        */
    int maxit;
    double lambda, tol;
    double arrival_s;           /* seconds after the start of the run       */
} nmf_request;

typedef struct {
    int64_t unknowns;           /* non-Dirichlet global DOFs                */
    int iterations, outer_iterations, converged;
    int reused_solver;          /* 1: served by a pooled solver              */
    int stream;
    double final_error;         /* relative FP64 residual                   */
    double l2_error;            /* vs the manufactured solution (validate)  */
    double solution_norm;       /* ||x|| over all global rows (validate)    */
    double arrival_s, admit_s, finish_s;
    double latency_ms;          /* finish - arrival                         */
    double queue_ms;            /* admit - arrival                          */
    double device_ms;           /* first to last device event of the solve  */
} nmf_result;

typedef struct {
    int64_t peak_bytes;
    int peak_active, solvers_created, solvers_reused, evictions;
    double makespan_s;
} nmf_server_stats;

int nmf_server_create(int64_t hbm_budget_bytes, int streams, int max_active,
                      nmf_server *out);
int nmf_server_destroy(nmf_server s);
/* Replays the arrival times and returns when every request has finished;
 * results[i] belongs to reqs[i].  validate != 0 also fills l2_error and
 * solution_norm (computed after finish_s is taken).  stats may be NULL. */
int nmf_server_run(nmf_server s, const nmf_request *reqs, int count,
                   int validate, nmf_result *results,
                   nmf_server_stats *stats);

/* The mesh generator used by nmf_solver_create_hex_box: (n+1)^3*3
 * vertex coordinates, vertex (ix,iy,iz) at ((iz*(n+1)+iy)*(n+1)+ix)*3. */
int nmf_hex_box_vertices(int n, double perturb, uint64_t seed, double *xyz);

/* ---------------------------------------------------------------------
 * General C0 assembly and iterative Helmholtz solve on any conforming mesh
 * (hex and tet collections sharing one map).
 * AssemblyMapCG (AssemblyMapCG.cpp:2804-2897): map[i] and sign[i] (+1/-1;
 * NULL: all +1) of every entry i of the concatenated element-local
 * coefficient vector (the operators' local vectors in the order they are
 * passed to nmf_linsys_create), as m_localToGlobalMap /
 * m_localToGlobalSign hold them; the ndir Dirichlet DOFs are numbered
 * first (m_numGlobalDirBndCoeffs).  Vectors are device FP64.  Assembly is
 * deterministic: every global value is summed in increasing local index.
 * ------------------------------------------------------------------- */
int nmf_assembly_create(const int *map, const double *sign, int64_t nloc,
                        int nglobal, int ndir, nmf_assembly *out);
int nmf_assembly_destroy(nmf_assembly a);
/* GlobalToLocal: loc[i] = sign[i] glob[map[i]] (:2804-2827) */
int nmf_assembly_global_to_local(nmf_assembly a, const double *glob,
                                 double *loc, void *stream);
/* Assemble: glob[g] = sum_{map[i]=g} sign[i] loc[i] (:2860-2886) */
int nmf_assembly_assemble(nmf_assembly a, const double *loc, double *glob,
                          void *stream);
/* GlobalLinSysIterativeFull + NekLinSysIterCG with the Diagonal
 * preconditioner (GlobalLinSysIterativeFull.cpp:96-383,
 * NekLinSysIterCG.cpp:118-372, PreconditionerDiagonal.cpp:93-189) over FP64
 * Helmholtz operators (hex and/or tet collections); Dirichlet data are zero
 * unless nmf_linsys_set_dirichlet_local gives their values.  The operators and the map must outlive the handle; the
 * diagonal is built at creation from the operators' current lambda and
 * geometry (exact element-matrix diagonals, one apply per local mode). */
int nmf_linsys_create(nmf_assembly a, const nmf_operator *ops, int nops,
                      void *stream, nmf_linsys *out);
/* The same with the reference's solver name: LinSysIterSolver
 * (GlobalLinSysIterative.cpp:86-88) "ConjugateGradient" (NekLinSysIterCG,
 * assembled vectors; what nmf_linsys_create builds) or
 * "ConjugateGradientLoc" (NekLinSysIterCGLoc.cpp:114-308, element-local
 * vectors: s = A w without assembly, the Diagonal preconditioner assembles
 * the residual, PreconditionerDiagonal.cpp:174-184); or GlobalSysSoln
 * "IterativeStaticCond" (GlobalLinSysStaticCond.cpp:98-230,
 * GlobalLinSysIterativeStaticCond.cpp:378-492: single-level condensation
 * of the interior modes, CG on the assembled Schur complement with its
 * diagonal as preconditioner; get_inv_diag then returns the inverse Schur
 * diagonal on the boundary DOFs).  Other names: NMF_ERR_UNSUPPORTED. */
int nmf_linsys_create_solver(nmf_assembly a, const nmf_operator *ops,
                             int nops, const char *solver, void *stream,
                             nmf_linsys *out);
int nmf_linsys_destroy(nmf_linsys l);
/* Right-hand side in the local layout (device; e.g. IProductWRTBase of f):
 * the global CG solves with b = Assemble(bloc), the local CG starts from
 * r = bloc */
int nmf_linsys_set_rhs_local(nmf_linsys l, const double *bloc);
/* Non-homogeneous Dirichlet data: dloc (device, the local layout) holds the
 * Dirichlet values on the local entries of Dirichlet DOFs and 0 elsewhere,
 * as ContField::v_ImposeDirichletConditions leaves the local coefficients
 * (ContField.cpp:557-608); NULL returns to homogeneous data.  Solves then
 * lift them as GlobalLinSysIterativeFull::v_Solve does
 * (GlobalLinSysIterativeFull.cpp:118-169): rhs = bloc - A dloc element by
 * element, CG for the perturbation (zero on the Dirichlet rows), solution =
 * perturbation + dloc (local) or + the Dirichlet values (global). */
int nmf_linsys_set_dirichlet_local(nmf_linsys l, const double *dloc);
/* Robin conditions du/dn + alpha u = g: the boundary mass that
 * GlobalLinSysIterativeFull adds after every element operator apply
 * (AddRobinTraceContribution, GlobalLinSysIterativeFull.cpp:134-150 and
 * 219-233; the robin BC info of ContField), as rows of the local vector:
 * y[rows[k]] += sum_{m = rowptr[k]}^{rowptr[k+1]-1} vals[m] x[cols[m]]
 * (host arrays; local indices; distinct rows; vals = int_face alpha
 * phi_i phi_j summed over an element's Robin faces).  It enters the
 * matrix multiply, the Dirichlet lifting and the Jacobi diagonal; g goes
 * into the right-hand side as for Neumann faces.  nrows = 0 removes it.
 * NMF_ERR_UNSUPPORTED with IterativeStaticCond. */
int nmf_linsys_set_robin_local(nmf_linsys l, int64_t nrows, const int64_t *rows,
                               const int64_t *rowptr, const int64_t *cols,
                               const double *vals);
/* Stops at ||r|| < tol ||b|| (unpreconditioned, Dirichlet rows excluded) */
int nmf_linsys_solve(nmf_linsys l, double tol, int maxit, nmf_solve_info *info);
int nmf_linsys_get_solution(nmf_linsys l, double *x_host);
/* The solution in the local layout (device, nloc entries), enqueued on the
 * solver's stream: GlobalToLocal of the global solution, or the local CG's
 * own vector (GlobalLinSysIterativeFull.cpp:367-371) */
int nmf_linsys_get_solution_local(nmf_linsys l, double *xloc_dev);
int nmf_linsys_get_inv_diag(nmf_linsys l, double *d_host);
/* The static-condensation diagonal into d_host (nglobal): every element's
 * Schur complement A_bb - A_bi A_ii^-1 A_ib of its interior modes (hex
 * p,q,r >= 2; tet p >= 2, q,r >= 1), its diagonal summed over the global
 * boundary DOFs -- PreconditionerDiagonal::StaticCondDiagonalPreconditionerSum
 * before the inversion (PreconditionerDiagonal.cpp:132-154,
 * Preconditioner.cpp:185-236).  Dirichlet rows and interior DOFs are 0.
 * FP64 on the device: element columns from operator applies to unit modes,
 * a Cholesky factorisation of A_ii per element.  At most 512 interior modes
 * per element; interior modes must not be shared; an element's boundary
 * modes must map to distinct global DOFs (NMF_ERR_UNSUPPORTED otherwise). */
int nmf_linsys_sc_diagonal(nmf_linsys l, double *d_host);
/* y = A x on global device vectors, Dirichlet rows zeroed */
int nmf_linsys_apply_global(nmf_linsys l, const double *x_dev, double *y_dev);

#ifdef __cplusplus
}
#endif
#endif
