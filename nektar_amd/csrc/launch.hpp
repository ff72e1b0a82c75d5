// Kernel launch dispatch for the matrix-free operators.  Each (shape, NM)
// pair is instantiated in its own translation unit (ops_<shape>_nm<NM>.hip)
// for NQ = NM+1 (the MODIFIED default Q = P+2) and NQ = NM+2.
#pragma once

#include <hip/hip_runtime.h>

namespace nektar_amd {

enum class OpType {
    BwdTrans = 0,
    IProduct = 1,
    PhysDeriv = 2,
    Helmholtz = 3,
    IProductDeriv = 4,
    AdvDiffReact = 5,
    PhysInterp = 6
};

struct KernelArgs {
    OpType op;
    bool fp32;
    bool deformed;
    int nelmt;
    const void *tab;
    const void *geo;
    double lambda;
    const void *in;
    void *out0, *out1, *out2;
    hipStream_t stream;
    int tile = 0;  // Helmholtz: elements per workgroup (0 = default)
    const void *in1 = nullptr, *in2 = nullptr;  // IProductDeriv: inputs 1, 2 (in = input 0)
    const void *adv = nullptr;  // AdvDiffReact: advection coefficients per point
    const int *menc = nullptr;  // hex Helmholtz: signed local-to-global map; `in` is then the global vector
};

// Order-templated kernels of the structured-hex C0 solver.
enum class SolverKernel { Apply = 0, Diag = 1, Rhs = 2, BwdGather = 3 };

struct SolverArgs {
    SolverKernel kind;
    bool fp32;      // Apply only: FP32 operator (vectors, tables, geometry)
    bool deformed;
    int n, N;       // elements and global nodes per direction
    const void *tab;
    const void *dtab;  // Diag: {B^2, B B', B'^2} (FP64)
    const void *geo;
    double lambda;
    const void *in;
    void *out;
    const int *done;   // Apply: early exit flag (NULL = always run)
    hipStream_t stream;
    int tile = 0;      // Apply: elements per workgroup (0 = default)
};

// Tile shapes (elements per workgroup) available for the hex Helmholtz
// kernels of a given order, in the order the autotuner tries them.
int hex_tile_candidates(int nm, int nq, bool fp32, int *out, int max);

// Grouped Helmholtz: several operators of one kind (shape, modes, points,
// precision, geometry type) in one launch.  Items are ordered with
// block0 ascending from 0; item i owns workgroups [block0_i, block0_{i+1}).
template <typename T>
struct GroupItem {
    const T *geo;
    const T *in;
    T *out;
    T lambda;
    int nelmt;
    int block0;
};

struct GroupArgs {
    bool fp32;
    bool deformed;
    int count;          // items
    const void *items;  // device array of GroupItem<float|double>
    int nblocks;        // total workgroups of the launch
    const void *tab;    // tables shared by every item of the kind
    hipStream_t stream;
    int tile = 0;       // hex: elements per workgroup (0 = default)
};

// Elements per workgroup: hex default fill (tile 0) or the given tile.
// A tet tile is the number of work-items per element: packed
// (nq0*(nq0-1), elements share waves) or wave-aligned (rounded up to a
// multiple of 64); workgroups hold 256/tile elements (at least one).  The
// kernels assert they use the same numbers.
constexpr int hex_fill_elems(int nq) { return 256 / (nq * nq) > 0 ? 256 / (nq * nq) : 1; }
// Hex Helmholtz tiles kMfmaTile + E select the matrix-core kernels with E
// elements per 256-work-item workgroup; they are compiled for NUMMODES >=
// kMfmaMinModes (where the per-direction products are dense enough).
constexpr int kMfmaTile = 1024, kMfmaMinModes = 5;
inline int hex_tile_elems(int nq, int tile)
{
    return tile >= kMfmaTile ? tile - kMfmaTile : (tile > 0 ? tile : hex_fill_elems(nq));
}
constexpr int tet_packed_tpe(int nq0) { return nq0 * (nq0 - 1); }
constexpr int tet_wave_tpe(int nq0) { return (nq0 * (nq0 - 1) + 63) / 64 * 64; }
// default: wave-aligned where one element fills most of a wave (Q = 8:
// 56 of 64 work-items), packed otherwise (measured, FP64 and FP32 apply)
constexpr int tet_default_tpe(int nq0) { return nq0 == 8 ? tet_wave_tpe(nq0) : tet_packed_tpe(nq0); }
constexpr int tet_tile_elems(int nq0, int tile)
{
    return 256 / (tile > 0 ? tile : tet_default_tpe(nq0)) > 0 ? 256 / (tile > 0 ? tile : tet_default_tpe(nq0)) : 1;
}
// tet tile candidates (work-items per element) in the order the autotuner
// tries them
int tet_tile_candidates(int nq0, int *out, int max);

hipError_t launch_hex_grouped(int nm, int nq, const GroupArgs &a);

// PhysInterp1DScaled: ni / no points per direction in and out, tables
// IT0 | IT1 | IT2 with ITd[i*ni[d] + p] = weight of input point p for
// output point i (k_interp.hip).
struct InterpArgs {
    bool fp32;
    int nelmt;
    int ni[3], no[3];
    const void *tab;
    const void *in;
    void *out;
    hipStream_t stream;
};
constexpr size_t kInterpMaxLds = 65536;  // dynamic LDS of one workgroup
size_t interp_lds_bytes(const int ni[3], const int no[3], bool fp32);
hipError_t launch_phys_interp(const InterpArgs &a);
hipError_t launch_tet_grouped(int nm, int nq0, const GroupArgs &a);

constexpr int kMinModes = 2, kMaxModes = 10;

bool hex_supported(int nm, int nq);
bool hex_op_supported(int nm, int nq);  // Collections operators: also Q = NUMMODES
bool tet_supported(int nm, int nq0);
hipError_t launch_hex(int nm, int nq, const KernelArgs &a);
hipError_t launch_tet(int nm, int nq0, const KernelArgs &a);
hipError_t launch_hex_solver(int nm, int nq, const SolverArgs &a);

}  // namespace nektar_amd
