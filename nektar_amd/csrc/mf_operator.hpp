// MatrixFree operator object behind the C ABI: one operator type for one
// collection of same-shape, same-order elements, exactly as one
// MatrixFree::*Template instance is (library/MatrixFreeOps/Operator.hpp,
// Helmholtz.h:62-110).  Host keeps the FP64 geometry as given; the device
// holds the 1-D tables and the geometry packed for the kernels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "basis.hpp"
#include "common.hpp"
#include "generic.hpp"
#include "launch.hpp"
#include "tune.hpp"

namespace nektar_amd {

struct MFOperator {
    OpType op = OpType::Helmholtz;
    Shape shape = Shape::Hex;
    bool deformed = false;
    bool fp32 = false;
    int nm = 0, nq[3] = {0, 0, 0}, nqtot = 0, ncoeff = 0, nelmt = 0;

    double lambda = 1.0;
    bool has_diff = false;
    double diff[6] = {1, 0, 1, 0, 0, 1};

    std::vector<double> jac, df;
    bool have_jac = false, have_df = false, geo_dirty = true;
    // LinearAdvectionDiffusionReaction: velocity per point, vel[(e*nqtot + q)*3 + j]
    std::vector<double> vel;
    bool have_vel = false, adv_dirty = true;

    // PhysInterp1DScaled: scaling factor, output points per direction and
    // the interpolation matrices I0 | I1 | I2 on the device
    double factor = 1.5;
    int nqo[3] = {0, 0, 0};

    DeviceBuffer tab, geo, adv, itab;
    int tile = 0;  // elements per workgroup of the Helmholtz kernel (0 = default)
    uint64_t version = 0;  // bumped by every change of geometry, factors or tile

    // Modes per direction.  `generic`: no compile-time kernel takes this
    // (shape, modes, points) -- anisotropic orders, point counts other than
    // the instantiated ones -- and the runtime-size kernels run it (the
    // reference's fallback, SwitchNodesPoints.h:302-343).
    int nmd[3] = {0, 0, 0};
    bool generic = false;
    bool equal_modes() const { return nmd[0] == nmd[1] && nmd[1] == nmd[2]; }
    GenLayout glay;              // generic: table offsets and sizes
    DeviceBuffer gitab, gws;     // generic: tet index tables, global workspace

    MFOperator(const std::string &key, int nummodes, int numpoints0, int nelmt, int precision);
    // Per-direction modes and points (MatrixFree::Operator's
    // std::vector<BasisSharedPtr>, Operator.hpp:59-63).
    MFOperator(const std::string &key, const int nummodes[3], const int numpoints[3], int nelmt, int precision);

    bool needs_jac() const
    {
        return op == OpType::IProduct || op == OpType::Helmholtz || op == OpType::IProductDeriv ||
               op == OpType::AdvDiffReact;
    }
    bool needs_df() const
    {
        return op == OpType::PhysDeriv || op == OpType::Helmholtz || op == OpType::IProductDeriv ||
               op == OpType::AdvDiffReact;
    }
    size_t geom_points() const { return deformed ? (size_t)nelmt * nqtot : (size_t)nelmt; }

    void set_jac(const double *j);
    void set_df(const double *d);
    void set_lambda(double l)
    {
        lambda = l;
        ++version;
    }
    void set_diffusion(const double *d6);
    // PhysInterp1DScaled::SetScalingFactor (Operator.hpp:151-155) and
    // SetUpInterp1D (:561-589): output points per direction and tables
    void set_scaling_factor(double f);
    // Collections::Operator::GetInputSize / GetOutputSize per element
    size_t in_per_elmt() const;
    size_t out_per_elmt() const;
    // LinearAdvectionDiffusionReaction::SetAdvectionVelocities
    // (Operator.hpp:476-481): three velocity components at every point
    void set_advection(const double *v);

    void apply(const void *in, void *out0, void *out1, void *out2, hipStream_t s);
    // FP64 hex Helmholtz only: out = A (GlobalToLocal of the global vector xg
    // through the signed map menc, this collection's nelmt * ncoeff entries)
    void apply_mapped(const void *xg, const int *menc, void *out, hipStream_t s);
    bool can_apply_mapped() const { return shape == Shape::Hex && op == OpType::Helmholtz && !fp32 && !generic; }
    // PhysDeriv operator()(dir, in, out, wsp) (Collections/PhysDeriv.cpp
    // :312-317): all three derivatives, the two not asked for into wsp
    // (wsp_size() values of the operator's precision)
    void apply_physderiv_dir(int dir, const void *in, void *out, void *wsp, hipStream_t s);
    // Collections::Operator::GetWspSize: workspace values of the
    // single-direction PhysDeriv form (0 for the other operators)
    size_t wsp_size() const { return op == OpType::PhysDeriv ? (size_t)2 * nelmt * nqtot : 0; }
    size_t value_bytes() const { return fp32 ? sizeof(float) : sizeof(double); }
    // the device geometry matches the host factors (no upload pending)
    bool geometry_current() const
    {
        const bool needs = (needs_jac() || needs_df()) && op != OpType::BwdTrans && op != OpType::PhysInterp;
        return !(needs && geo_dirty) && !(op == OpType::AdvDiffReact && adv_dirty);
    }
    // IProductWRTDerivBase: three inputs at the quadrature points, one output
    void apply_iproduct_deriv(const void *in0, const void *in1, const void *in2, void *out, hipStream_t s);

    // Tiles this operator's kernel is instantiated with (empty when the
    // kernel has a single shape).
    std::vector<int> tile_candidates() const;
    void set_tile(int t);
    // Time every candidate on this operator's geometry (scratch input and
    // output of the operator's size), keep the fastest and record it for
    // later operators of the same kind.  Returns the chosen tile (0 when
    // there is nothing to choose).
    int autotune(hipStream_t s, std::vector<double> *times_us = nullptr);

    // Packed geometry exactly as the kernels read it (host FP64); exposed
    // so solvers can reuse it.
    std::vector<double> pack_geometry() const;
    // Advection coefficients of the reference directions at every point,
    // adv[(e*3 + d)*nqtot + q] (LinearAdvectionDiffusionReaction only).
    std::vector<double> pack_advection() const;
    static std::vector<double> build_tables(Shape shape, int nm, int nq0);

    TuneKey tune_key() const;
    // Upload geometry if it changed (done by apply; batches call it before
    // they take the device pointers).
    // Runtime-size operators also allocate their global workspace here.
    void prepare();

private:
    void ensure_geometry();
    void init(const std::string &key, const int nummodes[3], const int numpoints[3], int n, int precision);
    // generic: arguments of the runtime-size kernels (allocates the global
    // workspace on first use when the buffers exceed one workgroup's LDS)
    GenArgs generic_args(const void *in0, const void *in1, const void *in2, void *o0, void *o1, void *o2,
                         hipStream_t s);
};

// Parse "<Op>_<Shape>_<Regular|Deformed>" (MatrixFree::GetOpstring,
// Operator.cpp:46-90).
void parse_operator_key(const std::string &key, OpType &op, Shape &shape, bool &deformed);

}  // namespace nektar_amd
