// Header-only C++ mirror of Nektar++'s matrix-free operator interfaces over
// the C ABI in nektar_mf.h, so that code written against the reference's
// MatrixFree / Collections classes can switch to the MI355X library.
//
//   Nektar::MatrixFree::Operator            <- MatrixFreeOps/Operator.hpp:70-366
//   Nektar::MatrixFree::IProductWRTDerivBase <- MatrixFreeOps/Operator.hpp:192-217
//   Nektar::MatrixFree::GetOperatorFactory  <- MatrixFreeOps/Operator.cpp:40-44
//   Nektar::MatrixFree::GetOpstring         <- MatrixFreeOps/Operator.cpp:46-90
//   Nektar::Collections::Helmholtz_MatrixFree::UpdateFactors
//                                           <- Collections/Helmholtz.cpp:615-691
//
// Arrays are std::vector<double> on the host (the reference's
// Array<OneD, NekDouble>) or raw device pointers; errors throw NekError
// (the reference throws ErrorUtil::NekError from NEKERROR/ASSERTL0).
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../nektar_mf.h"

namespace Nektar {

struct NekError : std::runtime_error {
    int code;
    NekError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void NmfCheck(int rc)
{
    if (rc != NMF_OK) throw NekError(rc, nmf_last_error());
}

namespace LibUtilities {
enum ShapeType { eHexahedron = 0, eTetrahedron = 1, ePrism = 2, ePyramid = 3 };

// The part of BasisKey the matrix-free operators read: modes and points
// of a MODIFIED expansion (Modified_A on GLL for hex; A/B/C on GLL and
// Gauss-Radau(1,0)/(2,0) for tets, MeshGraph.cpp:1688-1752).
struct BasisKey {
    int numModes;
    int numPoints;
};
}  // namespace LibUtilities

namespace MatrixFree {

inline std::string GetOpstring(LibUtilities::ShapeType shape, bool deformed = false)
{
    std::string s = "_";
    s += shape == LibUtilities::eHexahedron    ? "Hex"
         : shape == LibUtilities::eTetrahedron ? "Tet"
         : shape == LibUtilities::ePrism       ? "Prism"
                                               : "Pyr";
    s += deformed ? "_Deformed" : "_Regular";
    return s;
}

class Operator {
public:
    // One BasisKey per direction, as the reference factory takes
    // std::vector<BasisSharedPtr> (Operator.hpp:59-63); modes and points
    // may differ between directions (anisotropic expansions run the
    // library's runtime-size kernels).
    Operator(const std::string &key, const std::vector<LibUtilities::BasisKey> &basis, int nElmt,
             int precision = NMF_FP64)
    {
        if (basis.size() != 3)
            throw NekError(NMF_ERR_INVALID, "a 3-D operator takes one BasisKey per direction (" +
                                                std::to_string(basis.size()) + " given)");
        const int modes[3] = {basis[0].numModes, basis[1].numModes, basis[2].numModes};
        const int points[3] = {basis[0].numPoints, basis[1].numPoints, basis[2].numPoints};
        NmfCheck(nmf_operator_create_basis(key.c_str(), modes, points, nElmt, precision, &m_op));
        int nc, nq, ne, de, pr;
        NmfCheck(nmf_operator_info(m_op, &nc, &nq, &ne, &de, &pr));
        m_ncoeff = nc;
        m_nqtot = nq;
        m_nElmt = ne;
        m_deformed = de != 0;
        m_key = key;
    }
    virtual ~Operator() { nmf_operator_destroy(m_op); }
    Operator(const Operator &) = delete;
    Operator &operator=(const Operator &) = delete;

    bool NeedsDF() const
    {
        return m_key.rfind("PhysDeriv", 0) == 0 || m_key.rfind("Helmholtz", 0) == 0 ||
               m_key.rfind("IProductWRTDerivBase", 0) == 0 || IsAdvDiffReact();
    }
    bool NeedsJac() const
    {
        return m_key.rfind("IProduct", 0) == 0 || m_key.rfind("Helmholtz", 0) == 0 || IsAdvDiffReact();
    }

    // SetJac / SetDF take per-element (not SIMD-interleaved) factors.
    void SetJac(const std::vector<double> &jac)
    {
        CheckSize(jac.size(), GeomPoints(), "SetJac");
        NmfCheck(nmf_operator_set_jac(m_op, jac.data()));
    }
    void SetDF(const std::vector<double> &df)
    {
        CheckSize(df.size(), 9 * GeomPoints(), "SetDF");
        NmfCheck(nmf_operator_set_df(m_op, df.data()));
    }

    // operator()(input, output) on host arrays (BwdTrans, IProduct, Helmholtz)
    void operator()(const std::vector<double> &input, std::vector<double> &output)
    {
        output.resize(OutputSize());
        CheckSize(input.size(), InputSize(), "operator()");
        NmfCheck(nmf_operator_apply_host(m_op, input.data(), output.data()));
    }
    // device form: in/out are device pointers of the operator's precision
    void Apply(const void *in, void *out, void *stream = nullptr) { NmfCheck(nmf_operator_apply(m_op, in, out, stream)); }

    int GetInputSize() const { return (int)InputSize(); }
    int GetOutputSize() const { return (int)OutputSize(); }
    // the C-ABI handle (for nmf_linsys_* and the MultiRegions mirror)
    nmf_operator GetHandle() const { return m_op; }

protected:
    size_t GeomPoints() const { return (size_t)m_nElmt * (m_deformed ? m_nqtot : 1); }
    bool IsAdvDiffReact() const { return m_key.rfind("LinearAdvectionDiffusionReaction", 0) == 0; }
    size_t InputSize() const
    {
        int64_t in = 0;
        NmfCheck(nmf_operator_io_size(m_op, &in, nullptr));
        return (size_t)in;
    }
    size_t OutputSize() const
    {
        int64_t out = 0;
        NmfCheck(nmf_operator_io_size(m_op, nullptr, &out));
        return (size_t)out;
    }
    static void CheckSize(size_t got, size_t want, const char *what)
    {
        if (got != want)
            throw NekError(NMF_ERR_INVALID, std::string(what) + ": array has " + std::to_string(got) +
                                                " entries, expected " + std::to_string(want));
    }

    nmf_operator m_op = nullptr;
    std::string m_key;
    int m_ncoeff = 0, m_nqtot = 0, m_nElmt = 0;
    bool m_deformed = false;
};

class PhysDeriv : public Operator {
public:
    using Operator::Operator;
    // operator()(input, output[3]) (Operator.hpp:238-240)
    void operator()(const std::vector<double> &input, std::vector<std::vector<double>> &output)
    {
        CheckSize(input.size(), InputSize(), "PhysDeriv");
        output.resize(3);
        for (auto &o : output) o.resize(OutputSize());
        NmfCheck(nmf_operator_apply_physderiv_host(m_op, input.data(), output[0].data(), output[1].data(),
                                                   output[2].data()));
    }
};

class IProductWRTDerivBase : public Operator {
public:
    using Operator::Operator;
    // operator()(input[3], output) (Operator.hpp:213-215)
    void operator()(const std::vector<std::vector<double>> &input, std::vector<double> &output)
    {
        if (input.size() != 3) throw NekError(NMF_ERR_INVALID, "IProductWRTDerivBase takes three inputs");
        for (const auto &in : input) CheckSize(in.size(), (size_t)m_nElmt * m_nqtot, "IProductWRTDerivBase");
        output.resize(OutputSize());
        NmfCheck(nmf_operator_apply_iproduct_deriv_host(m_op, input[0].data(), input[1].data(), input[2].data(),
                                                       output.data()));
    }
};

class Helmholtz : public Operator {
public:
    using Operator::Operator;
    void SetLambda(double lambda) { NmfCheck(nmf_operator_set_lambda(m_op, lambda)); }
    // diff = {d00, d01, d11, d02, d12, d22}
    void SetConstVarDiffusion(const std::vector<double> &diff)
    {
        CheckSize(diff.size(), 6, "SetConstVarDiffusion");
        NmfCheck(nmf_operator_set_const_var_diffusion(m_op, diff.data()));
    }
};

// LinearAdvectionDiffusionReaction (Operator.hpp:368-497): out =
// lambda IProduct(u + V . grad u) + L u.  The Collections wrapper passes
// -lambda and V / (-lambda) (Collections/LinearAdvectionDiffusionReaction
// .cpp:578-589, 660-669).
class LinearAdvectionDiffusionReaction : public Operator {
public:
    using Operator::Operator;
    void SetLambda(double lambda) { NmfCheck(nmf_operator_set_lambda(m_op, lambda)); }
    void SetConstVarDiffusion(const std::vector<double> &diff)
    {
        CheckSize(diff.size(), 6, "SetConstVarDiffusion");
        NmfCheck(nmf_operator_set_const_var_diffusion(m_op, diff.data()));
    }
    // vel[(e*nqtot + q)*3 + j]: velocity component j at point q of element e
    void SetAdvectionVelocities(const std::vector<double> &vel)
    {
        CheckSize(vel.size(), (size_t)m_nElmt * m_nqtot * 3, "SetAdvectionVelocities");
        NmfCheck(nmf_operator_set_advection(m_op, vel.data()));
    }
};

// PhysInterp1DScaled (Operator.hpp:128-163): values at the quadrature
// points -> values at the scaled point sets; the Collections wrapper sets
// the factor from eFactorConst (Collections/PhysInterp1DScaled.cpp:158-186).
class PhysInterp1DScaled : public Operator {
public:
    using Operator::Operator;
    void SetScalingFactor(double factor) { NmfCheck(nmf_operator_set_scaling_factor(m_op, factor)); }
};

// String-keyed factory (NekFactory<std::string, Operator, basis, nElmt>).
class OperatorFactory {
public:
    std::shared_ptr<Operator> CreateInstance(const std::string &key, const std::vector<LibUtilities::BasisKey> &basis,
                                             int nElmt, int precision = NMF_FP64) const
    {
        if (key.rfind("PhysDeriv", 0) == 0) return std::make_shared<PhysDeriv>(key, basis, nElmt, precision);
        if (key.rfind("IProductWRTDerivBase", 0) == 0)
            return std::make_shared<IProductWRTDerivBase>(key, basis, nElmt, precision);
        if (key.rfind("Helmholtz", 0) == 0) return std::make_shared<Helmholtz>(key, basis, nElmt, precision);
        if (key.rfind("PhysInterp1DScaled", 0) == 0)
            return std::make_shared<PhysInterp1DScaled>(key, basis, nElmt, precision);
        if (key.rfind("LinearAdvectionDiffusionReaction", 0) == 0)
            return std::make_shared<LinearAdvectionDiffusionReaction>(key, basis, nElmt, precision);
        return std::make_shared<Operator>(key, basis, nElmt, precision);
    }
};

inline OperatorFactory &GetOperatorFactory()
{
    static OperatorFactory f;
    return f;
}

}  // namespace MatrixFree

namespace Collections {

enum ConstFactorType {
    eFactorLambda,
    eFactorCoeffD00,
    eFactorCoeffD01,
    eFactorCoeffD11,
    eFactorCoeffD02,
    eFactorCoeffD12,
    eFactorCoeffD22,
    eFactorConst  // PhysInterp1DScaled's scaling factor (Collections/PhysInterp1DScaled.cpp:158-186)
};
using FactorMap = std::map<ConstFactorType, double>;

// Helmholtz_MatrixFree::UpdateFactors (Collections/Helmholtz.cpp:615-691):
// lambda is required; non-default diffusion coefficients switch the
// operator to the constant-diffusion metric.
inline void UpdateFactors(MatrixFree::Helmholtz &op, const FactorMap &factors)
{
    auto x = factors.find(eFactorLambda);
    if (x == factors.end()) throw NekError(NMF_ERR_INVALID, "Constant factor not defined: FactorLambda");
    op.SetLambda(x->second);
    std::vector<double> diff = {1.0, 0.0, 1.0, 0.0, 0.0, 1.0};
    bool isConstVarDiff = false;
    const struct {
        ConstFactorType t;
        int idx;
        double dflt;
    } map[] = {{eFactorCoeffD00, 0, 1.0}, {eFactorCoeffD01, 1, 0.0}, {eFactorCoeffD11, 2, 1.0},
               {eFactorCoeffD02, 3, 0.0}, {eFactorCoeffD12, 4, 0.0}, {eFactorCoeffD22, 5, 1.0}};
    for (const auto &m : map) {
        auto it = factors.find(m.t);
        if (it != factors.end() && it->second != m.dflt) {
            isConstVarDiff = true;
            diff[m.idx] = it->second;
        }
    }
    if (isConstVarDiff) op.SetConstVarDiffusion(diff);
}

}  // namespace Collections
}  // namespace Nektar
