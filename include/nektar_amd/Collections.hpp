// Header-only C++ mirror of the reference's Collections::Operator interface
// (library/Collections/Operator.h:43-254, Collection.h:56-192) over the C
// ABI, so that code that applies operators through a Collection keeps its
// calls:
//
//   OperatorKey{shape, operator type, implementation, nodal}  <- Operator.h:118-126
//   GetOperatorFactory().CreateInstance(key, ...)             <- Operator.h:128-135
//   operator()(input, output0, output1, output2, wsp)          <- Operator.h:148-152
//   operator()(dir, input, output, wsp)                        <- Operator.h:154-158
//   UpdateFactors / UpdateVarcoeffs                            <- Operator.h:160-178
//   GetWspSize, GetNumElmt, GetInput/Output/Phys/CoeffSize     <- Operator.h:180-230
//
// The reference constructs an operator from the collection's expansions and
// their coalesced geometric factors; here CollectionExpansion carries the
// same information: one BasisKey per direction, the element count and the
// geometric factors per element (regular) or per quadrature point
// (deformed).  Only the MatrixFree implementation of hexahedra,
// tetrahedra and prisms (not nodal) is provided; other keys throw NekError with
// NMF_ERR_UNSUPPORTED, as the reference's factory fails for keys nobody
// registered.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "MatrixFreeOps.hpp"

namespace Nektar {

namespace StdRegions {
// the variable coefficients the matrix-free operators read
// (StdRegions/StdRegions.hpp VarCoeffType; advection velocities of
// LinearAdvectionDiffusionReaction, Collections/LinearAdvectionDiffusion-
// Reaction.cpp:598-690)
enum VarCoeffType { eVarCoeffVelX, eVarCoeffVelY, eVarCoeffVelZ };
using VarCoeffMap = std::map<VarCoeffType, std::vector<double>>;
}  // namespace StdRegions

namespace Collections {

enum OperatorType {
    eBwdTrans,
    eHelmholtz,
    eLinearAdvectionDiffusionReaction,
    eIProductWRTBase,
    eIProductWRTDerivBase,
    ePhysDeriv,
    ePhysInterp1DScaled,
    SIZE_OperatorType
};

const char *const OperatorTypeMap[] = {"BwdTrans",        "Helmholtz",            "LinearAdvectionDiffusionReaction",
                                       "IProductWRTBase", "IProductWRTDerivBase", "PhysDeriv",
                                       "PhysInterp1DScaled"};

enum ImplementationType { eNoImpType, eNoCollection, eIterPerExp, eStdMat, eSumFac, eMatrixFree, SIZE_ImplementationType };

const char *const ImplementationTypeMap[] = {"NoImplementationType", "NoCollection", "IterPerExp", "StdMat",
                                             "SumFac",               "MatrixFree"};

typedef bool ExpansionIsNodal;
typedef std::tuple<LibUtilities::ShapeType, OperatorType, ImplementationType, ExpansionIsNodal> OperatorKey;

// The collection an operator is built for (the reference's
// std::vector<ExpansionSharedPtr> and CoalescedGeomData).
struct CollectionExpansion {
    LibUtilities::ShapeType shape = LibUtilities::eHexahedron;
    std::vector<LibUtilities::BasisKey> basis;  // one per direction
    int nElmt = 0;
    bool deformed = false;
    std::vector<double> jac;  // jac[e] or jac[e*nqtot + q]
    std::vector<double> df;   // df[e*9 + d] or df[(e*nqtot + q)*9 + d], d = j*3 + k: d xi_k / d x_j
};

class Operator {
public:
    Operator(OperatorType type, const CollectionExpansion &exp, const FactorMap &factors) : m_type(type)
    {
        const std::string key = std::string(type == eIProductWRTBase ? "IProduct" : OperatorTypeMap[type]) +
                                MatrixFree::GetOpstring(exp.shape, exp.deformed && type != eBwdTrans &&
                                                                       type != ePhysInterp1DScaled);
        m_op = MatrixFree::GetOperatorFactory().CreateInstance(key, exp.basis, exp.nElmt);
        if (m_op->NeedsJac()) m_op->SetJac(exp.jac);
        if (m_op->NeedsDF()) m_op->SetDF(exp.df);
        int nc = 0, nq = 0, ne = 0, de = 0, pr = 0;
        NmfCheck(nmf_operator_info(m_op->GetHandle(), &nc, &nq, &ne, &de, &pr));
        m_ncoeff = nc;
        m_nqtot = nq;
        m_numElmt = ne;
        int64_t w = 0;
        NmfCheck(nmf_operator_wsp_size(m_op->GetHandle(), &w));
        m_wspSize = (unsigned)w;
        if (!factors.empty()) UpdateFactors(factors);
    }
    virtual ~Operator() = default;
    Operator(const Operator &) = delete;
    Operator &operator=(const Operator &) = delete;

    // operator()(input, output0, output1, output2, wsp): one input and
    // output0 (BwdTrans, IProductWRTBase, Helmholtz, LinearAdvection-
    // DiffusionReaction, PhysInterp1DScaled); three outputs (PhysDeriv);
    // IProductWRTDerivBase reads input, output0 and output1 and writes
    // output2 (Collections/IProductWRTDerivBase.cpp:285-318).
    virtual void operator()(const std::vector<double> &input, std::vector<double> &output0,
                            std::vector<double> &output1, std::vector<double> &output2, std::vector<double> &wsp)
    {
        (void)wsp;
        switch (m_type) {
        case ePhysDeriv: {
            std::vector<std::vector<double>> out;
            (*std::static_pointer_cast<MatrixFree::PhysDeriv>(m_op))(input, out);
            output0.swap(out[0]);
            output1.swap(out[1]);
            output2.swap(out[2]);
            break;
        }
        case eIProductWRTDerivBase: {
            std::vector<std::vector<double>> in = {input, output0, output1};
            (*std::static_pointer_cast<MatrixFree::IProductWRTDerivBase>(m_op))(in, output2);
            break;
        }
        default:
            (*m_op)(input, output0);
        }
    }

    // operator()(dir, input, output, wsp): PhysDeriv's derivative in one
    // direction (PhysDeriv.cpp:312-317); not valid for the other operators
    virtual void operator()(int dir, const std::vector<double> &input, std::vector<double> &output,
                            std::vector<double> &wsp)
    {
        (void)wsp;
        if (m_type != ePhysDeriv) throw NekError(NMF_ERR_INVALID, "operator()(dir, ...) is only valid for PhysDeriv");
        if (dir < 0 || dir > 2) throw NekError(NMF_ERR_INVALID, "PhysDeriv direction must be 0, 1 or 2");
        std::vector<double> o0, o1, o2, w;
        (*this)(input, o0, o1, o2, w);
        output.swap(dir == 0 ? o0 : dir == 1 ? o1 : o2);
    }

    // UpdateFactors: Helmholtz (lambda, constant diffusion:
    // Collections/Helmholtz.cpp:615-691), LinearAdvectionDiffusionReaction
    // (lambda, passed as -lambda: LinearAdvectionDiffusionReaction.cpp:
    // 578-590), PhysInterp1DScaled (eFactorConst: PhysInterp1DScaled.cpp:
    // 158-186); the other operators take no factors.
    virtual void UpdateFactors(const FactorMap &factors)
    {
        m_factors = factors;
        if (m_type == eHelmholtz) {
            Collections::UpdateFactors(*std::static_pointer_cast<MatrixFree::Helmholtz>(m_op), factors);
        } else if (m_type == eLinearAdvectionDiffusionReaction) {
            auto x = factors.find(eFactorLambda);
            if (x == factors.end()) throw NekError(NMF_ERR_INVALID, "Constant factor not defined: FactorLambda");
            std::static_pointer_cast<MatrixFree::LinearAdvectionDiffusionReaction>(m_op)->SetLambda(-x->second);
        } else if (m_type == ePhysInterp1DScaled) {
            auto x = factors.find(eFactorConst);
            if (x == factors.end()) throw NekError(NMF_ERR_INVALID, "Constant factor not defined: FactorConst");
            std::static_pointer_cast<MatrixFree::PhysInterp1DScaled>(m_op)->SetScalingFactor(x->second);
        }
    }

    // UpdateVarcoeffs: the advection velocities of LinearAdvectionDiffusion-
    // Reaction (eVarCoeffVelX/Y/Z, one value per quadrature point of the
    // collection), scaled by -1/lambda as the reference does
    // (LinearAdvectionDiffusionReaction.cpp:598-690); needs UpdateFactors
    // with the lambda first.
    virtual void UpdateVarcoeffs(StdRegions::VarCoeffMap &varcoeffs)
    {
        if (m_type != eLinearAdvectionDiffusionReaction)
            throw NekError(NMF_ERR_UNSUPPORTED, "UpdateVarcoeffs: only LinearAdvectionDiffusionReaction takes "
                                                "variable coefficients in this library");
        if (varcoeffs.empty()) return;
        auto lam = m_factors.find(eFactorLambda);
        if (lam == m_factors.end()) throw NekError(NMF_ERR_STATE, "UpdateVarcoeffs before UpdateFactors (lambda)");
        const size_t npts = (size_t)m_numElmt * m_nqtot;
        std::vector<double> vel(npts * 3, 0.0);
        const StdRegions::VarCoeffType dirs[3] = {StdRegions::eVarCoeffVelX, StdRegions::eVarCoeffVelY,
                                                  StdRegions::eVarCoeffVelZ};
        for (int j = 0; j < 3; ++j) {
            auto v = varcoeffs.find(dirs[j]);
            if (v == varcoeffs.end()) throw NekError(NMF_ERR_INVALID, "three advection velocities are needed");
            if (v->second.size() != npts) throw NekError(NMF_ERR_INVALID, "advection velocity of the wrong size");
            for (size_t p = 0; p < npts; ++p) vel[p * 3 + j] = -v->second[p] / lam->second;
        }
        std::static_pointer_cast<MatrixFree::LinearAdvectionDiffusionReaction>(m_op)->SetAdvectionVelocities(vel);
    }

    unsigned int GetWspSize() const { return m_wspSize; }
    unsigned int GetNumElmt() const { return m_numElmt; }
    unsigned int GetInputSize() const { return (unsigned)m_op->GetInputSize(); }
    unsigned int GetOutputSize() const { return (unsigned)m_op->GetOutputSize(); }
    unsigned int GetPhysSize() const { return (unsigned)m_numElmt * m_nqtot; }
    unsigned int GetCoeffSize() const { return (unsigned)m_numElmt * m_ncoeff; }
    // the C-ABI handle (for the linear-system mirror and device-pointer calls)
    nmf_operator GetHandle() const { return m_op->GetHandle(); }

private:
    OperatorType m_type;
    std::shared_ptr<MatrixFree::Operator> m_op;
    FactorMap m_factors;
    int m_ncoeff = 0, m_nqtot = 0, m_numElmt = 0;
    unsigned int m_wspSize = 0;
};

typedef std::shared_ptr<Operator> OperatorSharedPtr;

// Factory keyed by OperatorKey (Operator.h:128-135, OPERATOR_CREATE).
class OperatorFactory {
public:
    OperatorSharedPtr CreateInstance(const OperatorKey &key, const CollectionExpansion &exp,
                                     const FactorMap &factors = FactorMap()) const
    {
        const LibUtilities::ShapeType shape = std::get<0>(key);
        const OperatorType type = std::get<1>(key);
        const ImplementationType impl = std::get<2>(key);
        if (impl != eMatrixFree || std::get<3>(key))
            throw NekError(NMF_ERR_UNSUPPORTED, std::string("no ") + ImplementationTypeMap[impl] +
                                                    (std::get<3>(key) ? " nodal" : "") + " implementation of " +
                                                    OperatorTypeMap[type] + " in this library (MatrixFree only)");
        if (shape != exp.shape) throw NekError(NMF_ERR_INVALID, "the key's shape differs from the collection's");
        if (type < 0 || type >= SIZE_OperatorType) throw NekError(NMF_ERR_INVALID, "unknown operator type");
        return std::make_shared<Operator>(type, exp, factors);
    }
};

inline OperatorFactory &GetOperatorFactory()
{
    static OperatorFactory f;
    return f;
}

}  // namespace Collections
}  // namespace Nektar
