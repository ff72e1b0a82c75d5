// Header-only C++ mirror of Nektar++'s C0 assembly map and iterative global
// linear systems over the C ABI in nektar_mf.h, for the Helmholtz operators
// of include/nektar_amd/MatrixFreeOps.hpp:
//
//   Nektar::MultiRegions::AssemblyMapCG::GlobalToLocal / Assemble
//       <- MultiRegions/AssemblyMap/AssemblyMapCG.cpp:2804-2897
//   Nektar::MultiRegions::GlobalLinSysIterative::Solve
//       <- GlobalLinSysIterativeFull.cpp:96-383 (IterativeFull with
//          NekLinSysIterCG / NekLinSysIterCGLoc) and
//          GlobalLinSysStaticCond.cpp:98-230 (IterativeStaticCond), with the
//          Diagonal preconditioner (PreconditionerDiagonal.cpp:69-189)
//
// The map is the one AssemblyMapCG holds: for every entry of the
// concatenated element-local coefficients its global id and sign, the
// Dirichlet DOFs numbered first; Dirichlet data are homogeneous.  Host
// arrays are std::vector<double> (Array<OneD, NekDouble>); device forms take
// raw pointers.  Errors throw NekError.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "MatrixFreeOps.hpp"

namespace Nektar {
namespace MultiRegions {

namespace detail {
// a device array for host <-> device staging
class DeviceArray {
public:
    explicit DeviceArray(size_t n) : m_n(n)
    {
        if (n) NmfCheck(nmf_device_malloc(&m_p, n * sizeof(double)));
    }
    ~DeviceArray()
    {
        if (m_p) nmf_device_free(m_p);
    }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    double *get() const { return static_cast<double *>(m_p); }
    void upload(const std::vector<double> &h) const
    {
        if (m_n) NmfCheck(nmf_memcpy_h2d(m_p, h.data(), m_n * sizeof(double)));
    }
    void download(std::vector<double> &h) const
    {
        h.resize(m_n);
        if (m_n) NmfCheck(nmf_memcpy_d2h(h.data(), m_p, m_n * sizeof(double)));
    }

private:
    void *m_p = nullptr;
    size_t m_n;
};
}  // namespace detail

class AssemblyMapCG {
public:
    // localToGlobalSign may be empty (no sign changes, m_signChange false)
    AssemblyMapCG(const std::vector<int> &localToGlobalMap, const std::vector<double> &localToGlobalSign,
                  int numGlobalCoeffs, int numGlobalDirBndCoeffs)
        : m_nloc((int64_t)localToGlobalMap.size()), m_nglobal(numGlobalCoeffs), m_ndir(numGlobalDirBndCoeffs)
    {
        if (!localToGlobalSign.empty() && localToGlobalSign.size() != localToGlobalMap.size())
            throw NekError(NMF_ERR_INVALID, "AssemblyMapCG: map and sign sizes differ");
        NmfCheck(nmf_assembly_create(localToGlobalMap.data(),
                                     localToGlobalSign.empty() ? nullptr : localToGlobalSign.data(), m_nloc,
                                     m_nglobal, m_ndir, &m_a));
    }
    ~AssemblyMapCG() { nmf_assembly_destroy(m_a); }
    AssemblyMapCG(const AssemblyMapCG &) = delete;
    AssemblyMapCG &operator=(const AssemblyMapCG &) = delete;

    int GetNumLocalCoeffs() const { return (int)m_nloc; }
    int GetNumGlobalCoeffs() const { return m_nglobal; }
    int GetNumGlobalDirBndCoeffs() const { return m_ndir; }

    // device forms
    void GlobalToLocal(const double *global, double *loc, void *stream = nullptr) const
    {
        NmfCheck(nmf_assembly_global_to_local(m_a, global, loc, stream));
    }
    void Assemble(const double *loc, double *global, void *stream = nullptr) const
    {
        NmfCheck(nmf_assembly_assemble(m_a, loc, global, stream));
    }
    // host forms
    void GlobalToLocal(const std::vector<double> &global, std::vector<double> &loc) const
    {
        Check(global.size(), (size_t)m_nglobal, "GlobalToLocal");
        detail::DeviceArray g(m_nglobal), l(m_nloc);
        g.upload(global);
        GlobalToLocal(g.get(), l.get());
        NmfCheck(nmf_device_synchronize());
        l.download(loc);
    }
    void Assemble(const std::vector<double> &loc, std::vector<double> &global) const
    {
        Check(loc.size(), (size_t)m_nloc, "Assemble");
        detail::DeviceArray g(m_nglobal), l(m_nloc);
        l.upload(loc);
        Assemble(l.get(), g.get());
        NmfCheck(nmf_device_synchronize());
        g.download(global);
    }

    nmf_assembly GetHandle() const { return m_a; }

private:
    static void Check(size_t got, size_t want, const char *what)
    {
        if (got != want)
            throw NekError(NMF_ERR_INVALID, std::string(what) + ": array has " + std::to_string(got) +
                                                " entries, expected " + std::to_string(want));
    }
    nmf_assembly m_a = nullptr;
    int64_t m_nloc;
    int m_nglobal, m_ndir;
};

struct SolveInfo {
    int iterations = 0;
    bool converged = false;
    double finalError = 0.0;    // sqrt(eps / rhs magnitude), unpreconditioned
    double rhsMagnitude = 0.0;  // the magnitude the stopping test used
    double solveMs = 0.0;       // device time
};

// GlobalLinSysIterative for Helmholtz collections sharing one map.
// globalSysSoln: "IterativeFull" or "IterativeStaticCond" (SolverInfo
// GlobalSysSoln); linSysIterSolver: "ConjugateGradient" or
// "ConjugateGradientLoc" (SolverInfo LinSysIterSolver, IterativeFull only).
class GlobalLinSysIterative {
public:
    GlobalLinSysIterative(const AssemblyMapCG &map, const std::vector<std::shared_ptr<MatrixFree::Operator>> &ops,
                          const std::string &globalSysSoln = "IterativeFull",
                          const std::string &linSysIterSolver = "ConjugateGradient", void *stream = nullptr)
        : m_map(map), m_ops(ops)
    {
        if (globalSysSoln != "IterativeFull" && globalSysSoln != "IterativeStaticCond")
            throw NekError(NMF_ERR_UNSUPPORTED, "GlobalSysSoln '" + globalSysSoln + "' is not available");
        const std::string name = globalSysSoln == "IterativeStaticCond" ? globalSysSoln : linSysIterSolver;
        std::vector<nmf_operator> h;
        for (const auto &o : ops) h.push_back(o->GetHandle());
        NmfCheck(nmf_linsys_create_solver(map.GetHandle(), h.data(), (int)h.size(), name.c_str(), stream, &m_l));
    }
    ~GlobalLinSysIterative() { nmf_linsys_destroy(m_l); }
    GlobalLinSysIterative(const GlobalLinSysIterative &) = delete;
    GlobalLinSysIterative &operator=(const GlobalLinSysIterative &) = delete;

    // GlobalLinSys::Solve(pLocInput, pLocOutput, locToGloMap) on host arrays:
    // locInput is the local right-hand side (e.g. IProductWRTBase of the
    // forcing), locOutput the local solution (NekLinSysIter tolerance and
    // maximum iterations; the reference's defaults 1e-9 and 5000)
    void Solve(const std::vector<double> &locInput, std::vector<double> &locOutput, double tol = 1.0e-9,
               int maxit = 5000)
    {
        const size_t n = (size_t)m_map.GetNumLocalCoeffs();
        if (locInput.size() != n) throw NekError(NMF_ERR_INVALID, "Solve: local input has the wrong size");
        detail::DeviceArray in(n), out(n);
        in.upload(locInput);
        NmfCheck(nmf_linsys_set_rhs_local(m_l, in.get()));
        nmf_solve_info info{};
        NmfCheck(nmf_linsys_solve(m_l, tol, maxit, &info));
        NmfCheck(nmf_linsys_get_solution_local(m_l, out.get()));
        NmfCheck(nmf_device_synchronize());
        out.download(locOutput);
        m_info.iterations = info.iterations;
        m_info.converged = info.converged != 0;
        m_info.finalError = info.final_error;
        m_info.rhsMagnitude = info.rhs_magnitude;
        m_info.solveMs = info.solve_ms;
    }
    const SolveInfo &GetLastSolveInfo() const { return m_info; }

    // the global solution of the last solve (GetNumGlobalCoeffs entries)
    std::vector<double> GetSolution() const
    {
        std::vector<double> x(m_map.GetNumGlobalCoeffs());
        NmfCheck(nmf_linsys_get_solution(m_l, x.data()));
        return x;
    }
    // the preconditioner's inverse diagonal (m_diagonals, in the global numbering)
    std::vector<double> GetInverseDiagonal() const
    {
        std::vector<double> d(m_map.GetNumGlobalCoeffs());
        NmfCheck(nmf_linsys_get_inv_diag(m_l, d.data()));
        return d;
    }
    // Preconditioner::AssembleStaticCondGlobalDiagonals, in the global numbering
    std::vector<double> GetStaticCondDiagonal() const
    {
        std::vector<double> d(m_map.GetNumGlobalCoeffs());
        NmfCheck(nmf_linsys_sc_diagonal(m_l, d.data()));
        return d;
    }

private:
    const AssemblyMapCG &m_map;
    std::vector<std::shared_ptr<MatrixFree::Operator>> m_ops;  // kept alive while the system is
    nmf_linsys m_l = nullptr;
    SolveInfo m_info;
};

}  // namespace MultiRegions
}  // namespace Nektar
