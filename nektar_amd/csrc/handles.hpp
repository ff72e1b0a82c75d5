// Definitions of the opaque C-ABI handle types (include/nektar_mf.h),
// shared by the translation units that implement entry points.
#pragma once

#include <memory>

#include "assembly.hpp"
#include "batch.hpp"
#include "mf_operator.hpp"
#include "server.hpp"
#include "solver.hpp"

struct nmf_operator_s {
    std::unique_ptr<nektar_amd::MFOperator> impl;
};

struct nmf_batch_s {
    std::unique_ptr<nektar_amd::Batch> impl;
};

struct nmf_server_s {
    std::unique_ptr<nektar_amd::Server> impl;
};

struct nmf_solver_s {
    std::unique_ptr<nektar_amd::HexSolver> impl;
};

struct nmf_assembly_s {
    std::unique_ptr<nektar_amd::AssemblyMap> impl;
};

struct nmf_linsys_s {
    std::unique_ptr<nektar_amd::IterativeSolver> impl;
};
