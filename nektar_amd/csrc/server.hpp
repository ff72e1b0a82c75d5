// Serving loop for C0 Helmholtz solves (continuous batching at the
// granularity of captured 8-iteration chunks).
//
// Requests (structured hex box, order, lambda, forcing frequency, tolerance,
// precision) arrive at given times.  One host loop
//   * queues arrivals in arrival order,
//   * admits queued requests while a slot is free and their device memory
//     fits the HBM budget, first fit: a request that does not fit yet lets
//     later (smaller) ones pass, each at most kMaxOvertakes times, after
//     which it waits for memory with no one passing -- solvers (geometry,
//     diagonal, Krylov vectors, captured iterations) stay resident in a pool
//     keyed by mesh shape and are reused with a new lambda; idle ones are
//     evicted least-recently-used when memory is needed,
//   * builds new solvers on worker threads (allocation, geometry and the
//     diagonal end in a synchronisation of the solver's own stream), so the
//     loop keeps advancing the admitted solves meanwhile,
//   * advances every admitted solve without blocking (HexSolver::advance):
//     each solve's next chunk is enqueued as soon as the previous one has
//     reported its state, on one of a few streams, so the work of many
//     requests overlaps on the device,
//   * retires finished solves and records their latency.
// A solve's arithmetic does not depend on what runs beside it: results are
// identical alone and in the mix.
#pragma once

#include <cstdint>
#include <future>
#include <memory>
#include <vector>

#include <hip/hip_runtime.h>

#include "solver.hpp"

namespace nektar_amd {

struct ServeRequest {
    int n, nummodes, numpoints, k, mixed, maxit;
    double lambda, tol, arrival_s;
};

struct ServeResult {
    int64_t unknowns = 0;
    int iterations = 0, outer_iterations = 0, converged = 0, reused_solver = 0, stream = 0;
    double final_error = 0, l2_error = 0, solution_norm = 0;
    double arrival_s = 0, admit_s = 0, finish_s = 0, latency_ms = 0, queue_ms = 0, device_ms = 0;
};

struct ServeStats {
    int64_t peak_bytes = 0;
    int peak_active = 0, solvers_created = 0, solvers_reused = 0, evictions = 0;
    double makespan_s = 0;
};

class Server {
public:
    Server(int64_t budget_bytes, int streams, int max_active);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // Replays the arrival times (seconds from the call) and returns when
    // every request has finished; results in request order.  validate:
    // also compute each solution's L2 error against the manufactured
    // solution and its norm (after the finish time is taken).
    std::vector<ServeResult> run(const std::vector<ServeRequest> &reqs, bool validate, ServeStats *stats);

private:
    struct Pooled {
        std::unique_ptr<HexSolver> solver;   // null while it is being built
        std::future<std::unique_ptr<HexSolver>> building;
        int n, nm, nq;
        int64_t bytes;
        bool busy;
        uint64_t last_use;
        int stream;
    };
    int64_t used_bytes() const;
    bool make_room(int64_t need, ServeStats &st);

    static constexpr int kMaxOvertakes = 8;
    int64_t budget_;
    int max_active_;
    std::vector<hipStream_t> streams_;
    std::vector<int> stream_load_;
    std::vector<std::unique_ptr<Pooled>> pool_;  // entries do not move
    uint64_t tick_ = 0;
};

}  // namespace nektar_amd
