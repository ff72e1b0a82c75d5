#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <future>
#include <thread>

namespace nektar_amd {

namespace {
using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }
}  // namespace

Server::Server(int64_t budget_bytes, int streams, int max_active) : budget_(budget_bytes), max_active_(max_active)
{
    NMF_REQUIRE(budget_bytes > 0, NMF_ERR_INVALID, "HBM budget must be positive");
    NMF_REQUIRE(streams >= 1 && streams <= 16, NMF_ERR_INVALID, "streams must be 1..16");
    NMF_REQUIRE(max_active >= 1, NMF_ERR_INVALID, "max_active must be >= 1");
    for (int i = 0; i < streams; ++i) {
        hipStream_t s;
        NMF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        streams_.push_back(s);
    }
    stream_load_.assign(streams, 0);
}

Server::~Server()
{
    for (auto s : streams_) (void)hipStreamSynchronize(s);
    pool_.clear();
    for (auto s : streams_) (void)hipStreamDestroy(s);
}

int64_t Server::used_bytes() const
{
    int64_t b = 0;
    for (const auto &p : pool_) b += p->bytes;
    return b;
}

// Evict idle solvers (least recently used first) until `need` more bytes
// fit the budget.  Returns false if it cannot fit even with no idle ones.
bool Server::make_room(int64_t need, ServeStats &st)
{
    int64_t busy = 0;
    for (const auto &p : pool_)
        if (p->busy) busy += p->bytes;
    if (busy + need > budget_) return false;
    while (used_bytes() + need > budget_) {
        int victim = -1;
        for (int i = 0; i < (int)pool_.size(); ++i)
            if (!pool_[i]->busy && (victim < 0 || pool_[i]->last_use < pool_[victim]->last_use)) victim = i;
        if (victim < 0) return false;
        pool_.erase(pool_.begin() + victim);
        ++st.evictions;
    }
    return true;
}

std::vector<ServeResult> Server::run(const std::vector<ServeRequest> &reqs, bool validate, ServeStats *stats_out)
{
    for (const auto &r : reqs) {
        NMF_REQUIRE(r.n >= 1 && hex_supported(r.nummodes, r.numpoints), NMF_ERR_UNSUPPORTED,
                    "request mesh/order not supported");
        NMF_REQUIRE(r.tol > 0.0 && r.maxit > 0 && r.k >= 1, NMF_ERR_INVALID, "request needs tol > 0, maxit > 0, k >= 1");
        NMF_REQUIRE(r.arrival_s >= 0.0, NMF_ERR_INVALID, "arrival times must be >= 0");
        const int64_t need = (int64_t)HexSolver::estimate_bytes(r.n, r.nummodes, r.numpoints, false, r.mixed != 0);
        NMF_REQUIRE(need <= budget_, NMF_ERR_MEMORY, "a request needs more device memory than the whole budget");
    }
    ServeStats st;
    std::vector<ServeResult> out(reqs.size());
    std::vector<int> order(reqs.size());
    for (int i = 0; i < (int)reqs.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });

    struct Active {
        int req;
        Pooled *p;
    };
    std::deque<int> queue;
    std::vector<Active> active, starting;  // starting: solver still being built
    std::vector<int> overtaken(reqs.size(), 0);
    size_t next = 0, finished = 0;
    const auto t0 = Clock::now();

    // the set-up that follows a solver becoming available: forcing, first
    // chunk of the solve, bookkeeping
    auto launch = [&](int ri, Pooled *p) {
        const ServeRequest &r = reqs[ri];
        ServeResult &res = out[ri];
        p->solver->set_forcing_manufactured(r.k);
        p->solver->begin_solve(r.mixed != 0, r.tol, r.maxit);
        p->bytes = std::max<int64_t>(p->bytes, (int64_t)p->solver->device_bytes());
        res.unknowns = (int64_t)p->solver->nglobal() - (int64_t)p->solver->ndirichlet();
        active.push_back({ri, p});
        st.peak_active = std::max(st.peak_active, (int)active.size());
        st.peak_bytes = std::max(st.peak_bytes, used_bytes());
    };
    // admission of one request; false if its memory does not fit now
    auto admit = [&](int ri) {
        const ServeRequest &r = reqs[ri];
        const int64_t est = (int64_t)HexSolver::estimate_bytes(r.n, r.nummodes, r.numpoints, false, r.mixed != 0);
        Pooled *reuse = nullptr;
        for (auto &q : pool_)
            if (!q->busy && q->solver && q->n == r.n && q->nm == r.nummodes && q->nq == r.numpoints) {
                reuse = q.get();
                break;
            }
        const int64_t need = reuse ? std::max<int64_t>(0, est - reuse->bytes) : est;
        if (reuse) reuse->busy = true;  // keep it out of eviction
        if (!make_room(need, st)) {
            if (reuse) reuse->busy = false;
            return false;
        }
        ServeResult &res = out[ri];
        res.arrival_s = r.arrival_s;
        res.admit_s = seconds_since(t0);
        if (reuse) {
            reuse->last_use = ++tick_;
            ++stream_load_[reuse->stream];
            res.stream = reuse->stream;
            reuse->solver->set_lambda(r.lambda);
            res.reused_solver = 1;
            ++st.solvers_reused;
            launch(ri, reuse);
            return true;
        }
        // a new solver, built on a worker thread on the least-loaded stream;
        // its memory is reserved (the estimate) from now on
        const int s = (int)(std::min_element(stream_load_.begin(), stream_load_.end()) - stream_load_.begin());
        auto np = std::make_unique<Pooled>();
        np->n = r.n;
        np->nm = r.nummodes;
        np->nq = r.numpoints;
        np->busy = true;
        np->stream = s;
        np->bytes = est;
        np->last_use = ++tick_;
        const hipStream_t hs = streams_[s];
        const int n = r.n, nm = r.nummodes, nq = r.numpoints;
        const double lam = r.lambda;
        np->building = std::async(std::launch::async, [n, nm, nq, lam, hs] {
            // never beside a graph capture (which the allocations would break)
            std::lock_guard<std::mutex> lock(graph_capture_mutex());
            return std::make_unique<HexSolver>(HexBoxMesh::make(n, 0.0, 0), nm, nq, lam, hs);
        });
        ++stream_load_[s];
        res.stream = s;
        starting.push_back({ri, np.get()});
        pool_.push_back(std::move(np));
        ++st.solvers_created;
        return true;
    };

    while (finished < reqs.size()) {
        const double now = seconds_since(t0);
        while (next < order.size() && reqs[order[next]].arrival_s <= now) queue.push_back(order[next++]);

        // admission: first fit in arrival order while a slot is free; the
        // head can be passed at most kMaxOvertakes times
        for (auto it = queue.begin(); it != queue.end() && (int)(active.size() + starting.size()) < max_active_;) {
            const bool head = it == queue.begin();
            if (!head && overtaken[queue.front()] >= kMaxOvertakes) break;
            const int ri = *it;
            if (!admit(ri)) {
                ++it;
                continue;
            }
            if (!head) ++overtaken[queue.front()];
            it = queue.erase(it);
        }

        bool progressed = false;
        // solvers whose construction has finished start their solve
        for (size_t a = 0; a < starting.size();) {
            Pooled &p = *starting[a].p;
            if (p.building.wait_for(std::chrono::seconds(0)) != std::future_status::ready) {
                ++a;
                continue;
            }
            p.solver = p.building.get();  // rethrows a failed construction
            p.bytes = (int64_t)p.solver->device_bytes();
            launch(starting[a].req, &p);
            starting.erase(starting.begin() + a);
            progressed = true;
        }

        // progress: enqueue the next chunk of every solve whose state is in
        for (size_t a = 0; a < active.size();) {
            Pooled &p = *active[a].p;
            if (!p.solver->advance(false)) {
                ++a;
                continue;
            }
            progressed = true;
            const int ri = active[a].req;
            ServeResult &res = out[ri];
            res.finish_s = seconds_since(t0);
            const SolveInfo info = p.solver->finish();
            res.iterations = info.iterations;
            res.outer_iterations = info.outer_iterations;
            res.converged = info.converged;
            res.final_error = info.final_error;
            res.device_ms = info.solve_ms;
            res.latency_ms = 1e3 * (res.finish_s - res.arrival_s);
            res.queue_ms = 1e3 * (res.admit_s - res.arrival_s);
            if (validate) {
                res.l2_error = p.solver->l2_error_manufactured(reqs[ri].k);
                res.solution_norm = p.solver->solution_norm();
            }
            p.busy = false;
            --stream_load_[p.stream];
            active.erase(active.begin() + a);
            ++finished;
        }
        if (!progressed) {
            if (active.empty() && starting.empty() && queue.empty() && next < order.size()) {
                const double wait = reqs[order[next]].arrival_s - seconds_since(t0);
                if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
            } else {
                std::this_thread::yield();
            }
        }
    }
    st.makespan_s = seconds_since(t0);
    if (stats_out) *stats_out = st;
    return out;
}

}  // namespace nektar_amd
