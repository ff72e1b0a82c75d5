#include "tune.hpp"

#include <limits>
#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"

namespace nektar_amd {

bool TuneKey::operator<(const TuneKey &o) const
{
    return std::tie(shape, op, nm, nq, fp32, deformed, form) <
           std::tie(o.shape, o.op, o.nm, o.nq, o.fp32, o.deformed, o.form);
}

namespace {
std::mutex g_tune_mutex;
std::map<TuneKey, int> &tune_table()
{
    static std::map<TuneKey, int> t;
    return t;
}
}  // namespace

int tuned_tile(const TuneKey &k)
{
    std::lock_guard<std::mutex> l(g_tune_mutex);
    auto it = tune_table().find(k);
    return it == tune_table().end() ? 0 : it->second;
}

void record_tile(const TuneKey &k, int tile)
{
    std::lock_guard<std::mutex> l(g_tune_mutex);
    tune_table()[k] = tile;
}

void clear_tuned_tiles()
{
    std::lock_guard<std::mutex> l(g_tune_mutex);
    tune_table().clear();
}

int time_tiles(hipStream_t s, const std::vector<int> &candidates, const std::function<hipError_t(int)> &launch,
               int reps, std::vector<double> *times_us)
{
    NMF_REQUIRE(!candidates.empty(), NMF_ERR_UNSUPPORTED, "no tile candidates");
    NMF_REQUIRE(reps > 0, NMF_ERR_INVALID, "reps must be positive");
    hipEvent_t e0, e1;
    NMF_HIP(hipEventCreate(&e0));
    NMF_HIP(hipEventCreate(&e1));
    int best = 0;
    double best_t = std::numeric_limits<double>::infinity();
    if (times_us) times_us->assign(candidates.size(), std::numeric_limits<double>::infinity());
    for (size_t c = 0; c < candidates.size(); ++c) {
        const int t = candidates[c];
        if (launch(t) != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        for (int w = 0; w < 2; ++w) NMF_HIP(launch(t));
        NMF_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) NMF_HIP(launch(t));
        NMF_HIP(hipEventRecord(e1, s));
        NMF_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        NMF_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1e3 * ms / reps;
        if (times_us) (*times_us)[c] = us;
        if (us < best_t) {
            best_t = us;
            best = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    NMF_REQUIRE(best > 0, NMF_ERR_DEVICE, "every tile candidate failed to launch");
    return best;
}

}  // namespace nektar_amd
