// Tile autotuner: the workgroup shape (elements per workgroup) of the
// Helmholtz kernels is chosen per (shape, operator, order, points,
// precision, geometry, input form) by timing every instantiated tile on
// the caller's own data, and remembered for the rest of the process so
// later operators of the same kind start from the measured choice.
#pragma once

#include <functional>
#include <vector>

#include <hip/hip_runtime.h>

namespace nektar_amd {

struct TuneKey {
    int shape, op, nm, nq, fp32, deformed, form;  // form 0: element-local input, 1: gathered from a C0 vector
    bool operator<(const TuneKey &o) const;
};

// 0 when nothing has been measured for the key.
int tuned_tile(const TuneKey &k);
void record_tile(const TuneKey &k, int tile);
void clear_tuned_tiles();

// Times launch(tile) for every candidate (warm-up, then `reps` launches
// between two events on `s`) and returns the fastest tile; per-candidate
// mean times in microseconds go to times_us when given.  A candidate whose
// launch fails is skipped.
int time_tiles(hipStream_t s, const std::vector<int> &candidates, const std::function<hipError_t(int)> &launch,
               int reps, std::vector<double> *times_us = nullptr);

}  // namespace nektar_amd
