// Matrix-core (MFMA) hex Helmholtz kernels (Helmholtz.h:549-599) for
// gfx950: the Collections, grouped, structured-solver and mapped-solver
// forms of the per-workgroup body in hex_mfma_block.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include "hex_kernels.hpp"
#include "hex_mfma_block.hpp"

namespace nektar_amd {
namespace hexmf {

static_assert(Tab1<double, 7, 8>::OFF_D == hex::Tab<double, 7, 8>::OFF_D &&
                  Tab1<double, 7, 8>::OFF_W == hex::Tab<double, 7, 8>::OFF_W,
              "table layouts differ");

// Collections form: element-local coefficients in[e*NM^3 + m]
template <typename T, int NM, int NQ, bool DEF, int E>
__global__ __launch_bounds__(kBlock) void helmholtz_kernel(int nelmt, const T *__restrict__ tabp,
                                                           const T *__restrict__ geo, T lambda,
                                                           const T *__restrict__ in, T *__restrict__ out)
{
    constexpr int NM3 = NM * NM * NM;
    const int e0 = blockIdx.x * E;
    const auto load = [in](size_t e, int, int, int, int m) { return in[e * NM3 + m]; };
    helmholtz_block<T, NM, NQ, DEF, E>(tabp, geo, lambda, load, out, e0, min(E, nelmt - e0));
}

// Grouped form: several operators of one kind; block b belongs to the item
// whose [block0, next block0) range holds it
template <typename T, int NM, int NQ, bool DEF, int E>
__global__ __launch_bounds__(kBlock) void helmholtz_grouped_kernel(int count, const GroupItem<T> *__restrict__ items,
                                                                   const T *__restrict__ tabp)
{
    constexpr int NM3 = NM * NM * NM;
    const int b = blockIdx.x;
    int it = 0;
    while (it + 1 < count && b >= items[it + 1].block0) ++it;
    const GroupItem<T> I = items[it];
    const int e0 = (b - I.block0) * E;
    const T *in = I.in;
    const auto load = [in](size_t e, int, int, int, int m) { return in[e * NM3 + m]; };
    helmholtz_block<T, NM, NQ, DEF, E>(tabp, I.geo, I.lambda, load, I.out, e0, min(E, I.nelmt - e0));
}

// Structured solver form: coefficients gathered from the global C0 vector
// of the n^3 box (hex::StructuredFetch), results element-major
template <typename T, int NM, int NQ, bool DEF, int E>
__global__ __launch_bounds__(kBlock) void helmholtz_gather_kernel(int n, int N, const T *__restrict__ tabp,
                                                                  const T *__restrict__ geo, T lambda,
                                                                  const T *__restrict__ xg, T *__restrict__ out,
                                                                  const int *__restrict__ done)
{
    if (done && *done) return;
    const int nelmt = n * n * n, e0 = blockIdx.x * E;
    const auto load = [xg, n, N](size_t e, int p, int q, int r, int) {
        const int ex = (int)(e % n), ey = (int)((e / n) % n), ez = (int)(e / ((size_t)n * n));
        const hex::StructuredFetch<T, NM> f{xg, N, ex * (NM - 1), ey * (NM - 1), ez * (NM - 1)};
        return f(p, q, r);
    };
    helmholtz_block<T, NM, NQ, DEF, E>(tabp, geo, lambda, load, out, e0, min(E, nelmt - e0));
}

// General solver form: coefficients through the signed local-to-global map
// (hex::MappedFetch), results element-major
template <typename T, int NM, int NQ, bool DEF, int E>
__global__ __launch_bounds__(kBlock) void helmholtz_mapped_kernel(int nelmt, const T *__restrict__ tabp,
                                                                  const T *__restrict__ geo, T lambda,
                                                                  const T *__restrict__ xg,
                                                                  const int *__restrict__ menc, T *__restrict__ out)
{
    constexpr int NM3 = NM * NM * NM;
    const int e0 = blockIdx.x * E;
    const auto load = [xg, menc](size_t e, int, int, int, int m) {
        const int c = menc[e * NM3 + m];
        const T v = xg[c < 0 ? ~c : c];
        return c < 0 ? -v : v;
    };
    helmholtz_block<T, NM, NQ, DEF, E>(tabp, geo, lambda, load, out, e0, min(E, nelmt - e0));
}

}  // namespace hexmf
}  // namespace nektar_amd
