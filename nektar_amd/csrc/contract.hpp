// Table-row contractions shared by the hex and tet kernels.
//
// The 1-D basis and derivative tables are read through wave-uniform scalar
// loads, so every table value is an SGPR operand of the FMAs.  Two things
// keep that cheap: each contraction reads one contiguous table row per
// output (transposed copies of the tables are stored where a column would
// be needed), and rows are loaded one ahead of their use, with the
// optimiser kept from merging or hoisting loads of different rows.
#pragma once

#include <hip/hip_runtime.h>

namespace nektar_amd {

// Hides a (wave-uniform) table pointer from the optimiser: the asm emits
// nothing, but loads through the returned pointer can neither be merged
// with loads of the same entries elsewhere in the kernel nor be hoisted
// above this point.  Without it the compiler keeps whole tables live in
// scalar registers across all contractions and spills them to vector
// lanes (v_writelane/v_readlane), doubling the vector instruction count.
template <typename T>
using const_ptr = __attribute__((address_space(4))) const T *;

template <typename T>
__device__ __forceinline__ const_ptr<T> opaque(const T *p)
{
    const_ptr<T> q = (const_ptr<T>)p;
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" : "+s"(q));
    return q;
}

// Marks values as needed here: the compiler waits for their loads at this
// point (and not later, behind loads issued afterwards).
template <typename T, int L>
__device__ __forceinline__ void ready_s(const T (&x)[L])
{
#pragma unroll
    for (int l = 0; l < L; ++l) asm volatile("" ::"s"(x[l]));
}
template <typename T, int L>
__device__ __forceinline__ void ready_v(const T (&x)[L])
{
#pragma unroll
    for (int l = 0; l < L; ++l) asm volatile("" ::"v"(x[l]));
}

// out[o] (+)= sum_l v[l] * row(o)[l] for o < NO, row(o) a table row of L
// wave-uniform values.  Row o+1's scalar loads are issued once row o's
// values have arrived and stay in flight while row o's FMAs issue; a wait
// for scalar loads drains every outstanding LDS and scalar operation, so
// the loop itself touches no LDS (callers store `out` afterwards).
template <bool INIT, typename T, int NO, int L, typename RowFn>
__device__ __forceinline__ void contract(const T (&v)[L], RowFn row, T (&out)[NO])
{
    ready_v(v);
    T cur[L];
    {
        const auto r = row(0);
#pragma unroll
        for (int l = 0; l < L; ++l) cur[l] = r[l];
    }
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        ready_s(cur);
        T nxt[L];
        if (o + 1 < NO) {
            const auto r = row(o + 1);
#pragma unroll
            for (int l = 0; l < L; ++l) nxt[l] = r[l];
        }
        T s = INIT ? fma(v[0], cur[0], out[o]) : v[0] * cur[0];
#pragma unroll
        for (int l = 1; l < L; ++l) s = fma(v[l], cur[l], s);
        out[o] = s;
        if (o + 1 < NO) {
#pragma unroll
            for (int l = 0; l < L; ++l) cur[l] = nxt[l];
        }
    }
    // the FMAs stay here (not sunk towards the uses of out, which would
    // keep the table rows live in scalar registers)
    ready_v(out);
}

}  // namespace nektar_amd
