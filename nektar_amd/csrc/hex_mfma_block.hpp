// Matrix-core (MFMA) form of the fused hex Helmholtz operator: the
// per-workgroup body shared by the kernels of hex_mfma.hpp
// (Helmholtz.h:549-599, HelmholtzKernels.hpp:389-696) for gfx950.
//
// Every sum-factorisation step out[o] = sum_t M[o][t] in[t] along one tensor
// axis is a 16x16x4 MFMA product: A = the 1-D matrix M (16 output slots x 4
// input slots per step, held in registers for the whole kernel), B = the
// inputs of 16 tensor lines (4 input slots per step), D = 16 output slots of
// the same 16 lines.  Lanes: line slot n = lane % 16, group g = lane / 16.
//   A operand of step s, lane l: M[o(l % 16)][4s + g]
//   B operand of step s, lane l: line n, tensor index 4s + g
//   D register r of lane l:       line n, tensor index 4r + g
// with o(m) the tensor index of output row slot m (Mma<T>::slot).  The
// output register r of one product is therefore the B operand of step r of
// the next product along the same axis: u -> du/dxi2 and D2^T -> B^T run
// back to back in registers, and every other axis change goes through LDS.
// Slots past the tensor extent carry zeros (A rows and columns are zero
// there, B inputs are loaded as zero), so padding never adds NaNs.
//
// Layout: three LDS slabs of E elements; a slab holds an element's tensor
// at (a0, a1, a2) -> a0 + a1*NQ + a2*NQ^2 whatever its extents (NM or NQ
// per axis).  The 16 lines of a panel run over the E elements of the
// workgroup, so lines are never padded per element.  The axis-2 panels of
// each wave keep u and du/dxi2, then V2 and the mass term, in registers
// across the pointwise stage.
//
// Reference summation order is not reproduced (the products sum in fours);
// the FP64 result agrees with the reference to rounding (tested at 1e-12).
//
// This header also compiles as plain C++ for the host-side check of the
// indexing and arithmetic (tests/cpu_mfma): there 256 threads play the
// work-items, barriers are thread barriers and the two MFMA products are
// emulated with the lane layouts below.
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define NMF_MD __device__ __forceinline__
#define NMF_LDS __shared__
#else
#include <algorithm>
#include <cmath>
#define NMF_MD inline
#define NMF_LDS static
#endif

namespace nektar_amd {
namespace hexmf {

#ifdef __HIPCC__
NMF_MD int work_item() { return threadIdx.x; }
NMF_MD void barrier() { __syncthreads(); }
using mfma_f64x4 = double __attribute__((ext_vector_type(4)));
using mfma_f32x4 = float __attribute__((ext_vector_type(4)));
#else
// supplied by the host harness
int work_item();
void barrier();
template <typename T>
struct HostV4 {
    T v[4];
    T &operator[](int i) { return v[i]; }
    const T &operator[](int i) const { return v[i]; }
};
using mfma_f64x4 = HostV4<double>;
using mfma_f32x4 = HostV4<float>;
mfma_f64x4 host_mfma_f64(double a, double b, mfma_f64x4 c);
mfma_f32x4 host_mfma_f32(float a, float b, mfma_f32x4 c);
using std::fma;
using std::min;
#endif

template <typename T>
struct Mma;
// A[m][k] comes from lane m + 16k and B[k][n] from lane n + 16k in both
// precisions; row m of D sits in register m % 4 of lane group m / 4 (FP32)
// or in register m / 4 of lane group m % 4 (FP64), column n in lane n of
// the group.  slot(m) is the tensor index given to row m so that register
// r of lane group g always holds index 4r + g.
template <>
struct Mma<double> {
    using v4 = mfma_f64x4;
    static NMF_MD v4 op(double a, double b, v4 c)
    {
#ifdef __HIPCC__
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
#else
        return host_mfma_f64(a, b, c);
#endif
    }
    static NMF_MD int slot(int m) { return m; }
};
template <>
struct Mma<float> {
    using v4 = mfma_f32x4;
    static NMF_MD v4 op(float a, float b, v4 c)
    {
#ifdef __HIPCC__
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
#else
        return host_mfma_f32(a, b, c);
#endif
    }
    static NMF_MD int slot(int m) { return 4 * (m & 3) + (m >> 2); }
};

// 1-D tables in the layout of the operator's table buffer (the hex kernels'
// Tab): B[p*NQ + i] = phi_p(z_i), D[l*NQ + i] = l_l'(z_i), W[i]
template <typename T, int NM, int NQ>
struct Tab1 {
    static constexpr int OFF_D = NM * NQ, OFF_W = OFF_D + NQ * NQ;
    const T *base;
    NMF_MD T B(int p, int i) const { return base[p * NQ + i]; }
    NMF_MD T D(int l, int i) const { return base[OFF_D + l * NQ + i]; }
    NMF_MD T W(int i) const { return base[OFF_W + i]; }
};

constexpr int kWaves = 4, kBlock = 64 * kWaves;

// strides of a contraction along axis AX: SD along AX, SA and SB along the
// faster and the slower of the two line axes
template <int AX, int NQ>
struct Axis {
    static constexpr int SD = AX == 0 ? 1 : (AX == 1 ? NQ : NQ * NQ);
    static constexpr int SA = AX == 0 ? NQ : 1;
    static constexpr int SB = AX == 2 ? NQ : NQ * NQ;
};

// Panel it of a wave: lines pan*16 + lane%16 with pan = wave + kWaves*it of
// the E*NA*NB lines of axis AX (NA, NB: extents of the two line axes).
template <int AX, int NA, int NB, int E, int NQ>
struct Panel {
    static constexpr int NL = NA * NB, NLINES = E * NL, NPAN = (NLINES + 15) / 16;
    static constexpr int NPW = (NPAN + kWaves - 1) / kWaves;
    // panel it of the wave exists (wave-uniform; true for every it below
    // NPAN / kWaves, so only the last iteration can branch)
    static NMF_MD bool has(int wave, int it) { return it < NPAN / kWaves || wave + kWaves * it < NPAN; }
    bool lv;  // this lane's line exists (false on every lane of a panel past the last)
    int e, xa, xb, base;
    NMF_MD Panel(int wave, int lane, int it)
    {
        const int pan = wave + kWaves * it;
        const int line = pan * 16 + (lane & 15);
        lv = pan < NPAN && line < NLINES;
        const int ll = lv ? line : 0;
        e = ll / NL;
        const int rest = ll - e * NL;
        xb = rest / NA;
        xa = rest - xb * NA;
        base = e * NQ * NQ * NQ + xa * Axis<AX, NQ>::SA + xb * Axis<AX, NQ>::SB;
    }
};

// the B operands of a panel: inputs 4s + g of its lines, zero past NIN
template <typename T, int KS, int NIN, int SD>
NMF_MD void load_panel(const T *src, int base, bool lv, int g, T (&b)[KS])
{
    static_assert(KS <= 4 && NIN <= 4 * KS, "one tile covers at most 16 inputs");
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int t = 4 * s + g;
        b[s] = (lv && t < NIN) ? src[base + t * SD] : T(0);
    }
}

// acc += M * x, x the B operands of KS steps
template <typename T, int KS, typename X>
NMF_MD typename Mma<T>::v4 chain(const T (&A)[KS], const X &x, typename Mma<T>::v4 acc)
{
#pragma unroll
    for (int s = 0; s < KS; ++s) acc = Mma<T>::op(A[s], x[s], acc);
    return acc;
}

template <typename T, int NOUT, int SD>
NMF_MD void store(T *dst, int base, bool lv, int g, const typename Mma<T>::v4 &acc)
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int o = 4 * r + g;
        if (lv && o < NOUT) dst[base + o * SD] = acc[r];
    }
}

// One contraction along axis AX for every panel of the wave: src -> dst
// (src == dst allowed: a panel writes only its own lines).  All panels'
// inputs are loaded before the first product, so the loads overlap and
// the products of different panels interleave; a wave skips the products
// of a panel slot past the last panel.
template <typename T, int AX, int NA, int NB, int E, int NQ, int NIN, int NOUT, int KS>
NMF_MD void stage(const T (&M)[KS], const T *src, T *dst, int wave, int lane)
{
    using P = Panel<AX, NA, NB, E, NQ>;
    constexpr int SD = Axis<AX, NQ>::SD, NPW = P::NPW;
    const int g = lane >> 4;
    int base[NPW];
    bool lv[NPW];
    T b[NPW][KS];
#pragma unroll
    for (int it = 0; it < NPW; ++it) {
        const P pn(wave, lane, it);
        base[it] = pn.base;
        lv[it] = pn.lv;
        load_panel<T, KS, NIN, SD>(src, pn.base, pn.lv, g, b[it]);
    }
#pragma unroll
    for (int it = 0; it < NPW; ++it) {
        if (!P::has(wave, it)) break;
        const typename Mma<T>::v4 zero = {T(0), T(0), T(0), T(0)};
        store<T, NOUT, SD>(dst, base[it], lv[it], g, chain<T, KS>(M, b[it], zero));
    }
}

// LDS bytes of the kernel with E elements per workgroup
template <typename T, int NQ, int E>
constexpr int lds_bytes()
{
    return 3 * E * NQ * NQ * NQ * (int)sizeof(T);
}
template <typename T, int NM, int NQ, int E>
constexpr bool fits()
{
    return NQ <= 16 && NM <= NQ && E >= 1 && lds_bytes<T, NQ, E>() <= 65536;
}

// y = lambda M x + L x for elements e0 .. e0 + nval - 1 (nval <= E) of one
// operator, by a workgroup of 256 work-items.  load(e, p, q, r, m) returns
// coefficient (p, q, r) = m of element e; geo and out are indexed by the
// element number as in the hex Helmholtz kernels (hex_kernels.hpp).
template <typename T, int NM, int NQ, bool DEF, int E, typename Load>
NMF_MD void helmholtz_block(const T *__restrict__ tabp, const T *__restrict__ geo, T lambda,
                                                const Load &load, T *__restrict__ out, int e0, int nval)
{
#ifdef __HIPCC__
    // no contraction of separate multiplies and adds: every instantiation
    // (tile) rounds the same way, the fused multiply-adds are written out
#pragma clang fp contract(off)
#endif
    using V = typename Mma<T>::v4;
    constexpr int SL = NQ * NQ * NQ, NM3 = NM * NM * NM, NQ3 = NQ * NQ * NQ;
    constexpr int KM = (NM + 3) / 4, KQ = (NQ + 3) / 4;
    static_assert(fits<T, NM, NQ, E>(), "tile does not fit");
    NMF_LDS T A[E * SL];
    NMF_LDS T B[E * SL];
    NMF_LDS T C[E * SL];
    const Tab1<T, NM, NQ> tab{tabp};
    const int tid = work_item(), lane = tid & 63, wave = tid >> 6, g = lane >> 4;

    // the four 1-D matrices as A operands: Ab (p -> i), Ad (derivative),
    // At (derivative transposed), Ai (i -> p)
    T Ab[KM], Ad[KQ], At[KQ], Ai[KQ];
    {
        const int ro = Mma<T>::slot(lane & 15);
#pragma unroll
        for (int s = 0; s < KM; ++s) {
            const int p = 4 * s + g;
            Ab[s] = (p < NM && ro < NQ) ? tab.B(p, ro) : T(0);
        }
#pragma unroll
        for (int s = 0; s < KQ; ++s) {
            const int l = 4 * s + g;
            Ad[s] = (l < NQ && ro < NQ) ? tab.D(l, ro) : T(0);
            At[s] = (l < NQ && ro < NQ) ? tab.D(ro, l) : T(0);
            Ai[s] = (l < NQ && ro < NM) ? tab.B(ro, l) : T(0);
        }
    }
    // coefficients (p, q, r) -> A
    for (int t = tid; t < E * NM3; t += kBlock) {
        const int e = t / NM3, m = t - e * NM3;
        const int p = m % NM, q = (m / NM) % NM, r = m / (NM * NM);
        A[e * SL + p + q * NQ + r * NQ * NQ] = e < nval ? load((size_t)e0 + e, p, q, r, m) : T(0);
    }
    barrier();
    // p -> i along axis 0, lines (q, r): A -> B
    stage<T, 0, NM, NM, E, NQ, NM, NQ, KM>(Ab, A, B, wave, lane);
    barrier();
    // q -> j along axis 1, lines (i, r): B -> A
    stage<T, 1, NQ, NM, E, NQ, NM, NQ, KM>(Ab, B, A, wave, lane);
    barrier();
    // r -> k along axis 2 (u, to B and registers), then du/dxi2 in registers
    using P2 = Panel<2, NQ, NQ, E, NQ>;
    constexpr int N2 = P2::NPW;
    const V zero = {T(0), T(0), T(0), T(0)};
    int base2[N2];
    bool lv2[N2];
    V ureg[N2], greg[N2];
    {
        T b[N2][KM];
#pragma unroll
        for (int it = 0; it < N2; ++it) {
            const P2 pn(wave, lane, it);
            base2[it] = pn.base;
            lv2[it] = pn.lv;
            load_panel<T, KM, NM, NQ * NQ>(A, pn.base, pn.lv, g, b[it]);
        }
#pragma unroll
        for (int it = 0; it < N2; ++it) {
            ureg[it] = zero;
            greg[it] = zero;
        }
#pragma unroll
        for (int it = 0; it < N2; ++it) {
            if (!P2::has(wave, it)) break;
            ureg[it] = chain<T, KM>(Ab, b[it], zero);
        }
#pragma unroll
        for (int it = 0; it < N2; ++it) {
            if (!P2::has(wave, it)) break;
            greg[it] = chain<T, KQ>(Ad, ureg[it], zero);
            store<T, NQ, NQ * NQ>(B, base2[it], lv2[it], g, ureg[it]);
        }
    }
    barrier();
    // du/dxi0 (B -> A) and du/dxi1 (B -> C)
    stage<T, 0, NQ, NQ, E, NQ, NQ, NQ, KQ>(Ad, B, A, wave, lane);
    stage<T, 1, NQ, NQ, E, NQ, NQ, NQ, KQ>(Ad, B, C, wave, lane);
    barrier();
    // pointwise on the axis-2 panels: metric, weights, mass term; V0 -> A,
    // V1 -> C (in place), V2 and the mass term stay in registers
#pragma unroll
    for (int it = 0; it < N2; ++it) {
        const P2 pn(wave, lane, it);
        const int i = pn.xa, j = pn.xb;
        const bool ev = pn.lv && pn.e < nval;
        const size_t eg = (size_t)e0 + pn.e;
        T gg[7];
        T wij = T(1);
        if (!DEF) {
#pragma unroll
            for (int c = 0; c < 7; ++c) gg[c] = ev ? geo[eg * 7 + c] : T(0);
            wij = tab.W(i) * tab.W(j);
        }
        V v2 = zero, y = zero;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 4 * r + g;
            const bool pv = pn.lv && k < NQ;
            const int a = pn.base + k * NQ * NQ;
            T w = T(1);
            if (DEF) {
                const int c = (k * NQ + j) * NQ + i;
#pragma unroll
                for (int cc = 0; cc < 7; ++cc) gg[cc] = (ev && k < NQ) ? geo[(eg * 7 + cc) * NQ3 + c] : T(0);
            } else {
                w = wij * (k < NQ ? tab.W(k) : T(0));
            }
            const T a0 = pv ? A[a] : T(0), a1 = pv ? C[a] : T(0), a2 = greg[it][r];
            T t0 = gg[0] * a0;
            t0 = fma(gg[1], a1, t0);
            t0 = fma(gg[2], a2, t0);
            T t1 = gg[1] * a0;
            t1 = fma(gg[3], a1, t1);
            t1 = fma(gg[4], a2, t1);
            T t2 = gg[2] * a0;
            t2 = fma(gg[4], a1, t2);
            t2 = fma(gg[5], a2, t2);
            if (!DEF) {
                t0 *= w;
                t1 *= w;
                t2 *= w;
            }
            if (pv) {
                A[a] = t0;
                C[a] = t1;
            }
            v2[r] = pv ? t2 : T(0);
            const T m = ureg[it][r];
            y[r] = pv ? (lambda * gg[6]) * (DEF ? m : w * m) : T(0);
        }
        greg[it] = v2;
        ureg[it] = y;
    }
    barrier();
    // D0^T along axis 0, in place in A
    stage<T, 0, NQ, NQ, E, NQ, NQ, NQ, KQ>(At, A, A, wave, lane);
    barrier();
    // D1^T along axis 1 plus the axis-0 part at the same points, in place in C
    {
        using P = Panel<1, NQ, NQ, E, NQ>;
        constexpr int N1 = P::NPW;
        int base[N1];
        bool lv[N1];
        T b[N1][KQ];
        V acc[N1];
#pragma unroll
        for (int it = 0; it < N1; ++it) {
            const P pn(wave, lane, it);
            base[it] = pn.base;
            lv[it] = pn.lv;
            load_panel<T, KQ, NQ, NQ>(C, pn.base, pn.lv, g, b[it]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 4 * r + g;
                acc[it][r] = (pn.lv && o < NQ) ? A[pn.base + o * NQ] : T(0);
            }
        }
#pragma unroll
        for (int it = 0; it < N1; ++it) {
            if (!P::has(wave, it)) break;
            store<T, NQ, NQ>(C, base[it], lv[it], g, chain<T, KQ>(At, b[it], acc[it]));
        }
    }
    barrier();
    // axis 2: z = D2^T V2 + (axis-0 and axis-1 parts) + mass term, then
    // k -> r in registers: B as (i, j, r)
    {
        V z[N2];
#pragma unroll
        for (int it = 0; it < N2; ++it) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * r + g;
                z[it][r] = (lv2[it] && k < NQ) ? C[base2[it] + k * NQ * NQ] + ureg[it][r] : T(0);
            }
        }
#pragma unroll
        for (int it = 0; it < N2; ++it) {
            if (!P2::has(wave, it)) break;
            z[it] = chain<T, KQ>(At, greg[it], z[it]);
        }
#pragma unroll
        for (int it = 0; it < N2; ++it) {
            if (!P2::has(wave, it)) break;
            store<T, NM, NQ * NQ>(B, base2[it], lv2[it], g, chain<T, KQ>(Ai, z[it], zero));
        }
    }
    barrier();
    // j -> q along axis 1, lines (i, r): B -> A
    stage<T, 1, NQ, NM, E, NQ, NQ, NM, KQ>(Ai, B, A, wave, lane);
    barrier();
    // i -> p along axis 0, lines (q, r): A -> C
    stage<T, 0, NM, NM, E, NQ, NQ, NM, KQ>(Ai, A, C, wave, lane);
    barrier();
    for (int t = tid; t < nval * NM3; t += kBlock) {
        const int e = t / NM3, m = t - e * NM3;
        const int p = m % NM, q = (m / NM) % NM, r = m / (NM * NM);
        out[(size_t)(e0 + e) * NM3 + m] = C[e * SL + p + q * NQ + r * NQ * NQ];
    }
}

}  // namespace hexmf
}  // namespace nektar_amd
