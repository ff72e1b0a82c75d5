// Sum-factorised tetrahedron kernels for gfx950, collapsed coordinates.
//
// Points: NQ0 = NQ (GLL, Modified_A), NQ1 = NQ2 = NQ-1 (Gauss-Radau-Jacobi
// alpha=1 / alpha=2 with Modified_B / Modified_C), the MODIFIED defaults.
// Coefficients are ordered (p, q, r) with r fastest and triangular bounds
// q < NM-p, r < NM-p-q (BwdTransKernels.hpp:244-266).  Each element owns
// NQ0*NQ1 work-items; item (i, j) holds the k-column of the element's
// quadrature values in registers.  Triangular contractions run one
// work-item per (p,q) pair / (p,k) line, so every LDS read feeds a full
// line of fused multiply-adds.  The singular top/bottom vertex and edge
// corrections (BwdTransKernels.hpp:296-324, IProductKernels.hpp:456-513)
// are separable and are evaluated from the p=0 / p=1 partial sums.
//
// Table buffer (T): B0[NM*NQ0] B1[NMB*NQ1] B2[NC*NQ2] D0[NQ0^2] D1[NQ1^2]
//   D2[NQ2^2] W0[NQ0] W1[NQ1] W2[NQ2] H0[NQ0] H1[NQ1] H2[NQ1] H3[NQ2]
//   B0T[NQ0*NM] D0T[NQ0^2] D1T[NQ1^2] D2T[NQ2^2]
//   (W1, W2 carry the 0.5 / 0.25 Radau pre-scaling, Operator.hpp:600-616;
//    H* are the half-space factors of HelmholtzKernels.hpp:1368-1389;
//    the *T tables are transposed copies: B0T[i*NM + p] = B0[p*NQ0 + i],
//    DdT[i*NQd + l] = Dd[l*NQd + i].)
//   then, from a multiple of 4, zero-padded B1 and B2 rows of pitch RP that
//   each kernel stages into LDS (Dims::oB1P, Dims::LTAB).
// Geometry: IProduct jac[e] | jac[e*NQT+q]; PhysDeriv df[e*9+d] |
//   df[(e*9+d)*NQT+q]; Helmholtz geo[e*10+c] | geo[(e*10+c)*NQT+q] with
//   c = 0..8: df, c = 9: J.
#pragma once

#include <hip/hip_runtime.h>

#include "contract.hpp"
#include "launch.hpp"

namespace nektar_amd {
namespace tet {

template <int NM, int NQ>
struct Dims {
    static constexpr int NQ0 = NQ, NQ1 = NQ - 1, NQ2 = NQ - 1;
    static constexpr int NQT = NQ0 * NQ1 * NQ2;
    static constexpr int NMB = NM * (NM + 1) / 2;
    static constexpr int NC = NM * (NM + 1) * (NM + 2) / 6;
    static constexpr int oB0 = 0, oB1 = oB0 + NM * NQ0, oB2 = oB1 + NMB * NQ1;
    static constexpr int oD0 = oB2 + NC * NQ2, oD1 = oD0 + NQ0 * NQ0, oD2 = oD1 + NQ1 * NQ1;
    static constexpr int oW0 = oD2 + NQ2 * NQ2, oW1 = oW0 + NQ0, oW2 = oW1 + NQ1;
    static constexpr int oH0 = oW2 + NQ2, oH1 = oH0 + NQ0, oH2 = oH1 + NQ1, oH3 = oH2 + NQ1;
    static constexpr int oB0T = oH3 + NQ2, oD0T = oB0T + NQ0 * NM, oD1T = oD0T + NQ0 * NQ0;
    static constexpr int oD2T = oD1T + NQ1 * NQ1;
    static constexpr int TABSIZE = oD2T + NQ2 * NQ2;
    // Zero-padded copies of B1 and B2 after the tables above, staged into
    // LDS by every kernel that runs the collapsed contractions: rows of
    // pitch RP (a multiple of 4, so rows start on 16-byte boundaries in
    // either precision), each table followed by NM zero rows so that the
    // triangular loops can run to NM without predication.
    static constexpr int RP = (NQ - 1 + 3) / 4 * 4;
    static constexpr int NB1R = NMB + NM, NB2R = NC + NM;
    static constexpr int oB1P = (TABSIZE + 3) / 4 * 4;
    static constexpr int LTAB = (NB1R + NB2R) * RP;
};

// LDS slab of an element: quadrature values at (k*NQ1 + j)*PI + i, then
// 2*NQ2 values of vertex sums; the same slab holds the (pair, k) and
// (k, j, p) partial sums of the transforms (SIZE >= NMB*NQ2, NQ1*NQ2*NM).
// In FP64 the i-lines are padded to an odd pitch, so that work-items
// walking their own lines fall in different banks; in FP32 the padding
// costs a workgroup per CU and is measured slower, so it is left out.
template <typename T, int NM, int NQ>
struct Slab {
    using Dm = Dims<NM, NQ>;
    static constexpr int PI = (sizeof(T) == 8 && Dm::NQ0 % 2 == 0) ? Dm::NQ0 + 1 : Dm::NQ0;
    static constexpr int NQTL = Dm::NQ1 * Dm::NQ2 * PI;
    static constexpr int SIZE = NQTL + 2 * Dm::NQ2;
    static_assert((Dm::NMB + NM - 1) * Dm::NQ2 <= SIZE, "unpredicated pair reads must stay inside the slab");
};

// Work-items per element (TPE) of a launch.  Packed: TPE = NQ0*NQ1, so
// that elements share waves and every stage ends at a workgroup barrier.
// Wave-aligned: TPE rounded up to a multiple of 64; at TPE = 64 every
// element is one wave, its stages exchange data through LDS inside the
// wave only, and the waves of a workgroup never wait for each other (the
// work-items past NQ0*NQ1 idle in the per-column stages).
template <int NM, int NQ, int TPE_>
struct Tile {
    using Dm = Dims<NM, NQ>;
    static constexpr int TPE = TPE_;
    static constexpr int NCOL = Dm::NQ0 * Dm::NQ1;
    static_assert(TPE >= NCOL, "a tile needs one work-item per (i,j) column");
    static constexpr int EPB = 256 / TPE > 0 ? 256 / TPE : 1;
    static constexpr int BLOCK = EPB * TPE;
    static constexpr bool WAVE = TPE == 64;
};

// LDS ordering of an element's stages.  Inside one wave the LDS operations
// complete in issue order, so a wave-aligned element needs no hardware
// barrier, only a compiler fence that keeps its loads after its stores.
template <bool WAVE>
__device__ __forceinline__ void esync()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// (p, q) pair index t -> p, q (pairs ordered p-major, q < NM - p)
template <int NM>
__device__ __forceinline__ void pair_of(int t, int &p, int &q)
{
    p = 0;
    int rem = t;
#pragma unroll
    for (int pp = 0; pp < NM; ++pp) {
        if (rem >= NM - p) {
            rem -= NM - p;
            ++p;
        }
    }
    q = rem;
}
// first coefficient index of the (p, q) r-run
template <int NM>
__device__ __forceinline__ int coeff_base(int p, int q)
{
    int b = 0;
#pragma unroll
    for (int pp = 0; pp < NM; ++pp)
        if (pp < p) b += (NM - pp) * (NM - pp + 1) / 2;
    return b + q * (NM - p) - q * (q - 1) / 2;
}
// Modified_B mode index of (p, q)
template <int NM>
__device__ __forceinline__ int modeb(int p, int q)
{
    return p * NM - p * (p - 1) / 2 + q;
}

template <typename T, int NM, int NQ>
struct Tab {
    using Dm = Dims<NM, NQ>;
    const T *__restrict__ b;
    // contiguous table rows through opaque() (contract.hpp): one row per
    // output of a contraction, loaded one ahead by contract()
    __device__ __forceinline__ const_ptr<T> B0row(int p) const { return opaque(b + Dm::oB0 + p * Dm::NQ0); }
    __device__ __forceinline__ const_ptr<T> B0Trow(int i) const { return opaque(b + Dm::oB0T + i * NM); }
    __device__ __forceinline__ const_ptr<T> D0row(int i) const { return opaque(b + Dm::oD0 + i * Dm::NQ0); }
    __device__ __forceinline__ const_ptr<T> D1row(int j) const { return opaque(b + Dm::oD1 + j * Dm::NQ1); }
    __device__ __forceinline__ const_ptr<T> D2row(int k) const { return opaque(b + Dm::oD2 + k * Dm::NQ2); }
    __device__ __forceinline__ const_ptr<T> D0Trow(int i) const { return opaque(b + Dm::oD0T + i * Dm::NQ0); }
    __device__ __forceinline__ const_ptr<T> D1Trow(int j) const { return opaque(b + Dm::oD1T + j * Dm::NQ1); }
    __device__ __forceinline__ const_ptr<T> D2Trow(int k) const { return opaque(b + Dm::oD2T + k * Dm::NQ2); }
    __device__ __forceinline__ T B0(int p, int i) const { return b[Dm::oB0 + p * Dm::NQ0 + i]; }
    __device__ __forceinline__ T B1(int m, int j) const { return b[Dm::oB1 + m * Dm::NQ1 + j]; }
    __device__ __forceinline__ T B2(int m, int k) const { return b[Dm::oB2 + m * Dm::NQ2 + k]; }
    __device__ __forceinline__ T D0(int l, int i) const { return b[Dm::oD0 + l * Dm::NQ0 + i]; }
    __device__ __forceinline__ T D1(int l, int j) const { return b[Dm::oD1 + l * Dm::NQ1 + j]; }
    __device__ __forceinline__ T D2(int l, int k) const { return b[Dm::oD2 + l * Dm::NQ2 + k]; }
    __device__ __forceinline__ T W0(int i) const { return b[Dm::oW0 + i]; }
    __device__ __forceinline__ T W1(int j) const { return b[Dm::oW1 + j]; }
    __device__ __forceinline__ T W2(int k) const { return b[Dm::oW2 + k]; }
    __device__ __forceinline__ T H0(int i) const { return b[Dm::oH0 + i]; }
    __device__ __forceinline__ T H1(int j) const { return b[Dm::oH1 + j]; }
    __device__ __forceinline__ T H2(int j) const { return b[Dm::oH2 + j]; }
    __device__ __forceinline__ T H3(int k) const { return b[Dm::oH3 + k]; }
};

// The staged B1/B2 rows in LDS.  The collapsed contractions index these
// tables by a lane-dependent mode (the (p,q) pair or the p of a line), so
// they are read from LDS rather than as scattered global loads.  In FP64
// a row that does not fill whole 16-byte units is padded to an odd number
// of units (LP values), so that the row reads of neighbouring work-items
// start in different banks; rows that fill their units exactly, and FP32
// rows, keep the plain pitch (both measured faster).
// With G the Modified_C rows (NC + NM of them, the larger table) are read
// from the padded rows of the table buffer in global memory (pitch RP,
// cached) and only the B1 rows are staged, which leaves LDS for more
// resident workgroups.
template <typename T, int NM, int NQ, bool G = false>
struct LTab {
    using Dm = Dims<NM, NQ>;
    static constexpr int VW = 16 / sizeof(T);
    static constexpr int UNITS = (Dm::NQ2 + VW - 1) / VW;
    static constexpr bool PAD = sizeof(T) == 8 && Dm::NQ2 % VW != 0 && UNITS % 2 == 0;
    static constexpr int LP = (PAD ? UNITS + 1 : UNITS) * VW;
    // with G only B1 is staged, and without its last zero row: the pair
    // stages read rows pb + q <= modeb(NM-1, 0) + NM-1 = NB1R - 2
    static constexpr int NROWS = G ? Dm::NB1R - 1 : Dm::NB1R + Dm::NB2R;
    static constexpr int SIZE = NROWS * LP;
    const T *l;  // staged rows in LDS
    const T *g;  // the table buffer
    __device__ __forceinline__ const T *B1(int m) const { return l + m * LP; }
    __device__ __forceinline__ const T *B2(int m) const
    {
        if constexpr (G) return g + Dm::oB1P + (Dm::NB1R + m) * Dm::RP;
        return l + (Dm::NB1R + m) * LP;
    }
};

// The Helmholtz and inner-product kernels read the Modified_C rows from
// global memory from Q = 10 on in FP64 and Q = 11 in FP32: at P = 9 that
// takes FP64 Helmholtz from 59 to 41 KB of LDS per workgroup (2 -> 3
// waves per SIMD) and its time down by an eighth; at lower orders, and in
// BwdTrans, staging measured faster.
template <typename T, int NQ>
constexpr bool kB2Global = NQ >= (sizeof(T) == 8 ? 10 : 11);
// Those FP64 kernels then fit 4 workgroups per CU in LDS, and ask for the
// registers of 4 waves per SIMD (<= 128 VGPRs) to match.
// IProductWRTDerivBase gains from the global rows where the Helmholtz
// kernel loses: in FP32 from Q = 10 (69.5 -> 60 us at P = 8, 16,384
// tets), in FP64 from Q = 8 (P = 6: 131 -> 125 us, P = 7: 85 -> 81 us).
template <typename T, int NQ>
constexpr bool kB2GlobalIPD = kB2Global<T, NQ> || (sizeof(T) == 4 && NQ == 10) || (sizeof(T) == 8 && NQ >= 8);
template <typename T, int NQ>
constexpr int kTetWaves = (sizeof(T) == 8 && kB2Global<T, NQ>) ? 4 : 1;

// Work split of the (p,q) pair stages over TPE work-items.  With room for
// two work-items per pair, the forward r-contraction gives each a half of
// the k range (KW values from k = kg*KW, whole 16-byte units inside the
// staged row) and the transpose k-contraction gives each a half of the r
// range (RW rows from r = rg*RW); otherwise one work-item per pair does it
// all.
template <typename T, int NM, int NQ, int TPE>
struct PairSplit {
    using Dm = Dims<NM, NQ>;
    using L = LTab<T, NM, NQ>;
    static constexpr int KW2 = ((Dm::NQ2 + 1) / 2 + L::VW - 1) / L::VW * L::VW;
    static constexpr bool TWO = 2 * Dm::NMB <= TPE;
    static constexpr int KG = (TWO && 2 * KW2 <= L::LP) ? 2 : 1;
    static constexpr int KW = KG == 2 ? KW2 : Dm::NQ2;
    static constexpr int RG = TWO ? 2 : 1;
    static constexpr int RW = (NM + RG - 1) / RG;
    // reads of a B2 row stay inside its pitch in the table buffer too
    static_assert((KG == 2 ? 2 * KW : (Dm::NQ2 + L::VW - 1) / L::VW * L::VW) <= Dm::RP,
                  "B2 row reads must stay inside the table row");
};

// N consecutive values of a staged row, read as 16-byte vectors (rows are
// 16-byte aligned and padded to LP >= N rounded up to the vector width).
template <typename T, int N>
__device__ __forceinline__ void load_row(const T *row, T (&b)[N])
{
    constexpr int VW = 16 / sizeof(T);
    using V = T __attribute__((ext_vector_type(VW)));
#pragma unroll
    for (int c = 0; c < (N + VW - 1) / VW; ++c) {
        const V v = *reinterpret_cast<const V *>(row + c * VW);
#pragma unroll
        for (int w = 0; w < VW; ++w)
            if (c * VW + w < N) b[c * VW + w] = v[w];
    }
}

// Copy the padded B1/B2 rows of the table buffer (pitch RP) into LDS
// (pitch LP), whole block.  Units past RP are the zero padding.
template <typename T, int NM, int NQ, bool G = false>
__device__ __forceinline__ void stage_tables(const T *__restrict__ tabp, T *lt)
{
    using Dm = Dims<NM, NQ>;
    using L = LTab<T, NM, NQ, G>;
    constexpr int VW = L::VW;
    using V = T __attribute__((ext_vector_type(VW)));
    static_assert(Dm::RP % VW == 0 && Dm::oB1P % VW == 0, "table rows must be whole vectors");
    constexpr int UG = Dm::RP / VW, UL = L::LP / VW;
    const V *src = reinterpret_cast<const V *>(tabp + Dm::oB1P);
    V *dst = reinterpret_cast<V *>(lt);
    for (int x = threadIdx.x; x < L::NROWS * UL; x += blockDim.x) {
        const int r = x / UL, c = x % UL;
        V v = V(T(0));
        if (c < UG) v = src[r * UG + c];
        dst[x] = v;
    }
    __syncthreads();
}

// ---- BwdTrans: coefficients (global) -> k-column u of item (i,j)
// Stage 1 (pairs): fpq[pair][k]      -> S1[pair*NQ2 + k]
// Stage 2 (p,k)  : fp[k][j][p]       -> S2[(k*NQ1 + j)*NM + p]
// Stage 3 (j,k)  : u(i,j,k) + corr.  -> S3[(k*NQ1 + j)*PI + i]
// The triangular loops run over all NM rows: the coefficients (or partial
// sums) beyond a pair's length are zero, and the rows they meet are real
// modes of the next pair or the zero padding, so no lane needs a branch.
template <typename T, int NM, int NQ, int TPE, bool G>
__device__ __forceinline__ void bwd_to_column(const Tab<T, NM, NQ> &tab, const LTab<T, NM, NQ, G> &lt,
                                              const T *__restrict__ xin, bool valid, int t, T *S1, T *S2, T *S3,
                                              T (&u)[NQ - 1])
{
    using Dm = Dims<NM, NQ>;
    using Ti = Tile<NM, NQ, TPE>;
    using S = PairSplit<T, NM, NQ, TPE>;
    constexpr int NQ0 = Dm::NQ0, NQ1 = Dm::NQ1, NQ2 = Dm::NQ2;
    if (t < Dm::NMB * S::KG) {
        const int pr = t / S::KG, kg = t % S::KG;
        int p, q;
        pair_of<NM>(pr, p, q);
        const int base = coeff_base<NM>(p, q);
        const int len = NM - p - q;
        T x[NM];
#pragma unroll
        for (int r = 0; r < NM; ++r) x[r] = T(0);
        if (valid) {
#pragma unroll
            for (int r = 0; r < NM; ++r) {
                const T xv = xin[min(base + r, Dm::NC - 1)];
                x[r] = r < len ? xv : T(0);
            }
        }
        T s[S::KW];
#pragma unroll
        for (int k = 0; k < S::KW; ++k) s[k] = T(0);
#pragma unroll
        for (int r = 0; r < NM; ++r) {
            T b[S::KW];
            load_row(lt.B2(base + r) + kg * S::KW, b);
#pragma unroll
            for (int k = 0; k < S::KW; ++k) s[k] = fma(x[r], b[k], s[k]);
        }
#pragma unroll
        for (int kk = 0; kk < S::KW; ++kk) {
            const int k = kg * S::KW + kk;
            if (S::KG == 1 || k < NQ2) S1[pr * NQ2 + k] = s[kk];
        }
    }
    esync<Ti::WAVE>();
    if (t < NM * NQ2) {
        const int p = t % NM, k = t / NM;
        const int pb = modeb<NM>(p, 0);
        // rows pb+q < NMB+NM-1 stay inside the slab, so the reads need no
        // branch; what they find beyond the pair is discarded
        T a[NM];
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            const T av = S1[(pb + q) * NQ2 + k];
            a[q] = q < NM - p ? av : T(0);
        }
        T s[NQ1];
#pragma unroll
        for (int j = 0; j < NQ1; ++j) s[j] = T(0);
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            T b[NQ1];
            load_row(lt.B1(pb + q), b);
#pragma unroll
            for (int j = 0; j < NQ1; ++j) s[j] = fma(a[q], b[j], s[j]);
        }
#pragma unroll
        for (int j = 0; j < NQ1; ++j) S2[(k * NQ1 + j) * NM + p] = s[j];
    }
    esync<Ti::WAVE>();
    // singular-mode coefficients (broadcast reads)
    T c_top = T(0), c_bot = T(0), c_edge[NM > 2 ? NM - 2 : 1];
    if (valid) {
        c_top = xin[1];
        c_bot = xin[NM];
#pragma unroll
        for (int r = 1; r < NM - 1; ++r) c_edge[r - 1] = xin[NM + r];
    }
    if (t < NQ1 * NQ2) {
        const int j = t % NQ1, k = t / NQ1;
        T f[NM];
#pragma unroll
        for (int p = 0; p < NM; ++p) f[p] = S2[(k * NQ1 + j) * NM + p];
        // the corrections are separable in i: the top vertex adds
        // c_top B2_1 (B0_0 B1_1 + B0_1 (B1_0 + B1_1)), the bottom vertex and
        // the singular edge add B0_1 B1_1 edge(k), so they fold into the
        // p = 0 and p = 1 partial sums
        const T b10 = lt.B1(0)[j], b11 = lt.B1(1)[j];
        T edge = c_bot * lt.B2(0)[k];
#pragma unroll
        for (int r = 1; r < NM - 1; ++r) edge = fma(c_edge[r - 1], lt.B2(r + 1)[k], edge);
        const T tb = c_top * lt.B2(1)[k];
        f[0] = fma(tb, b11, f[0]);
        f[1] = fma(tb, b10 + b11, fma(edge, b11, f[1]));
        T o[NQ0];
        contract<false>(f, [&](int i) { return tab.B0Trow(i); }, o);
#pragma unroll
        for (int i = 0; i < NQ0; ++i) S3[(k * NQ1 + j) * Slab<T, NM, NQ>::PI + i] = o[i];
    }
    esync<Ti::WAVE>();
    if (t < Ti::NCOL) {
        const int i = t % NQ0, j = t / NQ0;
#pragma unroll
        for (int k = 0; k < NQ2; ++k) u[k] = S3[(k * NQ1 + j) * Slab<T, NM, NQ>::PI + i];
    } else {
#pragma unroll
        for (int k = 0; k < NQ2; ++k) u[k] = T(0);
    }
}

// ---- IProduct (unweighted) from the k-column y of item (i,j) (already
// multiplied by W*J) -> coefficients (global).
// Stage I1 (j,k): s[p] = sum_i B0(p,i) y  -> S2[(k*NQ1 + j)*NM + p]
// corrections  : TOP(k), BOT(k)          -> S3[k], S3[NQ2 + k]
// Stage I2 (p,k): sum_j B1 s            -> S1[pair*NQ2 + k]
// Stage I3 (pairs): sum_k B2            -> out
template <typename T, int NM, int NQ, int TPE, bool G>
__device__ __forceinline__ void iprod_from_column(const Tab<T, NM, NQ> &tab, const LTab<T, NM, NQ, G> &lt,
                                                  const T (&y)[NQ - 1], bool valid, int t, T *S1, T *S2, T *S3,
                                                  T *__restrict__ xout)
{
    using Dm = Dims<NM, NQ>;
    using Ti = Tile<NM, NQ, TPE>;
    using S = PairSplit<T, NM, NQ, TPE>;
    constexpr int NQ0 = Dm::NQ0, NQ1 = Dm::NQ1, NQ2 = Dm::NQ2;
    if (t < Ti::NCOL) {
        const int i = t % NQ0, j = t / NQ0;
#pragma unroll
        for (int k = 0; k < NQ2; ++k) S3[(k * NQ1 + j) * Slab<T, NM, NQ>::PI + i] = y[k];
    }
    esync<Ti::WAVE>();
    if (t < NQ1 * NQ2) {
        const int j = t % NQ1, k = t / NQ1;
        T v[NQ0];
#pragma unroll
        for (int i = 0; i < NQ0; ++i) v[i] = S3[(k * NQ1 + j) * Slab<T, NM, NQ>::PI + i];
        T o[NM];
        contract<false>(v, [&](int p) { return tab.B0row(p); }, o);
#pragma unroll
        for (int p = 0; p < NM; ++p) S2[(k * NQ1 + j) * NM + p] = o[p];
    }
    esync<Ti::WAVE>();
    if (t < NQ2) {
        const int k = t;
        T top = T(0), bot = T(0);
#pragma unroll
        for (int j = 0; j < NQ1; ++j) {
            const T s0 = S2[(k * NQ1 + j) * NM + 0], s1 = S2[(k * NQ1 + j) * NM + 1];
            const T b10 = tab.B1(0, j), b11 = tab.B1(1, j);
            top = fma(s0, b11, top);
            top = fma(s1, b10 + b11, top);
            bot = fma(s1, b11, bot);
        }
        S3[Slab<T, NM, NQ>::NQTL + k] = top;
        S3[Slab<T, NM, NQ>::NQTL + NQ2 + k] = bot;
    }
    if (t < NM * NQ2) {
        const int p = t % NM, k = t / NM;
        const int pb = modeb<NM>(p, 0);
        T v[NQ1];
#pragma unroll
        for (int j = 0; j < NQ1; ++j) v[j] = S2[(k * NQ1 + j) * NM + p];
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            T b[NQ1];
            load_row(lt.B1(pb + q), b);
            T s = v[0] * b[0];
#pragma unroll
            for (int j = 1; j < NQ1; ++j) s = fma(v[j], b[j], s);
            if (q < NM - p) S1[(pb + q) * NQ2 + k] = s;
        }
    }
    esync<Ti::WAVE>();
    if (t < Dm::NMB * S::RG && valid) {
        const int pr = t / S::RG, rg = t % S::RG;
        int p, q;
        pair_of<NM>(pr, p, q);
        const int base = coeff_base<NM>(p, q);
        const int len = NM - p - q;
        T v[NQ2];
#pragma unroll
        for (int k = 0; k < NQ2; ++k) v[k] = S1[pr * NQ2 + k];
        const bool p00 = (p == 0 && q == 0), p01 = (p == 0 && q == 1);
        // rows rg*RW .. rg*RW+RW-1 (past len: the next pair's rows or the
        // zero rows after the table, read but not stored)
#pragma unroll
        for (int rr = 0; rr < S::RW; ++rr) {
            const int r = rg * S::RW + rr;
            T b[NQ2];
            load_row(lt.B2(base + r), b);
            T s = v[0] * b[0];
#pragma unroll
            for (int k = 1; k < NQ2; ++k) s = fma(v[k], b[k], s);
            if (r < len) {
                if (p00 && r == 1) {
                    // top vertex: sum_k TOP(k) B2_1(k)
#pragma unroll
                    for (int k = 0; k < NQ2; ++k) s = fma(S3[Slab<T, NM, NQ>::NQTL + k], tab.B2(1, k), s);
                }
                if (p01 && r < NM - 1) {
                    // bottom vertex (r = 0) and singular edge (r >= 1): B2_{r+1}, B2_0
                    const int m2 = (r == 0) ? 0 : r + 1;
#pragma unroll
                    for (int k = 0; k < NQ2; ++k) s = fma(S3[Slab<T, NM, NQ>::NQTL + NQ2 + k], tab.B2(m2, k), s);
                }
                xout[base + r] = s;
            }
        }
    }
}

// ===================================================================
// Kernel entry points: TPE work-items per element (Tile), workgroups of
// Tile<NM, NQ, TPE>::BLOCK work-items.  Work-items past the NQ0*NQ1
// columns compute on the last column's indices (every read in bounds) and
// store nothing of it.
template <typename T, int NM, int NQ, int TPE>
__global__ __launch_bounds__((Tile<NM, NQ, TPE>::BLOCK)) void bwdtrans_kernel(int nelmt, const T *__restrict__ tabp,
                                                                           const T *__restrict__ in,
                                                                           T *__restrict__ out)
{
    using Dm = Dims<NM, NQ>;
    using Ti = Tile<NM, NQ, TPE>;
    __shared__ T lds[Ti::EPB][2][Slab<T, NM, NQ>::SIZE];
    __shared__ __align__(16) T ltab[LTab<T, NM, NQ>::SIZE];
    stage_tables<T, NM, NQ>(tabp, ltab);
    const Tab<T, NM, NQ> tab{tabp};
    const LTab<T, NM, NQ> lt{ltab, tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * Ti::EPB + le;
    const bool valid = e < nelmt;
    T u[NQ - 1];
    bwd_to_column<T, NM, NQ, TPE>(tab, lt, in + (size_t)e * Dm::NC, valid, t, lds[le][0], lds[le][1], lds[le][0],
                                  u);
    if (valid && t < Ti::NCOL) {
        const int i = t % Dm::NQ0, j = t / Dm::NQ0;
#pragma unroll
        for (int k = 0; k < Dm::NQ2; ++k) out[(size_t)e * Dm::NQT + (k * Dm::NQ1 + j) * Dm::NQ0 + i] = u[k];
    }
}

template <typename T, int NM, int NQ, bool DEF, int TPE>
__global__ __launch_bounds__((Tile<NM, NQ, TPE>::BLOCK), (kTetWaves<T, NQ>)) void iproduct_kernel(int nelmt, const T *__restrict__ tabp,
                                                                           const T *__restrict__ jac,
                                                                           const T *__restrict__ in,
                                                                           T *__restrict__ out)
{
    using Dm = Dims<NM, NQ>;
    using Ti = Tile<NM, NQ, TPE>;
    __shared__ T lds[Ti::EPB][2][Slab<T, NM, NQ>::SIZE];
    __shared__ __align__(16) T ltab[LTab<T, NM, NQ, kB2Global<T, NQ>>::SIZE];
    stage_tables<T, NM, NQ, kB2Global<T, NQ>>(tabp, ltab);
    const Tab<T, NM, NQ> tab{tabp};
    const LTab<T, NM, NQ, kB2Global<T, NQ>> lt{ltab, tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * Ti::EPB + le;
    const bool valid = e < nelmt;
    const int tc = min(t, Ti::NCOL - 1);
    const int i = tc % Dm::NQ0, j = tc / Dm::NQ0;
    const T wij = tab.W0(i) * tab.W1(j);
    T jv = T(0);
    if (!DEF && valid) jv = jac[e];
    T y[NQ - 1];
#pragma unroll
    for (int k = 0; k < Dm::NQ2; ++k) {
        const int c = (k * Dm::NQ1 + j) * Dm::NQ0 + i;
        const T fv = valid ? in[(size_t)e * Dm::NQT + c] : T(0);
        const T jq = DEF ? (valid ? jac[(size_t)e * Dm::NQT + c] : T(0)) : jv;
        y[k] = fv * jq * (wij * tab.W2(k));
    }
    iprod_from_column<T, NM, NQ, TPE>(tab, lt, y, valid, t, lds[le][0], lds[le][1], lds[le][0],
                                      out + (size_t)e * Dm::NC);
}

// collocation derivatives along i (lines (j,k)) and j (lines (i,k)) from A
template <typename T, int NM, int NQ>
__device__ __forceinline__ void deriv01(const Tab<T, NM, NQ> &tab, int t, const T *A, T *Bs, T *C)
{
    using Dm = Dims<NM, NQ>;
    constexpr int NQ0 = Dm::NQ0, NQ1 = Dm::NQ1, NQ2 = Dm::NQ2;
    if (t < NQ1 * NQ2) {
        const int j = t % NQ1, k = t / NQ1;
        T v[NQ0], o[NQ0];
#pragma unroll
        for (int l = 0; l < NQ0; ++l) v[l] = A[(k * NQ1 + j) * Slab<T, NM, NQ>::PI + l];
        contract<false>(v, [&](int i) { return tab.D0Trow(i); }, o);
#pragma unroll
        for (int i = 0; i < NQ0; ++i) Bs[(k * NQ1 + j) * Slab<T, NM, NQ>::PI + i] = o[i];
    }
    if (t < NQ0 * NQ2) {
        const int i = t % NQ0, k = t / NQ0;
        T v[NQ1], o[NQ1];
#pragma unroll
        for (int l = 0; l < NQ1; ++l) v[l] = A[(k * NQ1 + l) * Slab<T, NM, NQ>::PI + i];
        contract<false>(v, [&](int j) { return tab.D1Trow(j); }, o);
#pragma unroll
        for (int j = 0; j < NQ1; ++j) C[(k * NQ1 + j) * Slab<T, NM, NQ>::PI + i] = o[j];
    }
}

template <typename T, int NM, int NQ>
__device__ __forceinline__ void deriv2_column(const Tab<T, NM, NQ> &tab, const T (&u)[NQ - 1], T (&d2)[NQ - 1])
{
    contract<false>(u, [&](int k) { return tab.D2Trow(k); }, d2);
}

// PhysDeriv: PhysDerivKernels.hpp:436-507 (tensor) + :198-259 (collapsed
// chain rule) + :326-340 (df)
template <typename T, int NM, int NQ, bool DEF, int TPE>
__global__ __launch_bounds__((Tile<NM, NQ, TPE>::BLOCK)) void physderiv_kernel(
    int nelmt, const T *__restrict__ tabp, const T *__restrict__ df, const T *__restrict__ in,
    T *__restrict__ out0, T *__restrict__ out1, T *__restrict__ out2)
{
    using Dm = Dims<NM, NQ>;
    using Ti = Tile<NM, NQ, TPE>;
    __shared__ T lds[Ti::EPB][3][Slab<T, NM, NQ>::SIZE];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * Ti::EPB + le;
    const bool valid = e < nelmt, col = t < Ti::NCOL;
    const int tc = min(t, Ti::NCOL - 1);
    const int i = tc % Dm::NQ0, j = tc / Dm::NQ0;
    T *A = lds[le][0], *Bs = lds[le][1], *C = lds[le][2];
    T u[NQ - 1], d2[NQ - 1];
#pragma unroll
    for (int k = 0; k < Dm::NQ2; ++k) {
        const int c = (k * Dm::NQ1 + j) * Dm::NQ0 + i;
        u[k] = valid ? in[(size_t)e * Dm::NQT + c] : T(0);
        if (col) A[(k * Dm::NQ1 + j) * Slab<T, NM, NQ>::PI + i] = u[k];
    }
    deriv2_column<T, NM, NQ>(tab, u, d2);
    esync<Ti::WAVE>();
    deriv01<T, NM, NQ>(tab, t, A, Bs, C);
    esync<Ti::WAVE>();
    if (!valid || !col) return;
    T g[9];
    if (!DEF) {
#pragma unroll
        for (int d = 0; d < 9; ++d) g[d] = df[(size_t)e * 9 + d];
    }
    const T h0 = tab.H0(i), h1 = tab.H1(j), h2 = tab.H2(j);
#pragma unroll
    for (int k = 0; k < Dm::NQ2; ++k) {
        const int c = (k * Dm::NQ1 + j) * Dm::NQ0 + i;
        if (DEF) {
#pragma unroll
            for (int d = 0; d < 9; ++d) g[d] = df[((size_t)e * 9 + d) * Dm::NQT + c];
        }
        const T h3 = tab.H3(k);
        const int cl = (k * Dm::NQ1 + j) * Slab<T, NM, NQ>::PI + i;
        const T a0 = (h2 * h3) * Bs[cl];
        const T w0 = h0 * a0;
        const T w1 = h3 * C[cl];
        const T a1 = w0 + w1;
        const T a2 = fma(w1, h1, w0) + d2[k];
        T o0 = a0 * g[0];
        o0 = fma(a1, g[1], o0);
        o0 = fma(a2, g[2], o0);
        T o1 = a0 * g[3];
        o1 = fma(a1, g[4], o1);
        o1 = fma(a2, g[5], o1);
        T o2 = a0 * g[6];
        o2 = fma(a1, g[7], o2);
        o2 = fma(a2, g[8], o2);
        out0[(size_t)e * Dm::NQT + c] = o0;
        out1[(size_t)e * Dm::NQT + c] = o1;
        out2[(size_t)e * Dm::NQT + c] = o2;
    }
}

// Fused Helmholtz (Helmholtz.h:549-599 with DiffusionCoeffTetKernel,
// HelmholtzKernels.hpp:700-913).
// Second half shared by the fused Helmholtz apply and IProductWRTDerivBase:
// the weighted integrands of the collapsed directions 0 (in Y), 1 (in X)
// and 2 (register column v2) at the quadrature points; y (register column)
// gets sum_d D_d^T of them, then the unweighted transpose transform (with
// the collapsed-vertex corrections) gives the coefficients.  D0^T runs in
// place on the (j,k) lines of Y and D1^T in place on the (i,k) lines of X
// (each line belongs to one work-item, which reads it whole before writing
// it), so the element needs two slabs.  The caller's writes of X and Y
// must not be followed by a barrier yet.
template <typename T, int NM, int NQ, int TPE, bool G>
__device__ __forceinline__ void grad_transpose_to_coeffs(const Tab<T, NM, NQ> &tab, const LTab<T, NM, NQ, G> &lt,
                                                         bool valid, int t, int e, T (&y)[NQ - 1],
                                                         const T (&v2)[NQ - 1], T *X, T *Y, T *__restrict__ out)
{
    using Dm = Dims<NM, NQ>;
    using Ti = Tile<NM, NQ, TPE>;
    constexpr int NQ0 = Dm::NQ0, NQ1 = Dm::NQ1, NQ2 = Dm::NQ2;
    const int tc = min(t, Ti::NCOL - 1);
    const int i = tc % NQ0, j = tc / NQ0;
    contract<true>(v2, [&](int k) { return tab.D2row(k); }, y);
    esync<Ti::WAVE>();
    if (t < NQ1 * NQ2) {  // D0^T along i, lines (j,k) of Y, in place
        const int jj = t % NQ1, k = t / NQ1;
        T *line = Y + (k * NQ1 + jj) * Slab<T, NM, NQ>::PI;
        T v[NQ0], o[NQ0];
#pragma unroll
        for (int l = 0; l < NQ0; ++l) v[l] = line[l];
        contract<false>(v, [&](int ii) { return tab.D0row(ii); }, o);
#pragma unroll
        for (int ii = 0; ii < NQ0; ++ii) line[ii] = o[ii];
    }
    if (t < NQ0 * NQ2) {  // D1^T along j, lines (i,k) of X, in place
        const int ii = t % NQ0, k = t / NQ0;
        T v[NQ1], o[NQ1];
#pragma unroll
        for (int l = 0; l < NQ1; ++l) v[l] = X[(k * NQ1 + l) * Slab<T, NM, NQ>::PI + ii];
        contract<false>(v, [&](int jj) { return tab.D1row(jj); }, o);
#pragma unroll
        for (int jj = 0; jj < NQ1; ++jj) X[(k * NQ1 + jj) * Slab<T, NM, NQ>::PI + ii] = o[jj];
    }
    esync<Ti::WAVE>();
#pragma unroll
    for (int k = 0; k < NQ2; ++k) {
        const int c = (k * NQ1 + j) * Slab<T, NM, NQ>::PI + i;
        y[k] += Y[c] + X[c];
    }
    iprod_from_column<T, NM, NQ, TPE>(tab, lt, y, valid, t, X, Y, X, out + (size_t)e * Dm::NC);
}

// One workgroup of the fused Helmholtz apply: elements blk*EPB ..
// blk*EPB+EPB-1 of an operator with nelmt elements.
template <typename T, int NM, int NQ, bool DEF, int TPE, bool ADV = false>
__device__ __forceinline__ void helmholtz_block(int blk, int nelmt, const T *__restrict__ tabp,
                                                const T *__restrict__ geo, T lambda, const T *__restrict__ in,
                                                T *__restrict__ out, T (*lds)[2][Slab<T, NM, NQ>::SIZE], const T *ltab,
                                                const T *__restrict__ adv = nullptr)
{
    using Dm = Dims<NM, NQ>;
    using Ti = Tile<NM, NQ, TPE>;
    constexpr int NQ0 = Dm::NQ0, NQ1 = Dm::NQ1, NQ2 = Dm::NQ2, NQT = Dm::NQT;
    const Tab<T, NM, NQ> tab{tabp};
    const LTab<T, NM, NQ, kB2Global<T, NQ>> lt{ltab, tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blk * Ti::EPB + le;
    const bool valid = e < nelmt, col = t < Ti::NCOL;
    const int tc = min(t, Ti::NCOL - 1);
    const int i = tc % NQ0, j = tc / NQ0;
    T *X = lds[le][0], *Y = lds[le][1];

    // X holds the quadrature values after the transform; Y <- D0 X along
    // the (j,k) lines, then D1 in place on the (i,k) lines of X
    T u[NQ - 1], d2[NQ - 1];
    bwd_to_column<T, NM, NQ, TPE>(tab, lt, in + (size_t)e * Dm::NC, valid, t, X, Y, X, u);
    deriv2_column<T, NM, NQ>(tab, u, d2);
    if (t < NQ1 * NQ2) {
        const int jj = t % NQ1, k = t / NQ1;
        const T *line = X + (k * NQ1 + jj) * Slab<T, NM, NQ>::PI;
        T v[NQ0], o[NQ0];
#pragma unroll
        for (int l = 0; l < NQ0; ++l) v[l] = line[l];
        contract<false>(v, [&](int ii) { return tab.D0Trow(ii); }, o);
#pragma unroll
        for (int ii = 0; ii < NQ0; ++ii) Y[(k * NQ1 + jj) * Slab<T, NM, NQ>::PI + ii] = o[ii];
    }
    esync<Ti::WAVE>();
    // advection coefficients of this column: loads issued here, so that
    // their latency hides behind the direction-1 pass
    T cadv[3][ADV ? NQ2 : 1];
    if constexpr (ADV) {
#pragma unroll
        for (int k = 0; k < NQ2; ++k) {
            const T *ac = adv + (size_t)e * 3 * NQT + (k * NQ1 + j) * NQ0 + i;
            cadv[0][k] = valid ? ac[0] : T(0);
            cadv[1][k] = valid ? ac[NQT] : T(0);
            cadv[2][k] = valid ? ac[2 * NQT] : T(0);
        }
    }
    if (t < NQ0 * NQ2) {
        const int ii = t % NQ0, k = t / NQ0;
        T v[NQ1], o[NQ1];
#pragma unroll
        for (int l = 0; l < NQ1; ++l) v[l] = X[(k * NQ1 + l) * Slab<T, NM, NQ>::PI + ii];
        contract<false>(v, [&](int jj) { return tab.D1Trow(jj); }, o);
#pragma unroll
        for (int jj = 0; jj < NQ1; ++jj) X[(k * NQ1 + jj) * Slab<T, NM, NQ>::PI + ii] = o[jj];
    }
    esync<Ti::WAVE>();

    T y[NQ - 1], v2[NQ - 1];
    const T h0 = tab.H0(i), h1 = tab.H1(j), h2 = tab.H2(j);
    if constexpr (!DEF) {
        // Affine element: df and J are constants, and the metric terms of
        // the collapsed directions factor into a per-column part and the
        // direction-2 factors h3(k), W2(k):
        //   A_c = df[3c] + h0 (df[3c+1] + df[3c+2]),  E_c = df[3c+1] + h1 df[3c+2]
        //   g0 = h2^2 h3^2 sum A^2,  g3 = h2 h3^2 sum A E,  g4 = h2 h3 sum df[3c+2] A
        //   g1 = h3^2 sum E^2,       g5 = h3 sum df[3c+2] E, g2 = sum df[3c+2]^2
        // so each point needs the three products with (a0, a1, a2) only.
        T g[10];
#pragma unroll
        for (int c = 0; c < 10; ++c) g[c] = valid ? geo[(size_t)e * 10 + c] : T(0);
        T ga = T(0), gb = T(0), gc = T(0), gd = T(0), ge = T(0), gf = T(0);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T A = fma(h0, g[3 * c + 1] + g[3 * c + 2], g[3 * c]);
            const T E = fma(h1, g[3 * c + 2], g[3 * c + 1]);
            ga = fma(A, A, ga);
            gb = fma(A, E, gb);
            gc = fma(g[3 * c + 2], A, gc);
            gd = fma(E, E, gd);
            ge = fma(g[3 * c + 2], E, ge);
            gf = fma(g[3 * c + 2], g[3 * c + 2], gf);
        }
        const T cw = tab.W0(i) * tab.W1(j) * g[9];
        const T cwh2 = cw * h2;
        const T q00 = cwh2 * h2 * ga, q01 = cwh2 * gb, q02 = cwh2 * gc;
        const T q11 = cw * gd, q12 = cw * ge, q22 = cw * gf, lm = lambda * cw;
#pragma unroll
        for (int k = 0; k < NQ2; ++k) {
            const int cl = (k * NQ1 + j) * Slab<T, NM, NQ>::PI + i;
            const T w2 = tab.W2(k), h3 = tab.H3(k);
            const T s1 = w2 * h3, s2 = s1 * h3;
            const T a0 = Y[cl], a1 = X[cl], a2 = d2[k];
            T m0 = q00 * a0;
            m0 = fma(q01, a1, m0);
            T m1 = q01 * a0;
            m1 = fma(q11, a1, m1);
            T m2 = q02 * a0;
            m2 = fma(q12, a1, m2);
            const T n1 = s1 * a2, n0 = w2 * a2;
            const T o0 = fma(q02, n1, s2 * m0);
            const T o1 = fma(q12, n1, s2 * m1);
            v2[k] = fma(q22, n0, s1 * m2);
            if (col) {
                Y[cl] = o0;
                X[cl] = o1;
            }
            T m = u[k];
            if constexpr (ADV) {
                m = fma(cadv[0][k], a0, m);
                m = fma(cadv[1][k], a1, m);
                m = fma(cadv[2][k], a2, m);
            }
            y[k] = (lm * w2) * m;
        }
    } else {
        T g[10];
        const T wij = tab.W0(i) * tab.W1(j);
#pragma unroll
        for (int k = 0; k < NQ2; ++k) {
            const int c = (k * NQ1 + j) * NQ0 + i, cl = (k * NQ1 + j) * Slab<T, NM, NQ>::PI + i;
#pragma unroll
            for (int cc = 0; cc < 10; ++cc) g[cc] = valid ? geo[((size_t)e * 10 + cc) * NQT + c] : T(0);
            const T h3 = tab.H3(k);
            const T h2h3 = h2 * h3, h1h3 = h1 * h3, h0h2h3 = h0 * h2h3;
            T t1 = h0h2h3 * (g[1] + g[2]);
            t1 = fma(g[0], h2h3, t1);
            T t2 = h0h2h3 * (g[4] + g[5]);
            t2 = fma(g[3], h2h3, t2);
            T t3 = h0h2h3 * (g[7] + g[8]);
            t3 = fma(g[6], h2h3, t3);
            T t4 = g[1] * h3;
            t4 = fma(g[2], h1h3, t4);
            T t5 = g[4] * h3;
            t5 = fma(g[5], h1h3, t5);
            T t6 = g[7] * h3;
            t6 = fma(g[8], h1h3, t6);
            T g0 = t1 * t1;
            g0 = fma(t2, t2, g0);
            g0 = fma(t3, t3, g0);
            T g4 = g[2] * t1;
            g4 = fma(g[5], t2, g4);
            g4 = fma(g[8], t3, g4);
            T g3 = t1 * t4;
            g3 = fma(t2, t5, g3);
            g3 = fma(t3, t6, g3);
            T g1 = t4 * t4;
            g1 = fma(t5, t5, g1);
            g1 = fma(t6, t6, g1);
            T g5 = g[2] * t4;
            g5 = fma(g[5], t5, g5);
            g5 = fma(g[8], t6, g5);
            T g2 = g[2] * g[2];
            g2 = fma(g[5], g[5], g2);
            g2 = fma(g[8], g[8], g2);
            const T a0 = Y[cl], a1 = X[cl], a2 = d2[k];
            const T wj = wij * tab.W2(k) * g[9];
            T o0 = g0 * a0;
            o0 = fma(g3, a1, o0);
            o0 = fma(g4, a2, o0);
            T o1 = g3 * a0;
            o1 = fma(g1, a1, o1);
            o1 = fma(g5, a2, o1);
            T o2 = g4 * a0;
            o2 = fma(g5, a1, o2);
            o2 = fma(g2, a2, o2);
            if (col) {
                Y[cl] = wj * o0;
                X[cl] = wj * o1;
            }
            v2[k] = wj * o2;
            T m = u[k];
            if constexpr (ADV) {
                m = fma(cadv[0][k], a0, m);
                m = fma(cadv[1][k], a1, m);
                m = fma(cadv[2][k], a2, m);
            }
            y[k] = (lambda * wj) * m;
        }
    }
    grad_transpose_to_coeffs<T, NM, NQ, TPE>(tab, lt, valid, t, e, y, v2, X, Y, out);
}

template <typename T, int NM, int NQ, bool DEF, int TPE>
__global__ __launch_bounds__((Tile<NM, NQ, TPE>::BLOCK), (kTetWaves<T, NQ>)) void helmholtz_kernel(int nelmt, const T *__restrict__ tabp,
                                                                            const T *__restrict__ geo, T lambda,
                                                                            const T *__restrict__ in,
                                                                            T *__restrict__ out)
{
    using Ti = Tile<NM, NQ, TPE>;
    __shared__ T lds[Ti::EPB][2][Slab<T, NM, NQ>::SIZE];
    __shared__ __align__(16) T ltab[LTab<T, NM, NQ, kB2Global<T, NQ>>::SIZE];
    stage_tables<T, NM, NQ, kB2Global<T, NQ>>(tabp, ltab);
    helmholtz_block<T, NM, NQ, DEF, TPE>(blockIdx.x, nelmt, tabp, geo, lambda, in, out, lds, ltab);
}

// LinearAdvectionDiffusionReaction (LinearAdvectionDiffusionReaction.h
// :600-658, tet advection LinearAdvectionDiffusionReactionKernels.hpp
// :279-386): the mass integrand u becomes u + V . grad_x u, written as
// sum_d c_d du/deta_d over the collapsed derivatives, with
// c_0 = h2 h3 (a_0 + h0 (a_1 + a_2)), c_1 = h3 (a_1 + h1 a_2), c_2 = a_2,
// a_d = sum_j V_j df[j*3+d], packed per point at set-up:
// adv[(e*3 + d)*NQT + q].
template <typename T, int NM, int NQ, bool DEF, int TPE>
__global__ __launch_bounds__((Tile<NM, NQ, TPE>::BLOCK)) void adv_diff_react_kernel(
    int nelmt, const T *__restrict__ tabp, const T *__restrict__ geo, const T *__restrict__ adv, T lambda,
    const T *__restrict__ in, T *__restrict__ out)
{
    using Ti = Tile<NM, NQ, TPE>;
    __shared__ T lds[Ti::EPB][2][Slab<T, NM, NQ>::SIZE];
    __shared__ __align__(16) T ltab[LTab<T, NM, NQ, kB2Global<T, NQ>>::SIZE];
    stage_tables<T, NM, NQ, kB2Global<T, NQ>>(tabp, ltab);
    helmholtz_block<T, NM, NQ, DEF, TPE, true>(blockIdx.x, nelmt, tabp, geo, lambda, in, out, lds, ltab, adv);
}

// Grouped form: several operators of the same kind in one launch; block b
// belongs to the item whose [block0, next block0) range holds it.
template <typename T, int NM, int NQ, bool DEF, int TPE>
__global__ __launch_bounds__((Tile<NM, NQ, TPE>::BLOCK), (kTetWaves<T, NQ>)) void helmholtz_grouped_kernel(
    int count, const GroupItem<T> *__restrict__ items, const T *__restrict__ tabp)
{
    using Ti = Tile<NM, NQ, TPE>;
    __shared__ T lds[Ti::EPB][2][Slab<T, NM, NQ>::SIZE];
    __shared__ __align__(16) T ltab[LTab<T, NM, NQ, kB2Global<T, NQ>>::SIZE];
    stage_tables<T, NM, NQ, kB2Global<T, NQ>>(tabp, ltab);
    const int b = blockIdx.x;
    int it = 0;
    while (it + 1 < count && b >= items[it + 1].block0) ++it;
    const GroupItem<T> I = items[it];
    helmholtz_block<T, NM, NQ, DEF, TPE>(b - I.block0, I.nelmt, tabp, I.geo, I.lambda, I.in, I.out, lds, ltab);
}

// IProductWRTDerivBase for tets (IProductWRTDerivBase.h:751-788, pointwise
// step IProductWRTDerivBaseKernels.hpp:249-335): t_d = sum_j df[j*3+d] in_j,
// mapped to the collapsed directions t0 <- f0 (t0 + f1 (t1 + t2)),
// t1 <- f2 (t1 + f3 t2) with f0 = h2 h3, f1 = h0, f2 = h3, f3 = h1, then
// weighted by W J.  Geometry packed as for the Helmholtz kernel (df, J).
template <typename T, int NM, int NQ, bool DEF, int TPE>
__global__ __launch_bounds__((Tile<NM, NQ, TPE>::BLOCK), (kTetWaves<T, NQ>)) void iproduct_deriv_kernel(
    int nelmt, const T *__restrict__ tabp, const T *__restrict__ geo, const T *__restrict__ in0,
    const T *__restrict__ in1, const T *__restrict__ in2, T *__restrict__ out)
{
    using Dm = Dims<NM, NQ>;
    using Ti = Tile<NM, NQ, TPE>;
    constexpr int NQ0 = Dm::NQ0, NQ1 = Dm::NQ1, NQ2 = Dm::NQ2, NQT = Dm::NQT;
    __shared__ T lds[Ti::EPB][2][Slab<T, NM, NQ>::SIZE];
    __shared__ __align__(16) T ltab[LTab<T, NM, NQ, kB2GlobalIPD<T, NQ>>::SIZE];
    stage_tables<T, NM, NQ, kB2GlobalIPD<T, NQ>>(tabp, ltab);
    const Tab<T, NM, NQ> tab{tabp};
    const LTab<T, NM, NQ, kB2GlobalIPD<T, NQ>> lt{ltab, tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * Ti::EPB + le;
    const bool valid = e < nelmt, col = t < Ti::NCOL;
    const int tc = min(t, Ti::NCOL - 1);
    const int i = tc % NQ0, j = tc / NQ0;
    T *X = lds[le][0], *Y = lds[le][1];
    T g[10];
    if (!DEF) {
#pragma unroll
        for (int c = 0; c < 10; ++c) g[c] = valid ? geo[(size_t)e * 10 + c] : T(0);
    }
    const T h0 = tab.H0(i), h1 = tab.H1(j), h2 = tab.H2(j);
    const T wij = tab.W0(i) * tab.W1(j);
    T y[NQ - 1], v2[NQ - 1];
#pragma unroll
    for (int k = 0; k < NQ2; ++k) {
        const int c = (k * NQ1 + j) * NQ0 + i, cl = (k * NQ1 + j) * Slab<T, NM, NQ>::PI + i;
        if (DEF) {
#pragma unroll
            for (int cc = 0; cc < 10; ++cc) g[cc] = valid ? geo[((size_t)e * 10 + cc) * NQT + c] : T(0);
        }
        const size_t q = (size_t)e * NQT + c;
        const T f0 = valid ? in0[q] : T(0), f1 = valid ? in1[q] : T(0), f2 = valid ? in2[q] : T(0);
        T t0 = g[0] * f0;
        t0 = fma(g[3], f1, t0);
        t0 = fma(g[6], f2, t0);
        T t1 = g[1] * f0;
        t1 = fma(g[4], f1, t1);
        t1 = fma(g[7], f2, t1);
        T t2 = g[2] * f0;
        t2 = fma(g[5], f1, t2);
        t2 = fma(g[8], f2, t2);
        const T h3 = tab.H3(k);
        t0 = fma(t1 + t2, h0, t0) * (h2 * h3);
        t1 = fma(t2, h1, t1) * h3;
        const T wj = wij * tab.W2(k) * g[9];
        if (col) {
            Y[cl] = wj * t0;
            X[cl] = wj * t1;
        }
        v2[k] = wj * t2;
        y[k] = T(0);
    }
    grad_transpose_to_coeffs<T, NM, NQ, TPE>(tab, lt, valid, t, e, y, v2, X, Y, out);
}

}  // namespace tet
}  // namespace nektar_amd
