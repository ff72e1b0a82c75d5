// Sum-factorised hex kernels for gfx950 (CDNA4), templated on the scalar
// type T (double or float), modes NM and points NQ per direction.
//
// Work decomposition: a workgroup holds EPB elements; each element owns
// NQ*NQ work-items.  A contraction along one direction is done by one
// work-item per 1-D "line" of the tensor: it reads the line's NM (or NQ)
// inputs once, keeps them in registers and produces all outputs of the line,
// so every LDS read feeds NQ fused multiply-adds.  The 1-D tables (basis,
// derivative matrix, weights) are wave-uniform and live in scalar registers
// (loaded with scalar loads from the operator's table buffer).  Three LDS
// slabs of NQ^3 values per element carry the tensor between directions.
//
// Layouts (element-major, the Collections contract):
//   coefficients  in[e*NM^3 + (r*NM + q)*NM + p]     (p fastest)
//   quadrature    u[e*NQ^3 + (k*NQ + j)*NQ + i]      (i fastest)
//   table         B[p*NQ + i] | D[l*NQ + i] | W[i]    (D[l*NQ+i] = l_l'(z_i))
//   Helmholtz geometry, regular : g[e*7 + c]
//       c = 0..5: J*m_{00,01,02,11,12,22}, c = 6: J
//   Helmholtz geometry, deformed: g[(e*7 + c)*NQ^3 + q]  (weights folded in)
//   IProduct  jac: regular jac[e], deformed jac[e*NQ^3 + q]
//   PhysDeriv df : regular df[e*9 + d], deformed df[(e*9 + d)*NQ^3 + q]
//                  df[j*3 + k] = d xi_k / d x_j
//
// Reference: library/MatrixFreeOps/BwdTransKernels.hpp:160-227,
// IProductKernels.hpp:270-350, PhysDerivKernels.hpp:184-345 and 436-507,
// HelmholtzKernels.hpp:389-696, Helmholtz.h:549-599.  The fused Helmholtz
// kernel uses IProduct(dbdata_d, v) == IProduct(bdata, D_d^T (W J v)) (exact
// because dbdata = D * bdata, Basis.cpp:493-495) so that the four inner
// products of the reference collapse into one transpose transform.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "contract.hpp"
#include "launch.hpp"

namespace nektar_amd {
namespace hex {

template <int NQ>
constexpr int elements_per_block()
{
    return (256 / (NQ * NQ)) > 0 ? (256 / (NQ * NQ)) : 1;
}

// LDS row pitch (in values) of the per-element slabs.  Rows of an even
// number of doubles put the lanes of a line-wise access (lane stride = one
// row) on the same few banks; an odd pitch spreads them over all banks.
template <int NQ>
constexpr int pitch()
{
    return NQ % 2 == 0 ? NQ + 1 : NQ;
}

// values per slab: NQ*NQ rows of pitch<NQ>() (rows index (k,j), (r,q), ...)
template <int NQ>
constexpr int slab()
{
    return NQ * NQ * pitch<NQ>();
}

// 1-D tables of one direction: B[p*NQ+i] = phi_p(z_i), D[l*NQ+i] =
// l'_l(z_i) (column-major D(i,l)), W[i], and transposed copies BT, DT so
// that every contraction reads one contiguous row per output.
template <typename T, int NM, int NQ>
struct Tab {
    static constexpr int OFF_D = NM * NQ, OFF_W = OFF_D + NQ * NQ, OFF_BT = OFF_W + NQ, OFF_DT = OFF_BT + NQ * NM;
    const T *__restrict__ base;
    __device__ __forceinline__ T B(int p, int i) const { return base[p * NQ + i]; }
    __device__ __forceinline__ T D(int l, int i) const { return base[OFF_D + l * NQ + i]; }
    __device__ __forceinline__ T W(int i) const { return base[OFF_W + i]; }
    // B(p, 0..NQ-1), B(0..NM-1, i), D(l, 0..NQ-1), D(0..NQ-1, i)
    __device__ __forceinline__ const_ptr<T> Brow(int p) const { return opaque(base + p * NQ); }
    __device__ __forceinline__ const_ptr<T> BTrow(int i) const { return opaque(base + OFF_BT + i * NM); }
    __device__ __forceinline__ const_ptr<T> Drow(int l) const { return opaque(base + OFF_D + l * NQ); }
    __device__ __forceinline__ const_ptr<T> DTrow(int i) const { return opaque(base + OFF_DT + i * NQ); }
};

// Coefficient fetch for an element stored contiguously (the Collections
// layout): x(p, q, r) = xin[(r*NM + q)*NM + p].
template <typename T, int NM>
struct LocalFetch {
    const T *__restrict__ xin;
    __device__ __forceinline__ T operator()(int p, int q, int r) const { return xin[(r * NM + q) * NM + p]; }
    // the NM contiguous coefficients of line (q, r) with 16-byte loads (the
    // line starts on a value boundary only, which global loads allow)
    __device__ __forceinline__ void line(int q, int r, T (&x)[NM]) const
    {
        constexpr int VW = 16 / sizeof(T);
        using V = T __attribute__((ext_vector_type(VW), aligned(sizeof(T))));
        const T *l = xin + (r * NM + q) * NM;
#pragma unroll
        for (int c = 0; c < NM / VW; ++c) {
            const V v = *reinterpret_cast<const V *>(l + c * VW);
#pragma unroll
            for (int w = 0; w < VW; ++w) x[c * VW + w] = v[w];
        }
#pragma unroll
        for (int p = (NM / VW) * VW; p < NM; ++p) x[p] = l[p];
    }
};

// Coefficient fetch straight from a global C0 vector of the structured
// n^3 mesh (GlobalToLocal fused into the operator, AssemblyMapCG.cpp:
// 2804-2827): 1-D node of local mode p in element ex is ex*(NM-1) + {0 for
// p=0, NM-1 for p=1, p-1 for bubbles}; nodes per direction N = n*(NM-1)+1.
template <typename T, int NM>
struct StructuredFetch {
    const T *__restrict__ g;
    int N;
    int bx, by, bz;  // ex*(NM-1), ...
    __device__ __forceinline__ static int off(int p) { return p == 0 ? 0 : (p == 1 ? NM - 1 : p - 1); }
    __device__ __forceinline__ T operator()(int p, int q, int r) const
    {
        return g[((size_t)(bz + off(r)) * N + (by + off(q))) * N + (bx + off(p))];
    }
};

// Coefficient fetch through a signed local-to-global map (GlobalToLocal
// with orientation signs fused into the operator, AssemblyMapCG.cpp:
// 2804-2827): the element's entries me[(r*NM + q)*NM + p] hold the global
// index m (sign +1) or ~m (sign -1).
template <typename T, int NM>
struct MappedFetch {
    const T *__restrict__ g;
    const int *__restrict__ me;
    __device__ __forceinline__ T operator()(int p, int q, int r) const
    {
        const int m = me[(r * NM + q) * NM + p];
        const T v = g[m < 0 ? ~m : m];
        return m < 0 ? -v : v;
    }
    // line (q, r): its NM contiguous map entries with 16-byte loads, then
    // the NM gathers
    __device__ __forceinline__ void line(int q, int r, T (&x)[NM]) const
    {
        using V = int __attribute__((ext_vector_type(4), aligned(4)));
        const int *l = me + (r * NM + q) * NM;
        int m[NM];
#pragma unroll
        for (int c = 0; c < NM / 4; ++c) {
            const V v = *reinterpret_cast<const V *>(l + c * 4);
#pragma unroll
            for (int w = 0; w < 4; ++w) m[c * 4 + w] = v[w];
        }
#pragma unroll
        for (int p = (NM / 4) * 4; p < NM; ++p) m[p] = l[p];
#pragma unroll
        for (int p = 0; p < NM; ++p) {
            const T v = g[m[p] < 0 ? ~m[p] : m[p]];
            x[p] = m[p] < 0 ? -v : v;
        }
    }
};

// fetches that read a whole line (q, r) at once
template <typename F>
struct HasLine : std::false_type {};
template <typename T, int NM>
struct HasLine<LocalFetch<T, NM>> : std::true_type {};
template <typename T, int NM>
struct HasLine<MappedFetch<T, NM>> : std::true_type {};

// LDS row pitch of an element's coefficients staged as NM*NM rows (r,q)
// of NM values: odd, so that lanes reading one row each (lane stride one
// row) fall on different banks.
template <int NM>
constexpr int cpitch()
{
    return NM % 2 == 0 ? NM + 1 : NM;
}
// the staged coefficients fit one slab for every NQ >= NM
template <int NM, int NQ>
constexpr bool coeffs_fit_slab()
{
    return NM * NM * cpitch<NM>() <= slab<NQ>();
}

// ---- forward transform: coefficients -> u column (k) of item (i,j) in
//      registers.  (Copying element-local coefficients to LDS first, for
//      contiguous loads, measured slower: BwdTrans P=6 56 against 43 us,
//      Helmholtz 107 against 100 us; the barrier it needs keeps the
//      coefficient loads from overlapping the first table rows.)
template <typename T, int NM, int NQ, typename Fetch>
__device__ __forceinline__ void bwd_to_column(const Tab<T, NM, NQ> &tab, const Fetch &xin, bool valid, int t,
                                              T *A, T *Bs, T (&u)[NQ])
{
    constexpr int P = pitch<NQ>();
    // F1: contract p -> i along lines (q, r)
    if (t < NM * NM) {
        const int q = t % NM, r = t / NM;
        T x[NM], o[NQ];
        if constexpr (HasLine<Fetch>::value) {
            if (valid) {
                xin.line(q, r, x);
            } else {
#pragma unroll
                for (int p = 0; p < NM; ++p) x[p] = T(0);
            }
        } else {
#pragma unroll
            for (int p = 0; p < NM; ++p) x[p] = valid ? xin(p, q, r) : T(0);
        }
        contract<false>(x, [&](int i) { return tab.BTrow(i); }, o);
#pragma unroll
        for (int i = 0; i < NQ; ++i) A[(r * NM + q) * P + i] = o[i];
    }
    __syncthreads();
    // F2: contract q -> j along lines (i, r)
    if (t < NQ * NM) {
        const int i = t % NQ, r = t / NQ;
        T a[NM], o[NQ];
#pragma unroll
        for (int q = 0; q < NM; ++q) a[q] = A[(r * NM + q) * P + i];
        contract<false>(a, [&](int j) { return tab.BTrow(j); }, o);
#pragma unroll
        for (int j = 0; j < NQ; ++j) Bs[(r * NQ + j) * P + i] = o[j];
    }
    __syncthreads();
    // F3: contract r -> k along lines (i, j), result stays in registers
    {
        const int i = t % NQ, j = t / NQ;
        T b[NM];
#pragma unroll
        for (int r = 0; r < NM; ++r) b[r] = Bs[(r * NQ + j) * P + i];
        contract<false>(b, [&](int k) { return tab.BTrow(k); }, u);
    }
}

// ---- transpose transform: y column (registers, at quadrature points,
//      already weighted) -> coefficients (global).  In FP64 the
//      coefficients are staged in LDS and written by consecutive work-items
//      (contiguous stores per wave): IProduct P=6 44 against 54 us, the
//      fused Helmholtz unchanged.  FP32 stores them from the registers
//      (staged: Helmholtz 61 against 57 us, IProduct 23.5 against 22 us).
//      Per-work-item 16-byte vector stores instead of the staging measured
//      equal for Helmholtz and slower for FP64 IProduct (52-60 against
//      43 us).
template <typename T, int NM, int NQ>
__device__ __forceinline__ void iprod_from_column(const Tab<T, NM, NQ> &tab, const T (&y)[NQ], bool valid,
                                                  int t, T *Bs, T *C, T *__restrict__ xout)
{
    constexpr int P = pitch<NQ>(), PM = cpitch<NM>();
    static_assert(coeffs_fit_slab<NM, NQ>(), "staged coefficients exceed the slab");
    {
        const int i = t % NQ, j = t / NQ;
        T o[NM];
        contract<false>(y, [&](int r) { return tab.Brow(r); }, o);
#pragma unroll
        for (int r = 0; r < NM; ++r) Bs[(r * NQ + j) * P + i] = o[r];
    }
    __syncthreads();
    if (t < NQ * NM) {
        const int i = t % NQ, r = t / NQ;
        T z[NQ], o[NM];
#pragma unroll
        for (int j = 0; j < NQ; ++j) z[j] = Bs[(r * NQ + j) * P + i];
        contract<false>(z, [&](int q) { return tab.Brow(q); }, o);
#pragma unroll
        for (int q = 0; q < NM; ++q) C[(r * NM + q) * P + i] = o[q];
    }
    __syncthreads();
    if (t < NM * NM) {
        const int q = t % NM, r = t / NM;
        T z[NQ], o[NM];
#pragma unroll
        for (int i = 0; i < NQ; ++i) z[i] = C[(r * NM + q) * P + i];
        contract<false>(z, [&](int p) { return tab.Brow(p); }, o);
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int p = 0; p < NM; ++p) Bs[(r * NM + q) * PM + p] = o[p];
        } else if (valid) {
#pragma unroll
            for (int p = 0; p < NM; ++p) xout[(r * NM + q) * NM + p] = o[p];
        }
    }
    if constexpr (sizeof(T) == 8) {
        __syncthreads();
        if (valid) {
            for (int m = t; m < NM * NM * NM; m += NQ * NQ) xout[m] = Bs[(m / NM) * PM + m % NM];
        }
    }
}

// ---- collocation derivatives in directions 0 and 1 from slab A (u),
//      written to slabs Bs (d0) and C (d1).  Direction 2 is in registers.
template <typename T, int NM, int NQ>
__device__ __forceinline__ void deriv01(const Tab<T, NM, NQ> &tab, int t, const T *A, T *Bs, T *C)
{
    constexpr int P = pitch<NQ>();
    const int l0 = t % NQ, k = t / NQ;
    {   // line (j = l0, k) along i
        T v[NQ], o[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = A[(k * NQ + l0) * P + l];
        contract<false>(v, [&](int i) { return tab.DTrow(i); }, o);
#pragma unroll
        for (int i = 0; i < NQ; ++i) Bs[(k * NQ + l0) * P + i] = o[i];
    }
    {   // line (i = l0, k) along j
        T v[NQ], o[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = A[(k * NQ + l) * P + l0];
        contract<false>(v, [&](int j) { return tab.DTrow(j); }, o);
#pragma unroll
        for (int j = 0; j < NQ; ++j) C[(k * NQ + j) * P + l0] = o[j];
    }
}

template <typename T, int NM, int NQ>
__device__ __forceinline__ void deriv2_column(const Tab<T, NM, NQ> &tab, const T (&u)[NQ], T (&d2)[NQ])
{
    contract<false>(u, [&](int k) { return tab.DTrow(k); }, d2);
}

// ===================================================================
// BwdTrans (BwdTransKernels.hpp:160-227)
template <typename T, int NM, int NQ>
__global__ __launch_bounds__(elements_per_block<NQ>() * NQ * NQ) void bwdtrans_kernel(
    int nelmt, const T *__restrict__ tabp, const T *__restrict__ in, T *__restrict__ out)
{
    constexpr int EPB = elements_per_block<NQ>(), TPE = NQ * NQ, NQ3 = NQ * NQ * NQ;
    __shared__ T lds[EPB][2][slab<NQ>()];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * EPB + le;
    const bool valid = e < nelmt;
    T u[NQ];
    bwd_to_column<T, NM, NQ>(tab, LocalFetch<T, NM>{in + (size_t)e * NM * NM * NM}, valid, t, lds[le][0],
                             lds[le][1], u);
    if (valid) {
        const int i = t % NQ, j = t / NQ;
        T *o = out + (size_t)e * NQ3;
#pragma unroll
        for (int k = 0; k < NQ; ++k) o[(k * NQ + j) * NQ + i] = u[k];
    }
}

// IProductWRTBase (IProductKernels.hpp:270-350, SCALE=APPEND=false)
template <typename T, int NM, int NQ, bool DEF>
__global__ __launch_bounds__(elements_per_block<NQ>() * NQ * NQ) void iproduct_kernel(
    int nelmt, const T *__restrict__ tabp, const T *__restrict__ jac, const T *__restrict__ in,
    T *__restrict__ out)
{
    constexpr int EPB = elements_per_block<NQ>(), TPE = NQ * NQ, NQ3 = NQ * NQ * NQ;
    __shared__ T lds[EPB][2][slab<NQ>()];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * EPB + le;
    const bool valid = e < nelmt;
    const int i = t % NQ, j = t / NQ;
    T y[NQ];
    const T wij = tab.W(i) * tab.W(j);
    const T *f = in + (size_t)e * NQ3;
    T jv = T(0);
    if (!DEF && valid) jv = jac[e];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const int c = (k * NQ + j) * NQ + i;
        T fv = valid ? f[c] : T(0);
        T jq = DEF ? (valid ? jac[(size_t)e * NQ3 + c] : T(0)) : jv;
        y[k] = fv * jq * (wij * tab.W(k));
    }
    iprod_from_column<T, NM, NQ>(tab, y, valid, t, lds[le][0], lds[le][1], out + (size_t)e * NM * NM * NM);
}

// PhysDeriv (PhysDerivKernels.hpp:436-507 then :326-340)
template <typename T, int NM, int NQ, bool DEF>
__global__ __launch_bounds__(elements_per_block<NQ>() * NQ * NQ) void physderiv_kernel(
    int nelmt, const T *__restrict__ tabp, const T *__restrict__ df, const T *__restrict__ in,
    T *__restrict__ out0, T *__restrict__ out1, T *__restrict__ out2)
{
    constexpr int EPB = elements_per_block<NQ>(), TPE = NQ * NQ, NQ3 = NQ * NQ * NQ;
    __shared__ T lds[EPB][3][slab<NQ>()];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * EPB + le;
    const bool valid = e < nelmt;
    const int i = t % NQ, j = t / NQ;
    T *A = lds[le][0], *Bs = lds[le][1], *C = lds[le][2];
    T u[NQ], d2[NQ];
    const T *f = in + (size_t)e * NQ3;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const int c = (k * NQ + j) * NQ + i;
        u[k] = valid ? f[c] : T(0);
        A[(k * NQ + j) * pitch<NQ>() + i] = u[k];
    }
    deriv2_column<T, NM, NQ>(tab, u, d2);
    __syncthreads();
    deriv01<T, NM, NQ>(tab, t, A, Bs, C);
    __syncthreads();
    if (!valid) return;
    T g[9];
    if (!DEF) {
#pragma unroll
        for (int d = 0; d < 9; ++d) g[d] = df[(size_t)e * 9 + d];
    }
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const int c = (k * NQ + j) * NQ + i;
        if (DEF) {
#pragma unroll
            for (int d = 0; d < 9; ++d) g[d] = df[((size_t)e * 9 + d) * NQ3 + c];
        }
        const int lc = (k * NQ + j) * pitch<NQ>() + i;
        const T a0 = Bs[lc], a1 = C[lc], a2 = d2[k];
        T o0 = a0 * g[0];
        o0 = fma(a1, g[1], o0);
        o0 = fma(a2, g[2], o0);
        T o1 = a0 * g[3];
        o1 = fma(a1, g[4], o1);
        o1 = fma(a2, g[5], o1);
        T o2 = a0 * g[6];
        o2 = fma(a1, g[7], o2);
        o2 = fma(a2, g[8], o2);
        out0[(size_t)e * NQ3 + c] = o0;
        out1[(size_t)e * NQ3 + c] = o1;
        out2[(size_t)e * NQ3 + c] = o2;
    }
}

// Second half shared by the fused Helmholtz apply and IProductWRTDerivBase:
// the weighted integrands of the three reference directions are in B
// (direction 0), A (direction 1) and the register column v2 (direction 2)
// at the quadrature points; y (register column) gets sum_d D_d^T of them
// (IProduct w.r.t. dbdata_d = D_d bdata_d, Basis.cpp:493-495), and the
// unweighted transpose transform of y gives the coefficients.  The caller's
// writes of A and B must be followed by no barrier yet (one is taken here).
template <typename T, int NM, int NQ, typename Store>
__device__ __forceinline__ void grad_transpose_to_coeffs(const Tab<T, NM, NQ> &tab, bool valid, int t,
                                                         T (&y)[NQ], const T (&v2)[NQ], T *A, T *B,
                                                         const Store &xout)
{
    constexpr int P = pitch<NQ>();
    const int i = t % NQ, j = t / NQ;
    const int l0 = i, k0 = j;
    // D2^T v2 (in registers): (D^T v)(k) = sum_l D(l,k) v(l), D(out,in)=tab.D(in,out)
    contract<true>(v2, [&](int k) { return tab.Drow(k); }, y);
    __syncthreads();
    {   // D0^T along i on line (j = l0, k = k0), in place in B
        T v[NQ], o[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = B[(k0 * NQ + l0) * P + l];
        contract<false>(v, [&](int ii) { return tab.Drow(ii); }, o);
#pragma unroll
        for (int ii = 0; ii < NQ; ++ii) B[(k0 * NQ + l0) * P + ii] = o[ii];
    }
    __syncthreads();
    {   // D1^T along j on line (i = l0, k = k0): A <- B + D1^T A, in place in A
        T v[NQ], o[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = A[(k0 * NQ + l) * P + l0];
#pragma unroll
        for (int jj = 0; jj < NQ; ++jj) o[jj] = B[(k0 * NQ + jj) * P + l0];
        ready_v(o);
        contract<true>(v, [&](int jj) { return tab.Drow(jj); }, o);
#pragma unroll
        for (int jj = 0; jj < NQ; ++jj) A[(k0 * NQ + jj) * P + l0] = o[jj];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; ++k) y[k] += A[(k * NQ + j) * P + i];
    iprod_from_column<T, NM, NQ>(tab, y, valid, t, B, A, xout);
}


// Fused Helmholtz y = lambda*M x + L x (Helmholtz.h:549-599)
// Body shared by the Collections kernel (element-local input) and the
// solver kernel (input gathered from the global vector).  Two LDS slabs
// per element: the direction-1 passes work in place on lines that belong
// to one work-item each, so the third slab of a straightforward staging is
// not needed (more elements resident per CU).
// ADV: LinearAdvectionDiffusionReaction (LinearAdvectionDiffusionReaction.h
// :600-658): the mass integrand u becomes u + V . grad_x u, with the
// advection written as sum_d c_d du/dxi_d (c_d = sum_j V_j df[j*3+d], per
// point, packed at set-up: adv[(e*3 + d)*Q^3 + q]).
template <typename T, int NM, int NQ, bool DEF, typename Fetch, bool ADV = false>
__device__ __forceinline__ void helmholtz_element(const Tab<T, NM, NQ> &tab, const Fetch &fetch, bool valid, int e,
                                                  int t, const T *__restrict__ geo, T lambda, T *A, T *B,
                                                  T *__restrict__ xout, const T *__restrict__ adv = nullptr)
{
    constexpr int NQ3 = NQ * NQ * NQ, P = pitch<NQ>();
    // column (i, j) of this work-item; as a line index: (l0, k0) = (i, j)
    const int i = t % NQ, j = t / NQ;
    const int l0 = i, k0 = j;

    T u[NQ];
    bwd_to_column<T, NM, NQ>(tab, fetch, valid, t, A, B, u);
#pragma unroll
    for (int k = 0; k < NQ; ++k) A[(k * NQ + j) * P + i] = u[k];
    __syncthreads();
    {   // d0 along i on line (j = l0, k = k0): A -> B
        T v[NQ], o[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = A[(k0 * NQ + l0) * P + l];
        contract<false>(v, [&](int ii) { return tab.DTrow(ii); }, o);
#pragma unroll
        for (int ii = 0; ii < NQ; ++ii) B[(k0 * NQ + l0) * P + ii] = o[ii];
    }
    __syncthreads();
    // advection coefficients of this column: loads issued here, so that
    // their latency hides behind the direction-1 and direction-2 passes
    T cadv[3][ADV ? NQ : 1];
    if constexpr (ADV) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const T *ac = adv + (size_t)e * 3 * NQ3 + (k * NQ + j) * NQ + i;
            cadv[0][k] = valid ? ac[0] : T(0);
            cadv[1][k] = valid ? ac[NQ3] : T(0);
            cadv[2][k] = valid ? ac[2 * NQ3] : T(0);
        }
    }
    {   // d1 along j on line (i = l0, k = k0), in place in A
        T v[NQ], o[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = A[(k0 * NQ + l) * P + l0];
        contract<false>(v, [&](int jj) { return tab.DTrow(jj); }, o);
#pragma unroll
        for (int jj = 0; jj < NQ; ++jj) A[(k0 * NQ + jj) * P + l0] = o[jj];
    }
    __syncthreads();
    // direction-2 derivative of the register column (formed after the
    // barrier, next to its use, so no table row is held across barriers)
    T d2[NQ];
    deriv2_column<T, NM, NQ>(tab, u, d2);

    // pointwise: metric, Jacobian, weights, mass term (d0 in B, d1 in A)
    T y[NQ], v2[NQ];
    {
        T g[7];
        if (!DEF) {
#pragma unroll
            for (int c = 0; c < 7; ++c) g[c] = valid ? geo[(size_t)e * 7 + c] : T(0);
        }
        const T wij = tab.W(i) * tab.W(j);
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const int c = (k * NQ + j) * NQ + i;
            T w = T(1);
            if (DEF) {
#pragma unroll
                for (int cc = 0; cc < 7; ++cc) g[cc] = valid ? geo[((size_t)e * 7 + cc) * NQ3 + c] : T(0);
            } else {
                w = wij * tab.W(k);
            }
            const int lc = (k * NQ + j) * P + i;
            const T a0 = B[lc], a1 = A[lc], a2 = d2[k];
            T t0 = g[0] * a0;
            t0 = fma(g[1], a1, t0);
            t0 = fma(g[2], a2, t0);
            T t1 = g[1] * a0;
            t1 = fma(g[3], a1, t1);
            t1 = fma(g[4], a2, t1);
            T t2 = g[2] * a0;
            t2 = fma(g[4], a1, t2);
            t2 = fma(g[5], a2, t2);
            if (!DEF) {
                t0 *= w;
                t1 *= w;
                t2 *= w;
            }
            B[lc] = t0;
            A[lc] = t1;
            v2[k] = t2;
            T m = u[k];
            if constexpr (ADV) {
                m = fma(cadv[0][k], a0, m);
                m = fma(cadv[1][k], a1, m);
                m = fma(cadv[2][k], a2, m);
            }
            y[k] = (lambda * g[6]) * (DEF ? m : w * m);
        }
    }
    grad_transpose_to_coeffs<T, NM, NQ>(tab, valid, t, y, v2, A, B, xout);
}

// IProductWRTDerivBase (IProductWRTDerivBase.h:751-788, hex pointwise step
// IProductWRTDerivBaseKernels.hpp:194-247): out = sum_d IProduct(dbdata_d,
// t_d) with t_d = sum_j df[j*3+d] in_j.  Packed geometry: regular
// g[e*9 + j*3+d] = J df[j*3+d] (weights applied here); deformed
// g[(e*9 + j*3+d)*Q^3 + q] = W_q J_q df_q[j*3+d].
template <typename T, int NM, int NQ, bool DEF>
__global__ __launch_bounds__(elements_per_block<NQ>() * NQ * NQ) void iproduct_deriv_kernel(
    int nelmt, const T *__restrict__ tabp, const T *__restrict__ geo, const T *__restrict__ in0,
    const T *__restrict__ in1, const T *__restrict__ in2, T *__restrict__ out)
{
    constexpr int EPB = elements_per_block<NQ>(), TPE = NQ * NQ, NQ3 = NQ * NQ * NQ, P = pitch<NQ>();
    __shared__ T lds[EPB][2][slab<NQ>()];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * EPB + le;
    const bool valid = e < nelmt;
    const int i = t % NQ, j = t / NQ;
    T *A = lds[le][0], *B = lds[le][1];
    T g[9];
    if (!DEF) {
#pragma unroll
        for (int c = 0; c < 9; ++c) g[c] = valid ? geo[(size_t)e * 9 + c] : T(0);
    }
    const T wij = tab.W(i) * tab.W(j);
    T y[NQ], v2[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const int c = (k * NQ + j) * NQ + i;
        if (DEF) {
#pragma unroll
            for (int cc = 0; cc < 9; ++cc) g[cc] = valid ? geo[((size_t)e * 9 + cc) * NQ3 + c] : T(0);
        }
        const size_t q = (size_t)e * NQ3 + c;
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
        if (!DEF) {
            const T w = wij * tab.W(k);
            t0 *= w;
            t1 *= w;
            t2 *= w;
        }
        const int lc = (k * NQ + j) * P + i;
        B[lc] = t0;
        A[lc] = t1;
        v2[k] = t2;
        y[k] = T(0);
    }
    grad_transpose_to_coeffs<T, NM, NQ>(tab, valid, t, y, v2, A, B, out + (size_t)e * NM * NM * NM);
}

// Tile shapes (elements per workgroup) instantiated for the Helmholtz
// kernels; the autotuner times them per (shape, order, precision).
template <int NQ>
constexpr int tile_fill()
{
    return elements_per_block<NQ>();
}
template <typename T, int NQ, int EPB>
constexpr bool tile_fits()
{
    return EPB * NQ * NQ <= 1024 && EPB * 2 * slab<NQ>() * (int)sizeof(T) <= 65536;
}

// LinearAdvectionDiffusionReaction y = lambda*IProduct(u + V.grad u) + L x,
// one element per workgroup
template <typename T, int NM, int NQ, bool DEF>
__global__ __launch_bounds__(NQ * NQ) void adv_diff_react_kernel(int nelmt, const T *__restrict__ tabp,
                                                                 const T *__restrict__ geo,
                                                                 const T *__restrict__ adv, T lambda,
                                                                 const T *__restrict__ in, T *__restrict__ out)
{
    __shared__ T lds[2][slab<NQ>()];
    const Tab<T, NM, NQ> tab{tabp};
    const int e = blockIdx.x, t = threadIdx.x;
    const bool valid = e < nelmt;
    using F = LocalFetch<T, NM>;
    helmholtz_element<T, NM, NQ, DEF, F, true>(tab, F{in + (size_t)e * NM * NM * NM}, valid, e, t, geo, lambda,
                                               lds[0], lds[1], out + (size_t)e * NM * NM * NM, adv);
}

// Fused Helmholtz y = lambda*M x + L x (Helmholtz.h:549-599)
template <typename T, int NM, int NQ, bool DEF, int EPB>
__global__ __launch_bounds__(EPB * NQ * NQ) void helmholtz_kernel(int nelmt, const T *__restrict__ tabp,
                                                                  const T *__restrict__ geo, T lambda,
                                                                  const T *__restrict__ in, T *__restrict__ out)
{
    constexpr int TPE = NQ * NQ;
    __shared__ T lds[EPB][2][slab<NQ>()];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * EPB + le;
    const bool valid = e < nelmt;
    helmholtz_element<T, NM, NQ, DEF>(tab, LocalFetch<T, NM>{in + (size_t)e * NM * NM * NM}, valid, e, t, geo,
                                      lambda, lds[le][0], lds[le][1], out + (size_t)e * NM * NM * NM);
}

// Solver form on a general mesh: x gathered through the signed map from
// the global vector, element results written element-major
template <typename T, int NM, int NQ, bool DEF, int EPB>
__global__ __launch_bounds__(EPB * NQ * NQ) void helmholtz_mapped_kernel(int nelmt, const T *__restrict__ tabp,
                                                                         const T *__restrict__ geo, T lambda,
                                                                         const T *__restrict__ xg,
                                                                         const int *__restrict__ menc,
                                                                         T *__restrict__ out)
{
    constexpr int TPE = NQ * NQ;
    __shared__ T lds[EPB][2][slab<NQ>()];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * EPB + le;
    const bool valid = e < nelmt;
    const MappedFetch<T, NM> f{xg, menc + (size_t)(valid ? e : 0) * NM * NM * NM};
    helmholtz_element<T, NM, NQ, DEF>(tab, f, valid, e, t, geo, lambda, lds[le][0], lds[le][1],
                                      out + (size_t)e * NM * NM * NM);
}

// Grouped form: several operators of the same kind in one launch; block b
// belongs to the item whose [block0, next block0) range holds it.
template <typename T, int NM, int NQ, bool DEF, int EPB>
__global__ __launch_bounds__(EPB * NQ * NQ) void helmholtz_grouped_kernel(int count,
                                                                          const GroupItem<T> *__restrict__ items,
                                                                          const T *__restrict__ tabp)
{
    constexpr int TPE = NQ * NQ;
    __shared__ T lds[EPB][2][slab<NQ>()];
    const int b = blockIdx.x;
    int it = 0;
    while (it + 1 < count && b >= items[it + 1].block0) ++it;
    const GroupItem<T> I = items[it];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = (b - I.block0) * EPB + le;
    const bool valid = e < I.nelmt;
    helmholtz_element<T, NM, NQ, DEF>(tab, LocalFetch<T, NM>{I.in + (size_t)e * NM * NM * NM}, valid, e, t, I.geo,
                                      I.lambda, lds[le][0], lds[le][1], I.out + (size_t)e * NM * NM * NM);
}

// Solver form: x gathered from the global vector of the structured n^3
// mesh, element results written element-major (assembled afterwards).
template <typename T, int NM, int NQ, bool DEF, int EPB>
__global__ __launch_bounds__(EPB * NQ * NQ) void helmholtz_gather_kernel(int n, int N, const T *__restrict__ tabp,
                                                                         const T *__restrict__ geo, T lambda,
                                                                         const T *__restrict__ xg,
                                                                         T *__restrict__ out,
                                                                         const int *__restrict__ done)
{
    if (done && *done) return;
    constexpr int TPE = NQ * NQ;
    __shared__ T lds[EPB][2][slab<NQ>()];
    const Tab<T, NM, NQ> tab{tabp};
    const int le = threadIdx.x / TPE, t = threadIdx.x % TPE;
    const int e = blockIdx.x * EPB + le;
    const int nelmt = n * n * n;
    const bool valid = e < nelmt;
    const int ex = e % n, ey = (e / n) % n, ez = e / (n * n);
    const StructuredFetch<T, NM> f{xg, N, ex * (NM - 1), ey * (NM - 1), ez * (NM - 1)};
    helmholtz_element<T, NM, NQ, DEF>(tab, f, valid, e, t, geo, lambda, lds[le][0], lds[le][1],
                                      out + (size_t)e * NM * NM * NM);
}

}  // namespace hex
}  // namespace nektar_amd
