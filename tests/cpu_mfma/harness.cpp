// Host-side check of the matrix-core hex Helmholtz body
// (nektar_amd/csrc/hex_mfma_block.hpp) before it runs on a GPU: the same
// code compiled as C++ with AddressSanitizer and UndefinedBehaviorSanitizer.
// 256 threads play the work-items of one workgroup, the workgroup barrier
// is a thread barrier, and the 16x16x4 MFMA products are emulated per wave
// of 64 threads with the lane layouts measured on gfx950 (A[m][k] from lane
// m + 16k, B[k][n] from lane n + 16k; row m of D in register m % 4 of lane
// group m / 4 for FP32, in register m / 4 of lane group m % 4 for FP64).
// The three LDS slabs are separate static arrays, so an index outside any
// of them, the tables, the geometry or the element arrays aborts the run.
// Results are compared with the oracle's or_helmholtz (FP64 1e-12, FP32
// 2e-5 relative).  Prints "OK <max rel err FP64> <FP32>" on success.
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "basis.hpp"
#include "hex_mfma_block.hpp"

extern "C" void or_helmholtz(int shape, int nm, int nq0, int nelmt, int deformed, const double *jac,
                             const double *df, double lambda, const double *diff6, const double *in, double *out);

namespace {

class Barrier {
public:
    explicit Barrier(int n) : n_(n) {}
    void wait()
    {
        std::unique_lock<std::mutex> lk(m_);
        const long gen = gen_;
        if (++count_ == n_) {
            count_ = 0;
            ++gen_;
            cv_.notify_all();
        } else {
            cv_.wait(lk, [&] { return gen_ != gen; });
        }
    }

private:
    std::mutex m_;
    std::condition_variable cv_;
    int n_, count_ = 0;
    long gen_ = 0;
};

thread_local int tl_tid = 0;
Barrier g_block(nektar_amd::hexmf::kBlock);
Barrier g_wave[nektar_amd::hexmf::kWaves] = {Barrier(64), Barrier(64), Barrier(64), Barrier(64)};
struct Exchange {
    double a[64], b[64];
};
Exchange g_x[nektar_amd::hexmf::kWaves];

// D = A B + C for one wave; rowm(l, r) is the row held by register r of lane l
template <typename V, typename T, typename RowOf>
V emulate(T a, T b, V c, RowOf rowm)
{
    const int l = tl_tid & 63, w = tl_tid >> 6;
    g_x[w].a[l] = a;
    g_x[w].b[l] = b;
    g_wave[w].wait();
    V d;
    const int n = l & 15;
    for (int r = 0; r < 4; ++r) {
        const int m = rowm(l, r);
        double s = c[r];
        for (int k = 0; k < 4; ++k) s += g_x[w].a[m + 16 * k] * g_x[w].b[n + 16 * k];
        d[r] = (T)s;
    }
    g_wave[w].wait();  // every lane has read the operands before they change
    return d;
}

}  // namespace

namespace nektar_amd {
namespace hexmf {
int work_item() { return tl_tid; }
void barrier() { g_block.wait(); }
mfma_f64x4 host_mfma_f64(double a, double b, mfma_f64x4 c)
{
    return emulate<mfma_f64x4>(a, b, c, [](int l, int r) { return 4 * r + (l >> 4); });
}
mfma_f32x4 host_mfma_f32(float a, float b, mfma_f32x4 c)
{
    return emulate<mfma_f32x4>(a, b, c, [](int l, int r) { return 4 * (l >> 4) + r; });
}
}  // namespace hexmf
}  // namespace nektar_amd

using namespace nektar_amd;

namespace {

void metric6(const double *f, double *m)
{
    const int idx[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int c = 0; c < 6; ++c) {
        const int d = idx[c][0], e = idx[c][1];
        m[c] = f[d] * f[e] + f[3 + d] * f[3 + e] + f[6 + d] * f[6 + e];
    }
}

template <typename T, int NM, int NQ, bool DEF, int E>
double run_case(int nelmt, unsigned seed)
{
    constexpr int NM3 = NM * NM * NM, NQ3 = NQ * NQ * NQ;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0), J(0.5, 1.5);
    const Basis1D b = make_basis(BasisType::ModifiedA, NM, NQ);
    std::vector<double> tab64;  // B | D | W, the operator's table layout
    tab64.insert(tab64.end(), b.bdata.begin(), b.bdata.end());
    tab64.insert(tab64.end(), b.pts.D.begin(), b.pts.D.end());
    tab64.insert(tab64.end(), b.pts.w.begin(), b.pts.w.end());
    const size_t gp = DEF ? (size_t)nelmt * NQ3 : (size_t)nelmt;
    std::vector<double> jac(gp), df(gp * 9), in((size_t)nelmt * NM3), ref((size_t)nelmt * NM3);
    for (auto &v : jac) v = J(rng);
    for (size_t p = 0; p < gp; ++p)
        for (int c = 0; c < 9; ++c) df[p * 9 + c] = (c % 4 == 0 ? 1.0 : 0.0) + 0.3 * U(rng);
    for (auto &v : in) v = U(rng);
    const double lambda = 1.7;
    or_helmholtz(0, NM, NQ, nelmt, DEF ? 1 : 0, jac.data(), df.data(), lambda, nullptr, in.data(), ref.data());
    // packed geometry as the operator uploads it
    std::vector<double> geo64(gp * 7);
    for (int e = 0; e < nelmt; ++e) {
        if (!DEF) {
            double m[6];
            metric6(&df[(size_t)e * 9], m);
            for (int c = 0; c < 6; ++c) geo64[(size_t)e * 7 + c] = jac[e] * m[c];
            geo64[(size_t)e * 7 + 6] = jac[e];
            continue;
        }
        for (int k = 0, q = 0; k < NQ; ++k)
            for (int j = 0; j < NQ; ++j)
                for (int i = 0; i < NQ; ++i, ++q) {
                    const size_t pt = (size_t)e * NQ3 + q;
                    double m[6];
                    metric6(&df[pt * 9], m);
                    const double wj = b.pts.w[i] * b.pts.w[j] * b.pts.w[k] * jac[pt];
                    for (int c = 0; c < 6; ++c) geo64[((size_t)e * 7 + c) * NQ3 + q] = wj * m[c];
                    geo64[((size_t)e * 7 + 6) * NQ3 + q] = wj;
                }
    }
    std::vector<T> tab(tab64.begin(), tab64.end()), geo(geo64.begin(), geo64.end()), x(in.begin(), in.end());
    std::vector<T> out((size_t)nelmt * NM3, T(NAN));
    const int nblocks = (nelmt + E - 1) / E;
    const T *xp = x.data();
    auto work = [&](int tid) {
        tl_tid = tid;
        const auto load = [xp](size_t e, int, int, int, int m) { return xp[e * NM3 + m]; };
        for (int blk = 0; blk < nblocks; ++blk) {
            const int e0 = blk * E;
            hexmf::helmholtz_block<T, NM, NQ, DEF, E>(tab.data(), geo.data(), (T)lambda, load, out.data(), e0,
                                                      std::min(E, nelmt - e0));
            hexmf::barrier();  // one set of LDS arrays serves every workgroup here
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < hexmf::kBlock; ++t) th.emplace_back(work, t);
    for (auto &t : th) t.join();
    double m = 0, n = 0;
    for (size_t i = 0; i < out.size(); ++i) {
        m = std::fmax(m, std::fabs((double)out[i] - ref[i]));
        n = std::fmax(n, std::fabs(ref[i]));
        if (!std::isfinite((double)out[i])) return INFINITY;
    }
    return m / n;
}

template <typename T, int NM, int NQ, int E>
void both(double &worst, int nelmt, unsigned seed)
{
    if constexpr (hexmf::fits<T, NM, NQ, E>()) {
        const double a = run_case<T, NM, NQ, false, E>(nelmt, seed);
        const double d = run_case<T, NM, NQ, true, E>(nelmt, seed + 1);
        std::printf("%s nm %d nq %d E %d: regular %.3e deformed %.3e\n", sizeof(T) == 8 ? "fp64" : "fp32", NM, NQ, E,
                    a, d);
        worst = std::fmax(worst, std::fmax(a, d));
    }
}

template <int NM, int NQ>
void sizes(double &w64, double &w32, int nelmt)
{
    both<double, NM, NQ, 1>(w64, nelmt, 11);
    both<double, NM, NQ, 2>(w64, nelmt, 13);
    both<double, NM, NQ, 4>(w64, nelmt, 15);
    both<float, NM, NQ, 1>(w32, nelmt, 17);
    both<float, NM, NQ, 2>(w32, nelmt, 19);
    both<float, NM, NQ, 4>(w32, nelmt, 21);
}

}  // namespace

int main()
{
    double w64 = 0, w32 = 0;
    // element counts leave a ragged last workgroup for E = 2 and 4
    sizes<5, 6>(w64, w32, 5);
    sizes<7, 8>(w64, w32, 5);
    sizes<8, 9>(w64, w32, 3);
    sizes<8, 10>(w64, w32, 3);
    sizes<10, 11>(w64, w32, 3);
    sizes<10, 12>(w64, w32, 3);
    if (!(w64 <= 1e-12) || !(w32 <= 2e-5)) {
        std::printf("FAIL %.3e %.3e\n", w64, w32);
        return 1;
    }
    std::printf("OK %.3e %.3e\n", w64, w32);
    return 0;
}
