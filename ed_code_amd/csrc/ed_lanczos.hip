// ed_lanczos.hip -- device Lanczos around the H*v kernel (include/ed_amd.h,
// ed_lanczos).  Krylov vectors live in HBM for the whole solve; only alpha,
// beta and reorthogonalisation coefficients (FP64) cross to the host, where
// the tridiagonal matrix is diagonalised.
//
// Reductions are deterministic: every dot product is a fixed two-stage tree
// (per-block FP64 partials, then a compensated sum in block order).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#include <algorithm>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_internal.h"
#include "ed_tridiag.h"

using namespace edc;

namespace {

constexpr int kRedBlocks = 1024;   // blocks of the reduction kernels
constexpr int kRedThreads = 256;
constexpr int kMaxMulti = 64;      // vectors per multi-dot / multi-axpy pass

__device__ __forceinline__ cd2 ld(const cd2* p, int64_t i) { return p[i]; }
__device__ __forceinline__ cd2 ld(const cf2* p, int64_t i) { cf2 v = p[i]; return cd2{(double)v.re, (double)v.im}; }
__device__ __forceinline__ void st(cd2* p, int64_t i, double re, double im) { p[i] = cd2{re, im}; }
__device__ __forceinline__ void st(cf2* p, int64_t i, double re, double im) { p[i] = cf2{(float)re, (float)im}; }
// elements 2t and 2t+1 (p 16-byte aligned): one 16-byte load for complex64
__device__ __forceinline__ void ld2(const cf2* p, int64_t t, cd2& a, cd2& b) {
    const float4 q = ((const float4*)p)[t];
    a = cd2{(double)q.x, (double)q.y};
    b = cd2{(double)q.z, (double)q.w};
}
__device__ __forceinline__ void ld2(const cd2* p, int64_t t, cd2& a, cd2& b) { a = p[2 * t]; b = p[2 * t + 1]; }
// The same loads marked non-temporal, for the Krylov vectors a pass streams
// once (a read of 22.5 GB: 6.7-7.0 TB/s against 6.0-6.3 TB/s for plain loads)
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cd2 ldn(const cd2* p, int64_t i) {
    const nt_d2 q = __builtin_nontemporal_load((const nt_d2*)(p + i));
    return cd2{q.x, q.y};
}
__device__ __forceinline__ cd2 ldn(const cf2* p, int64_t i) {
    const nt_f2 q = __builtin_nontemporal_load((const nt_f2*)(p + i));
    return cd2{(double)q.x, (double)q.y};
}
__device__ __forceinline__ void ldn2(const cf2* p, int64_t t, cd2& a, cd2& b) {
    const nt_f4 q = __builtin_nontemporal_load((const nt_f4*)p + t);
    a = cd2{(double)q.x, (double)q.y};
    b = cd2{(double)q.z, (double)q.w};
}
__device__ __forceinline__ void ldn2(const cd2* p, int64_t t, cd2& a, cd2& b) { a = ldn(p, 2 * t); b = ldn(p, 2 * t + 1); }

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// (2u-1) + i(2u'-1), u from splitmix64(seed*K + 2i), u' from (seed*K + 2i + 1)
template <typename VT>
__global__ void k_start(int64_t D, uint64_t seed, VT* v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const double sc = 1.0 / 9007199254740992.0;
    uint64_t base = seed * 0x632BE59BD9B4E019ull;
    double a = (double)(splitmix64(base + 2 * (uint64_t)i) >> 11) * sc;
    double b = (double)(splitmix64(base + 2 * (uint64_t)i + 1) >> 11) * sc;
    st(v, i, 2.0 * a - 1.0, 2.0 * b - 1.0);
}

// partial[blk][k] = sum over the block's rows of conj(X_k[i]) * y[i] for the
// MB vectors X_k = X + k*ldx (MB a compile-time chunk, so the accumulators
// stay in registers).  Wave reduction by shuffles, then across the block's
// waves through LDS: one barrier per chunk.  PAIR: each thread takes element
// pairs (2t, 2t+1), 16-byte loads for complex64 (X, y and ldx keep every
// vector 16-byte aligned); an odd last element goes to thread 0 of block 0.
template <int MB, typename XT, typename YT, bool PAIR>
__global__ void __launch_bounds__(kRedThreads) k_multidot(int64_t D, const XT* __restrict__ X, int64_t ldx,
                                                           const YT* __restrict__ y, cd2* __restrict__ partial,
                                                           int kofs) {
    __shared__ double red[kRedThreads / 64][MB][2];
    double are[MB], aim[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) { are[k] = 0.0; aim[k] = 0.0; }
    if (PAIR) {
        const int64_t np = D >> 1;
        for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < np; t += (int64_t)gridDim.x * blockDim.x) {
            cd2 y0, y1;
            ld2(y, t, y0, y1);
#pragma unroll
            for (int k = 0; k < MB; k++) {
                cd2 x0, x1;
                ldn2(X + k * ldx, t, x0, x1);
                are[k] += x0.re * y0.re + x0.im * y0.im + x1.re * y1.re + x1.im * y1.im;
                aim[k] += x0.re * y0.im - x0.im * y0.re + x1.re * y1.im - x1.im * y1.re;
            }
        }
        if ((D & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
            const cd2 yv = ld(y, D - 1);
            for (int k = 0; k < MB; k++) {
                const cd2 x = ld(X + k * ldx, D - 1);
                are[k] += x.re * yv.re + x.im * yv.im;
                aim[k] += x.re * yv.im - x.im * yv.re;
            }
        }
    } else {
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
            cd2 yv = ld(y, i);
#pragma unroll
            for (int k = 0; k < MB; k++) {
                cd2 x = ldn(X + k * ldx, i);
                are[k] += x.re * yv.re + x.im * yv.im;
                aim[k] += x.re * yv.im - x.im * yv.re;
            }
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < MB; k++) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            are[k] += __shfl_xor(are[k], off, 64);
            aim[k] += __shfl_xor(aim[k], off, 64);
        }
        if (lane == 0) { red[wv][k][0] = are[k]; red[wv][k][1] = aim[k]; }
    }
    __syncthreads();
    if ((int)threadIdx.x < MB) {
        double r = 0.0, im = 0.0;
        for (int q = 0; q < kRedThreads / 64; q++) { r += red[q][threadIdx.x][0]; im += red[q][threadIdx.x][1]; }
        partial[(int64_t)blockIdx.x * kMaxMulti + kofs + threadIdx.x] = cd2{r, im};
    }
}

// y[i] = a*y[i] - sum_k c[k] * X_k[i]   (c FP64 complex, y updated in FP64),
// MB vectors per pass.
template <int MB, typename XT, typename YT>
__global__ void __launch_bounds__(kRedThreads) k_multiaxpy(int64_t D, const XT* __restrict__ X, int64_t ldx,
                                                            const cd2* __restrict__ c, double a, YT* __restrict__ y) {
    cd2 cs[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) cs[k] = c[k];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
        cd2 yv = ld(y, i);
        double re = a * yv.re, im = a * yv.im;
#pragma unroll
        for (int k = 0; k < MB; k++) {
            cd2 x = ldn(X + k * ldx, i);
            re -= cs[k].re * x.re - cs[k].im * x.im;
            im -= cs[k].re * x.im + cs[k].im * x.re;
        }
        st(y, i, re, im);
    }
}

// y = s * x (with conversion between storage types)
template <typename XT, typename YT>
__global__ void k_scale_copy(int64_t D, double s, const XT* x, YT* y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
        cd2 v = ld(x, i);
        st(y, i, s * v.re, s * v.im);
    }
}

// y += sum_k t[k] * X_k  with y in FP64 (Ritz vector assembly, real t)
template <int MB, typename XT>
__global__ void __launch_bounds__(kRedThreads) k_accumulate(int64_t D, const XT* __restrict__ X, int64_t ldx,
                                                             const double* __restrict__ t, cd2* __restrict__ y) {
    double ts[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) ts[k] = t[k];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
        cd2 yv = y[i];
#pragma unroll
        for (int k = 0; k < MB; k++) {
            cd2 x = ldn(X + k * ldx, i);
            yv.re += ts[k] * x.re;
            yv.im += ts[k] * x.im;
        }
        y[i] = yv;
    }
}

// Thick-restart change of basis in one pass over the Krylov vectors: for the
// outputs k0 <= k < k0 + ko (ko <= KO), W_k[e] = sum_{l < nl} t[(l kk + k) M
// + m] X_l[e], m = threadIdx.x % M the instance of element e (the grid stride
// is a multiple of M), summed in FP64 in the order l = 0, 1, ...; W_k = W +
// (k - k0) ldw.  W may be X itself when one launch writes every output: a
// thread reads all inputs of an element before it writes that element.
template <int KO, typename VT>
__global__ void __launch_bounds__(kRedThreads) k_restart(int64_t n, int M, const VT* X, int64_t ldx, int nl,
                                                          const double* __restrict__ t, int kk, int k0, int ko, VT* W,
                                                          int64_t ldw) {
    extern __shared__ double tsm[];   // [nl][KO][M]
    for (int q = threadIdx.x; q < nl * KO * M; q += blockDim.x) {
        const int l = q / (KO * M), r = q % (KO * M), k = r / M, m = r % M;
        tsm[q] = k < ko ? t[((int64_t)l * kk + k0 + k) * M + m] : 0.0;
    }
    __syncthreads();
    const int m = (int)(threadIdx.x % M);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double ar[KO], ai[KO];
#pragma unroll
        for (int k = 0; k < KO; k++) { ar[k] = 0.0; ai[k] = 0.0; }
        for (int l = 0; l < nl; l++) {
            const cd2 x = ldn(X + (int64_t)l * ldx, e);
            const double* tl = tsm + l * KO * M + m;
#pragma unroll
            for (int k = 0; k < KO; k++) {
                ar[k] = fma(tl[k * M], x.re, ar[k]);
                ai[k] = fma(tl[k * M], x.im, ai[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < KO; k++)
            if (k < ko) st(W + (int64_t)k * ldw, e, ar[k], ai[k]);
    }
}

struct Ctx {
    ed_op_s* op;
    int64_t D;
    hipStream_t st;
    cd2* d_partial = nullptr;   // [kRedBlocks][kMaxMulti]
    cd2* d_coef = nullptr;      // [kMaxMulti]
    double* d_t = nullptr;      // [kMaxMulti]
    cd2* d_coef_big = nullptr;  // [coef_cap]
    double* d_t_big = nullptr;  // [coef_cap]
    int coef_cap = 0;
    std::vector<cd2> h_partial;
    double hv_ms = 0.0;
    int64_t hv_calls = 0;
    hipEvent_t e0, e1;
};

unsigned grid_for(int64_t D) { return (unsigned)std::min<int64_t>(kRedBlocks, (D + kRedThreads - 1) / kRedThreads); }

// One element per thread for the multi-vector updates: a wave loads its
// elements, writes y and ends.  A grid-stride loop would make every wave wait
// for its own store (stores and loads share one counter) before the loads of
// its next element; the results are the same either way.
unsigned grid_stream(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kRedThreads - 1) / kRedThreads); }

// V_k <- sum_{l < nl} t[(l kk + k) M + m] V_l for k < kk (m the instance of
// each element, M of them interleaved; M = 1 for one instance): the kept
// Ritz vectors of a thick restart.  t has nl * kk * M entries (zeros past an
// instance's basis); d_t holds at least that many doubles, scratch at least
// kk * ldv elements.  One pass over the basis writes up to 16 outputs, in
// place when they are all of them.
template <typename VT>
int restart_basis(hipStream_t st, int64_t n, int M, VT* V, int64_t ldv, int nl, const std::vector<double>& t, int kk,
                  double* d_t, VT* scratch) {
    if (kk <= 0) return ED_OK;
    HIPCHK(hipMemcpyAsync(d_t, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, st));
    const bool wide = (size_t)nl * 16 * M * sizeof(double) <= 48 * 1024;
    const int ko_max = wide ? 16 : 8;
    const size_t sm = (size_t)nl * ko_max * M * sizeof(double);
    if (sm > 64 * 1024) return ed_fail(ED_EINVAL, "restart coefficients exceed the LDS (%d vectors, %d instances)", nl, M);
    const unsigned g = grid_for(n);
    const bool inplace = kk <= ko_max;
    for (int k0 = 0; k0 < kk; k0 += ko_max) {
        const int ko = std::min(ko_max, kk - k0);
        VT* out = inplace ? V : scratch + (int64_t)k0 * ldv;
        if (wide)
            hipLaunchKernelGGL((k_restart<16, VT>), dim3(g), dim3(kRedThreads), sm, st, n, M, V, ldv, nl, d_t, kk, k0, ko,
                               out, ldv);
        else
            hipLaunchKernelGGL((k_restart<8, VT>), dim3(g), dim3(kRedThreads), sm, st, n, M, V, ldv, nl, d_t, kk, k0, ko,
                               out, ldv);
        HIPCHK(hipGetLastError());
    }
    if (!inplace) HIPCHK(hipMemcpyAsync(V, scratch, (size_t)kk * ldv * sizeof(VT), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return ED_OK;
}
// Whether restart_basis can stage nl vectors' coefficients for M instances.
inline bool restart_fits(int nl, int M) { return (size_t)nl * 8 * M * sizeof(double) <= 64 * 1024; }


template <typename XT, typename YT, bool PAIR>
void launch_multidot_p(Ctx& c, int mm, const XT* X, int64_t ldx, const YT* y, unsigned g, int kofs) {
    switch (mm) {
        case 16: hipLaunchKernelGGL((k_multidot<16, XT, YT, PAIR>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, X, ldx, y, c.d_partial, kofs); break;
        case 8: hipLaunchKernelGGL((k_multidot<8, XT, YT, PAIR>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, X, ldx, y, c.d_partial, kofs); break;
        case 4: hipLaunchKernelGGL((k_multidot<4, XT, YT, PAIR>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, X, ldx, y, c.d_partial, kofs); break;
        case 2: hipLaunchKernelGGL((k_multidot<2, XT, YT, PAIR>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, X, ldx, y, c.d_partial, kofs); break;
        default: hipLaunchKernelGGL((k_multidot<1, XT, YT, PAIR>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, X, ldx, y, c.d_partial, kofs); break;
    }
}

// paired loads when every vector of the pass is 16-byte aligned
template <typename XT, typename YT>
void launch_multidot(Ctx& c, int mm, const XT* X, int64_t ldx, const YT* y, unsigned g, int kofs) {
    const bool aligned = ((uintptr_t)X % 16 == 0) && ((uintptr_t)y % 16 == 0) &&
                         (mm == 1 || ((size_t)ldx * sizeof(XT)) % 16 == 0);
    if (aligned) launch_multidot_p<XT, YT, true>(c, mm, X, ldx, y, g, kofs);
    else launch_multidot_p<XT, YT, false>(c, mm, X, ldx, y, g, kofs);
}

// out[k] = X_k^H y for k < m (any m), FP64, fixed reduction order.
template <typename XT, typename YT>
int multidot(Ctx& c, int m, const XT* X, int64_t ldx, const YT* y, std::vector<std::complex<double>>& out) {
    out.assign(m, 0.0);
    unsigned g = grid_for(c.D);
    for (int k0 = 0; k0 < m; k0 += kMaxMulti) {
        int mm = std::min(kMaxMulti, m - k0);
        // passes of 8, 4, 2, 1 vectors fill columns [0, mm) of the partials
        for (int q = 0; q < mm;) {
            int step = (mm - q) >= 16 ? 16 : (mm - q) >= 8 ? 8 : (mm - q) >= 4 ? 4 : (mm - q) >= 2 ? 2 : 1;
            launch_multidot<XT, YT>(c, step, X + (int64_t)(k0 + q) * ldx, ldx, y, g, q);
            HIPCHK(hipGetLastError());
            q += step;
        }
        HIPCHK(hipMemcpyAsync(c.h_partial.data(), c.d_partial, (size_t)g * kMaxMulti * sizeof(cd2),
                              hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        for (int k = 0; k < mm; k++) {
            // Kahan-compensated sum over the block partials, in block order
            double sr = 0, cr = 0, si = 0, ci = 0;
            for (unsigned b = 0; b < g; b++) {
                double yr = c.h_partial[(size_t)b * kMaxMulti + k].re - cr, ti = sr + yr;
                cr = (ti - sr) - yr; sr = ti;
                double yi = c.h_partial[(size_t)b * kMaxMulti + k].im - ci, tj = si + yi;
                ci = (tj - si) - yi; si = tj;
            }
            out[k0 + k] = std::complex<double>(sr, si);
        }
    }
    return ED_OK;
}

// y = a*y - sum_k coef[k] X_k
template <typename XT, typename YT>
int multiaxpy(Ctx& c, int m, const XT* X, int64_t ldx, const std::vector<std::complex<double>>& coef, double a, YT* y) {
    unsigned g = grid_stream(c.D);
    if (m <= 0) return ed_fail(ED_EINVAL, "multiaxpy needs at least one vector (m = %d)", m);
    std::vector<cd2> cc(m);
    for (int k = 0; k < m; k++) cc[k] = cd2{coef[k].real(), coef[k].imag()};
    cd2* dc = c.d_coef_big;
    if (m > c.coef_cap) return ed_fail(ED_EINVAL, "too many vectors (%d) in one update", m);
    HIPCHK(hipMemcpyAsync(dc, cc.data(), m * sizeof(cd2), hipMemcpyHostToDevice, c.st));
    bool first = true;
    for (int q = 0; q < m;) {
        int step = (m - q) >= 16 ? 16 : (m - q) >= 8 ? 8 : (m - q) >= 4 ? 4 : (m - q) >= 2 ? 2 : 1;
        double aa = first ? a : 1.0;
        const XT* Xq = X + (int64_t)q * ldx;
        switch (step) {
            case 16: hipLaunchKernelGGL((k_multiaxpy<16, XT, YT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, dc + q, aa, y); break;
            case 8: hipLaunchKernelGGL((k_multiaxpy<8, XT, YT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, dc + q, aa, y); break;
            case 4: hipLaunchKernelGGL((k_multiaxpy<4, XT, YT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, dc + q, aa, y); break;
            case 2: hipLaunchKernelGGL((k_multiaxpy<2, XT, YT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, dc + q, aa, y); break;
            default: hipLaunchKernelGGL((k_multiaxpy<1, XT, YT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, dc + q, aa, y); break;
        }
        HIPCHK(hipGetLastError());
        first = false;
        q += step;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    return ED_OK;
}

// y += sum_k t[k] X_k (Ritz vector), t real
template <typename XT>
int accumulate(Ctx& c, int m, const XT* X, int64_t ldx, const std::vector<double>& t, cd2* y) {
    unsigned g = grid_for(c.D);
    if (m > c.coef_cap) return ed_fail(ED_EINVAL, "too many vectors (%d) in one update", m);
    HIPCHK(hipMemcpyAsync(c.d_t_big, t.data(), m * sizeof(double), hipMemcpyHostToDevice, c.st));
    for (int q = 0; q < m;) {
        int step = (m - q) >= 16 ? 16 : (m - q) >= 8 ? 8 : (m - q) >= 4 ? 4 : (m - q) >= 2 ? 2 : 1;
        const XT* Xq = X + (int64_t)q * ldx;
        switch (step) {
            case 16: hipLaunchKernelGGL((k_accumulate<16, XT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, c.d_t_big + q, y); break;
            case 8: hipLaunchKernelGGL((k_accumulate<8, XT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, c.d_t_big + q, y); break;
            case 4: hipLaunchKernelGGL((k_accumulate<4, XT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, c.d_t_big + q, y); break;
            case 2: hipLaunchKernelGGL((k_accumulate<2, XT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, c.d_t_big + q, y); break;
            default: hipLaunchKernelGGL((k_accumulate<1, XT>), dim3(g), dim3(kRedThreads), 0, c.st, c.D, Xq, ldx, c.d_t_big + q, y); break;
        }
        HIPCHK(hipGetLastError());
        q += step;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    return ED_OK;
}

template <typename VT>
int hv(Ctx& c, const VT* v, VT* w) {
    HIPCHK(hipEventRecord(c.e0, c.st));
    int rc = op_hv(c.op, std::is_same<VT, cd2>::value ? ED_C128 : ED_C64, v, w, c.st);
    if (rc != ED_OK) return rc;
    HIPCHK(hipEventRecord(c.e1, c.st));
    HIPCHK(hipEventSynchronize(c.e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c.e0, c.e1));
    c.hv_ms += ms;
    c.hv_calls++;
    return ED_OK;
}

template <typename T>
int dnorm2(Ctx& c, const T* x, double& out) {
    std::vector<std::complex<double>> r;
    int rc = multidot<T, T>(c, 1, x, 0, x, r);
    out = r[0].real();
    return rc;
}

#define RC(x)                              \
    do {                                   \
        int rc_ = (x);                     \
        if (rc_ != ED_OK) return rc_;      \
    } while (0)

// Lowest eigenpair of the (m x m) tridiagonal (alphas, betas[1..m-1]).
bool lowest(const std::vector<double>& al, const std::vector<double>& be, double& e0, std::vector<double>* vec) {
    std::vector<double> d(al), e(be);
    std::vector<double> z;
    bool ok = tridiag_ql(d, e, vec ? &z : nullptr);
    e0 = d.empty() ? 0.0 : d[0];
    if (vec) vec->assign(z.begin(), z.begin() + al.size());
    return ok;
}

// lanczos.cpp:133-349, on the device.  VT = Krylov storage type.
template <typename VT>
int run_reference(Ctx& c, const ed_lanczos_params& p, std::vector<cd2*>& prev, double* evals) {
    const int64_t D = c.D;
    int how_many = (int)std::min<int64_t>(p.nev, D);
    int iterations = p.iterations;
    unsigned g = grid_for(D);
    VT *V = nullptr, *vp = nullptr, *w = nullptr;
    int maxit = (int)std::min<int64_t>(iterations, D);
    HIPCHK(hipMalloc(&V, (size_t)std::max(maxit, 1) * D * sizeof(VT)));
    HIPCHK(hipMalloc(&vp, D * sizeof(VT)));
    HIPCHK(hipMalloc(&w, D * sizeof(VT)));
    cd2* x = nullptr;
    HIPCHK(hipMalloc(&x, D * sizeof(cd2)));
    std::vector<std::complex<double>> q;
    for (int num = 0; num < how_many; num++) {
        iterations = (int)std::min<int64_t>(iterations, D - num);
        hipLaunchKernelGGL(k_start<VT>, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c.st, D, p.seed + (uint64_t)num, vp);
        HIPCHK(hipGetLastError());
        double previous = 1000.0, e0 = 0.0;
        std::vector<double> tvec;
        for (int cycle = 0; cycle < p.ncycles; cycle++) {
            // orth_wrt_previous_evecs + normalize (lanczos.cpp:199-200)
            for (size_t k = 0; k < prev.size(); k++) {
                RC((multidot<cd2, VT>(c, 1, prev[k], 0, vp, q)));
                RC((multiaxpy<cd2, VT>(c, 1, prev[k], 0, q, 1.0, vp)));
            }
            double nn;
            RC(dnorm2(c, vp, nn));
            hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), vp, vp);
            std::vector<double> alphas, betas(1, 0.0);
            int nvs = 0;
            for (int it = 0; it < iterations; it++) {
                RC(hv<VT>(c, vp, w));
                if (it != 0) {
                    std::vector<std::complex<double>> b1(1, betas.back());
                    RC((multiaxpy<VT, VT>(c, 1, V + (int64_t)(it - 1) * D, 0, b1, 1.0, w)));
                }
                RC((multidot<VT, VT>(c, 1, vp, 0, w, q)));          // alpha = v^H w
                std::complex<double> alpha = q[0];
                alphas.push_back(alpha.real());
                std::vector<std::complex<double>> a1(1, alpha);
                RC((multiaxpy<VT, VT>(c, 1, vp, 0, a1, 1.0, w)));
                hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0, vp, V + (int64_t)it * D);
                nvs = it + 1;
                double bb;
                RC(dnorm2(c, w, bb));
                double beta = std::sqrt(bb);
                hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / beta, w, vp);
                betas.push_back(beta);
                for (size_t k = 0; k < prev.size(); k++) {
                    RC((multidot<cd2, VT>(c, 1, prev[k], 0, vp, q)));
                    RC((multiaxpy<cd2, VT>(c, 1, prev[k], 0, q, 1.0, vp)));
                }
                if (cycle == p.ncycles - 1) {
                    // classical Gram-Schmidt against this cycle's vectors (lanczos.cpp:256-260)
                    RC((multidot<VT, VT>(c, nvs, V, D, vp, q)));
                    RC((multiaxpy<VT, VT>(c, nvs, V, D, q, 1.0, vp)));
                }
                RC(dnorm2(c, vp, nn));
                hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), vp, vp);
                std::vector<double> off(betas.begin() + 1, betas.end() - 1);
                if (!lowest(alphas, off, e0, nullptr)) return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
                if (cycle != 0 && cycle != p.ncycles - 1 && it > 1) {
                    if (std::fabs(e0 - previous) < p.tol) { it = iterations; cycle = p.ncycles - 2; }
                    previous = e0;
                } else if (cycle != p.ncycles - 1) {
                    previous = e0;
                }
            }
            // Ritz vector (lanczos.cpp:312-320), assembled in FP64
            std::vector<double> off(betas.begin() + 1, betas.end() - 1);
            off.resize(alphas.size() > 0 ? alphas.size() - 1 : 0);
            if (!lowest(alphas, off, e0, &tvec)) return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
            HIPCHK(hipMemsetAsync(x, 0, D * sizeof(cd2), c.st));
            RC((accumulate<VT>(c, nvs, V, D, tvec, x)));
            for (size_t k = 0; k < prev.size(); k++) {
                RC((multidot<cd2, cd2>(c, 1, prev[k], 0, x, q)));
                RC((multiaxpy<cd2, cd2>(c, 1, prev[k], 0, q, 1.0, x)));
            }
            RC(dnorm2(c, x, nn));
            hipLaunchKernelGGL((k_scale_copy<cd2, cd2>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), x, x);
            hipLaunchKernelGGL((k_scale_copy<cd2, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0, x, vp);
        }
        cd2* keep = nullptr;
        HIPCHK(hipMalloc(&keep, D * sizeof(cd2)));
        HIPCHK(hipMemcpyAsync(keep, x, D * sizeof(cd2), hipMemcpyDeviceToDevice, c.st));
        prev.push_back(keep);
        evals[num] = e0;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    dfree(V); dfree(vp); dfree(w); dfree(x);
    return ED_OK;
}

// One cycle with full (two-pass classical Gram-Schmidt) reorthogonalisation
// against every Krylov vector and every earlier eigenvector, each step.
template <typename VT>
int run_full(Ctx& c, const ed_lanczos_params& p, std::vector<cd2*>& prev, double* evals) {
    const int64_t D = c.D;
    int how_many = (int)std::min<int64_t>(p.nev, D);
    unsigned g = grid_for(D);
    int m = (int)std::min<int64_t>(p.iterations, D);
    VT *V = nullptr, *w = nullptr;
    const int64_t ldv = (D + 63) / 64 * 64;   // 16-byte aligned Krylov vectors (paired loads)
    HIPCHK(hipMalloc(&V, (size_t)(m + 1) * ldv * sizeof(VT)));
    HIPCHK(hipMalloc(&w, D * sizeof(VT)));
    cd2* x = nullptr;
    HIPCHK(hipMalloc(&x, D * sizeof(cd2)));
    std::vector<std::complex<double>> q;
    for (int num = 0; num < how_many; num++) {
        int mit = (int)std::min<int64_t>(m, D - num);
        VT* v0 = V;
        hipLaunchKernelGGL(k_start<VT>, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c.st, D, p.seed + (uint64_t)num, v0);
        for (int pass = 0; pass < 2; pass++)
            for (size_t k = 0; k < prev.size(); k++) {
                RC((multidot<cd2, VT>(c, 1, prev[k], 0, v0, q)));
                RC((multiaxpy<cd2, VT>(c, 1, prev[k], 0, q, 1.0, v0)));
            }
        double nn;
        RC(dnorm2(c, v0, nn));
        hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), v0, v0);
        std::vector<double> alphas, betas;
        double e0 = 0.0, eprev = 1e300;
        int k = 0;
        for (; k < mit; k++) {
            VT* vk = V + (int64_t)k * ldv;
            RC(hv<VT>(c, vk, w));
            RC((multidot<VT, VT>(c, 1, vk, 0, w, q)));
            double alpha = q[0].real();
            std::vector<std::complex<double>> cf{alpha};
            if (k > 0) {
                cf.insert(cf.begin(), betas.back());
                RC((multiaxpy<VT, VT>(c, 2, V + (int64_t)(k - 1) * ldv, ldv, cf, 1.0, w)));
            } else {
                RC((multiaxpy<VT, VT>(c, 1, vk, 0, cf, 1.0, w)));
            }
            const bool sel = (p.flags & ED_LZ_SELECTIVE) != 0;
            double nb2 = 0.0, bb = 0.0;
            bool have_bb = false;
            if (sel) RC(dnorm2(c, w, nb2));
            for (int pass = 0; pass < 2; pass++) {
                if (pass == 1 && sel) {
                    RC(dnorm2(c, w, bb));
                    if (bb > 0.5 * nb2) { have_bb = true; break; }
                }
                RC((multidot<VT, VT>(c, k + 1, V, ldv, w, q)));
                if (pass == 0) alpha += q[k].real();
                RC((multiaxpy<VT, VT>(c, k + 1, V, ldv, q, 1.0, w)));
                for (size_t e = 0; e < prev.size(); e++) {
                    RC((multidot<cd2, VT>(c, 1, prev[e], 0, w, q)));
                    RC((multiaxpy<cd2, VT>(c, 1, prev[e], 0, q, 1.0, w)));
                }
            }
            alphas.push_back(alpha);
            if (!have_bb) RC(dnorm2(c, w, bb));
            double beta = std::sqrt(bb);
            bool stop = false;
            if ((k + 1) % 5 == 0 || k + 1 == mit) {
                std::vector<double> off(betas);
                if (!lowest(alphas, off, e0, nullptr)) return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
                if (std::fabs(e0 - eprev) < p.tol) stop = true;
                eprev = e0;
            }
            if (beta < 1e-13 * std::max(1.0, std::fabs(alpha))) stop = true;   // invariant subspace
            if (stop || k + 1 == mit) { k++; break; }
            betas.push_back(beta);
            hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / beta, w, V + (int64_t)(k + 1) * ldv);
        }
        int nvs = k;
        std::vector<double> tvec, off(betas.begin(), betas.begin() + std::max(0, nvs - 1));
        std::vector<double> al(alphas.begin(), alphas.begin() + nvs);
        if (!lowest(al, off, e0, &tvec)) return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
        HIPCHK(hipMemsetAsync(x, 0, D * sizeof(cd2), c.st));
        RC((accumulate<VT>(c, nvs, V, ldv, tvec, x)));
        for (int pass = 0; pass < 2; pass++)
            for (size_t e = 0; e < prev.size(); e++) {
                RC((multidot<cd2, cd2>(c, 1, prev[e], 0, x, q)));
                RC((multiaxpy<cd2, cd2>(c, 1, prev[e], 0, q, 1.0, x)));
            }
        RC(dnorm2(c, x, nn));
        hipLaunchKernelGGL((k_scale_copy<cd2, cd2>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), x, x);
        cd2* keep = nullptr;
        HIPCHK(hipMalloc(&keep, D * sizeof(cd2)));
        HIPCHK(hipMemcpyAsync(keep, x, D * sizeof(cd2), hipMemcpyDeviceToDevice, c.st));
        prev.push_back(keep);
        evals[num] = e0;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    dfree(V); dfree(w); dfree(x);
    return ED_OK;
}

// Thick-restart Lanczos (Wu and Simon, SIAM J. Matrix Anal. Appl. 22 (2000)
// 602) for the nev lowest eigenpairs with at most m = iterations Krylov
// vectors: the ARPACK-like solver of the "SR" drivers (ARPACK_sector.cpp:
// 4-187 runs znaupd with ncv Arnoldi vectors).  Each expansion step
// removes the known part of H v_j (alpha v_j and beta v_{j-1}, or after a
// restart the couplings to the kept Ritz vectors), then reorthogonalises
// against all kept and new vectors (classical Gram-Schmidt, a second pass
// when the first removed more than 1 - 1/sqrt(2) of the norm); the projected matrix is the Lanczos tridiagonal bordered by the
// arrowhead row of the kept Ritz pairs.  After m vectors the Rayleigh-Ritz
// pairs are formed (symeig_jacobi); if the nev lowest have residual
// beta_m |y_m,i| <= tol max(1, |theta_i|) the run stops, else the
// nev + (m - nev)/2 lowest Ritz vectors and the normalised residual start
// the next cycle (at most ncycles cycles).
template <typename VT>
int run_trl(Ctx& c, const ed_lanczos_params& p, std::vector<cd2*>& prev, double* evals) {
    const int64_t D = c.D;
    const int nev = (int)std::min<int64_t>(p.nev, D);
    const int m = (int)std::min<int64_t>(p.iterations, D);
    if (m < nev + 1 && m < D) return ed_fail(ED_EINVAL, "thick restart needs iterations > nev (%d <= %d)", m, nev);
    unsigned g = grid_for(D);
    VT *V = nullptr, *w = nullptr;
    cd2* W = nullptr;
    double* Tr = nullptr;   // restart coefficients
    auto cleanup = [&]() { dfree(V); dfree(w); dfree(W); dfree(Tr); };
#define TRC1(expr)                 \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != ED_OK) {        \
            cleanup();             \
            return rc_;            \
        }                          \
    } while (0)
#define THIP1(expr) TRC1((expr) == hipSuccess ? ED_OK : ed_fail(ED_EDEVICE, "%s failed", #expr))
    THIP1(hipMalloc(&V, (size_t)(m + 1) * D * sizeof(VT)));
    THIP1(hipMalloc(&w, D * sizeof(VT)));
    THIP1(hipMalloc(&W, (size_t)m * D * sizeof(cd2)));
    THIP1(hipMalloc(&Tr, (size_t)m * m * sizeof(double)));
    auto Vk = [&](int k) { return V + (int64_t)k * D; };
    std::vector<std::complex<double>> q, q2;
    hipLaunchKernelGGL(k_start<VT>, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c.st, D, p.seed, Vk(0));
    double nn;
    TRC1(dnorm2(c, Vk(0), nn));
    hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), Vk(0), Vk(0));
    std::vector<double> T((size_t)m * m, 0.0), theta, Y;
    int start = 0, meff = m, ncyc = 0;
    double beta_m = 0.0;
    bool done = false;
    while (!done) {
        meff = m;
        for (int j = start; j < m; j++) {
            TRC1(hv<VT>(c, Vk(j), w));
            // the known part of H v_j first: alpha_j v_j and beta_{j-1} v_{j-1},
            // or after a restart the couplings to the kept Ritz vectors
            TRC1((multidot<VT, VT>(c, 1, Vk(j), 0, w, q)));
            double alpha = q[0].real();
            const int lo = j > start ? j - 1 : 0;
            std::vector<std::complex<double>> cf((size_t)(j + 1 - lo));
            for (int l = lo; l < j; l++) cf[l - lo] = T[(size_t)l * m + j];
            cf[j - lo] = alpha;
            TRC1((multiaxpy<VT, VT>(c, j + 1 - lo, Vk(lo), D, cf, 1.0, w)));
            // Gram-Schmidt against every kept and new vector; the second pass
            // only when the first removed more than 1 - 1/sqrt(2) of the norm
            double nb2, bb = 0.0;
            bool have_bb = false;
            TRC1(dnorm2(c, w, nb2));
            for (int pass = 0; pass < 2; pass++) {
                if (pass == 1) {
                    TRC1(dnorm2(c, w, bb));
                    if (bb > 0.5 * nb2) { have_bb = true; break; }
                }
                TRC1((multidot<VT, VT>(c, j + 1, V, D, w, q2)));
                alpha += q2[j].real();
                TRC1((multiaxpy<VT, VT>(c, j + 1, V, D, q2, 1.0, w)));
            }
            T[(size_t)j * m + j] = alpha;
            if (!have_bb) TRC1(dnorm2(c, w, bb));
            const double beta = std::sqrt(bb);
            if (beta < 1e-13 * std::max(1.0, std::fabs(T[(size_t)j * m + j]))) {   // invariant subspace
                meff = j + 1;
                beta_m = 0.0;
                break;
            }
            hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / beta, w, Vk(j + 1));
            if (j + 1 < m) {
                T[(size_t)j * m + j + 1] = T[(size_t)(j + 1) * m + j] = beta;
            } else {
                beta_m = beta;   // V[m] holds the normalised residual
            }
        }
        std::vector<double> Tm((size_t)meff * meff);
        for (int i = 0; i < meff; i++)
            for (int k = 0; k < meff; k++) Tm[(size_t)i * meff + k] = T[(size_t)i * m + k];
        if (!symeig_jacobi(Tm, meff, theta, Y)) TRC1(ed_fail(ED_ENOTCONV, "Rayleigh-Ritz eigensolver did not converge"));
        bool conv = true;
        for (int i = 0; i < std::min(nev, meff); i++)
            if (beta_m * std::fabs(Y[(meff - 1) + (size_t)i * meff]) > p.tol * std::max(1.0, std::fabs(theta[i])))
                conv = false;
        ncyc++;
        if (conv || beta_m == 0.0 || ncyc >= p.ncycles || meff < m) {
            done = true;
            break;
        }
        // restart with the kk lowest Ritz vectors and the residual
        const int kk = std::min(nev + (meff - nev) / 2, meff - 1);
        if (restart_fits(meff, 1)) {
            // the kept Ritz vectors in one pass over the basis
            std::vector<double> tr((size_t)meff * kk);
            for (int l = 0; l < meff; l++)
                for (int i = 0; i < kk; i++) tr[(size_t)l * kk + i] = Y[l + (size_t)i * meff];
            TRC1((restart_basis<VT>(c.st, D, 1, V, D, meff, tr, kk, Tr, (VT*)W)));
        } else {
            std::vector<double> y(meff);
            for (int i = 0; i < kk; i++) {
                cd2* Wi = W + (int64_t)i * D;
                THIP1(hipMemsetAsync(Wi, 0, D * sizeof(cd2), c.st));
                for (int l = 0; l < meff; l++) y[l] = Y[l + (size_t)i * meff];
                TRC1((accumulate<VT>(c, meff, V, D, y, Wi)));
            }
            for (int i = 0; i < kk; i++)
                hipLaunchKernelGGL((k_scale_copy<cd2, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0, W + (int64_t)i * D,
                                   Vk(i));
        }
        THIP1(hipMemcpyAsync(Vk(kk), Vk(m), D * sizeof(VT), hipMemcpyDeviceToDevice, c.st));
        std::fill(T.begin(), T.end(), 0.0);
        for (int i = 0; i < kk; i++) {
            T[(size_t)i * m + i] = theta[i];
            T[(size_t)i * m + kk] = T[(size_t)kk * m + i] = beta_m * Y[(meff - 1) + (size_t)i * meff];
        }
        start = kk;
    }
    // Ritz vectors of the nev lowest pairs
    std::vector<double> y(meff);
    for (int i = 0; i < std::min(nev, meff); i++) {
        cd2* x = nullptr;
        THIP1(hipMalloc(&x, D * sizeof(cd2)));
        THIP1(hipMemsetAsync(x, 0, D * sizeof(cd2), c.st));
        for (int l = 0; l < meff; l++) y[l] = Y[l + (size_t)i * meff];
        TRC1((accumulate<VT>(c, meff, V, D, y, x)));
        TRC1(dnorm2(c, x, nn));
        hipLaunchKernelGGL((k_scale_copy<cd2, cd2>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), x, x);
        prev.push_back(x);
        evals[i] = theta[i];
    }
    THIP1(hipStreamSynchronize(c.st));
    cleanup();
    return ED_OK;
#undef TRC1
#undef THIP1
}

// Memory-lean Lanczos: three Krylov vectors, no reorthogonalisation, lowest
// eigenvalue; a second pass regenerates the vectors from the stored alpha and
// beta to assemble the FP64 Ritz vector x (returned, unit norm).
template <typename VT>
int run_lean(Ctx& c, const ed_lanczos_params& p, cd2*& x_out, double& e0_out, int& steps) {
    const int64_t D = c.D;
    const unsigned g = grid_for(D);
    const bool progress = std::getenv("ED_PROGRESS") != nullptr;
    const int mit = (int)std::min<int64_t>(p.iterations, D);
    VT* buf[3] = {nullptr, nullptr, nullptr};
    for (auto*& b : buf) HIPCHK(hipMalloc(&b, D * sizeof(VT)));
    std::vector<double> alphas, betas;
    std::vector<std::complex<double>> q;
    auto start = [&](VT* v0) -> int {
        hipLaunchKernelGGL(k_start<VT>, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c.st, D, p.seed, v0);
        HIPCHK(hipGetLastError());
        double nn;
        RC(dnorm2(c, v0, nn));
        hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), v0, v0);
        HIPCHK(hipGetLastError());
        return ED_OK;
    };
    // pass 1: alpha, beta
    VT *vp = buf[0], *vk = buf[1], *w = buf[2];
    RC(start(vk));
    double e0 = 0.0, eprev = 1e300;
    int k = 0;
    for (; k < mit; k++) {
        RC(hv<VT>(c, vk, w));
        RC((multidot<VT, VT>(c, 1, vk, 0, w, q)));
        const double alpha = q[0].real();
        std::vector<std::complex<double>> a1(1, alpha);
        RC((multiaxpy<VT, VT>(c, 1, vk, 0, a1, 1.0, w)));
        if (k > 0) {
            std::vector<std::complex<double>> b1(1, betas.back());
            RC((multiaxpy<VT, VT>(c, 1, vp, 0, b1, 1.0, w)));
        }
        alphas.push_back(alpha);
        double bb;
        RC(dnorm2(c, w, bb));
        const double beta = std::sqrt(bb);
        bool stop = false;
        if ((k + 1) % 5 == 0 || k + 1 == mit) {
            if (!lowest(alphas, betas, e0, nullptr)) return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
            if (std::fabs(e0 - eprev) < p.tol * std::max(1.0, std::fabs(e0))) stop = true;
            eprev = e0;
        }
        if (progress) std::fprintf(stderr, "[lanczos-lean] pass 1 step %d alpha %.15g beta %.6g E0 %.15g\n", k, alpha, beta, e0);
        if (beta < 1e-13 * std::max(1.0, std::fabs(alpha))) stop = true;
        if (stop || k + 1 == mit) { k++; break; }
        betas.push_back(beta);
        hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / beta, w, w);
        HIPCHK(hipGetLastError());
        VT* t = vp; vp = vk; vk = w; w = t;
    }
    steps = k;
    std::vector<double> s;
    {
        std::vector<double> off(betas.begin(), betas.begin() + std::max(0, k - 1));
        std::vector<double> al(alphas.begin(), alphas.begin() + k);
        if (!lowest(al, off, e0, &s)) return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
    }
    e0_out = e0;
    // pass 2: the same vectors from the stored coefficients, x += s_j v_j
    cd2* x = nullptr;
    HIPCHK(hipMalloc(&x, D * sizeof(cd2)));
    HIPCHK(hipMemsetAsync(x, 0, D * sizeof(cd2), c.st));
    vp = buf[0]; vk = buf[1]; w = buf[2];
    RC(start(vk));
    for (int j = 0; j < k; j++) {
        std::vector<double> sj(1, s[j]);
        RC(accumulate<VT>(c, 1, vk, 0, sj, x));
        if (progress) std::fprintf(stderr, "[lanczos-lean] pass 2 step %d\n", j);
        if (j + 1 == k) break;
        RC(hv<VT>(c, vk, w));
        std::vector<std::complex<double>> a1(1, alphas[j]);
        RC((multiaxpy<VT, VT>(c, 1, vk, 0, a1, 1.0, w)));
        if (j > 0) {
            std::vector<std::complex<double>> b1(1, betas[j - 1]);
            RC((multiaxpy<VT, VT>(c, 1, vp, 0, b1, 1.0, w)));
        }
        hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / betas[j], w, w);
        HIPCHK(hipGetLastError());
        VT* t = vp; vp = vk; vk = w; w = t;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    for (auto* b : buf) dfree(b);
    double nn;
    RC(dnorm2(c, x, nn));
    hipLaunchKernelGGL((k_scale_copy<cd2, cd2>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), x, x);
    HIPCHK(hipGetLastError());
    x_out = x;
    return ED_OK;
}

// ---------------------------------------------------------------------------
// Graph-captured lean Lanczos.  Every scalar of a step stays on the device:
// alpha_k = Re<v_k, H v_k> and beta_{k+1} = ||w|| are reduced by a fixed-order
// tree into ab[] (alpha at 2k, beta at 2k+1), the update kernels read them
// from there, and a device step counter selects the slot.  Three steps (one
// full rotation of the three vector buffers) are captured once as a hipGraph
// and replayed; the host reads ab[] every kGraphCheck replays to test
// convergence.  Nothing in the captured work allocates or memsets.

constexpr int kGraphCheck = 2;   // replays (x 3 steps) between convergence checks

template <typename VT>
__global__ void __launch_bounds__(kRedThreads) k_dot_re_partial(int64_t D, const VT* __restrict__ a,
                                                                 const VT* __restrict__ b, double* __restrict__ part) {
    __shared__ double red[kRedThreads / 64];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
        const cd2 x = ld(a, i), y = ld(b, i);
        acc += x.re * y.re + x.im * y.im;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int q = 0; q < kRedThreads / 64; q++) r += red[q];
        part[blockIdx.x] = r;
    }
}

// out = fixed-order sum of part[0..g): one block, strided partial sums then a tree
__device__ double block_sum(const double* __restrict__ part, int g) {
    __shared__ double red[kRedThreads];
    double acc = 0.0;
    for (int b = threadIdx.x; b < g; b += blockDim.x) acc += part[b];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    return red[0];
}

// ab[2k] = alpha_k
__global__ void __launch_bounds__(kRedThreads) k_finish_alpha(const double* part, int g, double* ab, const int* step) {
    const double s = block_sum(part, g);
    if (threadIdx.x == 0) ab[2 * *step] = s;
}

// ab[2k+1] = beta_{k+1} = sqrt(sum)
__global__ void __launch_bounds__(kRedThreads) k_finish_beta(const double* part, int g, double* ab, const int* step) {
    const double s = block_sum(part, g);
    if (threadIdx.x == 0) ab[2 * *step + 1] = sqrt(s);
}

// w = w - alpha_k v - beta_k vp   (beta_0 = 0)
template <typename VT>
__global__ void k_lz_update(int64_t D, VT* __restrict__ w, const VT* __restrict__ v, const VT* __restrict__ vp,
                            const double* ab, const int* step) {
    const int k = *step;
    const double al = ab[2 * k], be = k > 0 ? ab[2 * k - 1] : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
        const cd2 x = ld(w, i), y = ld(v, i), z = ld(vp, i);
        st(w, i, x.re - al * y.re - be * z.re, x.im - al * y.im - be * z.im);
    }
}

// w = w / beta_{k+1} (0 once the Krylov space is exhausted or past nmax steps)
template <typename VT>
__global__ void k_lz_scale(int64_t D, VT* __restrict__ w, const double* ab, const int* step, int nmax) {
    const int k = *step;
    const double be = ab[2 * k + 1];
    const double f = (k + 1 < nmax && be > 1e-300) ? 1.0 / be : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
        const cd2 x = ld(w, i);
        st(w, i, f * x.re, f * x.im);
    }
}

// x += s_k v (k < n)
template <typename VT>
__global__ void k_lz_accum(int64_t D, cd2* __restrict__ x, const VT* __restrict__ v, const double* s, const int* step,
                           int n) {
    const int k = *step;
    if (k >= n) return;
    const double f = s[k];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
        const cd2 y = ld(v, i);
        cd2 a = x[i];
        a.re += f * y.re;
        a.im += f * y.im;
        x[i] = a;
    }
}

__global__ void k_step_inc(int* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

template <typename VT>
int run_lean_graph(Ctx& c, const ed_lanczos_params& p, cd2*& x_out, double& e0_out, int& steps) {
    const int64_t D = c.D;
    const unsigned g = grid_for(D);
    const int dt = std::is_same<VT, cd2>::value ? ED_C128 : ED_C64;
    const int mit = (int)std::min<int64_t>(p.iterations, D);
    const int cap = mit + 8;   // room for the steps of the last replay
    VT* buf[3] = {nullptr, nullptr, nullptr};
    for (auto*& b : buf) HIPCHK(hipMalloc(&b, D * sizeof(VT)));
    double *part = nullptr, *ab = nullptr, *sv = nullptr;
    int* step = nullptr;
    HIPCHK(hipMalloc(&part, g * sizeof(double)));
    HIPCHK(hipMalloc(&ab, 2 * cap * sizeof(double)));
    HIPCHK(hipMalloc(&sv, cap * sizeof(double)));
    HIPCHK(hipMalloc(&step, sizeof(int)));
    cd2* x = nullptr;
    HIPCHK(hipMalloc(&x, D * sizeof(cd2)));
    auto cleanup = [&]() {
        for (auto* b : buf) dfree(b);
        dfree(part); dfree(ab); dfree(sv); dfree(step);
    };
    // start vector, normalised, in buf[1]; warm H*v outside the capture
    // (allocates the gather stream and settles the kernel choice)
    auto start = [&]() -> int {
        // v_{-1} = 0 (the first update multiplies it by beta_0 = 0)
        HIPCHK(hipMemsetAsync(buf[0], 0, D * sizeof(VT), c.st));
        hipLaunchKernelGGL(k_start<VT>, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c.st, D, p.seed, buf[1]);
        HIPCHK(hipGetLastError());
        double nn;
        RC(dnorm2(c, buf[1], nn));
        hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), buf[1],
                           buf[1]);
        HIPCHK(hipGetLastError());
        const int zero = 0;
        HIPCHK(hipMemcpyAsync(step, &zero, sizeof(int), hipMemcpyHostToDevice, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        return ED_OK;
    };
    RC(start());
    RC(op_hv(c.op, dt, buf[1], buf[2], c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    // capture 3 steps of pass 1 (buffers rotate vp <- v <- w <- vp)
    auto capture = [&](bool pass2, int n, hipGraphExec_t& exec) -> int {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
        int r[3] = {0, 1, 2};
        for (int q = 0; q < 3; q++) {
            VT *vp = buf[r[0]], *v = buf[r[1]], *w = buf[r[2]];
            if (pass2)
                hipLaunchKernelGGL((k_lz_accum<VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, x, v, sv, step, n);
            int rc = op_hv(c.op, dt, v, w, c.st);
            if (rc != ED_OK) { (void)hipStreamEndCapture(c.st, &graph); return rc; }
            if (!pass2) {
                hipLaunchKernelGGL((k_dot_re_partial<VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, v, w, part);
                hipLaunchKernelGGL(k_finish_alpha, dim3(1), dim3(kRedThreads), 0, c.st, part, (int)g, ab, step);
            }
            hipLaunchKernelGGL((k_lz_update<VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, w, v, vp, ab, step);
            if (!pass2) {
                hipLaunchKernelGGL((k_dot_re_partial<VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, w, w, part);
                hipLaunchKernelGGL(k_finish_beta, dim3(1), dim3(kRedThreads), 0, c.st, part, (int)g, ab, step);
            }
            hipLaunchKernelGGL((k_lz_scale<VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, w, ab, step, pass2 ? n : cap - 4);
            hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, c.st, step);
            const int t = r[0]; r[0] = r[1]; r[1] = r[2]; r[2] = t;
        }
        HIPCHK(hipStreamEndCapture(c.st, &graph));
        HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        return ED_OK;
    };
    hipGraphExec_t g1;
    RC(capture(false, 0, g1));
    std::vector<double> hab(2 * cap);
    std::vector<double> alphas, betas;
    double e0 = 0.0, eprev = 1e300;
    int done = 0;   // steps completed on the device
    bool conv = false;
    while (!conv && done < mit) {
        for (int rep = 0; rep < kGraphCheck && done < mit && done + 3 <= cap - 4; rep++) {
            HIPCHK(hipGraphLaunch(g1, c.st));
            done += 3;
        }
        HIPCHK(hipMemcpyAsync(hab.data(), ab, 2 * done * sizeof(double), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        // alpha_k = hab[2k], beta_{k+1} = hab[2k+1]; stop at an exhausted Krylov space
        int valid = done;
        for (int k = 0; k < done; k++)
            if (hab[2 * k + 1] < 1e-13 * std::max(1.0, std::fabs(hab[2 * k]))) {
                valid = k + 1;
                conv = true;
                break;
            }
        done = valid;
        alphas.assign(done, 0.0);
        betas.assign(done - 1, 0.0);
        for (int k = 0; k < done; k++) {
            alphas[k] = hab[2 * k];
            if (k + 1 < done) betas[k] = hab[2 * k + 1];
        }
        if (!lowest(alphas, betas, e0, nullptr)) {
            (void)hipGraphExecDestroy(g1);
            cleanup();
            return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
        }
        if (std::fabs(e0 - eprev) < p.tol * std::max(1.0, std::fabs(e0))) conv = true;
        eprev = e0;
        if (std::getenv("ED_PROGRESS")) std::fprintf(stderr, "[lanczos-graph] pass 1 steps %d E0 %.15g\n", done, e0);
    }
    (void)hipGraphExecDestroy(g1);
    c.hv_calls += 1 + ((done + 2) / 3) * 3;   // warm-up plus the replayed steps
    steps = done;
    std::vector<double> sv_h;
    {
        std::vector<double> al(alphas.begin(), alphas.begin() + done), off(betas.begin(), betas.begin() + done - 1);
        if (!lowest(al, off, e0, &sv_h)) { cleanup(); return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge"); }
    }
    e0_out = e0;
    // pass 2 from the same start with the stored alpha/beta
    HIPCHK(hipMemcpyAsync(sv, sv_h.data(), done * sizeof(double), hipMemcpyHostToDevice, c.st));
    HIPCHK(hipMemsetAsync(x, 0, D * sizeof(cd2), c.st));
    RC(start());
    hipGraphExec_t g2;
    RC(capture(true, done, g2));
    for (int k = 0; k < done; k += 3) HIPCHK(hipGraphLaunch(g2, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    (void)hipGraphExecDestroy(g2);
    c.hv_calls += ((done + 2) / 3) * 3;
    cleanup();
    double nn;
    RC(dnorm2(c, x, nn));
    hipLaunchKernelGGL((k_scale_copy<cd2, cd2>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), x, x);
    HIPCHK(hipGetLastError());
    x_out = x;
    return ED_OK;
}

// FP64 residual ||H x - (x^H H x) x|| and Rayleigh quotient of unit vector x.
int fp64_check(Ctx& c, const cd2* x, double& rq, double& res) {
    cd2* y = nullptr;
    HIPCHK(hipMalloc(&y, c.D * sizeof(cd2)));
    RC(hv<cd2>(c, x, y));
    std::vector<std::complex<double>> q;
    RC((multidot<cd2, cd2>(c, 1, x, 0, y, q)));
    rq = q[0].real();
    std::vector<std::complex<double>> a1(1, rq);
    RC((multiaxpy<cd2, cd2>(c, 1, x, 0, a1, 1.0, y)));
    double rr;
    RC(dnorm2(c, y, rr));
    res = std::sqrt(rr);
    dfree(y);
    return ED_OK;
}

// ---------------------------------------------------------------------------
// Block Lanczos: M instances on one basis, interleaved vectors (element
// (i, m) at i*M + m), advanced together.  M divides 64, so with grid-stride
// loops over the D*M elements every thread stays on one instance m =
// threadIdx.x % M, and lanes with equal m are reduced together.

constexpr int kBCols = kMaxMulti * 8;   // partial columns: (vector, instance)

// partial[col * g + blk], col = (kofs + k) * M + m
template <int MB, typename XT, typename YT>
__global__ void __launch_bounds__(kRedThreads) k_bdot(int64_t n, int M, const XT* __restrict__ X, int64_t ldx,
                                                       const YT* __restrict__ y, cd2* __restrict__ partial,
                                                       int kofs) {
    __shared__ double red[kRedThreads / 64][MB][8][2];
    double are[MB], aim[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) { are[k] = 0.0; aim[k] = 0.0; }
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        cd2 yv = ld(y, e);
#pragma unroll
        for (int k = 0; k < MB; k++) {
            cd2 x = ldn(X + k * ldx, e);
            are[k] += x.re * yv.re + x.im * yv.im;
            aim[k] += x.re * yv.im - x.im * yv.re;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < MB; k++) {
        for (int off = 32; off >= M; off >>= 1) {
            are[k] += __shfl_xor(are[k], off, 64);
            aim[k] += __shfl_xor(aim[k], off, 64);
        }
        if (lane < M) { red[wv][k][lane][0] = are[k]; red[wv][k][lane][1] = aim[k]; }
    }
    __syncthreads();
    if ((int)threadIdx.x < MB * M) {
        const int k = threadIdx.x / M, m = threadIdx.x % M;
        double r = 0.0, im = 0.0;
        for (int q = 0; q < kRedThreads / 64; q++) { r += red[q][k][m][0]; im += red[q][k][m][1]; }
        partial[(int64_t)((kofs + k) * M + m) * gridDim.x + blockIdx.x] = cd2{r, im};
    }
}

// out[col] = compensated sum of partial[col * g + b], one wave per column:
// lane l sums b = l, l + 64, ... in order (Kahan), then lane 0 sums the 64
// lane totals in lane order (Kahan) -- a fixed order, so the result does not
// depend on scheduling.
__global__ void __launch_bounds__(64) k_colsum(const cd2* __restrict__ partial, int g, cd2* __restrict__ out) {
    __shared__ double red[64][2];
    const int col = blockIdx.x, lane = threadIdx.x;
    double sr = 0, cr = 0, si = 0, ci = 0;
    for (int b = lane; b < g; b += 64) {
        const cd2 p = partial[(int64_t)col * g + b];
        double yr = p.re - cr, tr = sr + yr;
        cr = (tr - sr) - yr; sr = tr;
        double yi = p.im - ci, tj = si + yi;
        ci = (tj - si) - yi; si = tj;
    }
    red[lane][0] = sr - cr;
    red[lane][1] = si - ci;
    __syncthreads();
    if (lane == 0) {
        sr = 0; cr = 0; si = 0; ci = 0;
        for (int l = 0; l < 64; l++) {
            double yr = red[l][0] - cr, tr = sr + yr;
            cr = (tr - sr) - yr; sr = tr;
            double yi = red[l][1] - ci, tj = si + yi;
            ci = (tj - si) - yi; si = tj;
        }
        out[col] = cd2{sr, si};
    }
}

// y[e] = a[m] * y[e] - sum_k c[k*M + m] * X_k[e]
template <int MB, typename XT, typename YT>
__global__ void __launch_bounds__(kRedThreads) k_baxpy(int64_t n, int M, const XT* __restrict__ X, int64_t ldx,
                                                        const cd2* __restrict__ c, const double* __restrict__ a,
                                                        YT* __restrict__ y) {
    const int m = threadIdx.x % M;
    cd2 cs[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) cs[k] = c[k * M + m];
    const double am = a ? a[m] : 1.0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        cd2 yv = ld(y, e);
        double re = am * yv.re, im = am * yv.im;
#pragma unroll
        for (int k = 0; k < MB; k++) {
            cd2 x = ldn(X + k * ldx, e);
            re -= cs[k].re * x.re - cs[k].im * x.im;
            im -= cs[k].re * x.im + cs[k].im * x.re;
        }
        st(y, e, re, im);
    }
}

// y[e] = s[m] * x[e]
template <typename XT, typename YT>
__global__ void k_bscale(int64_t n, int M, const double* __restrict__ sc, const XT* x, YT* y) {
    const int m = threadIdx.x % M;
    const double f = sc[m];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        cd2 v = ld(x, e);
        st(y, e, f * v.re, f * v.im);
    }
}

// y[e] += sum_k t[k*M + m] X_k[e]   (Ritz vectors, y FP64)
template <int MB, typename XT>
__global__ void __launch_bounds__(kRedThreads) k_baccumulate(int64_t n, int M, const XT* __restrict__ X, int64_t ldx,
                                                              const double* __restrict__ t, cd2* __restrict__ y) {
    const int m = threadIdx.x % M;
    double ts[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) ts[k] = t[k * M + m];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        cd2 yv = y[e];
#pragma unroll
        for (int k = 0; k < MB; k++) {
            cd2 x = ldn(X + k * ldx, e);
            yv.re += ts[k] * x.re;
            yv.im += ts[k] * x.im;
        }
        y[e] = yv;
    }
}

// start vector of instance m = the single-instance start vector of seed s0 + m
template <typename VT>
__global__ void k_bstart(int64_t D, int M, uint64_t s0, VT* v) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= D * M) return;
    const int64_t i = e / M;
    const uint64_t seed = s0 + (uint64_t)(e % M);
    const double sc = 1.0 / 9007199254740992.0;
    uint64_t base = seed * 0x632BE59BD9B4E019ull;
    double a = (double)(splitmix64(base + 2 * (uint64_t)i) >> 11) * sc;
    double b = (double)(splitmix64(base + 2 * (uint64_t)i + 1) >> 11) * sc;
    st(v, e, 2.0 * a - 1.0, 2.0 * b - 1.0);
}

struct BCtx {
    ed_block_s* blk;
    int64_t D, n;     // rows, D*M
    int M;
    hipStream_t st;
    unsigned g;       // reduction grid
    cd2* d_partial = nullptr;   // [kBCols * g]
    cd2* d_sum = nullptr;       // [kBCols]
    cd2* d_coef = nullptr;      // [cap * M]
    double* d_t = nullptr;      // [cap * M]
    double* d_s = nullptr;      // [8]
    int cap = 0;
    double hv_ms = 0.0;
    int64_t hv_calls = 0;
    hipEvent_t e0, e1;
};

template <int MB, typename XT, typename YT>
void launch_bdot(BCtx& c, const XT* X, int64_t ldx, const YT* y, int kofs) {
    hipLaunchKernelGGL((k_bdot<MB, XT, YT>), dim3(c.g), dim3(kRedThreads), 0, c.st, c.n, c.M, X, ldx, y, c.d_partial,
                       kofs);
}

// out[k*M + m] = <X_k, y>_m for k < m_vec
template <typename XT, typename YT>
int bdot(BCtx& c, int mv, const XT* X, int64_t ldx, const YT* y, std::vector<std::complex<double>>& out) {
    out.assign((size_t)mv * c.M, 0.0);
    std::vector<cd2> h;
    for (int k0 = 0; k0 < mv; k0 += kMaxMulti) {
        const int mm = std::min(kMaxMulti, mv - k0);
        for (int q = 0; q < mm;) {
            const int step = (mm - q) >= 16 ? 16 : (mm - q) >= 8 ? 8 : (mm - q) >= 4 ? 4 : (mm - q) >= 2 ? 2 : 1;
            const XT* Xq = X + (int64_t)(k0 + q) * ldx;
            switch (step) {
                case 16: launch_bdot<16, XT, YT>(c, Xq, ldx, y, q); break;
                case 8: launch_bdot<8, XT, YT>(c, Xq, ldx, y, q); break;
                case 4: launch_bdot<4, XT, YT>(c, Xq, ldx, y, q); break;
                case 2: launch_bdot<2, XT, YT>(c, Xq, ldx, y, q); break;
                default: launch_bdot<1, XT, YT>(c, Xq, ldx, y, q); break;
            }
            HIPCHK(hipGetLastError());
            q += step;
        }
        const int ncols = mm * c.M;
        hipLaunchKernelGGL(k_colsum, dim3(ncols), dim3(64), 0, c.st, c.d_partial, (int)c.g, c.d_sum);
        HIPCHK(hipGetLastError());
        h.resize(ncols);
        HIPCHK(hipMemcpyAsync(h.data(), c.d_sum, ncols * sizeof(cd2), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        for (int k = 0; k < ncols; k++) out[(size_t)k0 * c.M + k] = std::complex<double>(h[k].re, h[k].im);
    }
    return ED_OK;
}

// y = a .* y - sum_k coef[k*M + m] X_k   (a per instance; nullptr = 1)
template <typename XT, typename YT>
int baxpy(BCtx& c, int mv, const XT* X, int64_t ldx, const std::vector<std::complex<double>>& coef,
          const std::vector<double>* a, YT* y) {
    if (mv > c.cap) return ed_fail(ED_EINVAL, "too many vectors (%d) in one update", mv);
    std::vector<cd2> cc((size_t)mv * c.M);
    for (size_t k = 0; k < cc.size(); k++) cc[k] = cd2{coef[k].real(), coef[k].imag()};
    if (mv) HIPCHK(hipMemcpyAsync(c.d_coef, cc.data(), cc.size() * sizeof(cd2), hipMemcpyHostToDevice, c.st));
    double* da = nullptr;
    if (a) {
        HIPCHK(hipMemcpyAsync(c.d_s, a->data(), c.M * sizeof(double), hipMemcpyHostToDevice, c.st));
        da = c.d_s;
    }
    if (mv == 0) {
        if (da) hipLaunchKernelGGL((k_bscale<YT, YT>), dim3(c.g), dim3(kRedThreads), 0, c.st, c.n, c.M, da, y, y);
        HIPCHK(hipStreamSynchronize(c.st));
        return ED_OK;
    }
    for (int q = 0; q < mv;) {
        const int step = (mv - q) >= 16 ? 16 : (mv - q) >= 8 ? 8 : (mv - q) >= 4 ? 4 : (mv - q) >= 2 ? 2 : 1;
        const XT* Xq = X + (int64_t)q * ldx;
        const cd2* cq = c.d_coef + (size_t)q * c.M;
        double* aq = q == 0 ? da : nullptr;
        switch (step) {
            case 16: hipLaunchKernelGGL((k_baxpy<16, XT, YT>), dim3(grid_stream(c.n)), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, cq, aq, y); break;
            case 8: hipLaunchKernelGGL((k_baxpy<8, XT, YT>), dim3(grid_stream(c.n)), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, cq, aq, y); break;
            case 4: hipLaunchKernelGGL((k_baxpy<4, XT, YT>), dim3(grid_stream(c.n)), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, cq, aq, y); break;
            case 2: hipLaunchKernelGGL((k_baxpy<2, XT, YT>), dim3(grid_stream(c.n)), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, cq, aq, y); break;
            default: hipLaunchKernelGGL((k_baxpy<1, XT, YT>), dim3(grid_stream(c.n)), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, cq, aq, y); break;
        }
        HIPCHK(hipGetLastError());
        q += step;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    return ED_OK;
}

template <typename XT, typename YT>
int bscale(BCtx& c, const std::vector<double>& s, const XT* x, YT* y) {
    HIPCHK(hipMemcpyAsync(c.d_s, s.data(), c.M * sizeof(double), hipMemcpyHostToDevice, c.st));
    hipLaunchKernelGGL((k_bscale<XT, YT>), dim3(c.g), dim3(kRedThreads), 0, c.st, c.n, c.M, c.d_s, x, y);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.st));
    return ED_OK;
}

template <typename XT>
int baccumulate(BCtx& c, int mv, const XT* X, int64_t ldx, const std::vector<double>& t, cd2* y) {
    if (mv > c.cap) return ed_fail(ED_EINVAL, "too many vectors (%d) in one update", mv);
    HIPCHK(hipMemcpyAsync(c.d_t, t.data(), (size_t)mv * c.M * sizeof(double), hipMemcpyHostToDevice, c.st));
    for (int q = 0; q < mv;) {
        const int step = (mv - q) >= 16 ? 16 : (mv - q) >= 8 ? 8 : (mv - q) >= 4 ? 4 : (mv - q) >= 2 ? 2 : 1;
        const XT* Xq = X + (int64_t)q * ldx;
        const double* tq = c.d_t + (size_t)q * c.M;
        switch (step) {
            case 16: hipLaunchKernelGGL((k_baccumulate<16, XT>), dim3(c.g), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, tq, y); break;
            case 8: hipLaunchKernelGGL((k_baccumulate<8, XT>), dim3(c.g), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, tq, y); break;
            case 4: hipLaunchKernelGGL((k_baccumulate<4, XT>), dim3(c.g), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, tq, y); break;
            case 2: hipLaunchKernelGGL((k_baccumulate<2, XT>), dim3(c.g), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, tq, y); break;
            default: hipLaunchKernelGGL((k_baccumulate<1, XT>), dim3(c.g), dim3(kRedThreads), 0, c.st, c.n, c.M, Xq, ldx, tq, y); break;
        }
        HIPCHK(hipGetLastError());
        q += step;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    return ED_OK;
}

template <typename VT>
int bhv(BCtx& c, const VT* v, VT* w) {
    HIPCHK(hipEventRecord(c.e0, c.st));
    int rc = block_op_hv(c.blk, std::is_same<VT, cd2>::value ? ED_C128 : ED_C64, v, w, c.st);
    if (rc != ED_OK) return rc;
    HIPCHK(hipEventRecord(c.e1, c.st));
    HIPCHK(hipEventSynchronize(c.e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c.e0, c.e1));
    c.hv_ms += ms;
    c.hv_calls++;
    return ED_OK;
}

// Real parts of the per-instance diagonal <y, y>_m.
template <typename T>
int bnorm2(BCtx& c, const T* y, std::vector<double>& out) {
    std::vector<std::complex<double>> r;
    RC((bdot<T, T>(c, 1, y, 0, y, r)));
    out.resize(c.M);
    for (int m = 0; m < c.M; m++) out[m] = r[m].real();
    return ED_OK;
}

// x -= sum over earlier eigenvectors (two passes), per instance
int bdeflate(BCtx& c, std::vector<cd2*>& prev, cd2* x) {
    std::vector<std::complex<double>> q;
    for (int pass = 0; pass < 2; pass++)
        for (size_t e = 0; e < prev.size(); e++) {
            RC((bdot<cd2, cd2>(c, 1, prev[e], 0, x, q)));
            RC((baxpy<cd2, cd2>(c, 1, prev[e], 0, q, nullptr, x)));
        }
    return ED_OK;
}

template <typename VT>
int bdeflate_v(BCtx& c, std::vector<cd2*>& prev, VT* x) {
    std::vector<std::complex<double>> q;
    for (int pass = 0; pass < 2; pass++)
        for (size_t e = 0; e < prev.size(); e++) {
            RC((bdot<cd2, VT>(c, 1, prev[e], 0, x, q)));
            RC((baxpy<cd2, VT>(c, 1, prev[e], 0, q, nullptr, x)));
        }
    return ED_OK;
}

// The full-reorthogonalisation schedule of run_full for M instances at once.
// An instance stops (its next Krylov vectors are zero) once the residual
// estimate beta_{k+1} |s_k| of its lowest Ritz pair (s = eigenvector of the
// tridiagonal matrix) is below tol * max(1, |theta|), checked every 5 steps,
// or its Krylov space became invariant.  starts (optional): FP64 interleaved
// start vectors, one per eigenvalue; otherwise seeded random starts.
template <typename VT>
int run_block_full(BCtx& c, const ed_lanczos_params& p, double tol, const std::vector<cd2*>* starts,
                   std::vector<cd2*>& prev, double* evals) {
    const int64_t D = c.D, n = c.n;
    const int M = c.M;
    const int how_many = (int)std::min<int64_t>(p.nev, D);
    const int m_it = (int)std::min<int64_t>(p.iterations, D);
    VT *V = nullptr, *w = nullptr;
    HIPCHK(hipMalloc(&V, (size_t)(m_it + 1) * n * sizeof(VT)));
    HIPCHK(hipMalloc(&w, n * sizeof(VT)));
    cd2* x = nullptr;
    HIPCHK(hipMalloc(&x, n * sizeof(cd2)));
    std::vector<std::complex<double>> q;
    for (int num = 0; num < how_many; num++) {
        const int mit = (int)std::min<int64_t>(m_it, D - num);
        if (starts && num < (int)starts->size())
            hipLaunchKernelGGL((k_scale_copy<cd2, VT>), dim3(c.g), dim3(kRedThreads), 0, c.st, n, 1.0, (*starts)[num], V);
        else
            hipLaunchKernelGGL(k_bstart<VT>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.st, D, M,
                               p.seed + (uint64_t)num * 1000003ull, V);
        HIPCHK(hipGetLastError());
        RC(bdeflate_v<VT>(c, prev, V));
        std::vector<double> nn;
        RC(bnorm2(c, V, nn));
        std::vector<double> sc(M);
        for (int m = 0; m < M; m++) sc[m] = 1.0 / std::sqrt(nn[m]);
        RC((bscale<VT, VT>(c, sc, V, V)));
        std::vector<std::vector<double>> alphas(M), betas(M);
        std::vector<double> e0(M, 0.0);
        std::vector<int> nvs(M, 0);
        std::vector<char> active(M, 1);
        int k = 0;
        for (; k < mit; k++) {
            VT* vk = V + (int64_t)k * n;
            RC(bhv<VT>(c, vk, w));
            RC((bdot<VT, VT>(c, 1, vk, 0, w, q)));
            std::vector<double> alpha(M);
            for (int m = 0; m < M; m++) alpha[m] = q[m].real();
            std::vector<std::complex<double>> cf;
            if (k > 0) {
                cf.resize(2 * M);
                for (int m = 0; m < M; m++) {
                    cf[m] = betas[m].size() == (size_t)k ? betas[m].back() : 0.0;   // v_{k-1}
                    cf[M + m] = alpha[m];                                            // v_k
                }
                RC((baxpy<VT, VT>(c, 2, V + (int64_t)(k - 1) * n, n, cf, nullptr, w)));
            } else {
                cf.resize(M);
                for (int m = 0; m < M; m++) cf[m] = alpha[m];
                RC((baxpy<VT, VT>(c, 1, vk, 0, cf, nullptr, w)));
            }
            // two Gram-Schmidt passes; the second only when the first removed
            // more than 1 - 1/sqrt(2) of some instance's norm
            std::vector<double> nb2, bb;
            bool have_bb = false;
            RC(bnorm2(c, w, nb2));
            for (int pass = 0; pass < 2; pass++) {
                if (pass == 1) {
                    RC(bnorm2(c, w, bb));
                    bool need = false;
                    for (int m = 0; m < M; m++) need |= active[m] && bb[m] <= 0.5 * nb2[m];
                    if (!need) { have_bb = true; break; }
                }
                RC((bdot<VT, VT>(c, k + 1, V, n, w, q)));
                if (pass == 0)
                    for (int m = 0; m < M; m++) alpha[m] += q[(size_t)k * M + m].real();
                RC((baxpy<VT, VT>(c, k + 1, V, n, q, nullptr, w)));
                std::vector<std::complex<double>> qe;
                for (size_t e = 0; e < prev.size(); e++) {
                    RC((bdot<cd2, VT>(c, 1, prev[e], 0, w, qe)));
                    RC((baxpy<cd2, VT>(c, 1, prev[e], 0, qe, nullptr, w)));
                }
            }
            if (!have_bb) RC(bnorm2(c, w, bb));
            bool any = false;
            std::vector<double> scale(M, 0.0);
            for (int m = 0; m < M; m++) {
                if (!active[m]) continue;
                alphas[m].push_back(alpha[m]);
                nvs[m] = k + 1;
                const double beta = std::sqrt(bb[m]);
                bool stop = false;
                if ((k + 1) % 5 == 0 || k + 1 == mit) {
                    double ee;
                    std::vector<double> z;
                    if (!lowest(alphas[m], betas[m], ee, &z)) return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
                    e0[m] = ee;
                    if (beta * std::fabs(z.back()) < tol * std::max(1.0, std::fabs(ee))) stop = true;
                }
                if (beta < 1e-13 * std::max(1.0, std::fabs(alpha[m]))) stop = true;   // invariant subspace
                if (stop || k + 1 == mit) {
                    active[m] = 0;
                    continue;
                }
                betas[m].push_back(beta);
                scale[m] = 1.0 / beta;
                any = true;
            }
            if (!any) { k++; break; }
            RC((bscale<VT, VT>(c, scale, w, V + (int64_t)(k + 1) * n)));
        }
        // Ritz vectors: t[k*M + m] = lowest eigenvector of instance m's T
        const int kmax = *std::max_element(nvs.begin(), nvs.end());
        std::vector<double> t((size_t)kmax * M, 0.0);
        for (int m = 0; m < M; m++) {
            std::vector<double> tv, off(betas[m].begin(), betas[m].begin() + std::max(0, nvs[m] - 1));
            std::vector<double> al(alphas[m].begin(), alphas[m].begin() + nvs[m]);
            double ee;
            if (!lowest(al, off, ee, &tv)) return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
            e0[m] = ee;
            for (int j = 0; j < nvs[m]; j++) t[(size_t)j * M + m] = tv[j];
            evals[(size_t)m * how_many + num] = ee;
        }
        HIPCHK(hipMemsetAsync(x, 0, n * sizeof(cd2), c.st));
        RC(baccumulate<VT>(c, kmax, V, n, t, x));
        RC(bdeflate(c, prev, x));
        RC(bnorm2(c, x, nn));
        for (int m = 0; m < M; m++) sc[m] = 1.0 / std::sqrt(nn[m]);
        RC((bscale<cd2, cd2>(c, sc, x, x)));
        cd2* keep = nullptr;
        HIPCHK(hipMalloc(&keep, n * sizeof(cd2)));
        HIPCHK(hipMemcpyAsync(keep, x, n * sizeof(cd2), hipMemcpyDeviceToDevice, c.st));
        prev.push_back(keep);
    }
    HIPCHK(hipStreamSynchronize(c.st));
    dfree(V); dfree(w); dfree(x);
    return ED_OK;
}

// Thick-restart Lanczos (run_trl) for the M instances of a block operator in
// lock step: one block H*V per expansion step for all instances, the known
// part of H v_j removed first, then classical Gram-Schmidt against the
// instance's kept and new vectors (per-instance coefficients; a second pass
// when the first removed more than 1 - 1/sqrt(2) of any live instance's
// norm), the Rayleigh-Ritz of each instance's
// arrowhead-bordered tridiagonal, restart of every instance from its
// nev + (m - nev)/2 lowest Ritz vectors and its normalised residual, until
// every instance's nev lowest pairs have residual estimate
// beta_m |y_m,i| <= tol max(1, |theta_i|) or ncycles cycles have run.  An
// instance whose Krylov space becomes invariant (beta below the rounding
// level) keeps its projected matrix at that size: its later block columns
// are zero and add nothing.  With starts, instance m's start vector is the
// sum of the start vectors' columns m (the complex128 polish of complex64
// Ritz vectors); otherwise a seeded random vector.  Ritz vectors of the nev
// lowest pairs go to prev (interleaved, unit per instance), their values to
// evals[m * nev + i].
template <typename VT>
int run_block_trl(BCtx& c, const ed_lanczos_params& p, double tol, const std::vector<cd2*>* starts,
                  std::vector<cd2*>& prev, double* evals) {
    const int64_t D = c.D, n = c.n;
    const int M = c.M;
    const int nev = (int)std::min<int64_t>(p.nev, D);
    const int m = (int)std::min<int64_t>(p.iterations, D);
    if (m < nev + 1 && m < D) return ed_fail(ED_EINVAL, "thick restart needs iterations > nev (%d <= %d)", m, nev);
    if (m > c.cap) return ed_fail(ED_EINVAL, "too many Krylov vectors (%d)", m);
    const int kk = std::max(nev, std::min(nev + (m - nev) / 2, m - 1));
    VT *V = nullptr, *w = nullptr;
    cd2* W = nullptr;
    double* Tr = nullptr;   // restart coefficients
    auto cleanup = [&]() { dfree(V); dfree(w); dfree(W); dfree(Tr); };
#define TRC(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != ED_OK) {        \
            cleanup();             \
            return rc_;            \
        }                          \
    } while (0)
#define THIP(expr) TRC((expr) == hipSuccess ? ED_OK : ed_fail(ED_EDEVICE, "%s failed", #expr))
    THIP(hipMalloc(&V, (size_t)(m + 1) * n * sizeof(VT)));
    THIP(hipMalloc(&w, n * sizeof(VT)));
    THIP(hipMalloc(&W, (size_t)std::max(kk, nev) * n * sizeof(cd2)));
    THIP(hipMalloc(&Tr, (size_t)m * kk * M * sizeof(double)));
    auto Vk = [&](int k) { return V + (int64_t)k * n; };
    std::vector<std::complex<double>> q, q2;
    std::vector<double> nn, sc(M);
    // start vectors
    if (starts && !starts->empty()) {
        hipLaunchKernelGGL((k_scale_copy<cd2, VT>), dim3(c.g), dim3(kRedThreads), 0, c.st, n, 1.0, (*starts)[0], Vk(0));
        THIP(hipGetLastError());
        std::vector<std::complex<double>> minus1(M, std::complex<double>(-1.0, 0.0));
        for (size_t i = 1; i < starts->size(); i++) TRC((baxpy<cd2, VT>(c, 1, (*starts)[i], 0, minus1, nullptr, Vk(0))));
    } else {
        hipLaunchKernelGGL(k_bstart<VT>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.st, D, M, p.seed, Vk(0));
        THIP(hipGetLastError());
    }
    TRC(bnorm2(c, Vk(0), nn));
    for (int i = 0; i < M; i++) sc[i] = nn[i] > 0.0 ? 1.0 / std::sqrt(nn[i]) : 0.0;
    TRC((bscale<VT, VT>(c, sc, Vk(0), Vk(0))));
    // per instance: projected matrix (m x m, row-major), its valid size, whether
    // steps still extend it, the residual norm after the last step
    std::vector<std::vector<double>> T(M, std::vector<double>((size_t)m * m, 0.0));
    std::vector<int> meff(M, 0);
    std::vector<char> alive(M, 1);
    std::vector<double> beta_m(M, 0.0);
    std::vector<std::vector<double>> theta(M), Y(M);
    int start = 0, ncyc = 0;
    while (true) {
        for (int i = 0; i < M; i++)
            if (alive[i]) meff[i] = m;
        for (int j = start; j < m; j++) {
            TRC(bhv<VT>(c, Vk(j), w));
            // the known part of H v_j first (per instance): alpha_j v_j and
            // beta_{j-1} v_{j-1}, or after a restart the couplings to the kept
            // Ritz vectors
            TRC((bdot<VT, VT>(c, 1, Vk(j), 0, w, q)));
            std::vector<double> alpha(M);
            for (int i = 0; i < M; i++) alpha[i] = q[i].real();
            const int lo = j > start ? j - 1 : 0;
            std::vector<std::complex<double>> cf((size_t)(j + 1 - lo) * M);
            for (int i = 0; i < M; i++) {
                for (int l = lo; l < j; l++) cf[(size_t)(l - lo) * M + i] = T[i][(size_t)l * m + j];
                cf[(size_t)(j - lo) * M + i] = alpha[i];
            }
            TRC((baxpy<VT, VT>(c, j + 1 - lo, Vk(lo), n, cf, nullptr, w)));
            // Gram-Schmidt against every kept and new vector; the second pass
            // when the first removed more than 1 - 1/sqrt(2) of the norm of
            // any live instance
            std::vector<double> nb0;
            TRC(bnorm2(c, w, nb0));
            bool have_nn = false;
            for (int pass = 0; pass < 2; pass++) {
                if (pass == 1) {
                    TRC(bnorm2(c, w, nn));
                    bool need = false;
                    for (int i = 0; i < M; i++) need |= alive[i] && nn[i] <= 0.5 * nb0[i];
                    if (!need) { have_nn = true; break; }
                }
                TRC((bdot<VT, VT>(c, j + 1, V, n, w, q2)));
                for (int i = 0; i < M; i++) alpha[i] += q2[(size_t)j * M + i].real();
                TRC((baxpy<VT, VT>(c, j + 1, V, n, q2, nullptr, w)));
            }
            if (!have_nn) TRC(bnorm2(c, w, nn));
            for (int i = 0; i < M; i++) {
                sc[i] = 0.0;
                if (!alive[i]) continue;
                const double a = alpha[i];
                T[i][(size_t)j * m + j] = a;
                const double beta = std::sqrt(std::max(nn[i], 0.0));
                if (beta < 1e-13 * std::max(1.0, std::fabs(a))) {   // invariant subspace: the space is exact
                    meff[i] = j + 1;
                    beta_m[i] = 0.0;
                    alive[i] = 0;
                    continue;
                }
                sc[i] = 1.0 / beta;
                if (j + 1 < m) T[i][(size_t)j * m + j + 1] = T[i][(size_t)(j + 1) * m + j] = beta;
                else beta_m[i] = beta;   // V[m] holds the normalised residual
            }
            TRC((bscale<VT, VT>(c, sc, w, Vk(j + 1))));
        }
        // Rayleigh-Ritz per instance and the stopping rule
        bool conv = true;
        for (int i = 0; i < M; i++) {
            const int me = meff[i];
            std::vector<double> Tm((size_t)me * me);
            for (int a = 0; a < me; a++)
                for (int b = 0; b < me; b++) Tm[(size_t)a * me + b] = T[i][(size_t)a * m + b];
            if (!symeig_jacobi(Tm, me, theta[i], Y[i])) TRC(ed_fail(ED_ENOTCONV, "Rayleigh-Ritz eigensolver did not converge"));
            for (int k = 0; k < std::min(nev, me); k++)
                if (beta_m[i] * std::fabs(Y[i][(me - 1) + (size_t)k * me]) > tol * std::max(1.0, std::fabs(theta[i][k])))
                    conv = false;
        }
        ncyc++;
        if (conv || ncyc >= p.ncycles) break;
        // restart: W_k = sum_l Y[l, k] V_l per instance (k < kk; zero past an
        // exhausted instance's size), then V_k <- W_k and V_kk <- V_m
        const int kmax = *std::max_element(meff.begin(), meff.end());
        if (restart_fits(kmax, M)) {
            // the kept Ritz vectors of every instance in one pass over the basis
            std::vector<double> tr((size_t)kmax * kk * M, 0.0);
            for (int k = 0; k < kk; k++)
                for (int i = 0; i < M; i++)
                    if (k < meff[i])
                        for (int l = 0; l < meff[i]; l++)
                            tr[((size_t)l * kk + k) * M + i] = Y[i][l + (size_t)k * meff[i]];
            TRC((restart_basis<VT>(c.st, n, M, V, n, kmax, tr, kk, Tr, (VT*)W)));
        } else {
            std::vector<double> t((size_t)kmax * M);
            for (int k = 0; k < kk; k++) {
                std::fill(t.begin(), t.end(), 0.0);
                for (int i = 0; i < M; i++)
                    if (k < meff[i])
                        for (int l = 0; l < meff[i]; l++) t[(size_t)l * M + i] = Y[i][l + (size_t)k * meff[i]];
                cd2* Wk = W + (int64_t)k * n;
                THIP(hipMemsetAsync(Wk, 0, n * sizeof(cd2), c.st));
                TRC(baccumulate<VT>(c, kmax, V, n, t, Wk));
            }
            for (int k = 0; k < kk; k++)
                hipLaunchKernelGGL((k_scale_copy<cd2, VT>), dim3(c.g), dim3(kRedThreads), 0, c.st, n, 1.0,
                                   W + (int64_t)k * n, Vk(k));
            THIP(hipGetLastError());
        }
        THIP(hipMemcpyAsync(Vk(kk), Vk(m), n * sizeof(VT), hipMemcpyDeviceToDevice, c.st));
        for (int i = 0; i < M; i++) {
            std::vector<double>& Ti = T[i];
            std::fill(Ti.begin(), Ti.end(), 0.0);
            const int me = meff[i], kept = std::min(kk, me);
            for (int k = 0; k < kept; k++) {
                Ti[(size_t)k * m + k] = theta[i][k];
                if (alive[i]) Ti[(size_t)k * m + kk] = Ti[(size_t)kk * m + k] = beta_m[i] * Y[i][(me - 1) + (size_t)k * me];
            }
            if (!alive[i]) meff[i] = kept;   // an exhausted instance keeps its exact Ritz pairs only
        }
        start = kk;
    }
    // Ritz vectors of the nev lowest pairs, unit per instance
    const int kmax = *std::max_element(meff.begin(), meff.end());
    std::vector<double> t((size_t)kmax * M);
    for (int k = 0; k < nev; k++) {
        std::fill(t.begin(), t.end(), 0.0);
        for (int i = 0; i < M; i++)
            if (k < meff[i])
                for (int l = 0; l < meff[i]; l++) t[(size_t)l * M + i] = Y[i][l + (size_t)k * meff[i]];
        cd2* x = nullptr;
        THIP(hipMalloc(&x, n * sizeof(cd2)));
        prev.push_back(x);
        THIP(hipMemsetAsync(x, 0, n * sizeof(cd2), c.st));
        TRC(baccumulate<VT>(c, kmax, V, n, t, x));
        TRC(bnorm2(c, x, nn));
        for (int i = 0; i < M; i++) sc[i] = nn[i] > 0.0 ? 1.0 / std::sqrt(nn[i]) : 0.0;
        TRC((bscale<cd2, cd2>(c, sc, x, x)));
        for (int i = 0; i < M; i++) evals[(size_t)i * nev + k] = k < meff[i] ? theta[i][k] : 0.0;
    }
    THIP(hipStreamSynchronize(c.st));
    cleanup();
    return ED_OK;
#undef THIP
#undef TRC
}

// FP64 Rayleigh quotients and residual norms of unit interleaved vectors x.
int bfp64_check(BCtx& c, const cd2* x, std::vector<double>& rq, std::vector<double>& res) {
    cd2* y = nullptr;
    HIPCHK(hipMalloc(&y, c.n * sizeof(cd2)));
    RC(bhv<cd2>(c, x, y));
    std::vector<std::complex<double>> q;
    RC((bdot<cd2, cd2>(c, 1, x, 0, y, q)));
    rq.resize(c.M);
    for (int m = 0; m < c.M; m++) rq[m] = q[m].real();
    std::vector<std::complex<double>> a1(c.M);
    for (int m = 0; m < c.M; m++) a1[m] = rq[m];
    RC((baxpy<cd2, cd2>(c, 1, x, 0, a1, nullptr, y)));
    std::vector<double> rr;
    RC(bnorm2(c, y, rr));
    res.resize(c.M);
    for (int m = 0; m < c.M; m++) res[m] = std::sqrt(rr[m]);
    dfree(y);
    return ED_OK;
}

// ---- graph-captured full-reorthogonalisation Lanczos --------------------
// run_full with every scalar of a step on the device and the step captured
// once as a hipGraph, replayed for every k.  The step index k lives in device
// memory: kernels locate v_k = V + k D themselves and bound their loops by k,
// so one graph serves all steps.  H*v reads the fixed buffer vc, which the
// step's last kernel fills with v_{k+1} beside V[k+1].  Device scalars:
// ab[2k] = alpha_k, ab[2k+1] = beta_{k+1}; hc[0 .. nv) and hc[nv .. 2 nv) the
// two Gram-Schmidt passes' coefficients; hs[0] = <v_k, H v_k>, hs[1] = |w|^2
// before pass 1, hs[2] = |w|^2 after it, hs[3] = |w|^2 at the end, hs[4 + e]
// and hs[4 + np + e] the deflation coefficients of earlier eigenvector e in
// either pass.  With ED_LZ_SELECTIVE the second pass's kernels read hs[1],
// hs[2] and return at once unless hs[2] <= hs[1] / 2, the rule of run_full.
// The host reads ab[] every 5 steps, where run_full tests convergence, so the
// steps, stopping points and Ritz vectors are run_full's.
constexpr int kGChunk = 16;        // Krylov vectors per workgroup of k_gdot
constexpr int kFixed = INT32_MIN;  // a vector argument that does not move with k

__device__ __forceinline__ bool g_skip(const cd2* gate) { return gate && !(gate[2].re <= 0.5 * gate[1].re); }

// part[j * g + blk] = the block's share of <X_j, y>, j < n, X_j = X0 + j ldx,
// X0 = X + (k + xrel) ldx (xrel != kFixed), n = (nk ? k + 1 : 0) + nadd.
// Grid (g, chunks of kGChunk vectors); chunks at or past n exit at once.
// Each thread takes element pairs (2t, 2t+1): complex64 vectors then load
// 16 bytes per access, which streams at the HBM read rate where 8-byte
// accesses do not; X, y and ldx keep every vector 16-byte aligned.  An odd
// last element is added by thread 0 of workgroup column 0.
template <typename XT, typename YT>
__global__ void __launch_bounds__(kRedThreads) k_gdot(int64_t D, const XT* __restrict__ X, int64_t ldx, int xrel,
                                                       const YT* __restrict__ y, cd2* __restrict__ part,
                                                       const int* __restrict__ step, int nk, int nadd,
                                                       const cd2* gate) {
    if (g_skip(gate)) return;
    const int k = *step;
    const int n = (nk ? k + 1 : 0) + nadd;
    const int j0 = blockIdx.y * kGChunk;
    if (j0 >= n) return;
    const XT* X0 = (xrel == kFixed ? X : X + (int64_t)(k + xrel) * ldx) + (int64_t)j0 * ldx;
    const int nb = min(kGChunk, n - j0);
    __shared__ double red[kRedThreads / 64][kGChunk][2];
    double are[kGChunk], aim[kGChunk];
#pragma unroll
    for (int j = 0; j < kGChunk; j++) { are[j] = 0.0; aim[j] = 0.0; }
    const int64_t np = D >> 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < np; t += (int64_t)gridDim.x * blockDim.x) {
        cd2 y0, y1;
        ld2(y, t, y0, y1);
#pragma unroll
        for (int j = 0; j < kGChunk; j++) {
            if (j < nb) {
                cd2 x0, x1;
                ldn2(X0 + j * ldx, t, x0, x1);
                are[j] += x0.re * y0.re + x0.im * y0.im + x1.re * y1.re + x1.im * y1.im;
                aim[j] += x0.re * y0.im - x0.im * y0.re + x1.re * y1.im - x1.im * y1.re;
            }
        }
    }
    if ((D & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const cd2 yv = ld(y, D - 1);
        for (int j = 0; j < nb; j++) {
            const cd2 x = ld(X0 + j * ldx, D - 1);
            are[j] += x.re * yv.re + x.im * yv.im;
            aim[j] += x.re * yv.im - x.im * yv.re;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kGChunk; j++) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            are[j] += __shfl_xor(are[j], off, 64);
            aim[j] += __shfl_xor(aim[j], off, 64);
        }
        if (lane == 0) { red[wv][j][0] = are[j]; red[wv][j][1] = aim[j]; }
    }
    __syncthreads();
    if ((int)threadIdx.x < nb) {
        double r = 0.0, im = 0.0;
        for (int q = 0; q < kRedThreads / 64; q++) { r += red[q][threadIdx.x][0]; im += red[q][threadIdx.x][1]; }
        part[(int64_t)(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = cd2{r, im};
    }
}

// out[j] = compensated sum of part[j * g + b] over b (k_colsum's order),
// j < n as in k_gdot; one wave per column, columns at or past n exit.
__global__ void __launch_bounds__(64) k_gcolsum(const cd2* __restrict__ part, int g, cd2* __restrict__ out,
                                                const int* __restrict__ step, int nk, int nadd) {
    const int n = (nk ? *step + 1 : 0) + nadd;
    if ((int)blockIdx.x >= n) return;
    __shared__ double red[64][2];
    const int col = blockIdx.x, lane = threadIdx.x;
    double sr = 0, cr = 0, si = 0, ci = 0;
    for (int b = lane; b < g; b += 64) {
        const cd2 p = part[(int64_t)col * g + b];
        double yr = p.re - cr, tr = sr + yr;
        cr = (tr - sr) - yr; sr = tr;
        double yi = p.im - ci, tj = si + yi;
        ci = (tj - si) - yi; si = tj;
    }
    red[lane][0] = sr - cr;
    red[lane][1] = si - ci;
    __syncthreads();
    if (lane == 0) {
        sr = 0; cr = 0; si = 0; ci = 0;
        for (int l = 0; l < 64; l++) {
            double yr = red[l][0] - cr, tr = sr + yr;
            cr = (tr - sr) - yr; sr = tr;
            double yi = red[l][1] - ci, tj = si + yi;
            ci = (tj - si) - yi; si = tj;
        }
        out[col] = cd2{sr, si};
    }
}

// y -= sum_{j < n} c[j] X_j (n and X_j as in k_gdot), one element per thread
template <typename XT, typename YT>
__global__ void __launch_bounds__(kRedThreads) k_gaxpy(int64_t D, const XT* __restrict__ X, int64_t ldx, int xrel,
                                                        const cd2* __restrict__ c, YT* __restrict__ y,
                                                        const int* __restrict__ step, int nk, int nadd,
                                                        const cd2* gate) {
    if (g_skip(gate)) return;
    const int k = *step;
    const int n = (nk ? k + 1 : 0) + nadd;
    const XT* X0 = xrel == kFixed ? X : X + (int64_t)(k + xrel) * ldx;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const cd2 yv = ld(y, i);
    double re = yv.re, im = yv.im;
    int j = 0;
    for (; j + 4 <= n; j += 4) {
        cd2 x[4];
#pragma unroll
        for (int q = 0; q < 4; q++) x[q] = ldn(X0 + (int64_t)(j + q) * ldx, i);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const cd2 cf = c[j + q];
            re -= cf.re * x[q].re - cf.im * x[q].im;
            im -= cf.re * x[q].im + cf.im * x[q].re;
        }
    }
    for (; j < n; j++) {
        const cd2 x = ldn(X0 + (int64_t)j * ldx, i), cf = c[j];
        re -= cf.re * x.re - cf.im * x.im;
        im -= cf.re * x.im + cf.im * x.re;
    }
    st(y, i, re, im);
}

// w -= alpha0 v_k + beta_k v_{k-1}, alpha0 = Re hs[0], beta_0 = 0
template <typename VT>
__global__ void __launch_bounds__(kRedThreads) k_gthree(int64_t D, const VT* __restrict__ V, int64_t ldv,
                                                         VT* __restrict__ w, const cd2* __restrict__ hs,
                                                         const double* __restrict__ ab, const int* __restrict__ step) {
    const int k = *step;
    const double al = hs[0].re, be = k > 0 ? ab[2 * k - 1] : 0.0;
    const VT* vk = V + (int64_t)k * ldv;
    const VT* vm = V + (int64_t)(k > 0 ? k - 1 : 0) * ldv;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const cd2 x = ld(w, i), y = ld(vk, i), z = ld(vm, i);
    st(w, i, x.re - al * y.re - be * z.re, x.im - al * y.im - be * z.im);
}

// ab[2k] = alpha0 + Re(pass-1 coefficient of v_k), ab[2k+1] = |w| at the end
__global__ void k_gfinish(const cd2* __restrict__ hs, const cd2* __restrict__ h1, double* __restrict__ ab,
                          const int* __restrict__ step) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int k = *step;
    ab[2 * k] = hs[0].re + h1[k].re;
    ab[2 * k + 1] = sqrt(fmax(hs[3].re, 0.0));
}

// V[k+1] = vc = w / beta_{k+1}; zero once the Krylov space is invariant
// (beta < 1e-13 max(1, |alpha|), the stopping rule of run_full) or past the
// last vector slot (k + 1 >= nmax)
template <typename VT>
__global__ void __launch_bounds__(kRedThreads) k_gnext(int64_t D, const VT* __restrict__ w, VT* __restrict__ V,
                                                        int64_t ldv, VT* __restrict__ vc, const double* __restrict__ ab,
                                                        const int* __restrict__ step, int nmax) {
    const int k = *step;
    if (k + 1 >= nmax) return;
    const double al = ab[2 * k], be = ab[2 * k + 1];
    const double f = be >= 1e-13 * fmax(1.0, fabs(al)) ? 1.0 / be : 0.0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const cd2 x = ld(w, i);
    st(V + (int64_t)(k + 1) * ldv, i, f * x.re, f * x.im);
    st(vc, i, f * x.re, f * x.im);
}

template <typename VT>
int run_full_graph(Ctx& c, const ed_lanczos_params& p, std::vector<cd2*>& prev, double* evals) {
    const int64_t D = c.D;
    const int how_many = (int)std::min<int64_t>(p.nev, D);
    const unsigned g = grid_for(D), gs = grid_stream(D);
    const int m = (int)std::min<int64_t>(p.iterations, D);
    const int nv = ((m + 1 + kGChunk - 1) / kGChunk) * kGChunk;   // Krylov vector slots, rounded to chunks
    const int nchunk = nv / kGChunk;
    const int dt = std::is_same<VT, cd2>::value ? ED_C128 : ED_C64;
    const bool sel = (p.flags & ED_LZ_SELECTIVE) != 0;
    VT *V = nullptr, *w = nullptr, *vc = nullptr;
    cd2 *x = nullptr, *part = nullptr, *hc = nullptr, *hs = nullptr;
    double* ab = nullptr;
    int* step = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        dfree(V); dfree(w); dfree(vc); dfree(x); dfree(part); dfree(hc); dfree(hs); dfree(ab); dfree(step);
    };
#define GRC(expr)                      \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != ED_OK) {            \
            cleanup();                 \
            return rc_;                \
        }                              \
    } while (0)
#define GHIP(expr) GRC((expr) == hipSuccess ? ED_OK : ed_fail(ED_EDEVICE, "%s failed", #expr))
    // Krylov vector j at V + j ldv, ldv a multiple of 64 elements: every
    // vector 16-byte aligned for the paired loads of k_gdot
    const int64_t ldv = (D + 63) / 64 * 64;
    GHIP(hipMalloc(&V, (size_t)(m + 1) * ldv * sizeof(VT)));
    GHIP(hipMalloc(&w, D * sizeof(VT)));
    GHIP(hipMalloc(&vc, D * sizeof(VT)));
    GHIP(hipMalloc(&x, D * sizeof(cd2)));
    GHIP(hipMalloc(&part, (size_t)nv * g * sizeof(cd2)));
    GHIP(hipMalloc(&hc, 2 * (size_t)nv * sizeof(cd2)));
    GHIP(hipMalloc(&hs, (4 + 2 * (size_t)how_many) * sizeof(cd2)));
    GHIP(hipMalloc(&ab, 2 * (size_t)(m + 1) * sizeof(double)));
    GHIP(hipMalloc(&step, sizeof(int)));
    std::vector<std::complex<double>> q;
    std::vector<double> hab;
    for (int num = 0; num < how_many; num++) {
        const int mit = (int)std::min<int64_t>(m, D - num);
        const int np = (int)prev.size();
        hipLaunchKernelGGL(k_start<VT>, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c.st, D, p.seed + (uint64_t)num, V);
        GHIP(hipGetLastError());
        for (int pass = 0; pass < 2; pass++)
            for (int k = 0; k < np; k++) {
                GRC((multidot<cd2, VT>(c, 1, prev[k], 0, V, q)));
                GRC((multiaxpy<cd2, VT>(c, 1, prev[k], 0, q, 1.0, V)));
            }
        double nn;
        GRC(dnorm2(c, V, nn));
        hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), V, V);
        hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0, V, vc);
        GHIP(hipGetLastError());
        // two eager H*v: the first allocates the gather stream (and tunes the
        // kernel) outside the capture, the second times the product (the
        // graph's H*v share is this time per step)
        GRC(hv<VT>(c, vc, w));
        const double ms_before = c.hv_ms;
        GRC(hv<VT>(c, vc, w));
        const double hv_once = c.hv_ms - ms_before;
        const int zero = 0;
        GHIP(hipMemcpyAsync(step, &zero, sizeof(int), hipMemcpyHostToDevice, c.st));
        GHIP(hipStreamSynchronize(c.st));
        // capture one step
        hipGraph_t graph;
        GHIP(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
        {
            int rc = op_hv(c.op, dt, vc, w, c.st);
            auto dot1 = [&](const auto* X, int xrel, const auto* y, cd2* out, const cd2* gate) {
                using XT = std::remove_cv_t<std::remove_pointer_t<decltype(X)>>;
                using YT = std::remove_cv_t<std::remove_pointer_t<decltype(y)>>;
                hipLaunchKernelGGL((k_gdot<XT, YT>), dim3(g, 1), dim3(kRedThreads), 0, c.st, D, X, ldv, xrel, y, part,
                                   step, 0, 1, gate);
                hipLaunchKernelGGL(k_gcolsum, dim3(1), dim3(64), 0, c.st, part, (int)g, out, step, 0, 1);
            };
            // alpha0 = <v_k, w>, three-term update, |w|^2
            dot1(V, 0, w, hs + 0, nullptr);
            hipLaunchKernelGGL((k_gthree<VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, V, ldv, w, hs, ab, step);
            dot1(w, kFixed, w, hs + 1, nullptr);
            for (int pass = 0; pass < 2; pass++) {
                const cd2* gate = pass == 1 && sel ? hs : nullptr;
                cd2* h = hc + (size_t)pass * nv;
                hipLaunchKernelGGL((k_gdot<VT, VT>), dim3(g, nchunk), dim3(kRedThreads), 0, c.st, D, V, ldv, kFixed, w,
                                   part, step, 1, 0, gate);
                hipLaunchKernelGGL(k_gcolsum, dim3(nv), dim3(64), 0, c.st, part, (int)g, h, step, 1, 0);
                hipLaunchKernelGGL((k_gaxpy<VT, VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, V, ldv, kFixed, h, w, step,
                                   1, 0, gate);
                for (int e = 0; e < np; e++) {
                    cd2* he = hs + 4 + (size_t)pass * how_many + e;
                    dot1(prev[e], kFixed, w, he, gate);
                    hipLaunchKernelGGL((k_gaxpy<cd2, VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, prev[e], ldv, kFixed,
                                       he, w, step, 0, 1, gate);
                }
                if (pass == 0) dot1(w, kFixed, w, hs + 2, nullptr);
            }
            dot1(w, kFixed, w, hs + 3, nullptr);
            hipLaunchKernelGGL(k_gfinish, dim3(1), dim3(64), 0, c.st, hs, hc, ab, step);
            hipLaunchKernelGGL((k_gnext<VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, w, V, ldv, vc, ab, step, m + 1);
            hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, c.st, step);
            const hipError_t le = hipGetLastError();
            GHIP(hipStreamEndCapture(c.st, &graph));
            if (rc != ED_OK) { (void)hipGraphDestroy(graph); GRC(rc); }
            if (le != hipSuccess) { (void)hipGraphDestroy(graph); GRC(ed_fail(ED_EDEVICE, "launch in the captured step failed")); }
        }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        GHIP(ie);
        std::vector<double> alphas, betas;
        double e0 = 0.0, eprev = 1e300;
        int done = 0, nvs = 0;
        while (nvs == 0) {
            // replay up to the next multiple of 5 (or mit), where run_full checks
            const int from = done;
            const int upto = std::min(mit, (done / 5 + 1) * 5);
            for (; done < upto; done++) GHIP(hipGraphLaunch(exec, c.st));
            hab.resize(2 * (size_t)done);
            GHIP(hipMemcpyAsync(hab.data(), ab, hab.size() * sizeof(double), hipMemcpyDeviceToHost, c.st));
            GHIP(hipStreamSynchronize(c.st));
            alphas.assign(done, 0.0);
            betas.assign(done, 0.0);
            for (int k = 0; k < done; k++) { alphas[k] = hab[2 * k]; betas[k] = hab[2 * k + 1]; }
            for (int k = from; k < done && nvs == 0; k++)
                if (betas[k] < 1e-13 * std::max(1.0, std::fabs(alphas[k]))) nvs = k + 1;   // invariant subspace
            if (nvs == 0) {
                std::vector<double> off(betas.begin(), betas.begin() + done - 1);
                if (!lowest(alphas, off, e0, nullptr)) GRC(ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge"));
                if (std::fabs(e0 - eprev) < p.tol || done == mit) nvs = done;
                eprev = e0;
            }
            if (std::getenv("ED_PROGRESS")) std::fprintf(stderr, "[lanczos-graph-full] steps %d E0 %.15g\n", done, e0);
        }
        (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        c.hv_calls += done;
        c.hv_ms += hv_once * done;   // the replayed products, at the eager product's time
        std::vector<double> tvec, off(betas.begin(), betas.begin() + std::max(0, nvs - 1));
        std::vector<double> al(alphas.begin(), alphas.begin() + nvs);
        if (!lowest(al, off, e0, &tvec)) GRC(ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge"));
        GHIP(hipMemsetAsync(x, 0, D * sizeof(cd2), c.st));
        GRC((accumulate<VT>(c, nvs, V, ldv, tvec, x)));
        for (int pass = 0; pass < 2; pass++)
            for (int e = 0; e < np; e++) {
                GRC((multidot<cd2, cd2>(c, 1, prev[e], 0, x, q)));
                GRC((multiaxpy<cd2, cd2>(c, 1, prev[e], 0, q, 1.0, x)));
            }
        GRC(dnorm2(c, x, nn));
        hipLaunchKernelGGL((k_scale_copy<cd2, cd2>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), x, x);
        cd2* keep = nullptr;
        GHIP(hipMalloc(&keep, D * sizeof(cd2)));
        prev.push_back(keep);
        GHIP(hipMemcpyAsync(keep, x, D * sizeof(cd2), hipMemcpyDeviceToDevice, c.st));
        evals[num] = e0;
    }
    GHIP(hipStreamSynchronize(c.st));
#undef GHIP
#undef GRC
    cleanup();
    return ED_OK;
}

// ---- graph-captured reference-schedule Lanczos --------------------------
// run_reference with each step's scalars on the device and one step captured
// as a hipGraph (two graphs: the cycles without and the last cycle with the
// Gram-Schmidt pass over the cycle's vectors), replayed for every step.
// Device scalars: ab[2k] = Re alpha_k, ab[2k+1] = beta_k (|w| before the
// division); hs[0] = alpha_k (complex, as the reference subtracts it),
// hs[1] = |w|^2, hs[2] = |vp|^2 after the orthogonalisations, hs[4 + e] the
// deflation coefficient of earlier eigenvector e.  The first and the last
// cycle run all their steps without a host read; the other cycles are read
// every 5 steps, and the host applies the reference's test |E0 - E0_prev| <
// tol step by step to find where the cycle ends (the steps replayed past it
// are not used: the cycle's Ritz vector is built from the vectors up to that
// step and becomes the next start vector).

// w -= beta_{k-1} V[k-1]   (k > 0)
template <typename VT>
__global__ void __launch_bounds__(kRedThreads) k_rprev(int64_t D, const VT* __restrict__ V, int64_t ldv,
                                                        VT* __restrict__ w, const double* __restrict__ ab,
                                                        const int* __restrict__ step) {
    const int k = *step;
    if (k == 0) return;
    const double be = ab[2 * k - 1];
    const VT* vm = V + (int64_t)(k - 1) * ldv;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const cd2 x = ld(w, i), y = ld(vm, i);
    st(w, i, x.re - be * y.re, x.im - be * y.im);
}

// w -= alpha vp (alpha = hs[0], complex), V[k] = vp
template <typename VT>
__global__ void __launch_bounds__(kRedThreads) k_ralpha(int64_t D, const VT* __restrict__ vp, VT* __restrict__ w,
                                                         VT* __restrict__ V, int64_t ldv, const cd2* __restrict__ hs,
                                                         const int* __restrict__ step) {
    const int k = *step;
    const cd2 a = hs[0];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const cd2 x = ld(vp, i), y = ld(w, i);
    st(w, i, y.re - (a.re * x.re - a.im * x.im), y.im - (a.re * x.im + a.im * x.re));
    st(V + (int64_t)k * ldv, i, x.re, x.im);
}

// ab[2k] = Re alpha, ab[2k+1] = beta = |w|, vp = w / beta (0 if beta = 0)
template <typename VT>
__global__ void __launch_bounds__(kRedThreads) k_rbeta(int64_t D, const VT* __restrict__ w, VT* __restrict__ vp,
                                                        const cd2* __restrict__ hs, double* __restrict__ ab,
                                                        const int* __restrict__ step) {
    const double be = sqrt(fmax(hs[1].re, 0.0));
    const double f = be > 1e-300 ? 1.0 / be : 0.0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        const int k = *step;
        ab[2 * k] = hs[0].re;
        ab[2 * k + 1] = be;
    }
    if (i >= D) return;
    const cd2 x = ld(w, i);
    st(vp, i, f * x.re, f * x.im);
}

// vp /= |vp| (|vp|^2 = hs[2])
template <typename VT>
__global__ void __launch_bounds__(kRedThreads) k_rnorm(int64_t D, VT* __restrict__ vp, const cd2* __restrict__ hs) {
    const double nn = hs[2].re;
    const double f = nn > 1e-300 ? 1.0 / sqrt(nn) : 0.0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const cd2 x = ld(vp, i);
    st(vp, i, f * x.re, f * x.im);
}

template <typename VT>
int run_reference_graph(Ctx& c, const ed_lanczos_params& p, std::vector<cd2*>& prev, double* evals) {
    const int64_t D = c.D;
    const int how_many = (int)std::min<int64_t>(p.nev, D);
    int iterations = p.iterations;
    const unsigned g = grid_for(D), gs = grid_stream(D);
    const int maxit = (int)std::min<int64_t>(iterations, D);
    const int nv = ((std::max(maxit, 1) + kGChunk - 1) / kGChunk) * kGChunk;
    const int nchunk = nv / kGChunk;
    const int64_t ldv = (D + 63) / 64 * 64;   // 16-byte aligned vectors for the paired loads
    const int dt = std::is_same<VT, cd2>::value ? ED_C128 : ED_C64;
    VT *V = nullptr, *vp = nullptr, *w = nullptr;
    cd2 *x = nullptr, *part = nullptr, *hc = nullptr, *hs = nullptr;
    double* ab = nullptr;
    int* step = nullptr;
    hipGraphExec_t gx[2] = {nullptr, nullptr};   // [0] no Gram-Schmidt pass, [1] the last cycle's
    auto cleanup = [&]() {
        for (auto& e : gx)
            if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
        dfree(V); dfree(vp); dfree(w); dfree(x); dfree(part); dfree(hc); dfree(hs); dfree(ab); dfree(step);
    };
#define GRC(expr)                      \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != ED_OK) {            \
            cleanup();                 \
            return rc_;                \
        }                              \
    } while (0)
#define GHIP(expr) GRC((expr) == hipSuccess ? ED_OK : ed_fail(ED_EDEVICE, "%s failed", #expr))
    GHIP(hipMalloc(&V, (size_t)std::max(maxit, 1) * ldv * sizeof(VT)));
    GHIP(hipMalloc(&vp, D * sizeof(VT)));
    GHIP(hipMalloc(&w, D * sizeof(VT)));
    GHIP(hipMalloc(&x, D * sizeof(cd2)));
    GHIP(hipMalloc(&part, (size_t)nv * g * sizeof(cd2)));
    GHIP(hipMalloc(&hc, (size_t)nv * sizeof(cd2)));
    GHIP(hipMalloc(&hs, (4 + (size_t)how_many) * sizeof(cd2)));
    GHIP(hipMalloc(&ab, 2 * (size_t)std::max(maxit, 1) * sizeof(double)));
    GHIP(hipMalloc(&step, sizeof(int)));
    std::vector<std::complex<double>> q;
    std::vector<double> hab;
    auto dot1 = [&](const auto* X, const auto* y, cd2* out) {
        using XT = std::remove_cv_t<std::remove_pointer_t<decltype(X)>>;
        using YT = std::remove_cv_t<std::remove_pointer_t<decltype(y)>>;
        hipLaunchKernelGGL((k_gdot<XT, YT>), dim3(g, 1), dim3(kRedThreads), 0, c.st, D, X, ldv, kFixed, y, part, step,
                           0, 1, (const cd2*)nullptr);
        hipLaunchKernelGGL(k_gcolsum, dim3(1), dim3(64), 0, c.st, part, (int)g, out, step, 0, 1);
    };
    // T of the first n steps: diagonal ab[2j], off-diagonal ab[2j+1], j < n - 1
    auto tri = [&](int n, std::vector<double>& al, std::vector<double>& off) {
        al.resize(n);
        off.resize(std::max(0, n - 1));
        for (int j = 0; j < n; j++) al[j] = hab[2 * j];
        for (int j = 0; j + 1 < n; j++) off[j] = hab[2 * j + 1];
    };
    for (int num = 0; num < how_many; num++) {
        iterations = (int)std::min<int64_t>(iterations, D - num);
        const int np = (int)prev.size();
        hipLaunchKernelGGL(k_start<VT>, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, c.st, D, p.seed + (uint64_t)num, vp);
        GHIP(hipGetLastError());
        // two eager H*v: the first allocates the gather stream (and tunes the
        // kernel) outside the captures, the second times the product
        GRC(hv<VT>(c, vp, w));
        const double ms_before = c.hv_ms;
        GRC(hv<VT>(c, vp, w));
        const double hv_once = c.hv_ms - ms_before;
        for (int cg = 0; cg < 2; cg++) {
            if (cg == 0 && p.ncycles < 2) continue;
            hipGraph_t graph;
            GHIP(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
            const int rc = op_hv(c.op, dt, vp, w, c.st);
            hipLaunchKernelGGL((k_rprev<VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, V, ldv, w, ab, step);
            dot1(vp, w, hs + 0);
            hipLaunchKernelGGL((k_ralpha<VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, vp, w, V, ldv, hs, step);
            dot1(w, w, hs + 1);
            hipLaunchKernelGGL((k_rbeta<VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, w, vp, hs, ab, step);
            for (int e = 0; e < np; e++) {
                dot1(prev[e], vp, hs + 4 + e);
                hipLaunchKernelGGL((k_gaxpy<cd2, VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, prev[e], ldv, kFixed,
                                   hs + 4 + e, vp, step, 0, 1, (const cd2*)nullptr);
            }
            if (cg == 1) {
                // classical Gram-Schmidt against this cycle's vectors V[0..k]
                hipLaunchKernelGGL((k_gdot<VT, VT>), dim3(g, nchunk), dim3(kRedThreads), 0, c.st, D, V, ldv, kFixed, vp,
                                   part, step, 1, 0, (const cd2*)nullptr);
                hipLaunchKernelGGL(k_gcolsum, dim3(nv), dim3(64), 0, c.st, part, (int)g, hc, step, 1, 0);
                hipLaunchKernelGGL((k_gaxpy<VT, VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, V, ldv, kFixed, hc, vp,
                                   step, 1, 0, (const cd2*)nullptr);
            }
            dot1(vp, vp, hs + 2);
            hipLaunchKernelGGL((k_rnorm<VT>), dim3(gs), dim3(kRedThreads), 0, c.st, D, vp, hs);
            hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, c.st, step);
            const hipError_t le = hipGetLastError();
            GHIP(hipStreamEndCapture(c.st, &graph));
            if (rc != ED_OK) { (void)hipGraphDestroy(graph); GRC(rc); }
            if (le != hipSuccess) { (void)hipGraphDestroy(graph); GRC(ed_fail(ED_EDEVICE, "launch in the captured step failed")); }
            const hipError_t ie = hipGraphInstantiate(&gx[cg], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            GHIP(ie);
        }
        double previous = 1000.0, e0 = 0.0, nn = 0.0;
        std::vector<double> tvec, al, off;
        for (int cycle = 0; cycle < p.ncycles; cycle++) {
            // orth_wrt_previous_evecs + normalize (lanczos.cpp:199-200)
            for (int k = 0; k < np; k++) {
                GRC((multidot<cd2, VT>(c, 1, prev[k], 0, vp, q)));
                GRC((multiaxpy<cd2, VT>(c, 1, prev[k], 0, q, 1.0, vp)));
            }
            GRC(dnorm2(c, vp, nn));
            hipLaunchKernelGGL((k_scale_copy<VT, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), vp, vp);
            GHIP(hipGetLastError());
            const int zero = 0;
            GHIP(hipMemcpyAsync(step, &zero, sizeof(int), hipMemcpyHostToDevice, c.st));
            const bool last = cycle == p.ncycles - 1;
            hipGraphExec_t ex = gx[last ? 1 : 0];
            int done = 0, nvs = 0;
            bool converged = false;
            while (nvs == 0) {
                const int from = done;
                const int upto = (last || cycle == 0) ? iterations : std::min(iterations, done + 5);
                for (; done < upto; done++) GHIP(hipGraphLaunch(ex, c.st));
                hab.resize(2 * (size_t)done);
                GHIP(hipMemcpyAsync(hab.data(), ab, hab.size() * sizeof(double), hipMemcpyDeviceToHost, c.st));
                GHIP(hipStreamSynchronize(c.st));
                if (last || cycle == 0) {
                    nvs = done;
                    break;
                }
                // the reference's step-by-step test in a middle cycle
                for (int it = from; it < done && nvs == 0; it++) {
                    tri(it + 1, al, off);
                    if (!lowest(al, off, e0, nullptr)) GRC(ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge"));
                    if (it > 1 && std::fabs(e0 - previous) < p.tol) { nvs = it + 1; converged = true; }
                    previous = e0;
                }
                if (nvs == 0 && done == iterations) nvs = done;
            }
            c.hv_calls += done;
            c.hv_ms += hv_once * done;   // the replayed products, at the eager product's time
            if (std::getenv("ED_PROGRESS"))
                std::fprintf(stderr, "[lanczos-graph-reference] cycle %d steps %d (used %d)\n", cycle, done, nvs);
            // Ritz vector (lanczos.cpp:312-320), assembled in FP64
            tri(nvs, al, off);
            if (!lowest(al, off, e0, &tvec)) GRC(ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge"));
            GHIP(hipMemsetAsync(x, 0, D * sizeof(cd2), c.st));
            GRC((accumulate<VT>(c, nvs, V, ldv, tvec, x)));
            for (int k = 0; k < np; k++) {
                GRC((multidot<cd2, cd2>(c, 1, prev[k], 0, x, q)));
                GRC((multiaxpy<cd2, cd2>(c, 1, prev[k], 0, q, 1.0, x)));
            }
            GRC(dnorm2(c, x, nn));
            hipLaunchKernelGGL((k_scale_copy<cd2, cd2>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0 / std::sqrt(nn), x, x);
            hipLaunchKernelGGL((k_scale_copy<cd2, VT>), dim3(g), dim3(kRedThreads), 0, c.st, D, 1.0, x, vp);
            GHIP(hipGetLastError());
            if (converged) cycle = p.ncycles - 2;   // next: the last cycle
        }
        for (auto& e : gx)
            if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
        cd2* keep = nullptr;
        GHIP(hipMalloc(&keep, D * sizeof(cd2)));
        prev.push_back(keep);
        GHIP(hipMemcpyAsync(keep, x, D * sizeof(cd2), hipMemcpyDeviceToDevice, c.st));
        evals[num] = e0;
    }
    GHIP(hipStreamSynchronize(c.st));
#undef GHIP
#undef GRC
    cleanup();
    return ED_OK;
}

// ---- multi-sector block Lanczos ---------------------------------------
// run_block_full for every (sector, instance) of a multi-sector block in lock
// step.  Krylov vector k of all sectors is one contiguous n_all-element
// array (sector s at element offset off_s = row0_s * M, interleaved D_s x M).
// Each step is one H*V launch over the live sectors' row tiles, then the
// step's reductions and updates as segmented kernels: the live sectors' rows
// cut into chunks of kSegElems elements, one workgroup per chunk, per-chunk
// partial sums per (vector, instance) column, and one compensated sum per
// (sector, column) -- so a step costs the same number of launches and host
// round trips for 30 sectors as for one.  A sector leaves the launches once
// all its instances have stopped.  Same stopping rule and Ritz vectors as
// run_block_full; the second Gram-Schmidt pass runs for every live sector
// when one of them needs it.
constexpr int64_t kSegElems = int64_t(1) << 16;   // elements per chunk (a multiple of 64)

struct SegChunk {
    int32_t sec, pad;
    int64_t e0, e1;    // elements [e0, e1) of sector sec
};

template <int MB, typename XT, typename YT>
__global__ void __launch_bounds__(kRedThreads) k_sdot(const SegChunk* __restrict__ ch, int M,
                                                       const XT* __restrict__ X, int64_t ldx,
                                                       const YT* __restrict__ y, cd2* __restrict__ partial, int nch,
                                                       int kofs) {
    __shared__ double red[kRedThreads / 64][MB][8][2];
    const SegChunk c = ch[blockIdx.x];
    double are[MB], aim[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) { are[k] = 0.0; aim[k] = 0.0; }
    for (int64_t e = c.e0 + threadIdx.x; e < c.e1; e += blockDim.x) {
        cd2 yv = ld(y, e);
#pragma unroll
        for (int k = 0; k < MB; k++) {
            cd2 x = ldn(X + k * ldx, e);
            are[k] += x.re * yv.re + x.im * yv.im;
            aim[k] += x.re * yv.im - x.im * yv.re;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < MB; k++) {
        for (int off = 32; off >= M; off >>= 1) {
            are[k] += __shfl_xor(are[k], off, 64);
            aim[k] += __shfl_xor(aim[k], off, 64);
        }
        if (lane < M) { red[wv][k][lane][0] = are[k]; red[wv][k][lane][1] = aim[k]; }
    }
    __syncthreads();
    if ((int)threadIdx.x < MB * M) {
        const int k = threadIdx.x / M, m = threadIdx.x % M;
        double r = 0.0, im = 0.0;
        for (int q = 0; q < kRedThreads / 64; q++) { r += red[q][k][m][0]; im += red[q][k][m][1]; }
        partial[(int64_t)((kofs + k) * M + m) * nch + blockIdx.x] = cd2{r, im};
    }
}

// out[q * ncols + col] = compensated sum of partial[col * nch + c] over the
// chunks c0[q] <= c < c0[q + 1] of the q-th live sector (fixed order, as
// k_colsum).  Grid (ncols, live sectors), one wave each.
__global__ void __launch_bounds__(64) k_scolsum(const cd2* __restrict__ partial, int nch, const int* __restrict__ c0,
                                                int ncols, cd2* __restrict__ out) {
    __shared__ double red[64][2];
    const int col = blockIdx.x, q = blockIdx.y, lane = threadIdx.x;
    const int b0 = c0[q], b1 = c0[q + 1];
    double sr = 0, cr = 0, si = 0, ci = 0;
    for (int b = b0 + lane; b < b1; b += 64) {
        const cd2 p = partial[(int64_t)col * nch + b];
        double yr = p.re - cr, tr = sr + yr;
        cr = (tr - sr) - yr; sr = tr;
        double yi = p.im - ci, tj = si + yi;
        ci = (tj - si) - yi; si = tj;
    }
    red[lane][0] = sr - cr;
    red[lane][1] = si - ci;
    __syncthreads();
    if (lane == 0) {
        sr = 0; cr = 0; si = 0; ci = 0;
        for (int l = 0; l < 64; l++) {
            double yr = red[l][0] - cr, tr = sr + yr;
            cr = (tr - sr) - yr; sr = tr;
            double yi = red[l][1] - ci, tj = si + yi;
            ci = (tj - si) - yi; si = tj;
        }
        out[(int64_t)q * ncols + col] = cd2{sr, si};
    }
}

// y[e] = a[sec][m] y[e] - sum_k c[sec * cs + k*M + m] X_k[e] over the chunk
template <int MB, typename XT, typename YT>
__global__ void __launch_bounds__(kRedThreads) k_saxpy(const SegChunk* __restrict__ ch, int M,
                                                        const XT* __restrict__ X, int64_t ldx,
                                                        const cd2* __restrict__ c, int cs,
                                                        const double* __restrict__ a, YT* __restrict__ y) {
    const SegChunk sc = ch[blockIdx.x];
    const int m = threadIdx.x % M;
    cd2 cf[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) cf[k] = c[(int64_t)sc.sec * cs + k * M + m];
    const double am = a ? a[sc.sec * M + m] : 1.0;
    for (int64_t e = sc.e0 + threadIdx.x; e < sc.e1; e += blockDim.x) {
        cd2 yv = ld(y, e);
        double re = am * yv.re, im = am * yv.im;
#pragma unroll
        for (int k = 0; k < MB; k++) {
            cd2 x = ldn(X + k * ldx, e);
            re -= cf[k].re * x.re - cf[k].im * x.im;
            im -= cf[k].re * x.im + cf[k].im * x.re;
        }
        st(y, e, re, im);
    }
}

// y[e] = s[sec][m] x[e]
template <typename XT, typename YT>
__global__ void __launch_bounds__(kRedThreads) k_sscale(const SegChunk* __restrict__ ch, int M,
                                                         const double* __restrict__ sc, const XT* x, YT* y) {
    const SegChunk c = ch[blockIdx.x];
    const double f = sc[c.sec * M + threadIdx.x % M];
    for (int64_t e = c.e0 + threadIdx.x; e < c.e1; e += blockDim.x) {
        cd2 v = ld(x, e);
        st(y, e, f * v.re, f * v.im);
    }
}

// y[e] += sum_k t[sec * ts + k*M + m] X_k[e]   (Ritz vectors, y FP64)
template <int MB, typename XT>
__global__ void __launch_bounds__(kRedThreads) k_saccum(const SegChunk* __restrict__ ch, int M,
                                                         const XT* __restrict__ X, int64_t ldx,
                                                         const double* __restrict__ t, int ts, cd2* __restrict__ y) {
    const SegChunk c = ch[blockIdx.x];
    const int m = threadIdx.x % M;
    double tt[MB];
#pragma unroll
    for (int k = 0; k < MB; k++) tt[k] = t[(int64_t)c.sec * ts + k * M + m];
    for (int64_t e = c.e0 + threadIdx.x; e < c.e1; e += blockDim.x) {
        cd2 yv = y[e];
#pragma unroll
        for (int k = 0; k < MB; k++) {
            // a sector that stopped early never wrote its later Krylov
            // vectors: their coefficient is 0 and their memory is not read
            if (tt[k] == 0.0) continue;
            cd2 x = ldn(X + k * ldx, e);
            yv.re += tt[k] * x.re;
            yv.im += tt[k] * x.im;
        }
        y[e] = yv;
    }
}

// The segmented context: chunks of the live sectors (or of all sectors) and
// the buffers of the reductions.
struct SCtx {
    ed_mblock_s* mb;
    int S, M;
    std::vector<int64_t> off;        // [S + 1] element offsets
    int64_t n_all = 0;
    hipStream_t st;
    std::vector<SegChunk> ch;        // current chunk list (live sectors, in sector order)
    std::vector<int> c0;             // first chunk of each listed sector, [nlist + 1]
    std::vector<int> list;           // the listed sectors
    SegChunk* d_ch = nullptr;
    int* d_c0 = nullptr;
    int max_ch = 0;
    cd2* d_partial = nullptr;        // [kBCols * max_ch]
    cd2* d_sum = nullptr;            // [kBCols * S]
    cd2* d_coef = nullptr;           // [S * cap * M]
    double* d_t = nullptr;           // [S * cap * M]
    double* d_s = nullptr;           // [S * M]
    int cap = 0;
    double hv_ms = 0.0;
    int64_t hv_calls = 0;
    hipEvent_t e0, e1;
};

int s_list(SCtx& c, const std::vector<char>& live) {
    c.ch.clear();
    c.c0.clear();
    c.list.clear();
    for (int s = 0; s < c.S; s++) {
        if (!live[s]) continue;
        c.list.push_back(s);
        c.c0.push_back((int)c.ch.size());
        for (int64_t e = c.off[s]; e < c.off[s + 1]; e += kSegElems)
            c.ch.push_back(SegChunk{s, 0, e, std::min(e + kSegElems, c.off[s + 1])});
    }
    c.c0.push_back((int)c.ch.size());
    if ((int)c.ch.size() > c.max_ch) return ed_fail(ED_EINVAL, "chunk list exceeds its buffer");
    if (!c.ch.empty()) {
        HIPCHK(hipMemcpyAsync(c.d_ch, c.ch.data(), c.ch.size() * sizeof(SegChunk), hipMemcpyHostToDevice, c.st));
        HIPCHK(hipMemcpyAsync(c.d_c0, c.c0.data(), c.c0.size() * sizeof(int), hipMemcpyHostToDevice, c.st));
    }
    return ED_OK;
}

int pass_step(int left) { return left >= 16 ? 16 : left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1; }

// out[s][k*M + m] = <X_k, y>_(s, m) for k < mv and the listed sectors s
template <typename XT, typename YT>
int sdot(SCtx& c, int mv, const XT* X, int64_t ldx, const YT* y, std::vector<std::vector<std::complex<double>>>& out) {
    out.assign(c.S, std::vector<std::complex<double>>());
    for (int s : c.list) out[s].assign((size_t)mv * c.M, 0.0);
    const int nch = (int)c.ch.size();
    if (nch == 0) return ED_OK;
    std::vector<cd2> h;
    for (int k0 = 0; k0 < mv; k0 += kMaxMulti) {
        const int mm = std::min(kMaxMulti, mv - k0);
        for (int q = 0; q < mm;) {
            const int step = pass_step(mm - q);
            const XT* Xq = X + (int64_t)(k0 + q) * ldx;
#define ED_SDOT(MBV) hipLaunchKernelGGL((k_sdot<MBV, XT, YT>), dim3(nch), dim3(kRedThreads), 0, c.st, c.d_ch, c.M, Xq, ldx, y, c.d_partial, nch, q)
            switch (step) {
                case 16: ED_SDOT(16); break;
                case 8: ED_SDOT(8); break;
                case 4: ED_SDOT(4); break;
                case 2: ED_SDOT(2); break;
                default: ED_SDOT(1); break;
            }
#undef ED_SDOT
            HIPCHK(hipGetLastError());
            q += step;
        }
        const int ncols = mm * c.M;
        hipLaunchKernelGGL(k_scolsum, dim3(ncols, (unsigned)c.list.size()), dim3(64), 0, c.st, c.d_partial, nch, c.d_c0,
                           ncols, c.d_sum);
        HIPCHK(hipGetLastError());
        h.resize((size_t)ncols * c.list.size());
        HIPCHK(hipMemcpyAsync(h.data(), c.d_sum, h.size() * sizeof(cd2), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        for (size_t q = 0; q < c.list.size(); q++)
            for (int k = 0; k < ncols; k++)
                out[c.list[q]][(size_t)k0 * c.M + k] = std::complex<double>(h[q * ncols + k].re, h[q * ncols + k].im);
    }
    return ED_OK;
}

// y = a .* y - sum_k coef[s][k*M + m] X_k on the listed sectors (a per
// (sector, instance); nullptr = 1)
template <typename XT, typename YT>
int saxpy(SCtx& c, int mv, const XT* X, int64_t ldx, const std::vector<std::vector<std::complex<double>>>& coef,
          const std::vector<std::vector<double>>* a, YT* y) {
    const int nch = (int)c.ch.size();
    if (nch == 0 || mv == 0) return ED_OK;
    if (mv > c.cap) return ed_fail(ED_EINVAL, "too many vectors (%d) in one update", mv);
    const int cs = mv * c.M;
    std::vector<cd2> cc((size_t)c.S * cs, cd2{0.0, 0.0});
    for (int s : c.list)
        for (int k = 0; k < cs; k++) cc[(size_t)s * cs + k] = cd2{coef[s][k].real(), coef[s][k].imag()};
    HIPCHK(hipMemcpyAsync(c.d_coef, cc.data(), cc.size() * sizeof(cd2), hipMemcpyHostToDevice, c.st));
    double* da = nullptr;
    std::vector<double> ha;
    if (a) {
        ha.assign((size_t)c.S * c.M, 1.0);
        for (int s : c.list)
            for (int m = 0; m < c.M; m++) ha[(size_t)s * c.M + m] = (*a)[s][m];
        HIPCHK(hipMemcpyAsync(c.d_s, ha.data(), ha.size() * sizeof(double), hipMemcpyHostToDevice, c.st));
        da = c.d_s;
    }
    for (int q = 0; q < mv;) {
        const int step = pass_step(mv - q);
        const XT* Xq = X + (int64_t)q * ldx;
        const cd2* cq = c.d_coef + (size_t)q * c.M;
        double* aq = q == 0 ? da : nullptr;
#define ED_SAXPY(MBV) hipLaunchKernelGGL((k_saxpy<MBV, XT, YT>), dim3(nch), dim3(kRedThreads), 0, c.st, c.d_ch, c.M, Xq, ldx, cq, cs, aq, y)
        switch (step) {
            case 16: ED_SAXPY(16); break;
            case 8: ED_SAXPY(8); break;
            case 4: ED_SAXPY(4); break;
            case 2: ED_SAXPY(2); break;
            default: ED_SAXPY(1); break;
        }
#undef ED_SAXPY
        HIPCHK(hipGetLastError());
        q += step;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    return ED_OK;
}

template <typename XT, typename YT>
int sscale(SCtx& c, const std::vector<std::vector<double>>& sc, const XT* x, YT* y) {
    const int nch = (int)c.ch.size();
    if (nch == 0) return ED_OK;
    std::vector<double> hs((size_t)c.S * c.M, 0.0);
    for (int s : c.list)
        for (int m = 0; m < c.M; m++) hs[(size_t)s * c.M + m] = sc[s][m];
    HIPCHK(hipMemcpyAsync(c.d_s, hs.data(), hs.size() * sizeof(double), hipMemcpyHostToDevice, c.st));
    hipLaunchKernelGGL((k_sscale<XT, YT>), dim3(nch), dim3(kRedThreads), 0, c.st, c.d_ch, c.M, c.d_s, x, y);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.st));
    return ED_OK;
}

template <typename XT>
int saccum(SCtx& c, int mv, const XT* X, int64_t ldx, const std::vector<std::vector<double>>& t, cd2* y) {
    const int nch = (int)c.ch.size();
    if (nch == 0 || mv == 0) return ED_OK;
    if (mv > c.cap) return ed_fail(ED_EINVAL, "too many vectors (%d) in one update", mv);
    const int ts = mv * c.M;
    std::vector<double> ht((size_t)c.S * ts, 0.0);
    for (int s : c.list)
        for (int k = 0; k < ts && k < (int)t[s].size(); k++) ht[(size_t)s * ts + k] = t[s][k];
    HIPCHK(hipMemcpyAsync(c.d_t, ht.data(), ht.size() * sizeof(double), hipMemcpyHostToDevice, c.st));
    for (int q = 0; q < mv;) {
        const int step = pass_step(mv - q);
        const XT* Xq = X + (int64_t)q * ldx;
        const double* tq = c.d_t + (size_t)q * c.M;
#define ED_SACC(MBV) hipLaunchKernelGGL((k_saccum<MBV, XT>), dim3(nch), dim3(kRedThreads), 0, c.st, c.d_ch, c.M, Xq, ldx, tq, ts, y)
        switch (step) {
            case 16: ED_SACC(16); break;
            case 8: ED_SACC(8); break;
            case 4: ED_SACC(4); break;
            case 2: ED_SACC(2); break;
            default: ED_SACC(1); break;
        }
#undef ED_SACC
        HIPCHK(hipGetLastError());
        q += step;
    }
    HIPCHK(hipStreamSynchronize(c.st));
    return ED_OK;
}

template <typename T>
int snorm2(SCtx& c, const T* y, std::vector<std::vector<double>>& out) {
    std::vector<std::vector<std::complex<double>>> r;
    RC((sdot<T, T>(c, 1, y, 0, y, r)));
    out.assign(c.S, std::vector<double>(c.M, 0.0));
    for (int s : c.list)
        for (int m = 0; m < c.M; m++) out[s][m] = r[s][m].real();
    return ED_OK;
}

// x -= projections on the earlier eigenvectors (two passes), listed sectors
template <typename VT>
int sdeflate(SCtx& c, const std::vector<cd2*>& prev, VT* x) {
    std::vector<std::vector<std::complex<double>>> q;
    for (int pass = 0; pass < 2; pass++)
        for (size_t e = 0; e < prev.size(); e++) {
            RC((sdot<cd2, VT>(c, 1, prev[e], 0, x, q)));
            RC((saxpy<cd2, VT>(c, 1, prev[e], 0, q, nullptr, x)));
        }
    return ED_OK;
}

template <typename VT>
int shv(SCtx& c, const std::vector<char>& live, const VT* v, VT* w) {
    int rc = mblock_set_live(c.mb, live);
    if (rc != ED_OK) return rc;
    HIPCHK(hipEventRecord(c.e0, c.st));
    rc = mblock_op_hv(c.mb, std::is_same<VT, cd2>::value ? ED_C128 : ED_C64, v, w, c.st);
    if (rc != ED_OK) return rc;
    HIPCHK(hipEventRecord(c.e1, c.st));
    HIPCHK(hipEventSynchronize(c.e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c.e0, c.e1));
    c.hv_ms += ms;
    c.hv_calls++;
    return ED_OK;
}

template <typename VT>
int run_mblock_full(SCtx& c, const std::vector<int64_t>& Dsec, const ed_lanczos_params& p, double tol,
                    const std::vector<cd2*>* starts, std::vector<cd2*>& prev, double* evals) {
    const int S = c.S, M = c.M;
    const int64_t n_all = c.n_all;
    const int64_t Dmin = *std::min_element(Dsec.begin(), Dsec.end());
    const int64_t Dmax = *std::max_element(Dsec.begin(), Dsec.end());
    const int how_many = (int)std::min<int64_t>(p.nev, Dmin);
    const int m_it = (int)std::min<int64_t>(p.iterations, Dmax);
    const std::vector<char> all(S, 1);
    VT *V = nullptr, *w = nullptr;
    cd2* x = nullptr;
    HIPCHK(hipMalloc(&V, (size_t)(m_it + 1) * n_all * sizeof(VT)));
    HIPCHK(hipMalloc(&w, n_all * sizeof(VT)));
    HIPCHK(hipMalloc(&x, n_all * sizeof(cd2)));
    auto cleanup = [&]() { dfree(V); dfree(w); dfree(x); };
#define SRC(expr)                      \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != ED_OK) {            \
            cleanup();                 \
            return rc_;                \
        }                              \
    } while (0)
    std::vector<std::vector<std::complex<double>>> q;
    for (int num = 0; num < how_many; num++) {
        std::vector<int> mit(S);
        std::vector<std::vector<std::vector<double>>> alphas(S, std::vector<std::vector<double>>(M)),
            betas(S, std::vector<std::vector<double>>(M));
        std::vector<std::vector<int>> nvs(S, std::vector<int>(M, 0));
        std::vector<std::vector<char>> active(S, std::vector<char>(M, 1));
        std::vector<char> live(S, 1);
        SRC(s_list(c, all));
        for (int s = 0; s < S; s++) {
            mit[s] = (int)std::min<int64_t>(m_it, Dsec[s] - num);
            VT* v0 = V + c.off[s];
            const int64_t ns = c.off[s + 1] - c.off[s];
            if (starts && num < (int)starts->size())
                hipLaunchKernelGGL((k_scale_copy<cd2, VT>), dim3((unsigned)std::min<int64_t>(kRedBlocks, (ns + 255) / 256)),
                                   dim3(kRedThreads), 0, c.st, ns, 1.0, (*starts)[num] + c.off[s], v0);
            else
                hipLaunchKernelGGL(k_bstart<VT>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c.st, Dsec[s], M,
                                   p.seed + (uint64_t)num * 1000003ull, v0);
            SRC(hipGetLastError() == hipSuccess ? ED_OK : ed_fail(ED_EDEVICE, "start vector launch failed"));
        }
        SRC(sdeflate<VT>(c, prev, V));
        std::vector<std::vector<double>> nn, sc(S, std::vector<double>(M));
        SRC(snorm2(c, V, nn));
        for (int s = 0; s < S; s++)
            for (int m = 0; m < M; m++) sc[s][m] = 1.0 / std::sqrt(nn[s][m]);
        SRC((sscale<VT, VT>(c, sc, V, V)));
        for (int k = 0; k < m_it; k++) {
            if (std::find(live.begin(), live.end(), 1) == live.end()) break;
            SRC(s_list(c, live));
            VT* vk = V + (int64_t)k * n_all;
            SRC(shv<VT>(c, live, vk, w));
            SRC((sdot<VT, VT>(c, 1, vk, 0, w, q)));
            std::vector<std::vector<double>> alpha(S, std::vector<double>(M, 0.0));
            std::vector<std::vector<std::complex<double>>> cf(S);
            for (int s : c.list) {
                for (int m = 0; m < M; m++) alpha[s][m] = q[s][m].real();
                if (k > 0) {
                    cf[s].resize(2 * M);
                    for (int m = 0; m < M; m++) {
                        cf[s][m] = betas[s][m].size() == (size_t)k ? betas[s][m].back() : 0.0;   // v_{k-1}
                        cf[s][M + m] = alpha[s][m];                                                // v_k
                    }
                } else {
                    cf[s].resize(M);
                    for (int m = 0; m < M; m++) cf[s][m] = alpha[s][m];
                }
            }
            if (k > 0) SRC((saxpy<VT, VT>(c, 2, V + (int64_t)(k - 1) * n_all, n_all, cf, nullptr, w)));
            else SRC((saxpy<VT, VT>(c, 1, vk, 0, cf, nullptr, w)));
            // two Gram-Schmidt passes; the second only when the first removed
            // more than 1 - 1/sqrt(2) of some live instance's norm
            std::vector<std::vector<double>> nb2, bb;
            bool have_bb = false;
            SRC(snorm2(c, w, nb2));
            for (int pass = 0; pass < 2; pass++) {
                if (pass == 1) {
                    SRC(snorm2(c, w, bb));
                    bool need = false;
                    for (int s : c.list)
                        for (int m = 0; m < M; m++) need |= active[s][m] && bb[s][m] <= 0.5 * nb2[s][m];
                    if (!need) { have_bb = true; break; }
                }
                SRC((sdot<VT, VT>(c, k + 1, V, n_all, w, q)));
                if (pass == 0)
                    for (int s : c.list)
                        for (int m = 0; m < M; m++) alpha[s][m] += q[s][(size_t)k * M + m].real();
                SRC((saxpy<VT, VT>(c, k + 1, V, n_all, q, nullptr, w)));
                std::vector<std::vector<std::complex<double>>> qe;
                for (size_t e = 0; e < prev.size(); e++) {
                    SRC((sdot<cd2, VT>(c, 1, prev[e], 0, w, qe)));
                    SRC((saxpy<cd2, VT>(c, 1, prev[e], 0, qe, nullptr, w)));
                }
            }
            if (!have_bb) SRC(snorm2(c, w, bb));
            std::vector<std::vector<double>> scale(S, std::vector<double>(M, 0.0));
            for (int s : c.list) {
                bool any = false;
                for (int m = 0; m < M; m++) {
                    if (!active[s][m]) continue;
                    alphas[s][m].push_back(alpha[s][m]);
                    nvs[s][m] = k + 1;
                    const double beta = std::sqrt(bb[s][m]);
                    bool stop = false;
                    if ((k + 1) % 5 == 0 || k + 1 == mit[s]) {
                        double ee;
                        std::vector<double> z;
                        if (!lowest(alphas[s][m], betas[s][m], ee, &z)) {
                            cleanup();
                            return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
                        }
                        if (beta * std::fabs(z.back()) < tol * std::max(1.0, std::fabs(ee))) stop = true;
                    }
                    if (beta < 1e-13 * std::max(1.0, std::fabs(alpha[s][m]))) stop = true;   // invariant subspace
                    if (stop || k + 1 == mit[s]) {
                        active[s][m] = 0;
                        continue;
                    }
                    betas[s][m].push_back(beta);
                    scale[s][m] = 1.0 / beta;
                    any = true;
                }
                if (!any) live[s] = 0;
            }
            // the next Krylov vector of every sector still listed (zero for
            // its stopped instances)
            SRC((sscale<VT, VT>(c, scale, w, V + (int64_t)(k + 1) * n_all)));
        }
        // Ritz vectors: t[s][j*M + m] = lowest eigenvector of (s, m)'s T
        SRC(s_list(c, all));
        int kmax = 0;
        std::vector<std::vector<double>> t(S);
        for (int s = 0; s < S; s++) kmax = std::max(kmax, *std::max_element(nvs[s].begin(), nvs[s].end()));
        for (int s = 0; s < S; s++) {
            t[s].assign((size_t)kmax * M, 0.0);
            for (int m = 0; m < M; m++) {
                std::vector<double> tv, off(betas[s][m].begin(), betas[s][m].begin() + std::max(0, nvs[s][m] - 1));
                std::vector<double> al(alphas[s][m].begin(), alphas[s][m].begin() + nvs[s][m]);
                double ee;
                if (!lowest(al, off, ee, &tv)) {
                    cleanup();
                    return ed_fail(ED_ENOTCONV, "tridiagonal QL did not converge");
                }
                for (int j = 0; j < nvs[s][m]; j++) t[s][(size_t)j * M + m] = tv[j];
                evals[((size_t)s * M + m) * how_many + num] = ee;
            }
        }
        SRC(hipMemsetAsync(x, 0, n_all * sizeof(cd2), c.st) == hipSuccess ? ED_OK : ed_fail(ED_EDEVICE, "memset failed"));
        SRC(saccum<VT>(c, kmax, V, n_all, t, x));
        SRC(sdeflate<cd2>(c, prev, x));
        SRC(snorm2(c, x, nn));
        for (int s = 0; s < S; s++)
            for (int m = 0; m < M; m++) sc[s][m] = 1.0 / std::sqrt(nn[s][m]);
        SRC((sscale<cd2, cd2>(c, sc, x, x)));
        cd2* keep = nullptr;
        SRC(hipMalloc(&keep, n_all * sizeof(cd2)) == hipSuccess ? ED_OK : ed_fail(ED_ENOMEM, "Ritz vector allocation"));
        prev.push_back(keep);
        SRC(hipMemcpyAsync(keep, x, n_all * sizeof(cd2), hipMemcpyDeviceToDevice, c.st) == hipSuccess
                ? ED_OK : ed_fail(ED_EDEVICE, "copy failed"));
    }
#undef SRC
    HIPCHK(hipStreamSynchronize(c.st));
    cleanup();
    return ED_OK;
}

}  // namespace

extern "C" int ed_lanczos(ed_op op, const ed_lanczos_params* p, double* evals, double* evecs, ed_lanczos_stats* stats) {
    if (!op || !p || !evals) return ed_fail(ED_EINVAL, "null argument");
    if (p->nev < 1 || p->iterations < 1 || p->ncycles < 1) return ed_fail(ED_EINVAL, "nev, iterations, ncycles must be >= 1");
    if (p->precision != ED_C64 && p->precision != ED_C128) return ed_fail(ED_EINVAL, "bad precision %d", p->precision);
    const int64_t D = op->b->D;
    if (D == 0) return ed_fail(ED_EEMPTY, "No states in psector");
    auto t0 = std::chrono::steady_clock::now();
    Ctx c;
    c.op = op;
    c.D = D;
    // a private non-blocking stream: concurrent solves from several host
    // threads overlap instead of serialising on the legacy default stream
    HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    c.h_partial.resize((size_t)kRedBlocks * kMaxMulti);
    HIPCHK(hipMalloc(&c.d_partial, (size_t)kRedBlocks * kMaxMulti * sizeof(cd2)));
    HIPCHK(hipMalloc(&c.d_coef, kMaxMulti * sizeof(cd2)));
    HIPCHK(hipMalloc(&c.d_t, kMaxMulti * sizeof(double)));
    c.coef_cap = std::max(p->iterations, 1) + 64;
    HIPCHK(hipMalloc(&c.d_coef_big, c.coef_cap * sizeof(cd2)));
    HIPCHK(hipMalloc(&c.d_t_big, c.coef_cap * sizeof(double)));
    HIPCHK(hipEventCreate(&c.e0));
    HIPCHK(hipEventCreate(&c.e1));
    std::vector<cd2*> prev;
    int rc;
    if (p->mode == ED_LZ_LEAN) {
        if (p->nev != 1) {
            rc = ed_fail(ED_EINVAL, "the lean mode computes the lowest eigenvalue only (nev=1)");
        } else {
            cd2* x = nullptr;
            int steps = 0;
            if (p->flags & ED_LZ_GRAPH)
                rc = p->precision == ED_C128 ? run_lean_graph<cd2>(c, *p, x, evals[0], steps)
                                             : run_lean_graph<cf2>(c, *p, x, evals[0], steps);
            else
                rc = p->precision == ED_C128 ? run_lean<cd2>(c, *p, x, evals[0], steps)
                                             : run_lean<cf2>(c, *p, x, evals[0], steps);
            if (x) {
                // the complex64 gather stream is no longer needed: room for the
                // complex128 one of the FP64 check
                if (op->d_u64) { dfree(op->d_u64); op->d_u64 = nullptr; }
                if (op->d_img64) { dfree(op->d_img64); op->d_img64 = nullptr; }
                if (op->d_wimg64) { dfree(op->d_wimg64); op->d_wimg64 = nullptr; }
                prev.push_back(x);
            }
        }
    } else if (p->mode == ED_LZ_REFERENCE && (p->flags & ED_LZ_GRAPH))
        rc = p->precision == ED_C128 ? run_reference_graph<cd2>(c, *p, prev, evals)
                                     : run_reference_graph<cf2>(c, *p, prev, evals);
    else if (p->mode == ED_LZ_REFERENCE)
        rc = p->precision == ED_C128 ? run_reference<cd2>(c, *p, prev, evals) : run_reference<cf2>(c, *p, prev, evals);
    else if (p->mode == ED_LZ_THICK)
        rc = p->precision == ED_C128 ? run_trl<cd2>(c, *p, prev, evals) : run_trl<cf2>(c, *p, prev, evals);
    else if (p->flags & ED_LZ_GRAPH)
        rc = p->precision == ED_C128 ? run_full_graph<cd2>(c, *p, prev, evals)
                                     : run_full_graph<cf2>(c, *p, prev, evals);
    else
        rc = p->precision == ED_C128 ? run_full<cd2>(c, *p, prev, evals) : run_full<cf2>(c, *p, prev, evals);
    if (stats) std::memset(stats->eval_error, 0, sizeof(stats->eval_error));
    if (rc == ED_OK) {
        for (size_t k = 0; k < prev.size(); k++) {
            double rq = 0, res = 0;
            if (p->precision == ED_C64 || p->mode == ED_LZ_LEAN || (stats && k < 64)) {
                rc = fp64_check(c, prev[k], rq, res);
                if (rc != ED_OK) break;
                if (p->precision == ED_C64 || p->mode == ED_LZ_LEAN) evals[k] = rq;   // FP64 Rayleigh-quotient correction
                if (stats && k < 16) { stats->residual[k] = res; stats->rayleigh[k] = rq; }
                if (stats && k < 64) stats->eval_error[k] = res;
            }
            if (evecs) {
                // on the solve's stream, after the kernels that formed the vector
                hipError_t e = hipMemcpyAsync(evecs + 2 * D * k, prev[k], D * sizeof(cd2), hipMemcpyDeviceToHost, c.st);
                if (e == hipSuccess) e = hipStreamSynchronize(c.st);
                if (e != hipSuccess) { rc = ed_fail(ED_EDEVICE, "copy of eigenvector: %s", hipGetErrorString(e)); break; }
            }
        }
    }
    for (auto* v : prev) dfree(v);
    dfree(c.d_partial); dfree(c.d_coef); dfree(c.d_t); dfree(c.d_coef_big); dfree(c.d_t_big);
    (void)hipEventDestroy(c.e0);
    (void)hipEventDestroy(c.e1);
    (void)hipStreamDestroy(c.st);
    if (stats) {
        stats->hv_calls = c.hv_calls;
        stats->hv_seconds = c.hv_ms * 1e-3;
        stats->total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return rc;
}

extern "C" int ed_block_lanczos(ed_block blk, const ed_lanczos_params* p, double* evals, ed_lanczos_stats* stats) {
    if (!blk || !p || !evals) return ed_fail(ED_EINVAL, "null argument");
    if (p->nev < 1 || p->iterations < 1) return ed_fail(ED_EINVAL, "nev and iterations must be >= 1");
    if (p->precision != ED_C64 && p->precision != ED_C128) return ed_fail(ED_EINVAL, "bad precision %d", p->precision);
    const int M = block_size(blk);
    if (64 % M) return ed_fail(ED_EINVAL, "block Lanczos needs M in {1, 2, 4, 8}, got %d", M);
    const int64_t D = block_basis(blk)->D;
    if (D == 0) return ed_fail(ED_EEMPTY, "No states in psector");
    auto t0 = std::chrono::steady_clock::now();
    BCtx c;
    c.blk = blk;
    c.D = D;
    c.M = M;
    c.n = D * M;
    // a private non-blocking stream: concurrent solves from several host
    // threads overlap instead of serialising on the legacy default stream
    HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    c.g = (unsigned)std::min<int64_t>(kRedBlocks, (c.n + kRedThreads - 1) / kRedThreads);
    c.cap = std::max(p->iterations, 1) + 64;
    HIPCHK(hipMalloc(&c.d_partial, (size_t)kBCols * c.g * sizeof(cd2)));
    HIPCHK(hipMalloc(&c.d_sum, kBCols * sizeof(cd2)));
    HIPCHK(hipMalloc(&c.d_coef, (size_t)c.cap * M * sizeof(cd2)));
    HIPCHK(hipMalloc(&c.d_t, (size_t)c.cap * M * sizeof(double)));
    HIPCHK(hipMalloc(&c.d_s, 8 * sizeof(double)));
    HIPCHK(hipEventCreate(&c.e0));
    HIPCHK(hipEventCreate(&c.e1));
    std::vector<cd2*> prev;
    int rc;
    // ED_LZ_FULL: one full-reorthogonalisation run per eigenvalue, deflated
    // against the earlier ones; ED_LZ_THICK: the nev lowest pairs of every
    // instance from one thick-restarted Krylov space of `iterations` vectors
    const bool thick = p->mode == ED_LZ_THICK;
    auto run = [&](bool c128, double tl, const std::vector<cd2*>* st, std::vector<cd2*>& out) -> int {
        if (thick)
            return c128 ? run_block_trl<cd2>(c, *p, tl, st, out, evals) : run_block_trl<cf2>(c, *p, tl, st, out, evals);
        return c128 ? run_block_full<cd2>(c, *p, tl, st, out, evals) : run_block_full<cf2>(c, *p, tl, st, out, evals);
    };
    if (p->mode != ED_LZ_FULL && !thick) {
        rc = ed_fail(ED_EINVAL, "block Lanczos runs ED_LZ_FULL or ED_LZ_THICK, got mode %d", p->mode);
    } else if (p->precision == ED_C128) {
        rc = run(true, p->tol, nullptr, prev);
    } else {
        // complex64 Krylov vectors resolve residuals down to ~1e-6 relative;
        // when tol asks for less, the complex64 Ritz vectors start a short
        // complex128 run that polishes them to tol.
        const double floor64 = 1e-6;
        rc = run(false, std::max(p->tol, floor64), nullptr, prev);
        if (rc == ED_OK && p->tol < floor64) {
            std::vector<cd2*> coarse;
            coarse.swap(prev);
            rc = run(true, p->tol, &coarse, prev);
            for (auto* v : coarse) dfree(v);
        }
    }
    const int how_many = (int)std::min<int64_t>(p->nev, D);
    std::vector<double> err((size_t)M * how_many, 0.0);
    if (rc == ED_OK) {
        for (size_t k = 0; k < prev.size(); k++) {
            if (p->precision != ED_C64 && !stats) continue;
            std::vector<double> rq, res;
            rc = bfp64_check(c, prev[k], rq, res);
            if (rc != ED_OK) break;
            for (int m = 0; m < M; m++) {
                if (p->precision == ED_C64) evals[(size_t)m * how_many + k] = rq[m];
                if (stats && k == 0 && m < 16) { stats->residual[m] = res[m]; stats->rayleigh[m] = rq[m]; }
                if (k < (size_t)how_many) err[(size_t)m * how_many + k] = res[m];
            }
        }
        // each instance's eigenvalues in ascending order, their error bounds with them
        for (int m = 0; m < M; m++) {
            std::vector<std::pair<double, double>> pr(how_many);
            for (int k = 0; k < how_many; k++) pr[k] = {evals[(size_t)m * how_many + k], err[(size_t)m * how_many + k]};
            std::sort(pr.begin(), pr.end());
            for (int k = 0; k < how_many; k++) {
                evals[(size_t)m * how_many + k] = pr[k].first;
                err[(size_t)m * how_many + k] = pr[k].second;
            }
        }
    }
    if (stats) {
        std::memset(stats->eval_error, 0, sizeof(stats->eval_error));
        for (size_t q = 0; q < err.size() && q < 64; q++) stats->eval_error[q] = err[q];
    }
    for (auto* v : prev) dfree(v);
    dfree(c.d_partial); dfree(c.d_sum); dfree(c.d_coef); dfree(c.d_t); dfree(c.d_s);
    (void)hipEventDestroy(c.e0);
    (void)hipEventDestroy(c.e1);
    (void)hipStreamDestroy(c.st);
    if (stats) {
        stats->hv_calls = c.hv_calls;
        stats->hv_seconds = c.hv_ms * 1e-3;
        stats->total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return rc;
}

extern "C" int ed_mblock_lanczos(ed_mblock mb, const ed_lanczos_params* p, double* evals, double* errors,
                                 ed_lanczos_stats* stats) {
    if (!mb || !p || !evals) return ed_fail(ED_EINVAL, "null argument");
    if (p->nev < 1 || p->iterations < 1) return ed_fail(ED_EINVAL, "nev and iterations must be >= 1");
    if (p->precision != ED_C64 && p->precision != ED_C128) return ed_fail(ED_EINVAL, "bad precision %d", p->precision);
    const int S = mblock_nsec(mb), M = mblock_M(mb);
    if (64 % M) return ed_fail(ED_EINVAL, "block Lanczos needs M in {1, 2, 4, 8}, got %d", M);
    std::vector<int64_t> Dsec(S);
    for (int s = 0; s < S; s++) {
        Dsec[s] = block_basis(mblock_block(mb, s))->D;
        if (Dsec[s] == 0) return ed_fail(ED_EEMPTY, "sector %d has no states", s);
    }
    auto t0 = std::chrono::steady_clock::now();
    SCtx c;
    c.mb = mb;
    c.S = S;
    c.M = M;
    c.off.resize(S + 1);
    for (int s = 0; s <= S; s++) c.off[s] = mblock_row0(mb, s) * M;
    c.n_all = c.off[S];
    for (int s = 0; s < S; s++) c.max_ch += (int)((c.off[s + 1] - c.off[s] + kSegElems - 1) / kSegElems);
    c.cap = std::max(p->iterations, 1) + 64;
    HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c.e0));
    HIPCHK(hipEventCreate(&c.e1));
    int rc = ED_OK;
    std::vector<cd2*> prev;
    auto alloc = [&]() -> int {
        HIPCHK(hipMalloc(&c.d_ch, std::max(c.max_ch, 1) * sizeof(SegChunk)));
        HIPCHK(hipMalloc(&c.d_c0, (S + 1) * sizeof(int)));
        // also the partials of the per-sector FP64 checks (kRedBlocks workgroups)
        HIPCHK(hipMalloc(&c.d_partial, (size_t)kBCols * std::max(c.max_ch, kRedBlocks) * sizeof(cd2)));
        HIPCHK(hipMalloc(&c.d_sum, (size_t)kBCols * S * sizeof(cd2)));
        HIPCHK(hipMalloc(&c.d_coef, (size_t)S * c.cap * M * sizeof(cd2)));
        HIPCHK(hipMalloc(&c.d_t, (size_t)S * c.cap * M * sizeof(double)));
        HIPCHK(hipMalloc(&c.d_s, (size_t)S * M * sizeof(double)));
        return ED_OK;
    };
    rc = alloc();
    if (rc == ED_OK) {
        if (p->precision == ED_C128) {
            rc = run_mblock_full<cd2>(c, Dsec, *p, p->tol, nullptr, prev, evals);
        } else {
            // as ed_block_lanczos: complex64 to residual 1e-6, then a
            // complex128 run from its Ritz vectors when tol asks for less
            const double floor64 = 1e-6;
            rc = run_mblock_full<cf2>(c, Dsec, *p, std::max(p->tol, floor64), nullptr, prev, evals);
            if (rc == ED_OK && p->tol < floor64) {
                std::vector<cd2*> coarse;
                coarse.swap(prev);
                rc = run_mblock_full<cd2>(c, Dsec, *p, p->tol, &coarse, prev, evals);
                for (auto* v : coarse) dfree(v);
            }
        }
    }
    const int how_many = (int)std::min<int64_t>(p->nev, *std::min_element(Dsec.begin(), Dsec.end()));
    std::vector<double> err((size_t)S * M * how_many, 0.0);
    if (rc == ED_OK) {
        // the FP64 Rayleigh quotient and residual of every Ritz vector, sector
        // by sector through its block operator
        for (int s = 0; s < S && rc == ED_OK; s++) {
            BCtx b;
            b.blk = mblock_block(mb, s);
            b.D = Dsec[s];
            b.M = M;
            b.n = b.D * M;
            b.st = c.st;
            b.g = (unsigned)std::min<int64_t>(kRedBlocks, (b.n + kRedThreads - 1) / kRedThreads);
            b.cap = c.cap;
            b.d_partial = c.d_partial; b.d_sum = c.d_sum; b.d_coef = c.d_coef; b.d_t = c.d_t; b.d_s = c.d_s;
            b.e0 = c.e0; b.e1 = c.e1;
            for (size_t k = 0; k < prev.size(); k++) {
                std::vector<double> rq, res;
                rc = bfp64_check(b, prev[k] + c.off[s], rq, res);
                if (rc != ED_OK) break;
                for (int m = 0; m < M; m++) {
                    const size_t at = ((size_t)s * M + m) * how_many + k;
                    if (p->precision == ED_C64) evals[at] = rq[m];
                    err[at] = res[m];
                    const size_t sm = (size_t)s * M + m;
                    if (stats && k == 0 && sm < 16) { stats->residual[sm] = res[m]; stats->rayleigh[sm] = rq[m]; }
                }
            }
            c.hv_calls += b.hv_calls;
            c.hv_ms += b.hv_ms;
        }
        // each instance's eigenvalues ascending, their bounds with them
        for (size_t im = 0; rc == ED_OK && im < (size_t)S * M; im++) {
            std::vector<std::pair<double, double>> pr(how_many);
            for (int k = 0; k < how_many; k++) pr[k] = {evals[im * how_many + k], err[im * how_many + k]};
            std::sort(pr.begin(), pr.end());
            for (int k = 0; k < how_many; k++) {
                evals[im * how_many + k] = pr[k].first;
                err[im * how_many + k] = pr[k].second;
            }
        }
        if (errors) std::memcpy(errors, err.data(), err.size() * sizeof(double));
    }
    (void)mblock_set_live(mb, std::vector<char>(S, 1));
    for (auto* v : prev) dfree(v);
    for (void* q : {(void*)c.d_ch, (void*)c.d_c0, (void*)c.d_partial, (void*)c.d_sum, (void*)c.d_coef, (void*)c.d_t,
                    (void*)c.d_s})
        dfree(q);
    if (stats) {
        std::memset(stats->eval_error, 0, sizeof(stats->eval_error));
        for (size_t q = 0; q < err.size() && q < 64; q++) stats->eval_error[q] = err[q];
        stats->hv_calls = c.hv_calls;
        stats->hv_seconds = c.hv_ms * 1e-3;
        stats->total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    (void)hipEventDestroy(c.e0);
    (void)hipEventDestroy(c.e1);
    (void)hipStreamDestroy(c.st);
    return rc;
}
