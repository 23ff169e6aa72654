// ed_obs.hip -- what follows a solve, on the device: two-site spin
// correlations of momentum-basis vectors (without unwrapping) and the
// unwrapping of a momentum-basis vector to the full Sz basis.
//
// The reference's analysis programs (analysis/src/pspincorr.cpp:80-145 with
// szsz_op / sxsxsysy_op of analysis/src/chirality_exp.cpp:232-289) unwrap the
// eigenvectors to the full Sz basis and sum over its states for each pair
// (i, j).  For vectors of one momentum sector the pair expectation equals its
// average over the translation orbit of the pair, and that average is an
// operator of the sector:
//   Sz Sz:  diagonal, f(r) = (1/M) sum_{pairs p in the orbit} (b_i - 1/2)(b_j - 1/2)
//           of the representative r (the sum is the same on every orbit state);
//   (S+S- + S-S+)/2:  the XY part of a Hamiltonian whose bonds are the M
//           pairs of the orbit (J = 1, Delta = 0), divided by M,
// so one weighted dot product and one H*v with the sector's own kernels give
// <bra|O|ket> for a pair.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <utility>
#include <vector>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_internal.h"
#include "ed_tables.h"   // declares the C library's cexp

using namespace edc;

namespace {

constexpr int kObsBlock = 256;

// Per-block partial sums of conj(bra_a) ket_a w_a, where w_a is the Sz Sz
// orbit sum of representative a (MODE 0) or 1 (MODE 1, a plain dot product).
// Fixed reduction tree, so the result does not depend on timing.
template <int MODE>
__global__ void __launch_bounds__(kObsBlock) k_obs_dot(int64_t D, const uint64_t* __restrict__ rows,
                                                     const int2* __restrict__ pairs, int npairs,
                                                     const cd2* __restrict__ bra, const cd2* __restrict__ ket,
                                                     cd2* __restrict__ partial) {
    __shared__ double sr[kObsBlock], si[kObsBlock];
    double ar = 0.0, ai = 0.0;
    for (int64_t a = (int64_t)blockIdx.x * kObsBlock + threadIdx.x; a < D; a += (int64_t)gridDim.x * kObsBlock) {
        double wgt = 1.0;
        if (MODE == 0) {
            const uint64_t s = rows[a] & kStateMask;
            int acc = 0;   // sum of (2 b_i - 1)(2 b_j - 1) = 4 * sum (b_i - 1/2)(b_j - 1/2)
            for (int p = 0; p < npairs; p++) {
                const int bi = (int)((s >> pairs[p].x) & 1u), bj = (int)((s >> pairs[p].y) & 1u);
                acc += (bi == bj) ? 1 : -1;
            }
            wgt = 0.25 * (double)acc;
        }
        const cd2 b = bra[a], k = ket[a];
        // conj(b) * k * wgt
        ar += wgt * (b.re * k.re + b.im * k.im);
        ai += wgt * (b.re * k.im - b.im * k.re);
    }
    sr[threadIdx.x] = ar;
    si[threadIdx.x] = ai;
    __syncthreads();
    for (int h = kObsBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sr[threadIdx.x] += sr[threadIdx.x + h];
            si[threadIdx.x] += si[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = cd2{sr[0], si[0]};
}

std::vector<int> compose(const std::vector<int>& a, const std::vector<int>& b) {   // a after b
    std::vector<int> c(b.size());
    for (size_t i = 0; i < b.size(); i++) c[i] = a[b[i]];
    return c;
}

// Distinct unordered pairs {g(i), g(j)} over the translation grid.
std::vector<int2> pair_orbit(const ed_basis_s* B, int i, int j) {
    const int N = B->N;
    std::vector<int> g1(N);
    for (int q = 0; q < N; q++) g1[q] = q;
    std::set<std::pair<int, int>> seen;
    std::vector<int2> out;
    for (int a = 0; a < B->L1; a++) {
        std::vector<int> g2 = g1;
        for (int b = 0; b < B->L2; b++) {
            std::vector<int> g3 = g2;
            for (int c = 0; c < B->L3; c++) {
                int x = g3[i], y = g3[j];
                std::pair<int, int> key(std::min(x, y), std::max(x, y));
                if (seen.insert(key).second) out.push_back(make_int2(x, y));
                g3 = compose(B->T3, g3);
            }
            g2 = compose(B->T2, g2);
        }
        g1 = compose(B->T1, g1);
    }
    return out;
}

int obs_dot(const ed_basis_s* B, int mode, const int2* d_pairs, int npairs, const cd2* bra, const cd2* ket,
            cd2* d_partial, int nblk, double* out) {
    if (mode == 0)
        hipLaunchKernelGGL(k_obs_dot<0>, dim3(nblk), dim3(kObsBlock), 0, 0, B->D, B->d_rows, d_pairs, npairs, bra, ket,
                           d_partial);
    else
        hipLaunchKernelGGL(k_obs_dot<1>, dim3(nblk), dim3(kObsBlock), 0, 0, B->D, B->d_rows, d_pairs, npairs, bra, ket,
                           d_partial);
    HIPCHK(hipGetLastError());
    std::vector<cd2> h(nblk);
    HIPCHK(hipMemcpy(h.data(), d_partial, nblk * sizeof(cd2), hipMemcpyDeviceToHost));
    double re = 0.0, im = 0.0;
    for (const cd2& x : h) {
        re += x.re;
        im += x.im;
    }
    out[0] = re;
    out[1] = im;
    return ED_OK;
}

}  // namespace

extern "C" int ed_pair_correlations(ed_basis b, const void* bra, const void* ket, int32_t npairs, const int32_t* ij,
                                    double* out_zz, double* out_pm) {
    if (!b || npairs < 0 || (npairs > 0 && (!ij || !out_zz || !out_pm))) return ed_fail(ED_EINVAL, "null argument");
    const ed_basis_s* B = b;
    if (B->D > 0 && (!bra || !ket)) return ed_fail(ED_EINVAL, "null vector");
    for (int p = 0; p < npairs; p++) {
        if (ij[2 * p] < 0 || ij[2 * p] >= B->N || ij[2 * p + 1] < 0 || ij[2 * p + 1] >= B->N)
            return ed_fail(ED_EINVAL, "pair %d: site out of range", p);
        if (ij[2 * p] == ij[2 * p + 1]) return ed_fail(ED_EINVAL, "pair %d: the two sites coincide", p);
    }
    if (B->D == 0) {
        for (int p = 0; p < npairs; p++) out_zz[2 * p] = out_zz[2 * p + 1] = out_pm[2 * p] = out_pm[2 * p + 1] = 0.0;
        return ED_OK;
    }
    const int nblk = (int)std::min<int64_t>((B->D + kObsBlock - 1) / kObsBlock, 1024);
    int2* d_pairs = nullptr;
    cd2 *d_partial = nullptr, *d_w = nullptr;
    int rc = ED_OK;
    auto cleanup = [&]() {
        dfree(d_pairs);
        dfree(d_partial);
        dfree(d_w);
    };
    hipError_t e;
    if ((e = hipMalloc(&d_pairs, (size_t)B->L1 * B->L2 * B->L3 * sizeof(int2))) != hipSuccess ||
        (e = hipMalloc(&d_partial, nblk * sizeof(cd2))) != hipSuccess ||
        (e = hipMalloc(&d_w, B->D * sizeof(cd2))) != hipSuccess) {
        cleanup();
        return ed_fail(e == hipErrorOutOfMemory ? ED_ENOMEM : ED_EDEVICE, "ed_pair_correlations: %s",
                       hipGetErrorString(e));
    }
    for (int p = 0; p < npairs && rc == ED_OK; p++) {
        std::vector<int2> orb = pair_orbit(B, ij[2 * p], ij[2 * p + 1]);
        const int M = (int)orb.size();
        if ((e = hipMemcpy(d_pairs, orb.data(), M * sizeof(int2), hipMemcpyHostToDevice)) != hipSuccess) {
            rc = ed_fail(ED_EDEVICE, "ed_pair_correlations: %s", hipGetErrorString(e));
            break;
        }
        // Sz Sz: diagonal, orbit average
        double zz[2];
        rc = obs_dot(B, 0, d_pairs, M, (const cd2*)bra, (const cd2*)ket, d_partial, nblk, zz);
        if (rc != ED_OK) break;
        out_zz[2 * p] = zz[0] / M;
        out_zz[2 * p + 1] = zz[1] / M;
        // (S+S- + S-S+)/2: XY Hamiltonian of the orbit's bonds, then <bra|w>
        std::vector<int32_t> bxy(2 * M);
        std::vector<double> bJ(M, 1.0), bph(2 * M, 0.0);
        for (int q = 0; q < M; q++) {
            bxy[2 * q] = orb[q].x;
            bxy[2 * q + 1] = orb[q].y;
            bph[2 * q] = 1.0;
        }
        ed_terms t;
        std::memset(&t, 0, sizeof(t));
        t.lambda = 0.0;
        t.nbonds = M;
        t.bond_xy = bxy.data();
        t.bond_J = bJ.data();
        t.bond_phase = bph.data();
        ed_op op = nullptr;
        rc = ed_op_create(b, &t, &op);
        if (rc != ED_OK) break;
        rc = ed_hv(op, ED_C128, ket, d_w, nullptr);
        if (rc == ED_OK) {
            double pm[2];
            rc = obs_dot(B, 1, d_pairs, M, (const cd2*)bra, d_w, d_partial, nblk, pm);
            out_pm[2 * p] = pm[0] / M;
            out_pm[2 * p + 1] = pm[1] / M;
        }
        ed_op_destroy(op);
    }
    cleanup();
    return rc;
}

// --------------------------------------------------------------- unwrap ----
// Unwrap_pState (Momentum_Basis.cpp:978-1037; three generators :901-975):
// every state of each representative's orbit with coefficient
// v_a * ph / Norm_a, where ph is the running product of exp(-i px) per T1
// step, exp(-i py) per T2 step and exp(-i pz) per T3 step of the reference's
// walk (y cumulative over x, z cumulative over y), the first visit of a state
// kept (idsClass.cpp:33-53), entries with |re| or |im| above 1e-8, ascending
// by state.  On the device: pass 1 records the first walk step that reaches
// each state (atomicMin over the state's rank), pass 2 walks again and writes
// the coefficient at that step, pass 3 compacts in rank order, which is
// ascending state order.  Products and quotients are rounded one operation
// at a time, as the host's complex arithmetic does.
namespace {

constexpr int kNoVisit = 0x7fffffff;

struct Walk {
    TProg P1, P2, P3;
    int L1, L2, L3;
};

// Calls f(state, step) for every step of the reference walk from s.
template <class F>
__device__ void walk_orbit(uint64_t s, const Walk& W, F&& f) {
    uint64_t id = s;
    int k = 0;
    f(id, k++);
    for (int x = 0; x < W.L1; x++) {
        if (x > 0) {
            id = apply_T<uint64_t>(id, W.P1);
            f(id, k++);
        }
        for (int y = 0; y < W.L2; y++) {
            if (y > 0) {
                id = apply_T<uint64_t>(id, W.P2);
                f(id, k++);
            }
            for (int z = 1; z < W.L3; z++) {
                id = apply_T<uint64_t>(id, W.P3);
                f(id, k++);
            }
        }
    }
}

__global__ void k_fill_int(uint64_t n, int* p, int val) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

__global__ void k_unwrap_first(int64_t D, const uint64_t* __restrict__ rows, Walk W, RankInfo R,
                               const uint64_t* __restrict__ rtab, int* __restrict__ first) {
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < D; a += (int64_t)gridDim.x * blockDim.x) {
        walk_orbit(rows[a] & kStateMask, W, [&](uint64_t t, int k) {
            atomicMin(&first[rank_of<uint64_t>(t, R, rtab)], k);
        });
    }
}

// (a.re b.re - a.im b.im, a.re b.im + a.im b.re), each product and sum
// rounded on its own (no FMA), as the host's complex multiply.
__device__ inline cd2 cmul_rn(cd2 a, cd2 b) {
#pragma clang fp contract(off)
    const double rr = a.re * b.re, ii = a.im * b.im, ri = a.re * b.im, ir = a.im * b.re;
    return cd2{rr - ii, ri + ir};
}

__global__ void k_unwrap_fill(int64_t D, const uint64_t* __restrict__ rows, const double* __restrict__ ntab, Walk W,
                              RankInfo R, const uint64_t* __restrict__ rtab, cd2 ex, cd2 ey, cd2 ez,
                              const cd2* __restrict__ v, const int* __restrict__ first, cd2* __restrict__ coef,
                              int* __restrict__ keep) {
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < D; a += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t rw = rows[a];
        const double nrm = ntab[(int)(rw >> kLidShift)];
        const cd2 va = v[a];
        cd2 ph{1.0, 0.0};
        // the phase of step k: the product of the per-step factors so far
        uint64_t id = rw & kStateMask;
        int k = 0;
        auto emit = [&](uint64_t t, int step, bool first_step) {
            const uint64_t r = rank_of<uint64_t>(t, R, rtab);
            if (first[r] != step) return;
            cd2 c = first_step ? va : cmul_rn(va, ph);
            c.re = c.re / nrm;   // IEEE division, correctly rounded
            c.im = c.im / nrm;
            coef[r] = c;
            keep[r] = (fabs(c.re) > 1e-8 || fabs(c.im) > 1e-8) ? 1 : 0;
        };
        emit(id, k++, true);
        for (int x = 0; x < W.L1; x++) {
            if (x > 0) {
                id = apply_T<uint64_t>(id, W.P1);
                ph = cmul_rn(ph, ex);
                emit(id, k++, false);
            }
            for (int y = 0; y < W.L2; y++) {
                if (y > 0) {
                    id = apply_T<uint64_t>(id, W.P2);
                    ph = cmul_rn(ph, ey);
                    emit(id, k++, false);
                }
                for (int z = 1; z < W.L3; z++) {
                    id = apply_T<uint64_t>(id, W.P3);
                    ph = cmul_rn(ph, ez);
                    emit(id, k++, false);
                }
            }
        }
    }
}

__global__ void k_unwrap_scatter(uint64_t C, int N, int n, const uint64_t* __restrict__ binom,
                                 const int* __restrict__ keep, const unsigned long long* __restrict__ pos,
                                 const cd2* __restrict__ coef, int64_t* __restrict__ out_state,
                                 cd2* __restrict__ out_coef) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < C; r += (uint64_t)gridDim.x * blockDim.x) {
        if (!keep[r]) continue;
        const unsigned long long q = pos[r];
        out_state[q] = (int64_t)unrank(r, N, n, binom);
        out_coef[q] = coef[r];
    }
}

__global__ void k_to_u64(uint64_t n, const int* __restrict__ a, unsigned long long* __restrict__ b) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = (unsigned long long)a[i];
}

// exp(-i p) through the C library's cexp: std::exp(complex(0, -p)) of the
// reference's walk compiles to this call under g++, while clang lowers
// std::exp itself (cos and sin, which differ from cexp in the last bit).
cd2 cexp_neg(double p) {
    double _Complex z;
    __real__ z = 0.0;
    __imag__ z = -p;
    double _Complex e = cexp(z);
    return cd2{__real__ e, __imag__ e};
}

}  // namespace

extern "C" int ed_unwrap(ed_basis b, const void* v, int64_t capacity, int64_t* states, double* coeffs,
                         int64_t* count) {
    if (!b || !count) return ed_fail(ED_EINVAL, "null argument");
    const ed_basis_s* B = b;
    *count = 0;
    if (B->D == 0) return ED_OK;
    if (!v) return ed_fail(ED_EINVAL, "null vector");
    if (capacity > 0 && (!states || !coeffs)) return ed_fail(ED_EINVAL, "null output with capacity > 0");
    const uint64_t C = B->C;
    if (C > 0x7fffffffull) return ed_fail(ED_EINVAL, "unwrap of a sector with C(N,n) = %llu > 2^31 states",
                                          (unsigned long long)C);
    if ((uint64_t)B->L1 * B->L2 * B->L3 >= (uint64_t)kNoVisit) return ed_fail(ED_EINVAL, "walk too long");
    Walk W{B->lat_P1, B->lat_P2, B->lat_P3, B->L1, B->L2, B->L3};
    const cd2 ex = cexp_neg(B->px), ey = cexp_neg(B->py), ez = cexp_neg(B->pz);
    int *d_first = nullptr, *d_keep = nullptr;
    cd2 *d_coef = nullptr, *d_oc = nullptr;
    unsigned long long* d_pos = nullptr;
    int64_t* d_os = nullptr;
    void* d_tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = ED_OK;
    auto cleanup = [&]() {
        for (void* p : {(void*)d_first, (void*)d_keep, (void*)d_coef, (void*)d_oc, (void*)d_pos, (void*)d_os, d_tmp})
            dfree(p);
    };
#define ED_UW(x)                                                                                          \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess) {                                                                           \
            cleanup();                                                                                    \
            return ed_fail(e_ == hipErrorOutOfMemory ? ED_ENOMEM : ED_EDEVICE, "ed_unwrap: %s: %s", #x,    \
                           hipGetErrorString(e_));                                                        \
        }                                                                                                 \
    } while (0)
    ED_UW(hipMalloc(&d_first, C * sizeof(int)));
    ED_UW(hipMalloc(&d_keep, C * sizeof(int)));
    ED_UW(hipMalloc(&d_coef, C * sizeof(cd2)));
    ED_UW(hipMalloc(&d_pos, (C + 1) * sizeof(unsigned long long)));
    const unsigned gC = (unsigned)std::min<uint64_t>((C + 255) / 256, 1u << 20);
    const unsigned gD = (unsigned)std::min<int64_t>((B->D + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_fill_int, dim3(gC), dim3(256), 0, 0, C, d_first, kNoVisit);
    hipLaunchKernelGGL(k_fill_int, dim3(gC), dim3(256), 0, 0, C, d_keep, 0);
    ED_UW(hipGetLastError());
    hipLaunchKernelGGL(k_unwrap_first, dim3(gD), dim3(256), 0, 0, B->D, B->d_rows, W, B->rinfo, B->d_rtab, d_first);
    ED_UW(hipGetLastError());
    hipLaunchKernelGGL(k_unwrap_fill, dim3(gD), dim3(256), 0, 0, B->D, B->d_rows, B->d_ntab, W, B->rinfo, B->d_rtab,
                       ex, ey, ez, (const cd2*)v, d_first, d_coef, d_keep);
    ED_UW(hipGetLastError());
    // positions: exclusive sum of the keep flags (64-bit, C + 1 entries)
    hipLaunchKernelGGL(k_to_u64, dim3(gC), dim3(256), 0, 0, C, d_keep, d_pos);
    ED_UW(hipGetLastError());
    ED_UW(hipMemset(d_pos + C, 0, sizeof(unsigned long long)));
    ED_UW(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_pos, d_pos, (size_t)(C + 1)));
    ED_UW(hipMalloc(&d_tmp, tmp_bytes));
    ED_UW(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_pos, d_pos, (size_t)(C + 1)));
    unsigned long long total = 0;
    ED_UW(hipMemcpy(&total, d_pos + C, sizeof(total), hipMemcpyDeviceToHost));
    *count = (int64_t)total;
    if (capacity > 0 && total > 0) {
        if ((uint64_t)capacity < total) {
            cleanup();
            return ed_fail(ED_EINVAL, "ed_unwrap: capacity %lld < %llu entries", (long long)capacity, total);
        }
        ED_UW(hipMalloc(&d_os, total * sizeof(int64_t)));
        ED_UW(hipMalloc(&d_oc, total * sizeof(cd2)));
        hipLaunchKernelGGL(k_unwrap_scatter, dim3(gC), dim3(256), 0, 0, C, B->N, B->n, B->d_binom, d_keep, d_pos,
                           d_coef, d_os, d_oc);
        ED_UW(hipGetLastError());
        ED_UW(hipMemcpy(states, d_os, total * sizeof(int64_t), hipMemcpyDeviceToHost));
        ED_UW(hipMemcpy(coeffs, d_oc, total * sizeof(cd2), hipMemcpyDeviceToHost));
    }
#undef ED_UW
    cleanup();
    return rc;
}

// ------------------------------------------------------- density matrices ----
// Reduced density matrix of a cut A (sites cut[0..ncut)) from two
// momentum-basis vectors of one basis, the quantity of density_matrixij
// (analysis/src/Entropy.cpp:170-205, driven by analysis/src/dm.cpp:10-88):
//   dm(j, k) = sum_e psi1(j, e) conj(psi2(k, e)),
// j, k configurations of A (bit l of j = site cut[l]), e of the other sites,
// psi the unwrapped full-basis vectors with the unwrap threshold applied
// (entries with |re| and |im| <= 1e-8 are absent from the reference's EV
// files, so they count as zero).  The reference loops over 2^ncut x (unwrapped
// entries) with a binary search per term; here the unwrapped vector is laid
// out once as one matrix per count a of up spins in A,
//   Psi_a[row, col]:  row = rank of j among the C(ncut, a) weight-a
//                     configurations of A, col = rank of e among the
//                     C(N-ncut, n-a) configurations of the rest,
// (every weight-n state is one entry of one Psi_a), and each diagonal block
// of the matrix (states of equal Sz in A; the others vanish) is one complex
// GEMM, rho_a = Psi1_a Psi2_a^H, on rocBLAS.
namespace {

struct DmCut {
    int32_t ncut, nenv;
    int32_t cut[64];      // site of cut bit l
    int32_t env[64];      // the other sites, ascending
};

// combinatorial-number-system rank of the nb-bit word x: sum_t C(p_t, t)
__device__ __forceinline__ uint64_t cns_rank(uint64_t x, const uint64_t* binom) {
    uint64_t r = 0;
    int t = 0;
    while (x) {
        const int p = __ffsll((unsigned long long)x) - 1;
        t++;
        r += binom[p * 65 + t];
        x &= x - 1;
    }
    return r;
}

__global__ void k_dm_layout(uint64_t C, int N, int n, const uint64_t* __restrict__ binom, DmCut cut,
                            const uint64_t* __restrict__ off, const uint64_t* __restrict__ nrow,
                            const cd2* __restrict__ coef, const int* __restrict__ keep, cd2* __restrict__ psi) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < C; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = unrank(r, N, n, binom);
        uint64_t j = 0, e = 0;
        for (int l = 0; l < cut.ncut; l++) j |= ((s >> cut.cut[l]) & 1ull) << l;
        for (int l = 0; l < cut.nenv; l++) e |= ((s >> cut.env[l]) & 1ull) << l;
        const int a = __popcll(j);
        const uint64_t row = cns_rank(j, binom), col = cns_rank(e, binom);
        psi[off[a] + row + col * nrow[a]] = keep[r] ? coef[r] : cd2{0.0, 0.0};
    }
}

__global__ void k_zero_cd2(uint64_t n, cd2* p) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = cd2{0.0, 0.0};
}

// The unwrapped vector by state rank (coef and the threshold flags keep, [C]
// each): the first two passes of ed_unwrap.  States of orbits outside the
// sector keep coef 0, keep 0.
int unwrap_dense(const ed_basis_s* B, const cd2* v, int* d_first, cd2* d_coef, int* d_keep) {
    const uint64_t C = B->C;
    Walk W{B->lat_P1, B->lat_P2, B->lat_P3, B->L1, B->L2, B->L3};
    const cd2 ex = cexp_neg(B->px), ey = cexp_neg(B->py), ez = cexp_neg(B->pz);
    const unsigned gC = (unsigned)std::min<uint64_t>((C + 255) / 256, 1u << 20);
    const unsigned gD = (unsigned)std::min<int64_t>((B->D + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_fill_int, dim3(gC), dim3(256), 0, 0, C, d_first, kNoVisit);
    hipLaunchKernelGGL(k_fill_int, dim3(gC), dim3(256), 0, 0, C, d_keep, 0);
    hipLaunchKernelGGL(k_zero_cd2, dim3(gC), dim3(256), 0, 0, C, d_coef);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_unwrap_first, dim3(gD), dim3(256), 0, 0, B->D, B->d_rows, W, B->rinfo, B->d_rtab, d_first);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_unwrap_fill, dim3(gD), dim3(256), 0, 0, B->D, B->d_rows, B->d_ntab, W, B->rinfo, B->d_rtab,
                       ex, ey, ez, v, d_first, d_coef, d_keep);
    HIPCHK(hipGetLastError());
    return ED_OK;
}

}  // namespace

extern "C" int ed_density_matrix(ed_basis b, const void* v1, const void* v2, int32_t ncut, const int32_t* cut,
                                 int32_t n_cut_ones, double* dm) {
    if (!b || !cut || !dm) return ed_fail(ED_EINVAL, "null argument");
    const ed_basis_s* B = b;
    if (B->D > 0 && (!v1 || !v2)) return ed_fail(ED_EINVAL, "null vector");
    if (B->L3 != 1) return ed_fail(ED_EINVAL, "density matrices of three-generator sectors are not supported");
    const int N = B->N, n = B->n;
    if (ncut < 1 || ncut > N || ncut > 30) return ed_fail(ED_EINVAL, "cut of %d sites (1..min(N, 30))", ncut);
    DmCut dc;
    std::memset(&dc, 0, sizeof(dc));
    dc.ncut = ncut;
    std::vector<int> in(N, 0);
    for (int l = 0; l < ncut; l++) {
        if (cut[l] < 0 || cut[l] >= N || in[cut[l]]) return ed_fail(ED_EINVAL, "cut site %d out of range or repeated", cut[l]);
        in[cut[l]] = 1;
        dc.cut[l] = cut[l];
    }
    for (int s = 0; s < N; s++)
        if (!in[s]) dc.env[dc.nenv++] = s;
    if (n_cut_ones < -1 || n_cut_ones > ncut) return ed_fail(ED_EINVAL, "n_cut_ones %d out of range", n_cut_ones);
    // blocks: a up spins in A, rows C(ncut, a), columns C(N - ncut, n - a)
    std::vector<uint64_t> off(ncut + 2, 0), nrow(ncut + 1, 0), ncol(ncut + 1, 0);
    for (int a = 0; a <= ncut; a++) {
        nrow[a] = binom_u64(ncut, a);
        ncol[a] = (n - a >= 0 && n - a <= dc.nenv) ? binom_u64(dc.nenv, n - a) : 0;
        off[a + 1] = off[a] + nrow[a] * ncol[a];
    }
    const uint64_t C = B->C;
    if (off[ncut + 1] != C) return ed_fail(ED_EDEVICE, "density-matrix layout covers %llu of %llu states",
                                           (unsigned long long)off[ncut + 1], (unsigned long long)C);
    const uint64_t ndm = n_cut_ones < 0 ? (uint64_t(1) << ncut) : nrow[n_cut_ones];
    if (ndm > 32768) return ed_fail(ED_EINVAL, "density matrix of %llu x %llu exceeds 32768 rows", (unsigned long long)ndm,
                                    (unsigned long long)ndm);
    std::memset(dm, 0, ndm * ndm * 2 * sizeof(double));
    if (B->D == 0) return ED_OK;
    if (C > 0x7fffffffull) return ed_fail(ED_EINVAL, "density matrix of a sector with C(N,n) = %llu > 2^31 states",
                                          (unsigned long long)C);
    if ((uint64_t)B->L1 * B->L2 >= (uint64_t)kNoVisit) return ed_fail(ED_EINVAL, "walk too long");
    uint64_t maxblk = 0;
    for (int a = 0; a <= ncut; a++)
        if (n_cut_ones < 0 || a == n_cut_ones) maxblk = std::max(maxblk, nrow[a] * nrow[a]);
    int *d_first = nullptr, *d_keep = nullptr;
    cd2 *d_coef = nullptr, *d_psi1 = nullptr, *d_psi2 = nullptr, *d_rho = nullptr;
    uint64_t *d_off = nullptr, *d_nrow = nullptr;
    rocblas_handle h = nullptr;
    std::vector<cd2> blk;
    auto body = [&]() -> int {
        HIPCHK(hipMalloc(&d_first, C * sizeof(int)));
        HIPCHK(hipMalloc(&d_keep, C * sizeof(int)));
        HIPCHK(hipMalloc(&d_coef, C * sizeof(cd2)));
        HIPCHK(hipMalloc(&d_psi1, C * sizeof(cd2)));
        HIPCHK(hipMalloc(&d_psi2, C * sizeof(cd2)));
        HIPCHK(hipMalloc(&d_rho, std::max<uint64_t>(maxblk, 1) * sizeof(cd2)));
        HIPCHK(hipMalloc(&d_off, (ncut + 2) * sizeof(uint64_t)));
        HIPCHK(hipMalloc(&d_nrow, (ncut + 1) * sizeof(uint64_t)));
        HIPCHK(hipMemcpy(d_off, off.data(), (ncut + 2) * sizeof(uint64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_nrow, nrow.data(), (ncut + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
        const unsigned gC = (unsigned)std::min<uint64_t>((C + 255) / 256, 1u << 20);
        for (int q = 0; q < 2; q++) {
            int rc = unwrap_dense(B, (const cd2*)(q == 0 ? v1 : v2), d_first, d_coef, d_keep);
            if (rc != ED_OK) return rc;
            hipLaunchKernelGGL(k_dm_layout, dim3(gC), dim3(256), 0, 0, C, N, n, B->d_binom, dc, d_off, d_nrow, d_coef,
                               d_keep, q == 0 ? d_psi1 : d_psi2);
            HIPCHK(hipGetLastError());
        }
        if (rocblas_create_handle(&h) != rocblas_status_success) return ed_fail(ED_EDEVICE, "rocblas_create_handle failed");
        // the rank order of each weight's configurations (row r of block a = cfg[a][r])
        std::vector<std::vector<uint32_t>> cfg(ncut + 1);
        if (n_cut_ones < 0)
            for (uint64_t x = 0; x < (uint64_t(1) << ncut); x++) cfg[__builtin_popcountll(x)].push_back((uint32_t)x);
        const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
        for (int a = 0; a <= ncut; a++) {
            if (n_cut_ones >= 0 && a != n_cut_ones) continue;
            const uint64_t m = nrow[a], kk = ncol[a];
            if (m == 0 || kk == 0) continue;
            if (m > 0x7fffffffull || kk > 0x7fffffffull) return ed_fail(ED_EINVAL, "density-matrix block too large");
            // rho_a = Psi1_a Psi2_a^H (column-major, leading dimension m)
            rocblas_status rs = rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose,
                                              (rocblas_int)m, (rocblas_int)m, (rocblas_int)kk, &one,
                                              (const rocblas_double_complex*)(d_psi1 + off[a]), (rocblas_int)m,
                                              (const rocblas_double_complex*)(d_psi2 + off[a]), (rocblas_int)m, &zero,
                                              (rocblas_double_complex*)d_rho, (rocblas_int)m);
            if (rs != rocblas_status_success) return ed_fail(ED_EDEVICE, "rocblas_zgemm failed (%d)", (int)rs);
            blk.resize(m * m);
            HIPCHK(hipMemcpy(blk.data(), d_rho, m * m * sizeof(cd2), hipMemcpyDeviceToHost));
            if (n_cut_ones >= 0) {
                std::memcpy(dm, blk.data(), m * m * sizeof(cd2));
            } else {
                const uint64_t M2 = uint64_t(1) << ncut;
                for (uint64_t c = 0; c < m; c++)
                    for (uint64_t r = 0; r < m; r++) {
                        const uint64_t at = (uint64_t)cfg[a][r] + (uint64_t)cfg[a][c] * M2;
                        dm[2 * at] = blk[r + c * m].re;
                        dm[2 * at + 1] = blk[r + c * m].im;
                    }
            }
        }
        return ED_OK;
    };
    const int rc = body();
    if (h) rocblas_destroy_handle(h);
    for (void* p : {(void*)d_first, (void*)d_keep, (void*)d_coef, (void*)d_psi1, (void*)d_psi2, (void*)d_rho,
                    (void*)d_off, (void*)d_nrow})
        dfree(p);
    return rc;
}

extern "C" int ed_entanglement_entropies(const double* dm, int64_t n, double* renyi2, double* von_neumann) {
    if (!dm || n < 1 || !renyi2 || !von_neumann) return ed_fail(ED_EINVAL, "null argument or empty matrix");
    if (n > 32768) return ed_fail(ED_EINVAL, "matrix of %lld rows exceeds 32768", (long long)n);
    // Renyi_Entropy (Entropy.cpp:334-355): -log2 sum_ij |dm_ij|^2, in the
    // reference's summation order
    double sr = 0.0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) {
            const double re = dm[2 * (i + j * n)], im = dm[2 * (i + j * n) + 1];
            sr += re * re + im * im;
        }
    *renyi2 = -std::log2(sr);
    // Von_Neumann_Entropy (Entropy.cpp:307-331): -sum lambda log2 lambda over
    // the eigenvalues above 1e-8, here from rocSOLVER's Hermitian solver
    cd2* dA = nullptr;
    double *dW = nullptr, *dE = nullptr;
    int* dinfo = nullptr;
    rocblas_handle h = nullptr;
    std::vector<double> w(n);
    auto body = [&]() -> int {
        HIPCHK(hipMalloc(&dA, n * n * sizeof(cd2)));
        HIPCHK(hipMalloc(&dW, n * sizeof(double)));
        HIPCHK(hipMalloc(&dE, n * sizeof(double)));
        HIPCHK(hipMalloc(&dinfo, sizeof(int)));
        HIPCHK(hipMemcpy(dA, dm, n * n * sizeof(cd2), hipMemcpyHostToDevice));
        if (rocblas_create_handle(&h) != rocblas_status_success) return ed_fail(ED_EDEVICE, "rocblas_create_handle failed");
        rocblas_status rs = rocsolver_zheevd(h, rocblas_evect_none, rocblas_fill_upper, (rocblas_int)n,
                                             (rocblas_double_complex*)dA, (rocblas_int)n, dW, dE, dinfo);
        if (rs != rocblas_status_success) return ed_fail(ED_EDEVICE, "rocsolver_zheevd failed (%d)", (int)rs);
        int info = 0;
        HIPCHK(hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost));
        if (info != 0) return ed_fail(ED_ENOTCONV, "zheevd info=%d", info);
        HIPCHK(hipMemcpy(w.data(), dW, n * sizeof(double), hipMemcpyDeviceToHost));
        return ED_OK;
    };
    const int rc = body();
    if (h) rocblas_destroy_handle(h);
    for (void* p : {(void*)dA, (void*)dW, (void*)dE, (void*)dinfo}) dfree(p);
    if (rc != ED_OK) return rc;
    double s = 0.0;
    for (double l : w)
        if (std::fabs(l) > 1e-8) s += l * std::log2(l);
    *von_neumann = -s;
    return ED_OK;
}
