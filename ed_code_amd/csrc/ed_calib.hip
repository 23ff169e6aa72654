// ed_calib.hip -- random-gather calibration: the rate at which the chip
// serves independent random loads of one complex element per lane, the
// access pattern of the H*v gathers, as a measured ceiling next to the HBM
// peak.  No reference counterpart (a measurement helper, like ed_hv_profile).
//
// Each lane gathers per_lane elements from random 128-byte lines of a buffer
// of n_elem elements: lanes in groups of 2^gshift share each line (1 = every
// lane its own line), K loads in flight per lane, addresses from a
// counter-based hash (a few VALU instructions per load), so that the kernel
// measures the memory pipeline and not the address arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_internal.h"

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ void acc_add(edc::cf2& a, const edc::cf2& x) { a.re += x.re; a.im += x.im; }
__device__ __forceinline__ void acc_add(edc::cd2& a, const edc::cd2& x) { a.re += x.re; a.im += x.im; }

// nlines < 2^32 lines of EPL elements; line of load (group g, step b) =
// mulhi(hash(g, b), nlines).
template <typename T, int K>
__global__ void __launch_bounds__(512) k_gather_calib(const T* __restrict__ u, uint32_t nlines, uint32_t per_lane,
                                                      uint32_t gshift, uint32_t seed, T* __restrict__ out) {
    constexpr uint32_t EPL = 128 / sizeof(T);
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t grp = (uint32_t)(tid >> gshift);
    const uint32_t within = (uint32_t)tid & ((1u << gshift) - 1u) & (EPL - 1u);
    T acc{};
    for (uint32_t b = 0; b < per_lane; b += K) {
        T x[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const uint32_t h = mix32(grp * 0x9e3779b9u + (b + q) * 0x85ebca6bu + seed);
            const uint32_t line = (uint32_t)(((uint64_t)h * nlines) >> 32);
            x[q] = u[(uint64_t)line * EPL + within];
        }
#pragma unroll
        for (int q = 0; q < K; q++) acc_add(acc, x[q]);
    }
    out[tid] = acc;
}

template <typename T>
int run_calib(int64_t n_elem, int64_t n_gather, int group, int threads, int inflight, float* ms) {
    constexpr int64_t EPL = 128 / sizeof(T);
    const int64_t nlines = n_elem / EPL;
    if (nlines < 1 || nlines >= (int64_t(1) << 32)) return ed_fail(ED_EINVAL, "buffer of %lld elements", (long long)n_elem);
    if (group < 1 || group > 64 || (group & (group - 1))) return ed_fail(ED_EINVAL, "group must be a power of 2 <= 64");
    if (threads < 64 || threads > 512 || threads % 64) return ed_fail(ED_EINVAL, "threads must be 64..512, x64");
    if (inflight != 4 && inflight != 8) return ed_fail(ED_EINVAL, "inflight must be 4 or 8");
    const int64_t per_lane = 64;
    const int64_t nthr = (n_gather / per_lane + threads - 1) / threads * threads;
    if (nthr < threads || nthr >= (int64_t(1) << 31)) return ed_fail(ED_EINVAL, "gather count out of range");
    const uint32_t gshift = (uint32_t)__builtin_ctz((unsigned)group);
    T *u = nullptr, *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto body = [&]() -> int {
        HIPCHK(hipMalloc(&u, nlines * EPL * sizeof(T)));
        HIPCHK(hipMalloc(&out, nthr * sizeof(T)));
        HIPCHK(hipMemset(u, 0, nlines * EPL * sizeof(T)));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        float best = 1e30f;
        for (int rep = 0; rep < 4; rep++) {
            HIPCHK(hipEventRecord(e0, 0));
            if (inflight == 4)
                hipLaunchKernelGGL((k_gather_calib<T, 4>), dim3((unsigned)(nthr / threads)), dim3(threads), 0, 0, u,
                                   (uint32_t)nlines, (uint32_t)per_lane, gshift, 17u + rep, out);
            else
                hipLaunchKernelGGL((k_gather_calib<T, 8>), dim3((unsigned)(nthr / threads)), dim3(threads), 0, 0, u,
                                   (uint32_t)nlines, (uint32_t)per_lane, gshift, 17u + rep, out);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(e1, 0));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            if (rep > 0) best = std::min(best, t);
        }
        *ms = best;
        return ED_OK;
    };
    const int rc = body();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(u);
    dfree(out);
    return rc;
}

}  // namespace

extern "C" int ed_gather_calibrate(int64_t n_elem, int32_t elem_bytes, int64_t n_gather, int32_t group,
                                   int32_t threads, int32_t inflight, float* ms) {
    if (!ms) return ed_fail(ED_EINVAL, "null argument");
    if (elem_bytes == 8) return run_calib<edc::cf2>(n_elem, n_gather, group, threads, inflight, ms);
    if (elem_bytes == 16) return run_calib<edc::cd2>(n_elem, n_gather, group, threads, inflight, ms);
    return ed_fail(ED_EINVAL, "elem_bytes must be 8 or 16");
}
