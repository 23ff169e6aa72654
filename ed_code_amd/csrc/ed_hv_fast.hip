// ed_hv_fast.hip -- rotation fast path of H*v (ed_torus.h) for N <= 32
// lattices whose translations are bit rotations.
//
// One lane per row.  Per lane the antiparallel bonds of each family are the
// set bits of one XOR; each set bit is one gathered element: flip, orbit
// minimum by rotations (alignbit + min3 tree), byte-sliced rank from LDS,
// 12-byte directory entry, then the gather of u = v/norm.  The rank tables,
// family entries and phase table are staged in LDS once per block.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_internal.h"
#include "ed_torus.h"
#include "ed_torus_host.h"

using namespace edc;

template <bool TRIV, typename VT>
__global__ void __launch_bounds__(512) k_hv_torus(int64_t D, const uint64_t* __restrict__ rows, Torus T, int nfam,
                                                  const TFam* __restrict__ gfam, const TEntry* __restrict__ gent,
                                                  double ezz0, RankInfo R, const uint32_t* __restrict__ grtab,
                                                  int rtab_n, const DirEntry* __restrict__ dir,
                                                  const cd2* __restrict__ gphase, int nphase,
                                                  const double* __restrict__ ntab, const VT* __restrict__ u,
                                                  const VT* __restrict__ v, VT* __restrict__ w) {
    extern __shared__ __align__(16) unsigned char smem[];
    TEntry* ent = (TEntry*)smem;                              // kMaxFam*32
    TFam* fam = (TFam*)(ent + kMaxFam * 32);                  // kMaxFam
    cd2* phase = (cd2*)(fam + kMaxFam);                       // nphase
    uint32_t* rtab = (uint32_t*)(phase + nphase);             // rtab_n
    for (int q = threadIdx.x; q < nfam * 32; q += blockDim.x) ent[q] = gent[q];
    for (int q = threadIdx.x; q < nfam; q += blockDim.x) fam[q] = gfam[q];
    for (int q = threadIdx.x; q < nphase; q += blockDim.x) phase[q] = gphase[q];
    for (int q = threadIdx.x; q < rtab_n; q += blockDim.x) rtab[q] = grtab[q];
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    uint64_t rw = rows[i];
    uint32_t s = (uint32_t)(rw & kStateMask);
    double nrm = ntab[(int)(rw >> kLidShift)];
    double ar = 0.0, ai = 0.0;
    double ez = torus_row<TRIV>(s, T, nfam, fam, ent, ezz0, R, rtab, dir, phase, [&](int64_t j, double cr, double ci) {
        cd2 x = load_c(u, j);
        ar += cr * x.re - ci * x.im;
        ai += cr * x.im + ci * x.re;
    });
    cd2 vi = load_c(v, i);
    VT o;
    o.re = nrm * ar + ez * vi.re;
    o.im = nrm * ai + ez * vi.im;
    w[i] = o;
}

// LDS of the compiled-shape kernels, in this order: [cd2 x nfam*32*4
// (kModeMerged: coefficients)] [TEntry x nfam*32 (not kModeUniform)] [TFam x
// nfam, padded to even] [cd2 x nfam (kModeUniform: J/2, Delta J/2)] [cd2 phase
// x nphase] [u32 flip mask x nfam*32] [u32 z-sites x nfam*32 (kModeMerged)]
// [u32 rank tables x rtab_n] [u16 chunk map x nmap].  Stages the tables with
// the whole block and returns them as RowTabs.
template <int MODE>
__device__ __forceinline__ RowTabs stage_row_tabs(int nfam, const TFam* __restrict__ gfam,
                                                  const TEntry* __restrict__ gent, const uint32_t* __restrict__ gmki,
                                                  const cd2* __restrict__ gmco, const uint32_t* __restrict__ gfmk,
                                                  const cd2* __restrict__ gfco, const uint32_t* __restrict__ grtab,
                                                  int rtab_n, const uint16_t* __restrict__ gmap, int nmap,
                                                  const cd2* __restrict__ gphase, int nphase, uint16_t*& map) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool MG = MODE == kModeMerged;
    cd2* mco = (cd2*)smem;
    TEntry* ent = (TEntry*)(mco + (MG ? nfam * 32 * 4 : 0));
    TFam* fam = (TFam*)(ent + (MODE == kModeUniform ? 0 : nfam * 32));
    cd2* fco = (cd2*)(fam + nfam + (nfam & 1));
    cd2* phase = fco + (MODE == kModeUniform ? nfam : 0);
    uint32_t* fmk = (uint32_t*)(phase + nphase);
    uint32_t* mki = fmk + nfam * 32;
    uint32_t* rtab = mki + (MG ? nfam * 32 : 0);
    map = (uint16_t*)(rtab + rtab_n);
    if (MG) {
        for (int q = threadIdx.x; q < nfam * 32 * 4; q += blockDim.x) mco[q] = gmco[q];
        for (int q = threadIdx.x; q < nfam * 32; q += blockDim.x) mki[q] = gmki[q];
    }
    if (MODE != kModeUniform)
        for (int q = threadIdx.x; q < nfam * 32; q += blockDim.x) ent[q] = gent[q];
    for (int q = threadIdx.x; q < nfam; q += blockDim.x) fam[q] = gfam[q];
    if (MODE == kModeUniform)
        for (int q = threadIdx.x; q < nfam; q += blockDim.x) fco[q] = gfco[q];
    for (int q = threadIdx.x; q < nphase; q += blockDim.x) phase[q] = gphase[q];
    for (int q = threadIdx.x; q < nfam * 32; q += blockDim.x) fmk[q] = gfmk[q];
    for (int q = threadIdx.x; q < rtab_n; q += blockDim.x) rtab[q] = grtab[q];
    for (int q = threadIdx.x; q < nmap; q += blockDim.x) map[q] = gmap[q];
    __syncthreads();
    return RowTabs{nfam, fam, ent, mki, mco, fmk, fco, rtab, phase};
}

// Row body of the compiled-shape kernels: the lattice shape as compile-time
// constants (fully unrolled rotation search, masks as immediates), K gathered
// elements in flight per lane.  The order in which elements are accumulated
// does not depend on K, so every variant returns bit-identical results.
template <class TC, bool TRIV, int MODE, int K, bool PIPE, typename VT>
__device__ __forceinline__ void hv_torus_c_body(int64_t D, const uint64_t* __restrict__ rows, int nfam,
                                                const TFam* __restrict__ gfam, const TEntry* __restrict__ gent,
                                                const uint32_t* __restrict__ gmki, const cd2* __restrict__ gmco,
                                                const uint32_t* __restrict__ gfmk, const cd2* __restrict__ gfco,
                                                double ezz0, const RankInfo& R, const uint32_t* __restrict__ grtab,
                                                int rtab_n, const uint16_t* __restrict__ gmap, int nmap,
                                                int cb, uint32_t zidx, const cd2* __restrict__ gphase, int nphase,
                                                const double* __restrict__ ntab, const VT* __restrict__ u,
                                                const VT* __restrict__ v, VT* __restrict__ w) {
    uint16_t* map;
    const RowTabs T = stage_row_tabs<MODE>(nfam, gfam, gent, gmki, gmco, gfmk, gfco, grtab, rtab_n, gmap, nmap, gphase,
                                           nphase, map);
    // XCD tiles of consecutive rows, the 8 XCDs on adjacent tiles
    const unsigned lb = xcd_tile_block();
    int64_t i = (int64_t)lb * blockDim.x + threadIdx.x;
    if (i >= D) return;
    uint64_t rw = rows[i];
    uint32_t s = (uint32_t)(rw & kStateMask);
    double nrm = ntab[(int)(rw >> kLidShift)];
    double ar = 0.0, ai = 0.0;
    // u is the gather image: the rank of a target and the LDS chunk map give
    // its element, with no directory load
    auto jof = [&](uint32_t r, bool ok) { return image_index(map, cb, zidx, r, ok); };
    auto loadu = [&](uint32_t j) { return load_c32(u, j); };
    // explicit FMAs, so that the accumulation rounds identically in every
    // variant whatever the compiler contracts elsewhere
    auto acc = [&](const cd2& x, double cr, double ci) {
        ar = fma(cr, x.re, ar);
        ai = fma(cr, x.im, ai);
        if (MODE != kModeReal && MODE != kModeUniform) {
            ar = fma(-ci, x.im, ar);
            ai = fma(ci, x.re, ai);
        }
    };
    double ez = PIPE ? torus_row_pipelined<TC, TRIV, MODE, K>(s, T, ezz0, R, jof, loadu, acc)
                     : torus_row_batched<TC, TRIV, MODE, K>(s, T, ezz0, R, jof, loadu, acc);
    cd2 vi = load_c(v, i);
    VT o;
    o.re = fma(nrm, ar, ez * vi.re);
    o.im = fma(nrm, ai, ez * vi.im);
    w[i] = o;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int o) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, o), hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), o);
    return ((uint64_t)hi << 32) | lo;
}

// Row body of the split kernels: 2^LPRL lanes share a row (lanes 4 apart, so
// that each quad of lanes holds four rows), each lane q taking one
// consecutive share of its elements (torus_row_share), K in
// flight; the lanes' partial sums meet in a fixed tree (lane q with q ^ 1,
// then q ^ 2, ...).  A workgroup then covers blockDim / 2^LPRL rows, so the
// rows in flight on an XCD -- whose targets compete for its L2 -- are fewer
// for the same number of waves.  Rounds differently from the one-lane
// variants (the shares are summed separately); all split variants with the
// same 2^LPRL agree bit for bit.
template <class TC, bool TRIV, int MODE, int K, int LPRL, typename VT>
__device__ __forceinline__ void hv_torus_split_body(int64_t D, const uint64_t* __restrict__ rows, int nfam,
                                                    const TFam* __restrict__ gfam, const TEntry* __restrict__ gent,
                                                    const uint32_t* __restrict__ gmki, const cd2* __restrict__ gmco,
                                                    const uint32_t* __restrict__ gfmk, const cd2* __restrict__ gfco,
                                                    double ezz0, const RankInfo& R,
                                                    const uint32_t* __restrict__ grtab, int rtab_n,
                                                    const uint16_t* __restrict__ gmap, int nmap, int cb,
                                                    uint32_t zidx, const cd2* __restrict__ gphase, int nphase,
                                                    const double* __restrict__ ntab, const VT* __restrict__ u,
                                                    const VT* __restrict__ v, VT* __restrict__ w) {
    constexpr int LPR = 1 << LPRL;
    uint16_t* map;
    const RowTabs T = stage_row_tabs<MODE>(nfam, gfam, gent, gmki, gmco, gfmk, gfco, grtab, rtab_n, gmap, nmap, gphase,
                                           nphase, map);
    const unsigned rpb = blockDim.x >> LPRL;
    const unsigned lb = xcd_tile_block_rows(rpb);
    // lanes 4 apart share a row (lane q of the row in bits 2.., the row in
    // the others): a quad of lanes, whose gathers the vector L1 coalesces,
    // holds four consecutive rows at the same share
    const unsigned t = threadIdx.x;
    const int q = (int)((t >> 2) & (LPR - 1));
    const int64_t i = (int64_t)lb * rpb + (((t >> (2 + LPRL)) << 2) | (t & 3u));
    const bool live = i < D;
    uint32_t s = 0;
    double nrm = 0.0;
    if (live) {
        const uint64_t rw = rows[i];
        s = (uint32_t)(rw & kStateMask);
        nrm = ntab[(int)(rw >> kLidShift)];
    }
    // antiparallel count of every family: lane q counts families q, q + LPR, ...
    uint64_t counts = live ? split_counts<TC>(s, T, q, LPR) : 0u;
#pragma unroll
    for (int o = 1; o < LPR; o <<= 1) counts |= shfl_xor_u64(counts, 4 * o);
    double ar = 0.0, ai = 0.0, ez = 0.0;
    auto jof = [&](uint32_t r, bool ok) { return image_index(map, cb, zidx, r, ok); };
    auto loadu = [&](uint32_t j) { return load_c32(u, j); };
    auto acc = [&](const cd2& x, double cr, double ci) {
        ar = fma(cr, x.re, ar);
        ai = fma(cr, x.im, ai);
        if (MODE != kModeReal && MODE != kModeUniform) {
            ar = fma(-ci, x.im, ar);
            ai = fma(ci, x.re, ai);
        }
    };
    if (live) ez = torus_row_share<TC, TRIV, MODE, K>(s, q, LPRL, counts, T, ezz0, R, jof, loadu, acc);
    // lane q adds lane q ^ 1, then q ^ 2 (lanes 4 and 8 apart)
#pragma unroll
    for (int o = 1; o < LPR; o <<= 1) {
        ar += __shfl_xor(ar, 4 * o);
        ai += __shfl_xor(ai, 4 * o);
        ez += __shfl_xor(ez, 4 * o);
    }
    if (live && q == 0) {
        const cd2 vi = load_c(v, i);
        VT o;
        o.re = fma(nrm, ar, ez * vi.re);
        o.im = fma(nrm, ai, ez * vi.im);
        w[i] = o;
    }
}

#define ED_TORUS_C_ARGS                                                                                   \
    int64_t D, const uint64_t *__restrict__ rows, int nfam, const TFam *__restrict__ gfam,               \
        const TEntry *__restrict__ gent, const uint32_t *__restrict__ gmki, const cd2 *__restrict__ gmco,  \
        const uint32_t *__restrict__ gfmk,                                                                 \
        const cd2 *__restrict__ gfco, double ezz0, RankInfo R,                                             \
        const uint32_t *__restrict__ grtab,                                                               \
        int rtab_n, const uint16_t *__restrict__ gmap, int nmap, int cb, uint32_t zidx,                   \
        const cd2 *__restrict__ gphase, int nphase, const double *__restrict__ ntab, const VT *__restrict__ u, \
        const VT *__restrict__ v, VT *__restrict__ w
#define ED_TORUS_C_PASS \
    D, rows, nfam, gfam, gent, gmki, gmco, gfmk, gfco, ezz0, R, grtab, rtab_n, gmap, nmap, cb, zidx, gphase, nphase, ntab, u, v, w

// Variant A: four elements in flight, occupancy left to the register count.
template <class TC, bool TRIV, int MODE, typename VT>
__global__ void __launch_bounds__(512) k_hv_torus_c4(ED_TORUS_C_ARGS) {
    hv_torus_c_body<TC, TRIV, MODE, 4, false, VT>(ED_TORUS_C_PASS);
}

// Variant D: four elements in flight, batches software-pipelined.
template <class TC, bool TRIV, int MODE, typename VT>
__global__ void __launch_bounds__(512) k_hv_torus_p4(ED_TORUS_C_ARGS) {
    hv_torus_c_body<TC, TRIV, MODE, 4, true, VT>(ED_TORUS_C_PASS);
}

// Variant C: eight elements in flight.
template <class TC, bool TRIV, int MODE, typename VT>
__global__ void __launch_bounds__(512) k_hv_torus_c8(ED_TORUS_C_ARGS) {
    hv_torus_c_body<TC, TRIV, MODE, 8, false, VT>(ED_TORUS_C_PASS);
}

// Variant B: two elements in flight, registers capped for 8 waves per SIMD.
template <class TC, bool TRIV, int MODE, typename VT>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) k_hv_torus_c2w8(ED_TORUS_C_ARGS) {
    hv_torus_c_body<TC, TRIV, MODE, 2, false, VT>(ED_TORUS_C_PASS);
}

// Variant E: rows split over two lanes, eight elements in flight per lane.
template <class TC, bool TRIV, int MODE, typename VT>
__global__ void __launch_bounds__(512) k_hv_torus_s2(ED_TORUS_C_ARGS) {
    hv_torus_split_body<TC, TRIV, MODE, 8, 1, VT>(ED_TORUS_C_PASS);
}

// Variant F: rows split over four lanes, four elements in flight per lane.
template <class TC, bool TRIV, int MODE, typename VT>
__global__ void __launch_bounds__(512) k_hv_torus_s4(ED_TORUS_C_ARGS) {
    hv_torus_split_body<TC, TRIV, MODE, 4, 2, VT>(ED_TORUS_C_PASS);
}

// Variant G: rows split over two lanes, four elements in flight per lane
// (a lane's share of ~33 elements on the 32-site torus wastes fewer batch
// slots than with eight).
template <class TC, bool TRIV, int MODE, typename VT>
__global__ void __launch_bounds__(512) k_hv_torus_s2k4(ED_TORUS_C_ARGS) {
    hv_torus_split_body<TC, TRIV, MODE, 4, 1, VT>(ED_TORUS_C_PASS);
}

// Lanes per row of a variant (1 for the one-lane variants 0-3).
static int variant_lanes(int var) { return (var == 4 || var == 6) ? 2 : var == 5 ? 4 : 1; }

template <class TC, typename VT>
static void launch_c(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st, size_t sm, int var, int bs) {
    ed_basis_s* B = op->b;
    const int64_t D = B->D;
    const int64_t rpb = bs / variant_lanes(var);   // rows per workgroup
    dim3 grid((unsigned)((D + rpb - 1) / rpb)), blk(bs);
    int nph = B->L1 * B->L2;
    const uint32_t zidx = B->img_nslot << B->img_cb;   // first element of the zero slot
    // the merged families (bond pairs carrying their chiral entries) when built
    const bool mg = op->tmode == kModeMerged;
#define ED_GO(KER)                                                                                            \
    hipLaunchKernelGGL(KER, grid, blk, sm, st, D, B->d_rows, mg ? op->nmfam : op->nfam, mg ? op->d_mfam : op->d_tfam, \
                       mg ? op->d_ment : op->d_tent, op->d_mki, op->d_mco, mg ? op->d_mfmk : op->d_tfmk, op->d_tfco,  \
                       op->ezz0,                                                                                  \
                       B->rinfo, B->d_rtab32, B->rtab32_n, B->d_img_map, (int)B->img_nmap, B->img_cb, zidx,       \
                       B->d_phase, nph, B->d_ntab, u, v, w)
    // term class: k = 0 takes the general, real-bond or uniform-family kernel,
    // k != 0 the general or the bond-only one
#define ED_GO_MODES(KERNEL)                                                         \
    if (B->trivial_k) {                                                             \
        if (op->tmode == kModeUniform) ED_GO((KERNEL<TC, true, kModeUniform, VT>)); \
        else if (op->tmode == kModeReal) ED_GO((KERNEL<TC, true, kModeReal, VT>));  \
        else if (mg) ED_GO((KERNEL<TC, true, kModeMerged, VT>));                    \
        else ED_GO((KERNEL<TC, true, kModeGeneral, VT>));                           \
    } else {                                                                        \
        if (op->tmode == kModeBonds) ED_GO((KERNEL<TC, false, kModeBonds, VT>));    \
        else if (mg) ED_GO((KERNEL<TC, false, kModeMerged, VT>));                   \
        else ED_GO((KERNEL<TC, false, kModeGeneral, VT>));                          \
    }
    if (var == 0) {
        ED_GO_MODES(k_hv_torus_c4)
    } else if (var == 1) {
        ED_GO_MODES(k_hv_torus_c2w8)
    } else if (var == 2) {
        ED_GO_MODES(k_hv_torus_c8)
    } else if (var == 3) {
        ED_GO_MODES(k_hv_torus_p4)
    } else if (var == 4) {
        ED_GO_MODES(k_hv_torus_s2)
    } else if (var == 6) {
        ED_GO_MODES(k_hv_torus_s2k4)
    } else {
        ED_GO_MODES(k_hv_torus_s4)
    }
#undef ED_GO_MODES
#undef ED_GO
}

bool torus_shape_compiled(const Torus& T) {
    int L2 = T.kind == 2 ? 1 : T.L2;
#define ED_MATCH(n, wd, l2) if (T.N == n && T.w == wd && L2 == l2) return true;
    ED_TORUS_SHAPES(ED_MATCH)
#undef ED_MATCH
    return false;
}

static size_t torus_c_smem_bytes(const ed_op_s* op);

// The compiled-shape kernels gather from the sector's image (built with the
// basis when it fits the LDS map and the 32-bit byte offsets of the loads)
// and keep their tables in at most 64 KiB of LDS; other operators run the
// run-time-shape kernel.
#ifndef ED_TORUS_LDS_MAX
#define ED_TORUS_LDS_MAX (64 * 1024)
#endif
static bool compiled_fits(const ed_op_s* op) {
    return torus_shape_compiled(op->torus) && op->b->img_cb >= 0 && torus_c_smem_bytes(op) <= ED_TORUS_LDS_MAX;
}

// Whether the next H*v of this operator runs a compiled-shape kernel, whose
// gathered vector is the image rather than u by row.
bool torus_uses_image(const ed_op_s* op) {
    return op->use_torus && op->tune_path != 1 && op->tune_path != 3 && op->tune_path != 6 && op->tune_path != 8 &&
           compiled_fits(op);
}

size_t torus_smem_bytes(const ed_op_s* op) {
    const ed_basis_s* B = op->b;
    return kMaxFam * 32 * sizeof(TEntry) + kMaxFam * sizeof(TFam) + (size_t)B->L1 * B->L2 * sizeof(cd2) +
           (size_t)B->rtab32_n * sizeof(uint32_t);
}

// LDS of the compiled-shape kernel: entries and families sized by the
// operator's family count.
static size_t torus_c_smem_bytes(const ed_op_s* op) {
    const ed_basis_s* B = op->b;
    const bool mg = op->tmode == kModeMerged, un = op->tmode == kModeUniform;
    const int nf = mg ? op->nmfam : op->nfam;
    return (mg ? (size_t)nf * 32 * 4 * sizeof(cd2) : 0) + (un ? 0 : (size_t)nf * 32 * sizeof(TEntry)) +
           (size_t)(nf + (nf & 1)) * sizeof(TFam) + (un ? (size_t)nf * sizeof(cd2) : 0) +
           (size_t)B->L1 * B->L2 * sizeof(cd2) + (size_t)nf * 32 * sizeof(uint32_t) * (mg ? 2 : 1) +
           (size_t)B->rtab32_n * sizeof(uint32_t) + (size_t)B->img_nmap * sizeof(uint16_t);
}

// Launches the compiled-shape kernel for the operator's lattice; false when
// the shape is not one of the built-in ones.
template <typename VT>
static bool launch_compiled(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st, int var, int bs) {
    const Torus& T = op->torus;
    if (!compiled_fits(op)) return false;
    int L2 = T.kind == 2 ? 1 : T.L2;
#define ED_LAUNCH(n, wd, l2)                                                            \
    if (T.N == n && T.w == wd && L2 == l2) {                                            \
        launch_c<TorusC<n, wd, l2>, VT>(op, u, v, w, st, torus_c_smem_bytes(op), var, bs); \
        return true;                                                                    \
    }
    ED_TORUS_SHAPES(ED_LAUNCH)
#undef ED_LAUNCH
    return false;
}

template <typename VT>
static constexpr int dtype_slot() {
    return std::is_same<VT, cd2>::value ? 0 : 1;
}

// Candidate (variant, block) pairs of the compiled-shape kernel.  Which one is
// fastest depends on the register pressure of the shape, the vector precision
// and the momentum (measured on the 32-site torus: K=4 at 256 threads for
// complex64 at k=0, K=8 at 256 for complex128, K=2 at 8 waves for k != 0).
// The split variants (4, 5, 6) serve families of at most kSplitMaxFam.
static const int kTuneCfg[14][2] = {{0, 512}, {0, 256}, {1, 512}, {1, 256}, {2, 512}, {2, 256}, {3, 512},
                                    {3, 256}, {4, 512}, {4, 256}, {5, 512}, {5, 256}, {6, 512}, {6, 256}};

static bool variant_fits(const ed_op_s* op, int var) {
    const int nf = op->tmode == kModeMerged ? op->nmfam : op->nfam;
    return variant_lanes(var) == 1 || nf <= kSplitMaxFam;
}
static const int64_t kTuneMinRows = int64_t(1) << 20;

// Picks kept for the process: operators of the same compiled shape, term
// class, momentum class, family count, precision and size class (the bit
// length of D) take the pick the first of them measured, without timing the
// candidates again (a server builds an operator per request).
using TuneKey = std::tuple<int, int, int, int, int, int, int, int, int>;
static std::mutex g_tune_mu;
static std::map<TuneKey, std::pair<int, int>> g_tune_picks;

extern "C" int ed_tune_cache_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_picks.clear();
    return ED_OK;
}

static TuneKey tune_key(const ed_op_s* op, int dslot) {
    const ed_basis_s* B = op->b;
    const int nf = op->tmode == kModeMerged ? op->nmfam : op->nfam;
    int dlog = 0;
    for (int64_t d = B->D; d > 1; d >>= 1) dlog++;
    return TuneKey(B->torus.kind, B->torus.N, B->torus.w, B->torus.L2, op->tmode, B->trivial_k ? 1 : 0, nf, dslot, dlog);
}

// Chooses the configuration of the compiled-shape kernel for this operator
// and precision on the first call: below kTuneMinRows rows, or while the
// stream is being captured, the default (K=4, 512 threads, or the block fixed
// by ed_op_tune) is taken; above, the candidates run on the caller's vectors
// (one untimed launch, each candidate timed once, the three fastest twice
// more) and the fastest is kept.  Every candidate writes the same w.
template <typename VT>
int torus_prepare(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st) {
    if (!op->use_torus || op->tune_path != 0 || !compiled_fits(op)) return ED_OK;
    const int k = dtype_slot<VT>();
    if (op->sel_var[k] >= 0) return ED_OK;
    const int64_t D = op->b->D;
    int best_var = 0, best_bs = op->block_fixed ? op->tune_block : 512;
    const TuneKey key = tune_key(op, k);
    if (D >= kTuneMinRows && !op->block_fixed) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_picks.find(key);
        if (it != g_tune_picks.end()) {
            op->sel_var[k] = it->second.first;
            op->sel_block[k] = it->second.second;
            return ED_OK;
        }
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(st, &cs));
    bool timed = false;
    if (D >= kTuneMinRows && cs == hipStreamCaptureStatusNone) {
        timed = true;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        // one untimed launch, every candidate timed once, then the three
        // fastest twice more: the best of their three timings wins
        auto timed_launch = [&](int var, int bs, float& ms) -> int {
            HIPCHK(hipEventRecord(e0, st));
            launch_compiled<VT>(op, u, v, w, st, var, bs);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            return ED_OK;
        };
        std::vector<std::pair<float, int>> first;
        float ms = 0.f;
        bool warm = false;
        for (int ci = 0; ci < (int)(sizeof(kTuneCfg) / sizeof(kTuneCfg[0])); ci++) {
            const int* c = kTuneCfg[ci];
            if (op->block_fixed && c[1] != op->tune_block) continue;
            if (!variant_fits(op, c[0])) continue;
            int rc = ED_OK;
            if (!warm) rc = timed_launch(c[0], c[1], ms);
            warm = true;
            if (rc == ED_OK) rc = timed_launch(c[0], c[1], ms);
            if (rc != ED_OK) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                return rc;
            }
            first.emplace_back(ms, ci);
        }
        std::sort(first.begin(), first.end());
        float best = 1e30f;
        for (size_t r = 0; r < first.size() && r < 3; r++) {
            const int* c = kTuneCfg[first[r].second];
            float t = first[r].first;
            for (int rep = 0; rep < 2; rep++) {
                int rc = timed_launch(c[0], c[1], ms);
                if (rc != ED_OK) {
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    return rc;
                }
                t = std::min(t, ms);
            }
            if (t < best) {
                best = t;
                best_var = c[0];
                best_bs = c[1];
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    op->sel_var[k] = best_var;
    op->sel_block[k] = best_bs;
    if (timed && !op->block_fixed) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_picks.emplace(key, std::make_pair(best_var, best_bs));
    }
    return ED_OK;
}

// ---- stored element list (path 12) ----
// The off-diagonal elements of a sector with uniform real bond families at
// k = 0 (kModeUniform), generated once by the compiled-shape row machinery and
// kept in HBM: per element the gather-image element of its target and its
// family, packed in 32 bits.  The H*v kernel then streams these words and
// gathers, with no orbit search or rank: the instruction-free form of the
// same gathers, against which the matrix-free kernels are measured.  Rows
// are stored in groups of 64 (one wave): element q of row 64 g + l at
// ebase[g] + 64 q + l, so that a wave's loads of the list are coalesced;
// each group takes as many slots per row as its longest row.

constexpr int kEGroup = 64;

// Pass 1 (el null): the element count of each row into rl.  Pass 2: the
// packed elements at their slots and the row's diagonal, in the order of
// torus_row_batched (families in order, members by increasing site).
template <class TC>
__global__ void __launch_bounds__(256) k_elist(int64_t D, const uint64_t* __restrict__ rows, int nfam,
                                               const TFam* __restrict__ gfam, const uint32_t* __restrict__ gfmk,
                                               const cd2* __restrict__ gfco, double ezz0, RankInfo R,
                                               const uint32_t* __restrict__ grtab, int rtab_n,
                                               const uint16_t* __restrict__ gmap, int nmap, int cb, uint32_t zidx,
                                               const cd2* __restrict__ gphase, int nphase, int jbits,
                                               uint16_t* __restrict__ rl, const int64_t* __restrict__ ebase,
                                               uint32_t* __restrict__ el, double* __restrict__ diag) {
    uint16_t* map;
    const RowTabs T = stage_row_tabs<kModeUniform>(nfam, gfam, nullptr, nullptr, nullptr, gfmk, gfco, grtab, rtab_n,
                                                   gmap, nmap, gphase, nphase, map);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const uint32_t s = (uint32_t)(rows[i] & kStateMask);
    uint32_t* out = el ? el + ebase[i / kEGroup] + (i % kEGroup) : nullptr;
    double ezz = ezz0;
    int n = 0;
    for (int f = 0; f < T.nfam; f++) {
        uint32_t cur = fam_mask_c<TC>(s, T.fam[f]);
        ezz = fma(-(double)popc32(cur), T.fco[f].im, ezz);
        while (cur != 0u) {
            const int x = ctz32(cur);
            cur &= cur - 1u;
            if (out) {
                int code;
                const uint32_t rep = torus_rep_c<TC, kRepOnly>(s ^ T.fmk[f * 32 + x], code);
                const uint32_t r = rank_tab<uint32_t, uint32_t>(rep, R, T.rtab, (TC::N + 7) / 8);
                out[(int64_t)n * kEGroup] = image_index(map, cb, zidx, r, true) | ((uint32_t)f << jbits);
            }
            n++;
        }
    }
    if (out) diag[i] = ezz;
    else rl[i] = (uint16_t)n;
}

// Slots of each group: 64 times its longest row (the last group's missing
// rows count as empty).
__global__ void k_elist_groups(int64_t D, const uint16_t* __restrict__ rl, int64_t ng, int64_t* __restrict__ gs) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int m = 0;
    for (int l = 0; l < kEGroup; l++) {
        const int64_t i = g * kEGroup + l;
        if (i < D) m = max(m, (int)rl[i]);
    }
    gs[g] = (int64_t)m * kEGroup;
}

__global__ void k_elist_ctab(int nfam, const cd2* __restrict__ gfco, double* __restrict__ ct) {
    const int f = threadIdx.x;
    if (f < nfam) ct[f] = gfco[f].re;
}

// H*v from the stored list: K elements in flight per lane; the same FMAs in
// the same order as the compiled-shape kernels, with an empty slot adding
// 0 times the zero slot as theirs do.  The list's index is clamped to the
// row's last element so that every load is issued unconditionally.
template <int K, typename VT>
__global__ void __launch_bounds__(512) k_hv_elist(int64_t D, const uint64_t* __restrict__ rows,
                                                  const double* __restrict__ ntab, const int64_t* __restrict__ ebase,
                                                  const uint16_t* __restrict__ rl, const uint32_t* __restrict__ el,
                                                  const double* __restrict__ diag, const double* __restrict__ gct,
                                                  int nfam, int jbits, uint32_t zidx, const VT* __restrict__ u,
                                                  const VT* __restrict__ v, VT* __restrict__ w) {
    __shared__ double ct[kMaxFam];
    for (int q = threadIdx.x; q < nfam; q += blockDim.x) ct[q] = gct[q];
    __syncthreads();
    const unsigned lb = xcd_tile_block();
    const int64_t i = (int64_t)lb * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const double nrm = ntab[(int)(rows[i] >> kLidShift)];
    const uint32_t* p = el + ebase[i / kEGroup] + (i % kEGroup);
    const int n = rl[i];
    const uint32_t jm = (1u << jbits) - 1u;
    double ar = 0.0, ai = 0.0;
    for (int q = 0; q < n; q += K) {
        uint32_t j[K];
        double c[K];
#pragma unroll
        for (int h = 0; h < K; h++) {
            const uint32_t e = p[(int64_t)min(q + h, n - 1) * kEGroup];
            const bool ok = q + h < n;
            j[h] = ok ? (e & jm) : zidx;
            c[h] = ok ? ct[e >> jbits] : 0.0;
        }
        cd2 x[K];
#pragma unroll
        for (int h = 0; h < K; h++) x[h] = load_c32(u, j[h]);
#pragma unroll
        for (int h = 0; h < K; h++) {
            ar = fma(c[h], x[h].re, ar);
            ai = fma(c[h], x[h].im, ai);
        }
    }
    const cd2 vi = load_c(v, i);
    const double ez = diag[i];
    VT o;
    o.re = fma(nrm, ar, ez * vi.re);
    o.im = fma(nrm, ai, ez * vi.im);
    w[i] = o;
}

static int bit_length(uint64_t x) {
    int b = 0;
    while (x) {
        b++;
        x >>= 1;
    }
    return b;
}

// Whether the operator can take the stored element list.
int torus_elist_check(const ed_op_s* op) {
    const ed_basis_s* B = op->b;
    if (!op->use_torus || !torus_shape_compiled(op->torus) || !compiled_fits(op))
        return ed_fail(ED_EINVAL, "stored element list: rotation fast path with a compiled shape required");
    if (op->tmode != kModeUniform || !B->trivial_k)
        return ed_fail(ED_EINVAL, "stored element list: uniform real bond families at k = 0 only");
    const uint64_t ne = (uint64_t)(B->img_nslot + 1) << B->img_cb;
    if (bit_length(ne - 1) > 31 || bit_length(ne - 1) + bit_length((uint64_t)(op->nfam > 1 ? op->nfam - 1 : 0)) > 32)
        return ed_fail(ED_EINVAL, "stored element list: image element and family do not pack into 32 bits");
    return ED_OK;
}

template <class TC>
static void launch_elist_build(ed_op_s* op, hipStream_t st, size_t sm, bool emit) {
    ed_basis_s* B = op->b;
    const uint32_t zidx = B->img_nslot << B->img_cb;
    hipLaunchKernelGGL((k_elist<TC>), dim3((unsigned)((B->D + 255) / 256)), dim3(256), sm, st, B->D, B->d_rows,
                       op->nfam, op->d_tfam, op->d_tfmk, op->d_tfco, op->ezz0, B->rinfo, B->d_rtab32, B->rtab32_n,
                       B->d_img_map, (int)B->img_nmap, B->img_cb, zidx, B->d_phase, B->L1 * B->L2, op->ejbits,
                       op->d_erl, emit ? op->d_ebase : nullptr, emit ? op->d_el : nullptr, op->d_ediag);
}

static bool elist_build(ed_op_s* op, hipStream_t st, bool emit) {
    const Torus& T = op->torus;
    const int L2 = T.kind == 2 ? 1 : T.L2;
    const size_t sm = torus_c_smem_bytes(op);
#define ED_ELIST(n, wd, l2)                                        \
    if (T.N == n && T.w == wd && L2 == l2) {                       \
        launch_elist_build<TorusC<n, wd, l2>>(op, st, sm, emit);   \
        return true;                                               \
    }
    ED_TORUS_SHAPES(ED_ELIST)
#undef ED_ELIST
    return false;
}

// Builds the operator's stored element list (on the caller's stream, which
// is then synchronised once to size the list).
int torus_elist_prepare(ed_op_s* op, hipStream_t st) {
    int rc = torus_elist_check(op);
    if (rc != ED_OK) return rc;
    ed_basis_s* B = op->b;
    const int64_t D = B->D, ng = (D + kEGroup - 1) / kEGroup;
    const uint64_t ne = (uint64_t)(B->img_nslot + 1) << B->img_cb;
    op->ejbits = bit_length(ne - 1);
    auto fail = [&](int code) {
        for (void* q : {(void*)op->d_ebase, (void*)op->d_erl, (void*)op->d_el, (void*)op->d_ediag, (void*)op->d_ectab})
            dfree(q);
        op->d_ebase = nullptr;
        op->d_erl = nullptr;
        op->d_el = nullptr;
        op->d_ediag = nullptr;
        op->d_ectab = nullptr;
        return code;
    };
#define ED_EL(x)                                                                                            \
    do {                                                                                                    \
        hipError_t e_ = (x);                                                                                \
        if (e_ != hipSuccess)                                                                               \
            return fail(ed_fail(e_ == hipErrorOutOfMemory ? ED_ENOMEM : ED_EDEVICE, "%s: %s", #x,           \
                                hipGetErrorString(e_)));                                                    \
    } while (0)
    ED_EL(hipMalloc((void**)&op->d_erl, D * sizeof(uint16_t)));
    ED_EL(hipMalloc((void**)&op->d_ebase, (ng + 1) * sizeof(int64_t)));
    ED_EL(hipMalloc((void**)&op->d_ediag, D * sizeof(double)));
    ED_EL(hipMalloc((void**)&op->d_ectab, kMaxFam * sizeof(double)));
    if (!elist_build(op, st, false)) return fail(ed_fail(ED_EINVAL, "stored element list: shape not compiled"));
    ED_EL(hipGetLastError());
    int64_t* gs = nullptr;
    ED_EL(hipMalloc((void**)&gs, (ng + 1) * sizeof(int64_t)));
    hipLaunchKernelGGL(k_elist_groups, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, D, op->d_erl, ng, gs);
    hipLaunchKernelGGL(k_elist_ctab, dim3(1), dim3(kMaxFam), 0, st, op->nfam, op->d_tfco, op->d_ectab);
    size_t tmp_bytes = 0;
    void* tmp = nullptr;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, gs, op->d_ebase, (size_t)(ng + 1), st);
    if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, gs, op->d_ebase, (size_t)(ng + 1), st);
    int64_t slots = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&slots, op->d_ebase + ng, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dfree(tmp);
    dfree(gs);
    ED_EL(e);
    op->eslots = slots;
    ED_EL(hipMalloc((void**)&op->d_el, (size_t)std::max<int64_t>(slots, 1) * sizeof(uint32_t)));
    elist_build(op, st, true);
    ED_EL(hipGetLastError());
    int64_t nel = 0;
    {
        std::vector<uint16_t> h(D);
        ED_EL(hipMemcpyAsync(h.data(), op->d_erl, D * sizeof(uint16_t), hipMemcpyDeviceToHost, st));
        ED_EL(hipStreamSynchronize(st));
        for (uint16_t c : h) nel += c;
    }
    op->nelem = nel;
#undef ED_EL
    return ED_OK;
}

template <typename VT>
static int launch_elist(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st, int bs) {
    ed_basis_s* B = op->b;
    if (!op->d_el) return ed_fail(ED_EINVAL, "stored element list not built");
    const uint32_t zidx = B->img_nslot << B->img_cb;
    hipLaunchKernelGGL((k_hv_elist<4, VT>), dim3((unsigned)((B->D + bs - 1) / bs)), dim3(bs), 0, st, B->D, B->d_rows,
                       B->d_ntab, op->d_ebase, op->d_erl, op->d_el, op->d_ediag, op->d_ectab, op->nfam, op->ejbits,
                       zidx, u, v, w);
    HIPCHK(hipGetLastError());
    return ED_OK;
}

// Kernel configuration the next launch uses: path 2 (compiled shape),
// 3 (run-time shape); variant 0 (K=4), 1 (K=2 at 8 waves), 2 (K=8),
// 3 (K=4 pipelined), 4 (two lanes per row, K=8), 5 (four lanes per row,
// K=4), 6 (two lanes per row, K=4), -1 before tuning or for the run-time
// kernel.
void torus_config(const ed_op_s* op, int dslot, int& path, int& var, int& bs) {
    if (op->tune_path == 12) {
        path = 12;
        var = -1;
        bs = op->tune_block;
        return;
    }
    const bool compiled = op->tune_path != 3 && compiled_fits(op);
    path = compiled ? 2 : 3;
    bs = op->tune_block;
    var = -1;
    if (!compiled) return;
    if (op->tune_path == 2) var = 0;
    else if (op->tune_path == 4) var = 1;
    else if (op->tune_path == 5) var = 2;
    else if (op->tune_path == 7) var = 3;
    else if (op->tune_path == 9) var = variant_fits(op, 4) ? 4 : 0;
    else if (op->tune_path == 10) var = variant_fits(op, 5) ? 5 : 0;
    else if (op->tune_path == 11) var = variant_fits(op, 6) ? 6 : 0;
    else if (op->sel_var[dslot] >= 0) {
        var = op->sel_var[dslot];
        bs = op->sel_block[dslot];
    }
}

template <typename VT>
int launch_hv_torus(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st) {
    ed_basis_s* B = op->b;
    const int64_t D = B->D;
    int path, var, bs;
    torus_config(op, dtype_slot<VT>(), path, var, bs);
    if (path == 12) return launch_elist<VT>(op, u, v, w, st, bs);
    if (path == 2) {
        if (!launch_compiled<VT>(op, u, v, w, st, var < 0 ? 0 : var, bs))
            return ed_fail(ED_EINVAL, "compiled-shape kernel does not fit this sector");
        HIPCHK(hipGetLastError());
        return ED_OK;
    }
    dim3 grid((unsigned)((D + bs - 1) / bs)), blk(bs);
    size_t sm = torus_smem_bytes(op);
    int nph = B->L1 * B->L2;
    if (B->trivial_k)
        hipLaunchKernelGGL((k_hv_torus<true, VT>), grid, blk, sm, st, D, B->d_rows, op->torus, op->nfam, op->d_tfam,
                           op->d_tent, op->ezz0, B->rinfo, B->d_rtab32, B->rtab32_n, B->d_dir, B->d_phase, nph,
                           B->d_ntab, u, v, w);
    else
        hipLaunchKernelGGL((k_hv_torus<false, VT>), grid, blk, sm, st, D, B->d_rows, op->torus, op->nfam, op->d_tfam,
                           op->d_tent, op->ezz0, B->rinfo, B->d_rtab32, B->rtab32_n, B->d_dir, B->d_phase, nph,
                           B->d_ntab, u, v, w);
    HIPCHK(hipGetLastError());
    return ED_OK;
}

template int launch_hv_torus<cd2>(ed_op_s*, const cd2*, const cd2*, cd2*, hipStream_t);
template int launch_hv_torus<cf2>(ed_op_s*, const cf2*, const cf2*, cf2*, hipStream_t);
template int torus_prepare<cd2>(ed_op_s*, const cd2*, const cd2*, cd2*, hipStream_t);
template int torus_prepare<cf2>(ed_op_s*, const cf2*, const cf2*, cf2*, hipStream_t);
