// ed_hv_fast.hip -- rotation fast path of H*v (ed_torus.h) for N <= 32
// lattices whose translations are bit rotations.
//
// One lane per row.  Per lane the antiparallel bonds of each family are the
// set bits of one XOR; each set bit is one gathered element: flip, orbit
// minimum by rotations (alignbit + min3 tree), byte-sliced rank from LDS,
// 16-byte directory entry, then the gather of u = v/norm.  The rank tables,
// family entries and phase table are staged in LDS once per block.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_internal.h"
#include "ed_torus.h"
#include "ed_torus_host.h"

using namespace edc;

#ifndef ED_HV_BATCH
#define ED_HV_BATCH 4   // elements in flight per lane in the rotation kernel
#endif

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

// Same kernel with the lattice shape as compile-time constants (fully
// unrolled rotation search, masks as immediates).
template <class TC, bool TRIV, typename VT>
__global__ void __launch_bounds__(512) k_hv_torus_c(int64_t D, const uint64_t* __restrict__ rows, int nfam,
                                                    const TFam* __restrict__ gfam, const TEntry* __restrict__ gent,
                                                    double ezz0, RankInfo R, const uint32_t* __restrict__ grtab,
                                                    int rtab_n, const DirEntry* __restrict__ dir,
                                                    const cd2* __restrict__ gphase, int nphase,
                                                    const double* __restrict__ ntab, const VT* __restrict__ u,
                                                    const VT* __restrict__ v, VT* __restrict__ w) {
    extern __shared__ __align__(16) unsigned char smem[];
    TEntry* ent = (TEntry*)smem;
    TFam* fam = (TFam*)(ent + kMaxFam * 32);
    cd2* phase = (cd2*)(fam + kMaxFam);
    uint32_t* rtab = (uint32_t*)(phase + nphase);
    for (int q = threadIdx.x; q < nfam * 32; q += blockDim.x) ent[q] = gent[q];
    for (int q = threadIdx.x; q < nfam; q += blockDim.x) fam[q] = gfam[q];
    for (int q = threadIdx.x; q < nphase; q += blockDim.x) phase[q] = gphase[q];
    for (int q = threadIdx.x; q < rtab_n; q += blockDim.x) rtab[q] = grtab[q];
    __syncthreads();
    // Workgroups are dispatched round-robin over the 8 XCDs; remap so that
    // each XCD sweeps one contiguous range of rows and its L2 serves the
    // near-diagonal gathers of neighbouring rows.
    const unsigned nb = gridDim.x, q = nb / 8, rem = nb % 8;
    const unsigned xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const unsigned lb = xcd * q + (xcd < rem ? xcd : rem) + slot;
    {
        int64_t i = (int64_t)lb * blockDim.x + threadIdx.x;
        if (i >= D) return;
        uint64_t rw = rows[i];
        uint32_t s = (uint32_t)(rw & kStateMask);
        double nrm = ntab[(int)(rw >> kLidShift)];
        double ar = 0.0, ai = 0.0;
        double ez = torus_row_batched<TC, TRIV, ED_HV_BATCH>(
            s, nfam, fam, ent, ezz0, R, rtab, dir, phase, [&](int64_t j) { return load_c(u, j); },
            [&](const cd2& x, double cr, double ci) {
                ar += cr * x.re - ci * x.im;
                ai += cr * x.im + ci * x.re;
            });
        cd2 vi = load_c(v, i);
        VT o;
        o.re = nrm * ar + ez * vi.re;
        o.im = nrm * ai + ez * vi.im;
        w[i] = o;
    }
}

// Shapes compiled with constant geometry: the benchmark lattices and the
// small lattices the parity tests use.  Others run k_hv_torus.
#define ED_TORUS_SHAPES(X) \
    X(32, 4, 8)            \
    X(30, 6, 5)            \
    X(27, 3, 9)            \
    X(24, 24, 1)           \
    X(16, 4, 4)            \
    X(20, 4, 5)            \
    X(18, 6, 3)            \
    X(16, 16, 1)

template <class TC, typename VT>
static bool launch_c(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st, size_t sm) {
    ed_basis_s* B = op->b;
    const int64_t D = B->D;
    const int bs = op->tune_block;
    dim3 grid((unsigned)((D + bs - 1) / bs)), blk(bs);
    int nph = B->L1 * B->L2;
    if (B->trivial_k)
        hipLaunchKernelGGL((k_hv_torus_c<TC, true, VT>), grid, blk, sm, st, D, B->d_rows, op->nfam, op->d_tfam,
                           op->d_tent, op->ezz0, B->rinfo, B->d_rtab32, B->rtab32_n, B->d_dir, B->d_phase, nph,
                           B->d_ntab, u, v, w);
    else
        hipLaunchKernelGGL((k_hv_torus_c<TC, false, VT>), grid, blk, sm, st, D, B->d_rows, op->nfam, op->d_tfam,
                           op->d_tent, op->ezz0, B->rinfo, B->d_rtab32, B->rtab32_n, B->d_dir, B->d_phase, nph,
                           B->d_ntab, u, v, w);
    return true;
}

bool torus_shape_compiled(const Torus& T) {
    int L2 = T.kind == 2 ? 1 : T.L2;
#define ED_MATCH(n, wd, l2) if (T.N == n && T.w == wd && L2 == l2) return true;
    ED_TORUS_SHAPES(ED_MATCH)
#undef ED_MATCH
    return false;
}

size_t torus_smem_bytes(const ed_op_s* op) {
    const ed_basis_s* B = op->b;
    return kMaxFam * 32 * sizeof(TEntry) + kMaxFam * sizeof(TFam) + (size_t)B->L1 * B->L2 * sizeof(cd2) +
           (size_t)B->rtab32_n * sizeof(uint32_t);
}

template <typename VT>
int launch_hv_torus(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st) {
    ed_basis_s* B = op->b;
    const int64_t D = B->D;
    const int bs = op->tune_block;
    dim3 grid((unsigned)((D + bs - 1) / bs)), blk(bs);
    size_t sm = torus_smem_bytes(op);
    int nph = B->L1 * B->L2;
    if (op->tune_path != 3) {
        const Torus& T = op->torus;
        int L2 = T.kind == 2 ? 1 : T.L2;
#define ED_LAUNCH(n, wd, l2)                                               \
        if (T.N == n && T.w == wd && L2 == l2) {                           \
            launch_c<TorusC<n, wd, l2>, VT>(op, u, v, w, st, sm);          \
            HIPCHK(hipGetLastError());                                     \
            return ED_OK;                                                  \
        }
        ED_TORUS_SHAPES(ED_LAUNCH)
#undef ED_LAUNCH
    }
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
