// ed_hv64.hip -- H*v on two-row lattices with 64-bit states (ed_torus64.h):
// the relabelled tilted 40-site cluster and the 20-site test cluster.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_internal.h"
#include "ed_torus.h"
#include "ed_torus64.h"

using namespace edc;

// W when T1 rotates each of two W-site rows (site i + W*j -> (i+1)%W + W*j)
// and T2 exchanges the rows, W <= 26 (rows fit the exact double keys); else 0.
int rows2_detect(const std::vector<int>& T1, const std::vector<int>& T2) {
    const int N = (int)T1.size();
    if (N % 2 || N < 4 || N > 52) return 0;
    const int W = N / 2;
    for (int i = 0; i < N; i++) {
        if (T1[i] != (i / W) * W + (i % W + 1) % W) return 0;
        if (T2[i] != (i + W) % N) return 0;
    }
    return W;
}

bool rows2_shape_compiled(int N, int W) {
#define ED_R2MATCH(n, w) if (N == n && W == w) return true;
    ED_ROWS2_SHAPES(ED_R2MATCH)
#undef ED_R2MATCH
    return false;
}

// Split rank tables: every block but the last (the top byte's) as 32-bit.
int rows2_tables(ed_basis_s* B, const std::vector<uint64_t>& rt) {
    const RankInfo& R = B->rinfo;
    if (B->N > 40) return ED_OK;   // lower-byte sums would not fit 32 bits
    const size_t top = (size_t)R.off[R.nbytes - 1] * 256;
    std::vector<uint32_t> lo(top);
    for (size_t q = 0; q < top; q++) {
        if (rt[q] >> 32) return ed_fail(ED_EINVAL, "rank table entry exceeds 32 bits");
        lo[q] = (uint32_t)rt[q];
    }
    B->rlo_n = (int)top;
    HIPCHK(hipMalloc(&B->d_rlo, std::max<size_t>(top, 1) * sizeof(uint32_t)));
    if (top) HIPCHK(scopy(B->d_rlo, lo.data(), top * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&B->d_rhi, 256 * sizeof(uint64_t)));
    HIPCHK(scopy(B->d_rhi, rt.data() + top, 256 * sizeof(uint64_t), hipMemcpyHostToDevice));
    return ED_OK;
}

// Families of the two-row lattice (rows2_families), uploaded.
int rows2_op_setup(ed_op_s* op, const std::vector<Bond>& bonds, const std::vector<Chiral>& ch) {
    std::vector<TFam64> fam;
    std::vector<int> foff;
    std::vector<TEntry> ent;
    double ezz0 = 0;
    rows2_families(op->lam, op->Ch, op->b->rows2_w, bonds, ch, fam, foff, ent, ezz0);
    if (fam.empty()) return ED_OK;
    op->nfam64 = (int)fam.size();
    op->nterm64 = (int)ent.size();
    op->ezz0_64 = ezz0;
    HIPCHK(hipMalloc(&op->d_fam64, fam.size() * sizeof(TFam64)));
    HIPCHK(scopy(op->d_fam64, fam.data(), fam.size() * sizeof(TFam64), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&op->d_foff64, foff.size() * sizeof(int)));
    HIPCHK(scopy(op->d_foff64, foff.data(), foff.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&op->d_ent64, ent.size() * sizeof(TEntry)));
    HIPCHK(scopy(op->d_ent64, ent.data(), ent.size() * sizeof(TEntry), hipMemcpyHostToDevice));
    op->use_rows2 = true;
    return ED_OK;
}

namespace {

constexpr int kRows2K = 8;   // elements in flight per lane

template <class TC, bool TRIV, bool WIMG, typename VT, int BS>
__global__ void __launch_bounds__(BS) k_hv_rows2(int64_t D, const uint64_t* __restrict__ rows, int nfam,
                                                  const TFam64* __restrict__ gfam, const int* __restrict__ gfoff,
                                                  int nterm, const TEntry* __restrict__ gent, double ezz0, RankInfo R,
                                                  const uint32_t* __restrict__ grlo, int rlo_n,
                                                  const uint64_t* __restrict__ grhi, const DirEntry* __restrict__ dir,
                                                  const uint32_t* __restrict__ wmap, int cb,
                                                  const cd2* __restrict__ gphase, int nphase,
                                                  const double* __restrict__ ntab, const VT* __restrict__ u,
                                                  const VT* __restrict__ v, VT* __restrict__ w) {
    extern __shared__ __align__(16) unsigned char smem[];
    TEntry* ent = (TEntry*)smem;                         // nterm (32 B each)
    cd2* phase = (cd2*)(ent + nterm);                    // nphase
    TFam64* fam = (TFam64*)(phase + nphase);             // nfam (24 B each)
    uint64_t* rhi = (uint64_t*)(fam + nfam);             // 256
    int* foff = (int*)(rhi + 256);                       // nfam (+1 pad)
    uint32_t* rlo = (uint32_t*)(foff + nfam + (nfam & 1));   // rlo_n
    for (int q = threadIdx.x; q < nterm; q += blockDim.x) ent[q] = gent[q];
    for (int q = threadIdx.x; q < nphase; q += blockDim.x) phase[q] = gphase[q];
    for (int q = threadIdx.x; q < nfam; q += blockDim.x) { fam[q] = gfam[q]; foff[q] = gfoff[q]; }
    for (int q = threadIdx.x; q < 256; q += blockDim.x) rhi[q] = grhi[q];
    for (int q = threadIdx.x; q < rlo_n; q += blockDim.x) rlo[q] = grlo[q];
    __syncthreads();
    // XCD tiles of consecutive rows, the 8 XCDs on adjacent tiles
    const unsigned lb = xcd_tile_block();
    const int64_t i = (int64_t)lb * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const uint64_t rw = rows[i];
    const uint64_t s = rw & kStateMask;
    const double nrm = ntab[(int)(rw >> kLidShift)];
    double ar = 0.0, ai = 0.0;
    auto loadu = [&](int64_t j) { return load_c(u, j); };
    auto acc = [&](const cd2& x, double cr, double ci) {
        ar += cr * x.re - ci * x.im;
        ai += cr * x.im + ci * x.re;
    };
    double ez;
    if constexpr (WIMG) {
        // u is the wide gather image, addressed by rank through the chunk map
        // (cache resident)
        auto jget = [&](uint64_t r) -> uint32_t { return wmap[r >> cb]; };
        auto jres = [&](uint64_t r, uint32_t m) -> int64_t {
            return (int64_t)(((uint64_t)m << cb) | (r & ((uint64_t(1) << cb) - 1)));
        };
        ez = rows2_row<TC, TRIV, kRows2K>(s, nfam, fam, foff, ent, ezz0, R, rlo, rhi, phase, i, jget, jres, loadu, acc);
    } else {
        // u by row through the directory
        auto jget = [&](uint64_t r) -> DirEntry { return dir[r / kDirSpan]; };
        auto jres = [&](uint64_t r, const DirEntry& e) -> int64_t { return dir_decode_off(e, (uint32_t)(r % kDirSpan)); };
        ez = rows2_row<TC, TRIV, kRows2K>(s, nfam, fam, foff, ent, ezz0, R, rlo, rhi, phase, i, jget, jres, loadu, acc);
    }
    const cd2 vi = load_c(v, i);
    VT o;
    o.re = nrm * ar + ez * vi.re;
    o.im = nrm * ai + ez * vi.im;
    w[i] = o;
}

size_t rows2_smem(const ed_op_s* op) {
    const ed_basis_s* B = op->b;
    const int nf = op->nfam64;
    return (size_t)op->nterm64 * sizeof(TEntry) + (size_t)B->L1 * B->L2 * sizeof(cd2) + (size_t)nf * sizeof(TFam64) +
           256 * sizeof(uint64_t) + (size_t)(nf + (nf & 1)) * sizeof(int) + (size_t)B->rlo_n * sizeof(uint32_t);
}

// Workgroup size: the rank tables in LDS (about 53 KB for N = 40) are one
// copy per workgroup, so larger workgroups put more waves on a CU.
int rows2_block() {
    const char* e = std::getenv("ED_ROWS2_BLOCK");
    const int v = e ? std::atoi(e) : 0;
    return (v == 256 || v == 512 || v == 1024) ? v : 512;
}

template <class TC, typename VT, int BS>
void launch_r2_bs(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st, bool wimg) {
    ed_basis_s* B = op->b;
    dim3 grid((unsigned)((B->D + BS - 1) / BS)), blk(BS);
    const size_t sm = rows2_smem(op);
    const int nph = B->L1 * B->L2;
#define ED_R2GO(TRIV, WIMG)                                                                                        \
    hipLaunchKernelGGL((k_hv_rows2<TC, TRIV, WIMG, VT, BS>), grid, blk, sm, st, B->D, B->d_rows, op->nfam64,        \
                       (const TFam64*)op->d_fam64, op->d_foff64, op->nterm64, op->d_ent64, op->ezz0_64, B->rinfo,   \
                       B->d_rlo, B->rlo_n, B->d_rhi, B->d_dir, B->d_wimg_map, B->wimg_cb, B->d_phase, nph,          \
                       B->d_ntab, u, v, w)
    if (B->trivial_k) {
        if (wimg) ED_R2GO(true, true);
        else ED_R2GO(true, false);
    } else {
        if (wimg) ED_R2GO(false, true);
        else ED_R2GO(false, false);
    }
#undef ED_R2GO
}

template <class TC, typename VT>
void launch_r2(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st, bool wimg) {
    switch (rows2_block()) {
        case 256: launch_r2_bs<TC, VT, 256>(op, u, v, w, st, wimg); break;
        case 1024: launch_r2_bs<TC, VT, 1024>(op, u, v, w, st, wimg); break;
        default: launch_r2_bs<TC, VT, 512>(op, u, v, w, st, wimg); break;
    }
}

}  // namespace

template <typename VT>
int launch_hv_rows2(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st, bool wide_image) {
    const ed_basis_s* B = op->b;
    if (rows2_smem(op) > 64 * 1024) return ed_fail(ED_EINVAL, "two-row kernel tables exceed 64 KiB of LDS");
#define ED_R2LAUNCH(n, wd)                                   \
    if (B->N == n && B->rows2_w == wd) {                     \
        launch_r2<Rows2C<n, wd>, VT>(op, u, v, w, st, wide_image); \
        HIPCHK(hipGetLastError());                           \
        return ED_OK;                                        \
    }
    ED_ROWS2_SHAPES(ED_R2LAUNCH)
#undef ED_R2LAUNCH
    return ed_fail(ED_EINVAL, "two-row shape (%d, %d) is not compiled", B->N, B->rows2_w);
}

template int launch_hv_rows2<cd2>(ed_op_s*, const cd2*, const cd2*, cd2*, hipStream_t, bool);
template int launch_hv_rows2<cf2>(ed_op_s*, const cf2*, const cf2*, cf2*, hipStream_t, bool);
