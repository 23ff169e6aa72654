// ed_block.hip -- block H*V: M coupling instances on one momentum sector,
// W[:, m] = H_m V[:, m], in one pass over the rows.
//
// The instances share the term structure (the same bonds and triangles, in
// the same order; their couplings, anisotropy, chirality and phases may
// differ), so every gathered element -- flip, orbit minimum, rank, directory
// entry -- is found once and serves all M instances: the integer work and
// the directory traffic of a single H*v are paid once per block, the gather
// reads M adjacent values.  Vectors are interleaved: element (i, m) at
// i*M + m, i.e. a row-major D x M matrix, so that one target's M values are
// one contiguous load.
//
// Lattices without the compiled rotation fast path (or with bowtie terms)
// run each instance through its own operator on de-interleaved copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_internal.h"
#include "ed_torus.h"
#include "ed_torus_host.h"

using namespace edc;

namespace {

struct BEnt {            // per term: flip partner y, chiral z-site k0 (-1 for a bond)
    int32_t y, k0;
    int32_t k1, flags;   // merged terms: second chiral z-site; TChir flags
};
struct BCoef {           // per (term, instance): bond part (or the chiral entry)
    double cre, cim, dJ2;
};
struct alignas(16) BChir {   // per (term, instance), merged terms: the pair's coefficient at s_x = 0
    cd2 c[4];                // for the z-site bits s_k0 | s_k1 << 1 (merged_coef4; conjugated for s_x = 1)
};

inline unsigned nblk(int64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// The M instances' u = v / norm written into the interleaved gather image:
// image element pos[i] of row i holds its M values.
template <typename VT>
__global__ void k_prescale_block(int64_t D, int M, const uint64_t* rows, const double* nitab, const uint32_t* pos,
                                 const VT* v, VT* img) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= D * M) return;
    const int64_t i = e / M;
    double f = nitab[(int)(rows[i] >> kLidShift)];
    VT x = v[e];
    x.re = x.re * f;
    x.im = x.im * f;
    img[(int64_t)pos[i] * M + (e - i * M)] = x;
}

template <typename VT>
__global__ void k_zero_block(uint64_t n, VT* p) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
        VT z;
        z.re = 0;
        z.im = 0;
        p[e] = z;
    }
}

template <typename VT>
__global__ void k_deinterleave(int64_t D, int M, int m, const VT* V, VT* x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < D) x[i] = V[i * M + m];
}

template <typename VT>
__global__ void k_interleave(int64_t D, int M, int m, const VT* x, VT* V) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < D) V[i * M + m] = x[i];
}

__device__ __forceinline__ void st_c(cd2* p, double re, double im) { *p = cd2{re, im}; }
__device__ __forceinline__ void st_c(cf2* p, double re, double im) { *p = cf2{(float)re, (float)im}; }

// One lane per row; K elements in flight; M instances accumulated per element.
// Same element order, coefficient expressions and accumulation order as the
// single-instance rotation kernel (torus_row_batched).  Terms are numbered
// densely: term = foff[f] + (number of sites of family f below x).
template <class TC, bool TRIV, bool MG, int M, int K, typename VT>
__global__ void __launch_bounds__(512, M >= 8 ? 4 : 2) k_hv_block_c(int64_t D, const uint64_t* __restrict__ rows, int nfam,
                                                    const TFam* __restrict__ gfam, const int* __restrict__ gfoff,
                                                    int nterm, const BEnt* __restrict__ gent,
                                                    const BCoef* __restrict__ gcoef, const BChir* __restrict__ gchir,
                                                    const double* __restrict__ gezz0,
                                                    RankInfo R, const uint32_t* __restrict__ grtab, int rtab_n,
                                                    const uint16_t* __restrict__ gmap, int nmap, int cb,
                                                    uint32_t zidx, const cd2* __restrict__ gphase,
                                                    int nphase, const double* __restrict__ ntab,
                                                    const VT* __restrict__ u, const VT* __restrict__ v,
                                                    VT* __restrict__ w) {
    extern __shared__ __align__(16) unsigned char smem[];
    BChir* chir = (BChir*)smem;                                // nterm*M (merged terms)
    cd2* phase = (cd2*)(chir + (MG ? nterm * M : 0));          // nphase
    BCoef* coef = (BCoef*)(phase + nphase);                    // nterm*M (terms that are not merged)
    double* dj2 = (double*)coef;                               // nterm*M (merged terms: only dJ2)
    TFam* fam = MG ? (TFam*)(dj2 + nterm * M) : (TFam*)(coef + nterm * M);   // nfam
    int* foff = (int*)(fam + nfam);                            // nfam (+1 pad)
    BEnt* ent = (BEnt*)(foff + nfam + (nfam & 1));             // nterm
    double* ezz0 = (double*)(ent + nterm);                     // M
    uint32_t* rtab = (uint32_t*)(ezz0 + M);                    // rtab_n
    uint16_t* map = (uint16_t*)(rtab + rtab_n);                // nmap: gather-image chunk -> slot
    if (MG)
        for (int q = threadIdx.x; q < nterm * M; q += blockDim.x) chir[q] = gchir[q];
    for (int q = threadIdx.x; q < nphase; q += blockDim.x) phase[q] = gphase[q];
    if (MG)
        for (int q = threadIdx.x; q < nterm * M; q += blockDim.x) dj2[q] = gcoef[q].dJ2;
    else
        for (int q = threadIdx.x; q < nterm * M; q += blockDim.x) coef[q] = gcoef[q];
    for (int q = threadIdx.x; q < nfam; q += blockDim.x) { fam[q] = gfam[q]; foff[q] = gfoff[q]; }
    for (int q = threadIdx.x; q < nterm; q += blockDim.x) ent[q] = gent[q];
    for (int q = threadIdx.x; q < M; q += blockDim.x) ezz0[q] = gezz0[q];
    for (int q = threadIdx.x; q < rtab_n; q += blockDim.x) rtab[q] = grtab[q];
    for (int q = threadIdx.x; q < nmap; q += blockDim.x) map[q] = gmap[q];
    __syncthreads();
    // XCD tiles of consecutive rows, the 8 XCDs on adjacent tiles
    const unsigned lb = xcd_tile_block();
    const int64_t i = (int64_t)lb * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const uint64_t rw = rows[i];
    const uint32_t s = (uint32_t)(rw & kStateMask);
    const double nrm = ntab[(int)(rw >> kLidShift)];
    double ar[M], ai[M], ez[M];
#pragma unroll
    for (int m = 0; m < M; m++) { ar[m] = 0.0; ai[m] = 0.0; ez[m] = ezz0[m]; }
    int f = 0;
    uint32_t cur = nfam > 0 ? fam_mask_c<TC>(s, fam[0]) : 0u;
    bool more = true;
    while (more) {
        int ti[K], xs[K], code[K];
        uint32_t rk[K];
        bool ok[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            while (cur == 0u && f < nfam) {
                if (++f < nfam) cur = fam_mask_c<TC>(s, fam[f]);
            }
            ok[q] = f < nfam;
            ti[q] = 0; xs[q] = 0; rk[q] = 0; code[q] = 0;
            if (ok[q]) {
                const int x = ctz32(cur);
                cur &= cur - 1u;
                xs[q] = x;
                ti[q] = foff[f] + popc32(fam[f].S & ((1u << x) - 1u));
                const uint32_t t = s ^ (1u << x) ^ (1u << ent[ti[q]].y);
                int cd;
                const uint32_t rep = torus_rep_c<TC, TRIV ? kRepOnly : kRepAny>(t, cd);
                code[q] = cd;
                rk[q] = rank_tab<uint32_t, uint32_t>(rep, R, rtab, (TC::N + 7) / 8);
            }
        }
        more = ok[K - 1];
        // unconditional loads from the gather image (an empty slot reads the
        // zero slot; so does a target outside the sector)
        VT xv[K][M];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const VT* p = u + (int64_t)image_index(map, cb, zidx, rk[q], ok[q]) * M;
#pragma unroll
            for (int m = 0; m < M; m++) xv[q][m] = p[m];
        }
#pragma unroll
        for (int q = 0; q < K; q++) {
            const BEnt E = ent[ti[q]];
            const bool bond = MG || E.k0 < 0;
            const bool sx = (s >> xs[q]) & 1u, sy = (s >> E.y) & 1u, sk = (s >> (E.k0 & 31)) & 1u;
            // merged terms: the pair's coefficient by its two z-site bits
            const int cidx = (int)(((s >> (E.k0 & 31)) & 1u) | (((s >> (E.k1 & 31)) & 1u) << 1));
            cd2 ph{1.0, 0.0};
            if (!TRIV) ph = phase[code[q]];
#pragma unroll
            for (int m = 0; m < M; m++) {
                double cr, ci;
                if (MG) {
                    if (ok[q]) ez[m] -= dj2[ti[q] * M + m];
                    const cd2 co = chir[ti[q] * M + m].c[cidx];
                    cr = co.re;
                    ci = sx ? -co.im : co.im;
                } else if (bond) {
                    const BCoef C = coef[ti[q] * M + m];
                    if (ok[q]) ez[m] -= C.dJ2;
                    cr = C.cre;
                    ci = sx ? -C.cim : C.cim;
                } else {
                    const BCoef C = coef[ti[q] * M + m];
                    const double sg = sk ? 1.0 : -1.0;
                    cr = -C.cim * sg;
                    ci = (sy ? C.cre : -C.cre) * sg;
                }
                if (!TRIV) {
                    const double nr = cr * ph.re - ci * ph.im, ni = cr * ph.im + ci * ph.re;
                    cr = nr;
                    ci = ni;
                }
                if (!ok[q]) { cr = 0.0; ci = 0.0; }
                const cd2 xx = load_c(&xv[q][m], 0);
                ar[m] += cr * xx.re - ci * xx.im;
                ai[m] += cr * xx.im + ci * xx.re;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < M; m++) {
        const cd2 vi = load_c(v, i * M + m);
        st_c(w + i * M + m, nrm * ar[m] + ez[m] * vi.re, nrm * ai[m] + ez[m] * vi.im);
    }
}

}  // namespace

struct ed_block_s {
    ed_basis_s* b = nullptr;
    int M = 0;
    std::vector<ed_op_s*> ops;      // one operator per instance (fallback path)
    bool fast = false;              // compiled rotation block kernel applies
    int nfam = 0, nterm = 0;
    TFam* d_fam = nullptr;
    int* d_foff = nullptr;
    BEnt* d_ent = nullptr;
    BCoef* d_coef = nullptr;
    BChir* d_chir = nullptr;        // merged terms (chiral entries on the bond pairs)
    bool mg = false;
    double* d_ezz0 = nullptr;
    void* d_u64 = nullptr;          // prescaled interleaved gather image, complex64 [image elements * M]
    void* d_u128 = nullptr;         // complex128
    void* d_tmp = nullptr;          // fallback: two de-interleaved vectors, complex128-sized
};

namespace {

// Bond and chiral lists of instance m over the union of the instances' terms:
// a term that is nonzero in any instance is kept in every instance (with its
// own, possibly zero, coupling), so that all instances produce the same
// families; terms zero in all instances are dropped as ed_op_create drops them.
void union_lists(const ed_terms* T, int M, int m, std::vector<Bond>& bonds, std::vector<Chiral>& ch) {
    const ed_terms* t = &T[m];
    bonds.clear();
    ch.clear();
    for (int k = 0; k < t->nbonds; k++) {
        bool any = false;
        for (int q = 0; q < M; q++) any |= std::fabs(T[q].bond_J[k]) >= 1e-16;
        if (any)
            bonds.push_back(Bond{t->bond_xy[2 * k], t->bond_xy[2 * k + 1], t->bond_J[k], t->bond_phase[2 * k],
                                 t->bond_phase[2 * k + 1]});
    }
    bool anych = false;
    for (int q = 0; q < M; q++) anych |= std::fabs(T[q].Ch) > 1e-16;
    if (anych)
        for (int k = 0; k < t->nchiral; k++)
            ch.push_back(Chiral{t->chiral_ijk[3 * k], t->chiral_ijk[3 * k + 1], t->chiral_ijk[3 * k + 2], 0,
                                t->chiral_phase[2 * k], t->chiral_phase[2 * k + 1]});
}

size_t block_smem(const ed_block_s* K) {
    const ed_basis_s* B = K->b;
    const int nf = K->nfam;
    return (size_t)K->nterm * K->M * (K->mg ? sizeof(BChir) + sizeof(double) : sizeof(BCoef)) +
           (size_t)B->L1 * B->L2 * sizeof(cd2) + (size_t)nf * sizeof(TFam) +
           (size_t)(nf + (nf & 1)) * sizeof(int) + (size_t)K->nterm * sizeof(BEnt) + (size_t)K->M * sizeof(double) +
           (size_t)B->rtab32_n * sizeof(uint32_t) + (size_t)B->img_nmap * sizeof(uint16_t);
}

template <class TC, int M, typename VT>
void launch_block_c(ed_block_s* K, const VT* u, const VT* v, VT* w, hipStream_t st) {
    ed_basis_s* B = K->b;
    constexpr int KB = M >= 8 ? 1 : M >= 4 ? 2 : 4;
    // the tables in LDS are one copy per workgroup: for M <= 4 (59 KB with
    // chiral terms, within 128 VGPRs) 512 threads put twice the waves on a
    // CU; M = 8 needs 156-198 VGPRs, so a CU holds 2 waves per SIMD either way
    const unsigned bs = 512;
    dim3 grid(nblk(B->D, bs)), blk(bs);
    size_t sm = block_smem(K);
    int nph = B->L1 * B->L2;
    const uint32_t zidx = B->img_nslot << B->img_cb;   // first element of the zero slot
#define ED_BGO(TRIV, MG)                                                                                         \
    hipLaunchKernelGGL((k_hv_block_c<TC, TRIV, MG, M, KB, VT>), grid, blk, sm, st, B->D, B->d_rows, K->nfam, K->d_fam, \
                       K->d_foff, K->nterm, K->d_ent, K->d_coef, K->d_chir, K->d_ezz0, B->rinfo, B->d_rtab32,          \
                       B->rtab32_n, B->d_img_map, (int)B->img_nmap, B->img_cb, zidx, B->d_phase, nph, B->d_ntab, u, v, w)
    if (B->trivial_k) {
        if (K->mg) ED_BGO(true, true);
        else ED_BGO(true, false);
    } else {
        if (K->mg) ED_BGO(false, true);
        else ED_BGO(false, false);
    }
#undef ED_BGO
}

template <int M, typename VT>
bool dispatch_shape(ed_block_s* K, const VT* u, const VT* v, VT* w, hipStream_t st) {
    const Torus& T = K->b->torus;
    int L2 = T.kind == 2 ? 1 : T.L2;
#define ED_BLAUNCH(n, wd, l2)                                       \
    if (T.N == n && T.w == wd && L2 == l2) {                        \
        launch_block_c<TorusC<n, wd, l2>, M, VT>(K, u, v, w, st);   \
        return true;                                                \
    }
    ED_TORUS_SHAPES(ED_BLAUNCH)
#undef ED_BLAUNCH
    return false;
}

template <typename VT>
int block_hv(ed_block_s* K, const VT* v, VT* w, hipStream_t st) {
    ed_basis_s* B = K->b;
    const int64_t D = B->D, M = K->M;
    if (D == 0) return ED_OK;
    const int dt = std::is_same<VT, cd2>::value ? ED_C128 : ED_C64;
    if (M == 1) return op_hv(K->ops[0], dt, v, w, st);
    if (!K->fast) {
        if (!K->d_tmp) HIPCHK(hipMalloc(&K->d_tmp, 2 * D * sizeof(cd2)));
        VT* x = (VT*)K->d_tmp;
        VT* y = x + D;
        for (int m = 0; m < M; m++) {
            hipLaunchKernelGGL(k_deinterleave<VT>, dim3(nblk(D, 256)), dim3(256), 0, st, D, (int)M, m, v, x);
            HIPCHK(hipGetLastError());
            int rc = op_hv(K->ops[m], dt, x, y, st);
            if (rc != ED_OK) return rc;
            hipLaunchKernelGGL(k_interleave<VT>, dim3(nblk(D, 256)), dim3(256), 0, st, D, (int)M, m, y, w);
            HIPCHK(hipGetLastError());
        }
        return ED_OK;
    }
    VT*& u = std::is_same<VT, cd2>::value ? (VT*&)K->d_u128 : (VT*&)K->d_u64;
    if (!u) {
        // the zero positions of the image are written once, at allocation
        const uint64_t ne = ((uint64_t)(B->img_nslot + 1) << B->img_cb) * M;
        HIPCHK(hipMalloc((void**)&u, ne * sizeof(VT)));
        hipLaunchKernelGGL(k_zero_block<VT>, dim3((unsigned)std::min<uint64_t>((ne + 255) / 256, 1u << 20)), dim3(256),
                           0, st, ne, u);
    }
    hipLaunchKernelGGL(k_prescale_block<VT>, dim3(nblk(D * M, 256)), dim3(256), 0, st, D, (int)M, B->d_rows,
                       B->d_nitab, B->d_img_pos, v, u);
    HIPCHK(hipGetLastError());
    bool done = false;
    switch (M) {
        case 2: done = dispatch_shape<2, VT>(K, u, v, w, st); break;
        case 4: done = dispatch_shape<4, VT>(K, u, v, w, st); break;
        case 8: done = dispatch_shape<8, VT>(K, u, v, w, st); break;
        default: break;
    }
    if (!done) return ed_fail(ED_EINVAL, "no block kernel for this shape and M=%d", (int)M);
    HIPCHK(hipGetLastError());
    return ED_OK;
}

}  // namespace

int block_op_hv(ed_block_s* K, int dtype, const void* v, void* w, hipStream_t st) {
    if (dtype == ED_C128) return block_hv<cd2>(K, (const cd2*)v, (cd2*)w, st);
    if (dtype == ED_C64) return block_hv<cf2>(K, (const cf2*)v, (cf2*)w, st);
    return ed_fail(ED_EINVAL, "unknown dtype %d", dtype);
}

ed_basis_s* block_basis(const ed_block_s* K) { return K->b; }
int block_size(const ed_block_s* K) { return K->M; }

extern "C" void ed_block_destroy(ed_block blk) {
    if (!blk) return;
    for (auto* op : blk->ops) ed_op_destroy(op);
    for (void* p : {(void*)blk->d_fam, (void*)blk->d_foff, (void*)blk->d_ent, (void*)blk->d_coef, (void*)blk->d_chir,
                    (void*)blk->d_ezz0, blk->d_u64, blk->d_u128, blk->d_tmp})
        if (p) dfree(p);
    delete blk;
}

extern "C" int ed_block_create(ed_basis b, int32_t M, const ed_terms* terms, ed_block* out) {
    if (!b || !terms || !out) return ed_fail(ED_EINVAL, "null argument");
    if (M < 1 || M > 8) return ed_fail(ED_EINVAL, "M=%d not in 1..8", M);
    const ed_terms& t0 = terms[0];
    for (int m = 1; m < M; m++) {
        const ed_terms& t = terms[m];
        if (t.nbonds != t0.nbonds || t.nchiral != t0.nchiral || t.nbowtie != t0.nbowtie)
            return ed_fail(ED_EINVAL, "instance %d: term counts differ from instance 0", m);
        if (t.nbonds && std::memcmp(t.bond_xy, t0.bond_xy, 2 * sizeof(int32_t) * t.nbonds))
            return ed_fail(ED_EINVAL, "instance %d: bond sites differ from instance 0", m);
        if (t.nchiral && std::memcmp(t.chiral_ijk, t0.chiral_ijk, 3 * sizeof(int32_t) * t.nchiral))
            return ed_fail(ED_EINVAL, "instance %d: chiral sites differ from instance 0", m);
        if (t.nbowtie && std::memcmp(t.bowtie, t0.bowtie, 6 * sizeof(int32_t) * t.nbowtie))
            return ed_fail(ED_EINVAL, "instance %d: bowtie sites differ from instance 0", m);
    }
    auto* K = new ed_block_s();
    K->b = b;
    K->M = M;
    for (int m = 0; m < M; m++) {
        ed_op op = nullptr;
        int rc = ed_op_create(b, &terms[m], &op);
        if (rc != ED_OK) {
            ed_block_destroy(K);
            return rc;
        }
        K->ops.push_back(op);
    }
    // compiled rotation block kernel: rotational lattice with a compiled
    // shape, no bowtie terms in any instance, M in {2, 4, 8}
    bool nobow = true;
    for (int m = 0; m < M; m++)
        if (terms[m].nbowtie && std::fabs(terms[m].ChCh) > 1e-16) nobow = false;
    if (b->has_torus && b->N <= 32 && b->d_rtab32 && nobow && (M == 2 || M == 4 || M == 8) &&
        torus_shape_compiled(b->torus) && b->img_cb >= 0) {
        std::vector<std::vector<TFam>> fams(M);
        std::vector<std::vector<TEntry>> ents(M);
        std::vector<std::vector<TChir>> chxs(M);
        std::vector<double> ezz0(M);
        // with chiral terms, first the merged families (one term per flipped
        // pair, as the single-instance kernels); else one term per entry
        bool anych = false;
        for (int m = 0; m < M; m++) anych |= terms[m].nchiral > 0 && std::fabs(terms[m].Ch) > 1e-16;
        bool ok = false;
        for (int attempt = anych ? 0 : 1; attempt < 2 && !ok; attempt++) {
            const bool merged = attempt == 0;
            ok = true;
            for (int m = 0; m < M && ok; m++) {
                std::vector<Bond> bonds;
                std::vector<Chiral> ch;
                union_lists(terms, M, m, bonds, ch);
                ok = (merged ? torus_families_merged(terms[m].lambda, terms[m].Ch, b->torus, bonds, ch, fams[m],
                                                     ents[m], chxs[m], ezz0[m])
                             : torus_families(terms[m].lambda, terms[m].Ch, b->torus, bonds, ch, fams[m], ents[m],
                                              ezz0[m])) &&
                     !fams[m].empty();
            }
            for (int m = 1; m < M && ok; m++) {
                ok = fams[m].size() == fams[0].size();
                for (size_t f = 0; ok && f < fams[0].size(); f++) {
                    ok = fams[m][f].type == fams[0][f].type && fams[m][f].a == fams[0][f].a &&
                         fams[m][f].q == fams[0][f].q && fams[m][f].S == fams[0][f].S;
                    for (int x = 0; ok && merged && x < 32; x++) {
                        const TChir &A = chxs[m][f * 32 + x], &C0 = chxs[0][f * 32 + x];
                        ok = A.flags == C0.flags && A.k[0] == C0.k[0] && A.k[1] == C0.k[1];
                    }
                }
            }
            K->mg = ok && merged;
        }
        if (ok) {
            const int nf = (int)fams[0].size();
            std::vector<int> foff(nf);
            std::vector<BEnt> be;
            std::vector<BCoef> bc;
            std::vector<BChir> bx;
            for (int f = 0; f < nf; f++) {
                foff[f] = (int)be.size();
                for (int x = 0; x < 32; x++) {
                    if (!((fams[0][f].S >> x) & 1u)) continue;
                    const TEntry& E0 = ents[0][f * 32 + x];
                    if (K->mg) {
                        const TChir& X0 = chxs[0][f * 32 + x];
                        be.push_back(BEnt{E0.y, X0.k[0], X0.k[1], X0.flags});
                    } else {
                        be.push_back(BEnt{E0.y, fams[0][f].type == 0 ? -1 : E0.k, 0, 0});
                    }
                    for (int m = 0; m < M; m++) {
                        const TEntry& E = ents[m][f * 32 + x];
                        bc.push_back(BCoef{E.cre, E.cim, E.dJ2});
                        if (K->mg) {
                                BChir X;
                            merged_coef4(E, chxs[m][f * 32 + x], X.c);
                            bx.push_back(X);
                        }
                    }
                }
            }
            K->nfam = nf;
            K->nterm = (int)be.size();
            auto up = [&](void** d, const void* h, size_t n) -> int {
                HIPCHK(hipMalloc(d, n));
                HIPCHK(scopy(*d, h, n, hipMemcpyHostToDevice));
                return ED_OK;
            };
            int rc = up((void**)&K->d_fam, fams[0].data(), nf * sizeof(TFam));
            if (rc == ED_OK) rc = up((void**)&K->d_foff, foff.data(), nf * sizeof(int));
            if (rc == ED_OK) rc = up((void**)&K->d_ent, be.data(), be.size() * sizeof(BEnt));
            if (rc == ED_OK) rc = up((void**)&K->d_coef, bc.data(), bc.size() * sizeof(BCoef));
            if (rc == ED_OK && K->mg) rc = up((void**)&K->d_chir, bx.data(), bx.size() * sizeof(BChir));
            if (rc == ED_OK) rc = up((void**)&K->d_ezz0, ezz0.data(), M * sizeof(double));
            if (rc != ED_OK) {
                ed_block_destroy(K);
                return rc;
            }
            K->fast = block_smem(K) <= lds_per_block();
        }
    }
    *out = K;
    return ED_OK;
}

extern "C" int ed_block_hv(ed_block blk, int32_t dtype, const void* V, void* W, void* stream) {
    if (!blk || ((!V || !W) && blk->b->D)) return ed_fail(ED_EINVAL, "null argument");
    if (V == W && blk->b->D) return ed_fail(ED_EINVAL, "V and W must not alias");
    return block_op_hv(blk, dtype, V, W, (hipStream_t)stream);
}

extern "C" int ed_block_fast(ed_block blk) { return blk && blk->fast ? 1 : 0; }

// ------------------------------------------------ multi-sector block H*V ----
// Many momentum sectors of one lattice and one n_ones, M coupling instances
// each, in one launch: the work queue is every sector's row tiles of 256
// rows, in sector order; a workgroup maps its (XCD-tiled) queue position to
// (sector, tile), stages that sector's gather-image map and phase table in
// LDS next to the tables the sectors share (families, coefficients, rank
// tables), and runs the rows as the single-sector block kernel does.  The
// reference sweeps the sectors of a lattice one ARPACK run at a time
// (ARPACK_full_eigs.cpp:169-277).  The queue is a plain grid, not resident
// workgroups looping over tiles: on this kernel resident loops measured
// slower (the LDS tables are staged per workgroup either way, and a
// workgroup per tile keeps the occupancy the row kernel wants).
//
// Vectors: the sectors' interleaved D_s x M blocks concatenated in sector
// order; sector s occupies rows [row0_s, row0_s + D_s).  k = 0 sectors run the
// phase-tracking orbit search too, with their all-ones phase table, which
// gives the same sums.
namespace {

constexpr int kMTile = 256;

struct MSec {
    const uint64_t* rows;    // [D] representative | lid << 52
    const uint16_t* map;     // gather-image chunk -> slot
    const uint32_t* pos;     // [D] image element of each row
    const cd2* phase;        // [L1*L2]
    int64_t D, row0, img0, tile0;
    int32_t nmap, cb;
    uint32_t zidx;
    int32_t pad;
};

__device__ __forceinline__ int msec_of_tile(int nsec, const MSec* __restrict__ secs, int64_t t) {
    int lo = 0, hi = nsec - 1;
    while (lo < hi) {   // last sector whose first tile is <= t
        const int mid = (lo + hi + 1) >> 1;
        if (secs[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// u = v / norm into the concatenated image, for the rows of the sectors in
// the launch queue only: element e is instance e % M of tile slot e / M
// (slot = tile * kMTile + row in tile); slots past a sector's last row idle.
template <typename VT>
__global__ void k_prescale_mblock(int64_t ntiles, int M, int nsec, const MSec* __restrict__ secs,
                                  const double* __restrict__ nitab, const VT* __restrict__ v, VT* __restrict__ img) {
    const int64_t n = ntiles * kMTile * M;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = e / M;
        const int64_t t = g / kMTile;
        const MSec& S = secs[msec_of_tile(nsec, secs, t)];
        const int64_t i = (t - S.tile0) * kMTile + (g - t * kMTile);
        if (i >= S.D) continue;
        const int m = (int)(e - g * M);
        const double f = nitab[(int)(S.rows[i] >> kLidShift)];
        VT x = v[(S.row0 + i) * M + m];
        x.re = x.re * f;
        x.im = x.im * f;
        img[(S.img0 + (int64_t)S.pos[i]) * M + m] = x;
    }
}

template <class TC, bool MG, int M, int K, typename VT>
__global__ void __launch_bounds__(256) k_hv_mblock_c(int nsec, const MSec* __restrict__ secs, int nfam,
                                                     const TFam* __restrict__ gfam, const int* __restrict__ gfoff,
                                                     int nterm, const BEnt* __restrict__ gent,
                                                     const BCoef* __restrict__ gcoef, const BChir* __restrict__ gchir,
                                                     const double* __restrict__ gezz0, RankInfo R,
                                                     const uint32_t* __restrict__ grtab, int rtab_n, int nmap_max,
                                                     int nphase, const double* __restrict__ ntab,
                                                     const VT* __restrict__ u, const VT* __restrict__ v,
                                                     VT* __restrict__ w) {
    extern __shared__ __align__(16) unsigned char smem[];
    BChir* chir = (BChir*)smem;                                // nterm*M (merged terms)
    cd2* phase = (cd2*)(chir + (MG ? nterm * M : 0));          // nphase
    BCoef* coef = (BCoef*)(phase + nphase);                    // nterm*M (terms that are not merged)
    double* dj2 = (double*)coef;                               // nterm*M (merged terms: only dJ2)
    TFam* fam = MG ? (TFam*)(dj2 + nterm * M) : (TFam*)(coef + nterm * M);   // nfam
    int* foff = (int*)(fam + nfam);                            // nfam (+1 pad)
    BEnt* ent = (BEnt*)(foff + nfam + (nfam & 1));             // nterm
    double* ezz0 = (double*)(ent + nterm);                     // M
    uint32_t* rtab = (uint32_t*)(ezz0 + M);                    // rtab_n
    uint16_t* map = (uint16_t*)(rtab + rtab_n);                // nmap_max
    const int64_t t = (int64_t)xcd_tile_block();
    const MSec S = secs[msec_of_tile(nsec, secs, t)];
    if (MG)
        for (int q = threadIdx.x; q < nterm * M; q += blockDim.x) chir[q] = gchir[q];
    for (int q = threadIdx.x; q < nphase; q += blockDim.x) phase[q] = S.phase[q];
    if (MG)
        for (int q = threadIdx.x; q < nterm * M; q += blockDim.x) dj2[q] = gcoef[q].dJ2;
    else
        for (int q = threadIdx.x; q < nterm * M; q += blockDim.x) coef[q] = gcoef[q];
    for (int q = threadIdx.x; q < nfam; q += blockDim.x) { fam[q] = gfam[q]; foff[q] = gfoff[q]; }
    for (int q = threadIdx.x; q < nterm; q += blockDim.x) ent[q] = gent[q];
    for (int q = threadIdx.x; q < M; q += blockDim.x) ezz0[q] = gezz0[q];
    for (int q = threadIdx.x; q < rtab_n; q += blockDim.x) rtab[q] = grtab[q];
    for (int q = threadIdx.x; q < S.nmap; q += blockDim.x) map[q] = S.map[q];
    __syncthreads();
    const int64_t i = (t - S.tile0) * kMTile + threadIdx.x;
    if (i >= S.D) return;
    const uint64_t rw = S.rows[i];
    const uint32_t s = (uint32_t)(rw & kStateMask);
    const double nrm = ntab[(int)(rw >> kLidShift)];
    const VT* us = u + S.img0 * M;
    double ar[M], ai[M], ez[M];
#pragma unroll
    for (int m = 0; m < M; m++) { ar[m] = 0.0; ai[m] = 0.0; ez[m] = ezz0[m]; }
    int f = 0;
    uint32_t cur = nfam > 0 ? fam_mask_c<TC>(s, fam[0]) : 0u;
    bool more = true;
    while (more) {
        int ti[K], xs[K], code[K];
        uint32_t rk[K];
        bool ok[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            while (cur == 0u && f < nfam) {
                if (++f < nfam) cur = fam_mask_c<TC>(s, fam[f]);
            }
            ok[q] = f < nfam;
            ti[q] = 0; xs[q] = 0; rk[q] = 0; code[q] = 0;
            if (ok[q]) {
                const int x = ctz32(cur);
                cur &= cur - 1u;
                xs[q] = x;
                ti[q] = foff[f] + popc32(fam[f].S & ((1u << x) - 1u));
                const uint32_t tt = s ^ (1u << x) ^ (1u << ent[ti[q]].y);
                int cd;
                const uint32_t rep = torus_rep_c<TC, kRepAny>(tt, cd);
                code[q] = cd;
                rk[q] = rank_tab<uint32_t, uint32_t>(rep, R, rtab, (TC::N + 7) / 8);
            }
        }
        more = ok[K - 1];
        VT xv[K][M];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const VT* p = us + (int64_t)image_index(map, S.cb, S.zidx, rk[q], ok[q]) * M;
#pragma unroll
            for (int m = 0; m < M; m++) xv[q][m] = p[m];
        }
#pragma unroll
        for (int q = 0; q < K; q++) {
            const BEnt E = ent[ti[q]];
            const bool bond = MG || E.k0 < 0;
            const bool sx = (s >> xs[q]) & 1u, sy = (s >> E.y) & 1u, sk = (s >> (E.k0 & 31)) & 1u;
            // merged terms: the pair's coefficient by its two z-site bits
            const int cidx = (int)(((s >> (E.k0 & 31)) & 1u) | (((s >> (E.k1 & 31)) & 1u) << 1));
            const cd2 ph = phase[code[q]];
#pragma unroll
            for (int m = 0; m < M; m++) {
                double cr, ci;
                if (MG) {
                    if (ok[q]) ez[m] -= dj2[ti[q] * M + m];
                    const cd2 co = chir[ti[q] * M + m].c[cidx];
                    cr = co.re;
                    ci = sx ? -co.im : co.im;
                } else if (bond) {
                    const BCoef Cf = coef[ti[q] * M + m];
                    if (ok[q]) ez[m] -= Cf.dJ2;
                    cr = Cf.cre;
                    ci = sx ? -Cf.cim : Cf.cim;
                } else {
                    const BCoef Cf = coef[ti[q] * M + m];
                    const double sg = sk ? 1.0 : -1.0;
                    cr = -Cf.cim * sg;
                    ci = (sy ? Cf.cre : -Cf.cre) * sg;
                }
                const double nr = cr * ph.re - ci * ph.im, ni = cr * ph.im + ci * ph.re;
                cr = nr;
                ci = ni;
                if (!ok[q]) { cr = 0.0; ci = 0.0; }
                const cd2 xx = load_c(&xv[q][m], 0);
                ar[m] += cr * xx.re - ci * xx.im;
                ai[m] += cr * xx.im + ci * xx.re;
            }
        }
    }
    const int64_t g = S.row0 + i;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const cd2 vi = load_c(v, g * M + m);
        st_c(w + g * M + m, nrm * ar[m] + ez[m] * vi.re, nrm * ai[m] + ez[m] * vi.im);
    }
}

}  // namespace

struct ed_mblock_s {
    int M = 0;
    std::vector<ed_basis_s*> bases;
    std::vector<ed_block_s*> blocks;   // one block operator per sector (its fallback and its FP64 checks)
    std::vector<int64_t> row0;         // [nsec + 1]
    std::vector<MSec> secs_all;        // every non-empty sector, in order (tile0 unused)
    std::vector<int> sec_of;           // secs_all index -> sector index
    std::vector<char> live;            // per sector: in the launch queue
    std::vector<MSec> secs;            // the launch queue: live non-empty sectors, tile0 over them
    MSec* d_secs = nullptr;            // [secs_all.size()]
    int64_t ntiles = 0, img_elems = 0;
    int nmap_max = 0;
    bool fast = false;                 // one launch for all sectors
    void* d_u64 = nullptr;             // concatenated interleaved gather images
    void* d_u128 = nullptr;
};

namespace {

size_t mblock_smem(const ed_mblock_s* MB) {
    const ed_block_s* K = MB->blocks[0];
    const ed_basis_s* B = K->b;
    const int nf = K->nfam;
    return (size_t)K->nterm * K->M * (K->mg ? sizeof(BChir) + sizeof(double) : sizeof(BCoef)) +
           (size_t)B->L1 * B->L2 * sizeof(cd2) + (size_t)nf * sizeof(TFam) +
           (size_t)(nf + (nf & 1)) * sizeof(int) + (size_t)K->nterm * sizeof(BEnt) + (size_t)K->M * sizeof(double) +
           (size_t)B->rtab32_n * sizeof(uint32_t) + (size_t)MB->nmap_max * sizeof(uint16_t);
}

template <class TC, int M, typename VT>
void launch_mblock_c(ed_mblock_s* MB, const VT* u, const VT* v, VT* w, hipStream_t st) {
    const ed_block_s* K = MB->blocks[0];
    const ed_basis_s* B = K->b;
    constexpr int KB = M >= 4 ? 2 : 4;
    const int nsec = (int)MB->secs.size();
    const size_t sm = mblock_smem(MB);
    const int nph = B->L1 * B->L2;
    if (K->mg)
        hipLaunchKernelGGL((k_hv_mblock_c<TC, true, M, KB, VT>), dim3((unsigned)MB->ntiles), dim3(kMTile), sm, st, nsec,
                           MB->d_secs, K->nfam, K->d_fam, K->d_foff, K->nterm, K->d_ent, K->d_coef, K->d_chir,
                           K->d_ezz0, B->rinfo, B->d_rtab32, B->rtab32_n, MB->nmap_max, nph, B->d_ntab, u, v, w);
    else
        hipLaunchKernelGGL((k_hv_mblock_c<TC, false, M, KB, VT>), dim3((unsigned)MB->ntiles), dim3(kMTile), sm, st,
                           nsec, MB->d_secs, K->nfam, K->d_fam, K->d_foff, K->nterm, K->d_ent, K->d_coef, K->d_chir,
                           K->d_ezz0, B->rinfo, B->d_rtab32, B->rtab32_n, MB->nmap_max, nph, B->d_ntab, u, v, w);
}

template <int M, typename VT>
bool mdispatch_shape(ed_mblock_s* MB, const VT* u, const VT* v, VT* w, hipStream_t st) {
    const Torus& T = MB->bases[0]->torus;
    int L2 = T.kind == 2 ? 1 : T.L2;
#define ED_MLAUNCH(n, wd, l2)                                        \
    if (T.N == n && T.w == wd && L2 == l2) {                         \
        launch_mblock_c<TorusC<n, wd, l2>, M, VT>(MB, u, v, w, st);  \
        return true;                                                 \
    }
    ED_TORUS_SHAPES(ED_MLAUNCH)
#undef ED_MLAUNCH
    return false;
}

// The launch queue of the live sectors, their tiles numbered in sector order.
int mblock_upload_queue(ed_mblock_s* MB) {
    MB->secs.clear();
    int64_t tiles = 0;
    for (size_t q = 0; q < MB->secs_all.size(); q++) {
        if (!MB->live[MB->sec_of[q]]) continue;
        MSec S = MB->secs_all[q];
        S.tile0 = tiles;
        tiles += (S.D + kMTile - 1) / kMTile;
        MB->secs.push_back(S);
    }
    MB->ntiles = tiles;
    if (!MB->secs.empty())
        HIPCHK(scopy(MB->d_secs, MB->secs.data(), MB->secs.size() * sizeof(MSec), hipMemcpyHostToDevice));
    return ED_OK;
}

template <typename VT>
int mblock_hv(ed_mblock_s* MB, const VT* v, VT* w, hipStream_t st) {
    const int M = MB->M;
    const int nb = (int)MB->bases.size();   // all sectors (the launch queue holds the non-empty live ones)
    const int dt = std::is_same<VT, cd2>::value ? ED_C128 : ED_C64;
    if (!MB->fast) {
        // the sectors one after the other, each through its block operator
        for (int s = 0; s < nb; s++) {
            if (MB->bases[s]->D == 0 || !MB->live[s]) continue;
            int rc = block_op_hv(MB->blocks[s], dt, v + MB->row0[s] * M, w + MB->row0[s] * M, st);
            if (rc != ED_OK) return rc;
        }
        return ED_OK;
    }
    VT*& u = std::is_same<VT, cd2>::value ? (VT*&)MB->d_u128 : (VT*&)MB->d_u64;
    if (!u) {
        // the zero positions of the images are written once, at allocation
        const uint64_t ne = (uint64_t)MB->img_elems * M;
        HIPCHK(hipMalloc((void**)&u, ne * sizeof(VT)));
        hipLaunchKernelGGL(k_zero_block<VT>, dim3((unsigned)std::min<uint64_t>((ne + 255) / 256, 1u << 20)), dim3(256),
                           0, st, ne, u);
    }
    if (MB->secs.empty() || MB->ntiles == 0) return ED_OK;
    const int64_t n = MB->ntiles * kMTile * M;
    hipLaunchKernelGGL(k_prescale_mblock<VT>, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1 << 20)), dim3(256), 0,
                       st, MB->ntiles, M, (int)MB->secs.size(), MB->d_secs, MB->bases[0]->d_nitab, v, u);
    HIPCHK(hipGetLastError());
    bool done = false;
    switch (M) {
        case 2: done = mdispatch_shape<2, VT>(MB, u, v, w, st); break;
        case 4: done = mdispatch_shape<4, VT>(MB, u, v, w, st); break;
        case 8: done = mdispatch_shape<8, VT>(MB, u, v, w, st); break;
        default: break;
    }
    if (!done) return ed_fail(ED_EINVAL, "no multi-sector block kernel for this shape and M=%d", M);
    HIPCHK(hipGetLastError());
    return ED_OK;
}

}  // namespace

int mblock_op_hv(ed_mblock_s* MB, int dtype, const void* v, void* w, hipStream_t st) {
    if (dtype == ED_C128) return mblock_hv<cd2>(MB, (const cd2*)v, (cd2*)w, st);
    if (dtype == ED_C64) return mblock_hv<cf2>(MB, (const cf2*)v, (cf2*)w, st);
    return ed_fail(ED_EINVAL, "unknown dtype %d", dtype);
}

// Restricts the next H*V launches to the sectors with live[s] != 0 (rows of
// the other sectors are neither read nor written).
int mblock_set_live(ed_mblock_s* MB, const std::vector<char>& live) {
    if (live == MB->live) return ED_OK;
    MB->live = live;
    if (!MB->fast) return ED_OK;
    return mblock_upload_queue(MB);
}

int mblock_nsec(const ed_mblock_s* MB) { return (int)MB->bases.size(); }
int mblock_M(const ed_mblock_s* MB) { return MB->M; }
int64_t mblock_row0(const ed_mblock_s* MB, int s) { return MB->row0[s]; }
ed_block_s* mblock_block(const ed_mblock_s* MB, int s) { return MB->blocks[s]; }

extern "C" void ed_mblock_destroy(ed_mblock mb) {
    if (!mb) return;
    for (auto* K : mb->blocks) ed_block_destroy(K);
    for (void* p : {(void*)mb->d_secs, mb->d_u64, mb->d_u128})
        if (p) dfree(p);
    delete mb;
}

extern "C" int ed_mblock_create(const ed_basis* bases, int32_t nsec, int32_t M, const ed_terms* terms, ed_mblock* out) {
    if (!bases || !terms || !out) return ed_fail(ED_EINVAL, "null argument");
    if (nsec < 1 || nsec > 4096) return ed_fail(ED_EINVAL, "nsec=%d not in 1..4096", nsec);
    const ed_basis_s* B0 = bases[0];
    for (int s = 0; s < nsec; s++) {
        const ed_basis_s* B = bases[s];
        if (!B) return ed_fail(ED_EINVAL, "null basis %d", s);
        if (B->N != B0->N || B->n != B0->n || B->T1 != B0->T1 || B->T2 != B0->T2 || B->T3 != B0->T3)
            return ed_fail(ED_EINVAL, "basis %d: another lattice or n_ones than basis 0", s);
    }
    auto* MB = new ed_mblock_s();
    MB->M = M;
    MB->row0.assign(nsec + 1, 0);
    MB->live.assign(nsec, 1);
    for (int s = 0; s < nsec; s++) {
        ed_block K = nullptr;
        int rc = ed_block_create(bases[s], M, terms, &K);
        if (rc != ED_OK) {
            ed_mblock_destroy(MB);
            return rc;
        }
        MB->bases.push_back(bases[s]);
        MB->blocks.push_back(K);
        MB->row0[s + 1] = MB->row0[s] + bases[s]->D;
    }
    // one launch when every sector takes the shared block kernel with the
    // same term class, the sectors' LDS fits and the tile count fits a grid
    bool fast = true;
    int64_t tiles = 0, img = 0;
    std::vector<int> sec_of;
    for (int s = 0; s < nsec && fast; s++) {
        const ed_block_s* K = MB->blocks[s];
        const ed_basis_s* B = bases[s];
        fast = K->fast && K->mg == MB->blocks[0]->mg && B->img_cb >= 0;
        if (!fast) break;
        MSec S;
        std::memset(&S, 0, sizeof(S));
        S.rows = B->d_rows;
        S.map = B->d_img_map;
        S.pos = B->d_img_pos;
        S.phase = B->d_phase;
        S.D = B->D;
        S.row0 = MB->row0[s];
        S.img0 = img;
        S.tile0 = tiles;
        S.nmap = (int)B->img_nmap;
        S.cb = B->img_cb;
        S.zidx = B->img_nslot << B->img_cb;
        MB->secs.push_back(S);
        sec_of.push_back(s);
        MB->nmap_max = std::max(MB->nmap_max, (int)B->img_nmap);
        tiles += (B->D + kMTile - 1) / kMTile;
        img += (int64_t)(B->img_nslot + 1) << B->img_cb;
    }
    if (fast && tiles > 0 && tiles < (int64_t(1) << 31) && mblock_smem(MB) <= lds_per_block()) {
        // sectors with no rows take no tile; the search over first tiles
        // needs them strictly ordered, so they stay out of the queue
        for (size_t q = 0; q < MB->secs.size(); q++)
            if (MB->secs[q].D > 0) {
                MB->secs_all.push_back(MB->secs[q]);
                MB->sec_of.push_back(sec_of[q]);
            }
        if (!MB->secs_all.empty()) {
            int rc = hipMalloc(&MB->d_secs, MB->secs_all.size() * sizeof(MSec)) == hipSuccess ? ED_OK : ED_ENOMEM;
            MB->img_elems = img;
            MB->fast = true;
            if (rc == ED_OK) rc = mblock_upload_queue(MB);
            if (rc != ED_OK) {
                ed_mblock_destroy(MB);
                return ed_fail(ED_ENOMEM, "multi-sector table upload failed");
            }
        }
    }
    if (!MB->fast) MB->secs.clear();
    *out = MB;
    return ED_OK;
}

extern "C" int ed_mblock_rows(ed_mblock mb, int32_t sector, int64_t* row0, int64_t* D) {
    if (!mb || sector < 0 || sector > (int)mb->bases.size()) return ed_fail(ED_EINVAL, "bad sector index");
    const int ns = (int)mb->bases.size();
    if (row0) *row0 = mb->row0[std::min(sector, ns)];
    if (D) *D = sector == ns ? mb->row0[ns] : mb->bases[sector]->D;
    return ED_OK;
}

extern "C" int ed_mblock_fast(ed_mblock mb) { return mb && mb->fast ? 1 : 0; }

extern "C" int ed_mblock_hv(ed_mblock mb, int32_t dtype, const void* V, void* W, void* stream) {
    if (!mb || ((!V || !W) && mb->row0.back())) return ed_fail(ED_EINVAL, "null argument");
    if (V == W && mb->row0.back()) return ed_fail(ED_EINVAL, "V and W must not alias");
    int rc = mblock_set_live(mb, std::vector<char>(mb->bases.size(), 1));
    if (rc != ED_OK) return rc;
    return mblock_op_hv(mb, dtype, V, W, (hipStream_t)stream);
}
