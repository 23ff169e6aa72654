// ed_amd.hip -- gfx950 basis construction, H*v and the C ABI of include/ed_amd.h.
//
// Data layout in HBM per sector (N sites, n up spins, C = C(N,n) states,
// D representatives):
//   rows  : uint64 [D]   representative | (lid << 52), ascending; lid = number
//                        of translation-grid points fixing it, which gives
//                        norm = sqrt(L1*L2/lid) through a tiny table
//   dir   : 12 B [C/64]  {base, 64 occupancy bits} over the combinatorial
//                        rank of every weight-n state: index of a
//                        representative = base + popcount of the bits below
//   rtab  : uint64       byte-sliced rank tables (a few KB)
//   phase : complex128 [L1*L2] exp(-i(px lx + py ly)), computed on the host
//                        with the reference's expression (bit-equal)
// Nothing of size C survives the build except the 1.5 bit/state directory.
#include <hip/hip_runtime.h>

#include <mutex>
#include <hipcub/hipcub.hpp>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_internal.h"
#include "ed_tables.h"
#include "ed_torus_host.h"

using namespace edc;

// ----------------------------------------------------------------- errors --
static thread_local std::string g_err;

hipStream_t setup_stream() {
    thread_local hipStream_t st[16] = {};
    int dev = 0;
    // created once per thread and device and kept for the thread's life; a
    // device past the table or a failed creation uses the legacy stream (0)
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 0;
    if (!st[dev] && hipStreamCreateWithFlags(&st[dev], hipStreamNonBlocking) != hipSuccess) st[dev] = 0;
    return st[dev];
}

int ed_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char* ed_last_error(void) { return g_err.c_str(); }

// Gather stream of ed_hv_noalloc (u = v/norm by row, or the sector's gather
// image) in a file-scope device array: each GPU the library runs on holds its
// own copy until the process ends.  Host code never passes the array's name
// to a kernel (on the host it is only a stand-in for the symbol): it uses the
// device address that hipGetSymbolAddress returned for the current device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) by the first ed_hv_noalloc there: the
// lookup loads the code object and places the array, which processes that
// never call ed_hv_noalloc do not pay for.
int scratch_ptr(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return ed_fail(ED_EINVAL, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return ed_fail(ED_EDEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return ED_OK;
}
}  // namespace

extern "C" int ed_init(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return ed_fail(ED_EDEVICE, "no HIP device: %s", hipGetErrorString(e));
    if (device < 0 || device >= n) return ed_fail(ED_EINVAL, "device %d out of range (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    return ED_OK;
}

// --------------------------------------------------------------- lattice --
extern "C" int ed_lattice_create3(int32_t n_sites, const int32_t* T1, const int32_t* T2, const int32_t* T3,
                                  ed_lattice* out) {
    if (!out || !T1 || !T2) return ed_fail(ED_EINVAL, "null argument");
    if (n_sites < 1 || n_sites > 63) return ed_fail(ED_EINVAL, "n_sites=%d outside [1,63]", n_sites);
    std::vector<int> a(T1, T1 + n_sites), b(T2, T2 + n_sites), c(n_sites);
    for (int i = 0; i < n_sites; i++) c[i] = T3 ? T3[i] : i;
    for (auto* T : {&a, &b, &c}) {
        std::vector<char> seen(n_sites, 0);
        for (int i = 0; i < n_sites; i++) {
            int t = (*T)[i];
            if (t < 0 || t >= n_sites || seen[t]) return ed_fail(ED_EINVAL, "T1/T2/T3 must be permutations of 0..%d", n_sites - 1);
            seen[t] = 1;
        }
    }
    for (int i = 0; i < n_sites; i++)
        if (a[b[i]] != b[a[i]] || a[c[i]] != c[a[i]] || b[c[i]] != c[b[i]])
            return ed_fail(ED_EINVAL, "the translations do not commute");
    auto* L = new ed_lattice_s();
    L->N = n_sites;
    L->T1 = a;
    L->T2 = b;
    L->T3 = c;
    L->L1 = cycle_len0(a);
    L->L2 = cycle_len0(b);
    L->L3 = cycle_len0(c);
    // the rotation fast paths cover two-generator lattices
    if (L->L3 == 1) {
        L->has_torus = torus_detect(a, b, L->torus);
        L->rows2_w = rows2_detect(a, b);
    }
    if (!make_prog(a, L->P1) || !make_prog(b, L->P2) || !make_prog(c, L->P3)) {
        delete L;
        return ed_fail(ED_EINVAL, "translation has more than %d displacement groups", kMaxGroups);
    }
    *out = L;
    return ED_OK;
}

extern "C" int ed_lattice_create(int32_t n_sites, const int32_t* T1, const int32_t* T2, ed_lattice* out) {
    return ed_lattice_create3(n_sites, T1, T2, nullptr, out);
}

extern "C" int ed_lattice_orders(ed_lattice lat, int32_t* L1, int32_t* L2) {
    if (!lat) return ed_fail(ED_EINVAL, "null lattice");
    if (L1) *L1 = lat->L1;
    if (L2) *L2 = lat->L2;
    return ED_OK;
}

extern "C" int ed_lattice_order3(ed_lattice lat, int32_t* L3) {
    if (!lat || !L3) return ed_fail(ED_EINVAL, "null argument");
    *L3 = lat->L3;
    return ED_OK;
}

extern "C" void ed_lattice_destroy(ed_lattice lat) { delete lat; }

// ---------------------------------------------------------------- kernels --
// The basis kernels loop over their index range with a grid stride: one
// dispatch covers at most 2^32 work-items per dimension, and at N=40 the
// classification runs over 4.3e9 words.
#define ED_GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

template <typename S>
__global__ void k_classify(uint64_t C, int N, int n, Geo G, const uint64_t* binom, uint32_t* words, uint64_t nwords) {
    ED_GRID_STRIDE(w, nwords) {
        uint64_t r0 = w * 32;
        uint32_t bits = 0;
        if (r0 < C) {
            uint64_t s = unrank(r0, N, n, binom);
            for (int b = 0; b < 32; b++) {
                if (r0 + b >= C) break;
                if (classify<S>((S)s, G)) bits |= 1u << b;
                s = gosper_next(s);
            }
        }
        words[w] = bits;
    }
}

__global__ void k_group_counts(const uint32_t* words, uint64_t ngroups, uint64_t* cnt) {
    ED_GRID_STRIDE(g, ngroups) {
        cnt[g] = (uint64_t)(__popc(words[2 * g]) + __popc(words[2 * g + 1]));
    }
}

__global__ void k_dir_fill(const uint32_t* words, const uint64_t* base, uint64_t ngroups, DirEntry* dir) {
    ED_GRID_STRIDE(g, ngroups) {
        DirEntry e;
        e.base = (uint32_t)base[g];
        e.b0 = words[2 * g];
        e.b1 = words[2 * g + 1];
        dir[g] = e;
    }
}

// Marks the gather-image chunks (2^(6 + shift) ranks each) that hold a
// representative of the sector.  Threads of one chunk store the same byte.
__global__ void k_image_touch(const DirEntry* dir, uint64_t ngroups, int shift, uint8_t* touched) {
    ED_GRID_STRIDE(g, ngroups) {
        const DirEntry e = dir[g];
        if (e.b0 | e.b1) touched[g >> shift] = 1;
    }
}

// Image element of every row: the rank of its representative through the
// chunk map.
__global__ void k_image_pos(uint64_t D, const uint64_t* rows, RankInfo R, const uint64_t* rtab, const uint16_t* map,
                            int cb, uint32_t* pos) {
    ED_GRID_STRIDE(i, D) {
        const uint32_t r = (uint32_t)rank_of<uint32_t>((uint32_t)(rows[i] & kStateMask), R, rtab);
        pos[i] = ((uint32_t)map[r >> cb] << cb) | (r & ((1u << cb) - 1u));
    }
}

template <typename S>
__global__ void k_compact(uint64_t C, int N, int n, Geo G, const uint64_t* binom, const DirEntry* dir,
                          uint64_t ngroups, uint64_t* rows) {
    ED_GRID_STRIDE(g, ngroups) {
        DirEntry e = dir[g];
        if ((e.b0 | e.b1) == 0) continue;
        uint64_t r0 = g * kDirSpan;
        uint64_t s = unrank(r0, N, n, binom);
        uint64_t j = e.base;
        for (int b = 0; b < kDirSpan && r0 + b < C; b++) {
            uint32_t wd = b < 32 ? e.b0 : e.b1;
            if ((wd >> (b & 31)) & 1u) {
                int lid = classify<S>((S)s, G);
                rows[j++] = s | ((uint64_t)lid << kLidShift);
            }
            s = gosper_next(s);
        }
    }
}

__global__ void k_norms(const uint64_t* rows, int64_t D, const double* ntab, const double* nitab, double* norm,
                        double* norm_inv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    int lid = (int)(rows[i] >> kLidShift);
    if (norm) norm[i] = ntab[lid];
    if (norm_inv) norm_inv[i] = nitab[lid];
}

template <typename S>
__global__ void k_full_tables(uint64_t C, int N, int n, Geo G, const uint64_t* binom, RankInfo R, TablesView tv,
                              int64_t* rep_loc, cd2* state_ph) {
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t r0 = w * 32;
    if (r0 >= C) return;
    uint64_t s = unrank(r0, N, n, binom);
    for (int b = 0; b < 32 && r0 + b < C; b++) {
        int code;
        S rep = find_rep<S, true>((S)s, G, code);
        rep_loc[r0 + b] = dir_lookup(rank_of<S>(rep, R, tv.rtab), tv.dir);
        state_ph[r0 + b] = tv.phase[code];
        s = gosper_next(s);
    }
}

// u = v / norm, the gathered stream of the H*v kernels that address u by
// row through the directory.
template <typename VT>
__global__ void k_prescale(int64_t D, const uint64_t* rows, const double* nitab, const VT* v, VT* u) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    double f = nitab[(int)(rows[i] >> kLidShift)];
    VT x = v[i];
    x.re = x.re * f;
    x.im = x.im * f;
    u[i] = x;
}

// The same values written into the gather image (element pos[i] for row i;
// ascending with i, so the stores of a wave fall into few lines).
template <typename VT>
__global__ void k_prescale_img(int64_t D, const uint64_t* rows, const double* nitab, const uint32_t* pos, const VT* v,
                               VT* img) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    double f = nitab[(int)(rows[i] >> kLidShift)];
    VT x = v[i];
    x.re = x.re * f;
    x.im = x.im * f;
    img[pos[i]] = x;
}

// The same into the wide image of the two-row kernels: the rank of each
// row's representative through the global chunk map (64-bit positions).
template <typename VT>
__global__ void k_prescale_wimg(uint64_t D, const uint64_t* rows, const double* nitab, RankInfo R,
                                const uint64_t* rtab, const uint32_t* map, int cb, const VT* v, VT* img) {
    ED_GRID_STRIDE(i, D) {
        const uint64_t rw = rows[i];
        const uint64_t r = rank_of<uint64_t>(rw & kStateMask, R, rtab);
        const uint64_t e = ((uint64_t)map[r >> cb] << cb) | (r & ((uint64_t(1) << cb) - 1));
        double f = nitab[(int)(rw >> kLidShift)];
        VT x = v[i];
        x.re = x.re * f;
        x.im = x.im * f;
        img[e] = x;
    }
}

// Zero fill with a kernel (no memset node if the caller captures the stream).
template <typename VT>
__global__ void k_zero(uint64_t n, VT* p) {
    ED_GRID_STRIDE(i, n) {
        VT z;
        z.re = 0;
        z.im = 0;
        p[i] = z;
    }
}

template <typename VT>
__device__ __forceinline__ void store_c(VT* p, int64_t i, double re, double im);
template <>
__device__ __forceinline__ void store_c<cd2>(cd2* p, int64_t i, double re, double im) { p[i] = cd2{re, im}; }
template <>
__device__ __forceinline__ void store_c<cf2>(cf2* p, int64_t i, double re, double im) {
    p[i] = cf2{(float)re, (float)im};
}

// w[i] = norm[i] * sum_terms coef * u[j] + Ezz * v[i]: one lane per row.
template <typename S, bool TRIV, typename VT>
__global__ void __launch_bounds__(256) k_hv_rows(int64_t D, const uint64_t* rows, Geo G, RankInfo R, TablesView tv,
                                                 TermsView T, const double* ntab, const VT* u, const VT* v, VT* w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    uint64_t rw = rows[i];
    S s = (S)(rw & kStateMask);
    double nrm = ntab[(int)(rw >> kLidShift)];
    double ar = 0.0, ai = 0.0;
    double ez = visit_row<S, TRIV>(s, G, R, tv, T, [&](int64_t j, double cr, double ci) {
        cd2 x = load_c(u, j);
        ar += cr * x.re - ci * x.im;
        ai += cr * x.im + ci * x.re;
    });
    cd2 vi = load_c(v, i);
    store_c<VT>(w, i, nrm * ar + ez * vi.re, nrm * ai + ez * vi.im);
}

// Distinct targets of a row are collected in a per-lane list of at most
// kMaxRowTargets; a row with more (bowtie terms give hundreds) is counted
// exactly by extracting its distinct targets in ascending order, one pass over
// the row per value.
constexpr int kMaxRowTargets = 384;

template <typename S, bool TRIV>
__global__ void k_count_targets(int64_t D, const uint64_t* rows, Geo G, RankInfo R, TablesView tv, TermsView T,
                                unsigned long long* distinct, unsigned long long* total) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long dcount = 0, tcount = 0;
    if (i < D) {
        int64_t js[kMaxRowTargets];
        int nj = 0;
        bool over = false;
        S s = (S)(rows[i] & kStateMask);
        visit_row<S, TRIV>(s, G, R, tv, T, [&](int64_t j, double, double) {
            tcount++;
            for (int q = 0; q < nj; q++)
                if (js[q] == j) return;
            if (nj < kMaxRowTargets) js[nj++] = j; else over = true;
        });
        dcount = (unsigned long long)nj;
        if (over) {
            dcount = 0;
            int64_t cur = -1;
            while (true) {
                int64_t nxt = INT64_MAX;
                visit_row<S, TRIV>(s, G, R, tv, T, [&](int64_t j, double, double) {
                    if (j > cur && j < nxt) nxt = j;
                });
                if (nxt == INT64_MAX) break;
                dcount++;
                cur = nxt;
            }
        }
    }
    atomicAdd(distinct, dcount);
    atomicAdd(total, tcount);
}

template <typename S, bool TRIV>
__global__ void k_make_h(int64_t D, const uint64_t* rows, Geo G, RankInfo R, TablesView tv, TermsView T,
                         const double* ntab, const double* nitab, cd2* H) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    uint64_t rw = rows[i];
    S s = (S)(rw & kStateMask);
    double nrm = ntab[(int)(rw >> kLidShift)];
    double ez = visit_row<S, TRIV>(s, G, R, tv, T, [&](int64_t j, double cr, double ci) {
        double f = nrm * nitab[(int)(rows[j] >> kLidShift)];
        cd2 h = H[i + j * D];
        h.re += cr * f;
        h.im += ci * f;
        H[i + j * D] = h;
    });
    H[i + i * D].re += ez;
}

static inline unsigned nblocks(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }
// grid for the grid-stride kernels: at most 2^22 blocks (2^30 work-items)
// Block count of a grid-stride launch: one element per lane up to the cap,
// then every lane loops (at 40 sites the basis kernels see 2^31 work-items).
// The cap is 2^22 blocks; the environment variable ED_GRID_CAP lowers it
// (read once), so that small sectors run the same loops in tests.
static unsigned grid_cap() {
    static const unsigned cap = [] {
        const char* e = std::getenv("ED_GRID_CAP");
        unsigned long v = e ? std::strtoul(e, nullptr, 10) : 0ul;
        return (v >= 1 && v < (1ul << 22)) ? (unsigned)v : (1u << 22);
    }();
    return cap;
}
static inline unsigned sblocks(uint64_t n, unsigned bs) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + bs - 1) / bs, grid_cap()));
}

// ------------------------------------------------------------------ basis --
int basis_geo(const ed_basis_s* B, Geo& G) {
    std::memset(&G, 0, sizeof(G));
    G.N = B->N; G.L1 = B->L1; G.L2 = B->L2; G.L3 = B->L3; G.kx = B->kx; G.ky = B->ky; G.kz = B->kz;
    G.T1 = B->lat_P1;
    G.T2 = B->lat_P2;
    G.T3 = B->lat_P3;
    return ED_OK;
}

extern "C" int ed_basis_build3(ed_lattice lat, int32_t n_ones, int32_t kx, int32_t ky, int32_t kz, ed_basis* out) {
    if (!lat || !out) return ed_fail(ED_EINVAL, "null argument");
    const int N = lat->N;
    if (n_ones < 0 || n_ones > N) return ed_fail(ED_EINVAL, "n_ones=%d outside [0,%d] (Num sites < Num_ones)", n_ones, N);
    if (kx < 0 || ky < 0 || kz < 0 || kx >= lat->L1 || ky >= lat->L2 || kz >= lat->L3)
        return ed_fail(ED_EINVAL, "Invalid kx, ky or kz: (%d,%d,%d) with L1=%d L2=%d L3=%d", kx, ky, kz, lat->L1, lat->L2,
                       lat->L3);
    auto* B = new ed_basis_s();
    B->N = N; B->n = n_ones; B->kx = kx; B->ky = ky; B->kz = kz;
    B->L1 = lat->L1; B->L2 = lat->L2; B->L3 = lat->L3;
    B->T1 = lat->T1; B->T2 = lat->T2; B->T3 = lat->T3;
    B->lat_P1 = lat->P1; B->lat_P2 = lat->P2; B->lat_P3 = lat->P3;
    B->has_torus = lat->has_torus;
    B->torus = lat->torus;
    B->rows2_w = lat->rows2_w;
    // Lanczos_sector.cpp:57 (three generators :240)
    B->px = 2 * kx * M_PI / B->L1;
    B->py = 2 * ky * M_PI / B->L2;
    B->pz = 2 * kz * M_PI / B->L3;
    B->C = binom_u64(N, n_ones);
    B->trivial_k = (kx == 0 && ky == 0 && kz == 0);
    int rc = basis_build_device(B);
    if (rc != ED_OK) { delete_basis(B); return rc; }
    *out = B;
    return ED_OK;
}

extern "C" int ed_basis_build(ed_lattice lat, int32_t n_ones, int32_t kx, int32_t ky, ed_basis* out) {
    return ed_basis_build3(lat, n_ones, kx, ky, 0, out);
}

// Host-side tables + the device build (classify -> scan -> directory -> compact).
int basis_build_device(ed_basis_s* B) {
    const int N = B->N, n = B->n, L1 = B->L1, L2 = B->L2;
    std::vector<cd2> ph = phase_table(L1, L2, B->L3, B->px, B->py, B->pz);
    std::vector<double> nt, nit;
    norm_tables(L1, L2, B->L3, nt, nit);
    std::vector<uint64_t> bn = binom_table();
    std::vector<uint64_t> rt = rank_tables(N, n, B->rinfo);
    HIPCHK(hipMalloc(&B->d_phase, ph.size() * sizeof(cd2)));
    HIPCHK(hipMalloc(&B->d_ntab, nt.size() * sizeof(double)));
    HIPCHK(hipMalloc(&B->d_nitab, nit.size() * sizeof(double)));
    HIPCHK(hipMalloc(&B->d_binom, bn.size() * sizeof(uint64_t)));
    HIPCHK(hipMalloc(&B->d_rtab, rt.size() * sizeof(uint64_t)));
    HIPCHK(scopy(B->d_phase, ph.data(), ph.size() * sizeof(cd2), hipMemcpyHostToDevice));
    HIPCHK(scopy(B->d_ntab, nt.data(), nt.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(scopy(B->d_nitab, nit.data(), nit.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(scopy(B->d_binom, bn.data(), bn.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIPCHK(scopy(B->d_rtab, rt.data(), rt.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    if (B->rows2_w) {
        int rc = rows2_tables(B, rt);
        if (rc != ED_OK) return rc;
    }
    if (N <= 32) {
        std::vector<uint32_t> rt32(rt.size());
        for (size_t q = 0; q < rt.size(); q++) rt32[q] = (uint32_t)rt[q];   // C(N,n) < 2^32
        B->rtab32_n = (int)rt32.size();
        HIPCHK(hipMalloc(&B->d_rtab32, rt32.size() * sizeof(uint32_t)));
        HIPCHK(scopy(B->d_rtab32, rt32.data(), rt32.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }

    Geo G;
    basis_geo(B, G);
    const uint64_t C = B->C;
    const uint64_t ngroups = (C + kDirSpan - 1) / kDirSpan;
    const uint64_t nwords = ngroups * 2;
    B->ngroups = ngroups;
    uint32_t* words = nullptr;
    uint64_t *cnt = nullptr, *base = nullptr;
    void* tmp = nullptr;
    // the build's temporaries are freed on every exit path
    struct Temps {
        void** p[4];
        void release() { for (void** q : p) if (*q) { dfree(*q); *q = nullptr; } }
        ~Temps() { release(); }
    } temps{{(void**)&words, (void**)&cnt, (void**)&base, &tmp}};
    HIPCHK(hipMalloc(&words, nwords * sizeof(uint32_t)));
    HIPCHK(hipMalloc(&cnt, (ngroups + 1) * sizeof(uint64_t)));
    HIPCHK(hipMalloc(&base, (ngroups + 1) * sizeof(uint64_t)));
    HIPCHK(hipMalloc(&B->d_dir, ngroups * sizeof(DirEntry)));
    if (N <= 32)
        hipLaunchKernelGGL(k_classify<uint32_t>, dim3(sblocks(nwords, 256)), dim3(256), 0, setup_stream(), C, N, n, G, B->d_binom, words, nwords);
    else
        hipLaunchKernelGGL(k_classify<uint64_t>, dim3(sblocks(nwords, 256)), dim3(256), 0, setup_stream(), C, N, n, G, B->d_binom, words, nwords);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_group_counts, dim3(sblocks(ngroups, 256)), dim3(256), 0, setup_stream(), words, ngroups, cnt);
    HIPCHK(hipGetLastError());
    HIPCHK(szero(cnt + ngroups, sizeof(uint64_t)));
    // the group count exceeds 2^31 at N=40: the scan takes a 64-bit length
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, base, (size_t)(ngroups + 1), setup_stream()));
    HIPCHK(hipMalloc(&tmp, tmp_bytes + 16));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, base, (size_t)(ngroups + 1), setup_stream()));
    uint64_t D = 0;
    HIPCHK(scopy(&D, base + ngroups, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (D > 0xFFFFFFFFull) {
        return ed_fail(ED_EINVAL, "sector dimension %llu exceeds 2^32", (unsigned long long)D);
    }
    B->D = (int64_t)D;
    hipLaunchKernelGGL(k_dir_fill, dim3(sblocks(ngroups, 256)), dim3(256), 0, setup_stream(), words, base, ngroups, B->d_dir);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMalloc(&B->d_rows, std::max<uint64_t>(D, 1) * sizeof(uint64_t)));
    if (N <= 32)
        hipLaunchKernelGGL(k_compact<uint32_t>, dim3(sblocks(ngroups, 256)), dim3(256), 0, setup_stream(), C, N, n, G, B->d_binom, B->d_dir, ngroups, B->d_rows);
    else
        hipLaunchKernelGGL(k_compact<uint64_t>, dim3(sblocks(ngroups, 256)), dim3(256), 0, setup_stream(), C, N, n, G, B->d_binom, B->d_dir, ngroups, B->d_rows);
    HIPCHK(hipGetLastError());
    HIPCHK(ssync());
    temps.release();   // before the image layout: at N=40 the temporaries hold ~50 GB
    return basis_image_build(B);
}

// Gather image of the sector (ImageLayout): the smallest chunk size whose
// map fits in LDS and whose image stays within the size bounds.  Sectors of
// shapes without compiled kernels, or whose image would be too large, get
// none and run the directory kernels.
// Chooses the chunk size of a gather image: the smallest whose map fits
// max_entries and whose image stays within max_elems (up to three doublings
// tried), chunks marked from the directory on the device.
template <typename M>
static int image_choose(ed_basis_s* B, uint64_t max_entries, uint64_t max_elems, ImageLayoutT<M>& L) {
    const uint64_t C = B->C;
    const int cb0 = image_min_bits(C, max_entries);
    struct Flags {
        uint8_t* p = nullptr;
        ~Flags() { if (p) dfree(p); }
    } d_t;
    HIPCHK(hipMalloc(&d_t.p, ((C - 1) >> cb0) + 1));
    for (int cb = cb0; cb <= cb0 + 3 && L.cb < 0; cb++) {
        const uint64_t nmap = ((C - 1) >> cb) + 1;
        std::vector<uint8_t> t(nmap);
        HIPCHK(szero(d_t.p, nmap));
        hipLaunchKernelGGL(k_image_touch, dim3(sblocks(B->ngroups, 256)), dim3(256), 0, setup_stream(), B->d_dir, B->ngroups,
                           cb - 6, d_t.p);
        HIPCHK(hipGetLastError());
        HIPCHK(scopy(t.data(), d_t.p, nmap, hipMemcpyDeviceToHost));
        image_layout(C, cb, t, max_elems, L);
    }
    return ED_OK;
}

int basis_image_build(ed_basis_s* B) {
    B->img_cb = -1;
    B->wimg_cb = -1;
    if (B->D == 0) return ED_OK;
    if (B->rows2_w && B->N <= 40) {
        WideImageLayout L;
        int rc = image_choose(B, kWideMapMax, wide_image_max_elems(B->D), L);
        if (rc != ED_OK) return rc;
        if (L.cb >= 0) {
            HIPCHK(hipMalloc(&B->d_wimg_map, L.map.size() * sizeof(uint32_t)));
            HIPCHK(scopy(B->d_wimg_map, L.map.data(), L.map.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            B->wimg_cb = L.cb;
            B->wimg_nslot = L.nslot;
            B->wimg_nmap = L.map.size();
        }
    }
    if (B->N > 32 || !B->has_torus || !torus_shape_compiled(B->torus)) return ED_OK;
    ImageLayout L;
    int rc = image_choose(B, kImageMapMax, image_max_elems(B->D), L);
    if (rc != ED_OK) return rc;
    if (L.cb < 0) return ED_OK;
    HIPCHK(hipMalloc(&B->d_img_map, L.map.size() * sizeof(uint16_t)));
    HIPCHK(scopy(B->d_img_map, L.map.data(), L.map.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&B->d_img_pos, (size_t)B->D * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_image_pos, dim3(sblocks((uint64_t)B->D, 256)), dim3(256), 0, setup_stream(), (uint64_t)B->D, B->d_rows,
                       B->rinfo, B->d_rtab, B->d_img_map, L.cb, B->d_img_pos);
    HIPCHK(hipGetLastError());
    HIPCHK(ssync());
    B->img_cb = L.cb;
    B->img_nslot = (uint32_t)L.nslot;
    B->img_nmap = (uint32_t)L.map.size();
    return ED_OK;
}

void delete_basis(ed_basis_s* B) {
    if (!B) return;
    for (void* p : {(void*)B->d_rows, (void*)B->d_dir, (void*)B->d_rtab, (void*)B->d_binom, (void*)B->d_phase,
                    (void*)B->d_ntab, (void*)B->d_nitab, (void*)B->d_rtab32, (void*)B->d_rlo, (void*)B->d_rhi,
                    (void*)B->d_img_map, (void*)B->d_img_pos, (void*)B->d_wimg_map})
        if (p) dfree(p);
    delete B;
}

extern "C" void ed_basis_destroy(ed_basis b) { delete_basis(b); }

extern "C" int ed_basis_dim(ed_basis b, int64_t* D) {
    if (!b || !D) return ed_fail(ED_EINVAL, "null argument");
    *D = b->D;
    return ED_OK;
}

extern "C" int ed_basis_export(ed_basis b, int64_t* reps, double* norm, double* norm_inv) {
    if (!b) return ed_fail(ED_EINVAL, "null basis");
    const int64_t D = b->D;
    if (D == 0) return ED_OK;
    if (reps) {
        // the row words land in the caller's array and are masked in place
        // (no host staging copy: at 40 sites the array is 27.6 GB)
        HIPCHK(scopy(reps, b->d_rows, D * sizeof(uint64_t), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < D; i++) reps[i] = (int64_t)((uint64_t)reps[i] & kStateMask);
    }
    if (norm || norm_inv) {
        double *dn = nullptr, *dni = nullptr;
        auto body = [&]() -> int {
            HIPCHK(hipMalloc(&dn, D * sizeof(double)));
            HIPCHK(hipMalloc(&dni, D * sizeof(double)));
            hipLaunchKernelGGL(k_norms, dim3(nblocks(D, 256)), dim3(256), 0, setup_stream(), b->d_rows, D, b->d_ntab, b->d_nitab,
                               dn, dni);
            HIPCHK(hipGetLastError());
            if (norm) HIPCHK(scopy(norm, dn, D * sizeof(double), hipMemcpyDeviceToHost));
            if (norm_inv) HIPCHK(scopy(norm_inv, dni, D * sizeof(double), hipMemcpyDeviceToHost));
            return ED_OK;
        };
        const int rc = body();
        dfree(dn);
        dfree(dni);
        if (rc != ED_OK) return rc;
    }
    return ED_OK;
}

extern "C" int ed_basis_full_size(ed_basis b, int64_t* C, int64_t* nlin) {
    if (!b) return ed_fail(ED_EINVAL, "null basis");
    if (C) *C = (int64_t)b->C;
    if (nlin) *nlin = (int64_t)1 << (b->N / 2 + b->N % 2);
    return ED_OK;
}

extern "C" int ed_basis_export_full(ed_basis b, int64_t* rep_loc, double* state_ph, int64_t* Ia, int64_t* Ib,
                                    int32_t* bits_right) {
    if (!b) return ed_fail(ED_EINVAL, "null basis");
    const int N = b->N, n = b->n;
    const int half = N / 2 + N % 2;
    const uint64_t C = b->C;
    if (rep_loc || state_ph) {
        int64_t* d_loc = nullptr;
        cd2* d_ph = nullptr;
        HIPCHK(hipMalloc(&d_loc, std::max<uint64_t>(C, 1) * sizeof(int64_t)));
        HIPCHK(hipMalloc(&d_ph, std::max<uint64_t>(C, 1) * sizeof(cd2)));
        Geo G;
        basis_geo(b, G);
        TablesView tv{b->d_rtab, b->d_dir, b->d_phase};
        uint64_t nth = (C + 31) / 32;
        if (N <= 32)
            hipLaunchKernelGGL(k_full_tables<uint32_t>, dim3(nblocks(nth, 256)), dim3(256), 0, setup_stream(), C, N, n, G, b->d_binom, b->rinfo, tv, d_loc, d_ph);
        else
            hipLaunchKernelGGL(k_full_tables<uint64_t>, dim3(nblocks(nth, 256)), dim3(256), 0, setup_stream(), C, N, n, G, b->d_binom, b->rinfo, tv, d_loc, d_ph);
        HIPCHK(hipGetLastError());
        if (rep_loc) HIPCHK(scopy(rep_loc, d_loc, C * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (state_ph) HIPCHK(scopy(state_ph, d_ph, C * sizeof(cd2), hipMemcpyDeviceToHost));
        dfree(d_loc); dfree(d_ph);
    }
    // Lin tables (Lin_Tables.cpp:57-78) and their Ib completion
    // (Momentum_Basis.cpp:731-737): Ib[left] = rank(s) - Ia[right] + 1 for any
    // weight-n state s with that left part, -1 for left parts that never occur.
    const int64_t nl = (int64_t)1 << half;
    const int64_t br = nl - 1;
    if (bits_right) *bits_right = (int32_t)br;
    if (Ia || Ib) {
        std::vector<int64_t> ia(nl), ib(nl, -1);
        for (int64_t r = 0; r < nl; r++) {
            int w = __builtin_popcountll((unsigned long long)r);
            uint64_t rk = 0;
            int t = 0;
            for (int p = 0; p < half; p++)
                if ((r >> p) & 1) { t++; rk += binom_u64(p, t); }
            (void)w;
            ia[r] = (int64_t)rk + 1;
        }
        const int fl = N / 2;
        for (int64_t left = 0; left < nl && left < ((int64_t)1 << (N - fl)); left++) {
            int rest = n - __builtin_popcountll((unsigned long long)left);
            if (rest < 0 || rest > fl) continue;
            uint64_t s = ((uint64_t)left << fl) | (((uint64_t)1 << rest) - 1);
            uint64_t rk = 0;
            int t = 0;
            for (int p = 0; p < N; p++)
                if ((s >> p) & 1) { t++; rk += binom_u64(p, t); }
            ib[left] = (int64_t)rk - ia[s & br] + 1;
        }
        if (Ia) std::memcpy(Ia, ia.data(), nl * sizeof(int64_t));
        if (Ib) std::memcpy(Ib, ib.data(), nl * sizeof(int64_t));
    }
    return ED_OK;
}

// -------------------------------------------------------------- operator --
static int op_setup(ed_op_s* op, const std::vector<Bond>& bonds, const std::vector<Chiral>& ch,
                    const std::vector<Bowtie>& bw);

extern "C" int ed_op_create(ed_basis b, const ed_terms* t, ed_op* out) {
    if (!b || !t || !out) return ed_fail(ED_EINVAL, "null argument");
    std::vector<Bond> bonds;
    for (int k = 0; k < t->nbonds; k++) {
        int x = t->bond_xy[2 * k], y = t->bond_xy[2 * k + 1];
        if (x < 0 || y < 0 || x >= b->N || y >= b->N || x == y) return ed_fail(ED_EINVAL, "bond %d: bad sites (%d,%d)", k, x, y);
        if (std::fabs(t->bond_J[k]) < 1e-16) continue;   // H_v.cpp:232-233
        bonds.push_back(Bond{x, y, t->bond_J[k], t->bond_phase[2 * k], t->bond_phase[2 * k + 1]});
    }
    std::vector<Chiral> ch;
    if (std::fabs(t->Ch) > 1e-16)   // H_v.cpp:263
        for (int k = 0; k < t->nchiral; k++) {
            int i = t->chiral_ijk[3 * k], j = t->chiral_ijk[3 * k + 1], kk = t->chiral_ijk[3 * k + 2];
            if (i < 0 || j < 0 || kk < 0 || i >= b->N || j >= b->N || kk >= b->N || i == j)
                return ed_fail(ED_EINVAL, "chiral entry %d: bad sites", k);
            ch.push_back(Chiral{i, j, kk, 0, t->chiral_phase[2 * k], t->chiral_phase[2 * k + 1]});
        }
    std::vector<Bowtie> bw;
    if (std::fabs(t->ChCh) > 1e-16)   // H_v.cpp:305
        for (int k = 0; k < t->nbowtie; k++) {
            const int32_t* q = t->bowtie + 6 * k;
            for (int m = 0; m < 6; m++)
                if (q[m] < 0 || q[m] >= b->N) return ed_fail(ED_EINVAL, "bowtie %d: bad site", k);
            if (q[0] == q[1] || q[3] == q[4]) return ed_fail(ED_EINVAL, "bowtie %d: degenerate flip", k);
            bw.push_back(Bowtie{q[0], q[1], q[2], q[3], q[4], q[5]});
        }
    auto* op = new ed_op_s();
    op->b = b;
    op->lam = t->lambda; op->Ch = t->Ch; op->ChCh = t->ChCh;
    op->nb = (int)bonds.size(); op->nc = (int)ch.size(); op->nw = (int)bw.size();
    const int rc = op_setup(op, bonds, ch, bw);
    if (rc != ED_OK) {
        ed_op_destroy(op);   // frees whatever the setup allocated
        return rc;
    }
    *out = op;
    return ED_OK;
}

// Device copies of the terms and the tables of the kernel paths the
// operator can take.  On failure the caller destroys the operator.
static int op_setup(ed_op_s* op, const std::vector<Bond>& bonds, const std::vector<Chiral>& ch,
                    const std::vector<Bowtie>& bw) {
    ed_basis_s* b = op->b;
    if (!bonds.empty()) {
        HIPCHK(hipMalloc(&op->d_b, bonds.size() * sizeof(Bond)));
        HIPCHK(scopy(op->d_b, bonds.data(), bonds.size() * sizeof(Bond), hipMemcpyHostToDevice));
    }
    if (!ch.empty()) {
        HIPCHK(hipMalloc(&op->d_c, ch.size() * sizeof(Chiral)));
        HIPCHK(scopy(op->d_c, ch.data(), ch.size() * sizeof(Chiral), hipMemcpyHostToDevice));
    }
    if (!bw.empty()) {
        HIPCHK(hipMalloc(&op->d_w, bw.size() * sizeof(Bowtie)));
        HIPCHK(scopy(op->d_w, bw.data(), bw.size() * sizeof(Bowtie), hipMemcpyHostToDevice));
    }
    op->bonds = bonds;
    // rotation fast path: rotational lattice, no bowtie terms, families fit
    if (b->has_torus && bw.empty() && b->N <= 32 && b->d_rtab32) {
        std::vector<TFam> fam;
        std::vector<TEntry> ent;
        double ezz0 = 0;
        if (torus_families(op->lam, op->Ch, b->torus, bonds, ch, fam, ent, ezz0) && !fam.empty()) {
            op->use_torus = true;
            op->torus = b->torus;
            op->nfam = (int)fam.size();
            op->ezz0 = ezz0;
            bool bond_only = true, real = true;
            for (const TFam& F : fam) bond_only = bond_only && F.type == 0;
            for (const TEntry& E : ent) real = real && E.cim == 0.0;
            op->tmode = !bond_only ? kModeGeneral : (real && b->trivial_k ? kModeReal : kModeBonds);
            std::vector<uint32_t> fmk;
            std::vector<cd2> fco;
            if (torus_fam_extras(fam, ent, fmk, fco) && op->tmode == kModeReal) op->tmode = kModeUniform;
            op->tune_block = 512;
            HIPCHK(hipMalloc(&op->d_tfam, fam.size() * sizeof(TFam)));
            HIPCHK(scopy(op->d_tfam, fam.data(), fam.size() * sizeof(TFam), hipMemcpyHostToDevice));
            HIPCHK(hipMalloc(&op->d_tent, ent.size() * sizeof(TEntry)));
            HIPCHK(scopy(op->d_tent, ent.data(), ent.size() * sizeof(TEntry), hipMemcpyHostToDevice));
            HIPCHK(hipMalloc(&op->d_tfmk, fmk.size() * sizeof(uint32_t)));
            HIPCHK(scopy(op->d_tfmk, fmk.data(), fmk.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            HIPCHK(hipMalloc(&op->d_tfco, fco.size() * sizeof(cd2)));
            HIPCHK(scopy(op->d_tfco, fco.data(), fco.size() * sizeof(cd2), hipMemcpyHostToDevice));
            // chiral terms on a built-in shape: merge each pair's terms into one element
            std::vector<TFam> mfam;
            std::vector<TEntry> ment;
            std::vector<TChir> chx;
            double mezz0 = 0;
            if (!ch.empty() && torus_shape_compiled(b->torus) &&
                torus_families_merged(op->lam, op->Ch, b->torus, bonds, ch, mfam, ment, chx, mezz0)) {
                op->tmode = kModeMerged;
                op->nmfam = (int)mfam.size();
                HIPCHK(hipMalloc(&op->d_mfam, mfam.size() * sizeof(TFam)));
                HIPCHK(scopy(op->d_mfam, mfam.data(), mfam.size() * sizeof(TFam), hipMemcpyHostToDevice));
                HIPCHK(hipMalloc(&op->d_ment, ment.size() * sizeof(TEntry)));
                HIPCHK(scopy(op->d_ment, ment.data(), ment.size() * sizeof(TEntry), hipMemcpyHostToDevice));
                std::vector<uint32_t> mfmk;
                std::vector<cd2> mfco;
                (void)torus_fam_extras(mfam, ment, mfmk, mfco);
                HIPCHK(hipMalloc(&op->d_mfmk, mfmk.size() * sizeof(uint32_t)));
                HIPCHK(scopy(op->d_mfmk, mfmk.data(), mfmk.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
                std::vector<uint32_t> mki;
                std::vector<cd2> mco;
                torus_merged_tables(mfam, ment, chx, mki, mco);
                HIPCHK(hipMalloc(&op->d_mki, mki.size() * sizeof(uint32_t)));
                HIPCHK(scopy(op->d_mki, mki.data(), mki.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
                HIPCHK(hipMalloc(&op->d_mco, mco.size() * sizeof(cd2)));
                HIPCHK(scopy(op->d_mco, mco.data(), mco.size() * sizeof(cd2), hipMemcpyHostToDevice));
            }
        }
    }
    // two-row 64-bit rotation path: no bowtie terms, compiled shape
    if (b->rows2_w && bw.empty() && b->d_rlo && rows2_shape_compiled(b->N, b->rows2_w))
        return rows2_op_setup(op, bonds, ch);
    return ED_OK;
}

extern "C" int ed_op_tune(ed_op op, int32_t path, int32_t block) {
    if (!op) return ed_fail(ED_EINVAL, "null op");
    if (path < 0 || path > 12) return ed_fail(ED_EINVAL, "path %d not in 0..12", path);
    if ((path == 6 || path == 8) && !op->use_rows2)
        return ed_fail(ED_EINVAL, "two-row 64-bit path not available for this operator");
    if (path == 12) {
        int rc = torus_elist_check(op);
        if (rc != ED_OK) return rc;
    }
    if (((path >= 2 && path <= 5) || path == 7 || path >= 9) && !op->use_torus)
        return ed_fail(ED_EINVAL, "rotation fast path not available for this operator");
    if ((path == 2 || path == 4 || path == 5 || path == 7 || path >= 9) && !torus_shape_compiled(op->torus))
        return ed_fail(ED_EINVAL, "lattice shape is not one of the compiled shapes");
    if (block != 0 && (block < 64 || block > 512 || block % 64)) return ed_fail(ED_EINVAL, "block %d", block);
    op->tune_path = path;
    op->block_fixed = block != 0;
    if (block) op->tune_block = block;
    else if (op->use_torus) op->tune_block = 512;
    op->sel_var[0] = op->sel_var[1] = -1;
    return ED_OK;
}

extern "C" int ed_op_elist_stats(ed_op op, int64_t* elements, int64_t* slots) {
    if (!op || !elements || !slots) return ed_fail(ED_EINVAL, "null argument");
    *elements = op->d_el ? op->nelem : 0;
    *slots = op->d_el ? op->eslots : 0;
    return ED_OK;
}

extern "C" int ed_op_tuning(ed_op op, int32_t dtype, int32_t* path, int32_t* variant, int32_t* block) {
    if (!op || !path || !variant || !block) return ed_fail(ED_EINVAL, "null argument");
    if (dtype != ED_C128 && dtype != ED_C64) return ed_fail(ED_EINVAL, "unknown dtype %d", dtype);
    if (op->use_rows2 && (op->tune_path == 6 || op->tune_path == 8 || (op->tune_path == 0 && !op->use_torus))) {
        *path = op->tune_path == 8 ? 8 : 6;
        *variant = -1;
        *block = 256;
        return ED_OK;
    }
    if (!op->use_torus || op->tune_path == 1) {
        *path = 1;
        *variant = -1;
        *block = 256;
        return ED_OK;
    }
    int p, v, bs;
    torus_config(op, dtype == ED_C128 ? 0 : 1, p, v, bs);
    *path = p;
    *variant = v;
    *block = bs;
    return ED_OK;
}

extern "C" void ed_op_destroy(ed_op op) {
    if (!op) return;
    for (void* p : {(void*)op->d_b, (void*)op->d_c, (void*)op->d_w, op->d_u64, op->d_u128, (void*)op->d_tfam,
                    (void*)op->d_tent, op->d_fam64, (void*)op->d_foff64, (void*)op->d_ent64, (void*)op->d_mfam,
                    (void*)op->d_ment, (void*)op->d_tfmk, (void*)op->d_tfco,
                    (void*)op->d_mfmk, (void*)op->d_mki, (void*)op->d_mco, op->d_img64, op->d_img128, op->d_wimg64,
                    op->d_wimg128, (void*)op->d_ebase, (void*)op->d_erl, (void*)op->d_el, (void*)op->d_ediag,
                    (void*)op->d_ectab})
        if (p) dfree(p);
    delete op;
}

TermsView op_terms(const ed_op_s* op) {
    return TermsView{op->lam, op->Ch, op->ChCh, op->nb, op->nc, op->nw, op->d_b, op->d_c, op->d_w};
}

// ext (with ext_bytes) replaces the operator's gather buffers: the gather
// stream is built there by every call (zero fill where the layout has holes,
// then the prescale), nothing is allocated and the kernel variant is not
// timed (the tuned one, else the default).
template <typename VT>
static int launch_hv(ed_op_s* op, const VT* v, VT* w, hipStream_t st, hipEvent_t ev0 = nullptr,
                     hipEvent_t ev1 = nullptr, VT* ext = nullptr, size_t ext_bytes = 0) {
    ed_basis_s* B = op->b;
    const int64_t D = B->D;
    if (D == 0) return ED_OK;
    const bool rows2 =
        op->use_rows2 && (op->tune_path == 6 || op->tune_path == 8 || (op->tune_path == 0 && !op->use_torus));
    const bool img = !rows2 && torus_uses_image(op);
    // a gather buffer is allocated by the first H*v that needs it, which
    // therefore cannot be part of a stream capture
    auto may_alloc = [&]() -> int {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return ed_fail(ED_EINVAL, "the first H*v of a precision and kernel path allocates its gather buffer: "
                                      "run it before capturing");
        return ED_OK;
    };
    VT* u = nullptr;
    if (img && ext) {
        const uint64_t ne = (uint64_t)(B->img_nslot + 1) << B->img_cb;
        if (ne * sizeof(VT) > ext_bytes)
            return ed_fail(ED_ENOMEM, "gather image of %llu elements (%llu bytes) exceeds the %llu-byte scratch",
                           (unsigned long long)ne, (unsigned long long)(ne * sizeof(VT)),
                           (unsigned long long)ext_bytes);
        hipLaunchKernelGGL(k_zero<VT>, dim3(sblocks(ne, 256)), dim3(256), 0, st, ne, ext);
        hipLaunchKernelGGL(k_prescale_img<VT>, dim3(nblocks(D, 256)), dim3(256), 0, st, D, B->d_rows, B->d_nitab,
                           B->d_img_pos, v, ext);
        u = ext;
    } else if (img) {
        // the compiled rotation kernels gather from the image; its zero
        // positions are written once, at allocation
        VT*& im = std::is_same<VT, cd2>::value ? (VT*&)op->d_img128 : (VT*&)op->d_img64;
        const uint64_t ne = (uint64_t)(B->img_nslot + 1) << B->img_cb;
        if (!im) {
            int rc = may_alloc();
            if (rc != ED_OK) return rc;
            HIPCHK(hipMalloc((void**)&im, ne * sizeof(VT)));
            hipLaunchKernelGGL(k_zero<VT>, dim3(sblocks(ne, 256)), dim3(256), 0, st, ne, im);
        }
        hipLaunchKernelGGL(k_prescale_img<VT>, dim3(nblocks(D, 256)), dim3(256), 0, st, D, B->d_rows, B->d_nitab,
                           B->d_img_pos, v, im);
        u = im;
    }
    bool wimg = false;
    bool& wimg_off = op->wimg_off[std::is_same<VT, cd2>::value ? 0 : 1];
    if (ext && rows2 && B->wimg_cb >= 0 && op->tune_path != 8 &&
        ((B->wimg_nslot + 1) << B->wimg_cb) * sizeof(VT) <= ext_bytes) {
        // the wide image when it fits the scratch, else u by row below
        // (both layouts give the same sums)
        const uint64_t ne = (B->wimg_nslot + 1) << B->wimg_cb;
        hipLaunchKernelGGL(k_zero<VT>, dim3(sblocks(ne, 256)), dim3(256), 0, st, ne, ext);
        hipLaunchKernelGGL(k_prescale_wimg<VT>, dim3(sblocks((uint64_t)D, 256)), dim3(256), 0, st, (uint64_t)D,
                           B->d_rows, B->d_nitab, B->rinfo, B->d_rtab, B->d_wimg_map, B->wimg_cb, v, ext);
        u = ext;
        wimg = true;
    } else if (!ext && rows2 && B->wimg_cb >= 0 && !wimg_off && op->tune_path != 8) {
        // the two-row kernels gather from the wide image when it fits the
        // free device memory, else (for the operator's lifetime) from u by
        // row through the directory
        VT*& im = std::is_same<VT, cd2>::value ? (VT*&)op->d_wimg128 : (VT*&)op->d_wimg64;
        const uint64_t ne = (B->wimg_nslot + 1) << B->wimg_cb;
        if (!im) {
            int rc = may_alloc();
            if (rc != ED_OK) return rc;
            if (hipMalloc((void**)&im, ne * sizeof(VT)) != hipSuccess) {
                im = nullptr;
                wimg_off = true;
                (void)hipGetLastError();
            } else {
                hipLaunchKernelGGL(k_zero<VT>, dim3(sblocks(ne, 256)), dim3(256), 0, st, ne, im);
            }
        }
        if (im) {
            hipLaunchKernelGGL(k_prescale_wimg<VT>, dim3(sblocks((uint64_t)D, 256)), dim3(256), 0, st, (uint64_t)D,
                               B->d_rows, B->d_nitab, B->rinfo, B->d_rtab, B->d_wimg_map, B->wimg_cb, v, im);
            u = im;
            wimg = true;
        }
    }
    if (!img && !wimg && ext) {
        if ((uint64_t)D * sizeof(VT) > ext_bytes)
            return ed_fail(ED_ENOMEM, "gather stream of %lld elements exceeds the %llu-byte scratch", (long long)D,
                           (unsigned long long)ext_bytes);
        hipLaunchKernelGGL(k_prescale<VT>, dim3(nblocks(D, 256)), dim3(256), 0, st, D, B->d_rows, B->d_nitab, v, ext);
        u = ext;
    } else if (!img && !wimg) {
        VT*& uu = std::is_same<VT, cd2>::value ? (VT*&)op->d_u128 : (VT*&)op->d_u64;
        if (!uu) {
            int rc = may_alloc();
            if (rc != ED_OK) return rc;
            HIPCHK(hipMalloc((void**)&uu, D * sizeof(VT)));
        }
        hipLaunchKernelGGL(k_prescale<VT>, dim3(nblocks(D, 256)), dim3(256), 0, st, D, B->d_rows, B->d_nitab, v, uu);
        u = uu;
    }
    HIPCHK(hipGetLastError());
    Geo G;
    basis_geo(B, G);
    TablesView tv{B->d_rtab, B->d_dir, B->d_phase};
    TermsView T = op_terms(op);
    dim3 grid(nblocks(D, 256)), blk(256);
    if (!ext && op->use_torus && op->tune_path != 1 && op->tune_path != 6 && op->tune_path != 8) {
        int rc = torus_prepare<VT>(op, u, v, w, st);
        if (rc != ED_OK) return rc;
    }
    if (op->tune_path == 12 && !op->d_el) {
        // the stored element list is built by the first H*v on path 12
        int rc = ext ? ed_fail(ED_EINVAL, "the stored element list is built by a first ed_hv on path 12") : may_alloc();
        if (rc == ED_OK) rc = torus_elist_prepare(op, st);
        if (rc != ED_OK) return rc;
    }
    if (ev0) HIPCHK(hipEventRecord(ev0, st));
    if (rows2) {
        int rc = launch_hv_rows2<VT>(op, u, v, w, st, wimg);
        if (rc != ED_OK) return rc;
    } else if (op->use_torus && op->tune_path != 1) {
        int rc = launch_hv_torus<VT>(op, u, v, w, st);
        if (rc != ED_OK) return rc;
    } else if (B->N <= 32) {
        if (B->trivial_k)
            hipLaunchKernelGGL((k_hv_rows<uint32_t, true, VT>), grid, blk, 0, st, D, B->d_rows, G, B->rinfo, tv, T, B->d_ntab, u, v, w);
        else
            hipLaunchKernelGGL((k_hv_rows<uint32_t, false, VT>), grid, blk, 0, st, D, B->d_rows, G, B->rinfo, tv, T, B->d_ntab, u, v, w);
    } else {
        if (B->trivial_k)
            hipLaunchKernelGGL((k_hv_rows<uint64_t, true, VT>), grid, blk, 0, st, D, B->d_rows, G, B->rinfo, tv, T, B->d_ntab, u, v, w);
        else
            hipLaunchKernelGGL((k_hv_rows<uint64_t, false, VT>), grid, blk, 0, st, D, B->d_rows, G, B->rinfo, tv, T, B->d_ntab, u, v, w);
    }
    HIPCHK(hipGetLastError());
    if (ev1) HIPCHK(hipEventRecord(ev1, st));
    return ED_OK;
}

int op_hv(ed_op_s* op, int dtype, const void* v, void* w, hipStream_t st) {
    if (dtype == ED_C128) return launch_hv<cd2>(op, (const cd2*)v, (cd2*)w, st);
    if (dtype == ED_C64) return launch_hv<cf2>(op, (const cf2*)v, (cf2*)w, st);
    return ed_fail(ED_EINVAL, "unknown dtype %d", dtype);
}

extern "C" int ed_hv(ed_op op, int32_t dtype, const void* v, void* w, void* stream) {
    if (!op || (!v && op->b->D) || (!w && op->b->D)) return ed_fail(ED_EINVAL, "null argument");
    if (v == w && op->b->D) return ed_fail(ED_EINVAL, "v and w must not alias");
    return op_hv(op, dtype, v, w, (hipStream_t)stream);
}

extern "C" int ed_hv_noalloc(ed_op op, int32_t dtype, const void* v, void* w, void* stream) {
    if (!op || (!v && op->b->D) || (!w && op->b->D)) return ed_fail(ED_EINVAL, "null argument");
    if (v == w && op->b->D) return ed_fail(ED_EINVAL, "v and w must not alias");
    if (dtype != ED_C128 && dtype != ED_C64) return ed_fail(ED_EINVAL, "unknown dtype %d", dtype);
    if (op->b->D == 0) return ED_OK;
    void* ptr = nullptr;
    int rc = scratch_ptr(&ptr);
    if (rc != ED_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == ED_C128)
        return launch_hv<cd2>(op, (const cd2*)v, (cd2*)w, st, nullptr, nullptr, (cd2*)ptr, kNoAllocScratchBytes);
    return launch_hv<cf2>(op, (const cf2*)v, (cf2*)w, st, nullptr, nullptr, (cf2*)ptr, kNoAllocScratchBytes);
}

extern "C" int ed_hv_profile(ed_op op, int32_t dtype, const void* v, void* w, void* stream, float* kernel_ms) {
    if (!op || !kernel_ms) return ed_fail(ED_EINVAL, "null argument");
    if (v == w && op->b->D) return ed_fail(ED_EINVAL, "v and w must not alias");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    *kernel_ms = 0.f;
    // every exit after the events exist destroys them
    auto body = [&]() -> int {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        int rc;
        if (dtype == ED_C128) rc = launch_hv<cd2>(op, (const cd2*)v, (cd2*)w, st, e0, e1);
        else if (dtype == ED_C64) rc = launch_hv<cf2>(op, (const cf2*)v, (cf2*)w, st, e0, e1);
        else rc = ed_fail(ED_EINVAL, "unknown dtype %d", dtype);
        if (rc == ED_OK && op->b->D) {
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(kernel_ms, e0, e1));
        }
        return rc;
    };
    const int rc = body();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

extern "C" int ed_hv_host(ed_op op, const double* v, double* w) {
    if (!op) return ed_fail(ED_EINVAL, "null op");
    const int64_t D = op->b->D;
    if (D == 0) return ED_OK;
    if (!v || !w) return ed_fail(ED_EINVAL, "null argument");
    if (v == w) return ed_fail(ED_EINVAL, "v and w must not alias");
    void *dv = nullptr, *dw = nullptr;
    // every exit after the first allocation frees both buffers
    auto body = [&]() -> int {
        HIPCHK(hipMalloc(&dv, D * sizeof(cd2)));
        HIPCHK(hipMalloc(&dw, D * sizeof(cd2)));
        HIPCHK(scopy(dv, v, D * sizeof(cd2), hipMemcpyHostToDevice));
        int rc = op_hv(op, ED_C128, dv, dw, setup_stream());
        if (rc != ED_OK) return rc;
        HIPCHK(ssync());
        HIPCHK(scopy(w, dw, D * sizeof(cd2), hipMemcpyDeviceToHost));
        return ED_OK;
    };
    const int rc = body();
    dfree(dv);
    dfree(dw);
    return rc;
}

// w = A v for a CSR matrix, one lane per row, FP64 accumulation in the
// order of the row's entries (H_v.cpp:3-19).
__global__ void k_csr_hv(int64_t n, const int64_t* __restrict__ rowptr, const int64_t* __restrict__ col,
                         const cd2* __restrict__ val, const cd2* __restrict__ v, cd2* __restrict__ w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ar = 0.0, ai = 0.0;
    for (int64_t q = rowptr[i]; q < rowptr[i + 1]; q++) {
        const cd2 a = val[q], x = v[col[q]];
        ar += a.re * x.re - a.im * x.im;
        ai += a.re * x.im + a.im * x.re;
    }
    w[i] = cd2{ar, ai};
}

extern "C" int ed_csr_hv_host(int64_t n, int64_t ncols, const int64_t* rowptr, const int64_t* col, const double* val,
                              const double* v, double* w) {
    if (n < 0 || ncols < 0 || (n > 0 && (!rowptr || !v || !w))) return ed_fail(ED_EINVAL, "null argument");
    if (n == 0) return ED_OK;
    const int64_t nnz = rowptr[n];
    if (rowptr[0] != 0 || nnz < 0) return ed_fail(ED_EINVAL, "row pointers must start at 0");
    for (int64_t i = 0; i < n; i++)
        if (rowptr[i + 1] < rowptr[i]) return ed_fail(ED_EINVAL, "row pointers decrease at row %lld", (long long)i);
    for (int64_t q = 0; q < nnz; q++)
        if (col[q] < 0 || col[q] >= ncols) return ed_fail(ED_EINVAL, "column %lld out of range", (long long)col[q]);
    int64_t *drp = nullptr, *dcol = nullptr;
    cd2 *dval = nullptr, *dv = nullptr, *dw = nullptr;
    int rc = ED_OK;
    auto fail = [&](hipError_t e) {
        rc = ed_fail(e == hipErrorOutOfMemory ? ED_ENOMEM : ED_EDEVICE, "ed_csr_hv_host: %s", hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&drp, (n + 1) * sizeof(int64_t))) != hipSuccess ||
        (e = hipMalloc(&dcol, (nnz > 0 ? nnz : 1) * sizeof(int64_t))) != hipSuccess ||
        (e = hipMalloc(&dval, (nnz > 0 ? nnz : 1) * sizeof(cd2))) != hipSuccess ||
        (e = hipMalloc(&dv, (ncols > 0 ? ncols : 1) * sizeof(cd2))) != hipSuccess ||
        (e = hipMalloc(&dw, n * sizeof(cd2))) != hipSuccess ||
        (e = scopy(drp, rowptr, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess ||
        (nnz > 0 && (e = scopy(dcol, col, nnz * sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess) ||
        (nnz > 0 && (e = scopy(dval, val, nnz * sizeof(cd2), hipMemcpyHostToDevice)) != hipSuccess) ||
        (ncols > 0 && (e = scopy(dv, v, ncols * sizeof(cd2), hipMemcpyHostToDevice)) != hipSuccess)) {
        fail(e);
    } else {
        hipLaunchKernelGGL(k_csr_hv, dim3(nblocks(n, 256)), dim3(256), 0, setup_stream(), n, drp, dcol, dval, dv, dw);
        if ((e = hipGetLastError()) != hipSuccess || (e = ssync()) != hipSuccess ||
            (e = scopy(w, dw, n * sizeof(cd2), hipMemcpyDeviceToHost)) != hipSuccess)
            fail(e);
    }
    for (void* p : {(void*)drp, (void*)dcol, (void*)dval, (void*)dv, (void*)dw})
        if (p) dfree(p);
    return rc;
}

extern "C" int ed_op_count_targets(ed_op op, int64_t* distinct, int64_t* total) {
    if (!op) return ed_fail(ED_EINVAL, "null op");
    ed_basis_s* B = op->b;
    const int64_t D = B->D;
    unsigned long long* d = nullptr;
    HIPCHK(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0, 0};
    auto run = [&]() -> int {
        HIPCHK(szero(d, 2 * sizeof(unsigned long long)));
        Geo G;
        basis_geo(B, G);
        TablesView tv{B->d_rtab, B->d_dir, B->d_phase};
        TermsView T = op_terms(op);
        if (D > 0) {
            dim3 grid(nblocks(D, 128)), blk(128);
            if (B->N <= 32)
                hipLaunchKernelGGL((k_count_targets<uint32_t, true>), grid, blk, 0, setup_stream(), D, B->d_rows, G, B->rinfo, tv, T,
                                   d, d + 1);
            else
                hipLaunchKernelGGL((k_count_targets<uint64_t, true>), grid, blk, 0, setup_stream(), D, B->d_rows, G, B->rinfo, tv, T,
                                   d, d + 1);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(scopy(h, d, sizeof(h), hipMemcpyDeviceToHost));
        return ED_OK;
    };
    const int rc = run();
    dfree(d);
    if (rc != ED_OK) return rc;
    if (distinct) *distinct = (int64_t)h[0];
    if (total) *total = (int64_t)h[1];
    return ED_OK;
}

// The sector matrix, column-major D x D complex128, into dH on the setup
// stream (k_make_h, one row per lane).
static int fill_dense_h(ed_op_s* op, cd2* dH) {
    ed_basis_s* B = op->b;
    const int64_t D = B->D;
    HIPCHK(szero(dH, D * D * sizeof(cd2)));
    Geo G;
    basis_geo(B, G);
    TablesView tv{B->d_rtab, B->d_dir, B->d_phase};
    TermsView T = op_terms(op);
    dim3 grid(nblocks(D, 64)), blk(64);
    hipStream_t st = setup_stream();
    if (B->N <= 32) {
        if (B->trivial_k)
            hipLaunchKernelGGL((k_make_h<uint32_t, true>), grid, blk, 0, st, D, B->d_rows, G, B->rinfo, tv, T, B->d_ntab, B->d_nitab, dH);
        else
            hipLaunchKernelGGL((k_make_h<uint32_t, false>), grid, blk, 0, st, D, B->d_rows, G, B->rinfo, tv, T, B->d_ntab, B->d_nitab, dH);
    } else {
        if (B->trivial_k)
            hipLaunchKernelGGL((k_make_h<uint64_t, true>), grid, blk, 0, st, D, B->d_rows, G, B->rinfo, tv, T, B->d_ntab, B->d_nitab, dH);
        else
            hipLaunchKernelGGL((k_make_h<uint64_t, false>), grid, blk, 0, st, D, B->d_rows, G, B->rinfo, tv, T, B->d_ntab, B->d_nitab, dH);
    }
    HIPCHK(hipGetLastError());
    return ED_OK;
}

extern "C" int ed_make_h(ed_op op, double* H) {
    if (!op || !H) return ed_fail(ED_EINVAL, "null argument");
    const int64_t D = op->b->D;
    if (D > 20000) return ed_fail(ED_EINVAL, "dense H requested for D=%lld > 20000", (long long)D);
    if (D == 0) return ED_OK;
    cd2* dH = nullptr;
    // every exit after the allocation frees it
    auto body = [&]() -> int {
        HIPCHK(hipMalloc(&dH, D * D * sizeof(cd2)));
        const int frc = fill_dense_h(op, dH);
        if (frc != ED_OK) return frc;
        HIPCHK(scopy(H, dH, D * D * sizeof(cd2), hipMemcpyDeviceToHost));
        return ED_OK;
    };
    const int rc = body();
    dfree(dH);
    return rc;
}

// Dense Hermitian eigensolve of the sector matrix on the GPU (rocSOLVER
// zheevd): the small-sector path of exact_evecs (lanczos.cpp:81-131), which
// the reference runs through LAPACK zheev (Matrix_Functions.cpp:28-48).
extern "C" int ed_dense_eigh(ed_op op, double* evals, double* evecs) {
    if (!op || !evals) return ed_fail(ED_EINVAL, "null argument");
    const int64_t D = op->b->D;
    if (D == 0) return ed_fail(ED_EEMPTY, "No states in psector");
    if (D > 20000) return ed_fail(ED_EINVAL, "dense eigensolve requested for D=%lld > 20000", (long long)D);
    cd2* dH = nullptr;
    double *dW = nullptr, *dE = nullptr;
    int* dinfo = nullptr;
    rocblas_handle h = nullptr;
    // every exit frees the buffers and the handle
    auto body = [&]() -> int {
        HIPCHK(hipMalloc(&dH, D * D * sizeof(cd2)));
        HIPCHK(hipMalloc(&dW, D * sizeof(double)));
        HIPCHK(hipMalloc(&dE, D * sizeof(double)));
        HIPCHK(hipMalloc(&dinfo, sizeof(int)));
        const int frc = fill_dense_h(op, dH);
        if (frc != ED_OK) return frc;
        if (rocblas_create_handle(&h) != rocblas_status_success) {
            h = nullptr;
            return ed_fail(ED_EDEVICE, "rocblas_create_handle failed");
        }
        // the solver follows the matrix kernels on the setup stream
        if (rocblas_set_stream(h, setup_stream()) != rocblas_status_success)
            return ed_fail(ED_EDEVICE, "rocblas_set_stream failed");
        // column-major H, upper triangle referenced (the matrix is Hermitian)
        const rocblas_status rs = rocsolver_zheevd(h, evecs ? rocblas_evect_original : rocblas_evect_none,
                                                   rocblas_fill_upper, (rocblas_int)D, (rocblas_double_complex*)dH,
                                                   (rocblas_int)D, dW, dE, dinfo);
        if (rs != rocblas_status_success) return ed_fail(ED_EDEVICE, "rocsolver_zheevd failed (%d)", (int)rs);
        int info = 0;
        HIPCHK(scopy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost));
        if (info != 0) return ed_fail(ED_ENOTCONV, "zheevd info=%d", info);
        HIPCHK(scopy(evals, dW, D * sizeof(double), hipMemcpyDeviceToHost));
        if (evecs) HIPCHK(scopy(evecs, dH, D * D * sizeof(cd2), hipMemcpyDeviceToHost));
        return ED_OK;
    };
    const int rc = body();
    if (h) rocblas_destroy_handle(h);
    dfree(dH); dfree(dW); dfree(dE); dfree(dinfo);
    return rc;
}
