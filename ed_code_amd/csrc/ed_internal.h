// ed_internal.h -- handle structs shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <vector>

#include "ed_amd.h"
#include "ed_core.h"
#include "ed_torus.h"

int ed_fail(int code, const char* fmt, ...);

// Logical block index of a row kernel's workgroup.  Workgroups are dispatched
// round-robin over the 8 XCDs; this mapping hands each XCD a sequence of
// tiles of kXcdTileRows rows (tile t to XCD t mod 8), so that the 8 XCDs
// sweep adjacent tiles at the same time and their gathers meet in the shared
// Infinity Cache, while each XCD's L2 still sees consecutive rows.  The tail
// that does not fill a round of 8 tiles keeps the dispatch order.  A
// bijection of [0, gridDim.x) for any grid.
constexpr unsigned kXcdTileRows = 32768;
__device__ inline unsigned xcd_tile_block_rows(unsigned rows_per_block) {
    const unsigned nb = gridDim.x, xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    if ((rows_per_block & (rows_per_block - 1u)) == 0u) {
        // a power of two (every block size of the tuned kernels): the same
        // mapping with shifts in place of the divisions
        const unsigned lt = rows_per_block >= kXcdTileRows
                                ? 0u
                                : (unsigned)(__builtin_ctz(kXcdTileRows) - __builtin_ctz(rows_per_block));
        const unsigned full = nb >> (3u + lt);
        return blockIdx.x < (full << (3u + lt)) ? ((((slot >> lt) << 3) + xcd) << lt) + (slot & ((1u << lt) - 1u))
                                                 : blockIdx.x;
    }
    const unsigned T = rows_per_block >= kXcdTileRows ? 1u : kXcdTileRows / rows_per_block;
    const unsigned full = nb / (8 * T);
    return blockIdx.x < full * 8 * T ? ((slot / T) * 8 + xcd) * T + (slot % T) : blockIdx.x;
}
// one row per thread
__device__ inline unsigned xcd_tile_block() { return xcd_tile_block_rows(blockDim.x); }

inline void dfree(void* p) {
    if (p) (void)hipFree(p);
}

// The calling host thread's non-blocking stream (one per device) for the
// library's own setup work: basis builds, operator and image tables, host
// copies of tables, dense matrices.  Work on the legacy default stream would
// wait on other streams and is refused while another thread captures a
// graph (concurrent solves in the serving loop); a private non-blocking
// stream keeps the setup of one solve independent of the captures of others.
// None of this work reads device data a caller produced.
hipStream_t setup_stream();
// copy on the setup stream, then wait for it
inline hipError_t scopy(void* dst, const void* src, size_t n, hipMemcpyKind kind) {
    hipStream_t s = setup_stream();
    hipError_t e = hipMemcpyAsync(dst, src, n, kind, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}
// zero n bytes, ordered before later work on the setup stream
inline hipError_t szero(void* p, size_t n) { return hipMemsetAsync(p, 0, n, setup_stream()); }
inline hipError_t ssync() { return hipStreamSynchronize(setup_stream()); }

// LDS one workgroup may allocate on the current device: a whole CU's
// 160 KiB on gfx950; 64 KiB, the size of earlier CDNA parts, if the query fails
inline size_t lds_per_block() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) {
        (void)hipGetLastError();
        return 64 * 1024;
    }
    return (size_t)v;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            return ed_fail(e_ == hipErrorOutOfMemory ? ED_ENOMEM : ED_EDEVICE, "%s: %s (%s:%d)", #x, \
                           hipGetErrorString(e_), __FILE__, __LINE__);                             \
    } while (0)

struct ed_lattice_s {
    int N = 0, L1 = 0, L2 = 0, L3 = 1;
    std::vector<int> T1, T2, T3;     // T3 = identity for two-generator lattices
    edc::TProg P1, P2, P3;
    bool has_torus = false;   // translations are bit rotations (rotation fast path)
    edc::Torus torus;
    int rows2_w = 0;          // > 0: two rows of rows2_w sites (64-bit rotation path)
};

struct ed_basis_s {
    int N = 0, n = 0, kx = 0, ky = 0, kz = 0, L1 = 0, L2 = 0, L3 = 1;
    double px = 0, py = 0, pz = 0;
    uint64_t C = 0;
    int64_t D = 0;
    uint64_t ngroups = 0;
    bool trivial_k = false;
    std::vector<int> T1, T2, T3;
    edc::TProg lat_P1, lat_P2, lat_P3;
    edc::RankInfo rinfo;
    uint64_t* d_rows = nullptr;      // [D] representative | lid << 52
    edc::DirEntry* d_dir = nullptr;  // [ceil(C/64)]
    uint64_t* d_rtab = nullptr;      // rank tables
    uint32_t* d_rtab32 = nullptr;    // same, 32-bit (N <= 32)
    int rtab32_n = 0;
    bool has_torus = false;
    edc::Torus torus;
    int rows2_w = 0;                 // two-row rotation structure (see ed_torus64.h)
    uint32_t* d_rlo = nullptr;       // rank tables without the top-byte block, 32-bit (N <= 40)
    uint64_t* d_rhi = nullptr;       // top-byte block, 64-bit [256]
    int rlo_n = 0;
    uint64_t* d_binom = nullptr;     // C(p,t), 65 x 65
    edc::cd2* d_phase = nullptr;     // [L1*L2]
    double* d_ntab = nullptr;        // norm by lid
    double* d_nitab = nullptr;       // 1/norm by lid
    // gather image of u for the compiled rotation kernels (ImageLayout in
    // ed_tables.h); img_cb = -1 when the sector has none
    int img_cb = -1;
    uint32_t img_nslot = 0, img_nmap = 0;
    uint16_t* d_img_map = nullptr;   // [img_nmap] chunk -> slot
    uint32_t* d_img_pos = nullptr;   // [D] image element of row i
    // wide gather image of the two-row 64-bit kernels (WideImageLayout);
    // wimg_cb = -1 when the sector has none
    int wimg_cb = -1;
    uint64_t wimg_nslot = 0, wimg_nmap = 0;
    uint32_t* d_wimg_map = nullptr;  // [wimg_nmap] chunk -> slot
};

struct ed_op_s {
    ed_basis_s* b = nullptr;
    double lam = 1, Ch = 0, ChCh = 0;
    int nb = 0, nc = 0, nw = 0;
    edc::Bond* d_b = nullptr;
    edc::Chiral* d_c = nullptr;
    edc::Bowtie* d_w = nullptr;
    // rotation fast path
    bool use_torus = false;
    edc::Torus torus;
    int nfam = 0;
    edc::TFam* d_tfam = nullptr;
    edc::TEntry* d_tent = nullptr;
    double ezz0 = 0;
    int tmode = 0;             // term class of the families (kModeGeneral / kModeBonds / kModeReal / kModeMerged /
                               // kModeUniform)
    uint32_t* d_tfmk = nullptr;   // flip mask per family member [nfam * 32]
    edc::cd2* d_tfco = nullptr;   // kModeUniform: {J/2, Delta J/2} per family
    // kModeMerged: the compiled-shape kernels' bond-keyed families with the
    // chiral entries of each pair (d_tfam / d_tent stay for the other kernels)
    int nmfam = 0;
    edc::TFam* d_mfam = nullptr;
    edc::TEntry* d_ment = nullptr;
    uint32_t* d_mfmk = nullptr;   // flip mask per merged family member [nmfam * 32]
    uint32_t* d_mki = nullptr;    // z-sites of each merged member's chiral entries [nmfam * 32]
    edc::cd2* d_mco = nullptr;    // merged member coefficients by z-site bits [nmfam * 32 * 4]
    // two-row 64-bit rotation path (ed_hv64.hip)
    bool use_rows2 = false;
    int nfam64 = 0, nterm64 = 0;
    void* d_fam64 = nullptr;   // TFam64[nfam64]
    int* d_foff64 = nullptr;   // [nfam64]
    edc::TEntry* d_ent64 = nullptr;   // [nterm64], dense term order
    double ezz0_64 = 0;
    int tune_path = 0;        // 0 auto, 1 generic walk, 2/4/5 compiled shape K=4 / K=2 / K=8, 3 run-time shape,
                              // 6 two-row 64-bit kernel, 7 compiled shape K=4 pipelined, 8 as 6 with u by
                              // row through the directory, 9/10 compiled shape with rows split over 2 / 4
                              // lanes, 11 as 9 with K=4, 12 stored element list
    int tune_block = 256;
    bool block_fixed = false; // block set by ed_op_tune
    int sel_var[2] = {-1, -1};   // tuned variant per precision (0 complex128, 1 complex64)
    int sel_block[2] = {0, 0};
    void* d_u64 = nullptr;    // prescaled gather stream, complex64 [D]
    void* d_u128 = nullptr;   // prescaled gather stream, complex128 [D]
    void* d_img64 = nullptr;  // prescaled gather image (compiled rotation kernels), complex64
    void* d_img128 = nullptr; // complex128
    void* d_wimg64 = nullptr; // prescaled wide gather image (two-row 64-bit kernels), complex64
    void* d_wimg128 = nullptr;
    bool wimg_off[2] = {false, false};   // wide image did not fit (per precision: 0 complex128, 1 complex64)
    std::vector<edc::Bond> bonds;
    // stored element list (path 12, ed_hv_fast.hip): rows in groups of 64;
    // element q of row 64 g + l at el[ebase[g] + 64 q + l], packed as the
    // image element | family << ejbits; erl[i] elements in row i
    int64_t* d_ebase = nullptr;   // [ngroups + 1]
    uint16_t* d_erl = nullptr;    // [D]
    uint32_t* d_el = nullptr;     // [ebase[ngroups]]
    double* d_ediag = nullptr;    // [D] diagonal of each row
    double* d_ectab = nullptr;    // [nfam] coefficient of each family
    int ejbits = 0;
    int64_t eslots = 0;           // ebase[ngroups]: stored slots, padding included
    int64_t nelem = 0;            // elements
};

int basis_geo(const ed_basis_s* B, edc::Geo& G);
int basis_build_device(ed_basis_s* B);
void delete_basis(ed_basis_s* B);
edc::TermsView op_terms(const ed_op_s* op);
int op_hv(ed_op_s* op, int dtype, const void* v, void* w, hipStream_t st);
size_t torus_smem_bytes(const ed_op_s* op);
template <typename VT>
int launch_hv_torus(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st);
template <typename VT>
int torus_prepare(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st);
void torus_config(const ed_op_s* op, int dslot, int& path, int& var, int& bs);
int torus_elist_check(const ed_op_s* op);
int torus_elist_prepare(ed_op_s* op, hipStream_t st);
bool torus_shape_compiled(const edc::Torus& T);
bool torus_uses_image(const ed_op_s* op);
int basis_image_build(ed_basis_s* B);
int block_op_hv(ed_block_s* K, int dtype, const void* v, void* w, hipStream_t st);
int rows2_detect(const std::vector<int>& T1, const std::vector<int>& T2);
int rows2_tables(ed_basis_s* B, const std::vector<uint64_t>& rt);
int rows2_op_setup(ed_op_s* op, const std::vector<edc::Bond>& bonds, const std::vector<edc::Chiral>& ch);
bool rows2_shape_compiled(int N, int W);
template <typename VT>
int launch_hv_rows2(ed_op_s* op, const VT* u, const VT* v, VT* w, hipStream_t st, bool wide_image);
ed_basis_s* block_basis(const ed_block_s* K);
int block_size(const ed_block_s* K);
int mblock_op_hv(ed_mblock_s* MB, int dtype, const void* v, void* w, hipStream_t st);
int mblock_nsec(const ed_mblock_s* MB);
int mblock_set_live(ed_mblock_s* MB, const std::vector<char>& live);
int mblock_M(const ed_mblock_s* MB);
int64_t mblock_row0(const ed_mblock_s* MB, int s);
ed_block_s* mblock_block(const ed_mblock_s* MB, int s);
