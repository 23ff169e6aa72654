// ed_tables.h -- host-side construction of the small per-sector tables
// (translation programs, rank tables, phase and norm tables).  Plain C++; used
// by the HIP library and by the CPU unit-test build of ed_core.h.
#pragma once
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#include <algorithm>

#include "ed_core.h"

namespace edc {

inline uint64_t binom_u64(int n, int k) {
    if (k < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 r = 1;
    for (int i = 1; i <= k; i++) r = r * (unsigned)(n - k + i) / (unsigned)i;
    return (uint64_t)r;
}

// Groups the sites of permutation T by displacement d = T[i] - i.
inline bool make_prog(const std::vector<int>& T, TProg& P) {
    std::memset(&P, 0, sizeof(P));
    std::vector<int> ds;
    std::vector<uint64_t> ms;
    for (size_t i = 0; i < T.size(); i++) {
        int d = T[i] - (int)i;
        size_t g = 0;
        while (g < ds.size() && ds[g] != d) g++;
        if (g == ds.size()) { ds.push_back(d); ms.push_back(0); }
        ms[g] |= (uint64_t)1 << i;
    }
    if ((int)ds.size() > kMaxGroups) return false;
    P.ng = (int)ds.size();
    for (int g = 0; g < P.ng; g++) { P.mask[g] = ms[g]; P.sh[g] = ds[g]; }
    return true;
}

// Cycle length of site 0 (LxLy, Momentum_Basis.cpp:19-36).
inline int cycle_len0(const std::vector<int>& T) {
    int s = 0, n = 0;
    do { s = T[s]; n++; } while (s != 0 && n <= (int)T.size());
    return n;
}

}  // namespace edc

// C99 complex exponential from the C library.  The reference's
// std::exp(complex<double>) (Momentum_Basis.cpp:753) compiles to this call
// under g++; calling it directly keeps the phase table bit-equal under any
// host compiler (clang's own lowering of std::exp differs in the last bit).
extern "C" double _Complex cexp(double _Complex);

namespace edc {

// exp(-i(px lx + py ly)) for lx < L1, ly < L2, the expression of
// Momentum_Basis.cpp:753, indexed lx*L2 + ly.
inline std::vector<cd2> phase_table(int L1, int L2, int L3, double px, double py, double pz) {
#if defined(__clang__)
#pragma clang fp contract(off)   // no FMA: round -px*lx and py*ly separately, as g++ does
#endif
    std::vector<cd2> ph((size_t)L1 * L2 * L3);
    for (int lx = 0; lx < L1; lx++)
        for (int ly = 0; ly < L2; ly++)
            for (int lz = 0; lz < L3; lz++) {
                double _Complex z;
                __real__ z = 0.0;
                // the two-generator expression (Momentum_Basis.cpp:753) unless a
                // third generator is present (:872)
                __imag__ z = L3 == 1 ? -px * lx - py * ly : -px * lx - py * ly - pz * lz;
                double _Complex e = cexp(z);
                ph[((size_t)lx * L2 + ly) * L3 + lz] = cd2{__real__ e, __imag__ e};
            }
    return ph;
}

// norm = sqrt(L1*L2*L3/lid) with an integer quotient, 1/norm, as
// Momentum_Basis.cpp:764-766 (three generators :883-885), indexed by lid.
inline void norm_tables(int L1, int L2, int L3, std::vector<double>& nt, std::vector<double>& nit) {
    const int M = L1 * L2 * L3;
    nt.assign(M + 1, 0.0);
    nit.assign(M + 1, 0.0);
    for (int lid = 1; lid <= M; lid++) {
        int64_t orb = (int64_t)M / lid;
        nt[lid] = std::sqrt((double)orb);
        nit[lid] = 1 / std::sqrt((double)orb);
    }
}

inline std::vector<uint64_t> binom_table() {
    std::vector<uint64_t> bn(65 * 65);
    for (int p = 0; p < 65; p++)
        for (int t = 0; t < 65; t++) bn[p * 65 + t] = binom_u64(p, t);
    return bn;
}

// Byte-sliced combinatorial-number-system rank tables for weight-n words of
// N bits: entry (k, c, b) = sum over the set bits q of byte value b, taken in
// increasing order with local index i = 1, 2, ..., of C(8k + q, c + i), for
// every possible count c of set bits below byte k; the top byte has one block,
// c = n - popcount(b).
inline std::vector<uint64_t> rank_tables(int N, int n, RankInfo& R) {
    std::memset(&R, 0, sizeof(R));
    R.nbytes = (N + 7) / 8;
    std::vector<uint64_t> rt;
    int off = 0;
    for (int k = 0; k < R.nbytes; k++) {
        int cmin = std::max(0, n - (N - 8 * k));
        int cmax = std::min(n, 8 * k);
        R.off[k] = off;
        R.cmin[k] = cmin;
        if (k == R.nbytes - 1) {
            // the set bits below the top byte number n - popcount(byte)
            for (int b = 0; b < 256; b++) {
                int c = n - __builtin_popcount(b);
                uint64_t v = 0;
                if (c >= 0 && c <= 8 * k) {
                    int i = 0;
                    for (int q = 0; q < 8; q++)
                        if ((b >> q) & 1) { i++; v += binom_u64(8 * k + q, c + i); }
                }
                rt.push_back(v);
            }
            off++;
            continue;
        }
        for (int c = cmin; c <= cmax; c++) {
            for (int b = 0; b < 256; b++) {
                uint64_t v = 0;
                int i = 0;
                for (int q = 0; q < 8; q++)
                    if ((b >> q) & 1) { i++; v += binom_u64(8 * k + q, c + i); }
                rt.push_back(v);
            }
            off++;
        }
    }
    return rt;
}

// Gather image of the prescaled vector u for the compiled rotation kernels.
// u is laid out by rank instead of by row: the rank space [0, C) is cut into
// chunks of 2^cb ranks, each chunk that holds a representative of the
// sector gets a slot of 2^cb elements (slots in rank order), and one more
// slot, index nslot, holds zeros.  The target of rank r is element
//     (map[r >> cb] << cb) | (r & (2^cb - 1)),
// map sending every chunk without representatives to the zero slot; inside a
// slot the positions of ranks that are no representative of the sector
// (orbits outside the sector at k != 0, non-minimal words) are never written
// and stay zero.  The rank alone addresses u: a gather needs the small map
// (LDS or cache resident) instead of a directory entry.
template <typename M>
struct ImageLayoutT {
    int cb = -1;                 // chunk bits; -1: no image
    uint64_t nslot = 0;          // slots holding representatives; the zero slot
    std::vector<M> map;          // chunk -> slot, (C - 1) >> cb + 1 entries
    uint64_t elems() const { return cb < 0 ? 0 : (nslot + 1) << cb; }
};
// The compiled rotation kernels keep a u16 map of at most kImageMapMax
// entries (6 KiB) in LDS; the two-row 64-bit kernels (N <= 40) read a u32 map
// of at most kWideMapMax entries (16 MiB) from global memory, where it stays
// in L2 and the Infinity Cache.
using ImageLayout = ImageLayoutT<uint16_t>;
using WideImageLayout = ImageLayoutT<uint32_t>;
#ifndef ED_IMAGE_MAP_MAX
#define ED_IMAGE_MAP_MAX 3072
#endif
constexpr uint64_t kImageMapMax = ED_IMAGE_MAP_MAX;
constexpr uint64_t kWideMapMax = uint64_t(1) << 22;

// Smallest chunk size whose map fits max_entries (at least one 64-rank
// directory group per chunk).
inline int image_min_bits(uint64_t C, uint64_t max_entries) {
    int cb = 6;
    while (((C > 0 ? C - 1 : 0) >> cb) + 1 > max_entries) cb++;
    return cb;
}

// Builds the layout for chunk bits cb from touched[c] != 0 (chunk c holds a
// representative).  False when the image would exceed max_elems elements or
// the slot count the map type holds.
template <typename M>
inline bool image_layout(uint64_t C, int cb, const std::vector<uint8_t>& touched, uint64_t max_elems,
                         ImageLayoutT<M>& L) {
    const uint64_t nmap = ((C > 0 ? C - 1 : 0) >> cb) + 1;
    if (touched.size() < nmap) return false;
    uint64_t n = 0;
    for (uint64_t c = 0; c < nmap; c++) n += touched[c] ? 1u : 0u;
    if (n >= (uint64_t)(M)~M(0) || ((n + 1) << cb) > max_elems) return false;
    L.cb = cb;
    L.nslot = n;
    L.map.assign(nmap, (M)n);
    M sl = 0;
    for (uint64_t c = 0; c < nmap; c++)
        if (touched[c]) L.map[c] = sl++;
    return true;
}

// Element bound of the compiled kernels' image: 6 (D + 1) (at least 2^22) and
// 4 GiB of complex128, the reach of their 32-bit load offsets.
inline uint64_t image_max_elems(int64_t D) {
    return std::min<uint64_t>(std::max<uint64_t>(6 * (uint64_t)(D + 1), uint64_t(1) << 22), 0xffffffffull / 16);
}

// Element bound of the two-row kernels' image: 4 (D + 1) + 2^22.
inline uint64_t wide_image_max_elems(int64_t D) { return 4 * (uint64_t)(D + 1) + (uint64_t(1) << 22); }

}  // namespace edc
