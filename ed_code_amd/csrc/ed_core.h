// ed_core.h -- state-space arithmetic shared by the gfx950 kernels.
//
// Everything here is __host__ __device__ so that the exact code the kernels
// run can also be compiled by g++ for CPU unit tests of the logic (tests/
// builds it; the product library never runs it on the host).
//
// Conventions (reference semantics):
//   * bit s of a state is site s; bit = 1 is spin up (Basis_States.cpp:8-40);
//   * a translation T moves bit i to bit T[i] (Momentum_Basis.cpp:4-16);
//   * the representative of s is the minimum integer over the L1 x L2 grid
//     T1^x T2^y s, walked as Momentum_Basis.cpp:493-536 does, and the phase
//     index (lx, ly) is taken at the first strict new minimum;
//   * states of weight n are ranked in ascending order (Lin_Tables.cpp:42-78,
//     Momentum_Basis.cpp:732-737), which is the combinatorial number system
//     rank(s) = sum_t C(p_t, t) over the set-bit positions p_1 < p_2 < ....
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ED_HD __host__ __device__ __forceinline__
#else
#define ED_HD inline
#endif

namespace edc {

constexpr int kMaxGroups = 64;   // masked-shift groups per translation
constexpr int kDirSpan = 64;     // ranks covered by one directory entry
constexpr int kLidShift = 52;    // row word = state | (lid << kLidShift)
constexpr uint64_t kStateMask = (uint64_t(1) << kLidShift) - 1;

// A site permutation as a sum of masked shifts:
// T(s) = OR_g shift(s & mask[g], sh[g]) (sh > 0: left shift).
struct TProg {
    int32_t ng;
    int32_t pad;
    uint64_t mask[kMaxGroups];
    int32_t sh[kMaxGroups];
};

// Translation-grid geometry of one sector: the reference's L1 x L2 (x L3)
// walk; two-generator lattices have L3 = 1 and T3 = identity.
struct Geo {
    int32_t N, L1, L2, L3, kx, ky, kz;
    int32_t pad;
    TProg T1, T2, T3;
};

// Byte-sliced rank tables: rank(s) = sum_k tab[(off[k] + c_k - cmin[k])*256 + byte_k(s)]
// with c_k = number of set bits below byte k.
struct RankInfo {
    int32_t nbytes;
    int32_t off[8];
    int32_t cmin[8];
};

// Directory entry for ranks [64g, 64g+64): base = number of valid
// representatives of rank < 64g; b0, b1 = occupancy of the 64 ranks
// (12 bytes: 1.5 bit per state).
struct DirEntry {
    uint32_t base, b0, b1;
};

template <typename S>
ED_HD S apply_T(S s, const TProg& P) {
    S t = 0;
    for (int g = 0; g < P.ng; g++) {
        S m = s & (S)P.mask[g];
        int sh = P.sh[g];
        t |= (sh >= 0) ? (S)(m << sh) : (S)(m >> (-sh));
    }
    return t;
}

// Representative (orbit minimum) of s; code = (lx*L2 + ly)*L3 + lz of the
// first strict minimum in the reference walk (Momentum_Basis.cpp:493-536,
// 745-751; three generators :566-631, 860-870).
template <typename S, bool TRACK>
ED_HD S find_rep(S s, const Geo& G, int& code) {
    S t = s, best = s;
    int bx = 0, by = 0, bz = 0;
    for (int x = 0; x < G.L1; x++) {
        if (x > 0) {
            t = apply_T(t, G.T1);
            if (t < best) { best = t; if (TRACK) { bx = x; by = 0; bz = 0; } }
        }
        for (int y = 0; y < G.L2; y++) {
            if (y > 0) {
                t = apply_T(t, G.T2);
                if (t < best) { best = t; if (TRACK) { bx = x; by = y; bz = 0; } }
            }
            for (int z = 1; z < G.L3; z++) {
                t = apply_T(t, G.T3);
                if (t < best) { best = t; if (TRACK) { bx = x; by = y; bz = z; } }
            }
        }
    }
    if (TRACK) {
        // ty, tz count every T2 / T3 step of the walk (they are not reset)
        const int ty = bx * (G.L2 - 1) + by;
        const int tz = (bx * G.L2 + by) * (G.L3 - 1) + bz;
        const int lx = (G.L1 - bx) % G.L1;
        const int ly = ((G.L2 - ty) % G.L2 + G.L2) % G.L2;
        const int lz = ((G.L3 - tz) % G.L3 + G.L3) % G.L3;
        code = (lx * G.L2 + ly) * G.L3 + lz;
    } else {
        code = 0;
    }
    return best;
}

// Is s its own orbit minimum with a nonvanishing momentum projection?
// Returns lid (grid points fixing s, >= 1) or 0 if s is not a valid
// representative.  The projection test is the exact form of
// Momentum_Basis.cpp:538-540 (3-D: :633-636): the sum of
// exp(-i(px tx + py ty + pz tz)) over the fixing grid points is nonzero iff
// every one of them has kx*tx/L1 + ky*ty/L2 + kz*tz/L3 integral (they form a
// subgroup of Z_L1 x Z_L2 x Z_L3).
template <typename S>
ED_HD int classify(S s, const Geo& G) {
    S t = s;
    int lid = 1;
    bool ok = true;
    const int64_t M = (int64_t)G.L1 * G.L2 * G.L3;
    int ty = 0, tz = 0;
    auto fixed = [&](int x) {
        lid++;
        if (((int64_t)G.kx * x * G.L2 * G.L3 + (int64_t)G.ky * ty * G.L1 * G.L3 + (int64_t)G.kz * tz * G.L1 * G.L2) % M)
            ok = false;
    };
    for (int x = 0; x < G.L1; x++) {
        if (x > 0) {
            t = apply_T(t, G.T1);
            if (t < s) return 0;
            if (t == s) fixed(x);
        }
        for (int y = 0; y < G.L2; y++) {
            if (y > 0) {
                t = apply_T(t, G.T2);
                ty++;
                if (t < s) return 0;
                if (t == s) fixed(x);
            }
            for (int z = 1; z < G.L3; z++) {
                t = apply_T(t, G.T3);
                tz++;
                if (t < s) return 0;
                if (t == s) fixed(x);
            }
        }
    }
    return ok ? lid : 0;
}

// Rank of a weight-n word through the byte-sliced tables.  The last byte's
// table has a single block: its count of lower set bits is n - popcount(byte).
template <typename T, typename S>
ED_HD T rank_tab(S s, const RankInfo& R, const T* tab, int nbytes) {
    T r = 0;
    int c = 0;
    for (int k = 0; k < nbytes; k++) {
        unsigned b = (unsigned)(s >> (8 * k)) & 255u;
        int blk = (k == nbytes - 1) ? R.off[k] : (R.off[k] + c - R.cmin[k]);
        r += tab[blk * 256 + b];
#if defined(__HIP_DEVICE_COMPILE__)
        c += __popc(b);
#else
        c += __builtin_popcount(b);
#endif
    }
    return r;
}

template <typename S>
ED_HD uint64_t rank_of(S s, const RankInfo& R, const uint64_t* tab) {
    return rank_tab<uint64_t, S>(s, R, tab, R.nbytes);
}

ED_HD int popc32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __popc(x);
#else
    return __builtin_popcount(x);
#endif
}

ED_HD int ctz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ffsll((unsigned long long)x) - 1;
#else
    return __builtin_ctzll(x);
#endif
}

// Index of the representative with rank r, or -1 if r is not a valid
// representative.
ED_HD int64_t dir_lookup(uint64_t r, const DirEntry* dir) {
    DirEntry e = dir[r / kDirSpan];
    unsigned o = (unsigned)(r % kDirSpan);
    unsigned wsel = o >> 5, bit = o & 31u;
    uint32_t bits = wsel == 0 ? e.b0 : e.b1;
    if (!((bits >> bit) & 1u)) return -1;
    uint32_t j = e.base + popc32(bits & ((1u << bit) - 1u));
    if (wsel > 0) j += popc32(e.b0);
    return (int64_t)j;
}

// Next word of the same weight in ascending order (Gosper; the division of
// Basis_States.cpp:64-68 is by a power of two, done here as a shift).
ED_HD uint64_t gosper_next(uint64_t s) {
    uint64_t t = (s | (s - 1)) + 1;
    int sh = ctz64(s) + 1;
    return t | (((t & (~t + 1)) >> sh) - 1);
}

// State of rank r among N-bit words of weight n (inverse of rank_of);
// binom[p*65 + t] = C(p, t).
ED_HD uint64_t unrank(uint64_t r, int N, int n, const uint64_t* binom) {
    uint64_t s = 0;
    int p = N - 1;
    for (int t = n; t >= 1; t--) {
        while (binom[p * 65 + t] > r) p--;
        s |= (uint64_t)1 << p;
        r -= binom[p * 65 + t];
        p--;
    }
    return s;
}

}  // namespace edc

namespace edc {

struct alignas(16) cd2 { double re, im; };
struct alignas(8) cf2 { float re, im; };

ED_HD cd2 load_c(const cd2* p, int64_t j) { return p[j]; }
ED_HD cd2 load_c(const cf2* p, int64_t j) { cf2 v = p[j]; return cd2{(double)v.re, (double)v.im}; }
// Element j through a 32-bit byte offset (a scalar base plus a 32-bit vector
// offset on the device); the caller guarantees (j + 1) * sizeof(VT) < 2^32.
template <typename VT>
ED_HD cd2 load_c32(const VT* p, uint32_t j) {
    return load_c((const VT*)((const char*)p + (uint32_t)(j * (uint32_t)sizeof(VT))), 0);
}

// Device-side Hamiltonian terms.  Terms with |J| < 1e-16 and the chiral /
// bowtie lists when |Ch|, |ChCh| <= 1e-16 are dropped at upload, which is
// what H_v.cpp:232-233,263,305 skip at run time.
struct alignas(16) Bond { int32_t x, y; double J; double re, im; };
struct alignas(16) Chiral { int32_t i, j, k, pad; double re, im; };
struct Bowtie { int32_t x, y, z, p, q, r; };

struct TermsView {
    double lam, Ch, ChCh;
    int32_t nb, nc, nw;
    const Bond* b;
    const Chiral* c;
    const Bowtie* w;
};

struct TablesView {
    const uint64_t* rtab;   // rank tables
    const DirEntry* dir;    // representative directory
    const cd2* phase;       // [L1*L2] exp(-i(px lx + py ly)), host-computed
};

template <typename S, bool TRIV>
ED_HD int64_t locate(S t, const Geo& G, const RankInfo& R, const TablesView& tv, int& code) {
    S r = find_rep<S, !TRIV>(t, G, code);
    return dir_lookup(rank_of<S>(r, R, tv.rtab), tv.dir);
}

// Visits one row of H_v.cpp:218-355 in the reference's term order (bond
// terms, chiral terms, bowtie terms), calling f(j, cr, ci) with the target
// index j and the element coefficient without the norm ratio:
//     M(i, j) = (cr + i ci) * norm[i] / norm[j].
// Returns the diagonal Ezz.
template <typename S, bool TRIV, class F>
ED_HD double visit_row(S s, const Geo& G, const RankInfo& R, const TablesView& tv, const TermsView& T, F&& f) {
    double ez = 0.0;
    const double fac = 0.25 * T.lam;
    for (int k = 0; k < T.nb; k++) {
        const Bond bd = T.b[k];
        bool tx = (s >> bd.x) & 1, ty = (s >> bd.y) & 1;
        if (tx == ty) { ez += fac * bd.J; continue; }
        ez -= fac * bd.J;
        S t = s ^ (((S)1 << bd.x) | ((S)1 << bd.y));
        int code;
        int64_t j = locate<S, TRIV>(t, G, R, tv, code);
        if (j < 0) continue;
        // exp(i phi), conjugated when bit x is set (H_v.cpp:242)
        double cr = bd.re * bd.J * 0.5, ci = (tx ? -bd.im : bd.im) * bd.J * 0.5;
        if (!TRIV) {
            cd2 ph = tv.phase[code];
            double nr = cr * ph.re - ci * ph.im, ni = cr * ph.im + ci * ph.re;
            cr = nr; ci = ni;
        }
        f(j, cr, ci);
    }
    for (int k = 0; k < T.nc; k++) {
        const Chiral c = T.c[k];
        bool tx = (s >> c.i) & 1, ty = (s >> c.j) & 1;
        if (tx == ty) continue;
        S t = s ^ (((S)1 << c.i) | ((S)1 << c.j));
        int code;
        int64_t j = locate<S, TRIV>(t, G, R, tv, code);
        if (j < 0) continue;
        // ch_sgn = +i*ep if bit j set, else -i*conj(ep); negated if bit k clear (H_v.cpp:287-295)
        double sr = -c.im, si = ty ? c.re : -c.re;
        if (!((s >> c.k) & 1)) { sr = -sr; si = -si; }
        double cr = T.Ch * 0.25 * sr, ci = T.Ch * 0.25 * si;
        if (!TRIV) {
            cd2 ph = tv.phase[code];
            double nr = cr * ph.re - ci * ph.im, ni = cr * ph.im + ci * ph.re;
            cr = nr; ci = ni;
        }
        f(j, cr, ci);
    }
    for (int k = 0; k < T.nw; k++) {
        const Bowtie w = T.w[k];
        bool tx = (s >> w.x) & 1, ty = (s >> w.y) & 1;
        if (tx == ty) continue;
        S t = s ^ (((S)1 << w.x) | ((S)1 << w.y));
        double s1 = ty ? 1.0 : -1.0;            // ch_sgn1 = s1 * i   (H_v.cpp:318-328)
        if (!((t >> w.z) & 1)) s1 = -s1;
        bool tp = (t >> w.p) & 1, tq = (t >> w.q) & 1;
        if (tp == tq) continue;
        double s2 = tq ? 1.0 : -1.0;            // ch_sgn2 = s2 * i   (H_v.cpp:329-346)
        S t2 = t ^ (((S)1 << w.p) | ((S)1 << w.q));
        int code;
        int64_t j = locate<S, TRIV>(t2, G, R, tv, code);
        if (j < 0) continue;
        if (!((t2 >> w.r) & 1)) s2 = -s2;
        double cr = -T.ChCh * 0.0625 * s1 * s2, ci = 0.0;   // (i s1)(i s2) = -s1 s2
        if (!TRIV) {
            cd2 ph = tv.phase[code];
            ci = cr * ph.im;
            cr = cr * ph.re;
        }
        f(j, cr, ci);
    }
    return ez;
}

}  // namespace edc
