// ed_torus.h -- the rotation fast path for lattices whose translations are bit
// rotations (sites labelled x + w*y on a w x L2 torus, or a ring), N <= 32.
//
// On such lattices T1 rotates every w-bit row by one and T2 rotates the
// N-bit word by w, so every group element is a field rotation followed by a
// word rotation: a handful of ALU operations instead of a bit-permutation
// walk.  Bonds (and chiral entries) are grouped into families that share the
// group element h relating their two flipped sites (y = h(x)); the
// antiparallel bonds of a family are the set bits of (s ^ h^-1(s)) & S.
// __host__ __device__ like ed_core.h so that the CPU tests run this code.
#pragma once
#include "ed_core.h"

namespace edc {

constexpr int kMaxFam = 16;

struct Torus {
    int32_t kind;       // 0: not rotational; 1: w x L2 rows; 2: ring (T2 = identity)
    int32_t N, w, L1, L2;
    uint32_t nmask;     // (1 << N) - 1
    uint32_t hi[32];    // hi[r]: destinations of bits moved left by a field rotation by r
};

struct TFam {
    int32_t type;       // 0 bond, 1 chiral
    int32_t a, q;       // h^-1 = field rotation by a, then word rotation by q bits
    uint32_t S;         // sites x that carry a term of this family
};

struct alignas(16) TEntry {
    double cre, cim;    // bond: J/2 * exp(i phi); chiral: Ch/4 * exp(i theta)
    double dJ2;         // bond: Delta*J/2 (Ezz correction when antiparallel); chiral: 0
    int32_t y, k;       // flip partner, chiral z-site
};

// The chiral entries that flip the same pair (x, y) as a bond family member
// (kModeMerged): at most two, each with its coefficient Ch/4 exp(i theta),
// its z-site k, and whether its j-site is x (the entry lists the pair as
// (y, x)) rather than y.  flags: bit 2c = entry c present, bit 2c+1 = j is x.
struct alignas(16) TChir {
    double cre[2], cim[2];
    int32_t k[2];
    int32_t flags, pad;
};

ED_HD uint32_t alignbit32(uint32_t hi, uint32_t lo, uint32_t sh) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, sh);
#else
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (sh & 31));
#endif
}

// Index of the lowest set bit; x != 0 at every call (one v_ffbl, no guard
// for a zero argument).
ED_HD int ctz32(uint32_t x) { return __builtin_ctz(x); }

ED_HD uint32_t min3u(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t m = a < b ? a : b;
    return m < c ? m : c;
}

// every w-bit row rotated by r (0 <= r < w): site (x, y) -> (x + r mod w, y)
ED_HD uint32_t field_rot(uint32_t s, int r, const Torus& T) {
    if (r == 0) return s;
    uint32_t h = T.hi[r];
    return ((s << r) & h) | ((s >> (T.w - r)) & ~h);
}

// The N-bit word s as the concatenation hi:lo = (s << N) | s, so that a
// rotation left by q is alignbit(hi, lo, N - q) & nmask.
struct Cat { uint32_t hi, lo; };
ED_HD Cat cat_of(uint32_t s, const Torus& T) {
    if (T.N == 32) return Cat{s, s};
    return Cat{s >> (32 - T.N), (s << T.N) | s};
}
ED_HD uint32_t rot_cat(Cat c, int q, const Torus& T) {
    if (q == 0) return c.lo & T.nmask;
    return alignbit32(c.hi, c.lo, (uint32_t)(T.N - q)) & T.nmask;
}

ED_HD uint32_t torus_apply(uint32_t s, int a, int q, const Torus& T) {
    return rot_cat(cat_of(field_rot(s, a, T), T), q, T);
}

// Representative of t (orbit minimum).  With TRACK, code = lx*L2 + ly of the
// first minimum in the reference's walk (Momentum_Basis.cpp:493-536): the
// smallest T1 power x, then the smallest walk column y, which for T2 power b
// is y = (b + x) mod L2.
template <bool TRACK>
ED_HD uint32_t torus_rep(uint32_t t, const Torus& T, int& code) {
    const int L1 = T.L1, L2 = T.L2;
    uint32_t best = 0xffffffffu;
    int bx = 0;
    if (T.kind == 2) {
        Cat c = cat_of(t, T);
        for (int a = 0; a < L1; a += 2) {
            uint32_t c0 = rot_cat(c, a, T);
            uint32_t c1 = (a + 1 < L1) ? rot_cat(c, a + 1, T) : 0xffffffffu;
            best = min3u(best, c0, c1);
        }
        if (TRACK) {
            for (int a = L1 - 1; a >= 0; a--)
                if (rot_cat(c, a, T) == best) bx = a;
            code = ((L1 - bx) % L1) * L2;
        } else {
            code = 0;
        }
        return best;
    }
    uint32_t sa = t;
    uint32_t sbest = t;
    for (int a = 0; a < L1; a++) {
        if (a > 0) sa = field_rot(t, a, T);
        Cat c = cat_of(sa, T);
        uint32_t m = 0xffffffffu;
        for (int b = 0; b < L2; b += 2) {
            uint32_t c0 = rot_cat(c, b * T.w, T);
            uint32_t c1 = (b + 1 < L2) ? rot_cat(c, (b + 1) * T.w, T) : 0xffffffffu;
            m = min3u(m, c0, c1);
        }
        if (TRACK) {
            if (m < best) { best = m; bx = a; sbest = sa; }
        } else {
            best = m < best ? m : best;
        }
    }
    if (TRACK) {
        Cat c = cat_of(sbest, T);
        int by = 0;
        for (int y = L2 - 1; y >= 0; y--) {
            int b = ((y - bx) % L2 + L2) % L2;
            if (rot_cat(c, b * T.w, T) == best) by = y;
        }
        int ty = bx * (L2 - 1) + by;
        int lx = (L1 - bx) % L1;
        int ly = ((L2 - ty) % L2 + L2) % L2;
        code = lx * L2 + ly;
    } else {
        code = 0;
    }
    return best;
}

ED_HD uint32_t fam_mask(uint32_t s, const TFam& F, const Torus& T) {
    return (s ^ torus_apply(s, F.a, F.q, T)) & F.S;
}

// One row on the rotation fast path.  Calls g(j, cr, ci) per element as
// visit_row does; returns Ezz.  rtab is the 32-bit rank table.
template <bool TRIV, class Gf>
ED_HD double torus_row(uint32_t s, const Torus& T, int nfam, const TFam* fam, const TEntry* ent, double ezz0,
                       const RankInfo& R, const uint32_t* rtab, const DirEntry* dir, const cd2* phase, Gf&& g) {
    double ezz = ezz0;
    int f = 0;
    uint32_t cur = nfam > 0 ? fam_mask(s, fam[0], T) : 0u;
    while (true) {
        while (cur == 0u) {
            if (++f >= nfam) break;
            cur = fam_mask(s, fam[f], T);
        }
        if (f >= nfam) break;
        int x = ctz32(cur);
        cur &= cur - 1u;
        const TEntry E = ent[f * 32 + x];
        uint32_t t = s ^ (1u << x) ^ (1u << E.y);
        double cr, ci;
        if (fam[f].type == 0) {
            ezz -= E.dJ2;
            cr = E.cre;
            ci = ((s >> x) & 1u) ? -E.cim : E.cim;
        } else {
            // ch_sgn = +i*ep if bit j set else -i*conj(ep); negated if bit k clear
            double sg = ((s >> E.k) & 1u) ? 1.0 : -1.0;
            cr = -E.cim * sg;
            ci = (((s >> E.y) & 1u) ? E.cre : -E.cre) * sg;
        }
        int code;
        uint32_t rep = torus_rep<!TRIV>(t, T, code);
        // rank of the representative (byte-sliced tables)
        uint32_t r = rank_tab<uint32_t, uint32_t>(rep, R, rtab, R.nbytes);
        int64_t j = dir_lookup(r, dir);
        if (j < 0) continue;
        if (!TRIV) {
            cd2 ph = phase[code];
            double nr = cr * ph.re - ci * ph.im, ni = cr * ph.im + ci * ph.re;
            cr = nr; ci = ni;
        }
        g(j, cr, ci);
    }
    return ezz;
}

}  // namespace edc

namespace edc {

// Compile-time torus shape: N sites, rows of W sites, L2 rows (L2 == 1: ring,
// T1 = rotation of the whole word, T2 = identity).
template <int N_, int W_, int L2_>
struct TorusC {
    static constexpr int N = N_, W = W_, L2 = L2_;
    static constexpr bool ring = (L2_ == 1);
    static constexpr int L1 = ring ? N_ : W_;
    static constexpr uint32_t nmask = N_ == 32 ? 0xffffffffu : ((1u << N_) - 1u);
    static constexpr uint32_t rowones() {   // bit 0 of every row
        uint32_t m = 0;
        for (int y = 0; y < L2_; y++) m |= 1u << (y * W_);
        return m;
    }
    static constexpr uint32_t wmask = W_ >= 32 ? 0xffffffffu : ((1u << W_) - 1u);
    static constexpr uint32_t hi(int r) {    // sites with x >= r
        return rowones() * ((wmask >> r) << r);
    }
};

template <class TC>
ED_HD uint32_t rot_c(uint32_t hi, uint32_t lo, int q) {
    if (q == 0) return lo & TC::nmask;
    if (TC::N == 32) return alignbit32(lo, lo, (uint32_t)(32 - q));
    return alignbit32(hi, lo, (uint32_t)(TC::N - q)) & TC::nmask;
}

template <class TC>
ED_HD void cat_c(uint32_t s, uint32_t& hi, uint32_t& lo) {
    if (TC::N == 32) { hi = s; lo = s; }
    else { hi = s >> (32 - TC::N); lo = (s << TC::N) | s; }
}

template <class TC, int R>
ED_HD uint32_t field_rot_c(uint32_t s) {
    if (R == 0) return s;
    constexpr uint32_t h = TC::hi(R);
    return ((s << R) & h) | ((s >> (TC::W - R)) & ~h);
}

// runtime rotation amount (family masks): hi(r) computed arithmetically
template <class TC>
ED_HD uint32_t field_rot_rt(uint32_t s, int r) {
    if (r == 0) return s;
    uint32_t h = TC::rowones() * ((TC::wmask >> r) << r);
    return ((s << r) & h) | ((s >> (TC::W - r)) & ~h);
}

template <class TC>
ED_HD uint32_t fam_mask_c(uint32_t s, const TFam& F) {
    uint32_t t = TC::ring ? s : field_rot_rt<TC>(s, F.a);
    uint32_t hi, lo;
    cat_c<TC>(t, hi, lo);
    return (s ^ rot_c<TC>(hi, lo, F.q)) & F.S;
}

// Top-aligned doubled word for the orbit minimum of an N-bit field:
// Y = s 2^(64-N) + s 2^(64-2N), hi:lo = Y.  The left rotation of s by q is the
// top N bits of alignbit(hi, lo, 32 - q) (0 < q < 32) and of hi (q = 0), the
// same rotation rot_c gives.  The low 32 - N bits of such a candidate continue
// the cycle; they only order candidates whose top N bits are equal, so the
// minimum's top N bits are the exact minimum and no candidate needs a mask.
template <class TC>
ED_HD void cat_top(uint32_t s, uint32_t& hi, uint32_t& lo) {
    if (TC::N == 32) { hi = s; lo = s; return; }
    constexpr int S = 32 - TC::N;
    if (2 * TC::N > 32) {
        hi = (s << S) | (s >> (2 * TC::N - 32));
        lo = s << (64 - 2 * TC::N);
    } else {
        hi = (s << S) | (s << (32 - 2 * TC::N));
        lo = 0u;
    }
}

// top-aligned candidate of rotation q (see cat_top)
template <class TC>
ED_HD uint32_t rot_top(uint32_t hi, uint32_t lo, int q) {
    return q == 0 ? hi : alignbit32(hi, lo, (uint32_t)(32 - q));
}

// Whether two top-aligned candidates (rot_top) hold the same state.  A
// candidate's low 32 - N bits continue the doubled word: they repeat the
// top bits of the same rotated state when the rotation is at most 2N - 32
// (the window stays inside the two copies of the state), so then equal
// states are equal words; larger rotations of the shape reach into the
// zero padding and only the top N bits are compared.
template <class TC>
ED_HD bool top_eq(uint32_t a, uint32_t b) {
    constexpr int S = 32 - TC::N;
    constexpr int qmax = TC::ring ? TC::N - 1 : TC::N - TC::W;
    if (TC::N == 32 || qmax <= 2 * TC::N - 32) return a == b;
    return (a >> S) == (b >> S);
}

// The N-bit state rotated left by q bits (0 <= q < N, run-time q).
template <class TC>
ED_HD uint32_t rot_rt(uint32_t s, uint32_t q) {
    uint32_t hi, lo;
    cat_c<TC>(s, hi, lo);
    if (TC::N == 32) return alignbit32(lo, lo, (32u - q) & 31u);
    return alignbit32(hi, lo, (uint32_t)TC::N - q) & TC::nmask;
}

// What torus_rep_c reports besides the representative: nothing, the
// reference's phase code (the first minimum of its translation walk), or the
// code of some translation that maps t to its representative.  Two such
// translations differ by a stabiliser element of the representative; when
// the representative belongs to the sector every stabiliser element has
// phase 1, so both codes give the target the same phase (to rounding: the
// table evaluates each code's own angle), and a target outside the sector
// reads a zero of the gather image whatever its phase.
constexpr int kRepOnly = 0, kRepFirst = 1, kRepAny = 2;

// torus_rep for a compile-time shape; same result and, with kRepFirst, the
// same phase code.  The candidates are compared top-aligned (cat_top), one
// instruction each.
template <class TC, int TRACK>
ED_HD uint32_t torus_rep_c(uint32_t t, int& code) {
    constexpr int S = 32 - TC::N;
    uint32_t best = 0xffffffffu;
    if (TC::ring) {
        uint32_t hi, lo;
        cat_top<TC>(t, hi, lo);
#pragma unroll
        for (int a = 0; a < TC::L1; a += 2) {
            uint32_t c0 = rot_top<TC>(hi, lo, a);
            uint32_t c1 = (a + 1 < TC::L1) ? rot_top<TC>(hi, lo, a + 1) : 0xffffffffu;
            best = min3u(best, c0, c1);
        }
        const uint32_t rep = best >> S;
        if (TRACK) {
            int bx = 0;
            for (int a = TC::L1 - 1; a >= 0; a--)
                if (top_eq<TC>(rot_top<TC>(hi, lo, a), best)) bx = a;
            code = (TC::L1 - bx) % TC::L1;
        } else {
            code = 0;
        }
        return rep;
    }
    uint32_t mins[TC::L1], fr[TC::L1];
#pragma unroll
    for (int a = 0; a < TC::L1; a++) {
        uint32_t sa = field_rot_rt<TC>(t, a);   // a is a constant after unrolling
        fr[a] = sa;
        uint32_t hi, lo;
        cat_top<TC>(sa, hi, lo);
        uint32_t m = 0xffffffffu;
#pragma unroll
        for (int b = 0; b < TC::L2; b += 2) {
            uint32_t c0 = rot_top<TC>(hi, lo, b * TC::W);
            uint32_t c1 = (b + 1 < TC::L2) ? rot_top<TC>(hi, lo, (b + 1) * TC::W) : 0xffffffffu;
            m = min3u(m, c0, c1);
        }
        mins[a] = m;
        best = m < best ? m : best;
    }
    const uint32_t rep = best >> S;
    if (TRACK == kRepAny) {
        // the lowest field rotation bx whose row holds the minimum and the
        // lowest word rotation b of that row reaching it: the translation
        // lx = -bx mod L1, ly = -b mod L2
        int bx = 0;
        uint32_t sa = fr[0];
#pragma unroll
        for (int a = TC::L1 - 1; a >= 0; a--)
            if (top_eq<TC>(mins[a], best)) { bx = a; sa = fr[a]; }
        uint32_t hi, lo;
        cat_top<TC>(sa, hi, lo);
        int b = 0;
#pragma unroll
        for (int y = TC::L2 - 1; y > 0; y--)
            if (top_eq<TC>(rot_top<TC>(hi, lo, y * TC::W), best)) b = y;
        code = (bx == 0 ? 0 : TC::L1 - bx) * TC::L2 + (b == 0 ? 0 : TC::L2 - b);
    } else if (TRACK == kRepFirst) {
        // the reference's first minimum: the lowest field rotation bx whose
        // row holds the minimum, then the lowest step y of the walk's
        // cumulative T2 count, y = (b + bx) mod L2 for the word rotation b
        int bx = 0;
        uint32_t sa = fr[0];
#pragma unroll
        for (int a = TC::L1 - 1; a >= 0; a--)
            if (top_eq<TC>(mins[a], best)) { bx = a; sa = fr[a]; }
        const int bxm = bx % TC::L2;
        // step y is the word rotation b = (y - bxm) mod L2: with sa rotated
        // back by bxm rows once, step y is the rotation by y rows, a
        // compile-time amount
        const uint32_t sw = rot_rt<TC>(sa, (uint32_t)((bxm == 0 ? 0 : TC::L2 - bxm) * TC::W));
        uint32_t hi, lo;
        cat_top<TC>(sw, hi, lo);
        int by = 0;
#pragma unroll
        for (int y = TC::L2 - 1; y >= 0; y--)
            if (top_eq<TC>(rot_top<TC>(hi, lo, y * TC::W), best)) by = y;
        // lx = -bx mod L1; ly = -(bx (L2 - 1) + by) mod L2 = (bx - by) mod L2
        const int lx = bx == 0 ? 0 : TC::L1 - bx;
        const int d = bxm - by;
        const int ly = d < 0 ? d + TC::L2 : d;
        code = lx * TC::L2 + ly;
    } else {
        code = 0;
    }
    return rep;
}

// torus_row for a compile-time shape.
template <class TC, bool TRIV, class Gf>
ED_HD double torus_row_c(uint32_t s, int nfam, const TFam* fam, const TEntry* ent, double ezz0, const RankInfo& R,
                         const uint32_t* rtab, const DirEntry* dir, const cd2* phase, Gf&& g) {
    double ezz = ezz0;
    int f = 0;
    uint32_t cur = nfam > 0 ? fam_mask_c<TC>(s, fam[0]) : 0u;
    while (true) {
        while (cur == 0u) {
            if (++f >= nfam) break;
            cur = fam_mask_c<TC>(s, fam[f]);
        }
        if (f >= nfam) break;
        int x = ctz32(cur);
        cur &= cur - 1u;
        const TEntry E = ent[f * 32 + x];
        uint32_t t = s ^ (1u << x) ^ (1u << E.y);
        double cr, ci;
        if (fam[f].type == 0) {
            ezz -= E.dJ2;
            cr = E.cre;
            ci = ((s >> x) & 1u) ? -E.cim : E.cim;
        } else {
            double sg = ((s >> E.k) & 1u) ? 1.0 : -1.0;
            cr = -E.cim * sg;
            ci = (((s >> E.y) & 1u) ? E.cre : -E.cre) * sg;
        }
        int code;
        uint32_t rep = torus_rep_c<TC, TRIV ? kRepOnly : kRepAny>(t, code);
        uint32_t r = rank_tab<uint32_t, uint32_t>(rep, R, rtab, (TC::N + 7) / 8);
        int64_t j = dir_lookup(r, dir);
        if (j < 0) continue;
        if (!TRIV) {
            cd2 ph = phase[code];
            double nr = cr * ph.re - ci * ph.im, ni = cr * ph.im + ci * ph.re;
            cr = nr; ci = ni;
        }
        g(j, cr, ci);
    }
    return ezz;
}

}  // namespace edc

namespace edc {

// Decodes a directory entry fetched for rank r (see dir_lookup).
// Branch-free (word select by masks, so the entry stays in registers).
ED_HD int64_t dir_decode_off(const DirEntry& e, unsigned o) {
    const bool hi = (o >> 5) != 0;
    const unsigned bit = o & 31u;
    const uint32_t low = (1u << bit) - 1u;
    const uint32_t bits = hi ? e.b1 : e.b0;
    // ones below position o: all of b0 and the low bits of b1, or the low bits of b0
    const int64_t idx = (int64_t)(e.base + popc32(e.b0 & (hi ? 0xffffffffu : low)) + popc32(e.b1 & (hi ? low : 0u)));
    return ((bits >> bit) & 1u) ? idx : int64_t(-1);
}

ED_HD int64_t dir_decode(const DirEntry& e, uint32_t r) { return dir_decode_off(e, r % kDirSpan); }

// Term classes of the batched row (a compile-time switch, so the bond-only
// rows carry no branch for chiral entries):
//   kModeGeneral  bond and chiral families, complex coefficients;
//   kModeBonds    bond families only;
//   kModeReal     bond families only with real coefficients at k = 0 (the
//                 caller accumulates with two FMAs and ignores ci);
//   kModeMerged   bond-keyed families whose members also carry the chiral
//                 entries of the same flipped pair (TChir): one orbit search,
//                 rank and gather per pair instead of one per term;
//   kModeUniform  kModeReal where all members of a family share J/2 and
//                 Delta J/2: the coefficient is read once per family and the
//                 diagonal takes each family's antiparallel count at once, so
//                 an element costs no coefficient load and no FP64 add.
enum { kModeGeneral = 0, kModeBonds = 1, kModeReal = 2, kModeMerged = 3, kModeUniform = 4 };

// Tables of a compiled-shape row (staged in LDS by the kernels).
struct RowTabs {
    int nfam;
    const TFam* fam;
    const TEntry* ent;     // per member, [f * 32 + x] (not read under kModeUniform)
    const uint32_t* mki;   // kModeMerged: z-sites k0 | k1 << 8 of each member's chiral entries
    const cd2* mco;        // kModeMerged: coefficient per member at s_x = 0 and (s_k0, s_k1), [(f * 32 + x) * 4 + idx]
    const uint32_t* fmk;   // per member, [f * 32 + x]: the flip mask (1 << x) | (1 << y)
    const cd2* fco;        // kModeUniform: {J/2, Delta J/2} per family
    const uint32_t* rtab;  // 32-bit byte-sliced rank tables
    const cd2* phase;      // exp(-i k.r) per phase code
};

// Cursor over a row's elements: the current family f, its antiparallel
// members not yet visited (cur), and its coefficient under kModeUniform (cu).
struct RowCur {
    int f;
    uint32_t cur;
    double cu;
};

// Moves the cursor onto family c.f.  Under kModeUniform with EZ the diagonal
// takes the whole family at once: ezz -= popcount * Delta J / 2.
template <class TC, int MODE, bool EZ>
ED_HD void fam_enter(uint32_t s, const RowTabs& T, RowCur& c, double& ezz) {
    c.cur = fam_mask_c<TC>(s, T.fam[c.f]);
    if (MODE == kModeUniform) {
        c.cu = T.fco[c.f].re;
        if (EZ) ezz = fma(-(double)popc32(c.cur), T.fco[c.f].im, ezz);
    }
}

// One batch of up to K elements of a row: rank of the representative of each
// target and its coefficient (phase applied); ok[q] false for empty slots.
template <int K>
struct RowBatch {
    uint32_t rk[K];
    double cre[K], cim[K];
    bool ok[K];
};

// Fills the next batch of the row (the orbit search and rank of K
// elements), advancing the cursor and the diagonal ezz.  SPLIT: the row is
// shared by several lanes; at most `left` more elements belong to this lane,
// and the kModeUniform diagonal is not taken here (one lane computes it).
template <class TC, bool TRIV, int MODE, int K, bool SPLIT>
ED_HD void fill_batch(uint32_t s, const RowTabs& T, const RankInfo& R, RowCur& c, int& left, double& ezz,
                      RowBatch<K>& B) {
#pragma unroll
    for (int q = 0; q < K; q++) {
        while (c.cur == 0u && c.f < T.nfam && (!SPLIT || left > 0)) {
            if (++c.f < T.nfam) fam_enter<TC, MODE, !SPLIT>(s, T, c, ezz);
        }
        B.ok[q] = c.f < T.nfam && (!SPLIT || left > 0);
        B.rk[q] = 0; B.cre[q] = 0; B.cim[q] = 0;
        if (B.ok[q]) {
            if (SPLIT) left--;
            const int x = ctz32(c.cur);
            c.cur &= c.cur - 1u;
            const uint32_t t = s ^ T.fmk[c.f * 32 + x];
            double cr, ci = 0.0;
            if (MODE == kModeUniform) {
                cr = c.cu;
            } else if (MODE == kModeMerged) {
                // the bond and chiral terms of the pair, from the member's
                // table of its z-site bits, conjugated for s_x = 1
                // (torus_merged_tables)
                const int m = c.f * 32 + x;
                const uint32_t ki = T.mki[m];
                const uint32_t idx = ((s >> (ki & 31u)) & 1u) | (((s >> (ki >> 8)) & 1u) << 1);
                const cd2 co = T.mco[m * 4 + (int)idx];
                ezz -= T.ent[m].dJ2;
                cr = co.re;
                ci = ((s >> x) & 1u) ? -co.im : co.im;
            } else if (MODE != kModeGeneral || T.fam[c.f].type == 0) {
                const TEntry& E = T.ent[c.f * 32 + x];
                ezz -= E.dJ2;
                cr = E.cre;
                if (MODE != kModeReal) ci = ((s >> x) & 1u) ? -E.cim : E.cim;
            } else {
                const TEntry& E = T.ent[c.f * 32 + x];
                double sg = ((s >> E.k) & 1u) ? 1.0 : -1.0;
                cr = -E.cim * sg;
                ci = (((s >> E.y) & 1u) ? E.cre : -E.cre) * sg;
            }
            int code;
            uint32_t rep = torus_rep_c<TC, TRIV ? kRepOnly : kRepAny>(t, code);
            uint32_t r = rank_tab<uint32_t, uint32_t>(rep, R, T.rtab, (TC::N + 7) / 8);
            if (!TRIV) {
                // explicit FMAs: every variant rounds the same way
                cd2 ph = T.phase[code];
                double nr = fma(cr, ph.re, -(ci * ph.im)), ni = fma(cr, ph.im, ci * ph.re);
                cr = nr; ci = ni;
            }
            B.rk[q] = r; B.cre[q] = cr; B.cim[q] = ci;
        }
    }
}

// The row with K elements in flight per lane: the orbit searches and ranks of
// K elements are computed first, then their K gathers are issued back to
// back, so that a lane waits for memory once per K elements.  The loads are
// unconditional: a load under a lane mask would make the compiler wait for it
// inside the masked block and serialise the K loads.
// jof(r, ok) gives the element of u that holds the target of rank r; for an
// empty slot (ok false: no element left) it gives an element that holds
// zero, and so does every rank outside the sector, so such slots add nothing
// whatever their coefficient.  loadu(j) returns element j of u; g(x, cr, ci)
// accumulates.  Returns the diagonal.
template <class TC, bool TRIV, int MODE, int K, class JOf, class LoadU, class Gf>
ED_HD double torus_row_batched(uint32_t s, const RowTabs& T, double ezz0, const RankInfo& R, JOf&& jof,
                               LoadU&& loadu, Gf&& g) {
    double ezz = ezz0;
    RowCur c{0, 0u, 0.0};
    if (T.nfam > 0) fam_enter<TC, MODE, true>(s, T, c, ezz);
    int left = 0;
    bool more = true;
    while (more) {
        RowBatch<K> B;
        fill_batch<TC, TRIV, MODE, K, false>(s, T, R, c, left, ezz, B);
        more = B.ok[K - 1];
        cd2 xv[K];
#pragma unroll
        for (int q = 0; q < K; q++) xv[q] = loadu(jof(B.rk[q], B.ok[q]));
#pragma unroll
        for (int q = 0; q < K; q++) g(xv[q], B.cre[q], B.cim[q]);
    }
    return ezz;
}

// torus_row_batched with the batches software-pipelined: while the gathers
// of batch b are in flight the lane computes the orbit searches and ranks of
// batch b+1, then accumulates batch b and issues the gathers of b+1.  Each
// wait on memory thus overlaps a batch of integer work of the same lane.
// Elements are accumulated in the same order as torus_row_batched, so the
// result is bit-identical.
template <class TC, bool TRIV, int MODE, int K, class JOf, class LoadU, class Gf>
ED_HD double torus_row_pipelined(uint32_t s, const RowTabs& T, double ezz0, const RankInfo& R, JOf&& jof,
                                 LoadU&& loadu, Gf&& g) {
    double ezz = ezz0;
    RowCur c{0, 0u, 0.0};
    if (T.nfam > 0) fam_enter<TC, MODE, true>(s, T, c, ezz);
    int left = 0;
    RowBatch<K> A, B;
    fill_batch<TC, TRIV, MODE, K, false>(s, T, R, c, left, ezz, A);
    cd2 xa[K];
#pragma unroll
    for (int q = 0; q < K; q++) xa[q] = loadu(jof(A.rk[q], A.ok[q]));
    bool more = A.ok[K - 1];
    while (more) {
        fill_batch<TC, TRIV, MODE, K, false>(s, T, R, c, left, ezz, B);
#pragma unroll
        for (int q = 0; q < K; q++) g(xa[q], A.cre[q], A.cim[q]);
#pragma unroll
        for (int q = 0; q < K; q++) xa[q] = loadu(jof(B.rk[q], B.ok[q]));
#pragma unroll
        for (int q = 0; q < K; q++) {
            A.cre[q] = B.cre[q];
            A.cim[q] = B.cim[q];
        }
        more = B.ok[K - 1];
    }
#pragma unroll
    for (int q = 0; q < K; q++) g(xa[q], A.cre[q], A.cim[q]);
    return ezz;
}

// A row shared by LPR lanes (the split kernels): the n antiparallel elements
// of the row, in the order torus_row_batched visits them, are cut into LPR
// consecutive shares, elements [q n / LPR, (q + 1) n / LPR) to lane q.
// counts holds the antiparallel count of family f in bits [6f, 6f + 6)
// (nfam <= kSplitMaxFam).  The lane accumulates its share through g (K in
// flight) and returns its part of the diagonal: the per-element Delta J / 2
// of its share, plus, for lane 0, ezz0 (less every family's count times
// Delta J / 2 under kModeUniform, as torus_row_batched takes it).  The
// caller adds the lanes' sums as a tree, lane q with q ^ 1, then q ^ 2, ...
constexpr int kSplitMaxFam = 10;

ED_HD int split_count(uint64_t counts, int f) { return (int)((counts >> (6 * f)) & 63u); }

template <class TC>
ED_HD uint64_t split_counts(uint32_t s, const RowTabs& T, int first, int step) {
    uint64_t c = 0;
    for (int f = first; f < T.nfam; f += step) c |= (uint64_t)popc32(fam_mask_c<TC>(s, T.fam[f])) << (6 * f);
    return c;
}

// Lane q's part of the row's diagonal: ezz0 on lane 0 and, in the uniform
// class, the antiparallel counts of families q, q + step, ... times their
// Delta J/2 (the lanes' parts are added with the off-diagonal sums).
template <int MODE>
ED_HD double split_diag(uint64_t counts, const RowTabs& T, int q, int step, double ezz0) {
    double e = q == 0 ? ezz0 : 0.0;
    if (MODE == kModeUniform)
        for (int f = q; f < T.nfam; f += step) e = fma(-(double)split_count(counts, f), T.fco[f].im, e);
    return e;
}

// x without its k lowest set bits (k < popcount(x)): the position of the
// (k+1)-th set bit by halving the window five times.
ED_HD uint32_t drop_low_set_bits(uint32_t x, int k) {
    int p = 0;
#pragma unroll
    for (int w = 16; w >= 1; w >>= 1) {
        const int c = popc32((x >> p) & ((1u << w) - 1u));
        if (c <= k) {
            k -= c;
            p += w;
        }
    }
    return x & ~((1u << p) - 1u);
}

template <class TC, bool TRIV, int MODE, int K, class JOf, class LoadU, class Gf>
ED_HD double torus_row_share(uint32_t s, int q, int lpr_log2, uint64_t counts, const RowTabs& T, double ezz0,
                             const RankInfo& R, JOf&& jof, LoadU&& loadu, Gf&& g) {
    int n = 0;
    for (int f = 0; f < T.nfam; f++) n += split_count(counts, f);
    const int st = (q * n) >> lpr_log2, en = ((q + 1) * n) >> lpr_log2;
    double ezz = split_diag<MODE>(counts, T, q, 1 << lpr_log2, ezz0);
    int left = en - st;
    if (left <= 0) return ezz;
    // the family holding element st, then its (st - cum) earlier members skipped
    RowCur c{0, 0u, 0.0};
    int cum = 0;
    while (cum + split_count(counts, c.f) <= st) cum += split_count(counts, c.f++);
    fam_enter<TC, MODE, false>(s, T, c, ezz);
    if (st > cum) c.cur = drop_low_set_bits(c.cur, st - cum);
    bool more = true;
    while (more) {
        RowBatch<K> B;
        fill_batch<TC, TRIV, MODE, K, true>(s, T, R, c, left, ezz, B);
        more = B.ok[K - 1] && left > 0;
        cd2 xv[K];
#pragma unroll
        for (int h = 0; h < K; h++) xv[h] = loadu(jof(B.rk[h], B.ok[h]));
#pragma unroll
        for (int h = 0; h < K; h++) g(xv[h], B.cre[h], B.cim[h]);
    }
    return ezz;
}

// Element of the gather image (ImageLayout) holding the target of rank r;
// zidx (an element of the zero slot) for an empty slot.
ED_HD uint32_t image_index(const uint16_t* map, int cb, uint32_t zidx, uint32_t r, bool ok) {
    return ok ? (((uint32_t)map[r >> cb] << cb) | (r & ((1u << cb) - 1u))) : zidx;
}

}  // namespace edc

// Shapes compiled with constant geometry (N, row width, rows): the benchmark
// lattices and the small lattices the parity tests use.  Others run the
// kernels that take the shape at run time.
#define ED_TORUS_SHAPES(X) \
    X(32, 4, 8)            \
    X(30, 6, 5)            \
    X(27, 3, 9)            \
    X(24, 24, 1)           \
    X(16, 4, 4)            \
    X(20, 4, 5)            \
    X(18, 6, 3)            \
    X(16, 16, 1)

