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

ED_HD uint32_t alignbit32(uint32_t hi, uint32_t lo, uint32_t sh) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, sh);
#else
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (sh & 31));
#endif
}

ED_HD int ctz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ffs(x) - 1;
#else
    return __builtin_ctz(x);
#endif
}

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
        uint32_t r = 0;
        int c = 0;
        for (int k = 0; k < R.nbytes; k++) {
            unsigned b = (rep >> (8 * k)) & 255u;
            r += rtab[(R.off[k] + c - R.cmin[k]) * 256 + b];
            c += popc32(b);
        }
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
