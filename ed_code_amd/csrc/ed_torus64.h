// ed_torus64.h -- rotation fast path for two-row lattices up to 40 sites
// (64-bit states).
//
// A cluster whose translation group acts regularly, relabelled so that site
// i + W*j (0 <= i < W, j in {0, 1}) = a^i b^j (site 0), has T1 = rotation by
// one inside each W-site row and T2 = exchange of the two rows: the tilted
// 40-site square cluster (group Z20 x Z2, W = 20) is the case this serves.
// The state's two rows are W-bit fields f0, f1; the 2W translations of a
// state are (rot_a f1, rot_a f0) and (rot_a f0, rot_a f1) read as (high row,
// low row).  Each candidate is formed as the exact double hi * 2^W + lo
// (2W <= 52 bits), so the orbit minimum is a chain of v_min_f64 over 2W
// candidates with 32-bit rotations, instead of 64-bit shift/compare/select.
//
// Ranks reach C(40,20) = 1.4e11: the lower bytes' table entries fit 32 bits
// (their sum is a rank among 32-bit values), the top byte's entries are
// 64-bit.
#pragma once

#include <algorithm>
#include <utility>
#include <vector>

#include "ed_core.h"
#include "ed_torus.h"

namespace edc {

struct TFam64 {
    int32_t type;       // 0 bond, 1 chiral
    int32_t a, q;       // h^-1 = row rotation by a, then row exchange when q != 0
    int32_t pad;
    uint64_t S;         // sites x that carry a term of this family
};

template <int N_, int W_>
struct Rows2C {
    static constexpr int N = N_, W = W_, L1 = W_, L2 = 2;
    static_assert(N_ == 2 * W_ && W_ <= 26, "two rows of at most 26 sites");
    static constexpr uint32_t wmask = (1u << W_) - 1u;
    static constexpr uint64_t nmask = N_ == 64 ? ~0ull : ((1ull << N_) - 1ull);
    static constexpr int nbytes = (N_ + 7) / 8;
};

ED_HD int popc64(uint64_t x) { return popc32((uint32_t)x) + popc32((uint32_t)(x >> 32)); }

// f rotated left by r inside a W-bit row, from the doubled row X = f | f << W
template <class TC>
ED_HD uint32_t rowrot(uint64_t X, int r) {
    return r == 0 ? (uint32_t)X & TC::wmask : (uint32_t)(X >> (TC::W - r)) & TC::wmask;
}

template <class TC>
ED_HD uint64_t rows_rot_rt(uint64_t s, int a, bool swap) {
    const uint32_t f0 = (uint32_t)s & TC::wmask, f1 = (uint32_t)(s >> TC::W) & TC::wmask;
    const uint64_t X0 = (uint64_t)f0 | ((uint64_t)f0 << TC::W), X1 = (uint64_t)f1 | ((uint64_t)f1 << TC::W);
    const uint32_t A = rowrot<TC>(X0, a), B = rowrot<TC>(X1, a);
    return swap ? (((uint64_t)A << TC::W) | B) : (((uint64_t)B << TC::W) | A);
}

// family mask: sites x with s_x != s_{h(x)}, restricted to the family's sites
template <class TC>
ED_HD uint64_t fam_mask64(uint64_t s, const TFam64& F) {
    return (s ^ rows_rot_rt<TC>(s, F.a, F.q != 0)) & F.S;
}

// Orbit minimum of t over the 2W translations; with TRACK, code = lx*2 + ly
// of a translation T1^lx T2^ly taking the minimum to t.
template <class TC, bool TRACK>
ED_HD uint64_t rep64(uint64_t t, int& code) {
    const uint32_t f0 = (uint32_t)t & TC::wmask, f1 = (uint32_t)(t >> TC::W) & TC::wmask;
    const uint64_t X0 = (uint64_t)f0 | ((uint64_t)f0 << TC::W), X1 = (uint64_t)f1 | ((uint64_t)f1 << TC::W);
    constexpr double S = (double)(1ull << TC::W);
    double best = 1e300;
#pragma unroll
    for (int a = 0; a < TC::W; a++) {
        const double A = (double)rowrot<TC>(X0, a), B = (double)rowrot<TC>(X1, a);
        const double k0 = B * S + A, k1 = A * S + B;   // exact: products and sums below 2^52
        best = fmin(best, fmin(k0, k1));
    }
    const uint32_t hi = (uint32_t)(best * (1.0 / S));
    const uint32_t lo = (uint32_t)(best - (double)hi * S);
    if (TRACK) {
        int ba = 0, bb = 0;
        bool found = false;
        for (int a = 0; a < TC::W; a++) {
            const double A = (double)rowrot<TC>(X0, a), B = (double)rowrot<TC>(X1, a);
            if (!found && B * S + A == best) { ba = a; bb = 0; found = true; }
            if (!found && A * S + B == best) { ba = a; bb = 1; found = true; }
        }
        code = ((TC::W - ba) % TC::W) * 2 + (2 - bb) % 2;
    } else {
        code = 0;
    }
    return ((uint64_t)hi << TC::W) | lo;
}

// The same orbit minimum (and, with TRACK, the same code) from a few
// candidate translations.  The minimum's high row is the smallest of the 2W
// row rotations, so it starts with the longest cyclic run of zeros found in
// either row: a translation whose high row starts with fewer leading zeros
// cannot win.  Candidate translations are those whose high row starts with
// at least L = min(kPruneMax, that longest run) zeros -- the start positions
// p of zero runs of length >= L in f1 (no row exchange) or f0 (exchange),
// rotation a = W - 1 - p -- plus the identity; the minimum is taken over
// them as 64-bit keys (high row << 32 | low row, the same order as
// high * 2^W + low).  Random 40-site states keep 1.8 candidates on average
// instead of 40.  TRACK: among equal keys the smallest a, and at equal a no
// exchange first -- the translation rep64 picks.
constexpr int kPruneMax = 5;

template <class TC, bool TRACK>
ED_HD uint64_t rep64p(uint64_t t, int& code) {
    constexpr int W = TC::W;
    const uint32_t f0 = (uint32_t)t & TC::wmask, f1 = (uint32_t)(t >> W) & TC::wmask;
    const uint64_t X0 = (uint64_t)f0 | ((uint64_t)f0 << W), X1 = (uint64_t)f1 | ((uint64_t)f1 << W);
    // R_k[p] = 1 when bits p, p-1, ..., p-k+1 (cyclic) of the row are zero
    const uint32_t z0 = f0 ^ TC::wmask, z1 = f1 ^ TC::wmask;
    const uint64_t Z0 = (uint64_t)z0 | ((uint64_t)z0 << W), Z1 = (uint64_t)z1 | ((uint64_t)z1 << W);
    uint32_t r0 = z0, r1 = z1, m0 = z0, m1 = z1;
#pragma unroll
    for (int k = 1; k < kPruneMax; k++) {
        // bit p of (Z >> (W - k)) is z[p - k]; bits at W and above are cut by r
        r0 &= (uint32_t)(Z0 >> (W - k));
        r1 &= (uint32_t)(Z1 >> (W - k));
        if ((r0 | r1) != 0u) { m0 = r0; m1 = r1; }
    }
    // the identity translation first: key (f1, f0), a = 0, no exchange
    uint64_t best = ((uint64_t)f1 << 32) | f0;
    int bcode = 0;   // a * 2 + exchange
    auto consider = [&](uint32_t mask, bool swap) {
        while (mask) {
            const int p = ctz32(mask);
            mask &= mask - 1u;
            const int a = W - 1 - p;
            const uint32_t A = (uint32_t)(X0 >> (W - a)) & TC::wmask, B = (uint32_t)(X1 >> (W - a)) & TC::wmask;
            const uint64_t key = swap ? (((uint64_t)A << 32) | B) : (((uint64_t)B << 32) | A);
            if (TRACK) {
                const int sc = a * 2 + (swap ? 1 : 0);
                if (key < best || (key == best && sc < bcode)) { best = key; bcode = sc; }
            } else if (key < best) {
                best = key;
            }
        }
    };
    consider(m1, false);
    consider(m0, true);
    if (TRACK) {
        const int ba = bcode >> 1, bb = bcode & 1;
        code = ((W - ba) % W) * 2 + (2 - bb) % 2;
    } else {
        code = 0;
    }
    return ((best >> 32) << W) | (best & 0xffffffffull);
}

// Rank: lower bytes from 32-bit blocks (layout of rank_tables without its
// last block), top byte from its 64-bit block.
template <class TC>
ED_HD uint64_t rank64(uint64_t s, const RankInfo& R, const uint32_t* lo, const uint64_t* hi) {
    uint32_t r = 0;
    int c = 0;
#pragma unroll
    for (int k = 0; k < TC::nbytes - 1; k++) {
        const unsigned b = (unsigned)(s >> (8 * k)) & 255u;
        r += lo[(R.off[k] + c - R.cmin[k]) * 256 + b];
        c += popc32(b);
    }
    return (uint64_t)r + hi[(unsigned)(s >> (8 * (TC::nbytes - 1))) & 255u];
}

// Orders the scheduler not to move instructions across this point (device
// code); a no-op in the host build of the same header.
#if defined(__HIP_DEVICE_COMPILE__)
#define ED_SCHED_BARRIER() __builtin_amdgcn_sched_barrier(0)
#else
#define ED_SCHED_BARRIER() ((void)0)
#endif

// One row with K elements in flight (same scheme as torus_row_batched).
// The element of u holding the target of rank r is jres(r, jget(r)), or -1
// for a rank outside the sector; such elements, and the empty slots, load
// the row's own element (self) and add it with a zero coefficient.  jget
// does the lookup's memory access (a chunk-map or directory entry); the K
// lookups of a batch are all issued before any is used, so they are in
// flight together instead of one after the other.
template <class TC, bool TRIV, int K, class JGet, class JRes, class LoadU, class Gf>
ED_HD double rows2_row(uint64_t s, int nfam, const TFam64* fam, const int* foff, const TEntry* ent, double ezz0,
                       const RankInfo& R, const uint32_t* rlo, const uint64_t* rhi, const cd2* phase, int64_t self,
                       JGet&& jget, JRes&& jres, LoadU&& loadu, Gf&& g) {
    double ezz = ezz0;
    int f = 0;
    uint64_t cur = nfam > 0 ? fam_mask64<TC>(s, fam[0]) : 0ull;
    bool more = true;
    while (more) {
        uint64_t rk[K];
        double cre[K], cim[K];
        bool ok[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            while (cur == 0ull && f < nfam) {
                if (++f < nfam) cur = fam_mask64<TC>(s, fam[f]);
            }
            ok[q] = f < nfam;
            rk[q] = 0; cre[q] = 0; cim[q] = 0;
            if (ok[q]) {
                const int x = ctz64(cur);
                cur &= cur - 1ull;
                const TEntry E = ent[foff[f] + popc64(fam[f].S & ((1ull << x) - 1ull))];
                const uint64_t t = s ^ (1ull << x) ^ (1ull << E.y);
                double cr, ci;
                if (fam[f].type == 0) {
                    ezz -= E.dJ2;
                    cr = E.cre;
                    ci = ((s >> x) & 1ull) ? -E.cim : E.cim;
                } else {
                    const double sg = ((s >> E.k) & 1ull) ? 1.0 : -1.0;
                    cr = -E.cim * sg;
                    ci = (((s >> E.y) & 1ull) ? E.cre : -E.cre) * sg;
                }
                int code;
                const uint64_t rep = rep64p<TC, !TRIV>(t, code);
                rk[q] = rank64<TC>(rep, R, rlo, rhi);
                if (!TRIV) {
                    const cd2 ph = phase[code];
                    const double nr = cr * ph.re - ci * ph.im, ni = cr * ph.im + ci * ph.re;
                    cr = nr;
                    ci = ni;
                }
                cre[q] = cr;
                cim[q] = ci;
            }
        }
        more = ok[K - 1];
        using Tok = decltype(jget(uint64_t(0)));
        Tok tk[K];
#pragma unroll
        for (int q = 0; q < K; q++) tk[q] = jget(rk[q]);
        ED_SCHED_BARRIER();
        bool use[K];
        int64_t j[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            j[q] = jres(rk[q], tk[q]);
            use[q] = ok[q] && j[q] >= 0;
        }
        cd2 xv[K];
#pragma unroll
        for (int q = 0; q < K; q++) xv[q] = loadu(use[q] ? j[q] : self);
#pragma unroll
        for (int q = 0; q < K; q++) g(xv[q], use[q] ? cre[q] : 0.0, use[q] ? cim[q] : 0.0);
    }
    return ezz;
}

// Host: families of a two-row lattice of W-site rows.  A family holds the
// terms whose flip partner is y = h(x) for one translation h (and, for chiral
// terms, the same relation of the z-site to x), at most one per site x; terms
// are numbered densely, family by family in site order (foff[f] = first).
inline void rows2_families(double lam, double Ch, int W, const std::vector<Bond>& bonds, const std::vector<Chiral>& ch,
                           std::vector<TFam64>& fam, std::vector<int>& foff, std::vector<TEntry>& ent, double& ezz0) {
    struct Key { int type, a, q, krel; };
    std::vector<Key> keys;
    std::vector<std::vector<std::pair<int, TEntry>>> terms;
    fam.clear();
    foff.clear();
    ent.clear();
    ezz0 = 0;
    auto rel = [&](int x, int y, int& a_inv, int& q) {
        const int da = ((y % W - x % W) % W + W) % W, db = ((y / W - x / W) % 2 + 2) % 2;
        a_inv = (W - da) % W;
        q = db;
    };
    auto place = [&](int type, int x, int y, int krel, const TEntry& e) {
        int a, q;
        rel(x, y, a, q);
        for (size_t f = 0; f < fam.size(); f++)
            if (keys[f].type == type && keys[f].a == a && keys[f].q == q && keys[f].krel == krel &&
                !((fam[f].S >> x) & 1ull)) {
                fam[f].S |= 1ull << x;
                terms[f].push_back({x, e});
                return;
            }
        keys.push_back({type, a, q, krel});
        fam.push_back(TFam64{type, a, q, 0, 1ull << x});
        terms.push_back({{x, e}});
    };
    for (const Bond& bd : bonds) {
        TEntry e{0.5 * bd.J * bd.re, 0.5 * bd.J * bd.im, 0.5 * lam * bd.J, bd.y, 0};
        ezz0 += 0.25 * lam * bd.J;
        place(0, bd.x, bd.y, -1, e);
    }
    for (const Chiral& c : ch) {
        int a2, q2;
        rel(c.i, c.k, a2, q2);
        TEntry e{0.25 * Ch * c.re, 0.25 * Ch * c.im, 0.0, c.j, c.k};
        place(1, c.i, c.j, a2 * 64 + q2, e);
    }
    for (size_t f = 0; f < fam.size(); f++) {
        foff.push_back((int)ent.size());
        std::sort(terms[f].begin(), terms[f].end(),
                  [](const std::pair<int, TEntry>& p, const std::pair<int, TEntry>& q) { return p.first < q.first; });
        for (auto& t : terms[f]) ent.push_back(t.second);
    }
}

}  // namespace edc

// Compiled two-row shapes (N, W): the tilted 40-site cluster and a 20-site
// cluster for the parity tests.
#define ED_ROWS2_SHAPES(X) \
    X(40, 20)              \
    X(20, 10)
