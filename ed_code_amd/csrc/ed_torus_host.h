// ed_torus_host.h -- host-side detection of rotational lattices and grouping of
// the Hamiltonian terms into families for the rotation fast path (ed_torus.h).
// Plain C++: used by the HIP library and by the CPU unit-test build.
#pragma once
#include <cstring>
#include <vector>

#include "ed_core.h"
#include "ed_torus.h"

namespace edc {

inline bool torus_detect(const std::vector<int>& T1, const std::vector<int>& T2, Torus& T) {
    std::memset(&T, 0, sizeof(T));
    const int N = (int)T1.size();
    if (N > 32 || N < 2) return false;
    T.N = N;
    T.nmask = N == 32 ? 0xffffffffu : ((1u << N) - 1u);
    bool t2id = true, t1ring = true;
    for (int i = 0; i < N; i++) {
        if (T2[i] != i) t2id = false;
        if (T1[i] != (i + 1) % N) t1ring = false;
    }
    if (t2id && t1ring) {
        T.kind = 2; T.w = N; T.L1 = N; T.L2 = 1;
        return true;
    }
    for (int w = 2; w <= N; w++) {
        if (N % w) continue;
        bool ok = true;
        for (int i = 0; i < N && ok; i++) {
            if (T1[i] != (i / w) * w + (i % w + 1) % w) ok = false;
            if (T2[i] != (i + w) % N) ok = false;
        }
        if (!ok) continue;
        T.kind = 1; T.w = w; T.L1 = w; T.L2 = N / w;
        for (int r = 0; r < w; r++) {
            uint32_t h = 0;
            for (int i = 0; i < N; i++)
                if (i % w >= r) h |= 1u << i;
            T.hi[r] = h;
        }
        return T.L2 >= 1;
    }
    return false;
}

// group element (a, b) = T1^a T2^b with h(x) = y, as (field, word) rotations of h^-1
inline void rel_element(const Torus& T, int x, int y, int& a_inv, int& q_inv) {
    if (T.kind == 2) {
        int a = ((y - x) % T.N + T.N) % T.N;
        a_inv = 0;
        q_inv = (T.N - a) % T.N;
        return;
    }
    int w = T.w, L2 = T.L2;
    int a = ((y % w - x % w) % w + w) % w;
    int b = ((y / w - x / w) % L2 + L2) % L2;
    a_inv = (w - a) % w;
    q_inv = ((L2 - b) % L2) * w;
}

// Groups the terms into families; false if they do not fit the fast path.
inline bool torus_families(double lam, double Ch, const Torus& T, const std::vector<Bond>& bonds, const std::vector<Chiral>& ch,
                    std::vector<TFam>& fam, std::vector<TEntry>& ent, double& ezz0) {
    fam.clear();
    ent.assign(kMaxFam * 32, TEntry{0, 0, 0, 0, 0});
    ezz0 = 0;
    std::vector<int> keyk;   // chiral: relation of k to x (encoded), -1 for bonds
    auto place = [&](int type, int x, int y, int krel, const TEntry& e) -> bool {
        int a, q;
        rel_element(T, x, y, a, q);
        for (size_t f = 0; f < fam.size(); f++)
            if (fam[f].type == type && fam[f].a == a && fam[f].q == q && keyk[f] == krel && !((fam[f].S >> x) & 1u)) {
                fam[f].S |= 1u << x;
                ent[f * 32 + x] = e;
                return true;
            }
        if ((int)fam.size() >= kMaxFam) return false;
        fam.push_back(TFam{type, a, q, 1u << x});
        keyk.push_back(krel);
        ent[(fam.size() - 1) * 32 + x] = e;
        return true;
    };
    for (const Bond& b : bonds) {
        TEntry e{0.5 * b.J * b.re, 0.5 * b.J * b.im, 0.5 * lam * b.J, b.y, 0};
        ezz0 += 0.25 * lam * b.J;
        if (!place(0, b.x, b.y, -1, e)) return false;
    }
    for (const Chiral& c : ch) {
        int a2, q2;
        rel_element(T, c.i, c.k, a2, q2);
        TEntry e{0.25 * Ch * c.re, 0.25 * Ch * c.im, 0.0, c.j, c.k};
        if (!place(1, c.i, c.j, a2 * 64 + q2, e)) return false;
    }
    return true;
}

// Flip mask (1 << x) | (1 << y) of every family member ([f * 32 + x]) and
// each family's {J/2, Delta J/2}; true when every family is a bond family
// whose members all carry the same real J/2 and Delta J/2 (kModeUniform).
inline bool torus_fam_extras(const std::vector<TFam>& fam, const std::vector<TEntry>& ent,
                             std::vector<uint32_t>& fmk, std::vector<cd2>& fco) {
    fmk.assign(kMaxFam * 32, 0u);
    fco.assign(fam.size(), cd2{0.0, 0.0});
    bool uniform = true;
    for (size_t f = 0; f < fam.size(); f++) {
        bool first = true;
        for (int x = 0; x < 32; x++) {
            if (!((fam[f].S >> x) & 1u)) continue;
            const TEntry& E = ent[f * 32 + x];
            fmk[f * 32 + x] = (1u << x) | (1u << E.y);
            if (fam[f].type != 0 || E.cim != 0.0) uniform = false;
            if (first) {
                fco[f] = cd2{E.cre, E.dJ2};
                first = false;
            } else if (E.cre != fco[f].re || E.dJ2 != fco[f].im) {
                uniform = false;
            }
        }
    }
    return uniform;
}

// kModeMerged coefficient tables: a member (f, x) flipping the pair (x, y)
// carries its bond term and up to two chiral entries with z-sites k0, k1
// (TChir), so its coefficient is a function of the bits s_x, s_k0, s_k1 alone
// (s_y = 1 - s_x for an antiparallel pair).  Flipping s_x conjugates it: the
// real parts (J/2 cos phi, -cim sg) do not depend on s_x and every imaginary
// term changes sign, exactly in IEEE arithmetic.  mco[(f * 32 + x) * 4 + idx],
// idx = s_k0 | s_k1 << 1, holds the coefficient at s_x = 0, computed with the
// operations and order of the per-element formula (the bond part, then each
// chiral entry's -cim sg and +-cre sg); mki[f * 32 + x] = k0 | k1 << 8, an
// absent entry's z-site set to x.
// The four coefficients of one merged member (see torus_merged_tables).
inline void merged_coef4(const TEntry& E, const TChir& X, cd2 out[4]) {
    for (int idx = 0; idx < 4; idx++) {
        const int sx = 0, sk[2] = {idx & 1, (idx >> 1) & 1};
        double cr = E.cre, ci = sx ? -E.cim : E.cim;
        for (int h = 0; h < 2; h++) {
            if (!((X.flags >> (2 * h)) & 1)) continue;
            const double sg = sk[h] ? 1.0 : -1.0;
            const int jbit = ((X.flags >> (2 * h + 1)) & 1) ? sx : 1 - sx;
            cr += -X.cim[h] * sg;
            ci += (jbit ? X.cre[h] : -X.cre[h]) * sg;
        }
        out[idx] = cd2{cr, ci};
    }
}

inline void torus_merged_tables(const std::vector<TFam>& fam, const std::vector<TEntry>& ent,
                                const std::vector<TChir>& chx, std::vector<uint32_t>& mki, std::vector<cd2>& mco) {
    mki.assign(kMaxFam * 32, 0u);
    mco.assign((size_t)kMaxFam * 32 * 4, cd2{0.0, 0.0});
    for (size_t f = 0; f < fam.size(); f++)
        for (int x = 0; x < 32; x++) {
            if (!((fam[f].S >> x) & 1u)) continue;
            const size_t m = f * 32 + x;
            const TChir& X = chx[m];
            const int k0 = (X.flags & 1) ? X.k[0] : x, k1 = (X.flags & 4) ? X.k[1] : x;
            mki[m] = (uint32_t)k0 | ((uint32_t)k1 << 8);
            merged_coef4(ent[m], X, &mco[m * 4]);
        }
}

// Families for kModeMerged: keyed by the bonds (pair x -> y = h(x)); each
// chiral entry (i, j, k) joins the member whose pair is {i, j} (listed either
// way), or opens a member with a zero bond part if no bond flips that pair.
// False if a pair carries more than two chiral entries or the families do
// not fit.
inline bool torus_families_merged(double lam, double Ch, const Torus& T, const std::vector<Bond>& bonds,
                                  const std::vector<Chiral>& ch, std::vector<TFam>& fam, std::vector<TEntry>& ent,
                                  std::vector<TChir>& chx, double& ezz0) {
    fam.clear();
    ent.assign(kMaxFam * 32, TEntry{0, 0, 0, 0, 0});
    TChir none;
    std::memset(&none, 0, sizeof(none));
    chx.assign(kMaxFam * 32, none);
    ezz0 = 0;
    auto find = [&](int x, int y) -> int {   // family whose member x flips (x, y)
        int a, q;
        rel_element(T, x, y, a, q);
        for (size_t f = 0; f < fam.size(); f++)
            if (fam[f].a == a && fam[f].q == q && ((fam[f].S >> x) & 1u) && ent[f * 32 + x].y == y) return (int)f;
        return -1;
    };
    auto add = [&](int x, int y) -> int {    // new member x of a family of pair (x, y)'s key
        int a, q;
        rel_element(T, x, y, a, q);
        for (size_t f = 0; f < fam.size(); f++)
            if (fam[f].a == a && fam[f].q == q && !((fam[f].S >> x) & 1u)) {
                fam[f].S |= 1u << x;
                ent[f * 32 + x].y = y;
                return (int)f;
            }
        if ((int)fam.size() >= kMaxFam) return -1;
        fam.push_back(TFam{0, a, q, 1u << x});
        ent[(fam.size() - 1) * 32 + x].y = y;
        return (int)fam.size() - 1;
    };
    for (const Bond& b : bonds) {
        const int f = add(b.x, b.y);
        if (f < 0) return false;
        ent[f * 32 + b.x] = TEntry{0.5 * b.J * b.re, 0.5 * b.J * b.im, 0.5 * lam * b.J, b.y, 0};
        ezz0 += 0.25 * lam * b.J;
    }
    for (const Chiral& c : ch) {
        int x = c.i, f = find(c.i, c.j);
        bool rev = false;
        if (f < 0 && (f = find(c.j, c.i)) >= 0) {
            rev = true;
            x = c.j;
        }
        if (f < 0 && (f = add(c.i, c.j)) < 0) return false;
        TChir& X = chx[f * 32 + x];
        const int slot = !(X.flags & 1) ? 0 : !(X.flags & 4) ? 1 : -1;
        if (slot < 0) return false;
        X.cre[slot] = 0.25 * Ch * c.re;
        X.cim[slot] = 0.25 * Ch * c.im;
        X.k[slot] = c.k;
        X.flags |= (1 << (2 * slot)) | (rev ? 1 << (2 * slot + 1) : 0);
    }
    return true;
}

}  // namespace edc
