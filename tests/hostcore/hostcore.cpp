// hostcore.cpp -- CPU build of the kernel arithmetic in ed_code_amd/csrc/ed_core.h,
// for unit tests only (tests/conftest.py compiles it with g++).  It walks the
// same pipeline the device code runs (classify -> directory -> compaction ->
// visit_row) so that the kernels' logic is checked against the oracle without
// a GPU.  It is never loaded by the product.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ed_core.h"
#include "ed_tables.h"
#include "ed_tridiag.h"
#include "ed_torus.h"
#include "ed_torus_host.h"
#include "ed_torus64.h"

using namespace edc;

namespace {
struct HC {
    int N, n, kx, ky, L1, L2;
    double px, py;
    uint64_t C;
    Geo G;
    RankInfo R;
    std::vector<uint64_t> rt, bn, rows;
    std::vector<DirEntry> dir;
    std::vector<cd2> ph;
    std::vector<double> nt, nit;
    std::vector<Bond> b;
    std::vector<Chiral> c;
    std::vector<Bowtie> w;
    double lam = 1, Ch = 0, ChCh = 0;
};

struct Terms {
    double lambda;
    int32_t nbonds;
    const int32_t* bond_xy;
    const double* bond_J;
    const double* bond_phase;
    double Ch;
    int32_t nchiral;
    const int32_t* chiral_ijk;
    const double* chiral_phase;
    double ChCh;
    int32_t nbowtie;
    const int32_t* bowtie;
};

template <typename S>
void build(HC* h) {
    uint64_t ng = (h->C + kDirSpan - 1) / kDirSpan;
    std::vector<uint32_t> words(ng * 2, 0);
    for (uint64_t w = 0; w < ng * 2; w++) {
        uint64_t r0 = w * 32;
        if (r0 >= h->C) break;
        uint64_t s = unrank(r0, h->N, h->n, h->bn.data());
        uint32_t bits = 0;
        for (int b = 0; b < 32 && r0 + b < h->C; b++) {
            if (classify<S>((S)s, h->G)) bits |= 1u << b;
            s = gosper_next(s);
        }
        words[w] = bits;
    }
    h->dir.resize(ng);
    uint64_t base = 0;
    for (uint64_t g = 0; g < ng; g++) {
        h->dir[g] = DirEntry{(uint32_t)base, words[2 * g], words[2 * g + 1]};
        base += popc32(words[2 * g]) + popc32(words[2 * g + 1]);
    }
    h->rows.resize(base);
    for (uint64_t g = 0; g < ng; g++) {
        DirEntry e = h->dir[g];
        if ((e.b0 | e.b1) == 0) continue;
        uint64_t r0 = g * kDirSpan;
        uint64_t s = unrank(r0, h->N, h->n, h->bn.data());
        uint64_t j = e.base;
        for (int b = 0; b < kDirSpan && r0 + b < h->C; b++) {
            uint32_t wd = b < 32 ? e.b0 : e.b1;
            if ((wd >> (b & 31)) & 1u) h->rows[j++] = s | ((uint64_t)classify<S>((S)s, h->G) << kLidShift);
            s = gosper_next(s);
        }
    }
}

template <typename S, bool TRIV>
void hv(HC* h, const cd2* v, cd2* w) {
    TablesView tv{h->rt.data(), h->dir.data(), h->ph.data()};
    TermsView T{h->lam, h->Ch, h->ChCh, (int)h->b.size(), (int)h->c.size(), (int)h->w.size(), h->b.data(), h->c.data(),
                h->w.data()};
    std::vector<cd2> u(h->rows.size());
    for (size_t i = 0; i < u.size(); i++) {
        double f = h->nit[h->rows[i] >> kLidShift];
        u[i] = cd2{v[i].re * f, v[i].im * f};
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)h->rows.size(); i++) {
        S s = (S)(h->rows[i] & kStateMask);
        double nrm = h->nt[h->rows[i] >> kLidShift];
        double ar = 0, ai = 0;
        double ez = visit_row<S, TRIV>(s, h->G, h->R, tv, T, [&](int64_t j, double cr, double ci) {
            cd2 x = u[j];
            ar += cr * x.re - ci * x.im;
            ai += cr * x.im + ci * x.re;
        });
        w[i] = cd2{nrm * ar + ez * v[i].re, nrm * ai + ez * v[i].im};
    }
}
// layout: 0 = torus_row_batched, 1 = torus_row_pipelined, 2/4/8 = the row
// split over that many lanes (torus_row_share), the lanes' sums added in the
// split kernels' tree order (lane q with q ^ 1, then q ^ 2, ...).
template <class TC, bool TRIV, int MODE>
void hv_batched(HC* h, int layout, const RowTabs& T, double ezz0, const ImageLayout& L, const std::vector<cd2>& u,
                const cd2* v, cd2* w) {
    const uint32_t zidx = L.nslot << L.cb;   // the zero slot
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)h->rows.size(); i++) {
        uint32_t s = (uint32_t)(h->rows[i] & kStateMask);
        double nrm = h->nt[h->rows[i] >> kLidShift];
        auto loadu = [&](uint32_t j) { return u.at(j); };
        auto jof = [&](uint32_t r, bool ok) { return image_index(L.map.data(), L.cb, zidx, r, ok); };
        double ar = 0, ai = 0;
        auto acc = [&](const cd2& x, double cr, double ci) {
            ar = std::fma(cr, x.re, ar);
            ai = std::fma(cr, x.im, ai);
            if (MODE != kModeReal && MODE != kModeUniform) {
                ar = std::fma(-ci, x.im, ar);
                ai = std::fma(ci, x.re, ai);
            }
        };
        double ez;
        if (layout == 0) {
            ez = torus_row_batched<TC, TRIV, MODE, 4>(s, T, ezz0, h->R, jof, loadu, acc);
        } else if (layout == 1) {
            ez = torus_row_pipelined<TC, TRIV, MODE, 4>(s, T, ezz0, h->R, jof, loadu, acc);
        } else {
            const int lg = layout == 2 ? 1 : layout == 4 ? 2 : 3, n = 1 << lg;
            const uint64_t counts = split_counts<TC>(s, T, 0, 1);
            double pr[8], pi[8], pe[8];
            for (int q = 0; q < n; q++) {
                ar = ai = 0;
                pe[q] = torus_row_share<TC, TRIV, MODE, 4>(s, q, lg, counts, T, ezz0, h->R, jof, loadu, acc);
                pr[q] = ar;
                pi[q] = ai;
            }
            for (int o = 1; o < n; o <<= 1) {
                double tr[8], ti[8], te[8];
                for (int q = 0; q < n; q++) {
                    tr[q] = pr[q] + pr[q ^ o];
                    ti[q] = pi[q] + pi[q ^ o];
                    te[q] = pe[q] + pe[q ^ o];
                }
                for (int q = 0; q < n; q++) { pr[q] = tr[q]; pi[q] = ti[q]; pe[q] = te[q]; }
            }
            ar = pr[0];
            ai = pi[0];
            ez = pe[0];
        }
        w[i] = cd2{std::fma(nrm, ar, ez * v[i].re), std::fma(nrm, ai, ez * v[i].im)};
    }
}

template <class TC>
void hv_modes(HC* h, int layout, bool triv, int mode, const RowTabs& T, double ezz0, const ImageLayout& L,
              const std::vector<cd2>& u, const cd2* v, cd2* w) {
#define HC_RUN(TRIV, MODE) hv_batched<TC, TRIV, MODE>(h, layout, T, ezz0, L, u, v, w)
    if (mode == kModeMerged) {
        if (triv) HC_RUN(true, kModeMerged);
        else HC_RUN(false, kModeMerged);
    } else if (triv && mode == kModeUniform) HC_RUN(true, kModeUniform);
    else if (triv && mode == kModeReal) HC_RUN(true, kModeReal);
    else if (triv) HC_RUN(true, kModeGeneral);
    else if (mode != kModeGeneral) HC_RUN(false, kModeBonds);
    else HC_RUN(false, kModeGeneral);
#undef HC_RUN
}

// Gather image of the sector as the library chooses it (basis_image_build):
// chunks marked from the directory.
template <typename M>
void host_image(HC* h, uint64_t max_entries, uint64_t max_elems, ImageLayoutT<M>& L) {
    const int cb0 = image_min_bits(h->C, max_entries);
    for (int cb = cb0; cb <= cb0 + 3 && L.cb < 0; cb++) {
        std::vector<uint8_t> t(((h->C - 1) >> cb) + 1, 0);
        for (size_t g = 0; g < h->dir.size(); g++)
            if (h->dir[g].b0 | h->dir[g].b1) t[g >> (cb - 6)] = 1;
        image_layout(h->C, cb, t, max_elems, L);
    }
}

// u gathered from the wide image (the two-row kernels' default) when wide,
// else by row through the directory.
template <class TC, bool TRIV>
void hv_rows2(HC* h, const std::vector<TFam64>& fam, const std::vector<int>& foff, const std::vector<TEntry>& ent,
              double ezz0, const std::vector<uint32_t>& rlo, const uint64_t* rhi, const WideImageLayout& L, bool wide,
              const std::vector<cd2>& u, const cd2* v, cd2* w) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)h->rows.size(); i++) {
        uint64_t s = h->rows[i] & kStateMask;
        double nrm = h->nt[h->rows[i] >> kLidShift];
        double ar = 0, ai = 0;
        auto jof = [&](uint64_t r) -> int64_t {
            if (wide) return (int64_t)(((uint64_t)L.map.at(r >> L.cb) << L.cb) | (r & ((uint64_t(1) << L.cb) - 1)));
            return dir_decode_off(h->dir[r / kDirSpan], (uint32_t)(r % kDirSpan));
        };
        auto jres = [](uint64_t, int64_t j) { return j; };
        double ez = rows2_row<TC, TRIV, 4>(s, (int)fam.size(), fam.data(), foff.data(), ent.data(), ezz0, h->R,
                                            rlo.data(), rhi, h->ph.data(), i, jof, jres,
                                            [&](int64_t j) { return u.at(j); },
                                            [&](const cd2& x, double cr, double ci) {
                                                ar += cr * x.re - ci * x.im;
                                                ai += cr * x.im + ci * x.re;
                                            });
        w[i] = cd2{nrm * ar + ez * v[i].re, nrm * ai + ez * v[i].im};
    }
}

}  // namespace

extern "C" {

// T3 may be null: the identity (one or two generators).
void* hc_build3(int N, int n, const int32_t* T1, const int32_t* T2, const int32_t* T3, int kx, int ky, int kz) {
    HC* h = new HC();
    h->N = N; h->n = n; h->kx = kx; h->ky = ky;
    std::vector<int> a(T1, T1 + N), b(T2, T2 + N), c(N);
    for (int i = 0; i < N; i++) c[i] = T3 ? T3[i] : i;
    h->L1 = cycle_len0(a);
    h->L2 = cycle_len0(b);
    const int L3 = cycle_len0(c);
    h->px = 2 * kx * M_PI / h->L1;
    h->py = 2 * ky * M_PI / h->L2;
    const double pz = 2 * kz * M_PI / L3;
    h->C = binom_u64(N, n);
    std::memset(&h->G, 0, sizeof(h->G));
    h->G.N = N; h->G.L1 = h->L1; h->G.L2 = h->L2; h->G.L3 = L3; h->G.kx = kx; h->G.ky = ky; h->G.kz = kz;
    make_prog(a, h->G.T1);
    make_prog(b, h->G.T2);
    make_prog(c, h->G.T3);
    h->rt = rank_tables(N, n, h->R);
    h->bn = binom_table();
    h->ph = phase_table(h->L1, h->L2, L3, h->px, h->py, pz);
    norm_tables(h->L1, h->L2, L3, h->nt, h->nit);
    if (N <= 32) build<uint32_t>(h); else build<uint64_t>(h);
    return h;
}

void* hc_build(int N, int n, const int32_t* T1, const int32_t* T2, int kx, int ky) {
    return hc_build3(N, n, T1, T2, nullptr, kx, ky, 0);
}

void hc_free(void* p) { delete (HC*)p; }

int64_t hc_dim(void* p) { return (int64_t)((HC*)p)->rows.size(); }

void hc_export(void* p, int64_t* reps, double* norm, double* norm_inv) {
    HC* h = (HC*)p;
    for (size_t i = 0; i < h->rows.size(); i++) {
        int lid = (int)(h->rows[i] >> kLidShift);
        reps[i] = (int64_t)(h->rows[i] & kStateMask);
        norm[i] = h->nt[lid];
        norm_inv[i] = h->nit[lid];
    }
}

// rep_loc / state_ph of every weight-n state in ascending order.
void hc_full_tables(void* p, int64_t* rep_loc, double* state_ph) {
    HC* h = (HC*)p;
    TablesView tv{h->rt.data(), h->dir.data(), h->ph.data()};
    uint64_t s = unrank(0, h->N, h->n, h->bn.data());
    for (uint64_t r = 0; r < h->C; r++) {
        int code;
        uint64_t rep = h->N <= 32 ? (uint64_t)find_rep<uint32_t, true>((uint32_t)s, h->G, code)
                                  : find_rep<uint64_t, true>(s, h->G, code);
        rep_loc[r] = dir_lookup(rank_of<uint64_t>(rep, h->R, tv.rtab), tv.dir);
        state_ph[2 * r] = h->ph[code].re;
        state_ph[2 * r + 1] = h->ph[code].im;
        s = gosper_next(s);
    }
}

// Rank tables alone (no basis enumeration), for closed-form checks at any N.
struct Ranker {
    int N, n;
    RankInfo R;
    std::vector<uint64_t> rt, bn;
};

void* hc_ranker_new(int N, int n) {
    Ranker* r = new Ranker();
    r->N = N; r->n = n;
    r->rt = rank_tables(N, n, r->R);
    r->bn = binom_table();
    return r;
}

void hc_ranker_free(void* p) { delete (Ranker*)p; }

uint64_t hc_rank(void* p, uint64_t s) {
    Ranker* r = (Ranker*)p;
    return rank_of<uint64_t>(s, r->R, r->rt.data());
}

// Rank through the split tables of the two-row path (32-bit lower blocks,
// 64-bit top block), N <= 40.
uint64_t hc_rank_split(void* p, uint64_t s) {
    Ranker* r = (Ranker*)p;
    const size_t top = (size_t)r->R.off[r->R.nbytes - 1] * 256;
    std::vector<uint32_t> lo(r->rt.begin(), r->rt.begin() + top);
    uint32_t acc = 0;
    int c = 0;
    for (int k = 0; k < r->R.nbytes - 1; k++) {
        const unsigned b = (unsigned)(s >> (8 * k)) & 255u;
        acc += lo[(r->R.off[k] + c - r->R.cmin[k]) * 256 + b];
        c += __builtin_popcount(b);
    }
    return (uint64_t)acc + r->rt[top + ((s >> (8 * (r->R.nbytes - 1))) & 255u)];
}

uint64_t hc_unrank(void* p, uint64_t k) {
    Ranker* r = (Ranker*)p;
    return unrank(k, r->N, r->n, r->bn.data());
}

uint64_t hc_gosper_next(uint64_t s) { return gosper_next(s); }

void hc_set_terms(void* p, const Terms* t) {
    HC* h = (HC*)p;
    h->b.clear(); h->c.clear(); h->w.clear();
    h->lam = t->lambda; h->Ch = t->Ch; h->ChCh = t->ChCh;
    for (int k = 0; k < t->nbonds; k++)
        if (std::fabs(t->bond_J[k]) >= 1e-16)
            h->b.push_back(Bond{t->bond_xy[2 * k], t->bond_xy[2 * k + 1], t->bond_J[k], t->bond_phase[2 * k],
                                t->bond_phase[2 * k + 1]});
    if (std::fabs(t->Ch) > 1e-16)
        for (int k = 0; k < t->nchiral; k++)
            h->c.push_back(Chiral{t->chiral_ijk[3 * k], t->chiral_ijk[3 * k + 1], t->chiral_ijk[3 * k + 2], 0,
                                  t->chiral_phase[2 * k], t->chiral_phase[2 * k + 1]});
    if (std::fabs(t->ChCh) > 1e-16)
        for (int k = 0; k < t->nbowtie; k++) {
            const int32_t* q = t->bowtie + 6 * k;
            h->w.push_back(Bowtie{q[0], q[1], q[2], q[3], q[4], q[5]});
        }
}

void hc_hv(void* p, const double* v, double* w) {
    HC* h = (HC*)p;
    bool triv = h->kx == 0 && h->ky == 0;
    if (h->N <= 32) {
        if (triv) hv<uint32_t, true>(h, (const cd2*)v, (cd2*)w); else hv<uint32_t, false>(h, (const cd2*)v, (cd2*)w);
    } else {
        if (triv) hv<uint64_t, true>(h, (const cd2*)v, (cd2*)w); else hv<uint64_t, false>(h, (const cd2*)v, (cd2*)w);
    }
}

// H*v through the two-row 64-bit path (ed_torus64.h) for the compiled
// shapes, gathering from the wide image (wide != 0) or by row through the
// directory; 1 if the lattice is not a compiled two-row shape (w untouched).
int hc_hv_rows2w(void* p, const double* v_, double* w_, int wide) {
    HC* h = (HC*)p;
    const cd2* v = (const cd2*)v_;
    cd2* w = (cd2*)w_;
    if (h->G.L3 != 1) return 1;
    std::vector<int> a(h->N), b(h->N);
    for (int i = 0; i < h->N; i++) {
        uint64_t s1 = apply_T<uint64_t>(1ull << i, h->G.T1), s2 = apply_T<uint64_t>(1ull << i, h->G.T2);
        a[i] = ctz64(s1); b[i] = ctz64(s2);
    }
    const int N = h->N;
    int W = (N % 2 == 0) ? N / 2 : 0;
    for (int i = 0; i < N && W; i++)
        if (a[i] != (i / W) * W + (i % W + 1) % W || b[i] != (i + W) % N) W = 0;
    if (!W || !h->w.empty()) return 1;
    std::vector<TFam64> fam;
    std::vector<int> foff;
    std::vector<TEntry> ent;
    double ezz0;
    rows2_families(h->lam, h->Ch, W, h->b, h->c, fam, foff, ent, ezz0);
    const size_t top = (size_t)h->R.off[h->R.nbytes - 1] * 256;
    std::vector<uint32_t> rlo(h->rt.begin(), h->rt.begin() + top);
    const uint64_t* rhi = h->rt.data() + top;
    WideImageLayout L;
    if (wide) {
        host_image(h, kWideMapMax, wide_image_max_elems((int64_t)h->rows.size()), L);
        if (L.cb < 0) return 1;
    }
    std::vector<cd2> u(wide ? L.elems() : h->rows.size(), cd2{0.0, 0.0});
    for (size_t i = 0; i < h->rows.size(); i++) {
        double f = h->nit[h->rows[i] >> kLidShift];
        size_t e = i;
        if (wide) {
            const uint64_t r = rank_of<uint64_t>(h->rows[i] & kStateMask, h->R, h->rt.data());
            e = ((uint64_t)L.map[r >> L.cb] << L.cb) | (r & ((uint64_t(1) << L.cb) - 1));
        }
        u.at(e) = cd2{v[i].re * f, v[i].im * f};
    }
    const bool triv = h->kx == 0 && h->ky == 0;
#define HC_R2(n, wd)                                                                                           \
    if (N == n && W == wd) {                                                                                   \
        if (triv) hv_rows2<Rows2C<n, wd>, true>(h, fam, foff, ent, ezz0, rlo, rhi, L, wide, u, v, w);          \
        else hv_rows2<Rows2C<n, wd>, false>(h, fam, foff, ent, ezz0, rlo, rhi, L, wide, u, v, w);              \
        return 0;                                                                                              \
    }
    ED_ROWS2_SHAPES(HC_R2)
    HC_R2(12, 6)
    HC_R2(16, 8)
#undef HC_R2
    return 1;
}

int hc_hv_rows2(void* p, const double* v, double* w) { return hc_hv_rows2w(p, v, w, 1); }

// H*v through the rotation fast path (ed_torus.h).  Returns 1 if the
// lattice/terms do not fit the fast path (w untouched).
int hc_hv_torus(void* p, const double* v_, double* w_) {
    HC* h = (HC*)p;
    const cd2* v = (const cd2*)v_;
    cd2* w = (cd2*)w_;
    if (h->N > 32 || !h->w.empty() || h->G.L3 != 1) return 1;
    Torus T;
    std::vector<int> a(h->N), b(h->N);
    // recover the permutations from the translation programs
    for (int i = 0; i < h->N; i++) {
        uint64_t s1 = apply_T<uint64_t>(1ull << i, h->G.T1), s2 = apply_T<uint64_t>(1ull << i, h->G.T2);
        a[i] = ctz64(s1); b[i] = ctz64(s2);
    }
    if (!torus_detect(a, b, T)) return 1;
    std::vector<TFam> fam;
    std::vector<TEntry> ent;
    double ezz0;
    if (!torus_families(h->lam, h->Ch, T, h->b, h->c, fam, ent, ezz0)) return 1;
    std::vector<uint32_t> rt32(h->rt.begin(), h->rt.end());
    std::vector<cd2> u(h->rows.size());
    for (size_t i = 0; i < u.size(); i++) {
        double f = h->nit[h->rows[i] >> kLidShift];
        u[i] = cd2{v[i].re * f, v[i].im * f};
    }
    bool triv = h->kx == 0 && h->ky == 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)h->rows.size(); i++) {
        uint32_t s = (uint32_t)(h->rows[i] & kStateMask);
        double nrm = h->nt[h->rows[i] >> kLidShift];
        double ar = 0, ai = 0;
        auto g = [&](int64_t j, double cr, double ci) {
            cd2 x = u[j];
            ar += cr * x.re - ci * x.im;
            ai += cr * x.im + ci * x.re;
        };
        double ez = triv ? torus_row<true>(s, T, (int)fam.size(), fam.data(), ent.data(), ezz0, h->R, rt32.data(),
                                           h->dir.data(), h->ph.data(), g)
                         : torus_row<false>(s, T, (int)fam.size(), fam.data(), ent.data(), ezz0, h->R, rt32.data(),
                                            h->dir.data(), h->ph.data(), g);
        w[i] = cd2{nrm * ar + ez * v[i].re, nrm * ai + ez * v[i].im};
    }
    return 0;
}

// Representative and phase code of t through the rotation search (TRACK).
int hc_torus_rep(void* p, uint32_t t, uint32_t* rep, int* code) {
    HC* h = (HC*)p;
    Torus T;
    std::vector<int> a(h->N), b(h->N);
    for (int i = 0; i < h->N; i++) {
        a[i] = ctz64(apply_T<uint64_t>(1ull << i, h->G.T1));
        b[i] = ctz64(apply_T<uint64_t>(1ull << i, h->G.T2));
    }
    if (!torus_detect(a, b, T)) return 1;
    *rep = torus_rep<true>(t, T, *code);
    return 0;
}

// H*v through the compiled-shape, batched row function of the GPU kernel.
// Returns 1 if the lattice has none of the compiled shapes.
static int hv_torus_c_impl(void* p, const double* v_, double* w_, int layout) {
    HC* h = (HC*)p;
    const cd2* v = (const cd2*)v_;
    cd2* w = (cd2*)w_;
    if (h->N > 32 || !h->w.empty() || h->G.L3 != 1) return 1;
    Torus T;
    std::vector<int> a(h->N), b(h->N);
    for (int i = 0; i < h->N; i++) {
        a[i] = ctz64(apply_T<uint64_t>(1ull << i, h->G.T1));
        b[i] = ctz64(apply_T<uint64_t>(1ull << i, h->G.T2));
    }
    if (!torus_detect(a, b, T)) return 1;
    std::vector<TFam> fam;
    std::vector<TEntry> ent;
    double ezz0;
    if (!torus_families(h->lam, h->Ch, T, h->b, h->c, fam, ent, ezz0)) return 1;
    std::vector<uint32_t> rt32(h->rt.begin(), h->rt.end());
    // the gather image as the library lays it out (basis_image_build):
    // chunks marked from the directory, u written at the rows' positions
    ImageLayout L;
    host_image(h, kImageMapMax, image_max_elems((int64_t)h->rows.size()), L);
    if (L.cb < 0) return 1;
    std::vector<cd2> u(L.elems(), cd2{0.0, 0.0});
    for (size_t i = 0; i < h->rows.size(); i++) {
        double f = h->nit[h->rows[i] >> kLidShift];
        const uint32_t r = (uint32_t)rank_of<uint32_t>((uint32_t)(h->rows[i] & kStateMask), h->R, h->rt.data());
        u[image_index(L.map.data(), L.cb, 0, r, true)] = cd2{v[i].re * f, v[i].im * f};
    }
    bool triv = h->kx == 0 && h->ky == 0;
    // term class as the library picks it (op_create)
    bool bonds = true, real = true;
    for (const TFam& F : fam) bonds = bonds && F.type == 0;
    for (const TEntry& E : ent) real = real && E.cim == 0.0;
    int mode = !bonds ? kModeGeneral : (real && triv ? kModeReal : kModeBonds);
    std::vector<TChir> chx;
    if (!h->c.empty()) {
        std::vector<TFam> mf;
        std::vector<TEntry> me;
        double mz;
        if (torus_families_merged(h->lam, h->Ch, T, h->b, h->c, mf, me, chx, mz)) {
            fam.swap(mf);
            ent.swap(me);
            mode = kModeMerged;
        } else {
            chx.clear();
        }
    }
    std::vector<uint32_t> fmk;
    std::vector<cd2> fco;
    if (torus_fam_extras(fam, ent, fmk, fco) && mode == kModeReal) mode = kModeUniform;
    if (layout >= 2 && (int)fam.size() > kSplitMaxFam) return 1;
    std::vector<uint32_t> mki;
    std::vector<cd2> mco;
    if (mode == kModeMerged) torus_merged_tables(fam, ent, chx, mki, mco);
    const RowTabs RT{(int)fam.size(), fam.data(), ent.data(), mki.empty() ? nullptr : mki.data(),
                     mco.empty() ? nullptr : mco.data(), fmk.data(), fco.data(), rt32.data(), h->ph.data()};
    int L2 = T.kind == 2 ? 1 : T.L2;
#define HC_SHAPE(n, wd, l2)                                                                                  \
    if (T.N == n && T.w == wd && L2 == l2) {                                                                 \
        hv_modes<TorusC<n, wd, l2>>(h, layout, triv, mode, RT, ezz0, L, u, v, w);                            \
        return 0;                                                                                            \
    }
    HC_SHAPE(16, 4, 4) HC_SHAPE(20, 4, 5) HC_SHAPE(18, 6, 3) HC_SHAPE(16, 16, 1) HC_SHAPE(24, 24, 1)
#undef HC_SHAPE
    return 1;
}

// Term class the rotation fast path takes for this handle's terms (as
// op_create picks it), -1 if the lattice does not take that path.
int hc_torus_mode(void* p) {
    HC* h = (HC*)p;
    if (h->N > 32 || !h->w.empty() || h->G.L3 != 1) return -1;
    Torus T;
    std::vector<int> a(h->N), b(h->N);
    for (int i = 0; i < h->N; i++) {
        a[i] = ctz64(apply_T<uint64_t>(1ull << i, h->G.T1));
        b[i] = ctz64(apply_T<uint64_t>(1ull << i, h->G.T2));
    }
    if (!torus_detect(a, b, T)) return -1;
    std::vector<TFam> fam, mf;
    std::vector<TEntry> ent, me;
    std::vector<TChir> chx;
    double ezz0, mz;
    if (!torus_families(h->lam, h->Ch, T, h->b, h->c, fam, ent, ezz0)) return -1;
    if (!h->c.empty() && torus_families_merged(h->lam, h->Ch, T, h->b, h->c, mf, me, chx, mz)) return kModeMerged;
    bool bonds = true, real = true;
    for (const TFam& F : fam) bonds = bonds && F.type == 0;
    for (const TEntry& E : ent) real = real && E.cim == 0.0;
    const int mode = !bonds ? kModeGeneral : (real && h->kx == 0 && h->ky == 0 ? kModeReal : kModeBonds);
    std::vector<uint32_t> fmk;
    std::vector<cd2> fco;
    return torus_fam_extras(fam, ent, fmk, fco) && mode == kModeReal ? kModeUniform : mode;
}

int hc_hv_torus_c(void* p, const double* v, double* w) { return hv_torus_c_impl(p, v, w, 0); }

// Same through the software-pipelined row (torus_row_pipelined).
int hc_hv_torus_p(void* p, const double* v, double* w) { return hv_torus_c_impl(p, v, w, 1); }

// Same with every row split over `lanes` (2, 4 or 8) lanes as the split
// kernels do (torus_row_share, tree sum).
int hc_hv_torus_s(void* p, const double* v, double* w, int lanes) {
    if (lanes != 2 && lanes != 4 && lanes != 8) return 1;
    return hv_torus_c_impl(p, v, w, lanes);
}

// Compile-time-shape search (TorusC) for the shapes the tests use; returns 1
// if the lattice has none of them.
int hc_torus_rep_c_mode(void* p, uint32_t t, uint32_t* rep, int* code, int mode) {
    HC* h = (HC*)p;
    Torus T;
    std::vector<int> a(h->N), b(h->N);
    for (int i = 0; i < h->N; i++) {
        a[i] = ctz64(apply_T<uint64_t>(1ull << i, h->G.T1));
        b[i] = ctz64(apply_T<uint64_t>(1ull << i, h->G.T2));
    }
    if (!torus_detect(a, b, T)) return 1;
    int L2 = T.kind == 2 ? 1 : T.L2;
#define HC_SHAPE(n, wd, l2)                                                                                \
    if (T.N == n && T.w == wd && L2 == l2) {                                                               \
        *rep = mode == kRepAny ? torus_rep_c<TorusC<n, wd, l2>, kRepAny>(t, *code)                         \
                               : torus_rep_c<TorusC<n, wd, l2>, kRepFirst>(t, *code);                      \
        return 0;                                                                                          \
    }
    HC_SHAPE(32, 4, 8) HC_SHAPE(30, 6, 5) HC_SHAPE(27, 3, 9) HC_SHAPE(24, 24, 1)
    HC_SHAPE(16, 4, 4) HC_SHAPE(20, 4, 5) HC_SHAPE(18, 6, 3) HC_SHAPE(16, 16, 1)
#undef HC_SHAPE
    return 1;
}

// The representative and the reference's phase code (first minimum), or
// the code of some minimising translation (the kernels' search).
int hc_torus_rep_c(void* p, uint32_t t, uint32_t* rep, int* code) { return hc_torus_rep_c_mode(p, t, rep, code, kRepFirst); }
int hc_torus_rep_c_any(void* p, uint32_t t, uint32_t* rep, int* code) {
    return hc_torus_rep_c_mode(p, t, rep, code, kRepAny);
}
// The sector's phase for a translation code.
void hc_phase(void* p, int code, double* re, double* im) {
    HC* h = (HC*)p;
    *re = h->ph[code].re;
    *im = h->ph[code].im;
}

// Same through the generic walk of ed_core.h.
void hc_generic_rep(void* p, uint64_t t, uint64_t* rep, int* code) {
    HC* h = (HC*)p;
    *rep = find_rep<uint64_t, true>(t, h->G, *code);
}

// Eigenpairs of a dense symmetric matrix (row-major a[m*m]) with the
// library's Jacobi solver: d <- ascending eigenvalues, z <- column-major vectors.
// rep64p (pruned candidates) against rep64 (all 2W translations): the
// representative and, with TRACK, the translation code, on every state of
// the 20-site two-row cluster and on n random and structured 40-site
// states; returns the number of disagreements.
int hc_rep64_check(int64_t n, uint64_t seed) {
    int bad = 0;
    auto one = [&](auto tc, uint64_t t) {
        using TC = decltype(tc);
        int c0 = 0, c1 = 0;
        const uint64_t a = rep64<TC, true>(t, c0), b = rep64p<TC, true>(t, c1);
        int d0 = 0, d1 = 0;
        const uint64_t e = rep64<TC, false>(t, d0), f = rep64p<TC, false>(t, d1);
        if (a != b || c0 != c1 || e != f || a != e) bad++;
    };
    for (uint64_t t = 0; t < (1ull << 20); t++) one(Rows2C<20, 10>{}, t);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        return x;
    };
    const uint64_t m40 = (1ull << 40) - 1;
    for (int64_t i = 0; i < n; i++) {
        uint64_t t = next() & m40;
        switch (i % 4) {
            case 1: t &= next() & next(); break;                                       // sparse: long zero runs
            case 2: t |= next() | next(); t &= m40; break;                              // dense: short zero runs
            case 3: {                                                                   // periodic rows
                const int per = (int)(next() % 5) + 1;
                const int P[5] = {1, 2, 4, 5, 10};
                const int q = P[per - 1];
                uint64_t r0 = next() & ((1ull << q) - 1), r1 = next() & ((1ull << q) - 1), f0 = 0, f1 = 0;
                for (int j = 0; j < 20; j += q) { f0 |= r0 << j; f1 |= r1 << j; }
                t = (f0 & 0xfffffull) | ((next() & 1 ? f1 : f0) & 0xfffffull) << 20;
                break;
            }
            default: break;
        }
        one(Rows2C<40, 20>{}, t);
    }
    one(Rows2C<40, 20>{}, 0);
    one(Rows2C<40, 20>{}, m40);
    one(Rows2C<40, 20>{}, 0xfffffull);
    one(Rows2C<40, 20>{}, 0xfffffull << 20);
    return bad;
}

int hc_symeig(int m, const double* a, double* d, double* z) {
    std::vector<double> A(a, a + (size_t)m * m), dd, zz;
    bool ok = symeig_jacobi(A, m, dd, zz);
    std::memcpy(d, dd.data(), m * sizeof(double));
    std::memcpy(z, zz.data(), (size_t)m * m * sizeof(double));
    return ok ? 0 : 1;
}

// Eigenpairs of the symmetric tridiagonal (d[m], e[m-1]) with the library's QL
// solver: d <- ascending eigenvalues, z <- column-major eigenvectors.
int hc_tridiag(int m, double* d, const double* e, double* z) {
    std::vector<double> dd(d, d + m), ee(e, e + (m > 0 ? m - 1 : 0)), zz;
    bool ok = tridiag_ql(dd, ee, z ? &zz : nullptr);
    std::memcpy(d, dd.data(), m * sizeof(double));
    if (z) std::memcpy(z, zz.data(), (size_t)m * m * sizeof(double));
    return ok ? 0 : 1;
}

}  // extern "C"
