// sancheck.cpp -- the kernel arithmetic (hostcore.cpp, which compiles
// ed_code_amd/csrc/ed_core.h, ed_torus.h, ed_torus64.h and ed_tridiag.h for
// the CPU) built with -fsanitize=address,undefined into one executable.
// main() drives every code path on small lattices and checks them against
// each other: the rotation search and the compiled-shape batched and
// software-pipelined rows against the generic walk, the two-row 64-bit rows
// (wide image and directory) against the generic walk, rank/unrank round
// trips, the full-space tables against the basis, the tridiagonal QL and
// Jacobi eigensolvers against each other.  Exit status 0 = every check held
// and the sanitizers reported nothing (they abort on the first finding).
#include "hostcore.cpp"

#include <cstdio>
#include <cstdlib>
#include <random>

namespace {

int failures = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "sancheck: FAILED %s\n", what);
        failures++;
    }
}

struct Lat {
    int N;
    std::vector<int32_t> T1, T2;
    std::vector<int32_t> bxy;
    std::vector<double> bJ, bph;
    std::vector<int32_t> cijk;
    std::vector<double> cph;
    double Ch = 0;
    void bond(int x, int y, double J) {
        bxy.push_back(x); bxy.push_back(y); bJ.push_back(J); bph.push_back(1.0); bph.push_back(0.0);
    }
    Terms terms(double lam) const {
        return Terms{lam, (int32_t)bJ.size(), bxy.data(), bJ.data(), bph.data(), Ch, (int32_t)(cijk.size() / 3),
                     cijk.data(), cph.data(), 0.0, 0, nullptr};
    }
};

// w x h torus, site x + w y; T1: x+1, T2: y+1; nearest neighbours, J2
// diagonals, chiral up-triangles (x,y), (x+1,y), (x,y+1) when Ch != 0.
Lat torus(int w, int h, double J2, double Ch) {
    Lat L;
    L.N = w * h;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            L.T1.push_back(((x + 1) % w) + w * y);
            L.T2.push_back(x + w * ((y + 1) % h));
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int s = x + w * y, r = (x + 1) % w + w * y, u = x + w * ((y + 1) % h);
            L.bond(s, r, 1.0);
            L.bond(s, u, 1.0);
            if (J2 != 0) L.bond(s, (x + 1) % w + w * ((y + 1) % h), J2);
            if (Ch != 0) {
                L.cijk.push_back(s); L.cijk.push_back(r); L.cijk.push_back(u);
                L.cph.push_back(1.0); L.cph.push_back(0.0);
            }
        }
    L.Ch = Ch;
    return L;
}

// two rows of W sites: T1 rotates each row, T2 swaps the rows (a ladder)
Lat ladder(int W) {
    Lat L;
    L.N = 2 * W;
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < W; i++) {
            L.T1.push_back((i + 1) % W + W * j);
            L.T2.push_back(i + W * (1 - j));
        }
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < W; i++) {
            L.bond(i + W * j, (i + 1) % W + W * j, 1.0);
            if (j == 0) L.bond(i, i + W, 0.7);
            L.bond(i + W * j, (i + 2) % W + W * (1 - j), 0.3);
        }
    return L;
}

double relerr(const std::vector<double>& a, const std::vector<double>& b) {
    double n = 0, d = 0;
    for (size_t i = 0; i < a.size(); i++) { n += (a[i] - b[i]) * (a[i] - b[i]); d += b[i] * b[i]; }
    return d > 0 ? std::sqrt(n / d) : std::sqrt(n);
}

void run_sector(const Lat& L, int n, int kx, int ky, double lam, bool torus_paths, bool rows2) {
    void* h = hc_build(L.N, n, L.T1.data(), L.T2.data(), kx, ky);
    const int64_t D = hc_dim(h);
    check(D > 0, "non-empty sector");
    Terms t = L.terms(lam);
    hc_set_terms(h, &t);
    std::mt19937_64 rng(D);
    std::normal_distribution<double> nd;
    std::vector<double> v(2 * D), w0(2 * D), w1(2 * D);
    for (auto& x : v) x = nd(rng);
    hc_hv(h, v.data(), w0.data());
    if (torus_paths) {
        check(hc_hv_torus(h, v.data(), w1.data()) == 0 && relerr(w1, w0) < 1e-12, "rotation row vs generic walk");
        check(hc_hv_torus_c(h, v.data(), w1.data()) == 0 && relerr(w1, w0) < 1e-12, "batched row vs generic walk");
        check(hc_hv_torus_p(h, v.data(), w1.data()) == 0 && relerr(w1, w0) < 1e-12, "pipelined row vs generic walk");
        for (uint64_t r = 0; r < 200; r++) {
            uint32_t s = (uint32_t)(rng() & ((1ull << L.N) - 1)), rep = 0;
            int code = 0, gcode = 0;
            uint64_t grep = 0;
            hc_generic_rep(h, s, &grep, &gcode);
            check(hc_torus_rep(h, s, &rep, &code) == 0 && rep == grep && code == gcode, "rotation search = walk");
        }
    }
    if (rows2) {
        check(hc_hv_rows2w(h, v.data(), w1.data(), 1) == 0 && relerr(w1, w0) < 1e-12, "two-row wide image");
        check(hc_hv_rows2w(h, v.data(), w1.data(), 0) == 0 && relerr(w1, w0) < 1e-12, "two-row directory");
    }
    // the full-space tables point every representative at its own row
    const uint64_t C = binom_u64(L.N, n);
    std::vector<int64_t> rep_loc(C), reps(D);
    std::vector<double> ph(2 * C), nrm(D), nrmi(D);
    hc_full_tables(h, rep_loc.data(), ph.data());
    hc_export(h, reps.data(), nrm.data(), nrmi.data());
    void* rk = hc_ranker_new(L.N, n);
    for (int64_t i = 0; i < D; i++) check(rep_loc[hc_rank(rk, (uint64_t)reps[i])] == i, "rep_loc of a representative");
    hc_ranker_free(rk);
    hc_free(h);
}

}  // namespace

int main() {
    run_sector(torus(4, 4, 0.5, 0.0), 8, 1, 2, 0.8, true, false);
    run_sector(torus(4, 4, 0.5, 0.0), 8, 0, 0, 1.0, true, false);
    run_sector(torus(4, 4, 0.0, 0.25), 8, 1, 3, 1.0, true, false);
    run_sector(torus(4, 5, 0.3, 0.0), 10, 2, 1, 0.9, true, false);
    run_sector(ladder(10), 10, 0, 0, 1.0, false, true);
    run_sector(ladder(10), 9, 3, 1, 0.6, false, true);
    // rank / unrank round trips, including the split tables of the two-row path
    for (int N : {20, 27, 32, 40}) {
        const int n = N / 2;
        void* rk = hc_ranker_new(N, n);
        const uint64_t C = binom_u64(N, n);
        std::mt19937_64 rng(N);
        for (int q = 0; q < 20000; q++) {
            const uint64_t k = rng() % C, s = hc_unrank(rk, k);
            check(hc_rank(rk, s) == k, "rank(unrank(k)) == k");
            if (N > 32) check(hc_rank_split(rk, s) == k, "split rank");
            check(hc_rank(rk, hc_gosper_next(s)) == k + 1 || k + 1 == C, "Gosper successor");
        }
        hc_ranker_free(rk);
    }
    // the two eigensolvers on one symmetric tridiagonal matrix
    const int m = 40;
    std::vector<double> d(m), e(m - 1), a((size_t)m * m, 0.0), dq(m), dj(m), zq((size_t)m * m), zj((size_t)m * m);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> ud(-1, 1);
    for (int i = 0; i < m; i++) { d[i] = ud(rng); a[(size_t)i * m + i] = d[i]; }
    for (int i = 0; i + 1 < m; i++) {
        e[i] = ud(rng);
        a[(size_t)i * m + i + 1] = a[(size_t)(i + 1) * m + i] = e[i];
    }
    dq = d;
    check(hc_tridiag(m, dq.data(), e.data(), zq.data()) == 0, "tridiagonal QL converged");
    check(hc_symeig(m, a.data(), dj.data(), zj.data()) == 0, "Jacobi converged");
    double md = 0;
    for (int i = 0; i < m; i++) md = std::max(md, std::fabs(dq[i] - dj[i]));
    check(md < 1e-12, "QL and Jacobi eigenvalues agree");
    if (failures) return 1;
    std::printf("sancheck ok\n");
    return 0;
}
