// ed_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of hiteshjc/ED_Code's momentum/Sz-basis H*v path, used as the
// checker for the HIP library.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library; the product path never
// links, calls or falls back to it.
//
// Every function follows one reference routine; the file:line it restates is
// given in its comment (paths relative to the reference repository root).
//
// Pinning (what checks each function beyond this restatement; all in tests/):
//   or_n_choose_k          Pascal recursion + math.comb
//   or_gosper              ascending order, popcount, count = C(N,n); _ref bit-exact
//   or_lin_tables/or_rank  closed-form combinatorial-number-system rank; _ref bit-exact
//   or_translate/or_LxLy   permutation composition / orders on known lattices
//   or_representative      brute-force orbit minimum; _ref bit-exact (rep, lx, ly, lid)
//   or_representative3 /   three-generator forms: brute-force orbit minimum, Burnside
//   or_basis_build3        dimensions summing to C(N,n); _ref bit-exact
//   or_basis_*             Burnside sector dimensions, sum_k D_k = C(N,n); _ref bit-exact
//   or_hv                  _ref (<=1e-14), Hermiticity, dense brute-force projection
//   or_make_h              _ref bit-exact, numpy eigvalsh spectra, closed forms
//   or_hv_rows             or_hv on the same rows, zero elsewhere
//   or_count_targets       brute-force count over the reference's rep_loc (fixtures)
//   or_tables_new          or_hv on adopted reference tables = the fixtures' w
//   or_hv_rows_search /    the fixtures' w on the same rows (2-D sectors); the
//   or_rows_targets_search full-table oracle on the relabelled tilted cluster
//   or_density_matrix      numpy partial trace of random full-basis vectors, dimer
//                          product states (entropy = number of cut singlets),
//                          the Sz block = the block of the full matrix
//   or_lanczos             closed-form spectra (4-site ring, XY chain free fermions,
//                          Ising limit), numpy eigvalsh on dense H
#include <cstdint>
#include <cstring>
#include <cmath>
#include <complex>
#include <vector>
#include <algorithm>
#include <omp.h>

typedef std::complex<double> cd;

#define OR_PI 3.14159265358979323846  // global.h:19

#define OR_API extern "C"

// ---------------------------------------------------------------- terms ----
// Same memory layout as the product's ed_terms (include/ed_amd.h), declared
// independently so that the oracle shares no code with the product.
typedef struct {
    double lambda;             // XXZ anisotropy Delta ("Jz", ED.cpp)
    int32_t nbonds;
    const int32_t* bond_xy;    // [2*nbonds]
    const double* bond_J;      // [nbonds]
    const double* bond_phase;  // [2*nbonds] (re, im) of exp(i phi)
    double Ch;
    int32_t nchiral;
    const int32_t* chiral_ijk; // [3*nchiral]
    const double* chiral_phase;// [2*nchiral]
    double ChCh;
    int32_t nbowtie;
    const int32_t* bowtie;     // [6*nbowtie] x,y,z,p,q,r
} or_terms;

// ------------------------------------------------------ bit primitives -----
// Basis_States.cpp:8-13 (btest64), :17-18 (ibset64), :22-23 (ibclr64), :26-40 (flip)
static inline bool btest(int64_t a, int pos) { return (a >> pos) & 1; }
static inline int64_t ibset(int64_t a, int pos) { return a | (int64_t(1) << pos); }
static inline int64_t ibclr(int64_t a, int pos) { return a & ~(int64_t(1) << pos); }
static inline int64_t flip2(int64_t v, int i, int j, bool ti, bool tj) {
    v = ti ? ibclr(v, i) : ibset(v, i);
    v = tj ? ibclr(v, j) : ibset(v, j);
    return v;
}

// Simple_Math_Func.cpp:49-55
// exp(re + i im) through the C++ library (the C library's cexp), as the
// reference's walks form their step factors (Momentum_Basis.cpp:929,946,959).
OR_API void or_cexp(double re, double im, double* out) {
    std::complex<double> e = std::exp(std::complex<double>(re, im));
    out[0] = e.real();
    out[1] = e.imag();
}

OR_API int64_t or_n_choose_k(int n, int k) {
    if (k == 0) return 1;
    return (int64_t(n) * or_n_choose_k(n - 1, k - 1)) / k;
}

// Basis_States.cpp:42-73 (constrained_dets) == Lin_Tables.cpp:5-40 (gen_Sz_states):
// Gosper's next-combination walk; all N-bit words of weight n, ascending.
OR_API void or_gosper(int N, int n, int64_t* out) {
    int64_t C = or_n_choose_k(N, n);
    int64_t s = (int64_t)(std::pow(2.0, n)) - 1;
    for (int64_t i = 0; i < C; i++) {
        if (i > 0) {
            int64_t t = (s | (s - 1)) + 1;
            int64_t t2 = t & (-t);
            int64_t t3 = s & (-s);
            int64_t t4 = t2 / t3;
            int64_t t6 = (t4 >> 1) - 1;
            s = t | t6;
        }
        out[i] = s;
    }
}

// Lin_Tables.cpp:57-78.  Ia[r] = 1-based rank of r among `half`-bit words of
// the same weight; Ib[*] = -1 (completed by or_lin_complete).
OR_API void or_lin_tables(int half, int64_t* Ia, int64_t* Ib) {
    int64_t n = int64_t(1) << half;
    for (int w = 0; w <= half; w++) {
        int64_t C = or_n_choose_k(half, w);
        std::vector<int64_t> v(C);
        or_gosper(half, w, v.data());
        int64_t ia = 1;
        for (int64_t j = 0; j < C; j++) { Ia[v[j]] = ia++; Ib[v[j]] = -1; }
    }
    (void)n;
}

// Lin_Tables.cpp:42-54 (split_id): right = id & bits_right, left = id >> floor(N/2).
static inline void split_id(int64_t id, int64_t& left, int64_t& right, int64_t bits_right, int N) {
    right = id & bits_right;
    left = id >> (N / 2);
}

// Momentum_Basis.cpp:4-16 (translateT): bit i moves to bit T[i].
OR_API int64_t or_translate(int64_t s, const int32_t* T, int N) {
    int64_t t = 0;
    for (int i = 0; i < N; i++)
        if (btest(s, i)) t = ibset(t, T[i]);
    return t;
}

// Momentum_Basis.cpp:19-36 (LxLy): cycle length of site 0 under T1 and T2.
OR_API void or_LxLy(const int32_t* T1, const int32_t* T2, int N, int32_t* L1, int32_t* L2) {
    int a = 0, b = 0;
    int64_t t = 1;
    do { t = or_translate(t, T1, N); a++; } while (t != 1);
    t = 1;
    do { t = or_translate(t, T2, N); b++; } while (t != 1);
    *L1 = a; *L2 = b;
}

// Momentum_Basis.cpp:472-543 (representative_p).  Walks T1^x T2^y with the
// T2 count running on across rows; rep = min, (lx,ly) = (L1-tx, L2-ty) at the
// first strict new minimum; lid = number of grid points fixing s; p_id says
// whether sum of exp(-i k.g) over those points is nonzero.
OR_API int64_t or_representative(int64_t v, const int32_t* T1, const int32_t* T2, int N,
                          int L1, int L2, double px, double py,
                          int32_t* lx_out, int32_t* ly_out, int32_t* pid_out, int32_t* lid_out) {
    int lx = 0, ly = 0;
    int64_t rep = v, t = v;
    int tx = 0, ty = 0;
    cd ph(1.0, 0.0), idc(1.0, 0.0);
    int lid = 1;
    for (int x = 0; x < L1; x++) {
        if (x > 0) {
            t = or_translate(t, T1, N); tx++;
            ph = ph * std::exp(cd(0.0, -px));
            if (v == t) { idc = idc + ph; lid++; }
            if (rep > t) { rep = t; lx = L1 - tx; ly = L2 - ty; }
        }
        for (int y = 0; y < L2; y++) {
            if (y > 0) {
                t = or_translate(t, T2, N); ty++;
                ph = ph * std::exp(cd(0.0, -py));
                if (v == t) { idc = idc + ph; lid++; }
                if (rep > t) { rep = t; lx = L1 - tx; ly = L2 - ty; }
            }
        }
    }
    *pid_out = (std::fabs(idc.real()) > 1e-8 || std::fabs(idc.imag()) > 1e-8) ? 1 : 0;
    *lid_out = lid;
    *lx_out = lx; *ly_out = ly;
    return rep;
}

// Momentum_Basis.cpp:38-64 (LxLyLz).
OR_API void or_LxLyLz(const int32_t* T1, const int32_t* T2, const int32_t* T3, int N, int32_t* L1, int32_t* L2,
                      int32_t* L3) {
    or_LxLy(T1, T2, N, L1, L2);
    int c = 0;
    int64_t t = 1;
    do { t = or_translate(t, T3, N); c++; } while (t != 1);
    *L3 = c;
}

// Momentum_Basis.cpp:545-643 (representative_p, three generators): the walk
// nests T3 inside T2 inside T1 with the T2 and T3 counts running on; rep, lid,
// p_id as in the two-generator form, (lx,ly,lz) = (L1-tx, L2-ty, L3-tz) at the
// first strict new minimum.
OR_API int64_t or_representative3(int64_t v, const int32_t* T1, const int32_t* T2, const int32_t* T3, int N,
                                  int L1, int L2, int L3, double px, double py, double pz, int32_t* lx_out,
                                  int32_t* ly_out, int32_t* lz_out, int32_t* pid_out, int32_t* lid_out) {
    int lx = 0, ly = 0, lz = 0;
    int64_t rep = v, t = v;
    int tx = 0, ty = 0, tz = 0;
    cd ph(1.0, 0.0), idc(1.0, 0.0);
    int lid = 1;
    auto visit = [&]() {
        if (v == t) { idc = idc + ph; lid++; }
        if (rep > t) { rep = t; lx = L1 - tx; ly = L2 - ty; lz = L3 - tz; }
    };
    for (int x = 0; x < L1; x++) {
        if (x > 0) {
            t = or_translate(t, T1, N); tx++;
            ph = ph * std::exp(cd(0.0, -px));
            visit();
        }
        for (int y = 0; y < L2; y++) {
            if (y > 0) {
                t = or_translate(t, T2, N); ty++;
                ph = ph * std::exp(cd(0.0, -py));
                visit();
            }
            for (int z = 1; z < L3; z++) {
                t = or_translate(t, T3, N); tz++;
                ph = ph * std::exp(cd(0.0, -pz));
                visit();
            }
        }
    }
    *pid_out = (std::fabs(idc.real()) > 1e-8 || std::fabs(idc.imag()) > 1e-8) ? 1 : 0;
    *lid_out = lid;
    *lx_out = lx; *ly_out = ly; *lz_out = lz;
    return rep;
}

// ---------------------------------------------------------------- basis ----
struct OrBasis {
    int N, n, L1, L2, L3 = 1, half, bits_right;
    double px, py, pz = 0;
    int64_t C, D;
    std::vector<int64_t> reps;     // pblock_states
    std::vector<double> norm, norm_inv;
    std::vector<int64_t> rep_loc;  // [C]
    std::vector<cd> state_ph;      // [C]
    std::vector<int64_t> Ia, Ib;   // [2^half]
    int32_t lin_error;             // Momentum_Basis.cpp:736-737 consistency check
};

// Momentum_Basis.cpp:645-778 (pConstruct_States), 2-D form.
OR_API void* or_basis_build(int N, int n_ones, const int32_t* T1, const int32_t* T2, double px, double py) {
    OrBasis* B = new OrBasis();
    B->N = N; B->n = n_ones; B->px = px; B->py = py;
    B->half = N / 2 + (N % 2);
    or_LxLy(T1, T2, N, &B->L1, &B->L2);
    int L1 = B->L1, L2 = B->L2;
    int64_t C = or_n_choose_k(N, n_ones);
    B->C = C;
    std::vector<int64_t> all(C), repid(C);
    std::vector<char> pids(C);
    std::vector<int> lids(C);
    B->rep_loc.resize(C);
    B->state_ph.resize(C);
    B->Ia.resize(int64_t(1) << B->half);
    B->Ib.resize(int64_t(1) << B->half);
    or_lin_tables(B->half, B->Ia.data(), B->Ib.data());
    B->bits_right = (int)(std::pow(2.0, B->half)) - 1;
    or_gosper(N, n_ones, all.data());
    int err = 0;
    // Ib completion: serial so that the first writer wins deterministically
    // (the reference writes it inside the parallel loop; every writer of a
    // given `left` stores the same value).
    for (int64_t i = 0; i < C; i++) {
        int64_t left, right;
        split_id(all[i], left, right, B->bits_right, N);
        if (B->Ib[left] == -1) B->Ib[left] = i - B->Ia[right] + 1;
        else if (i != B->Ib[left] + B->Ia[right] - 1) err = 1;
    }
    B->lin_error = err;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < C; i++) {
        int32_t lx, ly, pid, lid;
        repid[i] = or_representative(all[i], T1, T2, N, L1, L2, px, py, &lx, &ly, &pid, &lid);
        pids[i] = (char)pid; lids[i] = lid;
        lx = lx % L1; ly = ly % L2;
        if (lx < 0) lx += L1;
        if (ly < 0) ly += L2;
        B->state_ph[i] = std::exp(cd(0.0, -px * lx - py * ly));
    }
    for (int64_t i = 0; i < C; i++) {
        if (repid[i] == all[i] && pids[i]) {
            B->reps.push_back(all[i]);
            int64_t orb = (int64_t)L1 * L2 / lids[i];
            B->norm.push_back(std::sqrt((double)orb));
            B->norm_inv.push_back(1 / std::sqrt((double)orb));
        }
    }
    B->D = (int64_t)B->reps.size();
    // Momentum_Basis.cpp:772-777 with findstate (:220-240): binary search.
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < C; i++) {
        int64_t r = repid[i], j = -1;
        int64_t lo = 0, hi = B->D - 1;
        while (hi >= lo) {
            int64_t m = lo + (hi - lo) / 2;
            if (r < B->reps[m]) hi = m - 1;
            else if (r > B->reps[m]) lo = m + 1;
            else { j = m; break; }
        }
        B->rep_loc[i] = j;
    }
    return B;
}

// Momentum_Basis.cpp:781-897 (pConstruct_States, three generators): as the
// two-generator form with representative_p over L1 x L2 x L3, state_ph =
// exp(i(-px lx - py ly - pz lz)) (:872) and norm = sqrt(L1 L2 L3 / lid) (:883).
OR_API void* or_basis_build3(int N, int n_ones, const int32_t* T1, const int32_t* T2, const int32_t* T3, double px,
                             double py, double pz) {
    OrBasis* B = new OrBasis();
    B->N = N; B->n = n_ones; B->px = px; B->py = py; B->pz = pz;
    B->half = N / 2 + (N % 2);
    or_LxLyLz(T1, T2, T3, N, &B->L1, &B->L2, &B->L3);
    const int L1 = B->L1, L2 = B->L2, L3 = B->L3;
    const int64_t C = or_n_choose_k(N, n_ones);
    B->C = C;
    std::vector<int64_t> all(C), repid(C);
    std::vector<char> pids(C);
    std::vector<int> lids(C);
    B->rep_loc.resize(C);
    B->state_ph.resize(C);
    B->Ia.resize(int64_t(1) << B->half);
    B->Ib.resize(int64_t(1) << B->half);
    or_lin_tables(B->half, B->Ia.data(), B->Ib.data());
    B->bits_right = (int)(std::pow(2.0, B->half)) - 1;
    or_gosper(N, n_ones, all.data());
    int err = 0;
    for (int64_t i = 0; i < C; i++) {
        int64_t left, right;
        split_id(all[i], left, right, B->bits_right, N);
        if (B->Ib[left] == -1) B->Ib[left] = i - B->Ia[right] + 1;
        else if (i != B->Ib[left] + B->Ia[right] - 1) err = 1;
    }
    B->lin_error = err;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < C; i++) {
        int32_t lx, ly, lz, pid, lid;
        repid[i] = or_representative3(all[i], T1, T2, T3, N, L1, L2, L3, px, py, pz, &lx, &ly, &lz, &pid, &lid);
        pids[i] = (char)pid; lids[i] = lid;
        lx = lx % L1; ly = ly % L2; lz = lz % L3;
        if (lx < 0) lx += L1;
        if (ly < 0) ly += L2;
        if (lz < 0) lz += L3;
        B->state_ph[i] = std::exp(cd(0.0, -px * lx - py * ly - pz * lz));
    }
    for (int64_t i = 0; i < C; i++) {
        if (repid[i] == all[i] && pids[i]) {
            B->reps.push_back(all[i]);
            int64_t orb = (int64_t)L1 * L2 * L3 / lids[i];
            B->norm.push_back(std::sqrt((double)orb));
            B->norm_inv.push_back(1 / std::sqrt((double)orb));
        }
    }
    B->D = (int64_t)B->reps.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < C; i++) {
        int64_t r = repid[i], j = -1;
        int64_t lo = 0, hi = B->D - 1;
        while (hi >= lo) {
            int64_t m = lo + (hi - lo) / 2;
            if (r < B->reps[m]) hi = m - 1;
            else if (r > B->reps[m]) lo = m + 1;
            else { j = m; break; }
        }
        B->rep_loc[i] = j;
    }
    return B;
}

OR_API int32_t or_basis_l3(void* h) { return ((OrBasis*)h)->L3; }

OR_API void or_basis_free(void* h) { delete (OrBasis*)h; }

OR_API void or_basis_info(void* h, int64_t* info /* [8]: N n L1 L2 half bits_right C D */) {
    OrBasis* B = (OrBasis*)h;
    info[0] = B->N; info[1] = B->n; info[2] = B->L1; info[3] = B->L2;
    info[4] = B->half; info[5] = B->bits_right; info[6] = B->C; info[7] = B->D;
}

OR_API int32_t or_basis_lin_error(void* h) { return ((OrBasis*)h)->lin_error; }

// Any output pointer may be null.  state_ph is interleaved (re, im).
OR_API void or_basis_copy(void* h, int64_t* reps, double* norm, double* norm_inv,
                   int64_t* rep_loc, double* state_ph, int64_t* Ia, int64_t* Ib) {
    OrBasis* B = (OrBasis*)h;
    if (reps) std::memcpy(reps, B->reps.data(), B->D * sizeof(int64_t));
    if (norm) std::memcpy(norm, B->norm.data(), B->D * sizeof(double));
    if (norm_inv) std::memcpy(norm_inv, B->norm_inv.data(), B->D * sizeof(double));
    if (rep_loc) std::memcpy(rep_loc, B->rep_loc.data(), B->C * sizeof(int64_t));
    if (state_ph) std::memcpy(state_ph, B->state_ph.data(), B->C * 2 * sizeof(double));
    int64_t nh = int64_t(1) << B->half;
    if (Ia) std::memcpy(Ia, B->Ia.data(), nh * sizeof(int64_t));
    if (Ib) std::memcpy(Ib, B->Ib.data(), nh * sizeof(int64_t));
}

// ------------------------------------------------------------------- H*v ---
// One row of H_v.cpp:218-355 (sHv, on-the-fly 2-D) for the state idi with
// norm norm_i.  look(tmp, j, ph) gives rep_loc and state_ph of the flipped
// state tmp; ninv(j) = norm_inv[j].  Calls emit(j, coef, kind) for every
// element in the reference's order: bond terms (kind 0), then the diagonal
// Ezz (kind 3, j = -2, coef = Ezz), then chiral (1) and bowtie (2).
template <class Look, class NInv, class Emit>
static inline void row_terms_g(int64_t idi, double norm_i, const or_terms* t, Look look, NInv ninv, Emit emit) {
    const double Ezzfac = 0.25 * t->lambda;
    double Ezz = 0.0;
    int64_t j;
    cd ph;
    for (int k = 0; k < t->nbonds; k++) {
        int x = t->bond_xy[2 * k], y = t->bond_xy[2 * k + 1];
        double J = t->bond_J[k];
        cd ep(t->bond_phase[2 * k], t->bond_phase[2 * k + 1]);
        if (std::fabs(J) < 1e-16) continue;
        bool tx = btest(idi, x), ty = btest(idi, y);
        if (tx == ty) Ezz += Ezzfac * J;
        else {
            if (tx) ep = std::conj(ep);
            Ezz += -(Ezzfac * J);
            int64_t tmp = flip2(idi, x, y, tx, ty);
            look(tmp, j, ph);
            if (j >= 0) emit(j, ep * J * 0.5 * ph * norm_i * ninv(j), 0);
        }
    }
    emit(int64_t(-2), cd(Ezz, 0.0), 3);
    if (std::fabs(t->Ch) > 1e-16) {
        for (int k = 0; k < t->nchiral; k++) {
            int x = t->chiral_ijk[3 * k], y = t->chiral_ijk[3 * k + 1], z = t->chiral_ijk[3 * k + 2];
            cd ep(t->chiral_phase[2 * k], t->chiral_phase[2 * k + 1]);
            bool tx = btest(idi, x), ty = btest(idi, y);
            if (tx != ty) {
                int64_t tmp = flip2(idi, x, y, tx, ty);
                look(tmp, j, ph);
                if (j >= 0) {
                    cd sg = ty ? cd(0, 1) * ep : cd(0, -1) * std::conj(ep);
                    if (!btest(idi, z)) sg = sg * cd(-1, 0);
                    emit(j, t->Ch * 0.25 * sg * ph * norm_i * ninv(j), 1);
                }
            }
        }
    }
    if (std::fabs(t->ChCh) > 1e-16) {
        for (int k = 0; k < t->nbowtie; k++) {
            const int32_t* b = t->bowtie + 6 * k;
            int x = b[0], y = b[1], z = b[2], p = b[3], q = b[4], r = b[5];
            bool tx = btest(idi, x), ty = btest(idi, y);
            cd s1 = ty ? cd(0, 1) : cd(0, -1);
            if (tx != ty) {
                int64_t tmp = flip2(idi, x, y, tx, ty);
                bool tz = btest(tmp, z), tp = btest(tmp, p), tq = btest(tmp, q);
                if (!tz) s1 = s1 * cd(-1, 0);
                cd s2 = tq ? cd(0, 1) : cd(0, -1);
                if (tp != tq) {
                    int64_t tmp2 = flip2(tmp, p, q, tp, tq);
                    bool tr = btest(tmp2, r);
                    look(tmp2, j, ph);
                    if (j >= 0) {
                        if (!tr) s2 = s2 * cd(-1, 0);
                        emit(j, t->ChCh * 0.25 * 0.25 * s1 * s2 * ph * norm_i * ninv(j), 2);
                    }
                }
            }
        }
    }
}

// Row i of the basis B through its full-space tables: rep_loc / state_ph at
// the Lin-table position Ia[right] + Ib[left] - 1 (H_v.cpp:229-234).  The
// diagonal element is emitted with j = i.
template <class Emit>
static inline void row_terms(const OrBasis* B, const or_terms* t, int64_t i, Emit emit) {
    const int N = B->N;
    auto look = [&](int64_t tmp, int64_t& j, cd& ph) {
        int64_t left, right;
        split_id(tmp, left, right, B->bits_right, N);
        int64_t loc = B->Ia[right] + B->Ib[left] - 1;
        j = B->rep_loc[loc];
        ph = B->state_ph[loc];
    };
    auto ninv = [&](int64_t j) { return B->norm_inv[j]; };
    row_terms_g(B->reps[i], B->norm[i], t, look, ninv,
                [&](int64_t j, cd c, int kind) { emit(kind == 3 ? i : j, c, kind); });
}

// H_v.cpp:192-357 (sHv).  v, w: interleaved complex128 [D].  w is overwritten.
OR_API void or_hv(void* h, const or_terms* t, const double* v_, double* w_) {
    const OrBasis* B = (const OrBasis*)h;
    const cd* v = (const cd*)v_;
    cd* w = (cd*)w_;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < B->D; i++) {
        cd acc(0.0, 0.0);
        row_terms(B, t, i, [&](int64_t j, cd c, int kind) {
            if (kind == 3) acc += v[i] * c.real(); else acc += c * v[j];
        });
        w[i] = acc;
    }
}

// Same as or_hv restricted to rows [r0, r1) (bounded CPU samples).
OR_API void or_hv_rows(void* h, const or_terms* t, const double* v_, double* w_, int64_t r0, int64_t r1) {
    const OrBasis* B = (const OrBasis*)h;
    const cd* v = (const cd*)v_;
    cd* w = (cd*)w_;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = r0; i < r1; i++) {
        cd acc(0.0, 0.0);
        row_terms(B, t, i, [&](int64_t j, cd c, int kind) {
            if (kind == 3) acc += v[i] * c.real(); else acc += c * v[j];
        });
        w[i] = acc;
    }
}

// Number of distinct (row, target) pairs with j >= 0 over all off-diagonal
// terms (the G of the algorithmic-bytes model), and the total element count.
OR_API void or_count_targets(void* h, const or_terms* t, int64_t* distinct, int64_t* total) {
    const OrBasis* B = (const OrBasis*)h;
    int64_t G = 0, T = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : G, T)
    for (int64_t i = 0; i < B->D; i++) {
        std::vector<int64_t> js;
        row_terms(B, t, i, [&](int64_t j, cd, int kind) { if (kind != 3) js.push_back(j); });
        T += (int64_t)js.size();
        std::sort(js.begin(), js.end());
        G += (int64_t)(std::unique(js.begin(), js.end()) - js.begin());
    }
    *distinct = G; *total = T;
}

// H_v.cpp:21-190 (makeH): dense column-major H(i,j) = H[i + j*D], interleaved.
OR_API void or_make_h(void* h, const or_terms* t, double* H_) {
    const OrBasis* B = (const OrBasis*)h;
    const int64_t D = B->D;
    cd* H = (cd*)H_;
    for (int64_t i = 0; i < D * D; i++) H[i] = 0.0;
    for (int64_t i = 0; i < D; i++) {
        row_terms(B, t, i, [&](int64_t j, cd c, int kind) {
            if (kind == 3) H[i + i * D] += c.real(); else H[i + j * D] += c;
        });
    }
}

// ------------------------------------------- tables supplied from outside ---
// An OrBasis over caller-supplied copies of pConstruct_States' outputs (reps,
// norm, norm_inv, rep_loc, state_ph, Ia, Ib, bits_right), so that or_hv,
// or_hv_rows and or_count_targets run on sectors whose oracle build would be
// too slow for a test.  The caller fills the arrays through or_tables_ptrs
// (order: reps, norm, norm_inv, rep_loc, state_ph, Ia, Ib) and pins the
// tables it supplies on its own (or_representative on sampled states).
OR_API void* or_tables_new(int N, int n_ones, const int32_t* T1, const int32_t* T2, double px, double py, int64_t C,
                           int64_t D, int64_t nlin, int bits_right) {
    OrBasis* B = new OrBasis();
    B->N = N; B->n = n_ones; B->px = px; B->py = py;
    B->half = N / 2 + (N % 2);
    or_LxLy(T1, T2, N, &B->L1, &B->L2);
    B->C = C; B->D = D;
    B->bits_right = bits_right;
    B->lin_error = 0;
    B->reps.resize(D); B->norm.resize(D); B->norm_inv.resize(D);
    B->rep_loc.resize(C); B->state_ph.resize(C);
    B->Ia.resize(nlin); B->Ib.resize(nlin);
    return B;
}

OR_API void or_tables_ptrs(void* h, void** out) {
    OrBasis* B = (OrBasis*)h;
    out[0] = B->reps.data(); out[1] = B->norm.data(); out[2] = B->norm_inv.data();
    out[3] = B->rep_loc.data(); out[4] = B->state_ph.data(); out[5] = B->Ia.data(); out[6] = B->Ib.data();
}

// ------------------------------------------------ rows without the tables ---
// H_v.cpp:218-355 for selected rows of a two-generator sector, without the
// full-space tables: each target's rep_loc / state_ph entry is recomputed the
// way pConstruct_States defines it (Momentum_Basis.cpp:645-778):
// representative_p of the target (:472-543; lx, ly reduced mod L1, L2 and
// state_ph = exp(i(-px lx - py ly)), :745-753), findstate of its
// representative in reps (:220-240, -1 when absent, which includes orbits
// whose momentum projection vanishes), and norm = sqrt(L1 L2 / lid) of each
// representative (:764-766; lid is the same for every state of an orbit).
// For sectors too large for the full-space tables (40 sites: C = 1.4e11).
template <class Fn>
static void rows_search(int N, const int32_t* T1, const int32_t* T2, double px, double py, int64_t D,
                        const int64_t* reps, const or_terms* t, int64_t nrows, const int64_t* rows, Fn fn) {
    int32_t L1, L2;
    or_LxLy(T1, T2, N, &L1, &L2);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t q = 0; q < nrows; q++) {
        const int64_t i = rows[q];
        int32_t lx, ly, pid, lid;
        or_representative(reps[i], T1, T2, N, L1, L2, px, py, &lx, &ly, &pid, &lid);
        const double ni = std::sqrt((double)((int64_t)L1 * L2 / lid));
        double last_ninv = 0.0;   // norm_inv of the target looked up last (row_terms_g asks right after)
        auto look = [&](int64_t tmp, int64_t& j, cd& ph) {
            int32_t tx, ty, tpid, tlid;
            int64_t r = or_representative(tmp, T1, T2, N, L1, L2, px, py, &tx, &ty, &tpid, &tlid);
            tx = tx % L1; ty = ty % L2;
            if (tx < 0) tx += L1;
            if (ty < 0) ty += L2;
            ph = std::exp(cd(0.0, -px * tx - py * ty));
            const int64_t* it = std::lower_bound(reps, reps + D, r);
            j = (it != reps + D && *it == r) ? (int64_t)(it - reps) : -1;
            if (j >= 0) last_ninv = 1 / std::sqrt((double)((int64_t)L1 * L2 / tlid));
        };
        auto ninv = [&](int64_t) { return last_ninv; };
        fn(q, i, ni, look, ninv);
    }
}

// The indices that rows[0..nrows) gather from (targets with j >= 0 and the
// rows themselves), sorted and unique: returns their number and writes the
// first cap of them to out.  Lets a caller supply only those entries of v.
OR_API int64_t or_rows_targets_search(int N, const int32_t* T1, const int32_t* T2, double px, double py, int64_t D,
                                      const int64_t* reps, const or_terms* t, int64_t nrows, const int64_t* rows,
                                      int64_t cap, int64_t* out) {
    std::vector<std::vector<int64_t>> per(nrows);
    rows_search(N, T1, T2, px, py, D, reps, t, nrows, rows, [&](int64_t q, int64_t i, double ni, auto& look, auto& ninv) {
        per[q].push_back(i);
        row_terms_g(reps[i], ni, t, look, ninv, [&](int64_t j, cd, int kind) { if (kind != 3) per[q].push_back(j); });
    });
    std::vector<int64_t> all;
    for (auto& p : per) all.insert(all.end(), p.begin(), p.end());
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    for (int64_t k = 0; k < (int64_t)all.size() && k < cap; k++) out[k] = all[k];
    return (int64_t)all.size();
}

// w_out[q] = (H v)[rows[q]], complex128 [nrows].  v is given by its entries
// at the sorted indices vidx[0..nv) (values vval, complex128 [nv]); an entry
// a row needs and vidx lacks makes that row NaN.
OR_API void or_hv_rows_search(int N, const int32_t* T1, const int32_t* T2, double px, double py, int64_t D,
                              const int64_t* reps, const or_terms* t, const int64_t* vidx, const double* vval_,
                              int64_t nv, int64_t nrows, const int64_t* rows, double* w_out) {
    const cd* vval = (const cd*)vval_;
    cd* w = (cd*)w_out;
    auto vat = [&](int64_t j) {
        const int64_t* it = std::lower_bound(vidx, vidx + nv, j);
        return (it != vidx + nv && *it == j) ? vval[it - vidx] : cd(NAN, NAN);
    };
    rows_search(N, T1, T2, px, py, D, reps, t, nrows, rows, [&](int64_t q, int64_t i, double ni, auto& look, auto& ninv) {
        cd acc(0.0, 0.0);
        const cd vi = vat(i);
        row_terms_g(reps[i], ni, t, look, ninv, [&](int64_t j, cd c, int kind) {
            if (kind == 3) acc += vi * c.real(); else acc += c * vat(j);
        });
        w[q] = acc;
    });
}

// ------------------------------------------------------- density matrices ---
// analysis/src/Entropy.cpp:57-82 (set_bits_Cut) and :84-106 (findstate_psi):
// the cut bits of a state replaced by the bits of a cut configuration; binary
// search of a state in an unwrapped vector sorted by state.
static int64_t set_bits_cut(int64_t big, int64_t small, int ncut, const int32_t* cut) {
    for (int i = 0; i < ncut; i++) big = btest(small, i) ? ibset(big, cut[i]) : ibclr(big, cut[i]);
    return big;
}

static int64_t findstate_psi(int64_t s, int64_t n, const int64_t* st) {
    int64_t lo = 0, hi = n - 1;
    while (hi >= lo) {
        int64_t m = lo + (hi - lo) / 2;
        if (s < st[m]) hi = m - 1;
        else if (s > st[m]) lo = m + 1;
        else return m;
    }
    return -1;
}

// Entropy.cpp:170-205 (density_matrixij): for every configuration k of the
// cut and every entry i of psi1, j = the cut bits of state i, the state with
// its cut bits set to k looked up in psi2, dm(j, k) += c1_i conj(c2).  psi1,
// psi2: unwrapped vectors sorted by state (states int64, coefficients
// interleaved complex128).  dm: column-major complex128 [2^ncut x 2^ncut].
// With n_cut_ones >= 0, Entropy.cpp:207-265 (density_matrixij_Sz): only
// entries whose cut has n_cut_ones up spins, k over those configurations in
// ascending order (constrained_dets), rows by the rank of j among them
// (findstate_basis); dm is [C(ncut, n_cut_ones)]^2.
OR_API void or_density_matrix(int64_t n1, const int64_t* st1, const double* c1_, int64_t n2, const int64_t* st2,
                              const double* c2_, int ncut, const int32_t* cut, int n_cut_ones, double* dm_) {
    const cd* c1 = (const cd*)c1_;
    const cd* c2 = (const cd*)c2_;
    cd* dm = (cd*)dm_;
    std::vector<int64_t> basis;
    if (n_cut_ones < 0) {
        for (int64_t k = 0; k < (int64_t(1) << ncut); k++) basis.push_back(k);
    } else {
        std::vector<int64_t> b(or_n_choose_k(ncut, n_cut_ones));
        or_gosper(ncut, n_cut_ones, b.data());
        basis = b;
    }
    const int64_t ndm = (int64_t)basis.size();
    for (int64_t q = 0; q < ndm * ndm; q++) dm[q] = 0.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 0; k < ndm; k++) {
        const int64_t sector_k = basis[k];
        for (int64_t i = 0; i < n1; i++) {
            const int64_t big = st1[i];
            int64_t j = 0;
            int ones = 0;
            for (int l = 0; l < ncut; l++)
                if (btest(big, cut[l])) { j = ibset(j, l); ones++; }
            if (n_cut_ones >= 0 && ones != n_cut_ones) continue;
            const int64_t sk = findstate_psi(set_bits_cut(big, sector_k, ncut, cut), n2, st2);
            if (sk < 0) continue;
            const int64_t row = n_cut_ones < 0 ? j : findstate_psi(j, ndm, basis.data());
            dm[row + k * ndm] += c1[i] * std::conj(c2[sk]);
        }
    }
}

// --------------------------------------------------------------- Lanczos ---
// Counter-based start vector (replaces the unseeded glibc rand() of
// lanczos.cpp:76-79,192): element i = (2u-1) + i(2u'-1) with u, u' from
// splitmix64 of (seed, 2i) and (seed, 2i+1).
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline double u01(uint64_t seed, uint64_t ctr) {
    return (double)(splitmix64(seed * 0x632BE59BD9B4E019ull + ctr) >> 11) * (1.0 / 9007199254740992.0);
}
OR_API void or_start_vector(uint64_t seed, int64_t D, double* v) {
    for (int64_t i = 0; i < D; i++) {
        v[2 * i] = 2.0 * u01(seed, 2 * (uint64_t)i) - 1.0;
        v[2 * i + 1] = 2.0 * u01(seed, 2 * (uint64_t)i + 1) - 1.0;
    }
}

// lanczos.cpp:29-48 (zdotc) with a fixed-order blocked reduction.
static cd zdotc(int64_t n, const cd* a, const cd* b) {
    const int64_t BS = 4096;
    int64_t nb = (n + BS - 1) / BS;
    std::vector<cd> part(nb);
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nb; blk++) {
        double sr = 0, si = 0;
        int64_t e = std::min(n, (blk + 1) * BS);
        for (int64_t i = blk * BS; i < e; i++) {
            cd p = std::conj(a[i]) * b[i];
            sr += p.real(); si += p.imag();
        }
        part[blk] = cd(sr, si);
    }
    double sr = 0, si = 0;
    for (auto& p : part) { sr += p.real(); si += p.imag(); }
    return cd(sr, si);
}
static void zaxpy(int64_t n, cd a, const cd* x, cd* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) y[i] += a * x[i];
}
static void zscal(int64_t n, cd a, cd* x) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) x[i] *= a;
}
static void normalize(int64_t n, cd* v) { zscal(n, 1.0 / std::sqrt(zdotc(n, v, v).real()), v); }
// lanczos.cpp:57-74 (classical Gram-Schmidt against a list)
static void orth_against(int64_t n, cd* v, const std::vector<std::vector<cd>>& list) {
    std::vector<cd> qs;
    for (auto& e : list) qs.push_back(std::conj(zdotc(n, v, e.data())));
    for (size_t i = 0; i < list.size(); i++) zaxpy(n, -qs[i], list[i].data(), v);
}

// Lowest eigenpair of the real symmetric tridiagonal (d, e) by Sturm-sequence
// bisection (eigenvalue) and inverse iteration (vector).  Stands in for the
// zheev_ call of lanczos.cpp:285 (LAPACK is not available here); a different
// algorithm from the product's implicit-QL solver on purpose.
static int sturm_count(const std::vector<double>& d, const std::vector<double>& e, double x) {
    int cnt = 0;
    double q = 1.0;
    for (size_t i = 0; i < d.size(); i++) {
        double off = (i == 0) ? 0.0 : e[i - 1] * e[i - 1];
        q = d[i] - x - (i == 0 ? 0.0 : off / q);
        if (q == 0.0) q = -1e-300;
        if (q < 0) cnt++;
    }
    return cnt;
}
static double tridiag_lowest(const std::vector<double>& d, const std::vector<double>& e, std::vector<double>* vec) {
    int m = (int)d.size();
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < m; i++) {
        double r = (i > 0 ? std::fabs(e[i - 1]) : 0) + (i < m - 1 ? std::fabs(e[i]) : 0);
        lo = std::min(lo, d[i] - r); hi = std::max(hi, d[i] + r);
    }
    for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(lo) + std::fabs(hi)); it++) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(d, e, mid) >= 1) hi = mid; else lo = mid;
    }
    double lam = 0.5 * (lo + hi);
    if (vec) {
        // Inverse iteration.  The shift sits just below lambda_min, so T - sh*I
        // is symmetric positive definite and LDL^T needs no pivoting.
        double sh = lam - 1e-10 * std::max(1.0, std::fabs(lam));
        std::vector<double> x(m, 1.0), y(m), L(m, 0.0), Dg(m);
        Dg[0] = d[0] - sh;
        for (int i = 1; i < m; i++) {
            L[i] = e[i - 1] / Dg[i - 1];
            Dg[i] = (d[i] - sh) - L[i] * e[i - 1];
        }
        for (int pass = 0; pass < 3; pass++) {
            y[0] = x[0];
            for (int i = 1; i < m; i++) y[i] = x[i] - L[i] * y[i - 1];
            for (int i = 0; i < m; i++) y[i] /= Dg[i];
            for (int i = m - 2; i >= 0; i--) y[i] -= L[i + 1] * y[i + 1];
            double nrm = 0;
            for (int i = 0; i < m; i++) nrm += y[i] * y[i];
            nrm = std::sqrt(nrm);
            for (int i = 0; i < m; i++) x[i] = y[i] / nrm;
        }
        *vec = x;
    }
    return lam;
}

// lanczos.cpp:133-349 (lanczos_evecs, 2-D), complex128 throughout.
// evals[num] receives the eigenvalue of the num-th deflated run.  When evecs
// is non-null it receives the Ritz vectors (interleaved, [nev][D]).
// Returns the number of H*v products.
OR_API int64_t or_lanczos(void* h, const or_terms* t, int nev, int iterations, int ncycles,
                   double tol, uint64_t seed, double* evals, double* evecs) {
    const OrBasis* B = (const OrBasis*)h;
    const int64_t D = B->D;
    int how_many = std::min<int64_t>(nev, D);
    std::vector<std::vector<cd>> prev;
    std::vector<cd> w(D), vp(D);
    int64_t hv_calls = 0;
    for (int num = 0; num < how_many; num++) {
        iterations = (int)std::min<int64_t>(iterations, D - num);
        or_start_vector(seed + (uint64_t)num, D, (double*)vp.data());
        double previous = 1000.0;
        double e0 = 0.0;
        for (int cycle = 0; cycle < ncycles; cycle++) {
            orth_against(D, vp.data(), prev);
            normalize(D, vp.data());
            std::fill(w.begin(), w.end(), cd(0, 0));
            std::vector<std::vector<cd>> vs;
            std::vector<double> alphas, betas(1, 0.0);
            std::vector<double> tvec;
            for (int it = 0; it < iterations; it++) {
                or_hv(h, t, (const double*)vp.data(), (double*)w.data());
                hv_calls++;
                if (it != 0) zaxpy(D, -betas.back(), vs[it - 1].data(), w.data());
                cd alpha = std::conj(zdotc(D, w.data(), vp.data()));
                alphas.push_back(alpha.real());
                zaxpy(D, -alpha, vp.data(), w.data());
                vs.push_back(vp);
                double beta = std::sqrt(zdotc(D, w.data(), w.data()).real());
                vp = w;
                zscal(D, 1.0 / beta, vp.data());
                betas.push_back(beta);
                std::fill(w.begin(), w.end(), cd(0, 0));
                orth_against(D, vp.data(), prev);
                if (cycle == ncycles - 1) orth_against(D, vp.data(), vs);
                normalize(D, vp.data());
                std::vector<double> dd(alphas), ee(betas.begin() + 1, betas.end() - 1);
                ee.resize(dd.size() > 0 ? dd.size() - 1 : 0);
                e0 = tridiag_lowest(dd, ee, &tvec);
                if (cycle != 0 && cycle != ncycles - 1 && it > 1) {
                    if (std::fabs(e0 - previous) < tol) { it = iterations; cycle = ncycles - 2; }
                    previous = e0;
                } else if (cycle == ncycles - 1) {
                } else {
                    previous = e0;
                }
            }
            // lanczos.cpp:312-320: Ritz vector from the lowest eigenvector of T
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < D; i++) {
                cd s(0, 0);
                for (size_t k = 0; k < vs.size(); k++) s += vs[k][i] * tvec[k];
                vp[i] = s;
            }
            orth_against(D, vp.data(), prev);
            normalize(D, vp.data());
        }
        prev.push_back(vp);
        evals[num] = e0;
        if (evecs) std::memcpy(evecs + 2 * D * num, vp.data(), D * sizeof(cd));
    }
    return hv_calls;
}

