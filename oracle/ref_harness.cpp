// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" glue around the reference's own basis + H*v slice, compiled from
// the reference sources where they lie (see oracle/Makefile, target `ref`).
// The result, oracle/_ref/libedref.so, is used by tests/ to pin the oracle
// restatement and by bench.py's cpu_baseline leg; it is never part of the
// product.  No reference source is copied into this repository: this file only
// includes the reference headers and forwards plain arrays to them.
#include "global.h"
#include "Basis_States.h"
#include "Lin_Tables.h"
#include "Momentum_Basis.h"
#include "H_v.h"
#include <cstring>

// Defined in Momentum_Basis.cpp:472-543 but not declared in its header.
int64_t representative_p(int& lx, int& ly, int& L1, int& L2, int64_t& v_id, std::vector<int>& T1,
                         std::vector<int>& T2, double& px, double& py, bool& p_id, int& lid);

namespace {
std::ofstream& devnull() {
    static std::ofstream f("/dev/null");
    return f;
}

struct RefBasis {
    std::vector<int64_t> reps;
    std::vector<double> norm, norm_inv;
    std::vector<int64_t> rep_loc;
    std::vector<std::complex<double>> state_ph;
    std::vector<int64_t> Ia, Ib;
    int bits_right;
    int n_sites;
    int64_t n_p;
    std::vector<int> T1, T2, T3;
    double px, py, pz = 0;
    bool three = false;
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

void unpack_terms(const Terms* t, std::vector<coordinates>& adj, std::vector<triangles>& ch,
                  std::vector<bowties>& bw) {
    for (int k = 0; k < t->nbonds; k++) {
        coordinates c;
        c.x = t->bond_xy[2 * k]; c.y = t->bond_xy[2 * k + 1]; c.J = t->bond_J[k];
        c.expphase = std::complex<double>(t->bond_phase[2 * k], t->bond_phase[2 * k + 1]);
        adj.push_back(c);
    }
    for (int k = 0; k < t->nchiral; k++) {
        triangles tr;
        tr.i = t->chiral_ijk[3 * k]; tr.j = t->chiral_ijk[3 * k + 1]; tr.k = t->chiral_ijk[3 * k + 2];
        tr.expphase = std::complex<double>(t->chiral_phase[2 * k], t->chiral_phase[2 * k + 1]);
        ch.push_back(tr);
    }
    for (int k = 0; k < t->nbowtie; k++) {
        const int32_t* b = t->bowtie + 6 * k;
        bowties bt;
        bt.x = b[0]; bt.y = b[1]; bt.z = b[2]; bt.p = b[3]; bt.q = b[4]; bt.r = b[5];
        bw.push_back(bt);
    }
}
}  // namespace

extern "C" {

int64_t ref_n_choose_k(int n, int k) { return n_choose_k(n, k); }

void ref_constrained_dets(int N, int n, int64_t* out) {
    std::vector<int64_t> d;
    constrained_dets(N, n, d);
    std::memcpy(out, d.data(), d.size() * sizeof(int64_t));
}

void ref_LxLy(const int32_t* T1, const int32_t* T2, int N, int32_t* L1, int32_t* L2) {
    std::vector<int> a(T1, T1 + N), b(T2, T2 + N);
    int l1, l2;
    LxLy(l1, l2, a, b);
    *L1 = l1; *L2 = l2;
}

int64_t ref_representative(int64_t s, const int32_t* T1, const int32_t* T2, int N, double px, double py,
                           int32_t* lx, int32_t* ly, int32_t* pid, int32_t* lid) {
    std::vector<int> a(T1, T1 + N), b(T2, T2 + N);
    int L1, L2;
    LxLy(L1, L2, a, b);
    int ilx, ily, ilid;
    bool p;
    int64_t v = s;
    int64_t r = representative_p(ilx, ily, L1, L2, v, a, b, px, py, p, ilid);
    *lx = ilx; *ly = ily; *pid = p ? 1 : 0; *lid = ilid;
    return r;
}

void* ref_basis_build(int N, int n_ones, const int32_t* T1, const int32_t* T2, double px, double py) {
    RefBasis* B = new RefBasis();
    B->T1.assign(T1, T1 + N);
    B->T2.assign(T2, T2 + N);
    B->px = px; B->py = py;
    B->n_sites = N;
    int nn = n_ones;
    // pConstruct_States stores p_ids in a bit-packed vector<bool> from inside
    // its OpenMP loop (Momentum_Basis.cpp:696,742): concurrent writes to bits
    // of one word race and can drop representatives.  Run it on one thread so
    // that the recorded result is the race-free one.
    int nt = omp_get_max_threads();
    omp_set_num_threads(1);
    pConstruct_States(B->reps, B->norm, B->norm_inv, B->T1, B->T2, nn, B->px, B->py, B->n_p,
                      B->rep_loc, B->state_ph, B->Ia, B->Ib, B->bits_right, devnull());
    omp_set_num_threads(nt);
    return B;
}

// The three-generator pConstruct_States (Momentum_Basis.cpp:781-897), on
// one thread for the same reason as above (:840,860).
void* ref_basis_build3(int N, int n_ones, const int32_t* T1, const int32_t* T2, const int32_t* T3, double px,
                       double py, double pz) {
    RefBasis* B = new RefBasis();
    B->T1.assign(T1, T1 + N);
    B->T2.assign(T2, T2 + N);
    B->T3.assign(T3, T3 + N);
    B->px = px; B->py = py; B->pz = pz;
    B->three = true;
    B->n_sites = N;
    int nn = n_ones;
    int nt = omp_get_max_threads();
    omp_set_num_threads(1);
    pConstruct_States(B->reps, B->norm, B->norm_inv, B->T1, B->T2, B->T3, nn, B->px, B->py, B->pz, B->n_p,
                      B->rep_loc, B->state_ph, B->Ia, B->Ib, B->bits_right, devnull());
    omp_set_num_threads(nt);
    return B;
}

// A RefBasis whose tables are filled from outside (e.g. the GPU library's
// full-space export, which tests show bit-equal to pConstruct_States), so that
// the reference sHv can be timed on sectors whose single-threaded build would
// take too long.  ref_tables_ptrs returns writable pointers to reps, norm,
// norm_inv, rep_loc, state_ph, Ia, Ib.
void* ref_tables_new(int N, const int32_t* T1, const int32_t* T2, double px, double py, int64_t C, int64_t D,
                     int64_t nlin, int bits_right) {
    RefBasis* B = new RefBasis();
    B->T1.assign(T1, T1 + N);
    B->T2.assign(T2, T2 + N);
    B->px = px; B->py = py;
    B->n_sites = N;
    B->n_p = D;
    B->bits_right = bits_right;
    B->reps.resize(D); B->norm.resize(D); B->norm_inv.resize(D);
    B->rep_loc.resize(C); B->state_ph.resize(C);
    B->Ia.resize(nlin); B->Ib.resize(nlin);
    return B;
}

void ref_tables_ptrs(void* h, void** out) {
    RefBasis* B = (RefBasis*)h;
    out[0] = B->reps.data(); out[1] = B->norm.data(); out[2] = B->norm_inv.data();
    out[3] = B->rep_loc.data(); out[4] = B->state_ph.data(); out[5] = B->Ia.data(); out[6] = B->Ib.data();
}

void ref_basis_free(void* h) { delete (RefBasis*)h; }

void ref_basis_info(void* h, int64_t* info /* [4]: D C nIa bits_right */) {
    RefBasis* B = (RefBasis*)h;
    info[0] = (int64_t)B->reps.size();
    info[1] = (int64_t)B->rep_loc.size();
    info[2] = (int64_t)B->Ia.size();
    info[3] = B->bits_right;
}

void ref_basis_copy(void* h, int64_t* reps, double* norm, double* norm_inv, int64_t* rep_loc,
                    double* state_ph, int64_t* Ia, int64_t* Ib) {
    RefBasis* B = (RefBasis*)h;
    if (reps) std::memcpy(reps, B->reps.data(), B->reps.size() * 8);
    if (norm) std::memcpy(norm, B->norm.data(), B->norm.size() * 8);
    if (norm_inv) std::memcpy(norm_inv, B->norm_inv.data(), B->norm_inv.size() * 8);
    if (rep_loc) std::memcpy(rep_loc, B->rep_loc.data(), B->rep_loc.size() * 8);
    if (state_ph) std::memcpy(state_ph, B->state_ph.data(), B->state_ph.size() * 16);
    if (Ia) std::memcpy(Ia, B->Ia.data(), B->Ia.size() * 8);
    if (Ib) std::memcpy(Ib, B->Ib.data(), B->Ib.size() * 8);
}

// Runs the reference sHv (H_v.cpp:192-357) on the first `rows` rows of the
// basis (rows < 0: all rows).  v, w interleaved complex128.
void ref_hv(void* h, const void* terms, const double* v, double* w, int64_t rows) {
    RefBasis* B = (RefBasis*)h;
    const Terms* t = (const Terms*)terms;
    std::vector<coordinates> adj;
    std::vector<triangles> ch;
    std::vector<bowties> bw;
    unpack_terms(t, adj, ch, bw);
    std::vector<int64_t> sub;
    const std::vector<int64_t>* rp = &B->reps;
    if (rows >= 0 && rows < (int64_t)B->reps.size()) {
        sub.assign(B->reps.begin(), B->reps.begin() + rows);
        rp = &sub;
    }
    double lam = t->lambda, Ch = t->Ch, ChCh = t->ChCh;
    if (B->three)   // the three-generator overload (H_v.cpp:360-527)
        sHv((std::complex<double>*)v, (std::complex<double>*)w, *rp, B->norm, B->norm_inv, B->rep_loc,
            B->state_ph, lam, adj, Ch, ch, ChCh, bw, B->n_sites, B->Ia, B->Ib, B->bits_right, B->T1, B->T2,
            B->T3, B->px, B->py, B->pz, devnull());
    else
        sHv((std::complex<double>*)v, (std::complex<double>*)w, *rp, B->norm, B->norm_inv, B->rep_loc,
            B->state_ph, lam, adj, Ch, ch, ChCh, bw, B->n_sites, B->Ia, B->Ib, B->bits_right, B->T1, B->T2,
            B->px, B->py, devnull());
}

// makeH (H_v.cpp:21-190): dense column-major D x D, interleaved complex128.
void ref_make_h(void* h, const void* terms, double* H) {
    RefBasis* B = (RefBasis*)h;
    const Terms* t = (const Terms*)terms;
    std::vector<coordinates> adj;
    std::vector<triangles> ch;
    std::vector<bowties> bw;
    unpack_terms(t, adj, ch, bw);
    double lam = t->lambda, Ch = t->Ch, ChCh = t->ChCh;
    zMatrix M;
    makeH(B->reps, B->norm, B->norm_inv, B->rep_loc, B->state_ph, lam, adj, Ch, ch, ChCh, bw, B->n_sites,
          B->Ia, B->Ib, B->bits_right, B->T1, B->T2, B->px, B->py, devnull(), M);
    std::memcpy(H, M.data(), M.size() * 16);
}

// Unwrap_pState (Momentum_Basis.cpp:978-1037, three generators :901-975) on
// the basis' representatives and norms; returns the number of entries and
// copies at most cap of them (states, complex128 coefficients).
int64_t ref_unwrap(void* h, const double* v, int64_t cap, int64_t* states, double* coeffs) {
    RefBasis* B = (RefBasis*)h;
    std::vector<std::complex<double>> vc(B->reps.size());
    for (size_t i = 0; i < vc.size(); i++) vc[i] = std::complex<double>(v[2 * i], v[2 * i + 1]);
    std::vector<ids> psi = B->three ? Unwrap_pState(vc, B->reps, B->norm, B->px, B->py, B->pz, B->T1, B->T2, B->T3)
                                    : Unwrap_pState(vc, B->reps, B->norm, B->px, B->py, B->T1, B->T2);
    for (int64_t q = 0; q < (int64_t)psi.size() && q < cap; q++) {
        states[q] = psi[q].stateID;
        coeffs[2 * q] = psi[q].coeff.real();
        coeffs[2 * q + 1] = psi[q].coeff.imag();
    }
    return (int64_t)psi.size();
}

}  // extern "C"
