// ed_refapi.cpp -- the reference's C++ sector interface (include/ed_refapi.h)
// on top of the C ABI (include/ed_amd.h).
//
// Sectors are built on the GPU and cached by (T1, T2, n_ones, kx, ky); the
// operator is cached by its term lists, so a reference Lanczos loop that calls
// sHv once per iteration pays one host<->device copy of v and w per call and
// nothing else.  Fatal conditions keep the reference's behaviour: a message to
// the log, then exit(1) (Lanczos_sector.cpp:51-55,93-97).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <memory>
#include <vector>

#include "ed_amd.h"
#include "ed_refapi.h"


namespace {

using cplx = std::complex<double>;

[[noreturn]] void fatal(std::ofstream& out, const std::string& what) {
    out << "\n" << what << "\n";
    out.flush();
    std::exit(1);
}

void check(int rc, std::ofstream& out, const char* where) {
    if (rc != ED_OK) fatal(out, std::string("ED GPU error in ") + where + ": " + ed_last_error());
}

int k_from_p(double p, int L) {
    // px = 2*pi*kx/L1 (Lanczos_sector.cpp:57); recover the integer
    long k = std::lround(p * L / (2.0 * M_PI));
    return (int)(((k % L) + L) % L);
}

struct TermsCopy {
    double lambda = 0, Ch = 0, ChCh = 0;
    std::vector<int32_t> bxy, cijk, bw;
    std::vector<double> bJ, bph, cph;
    bool operator==(const TermsCopy& o) const {
        return lambda == o.lambda && Ch == o.Ch && ChCh == o.ChCh && bxy == o.bxy && cijk == o.cijk && bw == o.bw &&
               bJ == o.bJ && bph == o.bph && cph == o.cph;
    }
    ed_terms view() const {
        ed_terms t;
        t.lambda = lambda;
        t.nbonds = (int32_t)bJ.size();
        t.bond_xy = bxy.data(); t.bond_J = bJ.data(); t.bond_phase = bph.data();
        t.Ch = Ch;
        t.nchiral = (int32_t)(cijk.size() / 3);
        t.chiral_ijk = cijk.data(); t.chiral_phase = cph.data();
        t.ChCh = ChCh;
        t.nbowtie = (int32_t)(bw.size() / 6);
        t.bowtie = bw.data();
        return t;
    }
};

TermsCopy terms_of(double lambda, const std::vector<coordinates>& adj, double Ch, const std::vector<triangles>& ch,
                   double ChCh, const std::vector<bowties>& chch) {
    TermsCopy t;
    t.lambda = lambda; t.Ch = Ch; t.ChCh = ChCh;
    for (const auto& c : adj) {
        t.bxy.push_back(c.x); t.bxy.push_back(c.y);
        t.bJ.push_back(c.J);
        t.bph.push_back(c.expphase.real()); t.bph.push_back(c.expphase.imag());
    }
    for (const auto& c : ch) {
        t.cijk.push_back(c.i); t.cijk.push_back(c.j); t.cijk.push_back(c.k);
        t.cph.push_back(c.expphase.real()); t.cph.push_back(c.expphase.imag());
    }
    for (const auto& b : chch) {
        int32_t q[6] = {b.x, b.y, b.z, b.p, b.q, b.r};
        t.bw.insert(t.bw.end(), q, q + 6);
    }
    return t;
}

struct Sector {
    std::vector<int> T1, T2, T3;   // T3 empty: two generators
    int n = -1, kx = -1, ky = -1, kz = 0;
    ed_lattice lat = nullptr;
    ed_basis basis = nullptr;
    int64_t D = 0;
    int L1 = 0, L2 = 0, L3 = 1;
    TermsCopy terms;
    ed_op op = nullptr;
    void* dv = nullptr;
    void* dw = nullptr;
    ~Sector() {
        if (dv) (void)hipFree(dv);
        if (dw) (void)hipFree(dw);
        if (op) ed_op_destroy(op);
        if (basis) ed_basis_destroy(basis);
        if (lat) ed_lattice_destroy(lat);
    }
};

std::vector<std::unique_ptr<Sector>>& cache() {
    static std::vector<std::unique_ptr<Sector>> c;
    return c;
}

// T3 empty (and pz ignored) for two generators.
Sector& sector(const std::vector<int>& T1, const std::vector<int>& T2, const std::vector<int>& T3, int n, double px,
               double py, double pz, std::ofstream& out) {
    static bool init = false;
    if (!init) { check(ed_init(0), out, "ed_init"); init = true; }
    if (T1.size() != T2.size() || (!T3.empty() && T3.size() != T1.size())) fatal(out, "T1, T2, T3 lengths differ");
    ed_lattice lat = nullptr;
    check(ed_lattice_create3((int32_t)T1.size(), T1.data(), T2.data(), T3.empty() ? nullptr : T3.data(), &lat), out,
          "ed_lattice_create3");
    int32_t L1, L2, L3;
    ed_lattice_orders(lat, &L1, &L2);
    ed_lattice_order3(lat, &L3);
    int kx = k_from_p(px, L1), ky = k_from_p(py, L2), kz = T3.empty() ? 0 : k_from_p(pz, L3);
    auto& C = cache();
    for (auto& s : C)
        if (s->T1 == T1 && s->T2 == T2 && s->T3 == T3 && s->n == n && s->kx == kx && s->ky == ky && s->kz == kz) {
            ed_lattice_destroy(lat);
            return *s;
        }
    std::unique_ptr<Sector> s(new Sector());
    s->T1 = T1; s->T2 = T2; s->T3 = T3; s->n = n; s->kx = kx; s->ky = ky; s->kz = kz;
    s->lat = lat; s->L1 = L1; s->L2 = L2; s->L3 = L3;
    check(ed_basis_build3(lat, n, kx, ky, kz, &s->basis), out, "ed_basis_build3");
    ed_basis_dim(s->basis, &s->D);
    if (C.size() >= 4) C.erase(C.begin());
    C.push_back(std::move(s));
    return *C.back();
}

Sector& sector(const std::vector<int>& T1, const std::vector<int>& T2, int n, double px, double py,
               std::ofstream& out) {
    static const std::vector<int> none;
    return sector(T1, T2, none, n, px, py, 0.0, out);
}

ed_op op_for(Sector& S, const TermsCopy& t, std::ofstream& out) {
    if (S.op && S.terms == t) return S.op;
    if (S.op) { ed_op_destroy(S.op); S.op = nullptr; }
    ed_terms tv = t.view();
    check(ed_op_create(S.basis, &tv, &S.op), out, "ed_op_create");
    S.terms = t;
    return S.op;
}

int popcount64(int64_t s) { return __builtin_popcountll((unsigned long long)s); }

// Momentum_Basis.cpp:978-1037 / :901-975 (Unwrap_pState), on the device
// (ed_unwrap): every state of each orbit with coefficient v_i exp(-i k.g) /
// Norm_i, first visit kept, entries with |re| or |im| > 1e-8, by state.
struct Ids {
    int64_t stateID;
    cplx coeff;
};

std::vector<Ids> unwrap(Sector& S, const std::vector<cplx>& v, std::ofstream& out) {
    std::vector<Ids> res;
    if (S.D == 0) return res;
    void* dv = nullptr;
    if (hipMalloc(&dv, S.D * sizeof(cplx)) != hipSuccess) fatal(out, "unwrap: device allocation failed");
    if (hipMemcpy(dv, v.data(), S.D * sizeof(cplx), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dv);
        fatal(out, "unwrap: copy failed");
    }
    int64_t n = 0;
    int rc = ed_unwrap(S.basis, dv, 0, nullptr, nullptr, &n);
    std::vector<int64_t> st(n);
    std::vector<cplx> co(n);
    if (rc == ED_OK && n > 0) rc = ed_unwrap(S.basis, dv, n, st.data(), (double*)co.data(), &n);
    (void)hipFree(dv);
    check(rc, out, "ed_unwrap");
    res.reserve(n);
    for (int64_t q = 0; q < n; q++) res.push_back(Ids{st[q], co[q]});
    return res;
}

void export_basis(Sector& S, std::vector<int64_t>& reps, std::vector<double>& norm, std::vector<double>& norm_inv,
                  std::ofstream& out) {
    reps.assign(S.D, 0);
    norm.assign(S.D, 0.0);
    norm_inv.assign(S.D, 0.0);
    check(ed_basis_export(S.basis, reps.data(), norm.data(), norm_inv.data()), out, "ed_basis_export");
}

// Writes the _EV<i>.txt / _pEV<i>.txt files of Lanczos_sector.cpp:136-197
// (two generators) and :306-366 (three: kz and pz columns added).
void write_vectors(const std::string& path, bool unwrap_peigvec, int n_sec_ev, int kx, int ky, int kz, double Sz,
                   double lambda, const std::vector<cplx>& eigs, const std::vector<std::vector<cplx>>& vecs,
                   Sector& S, double px, double py, double pz, std::ofstream& out) {
    const bool three = !S.T3.empty();
    std::vector<int64_t> reps;
    std::vector<double> norm, norm_inv;
    export_basis(S, reps, norm, norm_inv, out);
    for (int iib = 0; iib < std::min(n_sec_ev, (int)eigs.size()); iib++) {
        std::string fp = path + (unwrap_peigvec ? "_EV" : "_pEV") + std::to_string(iib) + ".txt";
        std::ofstream f(fp.c_str());
        f << "\n Eigenvector for state \n";
        if (three) {
            f << "Sz\t(kx,ky,kz)\tJz/lambda\n";
            f << Sz << "\t(" << kx << "," << ky << "," << kz << ")\t" << lambda << std::endl;
        } else {
            f << "Sz\t(kx,ky)\tJz/lambda\n";
            f << Sz << "\t(" << kx << "," << ky << ")\t" << lambda << std::endl;
        }
        f << "\n\n-------------------------------------------";
        f << "\nState Energy " << eigs[iib] << std::endl;
        if (unwrap_peigvec) {
            std::vector<Ids> psi = unwrap(S, vecs[iib], out);
            f << "\nEIGENVECTOR\n";
            for (auto& e : psi) f << e.stateID << "\t" << std::fixed << std::setprecision(15) << e.coeff << std::endl;
        } else {
            f << "\nMomentum EIGENVECTOR\n";
            if (three) {
                f << "\nnp \t\t px \t\tpy \t\tpz\n";
                f << S.D << "\t\t" << px << "\t\t" << py << "\t\t" << pz << std::endl;
            } else {
                f << "\nnp \t\t px \t\tpy\n";
                f << S.D << "\t\t" << px << "\t\t" << py << std::endl;
            }
            f << "\n\npblock_states \t\t Norm \t\t\t v_coeff\n\n";
            for (int64_t jb = 0; jb < S.D; jb++)
                f << reps[jb] << "\t\t\t" << norm[jb] << "\t\t\t" << std::fixed << std::setprecision(15)
                  << vecs[iib][jb] << std::endl;
        }
    }
}

// Krylov storage precision of the drop-in solvers: complex128 unless the
// environment variable ED_PRECISION is "c64", which stores the Krylov vectors
// in complex64 and keeps FP64 for alpha/beta, the reorthogonalisation
// coefficients, the tridiagonal eigensolve and a final FP64 residual /
// Rayleigh-quotient correction of every eigenvalue.
int krylov_precision(std::ofstream& out) {
    const char* e = std::getenv("ED_PRECISION");
    if (!e || !*e || !std::strcmp(e, "c128")) return ED_C128;
    if (!std::strcmp(e, "c64")) return ED_C64;
    fatal(out, std::string("ED_PRECISION must be c64 or c128, got ") + e);
}

// Lanczos steps as replayed hipGraphs (ED_LZ_GRAPH, the reference schedule's
// steps with their scalars on the device) unless the environment variable
// ED_GRAPH is "0"; either way the same steps, stopping points and eigenpairs.
int step_flags(std::ofstream& out) {
    const char* e = std::getenv("ED_GRAPH");
    if (!e || !*e || !std::strcmp(e, "1")) return ED_LZ_GRAPH;
    if (!std::strcmp(e, "0")) return 0;
    fatal(out, std::string("ED_GRAPH must be 0 or 1, got ") + e);
}

void run_lanczos(ed_op op, int64_t D, int n_ev, int iterations, int ncycles, double tol, int mode,
                 std::vector<cplx>& eigs, std::vector<std::vector<cplx>>& evecs, std::ofstream& out) {
    int nev = (int)std::min<int64_t>(n_ev, D);
    ed_lanczos_params p;
    std::memset(&p, 0, sizeof(p));
    p.nev = nev; p.iterations = iterations; p.ncycles = ncycles; p.precision = krylov_precision(out); p.mode = mode;
    p.flags = mode == ED_LZ_REFERENCE ? step_flags(out) : 0;
    p.tol = tol; p.seed = 1;
    std::vector<double> ev(nev);
    std::vector<cplx> vecs((size_t)nev * D);
    ed_lanczos_stats st;
    check(ed_lanczos(op, &p, ev.data(), (double*)vecs.data(), &st), out, "ed_lanczos");
    eigs.clear();
    for (int k = 0; k < nev; k++) {
        eigs.push_back(ev[k]);
        evecs.emplace_back(vecs.begin() + (size_t)k * D, vecs.begin() + (size_t)(k + 1) * D);
        out << "Eigenvalue number " << k << " = " << std::showpos << std::fixed << std::setprecision(15) << ev[k]
            << std::noshowpos << std::endl;
    }
    out << "\n\n--------------------------------------------------\n";
    out << "\n Lanczos SUMMARY \n" << std::endl;
    out << " Hv calls\t\t\t" << st.hv_calls << std::endl;
    out << " Hv total time \t\t\t" << st.hv_seconds / 60.0 << " minutes" << std::endl;
    // FP64 residual ||H x - E x|| of each returned Ritz vector: the error
    // bound of its eigenvalue (|E - E_exact| <= residual for Hermitian H)
    out << " Krylov storage\t\t\t" << (p.precision == ED_C64 ? "complex64" : "complex128") << std::endl;
    if (mode == ED_LZ_REFERENCE)
        out << " Lanczos steps\t\t\t" << ((p.flags & ED_LZ_GRAPH) ? "hipGraph replays" : "eager launches") << std::endl;
    for (int k = 0; k < std::min(nev, 16); k++)
        out << " FP64 residual " << k << "\t\t" << std::scientific << std::setprecision(3) << st.residual[k]
            << std::defaultfloat << std::endl;
}

}  // namespace

// ------------------------------------------------------------------ sHv ----
namespace {

void shv_impl(cplx* v, cplx* w, const std::vector<int64_t>& pblock_states, double lambda,
              const std::vector<coordinates>& adj_list, double Ch, const std::vector<triangles>& ch_list, double ChCh,
              const std::vector<bowties>& chch_list, const std::vector<int>& T1, const std::vector<int>& T2,
              const std::vector<int>& T3, double px, double py, double pz, std::ofstream& outfile) {
    if (pblock_states.empty()) return;
    Sector& S = sector(T1, T2, T3, popcount64(pblock_states[0]), px, py, pz, outfile);
    if (S.D != (int64_t)pblock_states.size())
        fatal(outfile, "sHv: pblock_states does not match the GPU basis of this sector");
    ed_op op = op_for(S, terms_of(lambda, adj_list, Ch, ch_list, ChCh, chch_list), outfile);
    const size_t bytes = (size_t)S.D * sizeof(cplx);
    if (!S.dv) {
        if (hipMalloc(&S.dv, bytes) != hipSuccess || hipMalloc(&S.dw, bytes) != hipSuccess)
            fatal(outfile, "sHv: device allocation failed");
    }
    if (hipMemcpy(S.dv, v, bytes, hipMemcpyHostToDevice) != hipSuccess) fatal(outfile, "sHv: copy in failed");
    check(ed_hv(op, ED_C128, S.dv, S.dw, nullptr), outfile, "ed_hv");
    if (hipMemcpy(w, S.dw, bytes, hipMemcpyDeviceToHost) != hipSuccess) fatal(outfile, "sHv: copy out failed");
}

}  // namespace

void sHv(std::complex<double>* v, std::complex<double>* w, const std::vector<int64_t>& pblock_states,
         const std::vector<double>&, const std::vector<double>&, const std::vector<int64_t>&,
         const std::vector<std::complex<double> >&, const double& lambda, const std::vector<coordinates>& adj_list,
         const double& Ch, const std::vector<triangles>& ch_list, const double& ChCh,
         const std::vector<bowties>& chch_list, const int&, const std::vector<int64_t>&,
         const std::vector<int64_t>&, const int&, const std::vector<int>& T1, const std::vector<int>& T2,
         const double& kx, const double& ky, std::ofstream& outfile) {
    static const std::vector<int> none;
    shv_impl(v, w, pblock_states, lambda, adj_list, Ch, ch_list, ChCh, chch_list, T1, T2, none, kx, ky, 0.0, outfile);
}

void sHv(std::complex<double>* v, std::complex<double>* w, const std::vector<int64_t>& pblock_states,
         const std::vector<double>&, const std::vector<double>&, const std::vector<int64_t>&,
         const std::vector<std::complex<double> >&, const double& lambda, const std::vector<coordinates>& adj_list,
         const double& Ch, const std::vector<triangles>& ch_list, const double& ChCh,
         const std::vector<bowties>& chch_list, const int&, const std::vector<int64_t>&,
         const std::vector<int64_t>&, const int&, const std::vector<int>& T1, const std::vector<int>& T2,
         const std::vector<int>& T3, const double& kx, const double& ky, const double& kz, std::ofstream& outfile) {
    shv_impl(v, w, pblock_states, lambda, adj_list, Ch, ch_list, ChCh, chch_list, T1, T2, T3, kx, ky, kz, outfile);
}

// ---------------------------------------------------- pConstruct_States ----
namespace {

void construct_impl(std::vector<int64_t>& v_sz, std::vector<double>& norm, std::vector<double>& norm_inv,
                    const std::vector<int>& T1, const std::vector<int>& T2, const std::vector<int>& T3, int num_ones,
                    double px, double py, double pz, int64_t& n_states, std::vector<int64_t>& rep_loc,
                    std::vector<cplx>& state_ph, std::vector<int64_t>& Ia, std::vector<int64_t>& Ib, int& bits_right,
                    std::ofstream& outfile) {
    if ((int)T1.size() < num_ones) fatal(outfile, "ERROR: Num sites < Num_ones");
    Sector& S = sector(T1, T2, T3, num_ones, px, py, pz, outfile);
    export_basis(S, v_sz, norm, norm_inv, outfile);
    n_states = S.D;
    int64_t C = 0, nl = 0;
    ed_basis_full_size(S.basis, &C, &nl);
    rep_loc.assign(C, -1);
    state_ph.assign(C, cplx(0, 0));
    Ia.assign(nl, 0);
    Ib.assign(nl, 0);
    int32_t br = 0;
    check(ed_basis_export_full(S.basis, rep_loc.data(), (double*)state_ph.data(), Ia.data(), Ib.data(), &br), outfile,
          "ed_basis_export_full");
    bits_right = br;
}

}  // namespace

void pConstruct_States(std::vector<int64_t>& v_sz, std::vector<double>& norm, std::vector<double>& norm_inv,
                       std::vector<int>& T1, std::vector<int>& T2, int& num_ones, double& px, double& py,
                       int64_t& n_states, std::vector<int64_t>& rep_loc,
                       std::vector<std::complex<double> >& state_ph, std::vector<int64_t>& Ia,
                       std::vector<int64_t>& Ib, int& bits_right, std::ofstream& outfile) {
    static const std::vector<int> none;
    construct_impl(v_sz, norm, norm_inv, T1, T2, none, num_ones, px, py, 0.0, n_states, rep_loc, state_ph, Ia, Ib,
                   bits_right, outfile);
}

void pConstruct_States(std::vector<int64_t>& v_sz, std::vector<double>& norm, std::vector<double>& norm_inv,
                       std::vector<int>& T1, std::vector<int>& T2, std::vector<int>& T3, int& num_ones, double& px,
                       double& py, double& pz, int64_t& n_states, std::vector<int64_t>& rep_loc,
                       std::vector<std::complex<double> >& state_ph, std::vector<int64_t>& Ia,
                       std::vector<int64_t>& Ib, int& bits_right, std::ofstream& outfile) {
    construct_impl(v_sz, norm, norm_inv, T1, T2, T3, num_ones, px, py, pz, n_states, rep_loc, state_ph, Ia, Ib,
                   bits_right, outfile);
}

// ---------------------------------------------------------------- makeH ----
void makeH(const std::vector<int64_t>& pblock_states, const std::vector<double>&, const std::vector<double>&,
           const std::vector<int64_t>&, const std::vector<std::complex<double> >&, const double& lambda,
           const std::vector<coordinates>& adj_list, const double& Ch, const std::vector<triangles>& ch_list,
           const double& ChCh, const std::vector<bowties>& chch_list, const int&, const std::vector<int64_t>&,
           const std::vector<int64_t>&, const int&, const std::vector<int>& T1, const std::vector<int>& T2,
           const double& kx, const double& ky, std::ofstream& outfile, zMatrix& H) {
    const int64_t n = (int64_t)pblock_states.size();
    H.resize(n, n);
    if (n == 0) return;
    Sector& S = sector(T1, T2, popcount64(pblock_states[0]), kx, ky, outfile);
    if (S.D != n) fatal(outfile, "makeH: pblock_states does not match the GPU basis of this sector");
    ed_op op = op_for(S, terms_of(lambda, adj_list, Ch, ch_list, ChCh, chch_list), outfile);
    check(ed_make_h(op, (double*)H.data()), outfile, "ed_make_h");
}

// ------------------------------------------------------- sparse sHv ----
void sHv(sMatrix& pBij, std::complex<double>* v, std::complex<double>* w) {
    const int64_t n = pBij.NRows();
    std::vector<int64_t> rowptr(n + 1, 0), col;
    std::vector<cplx> val;
    for (int64_t i = 0; i < n; i++) {
        const int64_t m = pBij.RowSize(i);
        for (int64_t j = 0; j < m; j++) {
            ids e = pBij.me(i, j);
            col.push_back(e.stateID);
            val.push_back(e.coeff);
        }
        rowptr[i + 1] = (int64_t)col.size();
    }
    // the columns a row refers to: at most the largest index used
    int64_t ncols = std::max<int64_t>(pBij.NCols(), n);
    for (int64_t c : col) ncols = std::max(ncols, c + 1);
    if (ed_csr_hv_host(n, ncols, rowptr.data(), col.data(), (const double*)val.data(), (const double*)v,
                       (double*)w) != ED_OK) {
        std::cerr << "sHv (sparse): " << ed_last_error() << std::endl;
        std::exit(1);
    }
}

// ------------------------------------------------- ARPACK_Iteration_EV ----
void ARPACK_Iteration_EV(std::vector<int64_t>& pblock_states, std::vector<double>&, std::vector<double>&,
                         std::vector<int64_t>&, std::vector<std::complex<double> >&, int&, int& n_it, double& px,
                         double& py, std::vector<std::complex<double> >& Evals, zMatrix& Evecs, bool find_ev,
                         double lambda, std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list,
                         double& ChCh, std::vector<bowties>& chch_list, double tol, int maxiter,
                         std::vector<int>& T1, std::vector<int>& T2, std::vector<int64_t>&, std::vector<int64_t>&,
                         int&, std::ofstream& outfile) {
    const int64_t n = (int64_t)pblock_states.size();
    int nev = n_it;
    if (nev >= n - 1 && n > 0) nev = (int)(n - 2);   // ARPACK_sector.cpp:34-35
    if (n == 0 || nev <= 0) fatal(outfile, "ARPACK_Iteration_EV: sector too small for the Arnoldi iteration");
    Sector& S = sector(T1, T2, popcount64(pblock_states[0]), px, py, outfile);
    if (S.D != n) fatal(outfile, "ARPACK_Iteration_EV: basis mismatch");
    ed_op op = op_for(S, terms_of(lambda, adj_list, Ch, ch_list, ChCh, chch_list), outfile);
    outfile << " ARPACK PARAMETERS " << std::endl;
    outfile << " nev = " << nev << std::endl;
    outfile << " n = " << n << std::endl;
    outfile << " n_it = " << n_it << std::endl;
    outfile << "----------------------\n" << std::endl;
    // ncv = 4 nev Arnoldi vectors (ARPACK_sector.cpp:42), at least 2 nev + 40
    // for the thick restart, maxiter restarts
    const int ncv = (int)std::min<int64_t>(n, std::max(4 * nev, 2 * nev + 40));
    std::vector<cplx> eigs;
    std::vector<std::vector<cplx>> vecs;
    run_lanczos(op, S.D, nev, ncv, std::max(maxiter, 1), tol, ED_LZ_THICK, eigs, vecs, outfile);
    if ((int)Evals.size() < nev) Evals.resize(nev);
    for (int i = 0; i < nev; i++) Evals[i] = eigs[nev - 1 - i];   // largest of the wanted first
    if (find_ev) {
        Evecs.resize(n_it, n);                                     // ARPACK_sector.cpp:176-181
        for (int i = 0; i < nev; i++)
            for (int64_t j = 0; j < n; j++) Evecs(i, j) = vecs[nev - 1 - i][j];
    }
}

// ---------------------------------------------------------- exact_evecs ----
void exact_evecs(std::vector<int64_t>& pblock_states, std::vector<double>&, std::vector<double>&,
                 std::vector<int64_t>&, std::vector<std::complex<double> >&, std::vector<std::complex<double> >& eigs,
                 std::vector<std::vector<std::complex<double> > >& previous_evecs, bool, double& lambda, int&,
                 double& px, double& py, int& n_ones, int64_t&, std::vector<coordinates>& adj_list, double& Ch,
                 std::vector<triangles>& ch_list, double& ChCh, std::vector<bowties>& chch_list,
                 std::vector<int>& T1, std::vector<int>& T2, int, double, int, int, std::vector<int64_t>&,
                 std::vector<int64_t>&, int, std::ofstream& outfile) {
    Sector& S = sector(T1, T2, n_ones, px, py, outfile);
    if (S.D != (int64_t)pblock_states.size()) fatal(outfile, "exact_evecs: basis mismatch");
    ed_op op = op_for(S, terms_of(lambda, adj_list, Ch, ch_list, ChCh, chch_list), outfile);
    std::vector<double> ev(S.D);
    std::vector<cplx> vecs((size_t)S.D * S.D);
    check(ed_dense_eigh(op, ev.data(), (double*)vecs.data()), outfile, "ed_dense_eigh");
    eigs.clear();
    for (double e : ev) eigs.push_back(e);
    for (int64_t i = 0; i < S.D; i++)
        previous_evecs.emplace_back(vecs.begin() + (size_t)i * S.D, vecs.begin() + (size_t)(i + 1) * S.D);
    outfile << "\n\n--------------------------------------------------\n";
    outfile << "\n EXACT SUMMARY \n" << std::endl;
}

// -------------------------------------------------------- lanczos_evecs ----
void lanczos_evecs(std::vector<int64_t>& pblock_states, std::vector<double>&, std::vector<double>&,
                   std::vector<int64_t>&, std::vector<std::complex<double> >&, std::vector<std::complex<double> >& eigs,
                   std::vector<std::vector<std::complex<double> > >& previous_evecs, bool, double& lambda, int&,
                   double& px, double& py, int& n_ones, int64_t&, std::vector<coordinates>& adj_list, double& Ch,
                   std::vector<triangles>& ch_list, double& ChCh, std::vector<bowties>& chch_list,
                   std::vector<int>& T1, std::vector<int>& T2, int n_ev, double tol, int iterations, int ncycles,
                   std::vector<int64_t>&, std::vector<int64_t>&, int, std::ofstream& outfile) {
    Sector& S = sector(T1, T2, n_ones, px, py, outfile);
    if (S.D != (int64_t)pblock_states.size()) fatal(outfile, "lanczos_evecs: basis mismatch");
    ed_op op = op_for(S, terms_of(lambda, adj_list, Ch, ch_list, ChCh, chch_list), outfile);
    run_lanczos(op, S.D, n_ev, iterations, ncycles, tol, ED_LZ_REFERENCE, eigs, previous_evecs, outfile);
}

// lanczos.h:38-69 (three generators)
void lanczos_evecs(std::vector<int64_t>& pblock_states, std::vector<double>&, std::vector<double>&,
                   std::vector<int64_t>&, std::vector<std::complex<double> >&, std::vector<std::complex<double> >& eigs,
                   std::vector<std::vector<std::complex<double> > >& previous_evecs, bool, double& lambda, int&,
                   double& px, double& py, double& pz, int& n_ones, int64_t&, std::vector<coordinates>& adj_list,
                   double& Ch, std::vector<triangles>& ch_list, double& ChCh, std::vector<bowties>& chch_list,
                   std::vector<int>& T1, std::vector<int>& T2, std::vector<int>& T3, int n_ev, double tol,
                   int iterations, int ncycles, std::vector<int64_t>&, std::vector<int64_t>&, int,
                   std::ofstream& outfile) {
    Sector& S = sector(T1, T2, T3, n_ones, px, py, pz, outfile);
    if (S.D != (int64_t)pblock_states.size()) fatal(outfile, "lanczos_evecs: basis mismatch");
    ed_op op = op_for(S, terms_of(lambda, adj_list, Ch, ch_list, ChCh, chch_list), outfile);
    run_lanczos(op, S.D, n_ev, iterations, ncycles, tol, ED_LZ_REFERENCE, eigs, previous_evecs, outfile);
}

// ------------------------------------------------------- sector_lanczos ----
namespace {

// Lanczos_sector.cpp:15-198 (two generators, T3 empty) and :201-367 (three).
void sector_lanczos_impl(int kx, int ky, int kz, double cSz, int n_spins, double lambda, double J2,
                         std::vector<coordinates>& adj_list, double Ch, std::vector<triangles>& ch_list, double ChCh,
                         std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2,
                         const std::vector<int>& T3, int n_sec_ev, bool find_ev, bool unwrap_peigvec,
                         const std::string& filepath_EV, double tol, int iterations, int ncycles,
                         std::ofstream& outfile) {
    const bool three = !T3.empty();
    int n_ones = (int)(cSz + n_spins / 2 + 1.0e-6);   // Lanczos_sector.cpp:45-46 (integer n_spins/2)
    ed_lattice lat = nullptr;
    if (T1.size() != T2.size() || (three && T3.size() != T1.size())) fatal(outfile, "T1, T2, T3 lengths differ");
    check(ed_lattice_create3((int32_t)T1.size(), T1.data(), T2.data(), three ? T3.data() : nullptr, &lat), outfile,
          "ed_lattice_create3");
    int32_t L1, L2, L3;
    ed_lattice_orders(lat, &L1, &L2);
    ed_lattice_order3(lat, &L3);
    ed_lattice_destroy(lat);
    if (kx >= L1 || ky >= L2 || kz >= L3 || kx < 0 || ky < 0 || kz < 0) fatal(outfile, "Invalid kx or ky");
    double px = 2 * kx * M_PI / L1, py = 2 * ky * M_PI / L2, pz = 2 * kz * M_PI / L3;
    double Sz = 0.5 * n_ones - 0.5 * (n_spins - n_ones);
    outfile << "\n--------------------------------------------------------------------\n";
    Sector& S = sector(T1, T2, T3, n_ones, px, py, pz, outfile);
    if (three)
        outfile << " SECTOR : " << Sz << "\t(" << px << "," << py << "," << pz << ")\t" << S.D << std::endl;
    else
        outfile << " SECTOR : " << Sz << "\t(" << px << "," << py << ")\t" << S.D << std::endl;
    if (S.D == 0) fatal(outfile, "No states in psector");
    outfile << "------------------------------------------------------------------" << std::endl;
    ed_op op = op_for(S, terms_of(lambda, adj_list, Ch, ch_list, ChCh, chch_list), outfile);
    std::vector<cplx> eigs;
    std::vector<std::vector<cplx>> vecs;
    // Lanczos_sector.cpp:105: dense below 10000 states (two generators); the
    // three-generator entry always runs Lanczos (:284-289)
    if (three || S.D > 10000) {
        run_lanczos(op, S.D, n_sec_ev, iterations, ncycles, tol, ED_LZ_REFERENCE, eigs, vecs, outfile);
    } else {
        std::vector<double> ev(S.D);
        std::vector<cplx> dv((size_t)S.D * S.D);
        check(ed_dense_eigh(op, ev.data(), find_ev ? (double*)dv.data() : nullptr), outfile, "ed_dense_eigh");
        for (double e : ev) eigs.push_back(e);
        if (find_ev)
            for (int64_t i = 0; i < std::min<int64_t>(n_sec_ev, S.D); i++)
                vecs.emplace_back(dv.begin() + (size_t)i * S.D, dv.begin() + (size_t)(i + 1) * S.D);
    }
    outfile << "pblock_eigs.size() = " << eigs.size() << std::endl;
    if (three)
        outfile << "\n\nEnergies\t\tEnergies/site\t\tSz\tkx\tky\tkz\tJz\tJ2\n\n";
    else
        outfile << "\n\nEnergies\t\tEnergies/site\t\tSz\tkx\tky\tJz\tJ2\n\n";
    for (size_t i = 0; i < eigs.size(); i++) {
        outfile << std::fixed << std::setprecision(15) << eigs[i].real() << "\t" << eigs[i].real() / double(n_spins)
                << "\t" << std::setprecision(3) << Sz << "\t" << px << "\t" << py << "\t";
        if (three) outfile << pz << "\t";
        outfile << lambda << "\t" << J2 << "\n";
    }
    outfile.flush();
    if (find_ev)
        write_vectors(filepath_EV, unwrap_peigvec, n_sec_ev, kx, ky, kz, Sz, lambda, eigs, vecs, S, px, py, pz, outfile);
}

}  // namespace

void sector_lanczos(int& kx, int& ky, double& cSz, int& n_spins, double& lambda, double& J2,
                    std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                    std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2, int& n_sec_ev,
                    bool& find_ev, bool& unwrap_peigvec, std::string& filepath_EV, double& tol, int& iterations,
                    int& ncycles, std::ofstream& outfile) {
    static const std::vector<int> none;
    sector_lanczos_impl(kx, ky, 0, cSz, n_spins, lambda, J2, adj_list, Ch, ch_list, ChCh, chch_list, T1, T2, none,
                        n_sec_ev, find_ev, unwrap_peigvec, filepath_EV, tol, iterations, ncycles, outfile);
}

void sector_lanczos(int& kx, int& ky, int& kz, double& cSz, int& n_spins, double& lambda, double& J2,
                    std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                    std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2,
                    std::vector<int>& T3, int& n_sec_ev, bool& find_ev, bool& unwrap_peigvec,
                    std::string& filepath_EV, double& tol, int& iterations, int& ncycles, std::ofstream& outfile) {
    sector_lanczos_impl(kx, ky, kz, cSz, n_spins, lambda, J2, adj_list, Ch, ch_list, ChCh, chch_list, T1, T2, T3,
                        n_sec_ev, find_ev, unwrap_peigvec, filepath_EV, tol, iterations, ncycles, outfile);
}

// ----------------------------------------------------------- ARPACK path ----
void ARPACK_sector(std::vector<int64_t>& pblock_states, std::vector<double>&, std::vector<double>&,
                   std::vector<int64_t>&, std::vector<std::complex<double> >&,
                   std::vector<std::complex<double> >& pblock_eigs, zMatrix& pblock_eigvecs, bool find_ev,
                   double& lambda, int&, double& px, double& py, int& n_ones, int64_t& n_p,
                   std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                   std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2, int n_ev,
                   double tol, int maxiter, std::vector<int64_t>&, std::vector<int64_t>&, int&,
                   std::ofstream& outfile) {
    Sector& S = sector(T1, T2, n_ones, px, py, outfile);
    if (S.D != (int64_t)pblock_states.size() || S.D != n_p) fatal(outfile, "ARPACK_sector: basis mismatch");
    ed_op op = op_for(S, terms_of(lambda, adj_list, Ch, ch_list, ChCh, chch_list), outfile);
    // ARPACK_sector.cpp:213-217: n_ev eigenvalues; n_p <= n_ev + 1 takes the
    // dense branch, which returns all n_p of them (:232-245)
    int it = n_ev > n_p - 1 ? (int)n_p : n_ev;
    std::vector<cplx> eigs;
    std::vector<std::vector<cplx>> vecs;
    if (n_p > n_ev + 1) {
        const int ncv = (int)std::min<int64_t>(n_p, std::max(4 * it, 2 * it + 40));
        run_lanczos(op, S.D, it, ncv, std::max(maxiter, 1), tol, ED_LZ_THICK, eigs, vecs, outfile);
        // ARPACK returns the wanted Ritz values largest first; the caller prints
        // them in reverse (ARPACK_sector.cpp:339-340), i.e. ascending.
        pblock_eigs.assign(eigs.rbegin(), eigs.rend());
        if (find_ev) {
            pblock_eigvecs.resize(it, S.D);   // Evecs(i, j) = component j of vector i (ARPACK_sector.cpp:176-181)
            for (int i = 0; i < it; i++)
                for (int64_t j = 0; j < S.D; j++) pblock_eigvecs(i, j) = vecs[it - 1 - i][j];
        }
    } else {
        // Dense branch (ARPACK_sector.cpp:232-251): it = n_p, every eigenpair
        // of the n_p x n_p sector matrix, the eigenvectors as the columns of
        // pblock_eigvecs (zgeev's VR: element (c, k) = component c of vector
        // k), the matrix printed to the log.  The reference fills this matrix
        // with its legacy builder (zXXZ_Heisenberg,
        // Heisenberg_Hamiltonian.cpp:92-200), whose diagonal omits J and which
        // drops bond phases and bowtie terms; here it is the sector matrix of
        // sHv, so that both branches diagonalise the same H.  zgeev leaves the
        // order unspecified; the values are stored largest first, as the
        // Arnoldi branch does, so that the caller prints them ascending.
        outfile << " Here \n\n";
        it = (int)n_p;
        std::vector<double> ev(S.D);
        std::vector<cplx> dv((size_t)S.D * S.D), H((size_t)S.D * S.D);
        check(ed_dense_eigh(op, ev.data(), (double*)dv.data()), outfile, "ed_dense_eigh");
        check(ed_make_h(op, (double*)H.data()), outfile, "ed_make_h");
        pblock_eigs.assign(it, cplx(0, 0));
        pblock_eigvecs.resize(n_p, n_p);
        for (int k = 0; k < it; k++) {
            const int src = it - 1 - k;
            pblock_eigs[k] = ev[src];
            for (int64_t c = 0; c < n_p; c++) pblock_eigvecs(c, k) = dv[(size_t)src * S.D + c];
        }
        for (int64_t ii = 0; ii < n_p; ii++) {
            for (int64_t jj = 0; jj < n_p; jj++)
                outfile << std::fixed << std::setprecision(5) << H[ii + jj * n_p] << "\t\t";
            outfile << std::endl;
        }
    }
    outfile << "\nEigenvalues\n";
    for (int kb = 0; kb < it; kb++) outfile << std::setprecision(10) << pblock_eigs[kb] << "\t";
}

void sector_eigs_eigvecs(int& kx, int& ky, double& cSz, int& n_spins, double& lambda, double& J2,
                         std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list,
                         double& ChCh, std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2,
                         int& n_sec_ev, bool& find_ev, bool& unwrap_peigvec, std::string& filepath_EV, double& tol,
                         int& maxiter, std::ofstream& outfile) {
    // ARPACK_sector.cpp:284-287: n_spins/2 plus one for every step of 1 below cSz
    int n_ones = n_spins / 2;
    for (double i = 0.0; i < cSz; i += 1) n_ones++;
    ed_lattice lat = nullptr;
    if (T1.size() != T2.size()) fatal(outfile, "T1 and T2 lengths differ");
    check(ed_lattice_create((int32_t)T1.size(), T1.data(), T2.data(), &lat), outfile, "ed_lattice_create");
    int32_t L1, L2;
    ed_lattice_orders(lat, &L1, &L2);
    ed_lattice_destroy(lat);
    if (kx >= L1 || ky >= L2 || kx < 0 || ky < 0) fatal(outfile, "Invalid kx or ky");
    double px = 2 * kx * M_PI / L1, py = 2 * ky * M_PI / L2;
    double Sz = 0.5 * n_ones - 0.5 * (n_spins - n_ones);
    outfile << "\n--------------------------------------------------------------------\n";
    Sector& S = sector(T1, T2, n_ones, px, py, outfile);
    outfile << " SECTOR : " << Sz << "\t(" << px << "," << py << ")\t" << S.D << std::endl;
    if (S.D == 0) fatal(outfile, "No states in psector");
    outfile << "-------------------------------------------------------\n" << std::endl;
    std::vector<int64_t> reps;
    std::vector<double> norm, norm_inv;
    export_basis(S, reps, norm, norm_inv, outfile);
    std::vector<int64_t> rep_loc, Ia, Ib;
    std::vector<cplx> state_ph, pblock_eigs;
    int bits_right = 0;
    int64_t n_p = S.D;
    zMatrix pblock_eigvecs;
    ARPACK_sector(reps, norm, norm_inv, rep_loc, state_ph, pblock_eigs, pblock_eigvecs, find_ev, lambda, n_spins, px,
                  py, n_ones, n_p, adj_list, Ch, ch_list, ChCh, chch_list, T1, T2, n_sec_ev, tol, maxiter, Ia, Ib,
                  bits_right, outfile);
    outfile << "\n\nEnergies\t\tEnergies/site\t\tSz\tkx\tky\tJz\tJ2\n\n";
    for (int i = (int)pblock_eigs.size() - 1; i >= 0; i--)
        outfile << std::fixed << std::setprecision(15) << pblock_eigs[i].real() << "\t"
                << pblock_eigs[i].real() / double(n_spins) << "\t" << std::setprecision(3) << Sz << "\t" << px << "\t"
                << py << "\t" << lambda << "\t" << J2 << "\n";
    if (!find_ev) return;
    // ARPACK_sector.cpp:345-397: files numbered from 1 by one character,
    // vector iib read as row iib of pblock_eigvecs; the reference reads past
    // the returned values when n_sec_ev exceeds them, the shim stops there
    for (int iib = 0; iib < std::min<int>(n_sec_ev, (int)pblock_eigs.size()); iib++) {
        const char tag = (char)(iib + 1 + (int)'0');
        std::string fp = filepath_EV + (unwrap_peigvec ? "_EV" : "_pEV") + tag + ".txt";
        std::ofstream f(fp.c_str());
        f << "\n Eigenvector for state \n";
        f << "Sz\t(kx,ky)\tJz/lambda\n";
        f << Sz << "\t(" << kx << "," << ky << ")\t" << lambda << std::endl;
        f << "\n\n-------------------------------------------";
        f << "\nState Energy " << pblock_eigs[iib] << std::endl;
        std::vector<cplx> v(n_p);
        for (int64_t jb = 0; jb < n_p; jb++) v[jb] = pblock_eigvecs(iib, jb);
        if (unwrap_peigvec) {
            std::vector<Ids> psi = unwrap(S, v, outfile);
            f << "\nEIGENVECTOR\n";
            for (auto& e : psi) f << e.stateID << "\t" << std::fixed << std::setprecision(15) << e.coeff << std::endl;
        } else {
            f << "\nMomentum EIGENVECTOR\n";
            f << "\nnp \t\t px \t\tpy\n";
            f << n_p << "\t\t" << px << "\t\t" << py << std::endl;
            f << "\n\npblock_states \t\t Norm \t\t\t v_coeff\n\n";
            for (int64_t jb = 0; jb < n_p; jb++)
                f << reps[jb] << "\t\t\t" << norm[jb] << "\t\t\t" << std::fixed << std::setprecision(15) << v[jb]
                  << std::endl;
        }
    }
}
