// ed_refapi.h -- the reference's C++ sector interface, implemented on top of
// the C ABI of ed_amd.h by ed_code_amd/csrc/ed_refapi.cpp.
//
// A caller written against hiteshjc/ED_Code (ED.cpp, read_input_file.cpp)
// links libed_amd.so instead of the reference's Momentum_Basis / H_v /
// lanczos / Lanczos_sector / ARPACK_sector objects.  The struct layouts below
// are the reference's (global.h:25-40, matrix.h:53-91) so that the caller's
// objects, compiled against its own global.h, share the mangled names and the
// memory layout; a translation unit must include either this header or the
// reference's global.h, not both.
//
// The Lanczos and ARPACK-like solvers behind these entries keep every Krylov
// vector in complex128 unless the environment variable ED_PRECISION is "c64":
// then the vectors are stored in complex64, alpha/beta, the
// reorthogonalisation coefficients and the tridiagonal eigensolve stay FP64,
// and each eigenvalue is replaced by the FP64 Rayleigh quotient of its Ritz
// vector, whose FP64 residual is written to the log.
#ifndef ED_REFAPI_H
#define ED_REFAPI_H

#include <complex>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

struct coordinates {            // global.h:25-30
    int x, y;
    double J;
    std::complex<double> expphase;
};

struct triangles {              // global.h:32-35
    int i, j, k;
    std::complex<double> expphase;
};

struct bowties {                // global.h:37-40
    int x, y, z;
    int p, q, r;
};

struct ids {                    // idsClass.h:8-11
    int64_t stateID;
    std::complex<double> coeff;
};

class sMatrix {                 // matrix.h:96-155: rows of (column, value) pairs
    std::vector<std::vector<ids> > states;
    int64_t Nr, Nc;

public:
    sMatrix() : Nr(0), Nc(0) {}
    sMatrix(int64_t m, int64_t n) : states(m), Nr(m), Nc(n) {}
    void clear() { states.clear(); }
    void resize(int64_t m, int64_t n) {
        states.clear();
        states.resize(m);
        Nr = m;
        Nc = n;
    }
    int64_t NRows() { return Nr; }
    int64_t NCols() { return Nc; }
    int64_t RowSize(int64_t i) { return (int64_t)states[i].size(); }
    ids me(int64_t i, int64_t j) { return states[i][j]; }
    void push_back(int64_t id, int64_t stateID, std::complex<double> coeff) {
        if (id < Nr) states[id].push_back(ids{stateID, coeff});
        else std::cout << "\nError in Sparse Matrix Push_Back \n";
    }
    std::vector<ids> return_sVector(int64_t id) { return states[id]; }
};

class zMatrix : public std::vector<std::complex<double> > {   // matrix.h:53-91, column-major
    int64_t m;
    int64_t n;

public:
    zMatrix() : m(0), n(0) {}
    zMatrix(int64_t m_, int64_t n_) : std::vector<std::complex<double> >(m_ * n_), m(m_), n(n_) {}
    void resize(int64_t m_, int64_t n_) {
        m = m_;
        n = n_;
        std::vector<std::complex<double> >::resize(m_ * n_);
    }
    int64_t NRows() const { return m; }
    int64_t NCols() const { return n; }
    std::complex<double>& operator()(int64_t i, int64_t j) { return operator[](i + j * m); }
    const std::complex<double>& operator()(int64_t i, int64_t j) const { return operator[](i + j * m); }
};

// H_v.h:9-28: the dense sector matrix H (D x D, D <= 20000), built on the GPU.
// kx, ky are px, py in radians; the tables rep_loc ... bits_right are not read.
void makeH(const std::vector<int64_t>& pblock_states, const std::vector<double>& norm,
           const std::vector<double>& norm_inv, const std::vector<int64_t>& rep_loc,
           const std::vector<std::complex<double> >& state_ph, const double& lambda,
           const std::vector<coordinates>& adj_list, const double& Ch, const std::vector<triangles>& ch_list,
           const double& ChCh, const std::vector<bowties>& chch_list, const int& n_sites,
           const std::vector<int64_t>& Ia, const std::vector<int64_t>& Ib, const int& bits_right,
           const std::vector<int>& T1, const std::vector<int>& T2, const double& kx, const double& ky,
           std::ofstream& outfile, zMatrix& H);

// H_v.h:31-33: w = pBij v for a caller-built sparse matrix, on the GPU.
void sHv(sMatrix& pBij, std::complex<double>* v, std::complex<double>* w);

// H_v.h:36-57.  kx, ky are px, py in radians (caller lanczos.cpp:232).  The
// tables rep_loc, state_ph, Ia, Ib, bits_right are accepted and not read: the
// GPU recomputes representatives and phases on the fly.
void sHv(std::complex<double>* v, std::complex<double>* w, const std::vector<int64_t>& pblock_states,
         const std::vector<double>& norm, const std::vector<double>& norm_inv, const std::vector<int64_t>& rep_loc,
         const std::vector<std::complex<double> >& state_ph, const double& lambda,
         const std::vector<coordinates>& adj_list, const double& Ch, const std::vector<triangles>& ch_list,
         const double& ChCh, const std::vector<bowties>& chch_list, const int& n_sites,
         const std::vector<int64_t>& Ia, const std::vector<int64_t>& Ib, const int& bits_right,
         const std::vector<int>& T1, const std::vector<int>& T2, const double& kx, const double& ky,
         std::ofstream& outfile);

// H_v.h:59-83 (three generators; kz is pz in radians).
void sHv(std::complex<double>* v, std::complex<double>* w, const std::vector<int64_t>& pblock_states,
         const std::vector<double>& norm, const std::vector<double>& norm_inv, const std::vector<int64_t>& rep_loc,
         const std::vector<std::complex<double> >& state_ph, const double& lambda,
         const std::vector<coordinates>& adj_list, const double& Ch, const std::vector<triangles>& ch_list,
         const double& ChCh, const std::vector<bowties>& chch_list, const int& n_sites,
         const std::vector<int64_t>& Ia, const std::vector<int64_t>& Ib, const int& bits_right,
         const std::vector<int>& T1, const std::vector<int>& T2, const std::vector<int>& T3, const double& kx,
         const double& ky, const double& kz, std::ofstream& outfile);

// Momentum_Basis.h:84-100 (three generators).
void pConstruct_States(std::vector<int64_t>& v_sz, std::vector<double>& norm, std::vector<double>& norm_inv,
                       std::vector<int>& T1, std::vector<int>& T2, std::vector<int>& T3, int& num_ones, double& px,
                       double& py, double& pz, int64_t& n_states, std::vector<int64_t>& rep_loc,
                       std::vector<std::complex<double> >& state_ph, std::vector<int64_t>& Ia,
                       std::vector<int64_t>& Ib, int& bits_right, std::ofstream& outfile);

// Momentum_Basis.h:151-165 (2-D).
void pConstruct_States(std::vector<int64_t>& v_sz, std::vector<double>& norm, std::vector<double>& norm_inv,
                       std::vector<int>& T1, std::vector<int>& T2, int& num_ones, double& px, double& py,
                       int64_t& n_states, std::vector<int64_t>& rep_loc,
                       std::vector<std::complex<double> >& state_ph, std::vector<int64_t>& Ia,
                       std::vector<int64_t>& Ib, int& bits_right, std::ofstream& outfile);

// lanczos.h:8-38
void exact_evecs(std::vector<int64_t>& pblock_states, std::vector<double>& Norm, std::vector<double>& Norm_inv,
                 std::vector<int64_t>& rep_loc, std::vector<std::complex<double> >& state_ph,
                 std::vector<std::complex<double> >& eigs, std::vector<std::vector<std::complex<double> > >& previous_evecs,
                 bool find_ev, double& lambda, int& n_spins, double& px, double& py, int& n_ones, int64_t& n_p,
                 std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                 std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2, int n_ev, double tol,
                 int iterations, int ncycles, std::vector<int64_t>& Ia, std::vector<int64_t>& Ib, int bits_right,
                 std::ofstream& outfile);

// lanczos.h:38-69 (three generators)
void lanczos_evecs(std::vector<int64_t>& pblock_states, std::vector<double>& Norm, std::vector<double>& Norm_inv,
                   std::vector<int64_t>& rep_loc, std::vector<std::complex<double> >& state_ph,
                   std::vector<std::complex<double> >& eigs,
                   std::vector<std::vector<std::complex<double> > >& previous_evecs, bool find_ev, double& lambda,
                   int& n_spins, double& px, double& py, double& pz, int& n_ones, int64_t& n_p,
                   std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                   std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2,
                   std::vector<int>& T3, int n_ev, double tol, int iterations, int ncycles,
                   std::vector<int64_t>& Ia, std::vector<int64_t>& Ib, int bits_right, std::ofstream& outfile);

// lanczos.h:71-99 (2-D)
void lanczos_evecs(std::vector<int64_t>& pblock_states, std::vector<double>& Norm, std::vector<double>& Norm_inv,
                   std::vector<int64_t>& rep_loc, std::vector<std::complex<double> >& state_ph,
                   std::vector<std::complex<double> >& eigs,
                   std::vector<std::vector<std::complex<double> > >& previous_evecs, bool find_ev, double& lambda,
                   int& n_spins, double& px, double& py, int& n_ones, int64_t& n_p,
                   std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                   std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2, int n_ev, double tol,
                   int iterations, int ncycles, std::vector<int64_t>& Ia, std::vector<int64_t>& Ib, int bits_right,
                   std::ofstream& outfile);

// Lanczos_sector.h:11-31 (2-D)
void sector_lanczos(int& kx, int& ky, double& cSz, int& n_spins, double& lambda, double& J2,
                    std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                    std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2, int& n_sec_ev,
                    bool& find_ev, bool& unwrap_peigvec, std::string& filepath_EV, double& tol, int& iterations,
                    int& ncycles, std::ofstream& outfile);

// Lanczos_sector.h:33-55 (3-D): the three-generator sector, always solved by
// Lanczos (Lanczos_sector.cpp:284-289); eigenvector files carry kz and pz.
void sector_lanczos(int& kx, int& ky, int& kz, double& cSz, int& n_spins, double& lambda, double& J2,
                    std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                    std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2,
                    std::vector<int>& T3, int& n_sec_ev, bool& find_ev, bool& unwrap_peigvec,
                    std::string& filepath_EV, double& tol, int& iterations, int& ncycles, std::ofstream& outfile);

// ARPACK_sector.h:12-37: the n_it lowest eigenvalues ("SR"; n_it is lowered
// to n-2 when n_it >= n-1) into Evals[0..n_it) largest first, as zneupd
// returns them, and with find_ev the Ritz vectors as Evecs(i, j) (n_it x n).
// Thick-restart Lanczos on the GPU (ED_LZ_THICK, ncv = max(4 n_it, 2 n_it + 40)
// vectors, maxiter restarts) in place of znaupd/zneupd.
void ARPACK_Iteration_EV(std::vector<int64_t>& pblock_states, std::vector<double>& norm,
                         std::vector<double>& norm_inv, std::vector<int64_t>& rep_loc,
                         std::vector<std::complex<double> >& state_ph, int& n_spins, int& n_it, double& px,
                         double& py, std::vector<std::complex<double> >& Evals, zMatrix& Evecs, bool find_ev,
                         double lambda, std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list,
                         double& ChCh, std::vector<bowties>& chch_list, double tol, int maxiter,
                         std::vector<int>& T1, std::vector<int>& T2, std::vector<int64_t>& Ia,
                         std::vector<int64_t>& Ib, int& bits_right, std::ofstream& outfile);

// ARPACK_sector.h:38-66: lowest n_ev eigenvalues ("SR") of the sector,
// largest first, with find_ev the vectors as rows of pblock_eigvecs
// (n_ev x n_p).  When n_p <= n_ev + 1 (ARPACK_sector.cpp:232-251) all n_p
// eigenpairs of the dense sector matrix, largest first, the vectors as the
// COLUMNS of pblock_eigvecs (n_p x n_p, zgeev's layout), and the matrix is
// printed to the log.
void ARPACK_sector(std::vector<int64_t>& pblock_states, std::vector<double>& norm, std::vector<double>& norm_inv,
                   std::vector<int64_t>& rep_loc, std::vector<std::complex<double> >& state_ph,
                   std::vector<std::complex<double> >& pblock_eigs, zMatrix& pblock_eigvecs, bool find_ev,
                   double& lambda, int& n_spins, double& px, double& py, int& n_ones, int64_t& n_p,
                   std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list, double& ChCh,
                   std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2, int n_ev,
                   double tol, int maxiter, std::vector<int64_t>& Ia, std::vector<int64_t>& Ib, int& bits_right,
                   std::ofstream& outfile);

// ARPACK_sector.h:67-86: n_ones = n_spins/2 plus one per unit step below cSz
// (ARPACK_sector.cpp:284-287), basis, ARPACK_sector, the energy table
// ascending, with find_ev the _EV<c>/_pEV<c> files numbered '1', '2', ...
// (:345-397).
void sector_eigs_eigvecs(int& kx, int& ky, double& cSz, int& n_spins, double& lambda, double& J2,
                         std::vector<coordinates>& adj_list, double& Ch, std::vector<triangles>& ch_list,
                         double& ChCh, std::vector<bowties>& chch_list, std::vector<int>& T1, std::vector<int>& T2,
                         int& n_sec_ev, bool& find_ev, bool& unwrap_peigvec, std::string& filepath_EV, double& tol,
                         int& maxiter, std::ofstream& outfile);

#endif  // ED_REFAPI_H
