// Drives entry points of include/ed_refapi.h that the reference's own main
// (ED.cpp, which fixes two generators at :69 and never calls makeH, the
// sparse sHv or ARPACK_Iteration_EV directly) does not reach, the way the
// reference's callers use them (Lanczos_sector.cpp:201-367 calls
// pConstruct_States and lanczos_evecs; lanczos.cpp:116 makeH; lanczos.cpp:232
// sHv; ARPACK_sector.cpp:226 ARPACK_Iteration_EV; ARPACK_full_eigs.cpp:86
// builds an sMatrix).  Linked against libed_amd.so.
//
//   driver hv     <spec> <v.bin> <w.bin> <reps.bin>  3-D pConstruct_States + sHv
//   driver sector <spec> <log> <vecpath> <unwrap>     3-D sector_lanczos
//   driver makeh  <spec> <H.bin>                      2-D pConstruct_States + makeH
//   driver sparse <spec> <v.bin> <w.bin>              makeH -> sMatrix -> sparse sHv
//   driver arpack <spec> <evals.bin> <evecs.bin>      2-D ARPACK_Iteration_EV
//   driver arpsec <spec> <evals.bin> <evecs.bin> <log> 2-D ARPACK_sector (both branches)
//   driver eigs   <spec> <log> <vecpath> <unwrap>     sector_eigs_eigvecs
//
// spec (whitespace separated): N n_ones kx ky kz lambda iterations ncycles
// n_ev tol, then T1[N], T2[N], T3[N], then nbonds and per bond x y J re im.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "ed_refapi.h"

namespace {

struct Spec {
    int N = 0, n = 0, kx = 0, ky = 0, kz = 0, iterations = 0, ncycles = 0, nev = 0;
    double lambda = 1, tol = 1e-10;
    std::vector<int> T1, T2, T3;
    std::vector<coordinates> adj;
};

Spec read_spec(const char* path) {
    std::ifstream f(path);
    Spec s;
    f >> s.N >> s.n >> s.kx >> s.ky >> s.kz >> s.lambda >> s.iterations >> s.ncycles >> s.nev >> s.tol;
    for (auto* T : {&s.T1, &s.T2, &s.T3}) {
        T->resize(s.N);
        for (int& t : *T) f >> t;
    }
    int nb = 0;
    f >> nb;
    for (int b = 0; b < nb; b++) {
        coordinates c;
        double re, im;
        f >> c.x >> c.y >> c.J >> re >> im;
        c.expphase = std::complex<double>(re, im);
        s.adj.push_back(c);
    }
    if (!f) {
        std::fprintf(stderr, "driver: bad spec %s\n", path);
        std::exit(2);
    }
    return s;
}

int orbit_len(const std::vector<int>& T) {
    // order of the permutation's cycle through site 0 (LxLyLz convention)
    int L = 1;
    for (int i = T[0]; i != 0; i = T[i]) L++;
    return L;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: driver hv|sector|makeh|sparse|arpack <spec> ...\n");
        return 2;
    }
    const std::string mode = argv[1];
    Spec s = read_spec(argv[2]);
    std::vector<triangles> ch;
    std::vector<bowties> chch;
    double Ch = 0, ChCh = 0;
    if (mode == "hv" && argc == 6) {
        std::ofstream log("/dev/null");
        const int L1 = orbit_len(s.T1), L2 = orbit_len(s.T2), L3 = orbit_len(s.T3);
        double px = 2 * s.kx * M_PI / L1, py = 2 * s.ky * M_PI / L2, pz = 2 * s.kz * M_PI / L3;
        std::vector<int64_t> reps, rep_loc, Ia, Ib;
        std::vector<double> norm, norm_inv;
        std::vector<std::complex<double> > state_ph;
        int64_t D = 0;
        int bits_right = 0;
        pConstruct_States(reps, norm, norm_inv, s.T1, s.T2, s.T3, s.n, px, py, pz, D, rep_loc, state_ph, Ia, Ib,
                          bits_right, log);
        std::vector<std::complex<double> > v(D), w(D);
        FILE* fv = std::fopen(argv[3], "rb");
        if (!fv || std::fread(v.data(), sizeof(v[0]), D, fv) != (size_t)D) {
            std::fprintf(stderr, "driver: v.bin must hold D=%lld complex128 values\n", (long long)D);
            return 2;
        }
        std::fclose(fv);
        sHv(v.data(), w.data(), reps, norm, norm_inv, rep_loc, state_ph, s.lambda, s.adj, Ch, ch, ChCh, chch, s.N, Ia,
            Ib, bits_right, s.T1, s.T2, s.T3, px, py, pz, log);
        FILE* fw = std::fopen(argv[4], "wb");
        std::fwrite(w.data(), sizeof(w[0]), D, fw);
        std::fclose(fw);
        FILE* fr = std::fopen(argv[5], "wb");
        std::fwrite(reps.data(), sizeof(reps[0]), D, fr);
        std::fclose(fr);
        std::printf("D %lld\n", (long long)D);
        return 0;
    }
    if (mode == "sector" && argc == 6) {
        std::ofstream log(argv[3]);
        double cSz = 0.5 * s.n - 0.5 * (s.N - s.n), J2 = 0;
        std::string vecpath = argv[4];
        bool find_ev = !vecpath.empty() && vecpath != "-", unwrap = std::atoi(argv[5]) != 0;
        sector_lanczos(s.kx, s.ky, s.kz, cSz, s.N, s.lambda, J2, s.adj, Ch, ch, ChCh, chch, s.T1, s.T2, s.T3, s.nev,
                       find_ev, unwrap, vecpath, s.tol, s.iterations, s.ncycles, log);
        return 0;
    }
    if (mode == "eigs" && argc == 6) {
        // sector_eigs_eigvecs (ARPACK_sector.cpp:262-401): maxiter = ncycles
        std::ofstream log(argv[3]);
        double cSz = 0.5 * s.n - 0.5 * (s.N - s.n), J2 = 0;
        std::string vecpath = argv[4];
        bool find_ev = !vecpath.empty() && vecpath != "-", unwrap = std::atoi(argv[5]) != 0;
        int maxiter = s.ncycles;
        sector_eigs_eigvecs(s.kx, s.ky, cSz, s.N, s.lambda, J2, s.adj, Ch, ch, ChCh, chch, s.T1, s.T2, s.nev, find_ev,
                            unwrap, vecpath, s.tol, maxiter, log);
        return 0;
    }
    if ((mode == "makeh" && argc == 4) || (mode == "sparse" && argc == 5) || (mode == "arpack" && argc == 5) ||
        (mode == "arpsec" && argc == 6)) {
        std::ofstream log("/dev/null");
        const int L1 = orbit_len(s.T1), L2 = orbit_len(s.T2);
        double px = 2 * s.kx * M_PI / L1, py = 2 * s.ky * M_PI / L2;
        std::vector<int64_t> reps, rep_loc, Ia, Ib;
        std::vector<double> norm, norm_inv;
        std::vector<std::complex<double> > state_ph;
        int64_t D = 0;
        int bits_right = 0, N = s.N;
        pConstruct_States(reps, norm, norm_inv, s.T1, s.T2, s.n, px, py, D, rep_loc, state_ph, Ia, Ib, bits_right,
                          log);
        if (mode == "arpsec") {
            // ARPACK_sector as sector_eigs_eigvecs calls it (ARPACK_sector.cpp:324-327)
            std::ofstream alog(argv[5]);
            std::vector<std::complex<double> > evals;
            zMatrix evecs;
            int n_ones = s.n;
            ARPACK_sector(reps, norm, norm_inv, rep_loc, state_ph, evals, evecs, true, s.lambda, N, px, py, n_ones, D,
                          s.adj, Ch, ch, ChCh, chch, s.T1, s.T2, s.nev, s.tol, s.ncycles, Ia, Ib, bits_right, alog);
            FILE* fe = std::fopen(argv[3], "wb");
            std::fwrite(evals.data(), sizeof(evals[0]), evals.size(), fe);
            std::fclose(fe);
            FILE* fv = std::fopen(argv[4], "wb");
            std::fwrite(evecs.data(), sizeof(evecs[0]), evecs.size(), fv);
            std::fclose(fv);
            std::printf("D %lld nev %zu rows %lld cols %lld\n", (long long)D, evals.size(), (long long)evecs.NRows(),
                        (long long)evecs.NCols());
            return 0;
        }
        if (mode == "arpack") {
            std::vector<std::complex<double> > evals;
            zMatrix evecs;
            int n_it = s.nev;
            ARPACK_Iteration_EV(reps, norm, norm_inv, rep_loc, state_ph, N, n_it, px, py, evals, evecs, true,
                                s.lambda, s.adj, Ch, ch, ChCh, chch, s.tol, s.iterations, s.T1, s.T2, Ia, Ib,
                                bits_right, log);
            FILE* fe = std::fopen(argv[3], "wb");
            std::fwrite(evals.data(), sizeof(evals[0]), evals.size(), fe);
            std::fclose(fe);
            FILE* fv = std::fopen(argv[4], "wb");
            std::fwrite(evecs.data(), sizeof(evecs[0]), evecs.size(), fv);
            std::fclose(fv);
            std::printf("D %lld nev %zu rows %lld\n", (long long)D, evals.size(), (long long)evecs.NRows());
            return 0;
        }
        zMatrix H;
        makeH(reps, norm, norm_inv, rep_loc, state_ph, s.lambda, s.adj, Ch, ch, ChCh, chch, N, Ia, Ib, bits_right,
              s.T1, s.T2, px, py, log, H);
        if (mode == "makeh") {
            FILE* fh = std::fopen(argv[3], "wb");
            std::fwrite(H.data(), sizeof(H[0]), H.size(), fh);
            std::fclose(fh);
            std::printf("D %lld\n", (long long)D);
            return 0;
        }
        // the sparse form of the same matrix, row by row, then w = H v
        sMatrix P(D, D);
        for (int64_t i = 0; i < D; i++)
            for (int64_t j = 0; j < D; j++)
                if (H(i, j) != std::complex<double>(0.0, 0.0)) P.push_back(i, j, H(i, j));
        std::vector<std::complex<double> > v(D), w(D);
        FILE* fv = std::fopen(argv[3], "rb");
        if (!fv || std::fread(v.data(), sizeof(v[0]), D, fv) != (size_t)D) {
            std::fprintf(stderr, "driver: v.bin must hold D=%lld complex128 values\n", (long long)D);
            return 2;
        }
        std::fclose(fv);
        sHv(P, v.data(), w.data());
        FILE* fw = std::fopen(argv[4], "wb");
        std::fwrite(w.data(), sizeof(w[0]), D, fw);
        std::fclose(fw);
        std::printf("D %lld\n", (long long)D);
        return 0;
    }
    std::fprintf(stderr, "driver: bad arguments\n");
    return 2;
}
