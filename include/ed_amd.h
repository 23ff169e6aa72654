/* ed_amd.h -- C ABI of the MI355X (gfx950) momentum/Sz-basis H*v library.
 *
 * Plain C: handles, raw pointers, sizes and status codes; no C++ or torch
 * types cross this boundary.  Each entry point names the reference interface
 * (hiteshjc/ED_Code, paths relative to its root) it replaces.  The reference
 * signatures themselves (std::vector by reference) are provided on top of this
 * ABI by the C++ shims in ed_code_amd/csrc/ed_refapi.cpp.
 *
 * Conventions: a state is a bit string, bit s = site s, bit 1 = spin up.  A
 * translation T moves bit i to bit T[i].  Complex vectors are interleaved
 * (re, im) pairs: complex64 = 2 x float, complex128 = 2 x double.  Every call
 * returns ED_OK or an error code; ed_last_error() gives the message of the
 * last failure on the calling thread.
 */
#ifndef ED_AMD_H
#define ED_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ed_status {
    ED_OK = 0,
    ED_EINVAL = 1,   /* bad argument (replaces the reference's exit(1) paths) */
    ED_ENOMEM = 2,   /* device allocation failed or footprint over budget */
    ED_EDEVICE = 3,  /* HIP runtime error */
    ED_EEMPTY = 4,   /* empty sector (Lanczos_sector.cpp:93-97) */
    ED_ENOTCONV = 5  /* eigensolver did not converge */
};

enum ed_dtype { ED_C64 = 0, ED_C128 = 1 };

/* Hamiltonian terms: the reference's adj_list / ch_list / chch_list
 * (global.h:25-40, built by read_input_file.cpp:193-319).  Same order and
 * meaning; lambda is Jz (Delta).  Arrays are host memory, read at
 * ed_op_create. */
typedef struct ed_terms {
    double lambda;
    int32_t nbonds;
    const int32_t* bond_xy;     /* [2*nbonds] sites x, y            */
    const double* bond_J;       /* [nbonds]                         */
    const double* bond_phase;   /* [2*nbonds] exp(i phi) as re, im  */
    double Ch;
    int32_t nchiral;
    const int32_t* chiral_ijk;  /* [3*nchiral] i, j, k              */
    const double* chiral_phase; /* [2*nchiral]                      */
    double ChCh;
    int32_t nbowtie;
    const int32_t* bowtie;      /* [6*nbowtie] x, y, z, p, q, r     */
} ed_terms;

typedef struct ed_lattice_s* ed_lattice;
typedef struct ed_basis_s* ed_basis;
typedef struct ed_op_s* ed_op;

const char* ed_last_error(void);

/* Selects the HIP device used by the calling thread. */
int ed_init(int device);

/* Translation group from T1/T2 (read_input_file.cpp T1=/T2= keys).
 * L1, L2 = cycle lengths of site 0 (LxLy, Momentum_Basis.cpp:19-36).
 * Fails with ED_EINVAL when T1/T2 are not permutations of the same length
 * (read_input_file.cpp:423-427) or do not commute. */
int ed_lattice_create(int32_t n_sites, const int32_t* T1, const int32_t* T2, ed_lattice* out);
/* Three generators (read_input_file.cpp T3= key, ED_3D_Lattice_Geometry;
 * LxLyLz, Momentum_Basis.cpp:38-64); T3 = NULL is the identity, which is
 * ed_lattice_create. */
int ed_lattice_create3(int32_t n_sites, const int32_t* T1, const int32_t* T2, const int32_t* T3, ed_lattice* out);
int ed_lattice_orders(ed_lattice lat, int32_t* L1, int32_t* L2);
/* L3 (1 for two-generator lattices). */
int ed_lattice_order3(ed_lattice lat, int32_t* L3);
void ed_lattice_destroy(ed_lattice lat);

/* Momentum basis of the sector (n_ones, kx, ky), built on the GPU.
 * Replaces pConstruct_States (Momentum_Basis.cpp:645-778, decl
 * Momentum_Basis.h:151-165): representatives, orbit norms, the index
 * directory and the gather-image layout (chunk map, row positions) stay
 * resident in HBM.  kx, ky are integers in [0,L1) x [0,L2)
 * (Lanczos_sector.cpp:51-57); px = 2 pi kx / L1, py = 2 pi ky / L2. */
int ed_basis_build(ed_lattice lat, int32_t n_ones, int32_t kx, int32_t ky, ed_basis* out);
/* Three-generator sector (pConstruct_States, Momentum_Basis.cpp:781-897,
 * decl Momentum_Basis.h:68-100): the reference's walk over L1 x L2 x L3 with
 * cumulative T2 and T3 counts, pz = 2 pi kz / L3 (Lanczos_sector.cpp:240).
 * H*v, make_h and the solvers take such a basis through the generic kernel. */
int ed_basis_build3(ed_lattice lat, int32_t n_ones, int32_t kx, int32_t ky, int32_t kz, ed_basis* out);
int ed_basis_dim(ed_basis b, int64_t* D);
/* Host copies of pblock_states, norm, norm_inv ([D] each; any may be NULL). */
int ed_basis_export(ed_basis b, int64_t* reps, double* norm, double* norm_inv);
/* The reference's full-space tables, [C] with C = C(N, n_ones): rep_loc (index
 * of the representative of each state, -1 if its orbit is not in the sector)
 * and state_ph (complex128) in ascending state order, plus the Lin tables
 * Ia/Ib ([2^ceil(N/2)]) and bits_right (Momentum_Basis.cpp:697-777,
 * Lin_Tables.cpp:57-78).  Only the reference-signature shims and small-sector
 * paths need them; the GPU operator never reads them. */
int ed_basis_full_size(ed_basis b, int64_t* C, int64_t* nlin);
int ed_basis_export_full(ed_basis b, int64_t* rep_loc, double* state_ph, int64_t* Ia, int64_t* Ib,
                         int32_t* bits_right);
void ed_basis_destroy(ed_basis b);

/* Operator = basis + terms, terms uploaded once. */
int ed_op_create(ed_basis b, const ed_terms* terms, ed_op* out);
void ed_op_destroy(ed_op op);
/* Kernel selection for ed_hv (no reference counterpart; used by tests that
 * compare the paths and by benchmarks): path 0 = automatic (rotation fast path
 * when available; for the built-in lattice shapes the first H*v of each
 * precision on a sector of at least 2^20 rows times the kernel variants and
 * block sizes and keeps the fastest; the pick is kept for the process, and
 * later operators of the same shape, term class, momentum class, family
 * count, precision and size class take it without timing), 1 = generic
 * translation walk,
 * 2 = rotation fast path with the lattice shape compiled in (variant K=4
 * elements in flight), 3 = rotation fast path with the shape as run-time
 * parameters, 4 = as 2 with K=2 at 8 waves per SIMD, 5 = as 2 with K=8,
 * 6 = two-row 64-bit kernel (lattices relabelled to two rows of W sites),
 * 7 = as 2 with K=4 and the batches software-pipelined, 8 = as 6 gathering u
 * by row through the directory (the layout the kernel falls back to when its
 * wide gather image does not fit in device memory), 9 = as 2 with every row
 * split over two lanes (K=8 each), 10 = as 2 with every row split over four
 * lanes (K=4 each), 11 = as 9 with K=4 per lane.  Variants 2, 4, 5 and 7 sum
 * a row's elements in one order and give identical results; 9, 10 and 11 sum
 * each lane's share separately and agree with them to rounding (identical
 * among themselves for one lane count: 9 and 11).  Paths 9-11 fall back to 2
 * for more than 10 term families.  12 = stored element list: the first H*v
 * on this path generates every off-diagonal element of the sector once
 * (image element of the target and bond family, 4 bytes each, rows in groups
 * of 64 padded to the group's longest row; about 5 GB for the 32-site
 * square torus at Sz = 0, k = 0) and keeps it with the operator; later H*v
 * stream the list and gather, with no orbit search or rank, and give the
 * results of path 2.  Only for uniform real bond families (every member of a
 * family with the same J) at k = 0, whose image element and family pack into
 * 32 bits; never chosen by path 0.
 * Paths 2-5, 7 and 9-12 need a lattice whose translations are bit rotations,
 * N <= 32 and no bowtie terms; all but 3 also a built-in shape; 6 and 8 a
 * compiled two-row shape and no bowtie terms.
 * block = threads per block (64..512, multiple of 64; 0 = default, tuned
 * under path 0).  Resets any earlier tuning. */
int ed_op_tune(ed_op op, int32_t path, int32_t block);
/* Size of the operator's stored element list (path 12): the elements and the
 * 4-byte slots kept, padding included; 0 and 0 before it is built.  No
 * counterpart in the reference, whose sHv is matrix-free. */
int ed_op_elist_stats(ed_op op, int64_t* elements, int64_t* slots);
/* The configuration the next ed_hv of precision dtype runs: path (1, 2, 3, 6,
 * 8 or 12 as above), variant (0 = K=4, 1 = K=2 at 8 waves, 2 = K=8, 3 = K=4
 * pipelined, 4 = two lanes per row, 5 = four lanes per row, 6 = two lanes
 * per row with K=4, -1 = not yet tuned or not applicable) and threads per
 * block. */
int ed_op_tuning(ed_op op, int32_t dtype, int32_t* path, int32_t* variant, int32_t* block);
/* Forgets the autotuner picks kept for the process (no reference
 * counterpart): later operators time their candidates again.  Operators
 * already tuned keep their pick. */
int ed_tune_cache_clear(void);
/* Number of distinct nonzero off-diagonal (row, target) pairs, G, and the
 * total number of off-diagonal elements (one counting pass on the GPU; no
 * reference counterpart: the G of the algorithmic-bytes model). */
int ed_op_count_targets(ed_op op, int64_t* distinct, int64_t* total);

/* w = H v (sHv, H_v.cpp:192-357, decl H_v.h:36-57).  v, w are DEVICE
 * pointers to D complex numbers of type dtype; w is overwritten; v and w must
 * not alias.  stream is a hipStream_t (NULL = default stream).  The call is
 * asynchronous with respect to the host.  The operator owns one gather buffer
 * per precision (u = v/norm, written by each call before its row kernel):
 * calls on one operator and precision that may run at the same time on
 * different streams race on it, so order them (one stream, or events), or
 * give each stream its own operator. */
int ed_hv(ed_op op, int32_t dtype, const void* v, void* w, void* stream);
/* ed_hv that allocates nothing (no counterpart in the reference: for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and result as ed_hv, but the gather stream (u = v/norm by row, or the
 * sector's gather image) is built by every call in a fixed 64 MiB device
 * array of the library (one per GPU, held until the process ends) instead
 * of the operator's buffers (its device address is looked up by the first
 * call on each device), and the kernel variant is not timed (the one an
 * earlier ed_hv tuned, else the default).  Fails with ED_ENOMEM, before
 * launching anything, when the stream exceeds 64 MiB: 4,194,304 complex128
 * or 8,388,608 complex64 elements of the layout (the gather image holds
 * about 2-3 D elements).  All calls on one GPU share the array: calls that
 * may overlap on one GPU race, so serialise them (one stream, or events). */
int ed_hv_noalloc(ed_op op, int32_t dtype, const void* v, void* w, void* stream);
/* ed_hv with the dominant kernel bracketed by HIP events on `stream`:
 * *kernel_ms receives the duration of the row kernel alone (the v/norm
 * prescale pass is excluded).  Synchronises the stream.  No reference
 * counterpart: measurement helper for roofline accounting. */
int ed_hv_profile(ed_op op, int32_t dtype, const void* v, void* w, void* stream, float* kernel_ms);
/* Random-gather calibration (no reference counterpart: a measurement helper
 * that gives the H*v roofline a measured ceiling).  One launch gathers
 * n_gather elements of elem_bytes (8 = complex64, 16 = complex128) from
 * random 128-byte lines of a device buffer of n_elem elements, allocated and
 * freed inside the call; lanes in groups of `group` (a power of 2, 1..64)
 * share each line, inflight (4 or 8) loads per lane are in flight, threads
 * per block 64..512.  *ms = the best of three timed launches (HIP events).
 * Synchronous. */
int ed_gather_calibrate(int64_t n_elem, int32_t elem_bytes, int64_t n_gather, int32_t group, int32_t threads,
                        int32_t inflight, float* ms);
/* Same with HOST complex128 pointers (the reference sHv calling convention);
 * copies in and out, synchronous. */
int ed_hv_host(ed_op op, const double* v, double* w);
/* Dense D x D column-major complex128 H (makeH, H_v.cpp:21-190), D <= 20000. */
int ed_make_h(ed_op op, double* H);
/* w = A v for a caller-built sparse matrix in CSR form (the reference's
 * sparse sHv, H_v.cpp:3-19, decl H_v.h:31-33): n rows, rowptr[n+1] starting
 * at 0, col[rowptr[n]] in [0, ncols), val complex128 [rowptr[n]]; v
 * complex128 [ncols], w complex128 [n]; HOST pointers, synchronous. */
int ed_csr_hv_host(int64_t n, int64_t ncols, const int64_t* rowptr, const int64_t* col, const double* val,
                   const double* v, double* w);

/* Two-site spin correlations of momentum-basis vectors, on the device and
 * without unwrapping (the reference unwraps the eigenvector files and sums
 * over the full Sz basis: analysis/src/pspincorr.cpp:80-145 with szsz_op and
 * sxsxsysy_op, analysis/src/chirality_exp.cpp:232-289).  For each pair p,
 * (ij[2p], ij[2p+1]): out_zz[p] = <bra| Sz_i Sz_j |ket> and
 * out_pm[p] = <bra| (S+_i S-_j + S-_i S+_j)/2 |ket> (complex128 each), with bra
 * and ket DEVICE complex128 [D] vectors of basis b.  Both vectors must lie in
 * the sector's momentum space (as every vector of the basis does); the value
 * is that of any pair in the translation orbit of (i, j). */
int ed_pair_correlations(ed_basis b, const void* bra, const void* ket, int32_t npairs, const int32_t* ij,
                         double* out_zz, double* out_pm);

/* Unwrap_pState (Momentum_Basis.cpp:978-1037; three generators :901-975) on
 * the device: the full-Sz-basis components of the momentum-basis vector v
 * (DEVICE complex128 [D]) -- every state of each orbit with coefficient
 * v_a * exp(-i p.g) / Norm_a along the reference's walk, first visit kept,
 * entries with |re| or |im| > 1e-8, ascending by state.  *count receives the
 * number of entries; with capacity >= *count they are written to the HOST
 * arrays states [count] and coeffs (complex128 [count]); capacity 0 only
 * counts.  Needs C(N, n_ones) < 2^31. */
int ed_unwrap(ed_basis b, const void* v, int64_t capacity, int64_t* states, double* coeffs, int64_t* count);

/* Reduced density matrix of the cut A = sites cut[0..ncut) from two
 * momentum-basis vectors (DEVICE complex128 [D]) of basis b: dm(j, k) =
 * sum_e psi1(j, e) conj(psi2(k, e)), j, k configurations of A (bit l = site
 * cut[l]), e of the other sites, psi1, psi2 the unwrapped vectors with the
 * unwrap threshold (entries with |re| and |im| <= 1e-8 count as zero) --
 * density_matrixij (analysis/src/Entropy.cpp:170-205, driver
 * analysis/src/dm.cpp:10-88; with psi1 = psi2 density_matrix, :135-168).
 * n_cut_ones = -1: the whole 2^ncut x 2^ncut matrix; n_cut_ones >= 0: its
 * block of configurations of A with that many up spins, rows and columns by
 * ascending configuration (density_matrixij_Sz, :207-265).  dm is HOST
 * complex128, column-major (element (j, k) at j + k rows), at most 32768
 * rows.  Needs C(N, n_ones) < 2^31 and two translation generators. */
int ed_density_matrix(ed_basis b, const void* v1, const void* v2, int32_t ncut, const int32_t* cut,
                      int32_t n_cut_ones, double* dm);
/* Entropies of a HOST complex128 density matrix (column-major, n x n):
 * *renyi2 = -log2 sum_ij |dm_ij|^2 (Renyi_Entropy, Entropy.cpp:334-355),
 * *von_neumann = -sum lambda log2 lambda over its eigenvalues with |lambda| >
 * 1e-8 (Von_Neumann_Entropy, Entropy.cpp:307-331; eigenvalues from rocSOLVER). */
int ed_entanglement_entropies(const double* dm, int64_t n, double* renyi2, double* von_neumann);

/* All D eigenvalues (ascending) and, if evecs is non-NULL, the eigenvectors
 * (column-major D x D complex128, column k = eigenvector k) of the sector
 * matrix, D <= 20000, on the GPU (exact_evecs, lanczos.cpp:81-131, decl
 * lanczos.h:8-38; the reference diagonalises with LAPACK zheev,
 * Matrix_Functions.cpp:28-48). */
int ed_dense_eigh(ed_op op, double* evals, double* evecs);

/* Lanczos (lanczos_evecs, lanczos.cpp:133-349, decl lanczos.h:71-99).
 * mode ED_LZ_REFERENCE keeps the reference's restart schedule: ncycles cycles
 * of `iterations` steps, deflation against earlier eigenvectors every step,
 * full Gram-Schmidt against the cycle's vectors in the last cycle only, early
 * jump to the last cycle when |E0 - E0_prev| < tol.  mode ED_LZ_FULL runs one
 * cycle with full reorthogonalisation every step.  precision ED_C128 keeps
 * every vector in complex128; ED_C64 stores the Krylov vectors in complex64
 * and keeps FP64 for alpha/beta, reorthogonalisation coefficients, the
 * tridiagonal eigensolve and a final FP64 residual / Rayleigh-quotient
 * correction.  The start vector is (2u-1) + i(2u'-1) with u, u' from
 * splitmix64(seed*K + 2i), (seed*K + 2i + 1) (the reference uses unseeded
 * rand(), lanczos.cpp:192); eigenvalue number m uses seed + m.  mode
 * ED_LZ_LEAN is the memory-lean solver for sectors whose Krylov basis does
 * not fit (N=40: 27.6 GB per complex64 vector): three vectors, no
 * reorthogonalisation, lowest eigenvalue only (nev must be 1), stopping when
 * it changes by less than tol*max(1,|E0|) over 5 steps; a second pass
 * regenerates the same vectors to assemble the FP64 Ritz vector, whose FP64
 * residual and Rayleigh quotient go to stats (the eigenvalue reported is the
 * Rayleigh quotient).  Setting the environment variable ED_PROGRESS prints
 * one line per step to stderr. */
/* ED_LZ_THICK is thick-restart Lanczos, the ARPACK-like solver of the "SR"
 * drivers (ARPACK_sector.cpp:4-187, znaupd with ncv Arnoldi vectors): the
 * nev lowest eigenpairs with at most iterations (= ncv) Krylov vectors, each
 * step reorthogonalised against all kept vectors, restarted from the
 * nev + (ncv - nev)/2 lowest Ritz vectors and the residual until every
 * wanted residual |beta y| <= tol max(1, |theta|) or ncycles restarts. */
enum ed_lanczos_mode { ED_LZ_REFERENCE = 0, ED_LZ_FULL = 1, ED_LZ_LEAN = 2, ED_LZ_THICK = 3 };

/* flags: ED_LZ_SELECTIVE makes the second Gram-Schmidt pass of ED_LZ_FULL
 * conditional: it runs only when the first pass removed more than
 * 1 - 1/sqrt(2) of the vector's norm (Daniel-Gragg-Kaufman-Stewart).
 * ED_LZ_GRAPH keeps every step's scalars on the device and replays captured
 * hipGraphs instead of launching each step's kernels from the host:
 * - with ED_LZ_LEAN, three steps are one graph; the host reads alpha/beta
 *   every 6 steps (so up to 5 steps past the convergence point run, and are
 *   used);
 * - with ED_LZ_FULL, one step (H*v, the three-term update, both Gram-Schmidt
 *   passes over all Krylov vectors and earlier eigenvectors, alpha, beta,
 *   the next vector) is one graph replayed for every k, the step index read
 *   by the kernels from device memory; the host reads alpha/beta every 5
 *   steps, where the eager solver tests convergence, so the steps, stopping
 *   point and eigenpairs are the eager solver's (to reduction rounding);
 *   with ED_LZ_SELECTIVE the second pass is skipped on the device. */
enum ed_lanczos_flags { ED_LZ_SELECTIVE = 1, ED_LZ_GRAPH = 2 };

typedef struct ed_lanczos_params {
    int32_t nev;
    int32_t iterations;
    int32_t ncycles;
    int32_t precision;   /* ed_dtype */
    int32_t mode;        /* ed_lanczos_mode */
    int32_t flags;       /* ed_lanczos_flags */
    double tol;
    uint64_t seed;
} ed_lanczos_params;

typedef struct ed_lanczos_stats {
    int64_t hv_calls;
    double hv_seconds;      /* device time in H*v kernels */
    double total_seconds;
    double residual[16];    /* FP64 ||H x - E x|| of the first 16 eigenvectors */
    double rayleigh[16];    /* FP64 Rayleigh quotient x^H H x / x^H x */
    /* Error bound of every returned eigenvalue against the exact one, in
     * FP64: the FP64 residual ||H x - E x|| of its unit Ritz vector x (for
     * Hermitian H some eigenvalue lies within that distance of E).  ed_lanczos:
     * [k] for k < min(nev, 64); ed_block_lanczos: [m*nev + k] in the order of
     * evals, for m*nev + k < 64.  0 where nothing is returned. */
    double eval_error[64];
} ed_lanczos_stats;

/* evals [nev]; evecs (host, complex128 [nev][D]) may be NULL; stats may be NULL. */
int ed_lanczos(ed_op op, const ed_lanczos_params* p, double* evals, double* evecs, ed_lanczos_stats* stats);

/* ---- Block operators: M coupling instances on one sector ----------------
 * The reference solves one coupling set per run (ED.cpp:118-128 ->
 * sector_lanczos) and sweeps sectors one at a time (ARPACK_full_eigs.cpp:
 * 169-277); a block holds M term sets on one basis and applies all M
 * Hamiltonians in one pass over the rows, sharing the orbit search, rank and
 * gather addressing of every gathered element.  The term sets must have the same
 * bond, chiral and bowtie site lists (same order); J, Jz, Ch, ChCh and phases
 * may differ per instance.  M in 1..8; M in {2, 4, 8} on a built-in torus
 * shape without bowtie terms, whose coefficient and rank tables fit the LDS
 * of one workgroup (160 KiB on gfx950), runs the shared kernel; other cases
 * apply the instances one by one. */
typedef struct ed_block_s* ed_block;
int ed_block_create(ed_basis basis, int32_t M, const ed_terms* terms, ed_block* out);
void ed_block_destroy(ed_block blk);
/* 1 when the shared block kernel applies, 0 when instances run one by one. */
int ed_block_fast(ed_block blk);
/* W = [H_0 V_0, ..., H_{M-1} V_{M-1}] for interleaved DEVICE vectors:
 * element (i, m) at index i*M + m (row-major D x M), complex of type dtype.
 * V and W must not alias.  Asynchronous on `stream`. */
int ed_block_hv(ed_block blk, int32_t dtype, const void* V, void* W, void* stream);
/* Lowest p->nev eigenvalues of each instance (evals[m*nev + k], ascending),
 * all instances advanced together.  p->mode ED_LZ_FULL: full-
 * reorthogonalisation Lanczos, one eigenvalue at a time with deflation
 * against the earlier Ritz vectors, at most p->iterations steps each; an
 * instance stops when the residual estimate of its lowest Ritz pair is below
 * p->tol * max(1, |E|).  ED_LZ_THICK: thick-restart Lanczos with
 * p->iterations Krylov vectors per instance and at most p->ncycles
 * restarts, the nev lowest pairs of each instance from one space, until
 * every instance's nev residual estimates are below p->tol * max(1, |E|).
 * Krylov vectors are stored in p->precision, alpha/beta and
 * reorthogonalisation coefficients are FP64.  With complex64 storage the run
 * stops at residual 1e-6 and, if p->tol is smaller, its Ritz vectors start a
 * complex128 run that reaches p->tol (ED_LZ_THICK: from the sum of each
 * instance's Ritz vectors); each eigenvalue is then the FP64 Rayleigh
 * quotient of its Ritz vector.  M must be 1, 2, 4 or 8.  stats (optional):
 * H*V calls and times, and the FP64 residual ||H x - E x|| and Rayleigh
 * quotient of the lowest Ritz vector of each instance (residual[m],
 * rayleigh[m]).  p->ncycles is ignored by ED_LZ_FULL. */
int ed_block_lanczos(ed_block blk, const ed_lanczos_params* p, double* evals, ed_lanczos_stats* stats);

/* ---- Multi-sector block operators: M coupling instances on each of nsec
 * sectors of one lattice and one n_ones in one launch ----------------------
 * The reference sweeps the sectors one ARPACK run at a time
 * (ARPACK_full_eigs.cpp:169-277, one coupling set per run).  A multi-sector
 * block applies the M Hamiltonians of every sector in one launch over a
 * queue of (sector, 256-row tile) items (uneven dimensions are uneven tile
 * counts); terms as ed_block_create (same site lists, couplings per
 * instance, the same for every sector).  Vectors: the sectors' interleaved
 * D_s x M blocks (element (i, m) at i*M + m) concatenated in sector order;
 * ed_mblock_rows gives each sector's first row and D_s (sector = nsec: the
 * total).  When a sector cannot take the shared block kernel (no compiled
 * torus shape, bowtie terms, M not in {2, 4, 8}) the sectors run one
 * launch each; ed_mblock_fast says which. */
typedef struct ed_mblock_s* ed_mblock;
int ed_mblock_create(const ed_basis* bases, int32_t nsec, int32_t M, const ed_terms* terms, ed_mblock* out);
void ed_mblock_destroy(ed_mblock mb);
int ed_mblock_rows(ed_mblock mb, int32_t sector, int64_t* row0, int64_t* D);
int ed_mblock_fast(ed_mblock mb);
/* W = H V for all sectors and instances (DEVICE vectors, layout above).
 * Asynchronous on `stream`. */
int ed_mblock_hv(ed_mblock mb, int32_t dtype, const void* V, void* W, void* stream);
/* ed_block_lanczos for every (sector, instance) at once, in lock step: each
 * Lanczos step is one multi-sector H*V launch, then every sector's
 * reorthogonalisation.  Krylov storage is (p->iterations + 1) vectors of all
 * sectors (the complex128 polish of a complex64 run at most 61): choose the
 * sector set to fit device memory.  evals[(s*M + m)*nev
 * + k] ascending per instance; errors (optional, same layout) the FP64
 * residual bound of each eigenvalue; stats as ed_block_lanczos (residual and
 * rayleigh of the lowest pair of the first 16 (sector, instance) pairs). */
int ed_mblock_lanczos(ed_mblock mb, const ed_lanczos_params* p, double* evals, double* errors,
                      ed_lanczos_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* ED_AMD_H */
