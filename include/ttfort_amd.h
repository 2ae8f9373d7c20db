/* ttfort_amd.h -- C ABI of the MI355X (gfx950) local-solver library for
 * tensor-train sweeps, a drop-in for the local-operator hot path of
 * oseledets/tt-fort.
 *
 * Two families of entry points:
 *
 *  1. Reference-compatible symbols.  The tt-fort `bind(c)` subroutines of
 *     tt_linalg.f90 (dbfun3, dqr, drow_add, drow_cut, dtransp, dperm1324,
 *     djac_gen, djac_apply, duchol_fort, dgmresr_hh_fort) under the same
 *     names with the same by-reference argument lists (Fortran default
 *     integers, host pointers, column-major arrays), and C versions of the
 *     local-operator callbacks the reference hands PRIMME and Expokit.  A
 *     program that links this library instead of tt_linalg.o reaches the
 *     GPU without source changes.  Workspace arguments (res1, res2, work,
 *     tau, work1) are accepted and ignored: the library keeps its own
 *     device scratch.  These calls stage host arrays through PCIe on every
 *     call.
 *
 *  2. Native `ttf_*` entry points.  Plain int sizes, device pointers,
 *     int status returns (0 = success), blocks of vectors, and the
 *     device-resident sweep drivers.  These are the fast path.
 *
 * Layout conventions (the reference's): TT core k is (r_{k-1}, n_k, r_k)
 * column-major; TT-matrix core k is (ra_{k-1}, n_k, m_k, ra_k) with n the
 * row (output) index; left interface phi1 is (ry, rx, ra), right interface
 * phi2 is (rx, ra, ry).
 *
 * Threading: one library state per process and device; calls from several
 * host threads must be serialised by the caller.  Every kernel runs on the
 * library's stream; native entry points are asynchronous unless stated.
 */
#ifndef TTFORT_AMD_H
#define TTFORT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fortran default INTEGER of the calling program: int32 for the reference's
 * i4 presets, int64 for its i8 presets (Makefile.in:7-15).  The library is
 * built once per kind: libttfort_amd.so (int32) and libttfort_amd_i8.so. */
#ifdef TTF_FINT64
typedef int64_t ttf_fint;
#else
typedef int32_t ttf_fint;
#endif

typedef struct { double re, im; } ttf_zdouble;
typedef struct { float re, im; } ttf_zfloat;

/* ------------------------------------------------------------------ runtime */
/* Select the device and create the library stream (idempotent).  Every other
 * entry point initialises lazily on device 0 if this was not called. */
int ttf_init(int device);
/* Release scratch arenas and the stream. */
int ttf_finalize(void);
/* Last error message ("" if none); ttf_clear_error resets it. */
const char* ttf_last_error(void);
void ttf_clear_error(void);
/* Library version string. */
const char* ttf_version(void);

/* Precision of the matvecs inside the iterative local solvers: 64
 * (default, FP64 throughout) or 32.  With 32, the local GMRES of
 * ttf_tt_amen_solve multiplies its Krylov directions on the fp32 matrix
 * cores with fp32 accumulation while every restart residual rhs - B x and
 * the solver's own residual checks stay FP64 (FP64 iterative refinement);
 * the block eigensolver of ttf_tt_eigb multiplies its search directions in
 * fp32 and recomputes A X in FP64 before any stopping decision, which is
 * always taken on FP64 residuals; the local exponentials of ttf_ztt_ksl
 * multiply the Krylov vectors by a complex64 copy of the local operator
 * (S-steps, norm estimates, Krylov basis, Hessenberg matrix and its
 * exponential stay complex128) at local tolerance max(1e-8, 1e-6), the
 * floor the complex64 rounding leaves the Krylov error estimate.  Not in
 * the reference. */
int ttf_set_local_precision(int bits);
int ttf_get_local_precision(void);

/* Device memory helpers (used by tests and benchmarks; any hipMalloc'ed
 * pointer is accepted by the native entry points). */
int ttf_malloc(void** ptr, size_t bytes);
int ttf_free(void* ptr);
int ttf_memcpy_h2d(void* dst, const void* src, size_t bytes);
int ttf_memcpy_d2h(void* dst, const void* src, size_t bytes);
int ttf_memcpy_d2d(void* dst, const void* src, size_t bytes);
int ttf_memset(void* dst, int value, size_t bytes);
/* Wait for all work on the library stream. */
int ttf_sync(void);
/* Events on the library stream, for timing: ttf_timer_start records a start
 * event, ttf_timer_stop records a stop event, waits for it and returns the
 * elapsed milliseconds. */
int ttf_timer_start(void);
int ttf_timer_stop(float* ms);
/* GEMM tile autotuning (on by default; env TTF_AUTOTUNE=0 disables).  The
 * first product of each new shape times the tile / split-K candidates on the
 * library stream and caches the fastest; disabling clears the cache and uses
 * a fixed heuristic.  ttf_autotune_count returns how many shapes were tuned.
 * With env TTF_TUNE_FILE set, tuned shapes are appended to that file and
 * read back by later processes, which then launch the same configurations. */
int ttf_set_autotune(int enable);
long ttf_autotune_count(void);

/* ------------------------------------------- local operator (bfun3 family) */
/* y(c1,i,c2) = sum phi1(c1,b1,a1) A(a1,i,j,a2) phi2(b2,a2,c2) x(b1,j,b2)
 * for nvec vectors; vector v of x starts at x + v*ldx, of y at y + v*ldy.
 * y may alias x.  Replaces ttals.f90:9-32 / tt_linalg.f90:127-159 (dbfun3),
 * one vector at a time in the reference (tt_eigb.f90:23-25). */
int ttf_dbfun3(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
               int nvec, const double* phi1, const double* A, const double* phi2,
               const double* x, int64_t ldx, double* y, int64_t ldy);
/* ttf_dbfun3 for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid allocation).
 * No counterpart in the reference.  Same arguments, kernels and results as
 * ttf_dbfun3, but the stage intermediates and split-K slices live in one
 * fixed device array of 64 MiB (8 M doubles) per GPU, held until the process
 * ends, and no allocation or tuning run happens in the call.  Products
 * needing more scratch than that are refused with an error before any
 * launch (a block of 8 vectors at (r, n, r, R) = (128, 2, 128, 5) takes
 * 4.7 M doubles).
 * All calls on one GPU share its array: calls that overlap on one GPU (from
 * several host threads or concurrent streams) race.  Device pointers;
 * asynchronous on the library stream. */
int ttf_dbfun3_noalloc(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                       int nvec, const double* phi1, const double* A, const double* phi2,
                       const double* x, int64_t ldx, double* y, int64_t ldy);
/* Times the three dbfun3 GEMM stages: each stage is launched `iters` times
 * back to back between two events on the library stream (after one launch
 * that settles its tuning) and stage_ms[k] receives the mean milliseconds
 * per launch (for the benchmark's per-kernel roofline).  Synchronous; y and
 * the contraction arena are overwritten. */
int ttf_dbfun3_profile(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                       int nvec, const double* phi1, const double* A, const double* phi2,
                       const double* x, int64_t ldx, double* y, int64_t ldy, int iters,
                       float* stage_ms);
/* Times the whole FP64 dbfun3 as the library runs it (the fused one-launch
 * kernel and its zero fill where the shape fits, else the GEMM chain):
 * `iters` back-to-back calls between two events on the library stream,
 * after one call that settles any tuning; *ms receives the mean
 * milliseconds per call and *fused (optional) 1 if the fused kernel ran.
 * Synchronous; y is overwritten. */
int ttf_dbfun3_chain_profile(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2, int nvec,
                             const double* phi1, const double* A, const double* phi2, const double* x,
                             int64_t ldx, double* y, int64_t ldy, int iters, float* ms, int* fused);
/* Same contraction in FP32 on the FP32 matrix cores (mixed-precision path). */
int ttf_sbfun3(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
               int nvec, const float* phi1, const float* A, const float* phi2,
               const float* x, int64_t ldx, float* y, int64_t ldy);
/* Complex128 version (ttals.f90:34-57, zbfun3). */
int ttf_zbfun3(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
               int nvec, const ttf_zdouble* phi1, const ttf_zdouble* A,
               const ttf_zdouble* phi2, const ttf_zdouble* x, int64_t ldx,
               ttf_zdouble* y, int64_t ldy);
/* Complex64 version (complex64 operands and intermediates on the fp32
 * matrix cores with fp32 accumulation): the reduced-precision matvec of the
 * KSL local exponentials (tt_ksl.f90:25-32 zmatvec). */
int ttf_cbfun3(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
               int nvec, const ttf_zfloat* phi1, const ttf_zfloat* A, const ttf_zfloat* phi2,
               const ttf_zfloat* x, int64_t ldx, ttf_zfloat* y, int64_t ldy);
/* y = B^H x for the zbfun3 operator B (ttals.f90:85-108, zbfun3_transp);
 * x has shape (ry1,n,ry2), y has shape (rx1,m,rx2). */
int ttf_zbfun3_transp(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                      int nvec, const ttf_zdouble* phi1, const ttf_zdouble* A,
                      const ttf_zdouble* phi2, const ttf_zdouble* x, int64_t ldx,
                      ttf_zdouble* y, int64_t ldy);

/* ------------------------------------------------------- interface updates */
/* phi1new(c2,b2,a2) = sum phi1(c1,b1,a1) y(c1,i,c2) A(a1,i,j,a2) x(b1,j,b2),
 * output (ry2, rx2, ra2).  Replaces tt_linalg.f90:287-312 / ttals.f90:203-227
 * (dphi_left). */
int ttf_dphi_left(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                  const double* phi1, const double* A, const double* x, const double* y,
                  double* phi1new);
/* phi2new(b1,a1,c1) = sum phi2(b2,a2,c2) x(b1,j,b2) A(a1,i,j,a2) y(c1,i,c2),
 * output (rx1, ra1, ry1).  Replaces tt_linalg.f90:259-284 / ttals.f90:150-173
 * (dphi_right). */
int ttf_dphi_right(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                   const double* phi2, const double* A, const double* x, const double* y,
                   double* phi2new);
/* Complex versions with conj(y) (ttals.f90:229-256 zphi_left, :176-200
 * zphi_right). */
int ttf_zphi_left(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                  const ttf_zdouble* phi1, const ttf_zdouble* A, const ttf_zdouble* x,
                  const ttf_zdouble* y, ttf_zdouble* phi1new);
int ttf_zphi_right(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                   const ttf_zdouble* phi2, const ttf_zdouble* A, const ttf_zdouble* x,
                   const ttf_zdouble* y, ttf_zdouble* phi2new);
/* Interfaces without an operator (y'x projections), tt_linalg.f90:329-340
 * (dphi2_left, output (ry2,rx2)) and :315-327 (dphi2_right, output (rx1,ry1)). */
int ttf_dphi2_left(int rx1, int rx2, int ry1, int n, int ry2, const double* phi1,
                   const double* x, const double* y, double* phi1new);
int ttf_dphi2_right(int rx1, int rx2, int ry1, int n, int ry2, const double* phi2,
                    const double* x, const double* y, double* phi2new);

/* Explicit local matrix B((c1,i,c2),(b1,j,b2)), N x N column-major with
 * N = rx1*m*rx2 (tt_eigb.f90:88-105, d2d_fullmat). */
int ttf_d2d_fullmat(int rx1, int m, int rx2, int ra1, int ra2, const double* phi1,
                    const double* A, const double* phi2, double* B);

/* ------------------------------------------------ dense linear algebra */
/* Thin SVD of an m x n column-major device matrix: A = U diag(S) VT, k =
 * min(m,n) singular values in descending order, U (m x k, ldu), VT (k x n,
 * ldvt).  Blocked one-sided Jacobi; replaces the dgesvd calls of the
 * truncation steps (tt_eigb.f90:312, :361; tt_adapt_als.f90:812, :930;
 * svd.f90:25).  *sweeps (optional) returns the Jacobi sweeps used.
 * Synchronous. */
int ttf_dgesvd(int m, int n, const double* A, int64_t lda, double* U, int64_t ldu, double* S,
               double* VT, int64_t ldvt, int* sweeps);
/* Singular values (S, k = min(m,n), descending) of an m x n device matrix
 * and the orthonormal singular vectors of one side only: side 'L' gives the
 * left vectors (m x k), side 'R' the right vectors (n x k), as the columns
 * of O (ldo).  The truncation steps need one side plus a product with A
 * (tt_eigb.f90:312-333 keeps V^T and absorbs U S; :361-376 keeps U and
 * absorbs S V^T).  When the wanted side is the shorter dimension the Jacobi
 * iteration runs on R^T of a Householder QR of the tall orientation.
 * Vectors of zero singular values are zero columns; *kpos (optional)
 * returns how many are nonzero, *sweeps (optional) the Jacobi sweeps.
 * Synchronous. */
int ttf_dsvd_side(int m, int n, const double* A, int64_t lda, char side, double* S, double* O,
                  int64_t ldo, int* kpos, int* sweeps);
/* Eigen-decomposition of `count` symmetric n x n device matrices (n <= 64)
 * stored consecutively (A is overwritten): eigenvalues to w (ascending, or
 * descending if desc), eigenvectors to Q.  Replaces the small dsyev calls
 * (tt_eigb.f90:238 for local sizes <= 64; tt_linalg.f90:630). */
/* Householder QR of A (n x m, ld n) on the device, in place: Q (n x k,
 * k = min(n, m)) into A's first k columns, R (k x m, ld ldr) upper
 * trapezoidal; LAPACK's reflector convention (tt_linalg.f90:30-51, dqr).
 * n < 16384. */
int ttf_dqr(int n, int m, double* A, double* R, int64_t ldr);
int ttf_dsyev_small(int n, int count, double* A, double* w, double* Q, int desc);
/* G (p x q, ldg) = X^T Y for tall X (N x p), Y (N x q); p, q <= 64. */
int ttf_dgram(int N, int p, int q, const double* X, int64_t ldx, const double* Y, int64_t ldy,
              double* G, int ldg);

/* ------------------------------------------------------- sweep drivers */
typedef struct {
  int sweeps;            /* sweeps performed */
  long total_matvecs;    /* local matvecs (vectors) over the run */
  double last_ermax;     /* max relative decrease of sum(lambda) over the last
                            completed sweep (tt_eigb.f90:294-295), the quantity
                            the stopping rule compares with eps */
  int max_rank;          /* largest TT rank at exit */
  double seconds;        /* wall time of the call */
  double solve_seconds;  /* local eigensolves */
  double trunc_seconds;  /* SVD truncations and core moves */
  double phi_seconds;    /* interface updates */
  int nrec;              /* sweeps recorded below (at most 64) */
  double sweep_seconds[64];   /* wall time of each sweep */
  double sweep_flops[64];     /* algorithmic local-matvec flops of each sweep */
  long sweep_matvecs[64];     /* local matvecs (vectors) of each sweep */
  double sweep_ermax[64];     /* ermax of each sweep */
  double sweep_lambda_sum[64];   /* sum of the B eigenvalues at the end of each sweep */
  int converged;              /* 1 if a sweep ended with ermax < eps (the
                                 reference's exit, tt_eigb.f90:401-403), 0 if
                                 the run stopped at nswp sweeps */
  long total_matvecs32;       /* of total_matvecs, those on the fp32 operator
                                 (ttf_set_local_precision(32)) */
  long refreshes;             /* FP64 recomputations of A X in the local
                                 eigensolves (mixed precision) */
  long sweep_matvecs32[64];   /* fp32 local matvecs of each sweep */
  long block_prec_solves;     /* local eigensolves preconditioned by the shifted
                                 diagonal blocks of the local matrix */
  long block_prec_fallbacks;  /* local eigensolves whose shifted blocks were not
                                 positive definite (diagonal preconditioner) */
} ttf_eigb_stats;

/* Block eigensolver in TT format: the B lowest eigenpairs of the symmetric
 * TT-matrix A.  Argument layout of tt_eigb (tt_eigb.f90:110-120): d modes,
 * n(d)=m(d) mode sizes, ra(d+1) operator ranks, crA the concatenated
 * operator cores (ra_k, n_k, m_k, ra_k+1), crY0 the initial block TT with
 * ranks ry(d+1) (ry(d+1) == B, in/out), eps the tolerance, rmax the rank
 * cap, lambda(B) out.  kickrank is accepted and ignored, as in the
 * reference; nswp <= 0 means 20, max_full_size <= 0 means 2000.  verb > 0
 * prints a line per local solve and per sweep on stderr (the reference's
 * progress lines); verb > 1 adds a line per truncation with the Jacobi
 * sweeps of its SVD and the kept rank.  The
 * result cores (what the reference leaves in tt_block_eig::result_core)
 * are written to `result` (capacity result_cap doubles; pass NULL to query
 * the size, returned in *result_size).  Host pointers; device-resident run. */
int ttf_tt_eigb(int d, const int* n, const int* m, const int* ra, const double* crA,
                const double* crY0, int* ry, double eps, int rmax, double* lambda, int B,
                int kickrank, int nswp, int max_full_size, int verb, double* result,
                int64_t result_cap, int64_t* result_size, ttf_eigb_stats* stats);

typedef struct {
  int sweeps;           /* sweeps performed */
  double last_err_max;  /* largest relative core change of the last sweep */
  int converged;        /* 1 if that was below eps (tt_adapt_als.f90:441-447) */
  int max_rank;         /* largest rank of y at exit */
  double seconds;       /* wall time of the call */
  double flops;         /* algorithmic flops of the local projections */
} ttf_mvk4_stats;

/* Approximate TT-matrix by TT-vector product y ~ A x by alternating
 * projections with enrichment (AMEn-mv), argument layout of tt_mvk4
 * (tt_adapt_als.f90:67-500): d modes, A (ra_k, n_k, m_k, ra_k+1) cores in
 * crA, x (rx_k, m_k, rx_k+1) cores in crX, the initial y0 (ry_k, n_k,
 * ry_k+1) in crY0 with ranks ry (in/out), eps the accuracy; the truncation
 * is my_chop3 at eps / sqrt(d), the enrichment min(kickrank, n ry) pivoted
 * Cholesky directions of A_k Phi_k+1 x_k per backward step.  rmax sets the
 * rank capacity (the reference uses it only to size work arrays): a rank
 * beyond max(rmax, max ry0) is an error.  nswp <= 0 means 20.  The cores of
 * y (the reference's tt_adapt_als::result_core) go to `result`.  Host
 * pointers; device-resident run. */
int ttf_tt_mvk4(int d, const int* n, const int* m, const int* rx, const int* ra, const double* crA,
                const double* crX, const double* crY0, int* ry, double eps, int rmax, int kickrank, int nswp,
                int verb, double* result, int64_t result_cap, int64_t* result_size, ttf_mvk4_stats* stats);

/* Many independent tt_eigb problems solved concurrently on one GPU (no
 * counterpart in the reference, which is not re-entrant: tt_eigb.f90:6-8).
 * Chain c has d[c] modes; its mode sizes (n = m) are n_all[sum_{c'<c} d],
 * its operator and block-TT ranks ra_all / ry_all[sum_{c'<c} (d+1)] (ry in
 * and out, last rank B), its operator and initial cores at
 * crA_all + crA_off[c] and crY0_all + crY0_off[c], tolerance eps[c] and
 * rank cap rmax[c]; all chains share B and nswp.  Results: lambda_all
 * (B per chain), cores at result_all + result_off[c] (capacity
 * result_off[c+1] - result_off[c] doubles; sizes in result_size[c] when
 * given), per-chain stats when given, per-chain status (0 = solved).
 * nworkers host threads (<= 0: 8), each with its own stream and scratch,
 * take the chains largest first (d rmax^3).  Host pointers. */
int ttf_tt_eigb_batch(int nchains, const int* d, const int* n_all, const int* ra_all, const double* crA_all,
                      const int64_t* crA_off, const double* crY0_all, const int64_t* crY0_off, int* ry_all,
                      const double* eps, const int* rmax, int B, int nswp, int nworkers, double* lambda_all,
                      double* result_all, const int64_t* result_off, int64_t* result_size, ttf_eigb_stats* stats,
                      int* status);

/* ------------------------------------------- integer decisions (test hooks)
 * The rank decisions of the truncations, callable on their own so that they
 * can be compared bit for bit with a restatement on identical inputs. */
/* Rank of the relative chop rule my_chop3(n, s, eps) (tt_adapt_als.f90:40-61
 * == matrix_util.f90:254-275) as ttf_tt_eigb applies it to the singular
 * values of a block core: the reference's loop, including its "+1" and the
 * strict comparison.  Host arrays; needs no GPU.  -1 on bad arguments. */
int ttf_my_chop3(int n, const double* s, double eps);
/* Residual-based truncation rank of tt_amen_solve (tt_adapt_als.f90:819-833,
 * :937-950) on a given sequence of candidate residuals: errs[c] is the
 * relative local residual of the core truncated to rank c, c = 1..k-1
 * (errs[0] unused).  Scans c = k-1 down to 1 and returns one more than the
 * first c with errs[c] > 2 res_new and errs[c] > eps2, or 1.  The same scan
 * ttf_tt_amen_solve runs (in batches of 8 candidates per block matvec).
 * Host arrays; needs no GPU.  -1 on bad arguments. */
int ttf_residual_rank_rule(int k, const double* errs, double res_new, double eps2);
/* The whole residual-based rank search of one local step on the device, as
 * ttf_tt_amen_solve runs it: SVD of the local core x (rx1, n, rx2) unfolded
 * as (rx1 n) x rx2 for dir = +1 or rx1 x (n rx2) for dir = -1, candidates
 * formed from the singular factors, multiplied by the local operator
 * (phi1, A, phi2 as ttf_dbfun3), residuals against rhs, and the scan above.
 * *rank receives the rank before any rank cap; errs (host, optional,
 * min(rows, cols) entries) the residual of each candidate evaluated and -1
 * for the others.  Device pointers except errs; synchronous. */
int ttf_amen_truncation_rank(int dir, int rx1, int n, int rx2, int ra1, int ra2, const double* phi1,
                             const double* A, const double* phi2, const double* x, const double* rhs,
                             double res_new, double eps2, int* rank, double* errs);

/* The Frobenius chop of the typed solvers' truncation (svd.f90:86-111):
 * the smallest rank r with sum_{k>r} s_k^2 < (tol ||s||)^2, found by the
 * reference's scan from the end; rmax > 0 caps r first (its tail counted
 * as discarded), tol < 0 skips the tolerance.  Host arrays; needs no GPU. */
int ttf_chop(int n, const double* s, double tol, int rmax);

/* Shifted block preconditioner of the local eigensolver: the inverses of
 * M_k - sigma I for the ry2 diagonal blocks M_k (order ry1 n <= 256) of the
 * local matrix over the left interface and the mode, i.e. the 'l' blocks
 * of djac_gen (tt_linalg.f90:434-462) shifted by sigma; no counterpart in
 * the reference, whose PRIMME call preconditions nothing.  inv receives
 * (ry1 n)^2 ry2 doubles (layout of ttf_djac_gen 'l'); *positive_definite
 * is 0 (and inv undefined) when some shifted block is not positive
 * definite.  Device pointers. */
int ttf_eig_block_prec(int ry1, int n, int ry2, int ra1, int ra2, const double* phi1, const double* A,
                       const double* phi2, double sigma, double* inv, int* positive_definite);

/* ------------------------------------------------- AMEn building blocks */
/* Block-Jacobi preconditioner of one local step (tt_linalg.f90:363-495,
 * djac_gen bind(c)): ptype 'c' rx1*rx2 inverted n x n blocks (layout
 * (n, n, rx2, rx1)), 'l' rx2 blocks of (rx1 n)^2, 'r' rx1 blocks of
 * (n rx2)^2; blocks up to 128.  Device pointers; jacs holds
 * ttf_djac_size() doubles.  A singular block is an error. */
int ttf_djac_gen(char ptype, int rx1, int n, int rx2, int ra1, int ra2, const double* phi1,
                 const double* A, const double* phi2, double* jacs);
int64_t ttf_djac_size(char ptype, int rx1, int n, int rx2);
/* y = P x with the blocks of ttf_djac_gen (tt_linalg.f90:497-533,
 * djac_apply bind(c)); y may alias x.  Device pointers. */
int ttf_djac_apply(char ptype, int rx1, int n, int rx2, const double* jacs, const double* x, double* y);
/* Right-preconditioned restarted GMRES for the local system B P sol = rhs
 * (tt_linalg.f90:651-845, dgmresr_hh_fort bind(c)): at most niters cycles
 * of GMRES(min(nrestart, N, 63)), stopping when the residual estimate falls
 * below tol times the initial residual; sol starts at zero.  Arnoldi by
 * classical Gram-Schmidt applied twice.  iters (optional, 2 ints) returns
 * (last cycle, Krylov dimension), relres (optional) the final estimate.
 * Device pointers; synchronous. */
int ttf_dgmresr(const double* phi1, const double* A, const double* phi2, const double* rhs, int rx1,
                int n, int rx2, int ra1, int ra2, int nrestart, double tol, int niters, char ptype,
                const double* jacs, double* sol, int* iters, double* relres);
/* Pivoted incomplete Cholesky of the Gram matrix of the n items of dA
 * (columns of an m x n dA for trans 'N', rows of an n x m dA for 'T'):
 * up to numvec orthonormal directions du (n x numvec), count in
 * *real_numvec, eigenvalues in lam (numvec, optional) (tt_linalg.f90:540-643,
 * duchol_fort bind(c)).  Device pointers; synchronous. */
int ttf_duchol(char trans, int m, int n, const double* dA, int numvec, double* du, double* lam,
               int* real_numvec);
/* Phi1 A x with the right interface open: y(c1, i, b2, a2), an (ry1 n) x
 * (rx2 ra2) block, the contraction tt_linalg.f90:242-256 (dbfun3_left) is
 * meant to compute.  Device pointers. */
int ttf_dbfun3_left(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                    const double* phi1, const double* A, const double* x, double* y);
/* A Phi2 x with the left interface open: y(b1, a1, i, c2), an (rx1 ra1) x
 * (n ry2) block (tt_linalg.f90:210-240, dbfun3_right).  Device pointers. */
int ttf_dbfun3_right(int rx1, int m, int rx2, int ry1, int n, int ry2, int ra1, int ra2,
                     const double* A, const double* phi2, const double* x, double* y);

typedef struct {
  int sweeps;            /* sweeps performed */
  long total_matvecs;    /* local matvecs (vectors): residuals, GMRES, truncation */
  double last_res_max;   /* largest local residual of the last sweep */
  double last_err_max;   /* largest relative local correction of the last sweep */
  int max_rank;          /* largest TT rank at exit */
  double seconds;        /* wall time of the call */
  double solve_seconds;  /* local residuals and GMRES solves */
  double trunc_seconds;  /* SVD, residual-based truncation, kick, core moves */
  double phi_seconds;    /* interface updates */
  double flops;          /* algorithmic flops of the local matvecs */
  long learned_solves;   /* local solves with the learned preconditioner (prec 'm') */
  long learned_fallbacks;/* of those, solves finished by plain GMRES */
} ttf_amen_stats;

/* AMEn solver for A x = y in TT format, argument layout of tt_amen_solve
 * (tt_adapt_als.f90:513-525): d modes, n(d) = m(d) mode sizes, ry(d+1) and
 * crY the right-hand side, ra(d+1) and crA the operator, crX0 with ranks
 * rx(d+1) (in/out) the initial guess, eps the tolerance.  The reference's
 * optional arguments are explicit: kickrank (4; -1 truncates without kick,
 * 0 only re-orthogonalises), nswp (20), verb (1), prec ('n', 'c', 'l',
 * 'r', or 'm': the learned block preconditioner below in flexible GMRES,
 * with plain GMRES finishing any solve it leaves above the tolerance; not
 * in the reference), nrestart (40), niters (2).  rmax > 0 caps the TT ranks (not in the
 * reference; <= 0 means a capacity of 256).  The result cores (the
 * reference's tt_adapt_als::result_core) go to `result` (capacity
 * result_cap doubles; size in *result_size).  Host pointers;
 * device-resident run. */
int ttf_tt_amen_solve(int d, const int* n, const int* m, const int* ry, const int* ra, const double* crA,
                      const double* crY, const double* crX0, int* rx, double eps, int kickrank, int nswp,
                      int verb, char prec, int nrestart, int niters, int rmax, double* result,
                      int64_t result_cap, int64_t* result_size, ttf_amen_stats* stats);
/* AMEn solve of A x = y with an auxiliary residual TT z, the typed solver
 * dtt_amen_solve (ttamen.f90:29-252) on flat arrays (its type(dtt)
 * arguments are not C-interoperable): A (ra_k, n_k, n_k, ra_k+1) cores in
 * crA, y (ry_k, n_k, ry_k+1) in crY, the initial x in crX0 with ranks rx
 * (in/out; the solution's ranks on return); defaults of the reference:
 * kickrank 4, nswp 20, local_prec 'n', local_iters 2, local_restart 40,
 * trunc_norm 1 (truncation by local residual; 0: Frobenius chop),
 * max_full_size 50 (dense local solves below that local size, at most 128).
 * z starts from N(0,1) cores drawn from `seed` (the reference uses its
 * Fortran random number generator).  rmax <= 0 means a rank capacity of
 * 256; ranks may reach rmax + kickrank between truncations, and beyond that
 * the call fails.  The cores of x go to `result`.  Host pointers;
 * device-resident run.  stats: learned_* stay zero. */
int ttf_dtt_amen_solve(int d, const int* n, const int* ra, const double* crA, const int* ry, const double* crY,
                       double tol, int* rx, const double* crX0, int kickrank, int nswp, char local_prec,
                       int local_iters, int local_restart, int trunc_norm, int max_full_size, int verb, int rmax,
                       unsigned seed, double* result, int64_t result_cap, int64_t* result_size,
                       ttf_amen_stats* stats);


/* ------------------------------------- learned block preconditioner (MLP) */
/* No counterpart in the reference: a replacement for the block-Jacobi 'c'
 * solve of the local GMRES (tt_linalg.f90:397-429, :497-533) by a small
 * network trained on the GPU.  Per diagonal block B_b (b = (b1, b2), mode
 * size n = 64) of the local matrix Phi1 A Phi2, the features are the three
 * central diagonals of B_b divided by s_b = mean |diag B_b| and the residual
 * slice r_b / |r_b| (256 values, one float column per block); the network
 * y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3 + x_r / x_d (256-256-256-64,
 * bf16 weights, fp32 activations and accumulation; the last term is the
 * point-Jacobi solution of the block, residual features over diagonal
 * features) predicts the normalised block correction, returned as
 * y_b |r_b| / s_b.  Parameters are one flat float
 * vector [W1 (256x256), b1, W2 (256x256), b2, W3 (256x64), b3], matrices
 * column-major; ttf_mlp_param_count() gives its length. */
long ttf_mlp_param_count(void);
/* Uniform(+-1/sqrt(fan_in)) W1, W2 from `seed`; W3 and the biases zero,
 * so the untrained network is the point-Jacobi solve. */
int ttf_mlp_init(unsigned seed);
/* Host parameter vector in/out; setting resets the Adam state. */
int ttf_mlp_set_params(const float* params);
int ttf_mlp_get_params(float* params);
/* Y (64 x M) = network(X (256 x M)); device pointers. */
int ttf_mlp_forward(int M, const float* X, float* Y);
/* Loss L = sum (Y - T)^2 / (64 M) over the batch and its gradient with
 * respect to the parameters (taken at the bf16 weights); X, T device,
 * loss and grad (param_count floats) host. */
int ttf_mlp_loss_grad(int M, const float* X, const float* T, double* loss, float* grad);
/* Adam training (lr, beta 0.9 / 0.999, eps 1e-8) over `epochs` passes of
 * shuffled minibatches of `batch` columns of the device data X (256 x n),
 * T (64 x n); the mean batch loss of each epoch goes to loss_hist. */
int ttf_mlp_train(long nsamples, const float* X, const float* T, int epochs, int batch, float lr, unsigned seed,
                  double* loss_hist, int hist_cap, int* hist_len);
/* Features X (256 x rx1 rx2) and output scales |r_b| / s_b of the local
 * vector v (rx1, n, rx2) for the operator Phi1 A Phi2 (device pointers). */
int ttf_mlp_features(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                     const double* phi2, const double* v, float* X, float* scale);
/* z = learned preconditioner applied to v (device pointers). */
int ttf_mlp_precond(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                    const double* phi2, const double* v, double* z);
/* Training-pair collection: while active, every block-Jacobi 'c' solve of
 * the local GMRES (any ttf_tt_amen_solve with prec 'c', or ttf_dgmresr)
 * records up to per_apply sampled blocks as (features, normalised exact
 * block correction) into library-owned device buffers of `capacity`
 * columns; ttf_mlp_collected returns them. */
int ttf_mlp_collect_begin(long capacity, int per_apply);
int ttf_mlp_collect_end(long* count);
int ttf_mlp_collected(float** X, float** T, long* count);

/* ------------------------------------------------------ KSL time stepping */
/* exp(t H) for a small complex m x m column-major device matrix (m <= 40):
 * the irreducible (6, 6) Pade form with scaling and squaring of
 * expokit.f:767-935 (ZGPADM), one work-group.  Device pointers. */
int ttf_zexpm(int m, double t, const ttf_zdouble* H, ttf_zdouble* E);
/* w = exp(tau B) v for the local operator B of zbfun3 (rx1, n, rx2, ra1,
 * ra2; phi1 (rx1, rx1, ra1), A (ra1, n, n, ra2), phi2 (rx2, ra2, rx2)):
 * explib.f90:53-125 (zexp_mv) with the adaptive Krylov steps of ZGEXPV
 * (expokit.f:2403-2791), Krylov size `order` (<= 38), tolerance tol, the
 * operator norm from Hager's 1-norm estimate (returned in *anorm_out) and
 * the matvec count in *nmult.  Device pointers; synchronous. */
int ttf_zexp_bfun3(int rx1, int n, int rx2, int ra1, int ra2, const ttf_zdouble* phi1, const ttf_zdouble* A,
                   const ttf_zdouble* phi2, const ttf_zdouble* v, ttf_zdouble* w, double tau, int order,
                   double tol, double* anorm_out, int* nmult);
/* Q (rows x k, orthonormal columns) and R = Q^H A (k x cols), k =
 * min(rows, cols), for a complex device matrix with rows * cols <= 8192 and
 * k <= 64 (zqr, matrix_util.f90:64-90); rank-deficient blocks get an
 * orthonormal completion. */
int ttf_zqr(int rows, int cols, const ttf_zdouble* A, ttf_zdouble* Q, ttf_zdouble* R);

typedef struct {
  long nsolves;       /* local exponentials (cores and S-steps) */
  long matvecs;       /* Krylov operator applications */
  long krylov_steps;  /* ZGEXPV integration steps */
  double seconds;     /* wall time of the call */
  double flops;       /* algorithmic real flops of those operator applications */
  long matvecs32;     /* of matvecs, the core products in complex64
                         (ttf_set_local_precision(32)) */
} ttf_ksl_stats;

/* One KSL time step for dY/dt = A Y in TT format, argument layout of
 * ztt_ksl (tt_ksl.f90:431-445): d modes, n = m mode sizes, ra and crA the
 * complex operator (the reference's convention: A carries the factor i of
 * Schroedinger dynamics), crY0 with ranks ry (in/out, kept fixed), tau the
 * time step.  rmax, kickrank and nswp are accepted and unused, as in the
 * reference; typ 1 (KSL) or 2 (symmetrised, default in the reference);
 * order the Krylov size (reference default 8, at most 38); usenrm 2 uses a
 * unit norm estimate, anything else Hager's 1-norm estimate.  The result
 * cores (the reference's dyn_tt::zresult_core) go to `result` (capacity in
 * complex entries; size in *result_size).  Host pointers; device-resident
 * run. */
int ttf_ztt_ksl(int d, const int* n, const int* m, const int* ra, const ttf_zdouble* crA,
                const ttf_zdouble* crY0, int* ry, double tau, int rmax, int kickrank, int nswp, int verb,
                int typ, int order, int usenrm, ttf_zdouble* result, int64_t result_cap,
                int64_t* result_size, ttf_ksl_stats* stats);

/* ---------------------------------------------------------------- SDV files */
/* Tensor trains in the reference's SDV stream format (ttio.f90:10-17
 * header; 29-192 dtt_write / ztt_write; 196-399 dtt_read / ztt_read): a
 * 128-byte header, then l, m, n(l:m), r(l-1:m) as int32, then the cores
 * (r(k-1), n(k), r(k)) column-major and back to back, float64 or complex128.
 * Host memory and host code only: these calls never touch the device and
 * work without ttf_init.  The reference's readers allocate the cores; here
 * the caller asks ttf_sdv_dims for the sizes first.
 *
 * Status: 0, or the reference readers' info codes -1 (no such file), -2
 * (not a TT file), -3 (version other than 1), positive errno values for
 * I/O failures; and, without a counterpart there, -4 (the file ends before
 * its sizes say), -5 (invalid sizes, or a caller buffer too small), -6 (a
 * real file given to ttf_ztt_read or a complex one to ttf_dtt_read; the
 * reference reads it without a check). */
/* Sizes of the tensor train in `path`: l, m, the d = m - l + 1 mode sizes
 * n and the d + 1 ranks r (n and r may both be NULL; otherwise dcap >= d),
 * the element count sum r(k-1) n(k) r(k) of the cores and whether they are
 * complex (inf(2) = 1 in the header). */
int ttf_sdv_dims(const char* path, int dcap, int* l, int* m, int* n, int* r, int64_t* count, int* is_complex);
/* ttio.f90:196-297, dtt_read: the cores into `cores` (cap elements). */
int ttf_dtt_read(const char* path, int64_t cap, double* cores);
/* ttio.f90:298-399, ztt_read. */
int ttf_ztt_read(const char* path, int64_t cap, ttf_zdouble* cores);
/* ttio.f90:29-109, dtt_write: cores l..m with mode sizes n (m - l + 1
 * entries) and ranks r (m - l + 2 entries).  Replaces an existing file. */
int ttf_dtt_write(const char* path, int l, int m, const int* n, const int* r, const double* cores);
/* ttio.f90:110-192, ztt_write. */
int ttf_ztt_write(const char* path, int l, int m, const int* n, const int* r, const ttf_zdouble* cores);

/* ------------------------------------------ reference-compatible symbols */
/* tt_linalg.f90:127-159, dbfun3 bind(c).  Host pointers; res1/res2 ignored. */
void dbfun3(const ttf_fint* rx1, const ttf_fint* m, const ttf_fint* rx2,
            const ttf_fint* ry1, const ttf_fint* n, const ttf_fint* ry2,
            const ttf_fint* ra1, const ttf_fint* ra2, const double* phi1,
            const double* A, const double* phi2, const double* x, double* y,
            double* res1, double* res2);
/* tt_linalg.f90:30-51, dqr bind(c): Householder QR in place (LAPACK's
 * reflector convention, so Q and R match dgeqrf + dorgqr to rounding).
 * A (n, m) returns Q in its first min(n, m) columns (and, for m > n, R's
 * trailing columns in the rest, as dgeqrf leaves them); R is the
 * reference's R(min(n, m), n): zeroed, then R(1:min(j, n), j) for
 * j = 1..m written with leading dimension min(n, m).  work, lwork, tau
 * ignored.  n < 16384. */
void dqr(const ttf_fint* n, const ttf_fint* m, double* A, double* R, double* work,
         const ttf_fint* lwork, double* tau);
/* tt_linalg.f90:54-69, drow_add bind(c): the k rows of B (k, n) appended
 * under A (m, n), in place (A must hold (m + k) n). */
void drow_add(const ttf_fint* m, const ttf_fint* n, const ttf_fint* k, double* A, const double* B);
/* tt_linalg.f90:71-79, drow_cut bind(c): A (m, n) -> its first k rows,
 * packed (k, n) in place. */
void drow_cut(const ttf_fint* m, const ttf_fint* n, const ttf_fint* k, double* A);
/* tt_linalg.f90:82-98, dtransp bind(c): B (m, n) = A (n, m)^T. */
void dtransp(const ttf_fint* n, const ttf_fint* m, const double* A, double* B);
/* tt_linalg.f90:101-121, dperm1324 bind(c): B (n1, n3, n2, n4) =
 * A (n1, n2, n3, n4). */
void dperm1324(const ttf_fint* n1, const ttf_fint* n2, const ttf_fint* n3, const ttf_fint* n4,
               const double* A, double* B);
/* tt_linalg.f90:363-495, djac_gen bind(c): block-Jacobi blocks ('c', 'l'
 * or 'r') of the local operator with Phi1 (rx1, rx1, ra1), A (ra1, n, n,
 * ra2), Phi2 (rx2, ra2, rx2); jacs in the reference's layout
 * (ttf_djac_size doubles).  Blocks up to 128. */
void djac_gen(const char* ptype, const ttf_fint* rx1, const ttf_fint* n, const ttf_fint* rx2,
              const ttf_fint* ra1, const ttf_fint* ra2, const double* Phi1, const double* A,
              const double* Phi2, double* jacs);
/* tt_linalg.f90:497-533, djac_apply bind(c): y = P x; work1 ignored. */
void djac_apply(const char* ptype, const ttf_fint* rx1, const ttf_fint* n, const ttf_fint* rx2,
                const double* jacs, const double* x, double* y, double* work1);
/* tt_linalg.f90:540-643, duchol_fort bind(c): pivoted incomplete Cholesky
 * of the Gram matrix of the n items of dA; du (n, numvec), lam (numvec),
 * real_numvec.  numvec <= 32. */
void duchol_fort(const char* trans, const ttf_fint* m, const ttf_fint* n, const double* dA,
                 const ttf_fint* numvec, double* du, double* lam, ttf_fint* real_numvec);
/* tt_linalg.f90:651-845, dgmresr_hh_fort bind(c): right-preconditioned
 * restarted GMRES of the local system (Phi1 (rx1, rx1, ra1), A (ra1, n, n,
 * ra2), Phi2 (rx2, ra2, rx2)); sol (rx1 n rx2) is returned before the
 * preconditioner is applied, as the reference returns it.  Krylov
 * dimension at most 63; verb ignored. */
void dgmresr_hh_fort(const double* Phi1, const double* A, const double* Phi2, const double* rhs,
                     const ttf_fint* rx1, const ttf_fint* n, const ttf_fint* rx2, const ttf_fint* ra1,
                     const ttf_fint* ra2, const ttf_fint* nrestart, const double* tol,
                     const ttf_fint* niters, const char* ptype, const double* jacs, double* sol,
                     const ttf_fint* verb);

/* ------------------------------- callbacks of the local eigen/exp solvers */
/* The reference binds one local step's operator for its callbacks with
 * init_bfun_sizes + init_bfun_main (tt_eigb.f90:29-49; tt_ksl.f90:55-90):
 * here the interfaces and operator core are uploaded once per step, so
 * each callback moves only x and y.  x is (rx1, m, rx2), y (ry1, n, ry2);
 * xsize / ysize are the leading dimensions of blocks of vectors. */
void ttf_init_bfun_sizes(const ttf_fint* rx1, const ttf_fint* m, const ttf_fint* rx2, const ttf_fint* ry1,
                         const ttf_fint* n, const ttf_fint* ry2, const ttf_fint* ra1, const ttf_fint* ra2,
                         const ttf_fint* xsize, const ttf_fint* ysize);
void ttf_init_bfun_main(const double* phi1, const double* A, const double* phi2);
void ttf_zinit_bfun_main(const ttf_zdouble* phi1, const ttf_zdouble* A, const ttf_zdouble* phi2);
/* PRIMME's matrixMatvec(x, y, blockSize, primme) (primme/primme.h:140-141)
 * as tt_eigb.f90:15-27 binds it: blockSize columns of host x and y, one
 * block product on the device; the PRIMME parameter block is not read. */
void ttf_primme_matvec(void* x, void* y, ttf_fint* blockSize, void* primme);
/* Expokit's matvec(x, y) (expm/expokit.f:2453-2457) and normest's
 * matvec_transp(x, y) (expm/normest.f90:146-151) as tt_ksl.f90:25-50 binds
 * them: y = B x and y = B^H x on host complex(8) arrays. */
void ttf_zmatvec(const ttf_zdouble* x, ttf_zdouble* y);
void ttf_zmatvec_transp(const ttf_zdouble* x, ttf_zdouble* y);

#ifdef __cplusplus
}
#endif

#endif /* TTFORT_AMD_H */
