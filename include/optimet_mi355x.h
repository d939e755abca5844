/*
 * optimet_mi355x.h -- C ABI of the MI355X multiple-scattering backend.
 *
 * The library sits behind OPTIMET's solver interface
 * optimet::solver::AbstractSolver (srcAna/Solver.h:62-144):
 *   omx_set_geometry / omx_set_excitation  <- the Geometry and Excitation
 *       objects handed to the solver's constructor and to
 *       AbstractSolver::update(geometry, incWave) (Solver.h:70-74, 94-99);
 *   omx_update                              <- AbstractSolver::update()
 *       (Solver.h:103; PreconditionedMatrixSolver.h:82-100);
 *   omx_solve                               <- AbstractSolver::solve(X_sca, X_int,
 *       X_sca_SH, X_int_SH, CGcoeff) (Solver.h:86-87; PreconditionedMatrixSolver.h:45-79);
 *   omx_set_cg_tables                       <- the CGcoeff argument of solve()
 *       (the nine tables of Geometry::Coefficients, Geometry.cpp:234-247);
 *   omx_cross_sections                      <- Result::getExtinctionCrossSection,
 *       getScatteringCrossSection, getAbsorptionCrossSection and
 *       getScatteringCrossSection_SH (Result.cpp:557-761);
 *   omx_absorption_cross_section_sh         <- Result::getAbsorptionCrossSection_SH
 *       (Result.cpp:763-794);
 *   omx_fields                              <- Result::getEHFields / setFields
 *       (Result.cpp:74-301, 797-933);
 *   omx_scattering_size                     <- AbstractSolver::scattering_size()
 *       (Solver.h:132-135).
 *
 * Conventions (the reference's, srcAna/Types.in.h:46-52, Reader.cpp:593-831):
 * complex numbers are interleaved (re, im) doubles; coefficient vectors hold,
 * per object, nf = nmax(nmax+2) TE ("b_n") coefficients followed by nf TM
 * ("a_n") ones, harmonic index p = n(n+1) - m - 1; lengths in metres, angles
 * in radians; permittivities and permeabilities relative to vacuum.
 *
 * Errors: every function returns OMX_OK (0) or a negative OMX_ERR_* code and
 * never throws; omx_last_error() returns the calling thread's last message.
 * A context is used by one host thread at a time; distinct contexts are
 * independent.
 */
#ifndef OPTIMET_MI355X_H
#define OPTIMET_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMX_OK 0
#define OMX_ERR_ARG -1       /* invalid argument or unsupported size */
#define OMX_ERR_DEVICE -2    /* HIP runtime error (no device, launch failure) */
#define OMX_ERR_MEMORY -3    /* device allocation failed */
#define OMX_ERR_SINGULAR -4  /* zero pivot in the LU factorisation */
#define OMX_ERR_STATE -5     /* call out of order (e.g. solve before update) */
#define OMX_ERR_NOCONV -6    /* iterative solver did not reach its tolerance */
#define OMX_ERR_OVERLAP -7   /* overlapping spheres (Geometry.cpp:39-51) */

/* Solver kinds: the backend selected from the XML input. */
#define OMX_SOLVER_LU64 0     /* complex FP64 LU (+ FP64 refinement after block pivoting) */
#define OMX_SOLVER_LU32_IR 1  /* complex FP32 LU on MFMA + FP64 GMRES-IR */
#define OMX_SOLVER_GMRES64 2  /* restarted GMRES, FP64 matrix streamed from HBM */
#define OMX_SOLVER_ACA_GMRES 3 /* <ACA compression="yes"/>: admissible blocks compressed by
                                  ACA at 1e-3, restarted GMRES on the compressed operator
                                  (PreconditionedMatrixSolver.h:47-56, 84-96; the
                                  reference's settings are tol 1e-6, 240 steps per cycle,
                                  2 cycles) */
#define OMX_SOLVER_GMRES32_IR 4 /* mixed Krylov: FP64 flexible GMRES on S whose preconditioner
                                   is one cycle of GMRES on the complex-FP32 copy of S (FP32
                                   matrix, FP64 vectors and sums) */
#define OMX_SOLVER_AUTO 5      /* by size: N >= 16384 starts with OMX_SOLVER_GMRES32_IR within a
                                  budget of the mixed LU's cost and falls back to
                                  OMX_SOLVER_LU32_IR on the same matrices; smaller systems use
                                  OMX_SOLVER_LU32_IR, except that with the second harmonic S
                                  is first tried with FP64 GMRES (within N / 125 products)
                                  while V's mixed LU factorises (from half of V's columns
                                  on); batched sweeps from N = 2048 with the second harmonic
                                  try S by lockstep FP64 GMRES (within 2 N / 125 steps)
                                  beside V's batched mixed LU */

/* Pivot search of the LU factorisations (omx_set_pivoting). */
#define OMX_PIVOT_BLOCK 0     /* among the rows of each 64 x 64 diagonal block, then
                                 FP64 GMRES-IR; falls back to OMX_PIVOT_PARTIAL on a
                                 zero pivot or a refinement that does not converge */
#define OMX_PIVOT_PARTIAL 1   /* partial pivoting over the whole column (tournament) */

/* Products of the complex-FP32 factorisation's bulk updates (omx_set_lu32_gemm). */
#define OMX_GEMM_BF16X3 0     /* each FP32 operand split in three bf16 terms, six
                                 bf16 MFMA products per real product, FP32
                                 accumulation: FP32 accuracy at the bf16 rate */
#define OMX_GEMM_FP32 1       /* the FP32 matrix cores */

typedef struct omx_context omx_context;

typedef struct {
  int N;                 /* unknowns per frequency, 2 P nf */
  int Np;                /* padded matrix order */
  int iterations;        /* GMRES / refinement inner iterations (fundamental) */
  int iterations_sh;     /* same, second harmonic */
  double residual;       /* final ||Q - S x|| / ||Q|| (fundamental) */
  double residual_sh;    /* final ||K - V x|| / ||K|| (second harmonic) */
  double t_update_ms;    /* last omx_update: assembly of Q, S (and V) */
  double t_factor_ms;    /* last omx_solve: factorisations */
  double t_solve_ms;     /* last omx_solve: triangular solves / iterations */
  double t_sh_source_ms; /* last omx_solve: second-harmonic source vector */
  double t_assemble_ms;  /* last omx_update: pair-coefficient kernels only */
  size_t device_bytes;   /* device memory held by the context */
  /* with omx_set_profile(ctx, 1), accumulated over the factorisations since:
   * HIP-event times on the streams the kernels run on */
  double t_lu_gemm_ms;   /* bulk trailing updates (K = 128) */
  double lu_gemm_flop;   /* their algorithmic flop, 8 m n k per launch */
  int lu_gemm_launches;  /* their kernel launches */
  double t_lu_panel_ms;  /* panel work of each pair up to its far-column step */
  double t_factor_total_ms; /* whole factorisations */
  int pivot_fallbacks;   /* refactorisations with partial pivoting since creation */
  int sweep_batch;       /* systems per batch of the last omx_solve_sweep */
  int matvec_calls;      /* FP64 matrix-vector products of the iterative solvers (since omx_set_profile) */
  double t_matvec_ms;    /* their HIP-event time (profiling on) */
  double matvec_bytes;   /* their algorithmic bytes, 16 Np^2 per system and product */
  int graph_captures;    /* factorisation graphs captured (omx_set_graphs) since creation */
  int graph_launches;    /* factorisations run as graph launches since creation */
} omx_stats;

const char *omx_last_error(void);

/* Context on HIP device `device`. */
int omx_create(omx_context **ctx, int device);
void omx_destroy(omx_context *ctx);

/* Geometry: nobj spheres with Cartesian centres xyz[3*nobj] (Reader.cpp:593-599
 * converts these to Scatterer::vR), radius[nobj], and per sphere the complex
 * eps_r, mu_r, eps_r_sh (relative permittivity at 2 omega), ksippp (chi_perp-perp-perp),
 * ksiparppar (chi_par-perp-par) and gamma (ElectroMagnetic::init_r,
 * ElectroMagnetic.cpp:34-54); bg_eps_r, bg_mu_r describe the background.
 * All spheres share nmax (PreconditionedMatrix.cpp:1162-1165); 1 <= nmax <= 10.
 * eps_r_sh, ksippp, ksiparppar, gamma may be NULL when no second harmonic is solved. */
int omx_set_geometry(omx_context *ctx, int nobj, int nmax, const double *xyz, const double *radius,
                     const double *eps_r, const double *mu_r, const double *eps_r_sh,
                     const double *ksippp, const double *ksiparppar, const double *gamma,
                     const double *bg_eps_r, const double *bg_mu_r);

/* Plane-wave excitation (Reader.cpp:782-831, Excitation.cpp:31-75): vacuum
 * wavelength (m), propagation direction theta, phi (rad), polarisation
 * E_theta, E_phi (complex), and the second-harmonic flag SH_cond. */
int omx_set_excitation(omx_context *ctx, double wavelength, double theta, double phi,
                       const double *etheta, const double *ephi, int sh);

/* Solver selection; tol is the relative residual target of the iterative
 * parts (GMRES-IR refinement, GMRES), maxit their iteration cap, restart the
 * GMRES restart length (<= 256). */
int omx_set_solver(omx_context *ctx, int kind, double tol, int maxit, int restart);

/* The Krylov budget of OMX_SOLVER_AUTO: outer steps of the mixed Krylov (each
 * at most 31 products) from N = 16384 on, FP64 GMRES products for S below it;
 * 0 (default) takes the cost model, what the mixed LU would cost (0.0143 N
 * FP32 products, N / 125 FP64 products).  No reference counterpart: a tuning
 * knob of the automatic choice. */
int omx_set_auto_budget(omx_context *ctx, int steps);

/* Autotuning of the batched sweeps (default on; no reference counterpart): per
 * shape (device, padded order Np, batch nb, FP64 or FP32 factors) the sweep
 * times each candidate bulk-update product of FP32 factors (split-bf16, FP32
 * MFMA) and each triangular-solve variant on the batch's own buffers before
 * its first chunk, and keeps the fastest (a candidate replaces the default
 * only when at least 5% faster).  Off: the products of omx_set_lu32_gemm and
 * the default triangular solves. */
int omx_set_autotune(omx_context *ctx, int on);

/* The tuned choice of a shape: *split (1 split-bf16, 0 FP32 MFMA), *trsv (the
 * triangular-solve variant, 0 = default), times[5] (ms: the two products'
 * probes, FP32 factors only, then the three triangular-solve variants).
 * OMX_ERR_STATE when the shape has not been tuned in this process. */
int omx_tune_query(int device, int Np, int nb, int fp64, int *split, int *trsv, double *times);

/* Forget every tuned choice of this process. */
int omx_tune_reset(void);

/* No reference counterpart: continuous batching across requests.  Solves the
 * n member contexts, each with its own geometry and excitation and already
 * assembled by omx_update (clusters of any P and nmax; one solver kind for all,
 * OMX_SOLVER_LU64 or OMX_SOLVER_LU32_IR, and one SH flag), as one batch:
 * their S (and V) are bordered by the identity to the largest padded order,
 * factored by one batched launch sequence and refined in lockstep in the
 * leader's sweep buffers; each member's x_int, SH source and x_int_SH come from
 * its own kernels.  Afterwards every member holds its solution as after
 * omx_solve (omx_read_solution, omx_cross_sections, omx_fields).  A member
 * whose block pivots or refinement fall short is solved alone by omx_solve.
 * The leader needs no geometry, and its own settings are left as they were.
 * All contexts must be on the leader's device and distinct from it. */
int omx_solve_group(omx_context *leader, int n, omx_context *const *members);

/* The solution of the last omx_solve or omx_solve_group (each pointer may be
 * NULL; N complex values each, interleaved). */
int omx_read_solution(omx_context *ctx, double *x_sca, double *x_int, double *x_sca_sh, double *x_int_sh);

/* No reference counterpart.  On: the block-pivoting factorisations of
 * omx_solve and of omx_solve_sweep's chunks are captured in HIP graphs once
 * per state (order, batch, solver kind, product, buffers) and replayed;
 * S's graph is cut at V's start mark so that V's factorisation still overlaps
 * S's.  The kernels and their order are those of the eager path, so results
 * are identical.  Off (default): eager launches, and the graphs are released. */
int omx_set_graphs(omx_context *ctx, int on);

/* Pivot search of the LU solvers (default OMX_PIVOT_BLOCK).  The reference
 * solves with Eigen's ColPivHouseholderQR (PreconditionedMatrixSolver.h:58,
 * 75) or ScaLAPACK pzgesv (scalapack/LinearSystemSolver.hpp:84-91); both
 * modes reach the same FP64 residual.  OMX_PIVOT_PARTIAL with
 * OMX_SOLVER_LU64 is the plain partial-pivoting FP64 LU solve. */
int omx_set_pivoting(omx_context *ctx, int mode);

/* Products of OMX_SOLVER_LU32_IR's bulk updates (default OMX_GEMM_BF16X3).
 * No reference counterpart: the reference factorises in FP64 only; both
 * modes give an FP32-accurate factorisation, refined to FP64 by GMRES-IR. */
int omx_set_lu32_gemm(omx_context *ctx, int mode);

/* The nine Clebsch-Gordan / Wigner tables of the second-harmonic source
 * (CGcoeff, Simulation.cpp:613-616), each nf*nf*nf doubles in the order
 * C_10m1, C_11m1, C_00m1, C_01m1, W_m1m1, W_11, W_00, W_10, W_01, indexed
 * [k][p][q].  Copied to the device; the pointers are not retained. */
int omx_set_cg_tables(omx_context *ctx, int nmax, const double *const *tables);
/* Builds the same nine tables on the device (Symbol.cpp:1036-1451). */
int omx_build_cg_tables(omx_context *ctx, int nmax);

/* AbstractSolver::update(): plane-wave coefficients, T-matrices, source
 * vector Q, scattering matrix S and, with SH_cond, V on the device. */
int omx_update(omx_context *ctx);

/* AbstractSolver::solve(): writes N complex values into each non-NULL output
 * (x_sh and x_int_sh only when SH_cond). */
int omx_solve(omx_context *ctx, double *x_sca, double *x_int, double *x_sca_sh, double *x_int_sh);

/* [C_ext, C_sca, C_abs, C_sca_SH] of the last solve, with the reference's
 * conventions (the incoherent C_sca sum of Result.cpp:629-631 included). */
int omx_cross_sections(omx_context *ctx, double *out4);

/* C_abs_SH of the last solve with SH_cond (Result::getAbsorptionCrossSection_SH,
 * Result.cpp:763-794, with AbsCSSHcoeff, Geometry.cpp:428-455): the radial
 * Gauss-Legendre integral of |E(2 omega)|^2 inside each sphere times the SH
 * conductivity -i eps0 2 omega (eps_r_SH - 1). */
int omx_absorption_cross_section_sh(omx_context *ctx, double *out);

/* Field maps of the last solve (Result::getEHFields with the projection off
 * and Result::setFields, Result.cpp:74-301, 797-933): E and H at omega, and
 * with E_sh and H_sh non-NULL (SH_cond) at 2 omega, at npts points xyz[3*npts]
 * (Cartesian, metres), each output 3*npts complex (x, y, z components).
 * Outside the spheres: incident plus scattered waves; inside a sphere: its
 * internal waves and, at 2 omega, the bulk-source term (Geometry::COEFFpartSH). */
int omx_fields(omx_context *ctx, int npts, const double *xyz, double *E, double *H, double *E_sh, double *H_sh);

/* 2 nmax (nmax+2) P, AbstractSolver::scattering_size(). */
int omx_scattering_size(const omx_context *ctx);

int omx_get_stats(const omx_context *ctx, omx_stats *out);

/* Per-kernel timing of the factorisations with HIP events (off by default:
 * it adds event records between launches). */
int omx_set_profile(omx_context *ctx, int on);

/* Parity access: copies the assembled N x N matrix (which 0: S, 1: V; FP64,
 * column-major) or a vector (which 0: Q, 1: K, 2: K1ana, 3: local incident
 * field of every object) to host memory. */
int omx_read_matrix(omx_context *ctx, int which, double *out);
int omx_read_vector(omx_context *ctx, int which, double *out);
/* Parity access to one block of the assembled matrix (which 0: S, 1: V):
 * rows row0 .. row0+nrows-1 and columns col0 .. col0+ncols-1, copied to out
 * as an nrows x ncols column-major complex128 array (no full-matrix copy). */
int omx_read_matrix_block(omx_context *ctx, int which, int row0, int col0, int nrows, int ncols, double *out);

/* The compressed operator of OMX_SOLVER_ACA_GMRES after omx_update, S (which
 * 0) or V (which 1), as Scattering_matrix_ACA_FF / _SH build it
 * (PreconditionedMatrix.cpp:489-551, 696-759).  ranks (P*P, may be NULL):
 * r of a compressed block (i, j) at ranks[i P + j], -1 for a dense block, 0
 * on the diagonal.  info (7 doubles, may be NULL): device bytes of the
 * operator's blocks, bytes of the dense FP64 matrix (16 N^2), the size the
 * reference reports in MB (16 bytes per stored entry, identity blocks
 * included, :547-549), compressed blocks, dense off-diagonal blocks, largest
 * and mean rank. */
int omx_aca_info(omx_context *ctx, int which, int *ranks, double *info);

/* Block (i, j) of the compressed operator: *rank = r with U (n x r,
 * column-major) and V (r x n, row-major) written to U, V (n*n complex
 * capacity each, n = 2 nmax (nmax + 2)); a dense block gives *rank = -1 and
 * the block in U (n x n column-major); the diagonal gives *rank = 0. */
int omx_aca_block(omx_context *ctx, int which, int i, int j, int *rank, double *U, double *V);

/* Wavelength sweep, batched: the current geometry and excitation solved at
 * nl vacuum wavelengths (m), as Simulation::scan_wavelengths does one
 * wavelength at a time (Simulation.cpp:604-685, update() and solve() per
 * wavelength).  eps_r and eps_r_sh (nl*P complex, wavelength-major) give each
 * sphere's relative permittivity at each wavelength (at 2 omega for eps_r_sh);
 * NULL keeps the geometry's values.  Up to `batch` systems (0: automatic, at
 * most 64 and half the free device memory) are assembled, factorised (block
 * pivoting) and refined together; a system with a zero pivot, or still short
 * of the tolerance after four restart cycles of the lockstep refinement, is
 * redone alone with partial pivoting.  Outputs (any
 * may be NULL): x_sca, x_int, x_sca_sh, x_int_sh nl*N complex each, xs nl*5
 * cross sections per wavelength, [C_ext, C_sca, C_abs, C_sca_SH, C_abs_SH]
 * (omx_cross_sections and omx_absorption_cross_section_sh; the scan writes
 * these, Simulation.cpp:657-671), iterations nl*2 (omega, 2 omega). */
int omx_solve_sweep(omx_context *ctx, int nl, const double *lambdas, const double *eps_r, const double *eps_r_sh,
                    int batch, double *x_sca, double *x_int, double *x_sca_sh, double *x_int_sh, double *xs,
                    int *iterations);

/* Device memory: free and total bytes of the context's device, bytes the
 * context holds (the footprint a scheduler admits requests against). */
int omx_mem_info(omx_context *ctx, size_t *free_bytes, size_t *total_bytes, size_t *held_bytes);
/* Releases the context's matrix-sized buffers (matrices, factors, Krylov
 * bases); the next update or sweep allocates them again. */
int omx_trim(omx_context *ctx);

/* Stand-alone dense complex solve A x = b on the device (A column-major,
 * N x N); kind OMX_SOLVER_LU64 or OMX_SOLVER_LU32_IR.  Uses the context's
 * device and stream. */
int omx_dense_solve(omx_context *ctx, int N, const double *A, const double *b, double *x, int kind);

/* Allocation-free variant of omx_dense_solve with OMX_SOLVER_LU64 and
 * OMX_PIVOT_PARTIAL (the plain FP64 LU with partial pivoting).  It has no
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Inputs and outputs are DEVICE pointers: A is an lda x Np
 * column-major complex128 array with Np = N rounded up to 64 and lda >= Np,
 * whose leading N x N block holds the matrix; the call overwrites A with the
 * identity-padded matrix's LU factors.  b and x hold N complex128 values.
 * Every work array (pivots, permutations, status words, diagonal-block inverses,
 * staged right-hand side) lives in one library-held scratch array per device
 * (about 8.6 MiB, held until the process ends), so N is limited to
 * 1 <= N <= 4096 (OMX_ERR_ARG beyond).  Calls that overlap on one device share
 * that array and race: issue them from one host thread at a time.  The call
 * runs on the context's stream and returns after x is written. */
int omx_dense_solve_noalloc(omx_context *ctx, int N, double *A, size_t lda, const double *b, double *x);

#ifdef __cplusplus
}
#endif
#endif
