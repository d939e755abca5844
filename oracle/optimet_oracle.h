/*
 * optimet_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the OPTIMET srcAna hot path (analytic spheres):
 * spherical Bessel/Hankel functions, spherical harmonics, the Stout (2002)
 * translation-addition recurrence, the A/B coupling blocks, Mie T-matrices,
 * the plane-wave source, the preconditioned scattering matrix S = I - T.A
 * (and its second-harmonic twin V), the dense solves, the second-harmonic
 * source vector with its Clebsch-Gordan / Wigner tables, and the cross
 * sections.  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark may load this library; the product path never does.
 *
 * Complex numbers cross this interface as interleaved (re, im) doubles.
 * Every function returns 0 on success and a negative code on error; the text
 * of the last error is available from ora_last_error().
 */
#ifndef OPTIMET_ORACLE_H
#define OPTIMET_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One cluster-solve problem, in the units the reference's reader produces
 * (Reader.cpp:593-605, 782-831): lengths in metres, angles in radians,
 * permittivities/permeabilities relative to vacuum. */
typedef struct {
  int nobj;
  int nmax;
  int sh;                 /* second-harmonic stage requested */
  const double *xyz;      /* 3*nobj Cartesian centres */
  const double *radius;   /* nobj */
  const double *eps_r;    /* 2*nobj complex, at omega */
  const double *mu_r;     /* 2*nobj complex */
  const double *eps_r_sh; /* 2*nobj complex, at 2*omega */
  const double *ksippp;   /* 2*nobj complex, chi_perp-perp-perp */
  const double *ksiparppar; /* 2*nobj complex, chi_par-perp-par */
  const double *gamma;    /* 2*nobj complex, bulk coefficient */
  double bg_eps_r[2];
  double bg_mu_r[2];
  double lambda;          /* vacuum wavelength, m */
  double theta_k, phi_k;  /* propagation direction */
  double etheta[2], ephi[2]; /* polarisation */
} ora_problem;

const char *ora_last_error(void);

/* type 0: spherical Bessel j_n; type 1: spherical Hankel h_n^(1).
 * data, ddata: (nmax+1) complex each. */
int ora_bessel(int type, double zr, double zi, int nmax, double *data, double *ddata);

/* Y_n^m(theta, phi) with the Condon-Shortley phase. */
int ora_ynm(int n, int m, double theta, double phi, double *out);

/* Translation-addition coefficient C(n,m,l,k) of a fresh
 * TranslationAdditionCoefficients(R, waveK, regular) instance. */
int ora_tac(const double *R_sph, double kr, double ki, int regular, int n, int m, int l, int k,
            double *out);

/* Coupling(R, waveK, nmax, regular): A ("diagonal") and B ("offdiagonal"),
 * each nf*nf complex, row-major A[p*nf+q] with nf = nmax(nmax+2). */
int ora_coupling(const double *R_sph, double kr, double ki, int nmax, int regular, double *A,
                 double *B);

/* Per-object diagonal vectors, 2*nf complex each (TE block first).
 * which: 0 T (getTLocal), 1 T_SH (getTLocalSH), 2 Iaux, 3 IauxSH1, 4 IauxSH2,
 *        5 TSH1_outer, 6 TSH2_outer */
int ora_object_vector(const ora_problem *pb, int obj, int which, double *out);

/* Plane-wave coefficients dataIncAp, dataIncBp: nf complex each. */
int ora_incident(const ora_problem *pb, double *ap, double *bp);

/* Excitation::getIncLocal at object obj: 2*nf complex. */
int ora_inc_local(const ora_problem *pb, int obj, double *out);

/* Q = source_vector(): N complex. */
int ora_source_vector(const ora_problem *pb, double *Q);

/* S (sh=0) or V (sh=1): N*N complex, column-major. */
int ora_scattering_matrix(const ora_problem *pb, int sh, double *S);

/* Blocks (i, j) of S (sh=0) or V (sh=1) for the nblk object pairs ij[2b], ij[2b+1]:
 * out holds nblk blocks of 2nf x 2nf complex, each column-major.  Pairs run on all
 * OpenMP threads. */
int ora_scattering_blocks(const ora_problem *pb, int sh, int nblk, const int *ij, double *out);

/* r = b - S x (sh=0) or b - V x (sh=1), N complex each, with the matrix assembled
 * one block row at a time (never held whole); block rows run on all OpenMP threads. */
int ora_residual(const ora_problem *pb, int sh, const double *x, const double *b, double *r);

/* Dense solve of S x = b (column-major, N*N). method 0: LU with partial
 * pivoting; 1: column-pivoted Householder QR. */
int ora_dense_solve(int N, const double *S, const double *b, double *x, int method);

/* The nine coupling tables of Geometry::Coefficients, each nfs*nf*nf doubles,
 * order C_10m1, C_11m1, C_00m1, C_01m1, W_m1m1, W_11, W_00, W_10, W_01. */
int ora_cg_tables(int nmax, int nmaxs, double *tables);

/* Wigner symbols (for pinning against exact rationals). */
double ora_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);
double ora_wigner6j(int j1, int j2, int j3, int j4, int j5, int j6);
double ora_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);

/* source_vectorSH and source_vectorSH_K1ana from conj(x_int): N complex each. */
int ora_source_sh(const ora_problem *pb, const double *tables, const double *x_int, double *K,
                  double *K1ana);

/* Full PreconditionedMatrix update()+solve(): x_sca, x_int (N complex each) and,
 * if pb->sh, x_sca_sh, x_int_sh.  xs: [Cext, Csca, Cabs, Csca_SH, Cabs_SH].
 * method 0: LU; 1: column-pivoted QR; 2: ACA-compressed operators + Gmres_Zcomp
 * (tol 1e-6, 240 steps, 2 cycles). */
int ora_solve_all(const ora_problem *pb, int method, double *x_sca, double *x_int,
                  double *x_sca_sh, double *x_int_sh, double *xs);

/* Cross sections from solved coefficients: xs = [Cext, Csca, Cabs, Csca_SH]. */
int ora_cross_sections(const ora_problem *pb, const double *x_sca, const double *x_sca_sh,
                       double *xs);

/* Result::getAbsorptionCrossSection_SH from the internal coefficients at omega and 2 omega
 * (N complex each) and the nine tables of ora_cg_tables. */
int ora_absorption_sh(const ora_problem *pb, const double *tables, const double *x_int,
                      const double *x_int_sh, double *out);

/* Result::getEHFields at npts Cartesian points xyz[3*npts] (projection off):
 * E, H at omega and Esh, Hsh at 2 omega, 3*npts complex each (x, y, z).
 * tables: the nine tables of ora_cg_tables (the SH bulk term inside spheres). */
int ora_fields(const ora_problem *pb, const double *tables, const double *x_sca, const double *x_int,
               const double *x_sca_sh, const double *x_int_sh, int npts, const double *xyz, double *E, double *H,
               double *Esh, double *Hsh);

/* ACA_compression of the n x n block A (column-major) at tolerance eps:
 * A ~ U V with U n x r (column-major, n x n capacity) and V r x n (row-major,
 * n x n capacity); *rank = r.  trace (3 n doubles, may be NULL): per step the
 * runner-up / largest candidate modulus of the pivot column's choice and of
 * the next row's choice, and the stopping ratio eps sqrt(estimate) / (|u| |v|). */
int ora_aca_block(int n, const double *A, double eps, int *rank, double *U, double *V, double *trace);

/* Scattering_matrix_ACA_FF (sh=0) / _SH (sh=1) expanded to the N x N
 * compressed operator S (column-major): admissible blocks (i != j,
 * |c_i - c_j| >= 2 (r_i + r_j)) replaced by U V.  rank: P*P entries, r of a
 * compressed block, -1 for a dense block, 0 on the diagonal (may be null). */
int ora_aca_operator(const ora_problem *pb, int sh, double eps, double *S, int *rank);

/* Gmres_Zcomp on the N x N operator A (column-major) from x = 0: tol on the
 * rotated residual estimate, maxit steps per cycle, `restarts` cycles.
 * info[0] = steps taken, info[1] = final relative residual estimate. */
int ora_gmres_zcomp(int N, const double *A, const double *b, double tol, int maxit, int restarts, double *x,
                    double *info);

/* CPU-baseline timing sample: out[0] s per pair coupling + block, out[1] s
 * of one dense solve of order nsolve (method as ora_dense_solve), out[2] s of
 * one object's second-harmonic source. */
int ora_time_sample(const ora_problem *pb, int npairs, int nsolve, int method, double *out);

#ifdef __cplusplus
}
#endif
#endif
