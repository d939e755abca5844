// omx_kernels.hpp -- launch interfaces shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "omx_math.hpp"

namespace omx {

// per-object vector kinds, layout ov[kind][obj][pol][n], pol 0 = TE, 1 = TM
enum {
  OV_T = 0,      // getTLocal
  OV_TSH,        // getTLocalSH
  OV_IAUX,       // getIaux
  OV_IAUXSH1,    // getIauxSH1
  OV_IAUXSH2,    // getIauxSH2
  OV_TSH1O,      // getTLocalSH1_outer
  OV_TSH2O,      // getTLocalSH2_outer
  OV_CABS,       // getCabsAux (real, in .x)
  OV_SHRAD,      // [0][n] = j_n(k_j R), [1][n] = k_j j_n'(k_j R) + j_n / R
  OV_COUNT
};

constexpr int kCoefN = 2 * kMaxNmax + 2;
constexpr int kCoefStride = (kCoefN + 1) * (kCoefN + 1);

struct ObjArgs {
  int P, nmax;
  double omega;
  cplx bg_eps, bg_mu;                // absolute
  const double *xyz, *radius;
  const cplx *eps, *mu, *eps_sh;     // absolute
  double *sph;                       // out: 3P spherical coordinates
  cplx *ov;                          // out: per-object vectors
  // batches (blockIdx.z = system z): omega, eps, eps_sh and ov of system z at
  // omz[z], eps + z P, eps_sh + z P, ov + z zov (omz null: one system)
  const double *omz = nullptr;
  size_t zov = 0;
};

struct IncArgs {
  int nmax;
  double theta, phi;
  cplx etheta, ephi;
  cplx *inc;                         // out: [a_inc (nf); b_inc (nf)]
};

struct PairArgs {
  int P, nmax, tkind;                // tkind: OV_T for S, OV_TSH for V
  cplx k;                            // background wavenumber (k or 2k)
  const double *sph;
  const cplx *ov;
  const double *coef;
  size_t lda;
  cplx *S64;                         // optional FP64 output
  float2 *S32;                       // optional complex-FP32 output
  // batches (blockIdx.z = system z): wavenumber kz[z], ov + z zov, matrices + z zS
  const cplx *kz = nullptr;
  size_t zov = 0, zS = 0;
  // pair list (plist non-null): nlist blocks, block b = pair plist[b] (i P + j)
  // written to S64 / S32 + b zblk with leading dimension lda
  const int *plist = nullptr;
  int nlist = 0;
  size_t zblk = 0;
};

struct OriginArgs {
  int P, nmax;
  cplx k;
  const double *sph;
  const cplx *ov;
  const double *coef;
  const cplx *inc;
  cplx *incloc;                      // out: N
  cplx *Q;                           // out: N
  double *w;                         // out: [2][P][nf]
  // batches (blockIdx.z = system z): wavenumber kz[z], ov + z zov,
  // incloc/Q + z zvec, w + z zw
  const cplx *kz = nullptr;
  size_t zov = 0, zvec = 0, zw = 0;
};

struct CrossArgs {
  int P, nmax;
  const cplx *x, *xsh, *incloc;
  const double *w;
  const cplx *ov;
  double *out;                       // [P][4]
};

// Result::getEHFields / setFields (Result.cpp:74-301, 797-933): E and H at
// omega and 2 omega at npts Cartesian points, Cartesian components.
struct FieldArgs {
  int P, nmax, sh, npts;
  double omega;
  cplx waveK, bg_eps, bg_mu;               // background wavenumber; absolute eps, mu
  const double *xyz, *radius;              // object centres (Cartesian) and radii
  const cplx *eps, *mu, *eps_sh, *gamma;   // absolute
  const double *tab;                       // nine tables (SH bulk term inside spheres)
  const cplx *inc;                         // [a_inc; b_inc]
  const cplx *x, *xint, *xsh, *xintsh;     // solution vectors (N each)
  const double *pts;                       // 3 npts
  cplx *E, *H, *Esh, *Hsh;                 // 3 npts each; Esh, Hsh unused without sh
};
void run_fields(const FieldArgs &a, hipStream_t st);

// nb > 1: a batch of nb systems in one launch (the batch fields of the args)
void run_objects(const ObjArgs &a, hipStream_t st, int nb = 1);
void run_incident(const IncArgs &a, hipStream_t st);
void run_pairs(const PairArgs &a, hipStream_t st, int nb = 1);
void run_origin(const OriginArgs &a, hipStream_t st, int nb = 1);
void run_cross(const CrossArgs &a, hipStream_t st);
// identity padding of nb matrices zs elements apart
void run_pad64(cplx *A, int N, int Np, size_t lda, hipStream_t st, int nb = 1, size_t zs = 0);
void run_pad32(float2 *A, int N, int Np, size_t lda, hipStream_t st, int nb = 1, size_t zs = 0);
void fill_coef(double *host);

// ---------------------------------------------------------------- ACA
struct AcaArgs {
  int n, nblk;       // block order (<= 256) and blocks in the launch
  double eps;        // compression tolerance
  const cplx *A;     // nblk blocks, n x n column-major, n^2 apart
  cplx *U, *V;       // nblk slots of n^2: U column p at U + p n, V row p at V + p n
  int *rank;         // out: nblk
};
void run_aca(const AcaArgs &a, hipStream_t st);
// block b's factors, U (n x r) then V (r x n), copied to pool + off[b]
void run_aca_pack(const cplx *U, const cplx *V, int n, int nblk, const int *rank, cplx *pool, const size_t *off,
                  hipStream_t st);

// One off-diagonal block of a compressed operator: rank -1 dense (n x n
// column-major at M), r > 0 compressed (U n x r at M, V r x n at M + n r,
// V x_j kept at t + toff).
struct AcaDesc {
  const cplx *M;
  int j, rank, toff;
};
struct AcaOp {
  int P, n, N, Np, G;          // G groups of a block row's blocks (0 when P = 1)
  const AcaDesc *desc;         // sorted by block row, rows at rowptr[i] .. rowptr[i+1]
  const int *rowptr, *acalist; // acalist: desc indices of the naca compressed blocks
  int naca;
  cplx *t, *part;              // V x products; partial rows [G][P][n]
};
// y = S x for the compressed operator S (identity diagonal blocks, identity
// on the padding rows N .. Np-1)
void aca_matvec(const AcaOp &op, const cplx *x, cplx *y, hipStream_t st);

// ---------------------------------------------------------------- LU
constexpr int kNB = 64;        // LU block size (panel width, padding granule)
constexpr int kPanelRows = 128; // rows of the panel held by one workgroup

// Distances between the systems of a batch (element counts of each buffer);
// all zero for a single system.  A batch runs with blockIdx.z = system.
struct BatchStrides {
  size_t a = 0;    // matrices (elements of T)
  size_t inv = 0;  // per-panel inverse sets (elements of T)
  size_t vec = 0;  // right-hand sides and solutions (complex FP64 elements)
  int perm = 0;    // permutations
  int map = 0;     // gather maps (mdst, msrc)
  int cnt = 0;     // map counts
  int status = 0;  // status words
};

struct LuWork {
  // device scratch of one factorization
  int *piv;             // Np: LAPACK-style swap sequence
  int *perm;            // Np: final row permutation (row i of P A is row perm[i] of A)
  int *cand[2];         // tournament candidates, ping-pong, [Gmax][NB]
  int *cnt[2];          // candidates per set, [Gmax]
  int *mdst, *msrc;     // gather maps: tournament 2 NB entries at the start; block
  int *mcnt;            // pivoting NB entries per panel (panel k at k NB, count mcnt[k])
  int *status;          // [0]: singular pivot column+1, [1]: wait timeout
  void *wlu;            // NB x NB complex: the tournament's selected-block factors
  void *dinv;           // per panel [L11^{-1} | U11^{-1}], 2 NB^2 elements of T, column-major
  hipStream_t aux;      // look-ahead stream of the panel pivoting
  hipEvent_t ev_next, ev_piv, ev_bulk, ev_pan;
  int gmax;
  int pivot;            // kPivotBlock or kPivotPartial
  int nb = 1;           // systems in the batch (block pivoting only when > 1)
  BatchStrides bs;      // their strides; the pointers above are system 0's
  const int *sys = nullptr;  // lu_solve on a subset: its nsys systems' indices (null: all nb)
  int nsys = 0;
  int trsv = 0;              // triangular-solve variant (kTrsvVariants; 0: the default)
  bool split = true;             // FP32 factors: bulk updates on the split-bf16 MFMA
  hipEvent_t ev_mark = nullptr;  // when set: recorded on the bulk stream once the
  int mark_col = 0;              // factorisation has passed column mark_col, and
  void (*on_mark)(void *) = nullptr;  // then on_mark(mark_arg) is called (on the host,
  void *mark_arg = nullptr;           // in the middle of the enqueueing)
};
// Elements of T in a factorisation's dinv: 2 NB^2 per panel.
inline size_t dinv_elems(int Np) { return 2 * (size_t)64 * Np; }

// Pivot search of a panel: among the rows of its diagonal block only
// (followed by iterative refinement), or over the whole column (tournament).
constexpr int kPivotBlock = 0;
constexpr int kPivotPartial = 1;

// Optional per-launch timing of the factorisation (events on the same stream).
struct LuProf {
  hipEvent_t *ev = nullptr;  // 4 events per panel pair: panel work start/end (look-ahead
                             // stream), bulk update start/end (main stream)
  int cap = 0;               // pairs the pool covers
  int used = 0;              // pairs recorded by the last factorisation
  double gemm_flop = 0.0;    // algorithmic flop of the recorded bulk updates (8 m n k each)
  int gemm_launches = 0;     // their kernel launches
};

// Factor the Np x Np complex matrix in place (P A = L U); T = double2 or float2.
template <typename T>
int lu_factor(T *A, int Np, size_t lda, LuWork &w, hipStream_t st, LuProf *prof = nullptr);
// Triangular-solve variants: super-block size (64-row blocks per workgroup)
// and when the near-phase tiles are loaded (0: as used, 1: at kernel start,
// 2: after the far phase).  Variant 0 is the default of each type.
constexpr int kTrsvVariants = 3;
// One bulk update of the shape the factorisation's first panel pair makes
// (M = N = Np - 128, K = 128) on nb matrices sA apart: a timing probe for the
// autotuner (A's contents are overwritten).
template <typename T>
void bulk_update_probe(T *A, int Np, size_t lda, int nb, size_t sA, bool split, hipStream_t st);
// Solve with the factors: x = A^{-1} b, vectors always complex FP64.
template <typename T>
void lu_solve(const T *A, int Np, size_t lda, LuWork &w, const cplx *b, cplx *x, hipStream_t st);

// ---------------------------------------------------------------- iterative
// (sys non-null: only the nsys systems it lists)
void gemv64(const cplx *A, size_t lda, int N, const cplx *x, cplx *y, cplx *part, hipStream_t st, int nb = 1,
            size_t sA = 0, size_t sv = 0, const int *sys = nullptr, int nsys = 0);
// y = A x with A the complex-FP32 copy (FP64 vectors and sums); part as gemv64's
void gemv32(const float2 *A, size_t lda, int N, const cplx *x, cplx *y, cplx *part, hipStream_t st);
constexpr int kGemvSplitMax = 32;

// level-1 helpers (omx_iter.hip); reductions return through small device arrays
void vcopy(cplx *y, const cplx *x, int n, hipStream_t st);
void demote(float2 *y, const cplx *x, size_t n, hipStream_t st);                 // y = (c64) x
// p[0..n) = 0 by a kernel (a memset node must not go into a captured graph)
void zero_ints(int *p, int n, hipStream_t st);
// dst (Nd x Nd) = src (Ns x Ns, Ns <= Nd) bordered by the identity, in FP64 or
// rounded to complex FP32; dst[0, Nd) = src[0, Ns) followed by zeros
void pad_matrix64(cplx *dst, int Nd, const cplx *src, int Ns, hipStream_t st);
void pad_matrix32(float2 *dst, int Nd, const cplx *src, int Ns, hipStream_t st);
void pad_vector(cplx *dst, int Nd, const cplx *src, int Ns, hipStream_t st);
void vaxpby(cplx *y, cplx a, const cplx *x, cplx b, int n, hipStream_t st);   // y = a x + b y
void vsub(cplx *z, const cplx *x, const cplx *y, int n, hipStream_t st);      // z = x - y
void vmul_ov(cplx *y, const cplx *x, const cplx *ov, int kind, int P, int nmax, const cplx *sub, int n,
             hipStream_t st);                                                  // y = x (.) ov[kind] - sub
void dotc_multi(const cplx *V, size_t ldv, int nv, const cplx *w, int n, cplx *part, cplx *out,
                hipStream_t st);                                               // out[k] = V_k^H w
void axpy_multi(const cplx *V, size_t ldv, int nv, const cplx *h, cplx *w, int n, hipStream_t st); // w -= V h
void gemv_multi(const cplx *V, size_t ldv, int nv, const cplx *h, cplx *z, int n, hipStream_t st); // z = V h
void nrm2sq(const cplx *x, int n, double *part, double *out, hipStream_t st);
// the same for nb systems side by side (vectors sv, sy, sx apart, bases sV apart,
// per-system scalars a[z], b[z], coefficient vectors sh apart, outputs so apart);
// act (may be null) skips the systems whose flag is 0
void vaxpby_b(cplx *y, size_t sy, const cplx *a, const cplx *x, size_t sx, const cplx *b, int n, int nb,
              const int *act, hipStream_t st);
void vsub_b(cplx *z, const cplx *x, const cplx *y, int n, int nb, size_t sv, hipStream_t st);
int red_blocks_b(int n, int nb);  // partial sums per system of the reductions below
void dotc_multi_b(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *w, size_t sv, int n, int nb, cplx *part,
                  cplx *out, int so, hipStream_t st);
void axpy_multi_b(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *h, int sh, cplx *w, size_t sv, int n,
                  int nb, hipStream_t st);
void gemv_multi_b(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *h, int sh, cplx *z, size_t sv, int n,
                  int nb, hipStream_t st);
void nrm2sq_b(const cplx *x, size_t sv, int n, int nb, double *part, double *out, hipStream_t st);

}  // namespace omx
