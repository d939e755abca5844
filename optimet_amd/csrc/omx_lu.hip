// omx_lu.hip -- blocked complex LU with partial pivoting for CDNA4.
//
// Right-looking blocked factorisation P A = L U of a column-major Np x Np
// complex matrix (Np a multiple of the block size NB), in complex FP64
// (double2) or complex FP32 (float2):
//   tslu_leaf/merge  panel pivoting by tournament (TSLU): register-resident
//                  partial-pivoting LUs of 128-row chunks nominate candidate
//                  rows, a binary tree of merges picks the NB pivot rows; the
//                  root also emits the LU factors of the selected block
//   tslu_finish    one wave turns the winners into the swap sequence and a
//                  gather map
//   laswp_map      the panel's row interchanges on every column, as one
//                  parallel gather/scatter
//   panel_inv      writes the selected block's factors and inverts L11, U11
//   blk_mul        the two panel triangular solves as in-place MFMA block
//                  products: U12 = L11^{-1} A12, L21 = A21 U11^{-1}
//   gemm_kernel    A22 -= L21 U12 as complex products on the matrix cores:
//                  v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32
//                  (FP32), four real MFMAs per complex tile product, planar
//                  re/im operands staged through LDS
//   trsv_kernel    sync-free blocked triangular solves: workgroup k owns
//                  block row k, consumes earlier blocks as their flags are
//                  published and solves its diagonal block in one wave
// Look-ahead: the tournament of panel k+1 runs on a second stream while the
// bulk of panel k's trailing update runs on the first.
// This replaces the ScaLAPACK pzgesv path (scalapack/LinearSystemSolver.hpp:84-91)
// and Eigen's ColPivHouseholderQR (PreconditionedMatrixSolver.h:58, 75).
#include <climits>
#include <vector>

#include "omx_kernels.hpp"

namespace omx {

template <typename T> struct CT;
// MFMA 16x16x4: lane l feeds A[l%16][l/16] and B[l/16][l%16]; accumulator
// register v of lane l holds D[row(l, v)][l%16] with
//   FP64 (v_mfma_f64_16x16x4f64): row = 4 v + l/16
//   FP32 (v_mfma_f32_16x16x4f32): row = 4 (l/16) + v
template <> struct CT<double2> {
  using R = double;
  typedef double V4 __attribute__((ext_vector_type(4)));
  static __device__ inline V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <> struct CT<float2> {
  using R = float;
  typedef float V4 __attribute__((ext_vector_type(4)));
  static __device__ inline V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Plain complex pair for register-resident arrays (the HIP vector types are
// unions that the compiler keeps in scratch memory when held in arrays).
template <typename R> struct Cx {
  R x, y;
};
template <typename T> using CxOf = Cx<decltype(T{}.x)>;
template <typename R> __device__ inline Cx<R> cx(R a, R b) { return Cx<R>{a, b}; }

template <typename T> __device__ inline T cm(T a, T b) {
  T r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T> __device__ inline T csub(T a, T b) {
  T r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
template <typename T> __device__ inline double mag2(T a) { return (double)a.x * a.x + (double)a.y * a.y; }
template <typename T> __device__ inline T crecip(T a) {
  // 1 / a, scaled to avoid overflow
  using R = decltype(a.x);
  T r;
  if(fabs((double)a.x) >= fabs((double)a.y)) {
    R q = a.y / a.x, d = a.x + a.y * q;
    r.x = R(1) / d;
    r.y = -q / d;
  } else {
    R q = a.x / a.y, d = a.x * q + a.y;
    r.x = q / d;
    r.y = R(-1) / d;
  }
  return r;
}

__global__ void iota_kernel(int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = i;
}

// ------------------------------------------------------------------ panel
// Tournament pivoting (TSLU, Grigori, Demmel and Xiang 2011): the panel's
// pivot rows are chosen by a reduction tree of local partial-pivoting LUs
// instead of one device-wide search per column.

template <typename T, int NB>
struct SelShared {
  CxOf<T> rowb[2][2][NB];  // [parity][0: old row c, 1: pivot row][column]
  CxOf<T> lcol[128];
  int rid[128];
  int sel[NB];
  double wv[2];
  int wi[2];
};

// Partial-pivoting LU of an R x 64 block (R <= 128) held in registers: thread
// t owns row r = t & 127, columns h*32 .. h*32+31 with h = t >> 7.  Rows carry
// global ids in s.rid; s.sel receives the ids of the pivot rows in order.
template <typename T>
__device__ __forceinline__ void reg_select(CxOf<T> (&v)[32], SelShared<T, 64> &s, int R) {
  using C = CxOf<T>;
  const int t = threadIdx.x, r = t & 127, h = t >> 7, lane = t & 63, w = t >> 6;
  const C zero{};
  const int nc = min(R, 64);
  // two halves of 32 columns; inside a half the column index is a
  // compile-time constant, so v[] stays in registers
#pragma unroll 1
  for(int hc = 0; hc < 2; ++hc) {
#pragma unroll
    for(int jc = 0; jc < 32; ++jc) {
      const int c = hc * 32 + jc, par = jc & 1;
      if(c < nc) {
        if(h == hc) {
          const bool ok = r >= c && r < R;
          double val = ok ? mag2(v[jc]) : -1.0;
          int idx = ok ? r : INT_MAX;
#pragma unroll
          for(int off = 32; off > 0; off >>= 1) {
            double ov = __shfl_xor(val, off, 64);
            int oi = __shfl_xor(idx, off, 64);
            if(ov > val || (ov == val && oi < idx)) val = ov, idx = oi;
          }
          if(lane == 0) {
            s.wv[w - 2 * hc] = val;
            s.wi[w - 2 * hc] = idx;
          }
        }
        __syncthreads();
        const int b = (s.wv[1] > s.wv[0] || (s.wv[1] == s.wv[0] && s.wi[1] < s.wi[0])) ? s.wi[1] : s.wi[0];
        if(r == b) {
#pragma unroll
          for(int j = 0; j < 32; ++j) s.rowb[par][1][h * 32 + j] = v[j];
        }
        if(r == c) {
#pragma unroll
          for(int j = 0; j < 32; ++j) s.rowb[par][0][h * 32 + j] = v[j];
        }
        if(t == 0) {
          const int id = s.rid[b];
          s.rid[b] = s.rid[c];
          s.rid[c] = id;
          s.sel[c] = id;
        }
        __syncthreads();
        if(b != c) {
          if(r == c) {
#pragma unroll
            for(int j = 0; j < 32; ++j) v[j] = s.rowb[par][1][h * 32 + j];
          } else if(r == b) {
#pragma unroll
            for(int j = 0; j < 32; ++j) v[j] = s.rowb[par][0][h * 32 + j];
          }
        }
        const C pv = s.rowb[par][1][c];
        const C ip = mag2(pv) > 0.0 ? crecip(pv) : zero;
        C l = zero;
        if(h == hc && r > c && r < R) {
          l = cm(v[jc], ip);
          v[jc] = l;
          s.lcol[r] = l;
        }
        __syncthreads();
        if(r > c && r < R) {
          if(h != hc) l = s.lcol[r];
          if(h == hc) {
#pragma unroll
            for(int j = jc + 1; j < 32; ++j) v[j] = csub(v[j], cm(l, s.rowb[par][1][h * 32 + j]));
          } else if(h > hc) {
#pragma unroll
            for(int j = 0; j < 32; ++j) v[j] = csub(v[j], cm(l, s.rowb[par][1][h * 32 + j]));
          }
        }
      }
    }
  }
}

// Root output: the LU factors of the selected rows (top 64 rows after the
// local pivoting), column-major 64 x 64.
template <typename T>
__device__ __forceinline__ void emit_factors(const CxOf<T> (&v)[32], T *wlu) {
  const int t = threadIdx.x, r = t & 127, h = t >> 7;
  if(r < 64) {
#pragma unroll
    for(int j = 0; j < 32; ++j) {
      T o;
      o.x = v[j].x;
      o.y = v[j].y;
      wlu[(h * 32 + j) * 64 + r] = o;
    }
  }
}

// Leaves: workgroup g factors panel rows [j0 + 128 g, ...) and nominates its
// pivot rows; with a single leaf it is also the root.
template <typename T>
__global__ __launch_bounds__(256) void tslu_leaf(const T *A, size_t lda, int Np, int j0, int *cand, int *cnt,
                                                 T *wlu) {
  __shared__ SelShared<T, 64> s;
  const int g = blockIdx.x, t = threadIdx.x, r = t & 127, h = t >> 7;
  const int r0 = j0 + g * 128, nr = min(128, Np - r0);
  CxOf<T> v[32];
#pragma unroll
  for(int j = 0; j < 32; ++j) {
    T a = r < nr ? A[(size_t)(j0 + h * 32 + j) * lda + r0 + r] : T{};
    v[j] = CxOf<T>{a.x, a.y};
  }
  if(h == 0) s.rid[r] = r0 + r;
  __syncthreads();
  reg_select<T>(v, s, nr);
  __syncthreads();
  const int k = min(nr, 64);
  if(t < k) cand[g * 64 + t] = s.sel[t];
  if(t == 0) cnt[g] = k;
  if(gridDim.x == 1) emit_factors<T>(v, wlu);
}

// One tree level: workgroup g plays candidate sets 2g and 2g+1.
template <typename T>
__global__ __launch_bounds__(256) void tslu_merge(const T *A, size_t lda, int j0, int gin, const int *cin,
                                                  const int *nin, int *cout, int *nout, T *wlu) {
  __shared__ SelShared<T, 64> s;
  const int g = blockIdx.x, t = threadIdx.x, r = t & 127, h = t >> 7;
  const int a = 2 * g, b = 2 * g + 1;
  const int na = nin[a], nb = b < gin ? nin[b] : 0;
  const int R = na + nb;
  if(h == 0 && r < R) s.rid[r] = r < na ? cin[a * 64 + r] : cin[b * 64 + (r - na)];
  __syncthreads();
  CxOf<T> v[32];
  const int id = r < R ? s.rid[r] : 0;
#pragma unroll
  for(int j = 0; j < 32; ++j) {
    T a = r < R ? A[(size_t)(j0 + h * 32 + j) * lda + id] : T{};
    v[j] = CxOf<T>{a.x, a.y};
  }
  reg_select<T>(v, s, R);
  __syncthreads();
  const int k = min(R, 64);
  if(t < k) cout[g * 64 + t] = s.sel[t];
  if(t == 0) nout[g] = k;
  if(gridDim.x == 1) emit_factors<T>(v, wlu);
}

// One wave: turns the winners sel[0..nsel) into the LAPACK swap sequence
// piv[j0..j0+nsel) and into a gather map (position dst[i] receives the row
// that was at src[i]); also applies the map to the running permutation.
__global__ __launch_bounds__(64) void tslu_finish(const int *sel, const int *nsel_p, int j0, int *piv, int *mdst,
                                                 int *msrc, int *mcnt, int *perm) {
  const int l = threadIdx.x;  // 64 lanes; NB == 64
  __shared__ int top[64];
  const int nsel = *nsel_p;
  top[l] = j0 + l;
  int opos = -1, ocont = -1;  // lane s: outside slot s
  int nout = 0;
  __syncthreads();
  for(int jj = 0; jj < nsel; ++jj) {
    const int X = sel[jj];
    // current position of row X
    unsigned long long mt = __ballot(top[l] == X);
    unsigned long long mo = __ballot(l < nout && ocont == X);
    int pos;
    if(mt) pos = j0 + __ffsll((long long)mt) - 1;
    else if(mo) pos = __shfl(opos, __ffsll((long long)mo) - 1, 64);
    else pos = X;
    const int cur = top[jj];
    if(l == 0) piv[j0 + jj] = pos;
    if(pos != j0 + jj) {
      if(pos >= j0 && pos < j0 + 64) {
        if(l == 0) {
          top[pos - j0] = cur;
          top[jj] = X;
        }
      } else {
        unsigned long long mp = __ballot(l < nout && opos == pos);
        int slot = mp ? __ffsll((long long)mp) - 1 : nout;
        if(l == slot) {
          opos = pos;
          ocont = cur;
        }
        if(!mp) ++nout;
        if(l == 0) top[jj] = X;
      }
    }
    __syncthreads();
  }
  // gather map: top positions, then outside slots; identity entries dropped
  const int t0 = top[l];
  const bool mv0 = t0 != j0 + l;
  const bool mv1 = l < nout;
  int v0 = mv0 ? perm[t0] : 0, v1 = mv1 ? perm[ocont] : 0;
  unsigned long long b0 = __ballot(mv0), b1 = __ballot(mv1);
  const int n0 = __popcll(b0);
  const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
  __syncthreads();
  if(mv0) {
    int i = __popcll(b0 & below);
    mdst[i] = j0 + l;
    msrc[i] = t0;
    perm[j0 + l] = v0;
  }
  if(mv1) {
    int i = n0 + __popcll(b1 & below);
    mdst[i] = opos;
    msrc[i] = ocont;
    perm[opos] = v1;
  }
  if(l == 0) *mcnt = n0 + __popcll(b1);
}

// Writes the selected block's factors into A and inverts L11 (unit lower)
// and U11 (thread j builds column j of each inverse, one after the other).  wlu holds [factors | Linv | Uinv], each 64 x 64 column-major.
template <typename T>
__global__ __launch_bounds__(256) void panel_inv(T *A, size_t lda, int j0, T *wlu, int *status) {
  __shared__ T F[64 * 65];
  __shared__ T X[64 * 64];   // row-major: X[i * 64 + j]; L11^{-1}, then U11^{-1}
  const int t = threadIdx.x;
  for(int idx = t; idx < 64 * 64; idx += 256) {
    int c = idx / 64, r = idx % 64;
    T v = wlu[idx];
    F[c * 65 + r] = v;
    A[(size_t)(j0 + c) * lda + j0 + r] = v;
  }
  __syncthreads();
  const T zero{};
  if(t < 64) {  // column j of L11^{-1} (unit lower)
    const int j = t;
    for(int i = 0; i < 64; ++i) {
      T x = (i == j) ? T{1, 0} : zero;
      if(i > j)
        for(int k = j; k < i; ++k) x = csub(x, cm(F[k * 65 + i], X[k * 64 + j]));
      X[i * 64 + j] = (i < j) ? zero : x;
    }
  }
  __syncthreads();
  for(int idx = t; idx < 64 * 64; idx += 256) {
    int c = idx / 64, r = idx % 64;
    wlu[4096 + idx] = X[r * 64 + c];
  }
  __syncthreads();
  if(t < 64) {  // column j of U11^{-1}
    const int j = t;
    for(int i = 63; i >= 0; --i) {
      T y = zero;
      if(i <= j) {
        y = (i == j) ? T{1, 0} : zero;
        for(int k = i + 1; k <= j; ++k) y = csub(y, cm(F[k * 65 + i], X[k * 64 + j]));
        const T d = F[i * 65 + i];
        if(!(mag2(d) > 0.0)) {
          atomicCAS(&status[0], 0, j0 + i + 1);
          y = zero;
        } else {
          y = cm(y, crecip(d));
        }
      }
      X[i * 64 + j] = y;
    }
  }
  __syncthreads();
  for(int idx = t; idx < 64 * 64; idx += 256) {
    int c = idx / 64, r = idx % 64;
    wlu[8192 + idx] = X[r * 64 + c];
  }
}

// Diagonal-block pivoting: LU with partial pivoting restricted to the 64 x 64
// diagonal block of the panel (rows are chosen among the block's rows; the
// interchanges are applied to whole rows), then the inverses of L11 and U11
// by Gauss-Jordan sweeps.  One workgroup: wave q holds columns 16q..16q+15 of
// every row, lane = row, and rows are not moved during the elimination (the
// pivot order is recorded instead).  Writes the factored block into A, the
// inverses into dinv ([L11^{-1} | U11^{-1}], column-major) and the block's row
// map (mdst <- msrc) for the remaining columns and the running permutation.
template <typename T>
__global__ __launch_bounds__(256) void diag_lu(T *A, size_t lda, int j0, T *dinv, int *mdst, int *msrc, int *mcnt,
                                               int *perm, int *status) {
  using C = CxOf<T>;
  __shared__ C pr[64];
  __shared__ C lc[64];
  __shared__ int pb;
  __shared__ int prow[64];
  __shared__ C F[64 * 65];  // factored block in pivoted order, column-major [c * 65 + r]
  __shared__ C X[64 * 64];  // inverse under construction, row-major [i * 64 + j]
  const int t = threadIdx.x, lane = t & 63, q = t >> 6;
  const C zero{};
  C v[16];
#pragma unroll
  for(int j = 0; j < 16; ++j) {
    T a = A[(size_t)(j0 + 16 * q + j) * lda + j0 + lane];
    v[j] = C{a.x, a.y};
  }
  bool used = false;
  int mypos = 0;
#pragma unroll 1
  for(int qc = 0; qc < 4; ++qc) {
#pragma unroll
    for(int jj = 0; jj < 16; ++jj) {
      const int c = 16 * qc + jj;
      if(q == qc) {
        double val = used ? -1.0 : mag2(v[jj]);
        int idx = used ? INT_MAX : lane;
#pragma unroll
        for(int off = 32; off > 0; off >>= 1) {
          double ov = __shfl_xor(val, off, 64);
          int oi = __shfl_xor(idx, off, 64);
          if(ov > val || (ov == val && oi < idx)) val = ov, idx = oi;
        }
        if(lane == 0) {
          pb = idx;
          prow[c] = idx;
        }
      }
      __syncthreads();
      const int b = pb;
      if(lane == b) {
#pragma unroll
        for(int j = 0; j < 16; ++j) pr[16 * q + j] = v[j];
        used = true;
        mypos = c;
      }
      if(q == qc) {
        C piv;
        piv.x = __shfl(v[jj].x, b, 64);
        piv.y = __shfl(v[jj].y, b, 64);
        const bool sing = !(mag2(piv) > 0.0);
        if(sing && lane == 0) atomicCAS(&status[0], 0, j0 + c + 1);
        const C ip = sing ? zero : crecip(piv);
        C l = zero;
        if(!used) {
          l = cm(v[jj], ip);
          v[jj] = l;
        }
        lc[lane] = l;
      }
      __syncthreads();
      if(!used) {
        const C l = (q == qc) ? v[jj] : lc[lane];
#pragma unroll
        for(int j = 0; j < 16; ++j)
          if(16 * q + j > c) v[j] = csub(v[j], cm(l, pr[16 * q + j]));
      }
    }
  }
  // factored block in pivoted order
#pragma unroll
  for(int j = 0; j < 16; ++j) {
    const int col = 16 * q + j;
    F[col * 65 + mypos] = v[j];
    T o;
    o.x = v[j].x;
    o.y = v[j].y;
    A[(size_t)(j0 + col) * lda + j0 + mypos] = o;
  }
  // row map and running permutation (wave 0)
  if(q == 0) {
    const int src = prow[lane];
    const bool mv = src != lane;
    const int old = perm[j0 + src];
    const unsigned long long bm = __ballot(mv);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if(mv) {
      const int i = __popcll(bm & below);
      mdst[i] = j0 + lane;
      msrc[i] = j0 + src;
    }
    if(lane == 0) *mcnt = __popcll(bm);
    __builtin_amdgcn_wave_barrier();
    perm[j0 + lane] = old;
  }
  __syncthreads();
  // L11^{-1}: X = I; for k: X[i][j] -= L[i][k] X[k][j], i > k, j <= k
  const int i = t & 63, jb = t >> 6;
  for(int m = 0; m < 16; ++m) {
    const int j = jb + 4 * m;
    X[i * 64 + j] = (i == j) ? C{1, 0} : zero;
  }
  __syncthreads();
  for(int k = 0; k < 63; ++k) {
    if(i > k) {
      const C lik = F[k * 65 + i];
      for(int m = 0; m < 16; ++m) {
        const int j = jb + 4 * m;
        if(j <= k) X[i * 64 + j] = csub(X[i * 64 + j], cm(lik, X[k * 64 + j]));
      }
    }
    __syncthreads();
  }
  for(int m = 0; m < 16; ++m) {
    const int j = jb + 4 * m;
    const C x = X[i * 64 + j];
    T o;
    o.x = x.x;
    o.y = x.y;
    dinv[j * 64 + i] = o;
  }
  __syncthreads();
  // U11^{-1}: Y = I; for k = 63..0: Y[k][:] /= U[k][k]; Y[i][:] -= U[i][k] Y[k][:], i < k
  for(int m = 0; m < 16; ++m) {
    const int j = jb + 4 * m;
    X[i * 64 + j] = (i == j) ? C{1, 0} : zero;
  }
  __syncthreads();
  for(int k = 63; k >= 0; --k) {
    const C d = F[k * 65 + k];
    const C id = mag2(d) > 0.0 ? crecip(d) : zero;
    C yk[16];
    for(int m = 0; m < 16; ++m) yk[m] = cm(X[k * 64 + jb + 4 * m], id);
    __syncthreads();
    if(i == k) {
      for(int m = 0; m < 16; ++m) X[k * 64 + jb + 4 * m] = yk[m];
    } else if(i < k) {
      const C uik = F[k * 65 + i];
      for(int m = 0; m < 16; ++m) {
        const int j = jb + 4 * m;
        if(j >= k) X[i * 64 + j] = csub(X[i * 64 + j], cm(uik, yk[m]));
      }
    }
    __syncthreads();
  }
  for(int m = 0; m < 16; ++m) {
    const int j = jb + 4 * m;
    const C x = X[i * 64 + j];
    T o;
    o.x = x.x;
    o.y = x.y;
    dinv[4096 + j * 64 + i] = o;
  }
}

// In-place block product with a 64 x 64 factor M (column-major, ld 64):
// LEFT: the 64 x 64 block at (j0, c0) becomes M * block (U12 = L11^{-1} A12);
// RIGHT: the block at (r0, j0) becomes block * M (L21 = A21 U11^{-1}).
template <typename T, bool LEFT>
__global__ __launch_bounds__(256) void blk_mul(T *A, size_t lda, int j0, int first, const T *M) {
  using R = typename CT<T>::R;
  using V4 = typename CT<T>::V4;
  constexpr int PAD = 4;
  __shared__ R As[2][64][64 + PAD];  // [plane][k][m]
  __shared__ R Bs[2][64][64 + PAD];  // [plane][k][n]
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, wm = wv & 1, wn = wv >> 1;
  const int off = first + blockIdx.x * 64;  // column (LEFT) or row (RIGHT) origin of the block
  for(int idx = t; idx < 64 * 64; idx += 256) {
    const int i = idx & 63, j = idx >> 6;  // i fastest
    if(LEFT) {
      T m = M[j * 64 + i];                              // M(i, k=j)
      As[0][j][i] = m.x;
      As[1][j][i] = m.y;
      T b = A[(size_t)(off + j) * lda + j0 + i];        // block(k=i, n=j)
      Bs[0][i][j] = b.x;
      Bs[1][i][j] = b.y;
    } else {
      T a = A[(size_t)(j0 + j) * lda + off + i];        // block(m=i, k=j)
      As[0][j][i] = a.x;
      As[1][j][i] = a.y;
      T m = M[j * 64 + i];                              // M(k=i, n=j)
      Bs[0][i][j] = m.x;
      Bs[1][i][j] = m.y;
    }
  }
  __syncthreads();
  V4 cr[2][2], ci[2][2];
#pragma unroll
  for(int a = 0; a < 2; ++a)
#pragma unroll
    for(int b = 0; b < 2; ++b) {
      cr[a][b] = V4{0, 0, 0, 0};
      ci[a][b] = V4{0, 0, 0, 0};
    }
#pragma unroll 4
  for(int ks = 0; ks < 16; ++ks) {
    const int kr = ks * 4 + (lane >> 4), li = lane & 15;
    R ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
    for(int a = 0; a < 2; ++a) {
      ar[a] = As[0][kr][wm * 32 + a * 16 + li];
      ai[a] = As[1][kr][wm * 32 + a * 16 + li];
      nai[a] = -ai[a];
      br[a] = Bs[0][kr][wn * 32 + a * 16 + li];
      bi[a] = Bs[1][kr][wn * 32 + a * 16 + li];
    }
#pragma unroll
    for(int a = 0; a < 2; ++a)
#pragma unroll
      for(int b = 0; b < 2; ++b) {
        cr[a][b] = CT<T>::mfma(ar[a], br[b], cr[a][b]);
        cr[a][b] = CT<T>::mfma(nai[a], bi[b], cr[a][b]);
        ci[a][b] = CT<T>::mfma(ar[a], bi[b], ci[a][b]);
        ci[a][b] = CT<T>::mfma(ai[a], br[b], ci[a][b]);
      }
  }
#pragma unroll
  for(int a = 0; a < 2; ++a)
#pragma unroll
    for(int b = 0; b < 2; ++b) {
      const int n = wn * 32 + b * 16 + (lane & 15);
#pragma unroll
      for(int v = 0; v < 4; ++v) {
        const int m = wm * 32 + a * 16 + CT<T>::row(lane, v);
        T o;
        o.x = cr[a][b][v];
        o.y = ci[a][b][v];
        if(LEFT)
          A[(size_t)(off + n) * lda + j0 + m] = o;
        else
          A[(size_t)(j0 + n) * lda + off + m] = o;
      }
    }
}

// ------------------------------------------------------------------ swaps
// Applies the panel's gather map to every column: each workgroup handles 8
// columns, reads all moved entries into registers, then writes them back.
template <typename T>
__global__ __launch_bounds__(256) void laswp_map(T *A, size_t lda, int Np, const int *__restrict__ mdst,
                                                 const int *__restrict__ msrc, const int *__restrict__ mcnt) {
  const int m = *mcnt;
  const int cl = threadIdx.x / 32, li = threadIdx.x % 32;
  const int col = blockIdx.x * 8 + cl;
  T v[4];
  int d[4];
  if(col < Np) {
    T *a = A + (size_t)col * lda;
#pragma unroll
    for(int k = 0; k < 4; ++k) {
      int i = li + 32 * k;
      d[k] = i < m ? mdst[i] : -1;
      if(i < m) v[k] = a[msrc[i]];
    }
  }
  __syncthreads();
  if(col < Np) {
    T *a = A + (size_t)col * lda;
#pragma unroll
    for(int k = 0; k < 4; ++k)
      if(d[k] >= 0) a[d[k]] = v[k];
  }
}

// ------------------------------------------------------------------ gemm
// C[row0 + m, col0 + n] -= sum_{k < NB} A[row0 + m, k0 + k] A[k0 + k, col0 + n]
// Tile 64 x 64 per workgroup, four waves in 2 x 2, each wave 2 x 2 MFMA
// 16 x 16 blocks with separate real and imaginary accumulators.
template <typename T, int NB>
__global__ __launch_bounds__(256) void gemm_kernel(T *A, size_t lda, int row0, int col0, int k0) {
  using R = typename CT<T>::R;
  using V4 = typename CT<T>::V4;
  constexpr int BK = 16, PAD = 4;
  __shared__ R As[2][BK][64 + PAD];
  __shared__ R Bs[2][BK][64 + PAD];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, wm = wv & 1, wn = wv >> 1;
  const int bm = row0 + blockIdx.x * 64, bn = col0 + blockIdx.y * 64;
  V4 cr[2][2], ci[2][2];
#pragma unroll
  for(int a = 0; a < 2; ++a)
#pragma unroll
    for(int b = 0; b < 2; ++b) {
      cr[a][b] = V4{0, 0, 0, 0};
      ci[a][b] = V4{0, 0, 0, 0};
    }
  for(int kc = 0; kc < NB; kc += BK) {
#pragma unroll
    for(int it = 0; it < (64 * BK) / 256; ++it) {
      int idx = t + it * 256;
      int r = idx & 63, kk = idx >> 6;
      T v = A[(size_t)(k0 + kc + kk) * lda + bm + r];
      As[0][kk][r] = v.x;
      As[1][kk][r] = v.y;
      int kb = idx % BK, c = idx / BK;
      T wv2 = A[(size_t)(bn + c) * lda + k0 + kc + kb];
      Bs[0][kb][c] = wv2.x;
      Bs[1][kb][c] = wv2.y;
    }
    __syncthreads();
#pragma unroll
    for(int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4), li = lane & 15;
      R ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for(int a = 0; a < 2; ++a) {
        ar[a] = As[0][kr][wm * 32 + a * 16 + li];
        ai[a] = As[1][kr][wm * 32 + a * 16 + li];
        nai[a] = -ai[a];
        br[a] = Bs[0][kr][wn * 32 + a * 16 + li];
        bi[a] = Bs[1][kr][wn * 32 + a * 16 + li];
      }
#pragma unroll
      for(int a = 0; a < 2; ++a)
#pragma unroll
        for(int b = 0; b < 2; ++b) {
          cr[a][b] = CT<T>::mfma(ar[a], br[b], cr[a][b]);
          cr[a][b] = CT<T>::mfma(nai[a], bi[b], cr[a][b]);
          ci[a][b] = CT<T>::mfma(ar[a], bi[b], ci[a][b]);
          ci[a][b] = CT<T>::mfma(ai[a], br[b], ci[a][b]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for(int a = 0; a < 2; ++a)
#pragma unroll
    for(int b = 0; b < 2; ++b) {
      int col = bn + wn * 32 + b * 16 + (lane & 15);
      int rowb = bm + wm * 32 + a * 16;
      T *cp = A + (size_t)col * lda + rowb;
#pragma unroll
      for(int v = 0; v < 4; ++v) {
        int rr = CT<T>::row(lane, v);
        T c = cp[rr];
        c.x -= cr[a][b][v];
        c.y -= ci[a][b][v];
        cp[rr] = c;
      }
    }
}

// ------------------------------------------------------------------ driver
template <typename T>
int lu_factor(T *A, int Np, size_t lda, LuWork &w, hipStream_t st, LuProf *prof) {
  if(Np % kNB) return -1;
  hipStream_t sb = w.aux;
  T *wlu = reinterpret_cast<T *>(w.wlu);
  hipLaunchKernelGGL(iota_kernel, dim3((Np + 255) / 256), dim3(256), 0, st, w.perm, Np);
  if(prof) {
    prof->used = 0;
    prof->gemm_flop = 0.0;
  }
  auto tournament = [&](int j0) {
    const int G = (Np - j0 + kPanelRows - 1) / kPanelRows;
    hipLaunchKernelGGL((tslu_leaf<T>), dim3(G), dim3(256), 0, sb, A, lda, Np, j0, w.cand[0], w.cnt[0], wlu);
    int gin = G, buf = 0;
    while(gin > 1) {
      const int gout = (gin + 1) / 2;
      hipLaunchKernelGGL((tslu_merge<T>), dim3(gout), dim3(256), 0, sb, A, lda, j0, gin, w.cand[buf], w.cnt[buf],
                         w.cand[buf ^ 1], w.cnt[buf ^ 1], wlu);
      buf ^= 1;
      gin = gout;
    }
    hipLaunchKernelGGL(tslu_finish, dim3(1), dim3(64), 0, sb, w.cand[buf], w.cnt[buf], j0, w.piv, w.mdst, w.msrc,
                       w.mcnt, w.perm);
    (void)hipEventRecord(w.ev_piv, sb);
  };
  if((Np - 0 + kPanelRows - 1) / kPanelRows > w.gmax) return -2;
  (void)hipEventRecord(w.ev_next, st);
  (void)hipStreamWaitEvent(sb, w.ev_next, 0);
  tournament(0);
  for(int j0 = 0; j0 < Np; j0 += kNB) {
    const int pi = j0 / kNB;
    const bool rec = prof && pi < prof->cap;
    const int rest = Np - j0 - kNB;
    (void)hipStreamWaitEvent(st, w.ev_piv, 0);
    if(rec) (void)hipEventRecord(prof->ev[4 * pi], st);
    hipLaunchKernelGGL((laswp_map<T>), dim3((Np + 7) / 8), dim3(256), 0, st, A, lda, Np, w.mdst, w.msrc, w.mcnt);
    hipLaunchKernelGGL((panel_inv<T>), dim3(1), dim3(256), 0, st, A, lda, j0, wlu, w.status);
    if(rest > 0) {
      hipLaunchKernelGGL((blk_mul<T, false>), dim3(rest / 64), dim3(256), 0, st, A, lda, j0, j0 + kNB,
                         wlu + 8192);
      hipLaunchKernelGGL((blk_mul<T, true>), dim3(rest / 64), dim3(256), 0, st, A, lda, j0, j0 + kNB, wlu + 4096);
    }
    if(rec) (void)hipEventRecord(prof->ev[4 * pi + 1], st);
    if(rest > 0) {
      if(rec) (void)hipEventRecord(prof->ev[4 * pi + 2], st);
      // next panel's columns first, then its tournament overlaps the rest
      hipLaunchKernelGGL((gemm_kernel<T, kNB>), dim3(rest / 64, 1), dim3(256), 0, st, A, lda, j0 + kNB, j0 + kNB,
                         j0);
      (void)hipEventRecord(w.ev_next, st);
      (void)hipStreamWaitEvent(sb, w.ev_next, 0);
      tournament(j0 + kNB);
      if(rest > 64)
        hipLaunchKernelGGL((gemm_kernel<T, kNB>), dim3(rest / 64, rest / 64 - 1), dim3(256), 0, st, A, lda,
                           j0 + kNB, j0 + 2 * kNB, j0);
      if(rec) {
        (void)hipEventRecord(prof->ev[4 * pi + 3], st);
        prof->gemm_flop += 8.0 * (double)rest * rest * kNB;
      }
    } else if(rec) {
      (void)hipEventRecord(prof->ev[4 * pi + 2], st);
      (void)hipEventRecord(prof->ev[4 * pi + 3], st);
    }
    if(rec) prof->used = pi + 1;
  }
  return 0;
}

// ------------------------------------------------------------------ solves
__global__ void gather_kernel(const cplx *b, const int *perm, cplx *y, int Np) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < Np) y[i] = b[perm[i]];
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsv_kernel(const T *__restrict__ A, size_t lda, int Np, cplx *x, int *flags,
                                                   int epoch, int *status) {
  constexpr int NB = kNB;
  const int nblk = Np / NB;
  const int k = LOWER ? blockIdx.x : nblk - 1 - blockIdx.x;
  __shared__ cplx xs[NB];
  __shared__ cplx part[4][NB];
  __shared__ T dg[NB * (NB + 1)];  // diagonal block, column-major, padded
  const int t = threadIdx.x, r = t & 63, cq = t >> 6;
  // the diagonal block does not depend on other workgroups: stage it first
  for(int idx = t; idx < NB * NB; idx += 256) {
    int c = idx / NB, rr = idx % NB;
    dg[c * (NB + 1) + rr] = A[(size_t)(k * NB + c) * lda + (size_t)k * NB + rr];
  }
  cplx acc = mk(0.0);
  const int ndep = LOWER ? k : nblk - 1 - k;
  for(int ii = 0; ii < ndep; ++ii) {
    const int b = LOWER ? ii : nblk - 1 - ii;
    // issue this tile's loads before waiting for x_b
    const T *a = A + (size_t)(b * NB + cq * 16) * lda + (size_t)k * NB + r;
    T tv[16];
#pragma unroll
    for(int cc = 0; cc < 16; ++cc) tv[cc] = a[(size_t)cc * lda];
    if(t == 0) {
      unsigned spins = 0;
      while(__hip_atomic_load(&flags[b], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch &&
            __hip_atomic_load(&status[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if(++spins > (1u << 24)) {
          __hip_atomic_store(&status[1], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    // x_b was published with a release before its flag: read it with
    // agent-scope atomic loads instead of invalidating the caches
    if(t < NB) {
      const unsigned long long *src = reinterpret_cast<const unsigned long long *>(x + (size_t)b * NB + t);
      unsigned long long re = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      unsigned long long im = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      xs[t] = mk(__longlong_as_double((long long)re), __longlong_as_double((long long)im));
    }
    __syncthreads();
#pragma unroll
    for(int cc = 0; cc < 16; ++cc) {
      T v = tv[cc];
      cplx xv = xs[cq * 16 + cc];
      acc.x += (double)v.x * xv.x - (double)v.y * xv.y;
      acc.y += (double)v.x * xv.y + (double)v.y * xv.x;
    }
    __syncthreads();
  }
  part[cq][r] = acc;
  __syncthreads();
  if(t < NB) {
    cplx s = x[(size_t)k * NB + t] - (part[0][t] + part[1][t] + part[2][t] + part[3][t]);
    if(LOWER) {
      for(int i = 0; i < NB; ++i) {
        cplx yi;
        yi.x = __shfl(s.x, i, 64);
        yi.y = __shfl(s.y, i, 64);
        if(t > i) {
          T v = dg[i * (NB + 1) + t];
          s.x -= (double)v.x * yi.x - (double)v.y * yi.y;
          s.y -= (double)v.x * yi.y + (double)v.y * yi.x;
        }
      }
    } else {
      for(int i = NB - 1; i >= 0; --i) {
        if(t == i) {
          T v = dg[i * (NB + 1) + i];
          s = s / mk((double)v.x, (double)v.y);
        }
        cplx yi;
        yi.x = __shfl(s.x, i, 64);
        yi.y = __shfl(s.y, i, 64);
        if(t < i) {
          T v = dg[i * (NB + 1) + t];
          s.x -= (double)v.x * yi.x - (double)v.y * yi.y;
          s.y -= (double)v.x * yi.y + (double)v.y * yi.x;
        }
      }
    }
    x[(size_t)k * NB + t] = s;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if(t == 0) __hip_atomic_store(&flags[k], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
void lu_solve(const T *A, int Np, size_t lda, LuWork &w, const cplx *b, cplx *x, hipStream_t st) {
  hipLaunchKernelGGL(gather_kernel, dim3((Np + 255) / 256), dim3(256), 0, st, b, w.perm, x, Np);
  int nblk = Np / kNB;
  int e1 = ++w.epoch;
  hipLaunchKernelGGL((trsv_kernel<T, true>), dim3(nblk), dim3(256), 0, st, A, lda, Np, x, w.flags, e1, w.status);
  int e2 = ++w.epoch;
  hipLaunchKernelGGL((trsv_kernel<T, false>), dim3(nblk), dim3(256), 0, st, A, lda, Np, x, w.flags, e2,
                     w.status);
}

template int lu_factor<double2>(double2 *, int, size_t, LuWork &, hipStream_t, LuProf *);
template int lu_factor<float2>(float2 *, int, size_t, LuWork &, hipStream_t, LuProf *);
template void lu_solve<double2>(const double2 *, int, size_t, LuWork &, const cplx *, cplx *, hipStream_t);
template void lu_solve<float2>(const float2 *, int, size_t, LuWork &, const cplx *, cplx *, hipStream_t);

// ------------------------------------------------------------------ gemv
// y = A x for the FP64 matrix (residuals and GMRES), two-stage column split.
constexpr int kGemvSplit = 32;
__global__ __launch_bounds__(256) void gemv_part_kernel(const cplx *__restrict__ A, size_t lda, int N,
                                                        const cplx *__restrict__ x, cplx *part, int chunk) {
  __shared__ cplx xs[512];
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * chunk, c1 = min(N, c0 + chunk);
  cplx acc = mk(0.0);
  for(int cb = c0; cb < c1; cb += 512) {
    int ce = min(c1, cb + 512);
    __syncthreads();
    for(int i = threadIdx.x; i < ce - cb; i += 256) xs[i] = x[cb + i];
    __syncthreads();
    if(r < N) {
      const cplx *a = A + (size_t)cb * lda + r;
      int n = ce - cb;
#pragma unroll 8
      for(int c = 0; c < n; ++c) {
        cplx v = a[(size_t)c * lda];
        acc.x += v.x * xs[c].x - v.y * xs[c].y;
        acc.y += v.x * xs[c].y + v.y * xs[c].x;
      }
    }
  }
  if(r < N) part[(size_t)blockIdx.y * N + r] = acc;
}
__global__ void gemv_reduce_kernel(const cplx *part, int ns, int N, cplx *y) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= N) return;
  cplx s = mk(0.0);
  for(int k = 0; k < ns; ++k) s = s + part[(size_t)k * N + r];
  y[r] = s;
}
void gemv64(const cplx *A, size_t lda, int N, const cplx *x, cplx *y, cplx *part, hipStream_t st) {
  int chunk = (N + kGemvSplit - 1) / kGemvSplit;
  hipLaunchKernelGGL(gemv_part_kernel, dim3((N + 255) / 256, kGemvSplit), dim3(256), 0, st, A, lda, N, x, part,
                     chunk);
  hipLaunchKernelGGL(gemv_reduce_kernel, dim3((N + 255) / 256), dim3(256), 0, st, part, kGemvSplit, N, y);
}

}  // namespace omx
