// omx_lu.hip -- blocked complex LU for CDNA4.
//
// Right-looking blocked factorisation P A = L U of a column-major Np x Np
// complex matrix (Np a multiple of the block size NB), in complex FP64
// (double2) or complex FP32 (float2).  Two pivot searches:
//   block (default)  diag_lu: partial pivoting among the rows of the panel's
//                  diagonal block, in one workgroup, which also inverts L11
//                  and U11; the solver recovers full accuracy by iterative
//                  refinement and refactors with the tournament if that fails
//   partial        tournament pivoting (TSLU): register-resident partial-
//                  pivoting LUs of 128-row chunks nominate candidate rows, a
//                  binary tree of merges picks the NB pivot rows (tslu_leaf,
//                  tslu_merge), tslu_finish turns them into a gather map and
//                  panel_inv writes the block's factors and inverses
// Shared by both:
//   laswp_map      the panel's row interchanges on the other columns, as one
//                  parallel gather/scatter
//   blk_mul        the two panel triangular solves as in-place MFMA block
//                  products: U12 = L11^{-1} A12, L21 = A21 U11^{-1}
//   gemm_tile      A22 -= L21 U12 as complex products on the matrix cores:
//                  v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32
//                  (FP32), four real MFMAs per complex tile product, planar
//                  re/im operands staged through LDS; panels are taken in
//                  pairs so that the bulk update runs with K = 2 NB
//   gemm_split     the FP32 bulk update (K = 2 NB) on the bf16 matrix cores:
//                  operands split in three bf16 terms, six products per
//                  real product, FP32 accumulation (FP32-accurate)
//   trsv_kernel    sync-free blocked triangular solves: workgroup k owns
//                  block row k, consumes earlier blocks as they are
//                  published (tagged words) and applies its stored
//                  diagonal-block inverse
// Look-ahead: the pivoting of panel k+1 runs on a second stream while the
// bulk of panel k's trailing update runs on the first.
// This replaces the ScaLAPACK pzgesv path (scalapack/LinearSystemSolver.hpp:84-91)
// and Eigen's ColPivHouseholderQR (PreconditionedMatrixSolver.h:58, 75).
#include <climits>
#include <type_traits>
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
// Reciprocal: the hardware estimate refined by Newton's iteration
// r <- r + r (1 - d r), which doubles the number of correct bits per step:
// one step after the single-precision estimate (within 1 ulp), two after the
// double-precision one (about 23 correct bits).
__device__ inline float rcp_nr(float d) {
  const float r = __builtin_amdgcn_rcpf(d);
  return __builtin_fmaf(r, __builtin_fmaf(-d, r, 1.0f), r);
}
__device__ inline double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}
__device__ inline float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename T> __device__ inline T crecip(T a) {
  // 1 / a by Smith's scaling: with s the smaller and b the larger component
  // in magnitude, q = s / b, 1 / a = (1, -q) / (b + s q) for |re| >= |im| and
  // (q, -1) / (b + s q) otherwise, so no intermediate overflows
  using R = decltype(a.x);
  const bool re = fabs(a.x) >= fabs(a.y);
  const R b = re ? a.x : a.y, sm = re ? a.y : a.x;
  const R q = sm * rcp_nr(b), id = rcp_nr(fmx(sm, q, b));
  T r;
  r.x = re ? id : q * id;
  r.y = re ? -q * id : -id;
  return r;
}
// v - a b with fused multiply-adds
template <typename R> __device__ inline Cx<R> cfms(Cx<R> v, Cx<R> a, Cx<R> b) {
  Cx<R> r;
  r.x = fmx(-a.x, b.x, fmx(a.y, b.y, v.x));
  r.y = fmx(-a.x, b.y, fmx(-a.y, b.x, v.y));
  return r;
}

// Componentwise select (a ternary on the aggregate would go through memory).
template <typename R> __device__ inline Cx<R> csel(bool p, Cx<R> a, Cx<R> b) {
  Cx<R> r;
  r.x = p ? a.x : b.x;
  r.y = p ? a.y : b.y;
  return r;
}

// Value of lane l (wave-uniform) of a complex register pair.
__device__ inline double rdl1(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ inline float rdl1(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
template <typename R> __device__ inline Cx<R> rdl(Cx<R> v, int l) { return Cx<R>{rdl1(v.x, l), rdl1(v.y, l)}; }

// Loads and stores with the nontemporal hint, for data a kernel streams once
// (a product's matrix, a bulk update's C tile): they stay out of the caches'
// replacement, which keeps the reused operands there.  In the FP64 product
// this raised 6.0 to 6.8 TB/s (N = 10240) and 5.6 to 6.0 TB/s (64 systems of
// N = 3840).
__device__ inline cplx ld_stream(const cplx *p) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(p));
  return cplx{v.x, v.y};
}
__device__ inline float2 ld_stream(const float2 *p) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 v = __builtin_nontemporal_load(reinterpret_cast<const f2 *>(p));
  float2 r;
  r.x = v.x;
  r.y = v.y;
  return r;
}
__device__ inline void st_stream(float2 *p, float2 c) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 v = {c.x, c.y};
  __builtin_nontemporal_store(v, reinterpret_cast<f2 *>(p));
}

__global__ void iota_kernel(int *p, int n, int sp) {
  p += (size_t)blockIdx.z * sp;
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
__global__ __launch_bounds__(256) void panel_inv(T *A, size_t lda, int j0, const T *wlu, T *dinv, int *status) {
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
    dinv[idx] = X[r * 64 + c];
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
    dinv[4096 + idx] = X[r * 64 + c];
  }
}

// Maximum of an unsigned key over the 64 lanes of a wave, returned to every
// lane: row shifts build each 16-lane row's maximum in its lane 15, two row
// broadcasts combine the rows into lane 63.
__device__ inline unsigned wave_max_u32(unsigned x) {
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true));  // row_shr:1
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true));  // row_shr:2
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true));  // row_shr:4
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true));  // row_shr:8
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x142, 0xA, 0xF, false));  // row_bcast:15
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x143, 0xC, 0xF, false));  // row_bcast:31
  return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}

// Diagonal-block pivoting: LU with partial pivoting restricted to the 64 x 64
// diagonal block of the panel (rows are chosen among the block's rows; the
// interchanges are applied to whole rows), then the inverses of L11 and U11.
// One workgroup of 8 waves.  Elimination: each wave holds one panel of 8
// columns of every row in registers, lane = row; rows are not moved (the
// pivot order is recorded instead).  The panel's owner runs its 8 pivot steps
// alone (pivot by a DPP maximum of a 32-bit key: max(|re|, |im|) as a float
// with the row in the low bits) and publishes the multipliers and pivot rows;
// after one barrier per panel the waves of later panels apply its 8 rank-1
// updates, each with the pivot row read from its own lane b.  Inversion: waves 0-3 build L11^{-1}, waves 4-7 U11^{-1}
// (lane = row, 16 columns per wave, row k broadcast by lane reads).  Writes
// the factored block into A, the inverses into dinv ([L11^{-1} | U11^{-1}],
// column-major), the block's row map (mdst <- msrc) and the permutation.
template <typename T>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(1, 2))) void diag_lu(T *A, size_t lda, int j0, T *dinv, int *mdst, int *msrc, int *mcnt,
                                               int *perm, int *status, BatchStrides bs) {
  using C = CxOf<T>;
  {
    const int z = blockIdx.z;
    A += z * bs.a;
    dinv += z * bs.inv;
    mdst += (size_t)z * bs.map;
    msrc += (size_t)z * bs.map;
    mcnt += (size_t)z * bs.cnt;
    perm += (size_t)z * bs.perm;
    status += (size_t)z * bs.status;
  }
  __shared__ int pb[2][8];  // a panel's pivot rows
  __shared__ int prow[64];
  __shared__ C F[64 * 65];  // factored block in pivoted order, column-major [c * 65 + r]
  // A panel's multipliers, one column per pivot step (0 on rows already
  // chosen), in two alternating buffers.  They live in F's storage: F is
  // written only after the last panel's barrier, when no wave reads them any
  // more, and the kernel's LDS stays small enough to start beside the bulk
  // update's workgroups.
  C(*lm)[8][64] = reinterpret_cast<C(*)[8][64]>(F);
  const int t = threadIdx.x, lane = t & 63, q = t >> 6;
  const C zero{}, one{1, 0};
  {
    // Wave q holds the columns of panel pq (8 columns); panels are taken in
    // order 0..7, and the two waves of a SIMD (q and q + 4) hold consecutive
    // panels, so every second panel's owner has its SIMD to itself.
    const int pq = 2 * (q & 3) + (q >> 2);
    C v[8];
#pragma unroll
    for(int j = 0; j < 8; ++j) {
      T a = A[(size_t)(j0 + 8 * pq + j) * lda + j0 + lane];
      v[j] = C{a.x, a.y};
    }
    bool used = false;
    int mypos = 0;
#pragma unroll 1
    for(int pc = 0; pc < 8; ++pc) {
      const int par = pc & 1;
      if(pq == pc) {
        // the owner's 8 pivot steps, without barriers: the pivot is the
        // largest max(|re|, |im|) among the rows not chosen yet (a DPP maximum
        // of a 32-bit key with the row in the low bits); the multipliers
        // replace the column below the pivot and update the panel's later
        // columns, and are published for the other waves
#pragma unroll
        for(int jj = 0; jj < 8; ++jj) {
          const int c = 8 * pc + jj;
          const float mx = (float)fmax(fabs((double)v[jj].x), fabs((double)v[jj].y));
          const unsigned key = used ? 0u : ((__float_as_uint(mx) & ~0x7Fu) | (unsigned)(127 - lane));
          const int b = __builtin_amdgcn_readfirstlane(127 - (int)(wave_max_u32(key) & 127));
          const C pv = rdl(v[jj], b);
          const bool sing = !(mag2(pv) > 0.0);
          if(sing && lane == b) atomicCAS(&status[0], 0, j0 + c + 1);
          const C ip = csel(sing, zero, crecip(pv));
          if(lane == b) {
            used = true;
            mypos = c;
          }
          const C l = csel(used, zero, cm(v[jj], ip));
          v[jj] = csel(used, v[jj], l);
          lm[par][jj][lane] = l;
          if(lane == 0) {
            pb[par][jj] = b;
            prow[c] = b;
          }
#pragma unroll
          for(int j = jj + 1; j < 8; ++j) v[j] = cfms(v[j], l, rdl(v[j], b));
        }
      }
      // one barrier per panel; the buffers alternate, so the next owner's
      // writes cannot reach a panel that a slower wave still reads
      __syncthreads();
      if(pq != pc) {
        int bb[8];
#pragma unroll
        for(int jj = 0; jj < 8; ++jj) {
          bb[jj] = __builtin_amdgcn_readfirstlane(pb[par][jj]);
          if(lane == bb[jj]) {
            used = true;
            mypos = 8 * pc + jj;
          }
        }
        if(pq > pc) {
          // the panel's 8 rank-1 updates, in the order of its pivot steps
#pragma unroll
          for(int jj = 0; jj < 8; ++jj) {
            const C l = lm[par][jj][lane];
#pragma unroll
            for(int j = 0; j < 8; ++j) v[j] = cfms(v[j], l, rdl(v[j], bb[jj]));
          }
        }
      }
    }
    // factored block in pivoted order
#pragma unroll
    for(int j = 0; j < 8; ++j) {
      const int col = 8 * pq + j;
      F[col * 65 + mypos] = v[j];
      T o;
      o.x = v[j].x;
      o.y = v[j].y;
      A[(size_t)(j0 + col) * lda + j0 + mypos] = o;
    }
  }
  __syncthreads();
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
  // Inversions: lane = row i; waves 0-3 hold columns 16 w .. 16 w + 15 of
  // L11^{-1}, waves 4-7 those of U11^{-1}; row k is broadcast within the wave
  // by lane reads, so the waves run without barriers.
  const int i = lane, w = q & 3;
  C x[16];
#pragma unroll
  for(int m = 0; m < 16; ++m) x[m] = csel(16 * w + m == i, one, zero);
  if(q < 4) {
    // L11^{-1} (unit lower): for k: row i -= L[i][k] row k, i > k.  Column
    // 16 w + m starts as e_{16 w + m} and its rows above 16 w + m stay zero,
    // so the steps k < 16 w leave this wave's columns as they are: skipped
    C fn = F[(16 * w) * 65 + i];  // column k of L, read one step ahead
#pragma unroll 1
    for(int k = 16 * w; k < 63; ++k) {
      const int ks = __builtin_amdgcn_readfirstlane(k);
      const C lik = csel(i > ks, fn, zero);
      fn = F[(ks + 1) * 65 + i];
#pragma unroll
      for(int m = 0; m < 16; ++m) x[m] = cfms(x[m], lik, rdl(x[m], ks));
    }
  } else {
    // U11^{-1} = (D^{-1} U11)^{-1} D^{-1} with D = diag(U11): a unit-upper
    // sweep (row i -= (U[i][k] / U[i][i]) row k, i < k, k = 63..0), then
    // column j scaled by 1 / U[j][j].  Rows below 16 w + 15 of this wave's
    // columns stay zero, so the sweep starts at k = 16 w + 15
    const C di = F[i * 65 + i];
    const C rinv = csel(mag2(di) > 0.0, crecip(di), zero);
    C fn = F[(16 * w + 15) * 65 + i];  // column k of U, read one step ahead
#pragma unroll 1
    for(int k = 16 * w + 15; k > 0; --k) {
      const int ks = __builtin_amdgcn_readfirstlane(k);
      const C uik = csel(i < ks, cm(fn, rinv), zero);
      fn = F[(ks - 1) * 65 + i];
#pragma unroll
      for(int m = 0; m < 16; ++m) x[m] = cfms(x[m], uik, rdl(x[m], ks));
    }
#pragma unroll
    for(int m = 0; m < 16; ++m) x[m] = cm(x[m], rdl(rinv, 16 * w + m));
  }
  T *out = dinv + (q < 4 ? 0 : 4096);
#pragma unroll
  for(int m = 0; m < 16; ++m) {
    T o;
    o.x = x[m].x;
    o.y = x[m].y;
    out[(16 * w + m) * 64 + i] = o;
  }
}

// In-place block product with a 64 x 64 factor M (column-major):
//   LEFT:  A[j0 + i, off + n] = sum_k M(i, k) A[j0 + k, off + n]
//   RIGHT: A[off + m, j0 + n] = sum_k A[off + m, j0 + k] M(k, n)
// with off = first + 64 blockIdx.x.  K is staged through LDS in chunks of
// 16 (a small LDS footprint lets these workgroups start beside the bulk
// update's); the block is overwritten only after all of it has been read.
template <typename T, bool LEFT>
__global__ __launch_bounds__(256) void blk_mul(T *A, size_t lda, int j0, int first, const T *M, BatchStrides bs) {
  A += blockIdx.z * bs.a;
  M += blockIdx.z * bs.inv;
  using R = typename CT<T>::R;
  using V4 = typename CT<T>::V4;
  constexpr int BK = 16, PAD = 16;
  __shared__ R As[2][BK][64 + PAD];  // [plane][k][m]
  __shared__ R Bs[2][BK][64 + PAD];  // [plane][k][n]
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, wm = wv & 1, wn = wv >> 1;
  const int off = first + blockIdx.x * 64;
  V4 cr[2][2], ci[2][2];
#pragma unroll
  for(int a = 0; a < 2; ++a)
#pragma unroll
    for(int b = 0; b < 2; ++b) {
      cr[a][b] = V4{0, 0, 0, 0};
      ci[a][b] = V4{0, 0, 0, 0};
    }
  // both operands in registers first (one round of load latency), then
  // staged through LDS chunk by chunk
  R ar_[4][4], ai_[4][4], br_[4][4], bi_[4][4];
#pragma unroll
  for(int ch = 0; ch < 4; ++ch)
#pragma unroll
    for(int it = 0; it < 4; ++it) {
      const int idx = t + 256 * it, kc = ch * BK;
      // left operand (m = idx % 64, k = idx / 64), right operand (k = idx % 16, n = idx / 16)
      const int m = idx & 63, ka = idx >> 6, kb = idx & 15, n = idx >> 4;
      T av, bv;
      if(LEFT) {
        av = M[(kc + ka) * 64 + m];                              // M(m, k)
        bv = A[(size_t)(off + n) * lda + j0 + kc + kb];          // block(k, n)
      } else {
        av = A[(size_t)(j0 + kc + ka) * lda + off + m];          // block(m, k)
        bv = M[n * 64 + kc + kb];                                // M(k, n)
      }
      ar_[ch][it] = av.x;
      ai_[ch][it] = av.y;
      br_[ch][it] = bv.x;
      bi_[ch][it] = bv.y;
    }
#pragma unroll
  for(int ch = 0; ch < 4; ++ch) {
    __syncthreads();
#pragma unroll
    for(int it = 0; it < 4; ++it) {
      const int idx = t + 256 * it;
      const int m = idx & 63, ka = idx >> 6, kb = idx & 15, n = idx >> 4;
      As[0][ka][m] = ar_[ch][it];
      As[1][ka][m] = ai_[ch][it];
      Bs[0][kb][n] = br_[ch][it];
      Bs[1][kb][n] = bi_[ch][it];
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
          ci[a][b] = CT<T>::mfma(ar[a], bi[b], ci[a][b]);
          cr[a][b] = CT<T>::mfma(nai[a], bi[b], cr[a][b]);
          ci[a][b] = CT<T>::mfma(ai[a], br[b], ci[a][b]);
        }
    }
  }
  __syncthreads();  // every workgroup thread has read the block before it is overwritten
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
// Applies a gather map (rows mdst[i] <- msrc[i], i < *mcnt <= 128) to the
// columns c0 .. cend-1, leaving out skip0 .. skip0 + NB - 1 when skip0 >= 0:
// each workgroup handles 8 columns, reads all moved entries into registers,
// then writes them back.
template <typename T>
__global__ __launch_bounds__(256) void laswp_map(T *A, size_t lda, int c0, int cend, const int *__restrict__ mdst,
                                                 const int *__restrict__ msrc, const int *__restrict__ mcnt,
                                                 int skip0, BatchStrides bs) {
  A += blockIdx.z * bs.a;
  mdst += (size_t)blockIdx.z * bs.map;
  msrc += (size_t)blockIdx.z * bs.map;
  mcnt += (size_t)blockIdx.z * bs.cnt;
  const int m = *mcnt;
  const int cl = threadIdx.x / 32, li = threadIdx.x % 32;
  int col = c0 + blockIdx.x * 8 + cl;
  if(skip0 >= 0 && col >= skip0) col += kNB;
  T v[4];
  int d[4];
  if(col < cend) {
    T *a = A + (size_t)col * lda;
#pragma unroll
    for(int k = 0; k < 4; ++k) {
      int i = li + 32 * k;
      d[k] = i < m ? mdst[i] : -1;
      if(i < m) v[k] = a[msrc[i]];
    }
  }
  __syncthreads();
  if(col < cend) {
    T *a = A + (size_t)col * lda;
#pragma unroll
    for(int k = 0; k < 4; ++k)
      if(d[k] >= 0) a[d[k]] = v[k];
  }
}
template <typename T>
void laswp(T *A, size_t lda, int c0, int cend, const int *mdst, const int *msrc, const int *mcnt, int skip0,
           hipStream_t st, int nb = 1, BatchStrides bs = BatchStrides{}) {
  const int n = cend - c0 - (skip0 >= 0 ? kNB : 0);
  if(n > 0)
    hipLaunchKernelGGL((laswp_map<T>), dim3((n + 7) / 8, 1, nb), dim3(256), 0, st, A, lda, c0, cend, mdst, msrc, mcnt,
                       skip0, bs);
}

// Deferred interchanges of block pivoting: panel k = blockIdx.y (>= 1)
// applies its map (NB entries at k NB) to the columns left of it.
template <typename T>
__global__ __launch_bounds__(256) void laswp_left(T *A, size_t lda, const int *__restrict__ mdst,
                                                  const int *__restrict__ msrc, const int *__restrict__ mcnt,
                                                  BatchStrides bs) {
  A += blockIdx.z * bs.a;
  mdst += (size_t)blockIdx.z * bs.map;
  msrc += (size_t)blockIdx.z * bs.map;
  mcnt += (size_t)blockIdx.z * bs.cnt;
  const int k = blockIdx.y + 1;
  const int cl = threadIdx.x / 32, li = threadIdx.x % 32;
  const int col = blockIdx.x * 8 + cl;
  if(blockIdx.x * 8 >= k * kNB) return;  // whole workgroup past the panel
  const int m = mcnt[k];
  const int *md = mdst + k * kNB, *ms = msrc + k * kNB;
  T v[2];
  int d[2];
  if(col < k * kNB) {
    const T *a = A + (size_t)col * lda;
#pragma unroll
    for(int q = 0; q < 2; ++q) {
      const int i = li + 32 * q;
      d[q] = i < m ? md[i] : -1;
      if(i < m) v[q] = a[ms[i]];
    }
  }
  __syncthreads();
  if(col < k * kNB) {
    T *a = A + (size_t)col * lda;
#pragma unroll
    for(int q = 0; q < 2; ++q)
      if(d[q] >= 0) a[d[q]] = v[q];
  }
}

// ------------------------------------------------------------------ gemm
// Trailing update:
// C[row0 + m, col0 + n] -= sum_{k < KD} A[row0 + m, k0 + k] A[k0 + k, col0 + n]
// for m < M, n < N.  A TM x TN tile per workgroup of WGM x WGN waves, each
// wave (TM/WGM) x (TN/WGN) as 16 x 16 MFMA blocks with separate real and
// imaginary accumulators; K is staged through LDS in chunks of BK (planar
// re/im, rows padded to 16 banks) and the next chunk is fetched into
// registers while the current one is multiplied.
template <typename T, int TM, int TN, int KD, int WGM, int WGN, int BK>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_tile(T *A, size_t lda, int row0, int col0, int k0, int M,
                                                           int N, size_t sA) {
  A += blockIdx.z * sA;
  using R = typename CT<T>::R;
  using V4 = typename CT<T>::V4;
  constexpr int NT = 64 * WGM * WGN, PAD = 16;
  constexpr int WM = TM / WGM, WN = TN / WGN, FM = WM / 16, FN = WN / 16;
  constexpr int LA = TM * BK / NT, LB = TN * BK / NT;
  static_assert(LA * NT == TM * BK && LB * NT == TN * BK, "tile not divisible");
  __shared__ R As[2][BK][TM + PAD];
  __shared__ R Bs[2][BK][TN + PAD];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, wm = wv % WGM, wn = wv / WGM;
  const int bm = blockIdx.x * TM, bn = blockIdx.y * TN;
  V4 cr[FM][FN], ci[FM][FN];
#pragma unroll
  for(int a = 0; a < FM; ++a)
#pragma unroll
    for(int b = 0; b < FN; ++b) {
      cr[a][b] = V4{0, 0, 0, 0};
      ci[a][b] = V4{0, 0, 0, 0};
    }
  R ar_[LA], ai_[LA], br_[LB], bi_[LB];
  auto gload = [&](int kc) {
#pragma unroll
    for(int i = 0; i < LA; ++i) {
      const int idx = t + NT * i, r = idx % TM, kk = idx / TM;
      ar_[i] = 0;
      ai_[i] = 0;
      if(bm + r < M) {
        const T v = A[(size_t)(k0 + kc + kk) * lda + row0 + bm + r];
        ar_[i] = v.x;
        ai_[i] = v.y;
      }
    }
#pragma unroll
    for(int i = 0; i < LB; ++i) {
      const int idx = t + NT * i, kb = idx % BK, c = idx / BK;
      br_[i] = 0;
      bi_[i] = 0;
      if(bn + c < N) {
        const T v = A[(size_t)(col0 + bn + c) * lda + k0 + kc + kb];
        br_[i] = v.x;
        bi_[i] = v.y;
      }
    }
  };
  gload(0);
#pragma unroll 1
  for(int kc = 0; kc < KD; kc += BK) {
    __syncthreads();  // the previous chunk has been read out of LDS
#pragma unroll
    for(int i = 0; i < LA; ++i) {
      const int idx = t + NT * i, r = idx % TM, kk = idx / TM;
      As[0][kk][r] = ar_[i];
      As[1][kk][r] = ai_[i];
    }
#pragma unroll
    for(int i = 0; i < LB; ++i) {
      const int idx = t + NT * i, kb = idx % BK, c = idx / BK;
      Bs[0][kb][c] = br_[i];
      Bs[1][kb][c] = bi_[i];
    }
    __syncthreads();
    if(kc + BK < KD) gload(kc + BK);
#pragma unroll
    for(int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4), li = lane & 15;
      R xr[FM], xi[FM], nxi[FM], yr[FN], yi[FN];
#pragma unroll
      for(int a = 0; a < FM; ++a) {
        xr[a] = As[0][kr][wm * WM + a * 16 + li];
        xi[a] = As[1][kr][wm * WM + a * 16 + li];
        nxi[a] = -xi[a];
      }
#pragma unroll
      for(int b = 0; b < FN; ++b) {
        yr[b] = Bs[0][kr][wn * WN + b * 16 + li];
        yi[b] = Bs[1][kr][wn * WN + b * 16 + li];
      }
#pragma unroll
      for(int a = 0; a < FM; ++a)
#pragma unroll
        for(int b = 0; b < FN; ++b) cr[a][b] = CT<T>::mfma(xr[a], yr[b], cr[a][b]);
#pragma unroll
      for(int a = 0; a < FM; ++a)
#pragma unroll
        for(int b = 0; b < FN; ++b) ci[a][b] = CT<T>::mfma(xr[a], yi[b], ci[a][b]);
#pragma unroll
      for(int a = 0; a < FM; ++a)
#pragma unroll
        for(int b = 0; b < FN; ++b) cr[a][b] = CT<T>::mfma(nxi[a], yi[b], cr[a][b]);
#pragma unroll
      for(int a = 0; a < FM; ++a)
#pragma unroll
        for(int b = 0; b < FN; ++b) ci[a][b] = CT<T>::mfma(xi[a], yr[b], ci[a][b]);
    }
  }
  // C -= the products: the loads of a group of EA block rows are all issued
  // before its first store (M and N are multiples of 16, so each 16 x 16
  // block is wholly in or out of range); FP64 takes one block row at a time
  // to stay within the register budget of four waves per SIMD
  constexpr int EA = sizeof(R) == 8 ? 1 : FM;
#pragma unroll
  for(int a0 = 0; a0 < FM; a0 += EA) {
    R vr[EA][FN][4], vi[EA][FN][4];
#pragma unroll
    for(int a = 0; a < EA; ++a)
#pragma unroll
      for(int b = 0; b < FN; ++b) {
        const int n = bn + wn * WN + b * 16 + (lane & 15), mb = bm + wm * WM + (a0 + a) * 16;
        const T *cp = A + (size_t)(col0 + n) * lda + row0;
#pragma unroll
        for(int v = 0; v < 4; ++v) {
          vr[a][b][v] = 0;
          vi[a][b][v] = 0;
          if(n < N && mb < M) {
            const T c = cp[mb + CT<T>::row(lane, v)];
            vr[a][b][v] = c.x;
            vi[a][b][v] = c.y;
          }
        }
      }
#pragma unroll
    for(int a = 0; a < EA; ++a)
#pragma unroll
      for(int b = 0; b < FN; ++b) {
        const int n = bn + wn * WN + b * 16 + (lane & 15), mb = bm + wm * WM + (a0 + a) * 16;
        if(n >= N || mb >= M) continue;
        T *cp = A + (size_t)(col0 + n) * lda + row0;
#pragma unroll
        for(int v = 0; v < 4; ++v) {
          T c;
          c.x = vr[a][b][v] - cr[a0 + a][b][v];
          c.y = vi[a][b][v] - ci[a0 + a][b][v];
          cp[mb + CT<T>::row(lane, v)] = c;
        }
      }
  }
}

// Tile configuration of the trailing updates (fastest of a sweep at
// M = N = 10112, K = 128: 103 TFLOP/s complex FP32, 55 TFLOP/s complex FP64).
template <typename T> struct UpdTile {
  static constexpr int TM = 128, TN = 64, WGM = 4, WGN = 2, BK = 16;
};

// Tile of the thin updates (K = NB, one of M and N = NB) on the panel
// chain: small tiles for many workgroups (fastest of a sweep at
// 10112 x 64 x 64: 7.9 us FP32, 15.7 us FP64, against 18.2 and 29.5 with
// UpdTile).
struct ThinTile {
  static constexpr int TM = 32, TN = 32, WGM = 2, WGN = 2, BK = 16;
};

// C -= A B over an M x N block (M, N multiples of 64) with K = KD.
template <typename T, int KD, typename U = UpdTile<T>>
void trailing_update(T *A, size_t lda, int row0, int col0, int k0, int M, int N, hipStream_t st, int nb = 1,
                     size_t sA = 0) {
  if(M <= 0 || N <= 0) return;
  hipLaunchKernelGGL((gemm_tile<T, U::TM, U::TN, KD, U::WGM, U::WGN, U::BK>),
                     dim3((M + U::TM - 1) / U::TM, (N + U::TN - 1) / U::TN, nb), dim3(64 * U::WGM * U::WGN), 0, st, A,
                     lda, row0, col0, k0, M, N, sA);
}

// ------------------------------------------------------------------ split-bf16 bulk update
// The complex-FP32 bulk update C -= A B (K = KD) on the bf16 matrix cores:
// every FP32 operand x is split in three bf16 terms x = x1 + x2 + x3 (each
// the round-to-nearest bf16 of what the previous ones leave; x - x1 - x2 - x3
// is below 2^-24 |x|) and the six products x_i y_j with i + j <= 4 (1-based)
// are accumulated in FP32 by v_mfma_f32_32x32x16_bf16: the result has the
// accuracy of an FP32 product (the dropped terms are below 2^-24 |x y|) at
// six bf16 instructions per real product, against the FP32 MFMA's 1/16 of the
// bf16 rate.
//
// The complex product is a real one over k' = 2 k + part:
//   D[n][m] = sum_k' X[n][k'] Y[k'][m],   X[n][2k] = Re B(k,n), X[n][2k+1] = Im B(k,n)
//   Re: Y[2k][m] = Re A(m,k), Y[2k+1][m] = -Im A(m,k)
//   Im: Y[2k][m] = Im A(m,k), Y[2k+1][m] =  Re A(m,k)
// X is B's column in its natural (re, im) storage; both Y are formed from the
// staged A by a sign flip or a swap of the bf16 halves of each 32-bit word.
// D's lanes run along m, so the epilogue reads and writes C's columns in
// 128-byte runs.
//
// LDS holds the three split planes of a K chunk as 16-byte units (4 complex k
// of one row of A or one column of B); the global loads of chunk c + 2 are in
// flight while chunk c is multiplied, and C's tile is fetched during the last
// chunk.  Workgroups are numbered so that each XCD takes a contiguous range of
// tiles, walked in groups of GM tile rows (operands reused from its L2).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x16 mfma_bf16_32(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// f[0..7] -> three words of 8 bf16 each (x1, x2, x3 of every value)
__device__ inline void split_bf16x3(const float *f, u32x4 &h, u32x4 &m, u32x4 &l) {
#pragma unroll
  for(int d = 0; d < 4; ++d) {
    unsigned hh[2], mm[2], ll[2];
#pragma unroll
    for(int e = 0; e < 2; ++e) {
      const float x = f[2 * d + e];
      const __bf16 b1 = (__bf16)x;
      const float r1 = x - (float)b1;
      const __bf16 b2 = (__bf16)r1;
      const float r2 = r1 - (float)b2;
      const __bf16 b3 = (__bf16)r2;
      hh[e] = __builtin_bit_cast(unsigned short, b1);
      mm[e] = __builtin_bit_cast(unsigned short, b2);
      ll[e] = __builtin_bit_cast(unsigned short, b3);
    }
    h[d] = hh[0] | (hh[1] << 16);
    m[d] = mm[0] | (mm[1] << 16);
    l[d] = ll[0] | (ll[1] << 16);
  }
}

// DB: the split planes double-buffered in LDS, so that each wave splits and
// stores chunk c + 1 while its products of chunk c run (one barrier per
// chunk); the products and their order are those of the single buffer.
template <int TM, int TN, int KD, int WGM, int WGN, int BK, int GM, bool DB = false, bool NTC = false>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_split(float2 *A, size_t lda, int row0, int col0, int k0, int M,
                                                            int N, size_t sA) {
  A += blockIdx.z * sA;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WM = TM / WGM, WN = TN / WGN, FM = WM / 32, FN = WN / 32;
  constexpr int KG = BK / 4, NC = KD / BK, NBUF = DB ? 2 : 1;
  constexpr int UA = TM * KG / NT, UB = TN * KG / NT;
  static_assert(UA * NT == TM * KG && UB * NT == TN * KG, "tile not divisible");
  static_assert(FM >= 1 && FN >= 1 && BK % 8 == 0 && KD % BK == 0, "tile shape");
  __shared__ u32x4 Ys[NBUF][3][KG][TM];
  __shared__ u32x4 Xs[NBUF][3][KG][TN];
  const int gx = (M + TM - 1) / TM, gy = (N + TN - 1) / TN, T = gx * gy;
  int L = blockIdx.x;
  const int full = T / 8 * 8;
  if(L < full) L = (L % 8) * (full / 8) + L / 8;
  const int per = GM * gy, g0 = (L / per) * GM, gsz = min(gx - g0, GM), r = L % per;
  const int bm = (g0 + r % gsz) * TM, bn = (r / gsz) * TN;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, wm = wv % WGM, wn = wv / WGM;
  f32x16 accR[FM][FN], accI[FM][FN];
#pragma unroll
  for(int a = 0; a < FM; ++a)
#pragma unroll
    for(int b = 0; b < FN; ++b) {
      accR[a][b] = f32x16{};
      accI[a][b] = f32x16{};
    }
  float ga[2][UA][8], gb[2][UB][8];
  auto gload = [&](int kc, int sl) {
#pragma unroll
    for(int i = 0; i < UA; ++i) {
      const int u = t + NT * i, m = u % TM, kg = u / TM;
#pragma unroll
      for(int j = 0; j < 4; ++j) {
        ga[sl][i][2 * j] = 0;
        ga[sl][i][2 * j + 1] = 0;
        if(bm + m < M) {
          const float2 v = A[(size_t)(k0 + kc + 4 * kg + j) * lda + row0 + bm + m];
          ga[sl][i][2 * j] = v.x;
          ga[sl][i][2 * j + 1] = v.y;
        }
      }
    }
#pragma unroll
    for(int i = 0; i < UB; ++i) {
      const int u = t + NT * i, kg = u % KG, n = u / KG;
#pragma unroll
      for(int j = 0; j < 8; ++j) gb[sl][i][j] = 0;
      if(bn + n < N) {
        const float4 *p = reinterpret_cast<const float4 *>(A + (size_t)(col0 + bn + n) * lda + k0 + kc + 4 * kg);
        const float4 p0 = p[0], p1 = p[1];
        gb[sl][i][0] = p0.x;
        gb[sl][i][1] = p0.y;
        gb[sl][i][2] = p0.z;
        gb[sl][i][3] = p0.w;
        gb[sl][i][4] = p1.x;
        gb[sl][i][5] = p1.y;
        gb[sl][i][6] = p1.z;
        gb[sl][i][7] = p1.w;
      }
    }
  };
  // C's tile: lane l, register v of block (a, b) holds D[n][m] with
  // m = 32 a + l % 32, n = 32 b + 8 (v / 4) + 4 (l / 32) + v % 4; M and N are
  // multiples of 32, so each 32 x 32 block is wholly in or out of range
  float vr[FM][FN][16], vi[FM][FN][16];
  auto cload = [&]() {
#pragma unroll
    for(int a = 0; a < FM; ++a)
#pragma unroll
      for(int b = 0; b < FN; ++b) {
        const int mb = bm + wm * WM + 32 * a, nb = bn + wn * WN + 32 * b;
        const bool ok = mb < M && nb < N;
#pragma unroll
        for(int v = 0; v < 16; ++v) {
          vr[a][b][v] = 0;
          vi[a][b][v] = 0;
          if(ok) {
            const int n = nb + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4), m = mb + (lane & 31);
            const float2 *cp = A + (size_t)(col0 + n) * lda + row0 + m;
            const float2 c = NTC ? ld_stream(cp) : *cp;
            vr[a][b][v] = c.x;
            vi[a][b][v] = c.y;
          }
        }
      }
  };
  // chunk c's operands from register slot sl, split into LDS buffer buf
  auto split_store = [&](int sl, int buf) {
#pragma unroll
    for(int i = 0; i < UA; ++i) {
      const int u = t + NT * i, m = u % TM, kg = u / TM;
      u32x4 h, md, lo;
      split_bf16x3(ga[sl][i], h, md, lo);
      Ys[buf][0][kg][m] = h;
      Ys[buf][1][kg][m] = md;
      Ys[buf][2][kg][m] = lo;
    }
#pragma unroll
    for(int i = 0; i < UB; ++i) {
      const int u = t + NT * i, kg = u % KG, n = u / KG;
      u32x4 h, md, lo;
      split_bf16x3(gb[sl][i], h, md, lo);
      Xs[buf][0][kg][n] = h;
      Xs[buf][1][kg][n] = md;
      Xs[buf][2][kg][n] = lo;
    }
  };
  gload(0, 0);
  if(NC > 1) gload(BK, 1);
  if constexpr(DB) {
    split_store(0, 0);
    __syncthreads();
    if(NC > 2) gload(2 * BK, 0);
  }
#pragma unroll
  for(int c = 0; c < NC; ++c) {
    const int sl = c & 1, buf = DB ? (c & 1) : 0;
    if constexpr(DB) {
      // chunk c + 1 into the other buffer (read by every wave before the
      // barrier that closed the previous chunk), then its slot refilled
      if(c + 1 < NC) split_store((c + 1) & 1, buf ^ 1);
      if(c + 3 < NC) gload((c + 3) * BK, (c + 1) & 1);
    } else {
      __syncthreads();  // the previous chunk has been read out of LDS
      split_store(sl, 0);
      __syncthreads();
      if(c + 2 < NC) gload((c + 2) * BK, sl);
    }
    if(c + 1 == NC) cload();
#pragma unroll
    for(int ks = 0; ks < BK / 8; ++ks) {
      // lane l feeds rows/columns l % 32 at k' = 16 ks + 8 (l / 32) .. + 7
      const int kg = 2 * ks + (lane >> 5), li = lane & 31;
      u32x4 x[3][FN], yr[3][FM], yi[3][FM];
#pragma unroll
      for(int s = 0; s < 3; ++s) {
#pragma unroll
        for(int b = 0; b < FN; ++b) x[s][b] = Xs[buf][s][kg][wn * WN + 32 * b + li];
#pragma unroll
        for(int a = 0; a < FM; ++a) {
          const u32x4 y = Ys[buf][s][kg][wm * WM + 32 * a + li];
#pragma unroll
          for(int d = 0; d < 4; ++d) {
            yr[s][a][d] = y[d] ^ 0x80000000u;
            yi[s][a][d] = (y[d] >> 16) | (y[d] << 16);
          }
        }
      }
      // the small products first
      constexpr int SY[6] = {2, 1, 0, 1, 0, 0}, SX[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
      for(int p = 0; p < 6; ++p)
#pragma unroll
        for(int a = 0; a < FM; ++a)
#pragma unroll
          for(int b = 0; b < FN; ++b) {
            accR[a][b] = mfma_bf16_32(x[SX[p]][b], yr[SY[p]][a], accR[a][b]);
            accI[a][b] = mfma_bf16_32(x[SX[p]][b], yi[SY[p]][a], accI[a][b]);
          }
    }
    if constexpr(DB) __syncthreads();  // chunk c read out; chunk c + 1 stored
  }
#pragma unroll
  for(int a = 0; a < FM; ++a)
#pragma unroll
    for(int b = 0; b < FN; ++b) {
      const int mb = bm + wm * WM + 32 * a, nb = bn + wn * WN + 32 * b;
      if(mb >= M || nb >= N) continue;
#pragma unroll
      for(int v = 0; v < 16; ++v) {
        const int n = nb + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4), m = mb + (lane & 31);
        float2 c;
        c.x = vr[a][b][v] - accR[a][b][v];
        c.y = vi[a][b][v] - accI[a][b][v];
        if constexpr(NTC)
          st_stream(A + (size_t)(col0 + n) * lda + row0 + m, c);
        else
          A[(size_t)(col0 + n) * lda + row0 + m] = c;
      }
    }
}

// Tile of the split-bf16 update (fastest of a sweep at M = N = 10112,
// K = 128: 145 TFLOP/s complex, against 87 for gemm_tile<float2> on the same
// random data).
struct SplitTile {
  static constexpr int TM = 128, TN = 128, WGM = 4, WGN = 2, BK = 16, GM = 8;
  // double-buffered split planes: bit-identical, 149 against 125 TFLOP/s at
  // M = N = 10112 (one matrix), 156 against 154 for 64 systems of 3712
  static constexpr bool DB = true;
  // C's tile read and written with the nontemporal hint: bit-identical, 5510
  // against 5644 us for 64 systems of 3840 (2.4%), 1-5% at one matrix of 10240
  static constexpr bool NTC = true;
};

// The bulk update with K = 2 NB: split-bf16 for FP32 factors when `split`,
// gemm_tile otherwise.  M and N are multiples of 64.
template <typename T>
void bulk_update(T *A, size_t lda, int row0, int col0, int k0, int M, int N, hipStream_t st, int nb, size_t sA,
                 bool split) {
  if(M <= 0 || N <= 0) return;
  if(N <= 2 * kNB) {
    // the next pair's columns (on the panel chain's critical path): small
    // tiles over many workgroups finish sooner than a few large split tiles
    trailing_update<T, 2 * kNB, ThinTile>(A, lda, row0, col0, k0, M, N, st, nb, sA);
    return;
  }
  if constexpr(std::is_same<T, float2>::value) {
    if(split) {
      using U = SplitTile;
      const int tiles = ((M + U::TM - 1) / U::TM) * ((N + U::TN - 1) / U::TN);
      hipLaunchKernelGGL((gemm_split<U::TM, U::TN, 2 * kNB, U::WGM, U::WGN, U::BK, U::GM, U::DB, U::NTC>), dim3(tiles, 1, nb),
                         dim3(64 * U::WGM * U::WGN), 0, st, A, lda, row0, col0, k0, M, N, sA);
      return;
    }
  }
  trailing_update<T, 2 * kNB>(A, lda, row0, col0, k0, M, N, st, nb, sA);
}

// ------------------------------------------------------------------ driver
// Panels of NB columns are taken in pairs (p0 = J, p1 = J + NB) so that the
// bulk of the trailing update runs with K = 2 NB.
//
// Block pivoting (interchanges stay inside a panel's 64 rows), two streams:
// sb (panel work) runs pair s+1 while st (bulk updates) finishes pair s.
//   sb  diag_lu p0; swaps on p1's columns; L21(p0); U12(p0) on p1's columns;
//       p1's columns -= L21(p0) U12(p0); diag_lu p1; L21(p1)
//   sb  after st's bulk update of pair s-1: swaps of p0 on the far columns,
//       U12(p0), block row p1 -= L(p1, p0) U12(p0), swaps of p1, U12(p1)
//   st  A22 -= [L21(p0) L21(p1)] [U12(p0); U12(p1)]: the next pair's
//       columns first (sb starts on them), then the rest
// Interchanges on the columns left of each panel are applied once at the end
// (panels touch disjoint rows, so their order does not matter).
//
// Tournament (partial) pivoting: one stream for the panel sequence, sb for
// the tournaments; every interchange is applied at once to all columns.
// Progress mark of a factorisation (see LuWork::ev_mark).
inline void mark(LuWork &w, hipStream_t st) {
  (void)hipEventRecord(w.ev_mark, st);
  if(w.on_mark) w.on_mark(w.mark_arg);
}

template <typename T>
int lu_factor(T *A, int Np, size_t lda, LuWork &w, hipStream_t st, LuProf *prof) {
  if(Np % kNB) return -1;
  const bool block = w.pivot == kPivotBlock;
  const int nb = w.nb;
  const BatchStrides bs = w.bs;
  if(!block && nb != 1) return -3;  // the tournament factors one system at a time
  hipStream_t sb = w.aux;
  T *wlu = reinterpret_cast<T *>(w.wlu);
  T *dinv = reinterpret_cast<T *>(w.dinv);
  auto dk = [&](int j0) { return dinv + (size_t)(j0 / kNB) * 2 * kNB * kNB; };
  hipLaunchKernelGGL(iota_kernel, dim3((Np + 255) / 256, 1, nb), dim3(256), 0, st, w.perm, Np, bs.perm);
  if(prof) {
    prof->used = 0;
    prof->gemm_flop = 0.0;
    prof->gemm_launches = 0;
  }
  if(block) {
    auto map = [&](int j0, int **d, int **s_, int **c) {
      const int k = j0 / kNB;
      *d = w.mdst + k * kNB;
      *s_ = w.msrc + k * kNB;
      *c = w.mcnt + k;
    };
    (void)hipEventRecord(w.ev_next, st);
    (void)hipEventRecord(w.ev_bulk, st);
    for(int J = 0, pi = 0; J < Np; J += 2 * kNB, ++pi) {
      const int p0 = J, p1 = J + kNB, far = J + 2 * kNB;
      const bool two = p1 < Np;
      const int rest0 = Np - p0 - kNB, restf = Np - far;
      const bool rec = prof && pi < prof->cap;
      hipEvent_t *pe = rec ? prof->ev + 4 * pi : nullptr;
      int *d0, *s0, *c0, *d1, *s1, *c1;
      map(p0, &d0, &s0, &c0);
      map(p1, &d1, &s1, &c1);
      (void)hipStreamWaitEvent(sb, w.ev_next, 0);
      if(rec) (void)hipEventRecord(pe[0], sb);
      hipLaunchKernelGGL((diag_lu<T>), dim3(1, 1, nb), dim3(512), 0, sb, A, lda, p0, dk(p0), d0, s0, c0, w.perm,
                         w.status, bs);
      if(rest0 > 0)
        hipLaunchKernelGGL((blk_mul<T, false>), dim3(rest0 / 64, 1, nb), dim3(256), 0, sb, A, lda, p0, p1,
                           dk(p0) + kNB * kNB, bs);
      if(two) {
        laswp<T>(A, lda, p1, p1 + kNB, d0, s0, c0, -1, sb, nb, bs);
        hipLaunchKernelGGL((blk_mul<T, true>), dim3(1, 1, nb), dim3(256), 0, sb, A, lda, p0, p1, dk(p0), bs);
        trailing_update<T, kNB, ThinTile>(A, lda, p1, p1, p0, Np - p1, kNB, sb, nb, bs.a);
        hipLaunchKernelGGL((diag_lu<T>), dim3(1, 1, nb), dim3(512), 0, sb, A, lda, p1, dk(p1), d1, s1, c1, w.perm,
                           w.status, bs);
        if(restf > 0)
          hipLaunchKernelGGL((blk_mul<T, false>), dim3(restf / 64, 1, nb), dim3(256), 0, sb, A, lda, p1, far,
                             dk(p1) + kNB * kNB, bs);
      }
      if(rec) (void)hipEventRecord(pe[1], sb);
      if(two && restf > 0) {
        (void)hipStreamWaitEvent(sb, w.ev_bulk, 0);
        laswp<T>(A, lda, far, Np, d0, s0, c0, -1, sb, nb, bs);
        hipLaunchKernelGGL((blk_mul<T, true>), dim3(restf / 64, 1, nb), dim3(256), 0, sb, A, lda, p0, far, dk(p0), bs);
        trailing_update<T, kNB, ThinTile>(A, lda, p1, far, p0, kNB, restf, sb, nb, bs.a);
        laswp<T>(A, lda, far, Np, d1, s1, c1, -1, sb, nb, bs);
        hipLaunchKernelGGL((blk_mul<T, true>), dim3(restf / 64, 1, nb), dim3(256), 0, sb, A, lda, p1, far, dk(p1), bs);
      }
      (void)hipEventRecord(w.ev_pan, sb);
      (void)hipStreamWaitEvent(st, w.ev_pan, 0);
      if(rec) (void)hipEventRecord(pe[2], st);
      if(two && restf > 0) {
        const int nn = restf < 2 * kNB ? restf : 2 * kNB;
        bulk_update<T>(A, lda, far, far, p0, restf, nn, st, nb, bs.a, w.split);
        (void)hipEventRecord(w.ev_next, st);
        bulk_update<T>(A, lda, far, far + nn, p0, restf, restf - nn, st, nb, bs.a, w.split);
        (void)hipEventRecord(w.ev_bulk, st);
        if(rec) {
          prof->gemm_flop += 8.0 * (double)restf * restf * (2 * kNB) * nb;
          prof->gemm_launches += restf > nn ? 2 : 1;
        }
      }
      if(rec) {
        (void)hipEventRecord(pe[3], st);
        prof->used = pi + 1;
      }
      if(w.ev_mark && far >= w.mark_col && far - 2 * kNB < w.mark_col) mark(w, st);
    }
    if(w.ev_mark && Np < w.mark_col) mark(w, st);
    if(Np > kNB)
      hipLaunchKernelGGL((laswp_left<T>), dim3((Np - kNB) / 8, Np / kNB - 1, nb), dim3(256), 0, st, A, lda, w.mdst,
                         w.msrc, w.mcnt, bs);
    return 0;
  }

  // ---- tournament pivoting
  auto pivoting = [&](int j0) {
    (void)hipEventRecord(w.ev_next, st);
    (void)hipStreamWaitEvent(sb, w.ev_next, 0);
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
  // interchanges on every column and the block's factors and inverses
  auto swaps = [&](int j0) {
    (void)hipStreamWaitEvent(st, w.ev_piv, 0);
    laswp<T>(A, lda, 0, Np, w.mdst, w.msrc, w.mcnt, -1, st);
    hipLaunchKernelGGL((panel_inv<T>), dim3(1), dim3(256), 0, st, A, lda, j0, wlu, dk(j0), w.status);
  };
  // L21 and U12 of panel j0
  auto solves = [&](int j0) {
    const int rest = Np - j0 - kNB;
    if(rest > 0) {
      hipLaunchKernelGGL((blk_mul<T, true>), dim3(rest / 64), dim3(256), 0, st, A, lda, j0, j0 + kNB, dk(j0),
                         BatchStrides{});
      hipLaunchKernelGGL((blk_mul<T, false>), dim3(rest / 64), dim3(256), 0, st, A, lda, j0, j0 + kNB,
                         dk(j0) + kNB * kNB, BatchStrides{});
    }
  };
  if((Np + kPanelRows - 1) / kPanelRows > w.gmax) return -2;
  pivoting(0);
  for(int J = 0, pi = 0; J < Np; J += 2 * kNB, ++pi) {
    const int p0 = J, p1 = J + kNB;
    const bool rec = prof && pi < prof->cap;
    hipEvent_t *pe = rec ? prof->ev + 4 * pi : nullptr;
    if(rec) (void)hipEventRecord(pe[0], st);
    swaps(p0);
    solves(p0);
    const int rest1 = Np - p1 - kNB;
    if(p1 < Np) {
      trailing_update<T, kNB, ThinTile>(A, lda, p1, p1, p0, Np - p1, kNB, st);
      pivoting(p1);
      // rows from below (not yet updated with p0 beyond p1) may move into
      // block row p1: interchange first, then update the block row
      swaps(p1);
      trailing_update<T, kNB, ThinTile>(A, lda, p1, p1 + kNB, p0, kNB, rest1, st);
      solves(p1);
    }
    if(rec) {
      (void)hipEventRecord(pe[1], st);
      (void)hipEventRecord(pe[2], st);
    }
    if(p1 < Np && rest1 > 0) {
      const int Jn = p1 + kNB;
      bulk_update<T>(A, lda, Jn, Jn, p0, rest1, kNB, st, 1, 0, w.split);
      pivoting(Jn);
      bulk_update<T>(A, lda, Jn, Jn + kNB, p0, rest1, rest1 - kNB, st, 1, 0, w.split);
      if(rec) {
        prof->gemm_flop += 8.0 * (double)rest1 * rest1 * (2 * kNB);
        prof->gemm_launches += rest1 > kNB ? 2 : 1;
      }
    }
    if(rec) {
      (void)hipEventRecord(pe[3], st);
      prof->used = pi + 1;
    }
    if(w.ev_mark && p1 + kNB >= w.mark_col && p0 < w.mark_col) mark(w, st);
  }
  if(w.ev_mark && Np < w.mark_col) mark(w, st);
  return 0;
}

// ------------------------------------------------------------------ solves
// Publication tags of the triangular solves: the low 4 bits of both 64-bit
// words of a published x entry hold 1 (the lower solve) or 2 (the upper
// solve), which changes the value by at most 2^-48 of itself.  The tags are
// fixed, so results do not depend on the history of launches.
constexpr unsigned long long kTagMask = 0xFull;
__device__ inline double with_tag(double v, unsigned tg) {
  return __longlong_as_double((long long)(((unsigned long long)__double_as_longlong(v) & ~kTagMask) | tg));
}
// The permuted right-hand side, with the tag bits cleared: a word the
// triangular solves have not yet published can never carry their tag.
__global__ void gather_kernel(const cplx *b, const int *perm, cplx *y, int Np, BatchStrides bs, const int *sys) {
  const size_t z = sys ? sys[blockIdx.z] : blockIdx.z;
  b += z * bs.vec;
  y += z * bs.vec;
  perm += z * bs.perm;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < Np) {
    const cplx v = b[perm[i]];
    y[i] = mk(with_tag(v.x, 0u), with_tag(v.y, 0u));
  }
}

// Sync-free blocked triangular solve.  Workgroup K owns SB consecutive
// 64-row blocks (a super-block).  Far phase: it consumes the blocks x_b of
// earlier super-blocks as they are published and accumulates
// L[super-block rows, b] x_b, with the next 64 x 16 tile slice of each thread
// prefetched into registers while the current one is multiplied.  Near phase: the
// super-block's own blocks are solved in sequence inside the workgroup -- the
// contributions of the blocks just solved, then the stored inverse of the
// diagonal block -- so only every SB-th link of the chain crosses
// workgroups.  With PRE the near phase's tiles (diagonal-block inverses and
// the intra-super-block L tiles) are loaded into registers at kernel start
// (PRE = 1) or after the far phase (PRE = 2).  Thread t handles row
// r = t % 64 and columns 16 (t / 64) .. 16 (t / 64) + 15 of every tile.
//
// Publication: a solved block is written with device-coherent (agent-scope)
// 64-bit stores whose low bits carry the solve's tag, and nothing else: a
// consumer polls the block's words themselves with device-coherent loads
// until every word carries the tag, so it needs no flag, no ordering between
// stores, and no second read.  Each 64-bit word is single-copy atomic, and
// the words it replaces carry another tag (0 in the right-hand side the
// gather wrote, 1 from the lower solve before the upper one), so a tagged
// word is always the new value.  Neither side writes back or invalidates the L2.
template <typename T, bool LOWER, int SB, int PRE>
__global__ __launch_bounds__(256) void trsv_kernel(const T *__restrict__ A, size_t lda, int Np,
                                                   const T *__restrict__ dinv, cplx *x,
                                                   int *status, BatchStrides bs, const int *sys) {
  constexpr int NB = kNB;
  constexpr int NI = SB * (SB - 1) / 2;
  const size_t zs = sys ? sys[blockIdx.z] : blockIdx.z;  // the system (a subset may be listed)
  A += zs * bs.a;
  dinv += zs * bs.inv;
  x += zs * bs.vec;
  constexpr unsigned tg = LOWER ? 1u : 2u;
  const int nblk = Np / NB, nsb = (nblk + SB - 1) / SB;
  const int K = LOWER ? blockIdx.x : nsb - 1 - blockIdx.x;
  const int kb0 = K * SB, nb = min(SB, nblk - kb0);
  const int t = threadIdx.x, r = t & 63;
  const int cq = __builtin_amdgcn_readfirstlane(t >> 6);  // wave-uniform
  constexpr int CH = 8;  // far blocks taken per round
  __shared__ cplx xs[CH * NB];
  __shared__ cplx part[4][NB];
  __shared__ cplx part2[SB][4][NB];
  __shared__ int nready_s;
  __shared__ cplx rhs[SB][NB];
  __shared__ cplx xsb[SB][NB];
  // right-hand sides of the super-block (only this workgroup writes them)
  cplx b0[SB];
#pragma unroll
  for(int s = 0; s < SB; ++s) b0[s] = (t < NB && s < nb) ? x[(size_t)(kb0 + s) * NB + t] : mk(0.0);
  // the thread's 16 entries of a 64 x 64 tile at base (column-major, leading
  // dimension ld): a wave-uniform column address plus the lane's row offset
  auto tile = [&](const T *base, size_t ld, T (&dst)[16]) {
    const T *col = base + (size_t)(cq * 16) * ld;
#pragma unroll
    for(int cc = 0; cc < 16; ++cc) dst[cc] = col[(size_t)cc * ld + (unsigned)r];
  };
  // diagonal-block inverse of block k, and the tile L[k, kc] (column-major in A)
  auto inv_base = [&](int k) { return dinv + (size_t)k * 2 * NB * NB + (LOWER ? 0 : NB * NB); };
  auto a_base = [&](int k, int kc) { return A + (size_t)kc * NB * lda + (size_t)k * NB; };
  // near-phase tiles: mv[s] = inverse of block kb0+s; lv[idx(s, s')] = L[kb0+s, kb0+s']
  T mv[PRE ? SB : 1][16], lv[PRE ? (NI > 0 ? NI : 1) : 1][16];
  auto preload = [&]() {
#pragma unroll
    for(int s = 0; s < SB; ++s)
      if(s < nb) tile(inv_base(kb0 + s), NB, mv[PRE ? s : 0]);
    int q = 0;
#pragma unroll
    for(int s = 0; s < SB; ++s)
#pragma unroll
      for(int s2 = 0; s2 < SB; ++s2) {
        if(LOWER ? (s2 < s) : (s2 > s)) {
          if(s < nb && s2 < nb) tile(a_base(kb0 + s, kb0 + s2), lda, lv[PRE ? q : 0]);
          ++q;
        }
      }
  };
  if(PRE == 1) preload();
  const int nfar = LOWER ? kb0 : nblk - (kb0 + nb);
  auto fblk = [&](int j) { return LOWER ? j : nblk - 1 - j; };  // j-th far block in order of publication
  // ---- far phase
  cplx acc[SB];
#pragma unroll
  for(int s = 0; s < SB; ++s) acc[s] = mk(0.0);
  T cur[16], nxt[16];
  if(nfar > 0) tile(a_base(kb0, fblk(0)), lda, cur);
  for(int j0 = 0; j0 < nfar;) {
    // the next far blocks, as far as they are published (wave 0, one entry
    // per lane): each poll reads a whole block and keeps it when every word
    // carries the tag; the blocks after a published one are tried at once
    if(t < 64) {
      unsigned spins = 0;
      int nr = 0;
      bool drain = false;
      while(nr < CH && j0 + nr < nfar) {
        const unsigned long long *src =
            reinterpret_cast<const unsigned long long *>(x + (size_t)fblk(j0 + nr) * NB + t);
        const unsigned long long re = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long im = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool ok = (re & kTagMask) == tg && (im & kTagMask) == tg;
        if(__ballot(ok) == ~0ull) {
          xs[nr * NB + t] = mk(__longlong_as_double((long long)re), __longlong_as_double((long long)im));
          ++nr;
          continue;
        }
        if(nr > 0) break;
        if(__builtin_amdgcn_readfirstlane(
               __hip_atomic_load(&status[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0) {
          drain = true;
          break;
        }
        if(++spins > (1u << 24)) {
          if(t == 0) __hip_atomic_store(&status[1], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          drain = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if(drain) {  // a timed-out wait: finish with zeros (the host reports the error)
        nr = min(CH, nfar - j0);
        for(int c = 0; c < nr; ++c) xs[c * NB + t] = mk(0.0);
      }
      if(t == 0) nready_s = nr;
    }
    __syncthreads();
    const int nr = nready_s;
    for(int c = 0; c < nr; ++c) {
      const int j = j0 + c, b = fblk(j);
#pragma unroll
      for(int s = 0; s < SB; ++s) {
        if(s < nb) {
          // prefetch the next slice: (s + 1, b) or (0, next far block)
          const bool inner = s + 1 < nb;
          if(inner || j + 1 < nfar) tile(inner ? a_base(kb0 + s + 1, b) : a_base(kb0, fblk(j + 1)), lda, nxt);
#pragma unroll
          for(int cc = 0; cc < 16; ++cc) {
            const T v = cur[cc];
            const cplx xv = xs[c * NB + cq * 16 + cc];
            acc[s].x += (double)v.x * xv.x - (double)v.y * xv.y;
            acc[s].y += (double)v.x * xv.y + (double)v.y * xv.x;
          }
#pragma unroll
          for(int cc = 0; cc < 16; ++cc) cur[cc] = nxt[cc];
        }
      }
    }
    __syncthreads();
    j0 += nr;
  }
  if(PRE == 2) preload();
  // right-hand sides of the super-block's blocks minus the far contributions
#pragma unroll
  for(int s = 0; s < SB; ++s) part2[s][cq][r] = acc[s];
  __syncthreads();
  if(t < NB) {
#pragma unroll
    for(int s = 0; s < SB; ++s)
      if(s < nb) rhs[s][t] = b0[s] - (part2[s][0][t] + part2[s][1][t] + part2[s][2][t] + part2[s][3][t]);
  }
  __syncthreads();
  // ---- near phase: the super-block's blocks in order
#pragma unroll
  for(int i = 0; i < SB; ++i) {
    const int s = LOWER ? i : SB - 1 - i;  // compile-time index; blocks past nb are skipped
    if(s < nb) {
      const int k = kb0 + s;
      // contributions of the blocks of this super-block solved before
      cplx a = mk(0.0);
      int q = 0;
      bool any = false;
#pragma unroll
      for(int ss = 0; ss < SB; ++ss)
#pragma unroll
        for(int s2 = 0; s2 < SB; ++s2) {
          if(LOWER ? (s2 < ss) : (s2 > ss)) {
            if(ss == s && s2 < nb) {
              any = true;
              T lt[16];
              if(PRE) {
#pragma unroll
                for(int cc = 0; cc < 16; ++cc) lt[cc] = lv[PRE ? q : 0][cc];
              } else {
                tile(a_base(k, kb0 + s2), lda, lt);
              }
#pragma unroll
              for(int cc = 0; cc < 16; ++cc) {
                const cplx xv = xsb[s2][cq * 16 + cc];
                a.x += (double)lt[cc].x * xv.x - (double)lt[cc].y * xv.y;
                a.y += (double)lt[cc].x * xv.y + (double)lt[cc].y * xv.x;
              }
            }
            ++q;
          }
        }
      if(any) {
        part[cq][r] = a;
        __syncthreads();
        if(t < NB) xs[t] = rhs[s][t] - (part[0][t] + part[1][t] + part[2][t] + part[3][t]);
      } else {
        if(t < NB) xs[t] = rhs[s][t];
      }
      __syncthreads();
      // x_k = (diagonal block)^{-1} r
      T m[16];
      if(PRE) {
#pragma unroll
        for(int cc = 0; cc < 16; ++cc) m[cc] = mv[PRE ? s : 0][cc];
      } else {
        tile(inv_base(k), NB, m);
      }
      cplx y = mk(0.0);
#pragma unroll
      for(int cc = 0; cc < 16; ++cc) {
        const cplx sv = xs[cq * 16 + cc];
        y.x += (double)m[cc].x * sv.x - (double)m[cc].y * sv.y;
        y.y += (double)m[cc].x * sv.y + (double)m[cc].y * sv.x;
      }
      part[cq][r] = y;
      __syncthreads();
      if(t < NB) {
        const cplx sum = part[0][t] + part[1][t] + part[2][t] + part[3][t];
        const cplx xk = mk(with_tag(sum.x, tg), with_tag(sum.y, tg));
        xsb[s][t] = xk;
        unsigned long long *dst = reinterpret_cast<unsigned long long *>(x + (size_t)k * NB + t);
        __hip_atomic_store(dst, (unsigned long long)__double_as_longlong(xk.x), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(dst + 1, (unsigned long long)__double_as_longlong(xk.y), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();  // part and xs are rewritten by the next block
    }
  }
}

// Super-block size and near-phase prefetch of the triangular solves (the
// complex-FP64 tiles of a prefetch would not fit in the register file):
// variant 0 is the default, 1 one block per workgroup (every link crosses
// workgroups, least work per link), 2 four blocks per workgroup with the
// near-phase tiles loaded as used (fewer cross-workgroup links).
template <typename T> struct TrsvCfg;
template <> struct TrsvCfg<float2> {
  static constexpr int SB[kTrsvVariants] = {2, 1, 4};
  static constexpr int PRE[kTrsvVariants] = {1, 1, 0};
};
template <> struct TrsvCfg<double2> {
  static constexpr int SB[kTrsvVariants] = {2, 1, 4};
  static constexpr int PRE[kTrsvVariants] = {2, 2, 0};
};

template <typename T, int V>
static void trsv_pair(const T *A, int Np, size_t lda, LuWork &w, cplx *x, int nz, hipStream_t st) {
  constexpr int SB = TrsvCfg<T>::SB[V];
  constexpr int PRE = TrsvCfg<T>::PRE[V];
  const T *dinv = reinterpret_cast<const T *>(w.dinv);
  const int nblk = Np / kNB, nsb = (nblk + SB - 1) / SB;
  hipLaunchKernelGGL((trsv_kernel<T, true, SB, PRE>), dim3(nsb, 1, nz), dim3(256), 0, st, A, lda, Np, dinv, x,
                     w.status, w.bs, w.sys);
  hipLaunchKernelGGL((trsv_kernel<T, false, SB, PRE>), dim3(nsb, 1, nz), dim3(256), 0, st, A, lda, Np, dinv, x,
                     w.status, w.bs, w.sys);
}

template <typename T>
void lu_solve(const T *A, int Np, size_t lda, LuWork &w, const cplx *b, cplx *x, hipStream_t st) {
  const int nz = w.sys ? w.nsys : w.nb;
  if(nz < 1) return;
  hipLaunchKernelGGL(gather_kernel, dim3((Np + 255) / 256, 1, nz), dim3(256), 0, st, b, w.perm, x, Np, w.bs, w.sys);
  switch(w.trsv) {
  case 1: trsv_pair<T, 1>(A, Np, lda, w, x, nz, st); break;
  case 2: trsv_pair<T, 2>(A, Np, lda, w, x, nz, st); break;
  default: trsv_pair<T, 0>(A, Np, lda, w, x, nz, st); break;
  }
}

template <typename T>
void bulk_update_probe(T *A, int Np, size_t lda, int nb, size_t sA, bool split, hipStream_t st) {
  if(Np <= 2 * kNB) return;
  bulk_update<T>(A, lda, 2 * kNB, 2 * kNB, 0, Np - 2 * kNB, Np - 2 * kNB, st, nb, sA, split);
}
template void bulk_update_probe<double2>(double2 *, int, size_t, int, size_t, bool, hipStream_t);
template void bulk_update_probe<float2>(float2 *, int, size_t, int, size_t, bool, hipStream_t);

template int lu_factor<double2>(double2 *, int, size_t, LuWork &, hipStream_t, LuProf *);
template int lu_factor<float2>(float2 *, int, size_t, LuWork &, hipStream_t, LuProf *);
template void lu_solve<double2>(const double2 *, int, size_t, LuWork &, const cplx *, cplx *, hipStream_t);
template void lu_solve<float2>(const float2 *, int, size_t, LuWork &, const cplx *, cplx *, hipStream_t);

// ------------------------------------------------------------------ gemv
// y = A x for the FP64 matrix (residuals and GMRES) or its complex-FP32 copy
// (entries widened to FP64, FP64 vectors and sums), two-stage column split.
constexpr int kGemvSplit = 32;
template <typename T>
__global__ __launch_bounds__(256) void gemv_part_kernel(const T *__restrict__ A, size_t lda, int N,
                                                        const cplx *__restrict__ x, cplx *part, int chunk, size_t sA,
                                                        size_t sv, const int *sys) {
  const size_t zs = sys ? sys[blockIdx.z] : blockIdx.z;
  A += zs * sA;
  x += zs * sv;
  part += (size_t)blockIdx.z * gridDim.y * N;
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
      const T *a = A + (size_t)cb * lda + r;
      int n = ce - cb;
#pragma unroll 8
      for(int c = 0; c < n; ++c) {
        const T e = ld_stream(a + (size_t)c * lda);
        const double vx = e.x, vy = e.y;
        acc.x += vx * xs[c].x - vy * xs[c].y;
        acc.y += vx * xs[c].y + vy * xs[c].x;
      }
    }
  }
  if(r < N) part[(size_t)blockIdx.y * N + r] = acc;
}
__global__ void gemv_reduce_kernel(const cplx *part, int ns, int N, cplx *y, size_t sv, const int *sys) {
  part += (size_t)blockIdx.z * ns * N;
  y += (sys ? sys[blockIdx.z] : blockIdx.z) * sv;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= N) return;
  cplx s = mk(0.0);
  for(int k = 0; k < ns; ++k) s = s + part[(size_t)k * N + r];
  y[r] = s;
}
// y = A x for nb systems (matrices sA, vectors sv apart); part holds
// kGemvSplit N partial vectors per system
// Column parts of a batched product: the fewest (the longest runs of
// columns, the fewest partial vectors) that still give 1280 workgroups, about
// five per CU.  Measured on 3840 x 3840 systems: 6.75 TB/s with 8 parts for 64
// systems against 5.94 with 32, 6.20 with 24 parts for 4 systems; one system of
// 10240 is fastest with kGemvSplit parts.
inline int gemv_parts(int N, int nz) {
  if(nz <= 1) return kGemvSplit;
  const long rb = (N + 255) / 256;
  for(int s : {8, 12, 16, 24})
    if(rb * nz * s >= 1280) return s;
  return kGemvSplit;
}
void gemv64(const cplx *A, size_t lda, int N, const cplx *x, cplx *y, cplx *part, hipStream_t st, int nb, size_t sA,
            size_t sv, const int *sys, int nsys) {
  const int nz = sys ? nsys : nb;
  if(nz < 1) return;
  const int ns = gemv_parts(N, nz);  // <= kGemvSplit: part holds kGemvSplit N vectors per system
  int chunk = (N + ns - 1) / ns;
  hipLaunchKernelGGL(gemv_part_kernel<cplx>, dim3((N + 255) / 256, ns, nz), dim3(256), 0, st, A, lda, N, x, part,
                     chunk, sA, sv, sys);
  hipLaunchKernelGGL(gemv_reduce_kernel, dim3((N + 255) / 256, 1, nz), dim3(256), 0, st, part, ns, N, y, sv, sys);
}
void gemv32(const float2 *A, size_t lda, int N, const cplx *x, cplx *y, cplx *part, hipStream_t st) {
  int chunk = (N + kGemvSplit - 1) / kGemvSplit;
  hipLaunchKernelGGL(gemv_part_kernel<float2>, dim3((N + 255) / 256, kGemvSplit, 1), dim3(256), 0, st, A, lda, N, x,
                     part, chunk, (size_t)0, (size_t)0, (const int *)nullptr);
  hipLaunchKernelGGL(gemv_reduce_kernel, dim3((N + 255) / 256, 1, 1), dim3(256), 0, st, part, kGemvSplit, N, y,
                     (size_t)0, (const int *)nullptr);
}

}  // namespace omx
