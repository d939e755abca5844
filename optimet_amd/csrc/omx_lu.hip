// omx_lu.hip -- blocked complex LU with partial pivoting for CDNA4.
//
// Right-looking blocked factorisation P A = L U of a column-major Np x Np
// complex matrix (Np a multiple of the block size NB), in complex FP64
// (double2) or complex FP32 (float2):
//   panel_kernel   cooperative factorisation of the NB-wide panel: every
//                  workgroup keeps 128 panel rows in LDS; per column one grid
//                  barrier exchanges the workgroups' pivot candidates
//                  (double-buffered, so one barrier per column suffices)
//   laswp_kernel   the panel's row interchanges on all other columns and on
//                  the running row permutation
//   trsm_kernel    U12 = L11^{-1} A12, one column per thread, the column held
//                  in registers and L11 in LDS
//   gemm_kernel    A22 -= L21 U12 as complex products on the matrix cores:
//                  v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32
//                  (FP32), four real MFMAs per complex tile product, planar
//                  re/im operands staged through LDS
//   trsv_kernel    sync-free blocked triangular solves: workgroup k owns
//                  block row k, consumes earlier blocks as their flags are
//                  published and solves its diagonal block in one wave
// This replaces the ScaLAPACK pzgesv path (scalapack/LinearSystemSolver.hpp:84-91)
// and Eigen's ColPivHouseholderQR (PreconditionedMatrixSolver.h:58, 75).
#include <climits>
#include <vector>

#include "omx_kernels.hpp"

namespace omx {

template <typename T> struct CT;
template <> struct CT<double2> {
  using R = double;
  typedef double V4 __attribute__((ext_vector_type(4)));
  static __device__ inline V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <> struct CT<float2> {
  using R = float;
  typedef float V4 __attribute__((ext_vector_type(4)));
  static __device__ inline V4 mfma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

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
  using R = typename CT<T>::R;
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

__device__ inline void grid_sync(unsigned *bar, unsigned target, int *status) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if(threadIdx.x == 0) {
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while(__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target &&
          __hip_atomic_load(&status[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      if(++spins > (1u << 24)) {  // bounded wait: report instead of hanging
        __hip_atomic_store(&status[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__global__ void zero_u32(unsigned *p) { *p = 0u; }
__global__ void iota_kernel(int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = i;
}

// ------------------------------------------------------------------ panel
template <typename T, int NB, int RW>
__global__ __launch_bounds__(256) void panel_kernel(T *A, size_t lda, int Np, int j0, int G, LuWork w,
                                                    unsigned base) {
  extern __shared__ unsigned char smem[];
  T *As = reinterpret_cast<T *>(smem);  // [NB][RW], column-major chunk
  T *u = As + NB * RW;                  // [NB]
  double *rv = reinterpret_cast<double *>(u + NB);
  int *ri = reinterpret_cast<int *>(rv + 256);
  int *rg = ri + 256;
  const int g = blockIdx.x, t = threadIdx.x;
  const int r0 = j0 + g * RW;
  const int nr = min(RW, Np - r0);
  T *cand_row = reinterpret_cast<T *>(w.cand_row);
  T *rowj = reinterpret_cast<T *>(w.rowj);
  const T zero{};
  for(int idx = t; idx < NB * RW; idx += 256) {
    int c = idx / RW, r = idx % RW;
    As[idx] = r < nr ? A[(size_t)(j0 + c) * lda + r0 + r] : zero;
  }
  __syncthreads();
  for(int c = 0; c < NB; ++c) {
    const int j = j0 + c, buf = c & 1;
    // (a) local pivot candidate among rows >= j
    double best = -1.0;
    int bidx = INT_MAX;
    for(int r = t; r < nr; r += 256) {
      int gr = r0 + r;
      if(gr >= j) {
        double v = mag2(As[c * RW + r]);
        if(v > best || (v == best && gr < bidx)) best = v, bidx = gr;
      }
    }
    rv[t] = best;
    ri[t] = bidx;
    __syncthreads();
    for(int s = 128; s > 0; s >>= 1) {
      if(t < s) {
        double v2 = rv[t + s];
        int i2 = ri[t + s];
        if(v2 > rv[t] || (v2 == rv[t] && i2 < ri[t])) rv[t] = v2, ri[t] = i2;
      }
      __syncthreads();
    }
    const double lbest = rv[0];
    const int lidx = ri[0];
    if(t < NB && lbest >= 0.0) cand_row[((size_t)buf * G + g) * NB + t] = As[t * RW + (lidx - r0)];
    if(t == 0) {
      w.cand_val[buf * G + g] = lbest;
      w.cand_idx[buf * G + g] = lidx;
    }
    if(j >= r0 && j < r0 + nr && t < NB) rowj[buf * NB + t] = As[t * RW + (j - r0)];
    // (b) exchange
    grid_sync(w.bar, base + (unsigned)G * (c + 1), w.status);
    // (c) global pivot
    double bv = -1.0;
    int bi = INT_MAX, bg = -1;
    for(int k = t; k < G; k += 256) {
      double v = w.cand_val[buf * G + k];
      int id = w.cand_idx[buf * G + k];
      if(v > bv || (v == bv && id < bi)) bv = v, bi = id, bg = k;
    }
    rv[t] = bv;
    ri[t] = bi;
    rg[t] = bg;
    __syncthreads();
    for(int s = 128; s > 0; s >>= 1) {
      if(t < s) {
        double v2 = rv[t + s];
        int i2 = ri[t + s];
        if(v2 > rv[t] || (v2 == rv[t] && i2 < ri[t])) rv[t] = v2, ri[t] = i2, rg[t] = rg[t + s];
      }
      __syncthreads();
    }
    const bool singular = !(rv[0] > 0.0);
    const int p = singular ? j : ri[0];
    const int win = rg[0];
    if(t < NB) u[t] = singular ? As[t * RW + 0] : cand_row[((size_t)buf * G + win) * NB + t];
    __syncthreads();
    if(singular) {
      if(t == 0 && g == 0) atomicCAS(&w.status[0], 0, j + 1);
      // keep row j in place; the column is left unscaled
      if(j >= r0 && j < r0 + nr && t < NB) u[t] = As[t * RW + (j - r0)];
      __syncthreads();
    }
    if(g == 0 && t == 0) w.piv[j] = p;
    if(t < NB) {
      if(j >= r0 && j < r0 + nr) As[t * RW + (j - r0)] = u[t];
      if(p != j && p >= r0 && p < r0 + nr) As[t * RW + (p - r0)] = rowj[buf * NB + t];
    }
    __syncthreads();
    // (d) eliminate below the pivot
    const T ip = singular ? zero : crecip(u[c]);
    const int r = t % RW, cgp = t / RW;
    constexpr int NCG = 256 / RW;
    T lval = zero;
    bool act = (r < nr) && (r0 + r > j);
    if(act) lval = cm(As[c * RW + r], ip);
    __syncthreads();
    if(act) {
      if(cgp == 0) As[c * RW + r] = lval;
      for(int cc = c + 1 + cgp; cc < NB; cc += NCG) As[cc * RW + r] = csub(As[cc * RW + r], cm(lval, u[cc]));
    }
    __syncthreads();
  }
  for(int idx = t; idx < NB * RW; idx += 256) {
    int c = idx / RW, r = idx % RW;
    if(r < nr) A[(size_t)(j0 + c) * lda + r0 + r] = As[idx];
  }
}

// ------------------------------------------------------------------ swaps
template <typename T>
__global__ void laswp_kernel(T *A, size_t lda, int Np, int j0, int nb, const int *__restrict__ piv, int *perm) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int ncol = Np - nb;
  if(c > ncol) return;
  if(c == ncol) {
    for(int jj = 0; jj < nb; ++jj) {
      int j = j0 + jj, p = piv[j];
      if(p != j) {
        int tmp = perm[j];
        perm[j] = perm[p];
        perm[p] = tmp;
      }
    }
    return;
  }
  int col = c < j0 ? c : c + nb;
  T *a = A + (size_t)col * lda;
  for(int jj = 0; jj < nb; ++jj) {
    int j = j0 + jj, p = piv[j];
    if(p != j) {
      T tmp = a[j];
      a[j] = a[p];
      a[p] = tmp;
    }
  }
}

// ------------------------------------------------------------------ trsm
template <typename T, int NB>
__global__ __launch_bounds__(256) void trsm_kernel(T *A, size_t lda, int Np, int j0) {
  extern __shared__ unsigned char smem[];
  T *L = reinterpret_cast<T *>(smem);  // [NB][NB] column-major
  for(int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
    int c = idx / NB, r = idx % NB;
    L[idx] = A[(size_t)(j0 + c) * lda + j0 + r];
  }
  __syncthreads();
  int col = j0 + NB + blockIdx.x * blockDim.x + threadIdx.x;
  if(col >= Np) return;
  T *a = A + (size_t)col * lda + j0;
  T x[NB];
#pragma unroll
  for(int i = 0; i < NB; ++i) x[i] = a[i];
#pragma unroll
  for(int jj = 0; jj < NB; ++jj) {
#pragma unroll
    for(int i = jj + 1; i < NB; ++i) x[i] = csub(x[i], cm(L[jj * NB + i], x[jj]));
  }
#pragma unroll
  for(int i = 0; i < NB; ++i) a[i] = x[i];
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
      int rowb = bm + wm * 32 + a * 16 + 4 * (lane >> 4);
      T *cp = A + (size_t)col * lda + rowb;
#pragma unroll
      for(int v = 0; v < 4; ++v) {
        T c = cp[v];
        c.x -= cr[a][b][v];
        c.y -= ci[a][b][v];
        cp[v] = c;
      }
    }
}

// ------------------------------------------------------------------ driver
template <typename T>
static size_t panel_lds() {
  return sizeof(T) * (kNB * kPanelRows + kNB) + 256 * sizeof(double) + 2 * 256 * sizeof(int);
}

template <typename T>
int lu_factor(T *A, int Np, size_t lda, LuWork &w, hipStream_t st) {
  if(Np % kNB) return -1;
  static bool attr = false;
  if(!attr) {
    (void)hipFuncSetAttribute((const void *)panel_kernel<T, kNB, kPanelRows>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)panel_lds<T>());
    (void)hipFuncSetAttribute((const void *)trsm_kernel<T, kNB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)(sizeof(T) * kNB * kNB));
    attr = true;
  }
  hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, st, w.bar);
  hipLaunchKernelGGL(iota_kernel, dim3((Np + 255) / 256), dim3(256), 0, st, w.perm, Np);
  unsigned base = 0;
  for(int j0 = 0; j0 < Np; j0 += kNB) {
    int G = (Np - j0 + kPanelRows - 1) / kPanelRows;
    if(G > w.gmax) return -2;
    void *args[] = {&A, &lda, &Np, &j0, &G, &w, &base};
    hipError_t e = hipLaunchCooperativeKernel((const void *)panel_kernel<T, kNB, kPanelRows>, dim3(G),
                                              dim3(256), args, panel_lds<T>(), st);
    if(e != hipSuccess) return -3;
    base += (unsigned)G * kNB;
    hipLaunchKernelGGL(laswp_kernel<T>, dim3((Np - kNB + 1 + 255) / 256), dim3(256), 0, st, A, lda, Np, j0,
                       kNB, w.piv, w.perm);
    int rest = Np - j0 - kNB;
    if(rest > 0) {
      hipLaunchKernelGGL((trsm_kernel<T, kNB>), dim3((rest + 255) / 256), dim3(256), sizeof(T) * kNB * kNB, st,
                         A, lda, Np, j0);
      hipLaunchKernelGGL((gemm_kernel<T, kNB>), dim3(rest / 64, rest / 64), dim3(256), 0, st, A, lda,
                         j0 + kNB, j0 + kNB, j0);
    }
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
  const int t = threadIdx.x, r = t & 63, cq = t >> 6;
  cplx acc = mk(0.0);
  const int ndep = LOWER ? k : nblk - 1 - k;
  for(int ii = 0; ii < ndep; ++ii) {
    const int b = LOWER ? ii : nblk - 1 - ii;
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
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if(t < NB) xs[t] = x[(size_t)b * NB + t];
    __syncthreads();
    const T *a = A + (size_t)(b * NB + cq * 16) * lda + (size_t)k * NB + r;
#pragma unroll 4
    for(int cc = 0; cc < 16; ++cc) {
      T v = a[(size_t)cc * lda];
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
    const T *d = A + (size_t)k * NB * lda + (size_t)k * NB;
    if(LOWER) {
      for(int i = 0; i < NB; ++i) {
        cplx yi;
        yi.x = __shfl(s.x, i, 64);
        yi.y = __shfl(s.y, i, 64);
        if(t > i) {
          T v = d[(size_t)i * lda + t];
          s.x -= (double)v.x * yi.x - (double)v.y * yi.y;
          s.y -= (double)v.x * yi.y + (double)v.y * yi.x;
        }
      }
    } else {
      for(int i = NB - 1; i >= 0; --i) {
        if(t == i) {
          T v = d[(size_t)i * lda + i];
          s = s / mk((double)v.x, (double)v.y);
        }
        cplx yi;
        yi.x = __shfl(s.x, i, 64);
        yi.y = __shfl(s.y, i, 64);
        if(t < i) {
          T v = d[(size_t)i * lda + t];
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

template int lu_factor<double2>(double2 *, int, size_t, LuWork &, hipStream_t);
template int lu_factor<float2>(float2 *, int, size_t, LuWork &, hipStream_t);
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
