// omx_iter.hip -- level-1 kernels of the iterative solvers (GMRES, GMRES-IR)
// and the element-wise unpreconditioning steps (Solver.cpp:57-116).
#include <type_traits>

#include "omx_kernels.hpp"

namespace omx {

__global__ void vcopy_k(cplx *y, const cplx *x, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) y[i] = x[i];
}
__global__ void vaxpby_k(cplx *y, cplx a, const cplx *x, cplx b, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) y[i] = a * x[i] + b * y[i];
}
__global__ void vsub_k(cplx *z, const cplx *x, const cplx *y, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) z[i] = x[i] - y[i];
}
// y[i] = x[i] * ov[kind][obj][pol][deg(i)] - sub[i], over the N = 2 nf P unknowns
__global__ void vmul_ov_k(cplx *y, const cplx *x, const cplx *ov, int kind, int P, int nmax, const cplx *sub,
                          int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int nf = nmax * (nmax + 2);
  int obj = i / (2 * nf), r = i % (2 * nf);
  int pol = r >= nf, q = r % nf, deg, m;
  unflat(q, deg, m);
  cplx v = x[i] * ov[(((size_t)kind * P + obj) * 2 + pol) * (nmax + 1) + deg];
  if(sub) v = v - sub[i];
  y[i] = v;
}

// out[k] = sum_i conj(V[k][i]) w[i]; one partial per block, then reduced.
constexpr int kRedBlocks = 128;
// Partial sums of v_k^H w, one workgroup row per vector k (blockIdx.y): the
// vectors' reductions run side by side instead of one after another.
__global__ __launch_bounds__(256) void dotc_multi_k(const cplx *V, size_t ldv, int nv, const cplx *w, int n,
                                                    cplx *part) {
  __shared__ cplx red[256];
  const int k = blockIdx.y;
  cplx s = mk(0.0);
  for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s = s + conj(V[k * ldv + i]) * w[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for(int st = 128; st > 0; st >>= 1) {
    if(threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[(size_t)k * gridDim.x + blockIdx.x] = red[0];
}
__global__ void reduce_rows_k(const cplx *part, int nb, int nv, cplx *out) {
  int k = blockIdx.x;
  __shared__ cplx red[128];
  cplx s = mk(0.0);
  for(int b = threadIdx.x; b < nb; b += 128) s = s + part[(size_t)k * nb + b];
  red[threadIdx.x] = s;
  __syncthreads();
  for(int st = 64; st > 0; st >>= 1) {
    if(threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if(threadIdx.x == 0) out[k] = red[0];
}
__global__ void axpy_multi_k(const cplx *V, size_t ldv, int nv, const cplx *h, cplx *w, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  cplx s = w[i];
  for(int k = 0; k < nv; ++k) s = s - V[k * ldv + i] * h[k];
  w[i] = s;
}
__global__ void gemv_multi_k(const cplx *V, size_t ldv, int nv, const cplx *h, cplx *z, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  cplx s = mk(0.0);
  for(int k = 0; k < nv; ++k) s = s + V[k * ldv + i] * h[k];
  z[i] = s;
}
__global__ __launch_bounds__(256) void nrm2_k(const cplx *x, int n, double *part) {
  __shared__ double red[256];
  double s = 0.0;
  for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += norm2(x[i]);
  red[threadIdx.x] = s;
  __syncthreads();
  for(int st = 128; st > 0; st >>= 1) {
    if(threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void nrm2_reduce_k(const double *part, int nb, double *out) {
  __shared__ double red[128];
  double s = 0.0;
  for(int b = threadIdx.x; b < nb; b += 128) s += part[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for(int st = 64; st > 0; st >>= 1) {
    if(threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if(threadIdx.x == 0) *out = red[0];
}

static inline dim3 g1(int n) { return dim3((n + 255) / 256); }

// y = (complex float) x over n elements (grid-stride: n may exceed 2^31)
__global__ void demote_k(float2 *y, const cplx *x, size_t n) {
  for(size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = make_float2((float)x[i].x, (float)x[i].y);
}
void demote(float2 *y, const cplx *x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(demote_k, dim3(4096), dim3(256), 0, st, y, x, n);
}
__global__ void zero_ints_k(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 0;
}
void zero_ints(int *p, int n, hipStream_t st) { hipLaunchKernelGGL(zero_ints_k, g1(n), dim3(256), 0, st, p, n); }

// dst (Nd x Nd, column-major) = src (Ns x Ns, leading dimension Ns) bordered
// by the identity (grid-stride: Nd^2 may exceed 2^31)
template <typename D>
__global__ void pad_matrix_k(D *dst, int Nd, const cplx *src, int Ns) {
  const size_t n = (size_t)Nd * Nd;
  for(size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / Nd), i = (int)(e % Nd);
    const cplx v = (i < Ns && j < Ns) ? src[(size_t)j * Ns + i] : mk(i == j ? 1.0 : 0.0);
    if constexpr(std::is_same<D, cplx>::value)
      dst[e] = v;
    else
      dst[e] = make_float2((float)v.x, (float)v.y);
  }
}
void pad_matrix64(cplx *dst, int Nd, const cplx *src, int Ns, hipStream_t st) {
  hipLaunchKernelGGL(pad_matrix_k<cplx>, dim3(4096), dim3(256), 0, st, dst, Nd, src, Ns);
}
void pad_matrix32(float2 *dst, int Nd, const cplx *src, int Ns, hipStream_t st) {
  hipLaunchKernelGGL(pad_matrix_k<float2>, dim3(4096), dim3(256), 0, st, dst, Nd, src, Ns);
}
// dst[0, Nd) = src[0, Ns) followed by zeros
__global__ void pad_vector_k(cplx *dst, int Nd, const cplx *src, int Ns) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < Nd) dst[i] = i < Ns ? src[i] : mk(0.0);
}
void pad_vector(cplx *dst, int Nd, const cplx *src, int Ns, hipStream_t st) {
  hipLaunchKernelGGL(pad_vector_k, g1(Nd), dim3(256), 0, st, dst, Nd, src, Ns);
}
void vcopy(cplx *y, const cplx *x, int n, hipStream_t st) { hipLaunchKernelGGL(vcopy_k, g1(n), dim3(256), 0, st, y, x, n); }
void vaxpby(cplx *y, cplx a, const cplx *x, cplx b, int n, hipStream_t st) {
  hipLaunchKernelGGL(vaxpby_k, g1(n), dim3(256), 0, st, y, a, x, b, n);
}
void vsub(cplx *z, const cplx *x, const cplx *y, int n, hipStream_t st) {
  hipLaunchKernelGGL(vsub_k, g1(n), dim3(256), 0, st, z, x, y, n);
}
void vmul_ov(cplx *y, const cplx *x, const cplx *ov, int kind, int P, int nmax, const cplx *sub, int n,
             hipStream_t st) {
  hipLaunchKernelGGL(vmul_ov_k, g1(n), dim3(256), 0, st, y, x, ov, kind, P, nmax, sub, n);
}
void dotc_multi(const cplx *V, size_t ldv, int nv, const cplx *w, int n, cplx *part, cplx *out, hipStream_t st) {
  int nb = std::min(kRedBlocks, (n + 255) / 256);
  hipLaunchKernelGGL(dotc_multi_k, dim3(nb, nv), dim3(256), 0, st, V, ldv, nv, w, n, part);
  hipLaunchKernelGGL(reduce_rows_k, dim3(nv), dim3(128), 0, st, part, nb, nv, out);
}
void axpy_multi(const cplx *V, size_t ldv, int nv, const cplx *h, cplx *w, int n, hipStream_t st) {
  hipLaunchKernelGGL(axpy_multi_k, g1(n), dim3(256), 0, st, V, ldv, nv, h, w, n);
}
void gemv_multi(const cplx *V, size_t ldv, int nv, const cplx *h, cplx *z, int n, hipStream_t st) {
  hipLaunchKernelGGL(gemv_multi_k, g1(n), dim3(256), 0, st, V, ldv, nv, h, z, n);
}
void nrm2sq(const cplx *x, int n, double *part, double *out, hipStream_t st) {
  int nb = std::min(kRedBlocks, (n + 255) / 256);
  hipLaunchKernelGGL(nrm2_k, dim3(nb), dim3(256), 0, st, x, n, part);
  hipLaunchKernelGGL(nrm2_reduce_k, dim3(1), dim3(128), 0, st, part, nb, out);
}

// ---------------------------------------------------------------- batches
// nb systems side by side (blockIdx.y = system): vectors sv apart, Krylov
// bases sV apart (vectors ldv apart inside a basis), per-system scalars and
// coefficient vectors in device arrays.
__global__ void vaxpby_b_k(cplx *y, size_t sy, const cplx *a, const cplx *x, size_t sx, const cplx *b, int n,
                           const int *act) {
  const int z = blockIdx.y;
  if(act && !act[z]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) y[z * sy + i] = a[z] * x[z * sx + i] + b[z] * y[z * sy + i];
}
__global__ void vsub_b_k(cplx *zz, const cplx *x, const cplx *y, int n, size_t sv) {
  const size_t o = blockIdx.y * sv;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) zz[o + i] = x[o + i] - y[o + i];
}
// The batched form: system z = blockIdx.y, vector k = blockIdx.z.
__global__ __launch_bounds__(256) void dotc_multi_b_k(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *w,
                                                      size_t sv, int n, cplx *part) {
  const int z = blockIdx.y, k = blockIdx.z;
  V += z * sV;
  w += z * sv;
  part += (size_t)z * nv * gridDim.x;
  __shared__ cplx red[256];
  cplx s = mk(0.0);
  for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s = s + conj(V[k * ldv + i]) * w[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for(int st = 128; st > 0; st >>= 1) {
    if(threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[(size_t)k * gridDim.x + blockIdx.x] = red[0];
}
__global__ void reduce_rows_b_k(const cplx *part, int nb, int nv, cplx *out, int so) {
  const int z = blockIdx.y, k = blockIdx.x;
  part += (size_t)z * nv * nb;
  __shared__ cplx red[128];
  cplx s = mk(0.0);
  for(int b = threadIdx.x; b < nb; b += 128) s = s + part[(size_t)k * nb + b];
  red[threadIdx.x] = s;
  __syncthreads();
  for(int st = 64; st > 0; st >>= 1) {
    if(threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if(threadIdx.x == 0) out[(size_t)z * so + k] = red[0];
}
__global__ void axpy_multi_b_k(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *h, int sh, cplx *w,
                               size_t sv, int n) {
  const int z = blockIdx.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  V += z * sV;
  h += (size_t)z * sh;
  cplx s = w[z * sv + i];
  for(int k = 0; k < nv; ++k) s = s - V[k * ldv + i] * h[k];
  w[z * sv + i] = s;
}
__global__ void gemv_multi_b_k(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *h, int sh, cplx *zz,
                               size_t sv, int n) {
  const int z = blockIdx.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  V += z * sV;
  h += (size_t)z * sh;
  cplx s = mk(0.0);
  for(int k = 0; k < nv; ++k) s = s + V[k * ldv + i] * h[k];
  zz[z * sv + i] = s;
}
__global__ __launch_bounds__(256) void nrm2_b_k(const cplx *x, size_t sv, int n, double *part) {
  const int z = blockIdx.y;
  x += z * sv;
  __shared__ double red[256];
  double s = 0.0;
  for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += norm2(x[i]);
  red[threadIdx.x] = s;
  __syncthreads();
  for(int st = 128; st > 0; st >>= 1) {
    if(threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[(size_t)z * gridDim.x + blockIdx.x] = red[0];
}
__global__ void nrm2_reduce_b_k(const double *part, int nb, double *out) {
  const int z = blockIdx.x;
  part += (size_t)z * nb;
  __shared__ double red[128];
  double s = 0.0;
  for(int b = threadIdx.x; b < nb; b += 128) s += part[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for(int st = 64; st > 0; st >>= 1) {
    if(threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if(threadIdx.x == 0) out[z] = red[0];
}

static inline dim3 gb(int n, int nb) { return dim3((n + 255) / 256, nb); }
void vaxpby_b(cplx *y, size_t sy, const cplx *a, const cplx *x, size_t sx, const cplx *b, int n, int nb,
              const int *act, hipStream_t st) {
  hipLaunchKernelGGL(vaxpby_b_k, gb(n, nb), dim3(256), 0, st, y, sy, a, x, sx, b, n, act);
}
void vsub_b(cplx *z, const cplx *x, const cplx *y, int n, int nb, size_t sv, hipStream_t st) {
  hipLaunchKernelGGL(vsub_b_k, gb(n, nb), dim3(256), 0, st, z, x, y, n, sv);
}
int red_blocks_b(int n, int nb) { return std::max(1, std::min(kRedBlocks / std::max(1, nb / 8), (n + 255) / 256)); }
void dotc_multi_b(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *w, size_t sv, int n, int nb, cplx *part,
                  cplx *out, int so, hipStream_t st) {
  const int rb = red_blocks_b(n, nb);
  hipLaunchKernelGGL(dotc_multi_b_k, dim3(rb, nb, nv), dim3(256), 0, st, V, ldv, sV, nv, w, sv, n, part);
  hipLaunchKernelGGL(reduce_rows_b_k, dim3(nv, nb), dim3(128), 0, st, part, rb, nv, out, so);
}
void axpy_multi_b(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *h, int sh, cplx *w, size_t sv, int n,
                  int nb, hipStream_t st) {
  hipLaunchKernelGGL(axpy_multi_b_k, gb(n, nb), dim3(256), 0, st, V, ldv, sV, nv, h, sh, w, sv, n);
}
void gemv_multi_b(const cplx *V, size_t ldv, size_t sV, int nv, const cplx *h, int sh, cplx *z, size_t sv, int n,
                  int nb, hipStream_t st) {
  hipLaunchKernelGGL(gemv_multi_b_k, gb(n, nb), dim3(256), 0, st, V, ldv, sV, nv, h, sh, z, sv, n);
}
void nrm2sq_b(const cplx *x, size_t sv, int n, int nb, double *part, double *out, hipStream_t st) {
  const int rb = red_blocks_b(n, nb);
  hipLaunchKernelGGL(nrm2_b_k, dim3(rb, nb), dim3(256), 0, st, x, sv, n, part);
  hipLaunchKernelGGL(nrm2_reduce_b_k, dim3(nb), dim3(128), 0, st, part, rb, out);
}

}  // namespace omx
