// omx_aca.hip -- adaptive cross approximation of the coupling blocks and the
// compressed matrix-vector product.
//
//   aca_kernel       ACA_compression (PreconditionedMatrix.cpp:763-862) with
//                    getMaxInd (:864-889): one workgroup per admissible block
//   aca_pack_kernel  the factors of each block, U (n x r) then V (r x n),
//                    copied from the n x n staging slots into a packed pool
//   aca_vt_kernel    t_b = V_b x_j for every compressed block b = (i, j)
//   aca_rows_kernel  partial y_i over a group of block row i's blocks:
//                    dense S_ij x_j or U_b t_b (matvec, :1058-1085)
//   aca_sum_kernel   y = x + sum of the partials (the identity diagonal blocks)
#include "omx_kernels.hpp"

namespace omx {

namespace {

__device__ inline double wave_sum(double v) {
  for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Largest value, first index on ties (getMaxInd keeps the first strict maximum).
__device__ inline void wave_argmax(double &v, int &i) {
  for(int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o);
    int oi = __shfl_xor(i, o);
    if(ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}

// Workgroup argmax of v over threads (v < 0: not a candidate); every thread
// returns the winning index, or -1 when no candidate has a non-zero modulus.
__device__ inline int block_argmax(double v, double (*wv)[3], int *wi, int *out) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int i = t;
  wave_argmax(v, i);
  if(lane == 0) {
    wv[w][0] = v;
    wi[w] = i;
  }
  __syncthreads();
  if(t == 0) {
    double bv = wv[0][0];
    int bi = wi[0];
    for(int k = 1; k < (int)(blockDim.x >> 6); ++k)
      if(wv[k][0] > bv || (wv[k][0] == bv && wi[k] < bi)) {
        bv = wv[k][0];
        bi = wi[k];
      }
    *out = bv > 0.0 ? bi : -1;
  }
  __syncthreads();
  return *out;
}

}  // namespace

// One workgroup (256 threads, thread t owns row / column t; n <= 256) per
// block.  Step k: the residual row I_k against the k terms so far, its
// largest unused entry J_k, v_k = row / pivot, the residual column J_k -> u_k,
// the reference's norm estimate (the cross terms over p <= k-2 and the first
// k entries, unconjugated, :826-843) and its stopping test for k >= 1, then
// I_{k+1} = largest unused entry of u_k.  A residual row or column with no
// non-zero candidate ends the iteration with the terms so far (they reproduce
// the block exactly).
__global__ __launch_bounds__(256) void aca_kernel(AcaArgs a) {
  const int b = blockIdx.x, t = threadIdx.x, n = a.n;
  const int lane = t & 63, w = t >> 6;
  const cplx *A = a.A + (size_t)b * n * n;
  cplx *U = a.U + (size_t)b * n * n, *V = a.V + (size_t)b * n * n;
  __shared__ cplx coef[256], uk[256], vk[256];
  __shared__ unsigned char usedR[256], usedC[256];
  __shared__ double wv[4][3];
  __shared__ int wi[4];
  __shared__ int s_idx, s_stop;
  __shared__ cplx s_piv;
  usedR[t] = 0;
  usedC[t] = 0;
  double norma = 0.0;  // thread 0's running estimate
  int ik = 0, r = 0;
  __syncthreads();
  for(int k = 0; k < n; ++k) {
    if(t == 0) usedR[ik] = 1;
    if(t < k) coef[t] = U[(size_t)t * n + ik];  // U(I_k, p)
    __syncthreads();
    cplx rv = mk(0.0);
    if(t < n) {
      cplx s = mk(0.0);
      for(int p = 0; p < k; ++p) s = s + coef[p] * V[(size_t)p * n + t];
      rv = A[(size_t)t * n + ik] - s;
    }
    const int jk = block_argmax((t < n && !usedC[t]) ? cabs(rv) : -1.0, wv, wi, &s_idx);
    if(jk < 0) break;
    if(t == jk) {
      s_piv = rv;
      usedC[jk] = 1;
    }
    __syncthreads();
    cplx v = mk(0.0);
    if(t < n) {
      v = rv / s_piv;
      V[(size_t)k * n + t] = v;
      vk[t] = v;
    }
    if(t < k) coef[t] = V[(size_t)t * n + jk];  // V(p, J_k)
    __syncthreads();
    cplx u = mk(0.0);
    if(t < n) {
      cplx s = mk(0.0);
      for(int p = 0; p < k; ++p) s = s + coef[p] * U[(size_t)p * n + t];
      u = A[(size_t)jk * n + t] - s;
      U[(size_t)k * n + t] = u;
      uk[t] = u;
    }
    __syncthreads();
    double cross = 0.0;
    if(t < k - 1) {  // p = t
      cplx p1 = mk(0.0), p2 = mk(0.0);
      for(int q = 0; q < k; ++q) {
        p1 = p1 + U[(size_t)t * n + q] * uk[q];
        p2 = p2 + V[(size_t)t * n + q] * vk[q];
      }
      cross = cabs(p1) * cabs(p2);
    }
    double su = wave_sum(norm2(u)), sv = wave_sum(norm2(v));
    cross = wave_sum(cross);
    if(lane == 0) {
      wv[w][0] = su;
      wv[w][1] = sv;
      wv[w][2] = cross;
    }
    __syncthreads();
    if(t == 0) {
      double SU = 0.0, SV = 0.0, CR = 0.0;
      for(int q = 0; q < 4; ++q) {
        SU += wv[q][0];
        SV += wv[q][1];
        CR += wv[q][2];
      }
      const double cn = sqrt(SU), rn = sqrt(SV);
      int stop = 0;
      if(k == 0) {
        norma = cn * cn * (rn * rn);
      } else {
        norma = norma + cn * cn * (rn * rn) + 2.0 * CR;
        stop = a.eps * sqrt(norma) >= cn * rn;
      }
      s_stop = stop;
    }
    __syncthreads();
    r = k + 1;
    if(s_stop) break;
    const int in = block_argmax((t < n && !usedR[t]) ? cabs(u) : -1.0, wv, wi, &s_idx);
    if(in < 0) break;
    ik = in;
  }
  if(t == 0) a.rank[b] = r;
}

__global__ void aca_pack_kernel(const cplx *U, const cplx *V, int n, const int *rank, cplx *pool, const size_t *off) {
  const int b = blockIdx.x;
  const size_t m = (size_t)n * rank[b];
  const cplx *u = U + (size_t)b * n * n, *v = V + (size_t)b * n * n;
  cplx *d = pool + off[b];
  for(size_t e = threadIdx.x; e < m; e += blockDim.x) {
    d[e] = u[e];
    d[m + e] = v[e];
  }
}

__global__ __launch_bounds__(256) void aca_vt_kernel(AcaOp op, const cplx *x) {
  const AcaDesc d = op.desc[op.acalist[blockIdx.x]];
  const int n = op.n, r = d.rank, t = threadIdx.x, lane = t & 63, w = t >> 6;
  __shared__ cplx xs[256];
  if(t < n) xs[t] = x[(size_t)d.j * n + t];
  __syncthreads();
  const cplx *Vb = d.M + (size_t)n * r;
  for(int p = w; p < r; p += 4) {
    cplx s = mk(0.0);
    for(int c = lane; c < n; c += 64) s = s + Vb[(size_t)p * n + c] * xs[c];
    s.x = wave_sum(s.x);
    s.y = wave_sum(s.y);
    if(lane == 0) op.t[d.toff + p] = s;
  }
}

__global__ __launch_bounds__(256) void aca_rows_kernel(AcaOp op, const cplx *x) {
  const int g = blockIdx.x, i = blockIdx.y, q = threadIdx.x, n = op.n;
  __shared__ cplx vs[256];
  cplx acc = mk(0.0);
  for(int e = op.rowptr[i] + g; e < op.rowptr[i + 1]; e += op.G) {
    const AcaDesc d = op.desc[e];
    const int m = d.rank < 0 ? n : d.rank;
    const cplx *src = d.rank < 0 ? x + (size_t)d.j * n : op.t + d.toff;
    __syncthreads();
    if(q < m) vs[q] = src[q];
    __syncthreads();
    if(q < n)
      for(int c = 0; c < m; ++c) acc = acc + d.M[(size_t)c * n + q] * vs[c];
  }
  if(q < n) op.part[((size_t)g * op.P + i) * n + q] = acc;
}

__global__ void aca_sum_kernel(AcaOp op, const cplx *x, cplx *y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= (size_t)op.Np) return;
  cplx s = x[e];
  if(e < (size_t)op.N) {
    const size_t i = e / op.n, q = e % op.n;
    for(int g = 0; g < op.G; ++g) s = s + op.part[((size_t)g * op.P + i) * op.n + q];
  }
  y[e] = s;
}

void run_aca(const AcaArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(aca_kernel, dim3(a.nblk), dim3(256), 0, st, a);
}

void run_aca_pack(const cplx *U, const cplx *V, int n, int nblk, const int *rank, cplx *pool, const size_t *off,
                  hipStream_t st) {
  hipLaunchKernelGGL(aca_pack_kernel, dim3(nblk), dim3(256), 0, st, U, V, n, rank, pool, off);
}

void aca_matvec(const AcaOp &op, const cplx *x, cplx *y, hipStream_t st) {
  if(op.naca > 0) hipLaunchKernelGGL(aca_vt_kernel, dim3(op.naca), dim3(256), 0, st, op, x);
  const int bd = ((op.n + 63) / 64) * 64;
  if(op.G > 0) hipLaunchKernelGGL(aca_rows_kernel, dim3(op.G, op.P), dim3(bd), 0, st, op, x);
  hipLaunchKernelGGL(aca_sum_kernel, dim3((op.Np + 255) / 256), dim3(256), 0, st, op, x, y);
}

}  // namespace omx
