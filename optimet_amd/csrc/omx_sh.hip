// omx_sh.hip -- second-harmonic source on the device.
//
//   cg_kernel      the nine Clebsch-Gordan / Wigner tables of
//                  Geometry::Coefficients (Symbol.cpp:1036-1451), one thread
//                  per (k, p, q) entry, Wigner 3j / 6j by Racah's single-sum
//                  formulas and 9j by the 6j-product sum (Edmonds 6.4.3)
//   shsrc_kernel   the bilinear source of Geometry::getIncLocalSH
//                  (Geometry.cpp:250-310; Symbol.cpp:155-353) for every
//                  object from conj(x_int), followed by source_vectorSH and
//                  source_vectorSH_K1ana (PreconditionedMatrix.cpp:1347-1436)
//   abssh_kernel   the SH absorption integrand of symbol::ACSshcoeff
//                  (Symbol.cpp:358-470) per object and output harmonic,
//                  reduced per object by abssh_sum_kernel into
//                  Result::getAbsorptionCrossSection_SH (Result.cpp:763-794)
#include "omx_kernels.hpp"
#include "omx_sh.hpp"

namespace omx {

__constant__ double c_fact[171];

__device__ inline bool tri(int a, int b, int c) { return a >= 0 && b >= 0 && c >= 0 && c >= abs(a - b) && c <= a + b; }
__device__ inline double fct(int n) { return c_fact[n]; }
__device__ inline double delta3(int a, int b, int c) {
  return sqrt(fct(a + b - c) * fct(a - b + c) * fct(-a + b + c) / fct(a + b + c + 1));
}
__device__ double w3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if(m1 + m2 + m3 != 0 || !tri(j1, j2, j3)) return 0.0;
  if(abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
  int tmin = max(0, max(j2 - j3 - m1, j1 - j3 + m2));
  int tmax = min(j1 + j2 - j3, min(j1 - m1, j2 + m2));
  double s = 0.0;
  for(int t = tmin; t <= tmax; ++t) {
    double d = fct(t) * fct(j3 - j2 + t + m1) * fct(j3 - j1 + t - m2) * fct(j1 + j2 - j3 - t) * fct(j1 - t - m1) *
               fct(j2 - t + m2);
    s += ((t & 1) ? -1.0 : 1.0) / d;
  }
  double pre = delta3(j1, j2, j3) *
               sqrt(fct(j1 + m1) * fct(j1 - m1) * fct(j2 + m2) * fct(j2 - m2) * fct(j3 + m3) * fct(j3 - m3));
  int e = j1 - j2 - m3;
  return ((e % 2 + 2) % 2 ? -1.0 : 1.0) * pre * s;
}
__device__ double w6j(int j1, int j2, int j3, int j4, int j5, int j6) {
  if(!tri(j1, j2, j3) || !tri(j1, j5, j6) || !tri(j4, j2, j6) || !tri(j4, j5, j3)) return 0.0;
  int a1 = j1 + j2 + j3, a2 = j1 + j5 + j6, a3 = j4 + j2 + j6, a4 = j4 + j5 + j3;
  int b1 = j1 + j2 + j4 + j5, b2 = j2 + j3 + j5 + j6, b3 = j3 + j1 + j6 + j4;
  int tmin = max(max(a1, a2), max(a3, a4)), tmax = min(b1, min(b2, b3));
  double s = 0.0;
  for(int t = tmin; t <= tmax; ++t) {
    double d = fct(t - a1) * fct(t - a2) * fct(t - a3) * fct(t - a4) * fct(b1 - t) * fct(b2 - t) * fct(b3 - t);
    s += ((t & 1) ? -1.0 : 1.0) * fct(t + 1) / d;
  }
  return delta3(j1, j2, j3) * delta3(j1, j5, j6) * delta3(j4, j2, j6) * delta3(j4, j5, j3) * s;
}
__device__ double w9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  if(!tri(a, b, c) || !tri(d, e, f) || !tri(g, h, i) || !tri(a, d, g) || !tri(b, e, h) || !tri(c, f, i)) return 0.0;
  int xmin = max(abs(a - i), max(abs(b - f), abs(d - h)));
  int xmax = min(a + i, min(b + f, d + h));
  double s = 0.0;
  for(int x = xmin; x <= xmax; ++x) s += (2 * x + 1) * w6j(a, b, c, f, i, x) * w6j(d, e, f, b, x, h) * w6j(g, h, i, x, a, d);
  return s;
}
// Symbol.cpp:44-49
__device__ inline double clegor(int j, int m, int j1, int m1, int j2, int m2) {
  return (((m + j1 - j2) % 2 + 2) % 2 ? -1.0 : 1.0) * sqrt(2.0 * j + 1.0) * w3j(j1, j2, j, m1, m2, -m);
}
// Symbol.cpp:143-151
__device__ inline double Wf(int L1, int J1, int M1, int L2, int J2, int M2, int L, int M) {
  double sg = (((J2 + L1 + L) % 2 + 2) % 2) ? -1.0 : 1.0;
  return sg *
         sqrt((2.0 * J1 + 1.0) * (2.0 * J2 + 1.0) * (2.0 * L1 + 1.0) * (2.0 * L2 + 1.0) / (4.0 * kPi * (2.0 * L + 1.0))) *
         w6j(L1, L2, L, J2, J1, 1) * clegor(L, 0, L1, 0, L2, 0) * clegor(L, M, J1, M1, J2, M2);
}

__global__ void cg_kernel(int nf, double *out) {
  const size_t sz = (size_t)nf * nf * nf;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(idx >= sz) return;
  int kk = (int)(idx / ((size_t)nf * nf)), p = (int)((idx / nf) % nf), q = (int)(idx % nf);
  int J, M, J1, M1, J2, M2;
  unflat(kk, J, M);
  unflat(p, J1, M1);
  unflat(q, J2, M2);
  // Every table carries CleGor(J, M, J1, M1, J2, M2) or the W factor built on
  // it: entries with M != M1 + M2 vanish.
  if(M != M1 + M2) {
    for(int t = 0; t < 9; ++t) out[t * sz + idx] = 0.0;
    return;
  }
  const double s32 = sqrt(3.0 / 2.0 / kPi);
  const double cg = clegor(J, M, J1, M1, J2, M2);
  const double sJ = sqrt(J / (2.0 * J + 1.0)), sJ1 = sqrt((J + 1) / (2.0 * J + 1.0));
  out[0 * sz + idx] =
      s32 * (2.0 * J1 + 1.0) * cg *
      (sqrt(J2 * (2.0 * J2 - 1.0)) * w9j(J1, J1, 1, J2, J2 - 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1, 0, J2 - 1, 0) * sJ -
       sqrt((J2 + 1.0) * (2.0 * J2 + 3.0)) * w9j(J1, J1, 1, J2, J2 + 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1, 0, J2 + 1, 0) * sJ +
       sqrt(J2 * (2.0 * J2 - 1.0)) * w9j(J1, J1, 1, J2, J2 - 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1, 0, J2 - 1, 0) * sJ1 -
       sqrt((J2 + 1) * (2.0 * J2 + 3.0)) * w9j(J1, J1, 1, J2, J2 + 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1, 0, J2 + 1, 0) * sJ1);
  const double t1 = sqrt((J1 + 1.0) * J2 * (2.0 * J1 - 1.0) * (2.0 * J2 - 1.0));
  const double t2 = sqrt((J1 + 1.0) * (J2 + 1) * (2.0 * J1 - 1.0) * (2.0 * J2 + 3.0));
  const double t3 = sqrt(J1 * J2 * (2.0 * J1 + 3.0) * (2.0 * J2 - 1.0));
  const double t4 = sqrt(J1 * (J2 + 1) * (2.0 * J1 + 3.0) * (2.0 * J2 + 3.0));
  out[1 * sz + idx] =
      s32 * cg *
      (t1 * w9j(J1, J1 - 1, 1, J2, J2 - 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1 - 1, 0, J2 - 1, 0) * sJ -
       t2 * w9j(J1, J1 - 1, 1, J2, J2 + 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1 - 1, 0, J2 + 1, 0) * sJ +
       t3 * w9j(J1, J1 + 1, 1, J2, J2 - 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1 + 1, 0, J2 - 1, 0) * sJ -
       t4 * w9j(J1, J1 + 1, 1, J2, J2 + 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1 + 1, 0, J2 + 1, 0) * sJ +
       t1 * w9j(J1, J1 - 1, 1, J2, J2 - 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1 - 1, 0, J2 - 1, 0) * sJ1 -
       t2 * w9j(J1, J1 - 1, 1, J2, J2 + 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1 - 1, 0, J2 + 1, 0) * sJ1 +
       t3 * w9j(J1, J1 + 1, 1, J2, J2 - 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1 + 1, 0, J2 - 1, 0) * sJ1 -
       t4 * w9j(J1, J1 + 1, 1, J2, J2 + 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1 + 1, 0, J2 + 1, 0) * sJ1);
  out[2 * sz + idx] =
      s32 * (2.0 * J1 + 1.0) * cg *
      (sqrt(J2 * (2.0 * J2 - 1.0)) * w9j(J1, J1, 1, J2, J2 - 1, 1, J, J, 1) * clegor(J, 0, J1, 0, J2 - 1, 0) -
       sqrt((J2 + 1.0) * (2.0 * J2 + 3.0)) * w9j(J1, J1, 1, J2, J2 + 1, 1, J, J, 1) * clegor(J, 0, J1, 0, J2 + 1, 0));
  out[3 * sz + idx] =
      s32 * cg *
      (t1 * w9j(J1, J1 - 1, 1, J2, J2 - 1, 1, J, J, 1) * clegor(J, 0, J1 - 1, 0, J2 - 1, 0) -
       sqrt((J1 + 1.0) * (J2 + 1.0) * (2.0 * J1 - 1.0) * (2.0 * J2 + 3.0)) * w9j(J1, J1 - 1, 1, J2, J2 + 1, 1, J, J, 1) *
           clegor(J, 0, J1 - 1, 0, J2 + 1, 0) +
       t3 * w9j(J1, J1 + 1, 1, J2, J2 - 1, 1, J, J, 1) * clegor(J, 0, J1 + 1, 0, J2 - 1, 0) -
       t4 * w9j(J1, J1 + 1, 1, J2, J2 + 1, 1, J, J, 1) * clegor(J, 0, J1 + 1, 0, J2 + 1, 0));
  const double a1 = sqrt(J1 / (2.0 * J1 + 1.0)), b1 = sqrt((J1 + 1) / (2.0 * J1 + 1.0));
  const double a2 = sqrt(J2 / (2.0 * J2 + 1.0)), b2 = sqrt((J2 + 1) / (2.0 * J2 + 1.0));
  out[4 * sz + idx] = a1 * a2 * Wf(J1 - 1, J1, M1, J2 - 1, J2, M2, J, M) + b1 * b2 * Wf(J1 + 1, J1, M1, J2 + 1, J2, M2, J, M) -
                      a1 * b2 * Wf(J1 - 1, J1, M1, J2 + 1, J2, M2, J, M) - b1 * a2 * Wf(J1 + 1, J1, M1, J2 - 1, J2, M2, J, M);
  out[5 * sz + idx] = b1 * b2 * Wf(J1 - 1, J1, M1, J2 - 1, J2, M2, J, M) + a1 * a2 * Wf(J1 + 1, J1, M1, J2 + 1, J2, M2, J, M) +
                      b1 * a2 * Wf(J1 - 1, J1, M1, J2 + 1, J2, M2, J, M) + a1 * b2 * Wf(J1 + 1, J1, M1, J2 - 1, J2, M2, J, M);
  out[6 * sz + idx] = Wf(J1, J1, M1, J2, J2, M2, J, M);
  out[7 * sz + idx] = b1 * Wf(J1 - 1, J1, M1, J2, J2, M2, J, M) + a1 * Wf(J1 + 1, J1, M1, J2, J2, M2, J, M);
  out[8 * sz + idx] = b2 * Wf(J1, J1, M1, J2 - 1, J2, M2, J, M) + a2 * Wf(J1, J1, M1, J2 + 1, J2, M2, J, M);
}

// Block (object j, output harmonic kk).
__global__ __launch_bounds__(256) void shsrc_kernel(ShArgs a) {
  const int j = blockIdx.x, kk = blockIdx.y, t = threadIdx.x;
  const int nmax = a.nmax, nf = nmax * (nmax + 2), S = nmax + 1, P = a.P;
  const size_t sz = (size_t)nf * nf * nf;
  __shared__ cplx A0[kMaxNmax * (kMaxNmax + 2)], A1[kMaxNmax * (kMaxNmax + 2)], Am1[kMaxNmax * (kMaxNmax + 2)];
  __shared__ cplx red[4][256];
  const double R = a.radius[j];
  const cplx kj = a.omega * csqrt_(a.eps[j] * a.mu[j]);
  const cplx I = mk(0.0, 1.0);
  for(int p = t; p < nf; p += blockDim.x) {
    int n, m;
    unflat(p, n, m);
    cplx c = conj(a.xint[(size_t)j * 2 * nf + p]), d = conj(a.xint[(size_t)j * 2 * nf + nf + p]);
    cplx jn = a.ov[(((size_t)OV_SHRAD * P + j) * 2 + 0) * S + n];
    cplx rd = a.ov[(((size_t)OV_SHRAD * P + j) * 2 + 1) * S + n];
    A0[p] = jn * c;                                                   // Symbol.cpp:52-58
    A1[p] = I * (mk(1.0) / kj) * rd * d;                              // :61-67
    Am1[p] = I * sqrt(n * (n + 1.0)) * ((mk(1.0) / kj) / R) * jn * d;  // :70-78
  }
  __syncthreads();
  cplx sv = mk(0.0), su = mk(0.0), g = mk(0.0), f = mk(0.0);
  const double *tb = a.tab + (size_t)kk * nf * nf;
  // every table entry carries the Clebsch-Gordan factor <J M | J1 M1 J2 M2>,
  // zero unless M1 + M2 = M: only q = (J2, M - M1) is visited for each p
  int nk, mk_;
  unflat(kk, nk, mk_);
  for(int e2 = t; e2 < nf * nmax; e2 += blockDim.x) {
    const int p = e2 / nmax, n2 = e2 % nmax + 1;
    int n1, m1;
    unflat(p, n1, m1);
    const int m2 = mk_ - m1;
    if(m2 < -n2 || m2 > n2) continue;
    const int q = flat(n2, m2), e = p * nf + q;
    double c10 = tb[e], c11 = tb[sz + e], c00 = tb[2 * sz + e], c01 = tb[3 * sz + e];
    double wm = tb[4 * sz + e], w11 = tb[5 * sz + e], w00 = tb[6 * sz + e], w10 = tb[7 * sz + e], w01 = tb[8 * sz + e];
    cplx a1p = A1[p], a0p = A0[p], amq = Am1[q];
    sv = sv + (a1p * amq) * c00 + (a0p * amq) * c01;
    su = su + (a1p * amq) * c10 + (a0p * amq) * c11;
    g = g + (Am1[p] * amq) * wm;
    f = f + (a1p * A1[q]) * w11 + (a0p * A0[q]) * w00 + (a1p * A0[q]) * w10 + (a0p * A1[q]) * w01;
  }
  red[0][t] = sv;
  red[1][t] = su;
  red[2][t] = g;
  red[3][t] = f;
  __syncthreads();
  for(int st = 128; st > 0; st >>= 1) {
    if(t < st)
      for(int e = 0; e < 4; ++e) red[e][t] = red[e][t] + red[e][t + st];
    __syncthreads();
  }
  if(t == 0) {
    int n, m;
    unflat(kk, n, m);
    const cplx mu0 = mk(4.0 * kPi * 1e-7);
    const cplx eps0 = mk(1.0 / (4.0 * kPi * 1e-7 * kC0 * kC0));
    const cplx eta = csqrt_(a.bg_mu / a.bg_eps) / csqrt_(mu0 / eps0);
    const cplx k01 = a.omega * csqrt_(eps0 * mu0);
    const cplx kpar = a.ksiparppar[j], kppp = a.ksippp[j], gam = a.gamma[j], epsj2 = a.eps_sh[j];
    cplx vp = red[0][0] * mk(2.0) * kpar * eta;                                   // Symbol.cpp:274-275
    cplx up = red[1][0] * mk(2.0) * I * kpar * eta;                               // :212-213
    double snn = sqrt((double)(n * (n + 1)));
    cplx upp = I * kppp * snn * red[2][0] / k01 / R +                              // :347-351
               I * gam * (eps0 / epsj2) * snn * (red[2][0] + red[3][0]) / k01 / R;
    auto ovv = [&](int kind, int pol) { return a.ov[(((size_t)kind * P + j) * 2 + pol) * S + n]; };
    size_t b = (size_t)j * 2 * nf;
    a.K[b + kk] = ovv(OV_TSH1O, 0) * vp + ovv(OV_TSH2O, 0) * mk(0.0);
    a.K[b + nf + kk] = ovv(OV_TSH1O, 1) * up + ovv(OV_TSH2O, 1) * upp;
    a.K1[b + kk] = ovv(OV_IAUXSH2, 0) * mk(0.0);
    a.K1[b + nf + kk] = ovv(OV_IAUXSH2, 1) * upp;
  }
}

// One workgroup per (object j, output harmonic kk): the 4-point
// Gauss-Legendre radial integral over [0, R] of |E(2 omega)|^2, the field of
// the bulk (gamma) source of the fundamental internal field (x_int,
// unconjugated) plus the internal SH field (x_int_sh).  Threads take the
// (p, q) pairs of the bilinear sums for all four radii at once; the Bessel
// arrays of the four radii sit in LDS.  The reference's literal forms are
// kept: the second derivative of bessel3der (Bessel.h:230-231, i in place of
// i + 1) and the SH X(+1) term j_n / r + j_n'(K r) without the factor K
// (Symbol.cpp:460).
constexpr int kAbsNQ = 4;
__device__ const double kAbsXi[kAbsNQ] = {-0.3399810435848563, 0.3399810435848563, -0.8611363115940526,
                                          0.8611363115940526};

// One thread per (object j, radius i, system z): the Bessel values of the
// integrand at the radius, for every order, into the object's table
// ([5][NQ][CAP]: j, j', j'' at k_j r, then j, j' at k_SH r).
__global__ void abssh_bessel_kernel(AbsShArgs a) {
  const int z = blockIdx.z, e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= a.P * kAbsNQ) return;
  const int j = e / kAbsNQ, i = e % kAbsNQ;
  constexpr int CAP = kMaxNmax + 2;
  const int nmax = a.nmax, S = nmax + 1;
  const double omega = a.omz ? a.omz[z] : a.omega;
  const cplx *eps = a.eps + (size_t)z * a.P, *eps_sh = a.eps_sh + (size_t)z * a.P;
  cplx *tb = a.bes + ((size_t)z * a.P + j) * kAbsBes;
  const double R = a.radius[j];
  const cplx kj = omega * csqrt_(eps[j] * a.mu[j]);
  const cplx ksh = (2.0 * omega) * csqrt_(eps_sh[j] * a.mu[j]);
  const double r = (R / 2.0) * kAbsXi[i] + R / 2.0;
  cplx f[CAP], d[CAP];
  const cplx x = r * kj;
  sph_bessel(0, x, nmax, f, d);
  for(int n = 0; n < S; ++n) {
    tb[(0 * kAbsNQ + i) * CAP + n] = f[n];
    tb[(1 * kAbsNQ + i) * CAP + n] = d[n];
    cplx v = mk(0.0);
    if(n >= 1 && hypot(x.x, x.y) > kErrEps) {
      v = -(mk((double)n) / x) * d[n] + (mk((double)n) / (x * x)) * f[n] + d[n - 1];
      if(x.y == 0.0) v.y = 0.0;
    }
    tb[(2 * kAbsNQ + i) * CAP + n] = v;
  }
  sph_bessel(0, r * ksh, nmax, f, d);
  for(int n = 0; n < S; ++n) {
    tb[(3 * kAbsNQ + i) * CAP + n] = f[n];
    tb[(4 * kAbsNQ + i) * CAP + n] = d[n];
  }
}

__global__ __launch_bounds__(256) void abssh_kernel(AbsShArgs a) {
  const int j = blockIdx.x, kk = blockIdx.y, t = threadIdx.x, z = blockIdx.z;
  const int nmax = a.nmax, nf = nmax * (nmax + 2), S = nmax + 1;
  const size_t sz = (size_t)nf * nf * nf;
  constexpr int NQ = kAbsNQ, CAP = kMaxNmax + 2;
  __shared__ cplx jb[NQ][CAP], djb[NQ][CAP], d2jb[NQ][CAP], jsh[NQ][CAP], djsh[NQ][CAP];
  __shared__ cplx cs[kMaxNmax * (kMaxNmax + 2)], ds[kMaxNmax * (kMaxNmax + 2)];
  __shared__ cplx red[2 * NQ][256];
  const double wi[NQ] = {0.6521451548625461, 0.6521451548625461, 0.3478548451374538, 0.3478548451374538};
  const double omega = a.omz ? a.omz[z] : a.omega;
  const cplx *eps = a.eps + (size_t)z * a.P, *eps_sh = a.eps_sh + (size_t)z * a.P;
  const cplx *xint = a.xint + z * a.zvec, *xintsh = a.xintsh + z * a.zvec;
  const double R = a.radius[j];
  const cplx kj = omega * csqrt_(eps[j] * a.mu[j]);
  const cplx ksh = (2.0 * omega) * csqrt_(eps_sh[j] * a.mu[j]);
  {
    const cplx *tb = a.bes + ((size_t)z * a.P + j) * kAbsBes;
    cplx(*dst[5])[CAP] = {jb, djb, d2jb, jsh, djsh};
    for(int e = t; e < 5 * NQ * S; e += blockDim.x) {
      const int w = e / (NQ * S), i = (e / S) % NQ, n = e % S;
      dst[w][i][n] = tb[(w * NQ + i) * CAP + n];
    }
  }
  for(int p = t; p < nf; p += blockDim.x) {
    cs[p] = xint[(size_t)j * 2 * nf + p];
    ds[p] = xint[(size_t)j * 2 * nf + nf + p];
  }
  __syncthreads();
  int n, m;
  unflat(kk, n, m);
  const double snn = sqrt((double)n * (n + 1));
  const cplx pre = (-mk(kEps0) / eps_sh[j]) * a.gamma[j];
  const cplx ik2 = mk(1.0) / (kj * kj);
  cplx xm1[NQ], xp1[NQ];
  for(int i = 0; i < NQ; ++i) xm1[i] = xp1[i] = mk(0.0);
  const double *W00 = a.tab + 6 * sz, *W11 = a.tab + 5 * sz, *Wm = a.tab + 4 * sz;
  const size_t base = (size_t)kk * nf * nf;
  // the W tables vanish unless M1 + M2 = M (their Clebsch-Gordan factor):
  // only q = (J2, M - M1) is visited for each p
  for(int e2 = t; e2 < nf * nmax; e2 += blockDim.x) {
    const int p = e2 / nmax, n2 = e2 % nmax + 1;
    int n1, m1;
    unflat(p, n1, m1);
    const int m2 = m - m1;
    if(m2 < -n2 || m2 > n2) continue;
    const int q = flat(n2, m2), e = p * nf + q;
    const double w00 = W00[base + e], w11 = W11[base + e], wm = Wm[base + e];
    const double s4 = sqrt((double)n1 * n2 * (n1 + 1) * (n2 + 1));
    const cplx cc = cs[p] * cs[q], dd = ds[p] * ds[q];
#pragma unroll
    for(int i = 0; i < NQ; ++i) {
      const double r = (R / 2.0) * kAbsXi[i] + R / 2.0, ir = 1.0 / r;
      const cplx j1 = jb[i][n1], j2 = jb[i][n2], dj1 = djb[i][n1], dj2 = djb[i][n2];
      // F_d00, F_d11, F_dm1m1, F_00, F_11, F_m1m1 (Symbol.cpp:80-141)
      const cplx fd00 = j1 * dj2 * kj + kj * dj1 * j2;
      const cplx b1 = kj * dj1 + j1 * ir, b2 = kj * dj2 + j2 * ir;
      const cplx a1 = kj * kj * d2jb[i][n1] - (ir * ir) * j1 + ir * kj * dj1;
      const cplx a2 = kj * kj * d2jb[i][n2] - (ir * ir) * j2 + ir * kj * dj2;
      const cplx fd11 = a1 * b2 + a2 * b1;
      const cplx fdm = (ir * ir) * (kj * dj1 * j2 + j1 * kj * dj2) - (2.0 * ir * ir * ir) * (j1 * j2);
      const cplx f00 = j1 * dj2, f11 = b2 * b1, fm = (ir * ir) * (j1 * j2);
      xm1[i] = xm1[i] + pre * (cc * w00 * fd00 + dd * ik2 * (w11 * fd11 + (wm * s4) * fdm));
      xp1[i] = xp1[i] + pre * (cc * (w00 * snn * ir) * f00 + dd * ik2 * ((w11 * snn * ir) * f11 + (wm * s4 * snn * ir) * fm));
    }
  }
#pragma unroll
  for(int i = 0; i < NQ; ++i) {
    red[i][t] = xm1[i];
    red[NQ + i][t] = xp1[i];
  }
  __syncthreads();
  for(int st = 128; st > 0; st >>= 1) {
    if(t < st)
      for(int e = 0; e < 2 * NQ; ++e) red[e][t] = red[e][t] + red[e][t + st];
    __syncthreads();
  }
  if(t == 0) {
    const cplx k01 = mk(omega * sqrt(kEps0 * kMu0));
    const cplx csh = xintsh[(size_t)j * 2 * nf + kk], dsh = xintsh[(size_t)j * 2 * nf + nf + kk];
    cplx integral = mk(0.0);
    for(int i = 0; i < NQ; ++i) {
      const double r = (R / 2.0) * kAbsXi[i] + R / 2.0;
      const cplx xm1sh = k01 * dsh * (mk(1.0) / ksh) * snn * (1.0 / r) * jsh[i][n];
      const cplx x0 = -(k01 * csh * jsh[i][n]);
      const cplx xp1sh = k01 * dsh * (mk(1.0) / ksh) * ((1.0 / r) * jsh[i][n] + djsh[i][n]);
      const cplx em = red[i][0] + xm1sh, ep = red[NQ + i][0] + xp1sh;
      integral = integral + (wi[i] * (r * r)) * (em * conj(em) + x0 * conj(x0) + ep * conj(ep));
    }
    a.coef[((size_t)z * a.P + j) * nf + kk] = integral * (R / 2.0);
  }
}

// Per object: Re((2 eta) 0.5 sigma_j sum_kk coefABS), harmonics summed in order.
__global__ void abssh_sum_kernel(AbsShArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.z;
  if(j >= a.P) return;
  const int nf = a.nmax * (a.nmax + 2);
  const double omega = a.omz ? a.omz[z] : a.omega;
  cplx s = mk(0.0);
  for(int kk = 0; kk < nf; ++kk) s = s + a.coef[((size_t)z * a.P + j) * nf + kk];
  const cplx eta = csqrt_(a.bg_mu / a.bg_eps);
  const cplx eps_r_sh = a.eps_sh[(size_t)z * a.P + j] / mk(kEps0);
  const cplx sigma = -(mk(0.0, 1.0) * mk(kEps0) * (2.0 * omega)) * (eps_r_sh - mk(1.0));
  a.out[(size_t)z * a.P + j] = ((mk(2.0) * eta) * mk(0.5) * sigma * s).x;
}

void upload_factorials() {
  double h[171];
  h[0] = 1.0;
  for(int i = 1; i < 171; ++i) h[i] = h[i - 1] * i;
  hipMemcpyToSymbol(HIP_SYMBOL(c_fact), h, sizeof(h));
}

void run_cg(int nmax, double *out, hipStream_t st) {
  int nf = nmax * (nmax + 2);
  size_t sz = (size_t)nf * nf * nf;
  hipLaunchKernelGGL(cg_kernel, dim3((unsigned)((sz + 255) / 256)), dim3(256), 0, st, nf, out);
}

void run_abssh(const AbsShArgs &a, hipStream_t st) {
  int nf = a.nmax * (a.nmax + 2);
  hipLaunchKernelGGL(abssh_bessel_kernel, dim3((a.P * kAbsNQ + 63) / 64, 1, a.nb), dim3(64), 0, st, a);
  hipLaunchKernelGGL(abssh_kernel, dim3(a.P, nf, a.nb), dim3(256), 0, st, a);
  hipLaunchKernelGGL(abssh_sum_kernel, dim3((a.P + 63) / 64, 1, a.nb), dim3(64), 0, st, a);
}

void run_shsrc(const ShArgs &a, hipStream_t st) {
  int nf = a.nmax * (a.nmax + 2);
  hipLaunchKernelGGL(shsrc_kernel, dim3(a.P, nf), dim3(256), 0, st, a);
}

}  // namespace omx
