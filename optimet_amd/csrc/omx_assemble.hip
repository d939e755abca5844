// omx_assemble.hip -- coefficient generation and matrix assembly on CDNA4.
//
// Kernels:
//   objects_kernel     Mie T-matrices and auxiliary per-object vectors
//                      (Scatterer.cpp:39-412, Geometry.cpp:86-145)
//   pair_kernel        Stout translation coefficients of one sphere pair
//                      (TranslationAdditionCoefficients.cpp:71-135,
//                      Coupling.cpp:30-87) fused with the block write
//                      S_ij = -T_i [[A^T,B^T],[B^T,A^T]] (PreconditionedMatrix.cpp:350-400)
//   origin_kernel      regular coupling from the origin per object: local
//                      incident field, Q_i = T_i inc_i (PreconditionedMatrix.cpp:1327-1345,
//                      Excitation.cpp:79-129) and the column weights of the
//                      scattering cross section (Result.cpp:579-647)
//   cross_kernel       per-object cross-section partial sums (Result.cpp:557-761)
//   pad_kernel         identity padding of the matrix up to the LU block size
//
// The Stout recurrence decouples into independent "rows": for a fixed
// instance (k or its negative-m twin), fixed m and fixed k = m + kappa, the
// diagonal recurrence (n = m) followed by the off-diagonal recurrence in n
// needs only the neighbours l-1, l+1 of the previous step.  Each row is
// mapped to 2*nmax+1 consecutive lanes of a wave (lane = l); neighbours come
// from cross-lane shuffles, so no LDS traffic or barrier sits inside the
// recurrence.  Every row takes exactly n steps to reach harmonic n.
#include "omx_kernels.hpp"

namespace omx {

// ------------------------------------------------------------------ objects
__device__ inline void riccati_set(cplx r0, cplx rho, int nmax, cplx *psi, cplx *dpsi, cplx *ksi,
                                   cplx *dksi, cplx *psr, cplx *dpsr) {
  cplx jf[kMaxNmax + 3], jd[kMaxNmax + 3], rf[kMaxNmax + 3], rd[kMaxNmax + 3], hf[kMaxNmax + 3],
      hd[kMaxNmax + 3];
  sph_bessel(0, r0, nmax, jf, jd);
  sph_bessel(0, rho * r0, nmax, rf, rd);
  sph_bessel(1, r0, nmax, hf, hd);
  for(int n = 1; n <= nmax; ++n) {
    psi[n] = r0 * jf[n];
    dpsi[n] = r0 * jd[n] + jf[n];
    ksi[n] = r0 * hf[n];
    dksi[n] = r0 * hd[n] + hf[n];
    psr[n] = (r0 * rho) * rf[n];
    dpsr[n] = (r0 * rho) * rd[n] + rf[n];
  }
}

// One thread per object.  Output ov[kind][obj][2][nmax+1] (TE at [0], TM at [1]).
__global__ void objects_kernel(ObjArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= a.P) return;
  if(a.omz) {  // system blockIdx.z of a batch
    const int z = blockIdx.z;
    a.omega = a.omz[z];
    a.eps += (size_t)z * a.P;
    a.eps_sh += (size_t)z * a.P;
    a.ov += z * a.zov;
    if(z > 0) a.sph = nullptr;  // the geometry is common: system 0 writes it
  }
  const int nmax = a.nmax, S = nmax + 1;
  auto OV = [&](int kind, int pol, int n) -> cplx & {
    return a.ov[(((size_t)kind * a.P + i) * 2 + pol) * S + n];
  };
  // spherical coordinates of the centre (Tools.cpp:250-258)
  double x = a.xyz[3 * i], y = a.xyz[3 * i + 1], z = a.xyz[3 * i + 2];
  double r = sqrt(x * x + y * y + z * z);
  if(a.sph) {
    if(r > 0.0) {
      a.sph[3 * i] = r;
      a.sph[3 * i + 1] = acos(z / r);
      a.sph[3 * i + 2] = atan2(y, x);
    } else {
      a.sph[3 * i] = a.sph[3 * i + 1] = a.sph[3 * i + 2] = 0.0;
    }
  }
  const double R = a.radius[i];
  const cplx eps = a.eps[i], mu = a.mu[i], eps_sh = a.eps_sh[i], mu_sh = mu;
  const cplx bge = a.bg_eps, bgm = a.bg_mu;
  cplx psi[kMaxNmax + 1], dpsi[kMaxNmax + 1], ksi[kMaxNmax + 1], dksi[kMaxNmax + 1],
      psr[kMaxNmax + 1], dpsr[kMaxNmax + 1];
  const cplx zeta_b2 = csqrt_(bgm / bge);
  for(int f = 0; f < 2; ++f) {  // f = 0: omega, f = 1: 2 omega
    double om = (f == 0 ? 1.0 : 2.0) * a.omega;
    cplx ks = om * csqrt_((f ? eps_sh : eps) * (f ? mu_sh : mu));
    cplx kb = om * csqrt_(bge * bgm);
    cplx rho = ks / kb, r0 = kb * R;
    cplx mu_s = f ? mu_sh : mu;
    cplx mu_sob = mu_s / bgm;
    riccati_set(r0, rho, nmax, psi, dpsi, ksi, dksi, psr, dpsr);
    cplx zeta_j2 = f ? csqrt_(mu_sh / eps_sh) : csqrt_(mu / eps);
    cplx zb = zeta_b2 / zeta_j2;
    cplx xb = kb * R, xi2 = ks * R;
    for(int n = 1; n <= nmax; ++n) {
      // T (Scatterer.cpp:70-84) / T_SH (:131-145)
      cplx te = (psi[n] / ksi[n]) * (mu_sob * dpsi[n] / psi[n] - rho * dpsr[n] / psr[n]) /
                (rho * dpsr[n] / psr[n] - mu_sob * dksi[n] / ksi[n]);
      cplx tm = (psi[n] / ksi[n]) * (mu_sob * dpsr[n] / psr[n] - rho * dpsi[n] / psi[n]) /
                (rho * dksi[n] / ksi[n] - mu_sob * dpsr[n] / psr[n]);
      OV(f ? OV_TSH : OV_T, 0, n) = te;
      OV(f ? OV_TSH : OV_T, 1, n) = tm;
      if(f == 0) {
        // Iaux (Scatterer.cpp:297-312)
        cplx mu_j = mu, mu_0 = bgm, I = mk(0.0, 1.0);
        OV(OV_IAUX, 0, n) = (mu_j * rho) / (mu_0 * rho * dpsr[n] * psi[n] - mu_j * psr[n] * dpsi[n]) * I;
        OV(OV_IAUX, 1, n) = (mu_j * rho) / (mu_j * psi[n] * dpsr[n] - mu_0 * rho * psr[n] * dpsi[n]) * I;
        // absorption weights (Geometry.cpp:126-140), real part kept in .x
        cplx t1 = I * rho * mu_0 * conj(mu_j) * conj(psr[n]) * dpsr[n];
        double t2 = cabs(mu_j * psr[n] * dpsi[n] - mu_0 * rho * dpsr[n] * psi[n]);
        OV(OV_CABS, 0, n) = mk(t1.x / (t2 * t2));
        t1 = I * conj(rho) * mu_0 * mu_j * conj(psr[n]) * dpsr[n];
        t2 = cabs(mu_0 * rho * psr[n] * dpsi[n] - mu_j * dpsr[n] * psi[n]);
        OV(OV_CABS, 1, n) = mk(t1.x / (t2 * t2));
      } else {
        // IauxSH1 (Scatterer.cpp:343-357), IauxSH2 (:390-404)
        OV(OV_IAUXSH1, 0, n) = (-xi2 * ksi[n]) / (xb * psr[n]);
        OV(OV_IAUXSH1, 1, n) = (-xi2 * dksi[n]) / (xb * dpsr[n]);
        cplx bnpp = (zb * xb * dpsr[n]) / (zb * ksi[n] * dpsr[n] - psr[n] * dksi[n]);
        cplx anpp = (zb * xb * psr[n]) / (zb * psr[n] * dksi[n] - ksi[n] * dpsr[n]);
        OV(OV_IAUXSH2, 0, n) =
            bnpp * ((-xi2 * ksi[n]) / (xb * psr[n]) + (xi2 * dksi[n]) / (zb * xb * dpsr[n]));
        OV(OV_IAUXSH2, 1, n) =
            anpp * ((-xi2 * dksi[n]) / (xb * dpsr[n]) + (xi2 * ksi[n]) / (zb * xb * psr[n]));
        // TSH1_outer (Scatterer.cpp:200-208), TSH2_outer (:250-260)
        OV(OV_TSH1O, 0, n) = -xb * psr[n] / (zb * ksi[n] * dpsr[n] - psr[n] * dksi[n]);
        OV(OV_TSH1O, 1, n) = -xb * dpsr[n] / (zb * psr[n] * dksi[n] - ksi[n] * dpsr[n]);
        OV(OV_TSH2O, 0, n) = zb * xb * dpsr[n] / (zb * ksi[n] * dpsr[n] - psr[n] * dksi[n]);
        OV(OV_TSH2O, 1, n) = zb * xb * psr[n] / (zb * psr[n] * dksi[n] - ksi[n] * dpsr[n]);
      }
    }
  }
  // radial data of the SH source (Symbol.cpp:52-78): j_n(k_j R) and k_j j_n' + j_n / R
  {
    cplx kj = a.omega * csqrt_(eps * mu);
    cplx jf[kMaxNmax + 3], jd[kMaxNmax + 3];
    sph_bessel(0, R * kj, nmax, jf, jd);
    for(int n = 0; n <= nmax; ++n) {
      OV(OV_SHRAD, 0, n) = jf[n];
      OV(OV_SHRAD, 1, n) = kj * jd[n] + jf[n] / R;
    }
  }
}

// ------------------------------------------------------------------ tables
// LDS workspace of one pair (or one object-origin coupling).
template <int NMAX>
struct PairLDS {
  static constexpr int L = 2 * NMAX + 1;           // lanes per recurrence row
  static constexpr int NL = (NMAX + 1) * (NMAX + 1); // (l, K) entries per (n, M), l <= NMAX
  static constexpr int NF = NMAX * (NMAX + 2);
  cplx h[2][L + 1];          // z_l(k R) for the +m and -m instances
  cplx Y[L * L];             // Y_l^k, index l*l + l + k
  cplx T[(2 * NMAX + 1) * NL]; // C(n, M, l, K) of the current n, index (M+n)*NL + l*l+l+K
  cplx Ti[2 * NF];           // row scale -T_i (assembly) or incident vector (origin)
  double w[NF];              // column weights (origin kernel)
};

// Computes z_l and Y_l^k for pair vector R (r, theta, phi) and wavenumber k.
template <int NMAX>
__device__ void pair_setup(PairLDS<NMAX> &s, double Rr, double Rt, double Rp, cplx k, bool hankel) {
  constexpr int L = PairLDS<NMAX>::L;
  int t = threadIdx.x;
  if(t < 2) {
    // +m instance: k; -m instance: conj(k) (regular) or -conj(k) (irregular)
    // (TranslationAdditionCoefficients.h:72-74)
    cplx kk = (t == 0) ? k : (hankel ? -conj(k) : conj(k));
    sph_values_to(hankel ? 1 : 0, Rr * kk, L + 1, s.h[t]);
  } else if(t >= 64 && t < 64 + L) {
    // normalised associated Legendre, column |k| = am, rows l = am .. L-1
    int am = t - 64;
    double x = cos(Rt), sx = sin(Rt);
    double p = 1.0 / sqrt(4.0 * kPi);
    for(int i = 1; i <= am; ++i) p *= -sqrt((2.0 * i + 1.0) / (2.0 * i)) * sx;
    double pm2 = 0.0, pm1 = 0.0;
    cplx e = mk(cos(am * Rp), sin(am * Rp));
    double sg = (am % 2) ? -1.0 : 1.0;
    for(int l = am; l < L; ++l) {
      double pl;
      if(l == am)
        pl = p;
      else if(l == am + 1)
        pl = sqrt(2.0 * am + 3.0) * x * p;
      else {
        double aa = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)am * am));
        double bb = sqrt(((l - 1.0) * (l - 1.0) - (double)am * am) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        pl = aa * (x * pm1 - bb * pm2);
      }
      pm2 = pm1;
      pm1 = pl;
      cplx y = pl * e;
      s.Y[l * l + l + am] = y;
      if(am > 0) s.Y[l * l + l - am] = sg * conj(y);
    }
  }
}

__device__ inline double shfl_d(double v, int src) { return __shfl(v, src, 64); }

// Fills s.T with C(n, M, l, K) for all |M| <= n, l <= NMAX, |K| <= l.
template <int NMAX>
__device__ void build_table(PairLDS<NMAX> &s, int n, bool hankel, const double *__restrict__ coef) {
  constexpr int L = PairLDS<NMAX>::L, NL = PairLDS<NMAX>::NL;
  constexpr int G = 64 / L;  // rows per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int rg = lane / L, l = lane % L;
  const int nrows = (2 * n + 1) * L;
  const int ngroups = (nrows + G - 1) / G;
  for(int i = threadIdx.x; i < (2 * n + 1) * NL; i += blockDim.x) s.T[i] = mk(0.0);
  __syncthreads();
  const double *CAP = coef, *CAM = coef + kCoefStride, *CBP = coef + 2 * kCoefStride,
               *CBM = coef + 3 * kCoefStride, *CIA = coef + 4 * kCoefStride, *CIB = coef + 5 * kCoefStride;
  auto ci = [](int nn, int mm) { return nn * nn + nn + mm; };
  const double s4pi = sqrt(4.0 * kPi);
  for(int g = wave; g < ngroups; g += nw) {
    int row = g * G + rg;
    bool act = (rg < G) && (row < nrows);
    int sm = act ? row / L : 0;
    int inst = (sm <= n) ? 0 : 1;
    int m = inst == 0 ? sm : sm - n;
    int k = act ? (row % L) - NMAX : 0;  // k = m + kappa
    int kap = k - m;
    // D_0(l) = C(0, 0, l, kappa) (TranslationAdditionCoefficients.cpp:102-111)
    cplx cur = mk(0.0), prev2 = mk(0.0);
    if(act && abs(kap) <= l) {
      cplx hb = s.h[inst][l];
      if(l == 0 && kap == 0)
        cur = hb;
      else {
        double fac = s4pi * (((l + kap) % 2 == 0) ? 1.0 : -1.0);
        cur = fac * (s.Y[l * l + l - kap] * hb);
      }
    }
    for(int t = 1; t <= n; ++t) {
      cplx lo, hi;
      lo.x = shfl_d(cur.x, lane - 1);
      lo.y = shfl_d(cur.y, lane - 1);
      hi.x = shfl_d(cur.x, lane + 1);
      hi.y = shfl_d(cur.y, lane + 1);
      if(l == 0) lo = mk(0.0);
      if(l == L - 1) hi = mk(0.0);
      cplx nw_;
      if(t <= m) {
        // diagonal step to (t, t, l, kappa + t) (:113-117)
        int kd = kap + t - 1;
        double bp = (l >= 1 && abs(kd) <= l - 1) ? CBP[ci(l - 1, kd)] : 0.0;
        double bm = (abs(kd) <= l + 1) ? CBM[ci(l + 1, kd)] : 0.0;
        nw_ = (lo * bp + hi * bm) * CIB[ci(t - 1, t - 1)];
        prev2 = mk(0.0);
      } else {
        // off-diagonal step to (t, m, l, k) (:119-124)
        double ap = (l >= 1 && abs(k) <= l - 1) ? CAP[ci(l - 1, k)] : 0.0;
        double amn = (abs(k) <= l + 1) ? CAM[ci(l + 1, k)] : 0.0;
        nw_ = (-(prev2 * CAM[ci(t - 1, m)]) + lo * ap + hi * amn) * CIA[ci(t - 1, m)];
        prev2 = cur;
      }
      cur = act ? nw_ : mk(0.0);
    }
    if(act && l <= NMAX && abs(k) <= l) {
      if(inst == 0) {
        s.T[(m + n) * NL + l * l + l + k] = cur;
      } else {
        int M = -m, K = -k;
        int e = hankel ? (n + M + l + K) : (K + M);
        cplx v = conj(cur);
        if(e % 2) v = -v;
        s.T[(M + n) * NL + l * l + l + K] = v;
      }
    }
  }
  __syncthreads();
}

template <int NMAX>
__device__ inline cplx tget(const PairLDS<NMAX> &s, int n, int M, int l, int K) {
  constexpr int NL = PairLDS<NMAX>::NL;
  if(abs(M) > n || l < 0 || abs(K) > l) return mk(0.0);
  return s.T[(M + n) * NL + l * l + l + K];
}

// Per-row (q = (l, K)) factors of the A / B formulas (Coupling.cpp:30-51).
struct QFac {
  int l, K;
  double lam, mu, b1, b2, g0, g1, g2;
};
__device__ inline QFac qfac(int q) {
  QFac f;
  unflat(q, f.l, f.K);
  int l = f.l, K = f.K;
  f.lam = 1.0 / sqrt((double)(l * (l + 1)));
  f.mu = sqrt((double)(2 * l + 1) / (double)((2 * l - 1) * l * (l + 1)));
  f.b1 = sqrt((double)((l - K) * (l + K + 1)));
  f.b2 = sqrt((double)((l + K) * (l - K + 1)));
  f.g0 = sqrt((double)((l - K) * (l + K)));
  f.g1 = sqrt((double)((l - K) * (l - K - 1)));
  f.g2 = sqrt((double)((l + K) * (l + K - 1)));
  return f;
}

// A(p, q) and B(p, q) for p = (n, M) (object j harmonic) and row q.
template <int NMAX>
__device__ inline void coeff_AB(const PairLDS<NMAX> &s, int n, int M, const QFac &f, cplx &A, cplx &B) {
  double nu = 1.0 / sqrt((double)(n * (n + 1)));
  double a1 = sqrt((double)((n - M) * (n + M + 1))), a2 = sqrt((double)((n + M) * (n - M + 1)));
  int l = f.l, K = f.K;
  cplx t0 = tget(s, n, M, l, K), tp = tget(s, n, M + 1, l, K + 1), tm = tget(s, n, M - 1, l, K - 1);
  A = (0.5 * f.lam * nu) * ((double)(2 * K * M) * t0 + (a1 * f.b1) * tp + (a2 * f.b2) * tm);
  cplx u0 = tget(s, n, M, l - 1, K), up = tget(s, n, M + 1, l - 1, K + 1), um = tget(s, n, M - 1, l - 1, K - 1);
  cplx v = ((double)(2 * M) * f.g0) * u0 + (a1 * f.g1) * up - (a2 * f.g2) * um;
  double fb = -0.5 * f.mu * nu;
  B = mk(-fb * v.y, fb * v.x);  // (0, fb) * v
}

// Pair vector r_i - r_j through Cartesian coordinates (Spherical.h:129-131).
__device__ inline void pair_vector(const double *sph, int i, int j, double &r, double &t, double &p) {
  double ri = sph[3 * i], ti = sph[3 * i + 1], pi_ = sph[3 * i + 2];
  double rj = sph[3 * j], tj = sph[3 * j + 1], pj = sph[3 * j + 2];
  double xi = ri * sin(ti) * cos(pi_), yi = ri * sin(ti) * sin(pi_), zi = ri * cos(ti);
  double xj = rj * sin(tj) * cos(pj), yj = rj * sin(tj) * sin(pj), zj = rj * cos(tj);
  double dx = xi - xj, dy = yi - yj, dz = zi - zj;
  r = sqrt(dx * dx + dy * dy + dz * dz);
  if(r > 0.0) {
    t = acos(dz / r);
    p = atan2(dy, dx);
  } else {
    t = p = 0.0;
  }
}

// ------------------------------------------------------------------ assembly
// One block per unordered pair i <= j (block rows i and j, triangular index
// blockIdx.x), or per listed ordered pair (pair-list mode, one block each, in
// its own slot).  The coefficients are always those of the pair (min, max);
// S_ji with j > i comes from them by the parity of the translation: A_pq(-R) = (-1)^(n+l) A_pq(R), B_pq(-R) = -(-1)^(n+l) B_pq(R)
// for p of degree n and q of degree l, scaled by -T_j (the reference computes
// the ordered pair j, i through its own pair vector, equal to rounding).
template <int NMAX>
__global__ __launch_bounds__(256) void pair_kernel(PairArgs a) {
  constexpr int NF = PairLDS<NMAX>::NF;
  __shared__ PairLDS<NMAX> s;
  const int P = a.P;
  if(a.kz) {  // system blockIdx.z of a batch
    const int z = blockIdx.z;
    a.k = a.kz[z];
    a.ov += z * a.zov;
    if(a.S64) a.S64 += z * a.zS;
    if(a.S32) a.S32 += z * a.zS;
  }
  int i, j;
  size_t x0, y0;
  if(a.plist) {  // block blockIdx.x of a pair list, in its own slot
    const int pid = a.plist[blockIdx.x];
    i = pid / P;
    j = pid % P;
    if(a.S64) a.S64 += blockIdx.x * a.zblk;
    if(a.S32) a.S32 += blockIdx.x * a.zblk;
  } else {  // triangular index k = j (j + 1) / 2 + i, i <= j
    const int k = blockIdx.x;
    int jj = (int)((sqrt(8.0 * k + 1.0) - 1.0) * 0.5);
    while((jj + 1) * (jj + 2) / 2 <= k) ++jj;
    while(jj * (jj + 1) / 2 > k) --jj;
    j = jj;
    i = k - jj * (jj + 1) / 2;
  }
  // the coefficients are those of the pair (ci, cj), ci <= cj; a listed block
  // with i > j is written as the mirror, as the full assembly writes it
  const int ci = min(i, j), cj = max(i, j);
  const bool direct = !a.plist || i < j;
  const bool mirror = a.plist ? i > j : i != j;
  size_t x1, y1;  // the mirrored block S_{cj ci}
  if(a.plist) {
    x0 = y0 = x1 = y1 = 0;
  } else {
    x0 = (size_t)ci * 2 * NF;
    y0 = (size_t)cj * 2 * NF;
    x1 = y0;
    y1 = x0;
  }
  const int t = threadIdx.x;
  const int cg = blockDim.x / NF;  // column groups
  const int q = t % NF, grp = t / NF;
  const bool qact = grp < cg;
  if(i == j) {  // identity diagonal block
    for(int c = grp; c < 2 * NF && qact; c += cg) {
      for(int h = 0; h < 2; ++h) {
        int r = q + h * NF;
        double v = (r == c) ? 1.0 : 0.0;
        size_t idx = (y0 + c) * a.lda + x0 + r;
        if(a.S64) a.S64[idx] = mk(v);
        if(a.S32) a.S32[idx] = make_float2((float)v, 0.0f);
      }
    }
    return;
  }
  // row scale -T_ci
  for(int r = t; r < 2 * NF; r += blockDim.x) {
    int nn, mm;
    unflat(r % NF, nn, mm);
    s.Ti[r] = -a.ov[(((size_t)a.tkind * P + ci) * 2 + (r >= NF)) * (NMAX + 1) + nn];
  }
  double Rr, Rt, Rp;
  pair_vector(a.sph, ci, cj, Rr, Rt, Rp);
  const bool ident = fabs(Rr) < kErrEps;  // Coupling.cpp:82-84
  if(!ident) pair_setup<NMAX>(s, Rr, Rt, Rp, a.k, true);
  __syncthreads();
  QFac f = qfac(q);
  const cplx Tq = s.Ti[q], Tq2 = s.Ti[q + NF];
  // row scale -T_cj of the mirrored block, and this row's parity (-1)^l
  cplx Uq = mk(0.0), Uq2 = mk(0.0);
  if(mirror && qact) {
    Uq = -a.ov[(((size_t)a.tkind * P + cj) * 2 + 0) * (NMAX + 1) + f.l];
    Uq2 = -a.ov[(((size_t)a.tkind * P + cj) * 2 + 1) * (NMAX + 1) + f.l];
  }
  const double sq = (f.l & 1) ? -1.0 : 1.0;
  for(int n = 1; n <= NMAX; ++n) {
    if(!ident) build_table<NMAX>(s, n, true, a.coef);
    if(qact) {
      for(int M = n - grp; M >= -n; M -= cg) {
        int p = flat(n, M);
        cplx A, B;
        if(ident) {
          A = mk(p == q ? 1.0 : 0.0);
          B = mk(0.0);
        } else {
          coeff_AB<NMAX>(s, n, M, f, A, B);
        }
        cplx v00 = Tq * A, v10 = Tq2 * B, v01 = Tq * B, v11 = Tq2 * A;
        size_t c0 = (y0 + p) * a.lda + x0, c1 = (y0 + NF + p) * a.lda + x0;
        if(direct && a.S64) {
          a.S64[c0 + q] = v00;
          a.S64[c0 + NF + q] = v10;
          a.S64[c1 + q] = v01;
          a.S64[c1 + NF + q] = v11;
        }
        if(direct && a.S32) {
          a.S32[c0 + q] = make_float2((float)v00.x, (float)v00.y);
          a.S32[c0 + NF + q] = make_float2((float)v10.x, (float)v10.y);
          a.S32[c1 + q] = make_float2((float)v01.x, (float)v01.y);
          a.S32[c1 + NF + q] = make_float2((float)v11.x, (float)v11.y);
        }
        if(mirror) {
          const double sa = (n & 1) ? -sq : sq;
          const cplx Am = sa * A, Bm = (-sa) * B;
          const cplx w00 = Uq * Am, w10 = Uq2 * Bm, w01 = Uq * Bm, w11 = Uq2 * Am;
          const size_t d0 = (y1 + p) * a.lda + x1, d1 = (y1 + NF + p) * a.lda + x1;
          if(a.S64) {
            a.S64[d0 + q] = w00;
            a.S64[d0 + NF + q] = w10;
            a.S64[d1 + q] = w01;
            a.S64[d1 + NF + q] = w11;
          }
          if(a.S32) {
            a.S32[d0 + q] = make_float2((float)w00.x, (float)w00.y);
            a.S32[d0 + NF + q] = make_float2((float)w10.x, (float)w10.y);
            a.S32[d1 + q] = make_float2((float)w01.x, (float)w01.y);
            a.S32[d1 + NF + q] = make_float2((float)w11.x, (float)w11.y);
          }
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ origin
// Block (j, f): object j, f = 0 at k (local incident field, Q, weights),
// f = 1 at 2k (weights for the second-harmonic scattering cross section).
template <int NMAX>
__global__ __launch_bounds__(256) void origin_kernel(OriginArgs a) {
  constexpr int NF = PairLDS<NMAX>::NF;
  __shared__ PairLDS<NMAX> s;
  const int j = blockIdx.x, fq = blockIdx.y;
  const int t = threadIdx.x;
  const int P = a.P;
  if(a.kz) {  // system blockIdx.z of a batch
    const int z = blockIdx.z;
    a.k = a.kz[z];
    a.ov += z * a.zov;
    a.incloc += z * a.zvec;
    a.Q += z * a.zvec;
    a.w += z * a.zw;
  }
  // incident vector [a_inc; b_inc] in s.Ti
  for(int r = t; r < 2 * NF; r += blockDim.x) s.Ti[r] = a.inc[r];
  for(int r = t; r < NF; r += blockDim.x) s.w[r] = 0.0;
  // Rrel = toSpherical(toCartesian(vR_j) - 0) (Excitation.cpp:81)
  double Rr, Rt, Rp;
  {
    double r = a.sph[3 * j], th = a.sph[3 * j + 1], ph = a.sph[3 * j + 2];
    double x = r * sin(th) * cos(ph), y = r * sin(th) * sin(ph), z = r * cos(th);
    Rr = sqrt(x * x + y * y + z * z);
    Rt = Rr > 0.0 ? acos(z / Rr) : 0.0;
    Rp = Rr > 0.0 ? atan2(y, x) : 0.0;
  }
  const bool ident = fabs(Rr) < kErrEps;
  cplx k = fq == 0 ? a.k : 2.0 * a.k;
  if(!ident) pair_setup<NMAX>(s, Rr, Rt, Rp, k, false);
  __syncthreads();
  const int q = t;
  QFac f = qfac(q < NF ? q : 0);
  cplx acc0 = mk(0.0), acc1 = mk(0.0);
  for(int n = 1; n <= NMAX; ++n) {
    if(!ident) build_table<NMAX>(s, n, false, a.coef);
    if(q < NF) {
      for(int M = n; M >= -n; --M) {
        int p = flat(n, M);
        cplx A, B;
        if(ident) {
          A = mk(p == q ? 1.0 : 0.0);
          B = mk(0.0);
        } else {
          coeff_AB<NMAX>(s, n, M, f, A, B);
        }
        // inc_local[q] = sum_p A(p,q) a_p + B(p,q) b_p ; [nf+q]: B a + A b
        acc0 = acc0 + A * s.Ti[p] + B * s.Ti[NF + p];
        acc1 = acc1 + B * s.Ti[p] + A * s.Ti[NF + p];
        atomicAdd(&s.w[p], norm2(A) + norm2(B));
      }
    }
    __syncthreads();
  }
  if(q < NF) {
    if(fq == 0) {
      int nn, mm;
      unflat(q, nn, mm);
      const cplx te = a.ov[(((size_t)OV_T * P + j) * 2 + 0) * (NMAX + 1) + nn];
      const cplx tm = a.ov[(((size_t)OV_T * P + j) * 2 + 1) * (NMAX + 1) + nn];
      a.incloc[(size_t)j * 2 * NF + q] = acc0;
      a.incloc[(size_t)j * 2 * NF + NF + q] = acc1;
      a.Q[(size_t)j * 2 * NF + q] = te * acc0;
      a.Q[(size_t)j * 2 * NF + NF + q] = tm * acc1;
    }
    a.w[((size_t)fq * P + j) * NF + q] = s.w[q];
  }
}

// ------------------------------------------------------------------ cross sections
// Block j: partial sums [Cext, Csca, Cabs, Csca_SH] of object j.
__global__ void cross_kernel(CrossArgs a) {
  const int j = blockIdx.x, t = threadIdx.x;
  const int nf = a.nmax * (a.nmax + 2), S = a.nmax + 1;
  __shared__ double red[4][256];
  double c[4] = {0, 0, 0, 0};
  for(int q = t; q < nf; q += blockDim.x) {
    int n, m;
    unflat(q, n, m);
    size_t b = (size_t)j * 2 * nf;
    cplx x0 = a.x[b + q], x1 = a.x[b + nf + q];
    cplx i0 = a.incloc[b + q], i1 = a.incloc[b + nf + q];
    c[0] += (conj(i0) * x0 + conj(i1) * x1).x;
    double w = a.w[(size_t)j * nf + q];
    c[1] += w * (norm2(x0) + norm2(x1));
    double ca0 = a.ov[(((size_t)OV_CABS * a.P + j) * 2 + 0) * S + n].x;
    double ca1 = a.ov[(((size_t)OV_CABS * a.P + j) * 2 + 1) * S + n].x;
    double m0 = cabs(x0), m1 = cabs(x1);
    c[2] += m0 * m0 * ca0 + m1 * m1 * ca1;
    if(a.xsh) {
      cplx y0 = a.xsh[b + q], y1 = a.xsh[b + nf + q];
      double w2 = a.w[((size_t)a.P + j) * nf + q];
      c[3] += w2 * (norm2(y0) + norm2(y1));
    }
  }
  for(int e = 0; e < 4; ++e) red[e][t] = c[e];
  __syncthreads();
  for(int st = blockDim.x / 2; st > 0; st >>= 1) {
    if(t < st)
      for(int e = 0; e < 4; ++e) red[e][t] += red[e][t + st];
    __syncthreads();
  }
  if(t < 4) a.out[(size_t)j * 4 + t] = red[t][0];
}

// ------------------------------------------------------------------ padding
// Rows/columns N..Np-1 of an Np x Np matrix: identity on the diagonal, zero
// elsewhere, so the padded system has the same solution in its first N rows.
template <typename T>
__global__ void pad_kernel(T *A, int N, int Np, size_t lda, size_t zs) {
  A += blockIdx.z * zs;
  size_t c = blockIdx.y;
  for(int r = blockIdx.x * blockDim.x + threadIdx.x; r < Np; r += gridDim.x * blockDim.x) {
    bool pad = (c >= (size_t)N) || (r >= N);
    if(!pad) continue;
    T v{};
    if((size_t)r == c) v.x = 1;
    A[c * lda + r] = v;
  }
}

// ------------------------------------------------------------------ launchers
template <int NMAX>
static void launch_pair(const PairArgs &a, hipStream_t st, int nb) {
  hipLaunchKernelGGL(pair_kernel<NMAX>, dim3(a.plist ? a.nlist : a.P * (a.P + 1) / 2, 1, nb), dim3(256), 0, st, a);
}
template <int NMAX>
static void launch_origin(const OriginArgs &a, hipStream_t st, int nb) {
  hipLaunchKernelGGL(origin_kernel<NMAX>, dim3(a.P, 2, nb), dim3(256), 0, st, a);
}

#define OMX_NMAX_SWITCH(nmax, F, ...) \
  switch(nmax) {                      \
  case 1: F<1>(__VA_ARGS__); break;   \
  case 2: F<2>(__VA_ARGS__); break;   \
  case 3: F<3>(__VA_ARGS__); break;   \
  case 4: F<4>(__VA_ARGS__); break;   \
  case 5: F<5>(__VA_ARGS__); break;   \
  case 6: F<6>(__VA_ARGS__); break;   \
  case 7: F<7>(__VA_ARGS__); break;   \
  case 8: F<8>(__VA_ARGS__); break;   \
  case 9: F<9>(__VA_ARGS__); break;   \
  case 10: F<10>(__VA_ARGS__); break; \
  default: break;                     \
  }

void run_objects(const ObjArgs &a, hipStream_t st, int nb) {
  hipLaunchKernelGGL(objects_kernel, dim3((a.P + 63) / 64, 1, nb), dim3(64), 0, st, a);
}
void run_pairs(const PairArgs &a, hipStream_t st, int nb) { OMX_NMAX_SWITCH(a.nmax, launch_pair, a, st, nb); }
void run_origin(const OriginArgs &a, hipStream_t st, int nb) { OMX_NMAX_SWITCH(a.nmax, launch_origin, a, st, nb); }
void run_cross(const CrossArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(cross_kernel, dim3(a.P), dim3(256), 0, st, a);
}
void run_pad64(cplx *A, int N, int Np, size_t lda, hipStream_t st, int nb, size_t zs) {
  if(Np == N) return;
  hipLaunchKernelGGL(pad_kernel<double2>, dim3((Np + 255) / 256, Np, nb), dim3(256), 0, st,
                     reinterpret_cast<double2 *>(A), N, Np, lda, zs);
}
void run_pad32(float2 *A, int N, int Np, size_t lda, hipStream_t st, int nb, size_t zs) {
  if(Np == N) return;
  hipLaunchKernelGGL(pad_kernel<float2>, dim3((Np + 255) / 256, Np, nb), dim3(256), 0, st, A, N, Np, lda, zs);
}

// Coefficient table: a_plus, a_minus, b_plus, b_minus, 1/a_plus, 1/b_plus at
// index n*n + n + m for n <= kCoefN.
void fill_coef(double *h) {
  for(int n = 0; n <= kCoefN; ++n)
    for(int m = -n; m <= n; ++m) {
      int i = n * n + n + m;
      h[i] = a_plus(n, m);
      h[kCoefStride + i] = a_minus(n, m);
      h[2 * kCoefStride + i] = b_plus(n, m);
      h[3 * kCoefStride + i] = b_minus(n, m);
      h[4 * kCoefStride + i] = a_plus(n, m) != 0.0 ? 1.0 / a_plus(n, m) : 0.0;
      h[5 * kCoefStride + i] = b_plus(n, m) != 0.0 ? 1.0 / b_plus(n, m) : 0.0;
    }
}

}  // namespace omx
