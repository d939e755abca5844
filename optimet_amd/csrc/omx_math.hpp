// omx_math.hpp -- device math for the MI355X multiple-scattering engine.
//
// Complex FP64 arithmetic, spherical Bessel / Hankel functions, normalised
// spherical harmonics and the Stout recurrence coefficients.  Algorithms:
//   * h_n^(1)(z): closed forms for n = 0, 1, then the upward three-term
//     recurrence h_{n+1} = (2n+1)/z h_n - h_{n-1} (dominant solution, stable);
//   * j_n(z): Miller's downward recurrence, rescaled against overflow and
//     normalised by least squares against the closed forms j_0 = sin z / z and
//     j_1 = sin z / z^2 - cos z / z;
//   * Y_l^m: normalised associated-Legendre recurrences with the
//     Condon-Shortley phase.
// These replace the AMOS zbesj/zbesh calls of Bessel.h:58-143 and the
// boost::math::spherical_harmonic call of TranslationAdditionCoefficients.cpp:64-68.
#pragma once
#include <hip/hip_runtime.h>

namespace omx {

constexpr int kMaxNmax = 10;                 // largest harmonic order supported
constexpr int kMaxL = 2 * kMaxNmax + 1;      // recurrence domain l <= 2 nmax (+1 guard)
constexpr double kPi = 3.14159265358979323846;
constexpr double kC0 = 299792458.0;
constexpr double kMu0 = 4.0 * kPi * 1e-7;           // constants.cpp:19-22
constexpr double kEps0 = 1.0 / (kMu0 * kC0 * kC0);
constexpr double kErrEps = 1e-10;

struct cplx {
  double x, y;
};

__host__ __device__ inline cplx mk(double a, double b = 0.0) { return cplx{a, b}; }
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cplx operator-(cplx a) { return {-a.x, -a.y}; }
__host__ __device__ inline cplx operator*(cplx a, cplx b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline cplx operator*(double s, cplx a) { return {s * a.x, s * a.y}; }
__host__ __device__ inline cplx operator*(cplx a, double s) { return {s * a.x, s * a.y}; }
__host__ __device__ inline cplx conj(cplx a) { return {a.x, -a.y}; }
__host__ __device__ inline double norm2(cplx a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline double cabs(cplx a) { return hypot(a.x, a.y); }
// Smith's algorithm: no intermediate overflow for large operands.
__host__ __device__ inline cplx operator/(cplx a, cplx b) {
  if(fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return {(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  }
  double r = b.x / b.y, d = b.x * r + b.y;
  return {(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}
__host__ __device__ inline cplx operator/(cplx a, double s) { return {a.x / s, a.y / s}; }
__host__ __device__ inline cplx cexp_(cplx a) {
  double e = exp(a.x);
  return {e * cos(a.y), e * sin(a.y)};
}
__host__ __device__ inline cplx csin_(cplx a) { return {sin(a.x) * cosh(a.y), cos(a.x) * sinh(a.y)}; }
__host__ __device__ inline cplx ccos_(cplx a) { return {cos(a.x) * cosh(a.y), -sin(a.x) * sinh(a.y)}; }
// principal square root
__host__ __device__ inline cplx csqrt_(cplx a) {
  if(a.x == 0.0 && a.y == 0.0) return {0.0, a.y};
  double t = sqrt((fabs(a.x) + hypot(a.x, a.y)) * 0.5);
  if(a.x >= 0.0) return {t, a.y / (2.0 * t)};
  return {fabs(a.y) / (2.0 * t), copysign(t, a.y)};
}

// Spherical Bessel (kind 0) or Hankel h^(1) (kind 1) values f[0..nmax+1]
// and derivatives d[0..nmax] with the conventions of Bessel.h:58-143:
// |z| <= 1e-10 gives zeros (j_0 = 1); real z gives real j_n.
template <int CAP>
__host__ __device__ inline void sph_bessel(int kind, cplx z, int nmax, cplx (&f)[CAP], cplx (&d)[CAP]) {
  for(int i = 0; i < CAP; ++i) f[i] = d[i] = mk(0.0);
  double az = hypot(z.x, z.y);
  if(az <= kErrEps) {
    if(kind == 0) f[0] = mk(1.0);
    return;
  }
  const int top = nmax + 1;
  if(kind == 1) {
    cplx eiz = cexp_(mk(-z.y, z.x));
    f[0] = mk(0.0, -1.0) * eiz / z;
    f[1] = -(eiz * (z + mk(0.0, 1.0))) / (z * z);
    for(int n = 1; n < top; ++n) f[n + 1] = ((2.0 * n + 1.0) * f[n]) / z - f[n - 1];
  } else {
    int start = (top > (int)ceil(az) ? top : (int)ceil(az)) + 30 + (int)ceil(4.0 * cbrt(az)) +
                (int)ceil(0.5 * az);
    cplx fp1 = mk(0.0), fn = mk(1e-300);
    for(int n = start; n > top; --n) {
      cplx fm1 = ((2.0 * n + 1.0) * fn) / z - fp1;
      fp1 = fn;
      fn = fm1;
      if(fabs(fn.x) + fabs(fn.y) > 1e200) {
        fn = 1e-200 * fn;
        fp1 = 1e-200 * fp1;
      }
    }
    f[top] = fn;
    cplx nxt = fp1;
    for(int n = top; n >= 1; --n) {
      cplx up = (n == top) ? nxt : f[n + 1];
      f[n - 1] = ((2.0 * n + 1.0) * f[n]) / z - up;
      if(fabs(f[n - 1].x) + fabs(f[n - 1].y) > 1e200) {
        for(int k = n - 1; k <= top; ++k) f[k] = 1e-200 * f[k];
        nxt = 1e-200 * nxt;
      }
    }
    double mx = fmax(cabs(f[0]), cabs(f[1]));
    for(int n = 0; n <= top; ++n) f[n] = f[n] / mx;
    cplx s = csin_(z), c = ccos_(z);
    cplx j0 = s / z, j1 = s / (z * z) - c / z;
    double w = norm2(f[0]) + norm2(f[1]);
    cplx sc = (j0 * conj(f[0]) + j1 * conj(f[1])) / w;
    for(int n = 0; n <= top; ++n) f[n] = sc * f[n];
  }
  for(int i = 0; i <= nmax; ++i) d[i] = -f[i + 1] + (((double)i) * f[i]) / z;
  if(kind == 0 && z.y == 0.0) {
    for(int i = 0; i <= top; ++i) f[i].y = 0.0;
    for(int i = 0; i <= nmax; ++i) d[i].y = 0.0;
  }
}

// Values only, written straight to memory (LDS) by one thread, without local
// arrays: out[0..nv) = j_l(z) (kind 0) or h_l^(1)(z) (kind 1), same
// algorithms and edge rules as sph_bessel.
__device__ inline void sph_values_to(int kind, cplx z, int nv, cplx *out) {
  double az = hypot(z.x, z.y);
  if(az <= kErrEps) {
    for(int i = 0; i < nv; ++i) out[i] = mk(0.0);
    if(kind == 0) out[0] = mk(1.0);
    return;
  }
  if(kind == 1) {
    cplx eiz = cexp_(mk(-z.y, z.x));
    cplx h0 = mk(0.0, -1.0) * eiz / z;
    cplx h1 = -(eiz * (z + mk(0.0, 1.0))) / (z * z);
    out[0] = h0;
    if(nv > 1) out[1] = h1;
    for(int n = 1; n + 1 < nv; ++n) {
      cplx h2 = ((2.0 * n + 1.0) * h1) / z - h0;
      out[n + 1] = h2;
      h0 = h1;
      h1 = h2;
    }
    return;
  }
  const int top = nv - 1;
  int start = ((top + 1) > (int)ceil(az) ? (top + 1) : (int)ceil(az)) + 30 + (int)ceil(4.0 * cbrt(az)) +
              (int)ceil(0.5 * az);
  cplx fp1 = mk(0.0), fn = mk(1e-300);
  for(int n = start; n > top; --n) {
    cplx fm1 = ((2.0 * n + 1.0) * fn) / z - fp1;
    fp1 = fn;
    fn = fm1;
    if(fabs(fn.x) + fabs(fn.y) > 1e200) {
      fn = 1e-200 * fn;
      fp1 = 1e-200 * fp1;
    }
  }
  out[top] = fn;
  cplx up = fp1, cur = fn;
  for(int n = top; n >= 1; --n) {
    cplx lo = ((2.0 * n + 1.0) * cur) / z - up;
    out[n - 1] = lo;
    up = cur;
    cur = lo;
    if(fabs(lo.x) + fabs(lo.y) > 1e200) {
      for(int k = n - 1; k <= top; ++k) out[k] = 1e-200 * out[k];
      up = 1e-200 * up;
      cur = 1e-200 * cur;
    }
  }
  cplx f0 = out[0], f1 = nv > 1 ? out[1] : mk(0.0);
  double mx = fmax(cabs(f0), cabs(f1));
  f0 = f0 / mx;
  f1 = f1 / mx;
  cplx s = csin_(z), c = ccos_(z);
  cplx j0 = s / z, j1 = s / (z * z) - c / z;
  double w = norm2(f0) + norm2(f1);
  cplx sc = ((j0 * conj(f0) + j1 * conj(f1)) / w) / mx;
  for(int n = 0; n <= top; ++n) {
    cplx v = sc * out[n];
    if(z.y == 0.0) v.y = 0.0;
    out[n] = v;
  }
}

// Stout recurrence coefficients (TranslationAdditionCoefficients.cpp:27-61);
// zero outside n >= 0, |m| <= n.
__host__ __device__ inline bool valid2(int n, int m) { return n >= 0 && abs(m) <= n; }
__host__ __device__ inline double a_plus(int n, int m) {
  if(!valid2(n, m)) return 0.0;
  return -sqrt((double)((n + m + 1) * (n - m + 1)) / (double)((2 * n + 1) * (2 * n + 3)));
}
__host__ __device__ inline double a_minus(int n, int m) {
  if(!valid2(n, m)) return 0.0;
  return sqrt((double)((n + m) * (n - m)) / (double)((2 * n + 1) * (2 * n - 1)));
}
__host__ __device__ inline double b_plus(int n, int m) {
  if(!valid2(n, m)) return 0.0;
  return sqrt((double)((n + m + 2) * (n + m + 1)) / (double)((2 * n + 1) * (2 * n + 3)));
}
__host__ __device__ inline double b_minus(int n, int m) {
  if(!valid2(n, m)) return 0.0;
  return sqrt((double)((n - m) * (n - m - 1)) / (double)((2 * n + 1) * (2 * n - 1)));
}

// Harmonic flat index p = n(n+1) - m - 1 (CompoundIterator.h:24-26).
__host__ __device__ inline int flat(int n, int m) { return n * (n + 1) - m - 1; }
__host__ __device__ inline void unflat(int p, int &n, int &m) {
  n = (int)sqrt(p + 1.0);
  m = -(p + 1) + n * (n + 1);
}

}  // namespace omx
