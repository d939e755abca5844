// optimet_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// CPU FP64 restatement of the OPTIMET srcAna hot path.  Each function cites
// the reference file:line whose arithmetic it follows.  The special functions
// that the reference takes from third-party libraries (AMOS zbesj/zbesh,
// boost::math::spherical_harmonic, GSL 3j/6j/9j, Eigen's
// ColPivHouseholderQR) are restated from their published mathematics:
//   * spherical Bessel j_n: Miller downward recurrence normalised by least
//     squares against the closed forms j_0 = sin z/z, j_1 = sin z/z^2 - cos z/z
//     (Abramowitz & Stegun 10.1.11, 10.1.19; Gautschi, SIAM Rev. 9 (1967));
//   * spherical Hankel h_n^(1): closed forms for n = 0, 1 and the upward
//     three-term recurrence (A&S 10.1.11, 10.1.19);
//   * Y_l^m: normalised associated-Legendre recurrences with the
//     Condon-Shortley phase;
//   * Wigner 3j/6j: Racah's single-sum formulas in long double;
//     9j: the 6j-product sum (Edmonds, eq. 6.4.3);
//   * dense solves: LU with partial pivoting (LAPACK zgetf2 algorithm) and
//     column-pivoted Householder QR (Businger-Golub).
// Parity of each piece is pinned by the tests in tests/test_oracle_*.py.
#include "optimet_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <exception>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using cd = std::complex<double>;

// constants.cpp:19-32
const double kPi = 3.14159265358979323846;
const double kC = 299792458;
const double kMu0 = 4.0 * kPi * 1e-7;
const double kEps0 = 1.0 / (kMu0 * kC * kC);
const double kErrEps = 1e-10;

thread_local std::string g_err;

struct OraError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline cd C(const double *p, int i) { return cd(p[2 * i], p[2 * i + 1]); }

// Runs body(i) for i in [0, n) on all OpenMP threads; the first exception
// thrown by any iteration is rethrown after the loop (none may escape a thread).
template <typename F> void parallel_for(int n, F &&body) {
  std::exception_ptr err = nullptr;
#pragma omp parallel for schedule(dynamic, 1)
  for(int i = 0; i < n; ++i) {
    try {
      body(i);
    } catch(...) {
#pragma omp critical(ora_err)
      if(!err) err = std::current_exception();
    }
  }
  if(err) std::rethrow_exception(err);
}

inline void put(double *p, size_t i, cd v) {
  p[2 * i] = v.real();
  p[2 * i + 1] = v.imag();
}

// ---------------------------------------------------------------- geometry
struct Sph {
  double r, the, phi;
};
struct Cart {
  double x, y, z;
};
// Tools.cpp:250-258 / Spherical.h:64-71
Sph to_spherical(Cart c) {
  double r = std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
  if(r > 0.0) return {r, std::acos(c.z / r), std::atan2(c.y, c.x)};
  return {0.0, 0.0, 0.0};
}
// Spherical.h:26-29
Cart to_cartesian(Sph s) {
  return {s.r * std::sin(s.the) * std::cos(s.phi), s.r * std::sin(s.the) * std::sin(s.phi),
          s.r * std::cos(s.the)};
}
// Spherical.h:129-131: difference taken through Cartesian coordinates
Sph sph_minus(Sph a, Sph b) {
  Cart ca = to_cartesian(a), cb = to_cartesian(b);
  return to_spherical({ca.x - cb.x, ca.y - cb.y, ca.z - cb.z});
}

// ---------------------------------------------------------------- Bessel
// Bessel.h:58-143.  Returns data[0..nmax], ddata[0..nmax].
enum { kBessel = 0, kHankel1 = 1 };

void bessel_raw(int type, cd z, int nmax, std::vector<cd> &f) {
  // f[0..nmax+1]
  f.assign(nmax + 2, cd(0, 0));
  const cd I(0, 1);
  if(type == kHankel1) {
    cd eiz = std::exp(I * z);
    f[0] = -I * eiz / z;
    f[1] = -eiz * (z + I) / (z * z);
    for(int n = 1; n <= nmax; ++n) f[n + 1] = (2.0 * n + 1.0) / z * f[n] - f[n - 1];
    return;
  }
  // Miller downward recurrence with rescaling.
  double az = std::abs(z);
  int top = nmax + 1;
  int start = std::max(top, (int)std::ceil(az)) + 30 + (int)std::ceil(4.0 * std::cbrt(az)) +
              (int)std::ceil(0.5 * az);
  cd fp1(0, 0), fn(1e-300, 0);
  // from start down to top+1: keep two values only
  for(int n = start; n > top; --n) {
    cd fm1 = (2.0 * n + 1.0) / z * fn - fp1;
    fp1 = fn;
    fn = fm1;
    if(std::abs(fn) > 1e200) {
      fn *= 1e-200;
      fp1 *= 1e-200;
    }
  }
  // fn is now f[top], fp1 is f[top+1]
  f[top] = fn;
  cd next = fp1;
  for(int n = top; n >= 1; --n) {
    cd fm1 = (2.0 * n + 1.0) / z * f[n] - (n == top ? next : f[n + 1]);
    f[n - 1] = fm1;
    if(std::abs(fm1) > 1e200) {
      for(int k = n - 1; k <= top; ++k) f[k] *= 1e-200;
      next *= 1e-200;
    }
  }
  double mx = std::max(std::abs(f[0]), std::abs(f[1]));
  for(int n = 0; n <= top; ++n) f[n] /= mx;
  cd j0 = std::sin(z) / z;
  cd j1 = std::sin(z) / (z * z) - std::cos(z) / z;
  double w = std::norm(f[0]) + std::norm(f[1]);
  cd s = (j0 * std::conj(f[0]) + j1 * std::conj(f[1])) / w;
  for(int n = 0; n <= top; ++n) f[n] *= s;
}

void bessel(int type, cd z, int nmax, std::vector<cd> &data, std::vector<cd> &ddata) {
  data.assign(nmax + 1, cd(0, 0));
  ddata.assign(nmax + 1, cd(0, 0));
  if(std::abs(z) <= kErrEps) {  // Bessel.h:70-74
    if(type == kBessel) data[0] = cd(1, 0);
  } else {
    std::vector<cd> f;
    bessel_raw(type, z, nmax, f);
    for(int i = 0; i <= nmax; ++i) data[i] = f[i];
    for(int i = 0; i <= nmax; ++i)  // Bessel.h:125-130
      ddata[i] = cd(-1.0, 0.0) * f[i + 1] + ((double)i / z) * f[i];
  }
  if(z.imag() == 0.0 && type == kBessel) {  // Bessel.h:133-140
    for(int i = 0; i <= nmax; ++i) {
      data[i] = data[i].real();
      ddata[i] = ddata[i].real();
    }
  }
}

// ---------------------------------------------------------------- Ynm
// TranslationAdditionCoefficients.cpp:64-68 (boost spherical_harmonic):
// Y_n^m = sqrt((2n+1)/(4pi) (n-m)!/(n+m)!) P_n^m(cos the) e^{i m phi},
// Condon-Shortley phase included, Y_n^{-m} = (-1)^m conj(Y_n^m).
cd ynm(int n, int m, double the, double phi) {
  if(n < 0 || std::abs(m) > n) return 0.0;
  int am = std::abs(m);
  double x = std::cos(the), sx = std::sin(the);
  // normalised P_m^m
  double p = 1.0 / std::sqrt(4.0 * kPi);
  for(int i = 1; i <= am; ++i) p *= -std::sqrt((2.0 * i + 1.0) / (2.0 * i)) * sx;
  double pmm = p;
  if(n > am) {
    double p1 = std::sqrt(2.0 * am + 3.0) * x * pmm;
    double p0 = pmm;
    for(int l = am + 2; l <= n; ++l) {
      double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)am * am));
      double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)am * am) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
      double p2 = a * (x * p1 - b * p0);
      p0 = p1;
      p1 = p2;
    }
    p = p1;
  }
  cd y = p * cd(std::cos(am * phi), std::sin(am * phi));
  if(m < 0) y = ((am % 2) ? -1.0 : 1.0) * std::conj(y);
  return y;
}

// ---------------------------------------------------------------- TAC
// TranslationAdditionCoefficients.cpp:27-61
inline bool valid2(int n, int m) { return n >= 0 && std::abs(m) <= n; }
inline double a_plus(int n, int m) {
  if(!valid2(n, m)) return 0e0;
  return -std::sqrt(static_cast<double>((n + m + 1) * (n - m + 1)) /
                    static_cast<double>((2 * n + 1) * (2 * n + 3)));
}
inline double a_minus(int n, int m) {
  if(!valid2(n, m)) return 0e0;
  return std::sqrt(static_cast<double>((n + m) * (n - m)) /
                   static_cast<double>((2 * n + 1) * (2 * n - 1)));
}
inline double b_plus(int n, int m) {
  if(!valid2(n, m)) return 0e0;
  return std::sqrt(static_cast<double>((n + m + 2) * (n + m + 1)) /
                   static_cast<double>((2 * n + 1) * (2 * n + 3)));
}
inline double b_minus(int n, int m) {
  if(!valid2(n, m)) return 0e0;
  return std::sqrt(static_cast<double>((n - m) * (n - m - 1)) /
                   static_cast<double>((2 * n + 1) * (2 * n - 1)));
}

// details::CachedRecurrence (TranslationAdditionCoefficients.h:34-64,
// .cpp:71-124).  The std::map memo of the reference becomes a dense memo; the
// arithmetic of every entry is unchanged.
class Recurrence {
public:
  Recurrence(Sph R, cd waveK, bool regular, int nmax)
      : dir_(R), k_(waveK), regular_(regular), nmax_(nmax), L_(2 * nmax + 3) {
    size_t sz = (size_t)(nmax + 1) * (nmax + 1) * L_ * (2 * L_ + 1);
    val_.assign(sz, cd(0, 0));
    done_.assign(sz, 0);
    // initial(l,k) needs z_l(kR) for l <= 2*nmax+1
    std::vector<cd> d, dd;
    bessel(regular ? kBessel : kHankel1, dir_.r * k_, L_, d, dd);
    zl_ = d;
  }
  bool is_regular() const { return regular_; }
  cd operator()(int n, int m, int l, int k) {
    if(!(valid2(n, m) && valid2(l, k))) return 0e0;
    if(n > nmax_ || l >= L_) throw OraError("recurrence index outside the memo domain");
    size_t idx = (((size_t)n * (nmax_ + 1) + m) * L_ + l) * (2 * L_ + 1) + (k + L_);
    if(done_[idx]) return val_[idx];
    cd r = recurrence(n, m, l, k);
    val_[idx] = r;
    done_[idx] = 1;
    return r;
  }

private:
  Sph dir_;
  cd k_;
  bool regular_;
  int nmax_, L_;
  std::vector<cd> val_, zl_;
  std::vector<char> done_;

  cd recurrence(int n, int m, int l, int k) {  // .cpp:93-100
    if(n == 0 && m == 0) return initial(l, k);
    if(n == m) return diagonal(n, l, k);
    return offdiagonal(n, m, l, k);
  }
  cd initial(int l, int k) {  // .cpp:102-111
    cd hb = zl_[l];
    if(l == 0 && k == 0) return hb;
    double factor = std::sqrt(4e0 * kPi) * ((l + k) % 2 == 0 ? 1 : -1);
    return factor * ynm(l, -k, dir_.the, dir_.phi) * hb;
  }
  cd diagonal(int n, int l, int k) {  // .cpp:113-117
    return ((*this)(n - 1, n - 1, l - 1, k - 1) * b_plus(l - 1, k - 1) +
            (*this)(n - 1, n - 1, l + 1, k - 1) * b_minus(l + 1, k - 1)) /
           b_plus(n - 1, n - 1);
  }
  cd offdiagonal(int n, int m, int l, int k) {  // .cpp:119-124
    return (-(*this)(n - 2, m, l, k) * a_minus(n - 1, m) +
            (*this)(n - 1, m, l - 1, k) * a_plus(l - 1, k) +
            (*this)(n - 1, m, l + 1, k) * a_minus(l + 1, k)) /
           a_plus(n - 1, m);
  }
};

// TranslationAdditionCoefficients (TranslationAdditionCoefficients.h:70-85,
// .cpp:128-135)
class TAC {
public:
  TAC(Sph R, cd waveK, bool regular, int nmax)
      : pos_(R, waveK, regular, nmax),
        neg_(R, regular ? std::conj(waveK) : -std::conj(waveK), regular, nmax) {}
  cd operator()(int n, int m, int l, int k) {
    if(m >= 0) return pos_(n, m, l, k);
    int sign = neg_.is_regular() ? k + m : n + m + l + k;
    cd r = std::conj(neg_(n, -m, l, -k));
    return sign % 2 == 0 ? r : -r;
  }

private:
  Recurrence pos_, neg_;
};

// ---------------------------------------------------------------- Coupling
inline int flat(int n, int m) { return n * (n + 1) - m - 1; }  // CompoundIterator.h:24-26

// Coupling.cpp:30-38
cd coeff_A(int n, int m, int l, int k, TAC &ta) {
  if(std::abs(k) > l) return 0e0;
  double factor = 0.5 / std::sqrt(l * (l + 1) * n * (n + 1));
  double c0 = 2 * k * m;
  double c1 = std::sqrt((n - m) * (n + m + 1) * (l - k) * (l + k + 1));
  double c2 = std::sqrt((n + m) * (n - m + 1) * (l + k) * (l - k + 1));
  return factor * (c0 * ta(n, m, l, k) + c1 * ta(n, m + 1, l, k + 1) + c2 * ta(n, m - 1, l, k - 1));
}
// Coupling.cpp:40-51
cd coeff_B(int n, int m, int l, int k, TAC &ta) {
  if(std::abs(k) > l) return 0e0;
  double a0 = 2 * l + 1;
  double a1 = (2 * l - 1) * l * (l + 1) * n * (n + 1);
  cd factor(0, -0.5 * std::sqrt(a0 / a1));
  double c0 = static_cast<double>(2 * m) * std::sqrt((l - k) * (l + k));
  double c1 = std::sqrt((n - m) * (n + m + 1) * (l - k) * (l - k - 1));
  double c2 = std::sqrt((n + m) * (n - m + 1) * (l + k) * (l + k - 1));
  return factor * (c0 * ta(n, m, l - 1, k) + c1 * ta(n, m + 1, l - 1, k + 1) -
                   c2 * ta(n, m - 1, l - 1, k - 1));
}

// Coupling::Coupling (Coupling.cpp:53-87).  `regular` has the Coupling
// meaning: true (the default) builds the irregular (Hankel) coefficients.
void coupling(Sph R, cd waveK, int nmax, bool regular, std::vector<cd> &A, std::vector<cd> &B) {
  int nf = nmax * (nmax + 2);
  A.assign((size_t)nf * nf, cd(0, 0));
  B.assign((size_t)nf * nf, cd(0, 0));
  if(std::abs(R.r) < kErrEps) {
    for(int p = 0; p < nf; ++p) A[(size_t)p * nf + p] = 1.0;
    return;
  }
  TAC ta(R, waveK, !regular, nmax);
  for(int n = 1; n <= nmax; ++n)
    for(int m = -n; m <= n; ++m) {
      int p = flat(n, m);
      for(int l = 1; l <= nmax; ++l)
        for(int k = -l; k <= l; ++k) {
          int q = flat(l, k);
          A[(size_t)p * nf + q] = coeff_A(n, m, l, k, ta);
          B[(size_t)p * nf + q] = coeff_B(n, m, l, k, ta);
        }
    }
}

// ---------------------------------------------------------------- problem
struct Object {
  Sph vR;
  double radius;
  cd eps, mu, eps_sh, mu_sh;  // absolute (ElectroMagnetic.cpp:34-54)
  cd eps_r_sh, ksippp, ksiparppar, gamma;
};

struct Problem {
  int nmax, nf, P;
  bool sh;
  std::vector<Object> obj;
  cd bg_eps, bg_mu, bg_eps_r, bg_mu_r;
  Sph vK;       // (k0, theta, phi)
  cd waveK;     // k0 * sqrt(eps_r mu_r) of the background (Excitation.cpp:31-34)
  double omega; // c * k0 (Excitation.h:103)
  cd Einc[3];   // Cartesian projection of (0, Etheta, Ephi) (Reader.cpp:820-827)
  std::vector<cd> ap, bp;
};

// Tools.cpp toProjection: spherical-basis vector at point -> Cartesian.
void to_projection(Sph pt, const cd v[3], cd out[3]) {
  double st = std::sin(pt.the), ct = std::cos(pt.the), sp = std::sin(pt.phi), cp = std::cos(pt.phi);
  out[0] = st * cp * v[0] + ct * cp * v[1] - sp * v[2];
  out[1] = st * sp * v[0] + ct * sp * v[1] + cp * v[2];
  out[2] = ct * v[0] - st * v[1];
}

// AuxCoefficients::VIGdVIG (AuxCoefficients.cpp:216-290)
void vigdvig(unsigned nMax, int m, Sph R, std::vector<double> &W, std::vector<double> &dW) {
  W.assign(nMax + 1, 0.0);
  dW.assign(nMax + 1, 0.0);
  bool neg = m < 0;
  unsigned n_min = static_cast<unsigned>(m = std::abs(m));
  double vig_the = neg ? kPi - R.the : R.the;
  if((std::abs(R.the) < 1e-10) || (std::abs(R.the) - kPi + 1e-10 > 0.0)) vig_the = vig_the + 1e-6;
  double x = std::cos(vig_the);
  double f2m = 1.0, fm = 1.0;
  for(int i = 2; i <= 2 * m; ++i) f2m *= i;
  for(int i = 2; i <= m; ++i) fm *= i;
  W[n_min] = std::pow(2.0, -m) * (std::sqrt(f2m) / fm) * std::pow(1.0 - x, m / 2.0) *
             std::pow(1.0 + x, m / 2.0);
  unsigned s = n_min;
  if(n_min == 0 && nMax > 0) {
    W[1] = x * W[0];
    s = 1;
  }
  unsigned mm = (unsigned)(m * m);
  for(; s < nMax; ++s) {
    W[s + 1] = ((2 * s + 1) * x * W[s] - std::sqrt((double)(s * s - mm)) * W[s - 1]) /
               std::sqrt((double)((s + 1) * (s + 1) - mm));
    dW[s] = (((s * std::sqrt((double)((s + 1) * (s + 1) - mm)) * W[s + 1]) / (2 * s + 1)) -
             (((s + 1) * std::sqrt((double)(s * s * (s * s - mm))) * W[s - 1]) / (s * (2 * s + 1)))) /
            std::sin(vig_the);
  }
  if(nMax > 0) {
    double Wn = ((2 * nMax + 1) * x * W[nMax] - std::sqrt((double)(nMax * nMax - mm)) * W[nMax - 1]) /
                std::sqrt((double)((nMax + 1) * (nMax + 1) - mm));
    dW[nMax] = (((nMax * std::sqrt((double)((nMax + 1) * (nMax + 1) - mm)) * Wn) / (2 * nMax + 1)) -
                (((nMax + 1) * std::sqrt((double)(nMax * nMax * (nMax * nMax - mm))) * W[nMax - 1]) /
                 (nMax * (2 * nMax + 1)))) /
               std::sin(vig_the);
  }
  if(neg) {
    for(unsigned i = 0; i <= nMax; ++i) {
      double c = 1.0 / std::pow(-1.0, i);
      W[i] *= c;
      dW[i] *= -c;
    }
  }
}

// Excitation::populate (Excitation.cpp:48-75) with the B, C vector spherical
// functions of AuxCoefficients (AuxCoefficients.cpp:31-39, 54-99, 292-343).
void populate(Problem &pb) {
  int nmax = pb.nmax, nf = pb.nf;
  pb.ap.assign(nf, 0.0);
  pb.bp.assign(nf, 0.0);
  Sph R{0.0, pb.vK.the, pb.vK.phi};
  std::vector<double> dn(nmax + 1);
  dn[0] = -1000;
  for(int i = 1; i <= nmax; ++i) dn[i] = std::sqrt((2.0 * i + 1.0) / (4.0 * kPi * (i * (i + 1))));
  std::vector<cd> Bp(3 * nf), Cp(3 * nf);
  for(int m = nmax; m >= -nmax; --m) {
    std::vector<double> W, dW;
    vigdvig(nmax, m, R, W, dW);
    for(int n = std::abs(m); n <= nmax; ++n) {
      if(n == 0) continue;
      double A;
      if(m == 0)
        A = 0.0;
      else if(std::abs(R.the) < 1e-10 || (std::abs(R.the) - kPi + 1e-10) > 0.0)
        A = m / std::cos(R.the) * dW[n];
      else
        A = m / std::sin(R.the) * W[n];
      cd cn[3] = {cd(0, 0), cd(0.0, A), cd(-dW[n], 0.0)};
      cd bn[3] = {cd(0, 0), cd(dW[n], 0.0), cd(0.0, A)};
      int p = flat(n, m);
      to_projection(R, cn, &Cp[3 * p]);
      to_projection(R, bn, &Bp[3 * p]);
    }
  }
  const cd I(0, 1);
  for(int p = 0; p < nf; ++p) {
    int n = (int)std::sqrt(p + 1.0);
    int m = -(p + 1) + n * (n + 1);
    cd dotC = std::conj(Cp[3 * p]) * pb.Einc[0] + std::conj(Cp[3 * p + 1]) * pb.Einc[1] +
              std::conj(Cp[3 * p + 2]) * pb.Einc[2];
    cd dotB = std::conj(Bp[3 * p]) * pb.Einc[0] + std::conj(Bp[3 * p + 1]) * pb.Einc[1] +
              std::conj(Bp[3 * p + 2]) * pb.Einc[2];
    // i^n exactly
    auto ipow = [](int e) {
      e = ((e % 4) + 4) % 4;
      return e == 0 ? cd(1, 0) : e == 1 ? cd(0, 1) : e == 2 ? cd(-1, 0) : cd(0, -1);
    };
    double sgn = (m % 2 == 0) ? 1.0 : -1.0;
    cd ph = std::exp(cd(0.0, -1.0) * (double)m * pb.vK.phi);
    pb.ap[p] = 4 * kPi * sgn * ipow(n) * dn[n] * dotC * ph;
    pb.bp[p] = 4 * kPi * sgn * ipow(n - 1) * dn[n] * dotB * ph;
  }
}

Problem make_problem(const ora_problem *in) {
  if(!in || in->nobj < 1 || in->nmax < 1) throw OraError("invalid problem");
  Problem pb;
  pb.nmax = in->nmax;
  pb.nf = in->nmax * (in->nmax + 2);
  pb.P = in->nobj;
  pb.sh = in->sh != 0;
  pb.bg_eps_r = C(in->bg_eps_r, 0);
  pb.bg_mu_r = C(in->bg_mu_r, 0);
  pb.bg_eps = pb.bg_eps_r * kEps0;
  pb.bg_mu = pb.bg_mu_r * kMu0;
  for(int i = 0; i < in->nobj; ++i) {
    Object o;
    o.vR = to_spherical({in->xyz[3 * i], in->xyz[3 * i + 1], in->xyz[3 * i + 2]});
    o.radius = in->radius[i];
    cd er = C(in->eps_r, i), mr = C(in->mu_r, i);
    o.eps = er * kEps0;
    o.mu = mr * kMu0;
    o.eps_r_sh = in->eps_r_sh ? C(in->eps_r_sh, i) : cd(0, 0);
    o.eps_sh = o.eps_r_sh * kEps0;
    o.mu_sh = mr * kMu0;
    o.ksippp = in->ksippp ? C(in->ksippp, i) : cd(0, 0);
    o.ksiparppar = in->ksiparppar ? C(in->ksiparppar, i) : cd(0, 0);
    o.gamma = in->gamma ? C(in->gamma, i) : cd(0, 0);
    pb.obj.push_back(o);
  }
  // Geometry::pushObject overlap rejection (Geometry.cpp:39-51)
  for(int i = 0; i < pb.P; ++i)
    for(int j = 0; j < i; ++j) {
      Cart a = to_cartesian(pb.obj[i].vR), b = to_cartesian(pb.obj[j].vR);
      double d = std::sqrt(std::pow(b.x - a.x, 2.0) + std::pow(b.y - a.y, 2.0) + std::pow(b.z - a.z, 2.0));
      if(d <= pb.obj[i].radius + pb.obj[j].radius) throw OraError("overlapping spheres");
    }
  // Reader.cpp:806-831, Excitation.cpp:31-34
  cd bgcoeff = std::sqrt(pb.bg_eps_r * pb.bg_mu_r);
  pb.vK = {2 * kPi / in->lambda, in->theta_k, in->phi_k};
  pb.waveK = pb.vK.r * bgcoeff;
  pb.omega = kC * pb.vK.r;
  cd Eaux[3] = {cd(0, 0), C(in->etheta, 0), C(in->ephi, 0)};
  to_projection({0.0, pb.vK.the, pb.vK.phi}, Eaux, pb.Einc);
  populate(pb);
  return pb;
}

// ---------------------------------------------------------------- Mie
struct Riccati {
  cd psi, dpsi, ksi, dksi, psirho, dpsirho;
};
// Scatterer.cpp:39-99 etc.: Riccati-Bessel functions at r_0 and rho*r_0.
std::vector<Riccati> riccati(cd r0, cd rho, int nmax) {
  std::vector<cd> Jd, Jdd, Hd, Hdd, Rd, Rdd;
  bessel(kBessel, r0, nmax, Jd, Jdd);
  bessel(kBessel, rho * r0, nmax, Rd, Rdd);
  bessel(kHankel1, r0, nmax, Hd, Hdd);
  std::vector<Riccati> out(nmax + 1);
  for(int n = 1; n <= nmax; ++n) {
    out[n].psi = r0 * Jd[n];
    out[n].dpsi = r0 * Jdd[n] + Jd[n];
    out[n].ksi = r0 * Hd[n];
    out[n].dksi = r0 * Hdd[n] + Hd[n];
    out[n].psirho = r0 * rho * Rd[n];
    out[n].dpsirho = r0 * rho * Rdd[n] + Rd[n];
  }
  return out;
}

// which: 0 T, 1 T_SH, 2 Iaux, 3 IauxSH1, 4 IauxSH2, 5 TSH1_outer, 6 TSH2_outer
std::vector<cd> object_vector(const Problem &pb, const Object &o, int which, const cd &bg_eps,
                              const cd &bg_mu) {
  int nmax = pb.nmax, nf = pb.nf;
  double omega = pb.omega;
  std::vector<cd> res(2 * nf, 0.0);
  bool shf = which != 0 && which != 2;
  cd ks = shf ? 2.0 * omega * std::sqrt(o.eps_sh * o.mu_sh) : omega * std::sqrt(o.eps * o.mu);
  cd kb = shf ? 2.0 * omega * std::sqrt(bg_eps * bg_mu) : omega * std::sqrt(bg_eps * bg_mu);
  cd rho = ks / kb;
  cd r0 = kb * o.radius;
  cd mu_s = shf ? o.mu_sh : o.mu;
  cd mu_sob = mu_s / bg_mu;
  auto rb = riccati(r0, rho, nmax);
  cd zeta_b2 = std::sqrt(bg_mu / bg_eps);
  cd zeta_j2 = shf ? std::sqrt(o.mu_sh / o.eps_sh) : std::sqrt(o.mu / o.eps);
  cd zeta_boj2 = zeta_b2 / zeta_j2;
  cd x_b2 = kb * o.radius;
  cd x_i2 = ks * o.radius;
  const cd I(0, 1);
  for(int n = 1, cur = 0; n <= nmax; cur += 2 * n + 1, ++n) {
    const Riccati &r = rb[n];
    cd te, tm;
    switch(which) {
    case 0:  // Scatterer.cpp:70-84
    case 1:  // Scatterer.cpp:131-145
      te = (r.psi / r.ksi) * (mu_sob * r.dpsi / r.psi - rho * r.dpsirho / r.psirho) /
           (rho * r.dpsirho / r.psirho - mu_sob * r.dksi / r.ksi);
      tm = (r.psi / r.ksi) * (mu_sob * r.dpsirho / r.psirho - rho * r.dpsi / r.psi) /
           (rho * r.dksi / r.ksi - mu_sob * r.dpsirho / r.psirho);
      break;
    case 2: {  // Scatterer.cpp:274-317
      cd mu_j = o.mu, mu_0 = bg_mu;
      te = (mu_j * rho) / (mu_0 * rho * r.dpsirho * r.psi - mu_j * r.psirho * r.dpsi) * I;
      tm = (mu_j * rho) / (mu_j * r.psi * r.dpsirho - mu_0 * rho * r.psirho * r.dpsi) * I;
      break;
    }
    case 3:  // Scatterer.cpp:320-364
      te = (-x_i2 * r.ksi) / (x_b2 * r.psirho);
      tm = (-x_i2 * r.dksi) / (x_b2 * r.dpsirho);
      break;
    case 4: {  // Scatterer.cpp:367-412
      cd bnpp = (zeta_boj2 * x_b2 * r.dpsirho) / (zeta_boj2 * r.ksi * r.dpsirho - r.psirho * r.dksi);
      cd anpp = (zeta_boj2 * x_b2 * r.psirho) / (zeta_boj2 * r.psirho * r.dksi - r.ksi * r.dpsirho);
      te = bnpp * ((-x_i2 * r.ksi) / (x_b2 * r.psirho) + (x_i2 * r.dksi) / (zeta_boj2 * x_b2 * r.dpsirho));
      tm = anpp * ((-x_i2 * r.dksi) / (x_b2 * r.dpsirho) + (x_i2 * r.ksi) / (zeta_boj2 * x_b2 * r.psirho));
      break;
    }
    case 5:  // Scatterer.cpp:200-208
      te = -x_b2 * r.psirho / (zeta_boj2 * r.ksi * r.dpsirho - r.psirho * r.dksi);
      tm = -x_b2 * r.dpsirho / (zeta_boj2 * r.psirho * r.dksi - r.ksi * r.dpsirho);
      break;
    case 6:  // Scatterer.cpp:250-260
      te = zeta_boj2 * x_b2 * r.dpsirho / (zeta_boj2 * r.ksi * r.dpsirho - r.psirho * r.dksi);
      tm = zeta_boj2 * x_b2 * r.psirho / (zeta_boj2 * r.psirho * r.dksi - r.ksi * r.dpsirho);
      break;
    default:
      throw OraError("bad object vector kind");
    }
    for(int i = 0; i < 2 * n + 1; ++i) {
      res[cur + i] = te;
      res[nf + cur + i] = tm;
    }
  }
  return res;
}

// Excitation::getIncLocal (Excitation.cpp:79-129): regular Coupling from the
// origin, T_AB = [[A^T, B^T], [B^T, A^T]], times [a_inc; b_inc].
std::vector<cd> inc_local_k(const Problem &pb, Sph point, cd waveK, const std::vector<cd> &ap,
                            const std::vector<cd> &bp) {
  int nf = pb.nf;
  Sph rrel = sph_minus(point, Sph{0.0, 0.0, 0.0});
  std::vector<cd> A, B;
  coupling(rrel, waveK, pb.nmax, false, A, B);
  std::vector<cd> out(2 * nf, 0.0);
  for(int p = 0; p < nf; ++p) {
    cd s0 = 0.0, s1 = 0.0;
    for(int q = 0; q < nf; ++q) {
      s0 += A[(size_t)q * nf + p] * ap[q] + B[(size_t)q * nf + p] * bp[q];
      s1 += B[(size_t)q * nf + p] * ap[q] + A[(size_t)q * nf + p] * bp[q];
    }
    out[p] = s0;
    out[nf + p] = s1;
  }
  return out;
}

std::vector<cd> source_vector(const Problem &pb) {  // PreconditionedMatrix.cpp:1327-1345
  int nf = pb.nf;
  std::vector<cd> Q((size_t)2 * nf * pb.P);
  for(int i = 0; i < pb.P; ++i) {
    auto inc = inc_local_k(pb, pb.obj[i].vR, pb.waveK, pb.ap, pb.bp);
    auto T = object_vector(pb, pb.obj[i], 0, pb.bg_eps, pb.bg_mu);
    for(int r = 0; r < 2 * nf; ++r) Q[(size_t)i * 2 * nf + r] = T[r] * inc[r];
  }
  return Q;
}

// Block (row object i, column object j) of S (sh = false) or V (sh = true),
// 2nf x 2nf column-major with leading dimension ld:
// S_ij(r, c) = -T_i[r] * [[A^T, B^T],[B^T, A^T]](r, c), S_ii = I
// (preconditioned_scattering_matrix(SH), PreconditionedMatrix.cpp:350-400, 555-610).
void scattering_block(const Problem &pb, bool sh, int i, int j, const std::vector<cd> &T, cd *out, size_t ld) {
  int nf = pb.nf;
  if(i == j) {
    for(int c = 0; c < 2 * nf; ++c)
      for(int r = 0; r < 2 * nf; ++r) out[(size_t)c * ld + r] = r == c ? 1.0 : 0.0;
    return;
  }
  cd k = sh ? 2.0 * pb.waveK : pb.waveK;
  std::vector<cd> A, B;
  coupling(sph_minus(pb.obj[i].vR, pb.obj[j].vR), k, pb.nmax, true, A, B);
  for(int c = 0; c < 2 * nf; ++c)
    for(int r = 0; r < 2 * nf; ++r) {
      int rr = r % nf, cc = c % nf;
      bool diag = (r < nf) == (c < nf);
      cd blk = diag ? A[(size_t)cc * nf + rr] : B[(size_t)cc * nf + rr];
      out[(size_t)c * ld + r] = -T[r] * blk;
    }
}

// The whole N x N matrix, column-major; block rows on all threads.
void scattering_matrix(const Problem &pb, bool sh, std::vector<cd> &S) {
  int nf = pb.nf, P = pb.P;
  size_t N = (size_t)2 * nf * P;
  S.assign(N * N, cd(0, 0));
  parallel_for(P, [&](int i) {
    auto T = object_vector(pb, pb.obj[i], sh ? 1 : 0, pb.bg_eps, pb.bg_mu);
    for(int j = 0; j < P; ++j)
      scattering_block(pb, sh, i, j, T, &S[(size_t)j * 2 * nf * N + (size_t)i * 2 * nf], N);
  });
}

// ---------------------------------------------------------------- solves
void lu_solve(int N, std::vector<cd> A, const cd *b, cd *x) {
  std::vector<int> piv(N);
  for(int j = 0; j < N; ++j) {
    int p = j;
    double best = std::abs(A[(size_t)j * N + j]);
    for(int i = j + 1; i < N; ++i) {
      double v = std::abs(A[(size_t)j * N + i]);
      if(v > best) best = v, p = i;
    }
    if(best == 0.0) throw OraError("singular matrix in LU");
    piv[j] = p;
    if(p != j)
      for(int c = 0; c < N; ++c) std::swap(A[(size_t)c * N + j], A[(size_t)c * N + p]);
    cd d = A[(size_t)j * N + j];
    for(int i = j + 1; i < N; ++i) A[(size_t)j * N + i] /= d;
    for(int c = j + 1; c < N; ++c) {
      cd u = A[(size_t)c * N + j];
      if(u == cd(0, 0)) continue;
      cd *col = &A[(size_t)c * N];
      const cd *lc = &A[(size_t)j * N];
      for(int i = j + 1; i < N; ++i) col[i] -= lc[i] * u;
    }
  }
  std::vector<cd> y(b, b + N);
  for(int j = 0; j < N; ++j) std::swap(y[j], y[piv[j]]);
  for(int j = 0; j < N; ++j)
    for(int i = j + 1; i < N; ++i) y[i] -= A[(size_t)j * N + i] * y[j];
  for(int j = N - 1; j >= 0; --j) {
    y[j] /= A[(size_t)j * N + j];
    for(int i = 0; i < j; ++i) y[i] -= A[(size_t)j * N + i] * y[j];
  }
  for(int i = 0; i < N; ++i) x[i] = y[i];
}

// Column-pivoted Householder QR (the algorithm of Eigen's ColPivHouseholderQR,
// PreconditionedMatrixSolver.h:58, 75).
void qr_solve(int N, std::vector<cd> A, const cd *b, cd *x) {
  std::vector<int> perm(N);
  for(int i = 0; i < N; ++i) perm[i] = i;
  std::vector<double> cn(N);
  for(int c = 0; c < N; ++c) {
    double s = 0;
    for(int i = 0; i < N; ++i) s += std::norm(A[(size_t)c * N + i]);
    cn[c] = s;
  }
  std::vector<cd> tau(N);
  for(int k = 0; k < N; ++k) {
    int best = k;
    for(int c = k + 1; c < N; ++c)
      if(cn[c] > cn[best]) best = c;
    if(best != k) {
      for(int i = 0; i < N; ++i) std::swap(A[(size_t)k * N + i], A[(size_t)best * N + i]);
      std::swap(cn[k], cn[best]);
      std::swap(perm[k], perm[best]);
    }
    cd *col = &A[(size_t)k * N];
    double xnorm = 0;
    for(int i = k + 1; i < N; ++i) xnorm += std::norm(col[i]);
    cd alpha = col[k];
    if(xnorm == 0.0 && alpha.imag() == 0.0) {
      tau[k] = 0.0;
    } else {
      double beta = -std::copysign(std::sqrt(std::norm(alpha) + xnorm), alpha.real());
      tau[k] = cd((beta - alpha.real()) / beta, -alpha.imag() / beta);
      cd scal = 1.0 / (alpha - beta);
      for(int i = k + 1; i < N; ++i) col[i] *= scal;
      col[k] = beta;
    }
    // apply H = I - tau v v^H (v[k] = 1) from the left to the remaining columns
    for(int c = k + 1; c < N; ++c) {
      cd *cc = &A[(size_t)c * N];
      cd s = cc[k];
      for(int i = k + 1; i < N; ++i) s += std::conj(col[i]) * cc[i];
      s *= std::conj(tau[k]);
      cc[k] -= s;
      for(int i = k + 1; i < N; ++i) cc[i] -= s * col[i];
      double nn = 0;
      for(int i = k + 1; i < N; ++i) nn += std::norm(cc[i]);
      cn[c] = nn;
    }
  }
  std::vector<cd> y(b, b + N);
  for(int k = 0; k < N; ++k) {  // y = Q^H b
    const cd *col = &A[(size_t)k * N];
    cd s = y[k];
    for(int i = k + 1; i < N; ++i) s += std::conj(col[i]) * y[i];
    s *= std::conj(tau[k]);
    y[k] -= s;
    for(int i = k + 1; i < N; ++i) y[i] -= s * col[i];
  }
  for(int j = N - 1; j >= 0; --j) {
    y[j] /= A[(size_t)j * N + j];
    for(int i = 0; i < j; ++i) y[i] -= A[(size_t)j * N + i] * y[j];
  }
  for(int i = 0; i < N; ++i) x[perm[i]] = y[i];
}

void dense_solve(int N, const std::vector<cd> &S, const std::vector<cd> &b, std::vector<cd> &x,
                 int method) {
  x.assign(N, 0.0);
  if(method == 0)
    lu_solve(N, S, b.data(), x.data());
  else
    qr_solve(N, S, b.data(), x.data());
}

// ---------------------------------------------------------------- ACA
// getMaxInd (PreconditionedMatrix.cpp:864-889): the first index of largest
// modulus among the entries whose index is not one of used[0 .. nused-1];
// -1 when no such entry is non-zero (the reference then returns an unset
// index).
int max_index(const cd *v, int n, const std::vector<int> &used, int nused, double *tie = nullptr) {
  double mx = 0.0, second = 0.0;
  int im = -1;
  for(int i = 0; i < n; ++i) {
    bool same = false;
    for(int j = 0; j < nused; ++j)
      if(used[j] == i) same = true;
    if(same) continue;
    double a = std::abs(v[i]);
    if(a > mx) {
      second = mx;
      mx = a;
      im = i;
    } else if(a > second) {
      second = a;
    }
  }
  if(tie) *tie = mx > 0.0 ? second / mx : 0.0;  // how close the runner-up came
  return im;
}

// ACA_compression (PreconditionedMatrix.cpp:763-862): partially pivoted
// adaptive cross approximation of the n x n block A (column-major),
// A ~ U V, U n x r (column p at U + p n), V r x n (row p at V + p n).
// Step k: residual row I_k, its largest entry J_k (columns not yet used),
// v_k = row / pivot, residual column J_k -> u_k; the squared Frobenius
// estimate grows by |u_k|^2 |v_k|^2 + 2 sum_{p = 0}^{k-2}
// |sum_{t < k} u_p[t] u_k[t]| |sum_{t < k} v_p[t] v_k[t]| (the reference's
// loop bounds and unconjugated products, :826-843); the iteration stops
// after step k >= 1 when eps sqrt(estimate) >= |u_k| |v_k|, else I_{k+1} is
// the largest entry of the residual column among the unused rows.  When a
// residual row or column has no non-zero candidate left the k terms
// reproduce A exactly and the iteration stops (the reference reads an unset
// index there).  Returns r.  trace (optional, 3 per step): the runner-up /
// largest modulus of the pivot column's choice, of the next row's choice,
// and eps sqrt(estimate) / (|u_k| |v_k|) (the test stops at >= 1).
int aca_block(int n, const cd *A, cd *U, cd *V, double eps, double *trace = nullptr) {
  std::vector<int> I(1, 0), J;
  std::vector<cd> row(n);
  auto nrm = [n](const cd *v) {
    double s = 0;
    for(int i = 0; i < n; ++i) s += std::norm(v[i]);
    return std::sqrt(s);
  };
  double norma = 0.0;
  int r = 0;
  for(int k = 0; k < n; ++k) {
    const int ik = I[k];
    for(int c = 0; c < n; ++c) {
      cd s = 0.0;
      for(int p = 0; p < k; ++p) s += U[(size_t)p * n + ik] * V[(size_t)p * n + c];
      row[c] = A[(size_t)c * n + ik] - s;
    }
    double tie = 0.0;
    const int jk = max_index(row.data(), n, J, k, &tie);
    if(trace) trace[3 * k] = tie, trace[3 * k + 1] = 0.0, trace[3 * k + 2] = 0.0;
    if(jk < 0) break;
    J.push_back(jk);
    const cd piv = row[jk];
    cd *vk = V + (size_t)k * n, *uk = U + (size_t)k * n;
    for(int c = 0; c < n; ++c) vk[c] = row[c] / piv;
    for(int q = 0; q < n; ++q) {
      cd s = 0.0;
      for(int p = 0; p < k; ++p) s += V[(size_t)p * n + jk] * U[(size_t)p * n + q];
      uk[q] = A[(size_t)jk * n + q] - s;
    }
    r = k + 1;
    const double cn = nrm(uk), rn = nrm(vk);
    if(k == 0) {
      norma = cn * cn * (rn * rn);
    } else {
      double sum = 0.0;
      for(int p = 0; p != k - 1; ++p) {
        cd p1 = 0.0, p2 = 0.0;
        for(int t = 0; t != k; ++t) {
          p1 += U[(size_t)p * n + t] * uk[t];
          p2 += V[(size_t)p * n + t] * vk[t];
        }
        sum += std::abs(p1) * std::abs(p2);
      }
      norma = norma + cn * cn * (rn * rn) + 2.0 * sum;
      if(trace) trace[3 * k + 2] = eps * std::sqrt(norma) / (cn * rn);
      if(eps * std::sqrt(norma) >= cn * rn) break;
    }
    const int in = max_index(uk, n, I, k + 1, trace ? &trace[3 * k + 1] : nullptr);
    if(in < 0) break;
    I.push_back(in);
  }
  return r;
}

// Admissibility of the pair (i, j) for compression: i != j and
// |c_i - c_j| >= 2 (r_i + r_j) (PreconditionedMatrix.cpp:526, the distance of
// Tools::findDistance, Tools.cpp:30-38).
bool admissible(const Problem &pb, int i, int j) {
  if(i == j) return false;
  Cart a = to_cartesian(pb.obj[i].vR), b = to_cartesian(pb.obj[j].vR);
  double d = std::sqrt(std::pow(b.x - a.x, 2.0) + std::pow(b.y - a.y, 2.0) + std::pow(b.z - a.z, 2.0));
  return d >= 2.0 * (pb.obj[i].radius + pb.obj[j].radius);
}

// Scattering_matrix_ACA_FF / _SH (PreconditionedMatrix.cpp:489-551,
// 696-759): every admissible block compressed, the others dense, expanded
// back to the N x N compressed operator (column-major) for the dense
// solvers; rank[i P + j] = r of a compressed block, -1 for a dense block,
// 0 on the diagonal.
void aca_operator(const Problem &pb, bool sh, double eps, std::vector<cd> &S, int *rank) {
  const int nf = pb.nf, P = pb.P, n2 = 2 * nf;
  const size_t N = (size_t)n2 * P;
  S.assign(N * N, cd(0, 0));
  parallel_for(P, [&](int i) {
    auto T = object_vector(pb, pb.obj[i], sh ? 1 : 0, pb.bg_eps, pb.bg_mu);
    std::vector<cd> blk((size_t)n2 * n2), U((size_t)n2 * n2), V((size_t)n2 * n2);
    for(int j = 0; j < P; ++j) {
      cd *dst = &S[(size_t)j * n2 * N + (size_t)i * n2];
      if(!admissible(pb, i, j)) {
        scattering_block(pb, sh, i, j, T, dst, N);
        if(rank) rank[(size_t)i * P + j] = i == j ? 0 : -1;
        continue;
      }
      scattering_block(pb, sh, i, j, T, blk.data(), n2);
      int r = aca_block(n2, blk.data(), U.data(), V.data(), eps);
      if(rank) rank[(size_t)i * P + j] = r;
      for(int c = 0; c < n2; ++c)
        for(int q = 0; q < n2; ++q) {
          cd s = 0.0;
          for(int p = 0; p < r; ++p) s += U[(size_t)p * n2 + q] * V[(size_t)p * n2 + c];
          dst[(size_t)c * N + q] = s;
        }
    }
  });
}

// Gmres_Zcomp (PreconditionedMatrix.cpp:892-985) with det_approx
// (:1087-1130) on the operator A (N x N column-major): restarted GMRES from
// x = 0, modified Gram-Schmidt, the Hessenberg column rotated by the
// reference's Givens rotations (c, s from the ratio of the smaller to the
// larger entry, W = [[conj c, -conj s], [s, c]]), convergence when
// |g_{n+1}| / |b| <= tol, at most maxit steps per cycle and `restarts`
// cycles; the small triangular system is solved by back substitution (the
// reference's column-pivoted QR of the same triangle).  info: steps,
// final estimate.
void gmres_zcomp(int N, const std::vector<cd> &A, const cd *b, double tol, int maxit, int restarts, cd *xout,
                 double *info) {
  auto mv = [&](const std::vector<cd> &x, std::vector<cd> &y) {
    y.assign(N, 0.0);
    for(int c = 0; c < N; ++c) {
      const cd xc = x[c];
      const cd *col = &A[(size_t)c * N];
      for(int i = 0; i < N; ++i) y[i] += col[i] * xc;
    }
  };
  auto norm = [N](const std::vector<cd> &v) {
    double s = 0;
    for(int i = 0; i < N; ++i) s += std::norm(v[i]);
    return std::sqrt(s);
  };
  std::vector<cd> x(N, 0.0), w, res(N);
  std::vector<std::vector<cd>> v;
  std::vector<cd> H, Ri, gi;
  double abs_y = 0;
  for(int i = 0; i < N; ++i) abs_y += std::norm(b[i]);
  abs_y = std::sqrt(abs_y);
  std::vector<double> err(1, 1.0);
  int n = 0, steps = 0;
  for(int rest = 1; rest <= restarts; ++rest) {
    if(err[n] <= tol) break;
    mv(x, w);
    for(int i = 0; i < N; ++i) res[i] = b[i] - w[i];
    const double beta = norm(res);
    v.assign(1, std::vector<cd>(N));
    for(int i = 0; i < N; ++i) v[0][i] = res[i] / beta;
    // Ri: the rotated Hessenberg matrix, (maxit + 1) x maxit column-major; gi the rotated beta e_1
    const int ld = maxit + 1;
    Ri.assign((size_t)ld * maxit, 0.0);
    gi.assign(ld, 0.0);
    gi[0] = beta;
    std::vector<cd> cs, sn;
    n = 0;
    err.resize(1);  // err(0) = 1 throughout (:902); the later entries are rewritten
    while(n < maxit && err[n] > tol) {
      mv(v[n], w);
      cd *h = &Ri[(size_t)n * ld];
      for(int t = 0; t <= n; ++t) {
        cd d = 0.0;
        for(int i = 0; i < N; ++i) d += std::conj(v[t][i]) * w[i];
        h[t] = d;
        for(int i = 0; i < N; ++i) w[i] -= d * v[t][i];
      }
      h[n + 1] = norm(w);
      v.emplace_back(N);
      for(int i = 0; i < N; ++i) v[n + 1][i] = w[i] / h[n + 1];
      for(int i = 0; i < n; ++i) {  // the earlier rotations on the new column
        cd a = h[i], bb = h[i + 1];
        h[i] = std::conj(cs[i]) * a + std::conj(-sn[i]) * bb;
        h[i + 1] = sn[i] * a + cs[i] * bb;
      }
      cd hi1 = h[n], hi2 = h[n + 1], c, s;
      if(std::abs(hi2) > std::abs(hi1)) {
        cd temp = hi1 / hi2;
        s = 1.0 / std::sqrt(1.0 + std::pow(std::abs(temp), 2));
        c = -temp * s;
      } else {
        cd temp = hi2 / hi1;
        c = 1.0 / std::sqrt(1.0 + std::pow(std::abs(temp), 2));
        s = -temp * c;
      }
      cs.push_back(c);
      sn.push_back(s);
      h[n] = std::conj(c) * hi1 + std::conj(-s) * hi2;
      h[n + 1] = s * hi1 + c * hi2;
      cd g0 = gi[n], g1 = gi[n + 1];
      gi[n] = std::conj(c) * g0 + std::conj(-s) * g1;
      gi[n + 1] = s * g0 + c * g1;
      err.push_back(std::abs(gi[n + 1]) / abs_y);
      ++n;
      ++steps;
    }
    std::vector<cd> y(n);
    for(int i = n - 1; i >= 0; --i) {
      cd s = gi[i];
      for(int k = i + 1; k < n; ++k) s -= Ri[(size_t)k * ld + i] * y[k];
      y[i] = s / Ri[(size_t)i * ld + i];
    }
    for(int j = 0; j < n; ++j)
      for(int i = 0; i < N; ++i) x[i] += y[j] * v[j][i];
  }
  std::copy(x.begin(), x.end(), xout);
  if(info) {
    info[0] = steps;
    info[1] = err[n];
  }
}

// ---------------------------------------------------------------- Wigner
long double fact(int n) {
  static const std::vector<long double> t = [] {  // initialised once, thread-safe
    std::vector<long double> v(171);
    v[0] = 1;
    for(int i = 1; i < 171; ++i) v[i] = v[i - 1] * i;
    return v;
  }();
  if(n < 0) throw OraError("negative factorial");
  return t[n];
}
bool triangle(int a, int b, int c) { return a >= 0 && b >= 0 && c >= 0 && c >= std::abs(a - b) && c <= a + b; }
long double delta(int a, int b, int c) {
  return std::sqrt(fact(a + b - c) * fact(a - b + c) * fact(-a + b + c) / fact(a + b + c + 1));
}
// gsl_sf_coupling_3j semantics (Symbol.cpp:28-31): zero outside selection rules.
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if(m1 + m2 + m3 != 0 || !triangle(j1, j2, j3)) return 0.0;
  if(std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
  int tmin = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
  int tmax = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
  long double s = 0;
  for(int t = tmin; t <= tmax; ++t) {
    long double d = fact(t) * fact(j3 - j2 + t + m1) * fact(j3 - j1 + t - m2) * fact(j1 + j2 - j3 - t) *
                    fact(j1 - t - m1) * fact(j2 - t + m2);
    s += ((t % 2) ? -1.0L : 1.0L) / d;
  }
  long double pre = delta(j1, j2, j3) * std::sqrt(fact(j1 + m1) * fact(j1 - m1) * fact(j2 + m2) *
                                                  fact(j2 - m2) * fact(j3 + m3) * fact(j3 - m3));
  int e = j1 - j2 - m3;
  return (double)(((e % 2 + 2) % 2 ? -1.0L : 1.0L) * pre * s);
}
long double wigner6j_l(int j1, int j2, int j3, int j4, int j5, int j6) {
  if(!triangle(j1, j2, j3) || !triangle(j1, j5, j6) || !triangle(j4, j2, j6) || !triangle(j4, j5, j3))
    return 0.0L;
  int a1 = j1 + j2 + j3, a2 = j1 + j5 + j6, a3 = j4 + j2 + j6, a4 = j4 + j5 + j3;
  int b1 = j1 + j2 + j4 + j5, b2 = j2 + j3 + j5 + j6, b3 = j3 + j1 + j6 + j4;
  int tmin = std::max({a1, a2, a3, a4}), tmax = std::min({b1, b2, b3});
  long double s = 0;
  for(int t = tmin; t <= tmax; ++t) {
    long double d = fact(t - a1) * fact(t - a2) * fact(t - a3) * fact(t - a4) * fact(b1 - t) *
                    fact(b2 - t) * fact(b3 - t);
    s += ((t % 2) ? -1.0L : 1.0L) * fact(t + 1) / d;
  }
  return delta(j1, j2, j3) * delta(j1, j5, j6) * delta(j4, j2, j6) * delta(j4, j5, j3) * s;
}
double wigner6j(int a, int b, int c, int d, int e, int f) { return (double)wigner6j_l(a, b, c, d, e, f); }
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  if(!triangle(a, b, c) || !triangle(d, e, f) || !triangle(g, h, i) || !triangle(a, d, g) ||
     !triangle(b, e, h) || !triangle(c, f, i))
    return 0.0;
  int xmin = std::max({std::abs(a - i), std::abs(b - f), std::abs(d - h)});
  int xmax = std::min({a + i, b + f, d + h});
  long double s = 0;
  for(int x = xmin; x <= xmax; ++x)
    s += (2 * x + 1) * wigner6j_l(a, b, c, f, i, x) * wigner6j_l(d, e, f, b, x, h) *
         wigner6j_l(g, h, i, x, a, d);
  return (double)s;
}
// Symbol.cpp:44-49
double clegor(int j, int m, int j1, int m1, int j2, int m2) {
  return std::pow(-1.0, m + j1 - j2) * std::sqrt(2.0 * j + 1.0) * wigner3j(j1, j2, j, m1, m2, -m);
}
// Symbol.cpp:143-151
double Wfun(int L1, int J1, int M1, int L2, int J2, int M2, int L, int M) {
  return std::pow(-1.0, J2 + L1 + L) *
         std::sqrt((2.0 * J1 + 1.0) * (2.0 * J2 + 1.0) * (2.0 * L1 + 1.0) * (2.0 * L2 + 1.0) /
                   (4.0 * kPi * (2.0 * L + 1.0))) *
         wigner6j(L1, L2, L, J2, J1, 1) * clegor(L, 0, L1, 0, L2, 0) * clegor(L, M, J1, M1, J2, M2);
}

// Symbol.cpp:1036-1451, the nine tables, index [k][p][q].
void cg_tables(int nmax, int nmaxs, double *out) {
  int nf = nmax * (nmax + 2), nfs = nmaxs * (nmaxs + 2);
  size_t sz = (size_t)nfs * nf * nf;
  const double s32 = std::sqrt(3.0 / 2.0 / kPi);
  parallel_for(nfs, [&](int kk) {  // output harmonics on all threads
    int J = (int)std::sqrt(kk + 1.0), M = -(kk + 1) + J * (J + 1);
    size_t idx = (size_t)kk * nf * nf;
    for(int p = 0; p < nf; ++p) {
      int J1 = (int)std::sqrt(p + 1.0), M1 = -(p + 1) + J1 * (J1 + 1);
      for(int q = 0; q < nf; ++q, ++idx) {
        int J2 = (int)std::sqrt(q + 1.0), M2 = -(q + 1) + J2 * (J2 + 1);
        double cg = clegor(J, M, J1, M1, J2, M2);
        double sJ = std::sqrt(J / (2.0 * J + 1.0)), sJ1 = std::sqrt((J + 1) / (2.0 * J + 1.0));
        // C_10m1 (Symbol.cpp:1059-1079)
        out[0 * sz + idx] =
            s32 * (2.0 * J1 + 1.0) * cg *
            (std::sqrt(J2 * (2.0 * J2 - 1.0)) * wigner9j(J1, J1, 1, J2, J2 - 1, 1, J, J + 1, 1) *
                 clegor(J + 1, 0, J1, 0, J2 - 1, 0) * sJ -
             std::sqrt((J2 + 1.0) * (2.0 * J2 + 3.0)) * wigner9j(J1, J1, 1, J2, J2 + 1, 1, J, J + 1, 1) *
                 clegor(J + 1, 0, J1, 0, J2 + 1, 0) * sJ +
             std::sqrt(J2 * (2.0 * J2 - 1.0)) * wigner9j(J1, J1, 1, J2, J2 - 1, 1, J, J - 1, 1) *
                 clegor(J - 1, 0, J1, 0, J2 - 1, 0) * sJ1 -
             std::sqrt((J2 + 1) * (2.0 * J2 + 3.0)) * wigner9j(J1, J1, 1, J2, J2 + 1, 1, J, J - 1, 1) *
                 clegor(J - 1, 0, J1, 0, J2 + 1, 0) * sJ1);
        // C_11m1 (Symbol.cpp:1111-1146)
        double t1 = std::sqrt((J1 + 1.0) * J2 * (2.0 * J1 - 1.0) * (2.0 * J2 - 1.0));
        double t2 = std::sqrt((J1 + 1.0) * (J2 + 1) * (2.0 * J1 - 1.0) * (2.0 * J2 + 3.0));
        double t3 = std::sqrt(J1 * J2 * (2.0 * J1 + 3.0) * (2.0 * J2 - 1.0));
        double t4 = std::sqrt(J1 * (J2 + 1) * (2.0 * J1 + 3.0) * (2.0 * J2 + 3.0));
        out[1 * sz + idx] =
            s32 * cg *
            (t1 * wigner9j(J1, J1 - 1, 1, J2, J2 - 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1 - 1, 0, J2 - 1, 0) * sJ -
             t2 * wigner9j(J1, J1 - 1, 1, J2, J2 + 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1 - 1, 0, J2 + 1, 0) * sJ +
             t3 * wigner9j(J1, J1 + 1, 1, J2, J2 - 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1 + 1, 0, J2 - 1, 0) * sJ -
             t4 * wigner9j(J1, J1 + 1, 1, J2, J2 + 1, 1, J, J + 1, 1) * clegor(J + 1, 0, J1 + 1, 0, J2 + 1, 0) * sJ +
             t1 * wigner9j(J1, J1 - 1, 1, J2, J2 - 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1 - 1, 0, J2 - 1, 0) * sJ1 -
             t2 * wigner9j(J1, J1 - 1, 1, J2, J2 + 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1 - 1, 0, J2 + 1, 0) * sJ1 +
             t3 * wigner9j(J1, J1 + 1, 1, J2, J2 - 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1 + 1, 0, J2 - 1, 0) * sJ1 -
             t4 * wigner9j(J1, J1 + 1, 1, J2, J2 + 1, 1, J, J - 1, 1) * clegor(J - 1, 0, J1 + 1, 0, J2 + 1, 0) * sJ1);
        // C_00m1 (Symbol.cpp:1177-1185)
        out[2 * sz + idx] =
            s32 * (2.0 * J1 + 1.0) * cg *
            (std::sqrt(J2 * (2.0 * J2 - 1.0)) * wigner9j(J1, J1, 1, J2, J2 - 1, 1, J, J, 1) *
                 clegor(J, 0, J1, 0, J2 - 1, 0) -
             std::sqrt((J2 + 1.0) * (2.0 * J2 + 3.0)) * wigner9j(J1, J1, 1, J2, J2 + 1, 1, J, J, 1) *
                 clegor(J, 0, J1, 0, J2 + 1, 0));
        // C_01m1 (Symbol.cpp:1216-1232)
        out[3 * sz + idx] =
            s32 * cg *
            (t1 * wigner9j(J1, J1 - 1, 1, J2, J2 - 1, 1, J, J, 1) * clegor(J, 0, J1 - 1, 0, J2 - 1, 0) -
             std::sqrt((J1 + 1.0) * (J2 + 1.0) * (2.0 * J1 - 1.0) * (2.0 * J2 + 3.0)) *
                 wigner9j(J1, J1 - 1, 1, J2, J2 + 1, 1, J, J, 1) * clegor(J, 0, J1 - 1, 0, J2 + 1, 0) +
             t3 * wigner9j(J1, J1 + 1, 1, J2, J2 - 1, 1, J, J, 1) * clegor(J, 0, J1 + 1, 0, J2 - 1, 0) -
             t4 * wigner9j(J1, J1 + 1, 1, J2, J2 + 1, 1, J, J, 1) * clegor(J, 0, J1 + 1, 0, J2 + 1, 0));
        double a1 = std::sqrt(J1 / (2.0 * J1 + 1.0)), b1 = std::sqrt((J1 + 1) / (2.0 * J1 + 1.0));
        double a2 = std::sqrt(J2 / (2.0 * J2 + 1.0)), b2 = std::sqrt((J2 + 1) / (2.0 * J2 + 1.0));
        // W_m1m1 (Symbol.cpp:1263-1279)
        out[4 * sz + idx] = a1 * a2 * Wfun(J1 - 1, J1, M1, J2 - 1, J2, M2, J, M) +
                            b1 * b2 * Wfun(J1 + 1, J1, M1, J2 + 1, J2, M2, J, M) -
                            a1 * b2 * Wfun(J1 - 1, J1, M1, J2 + 1, J2, M2, J, M) -
                            b1 * a2 * Wfun(J1 + 1, J1, M1, J2 - 1, J2, M2, J, M);
        // W_11 (Symbol.cpp:1313-1328)
        out[5 * sz + idx] = b1 * b2 * Wfun(J1 - 1, J1, M1, J2 - 1, J2, M2, J, M) +
                            a1 * a2 * Wfun(J1 + 1, J1, M1, J2 + 1, J2, M2, J, M) +
                            b1 * a2 * Wfun(J1 - 1, J1, M1, J2 + 1, J2, M2, J, M) +
                            a1 * b2 * Wfun(J1 + 1, J1, M1, J2 - 1, J2, M2, J, M);
        // W_00 (Symbol.cpp:1361)
        out[6 * sz + idx] = Wfun(J1, J1, M1, J2, J2, M2, J, M);
        // W_10 (Symbol.cpp:1394-1399)
        out[7 * sz + idx] = b1 * Wfun(J1 - 1, J1, M1, J2, J2, M2, J, M) + a1 * Wfun(J1 + 1, J1, M1, J2, J2, M2, J, M);
        // W_01 (Symbol.cpp:1433-1438)
        out[8 * sz + idx] = b2 * Wfun(J1, J1, M1, J2 - 1, J2, M2, J, M) + a2 * Wfun(J1, J1, M1, J2 + 1, J2, M2, J, M);
      }
    }
  });
}

// ---------------------------------------------------------------- SH source
// Geometry::getIncLocalSH (Geometry.cpp:250-310) with symbol::vp_mn, up_mn,
// upp_mn (Symbol.cpp:155-353) and the A_0 / A_1 / A_m1 helpers (:52-78).
// u = conj(x_int) (PreconditionedMatrixSolver.h:65-66).
void inc_local_sh(const Problem &pb, const double *tab, const std::vector<cd> &u, int obj, cd *out) {
  int nmax = pb.nmax, nf = pb.nf;
  size_t sz = (size_t)nf * nf * nf;
  const Object &o = pb.obj[obj];
  double R = o.radius;
  cd kj = pb.omega * std::sqrt(o.eps * o.mu);
  std::vector<cd> d, dd;
  bessel(kBessel, R * kj, nmax, d, dd);
  const cd I(0, 1);
  std::vector<cd> A0(nf), A1(nf), Am1(nf);
  for(int p = 0; p < nf; ++p) {
    int n = (int)std::sqrt(p + 1.0);
    cd c = u[(size_t)obj * 2 * nf + p], dm = u[(size_t)nf + obj * 2 * nf + p];
    A0[p] = d[n] * c;
    A1[p] = I * (1.0 / kj) * (kj * dd[n] + d[n] / R) * dm;
    Am1[p] = I * std::sqrt(n * (n + 1.0)) * (1.0 / kj / R) * d[n] * dm;
  }
  cd mu_0 = kMu0, eps_0 = kEps0;
  cd eta = std::sqrt(pb.bg_mu / pb.bg_eps) / std::sqrt(mu_0 / eps_0);
  cd k01 = pb.omega * std::sqrt(eps_0 * mu_0);
  for(int kk = 0; kk < nf; ++kk) {
    int n = (int)std::sqrt(kk + 1.0);
    cd sv = 0, su = 0, g = 0, f = 0;
    size_t base = (size_t)kk * nf * nf;
    for(int p = 0; p < nf; ++p)
      for(int q = 0; q < nf; ++q) {
        size_t t = base + (size_t)p * nf + q;
        sv += A1[p] * Am1[q] * tab[2 * sz + t] + A0[p] * Am1[q] * tab[3 * sz + t];
        su += A1[p] * Am1[q] * tab[0 * sz + t] + A0[p] * Am1[q] * tab[1 * sz + t];
        g += Am1[p] * Am1[q] * tab[4 * sz + t];
        f += A1[p] * A1[q] * tab[5 * sz + t] + A0[p] * A0[q] * tab[6 * sz + t] +
             A1[p] * A0[q] * tab[7 * sz + t] + A0[p] * A1[q] * tab[8 * sz + t];
      }
    out[kk] = sv * cd(2.0, 0.0) * o.ksiparppar * eta;                   // vp (Symbol.cpp:274-275)
    out[nf + kk] = su * cd(2.0, 0.0) * I * o.ksiparppar * eta;          // up (Symbol.cpp:212-213)
    out[2 * nf + kk] = 0.0;                                             // vpp (Geometry.cpp:296)
    out[3 * nf + kk] = I * o.ksippp * std::sqrt(n * (n + 1)) * g / k01 / R +  // upp (Symbol.cpp:347-351)
                       I * o.gamma * (eps_0 / o.eps_sh) * std::sqrt(n * (n + 1)) * (g + f) / k01 / R;
  }
}

// source_vectorSH / source_vectorSH_K1ana (PreconditionedMatrix.cpp:1347-1436)
void source_sh(const Problem &pb, const double *tab, const std::vector<cd> &x_int, std::vector<cd> &K,
               std::vector<cd> &K1) {
  int nf = pb.nf;
  std::vector<cd> u(x_int.size());
  for(size_t i = 0; i < u.size(); ++i) u[i] = std::conj(x_int[i]);
  K.assign((size_t)2 * nf * pb.P, 0.0);
  K1.assign((size_t)2 * nf * pb.P, 0.0);
  parallel_for(pb.P, [&](int j) {
    std::vector<cd> ana(4 * nf);
    inc_local_sh(pb, tab, u, j, ana.data());
    auto t1 = object_vector(pb, pb.obj[j], 5, pb.bg_eps, pb.bg_mu);
    auto t2 = object_vector(pb, pb.obj[j], 6, pb.bg_eps, pb.bg_mu);
    auto ia2 = object_vector(pb, pb.obj[j], 4, pb.bg_eps, pb.bg_mu);
    for(int r = 0; r < 2 * nf; ++r) {
      K[(size_t)j * 2 * nf + r] = t1[r] * ana[r] + t2[r] * ana[2 * nf + r];
      K1[(size_t)j * 2 * nf + r] = ia2[r] * ana[2 * nf + r];
    }
  });
}

// ---------------------------------------------------------------- SH absorption
// bessel3der (Bessel.h:147-245): the second-derivative array it returns,
// d2[i] = -(i/z) j_i' + (i/z^2) j_i + j_{i-1}' for i >= 1, d2[0] = 0, kept
// literally (the exact second derivative has i + 1 in place of i), real part
// only on the real axis, all zero for |z| <= 1e-10.
void bessel_d2(cd z, int nmax, std::vector<cd> &d, std::vector<cd> &dd, std::vector<cd> &d2) {
  bessel(kBessel, z, nmax, d, dd);
  d2.assign(nmax + 1, cd(0, 0));
  if(std::abs(z) <= kErrEps) return;
  for(int i = 1; i <= nmax; ++i) {
    d2[i] = -((double)i / z) * dd[i] + ((double)i / (z * z)) * d[i] + dd[i - 1];
    if(z.imag() == 0.0) d2[i] = d2[i].real();
  }
}

// symbol::CXm1 / CXp1 (Symbol.cpp:474-655), the sums COEFFXm1 / COEFFXp1 of
// ACSshcoeff (:430-455): coefficients of the X(-1) and X(+1) vector harmonics
// of the bulk (gamma) source of the fundamental internal field of object obj
// at radius r, for output harmonic kk.
void coeff_part_sh(const Problem &pb, const double *tab, const cd *x_int, int obj, int kk, double r, cd &xm1,
                   cd &xp1) {
  const int nmax = pb.nmax, nf = pb.nf;
  const size_t sz = (size_t)nf * nf * nf;
  const double *Wm1m1 = tab + 4 * sz, *W11 = tab + 5 * sz, *W00 = tab + 6 * sz;
  const Object &o = pb.obj[obj];
  const cd eps_0 = kEps0;
  const cd kj = pb.omega * std::sqrt(o.eps * o.mu);
  const cd pre = (-eps_0 / o.eps_sh) * o.gamma;
  const cd *c = x_int + (size_t)obj * 2 * nf, *d = c + nf;
  const int n = (int)std::sqrt(kk + 1.0);
  const double snn = std::sqrt((double)n * (n + 1));
  std::vector<cd> j, dj, d2j;
  bessel_d2(r * kj, nmax, j, dj, d2j);
  xm1 = 0.0;
  xp1 = 0.0;
  for(int p = 0; p < nf; ++p) {
    const int n1 = (int)std::sqrt(p + 1.0);
    for(int q = 0; q < nf; ++q) {
      const int n2 = (int)std::sqrt(q + 1.0);
      const size_t t = (size_t)kk * nf * nf + (size_t)p * nf + q;
      const double w00 = W00[t], w11 = W11[t], wm = Wm1m1[t];
      const double s4 = std::sqrt((double)n1 * n2 * (n1 + 1) * (n2 + 1));
      const cd cc = c[p] * c[q], ddp = d[p] * d[q];
      // F_d00, F_d11, F_dm1m1, F_00, F_11, F_m1m1 (Symbol.cpp:80-141)
      const cd fd00 = j[n1] * dj[n2] * kj + kj * dj[n1] * j[n2];
      const cd a1 = kj * kj * d2j[n1] - (1.0 / (r * r)) * j[n1] + (1.0 / r) * kj * dj[n1];
      const cd a2 = kj * kj * d2j[n2] - (1.0 / (r * r)) * j[n2] + (1.0 / r) * kj * dj[n2];
      const cd b1 = kj * dj[n1] + j[n1] / r, b2 = kj * dj[n2] + j[n2] / r;
      const cd fd11 = a1 * b2 + a2 * b1;
      const cd fdm1m1 = (1.0 / (r * r)) * (kj * dj[n1] * j[n2] + j[n1] * kj * dj[n2]) - (2.0 / (r * r * r)) * (j[n1] * j[n2]);
      const cd f00 = j[n1] * dj[n2];
      const cd f11 = b2 * b1;
      const cd fm1m1 = (1.0 / (r * r)) * (j[n1] * j[n2]);
      xm1 += pre * (cc * w00 * fd00 + ddp * (1.0 / (kj * kj)) * (w11 * fd11 + wm * s4 * fdm1m1));
      xp1 += pre * (cc * w00 * snn * (1.0 / r) * f00 +
                    ddp * (1.0 / (kj * kj)) * (w11 * snn * (1.0 / r) * f11 + wm * s4 * snn * (1.0 / r) * fm1m1));
    }
  }
}

// Geometry::AbsCSSHcoeff -> symbol::ACSshcoeff (Geometry.cpp:428-455,
// Symbol.cpp:358-470) for every output harmonic kk of object obj: the
// 4-point Gauss-Legendre radial integral over [0, R] of |E(2 omega)|^2 built
// from the bulk (gamma) source of the fundamental internal field and the
// internal SH field; summed over kk.  x_int and x_int_sh enter unconjugated.
// The SH term of the X(+1) component adds j_n'(K r) without its factor K
// (Symbol.cpp:460), kept literally.
cd abs_sh_object(const Problem &pb, const double *tab, const cd *x_int, const cd *x_int_sh, int obj) {
  const int nmax = pb.nmax, nf = pb.nf;
  const Object &o = pb.obj[obj];
  const double R = o.radius, omega = pb.omega;
  const cd eps_0 = kEps0, mu_0 = kMu0;
  const cd k01 = omega * std::sqrt(eps_0 * mu_0);
  const cd ksh = 2.0 * omega * std::sqrt(o.eps_sh * o.mu_sh);
  // Gauss-Legendre nodes and weights on [-1, 1] (Symbol.cpp:397-398)
  const double xi[4] = {-0.3399810435848563, 0.3399810435848563, -0.8611363115940526, 0.8611363115940526};
  const double wi[4] = {0.6521451548625461, 0.6521451548625461, 0.3478548451374538, 0.3478548451374538};
  const cd *csh = x_int_sh + (size_t)obj * 2 * nf, *dsh = csh + nf;
  cd total = 0.0;
  for(int kk = 0; kk < nf; ++kk) {
    const int n = (int)std::sqrt(kk + 1.0);
    const double snn = std::sqrt((double)n * (n + 1));
    cd integral = 0.0;
    for(int ii = 0; ii < 4; ++ii) {
      const double r = (R / 2.0) * xi[ii] + R / 2.0;
      std::vector<cd> jsh, djsh;
      bessel(kBessel, r * ksh, nmax, jsh, djsh);
      cd xm1, xp1;
      coeff_part_sh(pb, tab, x_int, obj, kk, r, xm1, xp1);
      const cd xm1sh = k01 * dsh[kk] * (1.0 / ksh) * snn * (1.0 / r) * jsh[n];
      const cd x0 = -k01 * csh[kk] * jsh[n];
      const cd xp1sh = k01 * dsh[kk] * (1.0 / ksh) * ((1.0 / r) * jsh[n] + djsh[n]);
      const cd em = xm1 + xm1sh, ep = xp1 + xp1sh;
      integral += wi[ii] * (r * r) * (em * std::conj(em) + x0 * std::conj(x0) + ep * std::conj(ep));
    }
    total += integral * (R / 2.0);
  }
  return total;
}

// Result::getAbsorptionCrossSection_SH (Result.cpp:763-794):
// sum over objects of Re(2 eta 0.5 sigma sum_kk coefABS), sigma = -i eps0 2 omega (eps_r_SH - 1),
// eta = sqrt(mu_b / eps_b).
double abs_sh(const Problem &pb, const double *tab, const cd *x_int, const cd *x_int_sh) {
  const cd eta = std::sqrt(pb.bg_mu / pb.bg_eps);
  std::vector<double> part(pb.P, 0.0);
  parallel_for(pb.P, [&](int j) {
    const cd sigma = -cd(0.0, 1.0) * kEps0 * 2.0 * pb.omega * (pb.obj[j].eps_r_sh - 1.0);
    part[j] = std::real((2.0 * eta) * 0.5 * sigma * abs_sh_object(pb, tab, x_int, x_int_sh, j));
  });
  double s = 0.0;
  for(double v : part) s += v;
  return s;
}

// ---------------------------------------------------------------- fields
// AuxCoefficients (AuxCoefficients.cpp:30-343): the vector spherical waves M,
// N and the vector harmonics X(-1), X(+1) of every harmonic p at the point R
// (spherical, relative to the expansion centre), projected on Cartesian axes
// (Tools::toProjection); regular (j_n) or radiating (h_n^(1)).
struct Vswf {
  std::vector<cd> M, N, Xm, Xp;  // [3 nf]: (x, y, z) of harmonic p at 3p
};
void vswf(Sph R, cd k, bool regular, int nmax, Vswf &v) {
  const int nf = nmax * (nmax + 2);
  v.M.assign(3 * nf, 0.0);
  v.N.assign(3 * nf, 0.0);
  v.Xm.assign(3 * nf, 0.0);
  v.Xp.assign(3 * nf, 0.0);
  std::vector<double> dn(nmax + 1);
  dn[0] = -1000;
  for(int i = 1; i <= nmax; ++i) dn[i] = std::sqrt((2.0 * i + 1.0) / (4.0 * kPi * (i * (i + 1))));
  std::vector<cd> data, ddata;
  bessel(regular ? kBessel : kHankel1, R.r * k, nmax, data, ddata);
  const bool axis = std::abs(R.the) < 1e-10 || (std::abs(R.the) - kPi + 1e-10) > 0.0;
  const cd Kr = k * R.r;
  for(int m = -nmax; m <= nmax; ++m) {
    std::vector<double> W, dW;
    vigdvig(nmax, m, R, W, dW);
    const double dm = (m % 2 == 0) ? 1.0 : -1.0;  // Legendre to Wigner
    const cd e(std::cos(m * R.phi), std::sin(m * R.phi));
    for(int n = std::max(1, std::abs(m)); n <= nmax; ++n) {
      double A;
      if(m == 0)
        A = 0.0;
      else if(axis)
        A = m / std::cos(R.the) * dW[n];
      else
        A = m / std::sin(R.the) * W[n];
      const cd Pn[3] = {W[n], 0.0, 0.0}, Cn[3] = {0.0, cd(0.0, A), cd(-dW[n], 0.0)},
               Bn[3] = {0.0, cd(dW[n], 0.0), cd(0.0, A)};
      const cd ct = dm * dn[n] * e;
      cd Mv[3], Nv[3], Xmv[3], Xpv[3];
      for(int c = 0; c < 3; ++c) {
        Mv[c] = ct * data[n] * Cn[c];
        Nv[c] = (1.0 / Kr) * dm * dn[n] * (((double)(n * (n + 1))) * data[n] * Pn[c] + (Kr * ddata[n] + data[n]) * Bn[c]) * e;
        Xmv[c] = dm * dn[n] * std::sqrt((double)(n * (n + 1))) * e * Pn[c];
        Xpv[c] = dm * dn[n] * e * Bn[c];
      }
      const int p = flat(n, m);
      to_projection(R, Mv, &v.M[3 * p]);
      to_projection(R, Nv, &v.N[3 * p]);
      to_projection(R, Xmv, &v.Xm[3 * p]);
      to_projection(R, Xpv, &v.Xp[3 * p]);
    }
  }
}

// Result::getEHFields with projection off (Result.cpp:74-301), and the
// particular solution coefficients of Result::setFields (Geometry.cpp:457-490)
// inside a sphere: E and H at omega and 2 omega at the point xyz (Cartesian),
// as Cartesian components.  Outside every sphere: the incident wave plus the
// radiating waves of every object (at 2 omega, the SH scattered waves, scaled
// by the vacuum wavenumber); inside the first sphere that holds the point
// (Geometry::checkInner): its internal waves (and, at 2 omega, the bulk source
// term X(-1) CXm1 + X(+1) CXp1).
void fields_at(const Problem &pb, const double *tab, const cd *xs, const cd *xi, const cd *xss, const cd *xis,
               const double xyz[3], cd E[3], cd H[3], cd Esh[3], cd Hsh[3]) {
  const int nmax = pb.nmax, nf = pb.nf;
  for(int c = 0; c < 3; ++c) E[c] = H[c] = Esh[c] = Hsh[c] = 0.0;
  const Sph R = to_spherical({xyz[0], xyz[1], xyz[2]});
  const cd mi(0.0, -1.0);
  const cd k0 = pb.omega * std::sqrt(cd(kEps0) * cd(kMu0));
  const cd iZ = mi / std::sqrt(pb.bg_mu / pb.bg_eps);
  int in = -1;
  for(int j = 0; j < pb.P && in < 0; ++j)
    if(sph_minus(R, pb.obj[j].vR).r <= pb.obj[j].radius) in = j;
  Vswf v;
  if(in < 0) {
    vswf(R, pb.waveK, true, nmax, v);
    for(int p = 0; p < nf; ++p)
      for(int c = 0; c < 3; ++c) {
        E[c] += v.M[3 * p + c] * pb.ap[p] + v.N[3 * p + c] * pb.bp[p];
        H[c] += (v.N[3 * p + c] * pb.ap[p] + v.M[3 * p + c] * pb.bp[p]) * iZ;
      }
    for(int j = 0; j < pb.P; ++j) {
      const Sph Rr = sph_minus(R, pb.obj[j].vR);
      vswf(Rr, pb.waveK, false, nmax, v);
      const cd *x = xs + (size_t)j * 2 * nf;
      for(int p = 0; p < nf; ++p)
        for(int c = 0; c < 3; ++c) {
          E[c] += v.M[3 * p + c] * x[p] + v.N[3 * p + c] * x[nf + p];
          H[c] += (v.N[3 * p + c] * x[p] + v.M[3 * p + c] * x[nf + p]) * iZ;
        }
      if(pb.sh && xss) {
        vswf(Rr, 2.0 * pb.waveK, false, nmax, v);
        const cd *y = xss + (size_t)j * 2 * nf;
        for(int p = 0; p < nf; ++p)
          for(int c = 0; c < 3; ++c) {
            Esh[c] += (v.M[3 * p + c] * y[p] + v.N[3 * p + c] * y[nf + p]) * k0;
            Hsh[c] += (v.N[3 * p + c] * y[p] + v.M[3 * p + c] * y[nf + p]) * (iZ * k0);
          }
      }
    }
    return;
  }
  const Object &o = pb.obj[in];
  const Sph Rr = sph_minus(R, o.vR);
  const cd ki = k0 * std::sqrt(o.eps / kEps0 * (o.mu / kMu0));
  const cd iZo = mi / std::sqrt(o.mu / o.eps);
  vswf(Rr, ki, true, nmax, v);
  const cd *x = xi + (size_t)in * 2 * nf;
  for(int p = 0; p < nf; ++p)
    for(int c = 0; c < 3; ++c) {
      E[c] += v.M[3 * p + c] * x[p] + v.N[3 * p + c] * x[nf + p];
      H[c] += (v.N[3 * p + c] * x[p] + v.M[3 * p + c] * x[nf + p]) * iZo;
    }
  if(pb.sh && xis) {
    const cd ksh = 2.0 * k0 * std::sqrt(o.eps_r_sh * (o.mu_sh / kMu0));
    const cd iZsh = mi / std::sqrt(o.mu_sh / o.eps_sh);
    vswf(Rr, ksh, true, nmax, v);
    const cd *y = xis + (size_t)in * 2 * nf;
    for(int p = 0; p < nf; ++p) {
      cd xm, xp;
      coeff_part_sh(pb, tab, xi, in, p, Rr.r, xm, xp);
      for(int c = 0; c < 3; ++c) {
        Esh[c] += (v.M[3 * p + c] * y[p] + v.N[3 * p + c] * y[nf + p]) * k0;
        Esh[c] += v.Xm[3 * p + c] * xm + v.Xp[3 * p + c] * xp;
        Hsh[c] += (v.N[3 * p + c] * y[p] + v.M[3 * p + c] * y[nf + p]) * iZsh * k0;
      }
    }
  }
}

// ---------------------------------------------------------------- cross sections
// Geometry::getCabsAux (Geometry.cpp:86-145)
void cabs_aux(const Problem &pb, const Object &o, std::vector<double> &out) {
  int nmax = pb.nmax, nf = pb.nf;
  cd ks = pb.omega * std::sqrt(o.eps * o.mu), kb = pb.omega * std::sqrt(pb.bg_eps * pb.bg_mu);
  cd rho = ks / kb, r0 = kb * o.radius, mu_j = o.mu, mu_0 = pb.bg_mu;
  std::vector<cd> Jd, Jdd, Rd, Rdd;
  bessel(kBessel, r0, nmax, Jd, Jdd);
  bessel(kBessel, rho * r0, nmax, Rd, Rdd);
  out.assign(2 * nf, 0.0);
  const cd I(0, 1);
  for(int p = 0; p < nf; ++p) {
    int n = (int)std::sqrt(p + 1.0);
    cd psi = r0 * Jd[n], dpsi = r0 * Jdd[n] + Jd[n];
    cd psirho = r0 * rho * Rd[n], dpsirho = r0 * rho * Rdd[n] + Rd[n];
    cd t1 = I * rho * mu_0 * std::conj(mu_j) * std::conj(psirho) * dpsirho;
    double t2 = std::abs(mu_j * psirho * dpsi - mu_0 * rho * dpsirho * psi);
    out[p] = std::real(t1) / (t2 * t2);
    t1 = I * std::conj(rho) * mu_0 * mu_j * std::conj(psirho) * dpsirho;
    t2 = std::abs(mu_0 * rho * psirho * dpsi - mu_j * dpsirho * psi);
    out[nf + p] = std::real(t1) / (t2 * t2);
  }
}

// Result.cpp:557-577, 579-647, 649-676, 678-761
void cross_sections(const Problem &pb, const cd *xs, const cd *xsh, double *out) {
  int nf = pb.nf;
  double rk = std::real(pb.waveK);
  double cext = 0, csca = 0, cabs = 0, cssh = 0;
  for(int j = 0; j < pb.P; ++j) {
    auto inc = inc_local_k(pb, pb.obj[j].vR, pb.waveK, pb.ap, pb.bp);
    const cd *x = xs + (size_t)j * 2 * nf;
    for(int p = 0; p < nf; ++p)
      cext += std::real(std::conj(inc[p]) * x[p] + std::conj(inc[p + nf]) * x[nf + p]);
    Sph rrel = sph_minus(pb.obj[j].vR, Sph{0, 0, 0});
    std::vector<cd> A, B;
    coupling(rrel, pb.waveK, pb.nmax, false, A, B);
    for(int p = 0; p < 2 * nf; ++p)
      for(int q = 0; q < 2 * nf; ++q) {
        bool diag = (p < nf) == (q < nf);
        cd t = diag ? A[(size_t)(q % nf) * nf + p % nf] : B[(size_t)(q % nf) * nf + p % nf];
        csca += std::real(t * std::conj(x[q]) * std::conj(t) * x[q]);
      }
    std::vector<double> ca;
    cabs_aux(pb, pb.obj[j], ca);
    for(int p = 0; p < nf; ++p) {
      double t1 = std::abs(x[p]);
      t1 *= t1;
      double t2 = std::abs(x[nf + p]);
      t2 *= t2;
      cabs += t1 * ca[p] + t2 * ca[nf + p];
    }
    if(xsh) {
      const cd *y = xsh + (size_t)j * 2 * nf;
      std::vector<cd> A2, B2;
      coupling(rrel, 2.0 * pb.waveK, pb.nmax, false, A2, B2);
      for(int p = 0; p < 2 * nf; ++p)
        for(int q = 0; q < 2 * nf; ++q) {
          bool diag = (p < nf) == (q < nf);
          cd t = diag ? A2[(size_t)(q % nf) * nf + p % nf] : B2[(size_t)(q % nf) * nf + p % nf];
          cssh += std::real(t * std::conj(y[q]) * std::conj(t) * y[q]);
        }
    }
  }
  out[0] = (-1. / (rk * rk)) * cext;
  out[1] = (1. / (rk * rk)) * csca;
  out[2] = (1. / (rk * rk)) * cabs;
  double arb = std::real(pb.bg_eps_r) * std::real(pb.bg_mu_r);
  out[3] = xsh ? (1.0 / (4.0 * arb)) * cssh : 0.0;
}

template <typename F> int guard(F &&f) {
  try {
    f();
    return 0;
  } catch(std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

}  // namespace

extern "C" {

const char *ora_last_error(void) { return g_err.c_str(); }

int ora_bessel(int type, double zr, double zi, int nmax, double *data, double *ddata) {
  return guard([&] {
    std::vector<cd> d, dd;
    bessel(type, cd(zr, zi), nmax, d, dd);
    for(int i = 0; i <= nmax; ++i) {
      put(data, i, d[i]);
      put(ddata, i, dd[i]);
    }
  });
}

int ora_ynm(int n, int m, double theta, double phi, double *out) {
  return guard([&] { put(out, 0, ynm(n, m, theta, phi)); });
}

int ora_tac(const double *R, double kr, double ki, int regular, int n, int m, int l, int k,
            double *out) {
  return guard([&] {
    int nm = std::max(n, l) + 1;
    TAC ta({R[0], R[1], R[2]}, cd(kr, ki), regular != 0, nm);
    put(out, 0, ta(n, m, l, k));
  });
}

int ora_coupling(const double *R, double kr, double ki, int nmax, int regular, double *A, double *B) {
  return guard([&] {
    std::vector<cd> a, b;
    coupling({R[0], R[1], R[2]}, cd(kr, ki), nmax, regular != 0, a, b);
    for(size_t i = 0; i < a.size(); ++i) {
      put(A, i, a[i]);
      put(B, i, b[i]);
    }
  });
}

int ora_object_vector(const ora_problem *in, int obj, int which, double *out) {
  return guard([&] {
    Problem pb = make_problem(in);
    if(obj < 0 || obj >= pb.P) throw OraError("object index");
    auto v = object_vector(pb, pb.obj[obj], which, pb.bg_eps, pb.bg_mu);
    for(size_t i = 0; i < v.size(); ++i) put(out, i, v[i]);
  });
}

int ora_incident(const ora_problem *in, double *ap, double *bp) {
  return guard([&] {
    Problem pb = make_problem(in);
    for(int i = 0; i < pb.nf; ++i) {
      put(ap, i, pb.ap[i]);
      put(bp, i, pb.bp[i]);
    }
  });
}

int ora_inc_local(const ora_problem *in, int obj, double *out) {
  return guard([&] {
    Problem pb = make_problem(in);
    auto v = inc_local_k(pb, pb.obj[obj].vR, pb.waveK, pb.ap, pb.bp);
    for(size_t i = 0; i < v.size(); ++i) put(out, i, v[i]);
  });
}

int ora_source_vector(const ora_problem *in, double *Q) {
  return guard([&] {
    Problem pb = make_problem(in);
    auto v = source_vector(pb);
    for(size_t i = 0; i < v.size(); ++i) put(Q, i, v[i]);
  });
}

int ora_scattering_matrix(const ora_problem *in, int sh, double *S) {
  return guard([&] {
    Problem pb = make_problem(in);
    std::vector<cd> s;
    scattering_matrix(pb, sh != 0, s);
    std::memcpy(S, s.data(), s.size() * sizeof(cd));
  });
}

int ora_scattering_blocks(const ora_problem *in, int sh, int nblk, const int *ij, double *out) {
  return guard([&] {
    Problem pb = make_problem(in);
    const int nf = pb.nf;
    const size_t bs = (size_t)4 * nf * nf;
    for(int b = 0; b < nblk; ++b)
      if(ij[2 * b] < 0 || ij[2 * b] >= pb.P || ij[2 * b + 1] < 0 || ij[2 * b + 1] >= pb.P)
        throw OraError("block index outside the cluster");
    parallel_for(nblk, [&](int b) {
      const int i = ij[2 * b], j = ij[2 * b + 1];
      auto T = object_vector(pb, pb.obj[i], sh ? 1 : 0, pb.bg_eps, pb.bg_mu);
      scattering_block(pb, sh != 0, i, j, T, reinterpret_cast<cd *>(out) + b * bs, (size_t)2 * nf);
    });
  });
}

int ora_residual(const ora_problem *in, int sh, const double *x, const double *b, double *r) {
  return guard([&] {
    Problem pb = make_problem(in);
    const int nf = pb.nf, P = pb.P, n2 = 2 * nf;
    const size_t N = (size_t)n2 * P;
    const cd *xx = reinterpret_cast<const cd *>(x), *bb = reinterpret_cast<const cd *>(b);
    cd *rr = reinterpret_cast<cd *>(r);
    parallel_for(P, [&](int i) {
      auto T = object_vector(pb, pb.obj[i], sh ? 1 : 0, pb.bg_eps, pb.bg_mu);
      std::vector<cd> blk((size_t)n2 * n2);
      std::vector<cd> acc(bb + (size_t)i * n2, bb + (size_t)(i + 1) * n2);
      for(int j = 0; j < P; ++j) {
        scattering_block(pb, sh != 0, i, j, T, blk.data(), n2);
        const cd *xj = xx + (size_t)j * n2;
        for(int c = 0; c < n2; ++c)
          for(int q = 0; q < n2; ++q) acc[q] -= blk[(size_t)c * n2 + q] * xj[c];
      }
      std::copy(acc.begin(), acc.end(), rr + (size_t)i * n2);
    });
    (void)N;
  });
}

int ora_dense_solve(int N, const double *S, const double *b, double *x, int method) {
  return guard([&] {
    std::vector<cd> s((const cd *)S, (const cd *)S + (size_t)N * N);
    std::vector<cd> bb((const cd *)b, (const cd *)b + N), xx;
    dense_solve(N, s, bb, xx, method);
    std::memcpy(x, xx.data(), (size_t)N * sizeof(cd));
  });
}

int ora_aca_block(int n, const double *A, double eps, int *rank, double *U, double *V, double *trace) {
  return guard([&] {
    if(n < 1) throw OraError("block order must be positive");
    *rank = aca_block(n, reinterpret_cast<const cd *>(A), reinterpret_cast<cd *>(U), reinterpret_cast<cd *>(V), eps,
                      trace);
  });
}

int ora_aca_operator(const ora_problem *in, int sh, double eps, double *S, int *rank) {
  return guard([&] {
    Problem pb = make_problem(in);
    std::vector<cd> s;
    aca_operator(pb, sh != 0, eps, s, rank);
    std::memcpy(S, s.data(), s.size() * sizeof(cd));
  });
}

int ora_gmres_zcomp(int N, const double *A, const double *b, double tol, int maxit, int restarts, double *x,
                    double *info) {
  return guard([&] {
    if(N < 1 || maxit < 1 || restarts < 1) throw OraError("invalid GMRES settings");
    std::vector<cd> a(reinterpret_cast<const cd *>(A), reinterpret_cast<const cd *>(A) + (size_t)N * N);
    gmres_zcomp(N, a, reinterpret_cast<const cd *>(b), tol, maxit, restarts, reinterpret_cast<cd *>(x), info);
  });
}

int ora_cg_tables(int nmax, int nmaxs, double *tables) {
  return guard([&] { cg_tables(nmax, nmaxs, tables); });
}

double ora_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) { return wigner3j(j1, j2, j3, m1, m2, m3); }
double ora_wigner6j(int a, int b, int c, int d, int e, int f) { return wigner6j(a, b, c, d, e, f); }
double ora_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  return wigner9j(a, b, c, d, e, f, g, h, i);
}

int ora_source_sh(const ora_problem *in, const double *tables, const double *x_int, double *K,
                  double *K1ana) {
  return guard([&] {
    Problem pb = make_problem(in);
    size_t N = (size_t)2 * pb.nf * pb.P;
    std::vector<cd> xi((const cd *)x_int, (const cd *)x_int + N), k, k1;
    source_sh(pb, tables, xi, k, k1);
    std::memcpy(K, k.data(), N * sizeof(cd));
    std::memcpy(K1ana, k1.data(), N * sizeof(cd));
  });
}

int ora_solve_all(const ora_problem *in, int method, double *x_sca, double *x_int, double *x_sca_sh,
                  double *x_int_sh, double *xs) {
  return guard([&] {
    Problem pb = make_problem(in);
    int nf = pb.nf;
    size_t N = (size_t)2 * nf * pb.P;
    // update(): PreconditionedMatrixSolver.h:82-100
    auto Q = source_vector(pb);
    std::vector<cd> S, X;
    // method 2: the ACA-compressed operators and Gmres_Zcomp with the
    // reference's settings tol 1e-6, 240 steps, 2 cycles
    // (PreconditionedMatrixSolver.h:47-56, 84-96)
    auto solve_sys = [&](bool sh, const std::vector<cd> &rhs, std::vector<cd> &out) {
      if(method == 2) {
        aca_operator(pb, sh, 1e-3, S, nullptr);
        out.assign(N, 0.0);
        gmres_zcomp((int)N, S, rhs.data(), 1e-6, 240, 2, out.data(), nullptr);
      } else {
        scattering_matrix(pb, sh, S);
        dense_solve((int)N, S, rhs, out, method);
      }
    };
    // solve(): PreconditionedMatrixSolver.h:45-79
    solve_sys(false, Q, X);
    std::vector<cd> XI(N);
    for(int j = 0; j < pb.P; ++j) {  // convertInternal (Solver.cpp:57-77)
      auto ia = object_vector(pb, pb.obj[j], 2, pb.bg_eps, pb.bg_mu);
      for(int r = 0; r < 2 * nf; ++r) XI[(size_t)j * 2 * nf + r] = X[(size_t)j * 2 * nf + r] * ia[r];
    }
    std::memcpy(x_sca, X.data(), N * sizeof(cd));
    std::memcpy(x_int, XI.data(), N * sizeof(cd));
    std::vector<cd> XS, XIS;
    if(pb.sh) {
      std::vector<double> tab((size_t)9 * nf * nf * nf);
      cg_tables(pb.nmax, pb.nmax, tab.data());
      std::vector<cd> K, K1;
      source_sh(pb, tab.data(), XI, K, K1);
      solve_sys(true, K, XS);
      XIS.resize(N);
      for(int j = 0; j < pb.P; ++j) {  // convertInternal_SH (Solver.cpp:95-116)
        auto ia = object_vector(pb, pb.obj[j], 3, pb.bg_eps, pb.bg_mu);
        for(int r = 0; r < 2 * nf; ++r) {
          size_t t = (size_t)j * 2 * nf + r;
          XIS[t] = XS[t] * ia[r] - K1[t];
        }
      }
      std::memcpy(x_sca_sh, XS.data(), N * sizeof(cd));
      std::memcpy(x_int_sh, XIS.data(), N * sizeof(cd));
    }
    if(xs) {
      cross_sections(pb, X.data(), pb.sh ? XS.data() : nullptr, xs);
      xs[4] = 0.0;
      if(pb.sh) {
        std::vector<double> tab((size_t)9 * nf * nf * nf);
        cg_tables(pb.nmax, pb.nmax, tab.data());
        xs[4] = abs_sh(pb, tab.data(), XI.data(), XIS.data());
      }
    }
  });
}

// Timing sample for the CPU baseline: out[0] seconds per ordered pair
// (irregular Coupling + the -T_i [[A^T,B^T],[B^T,A^T]] block, as
// PreconditionedMatrix.cpp:350-400), out[1] seconds of a dense solve of order
// nsolve with `method` (QR as PreconditionedMatrixSolver.h:58), out[2] seconds
// per object of the second-harmonic source (Geometry.cpp:250-310) with
// tables of the problem's size.
int ora_time_sample(const ora_problem *in, int npairs, int nsolve, int method, double *out) {
  return guard([&] {
    Problem pb = make_problem(in);
    int nf = pb.nf;
    auto now = [] { return std::chrono::steady_clock::now(); };
    std::vector<cd> blk((size_t)4 * nf * nf);
    auto T = object_vector(pb, pb.obj[0], 0, pb.bg_eps, pb.bg_mu);
    auto t0 = now();
    int done = 0;
    for(int i = 0; i < pb.P && done < npairs; ++i)
      for(int j = 0; j < pb.P && done < npairs; ++j) {
        if(i == j) continue;
        std::vector<cd> A, B;
        coupling(sph_minus(pb.obj[i].vR, pb.obj[j].vR), pb.waveK, pb.nmax, true, A, B);
        for(int c = 0; c < 2 * nf; ++c)
          for(int r = 0; r < 2 * nf; ++r) {
            int rr = r % nf, cc = c % nf;
            bool diag = (r < nf) == (c < nf);
            blk[(size_t)c * 2 * nf + r] = -T[r] * (diag ? A[(size_t)cc * nf + rr] : B[(size_t)cc * nf + rr]);
          }
        ++done;
      }
    out[0] = std::chrono::duration<double>(now() - t0).count() / std::max(done, 1);
    std::vector<cd> S((size_t)nsolve * nsolve), b(nsolve), x;
    unsigned long long st = 88172645463325252ull;
    auto rnd = [&st] {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      return (double)(st >> 11) / 9007199254740992.0 - 0.5;
    };
    for(auto &v : S) v = cd(rnd(), rnd());
    for(int i = 0; i < nsolve; ++i) S[(size_t)i * nsolve + i] += (double)nsolve;
    for(auto &v : b) v = cd(rnd(), rnd());
    t0 = now();
    dense_solve(nsolve, S, b, x, method);
    out[1] = std::chrono::duration<double>(now() - t0).count();
    std::vector<double> tab((size_t)9 * nf * nf * nf, 0.0);
    std::vector<cd> u((size_t)2 * nf * pb.P, cd(1.0, 0.5)), ana(4 * nf);
    t0 = now();
    inc_local_sh(pb, tab.data(), u, 0, ana.data());
    out[2] = std::chrono::duration<double>(now() - t0).count();
  });
}

int ora_absorption_sh(const ora_problem *in, const double *tables, const double *x_int, const double *x_int_sh,
                      double *out) {
  return guard([&] {
    Problem pb = make_problem(in);
    *out = abs_sh(pb, tables, (const cd *)x_int, (const cd *)x_int_sh);
  });
}

int ora_fields(const ora_problem *in, const double *tables, const double *x_sca, const double *x_int,
               const double *x_sca_sh, const double *x_int_sh, int npts, const double *xyz, double *E, double *H,
               double *Esh, double *Hsh) {
  return guard([&] {
    Problem pb = make_problem(in);
    parallel_for(npts, [&](int i) {
      cd e[3], h[3], es[3], hs[3];
      fields_at(pb, tables, (const cd *)x_sca, (const cd *)x_int, (const cd *)x_sca_sh, (const cd *)x_int_sh,
                xyz + 3 * i, e, h, es, hs);
      for(int c = 0; c < 3; ++c) {
        ((cd *)E)[3 * i + c] = e[c];
        ((cd *)H)[3 * i + c] = h[c];
        ((cd *)Esh)[3 * i + c] = es[c];
        ((cd *)Hsh)[3 * i + c] = hs[c];
      }
    });
  });
}

int ora_cross_sections(const ora_problem *in, const double *x_sca, const double *x_sca_sh, double *xs) {
  return guard([&] {
    Problem pb = make_problem(in);
    cross_sections(pb, (const cd *)x_sca, (const cd *)x_sca_sh, xs);
  });
}

}  // extern "C"
