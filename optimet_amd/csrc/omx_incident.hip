// omx_incident.hip -- plane-wave coefficients a_inc, b_inc of the excitation
// (Excitation.cpp:48-75 with AuxCoefficients.cpp:31-343).
//
// This translation unit is compiled without floating-point contraction
// (-ffp-contract=off, see optimet_amd/build.py).  For propagation along the z
// axis the reference shifts the polar angle to 1e-6 (AuxCoefficients.cpp:228)
// and divides the Wigner-d derivative, a difference of two O(1) terms that
// cancel to O(theta^2), by sin(1e-6).  The value it produces for the m = 0
// terms is then set by the rounding of each double operation, at about 1e-10
// of the dominant coefficients: fused multiply-adds alone move the result by
// that much.  Without contraction, and with the reference's operation order,
// the kernel reproduces the reference's double arithmetic (plain IEEE
// multiplies and adds, correctly rounded square roots) instead of a rounding
// of its own.
#include "omx_kernels.hpp"

namespace omx {

// ------------------------------------------------------------------ incident
// One block; thread t handles m = t - nmax.  Writes inc[0..nf) = a_inc and
// inc[nf..2nf) = b_inc.
__global__ void incident_kernel(IncArgs a) {
  const int nmax = a.nmax, nf = nmax * (nmax + 2);
  int t = threadIdx.x;
  if(t > 2 * nmax) return;
  int m = t - nmax;
  const double the = a.theta, phi = a.phi;
  // Einc = toProjection((0, theta, phi), (0, Etheta, Ephi)) (Reader.cpp:820-827)
  double st = sin(the), ct = cos(the), sp = sin(phi), cp = cos(phi);
  cplx E0 = ct * cp * a.etheta - sp * a.ephi;
  cplx E1 = ct * sp * a.etheta + cp * a.ephi;
  cplx E2 = -st * a.etheta;
  // VIGdVIG (AuxCoefficients.cpp:216-290)
  double W[kMaxNmax + 2], dW[kMaxNmax + 2];
  for(int i = 0; i <= nmax + 1; ++i) W[i] = dW[i] = 0.0;
  bool neg = m < 0;
  int am = abs(m);
  unsigned nM = (unsigned)nmax, mm = (unsigned)(am * am);
  double vig = neg ? kPi - the : the;
  bool axis = (fabs(the) < 1e-10) || (fabs(the) - kPi + 1e-10 > 0.0);
  if(axis) vig = vig + 1e-6;
  double x = cos(vig);
  double f2m = 1.0, fm = 1.0;
  for(int i = 2; i <= 2 * am; ++i) f2m *= i;
  for(int i = 2; i <= am; ++i) fm *= i;
  W[am] = pow(2.0, -am) * (sqrt(f2m) / fm) * pow(1.0 - x, am / 2.0) * pow(1.0 + x, am / 2.0);
  unsigned s = (unsigned)am;
  if(am == 0 && nM > 0) {
    W[1] = x * W[0];
    s = 1;
  }
  double sv = sin(vig);
  for(; s < nM; ++s) {
    W[s + 1] = ((2 * s + 1) * x * W[s] - sqrt((double)(s * s - mm)) * W[s - 1]) /
               sqrt((double)((s + 1) * (s + 1) - mm));
    dW[s] = (((s * sqrt((double)((s + 1) * (s + 1) - mm)) * W[s + 1]) / (2 * s + 1)) -
             (((s + 1) * sqrt((double)(s * s * (s * s - mm))) * W[s - 1]) / (s * (2 * s + 1)))) /
            sv;
  }
  {
    double Wn = ((2 * nM + 1) * x * W[nM] - sqrt((double)(nM * nM - mm)) * W[nM - 1]) /
                sqrt((double)((nM + 1) * (nM + 1) - mm));
    dW[nM] = (((nM * sqrt((double)((nM + 1) * (nM + 1) - mm)) * Wn) / (2 * nM + 1)) -
              (((nM + 1) * sqrt((double)(nM * nM * (nM * nM - mm))) * W[nM - 1]) / (nM * (2 * nM + 1)))) /
             sv;
  }
  if(neg)
    for(int i = 0; i <= nmax; ++i) {
      double c = (i % 2) ? -1.0 : 1.0;
      W[i] *= c;
      dW[i] *= -c;
    }
  double sgn = (am % 2 == 0) ? 1.0 : -1.0;
  cplx ph = mk(cos((double)m * phi), -sin((double)m * phi));
  for(int n = (am > 1 ? am : 1); n <= nmax; ++n) {
    double A;
    if(m == 0)
      A = 0.0;
    else if(axis)
      A = m / cos(the) * dW[n];
    else
      A = m / sin(the) * W[n];
    // C = (0, iA, -dW), B = (0, dW, iA) in the (r, theta, phi) basis, projected
    // to Cartesian (Tools.cpp toProjection), then conj(.) . Einc
    cplx c1 = mk(0.0, A), c2 = mk(-dW[n], 0.0);
    cplx b1 = mk(dW[n], 0.0), b2 = mk(0.0, A);
    cplx C0 = ct * cp * c1 - sp * c2, C1 = ct * sp * c1 + cp * c2, C2 = -st * c1;
    cplx B0 = ct * cp * b1 - sp * b2, B1 = ct * sp * b1 + cp * b2, B2 = -st * b1;
    cplx dotC = conj(C0) * E0 + conj(C1) * E1 + conj(C2) * E2;
    cplx dotB = conj(B0) * E0 + conj(B1) * E1 + conj(B2) * E2;
    double dn = sqrt((2.0 * n + 1.0) / (4.0 * kPi * (n * (n + 1))));
    int e = n % 4, e1 = (n + 3) % 4;
    cplx in = e == 0 ? mk(1) : e == 1 ? mk(0, 1) : e == 2 ? mk(-1) : mk(0, -1);
    cplx in1 = e1 == 0 ? mk(1) : e1 == 1 ? mk(0, 1) : e1 == 2 ? mk(-1) : mk(0, -1);
    int p = flat(n, m);
    a.inc[p] = ((((4 * kPi) * sgn) * in) * dn) * dotC * ph;
    a.inc[nf + p] = ((((4 * kPi) * sgn) * in1) * dn) * dotB * ph;
  }
}

void run_incident(const IncArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(incident_kernel, dim3(1), dim3(64), 0, st, a);
}

}  // namespace omx
