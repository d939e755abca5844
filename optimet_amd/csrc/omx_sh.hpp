// omx_sh.hpp -- second-harmonic source interfaces.
#pragma once
#include "omx_kernels.hpp"

namespace omx {

struct ShArgs {
  int P, nmax;
  double omega;
  cplx bg_eps, bg_mu;            // absolute
  const double *radius;
  const cplx *eps, *mu, *eps_sh; // absolute
  const cplx *ksippp, *ksiparppar, *gamma;
  const cplx *ov;
  const double *tab;             // nine tables [9][nf][nf][nf]
  const cplx *xint;              // x_int (N); the source uses conj(x_int)
  cplx *K, *K1;                  // out: N each
};

// Result::getAbsorptionCrossSection_SH (Result.cpp:763-794): per object the
// real part of eta sigma_j sum_kk coefABS(kk), sigma_j = -i eps0 2 omega
// (eps_r_SH - 1), written to out[j]; coef holds P nf complex scratch.
struct AbsShArgs {
  int P, nmax;
  double omega;
  cplx bg_eps, bg_mu;            // absolute
  const double *radius;
  const cplx *eps, *mu, *eps_sh; // absolute; mu also at 2 omega
  const cplx *gamma;
  const double *tab;             // nine tables [9][nf][nf][nf]
  const cplx *xint, *xintsh;     // internal coefficients at omega, 2 omega (N each)
  cplx *coef;                    // scratch, P nf
  double *out;                   // P
  // the Bessel values of the four quadrature radii per object (j, dj, d2j at
  // k_j r; j, dj at k_SH r), kAbsBes complex each, computed once per object
  cplx *bes = nullptr;
  // batches (blockIdx.z = system z): omega omz[z], eps, eps_sh + z P,
  // xint, xintsh + z zvec, coef + z P nf, out + z P, bes + z P kAbsBes
  const double *omz = nullptr;
  int nb = 1;
  size_t zvec = 0;
};
constexpr int kAbsBes = 4 * 5 * (kMaxNmax + 2);

void upload_factorials();
void run_cg(int nmax, double *out, hipStream_t st);
void run_shsrc(const ShArgs &a, hipStream_t st);
void run_abssh(const AbsShArgs &a, hipStream_t st);

}  // namespace omx
