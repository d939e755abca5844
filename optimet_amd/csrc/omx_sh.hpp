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

void upload_factorials();
void run_cg(int nmax, double *out, hipStream_t st);
void run_shsrc(const ShArgs &a, hipStream_t st);

}  // namespace omx
