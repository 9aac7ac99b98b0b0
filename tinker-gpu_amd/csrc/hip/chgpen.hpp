// HIPPO charge-penetration damping factors (device, FP64).
//
// A monopole is a point core charge Z plus a valence charge q - Z spread as
// a Slater density of exponent alpha, which also carries the dipole and
// quadrupole (include/seq/damp_hippo.h, include/seq/pair_mpole_chgpen.h).
// The damped kernels are lambda_(2n+1) rr_(2n+1) with
//   lambda_(2n+1) = 1 - sum_j e^-x_j P_nj(x_j) / (2n-1)!!,   x_j = alpha_j r,
// where P_0 is the leading polynomial of the kernel — core-valence 1 + x/2;
// valence-valence with equal exponents 1 + 11x/16 + 3x^2/16 + x^3/48; with
// exponents ai != ak the two parts ti^2 (1 + 2 tk + x/2) and
// tk^2 (1 + 2 ti + x/2), ti = ak^2 / (ak^2 - ai^2), tk = 1 - ti — and
// P_(n+1) = (2n+1) P_n - x P_n' + x P_n is the derivative series
// B_(n+1) = -(1/r) dB_n/dr written on the coefficients.  Exponents closer
// than 1e-3 use their mean (the two-part form cancels there).
#pragma once
#include <hip/hip_runtime.h>

namespace t9 {

template <int D0>
__device__ __forceinline__ void chgpen_terms(double x, const double* d0, double* t) {
  double d[D0 + 7];
#pragma unroll
  for (int k = 0; k < D0 + 7; ++k) d[k] = k <= D0 ? d0[k] : 0.0;
  const double ex = exp(-x);
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    double p = 0;
#pragma unroll
    for (int k = D0 + n; k >= 0; --k) p = fma(p, x, d[k]);
    t[n] = ex * p;
    const double m = 2 * n + 1;
#pragma unroll
    for (int k = D0 + n + 1; k >= 0; --k) d[k] = (k <= D0 + n ? (m - k) * d[k] : 0.0) + (k > 0 ? d[k - 1] : 0.0);
  }
}

__device__ __forceinline__ void chgpen_lambda(double r, double ai, double ak, double* lci, double* lck,
                                              double* lik) {
  const double dfac[6] = {1, 1, 3, 15, 105, 945};
  const double cv[2] = {1.0, 0.5};
  double ta[6], tb[6];
  chgpen_terms<1>(ai * r, cv, ta);
  chgpen_terms<1>(ak * r, cv, tb);
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    lci[n] = 1 - ta[n] / dfac[n];
    lck[n] = 1 - tb[n] / dfac[n];
  }
  if (fabs(ai - ak) < 1e-3) {
    const double vv[4] = {1.0, 11.0 / 16.0, 3.0 / 16.0, 1.0 / 48.0};
    chgpen_terms<3>(0.5 * (ai + ak) * r, vv, ta);
#pragma unroll
    for (int n = 0; n < 6; ++n) lik[n] = 1 - ta[n] / dfac[n];
  } else {
    const double ti = ak * ak / ((ak - ai) * (ak + ai)), tk = 1 - ti;
    const double pa[2] = {ti * ti * (1 + 2 * tk), 0.5 * ti * ti}, pb[2] = {tk * tk * (1 + 2 * ti), 0.5 * tk * tk};
    chgpen_terms<1>(ai * r, pa, ta);
    chgpen_terms<1>(ak * r, pb, tb);
#pragma unroll
    for (int n = 0; n < 6; ++n) lik[n] = 1 - (ta[n] + tb[n]) / dfac[n];
  }
}

}  // namespace t9
