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

// HIPPO Pauli repulsion radial series D_0..D_5 (FP64): D_0 = S^2, S the
// overlap of two normalized 1s Slater orbitals of exponents a = alpha_i / 2,
// b = alpha_k / 2 at distance r, and D_(n+1) = -(1/r) dD_n/dr.  In
// c = (a + b) r / 2, d = (b - a) r / 2:
//   S = P^(3/2) e^(-c) (f1(d) (1 + c) + f2(d) c^2),  P = 4 a b / (a + b)^2,
// f1(x) = sinh(x) / x = sum x^(2j) / (2j+1)!, f2(x) = (cosh x - f1) / x^2 =
// sum (2j+2) x^(2j) / (2j+3)!, both entire, so close exponents need no
// special case.  The derivatives come from Taylor arithmetic in u = r^2 / 2
// (-(1/r) d/dr = -d/du): D_n = (-1)^n n! [u^n] D_0(u0 + u).  Replaces the
// closed forms of damp_rep (include/seq/damp_hippo.h).
struct RepJet {
  double c[6];
};
__device__ __forceinline__ RepJet rj_mul(const RepJet& x, const RepJet& y) {
  RepJet z;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    z.c[q] = 0;
#pragma unroll
    for (int p = 0; p <= q; ++p) z.c[q] = fma(x.c[p], y.c[q - p], z.c[q]);
  }
  return z;
}

__device__ __forceinline__ void repel_series(double r, double ai, double ak, double* D) {
  const double a = 0.5 * ai, b = 0.5 * ak;
  const double P = 4 * a * b / ((a + b) * (a + b));
  const double P32 = P * sqrt(P);
  // r(u) = sqrt(2 u) about u0 = r^2 / 2: the square-root recurrence on the
  // jet of 2 u = r^2 + 2 e
  RepJet rj;
  rj.c[0] = r;
  {
    const double x[6] = {r * r, 2, 0, 0, 0, 0};
#pragma unroll
    for (int q = 1; q < 6; ++q) {
      double t = x[q];
#pragma unroll
      for (int p = 1; p < q; ++p) t -= rj.c[p] * rj.c[q - p];
      rj.c[q] = t / (2 * r);
    }
  }
  RepJet c, d;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    c.c[q] = 0.5 * (a + b) * rj.c[q];
    d.c[q] = 0.5 * (b - a) * rj.c[q];
  }
  const RepJet x2 = rj_mul(d, d);
  // f1 and f2 by Horner in x^2; 24 terms cover x^2 up to ~40
  constexpr int NT = 24;
  RepJet f1 = {}, f2 = {};
  {
    double fact = 1;  // (2j+1)!
    double k1[NT], k2[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      if (j > 0) fact *= (2.0 * j) * (2.0 * j + 1);
      k1[j] = 1 / fact;
      k2[j] = 1 / (fact * (2.0 * j + 3));
    }
    f1.c[0] = k1[NT - 1];
    f2.c[0] = k2[NT - 1];
    for (int j = NT - 2; j >= 0; --j) {
      f1 = rj_mul(f1, x2);
      f2 = rj_mul(f2, x2);
      f1.c[0] += k1[j];
      f2.c[0] += k2[j];
    }
  }
  // e^(-c): z' = -c' z
  RepJet ec;
  ec.c[0] = exp(-c.c[0]);
#pragma unroll
  for (int q = 1; q < 6; ++q) {
    double t = 0;
#pragma unroll
    for (int p = 1; p <= q; ++p) t -= p * c.c[p] * ec.c[q - p];
    ec.c[q] = t / q;
  }
  RepJet c1 = c;
  c1.c[0] += 1;
  RepJet poly = rj_mul(f1, c1);
  const RepJet f2c2 = rj_mul(f2, rj_mul(c, c));
#pragma unroll
  for (int q = 0; q < 6; ++q) poly.c[q] += f2c2.c[q];
  RepJet S = rj_mul(ec, poly);
#pragma unroll
  for (int q = 0; q < 6; ++q) S.c[q] *= P32;
  const RepJet D0 = rj_mul(S, S);
  double fn = 1;
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    if (n > 0) fn *= n;
    D[n] = ((n & 1) ? -fn : fn) * D0.c[n];
  }
}

}  // namespace t9
