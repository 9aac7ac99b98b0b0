// FP64 CPU oracle for the AMOEBA electrostatics path — TEST INFRASTRUCTURE.
//
// This file is the checker, never the product: only tests/, the smoke check and
// the CPU-baseline leg of bench.py load it.  It restates the reference's
// algorithms in plain double precision, written for clarity:
//   * local frame -> global multipoles       include/seq/rotpole.h:8-224
//   * Thole / Ewald damping                  include/seq/damp.h:8-151
//   * direct / mutual field pair terms       include/seq/pair_field.h:124-420
//   * exclusion handling (scale-1 re-add)    src/acc/amoeba/fieldewald.cpp:134-230
//   * PME spread / convolution / gather      src/acc/pme.cpp:10-979, src/pme.cpp:84-117
//   * reciprocal + self field                src/cu/amoeba/field.cu:12-118
//   * preconditioned CG                      src/acc/amoeba/induce.cpp:41-292
//   * diagonal / sparse preconditioner       src/acc/amoeba/induce.cpp:296-420
//   * multipole + polarization energy,      oracle_energy.inc (own formulation,
//     gradient, virial, torques             checked against empoleewald.cpp,
//                                           epolarewald.cpp, torque.cu results)
// Pinning: every function above is checked in tests/test_oracle_golden.py and
// tests/test_oracle_energy.py against the reference's golden values
// (test/localframe.cpp, test/localframe3.cpp, test/ref/tinkernist.*) and
// against closed forms (numpy FFT, Ewald-split invariance, finite-difference
// gradients, strain-derivative virial).  Parity unpinned: none of the
// restated functions; the reference's FP32 round-off and iteration counts are
// not reproducible without a run of its GPU build.
// Pairs are enumerated exhaustively (minimum image) or through a cell list;
// every pair is visited from both atoms, so each atom's sums are written by one
// thread only.  The 3D DFT is evaluated as three separable naive DFTs.
#include <cmath>
#include <map>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtPi = 1.77245385090551602730;
constexpr double kDebye = 4.803206802;

enum { FR_NONE = 0, FR_ZONLY = 1, FR_ZTHENX = 2, FR_BISECTOR = 3, FR_ZBISECT = 4, FR_3FOLD = 5 };

struct Scale { int k; double m, d, p, u; };

struct Oracle {
  int n = 0;
  std::vector<double> xyz, pole, rpole;
  std::vector<double> pole_eff;  // local-frame pole after the chirality check
  std::vector<int> frame, zaxis, xaxis, yaxis, ysign;
  std::vector<double> polarity, pdamp, thlval;
  std::vector<int> jpolar;
  int njpolar = 0;
  std::vector<std::vector<Scale>> scale;  // both directions
  bool periodic = false;
  double box[3] = {0, 0, 0};
  bool ewald = false;
  double aewald = 0, cut = 1e30;
  int grid[3] = {0, 0, 0};
  int order = 5;
  std::vector<double> bsmod[3];
  // vdW and valence (MD terms)
  bool has_md = false;
  std::vector<int> jvdw, vparent;
  int njvdw = 0;
  std::vector<double> vrmin, veps, vred;
  std::vector<std::vector<std::pair<int, double>>> vscale;  // both directions
  double vcut = 9, vtaper = 8.1;
  std::vector<int> bond_ij, angle_abc, urey_ac;
  std::vector<double> bond_kr, angle_kt, urey_kr;
  double cbnd = 0, qbnd = 0, cang = 0, qang = 0, pang = 0, sang = 0;
  // HIPPO charge penetration (core charge, valence charge, valence alpha)
  bool chgpen = false;
  std::vector<double> pcore, pval, palpha;
};

// ------------------------------------------------------------------ helpers

inline void image(const Oracle& o, double& x, double& y, double& z) {
  if (!o.periodic) return;
  x -= o.box[0] * std::nearbyint(x / o.box[0]);
  y -= o.box[1] * std::nearbyint(y / o.box[1]);
  z -= o.box[2] * std::nearbyint(z / o.box[2]);
}

const Scale* find_scale(const Oracle& o, int i, int k) {
  for (const auto& s : o.scale[i])
    if (s.k == k) return &s;
  return nullptr;
}

// Visit every ordered pair (i, k), k != i, with minimum-image distance within
// rc, calling f(i, k, xr, yr, zr, r2) where (xr,yr,zr) = x_k - x_i.  The call for
// a given i happens on one thread.
template <class F>
void for_pairs_at(const Oracle& o, const double* X, double rc, F f) {
  const int n = o.n;
  const double rc2 = rc * rc;
  bool use_cells = o.periodic && n > 3000;
  int nc[3] = {1, 1, 1};
  if (use_cells) {
    for (int d = 0; d < 3; ++d) {
      nc[d] = (int)std::floor(o.box[d] / rc);
      if (nc[d] < 3) use_cells = false;
    }
  }
  if (!use_cells) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
      for (int k = 0; k < n; ++k) {
        if (k == i) continue;
        double xr = X[3 * k] - X[3 * i];
        double yr = X[3 * k + 1] - X[3 * i + 1];
        double zr = X[3 * k + 2] - X[3 * i + 2];
        image(o, xr, yr, zr);
        double r2 = xr * xr + yr * yr + zr * zr;
        if (r2 <= rc2) f(i, k, xr, yr, zr, r2);
      }
    }
    return;
  }
  const int ncell = nc[0] * nc[1] * nc[2];
  std::vector<int> head(ncell, -1), next(n, -1), cell(n);
  for (int i = n - 1; i >= 0; --i) {
    int c[3];
    for (int d = 0; d < 3; ++d) {
      double f0 = X[3 * i + d] / o.box[d];
      f0 -= std::floor(f0);
      c[d] = std::min(nc[d] - 1, (int)(f0 * nc[d]));
    }
    int id = c[0] + nc[0] * (c[1] + nc[1] * c[2]);
    cell[i] = id;
    next[i] = head[id];
    head[id] = i;
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    int ci = cell[i];
    int c0 = ci % nc[0], c1 = (ci / nc[0]) % nc[1], c2 = ci / (nc[0] * nc[1]);
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          int x0 = (c0 + a + nc[0]) % nc[0], x1 = (c1 + b + nc[1]) % nc[1], x2 = (c2 + c + nc[2]) % nc[2];
          for (int k = head[x0 + nc[0] * (x1 + nc[1] * x2)]; k >= 0; k = next[k]) {
            if (k == i) continue;
            double xr = X[3 * k] - X[3 * i];
            double yr = X[3 * k + 1] - X[3 * i + 1];
            double zr = X[3 * k + 2] - X[3 * i + 2];
            image(o, xr, yr, zr);
            double r2 = xr * xr + yr * yr + zr * zr;
            if (r2 <= rc2) f(i, k, xr, yr, zr, r2);
          }
        }
  }
}

template <class F>
void for_pairs(const Oracle& o, double rc, F f) {
  for_pairs_at(o, o.xyz.data(), rc, f);
}

// Thole damping of the dipole-field tensors (damp.h dmpthole3_a), gamma from
// the pair table.
inline void thole3(double r, double pdi, double pdk, double pga, double& s3, double& s5, double& s7) {
  double damp = pdi * pdk;
  if (damp == 0) {
    s3 = s5 = s7 = 1.0;
    return;
  }
  double ratio = r / damp;
  double a = -pga * ratio * ratio * ratio;
  double e = std::exp(a);
  s3 = 1 - e;
  s5 = 1 - e * (1 - a);
  s7 = 1 - e * (1 - a + 0.6 * a * a);
}

// Ewald real-space radial functions bn[0..nb-1] (damp.h damp_ewald).
inline void ewald_bn(double* bn, int nb, double r, double aewald) {
  double rinv = 1 / r, rr2 = rinv * rinv;
  double ra = aewald * r;
  bn[0] = std::erfc(ra) * rinv;
  double alsq2 = 2 * aewald * aewald;
  double alsq2n = 1 / (kSqrtPi * aewald);
  double ex = std::exp(-ra * ra);
  for (int j = 1; j < nb; ++j) {
    alsq2n *= alsq2;
    bn[j] = ((2 * j - 1) * bn[j - 1] + alsq2n * ex) * rr2;
  }
}

// ------------------------------------------------------------------ rotpole

void rotpole(Oracle& o) {
  const int n = o.n;
  o.rpole.assign(10 * n, 0.0);
  std::vector<double> pole = o.pole;
  // chirality check for Z-then-X frames with a y-defining atom
  for (int i = 0; i < n; ++i) {
    if (o.frame[i] != FR_ZTHENX || o.yaxis[i] < 0 || o.ysign[i] == 0) continue;
    const double* X = o.xyz.data();
    int ia = i, ib = o.zaxis[i], ic = o.xaxis[i], id = o.yaxis[i];
    double xad = X[3 * ia] - X[3 * id], yad = X[3 * ia + 1] - X[3 * id + 1], zad = X[3 * ia + 2] - X[3 * id + 2];
    double xbd = X[3 * ib] - X[3 * id], ybd = X[3 * ib + 1] - X[3 * id + 1], zbd = X[3 * ib + 2] - X[3 * id + 2];
    double xcd = X[3 * ic] - X[3 * id], ycd = X[3 * ic + 1] - X[3 * id + 1], zcd = X[3 * ic + 2] - X[3 * id + 2];
    double c1 = ybd * zcd - zbd * ycd, c2 = ycd * zad - zcd * yad, c3 = yad * zbd - zad * ybd;
    double vol = xad * c1 + xbd * c2 + xcd * c3;
    int k = o.ysign[i];
    if ((k < 0 && vol > 0) || (k > 0 && vol < 0)) {
      pole[10 * i + 2] = -pole[10 * i + 2];
      pole[10 * i + 7] = -pole[10 * i + 7];
      pole[10 * i + 9] = -pole[10 * i + 9];
    }
  }
  o.pole_eff = pole;
  for (int i = 0; i < n; ++i) {
    double a[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int fr = o.frame[i];
    if (fr != FR_NONE) {
      auto vec = [&](int j, double* v) {
        v[0] = o.xyz[3 * j] - o.xyz[3 * i];
        v[1] = o.xyz[3 * j + 1] - o.xyz[3 * i + 1];
        v[2] = o.xyz[3 * j + 2] - o.xyz[3 * i + 2];
        double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= r;
        v[1] /= r;
        v[2] /= r;
      };
      auto norm = [](double* v) {
        double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= r;
        v[1] /= r;
        v[2] /= r;
      };
      double* xx = a[0];
      double* zz = a[2];
      double yy[3];
      vec(o.zaxis[i], zz);
      if (fr == FR_ZONLY) {
        bool ok = !(std::fabs(zz[0]) > 0.866);
        xx[0] = ok ? 1 : 0;
        xx[1] = ok ? 0 : 1;
        xx[2] = 0;
      } else {
        vec(o.xaxis[i], xx);
      }
      if (fr == FR_ZBISECT || fr == FR_3FOLD) vec(o.yaxis[i], yy);
      if (fr == FR_BISECTOR) {
        for (int d = 0; d < 3; ++d) zz[d] += xx[d];
        norm(zz);
      } else if (fr == FR_ZBISECT) {
        for (int d = 0; d < 3; ++d) xx[d] += yy[d];
        norm(xx);
      } else if (fr == FR_3FOLD) {
        for (int d = 0; d < 3; ++d) zz[d] += xx[d] + yy[d];
        norm(zz);
      }
      double dot = xx[0] * zz[0] + xx[1] * zz[1] + xx[2] * zz[2];
      for (int d = 0; d < 3; ++d) xx[d] -= dot * zz[d];
      norm(xx);
      a[1][0] = zz[1] * xx[2] - zz[2] * xx[1];
      a[1][1] = zz[2] * xx[0] - zz[0] * xx[2];
      a[1][2] = zz[0] * xx[1] - zz[1] * xx[0];
    }
    // rows of a are the local axes in the global frame: global = a^T local
    const double* p = &pole[10 * i];
    double* rp = &o.rpole[10 * i];
    rp[0] = p[0];
    for (int g = 0; g < 3; ++g) rp[1 + g] = a[0][g] * p[1] + a[1][g] * p[2] + a[2][g] * p[3];
    double m[3][3] = {{p[4], p[7], p[8]}, {p[7], p[5], p[9]}, {p[8], p[9], p[6]}};
    double r[3][3];
    for (int g = 0; g < 3; ++g)
      for (int h = 0; h < 3; ++h) {
        double s = 0;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) s += a[k][g] * a[l][h] * m[k][l];
        r[g][h] = s;
      }
    rp[4] = r[0][0];
    rp[5] = r[1][1];
    rp[6] = r[2][2];
    rp[7] = r[0][1];
    rp[8] = r[0][2];
    rp[9] = r[1][2];
  }
}

// ------------------------------------------------------------------ PME

// Cardinal B-spline M_p(x), support (0, p).
double bspline_M(int p, double x) {
  if (x <= 0 || x >= p) return 0.0;
  if (p == 1) return 1.0;
  if (p == 2) return 1.0 - std::fabs(x - 1.0);
  return (x * bspline_M(p - 1, x) + (p - x) * bspline_M(p - 1, x - 1.0)) / (p - 1);
}

// r-th derivative of M_p at x: sum_s (-1)^s C(r,s) M_{p-r}(x-s)
double bspline_dM(int p, int r, double x) {
  double c = 1, s = 0;
  for (int k = 0; k <= r; ++k) {
    if (k > 0) c = c * (r - k + 1) / k;
    s += ((k & 1) ? -c : c) * bspline_M(p - r, x - k);
  }
  return s;
}

void build_bsmod(Oracle& o) {
  const int p = o.order;
  for (int d = 0; d < 3; ++d) {
    const int K = o.grid[d];
    std::vector<double> arr(K, 0.0), mod(K, 0.0);
    for (int k = 1; k < p && k < K; ++k) arr[k] = bspline_M(p, (double)k);
    for (int m = 0; m < K; ++m) {
      double s1 = 0, s2 = 0;
      for (int j = 0; j < K; ++j) {
        double arg = 2 * kPi * (double)((long long)m * j % K) / K;
        s1 += arr[j] * std::cos(arg);
        s2 += arr[j] * std::sin(arg);
      }
      mod[m] = s1 * s1 + s2 * s2;
    }
    const double eps = 1e-7;
    if (mod[0] < eps) mod[0] = 0.5 * mod[1];
    for (int m = 1; m < K - 1; ++m)
      if (mod[m] < eps) mod[m] = 0.5 * (mod[m - 1] + mod[m + 1]);
    if (mod[K - 1] < eps) mod[K - 1] = 0.5 * mod[K - 2];
    // optimal zeta factor (exponential Euler spline interpolation)
    for (int m = 0; m < K; ++m) {
      int k = m;
      if (m + 1 > K / 2) k = m - K;
      double zeta = 1.0;
      if (k != 0) {
        double s1 = 1, s2 = 1, fac = kPi * k / K;
        for (int j = 1; j <= 50; ++j) {
          double a1 = fac / (fac + kPi * j);
          double a2 = fac / (fac - kPi * j);
          s1 += std::pow(a1, p) + std::pow(a2, p);
          s2 += std::pow(a1, 2 * p) + std::pow(a2, 2 * p);
        }
        zeta = s2 / s1;
      }
      mod[m] *= zeta * zeta;
    }
    o.bsmod[d] = mod;
  }
}

struct Spline {
  int base[3];
  double th[3][4][5];  // dim, derivative level 0..3, point 0..order-1
};

std::vector<Spline> make_splines(const Oracle& o) {
  std::vector<Spline> sp(o.n);
  const int p = o.order;
  for (int i = 0; i < o.n; ++i) {
    for (int d = 0; d < 3; ++d) {
      double w = o.xyz[3 * i + d] / o.box[d];
      w = w + 0.5 - std::floor(w + 0.5);  // [0,1)
      double fr = o.grid[d] * w;
      int ig = (int)std::floor(fr);
      double f = fr - ig;
      int b = ig - p + 1;
      if (b < 0) b += o.grid[d];
      sp[i].base[d] = b;
      for (int m = 0; m < p; ++m) {
        double x = f + p - 1 - m;
        for (int lv = 0; lv < 4; ++lv) sp[i].th[d][lv][m] = bspline_dM(p, lv, x);
      }
    }
  }
  return sp;
}

// Matrix taking Cartesian gradients to fractional-grid gradients for an
// orthorhombic box: a[d] = K_d / L_d.
void frac_scale(const Oracle& o, double* a) {
  for (int d = 0; d < 3; ++d) a[d] = o.grid[d] / o.box[d];
}

// In-place separable 3D DFT, sign -1 (forward) or +1 (backward, unnormalised).
void dft3(const Oracle& o, std::vector<cplx>& g, int sign) {
  const int K[3] = {o.grid[0], o.grid[1], o.grid[2]};
  for (int d = 0; d < 3; ++d) {
    const int Kd = K[d];
    std::vector<cplx> tw(Kd);
    for (int j = 0; j < Kd; ++j) tw[j] = std::polar(1.0, sign * 2 * kPi * j / Kd);
    const long stride = d == 0 ? 1 : (d == 1 ? K[0] : (long)K[0] * K[1]);
    const long nlines = (long)K[0] * K[1] * K[2] / Kd;
#pragma omp parallel for
    for (long l = 0; l < nlines; ++l) {
      long base;
      if (d == 0) base = l * K[0];
      else if (d == 1) base = (l % K[0]) + (l / K[0]) * (long)K[0] * K[1];
      else base = l;
      std::vector<cplx> in(Kd), out(Kd);
      for (int j = 0; j < Kd; ++j) in[j] = g[base + j * stride];
      for (int m = 0; m < Kd; ++m) {
        cplx s = 0;
        for (int j = 0; j < Kd; ++j) s += in[j] * tw[(long)m * j % Kd];
        out[m] = s;
      }
      for (int j = 0; j < Kd; ++j) g[base + j * stride] = out[j];
    }
  }
}

// Multiply by the reciprocal-space influence function; returns
// sum_k expterm |Q(k)|^2 * 0.5 (multiply by the Coulomb factor for energy).
double pme_conv(const Oracle& o, std::vector<cplx>& g, double* vir6) {
  const int K1 = o.grid[0], K2 = o.grid[1], K3 = o.grid[2];
  const double vol = o.box[0] * o.box[1] * o.box[2];
  const double pterm = (kPi / o.aewald) * (kPi / o.aewald);
  double e = 0, v[6] = {0, 0, 0, 0, 0, 0};
  for (int k3 = 0; k3 < K3; ++k3)
    for (int k2 = 0; k2 < K2; ++k2)
      for (int k1 = 0; k1 < K1; ++k1) {
        long idx = k1 + (long)K1 * (k2 + (long)K2 * k3);
        if (idx == 0) {
          g[0] = 0;
          continue;
        }
        int r1 = k1 < (K1 + 1) / 2 ? k1 : k1 - K1;
        int r2 = k2 < (K2 + 1) / 2 ? k2 : k2 - K2;
        int r3 = k3 < (K3 + 1) / 2 ? k3 : k3 - K3;
        double h1 = r1 / o.box[0], h2 = r2 / o.box[1], h3 = r3 / o.box[2];
        double hsq = h1 * h1 + h2 * h2 + h3 * h3;
        double term = -pterm * hsq;
        double ex = 0;
        if (term > -50) {
          double denom = hsq * kPi * vol * o.bsmod[0][k1] * o.bsmod[1][k2] * o.bsmod[2][k3];
          ex = std::exp(term) / denom;
          double struc2 = std::norm(g[idx]);
          double et = 0.5 * ex * struc2;
          e += et;
          double vt = (2 / hsq) * (1 - term) * et;
          v[0] += h1 * h1 * vt - et;
          v[1] += h1 * h2 * vt;
          v[2] += h1 * h3 * vt;
          v[3] += h2 * h2 * vt - et;
          v[4] += h2 * h3 * vt;
          v[5] += h3 * h3 * vt - et;
        }
        g[idx] *= ex;
      }
  if (vir6)
    for (int j = 0; j < 6; ++j) vir6[j] = v[j];
  return e;
}

// Spread fractional multipoles (10 comps, real part) or fractional dipoles
// (d -> real part, p -> imaginary part).
void spread(const Oracle& o, const std::vector<Spline>& sp, const double* fmp, const double* fud,
            const double* fup, std::vector<cplx>& g) {
  const int K1 = o.grid[0], K2 = o.grid[1], K3 = o.grid[2];
  const int p = o.order;
  g.assign((size_t)K1 * K2 * K3, 0.0);
  for (int i = 0; i < o.n; ++i) {
    const Spline& s = sp[i];
    for (int c = 0; c < p; ++c) {
      int z = (s.base[2] + c) % K3;
      double v0 = s.th[2][0][c], v1 = s.th[2][1][c], v2 = s.th[2][2][c];
      for (int b = 0; b < p; ++b) {
        int y = (s.base[1] + b) % K2;
        double u0 = s.th[1][0][b], u1 = s.th[1][1][b], u2 = s.th[1][2][b];
        for (int a = 0; a < p; ++a) {
          int x = (s.base[0] + a) % K1;
          double t0 = s.th[0][0][a], t1 = s.th[0][1][a], t2 = s.th[0][2][a];
          long idx = x + (long)K1 * (y + (long)K2 * z);
          if (fmp) {
            const double* f = fmp + 10 * i;
            // f: q, x, y, z, xx, yy, zz, xy, xz, yz (fractional)
            double val = f[0] * t0 * u0 * v0 + f[1] * t1 * u0 * v0 + f[2] * t0 * u1 * v0 + f[3] * t0 * u0 * v1 +
                         f[4] * t2 * u0 * v0 + f[5] * t0 * u2 * v0 + f[6] * t0 * u0 * v2 + f[7] * t1 * u1 * v0 +
                         f[8] * t1 * u0 * v1 + f[9] * t0 * u1 * v1;
            g[idx] += cplx(val, 0.0);
          } else {
            const double* fd = fud + 3 * i;
            const double* fp = fup + 3 * i;
            double vd = fd[0] * t1 * u0 * v0 + fd[1] * t0 * u1 * v0 + fd[2] * t0 * u0 * v1;
            double vp = fp[0] * t1 * u0 * v0 + fp[1] * t0 * u1 * v0 + fp[2] * t0 * u0 * v1;
            g[idx] += cplx(vd, vp);
          }
        }
      }
    }
  }
}

// Gather potential derivatives in fractional coordinates, 20 components in
// the order 0, x, y, z, xx, yy, zz, xy, xz, yz, xxx, yyy, zzz, xxy, xxz, xyy,
// yyz, xzz, yzz, xyz; `imag` selects the imaginary part of the grid.
void gather(const Oracle& o, const std::vector<Spline>& sp, const std::vector<cplx>& g, bool imag,
            double* fphi) {
  const int K1 = o.grid[0], K2 = o.grid[1], K3 = o.grid[2];
  const int p = o.order;
#pragma omp parallel for
  for (int i = 0; i < o.n; ++i) {
    const Spline& s = sp[i];
    double t[20] = {0};
    for (int c = 0; c < p; ++c) {
      int z = (s.base[2] + c) % K3;
      const double* V = &s.th[2][0][0];
      double v[4] = {s.th[2][0][c], s.th[2][1][c], s.th[2][2][c], s.th[2][3][c]};
      (void)V;
      for (int b = 0; b < p; ++b) {
        int y = (s.base[1] + b) % K2;
        double u[4] = {s.th[1][0][b], s.th[1][1][b], s.th[1][2][b], s.th[1][3][b]};
        for (int a = 0; a < p; ++a) {
          int x = (s.base[0] + a) % K1;
          double w[4] = {s.th[0][0][a], s.th[0][1][a], s.th[0][2][a], s.th[0][3][a]};
          long idx = x + (long)K1 * (y + (long)K2 * z);
          double q = imag ? g[idx].imag() : g[idx].real();
          t[0] += q * w[0] * u[0] * v[0];
          t[1] += q * w[1] * u[0] * v[0];
          t[2] += q * w[0] * u[1] * v[0];
          t[3] += q * w[0] * u[0] * v[1];
          t[4] += q * w[2] * u[0] * v[0];
          t[5] += q * w[0] * u[2] * v[0];
          t[6] += q * w[0] * u[0] * v[2];
          t[7] += q * w[1] * u[1] * v[0];
          t[8] += q * w[1] * u[0] * v[1];
          t[9] += q * w[0] * u[1] * v[1];
          t[10] += q * w[3] * u[0] * v[0];
          t[11] += q * w[0] * u[3] * v[0];
          t[12] += q * w[0] * u[0] * v[3];
          t[13] += q * w[2] * u[1] * v[0];
          t[14] += q * w[2] * u[0] * v[1];
          t[15] += q * w[1] * u[2] * v[0];
          t[16] += q * w[0] * u[2] * v[1];
          t[17] += q * w[1] * u[0] * v[2];
          t[18] += q * w[0] * u[1] * v[2];
          t[19] += q * w[1] * u[1] * v[1];
        }
      }
    }
    for (int j = 0; j < 20; ++j) fphi[20 * i + j] = t[j];
  }
}

// Cartesian multipoles -> fractional (orthorhombic): derivative order n picks
// up a^n; the Cartesian off-diagonal quadrupoles enter with a factor 2.
void cart_to_frac_mpole(const Oracle& o, const double* rp, double* fmp) {
  double a[3];
  frac_scale(o, a);
  for (int i = 0; i < o.n; ++i) {
    const double* c = rp + 10 * i;
    double* f = fmp + 10 * i;
    f[0] = c[0];
    f[1] = c[1] * a[0];
    f[2] = c[2] * a[1];
    f[3] = c[3] * a[2];
    f[4] = c[4] * a[0] * a[0];
    f[5] = c[5] * a[1] * a[1];
    f[6] = c[6] * a[2] * a[2];
    f[7] = 2 * c[7] * a[0] * a[1];
    f[8] = 2 * c[8] * a[0] * a[2];
    f[9] = 2 * c[9] * a[1] * a[2];
  }
}

// Reciprocal-space part of the permanent field (without the self term):
// field_i = -grad phi_i.  Also returns fphi (fractional, 20 per atom) and the
// reciprocal energy sum 0.5 * sum expterm |Q|^2 (unscaled by the Coulomb factor).
double recip_mpole(const Oracle& o, double* field, std::vector<double>* fphi_out) {
  auto sp = make_splines(o);
  std::vector<double> fmp(10 * o.n);
  cart_to_frac_mpole(o, o.rpole.data(), fmp.data());
  std::vector<cplx> g;
  spread(o, sp, fmp.data(), nullptr, nullptr, g);
  dft3(o, g, -1);
  double e = pme_conv(o, g, nullptr);
  dft3(o, g, +1);
  std::vector<double> fphi(20 * o.n);
  gather(o, sp, g, false, fphi.data());
  double a[3];
  frac_scale(o, a);
  if (field)
    for (int i = 0; i < o.n; ++i)
      for (int d = 0; d < 3; ++d) field[3 * i + d] = -a[d] * fphi[20 * i + 1 + d];
  if (fphi_out) *fphi_out = fphi;
  return e;
}

// Reciprocal-space field of two dipole sets (d real, p imaginary).
void recip_dipole(const Oracle& o, const double* ud, const double* up, double* fd, double* fp) {
  auto sp = make_splines(o);
  double a[3];
  frac_scale(o, a);
  std::vector<double> fud(3 * o.n), fup(3 * o.n);
  for (int i = 0; i < o.n; ++i)
    for (int d = 0; d < 3; ++d) {
      fud[3 * i + d] = ud[3 * i + d] * a[d];
      fup[3 * i + d] = up[3 * i + d] * a[d];
    }
  std::vector<cplx> g;
  spread(o, sp, nullptr, fud.data(), fup.data(), g);
  dft3(o, g, -1);
  pme_conv(o, g, nullptr);
  dft3(o, g, +1);
  std::vector<double> phd(20 * o.n), php(20 * o.n);
  gather(o, sp, g, false, phd.data());
  gather(o, sp, g, true, php.data());
  for (int i = 0; i < o.n; ++i)
    for (int d = 0; d < 3; ++d) {
      fd[3 * i + d] = -a[d] * phd[20 * i + 1 + d];
      fp[3 * i + d] = -a[d] * php[20 * i + 1 + d];
    }
}

// ------------------------------------------------------------ HIPPO damping

// Charge penetration (HIPPO; reference: include/seq/pair_mpole_chgpen.h,
// include/seq/damp_hippo.h).  A monopole is a point core charge Z plus a
// valence charge q - Z spread as a Slater density (alpha^3 / 8 pi)
// exp(-alpha r) that also carries the dipole and quadrupole.  The damped
// Coulomb kernels are, with a = alpha_i r, b = alpha_k r,
//   core-valence      f0 = [1 - (1 + a/2) e^-a] / r,
//   valence-valence   f0 = [1 - (1 + 11a/16 + 3a^2/16 + a^3/48) e^-a] / r
//                          (alpha_i = alpha_k),
//                     f0 = [1 - ti^2 (1 + 2 tk + a/2) e^-a
//                             - tk^2 (1 + 2 ti + b/2) e^-b] / r,
//                     ti = alpha_k^2 / (alpha_k^2 - alpha_i^2), tk = 1 - ti,
// (the electrostatic energy of a point charge, resp. a second density, in a
// Slater density).  Every higher kernel follows from the derivative series
// f_(n+1) = -(1/r) df_n/dr, computed here exactly on the polynomial
// coefficients: f_n = [(2n-1)!! - sum_j P_nj(r) exp(-alpha_j r)] / r^(2n+1)
// with P_(n+1),j = m P - r P' + alpha_j r P (m = 2n + 1), so
// lambda_(2n+1) = 1 - sum_j P_nj(r) exp(-alpha_j r) / (2n-1)!!.  Exponents
// closer than 1e-3 are replaced by their mean (the rule of Tinker's damppole:
// the two-exponent form cancels there).
struct ExpPart {
  double alpha;
  double c[16];
  int deg;
};

void damp_series(double r, int nj, ExpPart* parts, int nmax, double* lam) {
  double A = 1;
  for (int n = 0; n <= nmax; ++n) {
    double s = 0;
    for (int j = 0; j < nj; ++j) {
      double p = 0, rk = 1;
      for (int k = 0; k <= parts[j].deg; ++k) {
        p += parts[j].c[k] * rk;
        rk *= r;
      }
      s += p * std::exp(-parts[j].alpha * r);
    }
    lam[n] = 1 - s / A;
    const int m = 2 * n + 1;
    for (int j = 0; j < nj; ++j) {
      ExpPart& q = parts[j];
      for (int k = q.deg + 1; k >= 0; --k)
        q.c[k] = (k <= q.deg ? (m - k) * q.c[k] : 0.0) + (k > 0 ? q.alpha * q.c[k - 1] : 0.0);
      q.deg += 1;
    }
    A *= m;
  }
}

// lam[0..5] = lambda_1 .. lambda_11 of the core-valence kernel (exponent a)
void damp_core_valence(double r, double a, double* lam) {
  ExpPart p{a, {1.0, 0.5 * a}, 1};
  damp_series(r, 1, &p, 5, lam);
}

// lam[0..5] of the valence-valence kernel (exponents ai, ak)
void damp_valence_valence(double r, double ai, double ak, double* lam) {
  if (std::fabs(ai - ak) < 1e-3) {
    const double a = 0.5 * (ai + ak);
    ExpPart p{a, {1.0, 11.0 / 16.0 * a, 3.0 / 16.0 * a * a, a * a * a / 48.0}, 3};
    damp_series(r, 1, &p, 5, lam);
    return;
  }
  const double ti = ak * ak / (ak * ak - ai * ai), tk = 1 - ti;
  ExpPart p[2] = {{ai, {ti * ti * (1 + 2 * tk), ti * ti * 0.5 * ai}, 1},
                  {ak, {tk * tk * (1 + 2 * ti), tk * tk * 0.5 * ak}, 1}};
  damp_series(r, 2, p, 5, lam);
}

// ------------------------------------------------------------------ fields

double pair_thole(const Oracle& o, int i, int k) { return o.thlval[o.jpolar[i] * o.njpolar + o.jpolar[k]]; }

// Direct field (permanent multipoles): d and p scaling.
void dfield(const Oracle& o, double* fd, double* fp) {
  const int n = o.n;
  std::fill(fd, fd + 3 * n, 0.0);
  std::fill(fp, fp + 3 * n, 0.0);
  const double rc = o.ewald ? o.cut : o.cut;
  for_pairs(o, rc, [&](int i, int k, double xr, double yr, double zr, double r2) {
    const double* pk = &o.rpole[10 * k];
    double r = std::sqrt(r2), rinv = 1 / r, rr2 = rinv * rinv;
    double rr3 = rinv * rr2, rr5 = 3 * rr3 * rr2, rr7 = 5 * rr5 * rr2;
    double s3, s5, s7;
    thole3(r, o.pdamp[i], o.pdamp[k], pair_thole(o, i, k), s3, s5, s7);
    const Scale* sc = find_scale(o, i, k);
    double dsc = sc ? sc->d : 1.0, psc = sc ? sc->p : 1.0;
    double bn[4] = {0, 0, 0, 0};
    if (o.ewald) ewald_bn(bn, 4, r, o.aewald);
    double dkr = pk[1] * xr + pk[2] * yr + pk[3] * zr;
    double qkx = pk[4] * xr + pk[7] * yr + pk[8] * zr;
    double qky = pk[7] * xr + pk[5] * yr + pk[9] * zr;
    double qkz = pk[8] * xr + pk[9] * yr + pk[6] * zr;
    double qkr = qkx * xr + qky * yr + qkz * zr;
    // field of charge q, dipole and quadrupole (m = 1) or of q alone (m = 0)
    // with damping factors l3, l5, l7
    auto add = [&](double sc_, double* out, double q, double m, double l3, double l5, double l7) {
      double b1, b2, b3;
      if (o.ewald) {
        b1 = bn[1] - (1 - sc_ * l3) * rr3;
        b2 = bn[2] - (1 - sc_ * l5) * rr5;
        b3 = bn[3] - (1 - sc_ * l7) * rr7;
      } else {
        b1 = sc_ * l3 * rr3;
        b2 = sc_ * l5 * rr5;
        b3 = sc_ * l7 * rr7;
      }
      double c1 = -(b1 * q - m * (b2 * dkr - b3 * qkr));
      out[3 * i + 0] += c1 * xr - m * (b1 * pk[1] - 2 * b2 * qkx);
      out[3 * i + 1] += c1 * yr - m * (b1 * pk[2] - 2 * b2 * qky);
      out[3 * i + 2] += c1 * zr - m * (b1 * pk[3] - 2 * b2 * qkz);
    };
    if (o.chgpen) {
      // HIPPO: the induced dipole is a point probe; k's core charge acts
      // undamped, k's valence density (charge, dipole, quadrupole) with its
      // core-valence damping
      double lck[6];
      damp_core_valence(r, o.palpha[k], lck);
      for (int t = 0; t < 2; ++t) {
        const double sc_ = t == 0 ? dsc : psc;
        double* out = t == 0 ? fd : fp;
        add(sc_, out, o.pcore[k], 0.0, 1.0, 1.0, 1.0);
        add(sc_, out, o.pval[k], 1.0, lck[1], lck[2], lck[3]);
      }
      return;
    }
    add(dsc, fd, pk[0], 1.0, s3, s5, s7);
    add(psc, fp, pk[0], 1.0, s3, s5, s7);
  });
  if (o.ewald) {
    std::vector<double> rf(3 * n);
    recip_mpole(o, rf.data(), nullptr);
    const double term = 4.0 / 3.0 * o.aewald * o.aewald * o.aewald / kSqrtPi;
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        double add = rf[3 * i + d] + term * o.rpole[10 * i + 1 + d];
        fd[3 * i + d] += add;
        fp[3 * i + d] += add;
      }
  }
}

// Mutual field of induced dipoles (u scaling).
void ufield(const Oracle& o, const double* ud, const double* up, double* fd, double* fp) {
  const int n = o.n;
  std::fill(fd, fd + 3 * n, 0.0);
  std::fill(fp, fp + 3 * n, 0.0);
  for_pairs(o, o.cut, [&](int i, int k, double xr, double yr, double zr, double r2) {
    double r = std::sqrt(r2), rinv = 1 / r, rr2 = rinv * rinv;
    double rr3 = rinv * rr2, rr5 = 3 * rr3 * rr2, rr7 = 5 * rr5 * rr2;
    double s3, s5, s7;
    thole3(r, o.pdamp[i], o.pdamp[k], pair_thole(o, i, k), s3, s5, s7);
    (void)rr7;
    (void)s7;
    const Scale* sc = find_scale(o, i, k);
    double usc = sc ? sc->u : 1.0;
    if (o.chgpen) {  // HIPPO: valence-valence damping of the dipole-dipole tensor
      double lik[6];
      damp_valence_valence(r, o.palpha[i], o.palpha[k], lik);
      s3 = lik[1];
      s5 = lik[2];
    }
    double b1, b2;
    if (o.ewald) {
      double bn[3];
      ewald_bn(bn, 3, r, o.aewald);
      b1 = bn[1] - (1 - usc * s3) * rr3;
      b2 = bn[2] - (1 - usc * s5) * rr5;
    } else {
      b1 = usc * s3 * rr3;
      b2 = usc * s5 * rr5;
    }
    const double* dk = ud + 3 * k;
    const double* pk = up + 3 * k;
    double cd = b2 * (dk[0] * xr + dk[1] * yr + dk[2] * zr);
    double cp = b2 * (pk[0] * xr + pk[1] * yr + pk[2] * zr);
    fd[3 * i + 0] += cd * xr - b1 * dk[0];
    fd[3 * i + 1] += cd * yr - b1 * dk[1];
    fd[3 * i + 2] += cd * zr - b1 * dk[2];
    fp[3 * i + 0] += cp * xr - b1 * pk[0];
    fp[3 * i + 1] += cp * yr - b1 * pk[1];
    fp[3 * i + 2] += cp * zr - b1 * pk[2];
  });
  if (o.ewald) {
    std::vector<double> rd(3 * n), rp(3 * n);
    recip_dipole(o, ud, up, rd.data(), rp.data());
    const double term = 4.0 / 3.0 * o.aewald * o.aewald * o.aewald / kSqrtPi;
    for (int i = 0; i < 3 * n; ++i) {
      fd[i] += rd[i] + term * ud[i];
      fp[i] += rp[i] + term * up[i];
    }
  }
}

// Preconditioner z = M r.
void precond(const Oracle& o, bool sparse, double ucut, double udiag, const double* rd, const double* rp,
             double* zd, double* zp) {
  const int n = o.n;
  const double f = sparse ? udiag : 1.0;
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      zd[3 * i + d] = f * o.polarity[i] * rd[3 * i + d];
      zp[3 * i + d] = f * o.polarity[i] * rp[3 * i + d];
    }
  if (!sparse) return;
  std::vector<double> ad(3 * n, 0.0), ap(3 * n, 0.0);
  for_pairs(o, ucut, [&](int i, int k, double xr, double yr, double zr, double r2) {
    double r = std::sqrt(r2);
    double s3, s5, s7;
    thole3(r, o.pdamp[i], o.pdamp[k], pair_thole(o, i, k), s3, s5, s7);
    const Scale* sc = find_scale(o, i, k);
    double usc = sc ? sc->u : 1.0;
    s3 *= usc;
    s5 *= usc;
    double polik = o.polarity[i] * o.polarity[k];
    double rr3 = s3 * polik / (r * r2);
    double rr5 = 3 * s5 * polik / (r * r2 * r2);
    double m0 = rr5 * xr * xr - rr3, m1 = rr5 * xr * yr, m2 = rr5 * xr * zr;
    double m3 = rr5 * yr * yr - rr3, m4 = rr5 * yr * zr, m5 = rr5 * zr * zr - rr3;
    const double* a = rd + 3 * k;
    const double* b = rp + 3 * k;
    ad[3 * i + 0] += m0 * a[0] + m1 * a[1] + m2 * a[2];
    ad[3 * i + 1] += m1 * a[0] + m3 * a[1] + m4 * a[2];
    ad[3 * i + 2] += m2 * a[0] + m4 * a[1] + m5 * a[2];
    ap[3 * i + 0] += m0 * b[0] + m1 * b[1] + m2 * b[2];
    ap[3 * i + 1] += m1 * b[0] + m3 * b[1] + m4 * b[2];
    ap[3 * i + 2] += m2 * b[0] + m4 * b[1] + m5 * b[2];
  });
  for (int i = 0; i < 3 * n; ++i) {
    zd[i] += ad[i];
    zp[i] += ap[i];
  }
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

#include "oracle_energy.inc"
#include "oracle_md.inc"

}  // namespace

// ==================================================================== C ABI
// HIPPO dispersion damping of two Slater densities of exponents ai, ak at
// distance r, D and dD/dr.  With d = a r, t_i = a_k^2 / (a_k^2 - a_i^2),
// t_k = 1 - t_i and P(d, b) = b (1 + d + d^2/2) + d^3/4:
//   D = 1 - t_i^2 P(d_i, 1 + 2 t_k) e^(-d_i) - t_k^2 P(d_k, 1 + 2 t_i) e^(-d_k);
// for exponents within 1e-3, the limit at their mean exponent:
//   D = 1 - (1 + d + d^2/2 + 17 d^3/96 + 5 d^4/96 + d^5/96) e^(-d).
// The derivatives come from d/dd [P e^(-d)] = (P' - P) e^(-d).
void damp_disp(double r, double ai, double ak, double* D, double* dD) {
  if (std::fabs(ai - ak) > 1e-3) {
    const double ti = ak * ak / (ak * ak - ai * ai), tk = 1 - ti;
    const double ex[2] = {ai, ak}, tt[2] = {ti, tk}, bb[2] = {1 + 2 * tk, 1 + 2 * ti};
    double d0 = 1, d1 = 0;
    for (int j = 0; j < 2; ++j) {
      const double d = ex[j] * r, bj = bb[j];
      const double P = bj * (1 + d + 0.5 * d * d) + 0.25 * d * d * d;
      const double Pp = bj * (1 + d) + 0.75 * d * d;
      d0 -= tt[j] * tt[j] * P * std::exp(-d);
      d1 -= tt[j] * tt[j] * (Pp - P) * std::exp(-d) * ex[j];
    }
    *D = d0;
    *dD = d1;
  } else {
    const double a = 0.5 * (ai + ak), d = a * r;
    const double cP[6] = {1, 1, 0.5, 17.0 / 96, 5.0 / 96, 1.0 / 96};
    double P = 0, Pp = 0;
    for (int j = 5; j >= 0; --j) P = P * d + cP[j];
    for (int j = 5; j >= 1; --j) Pp = Pp * d + j * cP[j];
    *D = 1 - P * std::exp(-d);
    *dD = -(Pp - P) * std::exp(-d) * a;
  }
}

extern "C" {

void* orc_create(int n, const double* xyz, const double* pole, const int* frame, const int* zaxis,
                 const int* xaxis, const int* yaxis, const int* ysign, const double* polarity,
                 const double* pdamp, const int* jpolar, int njpolar, const double* thlval, int nscale,
                 const int* scale_ik, const double* scale_mdpu, const double* box, int ewald, double aewald,
                 double cut, const int* grid, int order) {
  Oracle* o = new Oracle;
  o->n = n;
  o->xyz.assign(xyz, xyz + 3 * n);
  o->pole.assign(pole, pole + 10 * n);
  o->frame.assign(frame, frame + n);
  o->zaxis.assign(zaxis, zaxis + n);
  o->xaxis.assign(xaxis, xaxis + n);
  o->yaxis.assign(yaxis, yaxis + n);
  o->ysign.assign(ysign, ysign + n);
  o->polarity.assign(polarity, polarity + n);
  o->pdamp.assign(pdamp, pdamp + n);
  o->jpolar.assign(jpolar, jpolar + n);
  o->njpolar = njpolar;
  o->thlval.assign(thlval, thlval + njpolar * njpolar);
  o->scale.assign(n, {});
  for (int s = 0; s < nscale; ++s) {
    int i = scale_ik[2 * s], k = scale_ik[2 * s + 1];
    const double* v = scale_mdpu + 4 * s;
    o->scale[i].push_back({k, v[0], v[1], v[2], v[3]});
    o->scale[k].push_back({i, v[0], v[1], v[2], v[3]});
  }
  if (box) {
    o->periodic = true;
    for (int d = 0; d < 3; ++d) o->box[d] = box[d];
  }
  o->ewald = ewald != 0;
  o->aewald = aewald;
  o->cut = cut;
  if (grid)
    for (int d = 0; d < 3; ++d) o->grid[d] = grid[d];
  o->order = order;
  if (o->ewald) build_bsmod(*o);
  rotpole(*o);
  return o;
}

void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

// HIPPO damping factors lambda_1 .. lambda_11 at distance r: core-valence
// with exponent ai (out[0..5]), with exponent ak (out[6..11]), valence-valence
// (out[12..17]).
void orc_damp_chgpen(double r, double ai, double ak, double* out) {
  damp_core_valence(r, ai, out);
  damp_core_valence(r, ak, out + 6);
  damp_valence_valence(r, ai, ak, out + 12);
}

// HIPPO charge penetration: per-atom core charge, valence charge and valence
// exponent; the multipole energy then uses the damped pair kernels.
void orc_set_chgpen(void* h, const double* pcore, const double* pval, const double* palpha) {
  Oracle* o = static_cast<Oracle*>(h);
  o->chgpen = true;
  o->pcore.assign(pcore, pcore + o->n);
  o->pval.assign(pval, pval + o->n);
  o->palpha.assign(palpha, palpha + o->n);
}

// HIPPO charge transfer (src/acc/hippo/echgtrn.cpp:13-183 with
// include/seq/pair_chgtrn.h): every pair i < k within the cutoff (minimum
// image), E = -f s (q_i exp(-a_k r) + q_k exp(-a_i r)), s the pair's
// multipole scale, times the fifth-order switch S(r) = sum c_j r^j between
// taper and cutoff, the polynomial fixed by S = 1 at the taper, S = 0 at the
// cutoff and zero first and second derivatives at both.  ct is [n][2]
// (charge, exponent).  Energy, dE/dx [n][3], virial [3][3] and the count of
// pairs with a nonzero unscaled energy and a nonzero scale.
void orc_chgtrn(void* h, const double* ct, double cut, double taper, double* e_out, double* grad, double* vir,
                long long* count) {
  const Oracle& o = *static_cast<Oracle*>(h);
  const int n = o.n;
  const double a = taper, b = cut, den = std::pow(b - a, 5);
  const double c[6] = {b * b * b * (b * b - 5 * a * b + 10 * a * a) / den, -30 * a * a * b * b / den,
                       30 * (a * b * b + a * a * b) / den, -10 * (a * a + 4 * a * b + b * b) / den,
                       15 * (a + b) / den, -6 / den};
  double e = 0, v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  long long cnt = 0;
  for (int q = 0; q < 3 * n; ++q) grad[q] = 0;
  for (int i = 0; i < n; ++i)
    for (int k = i + 1; k < n; ++k) {
      double xr[3] = {o.xyz[3 * k] - o.xyz[3 * i], o.xyz[3 * k + 1] - o.xyz[3 * i + 1],
                      o.xyz[3 * k + 2] - o.xyz[3 * i + 2]};
      image(o, xr[0], xr[1], xr[2]);
      const double r = std::sqrt(xr[0] * xr[0] + xr[1] * xr[1] + xr[2] * xr[2]);
      if (r > cut) continue;
      const Scale* sp = find_scale(o, i, k);
      const double sc = sp ? sp->m : 1.0;
      const double qi = ct[2 * i], ai = ct[2 * i + 1], qk = ct[2 * k], ak = ct[2 * k + 1];
      const double ei = std::exp(-ai * r), ek = std::exp(-ak * r);
      double e1 = -kCoulombF * (qi * ek + qk * ei);
      double d1 = kCoulombF * (qi * ak * ek + qk * ai * ei);
      if (r > taper) {
        double S = 0, dS = 0;
        for (int j = 5; j >= 0; --j) S = S * r + c[j];
        for (int j = 5; j >= 1; --j) dS = dS * r + j * c[j];
        d1 = e1 * dS + d1 * S;
        e1 *= S;
      }
      if (e1 != 0 && sc != 0) ++cnt;
      e += sc * e1;
      const double g = sc * d1 / r;
      for (int d = 0; d < 3; ++d) {
        grad[3 * i + d] -= g * xr[d];
        grad[3 * k + d] += g * xr[d];
      }
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) v[3 * p + q] += g * xr[p] * xr[q];
    }
  *e_out = e;
  for (int q = 0; q < 9; ++q) vir[q] = v[q];
  *count = cnt;
}

// HIPPO dispersion with a cutoff (src/acc/hippo/edisp.cpp:113-296,
// non-Ewald form, with include/seq/pair_disp.h): every pair i < k within the
// cutoff (minimum image), E = -s C_i C_k D^2 / r^6 times the same switch as
// charge transfer, s the pair's dispersion scale (pairs ik[q] with ds[q], 1
// otherwise), D the damping of the two Slater densities (damp_disp).  cd is
// [n][2] (C6, exponent).  The gradient differentiates E(r) numerically-free:
// D and dD/dr from the closed forms, the switch from its polynomial.
void orc_disp(void* h, const double* cd, double cut, double taper, int nds, const int* ik, const double* ds,
              double* e_out, double* grad, double* vir, long long* count) {
  const Oracle& o = *static_cast<Oracle*>(h);
  const int n = o.n;
  std::map<std::pair<int, int>, double> sc;
  for (int q = 0; q < nds; ++q) sc[{std::min(ik[2 * q], ik[2 * q + 1]), std::max(ik[2 * q], ik[2 * q + 1])}] = ds[q];
  const double a = taper, b = cut, den = std::pow(b - a, 5);
  const double c[6] = {b * b * b * (b * b - 5 * a * b + 10 * a * a) / den, -30 * a * a * b * b / den,
                       30 * (a * b * b + a * a * b) / den, -10 * (a * a + 4 * a * b + b * b) / den,
                       15 * (a + b) / den, -6 / den};
  double e = 0, v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  long long cnt = 0;
  for (int q = 0; q < 3 * n; ++q) grad[q] = 0;
  for (int i = 0; i < n; ++i)
    for (int k = i + 1; k < n; ++k) {
      double xr[3] = {o.xyz[3 * k] - o.xyz[3 * i], o.xyz[3 * k + 1] - o.xyz[3 * i + 1],
                      o.xyz[3 * k + 2] - o.xyz[3 * i + 2]};
      image(o, xr[0], xr[1], xr[2]);
      const double r = std::sqrt(xr[0] * xr[0] + xr[1] * xr[1] + xr[2] * xr[2]);
      if (r > cut) continue;
      auto it = sc.find({i, k});
      const double s = it == sc.end() ? 1.0 : it->second;
      double D, dD;
      damp_disp(r, cd[2 * i + 1], cd[2 * k + 1], &D, &dD);
      const double cc = cd[2 * i] * cd[2 * k];
      double e1 = -cc * D * D / std::pow(r, 6);
      double d1 = 6 * cc * D * D / std::pow(r, 7) - 2 * cc * D * dD / std::pow(r, 6);
      if (r > taper) {
        double S = 0, dS = 0;
        for (int j = 5; j >= 0; --j) S = S * r + c[j];
        for (int j = 5; j >= 1; --j) dS = dS * r + j * c[j];
        d1 = e1 * dS + d1 * S;
        e1 *= S;
      }
      if (e1 != 0 && s != 0) ++cnt;
      e += s * e1;
      const double g = s * d1 / r;
      for (int d = 0; d < 3; ++d) {
        grad[3 * i + d] -= g * xr[d];
        grad[3 * k + d] += g * xr[d];
      }
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) v[3 * p + q] += g * xr[p] * xr[q];
    }
  *e_out = e;
  for (int q = 0; q < 9; ++q) vir[q] = v[q];
  *count = cnt;
}

// HIPPO dispersion damping D(r) and dD/dr for exponents ai, ak (out[0..1]).
void orc_damp_disp(double r, double ai, double ak, double* out) { damp_disp(r, ai, ak, out, out + 1); }

// HIPPO Pauli repulsion (see repel): rp [n][3] size, exponent, valence count.
double orc_repel(void* h, const double* rp, double cut, double taper, int nrs, const int* ik, const double* rs,
                 double* grad, double* vir, long long* count) {
  return repel(*static_cast<Oracle*>(h), rp, cut, taper, nrs, ik, rs, grad, vir, count);
}

// Radial series B_0..B_5 of the repulsion kernel (before the sizes).
void orc_repel_series(double r, double ai, double ak, double* B) { repel_series(r, ai, ak, B); }

void orc_set_xyz(void* h, const double* xyz) {
  Oracle* o = static_cast<Oracle*>(h);
  o->xyz.assign(xyz, xyz + 3 * o->n);
  rotpole(*o);
}

void orc_get_rpole(void* h, double* rpole) {
  Oracle* o = static_cast<Oracle*>(h);
  std::copy(o->rpole.begin(), o->rpole.end(), rpole);
}

void orc_get_bsmod(void* h, int dim, double* out) {
  Oracle* o = static_cast<Oracle*>(h);
  std::copy(o->bsmod[dim].begin(), o->bsmod[dim].end(), out);
}

void orc_dfield(void* h, double* fd, double* fp) { dfield(*static_cast<Oracle*>(h), fd, fp); }

void orc_ufield(void* h, const double* ud, const double* up, double* fd, double* fp) {
  ufield(*static_cast<Oracle*>(h), ud, up, fd, fp);
}

// Reciprocal-space pieces, exposed for component-level parity checks.
void orc_recip_dipole(void* h, const double* ud, const double* up, double* fd, double* fp) {
  recip_dipole(*static_cast<Oracle*>(h), ud, up, fd, fp);
}

double orc_recip_mpole(void* h, double* field, double* fphi) {
  std::vector<double> f;
  double e = recip_mpole(*static_cast<Oracle*>(h), field, fphi ? &f : nullptr);
  if (fphi) std::copy(f.begin(), f.end(), fphi);
  return e;
}

// PME sub-steps on interleaved complex grids [K3][K2][K1][2]: spreading of
// the fractional multipoles (mode 0, real part) or fractional dipoles (mode
// 1, d real and p imaginary); the influence function in place; the Cartesian
// field -grad phi of a grid's real (or imaginary) part at each atom.
void orc_pme_grid(void* h, int mode, const double* ud, const double* up, double* grid) {
  const Oracle& o = *static_cast<Oracle*>(h);
  auto sp = make_splines(o);
  std::vector<cplx> g;
  if (mode == 0) {
    std::vector<double> fmp(10 * o.n);
    cart_to_frac_mpole(o, o.rpole.data(), fmp.data());
    spread(o, sp, fmp.data(), nullptr, nullptr, g);
  } else {
    double a[3];
    frac_scale(o, a);
    std::vector<double> fud(3 * o.n), fup(3 * o.n);
    for (int i = 0; i < o.n; ++i)
      for (int d = 0; d < 3; ++d) {
        fud[3 * i + d] = ud[3 * i + d] * a[d];
        fup[3 * i + d] = up[3 * i + d] * a[d];
      }
    spread(o, sp, nullptr, fud.data(), fup.data(), g);
  }
  for (size_t i = 0; i < g.size(); ++i) {
    grid[2 * i] = g[i].real();
    grid[2 * i + 1] = g[i].imag();
  }
}

void orc_pme_conv(void* h, double* grid) {
  const Oracle& o = *static_cast<Oracle*>(h);
  size_t tot = (size_t)o.grid[0] * o.grid[1] * o.grid[2];
  std::vector<cplx> g(tot);
  for (size_t i = 0; i < tot; ++i) g[i] = cplx(grid[2 * i], grid[2 * i + 1]);
  pme_conv(o, g, nullptr);
  for (size_t i = 0; i < tot; ++i) {
    grid[2 * i] = g[i].real();
    grid[2 * i + 1] = g[i].imag();
  }
}

void orc_pme_field(void* h, const double* grid, int imag, double* field) {
  const Oracle& o = *static_cast<Oracle*>(h);
  size_t tot = (size_t)o.grid[0] * o.grid[1] * o.grid[2];
  std::vector<cplx> g(tot);
  for (size_t i = 0; i < tot; ++i) g[i] = cplx(grid[2 * i], grid[2 * i + 1]);
  auto sp = make_splines(o);
  std::vector<double> fphi(20 * o.n);
  gather(o, sp, g, imag != 0, fphi.data());
  double a[3];
  frac_scale(o, a);
  for (int i = 0; i < o.n; ++i)
    for (int d = 0; d < 3; ++d) field[3 * i + d] = -a[d] * fphi[20 * i + 1 + d];
}

// Preconditioner z = M r (sparse != 0: the usolve-cutoff pair sum).
void orc_precond(void* h, int sparse, double ucut, double udiag, const double* rd, const double* rp, double* zd,
                 double* zp) {
  precond(*static_cast<Oracle*>(h), sparse != 0, ucut, udiag, rd, rp, zd, zp);
}

// 3D DFT of an interleaved complex grid (re, im) in place.
void orc_dft3(void* h, double* grid, int sign) {
  Oracle* o = static_cast<Oracle*>(h);
  size_t tot = (size_t)o->grid[0] * o->grid[1] * o->grid[2];
  std::vector<cplx> g(tot);
  for (size_t i = 0; i < tot; ++i) g[i] = cplx(grid[2 * i], grid[2 * i + 1]);
  dft3(*o, g, sign);
  for (size_t i = 0; i < tot; ++i) {
    grid[2 * i] = g[i].real();
    grid[2 * i + 1] = g[i].imag();
  }
}

// Induced dipoles by preconditioned CG (the reference's PCG, FP64).  Returns
// 0 on convergence, 1 if the iteration limit was reached.
// PCG solve as orc_induce, started from a given guess (u0d, u0p) when non-null:
// r0 = alpha^-1 (udir - u0) + field(u0) (src/cu/amoeba/pcg.cu:52-64).  guess_rms
// receives the RMS residual of the initial guess in Debye.
int orc_induce_from(void* h, double poleps, int politer, int prec, int guess, double peek, double ucut,
                    double udiag, const double* u0d, const double* u0p, double* ud_out, double* up_out,
                    int* iters_out, double* rms_out, double* guess_rms);

int orc_induce(void* h, double poleps, int politer, int prec, int guess, double peek, double ucut,
               double udiag, double* ud_out, double* up_out, int* iters_out, double* rms_out) {
  return orc_induce_from(h, poleps, politer, prec, guess, peek, ucut, udiag, nullptr, nullptr, ud_out, up_out,
                         iters_out, rms_out, nullptr);
}

int orc_induce_from(void* h, double poleps, int politer, int prec, int guess, double peek, double ucut,
                    double udiag, const double* u0d, const double* u0p, double* ud_out, double* up_out,
                    int* iters_out, double* rms_out, double* guess_rms) {
  Oracle& o = *static_cast<Oracle*>(h);
  const int n = o.n, n3 = 3 * n;
  std::vector<double> fd(n3), fp(n3), udir(n3), udirp(n3), ud(n3), up(n3);
  std::vector<double> rd(n3), rp(n3), zd(n3), zp(n3), cd(n3), cp(n3), vd(n3), vp(n3);
  std::vector<double> pinv(n);
  for (int i = 0; i < n; ++i) pinv[i] = 1.0 / std::max(o.polarity[i], 1e-16);
  dfield(o, fd.data(), fp.data());
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      udir[3 * i + d] = o.polarity[i] * fd[3 * i + d];
      udirp[3 * i + d] = o.polarity[i] * fp[3 * i + d];
    }
  if (u0d && u0p) {
    ud.assign(u0d, u0d + n3);
    up.assign(u0p, u0p + n3);
    ufield(o, ud.data(), up.data(), rd.data(), rp.data());
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        rd[3 * i + d] += pinv[i] * (udir[3 * i + d] - ud[3 * i + d]);
        rp[3 * i + d] += pinv[i] * (udirp[3 * i + d] - up[3 * i + d]);
      }
  } else if (guess) {
    ud = udir;
    up = udirp;
    ufield(o, ud.data(), up.data(), rd.data(), rp.data());
  } else {
    std::fill(ud.begin(), ud.end(), 0.0);
    std::fill(up.begin(), up.end(), 0.0);
    rd = fd;
    rp = fp;
  }
  for (int i = 0; i < n; ++i)
    if (o.polarity[i] == 0)
      for (int d = 0; d < 3; ++d) rd[3 * i + d] = rp[3 * i + d] = 0;
  if (guess_rms) *guess_rms = kDebye * std::sqrt(std::max(dot(rd, rd), dot(rp, rp)) / n);
  const bool sparse = prec && ucut > 0;
  precond(o, sparse, ucut, udiag, rd.data(), rp.data(), zd.data(), zp.data());
  cd = zd;
  cp = zp;
  double sum = dot(rd, zd), sump = dot(rp, zp);
  int iter = 0;
  double eps = 100;
  bool done = false;
  const int miniter = std::min(3, n);
  while (!done) {
    ++iter;
    ufield(o, cd.data(), cp.data(), fd.data(), fp.data());
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        vd[3 * i + d] = pinv[i] * cd[3 * i + d] - fd[3 * i + d];
        vp[3 * i + d] = pinv[i] * cp[3 * i + d] - fp[3 * i + d];
      }
    double a = dot(cd, vd), ap = dot(cp, vp);
    if (a != 0) a = sum / a;
    if (ap != 0) ap = sump / ap;
    for (int j = 0; j < n3; ++j) {
      ud[j] += a * cd[j];
      up[j] += ap * cp[j];
      rd[j] -= a * vd[j];
      rp[j] -= ap * vp[j];
    }
    for (int i = 0; i < n; ++i)
      if (o.polarity[i] == 0)
        for (int d = 0; d < 3; ++d) rd[3 * i + d] = rp[3 * i + d] = 0;
    precond(o, sparse, ucut, udiag, rd.data(), rp.data(), zd.data(), zp.data());
    double s1 = dot(rd, zd), sp1 = dot(rp, zp);
    double b = sum != 0 ? s1 / sum : 0, bp = sump != 0 ? sp1 / sump : 0;
    for (int j = 0; j < n3; ++j) {
      cd[j] = zd[j] + b * cd[j];
      cp[j] = zp[j] + bp * cp[j];
    }
    sum = s1;
    sump = sp1;
    double ed = dot(rd, rd), ep = dot(rp, rp);
    eps = kDebye * std::sqrt(std::max(ed, ep) / n);
    if (eps < poleps) done = true;
    if (iter < miniter) done = false;
    if (iter >= politer) done = true;
    if (done)
      for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
          ud[3 * i + d] += peek * o.polarity[i] * rd[3 * i + d];
          up[3 * i + d] += peek * o.polarity[i] * rp[3 * i + d];
        }
  }
  std::copy(ud.begin(), ud.end(), ud_out);
  std::copy(up.begin(), up.end(), up_out);
  if (iters_out) *iters_out = iter;
  if (rms_out) *rms_out = eps;
  return iter >= 100 ? 1 : 0;
}

void orc_set_md(void* h, const int* jvdw, int njvdw, const double* rmin, const double* eps, const double* red,
                const int* parent, int nvs, const int* vs_ik, const double* vs_s, double vcut, double vtaper,
                int nb, const int* bij, const double* bkr, int na, const int* aabc, const double* akt, int nu,
                const int* uac, const double* ukr, const double* poly6) {
  Oracle& o = *static_cast<Oracle*>(h);
  const int n = o.n;
  o.has_md = true;
  o.jvdw.assign(jvdw, jvdw + n);
  o.njvdw = njvdw;
  o.vrmin.assign(rmin, rmin + njvdw * njvdw);
  o.veps.assign(eps, eps + njvdw * njvdw);
  o.vred.assign(red, red + n);
  o.vparent.assign(parent, parent + n);
  o.vscale.assign(n, {});
  for (int q = 0; q < nvs; ++q) {
    const int i = vs_ik[2 * q], k = vs_ik[2 * q + 1];
    o.vscale[i].push_back({k, vs_s[q]});
    o.vscale[k].push_back({i, vs_s[q]});
  }
  o.vcut = vcut;
  o.vtaper = vtaper;
  o.bond_ij.assign(bij, bij + 2 * nb);
  o.bond_kr.assign(bkr, bkr + 2 * nb);
  o.angle_abc.assign(aabc, aabc + 3 * na);
  o.angle_kt.assign(akt, akt + 2 * na);
  o.urey_ac.assign(uac, uac + 2 * nu);
  o.urey_kr.assign(ukr, ukr + 2 * nu);
  o.cbnd = poly6[0];
  o.qbnd = poly6[1];
  o.cang = poly6[2];
  o.qang = poly6[3];
  o.pang = poly6[4];
  o.sang = poly6[5];
}

// vdW and valence energies e4 = {evdw, ebond, eangle, eurey}, gradient
// [n][3] (added to by the selected terms) and virial [3][3].
int orc_md_energy(void* h, int terms, double* e4, double* grad, double* vir) {
  const Oracle& o = *static_cast<Oracle*>(h);
  if (!o.has_md) return 1;
  std::vector<double> g(3 * o.n, 0.0);
  double W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  e4[0] = (terms & 1) ? vdw_energy(o, g.data(), W) : 0.0;
  double eb = 0, ea = 0, eu = 0;
  if (terms & 2) valence_energy(o, &eb, &ea, &eu, g.data(), W);
  e4[1] = eb;
  e4[2] = ea;
  e4[3] = eu;
  if (grad) std::copy(g.begin(), g.end(), grad);
  if (vir)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) vir[3 * a + b] = 0.5 * (W[3 * a + b] + W[3 * b + a]);
  return 0;
}

// Energy (e5 = {em, ep, em_real, em_recip, em_self}, kcal/mol), gradient
// (dE/dx, [n][3]) and virial ([3][3]) of the multipole (terms & 1) and
// polarization (terms & 2) energies at the given induced dipoles.
int orc_energy(void* h, int terms, const double* ud, const double* up, double* e5, double* grad, double* vir) {
  const Oracle& o = *static_cast<Oracle*>(h);
  EnergyOut e = energy_grad(o, terms & 1, terms & 2, ud, up, grad, vir);
  e5[0] = e.em;
  e5[1] = e.ep;
  e5[2] = e.em_real;
  e5[3] = e.em_recip;
  e5[4] = e.em_self;
  return 0;
}

}  // extern "C"
