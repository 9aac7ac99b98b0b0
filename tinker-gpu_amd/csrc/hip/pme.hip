// Particle-mesh Ewald for gfx950: order-5 B-splines, deterministic
// spreading, an own mixed-radix (2,3,4,5) 3D FFT and the gathers that turn
// the reciprocal potential into fields.
//
// Reference behaviour followed: src/acc/pme.cpp:10-979 and src/cu/pme.cu
// (fractional coordinates w -> (w + 1/2) - floor(w + 1/2), grid base
// floor(K w) - order + 1, x-fastest grid, d dipoles in the real part and p
// dipoles in the imaginary part of one complex grid), src/pme.cpp:84-117 and
// the Fortran dftmod rule (B-spline moduli with the optimal zeta factor),
// pmeConv (src/acc/pme.cpp:220-303, k = 0 zeroed, exp cut at -50), and the
// reciprocal + self fields of src/cu/amoeba/field.cu:12-118.
//
// Spreading ("owner computes"): the grid is cut into tiles whose edge divides
// K; each workgroup owns one tile, takes the atoms whose support base lies in
// its own tile or in the adjacent lower tiles, accumulates their
// contributions to its own points in LDS as 64-bit fixed point, and stores
// each point once.  Integer sums are order-independent, so the grid — and
// everything downstream — is bit-reproducible without global atomics.
//
// FFT: one workgroup transforms 16 (FP32) or 8 (FP64) grid lines in LDS with
// a Stockham auto-sort over radices 4, 2, 3, 5; the influence-function
// multiply is fused between the forward and inverse transforms of the z pass.
#include "ctx.hpp"

namespace t9 {
namespace {

constexpr double kPi = 3.14159265358979323846;

template <class real>
struct C2;
template <>
struct C2<float> {
  using T = float2;
};
template <>
struct C2<double> {
  using T = double2;
};
template <class real>
using cplx_t = typename C2<real>::T;

template <class real>
using R4 = typename RealTraits<real>::real4;

// Order-5 cardinal B-spline weights and derivatives for a point at fractional
// offset w in [0,1): th[lv][m] = d^lv/dx^lv M5(x) at x = w + 4 - m,
// m = 0..4 (grid point base + m).
template <int LV, class real>
__device__ __forceinline__ void bspline5(real w, real th[LV][5]) {
  real A2[2] = {w, 1 - w};
  real A3[3], A4[4], A5[5];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    real lo = k >= 1 ? A2[k - 1] : 0, hi = k <= 1 ? A2[k] : 0;
    A3[k] = ((w + k) * hi + (3 - w - k) * lo) * (real)0.5;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    real lo = k >= 1 ? A3[k - 1] : 0, hi = k <= 2 ? A3[k] : 0;
    A4[k] = ((w + k) * hi + (4 - w - k) * lo) * (real)(1.0 / 3.0);
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    real lo = k >= 1 ? A4[k - 1] : 0, hi = k <= 3 ? A4[k] : 0;
    A5[k] = ((w + k) * hi + (5 - w - k) * lo) * (real)0.25;
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    const int k = 4 - m;
    th[0][m] = A5[k];
    if (LV > 1) {
      real a = k <= 3 ? A4[k] : 0, b = k >= 1 ? A4[k - 1] : 0;
      th[LV > 1 ? 1 : 0][m] = a - b;
    }
    if (LV > 2) {
      real a = k <= 2 ? A3[k] : 0, b = (k >= 1 && k - 1 <= 2) ? A3[k - 1] : 0, c = k >= 2 ? A3[k - 2] : 0;
      th[LV > 2 ? 2 : 0][m] = a - 2 * b + c;
    }
    if (LV > 3) {
      real a = k <= 1 ? A2[k] : 0, b = (k >= 1 && k - 1 <= 1) ? A2[k - 1] : 0;
      real c = (k >= 2 && k - 2 <= 1) ? A2[k - 2] : 0, d = k >= 3 ? A2[k - 3] : 0;
      th[LV > 3 ? 3 : 0][m] = a - 3 * b + 3 * c - d;
    }
  }
}

// ------------------------------------------------------------ binning

template <class real>
__global__ void k_pme_bin(int n, const int* __restrict__ perm, const double* __restrict__ xyz, double lx, double ly,
                          double lz, int K1, int K2, int K3, int T1, int T2, int T3, int nt1, int nt2,
                          R4<real>* __restrict__ pw, int4* __restrict__ pbase, int* __restrict__ tile_count) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  const double L[3] = {lx, ly, lz};
  const int K[3] = {K1, K2, K3};
  int b[3];
  double w[3];
  for (int d = 0; d < 3; ++d) {
    double f = xyz[3 * i + d] / L[d];
    f = f + 0.5 - floor(f + 0.5);
    double u = K[d] * f;
    int ig = (int)floor(u);
    w[d] = u - ig;
    int bb = ig - 4;
    while (bb < 0) bb += K[d];
    while (bb >= K[d]) bb -= K[d];
    b[d] = bb;
  }
  const int tile = b[0] / T1 + nt1 * (b[1] / T2 + nt2 * (b[2] / T3));
  R4<real> ww;
  ww.x = (real)w[0];
  ww.y = (real)w[1];
  ww.z = (real)w[2];
  ww.w = 0;
  pw[s] = ww;
  pbase[s] = make_int4(b[0], b[1], b[2], tile);
  atomicAdd(&tile_count[tile], 1);
}

__global__ void k_tile_scatter(int n, const int4* __restrict__ pbase, const int* __restrict__ start,
                               int* __restrict__ cursor, int* __restrict__ atoms) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int t = pbase[s].w;
  atoms[start[t] + atomicAdd(&cursor[t], 1)] = s;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(const int* __restrict__ in, int* __restrict__ out, int m) {
  __shared__ int sh[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < m; base += 1024) {
    int idx = base + threadIdx.x;
    int v = idx < m ? in[idx] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < 1024; s <<= 1) {
      int t = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (idx < m) out[idx] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += sh[1023];
    __syncthreads();
  }
}

// ------------------------------------------------------------ spreading

struct TileGeom {
  int K[3], T[3], nt[3];
};

// Local index (0..T-1) of grid point base + m inside tile [lo, lo + T), or -1.
__device__ __forceinline__ int local_idx(int base, int m, int K, int lo, int T) {
  int g = base + m;
  if (g >= K) g -= K;
  int l = g - lo;
  return (l >= 0 && l < T) ? l : -1;
}

// MODE 0: induced dipoles d (real part) and p (imaginary part).
// MODE 1: permanent multipoles (real part; imaginary part zero).
template <class real, int MODE>
__global__ __launch_bounds__(256) void k_spread(TileGeom g, const int* __restrict__ tile_start,
                                                const int* __restrict__ tile_count,
                                                const int* __restrict__ tile_atoms, const R4<real>* __restrict__ pw,
                                                const int4* __restrict__ pbase, const real* __restrict__ src1,
                                                const real* __restrict__ src2, real a1, real a2, real a3,
                                                double scale, cplx_t<real>* __restrict__ grid) {
  extern __shared__ long long acc[];
  const int T1 = g.T[0], T2 = g.T[1], T3 = g.T[2];
  const int npt = T1 * T2 * T3;
  const int tt = blockIdx.x;
  const int t1 = tt % g.nt[0], t2 = (tt / g.nt[0]) % g.nt[1], t3 = tt / (g.nt[0] * g.nt[1]);
  const int lo1 = t1 * T1, lo2 = t2 * T2, lo3 = t3 * T3;
  for (int q = threadIdx.x; q < 2 * npt; q += blockDim.x) acc[q] = 0;
  __syncthreads();
  for (int c = 0; c < 8; ++c) {
    const int da = c & 1, db = (c >> 1) & 1, dc = c >> 2;
    if ((da && g.nt[0] == 1) || (db && g.nt[1] == 1) || (dc && g.nt[2] == 1)) continue;
    int b1 = t1 - da, b2 = t2 - db, b3 = t3 - dc;
    if (b1 < 0) b1 += g.nt[0];
    if (b2 < 0) b2 += g.nt[1];
    if (b3 < 0) b3 += g.nt[2];
    const int bin = b1 + g.nt[0] * (b2 + g.nt[1] * b3);
    const int st = tile_start[bin], ct = tile_count[bin];
    for (int q = threadIdx.x; q < ct; q += blockDim.x) {
      const int s = tile_atoms[st + q];
      const R4<real> w = pw[s];
      const int4 b = pbase[s];
      int l1[5], l2[5], l3[5];
      bool any1 = false, any2 = false, any3 = false;
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        l1[m] = local_idx(b.x, m, g.K[0], lo1, T1);
        l2[m] = local_idx(b.y, m, g.K[1], lo2, T2);
        l3[m] = local_idx(b.z, m, g.K[2], lo3, T3);
        any1 |= l1[m] >= 0;
        any2 |= l2[m] >= 0;
        any3 |= l3[m] >= 0;
      }
      if (!(any1 && any2 && any3)) continue;
      constexpr int LV = MODE == 0 ? 2 : 3;
      real th1[LV][5], th2[LV][5], th3[LV][5];
      bspline5<LV>(w.x, th1);
      bspline5<LV>(w.y, th2);
      bspline5<LV>(w.z, th3);
      real f[10];
      if (MODE == 0) {
        f[0] = src1[3 * s] * a1;
        f[1] = src1[3 * s + 1] * a2;
        f[2] = src1[3 * s + 2] * a3;
        f[3] = src2[3 * s] * a1;
        f[4] = src2[3 * s + 1] * a2;
        f[5] = src2[3 * s + 2] * a3;
      } else {
        const real* p = src1 + 10 * (size_t)s;
        f[0] = p[0];
        f[1] = p[1] * a1;
        f[2] = p[2] * a2;
        f[3] = p[3] * a3;
        f[4] = p[4] * a1 * a1;
        f[5] = p[5] * a2 * a2;
        f[6] = p[6] * a3 * a3;
        f[7] = 2 * p[7] * a1 * a2;
        f[8] = 2 * p[8] * a1 * a3;
        f[9] = 2 * p[9] * a2 * a3;
      }
      for (int m3 = 0; m3 < 5; ++m3) {
        if (l3[m3] < 0) continue;
        const real v0 = th3[0][m3], v1 = th3[1][m3];
        const real v2 = LV > 2 ? th3[LV > 2 ? 2 : 0][m3] : 0;
        for (int m2 = 0; m2 < 5; ++m2) {
          if (l2[m2] < 0) continue;
          const real u0 = th2[0][m2], u1 = th2[1][m2];
          const real u2 = LV > 2 ? th2[LV > 2 ? 2 : 0][m2] : 0;
          const int row = T1 * (l2[m2] + T2 * l3[m3]);
          for (int m1 = 0; m1 < 5; ++m1) {
            if (l1[m1] < 0) continue;
            const real t0 = th1[0][m1], t1v = th1[1][m1];
            const real t2v = LV > 2 ? th1[LV > 2 ? 2 : 0][m1] : 0;
            const int idx = row + l1[m1];
            if (MODE == 0) {
              real vd = f[0] * t1v * u0 * v0 + f[1] * t0 * u1 * v0 + f[2] * t0 * u0 * v1;
              real vp = f[3] * t1v * u0 * v0 + f[4] * t0 * u1 * v0 + f[5] * t0 * u0 * v1;
              atomicAdd((unsigned long long*)&acc[idx], (unsigned long long)llrint((double)vd * scale));
              atomicAdd((unsigned long long*)&acc[npt + idx], (unsigned long long)llrint((double)vp * scale));
            } else {
              real val = f[0] * t0 * u0 * v0 + f[1] * t1v * u0 * v0 + f[2] * t0 * u1 * v0 + f[3] * t0 * u0 * v1 +
                         f[4] * t2v * u0 * v0 + f[5] * t0 * u2 * v0 + f[6] * t0 * u0 * v2 + f[7] * t1v * u1 * v0 +
                         f[8] * t1v * u0 * v1 + f[9] * t0 * u1 * v1;
              atomicAdd((unsigned long long*)&acc[idx], (unsigned long long)llrint((double)val * scale));
            }
          }
        }
      }
    }
  }
  __syncthreads();
  const double inv = 1.0 / scale;
  for (int q = threadIdx.x; q < npt; q += blockDim.x) {
    const int l1 = q % T1, l2 = (q / T1) % T2, l3 = q / (T1 * T2);
    const long long gi = (lo1 + l1) + (long long)g.K[0] * ((lo2 + l2) + (long long)g.K[1] * (lo3 + l3));
    cplx_t<real> v;
    v.x = (real)((double)acc[q] * inv);
    v.y = (real)((double)acc[npt + q] * inv);
    grid[gi] = v;
  }
}

// ------------------------------------------------------------ FFT

struct FftPlan {
  int n;
  int nstage;
  int radix[16];
};

struct ConvArgs {
  int enabled;
  double pterm;        // (pi / alpha)^2
  double volterm;      // pi * V
  double il[3];        // 1 / box edge
  const double* bsmod;  // K1 + K2 + K3 moduli (FP64)
};

template <class real>
__device__ __forceinline__ cplx_t<real> cmul(cplx_t<real> a, cplx_t<real> b) {
  cplx_t<real> r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// Small DFT y_q = sum_r v_r exp(sign 2 pi i r q / R), R in {2,3,4,5}.
template <class real>
__device__ __forceinline__ void small_dft(int R, int sign, cplx_t<real>* v, cplx_t<real>* y) {
  if (R == 2) {
    y[0].x = v[0].x + v[1].x;
    y[0].y = v[0].y + v[1].y;
    y[1].x = v[0].x - v[1].x;
    y[1].y = v[0].y - v[1].y;
    return;
  }
  if (R == 4) {
    // exp(sign i pi/2) = sign * i
    cplx_t<real> a, b, c, d;
    a.x = v[0].x + v[2].x;
    a.y = v[0].y + v[2].y;
    b.x = v[0].x - v[2].x;
    b.y = v[0].y - v[2].y;
    c.x = v[1].x + v[3].x;
    c.y = v[1].y + v[3].y;
    d.x = v[1].x - v[3].x;
    d.y = v[1].y - v[3].y;
    // i*sign*d = (-sign d.y, sign d.x)
    const real sg = (real)sign;
    y[0].x = a.x + c.x;
    y[0].y = a.y + c.y;
    y[2].x = a.x - c.x;
    y[2].y = a.y - c.y;
    y[1].x = b.x - sg * d.y;
    y[1].y = b.y + sg * d.x;
    y[3].x = b.x + sg * d.y;
    y[3].y = b.y - sg * d.x;
    return;
  }
  // R = 3 or 5: direct sum with exact constants
  const double c3[3] = {1.0, -0.5, -0.5};
  const double s3[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
  const double c5[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                        0.30901699437494742410};
  const double s5[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                        -0.95105651629515357212};
  const double* cc = R == 3 ? c3 : c5;
  const double* ss = R == 3 ? s3 : s5;
  for (int q = 0; q < R; ++q) {
    real yx = 0, yy = 0;
    for (int r = 0; r < R; ++r) {
      const int k = (r * q) % R;
      const real cr = (real)cc[k], sr = (real)(sign * ss[k]);
      yx += v[r].x * cr - v[r].y * sr;
      yy += v[r].x * sr + v[r].y * cr;
    }
    y[q].x = yx;
    y[q].y = yy;
  }
}

template <class real>
__device__ void fft_lines_lds(cplx_t<real>* buf0, cplx_t<real>* buf1, int nlines, const FftPlan& plan,
                              const cplx_t<real>* __restrict__ tw, int sign, cplx_t<real>** result) {
  const int N = plan.n;
  cplx_t<real>* src = buf0;
  cplx_t<real>* dst = buf1;
  int Ns = 1;
  for (int st = 0; st < plan.nstage; ++st) {
    const int R = plan.radix[st];
    const int nb = N / R;
    const int tstep = N / (Ns * R);
    for (int q = threadIdx.x; q < nlines * nb; q += blockDim.x) {
      const int lb = q / nb, j = q - lb * nb;
      const int k = j % Ns;
      cplx_t<real> v[5], y[5];
      for (int r = 0; r < R; ++r) {
        cplx_t<real> x = src[lb * N + j + r * nb];
        int ti = (k * r * tstep) % N;
        cplx_t<real> t = tw[ti];
        if (sign > 0) t.y = -t.y;
        v[r] = r == 0 ? x : cmul<real>(x, t);
      }
      small_dft<real>(R, sign, v, y);
      const int ob = lb * N + (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) dst[ob + r * Ns] = y[r];
    }
    __syncthreads();
    cplx_t<real>* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  *result = src;
}

// One pass of 1D FFTs along `axis` over all grid lines.  With conv.enabled
// (z axis only), the forward transform is followed by the influence-function
// multiply and the inverse transform, all in LDS.
template <class real, int B>
__global__ __launch_bounds__(256) void k_fft_pass(cplx_t<real>* __restrict__ grid, int K1, int K2, int K3, int axis,
                                                  int sign, FftPlan plan, const cplx_t<real>* __restrict__ tw,
                                                  ConvArgs conv) {
  extern __shared__ char smem[];
  const int N = plan.n;
  cplx_t<real>* buf0 = (cplx_t<real>*)smem;
  cplx_t<real>* buf1 = buf0 + B * N;
  // line geometry
  long long base[B];
  long long stride;
  int nl = 0;
  if (axis == 0) {
    const int nlines = K2 * K3;
    const int l0 = blockIdx.x * B;
    nl = min(B, nlines - l0);
    stride = 1;
    for (int b = 0; b < B; ++b) base[b] = (long long)(l0 + b) * K1;
  } else {
    const int xb = (K1 + B - 1) / B;
    const int x0 = (blockIdx.x % xb) * B;
    const int other = blockIdx.x / xb;  // z for axis 1, y for axis 2
    nl = min(B, K1 - x0);
    if (axis == 1) {
      stride = K1;
      for (int b = 0; b < B; ++b) base[b] = (x0 + b) + (long long)K1 * K2 * other;
    } else {
      stride = (long long)K1 * K2;
      for (int b = 0; b < B; ++b) base[b] = (x0 + b) + (long long)K1 * other;
    }
  }
  // load: consecutive threads take consecutive lines of the same element for
  // axes 1/2 (adjacent x), consecutive elements for axis 0
  for (int q = threadIdx.x; q < nl * N; q += blockDim.x) {
    int lb, e;
    if (axis == 0) {
      lb = q / N;
      e = q - lb * N;
    } else {
      e = q / nl;
      lb = q - e * nl;
    }
    buf0[lb * N + e] = grid[base[lb] + e * stride];
  }
  __syncthreads();
  cplx_t<real>* res;
  fft_lines_lds<real>(buf0, buf1, nl, plan, tw, sign, &res);
  if (conv.enabled) {
    // res holds the z-direction forward transform of lines (x0 + lb, y)
    const int xb = (K1 + B - 1) / B;
    const int x0 = (blockIdx.x % xb) * B;
    const int y = blockIdx.x / xb;
    for (int q = threadIdx.x; q < nl * N; q += blockDim.x) {
      const int lb = q / N, k3 = q - lb * N;
      const int k1 = x0 + lb, k2 = y;
      double f = 0;
      if (k1 != 0 || k2 != 0 || k3 != 0) {
        const int r1 = k1 < (K1 + 1) / 2 ? k1 : k1 - K1;
        const int r2 = k2 < (K2 + 1) / 2 ? k2 : k2 - K2;
        const int r3 = k3 < (K3 + 1) / 2 ? k3 : k3 - K3;
        const double h1 = r1 * conv.il[0], h2 = r2 * conv.il[1], h3 = r3 * conv.il[2];
        const double hsq = h1 * h1 + h2 * h2 + h3 * h3;
        const double term = -conv.pterm * hsq;
        if (term > -50) {
          const double denom = hsq * conv.volterm * conv.bsmod[k1] * conv.bsmod[K1 + k2] * conv.bsmod[K1 + K2 + k3];
          f = exp(term) / denom;
        }
      }
      cplx_t<real> v = res[q];
      v.x = (real)(v.x * f);
      v.y = (real)(v.y * f);
      res[q] = v;
    }
    __syncthreads();
    cplx_t<real>* other = res == buf0 ? buf1 : buf0;
    fft_lines_lds<real>(res, other, nl, plan, tw, -sign, &res);
  }
  for (int q = threadIdx.x; q < nl * N; q += blockDim.x) {
    int lb, e;
    if (axis == 0) {
      lb = q / N;
      e = q - lb * N;
    } else {
      e = q / nl;
      lb = q - e * nl;
    }
    grid[base[lb] + e * stride] = res[lb * N + e];
  }
}

// ------------------------------------------------------------ gathers

// Field of the dipole grid at each atom: fd -= a * grad phi (real part),
// fp likewise from the imaginary part, plus the self term term * u.
template <class real>
__global__ __launch_bounds__(256) void k_gather_dip(int n, TileGeom g, const R4<real>* __restrict__ pw,
                                                    const int4* __restrict__ pbase,
                                                    const cplx_t<real>* __restrict__ grid, real a1, real a2, real a3,
                                                    real term, const real* __restrict__ ud,
                                                    const real* __restrict__ up, real* __restrict__ fd,
                                                    real* __restrict__ fp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const R4<real> w = pw[s];
  const int4 b = pbase[s];
  real th1[2][5], th2[2][5], th3[2][5];
  bspline5<2>(w.x, th1);
  bspline5<2>(w.y, th2);
  bspline5<2>(w.z, th3);
  real d1 = 0, d2 = 0, d3 = 0, p1 = 0, p2 = 0, p3 = 0;
  for (int m3 = 0; m3 < 5; ++m3) {
    int z = b.z + m3;
    if (z >= g.K[2]) z -= g.K[2];
    const real v0 = th3[0][m3], v1 = th3[1][m3];
    for (int m2 = 0; m2 < 5; ++m2) {
      int y = b.y + m2;
      if (y >= g.K[1]) y -= g.K[1];
      const real u0 = th2[0][m2], u1 = th2[1][m2];
      const long long row = (long long)g.K[0] * (y + (long long)g.K[1] * z);
      for (int m1 = 0; m1 < 5; ++m1) {
        int x = b.x + m1;
        if (x >= g.K[0]) x -= g.K[0];
        const cplx_t<real> q = grid[row + x];
        const real t0 = th1[0][m1], t1 = th1[1][m1];
        const real w1 = t1 * u0 * v0, w2 = t0 * u1 * v0, w3 = t0 * u0 * v1;
        d1 += q.x * w1;
        d2 += q.x * w2;
        d3 += q.x * w3;
        p1 += q.y * w1;
        p2 += q.y * w2;
        p3 += q.y * w3;
      }
    }
  }
  fd[3 * s] += term * ud[3 * s] - a1 * d1;
  fd[3 * s + 1] += term * ud[3 * s + 1] - a2 * d2;
  fd[3 * s + 2] += term * ud[3 * s + 2] - a3 * d3;
  fp[3 * s] += term * up[3 * s] - a1 * p1;
  fp[3 * s + 1] += term * up[3 * s + 1] - a2 * p2;
  fp[3 * s + 2] += term * up[3 * s + 2] - a3 * p3;
}

// Field of the permanent-multipole grid (real part) plus the self term,
// added to both the d and the p field.
template <class real>
__global__ __launch_bounds__(256) void k_gather_mpole_field(int n, TileGeom g, const R4<real>* __restrict__ pw,
                                                            const int4* __restrict__ pbase,
                                                            const cplx_t<real>* __restrict__ grid, real a1, real a2,
                                                            real a3, real term, const real* __restrict__ rpole,
                                                            real* __restrict__ fd, real* __restrict__ fp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const R4<real> w = pw[s];
  const int4 b = pbase[s];
  real th1[2][5], th2[2][5], th3[2][5];
  bspline5<2>(w.x, th1);
  bspline5<2>(w.y, th2);
  bspline5<2>(w.z, th3);
  real d1 = 0, d2 = 0, d3 = 0;
  for (int m3 = 0; m3 < 5; ++m3) {
    int z = b.z + m3;
    if (z >= g.K[2]) z -= g.K[2];
    const real v0 = th3[0][m3], v1 = th3[1][m3];
    for (int m2 = 0; m2 < 5; ++m2) {
      int y = b.y + m2;
      if (y >= g.K[1]) y -= g.K[1];
      const real u0 = th2[0][m2], u1 = th2[1][m2];
      const long long row = (long long)g.K[0] * (y + (long long)g.K[1] * z);
      for (int m1 = 0; m1 < 5; ++m1) {
        int x = b.x + m1;
        if (x >= g.K[0]) x -= g.K[0];
        const real q = grid[row + x].x;
        const real t0 = th1[0][m1], t1 = th1[1][m1];
        d1 += q * t1 * u0 * v0;
        d2 += q * t0 * u1 * v0;
        d3 += q * t0 * u0 * v1;
      }
    }
  }
  const real* p = rpole + 10 * (size_t)s;
  const real f1 = term * p[1] - a1 * d1, f2 = term * p[2] - a2 * d2, f3 = term * p[3] - a3 * d3;
  fd[3 * s] += f1;
  fd[3 * s + 1] += f2;
  fd[3 * s + 2] += f3;
  fp[3 * s] += f1;
  fp[3 * s + 1] += f2;
  fp[3 * s + 2] += f3;
}

// ------------------------------------------------------------ host side

FftPlan make_plan(int N) {
  FftPlan p{};
  p.n = N;
  int m = N;
  for (int R : {4, 2, 3, 5})
    while (m % R == 0) {
      if (p.nstage >= 16) throw ArgError("FFT plan too long");
      p.radix[p.nstage++] = R;
      m /= R;
    }
  if (m != 1) throw ArgError("PME grid size " + std::to_string(N) + " is not a product of 2, 3 and 5");
  return p;
}

TileGeom geom(const Ctx& c) {
  TileGeom g;
  for (int d = 0; d < 3; ++d) {
    g.K[d] = c.grid[d];
    g.T[d] = c.tsz[d];
    g.nt[d] = c.ntile[d];
  }
  return g;
}

template <class real>
constexpr int lines_per_block() {
  return sizeof(real) == 4 ? 16 : 8;
}

template <class real>
void fft_roundtrip(Ctx& c) {
  constexpr int B = lines_per_block<real>();
  const int K1 = c.grid[0], K2 = c.grid[1], K3 = c.grid[2];
  auto* grid = c.qgrid.as<cplx_t<real>>();
  const auto* tw = c.twiddle.as<cplx_t<real>>();
  FftPlan p1 = make_plan(K1), p2 = make_plan(K2), p3 = make_plan(K3);
  ConvArgs none{};
  ConvArgs conv{};
  conv.enabled = 1;
  conv.pterm = (kPi / c.aewald) * (kPi / c.aewald);
  conv.volterm = kPi * c.box[0] * c.box[1] * c.box[2];
  for (int d = 0; d < 3; ++d) conv.il[d] = 1.0 / c.box[d];
  conv.bsmod = c.bsmod.as<double>();
  const int xb = (K1 + B - 1) / B;
  auto launch = [&](int axis, int sign, const FftPlan& p, const cplx_t<real>* t, const ConvArgs& cv) {
    int nblk = axis == 0 ? ceil_div((long)K2 * K3, B) : (axis == 1 ? xb * K3 : xb * K2);
    size_t sh = 2 * (size_t)B * p.n * sizeof(cplx_t<real>);
    KScope ks_(c, KT_FFT);
    hipLaunchKernelGGL((k_fft_pass<real, B>), dim3(nblk), dim3(256), sh, c.stream, grid, K1, K2, K3, axis, sign, p, t,
                       cv);
    T9_LAUNCH_CHECK();
  };
  launch(0, -1, p1, tw, none);
  launch(1, -1, p2, tw + K1, none);
  launch(2, -1, p3, tw + K1 + K2, conv);
  launch(1, +1, p2, tw + K1, none);
  launch(0, +1, p1, tw, none);
}

template <class real>
void pme_bin_t(Ctx& c) {
  const int n = c.n;
  const int ntot = c.ntile[0] * c.ntile[1] * c.ntile[2];
  T9_HIP(hipMemsetAsync(c.tile_count.p, 0, sizeof(int) * ntot, c.stream));
  T9_HIP(hipMemsetAsync(c.tile_cursor.p, 0, sizeof(int) * ntot, c.stream));
  hipLaunchKernelGGL(k_pme_bin<real>, dim3(ceil_div(n, 256)), dim3(256), 0, c.stream, n, c.perm.as<int>(),
                     c.xyz.as<double>(), c.box[0], c.box[1], c.box[2], c.grid[0], c.grid[1], c.grid[2], c.tsz[0],
                     c.tsz[1], c.tsz[2], c.ntile[0], c.ntile[1], c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(),
                     c.tile_count.as<int>());
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, c.stream, c.tile_count.as<int>(), c.tile_start.as<int>(),
                     ntot);
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tile_scatter, dim3(ceil_div(n, 256)), dim3(256), 0, c.stream, n, c.pme_base.as<int4>(),
                     c.tile_start.as<int>(), c.tile_cursor.as<int>(), c.tile_atoms.as<int>());
  T9_LAUNCH_CHECK();
  c.bins_valid = true;
}

template <class real>
void spread_t(Ctx& c, int mode, const void* s1, const void* s2) {
  const int ntot = c.ntile[0] * c.ntile[1] * c.ntile[2];
  const size_t sh = 2 * sizeof(long long) * c.tsz[0] * c.tsz[1] * c.tsz[2];
  const real a1 = (real)(c.grid[0] / c.box[0]), a2 = (real)(c.grid[1] / c.box[1]), a3 = (real)(c.grid[2] / c.box[2]);
  auto kern = mode == 0 ? k_spread<real, 0> : k_spread<real, 1>;
  KScope ks_(c, KT_SPREAD);
  hipLaunchKernelGGL(kern, dim3(ntot), dim3(256), sh, c.stream, geom(c), c.tile_start.as<int>(),
                     c.tile_count.as<int>(), c.tile_atoms.as<int>(), c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(),
                     (const real*)s1, (const real*)s2, a1, a2, a3, RealTraits<real>::fixed_scale,
                     c.qgrid.as<cplx_t<real>>());
  T9_LAUNCH_CHECK();
}

template <class real>
void dipole_field_t(Ctx& c, const void* ud, const void* up, void* fd, void* fp) {
  if (!c.bins_valid) pme_bin_t<real>(c);
  spread_t<real>(c, 0, ud, up);
  fft_roundtrip<real>(c);
  const real a1 = (real)(c.grid[0] / c.box[0]), a2 = (real)(c.grid[1] / c.box[1]), a3 = (real)(c.grid[2] / c.box[2]);
  const real term = (real)(4.0 / 3.0 * c.aewald * c.aewald * c.aewald / 1.77245385090551602730);
  KScope ks_(c, KT_GATHER);
  hipLaunchKernelGGL(k_gather_dip<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, geom(c),
                     c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), c.qgrid.as<cplx_t<real>>(), a1, a2, a3, term,
                     (const real*)ud, (const real*)up, (real*)fd, (real*)fp);
  T9_LAUNCH_CHECK();
}

template <class real>
void mpole_field_t(Ctx& c, void* fd, void* fp) {
  if (!c.bins_valid) pme_bin_t<real>(c);
  spread_t<real>(c, 1, c.rpole.p, nullptr);
  fft_roundtrip<real>(c);
  const real a1 = (real)(c.grid[0] / c.box[0]), a2 = (real)(c.grid[1] / c.box[1]), a3 = (real)(c.grid[2] / c.box[2]);
  const real term = (real)(4.0 / 3.0 * c.aewald * c.aewald * c.aewald / 1.77245385090551602730);
  KScope ks_(c, KT_GATHER);
  hipLaunchKernelGGL(k_gather_mpole_field<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, geom(c),
                     c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), c.qgrid.as<cplx_t<real>>(), a1, a2, a3, term,
                     c.rpole.as<real>(), (real*)fd, (real*)fp);
  T9_LAUNCH_CHECK();
}

}  // namespace

void pme_bin(Ctx& c) {
  if (c.prec == 32) pme_bin_t<float>(c);
  else pme_bin_t<double>(c);
}

void pme_dipole_field(Ctx& c, const void* ud, const void* up, void* fd, void* fp) {
  if (c.prec == 32) dipole_field_t<float>(c, ud, up, fd, fp);
  else dipole_field_t<double>(c, ud, up, fd, fp);
}

void pme_mpole_field(Ctx& c, void* fd, void* fp) {
  if (c.prec == 32) mpole_field_t<float>(c, fd, fp);
  else mpole_field_t<double>(c, fd, fp);
}

// B-spline moduli of one axis (the dftmod rule, with the zeta correction),
// FP64 on the host.
std::vector<double> pme_bsmod(int K, int order) {
  auto M = [](int p, double x) {
    // cardinal B-spline by its explicit piecewise sum
    double s = 0, fact = 1;
    for (int k = 1; k < p; ++k) fact *= k;
    for (int k = 0; k <= p; ++k) {
      double t = x - k;
      if (t <= 0) continue;
      double c = 1;
      for (int q = 0; q < k; ++q) c = c * (p - q) / (q + 1);
      s += ((k & 1) ? -c : c) * std::pow(t, p - 1);
    }
    return (x <= 0 || x >= p) ? 0.0 : s / fact;
  };
  std::vector<double> arr(K, 0.0), mod(K, 0.0);
  for (int k = 1; k < order && k < K; ++k) arr[k] = M(order, (double)k);
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
  for (int m = 0; m < K; ++m) {
    int k = (m + 1 > K / 2) ? m - K : m;
    double zeta = 1.0;
    if (k != 0) {
      double s1 = 1, s2 = 1, fac = kPi * k / K;
      for (int j = 1; j <= 50; ++j) {
        double a1 = fac / (fac + kPi * j), a2 = fac / (fac - kPi * j);
        s1 += std::pow(a1, order) + std::pow(a2, order);
        s2 += std::pow(a1, 2 * order) + std::pow(a2, 2 * order);
      }
      zeta = s2 / s1;
    }
    mod[m] *= zeta * zeta;
  }
  return mod;
}

// Tile edge dividing K, closest to 8, between 4 and 16 (K <= 16: one tile).
int pme_tile_edge(int K) {
  if (K <= 16) return K;
  int best = -1;
  for (int t = 4; t <= 16; ++t)
    if (K % t == 0 && (best < 0 || std::abs(t - 8) < std::abs(best - 8))) best = t;
  if (best < 0) throw ArgError("PME grid size " + std::to_string(K) + " has no divisor between 4 and 16");
  return best;
}

}  // namespace t9
