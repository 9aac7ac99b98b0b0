// 3D FFT of the PME grid for gfx950 (replaces cuFFT C2C in place,
// src/cudart/fft.cpp:16-99).
//
// Line passes, five per reciprocal-space round trip: x forward, y forward,
// z forward + influence function + z inverse (one kernel, data stays in
// LDS), y inverse, x inverse; FP32 square planes take the x and y passes in
// one plane kernel each way (three grid passes).  Each workgroup loads 16 (FP32) or 8 (FP64)
// lines — contiguous lines for x, lines at consecutive x for y and z so every
// global row access is a contiguous 128-byte segment — transforms them in
// LDS with the Stockham stages of fft_core.hpp and writes them back.
// The influence function is the reference's pmeConv
// (src/acc/pme.cpp:220-303): exp(-pi^2 h^2 / a^2) / (pi V h^2 b1 b2 b3),
// k = 0 set to zero, terms with exponent below -50 dropped; it is evaluated
// in FP64 once per box into a table (z-pass order) and on the fly when the
// virial sums are wanted.  Transform lengths used by the PME grids of
// interest (24 ... 160) get kernels with compile-time index arithmetic.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "ctx.hpp"
#include "fft_core.hpp"

namespace t9 {
namespace {

struct ConvArgs {
  int enabled;
  const double* etab;  // separable factors of the influence function ([K1 + K2 + K3]), or null
  double pterm;    // (pi / alpha)^2
  double volterm;  // pi * V
  double il[3];    // 1 / box edge
  const double* bsmod;
  int fwd_only;    // stop after the forward transform (virial sums only)
  double vscale;   // weight of the convolution virial sums
  double* vrows;   // per-block partial rows, or null
};

// Influence function of reciprocal point (k1, k2, k3) (the reference's
// pmeConv, src/acc/pme.cpp:220-303) from its separable factors:
// exp(-pterm h^2) / (pi V bsmod1 bsmod2 bsmod3 h^2) = e1[k1] e2[k2] e3[k3] / h^2,
// e_d[k] = exp(-pterm h_d^2) / bsmod_d[k] (and 1 / (pi V) in e1); k = 0 and
// exponents below -50 give 0.  FP64.
__device__ __forceinline__ double conv_factor(const ConvArgs& cv, int K1, int K2, int K3, int k1, int k2, int k3) {
  if (k1 == 0 && k2 == 0 && k3 == 0) return 0.0;
  const int r1 = k1 < (K1 + 1) / 2 ? k1 : k1 - K1;
  const int r2 = k2 < (K2 + 1) / 2 ? k2 : k2 - K2;
  const int r3 = k3 < (K3 + 1) / 2 ? k3 : k3 - K3;
  const double h1 = r1 * cv.il[0], h2 = r2 * cv.il[1], h3 = r3 * cv.il[2];
  const double hsq = h1 * h1 + h2 * h2 + h3 * h3;
  if (!(-cv.pterm * hsq > -50)) return 0.0;
  return cv.etab[k1] * cv.etab[K1 + k2] * cv.etab[K1 + K2 + k3] / hsq;
}

struct DevBarrier {
  __device__ void operator()() const { __syncthreads(); }
};

__device__ __forceinline__ long long elem_addr(int axis, int K1, int K2, int lb, int e, long long l0, int x0,
                                               int other) {
  if (axis == 0) return (l0 + lb) * K1 + e;
  if (axis == 1) return (x0 + lb) + (long long)K1 * (e + (long long)K2 * other);
  return (x0 + lb) + (long long)K1 * (other + (long long)K2 * e);
}

// FP32 passes with a compile-time length issue all of a thread's loads
// before the first LDS write and keep a copy of the twiddles in LDS; the
// FP64 passes keep the plain load loop (their prefetch arrays would go to
// scratch memory) and read the twiddles from global memory.
template <class real>
__host__ __device__ constexpr bool twiddles_in_lds() {
  return sizeof(real) == 4;
}

// transform lengths with a compile-time kernel (the switch in roundtrip_t)
inline bool ct_len(int n) {
  switch (n) {
    case 24: case 32: case 40: case 48: case 64: case 80: case 96: case 120: case 128: case 160: return true;
    default: return false;
  }
}

// NC > 0: the transform length is the compile-time NC (fast index
// arithmetic); NC == 0: any 2,3,5-smooth length from `plan`.
template <class real, int B, int NC>
__global__ __launch_bounds__(256) void k_fft_pass(Cplx<real>* __restrict__ grid, int K1, int K2, int K3, int axis,
                                                  int sign, FftPlan plan, const Cplx<real>* __restrict__ tw,
                                                  ConvArgs conv, Gate gate) {
  if (gate_closed(gate)) return;
  extern __shared__ __align__(16) char smem[];
  const int N = NC > 0 ? NC : plan.n;
  Cplx<real>* buf0 = reinterpret_cast<Cplx<real>*>(smem);
  const int LP = N + 1;  // padded LDS line pitch: adjacent lines start on different banks
  Cplx<real>* buf1 = buf0 + B * LP;
  int nl, x0 = 0, other = 0;
  long long l0 = 0;
  if (axis == 0) {
    const int nlines = K2 * K3;
    l0 = (long long)blockIdx.x * B;
    nl = min(B, nlines - (int)l0);
  } else {
    const int xb = (K1 + B - 1) / B;
    x0 = (blockIdx.x % xb) * B;
    other = blockIdx.x / xb;
    nl = min(B, K1 - x0);
  }
  const int tot = nl * N;
  // element q of the block's lines -> (line, index along the transform)
  auto split = [&](int q, int& lb, int& e) {
    if (axis == 0) {
      lb = q / N;
      e = q - lb * N;
    } else if (nl == B) {
      e = q / B;
      lb = q - e * B;
    } else {
      e = q / nl;
      lb = q - e * nl;
    }
  };
  // z pass with the influence function computed from its separable factors
  const bool tabconv = conv.enabled && conv.etab && !conv.vrows;
  const Cplx<real>* twk = tw;
  constexpr bool kPipe = NC > 0 && twiddles_in_lds<real>();
  constexpr int PER = kPipe ? (B * NC + 255) / 256 : 1;  // elements per thread (256 threads)
  if constexpr (kPipe) {
    // all of a thread's loads are issued before the first LDS write; the
    // twiddles (and the z pass's influence-table values) come in the same phase
    Cplx<real> v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int q = threadIdx.x + 256 * k;
      if (q < tot) {
        int lb, e;
        split(q, lb, e);
        v[k] = grid[elem_addr(axis, K1, K2, lb, e, l0, x0, other)];
      }
    }
    Cplx<real>* twl = buf1 + B * LP;
    if (threadIdx.x < NC) twl[threadIdx.x] = tw[threadIdx.x];
    twk = twl;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int q = threadIdx.x + 256 * k;
      if (q < tot) {
        int lb, e;
        split(q, lb, e);
        buf0[lb * LP + e] = v[k];
      }
    }
  } else {
    for (int q = threadIdx.x; q < tot; q += blockDim.x) {
      int lb, e;
      split(q, lb, e);
      buf0[lb * LP + e] = grid[elem_addr(axis, K1, K2, lb, e, l0, x0, other)];
    }
  }
  __syncthreads();
  Cplx<real>* res;
  if constexpr (NC > 0) res = fft_stages_ct<real, NC>(buf0, buf1, nl, twk, sign, threadIdx.x, blockDim.x);
  else res = fft_stages<real>(buf0, buf1, nl, LP, plan, tw, sign, threadIdx.x, blockDim.x, DevBarrier());
  if (tabconv) {
    // convolution with the influence function: lines are (x0 + lb, y =
    // other), element index k3
    for (int q = threadIdx.x; q < tot; q += blockDim.x) {
      const int lb = q / N, k3 = q - lb * N;
      const real f = (real)conv_factor(conv, K1, K2, K3, x0 + lb, other, k3);
      Cplx<real> v = res[lb * LP + k3];
      v.x *= f;
      v.y *= f;
      res[lb * LP + k3] = v;
    }
    __syncthreads();
    Cplx<real>* spare = res == buf0 ? buf1 : buf0;
    if constexpr (NC > 0) res = fft_stages_ct<real, NC>(res, spare, nl, twk, -sign, threadIdx.x, blockDim.x);
    else res = fft_stages<real>(res, spare, nl, LP, plan, tw, -sign, threadIdx.x, blockDim.x, DevBarrier());
  } else if (conv.enabled) {
    // z pass: lines are (x0 + lb, y = other), element index is k3.  The
    // strain derivative of 1/2 sum_k G(k) |S(k)|^2 (G the influence function)
    // is accumulated from the transformed values before they are scaled.
    double vs[6] = {0, 0, 0, 0, 0, 0};
    for (int q = threadIdx.x; q < tot; q += blockDim.x) {
      const int lb = q / N, k3 = q - lb * N;
      const int k1 = x0 + lb, k2 = other;
      double f = 0, term = 0, hsq = 1, h1 = 0, h2 = 0, h3 = 0;
      if (k1 != 0 || k2 != 0 || k3 != 0) {
        const int r1 = k1 < (K1 + 1) / 2 ? k1 : k1 - K1;
        const int r2 = k2 < (K2 + 1) / 2 ? k2 : k2 - K2;
        const int r3 = k3 < (K3 + 1) / 2 ? k3 : k3 - K3;
        h1 = r1 * conv.il[0];
        h2 = r2 * conv.il[1];
        h3 = r3 * conv.il[2];
        hsq = h1 * h1 + h2 * h2 + h3 * h3;
        term = -conv.pterm * hsq;
        if (term > -50)
          f = exp(term) / (hsq * conv.volterm * conv.bsmod[k1] * conv.bsmod[K1 + k2] * conv.bsmod[K1 + K2 + k3]);
      }
      Cplx<real> v = res[lb * LP + k3];
      if (conv.vrows && f != 0) {
        const double et = 0.5 * f * ((double)v.x * v.x + (double)v.y * v.y);
        const double vt = 2.0 * (1.0 - term) / hsq * et;
        vs[0] += h1 * h1 * vt - et;
        vs[1] += h2 * h2 * vt - et;
        vs[2] += h3 * h3 * vt - et;
        vs[3] += h1 * h2 * vt;
        vs[4] += h1 * h3 * vt;
        vs[5] += h2 * h3 * vt;
      }
      if (!conv.fwd_only) {
        v.x = (real)(v.x * f);
        v.y = (real)(v.y * f);
        res[lb * LP + k3] = v;
      }
    }
    if (conv.vrows) {
      double red[EP_N];
      for (int q = 0; q < EP_N; ++q) red[q] = 0;
      for (int q = 0; q < 6; ++q) red[EP_VXX + q] = conv.vscale * vs[q];
      block_row<EP_N>(red, conv.vrows + (size_t)blockIdx.x * kEnergyRow);
    }
    if (conv.fwd_only) return;
    __syncthreads();
    Cplx<real>* spare = res == buf0 ? buf1 : buf0;
    if constexpr (NC > 0) res = fft_stages_ct<real, NC>(res, spare, nl, twk, -sign, threadIdx.x, blockDim.x);
    else res = fft_stages<real>(res, spare, nl, LP, plan, tw, -sign, threadIdx.x, blockDim.x, DevBarrier());
  }
  for (int q = threadIdx.x; q < tot; q += blockDim.x) {
    int lb, e;
    if (axis == 0) {
      lb = q / N;
      e = q - lb * N;
    } else {
      e = q / nl;
      lb = q - e * nl;
    }
    grid[elem_addr(axis, K1, K2, lb, e, l0, x0, other)] = res[lb * LP + e];
  }
}

// Separable factors of the influence function, FP64 (see conv_factor):
// e_d[k] = exp(-pterm h_d^2) / bsmod_d[k], h_d the signed frequency over the
// box edge, with 1 / (pi V) folded into axis 0.
__global__ void k_conv_etab(int K1, int K2, int K3, double pterm, double volterm, double il0, double il1, double il2,
                            const double* __restrict__ bsmod, double* __restrict__ etab) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K1 + K2 + K3) return;
  int K = K1, k = t, d = 0;
  if (t >= K1 + K2) {
    K = K3;
    k = t - K1 - K2;
    d = 2;
  } else if (t >= K1) {
    K = K2;
    k = t - K1;
    d = 1;
  }
  const int r = k < (K + 1) / 2 ? k : k - K;
  const double h = r * (d == 0 ? il0 : (d == 1 ? il1 : il2));
  double e = exp(-pterm * h * h) / bsmod[t];
  if (d == 0) e /= volterm;
  etab[t] = e;
}

// ---------------------------------------------------------------- planes
// FP32 x and y transforms of one z-plane (N x N, contiguous in the grid) in
// LDS: one 1024-thread workgroup per plane loads it with 16-byte loads,
// runs the Stockham stages of every row, then of every column, in place, and
// stores it back.  A stage in place: each thread reads its butterflies'
// inputs into registers and computes them, a barrier, then it writes the
// outputs.  Each butterfly is the arithmetic of the line passes (same
// indices, twiddles and radix order), so the result is the same bits; the
// round trip takes 3 grid passes (plane, z with the influence function,
// plane) instead of 5.
constexpr int kPlaneThreads = 1024;

// Element (line lb, index e) of the plane at lb * LS + e * ES: rows have
// LS = N + 1, ES = 1; columns LS = 1, ES = N + 1.  Threads take butterflies
// line-fastest, so a wavefront's column accesses are contiguous.
template <int N, int LS, int ES, int ST = 0, int NS = 1>
__device__ __forceinline__ void plane_stages(Cplx<float>* buf, const Cplx<float>* tw, float sg) {
  if constexpr (ST < ct_nstage(N)) {
    constexpr int R = ct_radix(N, ST);
    constexpr int nb = N / R, total = N * nb, PER = (total + kPlaneThreads - 1) / kPlaneThreads;
    constexpr int tstep = N / (NS * R);
    Cplx<float> v[PER][R];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = threadIdx.x + kPlaneThreads * u;
      if (q < total) {
        const int j = q / N, lb = q - j * N;
#pragma unroll
        for (int r = 0; r < R; ++r) v[u][r] = buf[lb * LS + (j + r * nb) * ES];
        if constexpr (NS > 1) {
          const int k = j - (j / NS) * NS;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            Cplx<float> t = tw[k * r * tstep];
            t.y = sg > 0 ? -t.y : t.y;
            v[u][r] = cmulc(v[u][r], t);
          }
        }
        small_dft<R>(v[u], sg);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = threadIdx.x + kPlaneThreads * u;
      if (q < total) {
        const int j = q / N, lb = q - j * N;
        const int jd = j / NS, k = j - jd * NS;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[lb * LS + (jd * NS * R + k + r * NS) * ES] = v[u][r];
      }
    }
    __syncthreads();
    plane_stages<N, LS, ES, ST + 1, NS * R>(buf, tw, sg);
  }
}

template <int N>
__global__ __launch_bounds__(kPlaneThreads) void k_fft_plane(Cplx<float>* __restrict__ grid, int sign,
                                                             const Cplx<float>* __restrict__ twx,
                                                             const Cplx<float>* __restrict__ twy, Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int P = N + 1;
  __shared__ Cplx<float> pl[N * P];
  __shared__ Cplx<float> tws[2 * N];
  float4* g = reinterpret_cast<float4*>(grid + (size_t)blockIdx.x * N * N);
  constexpr int NE = N * N / 2;  // two complex values per 16-byte load
  constexpr int PER = (NE + kPlaneThreads - 1) / kPlaneThreads;
  float4 v[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int q = threadIdx.x + kPlaneThreads * u;
    if (q < NE) v[u] = g[q];
  }
  if (threadIdx.x < N) tws[threadIdx.x] = twx[threadIdx.x];
  else if (threadIdx.x < 2 * N) tws[threadIdx.x] = twy[threadIdx.x - N];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int q = threadIdx.x + kPlaneThreads * u;
    if (q < NE) {
      const int y = (2 * q) / N, x = 2 * q - y * N;
      pl[y * P + x] = Cplx<float>{v[u].x, v[u].y};
      pl[y * P + x + 1] = Cplx<float>{v[u].z, v[u].w};
    }
  }
  __syncthreads();
  const float sg = (float)sign;
  if (sign < 0) {  // forward: x, then y (the order of the line passes)
    plane_stages<N, P, 1>(pl, tws, sg);
    plane_stages<N, 1, P>(pl, tws + N, sg);
  } else {         // inverse: y, then x
    plane_stages<N, 1, P>(pl, tws + N, sg);
    plane_stages<N, P, 1>(pl, tws, sg);
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int q = threadIdx.x + kPlaneThreads * u;
    if (q < NE) {
      const int y = (2 * q) / N, x = 2 * q - y * N;
      const Cplx<float> a = pl[y * P + x], b = pl[y * P + x + 1];
      g[q] = float4{a.x, a.y, b.x, b.y};
    }
  }
}

// plane lengths with a kernel (LDS: N (N + 1) + 2 N complex values)
inline bool plane_len(int n) {
  switch (n) {
    case 24: case 32: case 40: case 48: case 64: case 80: case 96: case 120: case 128: return true;
    default: return false;
  }
}

// 8 lines per 256-thread workgroup in both precisions: for K = 120 FP32 the
// z pass then runs 1,800 workgroups of 15.5 KB LDS, several per CU, so one
// workgroup's loads overlap another's LDS stages (16 lines: 0.7 % slower, 4
// lines: 3 % slower on the FFT family in the MD step)
template <class real>
constexpr int lines_per_block() {
  return 8;
}

// One line pass of the 3D transform along `axis` (sign -1 forward, +1
// inverse), with the influence function when cv.enabled (z pass only).
template <class real>
void launch_pass(Cplx<real>* grid, int K1, int K2, int K3, int axis, int sign, const FftPlan& plan,
                 const Cplx<real>* tw, const ConvArgs& cv, Gate gate, hipStream_t st) {
  constexpr int B = lines_per_block<real>();
  const int xb = (K1 + B - 1) / B;
  const int nblk = axis == 0 ? ceil_div((long)K2 * K3, B) : (axis == 1 ? xb * K3 : xb * K2);
  // two line buffers + the twiddles (compile-time lengths stage them in LDS)
  const size_t sh =
      (2 * (size_t)B * (plan.n + 1) + (twiddles_in_lds<real>() && ct_len(plan.n) ? plan.n : 0)) * sizeof(Cplx<real>);
  void (*kern)(Cplx<real>*, int, int, int, int, int, FftPlan, const Cplx<real>*, ConvArgs, Gate);
  switch (plan.n) {
    case 24: kern = k_fft_pass<real, B, 24>; break;
    case 32: kern = k_fft_pass<real, B, 32>; break;
    case 40: kern = k_fft_pass<real, B, 40>; break;
    case 48: kern = k_fft_pass<real, B, 48>; break;
    case 64: kern = k_fft_pass<real, B, 64>; break;
    case 80: kern = k_fft_pass<real, B, 80>; break;
    case 96: kern = k_fft_pass<real, B, 96>; break;
    case 120: kern = k_fft_pass<real, B, 120>; break;
    case 128: kern = k_fft_pass<real, B, 128>; break;
    case 160: kern = k_fft_pass<real, B, 160>; break;
    default: kern = k_fft_pass<real, B, 0>; break;
  }
  hipLaunchKernelGGL(kern, dim3(nblk), dim3(256), sh, st, grid, K1, K2, K3, axis, sign, plan, tw, cv, gate);
  T9_LAUNCH_CHECK();
}

// Unnormalised forward (sign -1: x, y, z) or inverse (+1: z, y, x) 3D
// transform in place, line passes only; tw holds the forward twiddles of
// each axis back to back (K1 + K2 + K3 values).
template <class real>
void fft3d_lines(Cplx<real>* grid, const int* K, int sign, const Cplx<real>* tw, hipStream_t st) {
  const FftPlan p[3] = {make_fft_plan(K[0]), make_fft_plan(K[1]), make_fft_plan(K[2])};
  const Cplx<real>* twa[3] = {tw, tw + K[0], tw + K[0] + K[1]};
  ConvArgs none{};
  for (int q = 0; q < 3; ++q) {
    const int axis = sign < 0 ? q : 2 - q;
    launch_pass<real>(grid, K[0], K[1], K[2], axis, sign, p[axis], twa[axis], none, Gate{}, st);
  }
}

// Reciprocal-space influence function applied point by point (the
// reference's pmeConv, src/acc/pme.cpp:220-303): grid index
// k1 + K1 (k2 + K2 k3), FP64 factor, k = 0 and exponents below -50 zeroed.
template <class real>
__global__ void k_pme_conv(int K1, int K2, int K3, double pterm, double volterm, double il0, double il1, double il2,
                           const double* __restrict__ bsmod, Cplx<real>* __restrict__ grid) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= (long long)K1 * K2 * K3) return;
  const int k1 = (int)(e % K1);
  const long long l = e / K1;
  const int k2 = (int)(l % K2), k3 = (int)(l / K2);
  double f = 0;
  if (k1 != 0 || k2 != 0 || k3 != 0) {
    const int r1 = k1 < (K1 + 1) / 2 ? k1 : k1 - K1;
    const int r2 = k2 < (K2 + 1) / 2 ? k2 : k2 - K2;
    const int r3 = k3 < (K3 + 1) / 2 ? k3 : k3 - K3;
    const double h1 = r1 * il0, h2 = r2 * il1, h3 = r3 * il2;
    const double hsq = h1 * h1 + h2 * h2 + h3 * h3;
    const double term = -pterm * hsq;
    if (term > -50) f = exp(term) / (hsq * volterm * bsmod[k1] * bsmod[K1 + k2] * bsmod[K1 + K2 + k3]);
  }
  Cplx<real> v = grid[e];
  v.x = (real)(v.x * f);
  v.y = (real)(v.y * f);
  grid[e] = v;
}

template <class real>
void roundtrip_t(Ctx& c, bool fwd_only, double* vrows, double vscale, bool lines = false) {
  constexpr int B = lines_per_block<real>();
  const int K1 = c.grid[0], K2 = c.grid[1], K3 = c.grid[2];
  auto* grid = c.qgrid.as<Cplx<real>>();
  const auto* tw = c.twiddle.as<Cplx<real>>();
  FftPlan p[3] = {make_fft_plan(K1), make_fft_plan(K2), make_fft_plan(K3)};
  ConvArgs none{};
  ConvArgs conv{};
  conv.enabled = 1;
  const double pi = 3.14159265358979323846;
  conv.pterm = (pi / c.aewald) * (pi / c.aewald);
  conv.volterm = pi * c.box[0] * c.box[1] * c.box[2];
  for (int d = 0; d < 3; ++d) conv.il[d] = 1.0 / c.box[d];
  conv.bsmod = c.bsmod.as<double>();
  conv.fwd_only = fwd_only ? 1 : 0;
  conv.vscale = vscale;
  conv.vrows = vrows;
  const Cplx<real>* twa[3] = {tw, tw + K1, tw + K1 + K2};
  if (!c.conv_table_valid) {
    c.conv_table.alloc(sizeof(double) * (size_t)(K1 + K2 + K3));
    hipLaunchKernelGGL(k_conv_etab, dim3(ceil_div(K1 + K2 + K3, 256)), dim3(256), 0, c.stream, K1, K2, K3,
                       conv.pterm, conv.volterm, conv.il[0], conv.il[1], conv.il[2], conv.bsmod,
                       c.conv_table.as<double>());
    T9_LAUNCH_CHECK();
    c.conv_table_valid = true;
  }
  conv.etab = c.conv_table.as<double>();
  auto launch = [&](int axis, int sign, const ConvArgs& cv) {
    launch_pass<real>(grid, K1, K2, K3, axis, sign, p[axis], twa[axis], cv, c.gate, c.stream);
  };
  if (c.fft_plane < 0) {
    const char* fp = std::getenv("T9HIP_FFT");
    c.fft_plane = sizeof(real) == 4 && K1 == K2 && plane_len(K1) && !(fp && !std::strcmp(fp, "lines")) ? 1 : 0;
  }
  if constexpr (sizeof(real) == 4) {
    if (c.fft_plane == 1 && !lines) {
      void (*pk)(Cplx<float>*, int, const Cplx<float>*, const Cplx<float>*, Gate);
      switch (K1) {
        case 24: pk = k_fft_plane<24>; break;
        case 32: pk = k_fft_plane<32>; break;
        case 40: pk = k_fft_plane<40>; break;
        case 48: pk = k_fft_plane<48>; break;
        case 64: pk = k_fft_plane<64>; break;
        case 80: pk = k_fft_plane<80>; break;
        case 96: pk = k_fft_plane<96>; break;
        case 120: pk = k_fft_plane<120>; break;
        default: pk = k_fft_plane<128>; break;
      }
      KScope ks_(c, KT_FFT, fwd_only ? 2 : 3);
      hipLaunchKernelGGL(pk, dim3(K3), dim3(kPlaneThreads), 0, c.stream, grid, -1, twa[0], twa[1], c.gate);
      T9_LAUNCH_CHECK();
      launch(2, -1, conv);
      if (fwd_only) return;
      hipLaunchKernelGGL(pk, dim3(K3), dim3(kPlaneThreads), 0, c.stream, grid, +1, twa[0], twa[1], c.gate);
      T9_LAUNCH_CHECK();
      return;
    }
  }
  KScope ks_(c, KT_FFT, fwd_only ? 3 : 5);  // the passes run back to back
  launch(0, -1, none);
  launch(1, -1, none);
  launch(2, -1, conv);
  if (fwd_only) return;
  launch(1, +1, none);
  launch(0, +1, none);
}

template <class real>
int zpass_blocks(const Ctx& c) {
  constexpr int B = lines_per_block<real>();
  return ((c.grid[0] + B - 1) / B) * c.grid[1];
}

}  // namespace

void pme_fft_roundtrip(Ctx& c, bool lines) {
  if (c.prec == 32) roundtrip_t<float>(c, false, nullptr, 0.0, lines);
  else roundtrip_t<double>(c, false, nullptr, 0.0, lines);
}

void pme_fft_virial(Ctx& c, int roundtrip, double vscale) {
  if (!roundtrip && vscale == 0) return;
  const int nz = c.prec == 32 ? zpass_blocks<float>(c) : zpass_blocks<double>(c);
  double* rows = vscale != 0 ? energy_rows(c, nz) : nullptr;
  if (c.prec == 32) roundtrip_t<float>(c, !roundtrip, rows, vscale);
  else roundtrip_t<double>(c, !roundtrip, rows, vscale);
}

void pme_check_fft_sizes(const Ctx& c) {
  const size_t esize = c.prec == 32 ? 8 : 16;
  const int B = c.prec == 32 ? 16 : 8;
  for (int d = 0; d < 3; ++d) {
    if (make_fft_plan(c.grid[d]).nstage < 0)
      throw ArgError("PME grid size " + std::to_string(c.grid[d]) + " is not a product of 2, 3 and 5");
    if (2 * (size_t)B * (c.grid[d] + 1) * esize > 65536)
      throw ArgError("PME grid dimension too large for the LDS FFT");
  }
}

}  // namespace t9

namespace t9 {

void pme_fft_dir(Ctx& c, int sign) {
  if (c.prec == 32)
    fft3d_lines<float>(c.qgrid.as<Cplx<float>>(), c.grid, sign, c.twiddle.as<Cplx<float>>(), c.stream);
  else
    fft3d_lines<double>(c.qgrid.as<Cplx<double>>(), c.grid, sign, c.twiddle.as<Cplx<double>>(), c.stream);
}

void pme_conv_grid(Ctx& c) {
  const int K1 = c.grid[0], K2 = c.grid[1], K3 = c.grid[2];
  const double pi = 3.14159265358979323846;
  const double pterm = (pi / c.aewald) * (pi / c.aewald), volterm = pi * c.box[0] * c.box[1] * c.box[2];
  const long long ng = (long long)K1 * K2 * K3;
  if (c.prec == 32)
    hipLaunchKernelGGL(k_pme_conv<float>, dim3(ceil_div(ng, 256)), dim3(256), 0, c.stream, K1, K2, K3, pterm, volterm,
                       1.0 / c.box[0], 1.0 / c.box[1], 1.0 / c.box[2], c.bsmod.as<double>(),
                       c.qgrid.as<Cplx<float>>());
  else
    hipLaunchKernelGGL(k_pme_conv<double>, dim3(ceil_div(ng, 256)), dim3(256), 0, c.stream, K1, K2, K3, pterm,
                       volterm, 1.0 / c.box[0], 1.0 / c.box[1], 1.0 / c.box[2], c.bsmod.as<double>(),
                       c.qgrid.as<Cplx<double>>());
  T9_LAUNCH_CHECK();
}

void fft_check_size(int K, int prec) {
  const size_t esize = prec == 32 ? 8 : 16;
  const int B = prec == 32 ? 16 : 8;
  if (K < 1 || make_fft_plan(K).nstage < 0)
    throw ArgError("FFT length " + std::to_string(K) + " is not a product of 2, 3 and 5");
  if (2 * (size_t)B * (K + 1) * esize > 65536) throw ArgError("FFT length too large for the LDS line pass");
}

namespace {
// Forward twiddles of the three axes back to back: (cos, -sin)(2 pi k / K),
// evaluated in FP64 (sincospi: the angle's argument reduction is exact).
template <class real>
__global__ void k_fft_twiddles(int K1, int K2, int K3, Cplx<real>* __restrict__ tw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K1 + K2 + K3) return;
  int K = K1, k = t;
  if (t >= K1 + K2) {
    K = K3;
    k = t - K1 - K2;
  } else if (t >= K1) {
    K = K2;
    k = t - K1;
  }
  double sn, cs;
  sincospi(2.0 * k / K, &sn, &cs);
  tw[t] = Cplx<real>{(real)cs, (real)-sn};
}
}  // namespace

void fft_twiddles_device(void* tw, int prec, const int* K, hipStream_t st) {
  const int tot = K[0] + K[1] + K[2];
  if (prec == 32)
    hipLaunchKernelGGL(k_fft_twiddles<float>, dim3(ceil_div(tot, 256)), dim3(256), 0, st, K[0], K[1], K[2],
                       (Cplx<float>*)tw);
  else
    hipLaunchKernelGGL(k_fft_twiddles<double>, dim3(ceil_div(tot, 256)), dim3(256), 0, st, K[0], K[1], K[2],
                       (Cplx<double>*)tw);
  T9_LAUNCH_CHECK();
}

void fft3d_device(void* grid, int prec, const int* K, int sign, const void* tw, hipStream_t st) {
  if (prec == 32) fft3d_lines<float>((Cplx<float>*)grid, K, sign, (const Cplx<float>*)tw, st);
  else fft3d_lines<double>((Cplx<double>*)grid, K, sign, (const Cplx<double>*)tw, st);
}

// Scratch of the allocation-free transform: the twiddle table of the three
// axes, (K1 + K2 + K3) complex values of at most 16 bytes each, K <= 255 per
// axis.  One copy per GPU, held until the process ends; every call on a GPU
// rewrites it, so calls that overlap on one GPU race.
constexpr size_t kFftScratchBytes = 3 * 256 * 16;
static __device__ __align__(256) unsigned char scratch[kFftScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};

// Device address of `scratch` on the current GPU, looked up once per device
// with that device current.
void* fft_scratch_ptr() {
  int dev = -1;
  T9_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) throw ArgError("device index out of range for the FFT scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    T9_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

void fft3d_noalloc(void* grid, int prec, const int* K, int sign) {
  const size_t need = (prec == 32 ? 8 : 16) * ((size_t)K[0] + K[1] + K[2]);
  if (need > kFftScratchBytes)
    throw ArgError("FFT lengths " + std::to_string(K[0]) + "+" + std::to_string(K[1]) + "+" + std::to_string(K[2]) +
                   " need " + std::to_string(need) + " twiddle bytes, above the " +
                   std::to_string(kFftScratchBytes) + "-byte scratch table of the allocation-free transform");
  void* tw = fft_scratch_ptr();
  fft_twiddles_device(tw, prec, K, nullptr);
  fft3d_device(grid, prec, K, sign, tw, nullptr);
}

}  // namespace t9
