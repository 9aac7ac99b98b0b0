// 3D FFT of the PME grid for gfx950 (replaces cuFFT C2C in place,
// src/cudart/fft.cpp:16-99).
//
// Five passes per reciprocal-space round trip: x forward, y forward,
// z forward + influence function + z inverse (one kernel, data stays in
// LDS), y inverse, x inverse.  Each workgroup loads 16 (FP32) or 8 (FP64)
// lines — contiguous lines for x, lines at consecutive x for y and z so every
// global row access is a contiguous 128-byte segment — transforms them in
// LDS with the Stockham stages of fft_core.hpp and writes them back.
// The influence function is the reference's pmeConv
// (src/acc/pme.cpp:220-303): exp(-pi^2 h^2 / a^2) / (pi V h^2 b1 b2 b3),
// k = 0 set to zero, terms with exponent below -50 dropped; it is evaluated
// in FP64 once per box into a table (z-pass order) and on the fly when the
// virial sums are wanted.  Transform lengths used by the PME grids of
// interest (24 ... 160) get kernels with compile-time index arithmetic.
#include "ctx.hpp"
#include "fft_core.hpp"

namespace t9 {
namespace {

struct ConvArgs {
  int enabled;
  const void* table;  // precomputed influence function (real [K1][K2][K3] in z-pass order), or null
  double pterm;    // (pi / alpha)^2
  double volterm;  // pi * V
  double il[3];    // 1 / box edge
  const double* bsmod;
  int fwd_only;    // stop after the forward transform (virial sums only)
  double vscale;   // weight of the convolution virial sums
  double* vrows;   // per-block partial rows, or null
};

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
  const bool tabconv = conv.enabled && conv.table && !conv.vrows;
  const real* tb = (const real*)conv.table + ((long long)other * K1 + x0) * N;
  const Cplx<real>* twk = tw;
  constexpr bool kPipe = NC > 0 && twiddles_in_lds<real>();
  constexpr int PER = kPipe ? (B * NC + 255) / 256 : 1;  // elements per thread (256 threads)
  real ftab[PER];
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
    if (tabconv)
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int q = threadIdx.x + 256 * k;
        ftab[k] = q < tot ? tb[q] : (real)0;
      }
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
    // convolution with the tabulated influence function
    if constexpr (kPipe) {
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int q = threadIdx.x + 256 * k;
        if (q < tot) {
          Cplx<real> v = res[(q / N) * LP + (q % N)];
          v.x *= ftab[k];
          v.y *= ftab[k];
          res[(q / N) * LP + (q % N)] = v;
        }
      }
    } else {
      for (int q = threadIdx.x; q < tot; q += blockDim.x) {
        const real f = tb[q];
        Cplx<real> v = res[(q / N) * LP + (q % N)];
        v.x *= f;
        v.y *= f;
        res[(q / N) * LP + (q % N)] = v;
      }
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

// Influence function in z-pass order: table[(k2 * K1 + k1) * K3 + k3]
// (lines at consecutive x, element index k3), the same FP64 expression as
// the on-the-fly path.
template <class real>
__global__ void k_conv_table(int K1, int K2, int K3, double pterm, double volterm, double il0, double il1, double il2,
                             const double* __restrict__ bsmod, real* __restrict__ table) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= (long long)K1 * K2 * K3) return;
  const int k3 = (int)(e % K3);
  const long long l = e / K3;
  const int k1 = (int)(l % K1), k2 = (int)(l / K1);
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
  table[e] = (real)f;
}

template <class real>
constexpr int lines_per_block() {
  return sizeof(real) == 4 ? 16 : 8;
}

template <class real>
void roundtrip_t(Ctx& c, bool fwd_only, double* vrows, double vscale) {
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
  const int xb = (K1 + B - 1) / B;
  const Cplx<real>* twa[3] = {tw, tw + K1, tw + K1 + K2};
  if (!c.conv_table_valid) {
    c.conv_table.alloc(sizeof(real) * (size_t)K1 * K2 * K3);
    hipLaunchKernelGGL(k_conv_table<real>, dim3(ceil_div((long)K1 * K2 * K3, 256)), dim3(256), 0, c.stream, K1, K2, K3,
                       conv.pterm, conv.volterm, conv.il[0], conv.il[1], conv.il[2], conv.bsmod,
                       c.conv_table.as<real>());
    T9_LAUNCH_CHECK();
    c.conv_table_valid = true;
  }
  conv.table = c.conv_table.p;
  auto launch = [&](int axis, int sign, const ConvArgs& cv) {
    const int nblk = axis == 0 ? ceil_div((long)K2 * K3, B) : (axis == 1 ? xb * K3 : xb * K2);
    // two line buffers + the twiddles (compile-time lengths stage them in LDS)
    const size_t sh = (2 * (size_t)B * (p[axis].n + 1) + (twiddles_in_lds<real>() && ct_len(p[axis].n) ? p[axis].n : 0)) *
                      sizeof(Cplx<real>);
    void (*kern)(Cplx<real>*, int, int, int, int, int, FftPlan, const Cplx<real>*, ConvArgs, Gate);
    switch (p[axis].n) {
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
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(256), sh, c.stream, grid, K1, K2, K3, axis, sign, p[axis], twa[axis],
                       cv, c.gate);
    T9_LAUNCH_CHECK();
  };
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

void pme_fft_roundtrip(Ctx& c) {
  if (c.prec == 32) roundtrip_t<float>(c, false, nullptr, 0.0);
  else roundtrip_t<double>(c, false, nullptr, 0.0);
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
