// Device-side building blocks shared by the gfx950 kernels: precision
// traits, minimum image, Thole/Ewald radial functions, deterministic block
// reductions, and the launch-error check.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "host/errors.hpp"

namespace t9 {

#define T9_HIP(call)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      throw ::t9::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " at " __FILE__ \
                           ":" + std::to_string(__LINE__));                                     \
  } while (0)

#define T9_LAUNCH_CHECK() T9_HIP(hipGetLastError())

constexpr int kWave = 64;
constexpr int kCodeBits = 5;  // neighbor entry = (j << kCodeBits) | scale class
constexpr int kMaxCodes = 1 << kCodeBits;

template <class real>
struct RealTraits;
template <>
struct RealTraits<float> {
  using real3 = float3;
  using real4 = float4;
  static constexpr double fixed_scale = 4294967296.0;  // 2^32 for PME grid accumulation
};
template <>
struct RealTraits<double> {
  using real3 = double3;
  using real4 = double4;
  static constexpr double fixed_scale = 17592186044416.0;  // 2^44
};

// The dispatcher hands workgroup w to XCD w mod 8, and each of the 8 XCDs has
// its own L2.  xcd_block() renumbers the workgroups of a launch so that XCD x
// runs one contiguous range of logical blocks: neighbouring atoms (Morton
// order) or grid tiles then share an L2 instead of being spread over all
// eight.  It is a bijection on [0, gridDim.x); a kernel that uses it for every
// block-indexed access computes the same values as with blockIdx.x.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, nb = gridDim.x;
  const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds, k = b / kXcds;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// In-cutoff compaction: the first rows of each wavefront's output slabs are
// assembled in LDS (in-cutoff list, preconditioner list) and written whole.
constexpr int kCompactRowsA = 176, kCompactRowsB = 64;

__device__ __forceinline__ float t9_rint(float x) { return rintf(x); }
__device__ __forceinline__ double t9_rint(double x) { return rint(x); }
__device__ __forceinline__ float t9_exp(float x) { return expf(x); }
__device__ __forceinline__ double t9_exp(double x) { return exp(x); }
__device__ __forceinline__ float t9_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double t9_sqrt(double x) { return sqrt(x); }
// 1/sqrt(x) of a pair distance squared: the hardware reciprocal square root
// (1 ulp) without the library's denormal-input scaling, which a squared
// distance of two atoms never needs.
__device__ __forceinline__ float t9_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double t9_rsqrt(double x) { return rsqrt(x); }

// erfc: the device library's erfcf / erfc.  The reference's FP32 build uses
// the 5-term A&S 7.1.26 form (include/math/libfunc.h:8-29), whose relative
// error reaches 2e-3 near the 7 Å cutoff (x ~ 3.8); that bias is systematic
// over millions of pairs, so the accurate routine is used instead.
__device__ __forceinline__ float t9_erfc_ex(float x, float /*expmx2*/) { return erfcf(x); }
__device__ __forceinline__ double t9_erfc_ex(double x, double /*expmx2*/) { return erfc(x); }

// exp(x) for the Thole damping factors, x <= 0: FP32 uses the hardware
// base-2 exponential of x log2(e) directly.  Its relative error grows as
// |x| 6e-8, so the damping terms 1 - e, 1 - e (1 - x), ... carry an absolute
// error below max |x| e^x 6e-8 = 2.2e-8; results below 2^-126 flush to 0.
__device__ __forceinline__ float t9_exp_damp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ double t9_exp_damp(double x) { return exp(x); }

// erfc(z) together with exp(-z^2) for z >= 0, the pair needed by the Ewald
// real-space recursion.  FP32 (one hardware exponential):
//  * exp(-z^2): z^2 = hi + lo exactly (FMA), hi log2(e) = y + ylo in two
//    floats, exp(-z^2) = 2^-y (1 - ylo ln 2 - lo); relative error 1.2e-7;
//  * erfc(z) = exp(-z^2) t Q(s), t = 1 / (1 + z/2), s = (8t - 5) / 3 in
//    [-1, 1] for 0 <= z <= 6, Q a degree-9 polynomial for erfc(z) exp(z^2)/t
//    (smooth in t down to t = 0): least-squares fit on Chebyshev nodes
//    reweighted toward the largest relative error, then its FP32
//    coefficients moved by a few ulps to cancel the mean error of the FP32
//    evaluation over the pair distances of a 7 Å Ewald cutoff.  Error of the
//    FP32 evaluation: 3.6e-7 at most, 9.3e-8 RMS over 0 <= z <= 6, mean
//    (weighted by erfc(z) z, the energy of a shell of pairs) below 1e-8 over
//    0.4 <= z <= 3.9; beyond z = 6 (erfc < 2e-17) the relative error grows.
// FP64: library routines.
__device__ __forceinline__ float t9_erfc_pair(float z, float& expmz2) {
  const float hi = z * z;
  const float lo = fmaf(z, z, -hi);
  const float y = hi * 1.44269504088896341f;
  float ylo = fmaf(hi, 1.44269504088896341f, -y);
  ylo = fmaf(hi, 1.92596299e-8f, ylo);  // log2(e) - (float)log2(e)
  const float e0 = __builtin_amdgcn_exp2f(-y);
  expmz2 = fmaf(e0, fmaf(-ylo, 0.693147181f, -lo), e0);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, z, 1.0f));
  const float q = fmaf(2.666666746e+00f, t, -1.666666627e+00f);
  float p = 4.076925961e-06f;
  p = fmaf(p, q, -2.033701276e-05f);
  p = fmaf(p, q, -2.042249071e-05f);
  p = fmaf(p, q, 3.052748507e-04f);
  p = fmaf(p, q, -1.570393651e-04f);
  p = fmaf(p, q, -4.181238357e-03f);
  p = fmaf(p, q, 2.516048728e-03f);
  p = fmaf(p, q, 8.378951252e-02f);
  p = fmaf(p, q, 3.121041656e-01f);
  p = fmaf(p, q, 6.056598425e-01f);
  return t * expmz2 * p;
}
__device__ __forceinline__ double t9_erfc_pair(double z, double& expmz2) {
  expmz2 = exp(-z * z);
  return erfc(z);
}

// sqrt(x), 1/x and a/d corrected by one residual step on the hardware
// estimate (FP32; Markstein's forms): the correction leaves an error within
// about half an ulp either way, so the rounding is unbiased, at a quarter of
// the cost of the IEEE sequences.  FP64: the exact operations.
__device__ __forceinline__ float t9_sqrt_nr(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  const float s = x * y;
  return fmaf(0.5f * y, fmaf(-s, s, x), s);
}
__device__ __forceinline__ double t9_sqrt_nr(double x) { return sqrt(x); }
__device__ __forceinline__ float t9_rcp_nr(float d) {
  const float y = __builtin_amdgcn_rcpf(d);
  return fmaf(y, fmaf(-d, y, 1.0f), y);
}
__device__ __forceinline__ double t9_rcp_nr(double d) { return 1.0 / d; }
// a / d from the refined reciprocal y of d: q = a y, corrected by the
// residual a - d q (Markstein)
__device__ __forceinline__ float t9_div_nr(float a, float d, float y) {
  const float q = a * y;
  return fmaf(fmaf(-d, q, a), y, q);
}
__device__ __forceinline__ double t9_div_nr(double a, double d, double) { return a / d; }

// Minimum image in an orthorhombic box.
template <class real>
__device__ __forceinline__ void min_image(real& x, real& y, real& z, real lx, real ly, real lz, real ilx, real ily,
                                          real ilz) {
  x -= lx * t9_rint(x * ilx);
  y -= ly * t9_rint(y * ily);
  z -= lz * t9_rint(z * ilz);
}

// Box description passed by value to kernels.  q + qlo = l / 2^32: the
// length of one unit of the FP32 path's fixed-point coordinates, split so
// that the product with a displacement carries no systematic scale error.
template <class real>
struct BoxT {
  real l[3];
  real il[3];
  real q[3], qlo[3];
  int periodic;
};

// Sorted-order position record of the pair kernels.
//  * FP32: fixed-point fractional coordinates, (x - origin) / l mod 1 in
//    units of 2^-32.  A displacement is the exact integer difference; its
//    32-bit wrap-around is the minimum image (orthorhombic box; a virtual box
//    several times the system's extent for non-periodic systems), and only
//    the final conversion to float rounds.
//  * FP64: wrapped Cartesian coordinates, minimum image by rint.
// w = 1 / Thole radius (0 for sites without damping).
template <class real>
struct PosRec;
template <>
struct alignas(16) PosRec<float> {
  int x, y, z;
  float w;
};
template <>
struct alignas(32) PosRec<double> {
  double x, y, z, w;
};

// (x, y, z) = minimum-image displacement b - a.
__device__ __forceinline__ void pair_disp(const PosRec<float>& a, const PosRec<float>& b, const BoxT<float>& box,
                                          float& x, float& y, float& z) {
  const float dx = (float)(int)((unsigned)b.x - (unsigned)a.x);
  const float dy = (float)(int)((unsigned)b.y - (unsigned)a.y);
  const float dz = (float)(int)((unsigned)b.z - (unsigned)a.z);
  x = fmaf(dx, box.q[0], dx * box.qlo[0]);
  y = fmaf(dy, box.q[1], dy * box.qlo[1]);
  z = fmaf(dz, box.q[2], dz * box.qlo[2]);
}
__device__ __forceinline__ void pair_disp(const PosRec<double>& a, const PosRec<double>& b, const BoxT<double>& box,
                                          double& x, double& y, double& z) {
  x = b.x - a.x;
  y = b.y - a.y;
  z = b.z - a.z;
  if (box.periodic) min_image(x, y, z, box.l[0], box.l[1], box.l[2], box.il[0], box.il[1], box.il[2]);
}
// The same displacement b - a in FP64: exact for the fixed-point form (the
// integer difference times the quantum hi + lo), the plain one otherwise.
__device__ __forceinline__ void pair_disp64(const PosRec<float>& a, const PosRec<float>& b, const BoxT<float>& box,
                                            double* R) {
  R[0] = (double)(int)((unsigned)b.x - (unsigned)a.x) * ((double)box.q[0] + (double)box.qlo[0]);
  R[1] = (double)(int)((unsigned)b.y - (unsigned)a.y) * ((double)box.q[1] + (double)box.qlo[1]);
  R[2] = (double)(int)((unsigned)b.z - (unsigned)a.z) * ((double)box.q[2] + (double)box.qlo[2]);
}
__device__ __forceinline__ void pair_disp64(const PosRec<double>& a, const PosRec<double>& b, const BoxT<double>& box,
                                            double* R) {
  pair_disp(a, b, box, R[0], R[1], R[2]);
}

// Position record of a site from FP64 coordinates; (org, len) is the box
// (periodic) or the virtual box (non-periodic).
template <class real>
__device__ __forceinline__ PosRec<real> make_pos(double x, double y, double z, double w, const double* org,
                                                 const double* len, int periodic);
template <>
__device__ __forceinline__ PosRec<float> make_pos<float>(double x, double y, double z, double w, const double* org,
                                                         const double* len, int /*periodic*/) {
  const double v[3] = {x, y, z};
  int o[3];
  for (int d = 0; d < 3; ++d) {
    double f = (v[d] - org[d]) / len[d];
    f -= floor(f);
    o[d] = (int)(unsigned)(unsigned long long)(f * 4294967296.0 + 0.5);
  }
  PosRec<float> p;
  p.x = o[0];
  p.y = o[1];
  p.z = o[2];
  p.w = (float)w;
  return p;
}
template <>
__device__ __forceinline__ PosRec<double> make_pos<double>(double x, double y, double z, double w, const double* org,
                                                           const double* len, int periodic) {
  const double v[3] = {x, y, z};
  double o[3];
  for (int d = 0; d < 3; ++d) {
    double f = v[d];
    if (periodic) {
      f = v[d] - len[d] * floor(v[d] / len[d]);
      if (f >= len[d]) f -= len[d];
    }
    o[d] = f;
  }
  PosRec<double> p;
  p.x = o[0];
  p.y = o[1];
  p.z = o[2];
  p.w = w;
  return p;
}

// Per scale-class factors (m, d, p, u, and the HIPPO dispersion and
// repulsion scales ds, rs); class 0 is the all-ones class.
struct ScaleTable {
  double m[kMaxCodes], d[kMaxCodes], p[kMaxCodes], u[kMaxCodes], ds[kMaxCodes], rs[kMaxCodes];
};

// Thole damping of the 1/r^3 and 1/r^5 (and 1/r^7) tensors:
// lambda3 = 1 - e^{-a u^3}, lambda5 = 1 - (1 + a u^3) e^{-a u^3},
// lambda7 = 1 - (1 + a u^3 + 0.6 a^2 u^6) e^{-a u^3}, u = r / (pdi pdk)^{...}
// (include/seq/damp.h:8-60).
template <class real>
__device__ __forceinline__ void thole_damp(real r, real pdik, real pga, real& s3, real& s5, real& s7) {
  if (pdik == 0) {
    s3 = s5 = s7 = 1;
    return;
  }
  real ratio = r / pdik;
  real damp = -pga * ratio * ratio * ratio;
  real e = t9_exp(damp);
  s3 = 1 - e;
  s5 = 1 - e * (1 - damp);
  s7 = 1 - e * (1 - damp + (real)0.6 * damp * damp);
}

// Ewald real-space radial functions b0..b{NB-1} (include/seq/damp.h:137-151).
template <int NB, class real>
__device__ __forceinline__ void ewald_bn(real* bn, real r, real rinv, real rr2, real aewald) {
  real ra = aewald * r;
  real ex = t9_exp(-ra * ra);
  bn[0] = t9_erfc_ex(ra, ex) * rinv;
  real alsq2 = 2 * aewald * aewald;
  real alsq2n = (real)0.56418958354775628695 / aewald;  // 1/(sqrt(pi) a)
#pragma unroll
  for (int j = 1; j < NB; ++j) {
    alsq2n *= alsq2;
    bn[j] = ((2 * j - 1) * bn[j - 1] + alsq2n * ex) * rr2;
  }
}

// Pair kernels run kPairParts wavefronts per 64 atoms: wave p of a
// workgroup walks list entries p, p + kPairParts, ... of each of the 64
// atoms, and the parts are summed in part order through LDS (parts_reduce),
// so results do not depend on scheduling.  Eight waves per 64 atoms keep
// ~12 waves per SIMD in flight on 96k atoms instead of ~1.5.
constexpr int kPairParts = 8;

// Early exit for the kernels of one CG iteration.  The solver keeps the
// iteration at which it converged in CG scalar slot kDoneSlot (0 while it
// runs); kernels enqueued for a later iteration return at once, so the host
// can enqueue several iterations without reading the residual back.
constexpr int kDoneSlot = 15;
constexpr int kTraceCap = 256;  // per-iteration residual trace of the last CG solve
// a . b over the 3 components of element s, in FP64 (CG dot products)
template <class real>
__device__ __forceinline__ double dot3(const real* __restrict__ a, const real* __restrict__ b, int s) {
  double v = 0;
  for (int d = 0; d < 3; ++d) v += (double)a[3 * s + d] * (double)b[3 * s + d];
  return v;
}
// sum over the 64 lanes of a wavefront by a fixed butterfly (every lane gets it)
__device__ __forceinline__ double wave_bfly_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
constexpr int kDispSlot = 44;   // pinned readback slot of the list check's flag (40: kinetic energy; 45-47: list totals)
struct Gate {
  const double* sc = nullptr;  // CG scalars; nullptr: no gate
  int iter = 0;                // iteration the launch belongs to
};
__device__ __forceinline__ bool gate_closed(const Gate& g) {
  if (!g.sc) return false;
  const double d = g.sc[kDoneSlot];
  return d != 0 && d < (double)g.iter;
}

// v[0..NV) of wave 0 = the sum of the NP waves' values (NP a power of 2), by
// a fixed tree: wave p += wave p + h for h = NP/2, NP/4, ..., 1.  sh:
// parts_lds<NP, NV>() doubles (half the waves' values at a time, so a
// workgroup of pair kernels keeps its LDS small enough to share a CU with an
// FFT plane workgroup).  Every thread of the workgroup calls it.
template <int NP, int NV>
constexpr int parts_lds() {
  return NP / 2 * NV * 64;
}
template <int NP, int NV>
__device__ __forceinline__ void parts_reduce(double* v, double* sh) {
  static_assert((NP & (NP - 1)) == 0, "parts_reduce: NP must be a power of 2");
  const int lane = threadIdx.x & 63, part = threadIdx.x >> 6;
#pragma unroll
  for (int h = NP / 2; h >= 1; h >>= 1) {
    if (part >= h && part < 2 * h)
#pragma unroll
      for (int q = 0; q < NV; ++q) sh[((part - h) * NV + q) * 64 + lane] = v[q];
    __syncthreads();
    if (part < h)
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] += sh[(part * NV + q) * 64 + lane];
    __syncthreads();
  }
}

// Deterministic block sum of doubles: fixed tree over a fixed block size.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// Deterministic sum of NV per-thread doubles over a 256-thread block (wave
// butterflies, then the four waves in a fixed order); thread 0 writes
// row[0..NV).  Every thread of the block must call it.
template <int NV>
__device__ __forceinline__ void block_row(double* v, double* row) {
  __shared__ double sh[4][NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = v[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[q] = x;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[w][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) row[q] = (sh[0][q] + sh[1][q]) + (sh[2][q] + sh[3][q]);
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace t9
