// Learned initial guess for the induced dipoles (gfx950, BF16 MFMA).
//
// No counterpart in the reference: it replaces the predictor at the plug
// point of src/amoeba/induce.cpp:29-69 / src/cu/upredict.cu:34-201 (ASPC),
// which stays as the fallback.  The guess is
//     u0 = u_ASPC + R_i^T NN(x_i)          (d and p dipoles)
// where R_i is the atom's multipole local frame and x_i 32 features in that
// frame: the extrapolation error of the direct field (current field minus the
// ASPC combination of the previous 16 fields) at the atom and at its two
// frame atoms, the current field, the last dipole and its change, the
// polarizability and a type one-hot.  Everything rotates with the frame, so
// the guess is rotation-equivariant.
//
// Network: 32 -> 128 (SiLU) -> 128 (SiLU) -> 6, the last layer also seeing
// the inputs (linear skip).  Inputs, weights and activations are BF16, every
// product accumulates in FP32 on v_mfma_f32_32x32x16_bf16; one wavefront
// owns 32 rows and stages its activations in LDS.  Training: the backward
// pass in the same arithmetic, per-workgroup partial weight gradients
// reduced in a fixed order (bit-reproducible), Adam on FP32 master weights.
// Training data: during MD, every stride-th atom's raw features and its
// converged correction R_i (u - u_ASPC) are appended to a device buffer.
#include <algorithm>
#include <cmath>
#include <random>

#include "ctx.hpp"

namespace t9 {
namespace {

constexpr int NF = 32, NH = 128, NO = 6, NSK = NH + NF;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NF * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NSK * NO, NPARAM = OFF_B3 + NO;
constexpr int kNorm = 2 * NF + 2 * NO;  // feature mean, 1/std; target mean, std

typedef unsigned short u16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ u16 f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}

// SiLU and its derivative from the hardware base-2 exponential and
// reciprocal (relative error near 1e-7, far below the BF16 rounding that
// follows); z below -87 gives exp2 = inf, sigmoid 0.
__device__ __forceinline__ float sigmoid_hw(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * 1.44269504088896341f));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_hw(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoid_hw(z);
  return s * (1.0f + z * (1.0f - s));
}

// One wavefront: acc += A(32 x K) B(K x 32) with A(m, k) = A[m*ars + k*acs],
// B(k, n) = B[k*brs + n*bcs]; entries with k >= kv or n >= nv read as zero.
// K is a multiple of 16.  Operand layout of v_mfma_f32_32x32x16_bf16: lane l
// holds A(l % 32, 8 (l / 32) + i) and B(8 (l / 32) + i, l % 32), i = 0..7.
// Result: lane l, register j holds C(8 (j / 4) + 4 (l / 32) + j % 4, l % 32).
__device__ __forceinline__ f32x16 mm32(const u16* A, int ars, int acs, const u16* B, int brs, int bcs, int K, int kv,
                                       int nv, f32x16 acc) {
  const int l = threadIdx.x & 63, r = l & 31, kh = (l >> 5) << 3;
  for (int k0 = 0; k0 < K; k0 += 16) {
    s16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = k0 + kh + i;
      a[i] = k < kv ? (short)A[r * ars + k * acs] : (short)0;
      b[i] = (k < kv && r < nv) ? (short)B[k * brs + r * bcs] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                  0, 0, 0);
  }
  return acc;
}

__device__ __forceinline__ int mrow(int j) { return 8 * (j >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (j & 3); }
__device__ __forceinline__ int ncol() { return threadIdx.x & 31; }

struct Net {
  const u16* w1;  // [32][128] bf16
  const float* b1;
  const u16* w2;  // [128][128]
  const float* b2;
  const u16* w3;  // [160][6]
  const float* b3;
};

__device__ __forceinline__ Net net_of(const u16* pb, const float* p) {
  Net t;
  t.w1 = pb + OFF_W1;
  t.b1 = p + OFF_B1;
  t.w2 = pb + OFF_W2;
  t.b2 = p + OFF_B2;
  t.w3 = pb + OFF_W3;
  t.b3 = p + OFF_B3;
  return t;
}

// Forward pass of the wave's 32 rows from sX (bf16 [32][32]); activations to
// sH1 / sH2 (bf16 [32][128]), pre-activations to sZ1 / sZ2 when given.
// Returns the output accumulator (columns 0..5 valid, bias added).
__device__ f32x16 wave_forward(const Net& w, const u16* sX, u16* sH1, u16* sH2, float* sZ1, float* sZ2) {
  const f32x16 zero = {0};
  for (int t = 0; t < 4; ++t) {
    const f32x16 acc = mm32(sX, NF, 1, w.w1 + 32 * t, NH, 1, NF, NF, 32, zero);
    const int n = 32 * t + ncol();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float z = acc[j] + w.b1[n];
      if (sZ1) sZ1[mrow(j) * NH + n] = z;
      sH1[mrow(j) * NH + n] = f2bf(silu(z));
    }
  }
  __syncthreads();
  for (int t = 0; t < 4; ++t) {
    const f32x16 acc = mm32(sH1, NH, 1, w.w2 + 32 * t, NH, 1, NH, NH, 32, zero);
    const int n = 32 * t + ncol();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float z = acc[j] + w.b2[n];
      if (sZ2) sZ2[mrow(j) * NH + n] = z;
      sH2[mrow(j) * NH + n] = f2bf(silu(z));
    }
  }
  __syncthreads();
  f32x16 acc = mm32(sH2, NH, 1, w.w3, NO, 1, NH, NH, NO, zero);
  acc = mm32(sX, NF, 1, w.w3 + NH * NO, NO, 1, NF, NF, NO, acc);
  const float b = ncol() < NO ? w.b3[ncol()] : 0.0f;
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] += b;
  return acc;
}

// ---- inference: weights in MFMA operand order, layers computed transposed
//
// Each layer is computed as out^T (32 units x 32 samples) = W^T (32 x K)
// act^T (K x 32 samples), so that samples run along the MFMA columns.  In
// that form the output registers of one layer are the B operand of the next
// once the hidden units are renumbered: output tile t, lane half h = l / 32,
// register j holds unit p = 32 t + 8 (j / 4) + 4 h + j % 4, and is used as
// B row k = 8 h + (j % 8) of k-step 2 t + j / 8 of the next layer.  The
// weights are packed once per parameter update as A-operand fragments
// (fragment f, lane l: 8 bf16), 50 fragments of 1 KB:
//   f  0.. 7: W1^T, tile t = f / 2, k-step f % 2 (features in natural order)
//   f  8..39: W2^T, tile (f - 8) / 8, k-step (f - 8) % 8 (renumbered units)
//   f 40..47: W3^T over the hidden units (renumbered), rows >= 6 zero
//   f 48..49: W3^T over the skip features, rows >= 6 zero
constexpr int kFrags = 50;

__host__ __device__ __forceinline__ int hidden_of(int ks, int h, int i) {  // k-step, lane half, element -> unit
  const int t = ks >> 1, j = 8 * (ks & 1) + i;
  return 32 * t + 8 * (j >> 2) + 4 * h + (j & 3);
}

__global__ void k_pack_weights(const u16* __restrict__ pb, u16* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kFrags * 64 * 8) return;
  const int f = e >> 9, l = (e >> 3) & 63, i = e & 7, h = l >> 5, r = l & 31;
  u16 v = 0;
  if (f < 8) {
    const int t = f >> 1, k = 16 * (f & 1) + 8 * h + i;
    v = pb[OFF_W1 + k * NH + 32 * t + r];
  } else if (f < 40) {
    const int t2 = (f - 8) >> 3, ks = (f - 8) & 7;
    v = pb[OFF_W2 + hidden_of(ks, h, i) * NH + 32 * t2 + r];
  } else if (f < 48) {
    if (r < NO) v = pb[OFF_W3 + hidden_of(f - 40, h, i) * NO + r];
  } else {
    if (r < NO) v = pb[OFF_W3 + (NH + 16 * (f - 48) + 8 * h + i) * NO + r];
  }
  out[e] = v;
}

__device__ __forceinline__ f32x16 mfma16(const uint4 a, const bf16x8 b, f32x16 acc) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), b, acc, 0, 0, 0);
}

// One wavefront, 32 samples: bX = the two B fragments of the normalized
// features (k-steps 0, 1).  Returns the output tile: lanes 0..31 hold
// outputs 0..3 of sample l in registers 0..3, lanes 32..63 outputs 4, 5 of
// sample l - 32 in registers 0, 1 (bias added).
__device__ __forceinline__ f32x16 tile_forward(const uint4* __restrict__ sW, const float* __restrict__ sb,
                                               const bf16x8 bX[2]) {
  const int l = threadIdx.x & 63, h = l >> 5;
  const f32x16 zero = {0};
  bf16x8 hB[8], h2B[8];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x16 acc = mfma16(sW[(2 * t) * 64 + l], bX[0], zero);
    acc = mfma16(sW[(2 * t + 1) * 64 + l], bX[1], acc);
    s16x8 lo, hi;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float z = acc[j] + sb[32 * t + 8 * (j >> 2) + 4 * h + (j & 3)];
      const short v = (short)f2bf(silu(z));
      if (j < 8) lo[j] = v;
      else hi[j - 8] = v;
    }
    hB[2 * t] = __builtin_bit_cast(bf16x8, lo);
    hB[2 * t + 1] = __builtin_bit_cast(bf16x8, hi);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x16 acc = zero;
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) acc = mfma16(sW[(8 + 8 * t + ks) * 64 + l], hB[ks], acc);
    s16x8 lo, hi;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float z = acc[j] + sb[NH + 32 * t + 8 * (j >> 2) + 4 * h + (j & 3)];
      const short v = (short)f2bf(silu(z));
      if (j < 8) lo[j] = v;
      else hi[j - 8] = v;
    }
    h2B[2 * t] = __builtin_bit_cast(bf16x8, lo);
    h2B[2 * t + 1] = __builtin_bit_cast(bf16x8, hi);
  }
  f32x16 acc = zero;
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) acc = mfma16(sW[(40 + ks) * 64 + l], h2B[ks], acc);
  acc = mfma16(sW[48 * 64 + l], bX[0], acc);
  acc = mfma16(sW[49 * 64 + l], bX[1], acc);
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] += sb[2 * NH + 4 * h + j];
  return acc;
}

// Stage the packed weights, the biases (and the normalization) in LDS.
__device__ __forceinline__ void stage_net(const uint4* __restrict__ packed, const float* __restrict__ p,
                                          const float* __restrict__ norm, uint4* sW, float* sb, float* sn) {
  for (int e = threadIdx.x; e < kFrags * 64; e += blockDim.x) sW[e] = packed[e];
  for (int e = threadIdx.x; e < 2 * NH + 8; e += blockDim.x)
    sb[e] = e < NH ? p[OFF_B1 + e] : (e < 2 * NH ? p[OFF_B2 + e - NH] : (e - 2 * NH < NO ? p[OFF_B3 + e - 2 * NH] : 0.f));
  if (sn)
    for (int e = threadIdx.x; e < kNorm; e += blockDim.x) sn[e] = norm[e];
  __syncthreads();
}

// Y[m][8] (columns 0..5) = network(X[m][32]), X normalized bf16.  Four
// wavefronts per workgroup, each looping over 32-row tiles.
__global__ __launch_bounds__(256) void k_mlp_infer_y(int m, const u16* __restrict__ X, const uint4* __restrict__ packed,
                                                     const float* __restrict__ p, float* __restrict__ Y) {
  __shared__ uint4 sW[kFrags * 64];
  __shared__ float sb[2 * NH + 8];
  stage_net(packed, p, nullptr, sW, sb, nullptr);
  const int l = threadIdx.x & 63, h = l >> 5;
  const int ntile = (m + 31) / 32;
  for (int tile = blockIdx.x * 4 + (threadIdx.x >> 6); tile < ntile; tile += gridDim.x * 4) {
    // weights are re-read from LDS per tile rather than held in registers
    asm volatile("" ::: "memory");
    const int s = 32 * tile + (l & 31);
    bf16x8 bX[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      uint4 v = {0, 0, 0, 0};
      if (s < m) v = *reinterpret_cast<const uint4*>(X + (size_t)s * NF + 16 * ks + 8 * h);
      bX[ks] = __builtin_bit_cast(bf16x8, v);
    }
    const f32x16 acc = tile_forward(sW, sb, bX);
    if (s < m) {
      if (h == 0)
        for (int j = 0; j < 4; ++j) Y[(size_t)s * 8 + j] = acc[j];
      else
        for (int j = 0; j < 2; ++j) Y[(size_t)s * 8 + 4 + j] = acc[j];
    }
  }
}

// The learned correction applied in place: normalize the raw features F
// [n][32], run the network, u += R^T (y * std + mean) for d (outputs 0..2)
// and p (3..5) with R the atom's frame rotation.
template <class real>
__global__ __launch_bounds__(256) void k_mlp_apply(int n, const float* __restrict__ F, const uint4* __restrict__ packed,
                                                   const float* __restrict__ p, const float* __restrict__ norm,
                                                   const float* __restrict__ frot, real* __restrict__ ud,
                                                   real* __restrict__ up) {
  __shared__ uint4 sW[kFrags * 64];
  __shared__ float sb[2 * NH + 8], sn[kNorm];
  stage_net(packed, p, norm, sW, sb, sn);
  const int l = threadIdx.x & 63, h = l >> 5;
  const int ntile = (n + 31) / 32;
  for (int tile = blockIdx.x * 4 + (threadIdx.x >> 6); tile < ntile; tile += gridDim.x * 4) {
    // weights are re-read from LDS per tile rather than held in registers
    asm volatile("" ::: "memory");
    const int s = 32 * tile + (l & 31);
    bf16x8 bX[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      s16x8 v;
      const int k0 = 16 * ks + 8 * h;
      if (s < n) {
        const float4 f0 = *reinterpret_cast<const float4*>(F + (size_t)s * NF + k0);
        const float4 f1 = *reinterpret_cast<const float4*>(F + (size_t)s * NF + k0 + 4);
        const float f[8] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w};
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (short)f2bf((f[i] - sn[k0 + i]) * sn[NF + k0 + i]);
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = 0;
      }
      bX[ks] = __builtin_bit_cast(bf16x8, v);
    }
    const f32x16 acc = tile_forward(sW, sb, bX);
    // lanes 0..31: outputs 0..3; lanes 32..63: outputs 4, 5 (same sample)
    const float x3 = __shfl_xor(acc[3], 32);
    if (s < n) {
      float y0, y1, y2;
      if (h == 0) {
        y0 = acc[0] * sn[2 * NF + NO + 0] + sn[2 * NF + 0];
        y1 = acc[1] * sn[2 * NF + NO + 1] + sn[2 * NF + 1];
        y2 = acc[2] * sn[2 * NF + NO + 2] + sn[2 * NF + 2];
      } else {
        y0 = x3 * sn[2 * NF + NO + 3] + sn[2 * NF + 3];
        y1 = acc[0] * sn[2 * NF + NO + 4] + sn[2 * NF + 4];
        y2 = acc[1] * sn[2 * NF + NO + 5] + sn[2 * NF + 5];
      }
      const float* R = frot + 9 * (size_t)s;
      real* u = h == 0 ? ud : up;
#pragma unroll
      for (int g = 0; g < 3; ++g) u[3 * s + g] += (real)(R[g] * y0 + R[3 + g] * y1 + R[6 + g] * y2);
    }
  }
}

// Per-workgroup partial gradient of the mean squared error over `mtot`
// rows (this workgroup: rows row0 .. row0+31 of X, T) and the partial sum of
// squared errors, written to part[blockIdx.x][NPARAM + 1].
__global__ __launch_bounds__(64) void k_mlp_bwd(int m, int mtot, const u16* __restrict__ X,
                                                const float* __restrict__ T, const u16* __restrict__ pb,
                                                const float* __restrict__ p, float* __restrict__ part) {
  __shared__ u16 sX[32 * NF], sH1[32 * NH], sH2[32 * NH], sdY[32 * 32];
  __shared__ float sZ1[32 * NH], sZ2[32 * NH], sred[64];
  const int row0 = blockIdx.x * 32;
  const int l = threadIdx.x;
  float* gp = part + (size_t)blockIdx.x * (NPARAM + 1);
  for (int e = l; e < 32 * NF; e += 64) {
    const int rr = e / NF;
    sX[e] = row0 + rr < m ? X[(size_t)(row0 + rr) * NF + (e % NF)] : (u16)0;
  }
  __syncthreads();
  const Net w = net_of(pb, p);
  const f32x16 y = wave_forward(w, sX, sH1, sH2, sZ1, sZ2);
  const f32x16 zero = {0};
  const int n = ncol();
  // dY = 2 (Y - T) / (mtot * 6); loss partial
  const float sc = 2.0f / ((float)mtot * NO);
  float lsum = 0, bsum = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int r = row0 + mrow(j);
    float d = 0;
    if (n < NO && r < m) {
      const float e = y[j] - T[(size_t)r * NO + n];
      lsum += e * e;
      d = sc * e;
    }
    bsum += d;
    sdY[mrow(j) * 32 + n] = f2bf(d);
  }
  // bias 3 and loss: column sums over the two lane halves
  sred[l] = bsum;
  __syncthreads();
  if (l < NO) gp[OFF_B3 + l] = sred[l] + sred[l + 32];
  __syncthreads();
  sred[l] = lsum;
  __syncthreads();
  if (l == 0) {
    float t = 0;
    for (int q = 0; q < 64; ++q) t += sred[q];
    gp[NPARAM] = t;
  }
  // dW3 = [H2 X]^T dY  (160 x 6)
  for (int t = 0; t < 5; ++t) {
    const f32x16 acc = t < 4 ? mm32(sH2 + 32 * t, 1, NH, sdY, 32, 1, 32, 32, 32, zero)
                             : mm32(sX, 1, NF, sdY, 32, 1, 32, 32, 32, zero);
    if (n < NO)
#pragma unroll
      for (int j = 0; j < 16; ++j) gp[OFF_W3 + (32 * t + mrow(j)) * NO + n] = acc[j];
  }
  __syncthreads();
  // dZ2 = (dY W3[:128]^T) * silu'(Z2) -> sH2 (bf16); bias 2
  for (int t = 0; t < 4; ++t) {
    const f32x16 acc = mm32(sdY, 32, 1, w.w3 + 32 * t * NO, 1, NO, 32, NO, 32, zero);
    const int h = 32 * t + n;
    float bs = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float d = acc[j] * dsilu(sZ2[mrow(j) * NH + h]);
      bs += d;
      sZ2[mrow(j) * NH + h] = d;
    }
    sred[l] = bs;
    __syncthreads();
    if (l < 32) gp[OFF_B2 + 32 * t + l] = sred[l] + sred[l + 32];
    __syncthreads();
  }
  for (int e = l; e < 32 * NH; e += 64) sH2[e] = f2bf(sZ2[e]);
  __syncthreads();
  // dW2 = H1^T dZ2  (128 x 128)
  for (int tm = 0; tm < 4; ++tm)
    for (int tn = 0; tn < 4; ++tn) {
      const f32x16 acc = mm32(sH1 + 32 * tm, 1, NH, sH2 + 32 * tn, NH, 1, 32, 32, 32, zero);
#pragma unroll
      for (int j = 0; j < 16; ++j) gp[OFF_W2 + (32 * tm + mrow(j)) * NH + 32 * tn + n] = acc[j];
    }
  __syncthreads();
  // dZ1 = (dZ2 W2^T) * silu'(Z1) -> sH1 (bf16); bias 1
  for (int t = 0; t < 4; ++t) {
    const f32x16 acc = mm32(sH2, NH, 1, w.w2 + 32 * t * NH, 1, NH, NH, NH, 32, zero);
    const int h = 32 * t + n;
    float bs = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float d = acc[j] * dsilu(sZ1[mrow(j) * NH + h]);
      bs += d;
      sZ1[mrow(j) * NH + h] = d;
    }
    sred[l] = bs;
    __syncthreads();
    if (l < 32) gp[OFF_B1 + 32 * t + l] = sred[l] + sred[l + 32];
    __syncthreads();
  }
  for (int e = l; e < 32 * NH; e += 64) sH1[e] = f2bf(sZ1[e]);
  __syncthreads();
  // dW1 = X^T dZ1  (32 x 128)
  for (int tn = 0; tn < 4; ++tn) {
    const f32x16 acc = mm32(sX, 1, NF, sH1 + 32 * tn, NH, 1, 32, 32, 32, zero);
#pragma unroll
    for (int j = 0; j < 16; ++j) gp[OFF_W1 + mrow(j) * NH + 32 * tn + n] = acc[j];
  }
}

// g[q] = sum over workgroups of part[w][q], in workgroup order (q = NPARAM: loss)
__global__ void k_grad_reduce(int nwg, const float* __restrict__ part, float* __restrict__ g) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > NPARAM) return;
  float s = 0;
  for (int w = 0; w < nwg; ++w) s += part[(size_t)w * (NPARAM + 1) + q];
  g[q] = s;
}

// Adam (bias-corrected), FP32 master parameters; refreshes the bf16 copy.
__global__ void k_adam(const float* __restrict__ g, float* __restrict__ p, float* __restrict__ mo,
                       float* __restrict__ ve, u16* __restrict__ pb, float lr, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NPARAM) return;
  const float gq = g[q];
  const float m = 0.9f * mo[q] + 0.1f * gq;
  const float v = 0.999f * ve[q] + 0.001f * gq * gq;
  mo[q] = m;
  ve[q] = v;
  const float mh = m / c1, vh = v / c2;
  const float pq = p[q] - lr * mh / (sqrtf(vh) + 1e-8f);
  p[q] = pq;
  pb[q] = f2bf(pq);
}

__global__ void k_to_bf16(int np, const float* __restrict__ p, u16* __restrict__ pb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np) pb[q] = f2bf(p[q]);
}

// ------------------------------------------------------------ features

struct Coef16f {
  double c[16];
};

__device__ __forceinline__ void rot3(const float* R, const float* v, float* o) {
  for (int k = 0; k < 3; ++k) o[k] = R[3 * k] * v[0] + R[3 * k + 1] * v[1] + R[3 * k + 2] * v[2];
}

// dE (d and p) of each caller-order atom i: the current field (sorted slot
// iperm[i]) minus the ASPC combination of its history (slots weighted by ce);
// lanes in caller order so the history rows are read contiguously.
template <class real>
__global__ __launch_bounds__(256) void k_field_err(int n, const int* __restrict__ iperm, const real* __restrict__ fd,
                                                   const real* __restrict__ fp, const real* __restrict__ ealt,
                                                   Coef16f ce, float* __restrict__ de) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int si = iperm[i];
  const size_t slot = 6 * (size_t)n;
  double a[3], b[3];
  for (int d = 0; d < 3; ++d) {
    a[d] = (double)fd[3 * si + d];
    b[d] = (double)fp[3 * si + d];
  }
  for (int k = 0; k < 16; ++k) {
    const real* h = ealt + k * slot;
    for (int d = 0; d < 3; ++d) {
      a[d] -= ce.c[k] * (double)h[3 * (size_t)i + d];
      b[d] -= ce.c[k] * (double)h[3 * (size_t)n + 3 * (size_t)i + d];
    }
  }
  for (int d = 0; d < 3; ++d) {
    de[6 * (size_t)i + d] = (float)a[d];
    de[6 * (size_t)i + 3 + d] = (float)b[d];
  }
}

template <class real>
__global__ __launch_bounds__(256) void k_features(int n, const int* __restrict__ perm, const int* __restrict__ frame,
                                                  const int* __restrict__ zaxis, const int* __restrict__ xaxis,
                                                  const float* __restrict__ frot, const real* __restrict__ polarity,
                                                  const int* __restrict__ jpolar, const real* __restrict__ fd,
                                                  const float* __restrict__ de, const real* __restrict__ ualt,
                                                  int u0, int u1, float* __restrict__ F) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  const int fr = frame[i];
  const int zi = fr > 0 ? zaxis[i] : i;
  const int xi = (fr > 1 && xaxis[i] >= 0) ? xaxis[i] : i;
  float R[9];
  for (int q = 0; q < 9; ++q) R[q] = frot[9 * (size_t)s + q];
  float* f = F + (size_t)s * NF;
  const int js[3] = {i, zi, xi};
  for (int t = 0; t < 3; ++t) {
    const float* e = de + 6 * (size_t)js[t];
    rot3(R, e, f + 6 * t);
    rot3(R, e + 3, f + 6 * t + 3);
  }
  float e[3];
  for (int d = 0; d < 3; ++d) e[d] = (float)fd[3 * s + d];
  rot3(R, e, f + 18);
  const size_t slot = 6 * (size_t)n;
  float du[3], u[3];
  for (int d = 0; d < 3; ++d) {
    const float x0 = (float)ualt[u0 * slot + 3 * (size_t)i + d], x1 = (float)ualt[u1 * slot + 3 * (size_t)i + d];
    du[d] = x0 - x1;
    u[d] = x0;
  }
  rot3(R, du, f + 21);
  rot3(R, u, f + 24);
  f[27] = (float)polarity[s];
  for (int q = 0; q < 4; ++q) f[28 + q] = 0.0f;
  f[28 + min(jpolar[s], 3)] = 1.0f;
}

// history slot <- current direct field (sorted -> caller order)
template <class real>
__global__ void k_save_field(int n, const int* __restrict__ perm, const real* __restrict__ fd,
                             const real* __restrict__ fp, real* __restrict__ slot) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  for (int d = 0; d < 3; ++d) {
    slot[3 * (size_t)i + d] = fd[3 * s + d];
    slot[3 * (size_t)n + 3 * (size_t)i + d] = fp[3 * s + d];
  }
}

// capture every stride-th caller atom: raw features and R (u - u_aspc)
template <class real>
__global__ void k_capture(int n, int stride, long long base, const int* __restrict__ perm,
                          const float* __restrict__ F, const float* __restrict__ frot, const real* __restrict__ ud,
                          const real* __restrict__ up, const real* __restrict__ ua, float* __restrict__ cx,
                          float* __restrict__ cy) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  if (i % stride) return;
  const long long slot = base + i / stride;
  for (int q = 0; q < NF; ++q) cx[slot * NF + q] = F[(size_t)s * NF + q];
  const float* R = frot + 9 * (size_t)s;
  float a[3], b[3];
  for (int d = 0; d < 3; ++d) {
    a[d] = (float)((double)ud[3 * s + d] - (double)ua[3 * s + d]);
    b[d] = (float)((double)up[3 * s + d] - (double)ua[3 * (size_t)n + 3 * s + d]);
  }
  rot3(R, a, cy + slot * NO);
  rot3(R, b, cy + slot * NO + 3);
}

// column sums and sums of squares of rows [r0, r1) of a [rows][w] matrix,
// one block per chunk: out[blk][2w]
__global__ __launch_bounds__(256) void k_colstats(long long rows, int w, const float* __restrict__ A,
                                                  long long chunk, double* __restrict__ out) {
  __shared__ double sh[256];
  const long long r0 = blockIdx.x * chunk, r1 = min(rows, r0 + chunk);
  for (int c = 0; c < w; ++c) {
    double s1 = 0, s2 = 0;
    for (long long r = r0 + threadIdx.x; r < r1; r += 256) {
      const double v = A[r * w + c];
      s1 += v;
      s2 += v * v;
    }
    double t1 = block_sum<256>(s1, sh);
    double t2 = block_sum<256>(s2, sh);
    if (threadIdx.x == 0) {
      out[(size_t)blockIdx.x * 2 * w + c] = t1;
      out[(size_t)blockIdx.x * 2 * w + w + c] = t2;
    }
  }
}

// training batch: rows idx = (a * (start + b) + c) mod ncap
__global__ void k_make_batch(int m, long long start, unsigned long long a, unsigned long long c, long long ncap,
                             const float* __restrict__ cx, const float* __restrict__ cy,
                             const float* __restrict__ norm, u16* __restrict__ X, float* __restrict__ T) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= m) return;
  const unsigned long long idx = (unsigned long long)(((unsigned __int128)a * (unsigned long long)(start + b) + c) %
                                                      (unsigned long long)ncap);
  for (int q = 0; q < NF; ++q) X[(size_t)b * NF + q] = f2bf((cx[idx * NF + q] - norm[q]) * norm[NF + q]);
  for (int q = 0; q < NO; ++q)
    T[(size_t)b * NO + q] = (cy[idx * NO + q] - norm[2 * NF + q]) / norm[2 * NF + NO + q];
}

template <class real>
void features_t(Ctx& c, const double* ce, const double* cu) {
  Coef16f cf;
  for (int k = 0; k < 16; ++k) cf.c[k] = ce[k];
  const int u0 = (int)cu[0], u1 = (int)cu[1];
  c.gfeat.alloc(sizeof(float) * NF * (size_t)c.n);
  c.gde.alloc(sizeof(float) * 6 * (size_t)c.n);
  hipLaunchKernelGGL(k_field_err<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.iperm.as<int>(),
                     c.field.as<real>(), c.fieldp.as<real>(), c.ealt.as<real>(), cf, c.gde.as<float>());
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_features<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.as<int>(),
                     c.frame.as<int>(), c.zaxis.as<int>(), c.xaxis.as<int>(), c.frot.as<float>(),
                     c.polarity.as<real>(), c.jpolar.as<int>(), c.field.as<real>(), c.gde.as<float>(),
                     c.ualt.as<real>(), u0, u1, c.gfeat.as<float>());
  T9_LAUNCH_CHECK();
}

}  // namespace

int guess_nparam() { return NPARAM; }

void guess_setup(Ctx& c) {
  c.gp.alloc(sizeof(float) * (NPARAM + 1));
  c.gm.alloc(sizeof(float) * NPARAM);
  c.gv.alloc(sizeof(float) * NPARAM);
  c.gpb.alloc(sizeof(u16) * NPARAM);
  c.ggrad.alloc(sizeof(float) * (NPARAM + 1));
  c.gnorm.alloc(sizeof(float) * kNorm);
  c.ealt.alloc(rsize(c) * 6 * (size_t)c.n * 16);
  c.uaspc.alloc(rsize(c) * 6 * (size_t)c.n);
  c.raspc.alloc(rsize(c) * 6 * (size_t)c.n);
}

bool guess_wants_features(const Ctx& c) { return c.cap_stride > 0 || (c.guess_on && c.guess_ready); }

// coef_e[k]: weight of history slot k of the fields; coef_u[0], [1]: the
// slots of u(t-1), u(t-2)
void guess_features(Ctx& c, const double* coef_e, const double* coef_u) {
  if (c.prec == 32) features_t<float>(c, coef_e, coef_u);
  else features_t<double>(c, coef_e, coef_u);
}

void guess_save_fields(Ctx& c) {
  const size_t slot = 6 * (size_t)c.n;
  const int k = c.nealt % 16;
  if (c.prec == 32)
    hipLaunchKernelGGL(k_save_field<float>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.as<int>(),
                       c.field.as<float>(), c.fieldp.as<float>(), c.ealt.as<float>() + k * slot);
  else
    hipLaunchKernelGGL(k_save_field<double>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                       c.perm.as<int>(), c.field.as<double>(), c.fieldp.as<double>(), c.ealt.as<double>() + k * slot);
  T9_LAUNCH_CHECK();
  c.nealt += 1;
  if (c.nealt > 32) c.nealt -= 16;
}

void guess_to_bf16(Ctx& c, long long count, const float* src, void* dst) {
  hipLaunchKernelGGL(k_to_bf16, dim3(ceil_div(count, 256)), dim3(256), 0, c.stream, (int)count, src, (u16*)dst);
  T9_LAUNCH_CHECK();
}

// Inference weights in MFMA operand order, rebuilt after a parameter change.
void pack_weights(Ctx& c) {
  if (c.gpack_valid) return;
  c.gpack.alloc(sizeof(u16) * kFrags * 64 * 8);
  hipLaunchKernelGGL(k_pack_weights, dim3(ceil_div(kFrags * 64 * 8, 256)), dim3(256), 0, c.stream, c.gpb.as<u16>(),
                     c.gpack.as<u16>());
  T9_LAUNCH_CHECK();
  c.gpack_valid = true;
}

// Workgroups of the inference kernels: four 32-row tiles per pass, at most
// two workgroups per CU (each stages the 52 KB of weights in LDS).
int infer_blocks(int m) { return std::max(1, std::min(ceil_div(ceil_div(m, 32), 4), 512)); }

void mlp_forward(Ctx& c, int m, const void* X, float* Y) {
  pack_weights(c);
  hipLaunchKernelGGL(k_mlp_infer_y, dim3(infer_blocks(m)), dim3(256), 0, c.stream, m, (const u16*)X,
                     c.gpack.as<uint4>(), c.gp.as<float>(), Y);
  T9_LAUNCH_CHECK();
}

void guess_apply(Ctx& c, void* ud, void* up) {
  const int n = c.n;
  pack_weights(c);
  if (c.prec == 32)
    hipLaunchKernelGGL(k_mlp_apply<float>, dim3(infer_blocks(n)), dim3(256), 0, c.stream, n, c.gfeat.as<float>(),
                       c.gpack.as<uint4>(), c.gp.as<float>(), c.gnorm.as<float>(), c.frot.as<float>(), (float*)ud,
                       (float*)up);
  else
    hipLaunchKernelGGL(k_mlp_apply<double>, dim3(infer_blocks(n)), dim3(256), 0, c.stream, n, c.gfeat.as<float>(),
                       c.gpack.as<uint4>(), c.gp.as<float>(), c.gnorm.as<float>(), c.frot.as<float>(), (double*)ud,
                       (double*)up);
  T9_LAUNCH_CHECK();
}

void guess_capture(Ctx& c) {
  if (c.cap_stride <= 0) return;
  const int n = c.n;
  const long long per = (n + c.cap_stride - 1) / c.cap_stride;
  if (c.cap_n + per > c.cap_max) {  // buffer full: capture ends
    c.cap_stride = 0;
    return;
  }
  if (c.prec == 32)
    hipLaunchKernelGGL(k_capture<float>, dim3(ceil_div(n, 256)), dim3(256), 0, c.stream, n, c.cap_stride, c.cap_n,
                       c.perm.as<int>(), c.gfeat.as<float>(), c.frot.as<float>(), c.uind.as<float>(),
                       c.uinp.as<float>(), c.uaspc.as<float>(), c.cap_x.as<float>(), c.cap_y.as<float>());
  else
    hipLaunchKernelGGL(k_capture<double>, dim3(ceil_div(n, 256)), dim3(256), 0, c.stream, n, c.cap_stride, c.cap_n,
                       c.perm.as<int>(), c.gfeat.as<float>(), c.frot.as<float>(), c.uind.as<double>(),
                       c.uinp.as<double>(), c.uaspc.as<double>(), c.cap_x.as<float>(), c.cap_y.as<float>());
  T9_LAUNCH_CHECK();
  c.cap_n += per;
}

// ---- host-side operations behind the C ABI

void guess_set_params(Ctx& c, const float* params, const float* norm, uint64_t seed) {
  guess_setup(c);
  c.gpack_valid = false;
  std::vector<float> p(NPARAM);
  if (params) {
    std::copy(params, params + NPARAM, p.begin());
  } else {  // seeded uniform init, fan-in scaled; last layer small
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    auto fill = [&](int off, int fan_in, int count, double scale) {
      const double lim = std::sqrt(3.0 / fan_in) * scale;
      for (int q = 0; q < count; ++q) p[off + q] = (float)(lim * u(g));
    };
    fill(OFF_W1, NF, NF * NH, 1.0);
    fill(OFF_W2, NH, NH * NH, 1.0);
    fill(OFF_W3, NSK, NSK * NO, 0.1);
  }
  T9_HIP(hipMemcpyAsync(c.gp.p, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, c.stream));
  T9_HIP(hipMemsetAsync(c.gm.p, 0, sizeof(float) * NPARAM, c.stream));
  T9_HIP(hipMemsetAsync(c.gv.p, 0, sizeof(float) * NPARAM, c.stream));
  hipLaunchKernelGGL(k_to_bf16, dim3(ceil_div(NPARAM, 256)), dim3(256), 0, c.stream, NPARAM, c.gp.as<float>(),
                     c.gpb.as<u16>());
  T9_LAUNCH_CHECK();
  c.adam_t = 0;
  c.guess_params = true;
  if (norm) {
    c.gnorm_h.assign(norm, norm + kNorm);
    T9_HIP(hipMemcpyAsync(c.gnorm.p, c.gnorm_h.data(), sizeof(float) * kNorm, hipMemcpyHostToDevice, c.stream));
    c.guess_ready = true;
  }
  T9_HIP(hipStreamSynchronize(c.stream));
}

void guess_get_params(Ctx& c, float* params, float* norm) {
  if (!c.guess_params) throw ArgError("the learned guess has no parameters");
  if (params)
    T9_HIP(hipMemcpyAsync(params, c.gp.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  if (norm) {
    if (c.gnorm_h.size() != (size_t)kNorm) throw ArgError("the learned guess has no normalization yet");
    std::copy(c.gnorm_h.begin(), c.gnorm_h.end(), norm);
  }
}

// Mean squared error and its gradient on rows X (bf16 bits) / T (normalized).
double mlp_grad(Ctx& c, int m, const void* X, const float* T, float* grad_dev) {
  const int nwg = ceil_div(m, 32);
  c.gpart.alloc(sizeof(float) * (size_t)nwg * (NPARAM + 1));
  hipLaunchKernelGGL(k_mlp_bwd, dim3(nwg), dim3(64), 0, c.stream, m, m, (const u16*)X, T, c.gpb.as<u16>(),
                     c.gp.as<float>(), c.gpart.as<float>());
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_grad_reduce, dim3(ceil_div(NPARAM + 1, 256)), dim3(256), 0, c.stream, nwg,
                     c.gpart.as<float>(), grad_dev);
  T9_LAUNCH_CHECK();
  return 0.0;
}

void adam_step(Ctx& c, const float* grad_dev, double lr) {
  c.adam_t += 1;
  c.gpack_valid = false;
  const float c1 = (float)(1.0 - std::pow(0.9, c.adam_t)), c2 = (float)(1.0 - std::pow(0.999, c.adam_t));
  hipLaunchKernelGGL(k_adam, dim3(ceil_div(NPARAM, 256)), dim3(256), 0, c.stream, grad_dev, c.gp.as<float>(),
                     c.gm.as<float>(), c.gv.as<float>(), c.gpb.as<u16>(), (float)lr, c1, c2);
  T9_LAUNCH_CHECK();
}

void guess_capture_start(Ctx& c, int stride, long long max_samples) {
  if (stride < 0 || max_samples < 0) throw ArgError("bad capture arguments");
  c.cap_stride = stride;
  if (stride == 0) return;
  guess_setup(c);
  if (max_samples > c.cap_max) {
    c.cap_x.alloc(sizeof(float) * NF * (size_t)max_samples);
    c.cap_y.alloc(sizeof(float) * NO * (size_t)max_samples);
    c.cap_max = max_samples;
    c.cap_n = 0;
  }
}

// Normalization from the captured samples (mean, 1/std of the features; mean,
// std of the targets), fixed-order reductions.
void guess_fit_norm(Ctx& c) {
  if (c.cap_n < 2) throw ArgError("no captured samples to normalize");
  const int nblk = 256;
  const long long chunk = (c.cap_n + nblk - 1) / nblk;
  DBuf sx, sy;
  sx.alloc(sizeof(double) * nblk * 2 * NF);
  sy.alloc(sizeof(double) * nblk * 2 * NO);
  hipLaunchKernelGGL(k_colstats, dim3(nblk), dim3(256), 0, c.stream, c.cap_n, NF, c.cap_x.as<float>(), chunk,
                     sx.as<double>());
  hipLaunchKernelGGL(k_colstats, dim3(nblk), dim3(256), 0, c.stream, c.cap_n, NO, c.cap_y.as<float>(), chunk,
                     sy.as<double>());
  T9_LAUNCH_CHECK();
  std::vector<double> hx(nblk * 2 * NF), hy(nblk * 2 * NO);
  T9_HIP(hipMemcpyAsync(hx.data(), sx.p, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipMemcpyAsync(hy.data(), sy.p, sizeof(double) * hy.size(), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  std::vector<float> nm(kNorm);
  const double N = (double)c.cap_n;
  auto stats = [&](const std::vector<double>& h, int w, int col, double& mean, double& sd) {
    double s1 = 0, s2 = 0;
    for (int b = 0; b < nblk; ++b) {
      s1 += h[(size_t)b * 2 * w + col];
      s2 += h[(size_t)b * 2 * w + w + col];
    }
    mean = s1 / N;
    sd = std::sqrt(std::max(0.0, s2 / N - mean * mean));
  };
  for (int q = 0; q < NF; ++q) {
    double mu, sd;
    stats(hx, NF, q, mu, sd);
    nm[q] = (float)mu;
    nm[NF + q] = (float)(sd > 1e-12 ? 1.0 / sd : 1.0);
  }
  for (int q = 0; q < NO; ++q) {
    double mu, sd;
    stats(hy, NO, q, mu, sd);
    nm[2 * NF + q] = (float)mu;
    nm[2 * NF + NO + q] = (float)(sd > 1e-30 ? sd : 1.0);
  }
  c.gnorm_h = nm;
  T9_HIP(hipMemcpyAsync(c.gnorm.p, nm.data(), sizeof(float) * kNorm, hipMemcpyHostToDevice, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  c.guess_ready = true;
}

// Adam over the captured samples: `epochs` passes in a seeded pseudo-random
// order (i -> (a i + b) mod N with a coprime to N), batches of `batch` rows.
// Returns the mean loss of the first and the last epoch.
void guess_train(Ctx& c, int epochs, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last) {
  if (epochs < 1 || batch < 32) throw ArgError("bad training arguments");
  if (!c.guess_params) guess_set_params(c, nullptr, nullptr, seed);
  if (!c.guess_ready) guess_fit_norm(c);
  const long long N = c.cap_n;
  if (N < batch) throw ArgError("fewer captured samples than one batch");
  DBuf X, T, L;
  X.alloc(sizeof(u16) * NF * (size_t)batch);
  T.alloc(sizeof(float) * NO * (size_t)batch);
  const long long nb = N / batch;
  L.alloc(sizeof(float) * (size_t)nb);
  std::mt19937_64 g(seed);
  auto gcd = [](unsigned long long a, unsigned long long b) {
    while (b) {
      const unsigned long long t = a % b;
      a = b;
      b = t;
    }
    return a;
  };
  for (int ep = 0; ep < epochs; ++ep) {
    unsigned long long a = 0;
    do a = g() % (unsigned long long)N; while (a == 0 || gcd(a, (unsigned long long)N) != 1);
    const unsigned long long b = g() % (unsigned long long)N;
    for (long long q = 0; q < nb; ++q) {
      hipLaunchKernelGGL(k_make_batch, dim3(ceil_div(batch, 256)), dim3(256), 0, c.stream, batch, q * batch, a, b, N,
                         c.cap_x.as<float>(), c.cap_y.as<float>(), c.gnorm.as<float>(), X.as<u16>(), T.as<float>());
      T9_LAUNCH_CHECK();
      mlp_grad(c, batch, X.p, T.as<float>(), c.ggrad.as<float>());
      T9_HIP(hipMemcpyAsync(L.as<float>() + q, c.ggrad.as<float>() + NPARAM, sizeof(float), hipMemcpyDeviceToDevice,
                            c.stream));
      adam_step(c, c.ggrad.as<float>(), lr);
    }
    std::vector<float> hl(nb);
    T9_HIP(hipMemcpyAsync(hl.data(), L.p, sizeof(float) * nb, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    double s = 0;
    for (float v : hl) s += v;
    s /= (double)nb * batch * NO;
    if (!std::isfinite(s)) throw ArgError("training diverged (non-finite loss)");
    if (ep == 0 && loss_first) *loss_first = s;
    if (loss_last) *loss_last = s;
  }
}

}  // namespace t9
