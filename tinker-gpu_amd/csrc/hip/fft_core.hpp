// Mixed-radix Stockham FFT building blocks (radices 2, 3, 4, 5, 8), usable from
// device kernels and from host code (the host path exists only so the same
// arithmetic can be exercised on a CPU).
//
// A transform of length N runs one stage per factor R with Ns = product of
// the factors already applied: butterfly j (0 <= j < N/R), k = j mod Ns, reads
// x[j + r N/R], multiplies by W_N^{k r N/(Ns R)} (no modulo needed since
// k r N/(Ns R) < N), applies the R-point DFT and writes
// y[(j / Ns) Ns R + k + r Ns].  The output is in natural order.
#pragma once
#include <hip/hip_runtime.h>

namespace t9 {

template <class real>
struct Cplx {
  real x, y;
};

template <class real>
__host__ __device__ __forceinline__ Cplx<real> cadd(Cplx<real> a, Cplx<real> b) {
  return {a.x + b.x, a.y + b.y};
}
template <class real>
__host__ __device__ __forceinline__ Cplx<real> csub(Cplx<real> a, Cplx<real> b) {
  return {a.x - b.x, a.y - b.y};
}
template <class real>
__host__ __device__ __forceinline__ Cplx<real> cmulc(Cplx<real> a, Cplx<real> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class real>
__host__ __device__ __forceinline__ Cplx<real> cscale(Cplx<real> a, real s) {
  return {a.x * s, a.y * s};
}
// multiply by s*i (s = +1 or -1)
template <class real>
__host__ __device__ __forceinline__ Cplx<real> ctimes_i(Cplx<real> a, real s) {
  return {-s * a.y, s * a.x};
}

// In-place R-point DFT, y_q = sum_r v_r exp(sign 2 pi i r q / R).
template <int R, class real>
__host__ __device__ __forceinline__ void small_dft(Cplx<real>* v, real sg) {
  if constexpr (R == 2) {
    Cplx<real> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const real h = (real)0.86602540378443864676;
    Cplx<real> t1 = cadd(v[1], v[2]);
    Cplx<real> t2 = ctimes_i(csub(v[1], v[2]), sg * h);
    Cplx<real> m = csub(v[0], cscale(t1, (real)0.5));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, t2);
    v[2] = csub(m, t2);
  } else if constexpr (R == 4) {
    Cplx<real> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    Cplx<real> c = cadd(v[1], v[3]), d = ctimes_i(csub(v[1], v[3]), sg);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  } else if constexpr (R == 8) {
    // two 4-point DFTs (even, odd inputs) and the eighth roots of unity
    Cplx<real> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    small_dft<4, real>(e, sg);
    small_dft<4, real>(o, sg);
    const real h = (real)0.70710678118654752440;
    const Cplx<real> w1 = {h * (o[1].x - sg * o[1].y), h * (o[1].y + sg * o[1].x)};   // (1 + sg i)/sqrt2
    const Cplx<real> w2 = ctimes_i(o[2], sg);                                           // sg i
    const Cplx<real> w3 = {h * (-o[3].x - sg * o[3].y), h * (-o[3].y + sg * o[3].x)};  // (-1 + sg i)/sqrt2
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
  } else {
    static_assert(R == 5, "radix");
    const real c1 = (real)0.30901699437494742410, c2 = (real)-0.80901699437494742410;
    const real s1 = (real)0.95105651629515357212, s2 = (real)0.58778525229247312917;
    Cplx<real> a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
    Cplx<real> b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
    Cplx<real> x0 = v[0];
    Cplx<real> m1 = cadd(x0, cadd(cscale(a1, c1), cscale(a2, c2)));
    Cplx<real> m2 = cadd(x0, cadd(cscale(a1, c2), cscale(a2, c1)));
    Cplx<real> n1 = ctimes_i(cadd(cscale(b1, s1), cscale(b2, s2)), sg);
    Cplx<real> n2 = ctimes_i(csub(cscale(b1, s2), cscale(b2, s1)), sg);
    v[0] = cadd(x0, cadd(a1, a2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
}

// One Stockham stage over `nl` lines of length N stored line-major
// (src[l*N + e]); butterflies are distributed round-robin over `nthr`
// workers, this call doing the ones of worker `tid`.  tw holds
// exp(-2 pi i m / N), m = 0..N-1 (sign > 0 uses the conjugate).
template <int R, class real>
__host__ __device__ __forceinline__ void stockham_stage(const Cplx<real>* src, Cplx<real>* dst, int N, int pitch,
                                                        int Ns, int nl, const Cplx<real>* tw, real sg, int tid,
                                                        int nthr) {
  const int nb = N / R;
  const int tstep = N / (Ns * R);
  const int total = nl * nb;
  for (int q = tid; q < total; q += nthr) {
    const int lb = q / nb;
    const int j = q - lb * nb;
    const int jd = j / Ns;
    const int k = j - jd * Ns;
    const Cplx<real>* s = src + lb * pitch + j;
    Cplx<real> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[r * nb];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        Cplx<real> t = tw[k * r * tstep];
        if (sg > 0) t.y = -t.y;  // inverse transform: conjugate twiddle
        v[r] = cmulc(v[r], t);
      }
    }
    small_dft<R>(v, sg);
    Cplx<real>* d = dst + lb * pitch + jd * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * Ns] = v[r];
  }
}

struct FftPlan {
  int n;
  int nstage;
  int radix[16];
};

// Run all stages of `plan` on nl lines (line l at buf + l*pitch); returns the
// buffer holding the result (buf0 or buf1).  `barrier` separates stages (a
// workgroup barrier on the device, a no-op when one host worker does
// everything).
template <class real, class Barrier>
__host__ __device__ __forceinline__ Cplx<real>* fft_stages(Cplx<real>* buf0, Cplx<real>* buf1, int nl, int pitch,
                                                          const FftPlan& plan, const Cplx<real>* tw, int sign,
                                                          int tid, int nthr, Barrier barrier) {
  Cplx<real>* src = buf0;
  Cplx<real>* dst = buf1;
  const real sg = (real)sign;
  const int N = plan.n;
  int Ns = 1;
  for (int st = 0; st < plan.nstage; ++st) {
    const int R = plan.radix[st];
    switch (R) {
      case 2: stockham_stage<2, real>(src, dst, N, pitch, Ns, nl, tw, sg, tid, nthr); break;
      case 3: stockham_stage<3, real>(src, dst, N, pitch, Ns, nl, tw, sg, tid, nthr); break;
      case 4: stockham_stage<4, real>(src, dst, N, pitch, Ns, nl, tw, sg, tid, nthr); break;
      case 8: stockham_stage<8, real>(src, dst, N, pitch, Ns, nl, tw, sg, tid, nthr); break;
      default: stockham_stage<5, real>(src, dst, N, pitch, Ns, nl, tw, sg, tid, nthr); break;
    }
    barrier();
    Cplx<real>* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

// Compile-time plan: the st-th radix of N, radix 8 first (fewer stages and
// barriers), then 4, 2, 3, 5.
__host__ __device__ constexpr int ct_radix(int N, int st) {
  int m = N, c = 0;
  const int rs[5] = {8, 4, 2, 3, 5};
  for (int a = 0; a < 5; ++a)
    while (m % rs[a] == 0) {
      if (c == st) return rs[a];
      m /= rs[a];
      ++c;
    }
  return 0;
}
__host__ __device__ constexpr int ct_nstage(int N) {
  int c = 0;
  while (ct_radix(N, c) != 0) ++c;
  return c;
}

// One Stockham stage with N, R and Ns known at compile time (the index
// arithmetic reduces to multiplications and shifts); lines of pitch N + 1.
template <int R, int N, int NS, class real>
__device__ __forceinline__ void stockham_stage_ct(const Cplx<real>* src, Cplx<real>* dst, int nl,
                                                  const Cplx<real>* tw, real sg, int tid, int nthr) {
  constexpr int nb = N / R;
  constexpr int tstep = N / (NS * R);
  constexpr int P = N + 1;
  const int total = nl * nb;
  for (int q = tid; q < total; q += nthr) {
    const int lb = q / nb;
    const int j = q - lb * nb;
    const int jd = j / NS;
    const int k = j - jd * NS;
    const Cplx<real>* s = src + lb * P + j;
    Cplx<real> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[r * nb];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        Cplx<real> t = tw[k * r * tstep];
        t.y = sg > 0 ? -t.y : t.y;
        v[r] = cmulc(v[r], t);
      }
    }
    small_dft<R>(v, sg);
    Cplx<real>* d = dst + lb * P + jd * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * NS] = v[r];
  }
}

template <class real, int N, int ST = 0, int NS = 1>
__device__ __forceinline__ Cplx<real>* fft_stages_ct(Cplx<real>* src, Cplx<real>* dst, int nl, const Cplx<real>* tw,
                                                    int sign, int tid, int nthr) {
  if constexpr (ST == ct_nstage(N)) {
    return src;
  } else {
    constexpr int R = ct_radix(N, ST);
    stockham_stage_ct<R, N, NS, real>(src, dst, nl, tw, (real)sign, tid, nthr);
    __syncthreads();
    return fft_stages_ct<real, N, ST + 1, NS * R>(dst, src, nl, tw, sign, tid, nthr);
  }
}

inline FftPlan make_fft_plan(int N) {
  FftPlan p{};
  p.n = N;
  p.nstage = 0;
  int m = N;
  const int rs[4] = {4, 2, 3, 5};
  for (int a = 0; a < 4; ++a)
    while (m % rs[a] == 0 && p.nstage < 16) {
      p.radix[p.nstage++] = rs[a];
      m /= rs[a];
    }
  if (m != 1) p.nstage = -1;  // not a 2,3,5-smooth length
  return p;
}

}  // namespace t9
