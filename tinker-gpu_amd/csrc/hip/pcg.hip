// Preconditioned conjugate gradient for the mutual induced dipoles.
//
// Algorithm of the reference (include/ff/cuinduce.h:8-38,
// src/acc/amoeba/induce.cpp:41-292, src/cu/amoeba/pcg.cu:14-185): solve
// (alpha^-1 + T) u = E for the d and p dipole sets simultaneously, direct-field
// guess, zeroed residual on unpolarizable sites, min 3 iterations, stop when
// debye * sqrt(max(r_d.r_d, r_p.r_p) / n) < polar-eps, then one "peek" step
// u += peek * alpha * r.  Differences by design: the CG scalars (dot products,
// step lengths, residual norm) are FP64 and live on the device — each dot is
// a fixed-order two-level reduction, so iterates are bit-reproducible — and
// the host reads back one residual per iteration to decide convergence.
//
// Predictor (polar-predict ASPC, src/amoeba/induce.cpp:29-69,
// src/cu/upredict.cu:34-201): after 16 converged solves the initial guess is
// sum_j B_j u(t - j) over the last 16 solutions with Kolafa's always-stable
// predictor coefficients (k = 14); before that, and without the keyword, the
// direct-field guess.  The history is kept in caller order so a re-sort of
// the atoms at a list rebuild does not disturb it.  With the learned guess on
// (guess.hip) the ASPC guess is corrected by the network; the corrected guess
// is kept only when its initial residual is not worse than the ASPC guess's
// (re-measured every 20 solves), else the solve starts from ASPC.
#include <array>

#include <algorithm>

#include "ctx.hpp"

namespace t9 {
namespace {

constexpr int kB = 256;
// device scalar slots
enum { S_SUM = 0, S_SUMP, S_A, S_AP, S_B, S_BP, S_EPSD, S_EPSP, S_EPS, S_EPS0, S_DONE = kDoneSlot, S_NSLOT = 16 };
constexpr int kMaxUalt = 16;

struct Coef16 {
  double c[kMaxUalt];
};

// u0 = sum_k c_k hist_k (caller-order history, slot k: [2][n][3]); lanes in
// caller order so the 16 history rows are read contiguously
template <class real>
__global__ __launch_bounds__(kB) void k_aspc_guess(int n, const int* __restrict__ iperm, Coef16 cf,
                                                  const real* __restrict__ hist, real* __restrict__ ud,
                                                  real* __restrict__ up) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = iperm[i];
  const size_t slot = 6 * (size_t)n;
  for (int d = 0; d < 3; ++d) {
    double a = 0, b = 0;
    for (int k = 0; k < kMaxUalt; ++k) {
      a += cf.c[k] * (double)hist[k * slot + 3 * (size_t)i + d];
      b += cf.c[k] * (double)hist[k * slot + 3 * (size_t)n + 3 * (size_t)i + d];
    }
    ud[3 * s + d] = (real)a;
    up[3 * s + d] = (real)b;
  }
}

// r0 = alpha^-1 (udir - u0) + (field of u0), zero on unpolarizable sites
template <class real>
__global__ __launch_bounds__(kB) void k_rsd_guess(int n, const real* __restrict__ pol, const real* __restrict__ polinv,
                                                 const real* __restrict__ udir, const real* __restrict__ udirp,
                                                 const real* __restrict__ ud, const real* __restrict__ up,
                                                 real* __restrict__ rd, real* __restrict__ rp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const bool zero = pol[s] == 0;
  const real pi = polinv[s];
  for (int d = 0; d < 3; ++d) {
    const int k = 3 * s + d;
    rd[k] = zero ? (real)0 : pi * (udir[k] - ud[k]) + rd[k];
    rp[k] = zero ? (real)0 : pi * (udirp[k] - up[k]) + rp[k];
  }
}

template <class real>
__global__ __launch_bounds__(kB) void k_save_hist(int n, const int* __restrict__ perm, const real* __restrict__ ud,
                                                 const real* __restrict__ up, real* __restrict__ slot) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  for (int d = 0; d < 3; ++d) {
    slot[3 * (size_t)i + d] = ud[3 * s + d];
    slot[3 * (size_t)n + 3 * (size_t)i + d] = up[3 * s + d];
  }
}

template <class real>
__global__ __launch_bounds__(kB) void k_udir(int n, const real* __restrict__ pol, const real* __restrict__ fd,
                                            const real* __restrict__ fp, real* __restrict__ udir,
                                            real* __restrict__ udirp, real* __restrict__ ud, real* __restrict__ up,
                                            int guess) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const real a = pol[s];
  for (int d = 0; d < 3; ++d) {
    const real x = a * fd[3 * s + d], y = a * fp[3 * s + d];
    udir[3 * s + d] = x;
    udirp[3 * s + d] = y;
    ud[3 * s + d] = guess ? x : 0;
    up[3 * s + d] = guess ? y : 0;
  }
}

template <class real>
__global__ __launch_bounds__(kB) void k_copy3_zero_nopol(int n, const real* __restrict__ pol,
                                                        const real* __restrict__ a, const real* __restrict__ b,
                                                        real* __restrict__ ra, real* __restrict__ rb) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const bool zero = pol[s] == 0;
  for (int d = 0; d < 3; ++d) {
    ra[3 * s + d] = zero ? (real)0 : a[3 * s + d];
    rb[3 * s + d] = zero ? (real)0 : b[3 * s + d];
  }
}

struct DotOut {
  double* partials;   // [nblk][4]
  unsigned* counter;  // workgroups finished (reset by the last one)
  double* sc;         // CG scalars
  int stage, n;
  double debye;
  // stage 2: convergence test of iteration `iter` (src/cu/amoeba/pcg.cu:147-170):
  // eps < tol once iter >= iter_min, or iter >= iter_max; the converged
  // iteration is stored in slot S_DONE
  double tol;
  int iter, iter_min, iter_max;
  double* trace;  // RMS residual per iteration (entry 0: initial guess)
};

// Finalize: sum the per-block partials in fixed order, then update the CG
// scalars.  stage 0: initial sum = r.z (slots SUM, SUMP); 1: a = sum / p.Tp
// (slots A, AP); 2: b = sum1 / sum, sum = sum1, eps from r.r, convergence
// test; 3: eps of the initial residual.  Run by the last workgroup of the
// dot kernel.
__device__ void finalize_block(int nblk, int ndots, const DotOut& o, const double* __restrict__ partials,
                               double* __restrict__ sc, double* sh) {
  const int stage = o.stage, n = o.n;
  const double debye = o.debye;
  double tot[4] = {0, 0, 0, 0};
  for (int q = 0; q < ndots; ++q) {
    double v = 0;
    for (int b = threadIdx.x; b < nblk; b += kB) v += partials[b * 4 + q];
    tot[q] = block_sum<kB>(v, sh);
  }
  if (threadIdx.x != 0) return;
  if (stage == 0) {
    sc[S_SUM] = tot[0];
    sc[S_SUMP] = tot[1];
  } else if (stage == 3) {
    sc[S_EPS0] = debye * sqrt(fmax(tot[0], tot[1]) / n);
    o.trace[0] = sc[S_EPS0];
  } else if (stage == 1) {
    sc[S_A] = tot[0] != 0 ? sc[S_SUM] / tot[0] : 0;
    sc[S_AP] = tot[1] != 0 ? sc[S_SUMP] / tot[1] : 0;
  } else {
    sc[S_B] = sc[S_SUM] != 0 ? tot[0] / sc[S_SUM] : 0;
    sc[S_BP] = sc[S_SUMP] != 0 ? tot[1] / sc[S_SUMP] : 0;
    sc[S_SUM] = tot[0];
    sc[S_SUMP] = tot[1];
    sc[S_EPSD] = tot[2];
    sc[S_EPSP] = tot[3];
    const double eps = debye * sqrt(fmax(tot[2], tot[3]) / n);
    sc[S_EPS] = eps;
    if (o.iter < kTraceCap) o.trace[o.iter] = eps;
    bool done = eps < o.tol;
    if (o.iter < o.iter_min) done = false;
    if (o.iter >= o.iter_max) done = true;
    if (done && sc[S_DONE] == 0) sc[S_DONE] = o.iter;
  }
}

// Every workgroup writes its partial sums; the last one to finish (counted
// with a device-scope atomic after a release fence) sums all partials in
// workgroup order and updates the scalars, so the result does not depend on
// which workgroup finishes last.
__device__ __forceinline__ void dots_epilogue(const double* tot, int ndots, const DotOut& o, double* sh) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    for (int q = 0; q < ndots; ++q) o.partials[blockIdx.x * 4 + q] = tot[q];
    __threadfence();
    last = atomicAdd(o.counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  finalize_block(gridDim.x, ndots, o, o.partials, o.sc, sh);
  if (threadIdx.x == 0) atomicExch(o.counter, 0u);
}

// two dot products a1.b1 and a2.b2 and optionally a3.b3, a4.b4, finalized
template <class real>
__global__ __launch_bounds__(kB) void k_dots(int n, int ndots, const real* __restrict__ a1,
                                            const real* __restrict__ b1, const real* __restrict__ a2,
                                            const real* __restrict__ b2, const real* __restrict__ a3,
                                            const real* __restrict__ b3, const real* __restrict__ a4,
                                            const real* __restrict__ b4, DotOut o, Gate gate = Gate()) {
  if (gate_closed(gate)) return;
  __shared__ double sh[kB];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const real* A[4] = {a1, a2, a3, a4};
  const real* Bv[4] = {b1, b2, b3, b4};
  double tot[4] = {0, 0, 0, 0};
  for (int q = 0; q < ndots; ++q) {
    double v = 0;
    if (s < n)
      for (int d = 0; d < 3; ++d) v += (double)A[q][3 * s + d] * (double)Bv[q][3 * s + d];
    tot[q] = block_sum<kB>(v, sh);
  }
  dots_epilogue(tot, ndots, o, sh);
}

// vec = conj / alpha - field (reference pcgP1) with the dots conj.vec,
// finalized into a = sum / p.Tp
template <class real>
__global__ __launch_bounds__(kB) void k_p1_dots(int n, const real* __restrict__ polinv, const real* __restrict__ cd,
                                               const real* __restrict__ cp, const real* __restrict__ fd,
                                               const real* __restrict__ fp, real* __restrict__ vd,
                                               real* __restrict__ vp, DotOut o, Gate gate) {
  if (gate_closed(gate)) return;
  __shared__ double sh[kB];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double x = 0, y = 0;
  if (s < n) {
    const real pi = polinv[s];
    for (int d = 0; d < 3; ++d) {
      const real a = pi * cd[3 * s + d] - fd[3 * s + d];
      const real b = pi * cp[3 * s + d] - fp[3 * s + d];
      vd[3 * s + d] = a;
      vp[3 * s + d] = b;
      x += (double)cd[3 * s + d] * (double)a;
      y += (double)cp[3 * s + d] * (double)b;
    }
  }
  double tot[4] = {0, 0, 0, 0};
  tot[0] = block_sum<kB>(x, sh);
  tot[1] = block_sum<kB>(y, sh);
  dots_epilogue(tot, 2, o, sh);
}

// u += a p; r -= a vec; r = 0 on unpolarizable sites   (reference pcgP2)
template <class real>
__global__ __launch_bounds__(kB) void k_p2(int n, const double* __restrict__ sc, const real* __restrict__ pol,
                                          const real* __restrict__ cd, const real* __restrict__ cp,
                                          const real* __restrict__ vd, const real* __restrict__ vp,
                                          real* __restrict__ ud, real* __restrict__ up, real* __restrict__ rd,
                                          real* __restrict__ rp, Gate gate) {
  if (gate_closed(gate)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const real a = (real)sc[S_A], ap = (real)sc[S_AP];
  const bool zero = pol[s] == 0;
  for (int d = 0; d < 3; ++d) {
    const int k = 3 * s + d;
    ud[k] += a * cd[k];
    up[k] += ap * cp[k];
    rd[k] = zero ? (real)0 : rd[k] - a * vd[k];
    rp[k] = zero ? (real)0 : rp[k] - ap * vp[k];
  }
}

// p = z + b p   (reference pcgP3); in the iteration that converged, also
// the peek step u += peek * alpha * r (reference pcgPeek)
template <class real>
__global__ __launch_bounds__(kB) void k_p3(int n, const double* __restrict__ sc, const real* __restrict__ zd,
                                          const real* __restrict__ zp, real* __restrict__ cd,
                                          real* __restrict__ cp, real peek, const real* __restrict__ pol,
                                          const real* __restrict__ rd, const real* __restrict__ rp,
                                          real* __restrict__ ud, real* __restrict__ up, Gate gate) {
  if (gate_closed(gate)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const real b = (real)sc[S_B], bp = (real)sc[S_BP];
  for (int d = 0; d < 3; ++d) {
    const int k = 3 * s + d;
    cd[k] = zd[k] + b * cd[k];
    cp[k] = zp[k] + bp * cp[k];
  }
  if (sc[S_DONE] == (double)gate.iter) {
    const real t = peek * pol[s];
    for (int d = 0; d < 3; ++d) {
      ud[3 * s + d] += t * rd[3 * s + d];
      up[3 * s + d] += t * rp[3 * s + d];
    }
  }
}

__global__ void k_set_scalar(double* sc, int slot, double v) {
  if (threadIdx.x == 0) sc[slot] = v;
}

template <class real>
int induce_t(Ctx& c, int* iters, double* rms) {
  const int n = c.n;
  const int nblk = ceil_div(n, kB);
  const dim3 G(nblk), Bk(kB);
  hipStream_t st = c.stream;
  real* field = c.field.as<real>();
  real* fieldp = c.fieldp.as<real>();
  real* ud = c.uind.as<real>();
  real* up = c.uinp.as<real>();
  real* rd = c.rsd.as<real>();
  real* rp = c.rsdp.as<real>();
  real* zd = c.zrsd.as<real>();
  real* zp = c.zrsdp.as<real>();
  real* cd = c.conj.as<real>();
  real* cp = c.conjp.as<real>();
  real* vd = c.vec.as<real>();
  real* vp = c.vecp.as<real>();
  double* part = c.partials.as<double>();
  double* sc = c.scalars.as<double>();
  const double debye = kDebye;
  const bool guess = c.opt.pcg_guess;
  const int maxiter = 100;
  const int miniter = n < 3 ? n : 3;
  int iter = 0, iter0 = 0;  // iter0: iterations spent before a restart from ASPC
  auto dout = [&](int stage, int it = 0) {
    DotOut o;
    o.partials = part;
    o.counter = c.dot_counter.as<unsigned>();
    o.sc = sc;
    o.stage = stage;
    o.n = n;
    o.debye = debye;
    o.tol = c.opt.polar_eps;
    o.iter = it;
    o.iter_min = iter0 + miniter;
    o.iter_max = c.opt.polar_iter;
    o.trace = c.pcg_trace.as<double>();
    return o;
  };
  // slot S_DONE = 0: no iteration has converged yet
  auto reset_done = [&]() {
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, st, sc, (int)S_DONE, 0.0);
    T9_LAUNCH_CHECK();
  };

  dfield_all(c);
  // ASPC once 16 solutions are stored (src/cu/amoeba/pcg.cu:30-33)
  const bool aspc = c.opt.predictor == "ASPC";
  const bool predict = aspc && c.nualt >= kMaxUalt;
  // learned-guess features need 16 stored fields as well
  const bool feats = predict && c.nealt >= kMaxUalt && guess_wants_features(c);
  const bool learned = feats && c.guess_on && c.guess_ready;
  hipLaunchKernelGGL(k_udir<real>, G, Bk, 0, st, n, c.polarity.as<real>(), field, fieldp, c.udir.as<real>(),
                     c.udirp.as<real>(), ud, up, (guess || predict) ? 1 : 0);
  T9_LAUNCH_CHECK();
  auto slot_of_age = [](int count, int age) { return ((count - 1 - age) % kMaxUalt + kMaxUalt) % kMaxUalt; };
  // r0 of the guess in ud/up
  auto residual = [&]() {
    ufield_all(c, ud, up, rd, rp);
    hipLaunchKernelGGL(k_rsd_guess<real>, G, Bk, 0, st, n, c.polarity.as<real>(), c.polinv.as<real>(),
                       c.udir.as<real>(), c.udirp.as<real>(), ud, up, rd, rp);
    T9_LAUNCH_CHECK();
  };
  // RMS of r0 in Debye, read back now
  auto residual_rms = [&]() {
    hipLaunchKernelGGL(k_dots<real>, G, Bk, 0, st, n, 2, rd, rd, rp, rp, (const real*)nullptr,
                       (const real*)nullptr, (const real*)nullptr, (const real*)nullptr, dout(3));
    T9_LAUNCH_CHECK();
    T9_HIP(hipMemcpyAsync(c.h_scalars, sc, sizeof(double) * S_NSLOT, hipMemcpyDeviceToHost, st));
    T9_HIP(hipStreamSynchronize(st));
    return c.h_scalars[S_EPS0];
  };
  const size_t vbytes = sizeof(real) * 3 * (size_t)n;
  real* ua = c.uaspc.as<real>();
  c.last_predictor = predict ? 1 : 0;
  if (predict) {
    const auto B = aspc_coefficients();
    Coef16 cf;
    for (int k = 0; k < kMaxUalt; ++k) cf.c[k] = B[((c.nualt - 1 - k) % kMaxUalt + kMaxUalt) % kMaxUalt];
    hipLaunchKernelGGL(k_aspc_guess<real>, G, Bk, 0, st, n, c.iperm.as<int>(), cf, c.ualt.as<real>(), ud, up);
    T9_LAUNCH_CHECK();
    if (feats) {
      T9_HIP(hipMemcpyAsync(ua, ud, vbytes, hipMemcpyDeviceToDevice, st));
      T9_HIP(hipMemcpyAsync(ua + 3 * (size_t)n, up, vbytes, hipMemcpyDeviceToDevice, st));
      double ce[kMaxUalt], cu[2];
      for (int k = 0; k < kMaxUalt; ++k) ce[k] = B[((c.nealt - 1 - k) % kMaxUalt + kMaxUalt) % kMaxUalt];
      cu[0] = slot_of_age(c.nualt, 0);
      cu[1] = slot_of_age(c.nualt, 1);
      guess_features(c, ce, cu);
    }
  }
  if (aspc) guess_save_fields(c);
  bool learned_try = false;  // the learned guess is in use; judged at the first readback
  if (learned) {
    // the ASPC residual is re-measured every 50 learned solves (one extra
    // matrix-vector product); between checks the learned guess is compared
    // with that reference at the first iteration's readback
    if (c.guess_steps % 50 == 0) {
      residual();
      c.aspc_eps0_ref = residual_rms();
    }
    ++c.guess_steps;
    guess_apply(c, ud, up);
    residual();
    learned_try = true;
    c.last_predictor = 2;
  } else if (predict) {
    residual();
  } else if (guess) {
    ufield_all(c, ud, up, rd, rp);
    hipLaunchKernelGGL(k_copy3_zero_nopol<real>, G, Bk, 0, st, n, c.polarity.as<real>(), rd, rp, rd, rp);
  } else {
    hipLaunchKernelGGL(k_copy3_zero_nopol<real>, G, Bk, 0, st, n, c.polarity.as<real>(), field, fieldp, rd, rp);
  }
  T9_LAUNCH_CHECK();
  // residual norm of the initial guess (Debye, read back with the first
  // iteration), first preconditioned search direction
  auto start_cg = [&]() {
    hipLaunchKernelGGL(k_dots<real>, G, Bk, 0, st, n, 2, rd, rd, rp, rp, (const real*)nullptr,
                       (const real*)nullptr, (const real*)nullptr, (const real*)nullptr, dout(3));
    T9_LAUNCH_CHECK();
    precond_apply(c, rd, rp, zd, zp);
    T9_HIP(hipMemcpyAsync(cd, zd, sizeof(real) * 3 * n, hipMemcpyDeviceToDevice, st));
    T9_HIP(hipMemcpyAsync(cp, zp, sizeof(real) * 3 * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_dots<real>, G, Bk, 0, st, n, 2, rd, zd, rp, zp, (const real*)nullptr,
                       (const real*)nullptr, (const real*)nullptr, (const real*)nullptr, dout(0));
    T9_LAUNCH_CHECK();
  };
  reset_done();
  start_cg();

  // One CG iteration, gated: its kernels return at once if an earlier
  // iteration converged (the test runs in the last workgroup of the r.r dot
  // product; the peek step is applied by that iteration's p3).
  auto enqueue = [&](int it) {
    c.gate.sc = sc;
    c.gate.iter = it;
    const Gate g = c.gate;
    ufield_all(c, cd, cp, field, fieldp);
    hipLaunchKernelGGL(k_p1_dots<real>, G, Bk, 0, st, n, c.polinv.as<real>(), cd, cp, field, fieldp, vd, vp,
                       dout(1, it), g);
    hipLaunchKernelGGL(k_p2<real>, G, Bk, 0, st, n, sc, c.polarity.as<real>(), cd, cp, vd, vp, ud, up, rd, rp, g);
    T9_LAUNCH_CHECK();
    precond_apply(c, rd, rp, zd, zp);
    hipLaunchKernelGGL(k_dots<real>, G, Bk, 0, st, n, 4, rd, zd, rp, zp, rd, rd, rp, rp, dout(2, it), g);
    hipLaunchKernelGGL(k_p3<real>, G, Bk, 0, st, n, sc, zd, zp, cd, cp, (real)c.opt.pcg_peek,
                       c.polarity.as<real>(), rd, rp, ud, up, g);
    T9_LAUNCH_CHECK();
    c.gate = Gate();
  };
  // The host enqueues as many iterations as the previous solve needed, then
  // reads the scalars back once; after that one iteration per readback.  A
  // learned guess is judged after its first iteration.
  double eps = 100;
  int enq = 0;
  for (;;) {
    const int upto = learned_try ? enq + 1 : std::max(enq + 1, iter0 + c.pcg_planned);
    while (enq < upto) enqueue(++enq);
    T9_HIP(hipMemcpyAsync(c.h_scalars, sc, sizeof(double) * S_NSLOT, hipMemcpyDeviceToHost, st));
    T9_HIP(hipStreamSynchronize(st));
    if (learned_try) {
      learned_try = false;
      if (!(c.h_scalars[S_EPS0] <= c.aspc_eps0_ref)) {  // worse than ASPC, or not finite: restart
        T9_HIP(hipMemcpyAsync(ud, ua, vbytes, hipMemcpyDeviceToDevice, st));
        T9_HIP(hipMemcpyAsync(up, ua + 3 * (size_t)n, vbytes, hipMemcpyDeviceToDevice, st));
        residual();
        reset_done();
        start_cg();
        ++c.guess_fallback;
        c.last_predictor = 1;
        iter0 = enq;
        continue;
      }
      ++c.guess_used;
    }
    const double done_at = c.h_scalars[S_DONE];
    if (done_at != 0) {
      iter = (int)done_at;
      eps = c.h_scalars[S_EPS];
      break;
    }
  }
  c.pcg_planned = iter - iter0;
  c.pcg_trace_len = std::min(iter + 1, kTraceCap);
  c.uind_valid = true;
  c.last_guess_rms = c.h_scalars[S_EPS0];
  if (c.last_predictor == 1 && !learned) c.aspc_eps0_ref = c.last_guess_rms;
  if (feats) guess_capture(c);
  if (c.opt.predictor == "ASPC") {  // store the solution (src/amoeba/induce.cpp:29-62)
    const size_t slot = 6 * (size_t)n;
    hipLaunchKernelGGL(k_save_hist<real>, G, Bk, 0, st, n, c.perm.as<int>(), ud, up,
                       c.ualt.as<real>() + (size_t)(c.nualt % kMaxUalt) * slot);
    T9_LAUNCH_CHECK();
    c.nualt += 1;
    if (c.nualt > 2 * kMaxUalt) c.nualt -= kMaxUalt;
  }
  if (iters) *iters = iter;
  if (rms) *rms = eps;
  return iter >= maxiter ? 1 : 0;
}

}  // namespace

// Kolafa's always-stable predictor-corrector coefficients for k = 14
// (J. Kolafa, J. Comput. Chem. 25, 335 (2004)):
//   B_j = (-1)^(j+1) j C(2k+4, k+2-j) / C(2k+2, k+1),  j = 1..k+2,
// B[0] multiplies the most recent solution.
std::array<double, 16> aspc_coefficients() {
  const int k = 14;
  auto binom = [](int nn, int r) {
    double v = 1;
    for (int q = 1; q <= r; ++q) v = v * (nn - r + q) / q;
    return v;
  };
  std::array<double, 16> B{};
  const double den = binom(2 * k + 2, k + 1);
  for (int j = 1; j <= k + 2; ++j) B[j - 1] = ((j & 1) ? 1.0 : -1.0) * j * binom(2 * k + 4, k + 2 - j) / den;
  return B;
}

int induce_pcg(Ctx& c, int* iters, double* rms) {
  return c.prec == 32 ? induce_t<float>(c, iters, rms) : induce_t<double>(c, iters, rms);
}

}  // namespace t9
