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
// (guess.hip) the ASPC guess is corrected by the network, and every solve
// starts from the better of the two by RMS residual: the ASPC guess's mutual
// field is computed beside the direct field (it needs only the stored
// solutions), both residuals are formed, and a device kernel keeps the
// learned guess only when its residual is not larger (else the ASPC guess and
// its residual are copied in) before the first CG iteration.
#include <array>

#include <algorithm>

#include "ctx.hpp"

namespace t9 {
namespace {

constexpr int kB = 256;
// device scalar slots
// slots 0..3: r.z sums SUM[0] (d, p), SUM[1] (d, p) by iteration parity
enum { S_EPS = 8, S_EPS0 = 9, S_PICK = 10, S_EPSA = 11, S_DONE = kDoneSlot, S_NSLOT = 16 };
constexpr int kMaxUalt = 16;

struct Coef16 {
  double c[kMaxUalt];
};

// u0 = sum_k c_k hist_k (caller-order history, slot k: [2][n][3]); lanes in
// caller order so the 16 history rows are read contiguously
template <class real>
__global__ __launch_bounds__(kB) void k_aspc_guess(int n, const int* __restrict__ iperm, Coef16 cf,
                                                  const real* __restrict__ hist, real* __restrict__ ud,
                                                  real* __restrict__ up, real* __restrict__ keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = iperm[i];
  const size_t slot = 6 * (size_t)n;
  // one pass over the history: each slot's three d and three p components
  // of the atom are read together (the per-component sums keep slot order)
  double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
  for (int k = 0; k < kMaxUalt; ++k) {
    const real* hd = hist + k * slot + 3 * (size_t)i;
    const real* hp = hd + 3 * (size_t)n;
    const double c = cf.c[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      a[d] += c * (double)hd[d];
      b[d] += c * (double)hp[d];
    }
  }
  for (int d = 0; d < 3; ++d) {
    ud[3 * s + d] = (real)a[d];
    up[3 * s + d] = (real)b[d];
    if (keep) {  // a copy of the ASPC guess (d, then p) for the learned correction's fallback
      keep[3 * s + d] = (real)a[d];
      keep[3 * (size_t)n + 3 * s + d] = (real)b[d];
    }
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

// udir = alpha E (d, p) and the initial guess: a copy of `from` ([2][n][3]:
// d, then p) when given, else udir (guess != 0) or zero
template <class real>
__global__ __launch_bounds__(kB) void k_udir(int n, const real* __restrict__ pol, const real* __restrict__ fd,
                                            const real* __restrict__ fp, real* __restrict__ udir,
                                            real* __restrict__ udirp, real* __restrict__ ud, real* __restrict__ up,
                                            int guess, const real* __restrict__ from) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const real a = pol[s];
  for (int d = 0; d < 3; ++d) {
    const real x = a * fd[3 * s + d], y = a * fp[3 * s + d];
    udir[3 * s + d] = x;
    udirp[3 * s + d] = y;
    ud[3 * s + d] = from ? from[3 * s + d] : (guess ? x : 0);
    up[3 * s + d] = from ? from[3 * (size_t)n + 3 * s + d] : (guess ? y : 0);
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

// ---- CG scalars
//
// Dot products are summed per workgroup (partials[block][4]); the kernel that
// needs a scalar sums all partials itself, in every workgroup, in the same
// fixed order (thread t takes blocks t, t + 256, ..., then a fixed-order
// block reduction), so every workgroup holds the same bits without a fence,
// an atomic or an extra launch.  The running r.z sum of iteration `it` lives
// in slots SUM[it & 1] (d, p); iteration it writes SUM[(it + 1) & 1], so no
// workgroup reads a slot that another workgroup of the same kernel writes.

// Fixed-order sum over a 256-thread block of NV values per thread: wave
// butterflies, lane 0 of each wave to LDS, the four wave sums in order.
// Every thread gets the same result.
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV]) {
  __shared__ double sh[4][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double x = wave_bfly_sum(v[q]);
    if (lane == 0) sh[w][q] = x;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = (sh[0][q] + sh[1][q]) + (sh[2][q] + sh[3][q]);
  __syncthreads();
}

// this workgroup's sums of NV per-thread values -> partials[blockIdx.x][q]
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* __restrict__ part) {
  block_sums<NV>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) part[blockIdx.x * 4 + q] = v[q];
}

// totals of the NV partial columns over nblk workgroups
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nblk, double (&t)[NV]) {
#pragma unroll
  for (int q = 0; q < NV; ++q) t[q] = 0;
  for (int b = threadIdx.x; b < nblk; b += kB)
#pragma unroll
    for (int q = 0; q < NV; ++q) t[q] += part[b * 4 + q];
  block_sums<NV>(t);
}

// The same totals from partials written per 64 elements (one wavefront's
// butterfly sum each, by k_precond): each group of 4 is combined as
// (w0 + w1) + (w2 + w3), the sum block_sums forms in a 256-thread workgroup,
// so the totals are those of k_dots' partials, bit for bit.
template <int NV>
__device__ __forceinline__ void reduce_partials64(const double* __restrict__ part, int nblk64, double (&t)[NV]) {
#pragma unroll
  for (int q = 0; q < NV; ++q) t[q] = 0;
  const int ng = (nblk64 + 3) / 4;
  for (int b = threadIdx.x; b < ng; b += kB)
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) w[u] = 4 * b + u < nblk64 ? part[(4 * b + u) * 4 + q] : 0.0;
      t[q] += (w[0] + w[1]) + (w[2] + w[3]);
    }
  block_sums<NV>(t);
}

__device__ __forceinline__ double sum_slot(int it, int dp) { return 2 * (it & 1) + dp; }

// NV (2 or 4) dot products a_q . b_q -> partials
template <class real, int NV>
__global__ __launch_bounds__(kB) void k_dots(int n, const real* __restrict__ a1, const real* __restrict__ b1,
                                            const real* __restrict__ a2, const real* __restrict__ b2,
                                            const real* __restrict__ a3, const real* __restrict__ b3,
                                            const real* __restrict__ a4, const real* __restrict__ b4,
                                            double* __restrict__ part, Gate gate = Gate()) {
  if (gate_closed(gate)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const real* A[4] = {a1, a2, a3, a4};
  const real* Bv[4] = {b1, b2, b3, b4};
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = s < n ? dot3(A[q], Bv[q], s) : 0.0;
  write_partials<NV>(v, part);
}

// Start of a solve, one workgroup.  mode 1: partials hold r.z (d, p) and r.r
// (d, p), per 64 elements (k_precond): SUM[parity] = r.z, initial RMS residual
// from r.r; mode 0: partials hold r.r only (k_dots, the residual of a trial
// guess).
__global__ __launch_bounds__(kB) void k_cg_start(int nblk, int mode, int parity, int n, double debye,
                                                const double* __restrict__ part, double* __restrict__ sc,
                                                double* __restrict__ trace) {
  double t[4];
  if (mode) reduce_partials64<4>(part, nblk, t);
  else reduce_partials<4>(part, nblk, t);
  if (threadIdx.x != 0) return;
  const double rr_d = mode ? t[2] : t[0], rr_p = mode ? t[3] : t[1];
  const double eps0 = debye * sqrt(fmax(rr_d, rr_p) / n);
  sc[S_EPS0] = eps0;
  trace[0] = eps0;
  if (mode) {
    sc[2 * parity] = t[0];
    sc[2 * parity + 1] = t[1];
  }
}

// vec = conj / alpha - field (reference pcgP1), partials of conj.vec
template <class real>
__global__ __launch_bounds__(kB) void k_p1_dots(int n, const real* __restrict__ polinv, const real* __restrict__ cd,
                                               const real* __restrict__ cp, const real* __restrict__ fd,
                                               const real* __restrict__ fp, real* __restrict__ vd,
                                               real* __restrict__ vp, double* __restrict__ part, Gate gate) {
  if (gate_closed(gate)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0, 0};
  if (s < n) {
    const real pi = polinv[s];
    for (int d = 0; d < 3; ++d) {
      const real a = pi * cd[3 * s + d] - fd[3 * s + d];
      const real b = pi * cp[3 * s + d] - fp[3 * s + d];
      vd[3 * s + d] = a;
      vp[3 * s + d] = b;
      v[0] += (double)cd[3 * s + d] * (double)a;
      v[1] += (double)cp[3 * s + d] * (double)b;
    }
  }
  write_partials<2>(v, part);
}

// a = sum / p.vec from the p1 partials; u += a p; r -= a vec; r = 0 on
// unpolarizable sites   (reference pcgP2)
template <class real>
__global__ __launch_bounds__(kB) void k_p2(int n, int nblk, const double* __restrict__ part,
                                          const double* __restrict__ sc, const real* __restrict__ pol,
                                          const real* __restrict__ cd, const real* __restrict__ cp,
                                          const real* __restrict__ vd, const real* __restrict__ vp,
                                          real* __restrict__ ud, real* __restrict__ up, real* __restrict__ rd,
                                          real* __restrict__ rp, Gate gate) {
  if (gate_closed(gate)) return;
  double t[2];
  reduce_partials<2>(part, nblk, t);
  const double sd = sc[(int)sum_slot(gate.iter, 0)], sp = sc[(int)sum_slot(gate.iter, 1)];
  const real a = (real)(t[0] != 0 ? sd / t[0] : 0), ap = (real)(t[1] != 0 ? sp / t[1] : 0);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const bool zero = pol[s] == 0;
  for (int d = 0; d < 3; ++d) {
    const int k = 3 * s + d;
    ud[k] += a * cd[k];
    up[k] += ap * cp[k];
    rd[k] = zero ? (real)0 : rd[k] - a * vd[k];
    rp[k] = zero ? (real)0 : rp[k] - ap * vp[k];
  }
}

// From the partials of r.z (d, p) and r.r (d, p): b = sum1 / sum, the RMS
// residual and the convergence test (src/cu/amoeba/pcg.cu:147-170: eps < tol
// once iter >= iter_min, or iter >= iter_max); p = z + b p (reference
// pcgP3); in the iteration that converged also the peek step u += peek *
// alpha * r (reference pcgPeek).  Workgroup 0 stores sum1 for the next
// iteration, eps, the trace entry and the converged iteration (slot S_DONE).
template <class real>
__global__ __launch_bounds__(kB) void k_p3(int n, int nblk, const double* __restrict__ part, double* __restrict__ sc,
                                          double tol, int iter_min, int iter_max, double debye,
                                          double* __restrict__ trace, const real* __restrict__ zd,
                                          const real* __restrict__ zp, real* __restrict__ cd,
                                          real* __restrict__ cp, real peek, const real* __restrict__ pol,
                                          const real* __restrict__ rd, const real* __restrict__ rp,
                                          real* __restrict__ ud, real* __restrict__ up, Gate gate) {
  if (gate_closed(gate)) return;
  const int it = gate.iter;
  double t[4];
  reduce_partials64<4>(part, nblk, t);  // r.z and r.r, per 64 elements from k_precond
  const double sd = sc[(int)sum_slot(it, 0)], sp = sc[(int)sum_slot(it, 1)];
  const real b = (real)(sd != 0 ? t[0] / sd : 0), bp = (real)(sp != 0 ? t[1] / sp : 0);
  const double eps = debye * sqrt(fmax(t[2], t[3]) / n);
  bool done = eps < tol;
  if (it < iter_min) done = false;
  if (it >= iter_max) done = true;
  // slot S_DONE reads 0, or `it` once workgroup 0 has stored it below
  const bool converged_here = done && (sc[S_DONE] == 0 || sc[S_DONE] == (double)it);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[(int)sum_slot(it + 1, 0)] = t[0];
    sc[(int)sum_slot(it + 1, 1)] = t[1];
    sc[S_EPS] = eps;
    if (it < kTraceCap) trace[it] = eps;
    if (converged_here) sc[S_DONE] = it;
  }
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  for (int d = 0; d < 3; ++d) {
    const int k = 3 * s + d;
    cd[k] = zd[k] + b * cd[k];
    cp[k] = zp[k] + bp * cp[k];
  }
  if (converged_here) {
    const real tp = peek * pol[s];
    for (int d = 0; d < 3; ++d) {
      ud[3 * s + d] += tp * rd[3 * s + d];
      up[3 * s + d] += tp * rp[3 * s + d];
    }
  }
}

__global__ void k_set_scalar(double* sc, int slot, double v) {
  if (threadIdx.x == 0) sc[slot] = v;
}

// The starting point of a learned solve: the learned guess (ud, up; residual
// rd, rp) unless its RMS residual exceeds the ASPC guess's (ua; residual ra,
// each [2][n][3]: d, then p) or is not finite; then the ASPC guess and its
// residual are copied in.  The partials hold r.r of rd, rp, ra_d, ra_p (k_dots);
// every workgroup sums them in the same order, so all make the same choice.
// Workgroup 0 records it (slot S_PICK: 2 learned, 1 ASPC) and the ASPC
// residual (slot S_EPSA, Debye).
template <class real>
__global__ __launch_bounds__(kB) void k_guess_select(int n, int nblk, double debye, const double* __restrict__ part,
                                                    const real* __restrict__ ua, const real* __restrict__ ra,
                                                    real* __restrict__ ud, real* __restrict__ up,
                                                    real* __restrict__ rd, real* __restrict__ rp,
                                                    double* __restrict__ sc) {
  double t[4];
  reduce_partials<4>(part, nblk, t);
  const double el = fmax(t[0], t[1]), ea = fmax(t[2], t[3]);
  const bool use_aspc = !(el <= ea);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_PICK] = use_aspc ? 1.0 : 2.0;
    sc[S_EPSA] = debye * sqrt(ea / n);
  }
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (!use_aspc || s >= n) return;
  const size_t h = 3 * (size_t)n;
  for (int d = 0; d < 3; ++d) {
    const int k = 3 * s + d;
    ud[k] = ua[k];
    up[k] = ua[h + k];
    rd[k] = ra[k];
    rp[k] = ra[h + k];
  }
}

template <class real>
int induce_t(Ctx& c, int* iters, double* rms) {
  const int n = c.n;
  const int nblk = ceil_div(n, kB), nblk64 = ceil_div(n, 64);
  double* part64 = c.partials64.as<double>();
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
  int iter = 0;
  // slot S_DONE = 0: no iteration has converged yet
  auto reset_done = [&]() {
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, st, sc, (int)S_DONE, 0.0);
    T9_LAUNCH_CHECK();
  };

  // ASPC once 16 solutions are stored (src/cu/amoeba/pcg.cu:30-33)
  const bool aspc = c.opt.predictor == "ASPC";
  const bool predict = aspc && c.nualt >= kMaxUalt;
  // learned-guess features need 16 stored fields as well
  const bool feats = predict && c.nealt >= kMaxUalt && guess_wants_features(c);
  const bool learned = feats && c.guess_on && c.guess_ready;
  const auto B = aspc_coefficients();
  Coef16 cf;
  if (predict)
    for (int k = 0; k < kMaxUalt; ++k) cf.c[k] = B[((c.nualt - 1 - k) % kMaxUalt + kMaxUalt) % kMaxUalt];
  real* ua = c.uaspc.as<real>();
  real* ra = learned ? c.raspc.as<real>() : nullptr;
  if (learned) {
    // the ASPC guess first: its mutual field (into ra) is computed beside
    // the direct field
    hipLaunchKernelGGL(k_aspc_guess<real>, G, Bk, 0, st, n, c.iperm.as<int>(), cf, c.ualt.as<real>(), ua,
                       ua + 3 * (size_t)n, (real*)nullptr);
    T9_LAUNCH_CHECK();
    dfield_all(c, ua, ra);
  } else {
    dfield_all(c);
  }
  // the learned solve starts from the ASPC guess (computed above), the
  // corrections are added below
  hipLaunchKernelGGL(k_udir<real>, G, Bk, 0, st, n, c.polarity.as<real>(), field, fieldp, c.udir.as<real>(),
                     c.udirp.as<real>(), ud, up, (guess || predict) ? 1 : 0, learned ? (const real*)ua : nullptr);
  T9_LAUNCH_CHECK();
  auto slot_of_age = [](int count, int age) { return ((count - 1 - age) % kMaxUalt + kMaxUalt) % kMaxUalt; };
  // r0 of the guess in ud/up
  auto residual = [&]() {
    ufield_all(c, ud, up, rd, rp);
    hipLaunchKernelGGL(k_rsd_guess<real>, G, Bk, 0, st, n, c.polarity.as<real>(), c.polinv.as<real>(),
                       c.udir.as<real>(), c.udirp.as<real>(), ud, up, rd, rp);
    T9_LAUNCH_CHECK();
  };
  double* trace = c.pcg_trace.as<double>();
  c.last_predictor = predict ? 1 : 0;
  if (predict) {
    if (!learned) {
      hipLaunchKernelGGL(k_aspc_guess<real>, G, Bk, 0, st, n, c.iperm.as<int>(), cf, c.ualt.as<real>(), ud, up,
                         feats ? ua : nullptr);
      T9_LAUNCH_CHECK();
    }
    if (feats) {
      double ce[kMaxUalt], cu[2];
      for (int k = 0; k < kMaxUalt; ++k) ce[k] = B[((c.nealt - 1 - k) % kMaxUalt + kMaxUalt) % kMaxUalt];
      cu[0] = slot_of_age(c.nualt, 0);
      cu[1] = slot_of_age(c.nualt, 1);
      guess_features(c, ce, cu);
    }
  }
  if (aspc) guess_save_fields(c);
  if (learned) {
    // the ASPC residual alpha^-1 (udir - ua) + T ua, in place of T ua; the
    // learned guess and its residual; the better of the two
    hipLaunchKernelGGL(k_rsd_guess<real>, G, Bk, 0, st, n, c.polarity.as<real>(), c.polinv.as<real>(),
                       c.udir.as<real>(), c.udirp.as<real>(), ua, ua + 3 * (size_t)n, ra, ra + 3 * (size_t)n);
    T9_LAUNCH_CHECK();
    guess_apply(c, ud, up);
    residual();
    hipLaunchKernelGGL((k_dots<real, 4>), G, Bk, 0, st, n, rd, rd, rp, rp, ra, ra, ra + 3 * (size_t)n,
                       ra + 3 * (size_t)n, part, Gate());
    hipLaunchKernelGGL(k_guess_select<real>, G, Bk, 0, st, n, nblk, debye, part, ua, ra, ud, up, rd, rp, sc);
    T9_LAUNCH_CHECK();
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
  // first preconditioned search direction; r.z of the first iteration and
  // the residual norm of the initial guess (Debye, read back with the first
  // iteration)
  auto start_cg = [&]() {
    // z, the first search direction p = z, and r.z, r.r per 64 elements
    precond_apply(c, rd, rp, zd, zp, cd, cp, part64);
    hipLaunchKernelGGL(k_cg_start, dim3(1), Bk, 0, st, nblk64, 1, 1, n, debye, part64, sc, trace);
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
                       part, g);
    hipLaunchKernelGGL(k_p2<real>, G, Bk, 0, st, n, nblk, part, sc, c.polarity.as<real>(), cd, cp, vd, vp, ud, up,
                       rd, rp, g);
    T9_LAUNCH_CHECK();
    precond_apply(c, rd, rp, zd, zp, nullptr, nullptr, part64);  // z, and r.z, r.r per 64 elements
    hipLaunchKernelGGL(k_p3<real>, G, Bk, 0, st, n, nblk64, part64, sc, c.opt.polar_eps, miniter,
                       c.opt.polar_iter, debye, trace, zd, zp, cd, cp, (real)c.opt.pcg_peek, c.polarity.as<real>(),
                       rd, rp, ud, up, g);
    T9_LAUNCH_CHECK();
    c.gate = Gate();
  };
  // The host enqueues as many iterations as the previous solve needed, then
  // reads the scalars back once; after that one iteration per readback.
  double eps = 100;
  int enq = 0;
  for (;;) {
    const int upto = std::max(enq + 1, c.pcg_planned);
    while (enq < upto) enqueue(++enq);
    T9_HIP(hipMemcpyAsync(c.h_scalars, sc, sizeof(double) * S_NSLOT, hipMemcpyDeviceToHost, st));
    mark_readback(c);
    wait_marked(c);
    const double done_at = c.h_scalars[S_DONE];
    if (done_at != 0) {
      iter = (int)done_at;
      eps = c.h_scalars[S_EPS];
      break;
    }
  }
  if (learned) {  // the choice of k_guess_select
    c.last_predictor = (int)c.h_scalars[S_PICK];
    if (c.last_predictor == 2) ++c.guess_used;
    else ++c.guess_fallback;
  }
  c.pcg_planned = iter;
  c.pcg_trace_len = std::min(iter + 1, kTraceCap);
  c.uind_valid = true;
  c.last_guess_rms = c.h_scalars[S_EPS0];
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
