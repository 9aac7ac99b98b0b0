// Multipole + polarization energy, gradient and virial for gfx950 (replaces
// the reference's emplar path: src/amoeba/emplar.cpp:10-28,
// src/cu/amoeba/emplar.cu:77-605, src/cu/hippo/empole.cu:195-370,
// src/cu/epolarrecip.cu:9-530, src/cu/amoeba/epolar.cu:15-30,
// src/cu/amoeba/torque.cu:9-380).
//
// Formulation.  For lane i and neighbour k (R = r_i - r_k) the pair terms are
// written with the radial derivative series B_n (B_{n+1} = -(1/r) dB_n/dr):
// Ewald b_n for the multipole energy, b_n - (1 - lambda_n) rr_n for the
// Thole-damped polarization terms.  The lane evaluates
//   * the multipole pair energy E(M_i, M_k) and its R-gradient,
//   * the potential derivatives (grad phi, grad grad phi) of M_k at r_i,
//     which are dE/dM_i and later give the torque,
//   * the R-gradient of the polarization pair energy at fixed dipoles,
//     1/2 [u_i.grad grad phi_Mk + M_i.grad phi_Uk + ud_i.grad grad phi_upk
//          + up_i.grad grad phi_udk]  (u = ud + up),
//     and 1/2 of the potential derivatives of U_k for the torque,
// with E_pol = -1/2 f sum ud.E_p taken as a dot product after the solve.
// Every ordered pair is visited from both atoms (full lists), so each lane
// owns its atom's gradient, torque terms, half the pair energy and half the
// pair virial, accumulated in FP64 registers: no atomics, bit-reproducible.
// Pairs with scale factors are skipped in the list loop; the same lane then
// walks its scaled partners and evaluates each whole pair term (unscaled
// Ewald part plus the (s - 1) non-Ewald correction, kernel lambda_n rr_n) in
// FP64, whatever the working precision.
//
// Reciprocal space: the permanent-multipole grid gives 20 potential
// derivatives per atom (energy 1/2 M.phi, gradient M.grad phi, torque
// terms); the induced-dipole grid (ud real, up imaginary) gives the
// polarization gradient 1/2 [u.grad grad phi_M + M.grad phi_u + ud.grad grad
// phi_up + up.grad grad phi_ud] and torque terms.  The virial is the strain
// derivative: convolution sums over k (cross terms of the polarization via
// two packed grids and the polarization identity) plus the change of the
// Cartesian moments' coupling as k deforms.
//
// Torques: tau = d x grad phi + 2 axial(Q grad grad phi) is spread over the
// frame atoms with the derivative of each frame's rotation (see k_torque);
// each atom then gathers the contributions that name it, in a fixed order.
#include <algorithm>

#include "bspline.hpp"
#include "chgpen.hpp"
#include "ctx.hpp"

namespace t9 {

namespace {

constexpr double kCoul = 332.0637133;
constexpr double kSqrtPiD = 1.77245385090551602730;

template <class real>
using R4 = typename RealTraits<real>::real4;

constexpr int kRow = kEnergyRow;
// the fused energy kernel holds more registers per lane: four waves per 64
// atoms balance latency hiding against occupancy
constexpr int kEmplarParts = 4;

template <class real>
struct Mp {
  real q, d[3], Q[6];  // Q: xx yy zz xy xz yz (traceless)
};

template <class real>
__device__ __forceinline__ Mp<real> load_mp(const real* __restrict__ p) {
  Mp<real> m;
  m.q = p[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) m.d[k] = p[1 + k];
#pragma unroll
  for (int k = 0; k < 6; ++k) m.Q[k] = p[4 + k];
  return m;
}

template <class real>
__device__ __forceinline__ real dot3(const real* a, const real* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <class real>
__device__ __forceinline__ void qv(const real* Q, const real* v, real* o) {
  o[0] = Q[0] * v[0] + Q[3] * v[1] + Q[4] * v[2];
  o[1] = Q[3] * v[0] + Q[1] * v[1] + Q[5] * v[2];
  o[2] = Q[4] * v[0] + Q[5] * v[1] + Q[2] * v[2];
}

template <class real>
struct PairOut {
  real em;     // multipole pair energy
  real de[3];  // dE/dR of the pair (multipole + polarization)
  real g[3];   // grad phi at i (multipole source + 1/2 dipole source)
  real h[6];   // grad grad phi at i
};

// Pair terms for field site i and source k, R = r_i - r_k.  BM[0..5]:
// multipole kernel; BT[1..4]: polarization kernel; uie / uke: dipoles of i
// and k entering the dipole-multipole terms; ud/up: dipoles entering the
// dipole-dipole terms with weight wu.
template <class real, bool DM, bool DP>
__device__ __forceinline__ void pair_terms(const real* R, const real* BM, const real* BT, const Mp<real>& mi,
                                           const Mp<real>& mk, const real* uie, const real* uke, const real* udi,
                                           const real* upi, const real* udk, const real* upk, real wu,
                                           PairOut<real>& o) {
  real QkR[3], QiR[3];
  qv(mk.Q, R, QkR);
  qv(mi.Q, R, QiR);
  const real dkR = dot3(mk.d, R), diR = dot3(mi.d, R);
  const real RQkR = dot3(R, QkR), RQiR = dot3(R, QiR);
#pragma unroll
  for (int a = 0; a < 3; ++a) o.de[a] = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) o.g[a] = 0;
#pragma unroll
  for (int a = 0; a < 6; ++a) o.h[a] = 0;
  o.em = 0;
  if (DM) {
    const real didk = dot3(mi.d, mk.d), diQkR = dot3(mi.d, QkR), dkQiR = dot3(mk.d, QiR);
    const real QiRQkR = dot3(QiR, QkR);
    const real QiQk = mi.Q[0] * mk.Q[0] + mi.Q[1] * mk.Q[1] + mi.Q[2] * mk.Q[2] +
                      2 * (mi.Q[3] * mk.Q[3] + mi.Q[4] * mk.Q[4] + mi.Q[5] * mk.Q[5]);
    const real G0 = mi.q * mk.q;
    const real G1 = mi.q * dkR - mk.q * diR + didk;
    const real G2 = mi.q * RQkR + mk.q * RQiR - diR * dkR + 2 * (diQkR - dkQiR + QiQk);
    const real G3 = RQiR * dkR - diR * RQkR - 4 * QiRQkR;
    const real G4 = RQiR * RQkR;
    o.em = BM[0] * G0 + BM[1] * G1 + BM[2] * G2 + BM[3] * G3 + BM[4] * G4;
    const real cR = -(BM[1] * G0 + BM[2] * G1 + BM[3] * G2 + BM[4] * G3 + BM[5] * G4);
    real Qkdi[3], Qidk[3], QiQkR[3], QkQiR[3];
    qv(mk.Q, mi.d, Qkdi);
    qv(mi.Q, mk.d, Qidk);
    qv(mi.Q, QkR, QiQkR);
    qv(mk.Q, QiR, QkQiR);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      o.de[a] = cR * R[a] + BM[1] * (mi.q * mk.d[a] - mk.q * mi.d[a]) +
                BM[2] * (2 * (mi.q * QkR[a] + mk.q * QiR[a] + Qkdi[a] - Qidk[a]) - dkR * mi.d[a] - diR * mk.d[a]) +
                BM[3] * (RQiR * mk.d[a] - RQkR * mi.d[a] + 2 * (dkR * QiR[a] - diR * QkR[a]) -
                         4 * (QiQkR[a] + QkQiR[a])) +
                BM[4] * 2 * (RQkR * QiR[a] + RQiR * QkR[a]);
    // potential derivatives of M_k at r_i
    const real g1 = -(mk.q * BM[1] + dkR * BM[2] + RQkR * BM[3]);
    const real g2 = mk.q * BM[2] + dkR * BM[3] + RQkR * BM[4];
#pragma unroll
    for (int a = 0; a < 3; ++a) o.g[a] = g1 * R[a] + BM[1] * mk.d[a] + 2 * BM[2] * QkR[a];
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      const int a = ia[t], b = ib[t];
      o.h[t] = g2 * R[a] * R[b] - BM[2] * (R[a] * mk.d[b] + mk.d[a] * R[b]) -
               2 * BM[3] * (R[a] * QkR[b] + QkR[a] * R[b]) + 2 * BM[2] * mk.Q[t];
      if (t < 3) o.h[t] += g1;
    }
  }
  if (DP) {
    // P1: u_i . grad grad phi_Mk
    const real g1 = -(mk.q * BT[1] + dkR * BT[2] + RQkR * BT[3]);
    const real g2 = mk.q * BT[2] + dkR * BT[3] + RQkR * BT[4];
    const real uiR = dot3(uie, R), dkui = dot3(mk.d, uie), QkRui = dot3(QkR, uie);
    real Qkui[3];
    qv(mk.Q, uie, Qkui);
    // P2: M_i . grad phi_Uk (orders 1..3)
    const real ukR = dot3(uke, R), diuk = dot3(mi.d, uke), QiRuk = dot3(QiR, uke);
    real Qiuk[3];
    qv(mi.Q, uke, Qiuk);
    const real c2 = BT[2] * (mi.q * ukR + diuk) - BT[3] * (diR * ukR + 2 * QiRuk) + BT[4] * RQiR * ukR;
    // P3: ud_i . grad grad phi_up,k + up_i . grad grad phi_ud,k
    const real udiR = dot3(udi, R), upiR = dot3(upi, R), udkR = dot3(udk, R), upkR = dot3(upk, R);
    const real c3 = wu * (BT[3] * (upkR * udiR + udkR * upiR) - BT[2] * (dot3(udi, upk) + dot3(upi, udk)));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const real p1 = g2 * R[a] * uiR + g1 * uie[a] - BT[2] * (R[a] * dkui + mk.d[a] * uiR) -
                      2 * BT[3] * (R[a] * QkRui + QkR[a] * uiR) + 2 * BT[2] * Qkui[a];
      const real p2 = -c2 * R[a] + BT[1] * mi.q * uke[a] - BT[2] * (ukR * mi.d[a] + diR * uke[a] + 2 * Qiuk[a]) +
                      BT[3] * (2 * ukR * QiR[a] + RQiR * uke[a]);
      const real p3 = c3 * R[a] - wu * BT[2] * (upk[a] * udiR + udk[a] * upiR + upkR * udi[a] + udkR * upi[a]);
      o.de[a] += (real)0.5 * (p1 + p2 + p3);
    }
    // 1/2 potential derivatives of U_k at r_i
#pragma unroll
    for (int a = 0; a < 3; ++a) o.g[a] += (real)0.5 * (BT[1] * uke[a] - ukR * BT[2] * R[a]);
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      const int a = ia[t], b = ib[t];
      real v = ukR * BT[3] * R[a] * R[b] - BT[2] * (R[a] * uke[b] + uke[a] * R[b]);
      if (t < 3) v -= ukR * BT[2];
      o.h[t] += (real)0.5 * v;
    }
  }
}

template <class real>
struct EwK6 {
  real a;
  real e[6];  // e[n] = 2^n a^(2n-1) / sqrt(pi), n = 1..5
};

// Thole lambda_3,5,7,9 (damp.h:8-60 series; lambda_9 completes the
// derivative series needed by the dipole-quadrupole gradient).
template <class real>
__device__ __forceinline__ void thole4(real r, real rpdik, real pga, real* lam) {
  const real ratio = r * rpdik;
  const real a = rpdik == 0 ? (real)1e16 : pga * ratio * ratio * ratio;
  const real e = t9_exp_damp(-a);
  lam[1] = 1 - e;
  lam[2] = 1 - e * (1 + a);
  lam[3] = 1 - e * (1 + a + (real)0.6 * a * a);
  lam[4] = 1 - e * (1 + a + a * a * ((real)(18.0 / 35.0) + (real)(9.0 / 35.0) * a));
}

template <class real>
struct Acc {
  double e, gr[3], tg[3], th[6], v[6];
  __device__ __forceinline__ void pack(double* o) const {
    o[0] = e;
    for (int k = 0; k < 3; ++k) o[1 + k] = gr[k];
    for (int k = 0; k < 3; ++k) o[4 + k] = tg[k];
    for (int k = 0; k < 6; ++k) o[7 + k] = th[k];
    for (int k = 0; k < 6; ++k) o[13 + k] = v[k];
  }
  __device__ __forceinline__ void unpack(const double* o) {
    e = o[0];
    for (int k = 0; k < 3; ++k) gr[k] = o[1 + k];
    for (int k = 0; k < 3; ++k) tg[k] = o[4 + k];
    for (int k = 0; k < 6; ++k) th[k] = o[7 + k];
    for (int k = 0; k < 6; ++k) v[k] = o[13 + k];
  }
  __device__ void zero() {
    e = 0;
    for (int k = 0; k < 3; ++k) gr[k] = tg[k] = 0;
    for (int k = 0; k < 6; ++k) th[k] = v[k] = 0;
  }
  template <class U>
  __device__ __forceinline__ void add(const PairOut<U>& o, const U* R) {
    e += (double)o.em;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gr[k] += (double)o.de[k];
      tg[k] += (double)o.g[k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) th[k] += (double)o.h[k];
    v[0] += (double)(o.de[0] * R[0]);
    v[1] += (double)(o.de[1] * R[1]);
    v[2] += (double)(o.de[2] * R[2]);
    v[3] += (double)((U)0.5 * (o.de[0] * R[1] + o.de[1] * R[0]));
    v[4] += (double)((U)0.5 * (o.de[0] * R[2] + o.de[2] * R[0]));
    v[5] += (double)((U)0.5 * (o.de[1] * R[2] + o.de[2] * R[1]));
  }
};

template <class real>
struct EmplarArgs {
  int n;
  const int* ncut;
  const long long* woff;
  const uint32_t* ecut;
  const PosRec<real>* pos;
  const real* rpole;
  const real* ud;
  const real* up;
  const int* jpolar;
  const real* thlval;
  int nj;
  int uth;
  real pga_u;
  const int* perm;
  const int* iperm;
  const int* excl_off;
  const int* excl_j;
  const int* excl_code;
  ScaleTable sc;
  BoxT<real> box;
  real cut2;
  EwK6<real> ek;
  EwK6<double> ek64;  // the same constants for the FP64 scaled-pair terms
  double* grad;  // [n][3]
  double* tq;    // [n][9]
  double* rows;  // [gridDim][kRow]
};

// MODE bit 0: multipole terms, bit 1: polarization terms.
template <class real, int EWALD, int MODE>
__global__ __launch_bounds__(64 * kEmplarParts) void k_emplar(EmplarArgs<real> A) {
  constexpr bool DM = MODE & 1, DP = (MODE & 2) != 0;
  __shared__ double sh[parts_lds<kEmplarParts, 19>()];
  const int part = threadIdx.x >> 6;
  const int blk = xcd_block();
  const int s = blk * 64 + (threadIdx.x & 63);
  Acc<real> acc;
  acc.zero();
  if (s < A.n) {
    const PosRec<real> pi = A.pos[s];
    const Mp<real> mi = load_mp(A.rpole + 10 * (size_t)s);
    real udi[3] = {0, 0, 0}, upi[3] = {0, 0, 0}, ui[3] = {0, 0, 0};
    if (DP)
      for (int k = 0; k < 3; ++k) {
        udi[k] = A.ud[3 * s + k];
        upi[k] = A.up[3 * s + k];
        ui[k] = udi[k] + upi[k];
      }
    const int jpi = A.uth ? 0 : A.jpolar[s] * A.nj;
    const int len = A.ncut[s];
    const uint32_t* e = A.ecut + A.woff[s >> 6] + (s & 63);
    uint32_t nxt = part < len ? e[(long long)part * 64] : 0;
    for (int kk = part; kk < len; kk += kEmplarParts) {
      const uint32_t ent = nxt;
      if (kk + kEmplarParts < len) nxt = e[(long long)(kk + kEmplarParts) * 64];
      if (ent & (kMaxCodes - 1)) continue;  // scaled pairs: k_emplar_scaled
      const int j = (int)(ent >> kCodeBits);
      const PosRec<real> pj = A.pos[j];
      real R[3];
      pair_disp(pj, pi, A.box, R[0], R[1], R[2]);
      const real r2 = dot3(R, R);
      const real rinv = t9_rsqrt(r2);
      const real r = r2 * rinv;
      const real rr2 = rinv * rinv;
      real rr[6];
      rr[0] = rinv;
#pragma unroll
      for (int q = 1; q < 6; ++q) rr[q] = rr[q - 1] * (real)(2 * q - 1) * rr2;
      real bn[6];
      if (EWALD) {
        real ex;
        bn[0] = t9_erfc_pair(A.ek.a * r, ex) * rinv;
#pragma unroll
        for (int q = 1; q < 6; ++q) bn[q] = ((real)(2 * q - 1) * bn[q - 1] + A.ek.e[q] * ex) * rr2;
      } else {
#pragma unroll
        for (int q = 0; q < 6; ++q) bn[q] = rr[q];
      }
      const Mp<real> mk = load_mp(A.rpole + 10 * (size_t)j);
      real BT[5] = {0, 0, 0, 0, 0};
      real udk[3] = {0, 0, 0}, upk[3] = {0, 0, 0}, uk[3] = {0, 0, 0};
      if (DP) {
        const real pga = A.uth ? A.pga_u : A.thlval[jpi + A.jpolar[j]];
        real lam[5];
        thole4(r, pi.w * pj.w, pga, lam);
#pragma unroll
        for (int q = 1; q < 5; ++q) BT[q] = bn[q] - (1 - lam[q]) * rr[q];
        for (int q = 0; q < 3; ++q) {
          udk[q] = A.ud[3 * j + q];
          upk[q] = A.up[3 * j + q];
          uk[q] = udk[q] + upk[q];
        }
      }
      PairOut<real> o;
      pair_terms<real, DM, DP>(R, bn, BT, mi, mk, ui, uk, udi, upi, udk, upk, (real)1, o);
      acc.add(o, R);
    }
  }
  double av[19];
  acc.pack(av);
  parts_reduce<kEmplarParts, 19>(av, sh);
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (part == 0) {
    acc.unpack(av);
    if (s < A.n) {
      const double f = kCoul;
      for (int k = 0; k < 3; ++k) A.grad[3 * (size_t)s + k] = f * acc.gr[k];
      for (int k = 0; k < 3; ++k) A.tq[9 * (size_t)s + k] = f * acc.tg[k];
      for (int k = 0; k < 6; ++k) A.tq[9 * (size_t)s + 3 + k] = f * acc.th[k];
    }
    red[EP_MREAL] = 0.5 * kCoul * acc.e;
    for (int k = 0; k < 6; ++k) red[EP_VXX + k] = 0.5 * kCoul * acc.v[k];
  }
  block_row<EP_N>(red, A.rows + (size_t)blk * kRow);
}

// ------------------------------------------------------------ HIPPO

// Charge penetration (HIPPO; replaces the multipole pair of
// src/cu/hippo/empoleChgpen_cu1.cc with include/seq/pair_mpole_chgpen.h and
// include/seq/damp_hippo.h).  A monopole is a point core charge Z plus a
// valence charge q - Z spread as a Slater density of exponent alpha, which
// also carries the dipole and quadrupole.  A pair is four charge-distribution
// pairs, each the plain multipole contraction with its own radial series:
// core-core undamped, core of i in the valence of k (k's damping), valence of
// i at the core of k (i's damping), valence-valence; scale factor s and Ewald
// enter as B_n = b_n - (1 - s lambda_n) rr_n (s lambda_n rr_n without Ewald).
// Only the valence parts have orientation, so only they give torque terms.
// Polarization (src/cu/hippo/epolarChgpen_cu1.cc, pair_polar_chgpen.h): an
// induced dipole is a point probe — u_i sees the core of k undamped and the
// valence multipole of k with k's damping, the multipole of i is seen by u_k
// with i's damping (p scale, which HIPPO also uses as the d scale), and the
// dipole-dipole term has the valence-valence damping (u scale); the terms
// are those of k_emplar's polarization gradient with these kernels.
//
// Damping factors: chgpen_lambda (chgpen.hpp), in FP64 whatever the working
// precision.

// kEmplarParts waves per 64 atoms over the in-cutoff list including the
// scaled pairs (the entry's class gives the scales), both sides of every
// pair, FP64 arithmetic; the outputs of k_emplar (lane-owned gradient and
// torque terms, half pair energy and virial).  MODE bit 0: multipole terms,
// bit 1: polarization gradient.
template <class real, int EWALD, int MODE>
__global__ __launch_bounds__(64 * kEmplarParts) void k_empole_chgpen(EmplarArgs<real> A,
                                                                     const double* __restrict__ cp) {
  constexpr bool DM = MODE & 1, DP = (MODE & 2) != 0;
  __shared__ double sh[parts_lds<kEmplarParts, 19>()];
  const int part = threadIdx.x >> 6;
  const int blk = xcd_block();
  const int s = blk * 64 + (threadIdx.x & 63);
  Acc<double> acc;
  acc.zero();
  if (s < A.n) {
    const PosRec<real> pi = A.pos[s];
    const int io = A.perm[s];
    const double alpha_i = cp[3 * io + 2];
    Mp<double> ci = {}, vi;
    ci.q = cp[3 * io];
    vi.q = cp[3 * io + 1];
    for (int t = 0; t < 3; ++t) vi.d[t] = A.rpole[10 * (size_t)s + 1 + t];
    for (int t = 0; t < 6; ++t) vi.Q[t] = A.rpole[10 * (size_t)s + 4 + t];
    const int len = A.ncut[s];
    const uint32_t* e = A.ecut + A.woff[s >> 6] + (s & 63);
    const double z3[3] = {0, 0, 0};
    double udi[3] = {0, 0, 0}, upi[3] = {0, 0, 0}, ui[3] = {0, 0, 0};
    if (DP)
      for (int t = 0; t < 3; ++t) {
        udi[t] = A.ud[3 * s + t];
        upi[t] = A.up[3 * s + t];
        ui[t] = udi[t] + upi[t];
      }
    for (int kk = part; kk < len; kk += kEmplarParts) {
      const uint32_t ent = e[(long long)kk * 64];
      const int code = (int)(ent & (kMaxCodes - 1));
      const int j = (int)(ent >> kCodeBits);
      const double sm = code ? (double)A.sc.m[code] : 1.0;
      const double sp = code ? (double)A.sc.p[code] : 1.0, su = code ? (double)A.sc.u[code] : 1.0;
      double R[3];
      pair_disp64(A.pos[j], pi, A.box, R);  // r_i - r_k
      const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      const double r = sqrt(r2), rinv = 1 / r, rr2 = rinv * rinv;
      double rr[6], bn[6] = {0, 0, 0, 0, 0, 0};
      rr[0] = rinv;
#pragma unroll
      for (int q = 1; q < 6; ++q) rr[q] = rr[q - 1] * (double)(2 * q - 1) * rr2;
      if (EWALD) {
        double ex;
        bn[0] = t9_erfc_pair(A.ek64.a * r, ex) * rinv;
#pragma unroll
        for (int q = 1; q < 6; ++q) bn[q] = ((double)(2 * q - 1) * bn[q - 1] + A.ek64.e[q] * ex) * rr2;
      }
      const int jo = A.perm[j];
      double lci[6], lck[6], lik[6];
      chgpen_lambda(r, alpha_i, cp[3 * jo + 2], lci, lck, lik);
      double Bcc[6], Bi[6], Bk[6], Bik[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const double b = EWALD ? bn[q] - rr[q] : 0.0;  // Ewald part less the bare kernel
        Bcc[q] = b + sm * rr[q];
        Bi[q] = b + sm * lci[q] * rr[q];
        Bk[q] = b + sm * lck[q] * rr[q];
        Bik[q] = b + sm * lik[q] * rr[q];
      }
      Mp<double> ck = {}, vk;
      ck.q = cp[3 * jo];
      vk.q = cp[3 * jo + 1];
      for (int t = 0; t < 3; ++t) vk.d[t] = A.rpole[10 * (size_t)j + 1 + t];
      for (int t = 0; t < 6; ++t) vk.Q[t] = A.rpole[10 * (size_t)j + 4 + t];
      PairOut<double> o;
      if (DM) {
        // core of i: energy and gradient only (a point charge has no torque)
        pair_terms<double, true, false>(R, Bcc, nullptr, ci, ck, z3, z3, z3, z3, z3, z3, 1.0, o);
        for (int t = 0; t < 3; ++t) o.g[t] = 0;
        for (int t = 0; t < 6; ++t) o.h[t] = 0;
        acc.add(o, R);
        pair_terms<double, true, false>(R, Bk, nullptr, ci, vk, z3, z3, z3, z3, z3, z3, 1.0, o);
        for (int t = 0; t < 3; ++t) o.g[t] = 0;
        for (int t = 0; t < 6; ++t) o.h[t] = 0;
        acc.add(o, R);
        // valence of i: also the potential derivatives that turn into its torque
        pair_terms<double, true, false>(R, Bi, nullptr, vi, ck, z3, z3, z3, z3, z3, z3, 1.0, o);
        acc.add(o, R);
        pair_terms<double, true, false>(R, Bik, nullptr, vi, vk, z3, z3, z3, z3, z3, z3, 1.0, o);
        acc.add(o, R);
      }
      if (DP) {
        double udk[3], upk[3], uk[3];
        for (int t = 0; t < 3; ++t) {
          udk[t] = A.ud[3 * j + t];
          upk[t] = A.up[3 * j + t];
          uk[t] = udk[t] + upk[t];
        }
        double Pc[6], Pk[6], Pi[6], Pu[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          const double b = EWALD ? bn[q] - rr[q] : 0.0;
          Pc[q] = b + sp * rr[q];
          Pk[q] = b + sp * lck[q] * rr[q];
          Pi[q] = b + sp * lci[q] * rr[q];
          Pu[q] = b + su * lik[q] * rr[q];
        }
        // u_i at the core of k and the core of i in the field of u_k
        // (undamped; the core has no torque)
        pair_terms<double, false, true>(R, nullptr, Pc, ci, ck, ui, uk, z3, z3, z3, z3, 0.0, o);
        for (int t = 0; t < 3; ++t) o.g[t] = 0;
        for (int t = 0; t < 6; ++t) o.h[t] = 0;
        acc.add(o, R);
        // u_i in the valence multipole of k (k's damping)
        pair_terms<double, false, true>(R, nullptr, Pk, ci, vk, ui, z3, z3, z3, z3, z3, 0.0, o);
        acc.add(o, R);
        // the valence multipole of i in the field of u_k (i's damping; torque)
        pair_terms<double, false, true>(R, nullptr, Pi, vi, ck, z3, uk, z3, z3, z3, z3, 0.0, o);
        acc.add(o, R);
        // dipole-dipole (valence-valence damping)
        pair_terms<double, false, true>(R, nullptr, Pu, ci, ck, z3, z3, udi, upi, udk, upk, 1.0, o);
        acc.add(o, R);
      }
    }
  }
  double av[19];
  acc.pack(av);
  parts_reduce<kEmplarParts, 19>(av, sh);
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (part == 0) {
    acc.unpack(av);
    if (s < A.n) {
      const double f = kCoul;
      for (int k = 0; k < 3; ++k) A.grad[3 * (size_t)s + k] = f * acc.gr[k];
      for (int k = 0; k < 3; ++k) A.tq[9 * (size_t)s + k] = f * acc.tg[k];
      for (int k = 0; k < 6; ++k) A.tq[9 * (size_t)s + 3 + k] = f * acc.th[k];
    }
    red[EP_MREAL] = 0.5 * kCoul * acc.e;
    for (int k = 0; k < 6; ++k) red[EP_VXX + k] = 0.5 * kCoul * acc.v[k];
  }
  block_row<EP_N>(red, A.rows + (size_t)blk * kRow);
}

// Pauli repulsion (HIPPO; replaces src/acc/hippo/erepel.cpp:13-215 and
// src/cu/hippo/erepel_cu1.cc with include/seq/pair_repel.h): the multipole
// contraction of the two valence electron clouds (-v_i, d_i, Q_i) and
// (-v_k, d_k, Q_k) (v the valence count, d and Q the rotated multipoles)
// with the radial series D_n of repel_series, times K_i K_k s / r (sizes,
// the pair's repulsion scale) and the fifth-order switch between taper and
// cutoff; the radial factor F = Sw / r multiplies the energy, its potential
// derivatives (torques) and, with F', the pair gradient.  Owner computes
// over the buffered multipole list; egrad and etq are added to, the energy
// and virial halved.  FP64 pair arithmetic in both precisions.
struct RepelArgs {
  int n;
  const int* nlist;
  const long long* woff;
  const uint32_t* entries;
  const int* perm;
  const double* rp;  // caller order [n][3]: size, exponent, valence count
  double sc[kMaxCodes];
  double off2, taper, inv_w;
  double* grad;      // sorted [n][3], +=
  double* tq;        // sorted [n][9], +=
  double* rows;
};

template <class real>
__device__ __forceinline__ void repel_pair(const double* R, double r, const Mp<double>& mi, const Mp<double>& mk,
                                           double ai, double ak, double kk, const RepelArgs& A,
                                           PairOut<double>& o) {
  const double z3[3] = {0, 0, 0};
  double D[6];
  repel_series(r, ai, ak, D);
#pragma unroll
  for (int q = 0; q < 6; ++q) D[q] *= kk;
  pair_terms<double, true, false>(R, D, nullptr, mi, mk, z3, z3, z3, z3, z3, z3, 1.0, o);
  double Sw = 1, dSw = 0;
  if (r > A.taper) {
    const double x = (r - A.taper) * A.inv_w;
    Sw = 1 - x * x * x * (10 + x * (6 * x - 15));
    dSw = -30 * x * x * (1 - x) * (1 - x) * A.inv_w;
  }
  const double F = Sw / r, dF = dSw / r - Sw / (r * r);
  const double e0 = o.em;
  o.em *= F;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    o.de[t] = F * o.de[t] + e0 * dF * R[t] / r;
    o.g[t] *= F;
  }
#pragma unroll
  for (int t = 0; t < 6; ++t) o.h[t] *= F;
}

template <class real>
__device__ __forceinline__ Mp<double> repel_mp(const real* __restrict__ rpole, int s, double v) {
  Mp<double> m;
  m.q = -v;  // the valence electrons
  for (int t = 0; t < 3; ++t) m.d[t] = rpole[10 * (size_t)s + 1 + t];
  for (int t = 0; t < 6; ++t) m.Q[t] = rpole[10 * (size_t)s + 4 + t];
  return m;
}

template <class real>
__global__ __launch_bounds__(64 * kEmplarParts) void k_repel(RepelArgs A, const PosRec<real>* __restrict__ pos,
                                                             BoxT<real> box, const real* __restrict__ rpole) {
  __shared__ double sh[parts_lds<kEmplarParts, 19>()];
  const int part = threadIdx.x >> 6;
  const int blk = xcd_block();
  const int s = blk * 64 + (threadIdx.x & 63);
  Acc<double> acc;
  acc.zero();
  if (s < A.n) {
    const PosRec<real> pi = pos[s];
    const int io = A.perm[s];
    const double Ki = A.rp[3 * io], ai = A.rp[3 * io + 1];
    const Mp<double> mi = repel_mp(rpole, s, A.rp[3 * io + 2]);
    const int len = A.nlist[s];
    const uint32_t* en = A.entries + A.woff[s >> 6] + (s & 63);
    for (int kk = part; kk < len; kk += kEmplarParts) {
      const uint32_t ent = en[(long long)kk * 64];
      const int k = (int)(ent >> kCodeBits);
      const int code = (int)(ent & (kMaxCodes - 1));
      const double sc = A.sc[code];
      if (sc == 0) continue;
      double R[3];
      pair_disp64(pos[k], pi, box, R);  // r_i - r_k
      const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      if (!(r2 <= A.off2)) continue;
      const int ko = A.perm[k];
      const Mp<double> mk = repel_mp(rpole, k, A.rp[3 * ko + 2]);
      PairOut<double> o;
      repel_pair<real>(R, sqrt(r2), mi, mk, ai, A.rp[3 * ko + 1], Ki * A.rp[3 * ko] * sc, A, o);
      acc.add(o, R);
    }
  }
  double av[19];
  acc.pack(av);
  parts_reduce<kEmplarParts, 19>(av, sh);
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (part == 0) {
    acc.unpack(av);
    if (s < A.n) {
      for (int k = 0; k < 3; ++k) A.grad[3 * (size_t)s + k] += acc.gr[k];
      for (int k = 0; k < 3; ++k) A.tq[9 * (size_t)s + k] += acc.tg[k];
      for (int k = 0; k < 6; ++k) A.tq[9 * (size_t)s + 3 + k] += acc.th[k];
    }
    red[EP_EREP] = 0.5 * acc.e;
    for (int k = 0; k < 6; ++k) red[EP_VXX + k] = 0.5 * acc.v[k];
  }
  block_row<EP_N>(red, A.rows + (size_t)blk * kRow);
}

// Analyze-mode count (the reference's nrep): pairs within the cutoff, each
// once, whose unscaled energy is not zero, less the pairs of scale 0.
template <class real>
__global__ __launch_bounds__(256) void k_repel_count(RepelArgs A, const PosRec<real>* __restrict__ pos,
                                                     BoxT<real> box, const real* __restrict__ rpole,
                                                     unsigned long long* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0;
  if (s < A.n) {
    const PosRec<real> pi = pos[s];
    const int io = A.perm[s];
    const Mp<double> mi = repel_mp(rpole, s, A.rp[3 * io + 2]);
    const int len = A.nlist[s];
    const uint32_t* en = A.entries + A.woff[s >> 6] + (s & 63);
    for (int kk = 0; kk < len; ++kk) {
      const uint32_t ent = en[(long long)kk * 64];
      const int k = (int)(ent >> kCodeBits);
      if (k <= s) continue;
      const int code = (int)(ent & (kMaxCodes - 1));
      if (A.sc[code] == 0) continue;
      double R[3];
      pair_disp64(pos[k], pi, box, R);
      const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      if (!(r2 <= A.off2)) continue;
      const int ko = A.perm[k];
      PairOut<double> o;
      repel_pair<real>(R, sqrt(r2), mi, repel_mp(rpole, k, A.rp[3 * ko + 2]), A.rp[3 * io + 1], A.rp[3 * ko + 1],
                       A.rp[3 * io] * A.rp[3 * ko], A, o);
      if (o.em != 0) ++cnt;
    }
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, cnt);
}

// Charge transfer and dispersion (HIPPO) share one pair-kernel skeleton:
// owner computes over the buffered multipole list (both directions), each
// atom sums its own gradient, the energy and virial are halved; FP64 pair
// arithmetic on the exact FP64 displacement of the position records, in
// both precisions; the fifth-order switch S(x) = 1 - x^3 (10 - 15 x + 6 x^2),
// x = (r - taper) / (cut - taper) (S = 1, S' = S'' = 0 at the taper; S = 0,
// S' = S'' = 0 at the cutoff) between taper and cutoff.
//
// Charge transfer (src/acc/hippo/echgtrn.cpp:13-183, src/cu/hippo/
// echgtrn_cu1.cc with include/seq/pair_chgtrn.h): E = -f s (q_i e^(-a_k r) +
// q_k e^(-a_i r)), f the Coulomb constant and s the pair's multipole scale.
//
// Dispersion with a cutoff (src/acc/hippo/edisp.cpp:113-296 in its non-Ewald
// form, include/seq/pair_disp.h): E = -s C_i C_k D(r)^2 / r^6, s the pair's
// dispersion scale and D the damping of two Slater densities: with
// d_i = a_i r, t_i = a_k^2 / (a_k^2 - a_i^2), t_k = 1 - t_i,
//   D = 1 - t_i^2 P(d_i, 2 t_k + 1) e^(-d_i) - t_k^2 P(d_k, 2 t_i + 1) e^(-d_k),
//   P(d, b) = b (1 + d + d^2 / 2) + d^3 / 4,
// and for exponents within 1e-3 the limit at their mean a, d = a r:
//   D = 1 - (1 + d + d^2/2 + 17 d^3/96 + 5 d^4/96 + d^5/96) e^(-d).
struct HippoPairArgs {
  int n;
  const int* nlist;
  const long long* woff;
  const uint32_t* entries;
  const int* perm;
  const double* par;  // caller order [n][2]
  double sc[kMaxCodes];
  double off2, taper, inv_w;
  int slot;           // energy row slot
  double* grad;       // sorted [n][3], +=
  double* rows;
};

__device__ __forceinline__ void hippo_switch(double r, double taper, double inv_w, double& e, double& de) {
  if (r > taper) {
    const double x = (r - taper) * inv_w;
    const double S = 1 - x * x * x * (10 + x * (6 * x - 15));
    const double dS = -30 * x * x * (1 - x) * (1 - x) * inv_w;
    de = e * dS + de * S;
    e *= S;
  }
}

__device__ __forceinline__ double chgtrn_pair(double r, double qi, double ai, double qk, double ak, double f,
                                              double taper, double inv_w, double& de) {
  const double ei = exp(-ai * r), ek = exp(-ak * r);
  double e = -f * (qi * ek + qk * ei);
  de = f * (qi * ak * ek + qk * ai * ei);
  hippo_switch(r, taper, inv_w, e, de);
  return e;
}

__device__ __forceinline__ double disp_pair(double r, double ci, double ai, double ck, double ak, double s,
                                            double taper, double inv_w, double& de) {
  double D, dD;  // damping and its r derivative
  if (fabs(ai - ak) > 1e-3) {
    const double di = ai * r, dk = ak * r, ei = exp(-di), ek = exp(-dk);
    const double ti = ak * ak / ((ak + ai) * (ak - ai)), tk = 1 - ti;
    const double bi = 2 * tk + 1, bk = 2 * ti + 1;
    const double Pi = bi * (1 + di + 0.5 * di * di) + 0.25 * di * di * di;
    const double Pk = bk * (1 + dk + 0.5 * dk * dk) + 0.25 * dk * dk * dk;
    D = 1 - ti * ti * Pi * ei - tk * tk * Pk * ek;
    dD = ti * ti * ai * di * di * (0.25 * (di - 1) + tk) * ei + tk * tk * ak * dk * dk * (0.25 * (dk - 1) + ti) * ek;
  } else {
    const double a = 0.5 * (ai + ak), d = a * r, ed = exp(-d);
    const double P = 1 + d * (1 + d * (0.5 + d * (17.0 / 96 + d * (5.0 / 96 + d / 96))));
    D = 1 - P * ed;
    dD = a * ed * d * d * (d * d * d - 3 * d - 3) / 96;
  }
  const double r2 = r * r, rr6 = 1 / (r2 * r2 * r2);
  double e = -ci * ck * rr6;
  de = (-6 * e / r) * D * D + 2 * e * D * dD;
  e *= D * D;
  hippo_switch(r, taper, inv_w, e, de);
  de *= s;
  return e * s;
}

// KIND 0: charge transfer (scale m); 1: dispersion (scale ds)
template <int KIND>
__device__ __forceinline__ double hippo_pair(double r, const double* pi, const double* pk, double s, const HippoPairArgs& A,
                                             double& de) {
  if (KIND == 0) return chgtrn_pair(r, pi[0], pi[1], pk[0], pk[1], kCoul * s, A.taper, A.inv_w, de);
  return disp_pair(r, pi[0], pi[1], pk[0], pk[1], s, A.taper, A.inv_w, de);
}

template <class real, bool VIR, int KIND>
__global__ __launch_bounds__(64 * kPairParts) void k_hippo_pair(HippoPairArgs A, const PosRec<real>* __restrict__ pos,
                                                                BoxT<real> box) {
  __shared__ double sh[parts_lds<kPairParts, 10>()];
  const int part = threadIdx.x >> 6;
  const int blk = xcd_block();
  const int s = blk * 64 + (threadIdx.x & 63);
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // e, g[3], v[6]
  if (s < A.n) {
    const PosRec<real> pi = pos[s];
    const int oi = A.perm[s];
    const double pari[2] = {A.par[2 * oi], A.par[2 * oi + 1]};
    const int len = A.nlist[s];
    const uint32_t* en = A.entries + A.woff[s >> 6] + (s & 63);
    for (int kk = part; kk < len; kk += kPairParts) {
      const uint32_t ent = en[(long long)kk * 64];
      const int k = (int)(ent >> kCodeBits);
      const int code = (int)(ent & (kMaxCodes - 1));
      const double sc = A.sc[code];
      if (sc == 0) continue;
      double R[3];
      pair_disp64(pos[k], pi, box, R);  // x_i - x_k
      const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      if (!(r2 <= A.off2)) continue;
      const int ok = A.perm[k];
      const double park[2] = {A.par[2 * ok], A.par[2 * ok + 1]};
      const double r = sqrt(r2);
      double de;
      const double e = hippo_pair<KIND>(r, pari, park, sc, A, de);
      const double g = de / r;
      acc[0] += e;
      for (int d = 0; d < 3; ++d) acc[1 + d] += g * R[d];
      if (VIR) {
        acc[4] += g * R[0] * R[0];
        acc[5] += g * R[1] * R[1];
        acc[6] += g * R[2] * R[2];
        acc[7] += g * R[0] * R[1];
        acc[8] += g * R[0] * R[2];
        acc[9] += g * R[1] * R[2];
      }
    }
  }
  parts_reduce<kPairParts, 10>(acc, sh);
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (part == 0) {
    if (s < A.n)
      for (int d = 0; d < 3; ++d) A.grad[3 * (size_t)s + d] += acc[1 + d];
    red[A.slot] = 0.5 * acc[0];
    for (int q = 0; q < 6; ++q) red[EP_VXX + q] = 0.5 * acc[4 + q];
  }
  block_row<EP_N>(red, A.rows + (size_t)blk * kRow);
}

// Analyze-mode count (the reference's nct / ndisp): pairs within the cutoff,
// each once, whose unscaled energy is not zero, less the pairs of scale 0.
template <class real, int KIND>
__global__ __launch_bounds__(256) void k_hippo_count(HippoPairArgs A, const PosRec<real>* __restrict__ pos,
                                                     BoxT<real> box, unsigned long long* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0;
  if (s < A.n) {
    const PosRec<real> pi = pos[s];
    const int oi = A.perm[s];
    const double pari[2] = {A.par[2 * oi], A.par[2 * oi + 1]};
    const int len = A.nlist[s];
    const uint32_t* en = A.entries + A.woff[s >> 6] + (s & 63);
    for (int kk = 0; kk < len; ++kk) {
      const uint32_t ent = en[(long long)kk * 64];
      const int k = (int)(ent >> kCodeBits);
      if (k <= s) continue;
      const int code = (int)(ent & (kMaxCodes - 1));
      if (A.sc[code] == 0) continue;
      double R[3];
      pair_disp64(pos[k], pi, box, R);
      const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      if (!(r2 <= A.off2)) continue;
      const int ok = A.perm[k];
      const double park[2] = {A.par[2 * ok], A.par[2 * ok + 1]};
      double de;
      if (hippo_pair<KIND>(sqrt(r2), pari, park, 1.0, A, de) != 0) ++cnt;
    }
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, cnt);
}

// ------------------------------------------------------------ PME

struct GridGeom {
  int K[3];
  double a[3];  // K / L
};

// 20 fractional potential derivatives of one grid component at an atom
// (order 0, x, y, z, xx, yy, zz, xy, xz, yz, xxx, yyy, zzz, xxy, xxz, xyy, yyz,
// xzz, yzz, xyz), converted to Cartesian.
template <class real, int PART>
__device__ __forceinline__ void gather20(const GridGeom& g, const R4<real>& w, const int4& b, const void* gridv,
                                         double* c) {
  using C = typename std::conditional<sizeof(real) == 4, float2, double2>::type;
  const C* grid = static_cast<const C*>(gridv);
  real th1[4][5], th2[4][5], th3[4][5];
  bspline5<4>(w.x, th1);
  bspline5<4>(w.y, th2);
  bspline5<4>(w.z, th3);
  real t[20];
#pragma unroll
  for (int k = 0; k < 20; ++k) t[k] = 0;
  for (int m3 = 0; m3 < 5; ++m3) {
    int z = b.z + m3;
    if (z >= g.K[2]) z -= g.K[2];
    const real v0 = th3[0][m3], v1 = th3[1][m3], v2 = th3[2][m3], v3 = th3[3][m3];
    // the FP64 variant keeps this loop rolled (unrolled, it spills to scratch)
#pragma unroll(sizeof(real) == 4 ? 5 : 1)
    for (int m2 = 0; m2 < 5; ++m2) {
      int y = b.y + m2;
      if (y >= g.K[1]) y -= g.K[1];
      const real u0 = th2[0][m2], u1 = th2[1][m2], u2 = th2[2][m2], u3 = th2[3][m2];
      const long long row = (long long)g.K[0] * (y + (long long)g.K[1] * z);
      // row sums over x of q * th1[l][m1]
      real s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
      for (int m1 = 0; m1 < 5; ++m1) {
        int x = b.x + m1;
        if (x >= g.K[0]) x -= g.K[0];
        const C qq = grid[row + x];
        const real q = PART == 0 ? (real)qq.x : (real)qq.y;
        s0 += q * th1[0][m1];
        s1 += q * th1[1][m1];
        s2 += q * th1[2][m1];
        s3 += q * th1[3][m1];
      }
      const real u0v0 = u0 * v0, u1v0 = u1 * v0, u0v1 = u0 * v1;
      t[0] += s0 * u0v0;
      t[1] += s1 * u0v0;
      t[2] += s0 * u1v0;
      t[3] += s0 * u0v1;
      t[4] += s2 * u0v0;
      t[5] += s0 * u2 * v0;
      t[6] += s0 * u0 * v2;
      t[7] += s1 * u1v0;
      t[8] += s1 * u0v1;
      t[9] += s0 * u1 * v1;
      t[10] += s3 * u0v0;
      t[11] += s0 * u3 * v0;
      t[12] += s0 * u0 * v3;
      t[13] += s2 * u1v0;
      t[14] += s2 * u0v1;
      t[15] += s1 * u2 * v0;
      t[16] += s0 * u2 * v1;
      t[17] += s1 * u0 * v2;
      t[18] += s0 * u1 * v2;
      t[19] += s1 * u1 * v1;
    }
  }
  const double a1 = g.a[0], a2 = g.a[1], a3 = g.a[2];
  c[0] = t[0];
  c[1] = t[1] * a1;
  c[2] = t[2] * a2;
  c[3] = t[3] * a3;
  c[4] = t[4] * a1 * a1;
  c[5] = t[5] * a2 * a2;
  c[6] = t[6] * a3 * a3;
  c[7] = t[7] * a1 * a2;
  c[8] = t[8] * a1 * a3;
  c[9] = t[9] * a2 * a3;
  c[10] = t[10] * a1 * a1 * a1;
  c[11] = t[11] * a2 * a2 * a2;
  c[12] = t[12] * a3 * a3 * a3;
  c[13] = t[13] * a1 * a1 * a2;
  c[14] = t[14] * a1 * a1 * a3;
  c[15] = t[15] * a1 * a2 * a2;
  c[16] = t[16] * a2 * a2 * a3;
  c[17] = t[17] * a1 * a3 * a3;
  c[18] = t[18] * a2 * a3 * a3;
  c[19] = t[19] * a1 * a2 * a3;
}

// Multipole m (q, d, Q as xx yy zz xy xz yz) in the potential derivatives c:
// energy, gradient with respect to its position.
__device__ __forceinline__ double contract_e(const double* m, const double* c) {
  return m[0] * c[0] + m[1] * c[1] + m[2] * c[2] + m[3] * c[3] + m[4] * c[4] + m[5] * c[5] + m[6] * c[6] +
         2 * (m[7] * c[7] + m[8] * c[8] + m[9] * c[9]);
}
__device__ __forceinline__ void contract_g(const double* m, const double* c, double* g) {
  g[0] = m[0] * c[1] + m[1] * c[4] + m[2] * c[7] + m[3] * c[8] + m[4] * c[10] + m[5] * c[15] + m[6] * c[17] +
         2 * (m[7] * c[13] + m[8] * c[14] + m[9] * c[19]);
  g[1] = m[0] * c[2] + m[1] * c[7] + m[2] * c[5] + m[3] * c[9] + m[4] * c[13] + m[5] * c[11] + m[6] * c[18] +
         2 * (m[7] * c[15] + m[8] * c[19] + m[9] * c[16]);
  g[2] = m[0] * c[3] + m[1] * c[8] + m[2] * c[9] + m[3] * c[6] + m[4] * c[14] + m[5] * c[16] + m[6] * c[12] +
         2 * (m[7] * c[19] + m[8] * c[17] + m[9] * c[18]);
}
// dipole u in the second derivatives h (xx yy zz xy xz yz): (H u)
__device__ __forceinline__ void hu(const double* h, const double* u, double* o) {
  o[0] = h[0] * u[0] + h[3] * u[1] + h[4] * u[2];
  o[1] = h[3] * u[0] + h[1] * u[1] + h[5] * u[2];
  o[2] = h[4] * u[0] + h[5] * u[1] + h[2] * u[2];
}
// symmetric part of phi_a d_b + 2 (Q H)_ba, accumulated with weight w into v6
__device__ __forceinline__ void vir_moments(const double* d, const double* Q, const double* phi1, const double* H,
                                            double w, double* v6) {
  double M[3][3];
  const double Qm[3][3] = {{Q[0], Q[3], Q[4]}, {Q[3], Q[1], Q[5]}, {Q[4], Q[5], Q[2]}};
  const double Hm[3][3] = {{H[0], H[3], H[4]}, {H[3], H[1], H[5]}, {H[4], H[5], H[2]}};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double qh = 0;
      for (int e = 0; e < 3; ++e) qh += Qm[b][e] * Hm[e][a];
      M[a][b] = phi1[a] * d[b] + 2 * qh;
    }
  v6[0] += w * M[0][0];
  v6[1] += w * M[1][1];
  v6[2] += w * M[2][2];
  v6[3] += w * 0.5 * (M[0][1] + M[1][0]);
  v6[4] += w * 0.5 * (M[0][2] + M[2][0]);
  v6[5] += w * 0.5 * (M[1][2] + M[2][1]);
}

template <class real>
__global__ __launch_bounds__(256) void k_recip_mpole(int n, GridGeom g, const R4<real>* __restrict__ pw,
                                                     const int4* __restrict__ pbase, const void* __restrict__ grid,
                                                     const real* __restrict__ rpole, int use_m, double* __restrict__ grad,
                                                     double* __restrict__ tq, double* __restrict__ cphim,
                                                     double* __restrict__ rows) {
  const int blk = xcd_block();
  const int s = blk * blockDim.x + threadIdx.x;
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (s < n) {
    double c[20];
    gather20<real, 0>(g, pw[s], pbase[s], grid, c);
    if (cphim)
      for (int k = 0; k < 10; ++k) cphim[10 * (size_t)s + k] = c[k];
    if (use_m) {
      double m[10];
      for (int k = 0; k < 10; ++k) m[k] = (double)rpole[10 * (size_t)s + k];
      const double f = kCoul;
      red[EP_MRECIP] = 0.5 * f * contract_e(m, c);
      double gg[3];
      contract_g(m, c, gg);
      for (int k = 0; k < 3; ++k) grad[3 * (size_t)s + k] += f * gg[k];
      for (int k = 0; k < 3; ++k) tq[9 * (size_t)s + k] += f * c[1 + k];
      for (int k = 0; k < 6; ++k) tq[9 * (size_t)s + 3 + k] += f * c[4 + k];
      vir_moments(m + 1, m + 4, c + 1, c + 4, -f, red + EP_VXX);
    }
  }
  block_row<EP_N>(red, rows + (size_t)blk * kRow);
}

template <class real>
__global__ __launch_bounds__(256) void k_recip_polar(int n, GridGeom g, const R4<real>* __restrict__ pw,
                                                     const int4* __restrict__ pbase, const void* __restrict__ grid,
                                                     const real* __restrict__ rpole, const real* __restrict__ ud,
                                                     const real* __restrict__ up, const double* __restrict__ cphim,
                                                     double self3, double* __restrict__ grad,
                                                     double* __restrict__ tq, double* __restrict__ rows) {
  const int blk = xcd_block();
  const int s = blk * blockDim.x + threadIdx.x;
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (s < n) {
    double cd[20], cp[20], cx[20];
    gather20<real, 0>(g, pw[s], pbase[s], grid, cd);
    gather20<real, 1>(g, pw[s], pbase[s], grid, cp);
    for (int k = 0; k < 20; ++k) cx[k] = cd[k] + cp[k];
    double m[10], u[3], v[3], x[3], cm[10];
    for (int k = 0; k < 10; ++k) m[k] = (double)rpole[10 * (size_t)s + k];
    for (int k = 0; k < 10; ++k) cm[k] = cphim[10 * (size_t)s + k];
    for (int k = 0; k < 3; ++k) {
      u[k] = (double)ud[3 * (size_t)s + k];
      v[k] = (double)up[3 * (size_t)s + k];
      x[k] = u[k] + v[k];
    }
    const double f = kCoul;
    double t1[3], t2[3], t3[3], t4[3];
    hu(cm + 4, x, t1);   // (ud + up) . grad grad phi_M
    contract_g(m, cx, t2);  // M . grad phi_(ud+up)
    hu(cp + 4, u, t3);   // ud . grad grad phi_up
    hu(cd + 4, v, t4);   // up . grad grad phi_ud
    for (int k = 0; k < 3; ++k) grad[3 * (size_t)s + k] += 0.5 * f * (t1[k] + t2[k] + t3[k] + t4[k]);
    for (int k = 0; k < 3; ++k) tq[9 * (size_t)s + k] += 0.5 * f * (cx[1 + k] - self3 * x[k]);
    for (int k = 0; k < 6; ++k) tq[9 * (size_t)s + 3 + k] += 0.5 * f * cx[4 + k];
    // virial: moments of the cross forms 1/2 E(u_d+u_p, M) + 1/2 E(u_d, u_p)
    const double zq[6] = {0, 0, 0, 0, 0, 0};
    vir_moments(x, zq, cm + 1, cm + 4, -0.5 * f, red + EP_VXX);
    vir_moments(m + 1, m + 4, cx + 1, cx + 4, -0.5 * f, red + EP_VXX);
    vir_moments(u, zq, cp + 1, cp + 4, -0.5 * f, red + EP_VXX);
    vir_moments(v, zq, cd + 1, cd + 4, -0.5 * f, red + EP_VXX);
  }
  block_row<EP_N>(red, rows + (size_t)blk * kRow);
}

// Ewald self energy of the multipoles and the dot-product polarization energy
// -1/2 f sum ud . E_p with E_p = udirp / alpha (src/cu/amoeba/epolar.cu:15-30).
template <class real>
__global__ __launch_bounds__(256) void k_self_polar(int n, int use_m, int use_p, int ewald, double aewald,
                                                    const real* __restrict__ rpole, const real* __restrict__ ud,
                                                    const real* __restrict__ udirp, const real* __restrict__ polinv,
                                                    double* __restrict__ rows) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (s < n) {
    if (use_m && ewald) {
      const real* p = rpole + 10 * (size_t)s;
      double v[10];
      for (int k = 0; k < 10; ++k) v[k] = (double)p[k];
      const double cii = v[0] * v[0];
      const double dii = v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
      const double qii = 2 * (v[7] * v[7] + v[8] * v[8] + v[9] * v[9]) + v[4] * v[4] + v[5] * v[5] + v[6] * v[6];
      const double a2 = 2 * aewald * aewald;
      red[EP_MSELF] = -kCoul * aewald / kSqrtPiD * (cii + a2 * (dii / 3 + 2 * a2 * qii * 0.2));
    }
    if (use_p) {
      double sdot = 0;
      for (int k = 0; k < 3; ++k) sdot += (double)ud[3 * (size_t)s + k] * (double)udirp[3 * (size_t)s + k];
      red[EP_POLAR] = -0.5 * kCoul * (double)polinv[s] * sdot;
    }
  }
  block_row<EP_N>(red, rows + (size_t)blockIdx.x * kRow);
}

// ------------------------------------------------------------ torques

__device__ __forceinline__ void dvec(const double* xyz, int i, int j, const double* L, int periodic, double* v) {
  for (int d = 0; d < 3; ++d) {
    double x = xyz[3 * j + d] - xyz[3 * i + d];
    if (periodic) x -= L[d] * rint(x / L[d]);
    v[d] = x;
  }
}
__device__ __forceinline__ double nrm(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
__device__ __forceinline__ void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
// (I - e e^T) v / |x|, the derivative of a unit vector e = x/|x| applied to v
__device__ __forceinline__ void proj_div(const double* e, double len, const double* v, double* o) {
  const double ev = e[0] * v[0] + e[1] * v[1] + e[2] * v[2];
  for (int d = 0; d < 3; ++d) o[d] = (v[d] - ev * e[d]) / len;
}

// Torque of atom i's multipole and its distribution over the frame atoms.
// The frame is e_z = Z/|Z|, e_x = W/|W| with W = X - (X.e_z) e_z, e_y = e_z x
// e_x.  A rotation by dtheta moves the multipole; the frame's rotation is
//   dtheta = e_z x de_z + (dX.e_y - (X.e_z)(de_z.e_y)) / |W| e_z,
// so for tau = dE/dtheta
//   dE/dZ = (tau x e_z)/|Z| - tau_z (X.e_z)/(|W| |Z|) e_y,  dE/dX = tau_z/|W| e_y,
// and Z, X are the frame vectors of each frame type (u = r_z - r_i,
// v = r_x - r_i, w = r_y - r_i):  Z-then-X Z = u, X = v;  Bisector
// Z = u^ + v^, X = v;  Z-only Z = u, X fixed;  Z-Bisect Z = u, X = v^ + w^;
// 3-Fold Z = u^ + v^ + w^, X = v.  Frame construction as in mpole.hip.
template <class real>
__global__ __launch_bounds__(256) void k_torque(int n, const int* __restrict__ perm, const double* __restrict__ xyz,
                                                double lx, double ly, double lz, int periodic,
                                                const int* __restrict__ frame, const int* __restrict__ zaxis,
                                                const int* __restrict__ xaxis, const int* __restrict__ yaxis,
                                                const real* __restrict__ rpole, const double* __restrict__ tq,
                                                double* __restrict__ fout, double* __restrict__ rows) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (s < n) {
    double* out = fout + 12 * (size_t)s;
    for (int k = 0; k < 12; ++k) out[k] = 0;
    const int i = perm[s];
    const int fr = frame[i];
    if (fr != 0) {
      const double L[3] = {lx, ly, lz};
      const real* p = rpole + 10 * (size_t)s;
      const double d[3] = {(double)p[1], (double)p[2], (double)p[3]};
      const double Q[3][3] = {{(double)p[4], (double)p[7], (double)p[8]},
                              {(double)p[7], (double)p[5], (double)p[9]},
                              {(double)p[8], (double)p[9], (double)p[6]}};
      const double* t = tq + 9 * (size_t)s;
      const double g[3] = {t[0], t[1], t[2]};
      const double H[3][3] = {{t[3], t[6], t[7]}, {t[6], t[4], t[8]}, {t[7], t[8], t[5]}};
      double P[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[a][b] = Q[a][0] * H[0][b] + Q[a][1] * H[1][b] + Q[a][2] * H[2][b];
      double tau[3];
      cross(d, g, tau);
      tau[0] += 2 * (P[1][2] - P[2][1]);
      tau[1] += 2 * (P[2][0] - P[0][2]);
      tau[2] += 2 * (P[0][1] - P[1][0]);
      double u[3], v[3] = {0, 0, 0}, w[3] = {0, 0, 0};
      dvec(xyz, i, zaxis[i], L, periodic, u);
      const double ul = nrm(u);
      double uh[3] = {u[0] / ul, u[1] / ul, u[2] / ul};
      double vl = 1, wl = 1, vh[3] = {0, 0, 0}, wh[3] = {0, 0, 0};
      if (fr != 1) {
        dvec(xyz, i, xaxis[i], L, periodic, v);
        vl = nrm(v);
        for (int k = 0; k < 3; ++k) vh[k] = v[k] / vl;
      }
      if (fr == 4 || fr == 5) {
        dvec(xyz, i, yaxis[i], L, periodic, w);
        wl = nrm(w);
        for (int k = 0; k < 3; ++k) wh[k] = w[k] / wl;
      }
      double Z[3], X[3];
      for (int k = 0; k < 3; ++k) {
        if (fr == 3) Z[k] = uh[k] + vh[k];
        else if (fr == 5) Z[k] = uh[k] + vh[k] + wh[k];
        else Z[k] = u[k];
        if (fr == 4) X[k] = vh[k] + wh[k];
        else X[k] = v[k];
      }
      if (fr == 1) {
        const bool ok = !(fabs(uh[0]) > 0.866);
        X[0] = ok ? 1 : 0;
        X[1] = ok ? 0 : 1;
        X[2] = 0;
      }
      const double zl = nrm(Z);
      const double ez[3] = {Z[0] / zl, Z[1] / zl, Z[2] / zl};
      const double xz = X[0] * ez[0] + X[1] * ez[1] + X[2] * ez[2];
      double W[3] = {X[0] - xz * ez[0], X[1] - xz * ez[1], X[2] - xz * ez[2]};
      const double wlen = nrm(W);
      const double ex[3] = {W[0] / wlen, W[1] / wlen, W[2] / wlen};
      double ey[3];
      cross(ez, ex, ey);
      const double tz = tau[0] * ez[0] + tau[1] * ez[1] + tau[2] * ez[2];
      double tez[3];
      cross(tau, ez, tez);
      double dZ[3], dX[3];
      for (int k = 0; k < 3; ++k) {
        dZ[k] = tez[k] / zl - tz * xz / (wlen * zl) * ey[k];
        dX[k] = tz / wlen * ey[k];
      }
      double gu[3] = {0, 0, 0}, gv[3] = {0, 0, 0}, gw[3] = {0, 0, 0};
      if (fr == 2 || fr == 1) {
        for (int k = 0; k < 3; ++k) gu[k] = dZ[k];
        if (fr == 2)
          for (int k = 0; k < 3; ++k) gv[k] = dX[k];
      } else if (fr == 3) {
        proj_div(uh, ul, dZ, gu);
        proj_div(vh, vl, dZ, gv);
        for (int k = 0; k < 3; ++k) gv[k] += dX[k];
      } else if (fr == 4) {
        for (int k = 0; k < 3; ++k) gu[k] = dZ[k];
        proj_div(vh, vl, dX, gv);
        proj_div(wh, wl, dX, gw);
      } else if (fr == 5) {
        proj_div(uh, ul, dZ, gu);
        proj_div(vh, vl, dZ, gv);
        proj_div(wh, wl, dZ, gw);
        for (int k = 0; k < 3; ++k) gv[k] += dX[k];
      }
      for (int k = 0; k < 3; ++k) {
        out[k] = -(gu[k] + gv[k] + gw[k]);
        out[3 + k] = gu[k];
        out[6 + k] = gv[k];
        out[9 + k] = gw[k];
      }
      // strain moves the frame atoms relative to atom i
      double Wv[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Wv[a][b] = gu[a] * u[b] + gv[a] * v[b] + gw[a] * w[b];
      red[EP_VXX] = Wv[0][0];
      red[EP_VYY] = Wv[1][1];
      red[EP_VZZ] = Wv[2][2];
      red[EP_VXY] = 0.5 * (Wv[0][1] + Wv[1][0]);
      red[EP_VXZ] = 0.5 * (Wv[0][2] + Wv[2][0]);
      red[EP_VYZ] = 0.5 * (Wv[1][2] + Wv[2][1]);
    }
  }
  block_row<EP_N>(red, rows + (size_t)blockIdx.x * kRow);
}

// grad[t] += contributions of every frame that names atom perm[t], in the
// fixed order of the reverse table.
__global__ void k_frame_gather(int n, const int* __restrict__ perm, const int* __restrict__ iperm,
                               const int* __restrict__ rev_off, const int* __restrict__ rev_ent,
                               const double* __restrict__ fout, double* __restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int j = perm[t];
  double g0 = 0, g1 = 0, g2 = 0;
  for (int q = rev_off[j]; q < rev_off[j + 1]; ++q) {
    const int ent = rev_ent[q];
    const int s = iperm[ent >> 2], slot = ent & 3;
    const double* o = fout + 12 * (size_t)s + 3 * slot;
    g0 += o[0];
    g1 += o[1];
    g2 += o[2];
  }
  grad[3 * (size_t)t] += g0;
  grad[3 * (size_t)t + 1] += g1;
  grad[3 * (size_t)t + 2] += g2;
}

// Dipoles + signed multipoles for the virial's packed grids:
// z10 = (u_d + u_p) + sgn M, z3 = u_d + sgn u_p.
template <class real>
__global__ void k_make_z(int n, real sgn, const real* __restrict__ rpole, const real* __restrict__ ud,
                         const real* __restrict__ up, real* __restrict__ z10, real* __restrict__ z3) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  for (int k = 0; k < 10; ++k) z10[10 * (size_t)s + k] = sgn * rpole[10 * (size_t)s + k];
  for (int k = 0; k < 3; ++k) {
    const real a = ud[3 * (size_t)s + k], b = up[3 * (size_t)s + k];
    z10[10 * (size_t)s + 1 + k] += a + b;
    z3[3 * (size_t)s + k] = a + sgn * b;
  }
}

__global__ __launch_bounds__(256) void k_rows_reduce(int nrows, const double* __restrict__ rows,
                                                     double* __restrict__ out) {
  double v[EP_N];
  for (int q = 0; q < EP_N; ++q) v[q] = 0;
  for (int r = threadIdx.x; r < nrows; r += 256)
    for (int q = 0; q < EP_N; ++q) v[q] += rows[(size_t)r * kRow + q];
  double res[EP_N];
  for (int q = 0; q < EP_N; ++q) res[q] = v[q];
  block_row<EP_N>(res, out);
}

// dst (caller order) = src (sorted) + add (caller order, may be null)
__global__ void k_grad_to_orig(int n, const int* __restrict__ perm, const double* __restrict__ src,
                               const double* __restrict__ add, double* __restrict__ dst) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  for (int d = 0; d < 3; ++d)
    dst[3 * (size_t)i + d] = src[3 * (size_t)s + d] + (add ? add[3 * (size_t)i + d] : 0.0);
}

// The scaled partners of each atom (inside the cutoff): each whole pair term,
// unscaled Ewald part plus the (s - 1) non-Ewald correction, in FP64 from the
// exact displacement, whatever the working precision.  These are the large
// intramolecular terms, alike from one molecule to the next, so FP32 rounding
// in them would not average out over the system.  Runs after k_emplar (a
// kernel of its own so the FP64 registers do not lower the occupancy of the
// list loop) and adds to the same lane-owned gradient and torque terms.
template <class real, int EWALD, int MODE>
__global__ __launch_bounds__(256) void k_emplar_scaled(EmplarArgs<real> A) {
  constexpr bool DM = MODE & 1, DP = (MODE & 2) != 0;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  Acc<real> acc;
  acc.zero();
  if (s < A.n) {
    const int io = A.perm[s];
    const PosRec<real> pi = A.pos[s];
    const real* pp = A.rpole + 10 * (size_t)s;
    Mp<double> mid;
    mid.q = pp[0];
    for (int t = 0; t < 3; ++t) mid.d[t] = pp[1 + t];
    for (int t = 0; t < 6; ++t) mid.Q[t] = pp[4 + t];
    double udd[3] = {0, 0, 0}, upd[3] = {0, 0, 0}, uid[3] = {0, 0, 0};
    if (DP)
      for (int t = 0; t < 3; ++t) {
        udd[t] = A.ud[3 * s + t];
        upd[t] = A.up[3 * s + t];
        uid[t] = udd[t] + upd[t];
      }
    const int jpi = A.uth ? 0 : A.jpolar[s] * A.nj;
    for (int q = A.excl_off[io]; q < A.excl_off[io + 1]; ++q) {
      const int code = A.excl_code[q];
      const double wm = A.sc.m[code] - 1, wd = A.sc.d[code] - 1, wp = A.sc.p[code] - 1, wu = A.sc.u[code] - 1;
      const int j = A.iperm[A.excl_j[q]];
      const PosRec<real> pj = A.pos[j];
      // same displacement and test as the in-cutoff compaction
      real xr, yr, zr;
      pair_disp(pi, pj, A.box, xr, yr, zr);
      if (!(xr * xr + yr * yr + zr * zr <= A.cut2)) continue;
      double R[3];
      pair_disp64(pj, pi, A.box, R);  // r_i - r_k
      const double r2 = dot3(R, R);
      const double r = sqrt(r2), rinv = 1 / r, rr2 = rinv * rinv;
      double rr[6], bn[6], BM[6];
      rr[0] = rinv;
#pragma unroll
      for (int t = 1; t < 6; ++t) rr[t] = rr[t - 1] * (double)(2 * t - 1) * rr2;
      if (EWALD) {
        double ex;
        bn[0] = t9_erfc_pair(A.ek64.a * r, ex) * rinv;
#pragma unroll
        for (int t = 1; t < 6; ++t) bn[t] = ((double)(2 * t - 1) * bn[t - 1] + A.ek64.e[t] * ex) * rr2;
      } else {
#pragma unroll
        for (int t = 0; t < 6; ++t) bn[t] = rr[t];
      }
#pragma unroll
      for (int t = 0; t < 6; ++t) BM[t] = bn[t] + wm * rr[t];
      const real* pk = A.rpole + 10 * (size_t)j;
      Mp<double> mkd;
      mkd.q = pk[0];
      for (int t = 0; t < 3; ++t) mkd.d[t] = pk[1 + t];
      for (int t = 0; t < 6; ++t) mkd.Q[t] = pk[4 + t];
      double BT[5] = {0, 0, 0, 0, 0}, BTs[5] = {0, 0, 0, 0, 0};
      double udk[3] = {0, 0, 0}, upk[3] = {0, 0, 0}, ukd[3] = {0, 0, 0}, uie[3] = {0, 0, 0}, uke[3] = {0, 0, 0};
      if (DP) {
        const double pga = A.uth ? (double)A.pga_u : (double)A.thlval[jpi + A.jpolar[j]];
        double lam[5];
        thole4(r, (double)pi.w * (double)pj.w, pga, lam);
#pragma unroll
        for (int t = 1; t < 5; ++t) {
          BT[t] = bn[t] - (1 - lam[t]) * rr[t];
          BTs[t] = lam[t] * rr[t];
        }
        for (int t = 0; t < 3; ++t) {
          udk[t] = A.ud[3 * j + t];
          upk[t] = A.up[3 * j + t];
          ukd[t] = udk[t] + upk[t];
          uie[t] = wd * upd[t] + wp * udd[t];
          uke[t] = wd * upk[t] + wp * udk[t];
        }
      }
      PairOut<double> o;
      pair_terms<double, DM, DP>(R, BM, BT, mid, mkd, uid, ukd, udd, upd, udk, upk, 1.0, o);
      acc.add(o, R);
      if (DP && (wd != 0 || wp != 0 || wu != 0)) {
        const double Z[6] = {0, 0, 0, 0, 0, 0};
        pair_terms<double, DM, DP>(R, Z, BTs, mid, mkd, uie, uke, udd, upd, udk, upk, wu, o);
        acc.add(o, R);
      }
    }
    const double f = kCoul;
    for (int k = 0; k < 3; ++k) A.grad[3 * (size_t)s + k] += f * acc.gr[k];
    for (int k = 0; k < 3; ++k) A.tq[9 * (size_t)s + k] += f * acc.tg[k];
    for (int k = 0; k < 6; ++k) A.tq[9 * (size_t)s + 3 + k] += f * acc.th[k];
  }
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  red[EP_MREAL] = 0.5 * kCoul * acc.e;
  for (int k = 0; k < 6; ++k) red[EP_VXX + k] = 0.5 * kCoul * acc.v[k];
  block_row<EP_N>(red, A.rows + (size_t)blockIdx.x * kRow);
}

// ------------------------------------------------------------ host side

template <class real>
EwK6<real> ewk6(const Ctx& c) {
  EwK6<real> k;
  const double a = c.aewald, isp = 0.56418958354775628695;
  k.a = (real)a;
  double t = 2 * a * isp;  // 2 a / sqrt(pi)
  k.e[0] = 0;
  for (int q = 1; q < 6; ++q) {
    k.e[q] = (real)t;
    t *= 2 * a * a;
  }
  return k;
}


template <class real>
void real_space_t(Ctx& c, bool use_m, bool use_p) {
  EmplarArgs<real> A;
  const auto& L = c.mlist;
  if (!L.cut_valid) throw ArgError("in-cutoff neighbor list is stale (internal ordering error)");
  A.n = c.n;
  A.ncut = L.ncut.as<int>();
  A.woff = L.off.as<long long>();
  A.ecut = L.ecut.as<uint32_t>();
  A.pos = c.pos.as<PosRec<real>>();
  A.rpole = c.rpole.as<real>();
  A.ud = c.uind.as<real>();
  A.up = c.uinp.as<real>();
  A.jpolar = c.jpolar.as<int>();
  A.thlval = c.thlval.as<real>();
  A.nj = c.nj;
  A.uth = c.uniform_thole ? 1 : 0;
  A.pga_u = (real)c.thole_value;
  A.perm = c.perm.as<int>();
  A.iperm = c.iperm.as<int>();
  A.excl_off = c.excl_off.as<int>();
  A.excl_j = c.excl_j.as<int>();
  A.excl_code = c.excl_code.as<int>();
  A.sc = c.scale;
  A.box = make_boxT<real>(c);
  const double c2 = L.cut * L.cut;
  A.cut2 = (real)(c2 > 1e30 ? 1e30 : c2);
  A.ek = ewk6<real>(c);
  A.ek64 = ewk6<double>(c);
  A.grad = c.egrad.as<double>();
  A.tq = c.etq.as<double>();
  const int nb = ceil_div(c.n, 256);
  const int mode = (use_m ? 1 : 0) | (use_p ? 2 : 0);
  void (*kern)(EmplarArgs<real>) = nullptr;
  if (c.ewald) kern = mode == 1 ? k_emplar<real, 1, 1> : (mode == 2 ? k_emplar<real, 1, 2> : k_emplar<real, 1, 3>);
  else kern = mode == 1 ? k_emplar<real, 0, 1> : (mode == 2 ? k_emplar<real, 0, 2> : k_emplar<real, 0, 3>);
  void (*kscaled)(EmplarArgs<real>) = nullptr;
  if (c.ewald)
    kscaled = mode == 1 ? k_emplar_scaled<real, 1, 1>
                        : (mode == 2 ? k_emplar_scaled<real, 1, 2> : k_emplar_scaled<real, 1, 3>);
  else
    kscaled = mode == 1 ? k_emplar_scaled<real, 0, 1>
                        : (mode == 2 ? k_emplar_scaled<real, 0, 2> : k_emplar_scaled<real, 0, 3>);
  A.rows = energy_rows(c, c.nwave);
  if (c.chgpen) {  // HIPPO: every pair, scaled ones included, in one kernel
    KScope ks_(c, KT_EMPLAR);
    void (*kc)(EmplarArgs<real>, const double*) = nullptr;
    if (c.ewald)
      kc = mode == 1 ? k_empole_chgpen<real, 1, 1> : (mode == 2 ? k_empole_chgpen<real, 1, 2> : k_empole_chgpen<real, 1, 3>);
    else
      kc = mode == 1 ? k_empole_chgpen<real, 0, 1> : (mode == 2 ? k_empole_chgpen<real, 0, 2> : k_empole_chgpen<real, 0, 3>);
    hipLaunchKernelGGL(kc, dim3(c.nwave), dim3(64 * kEmplarParts), 0, c.stream, A, c.chgpen_o.as<double>());
    T9_LAUNCH_CHECK();
    return;
  }
  {
    KScope ks_(c, KT_EMPLAR);
    hipLaunchKernelGGL(kern, dim3(c.nwave), dim3(64 * kEmplarParts), 0, c.stream, A);
    T9_LAUNCH_CHECK();
  }
  A.rows = energy_rows(c, nb);
  hipLaunchKernelGGL(kscaled, dim3(nb), dim3(256), 0, c.stream, A);
  T9_LAUNCH_CHECK();
}

template <class real>
GridGeom ggeom(const Ctx& c) {
  GridGeom g;
  for (int d = 0; d < 3; ++d) {
    g.K[d] = c.grid[d];
    g.a[d] = c.grid[d] / c.box[d];
  }
  return g;
}

template <class real>
void recip_t(Ctx& c, bool use_m, bool use_p, bool want_v) {
  const int n = c.n, nb = ceil_div(n, 256);
  const double f = kCoul;
  pme_bin(c);  // B-spline bases of the current coordinates (valid, and waited for)
  // permanent multipoles: the direct field of the same multipoles left its
  // convolved grid; spread and transform again only for the virial sums
  const void* mgrid = c.qgrid.p;
  if (!(want_v && use_m) && c.qgrid_m_stamp == c.rpole_stamp && c.qgrid_m.p) {
    mgrid = c.qgrid_m.p;
  } else {
    pme_spread(c, 1, c.rpole.p, nullptr);
    pme_fft_virial(c, 1, (want_v && use_m) ? f : 0.0);
  }
  {
    KScope ks_(c, KT_GATHER);
    hipLaunchKernelGGL(k_recip_mpole<real>, dim3(nb), dim3(256), 0, c.stream, n, ggeom<real>(c),
                       c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), mgrid, c.rpole.as<real>(),
                       use_m ? 1 : 0, c.egrad.as<double>(), c.etq.as<double>(),
                       use_p ? c.ecphim.as<double>() : nullptr, energy_rows(c, nb));
    T9_LAUNCH_CHECK();
  }
  if (!use_p) return;
  // induced dipoles (d real, p imaginary)
  pme_spread(c, 0, c.uind.p, c.uinp.p);
  pme_fft_virial(c, 1, 0.0);
  {
    const double self3 = 4.0 / 3.0 * c.aewald * c.aewald * c.aewald / kSqrtPiD;
    KScope ks_(c, KT_GATHER);
    hipLaunchKernelGGL(k_recip_polar<real>, dim3(nb), dim3(256), 0, c.stream, n, ggeom<real>(c),
                       c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), (const void*)c.qgrid.p, c.rpole.as<real>(),
                       c.uind.as<real>(), c.uinp.as<real>(), c.ecphim.as<double>(), self3, c.egrad.as<double>(),
                       c.etq.as<double>(), energy_rows(c, nb));
    T9_LAUNCH_CHECK();
  }
  if (!want_v) return;
  // convolution virial of 1/2 E(u_d+u_p, M) + 1/2 E(u_d, u_p) by the
  // polarization identity: 1/4 [V(X+M) + V(u_d+u_p)] - 1/4 [V(X-M) + V(u_d-u_p)]
  real* z10 = c.escr.as<real>();
  real* z3 = z10 + 10 * (size_t)n;
  for (int sg = 1; sg >= -1; sg -= 2) {
    hipLaunchKernelGGL(k_make_z<real>, dim3(nb), dim3(256), 0, c.stream, n, (real)sg, c.rpole.as<real>(),
                       c.uind.as<real>(), c.uinp.as<real>(), z10, z3);
    T9_LAUNCH_CHECK();
    pme_spread(c, 2, z10, z3);
    pme_fft_virial(c, 0, 0.25 * f * sg);
  }
}

template <class real>
void energy_t(Ctx& c, int vers, int terms, double* out16) {
  const bool use_m = c.opt.use_mpole && (terms & TERM_ELEC), use_p = c.opt.use_polar && (terms & TERM_ELEC);
  const bool want_v = (vers & 4) != 0;
  const int n = c.n, nb = ceil_div(n, 256);
  if (!c.vdw_pending) c.erows = 0;  // else the rows after the pending vdW kernel's
  const bool rep = (terms & TERM_REPEL) && c.use_repel;
  if (use_m || use_p) {
    real_space_t<real>(c, use_m, use_p);
    if (c.ewald) recip_t<real>(c, use_m, use_p, want_v);
    hipLaunchKernelGGL(k_self_polar<real>, dim3(nb), dim3(256), 0, c.stream, n, use_m ? 1 : 0, use_p ? 1 : 0,
                       c.ewald ? 1 : 0, c.aewald, c.rpole.as<real>(), c.uind.as<real>(), c.udirp.as<real>(),
                       c.polinv.as<real>(), energy_rows(c, nb));
    T9_LAUNCH_CHECK();
  } else {
    T9_HIP(hipMemsetAsync(c.egrad.p, 0, sizeof(double) * 3 * n, c.stream));
    if (rep) T9_HIP(hipMemsetAsync(c.etq.p, 0, sizeof(double) * 9 * n, c.stream));
  }
  if (rep) repel_run(c, want_v);
  if (use_m || use_p || rep) {
    const double lx = c.periodic ? c.box[0] : 1.0, ly = c.periodic ? c.box[1] : 1.0,
                 lz = c.periodic ? c.box[2] : 1.0;
    hipLaunchKernelGGL(k_torque<real>, dim3(nb), dim3(256), 0, c.stream, n, c.perm.as<int>(), c.xyz.as<double>(),
                       lx, ly, lz, c.periodic ? 1 : 0, c.frame.as<int>(), c.zaxis.as<int>(), c.xaxis.as<int>(),
                       c.yaxis.as<int>(), c.rpole.as<real>(), c.etq.as<double>(), c.eframe.as<double>(),
                       energy_rows(c, nb));
    T9_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_frame_gather, dim3(nb), dim3(256), 0, c.stream, n, c.perm.as<int>(), c.iperm.as<int>(),
                       c.rev_off.as<int>(), c.rev_ent.as<int>(), c.eframe.as<double>(), c.egrad.as<double>());
    T9_LAUNCH_CHECK();
  }
  if ((terms & TERM_CHGTRN) && c.use_chgtrn) chgtrn_run(c, want_v);
  if ((terms & TERM_DISP) && c.use_disp) disp_run(c, want_v);
  if (c.vdw_pending) {  // join the vdW pair kernel of the second stream
    T9_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
    vdw_finish(c);
    c.vdw_pending = false;
  } else if ((terms & TERM_VDW) && c.use_vdw) {
    vdw_run(c, want_v);
  }
  if ((terms & TERM_VALENCE) && c.use_valence) valence_run(c, c.g_fast.as<double>(), true);
  // without CALC_ENERGY / CALC_VIRIAL nothing comes back to the host (no sync)
  if (c.erows == 0 || !(vers & 5)) {
    for (int q = 0; q < kEnergyRow; ++q) out16[q] = 0;
    return;
  }
  double* dsum = c.scalars.as<double>() + 16;
  hipLaunchKernelGGL(k_rows_reduce, dim3(1), dim3(256), 0, c.stream, c.erows, c.epart.as<double>(), dsum);
  T9_LAUNCH_CHECK();
  T9_HIP(hipMemcpyAsync(c.h_scalars + 16, dsum, sizeof(double) * EP_N, hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  for (int q = 0; q < kEnergyRow; ++q) out16[q] = q < EP_N ? c.h_scalars[16 + q] : 0.0;
}

}  // namespace

double* energy_rows(Ctx& c, int nrows) {
  if (c.erows + nrows > c.ecap) throw ArgError("energy reduction buffer too small (internal sizing error)");
  double* p = c.epart.as<double>() + (size_t)c.erows * kRow;
  c.erows += nrows;
  return p;
}

// Allocates the energy-path buffers and the reverse frame table.
void energy_setup(Ctx& c, const System& s) {
  const int n = c.n;
  c.egrad.alloc(sizeof(double) * 3 * n);
  c.etq.alloc(sizeof(double) * 9 * n);
  c.eframe.alloc(sizeof(double) * 12 * n);
  if (c.ewald) {
    c.ecphim.alloc(sizeof(double) * 10 * n);
    c.escr.alloc(rsize(c) * 13 * (size_t)n);
  }
  // reverse table: for each atom, the (frame owner, slot) pairs naming it
  std::vector<std::vector<int>> rev(n);
  for (int i = 0; i < n; ++i) {
    const int fr = s.frame[i];
    if (fr == 0) continue;
    rev[i].push_back(i << 2);
    rev[s.zaxis[i]].push_back((i << 2) | 1);
    if (fr != 1 && s.xaxis[i] >= 0) rev[s.xaxis[i]].push_back((i << 2) | 2);
    if ((fr == 4 || fr == 5) && s.yaxis[i] >= 0) rev[s.yaxis[i]].push_back((i << 2) | 3);
  }
  std::vector<int> off(n + 1, 0), ent;
  for (int j = 0; j < n; ++j) {
    std::sort(rev[j].begin(), rev[j].end());
    off[j + 1] = off[j] + (int)rev[j].size();
    ent.insert(ent.end(), rev[j].begin(), rev[j].end());
  }
  if (ent.empty()) ent.push_back(0);
  c.rev_off.alloc(sizeof(int) * (n + 1));
  c.rev_ent.alloc(sizeof(int) * ent.size());
  T9_HIP(hipMemcpyAsync(c.rev_off.p, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c.stream));
  T9_HIP(hipMemcpyAsync(c.rev_ent.p, ent.data(), sizeof(int) * ent.size(), hipMemcpyHostToDevice, c.stream));
  // reduction rows: per-atom kernels (real space list and scaled pairs, 2
  // recip, self, torque, vdW, valence) plus up to three convolution passes of one row per z-pass
  // workgroup
  const int nb = ceil_div(n, 256);
  long long lines = 0;
  if (c.ewald) lines = 3LL * ((c.grid[0] + 7) / 8) * c.grid[1] + 16;
  const long long nterm = (long long)(s.bond_kr.size() + s.angle_kt.size() + s.urey_kr.size()) / 2;
  // list loops: one row per 64 atoms; valence: one row per 256 terms
  c.ecap = (int)(9LL * nb + 5LL * ceil_div(n, 64) + (nterm + 255) / 256 + lines + 16);
  c.epart.alloc(sizeof(double) * kRow * (size_t)c.ecap);
  T9_HIP(hipStreamSynchronize(c.stream));
}

namespace {

// Interaction counts of the analyze mode over the in-cutoff list, each pair
// once (j > i): multipole pairs unless their m scale is 0; polarization
// pairs unless their p scale is 0 or neither atom is polarizable (the pair
// energy is then exactly zero, which the reference does not count).
template <class real>
__global__ __launch_bounds__(256) void k_counts(int n, const int* __restrict__ ncut,
                                                const long long* __restrict__ woff,
                                                const uint32_t* __restrict__ ecut,
                                                const real* __restrict__ polarity, ScaleTable sc,
                                                unsigned long long* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cm = 0, cp = 0;
  if (s < n) {
    const int len = ncut[s];
    const uint32_t* e = ecut + woff[s >> 6] + (s & 63);
    const bool pol_i = polarity[s] != 0;
    for (int k = 0; k < len; ++k) {
      const uint32_t ent = e[(long long)k * 64];
      const int j = (int)(ent >> kCodeBits);
      if (j <= s) continue;
      const int code = (int)(ent & (kMaxCodes - 1));
      if (code == 0 || sc.m[code] != 0) ++cm;
      if ((code == 0 || sc.p[code] != 0) && (pol_i || polarity[j] != 0)) ++cp;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    cm += __shfl_xor(cm, o, 64);
    cp += __shfl_xor(cp, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], cm);
    atomicAdd(&out[1], cp);
  }
}

}  // namespace

// Analyze-mode interaction counts of the multipole and polarization terms
// (the reference's nem / nep of energy(calc::v3): src/acc/amoeba/
// empoleewald.cpp:78-79,132-133,177-180 and epolarewald.cpp:128-130,
// 280-282,499-500): pairs within the cutoff, less excluded (scale 0) pairs,
// plus one self term per atom with Ewald.
void count_interactions(Ctx& c, long long* nem, long long* nep) {
  const auto& L = c.mlist;
  if (!L.cut_valid) throw ArgError("in-cutoff neighbor list is stale (internal ordering error)");
  DBuf out;
  out.alloc(sizeof(unsigned long long) * 2);
  T9_HIP(hipMemsetAsync(out.p, 0, sizeof(unsigned long long) * 2, c.stream));
  if (c.prec == 32)
    hipLaunchKernelGGL(k_counts<float>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, L.ncut.as<int>(),
                       L.off.as<long long>(), L.ecut.as<uint32_t>(), c.polarity.as<float>(), c.scale,
                       out.as<unsigned long long>());
  else
    hipLaunchKernelGGL(k_counts<double>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, L.ncut.as<int>(),
                       L.off.as<long long>(), L.ecut.as<uint32_t>(), c.polarity.as<double>(), c.scale,
                       out.as<unsigned long long>());
  T9_LAUNCH_CHECK();
  unsigned long long h[2] = {0, 0};
  T9_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  const long long self = c.ewald ? c.n : 0;
  *nem = c.opt.use_mpole ? (long long)h[0] + self : 0;
  *nep = c.opt.use_polar ? (long long)h[1] + self : 0;
}

namespace {
HippoPairArgs hippo_args(Ctx& c, int kind) {
  const auto& L = c.mlist;
  HippoPairArgs A;
  A.n = c.n;
  A.nlist = L.nlist.as<int>();
  A.woff = L.off.as<long long>();
  A.entries = L.entries.as<uint32_t>();
  A.perm = c.perm.as<int>();
  const double cut = kind == 0 ? c.ct_cut : c.disp_cut, taper = kind == 0 ? c.ct_taper : c.disp_taper;
  A.par = kind == 0 ? c.chgtrn_o.as<double>() : c.disp_o.as<double>();
  for (int k = 0; k < kMaxCodes; ++k) A.sc[k] = k == 0 ? 1.0 : (kind == 0 ? c.scale.m[k] : c.scale.ds[k]);
  A.off2 = cut * cut;
  A.taper = taper;
  A.inv_w = cut > taper ? 1.0 / (cut - taper) : 0.0;
  A.slot = kind == 0 ? EP_ECT : EP_EDISP;
  A.grad = c.egrad.as<double>();
  A.rows = nullptr;
  return A;
}

template <class real, int KIND>
void hippo_run_t(Ctx& c, bool vir) {
  HippoPairArgs A = hippo_args(c, KIND);
  A.rows = energy_rows(c, c.nwave);
  const PosRec<real>* pos = c.pos.as<PosRec<real>>();
  const BoxT<real> box = make_boxT<real>(c);
  if (vir)
    hipLaunchKernelGGL((k_hippo_pair<real, true, KIND>), dim3(c.nwave), dim3(64 * kPairParts), 0, c.stream, A, pos, box);
  else
    hipLaunchKernelGGL((k_hippo_pair<real, false, KIND>), dim3(c.nwave), dim3(64 * kPairParts), 0, c.stream, A, pos,
                       box);
  T9_LAUNCH_CHECK();
}

template <class real, int KIND>
long long hippo_count_t(Ctx& c) {
  DBuf out;
  out.alloc(sizeof(unsigned long long));
  T9_HIP(hipMemsetAsync(out.p, 0, sizeof(unsigned long long), c.stream));
  hipLaunchKernelGGL((k_hippo_count<real, KIND>), dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream,
                     hippo_args(c, KIND), c.pos.as<PosRec<real>>(), make_boxT<real>(c), out.as<unsigned long long>());
  T9_LAUNCH_CHECK();
  unsigned long long h = 0;
  T9_HIP(hipMemcpyAsync(&h, out.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  return (long long)h;
}
}  // namespace

// HIPPO charge transfer and dispersion after the electrostatic terms (which
// set egrad).
void chgtrn_run(Ctx& c, bool vir) {
  if (c.prec == 32) hippo_run_t<float, 0>(c, vir);
  else hippo_run_t<double, 0>(c, vir);
}

long long chgtrn_count(Ctx& c) {
  if (!c.use_chgtrn) return 0;
  return c.prec == 32 ? hippo_count_t<float, 0>(c) : hippo_count_t<double, 0>(c);
}

namespace {
template <class real>
RepelArgs repel_args(Ctx& c) {
  const auto& L = c.mlist;
  RepelArgs A;
  A.n = c.n;
  A.nlist = L.nlist.as<int>();
  A.woff = L.off.as<long long>();
  A.entries = L.entries.as<uint32_t>();
  A.perm = c.perm.as<int>();
  A.rp = c.repel_o.as<double>();
  for (int k = 0; k < kMaxCodes; ++k) A.sc[k] = k == 0 ? 1.0 : c.scale.rs[k];
  A.off2 = c.rep_cut * c.rep_cut;
  A.taper = c.rep_taper;
  A.inv_w = c.rep_cut > c.rep_taper ? 1.0 / (c.rep_cut - c.rep_taper) : 0.0;
  A.grad = c.egrad.as<double>();
  A.tq = c.etq.as<double>();
  A.rows = nullptr;
  return A;
}

template <class real>
void repel_run_t(Ctx& c) {
  RepelArgs A = repel_args<real>(c);
  A.rows = energy_rows(c, c.nwave);
  hipLaunchKernelGGL(k_repel<real>, dim3(c.nwave), dim3(64 * kEmplarParts), 0, c.stream, A,
                     c.pos.as<PosRec<real>>(), make_boxT<real>(c), c.rpole.as<real>());
  T9_LAUNCH_CHECK();
}

template <class real>
long long repel_count_t(Ctx& c) {
  DBuf out;
  out.alloc(sizeof(unsigned long long));
  T9_HIP(hipMemsetAsync(out.p, 0, sizeof(unsigned long long), c.stream));
  hipLaunchKernelGGL(k_repel_count<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, repel_args<real>(c),
                     c.pos.as<PosRec<real>>(), make_boxT<real>(c), c.rpole.as<real>(), out.as<unsigned long long>());
  T9_LAUNCH_CHECK();
  unsigned long long h = 0;
  T9_HIP(hipMemcpyAsync(&h, out.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  return (long long)h;
}
}  // namespace

// HIPPO repulsion (the virial rows are written in both forms; the torque
// kernel adds the frame part).
void repel_run(Ctx& c, bool /*vir*/) {
  if (c.prec == 32) repel_run_t<float>(c);
  else repel_run_t<double>(c);
}

long long repel_count(Ctx& c) {
  if (!c.use_repel) return 0;
  return c.prec == 32 ? repel_count_t<float>(c) : repel_count_t<double>(c);
}

void disp_run(Ctx& c, bool vir) {
  if (c.prec == 32) hippo_run_t<float, 1>(c, vir);
  else hippo_run_t<double, 1>(c, vir);
}

long long disp_count(Ctx& c) {
  if (!c.use_disp) return 0;
  return c.prec == 32 ? hippo_count_t<float, 1>(c) : hippo_count_t<double, 1>(c);
}

void energy_run(Ctx& c, int vers, int terms, double* out16, double* grad_orig, bool grad_on_device) {
  if (c.prec == 32) energy_t<float>(c, vers, terms, out16);
  else energy_t<double>(c, vers, terms, out16);
  if (grad_orig) {
    const bool val = (terms & TERM_VALENCE) && c.use_valence;
    DBuf tmp;
    if (!grad_on_device) tmp.alloc(sizeof(double) * 3 * c.n);
    double* dst = grad_on_device ? grad_orig : tmp.as<double>();
    hipLaunchKernelGGL(k_grad_to_orig, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.as<int>(),
                       c.egrad.as<double>(), val ? c.g_fast.as<double>() : nullptr, dst);
    T9_LAUNCH_CHECK();
    if (!grad_on_device)
      T9_HIP(hipMemcpyAsync(grad_orig, tmp.p, sizeof(double) * 3 * c.n, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
  }
}

}  // namespace t9
