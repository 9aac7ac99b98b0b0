// Molecular-dynamics terms around the electrostatics, for gfx950: AMOEBA
// buffered 14-7 vdW on reduced hydrogen sites, the valence terms of the water
// model, and a RESPA integrator with the Bussi thermostat.
//
// Reference behaviour followed: include/seq/pair_hal.h (buffered 14-7),
// include/math/switch.h (5th-order taper), src/cu/evdw.cu / ehal (site
// reduction), include/seq/bond.h, angle.h, urey.h (valence polynomials),
// src/md/integrator.cpp:70-165 and src/md/propagator.cpp (RESPA splitting:
// valence terms in the inner loop), src/md/thermostat.cpp (Bussi).
//
// vdW: one lane per sorted site over the site list (full, both directions),
// pair scale by the entry's code, FP64 lane accumulation of the site
// gradient, half the pair energy and half the pair virial; sites then hand
// their gradient to the atoms (reduced hydrogen: red to itself, 1 - red to the
// parent) through a fixed-order gather.  Valence: one thread per atom walks the
// terms that contain it and keeps its own share; the term's first atom adds
// the energy and virial.  No atomics anywhere, bit-reproducible.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>

#include "ctx.hpp"

namespace t9 {
namespace {

constexpr double kEkcalD = kEkcal;
constexpr double kGasR = kGasConst;

template <class real>
__global__ void k_vsite_pos(int n, const int* __restrict__ perm, const double* __restrict__ xyz,
                            const int* __restrict__ vparent, const double* __restrict__ vred,
                            const int* __restrict__ jvdw_o, double3 vorg, double3 vlen, double3 box, int periodic,
                            PosRec<real>* __restrict__ vpos, int* __restrict__ jvdw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s], p = vparent[i];
  const double L[3] = {box.x, box.y, box.z};
  double x[3];
  for (int d = 0; d < 3; ++d) {
    x[d] = xyz[3 * i + d];
    if (p != i) {
      double dd = xyz[3 * i + d] - xyz[3 * p + d];
      if (periodic) dd -= L[d] * rint(dd / L[d]);
      x[d] = xyz[3 * p + d] + vred[i] * dd;
    }
  }
  const double org[3] = {vorg.x, vorg.y, vorg.z}, len[3] = {vlen.x, vlen.y, vlen.z};
  vpos[s] = make_pos<real>(x[0], x[1], x[2], 0.0, org, len, periodic);
  jvdw[s] = jvdw_o[i];
}

struct VScale {
  double s[kMaxCodes];
};

template <class real>
struct VdwArgs {
  int n;
  const int* nlist;
  const long long* woff;
  const uint32_t* entries;
  const PosRec<real>* pos;
  const int* jvdw;
  const real* irv;  // [nj*nj][2] 1 / R_min as hi + lo
  const real* eps;  // [nj*nj][2] epsilon as hi + lo
  int nj;
  double vscale[kMaxCodes];
  BoxT<real> box;
  real cut2, cut, taper;
  double* gsite;  // [n][3] sorted
  double* rows;
};

// VIR: accumulate the virial (only when it is requested).  The site gradient
// is summed in FP32 over runs of 16 list entries and flushed to FP64.  The
// kPairParts waves of a workgroup split each site's list (entries
// p, p + kPairParts, ...) and their sums are added in part order.
template <class real, bool VIR>
__global__ __launch_bounds__(64 * kPairParts) void k_vdw(VdwArgs<real> A) {
  __shared__ double sh[parts_lds<kPairParts, 10>()];
  const int part = threadIdx.x >> 6;
  const int blk = xcd_block();
  const int s = blk * 64 + (threadIdx.x & 63);
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // e, g[3], v[6]
  double& e = acc[0];
  double* g = acc + 1;
  double* v = acc + 4;
  if (s < A.n) {
    const PosRec<real> pi = A.pos[s];
    const int ji = A.jvdw[s] * A.nj;
    const int len = A.nlist[s];
    const uint32_t* en = A.entries + A.woff[s >> 6] + (s & 63);
    const real inv_tc = 1 / (A.taper - A.cut);
    real gs[3] = {0, 0, 0};
    uint32_t nxt = part < len ? en[(long long)part * 64] : 0;
    for (int kk = part; kk < len; kk += kPairParts) {
      const uint32_t ent = nxt;
      if (kk + kPairParts < len) nxt = en[(long long)(kk + kPairParts) * 64];
      if ((kk & 63) >= 64 - kPairParts) {  // every 16 entries of this part
        for (int d = 0; d < 3; ++d) {
          g[d] += (double)gs[d];
          gs[d] = 0;
        }
      }
      const int k = (int)(ent >> kCodeBits);
      const int code = (int)(ent & (kMaxCodes - 1));
      const real sc = code ? (real)A.vscale[code] : (real)1;
      if (sc == 0) continue;
      const PosRec<real> pk = A.pos[k];
      real R[3];
      pair_disp(pk, pi, A.box, R[0], R[1], R[2]);  // site i - site k
      const real r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      if (!(r2 <= A.cut2)) continue;
      // refined hardware reciprocal square root and reciprocals: errors of
      // about half an ulp either way, so the per-pair rounding stays unbiased
      // (the raw hardware estimates are not)
      const real r = t9_sqrt_nr(r2);
      const real rinv = t9_rcp_nr(r);
      const int t = ji + A.jvdw[k];
      // the pair parameters are shared by every pair of two types, so their
      // rounding would bias the whole sum: hi + lo split, as for coordinates
      const real ivh = A.irv[2 * t], ivl = A.irv[2 * t + 1];
      const real eh = A.eps[2 * t] * sc, el = A.eps[2 * t + 1] * sc;
      const real rho = fma(r, ivh, r * ivl);
      const real rho6 = rho * rho * rho * rho * rho * rho, rho7 = rho6 * rho;
      const real d7 = rho + (real)0.07;
      const real y7 = t9_rcp_nr(d7);
      const real b = t9_div_nr((real)1.07, d7, y7);
      const real b2 = b * b, b3 = b2 * b;
      const real t1 = b3 * b3 * b;
      const real s2 = t9_rcp_nr(rho7 + (real)0.12);
      const real t2 = (real)1.12 * s2;
      const real u = t1 * (t2 - 2);
      real en_ = fma(u, eh, u * el);
      real de = eh * ivh * (-7 * t1 * y7 * (t2 - 2) - 7 * t1 * rho6 * t2 * s2);
      if (r > A.taper) {
        const real x = (r - A.cut) * inv_tc;
        const real S = x * x * x * (x * (6 * x - 15) + 10);
        const real dS = 30 * (x * (1 - x)) * (x * (1 - x)) * inv_tc;
        de = de * S + en_ * dS;
        en_ *= S;
      }
      e += (double)en_;
      const real f = de * rinv;
      for (int d = 0; d < 3; ++d) gs[d] += f * R[d];
      if (VIR) {
        v[0] += (double)(f * R[0] * R[0]);
        v[1] += (double)(f * R[1] * R[1]);
        v[2] += (double)(f * R[2] * R[2]);
        v[3] += (double)(f * R[0] * R[1]);
        v[4] += (double)(f * R[0] * R[2]);
        v[5] += (double)(f * R[1] * R[2]);
      }
    }
    for (int d = 0; d < 3; ++d) g[d] += (double)gs[d];
  }
  parts_reduce<kPairParts, 10>(acc, sh);
  double red[EP_N];
  for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (part == 0) {
    if (s < A.n)
      for (int d = 0; d < 3; ++d) A.gsite[3 * (size_t)s + d] = g[d];
    red[EP_EVDW] = 0.5 * e;
    for (int q = 0; q < 6; ++q) red[EP_VXX + q] = 0.5 * v[q];
  }
  block_row<EP_N>(red, A.rows + (size_t)blk * kEnergyRow);
}

// atom gradient (sorted egrad) += own site share + shares of the reduced sites
// that hang on it
__global__ void k_vdw_to_atoms(int n, const int* __restrict__ perm, const int* __restrict__ iperm,
                               const int* __restrict__ vparent, const double* __restrict__ vred,
                               const int* __restrict__ coff, const int* __restrict__ child,
                               const double* __restrict__ gsite, double* __restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int j = perm[t];
  const double w = vparent[j] != j ? vred[j] : 1.0;
  double g[3];
  for (int d = 0; d < 3; ++d) g[d] = w * gsite[3 * (size_t)t + d];
  for (int q = coff[j]; q < coff[j + 1]; ++q) {
    const int c = child[q];
    const int sc = iperm[c];
    const double wc = 1.0 - vred[c];
    for (int d = 0; d < 3; ++d) g[d] += wc * gsite[3 * (size_t)sc + d];
  }
  for (int d = 0; d < 3; ++d) grad[3 * (size_t)t + d] += g[d];
}

// positions by atom index: the coordinate array, or a molecule staged in LDS
struct GPos {
  const double* xyz;
  __device__ __forceinline__ double operator()(int a, int d) const { return xyz[3 * a + d]; }
};
struct LdsPos {  // a molecule's positions in LDS, slot k holding atom id[k]
  const double* s;
  int id0, id1, id2;
  __device__ __forceinline__ double operator()(int a, int d) const {
    const int k = a == id0 ? 0 : (a == id1 ? 1 : (a == id2 ? 2 : 3));
    return s[3 * k + d];
  }
};

template <class Pos>
__device__ __forceinline__ void mdiff(const Pos& P, int a, int b, const double* L, int periodic, double* v) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double x = P(a, d) - P(b, d);
    if (periodic) x -= L[d] * rint(x / L[d]);
    v[d] = x;
  }
}

struct ValArgs {
  int n;
  const double* xyz;
  double L[3];
  int periodic;
  const int* toff;
  const int* term;  // (kind << 28) | (index << 2) | role
  const int* bij;
  const double* bkr;
  const int* aabc;
  const double* akt;
  const int* uac;
  const double* ukr;
  double cb, qb, ca, qa, pa, sa;
  double* grad;  // caller order [n][3]
  double* rows;
};

// Gradient of one valence term for its atoms (up to 3), energy, and the
// relative positions used by the virial.
template <class Pos>
__device__ __forceinline__ void valence_term(const ValArgs& A, const Pos& P, int kind, int idx, double* e,
                                             double g[3][3], double rel[3][3], int* at, int* nat) {
  if (kind == 0 || kind == 2) {  // bond or Urey-Bradley: stretch polynomial
    const int* ij = kind == 0 ? A.bij + 2 * idx : A.uac + 2 * idx;
    const double* kr = kind == 0 ? A.bkr + 2 * idx : A.ukr + 2 * idx;
    const double c = kind == 0 ? A.cb : 0.0, q = kind == 0 ? A.qb : 0.0;
    at[0] = ij[0];
    at[1] = ij[1];
    *nat = 2;
    double R[3];
    mdiff(P, ij[0], ij[1], A.L, A.periodic, R);
    const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    const double d = r - kr[1];
    *e = kr[0] * d * d * (1 + c * d + q * d * d);
    const double de = kr[0] * d * (2 + 3 * c * d + 4 * q * d * d) / r;
    for (int a = 0; a < 3; ++a) {
      g[0][a] = de * R[a];
      g[1][a] = -de * R[a];
      rel[0][a] = R[a];  // relative to atom 1
      rel[1][a] = 0;
    }
  } else {  // angle a-b-c, b the vertex
    const int* abc = A.aabc + 3 * idx;
    at[0] = abc[0];
    at[1] = abc[1];
    at[2] = abc[2];
    *nat = 3;
    double u[3], v[3];
    mdiff(P, abc[0], abc[1], A.L, A.periodic, u);
    mdiff(P, abc[2], abc[1], A.L, A.periodic, v);
    const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double lv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double cs = (u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) / (lu * lv);
    cs = fmin(1.0, fmax(-1.0, cs));
    const double th = acos(cs), sn = sin(th);
    const double deg = 57.29577951308232087680;
    const double d = th * deg - A.akt[2 * idx + 1];
    const double k = A.akt[2 * idx] / (deg * deg);
    *e = k * d * d * (1 + A.ca * d + A.qa * d * d + A.pa * d * d * d + A.sa * d * d * d * d);
    const double dedth =
        k * d * (2 + 3 * A.ca * d + 4 * A.qa * d * d + 5 * A.pa * d * d * d + 6 * A.sa * d * d * d * d) * deg;
    for (int q = 0; q < 3; ++q) {
      g[0][q] = dedth * (cs * u[q] / lu - v[q] / lv) / (lu * sn);
      g[2][q] = dedth * (cs * v[q] / lv - u[q] / lu) / (lv * sn);
      g[1][q] = -(g[0][q] + g[2][q]);
      rel[0][q] = u[q];  // relative to the vertex
      rel[1][q] = 0;
      rel[2][q] = v[q];
    }
  }
}

// One thread per term (bonds, then angles, then Urey-Bradley): the term's
// gradient on each of its atoms goes to tg[term][slot][3]; with ENE the
// term's energy and virial go to the block's partial row.
template <bool ENE>
__global__ __launch_bounds__(256) void k_valence_terms(ValArgs A, int nb, int na, int nu, double* __restrict__ tg) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double red[EP_N];
  if (ENE)
    for (int q = 0; q < EP_N; ++q) red[q] = 0;
  if (t < nb + na + nu) {
    const int kind = t < nb ? 0 : (t < nb + na ? 1 : 2);
    const int idx = kind == 0 ? t : (kind == 1 ? t - nb : t - nb - na);
    double e, g[3][3], rel[3][3];
    int at[3], nat;
    valence_term(A, GPos{A.xyz}, kind, idx, &e, g, rel, at, &nat);
    for (int m = 0; m < nat; ++m)
      for (int d = 0; d < 3; ++d) tg[9 * (size_t)t + 3 * m + d] = g[m][d];
    if (ENE) {
      red[kind == 0 ? EP_EBOND : (kind == 1 ? EP_EANGLE : EP_EUREY)] = e;
      double W[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int m = 0; m < nat; ++m)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) W[a][b] += g[m][a] * rel[m][b];
      red[EP_VXX] = W[0][0];
      red[EP_VYY] = W[1][1];
      red[EP_VZZ] = W[2][2];
      red[EP_VXY] = 0.5 * (W[0][1] + W[1][0]);
      red[EP_VXZ] = 0.5 * (W[0][2] + W[2][0]);
      red[EP_VYZ] = 0.5 * (W[1][2] + W[2][1]);
    }
  }
  if (ENE) block_row<EP_N>(red, A.rows + (size_t)blockIdx.x * kEnergyRow);
}

// Each atom sums the gradients its terms left for it, in its term order.
__global__ __launch_bounds__(256) void k_valence_gather(int n, const int* __restrict__ toff,
                                                        const int* __restrict__ term, int nb, int na,
                                                        const double* __restrict__ tg, double* __restrict__ grad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double g0 = 0, g1 = 0, g2 = 0;
  for (int q = toff[j]; q < toff[j + 1]; ++q) {
    const int tt = term[q];
    const int kind = tt >> 28, idx = (tt >> 2) & 0x3ffffff, role = tt & 3;
    const int t = kind == 0 ? idx : (kind == 1 ? nb + idx : nb + na + idx);
    const double* p = tg + 9 * (size_t)t + 3 * role;
    g0 += p[0];
    g1 += p[1];
    g2 += p[2];
  }
  grad[3 * (size_t)j] = g0;
  grad[3 * (size_t)j + 1] = g1;
  grad[3 * (size_t)j + 2] = g2;
}

// Inner RESPA step after the term kernel: gather the valence gradient of
// each atom (as k_valence_gather), kick v -= h g / m, then drift x += h v.
__global__ __launch_bounds__(256) void k_valence_gather_kick_drift(int n, const int* __restrict__ toff,
                                                                   const int* __restrict__ term, int nb, int na,
                                                                   const double* __restrict__ tg,
                                                                   double* __restrict__ grad,
                                                                   const double* __restrict__ mass, double h,
                                                                   double* __restrict__ v, double* __restrict__ x) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double g[3] = {0, 0, 0};
  for (int q = toff[j]; q < toff[j + 1]; ++q) {
    const int tt = term[q];
    const int kind = tt >> 28, idx = (tt >> 2) & 0x3ffffff, role = tt & 3;
    const int t = kind == 0 ? idx : (kind == 1 ? nb + idx : nb + na + idx);
    const double* p = tg + 9 * (size_t)t + 3 * role;
    for (int d = 0; d < 3; ++d) g[d] += p[d];
  }
  const double m = mass[j];
  const double c = m > 0 ? h * kEkcal / m : 0.0;
  for (int d = 0; d < 3; ++d) {
    grad[3 * (size_t)j + d] = g[d];
    const double vv = v[3 * (size_t)j + d] - c * g[d];
    v[3 * (size_t)j + d] = vv;
    x[3 * (size_t)j + d] += h * vv;
  }
}

// The whole RESPA inner loop of one outer step, one thread per molecule
// (the fast forces couple only atoms of one molecule): drift, then nrep times
// the molecule's valence terms, the gather of each atom's term shares, the
// inner kick and the next drift.  The same arithmetic, in the same order, as
// k_valence_terms<false> + k_valence_gather_kick_drift launched per inner
// step; the thread reads the coordinates it has just written, so x and v
// carry no restrict.
__global__ __launch_bounds__(256) void k_respa_inner(ValArgs A, int nmol, const int* __restrict__ mol_off,
                                                     const int* __restrict__ mol_atom,
                                                     const int* __restrict__ mol_toff,
                                                     const int* __restrict__ mol_term, int nb, int na, double* tg,
                                                     double* grad, const double* __restrict__ mass, double h, int nrep,
                                                     double* v, double* x) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  const int a0 = mol_off[m], a1 = mol_off[m + 1];
  for (int q = a0; q < a1; ++q) {
    const int i = mol_atom[q];
    for (int d = 0; d < 3; ++d) x[3 * i + d] += h * v[3 * i + d];
  }
  const int t0 = mol_toff[m], t1 = mol_toff[m + 1];
  for (int rep = 0; rep < nrep; ++rep) {
    for (int q = t0; q < t1; ++q) {
      const int t = mol_term[q];
      const int kind = t < nb ? 0 : (t < nb + na ? 1 : 2);
      const int idx = kind == 0 ? t : (kind == 1 ? t - nb : t - nb - na);
      double e, g[3][3], rel[3][3];
      int at[3], nat;
      valence_term(A, GPos{A.xyz}, kind, idx, &e, g, rel, at, &nat);
      for (int r = 0; r < nat; ++r)
        for (int d = 0; d < 3; ++d) tg[9 * (size_t)t + 3 * r + d] = g[r][d];
    }
    for (int q = a0; q < a1; ++q) {
      const int j = mol_atom[q];
      double gj[3] = {0, 0, 0};
      for (int p = A.toff[j]; p < A.toff[j + 1]; ++p) {
        const int tt = A.term[p];
        const int kind = tt >> 28, idx = (tt >> 2) & 0x3ffffff, role = tt & 3;
        const int t = kind == 0 ? idx : (kind == 1 ? nb + idx : nb + na + idx);
        const double* pp = tg + 9 * (size_t)t + 3 * role;
        for (int d = 0; d < 3; ++d) gj[d] += pp[d];
      }
      const double mj = mass[j];
      const double c = mj > 0 ? h * kEkcal / mj : 0.0;
      for (int d = 0; d < 3; ++d) {
        grad[3 * (size_t)j + d] = gj[d];
        const double vv = v[3 * (size_t)j + d] - c * gj[d];
        v[3 * (size_t)j + d] = vv;
        x[3 * (size_t)j + d] += h * vv;
      }
    }
  }
}

// The same loop for molecules of at most 4 atoms and 6 terms (water, ions):
// a group of 8 lanes per molecule.  Lanes 0..3 hold the atoms' positions and
// velocities in registers for all inner steps; lanes 0..5 each evaluate one
// term.  Per step the atom lanes stage positions in LDS, the term lanes write
// their per-atom gradients to LDS, and each atom lane sums its own terms in
// the order of its term list, as the gather does, then kicks and drifts.
// Positions and velocities are written once at the end.
constexpr int kRespaGrp = 8, kRespaMols = 256 / kRespaGrp;
__global__ __launch_bounds__(256) void k_respa_inner_grp(ValArgs A, int nmol, const int* __restrict__ mol_off,
                                                         const int* __restrict__ mol_atom,
                                                         const int* __restrict__ mol_toff,
                                                         const int* __restrict__ mol_term, int nb, int na,
                                                         double* __restrict__ grad, const double* __restrict__ mass,
                                                         double h, int nrep, double* __restrict__ v,
                                                         double* __restrict__ x) {
  __shared__ double spos[kRespaMols][4][3];
  __shared__ double sg[kRespaMols][6][3][3];
  const int lm = threadIdx.x / kRespaGrp, lane = threadIdx.x % kRespaGrp;
  const int m = blockIdx.x * kRespaMols + lm;
  int a0 = 0, nat_m = 0, t0 = 0, nt_m = 0;
  if (m < nmol) {
    a0 = mol_off[m];
    nat_m = mol_off[m + 1] - a0;
    t0 = mol_toff[m];
    nt_m = mol_toff[m + 1] - t0;
  }
  // atom lane: position, velocity, and the LDS offsets of its terms' rows
  const bool is_atom = lane < nat_m;
  const int i = is_atom ? mol_atom[a0 + lane] : 0;
  double xl[3] = {0, 0, 0}, vl[3] = {0, 0, 0}, cm = 0;
  int ncsr = 0, soff[6];
  if (is_atom) {
    const double mi = mass[i];
    cm = mi > 0 ? h * kEkcal / mi : 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      vl[d] = v[3 * (size_t)i + d];
      xl[d] = x[3 * (size_t)i + d] + h * vl[d];  // the first drift
    }
    const int p0 = A.toff[i];
    ncsr = A.toff[i + 1] - p0;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      soff[p] = 0;
      if (p < ncsr) {
        const int tt = A.term[p0 + p];
        const int kind = tt >> 28, idx = (tt >> 2) & 0x3ffffff, role = tt & 3;
        const int t = kind == 0 ? idx : (kind == 1 ? nb + idx : nb + na + idx);
        int q = 0;
        for (int u = 0; u < nt_m; ++u)
          if (mol_term[t0 + u] == t) q = u;
        soff[p] = 9 * q + 3 * role;
      }
    }
  }
  // term lane: which term, and the atom ids of the molecule's slots
  const bool is_term = lane < nt_m;
  int kind = 0, idx = 0;
  LdsPos P{&spos[lm][0][0], -1, -1, -1};
  if (is_term) {
    const int t = mol_term[t0 + lane];
    kind = t < nb ? 0 : (t < nb + na ? 1 : 2);
    idx = kind == 0 ? t : (kind == 1 ? t - nb : t - nb - na);
    P.id0 = mol_atom[a0];
    P.id1 = nat_m > 1 ? mol_atom[a0 + 1] : -1;
    P.id2 = nat_m > 2 ? mol_atom[a0 + 2] : -1;
  }
  const double* gmol = &sg[lm][0][0][0];
  double gl[3] = {0, 0, 0};
  for (int rep = 0; rep < nrep; ++rep) {
    if (is_atom)
#pragma unroll
      for (int d = 0; d < 3; ++d) spos[lm][lane][d] = xl[d];
    __syncthreads();
    if (is_term) {
      double e, g[3][3], rel[3][3];
      int at[3], nat;
      valence_term(A, P, kind, idx, &e, g, rel, at, &nat);
#pragma unroll
      for (int r = 0; r < 3; ++r)
        if (r < nat)
#pragma unroll
          for (int d = 0; d < 3; ++d) sg[lm][lane][r][d] = g[r][d];
    }
    __syncthreads();
    if (is_atom) {
      gl[0] = gl[1] = gl[2] = 0;
#pragma unroll
      for (int p = 0; p < 6; ++p)
        if (p < ncsr)
#pragma unroll
          for (int d = 0; d < 3; ++d) gl[d] += gmol[soff[p] + d];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        vl[d] -= cm * gl[d];
        xl[d] += h * vl[d];
      }
    }
  }
  if (is_atom) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      x[3 * (size_t)i + d] = xl[d];
      v[3 * (size_t)i + d] = vl[d];
      if (nrep > 0) grad[3 * (size_t)i + d] = gl[d];
    }
  }
}

// ------------------------------------------------------------ integrator

// v -= (h1 g1 + h2 g2) / m (h in ps, g in kcal/mol/Å, v in Å/ps); g2 may be null
__global__ void k_kick(int n, double h1, const double* __restrict__ g1, double h2, const double* __restrict__ g2,
                       const double* __restrict__ mass, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = mass[i];
  if (m <= 0) return;
  const double c = kEkcalD / m;
  for (int d = 0; d < 3; ++d) {
    double a = h1 * g1[3 * i + d];
    if (g2) a += h2 * g2[3 * i + d];
    v[3 * i + d] -= c * a;
  }
}

__global__ void k_drift(int n, double h, const double* __restrict__ v, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) x[3 * i + d] += h * v[3 * i + d];
}

__global__ void k_scale_v(int n, double a, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) v[3 * i + d] *= a;
}

// 2 K (g/mol Å^2/ps^2) partial sums
__global__ __launch_bounds__(256) void k_ekin(int n, const double* __restrict__ mass, const double* __restrict__ v,
                                              double* __restrict__ rows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double red[1] = {0};
  if (i < n) red[0] = mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  block_row<1>(red, rows + blockIdx.x);
}

// out[0] = sum of rows (fixed order: strided partial sums, fixed tree)
__global__ __launch_bounds__(256) void k_sum_rows1(int nrows, const double* __restrict__ rows,
                                                   double* __restrict__ out) {
  __shared__ double sh[256];
  double v = 0;
  for (int r = threadIdx.x; r < nrows; r += 256) v += rows[r];
  const double t = block_sum<256>(v, sh);
  if (threadIdx.x == 0) out[0] = t;
}

// Bussi factor from the kinetic energy rows (src/mdpt.cpp:41-71): the
// host supplies the normal deviate r and the chi-square deviate s2 of
// nfree - 1 degrees of freedom; the factor goes to out[0].
__global__ __launch_bounds__(256) void k_bussi_factor(int nrows, const double* __restrict__ rows, int nfree,
                                                      double cc, double kelvin, double r, double s2,
                                                      double* __restrict__ out) {
  __shared__ double sh[256];
  double v = 0;
  for (int q = threadIdx.x; q < nrows; q += 256) v += rows[q];
  const double k2 = block_sum<256>(v, sh);  // 2 K in g/mol A^2/ps^2
  if (threadIdx.x != 0) return;
  const double ek = 0.5 * k2 / kEkcal;
  double temp = 2.0 * ek / (nfree * kGasConst);
  if (temp == 0) temp = 0.1;
  const double d = (1 - cc) * (kelvin / temp) / nfree;
  double sc = sqrt(cc + (s2 + r * r) * d + 2 * r * sqrt(cc * d));
  if (r + sqrt(cc / d) < 0) sc = -sc;
  out[0] = sc;
}

__global__ void k_scale_v_dev(int n, const double* __restrict__ a, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = a[0];
  for (int d = 0; d < 3; ++d) v[3 * i + d] *= f;
}

__global__ void k_sorted_grad_to_orig(int n, const int* __restrict__ perm, const double* __restrict__ src,
                                      double* __restrict__ dst) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  for (int d = 0; d < 3; ++d) dst[3 * (size_t)i + d] = src[3 * (size_t)s + d];
}

template <class T>
void upload(DBuf& b, const std::vector<T>& v, hipStream_t st) {
  b.alloc(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) T9_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

template <class real>
void vdw_sites_t(Ctx& c) {
  const double3 box = make_double3(c.periodic ? c.box[0] : 1.0, c.periodic ? c.box[1] : 1.0,
                                   c.periodic ? c.box[2] : 1.0);
  hipLaunchKernelGGL(k_vsite_pos<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.as<int>(),
                     c.xyz.as<double>(), c.vparent.as<int>(), c.vred.as<double>(), c.jvdw_o.as<int>(),
                     make_double3(c.vorg[0], c.vorg[1], c.vorg[2]), make_double3(c.vlen[0], c.vlen[1], c.vlen[2]),
                     box, c.periodic ? 1 : 0, c.vpos.as<PosRec<real>>(), c.jvdw.as<int>());
  T9_LAUNCH_CHECK();
}

template <class real>
void vdw_launch_t(Ctx& c, bool vir, hipStream_t st, double* rows) {
  VdwArgs<real> A;
  A.n = c.n;
  A.nlist = c.vlist.nlist.as<int>();
  A.woff = c.vlist.off.as<long long>();
  A.entries = c.vlist.entries.as<uint32_t>();
  A.pos = c.vpos.as<PosRec<real>>();
  A.jvdw = c.jvdw.as<int>();
  A.irv = c.vrmin.as<real>();
  A.eps = c.veps.as<real>();
  A.nj = c.njvdw;
  for (int k = 0; k < kMaxCodes; ++k) A.vscale[k] = c.vscale[k];
  A.box = make_boxT<real>(c);
  A.cut = (real)c.vdw_cut;
  A.cut2 = (real)(c.vdw_cut * c.vdw_cut);
  A.taper = (real)c.vdw_taper;
  A.gsite = c.vsite_grad.as<double>();
  A.rows = rows;
  KScope ks_(c, KT_VDW, 1, st);
  if (vir) hipLaunchKernelGGL((k_vdw<real, true>), dim3(c.nwave), dim3(64 * kPairParts), 0, st, A);
  else hipLaunchKernelGGL((k_vdw<real, false>), dim3(c.nwave), dim3(64 * kPairParts), 0, st, A);
  T9_LAUNCH_CHECK();
}

}  // namespace

void vdw_sites(Ctx& c) {
  if (c.prec == 32) vdw_sites_t<float>(c);
  else vdw_sites_t<double>(c);
}

namespace {

// Analyze-mode count of the vdW term (the reference's nev, src/acc/ehal.cpp:
// 127,215-216): site pairs within the cutoff, each once, less the pairs with
// vdW scale 0; the same distance test as the pair kernel.
template <class real>
__global__ __launch_bounds__(256) void k_vdw_count(int n, const int* __restrict__ nlist,
                                                   const long long* __restrict__ woff,
                                                   const uint32_t* __restrict__ entries,
                                                   const PosRec<real>* __restrict__ pos, BoxT<real> box, real cut2,
                                                   VScale vs, unsigned long long* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0;
  if (s < n) {
    const PosRec<real> pi = pos[s];
    const int len = nlist[s];
    const uint32_t* en = entries + woff[s >> 6] + (s & 63);
    for (int kk = 0; kk < len; ++kk) {
      const uint32_t ent = en[(long long)kk * 64];
      const int k = (int)(ent >> kCodeBits);
      if (k <= s) continue;
      const int code = (int)(ent & (kMaxCodes - 1));
      if (code && vs.s[code] == 0) continue;
      real R[3];
      pair_disp(pos[k], pi, box, R[0], R[1], R[2]);
      if (R[0] * R[0] + R[1] * R[1] + R[2] * R[2] <= cut2) ++cnt;
    }
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, cnt);
}

template <class real>
long long vdw_count_t(Ctx& c) {
  DBuf out;
  out.alloc(sizeof(unsigned long long));
  T9_HIP(hipMemsetAsync(out.p, 0, sizeof(unsigned long long), c.stream));
  VScale vs;
  for (int k = 0; k < kMaxCodes; ++k) vs.s[k] = c.vscale[k];
  hipLaunchKernelGGL(k_vdw_count<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                     c.vlist.nlist.as<int>(), c.vlist.off.as<long long>(), c.vlist.entries.as<uint32_t>(),
                     c.vpos.as<PosRec<real>>(), make_boxT<real>(c), (real)(c.vdw_cut * c.vdw_cut), vs,
                     out.as<unsigned long long>());
  T9_LAUNCH_CHECK();
  unsigned long long h = 0;
  T9_HIP(hipMemcpyAsync(&h, out.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  return (long long)h;
}

}  // namespace

long long vdw_count(Ctx& c) {
  if (!c.use_vdw) return 0;
  return c.prec == 32 ? vdw_count_t<float>(c) : vdw_count_t<double>(c);
}

void vdw_launch(Ctx& c, bool vir, hipStream_t st, double* rows) {
  if (c.prec == 32) vdw_launch_t<float>(c, vir, st, rows);
  else vdw_launch_t<double>(c, vir, st, rows);
}

// per-site gradients (reduced hydrogen sites) onto the atoms, into the
// energy-path gradient
void vdw_finish(Ctx& c) {
  hipLaunchKernelGGL(k_vdw_to_atoms, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.as<int>(),
                     c.iperm.as<int>(), c.vparent.as<int>(), c.vred.as<double>(), c.vchild_off.as<int>(),
                     c.vchild.as<int>(), c.vsite_grad.as<double>(), c.egrad.as<double>());
  T9_LAUNCH_CHECK();
}

void vdw_run(Ctx& c, bool vir) {
  if (!c.use_vdw) return;
  vdw_launch(c, vir, c.stream, energy_rows(c, c.nwave));
  vdw_finish(c);
}

void valence_run(Ctx& c, double* gval, bool energy) {
  if (!c.use_valence) return;
  ValArgs A;
  A.n = c.n;
  A.xyz = c.xyz.as<double>();
  for (int d = 0; d < 3; ++d) A.L[d] = c.periodic ? c.box[d] : 1.0;
  A.periodic = c.periodic ? 1 : 0;
  A.toff = c.vterm_off.as<int>();
  A.term = c.vterm.as<int>();
  A.bij = c.bond_ij.as<int>();
  A.bkr = c.bond_kr.as<double>();
  A.aabc = c.angle_abc.as<int>();
  A.akt = c.angle_kt.as<double>();
  A.uac = c.urey_ac.as<int>();
  A.ukr = c.urey_kr.as<double>();
  A.cb = c.vpoly[0];
  A.qb = c.vpoly[1];
  A.ca = c.vpoly[2];
  A.qa = c.vpoly[3];
  A.pa = c.vpoly[4];
  A.sa = c.vpoly[5];
  A.grad = gval;
  const int nt = c.nbond + c.nangle + c.nurey;
  const int nb = ceil_div(std::max(nt, 1), 256);
  c.vterm_grad.alloc(sizeof(double) * 9 * (size_t)std::max(nt, 1));
  KScope ks_(c, KT_VALENCE);
  if (energy) {
    A.rows = energy_rows(c, nb);
    hipLaunchKernelGGL(k_valence_terms<true>, dim3(nb), dim3(256), 0, c.stream, A, c.nbond, c.nangle, c.nurey,
                       c.vterm_grad.as<double>());
  } else {
    A.rows = nullptr;
    hipLaunchKernelGGL(k_valence_terms<false>, dim3(nb), dim3(256), 0, c.stream, A, c.nbond, c.nangle, c.nurey,
                       c.vterm_grad.as<double>());
  }
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_valence_gather, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                     c.vterm_off.as<int>(), c.vterm.as<int>(), c.nbond, c.nangle, c.vterm_grad.as<double>(), gval);
  T9_LAUNCH_CHECK();
}

// One inner RESPA step of the fast (valence) group on the current
// coordinates: gradient to gval, kick by h, then drift by h.
void valence_inner(Ctx& c, double* gval, double h) {
  ValArgs A;
  A.n = c.n;
  A.xyz = c.xyz.as<double>();
  for (int d = 0; d < 3; ++d) A.L[d] = c.periodic ? c.box[d] : 1.0;
  A.periodic = c.periodic ? 1 : 0;
  A.toff = c.vterm_off.as<int>();
  A.term = c.vterm.as<int>();
  A.bij = c.bond_ij.as<int>();
  A.bkr = c.bond_kr.as<double>();
  A.aabc = c.angle_abc.as<int>();
  A.akt = c.angle_kt.as<double>();
  A.uac = c.urey_ac.as<int>();
  A.ukr = c.urey_kr.as<double>();
  A.cb = c.vpoly[0];
  A.qb = c.vpoly[1];
  A.ca = c.vpoly[2];
  A.qa = c.vpoly[3];
  A.pa = c.vpoly[4];
  A.sa = c.vpoly[5];
  A.grad = gval;
  A.rows = nullptr;
  const int nt = c.nbond + c.nangle + c.nurey;
  c.vterm_grad.alloc(sizeof(double) * 9 * (size_t)std::max(nt, 1));
  KScope ks_(c, KT_VALENCE);
  hipLaunchKernelGGL(k_valence_terms<false>, dim3(ceil_div(std::max(nt, 1), 256)), dim3(256), 0, c.stream, A,
                     c.nbond, c.nangle, c.nurey, c.vterm_grad.as<double>());
  hipLaunchKernelGGL(k_valence_gather_kick_drift, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                     c.vterm_off.as<int>(), c.vterm.as<int>(), c.nbond, c.nangle, c.vterm_grad.as<double>(), gval,
                     c.mass.as<double>(), h, c.vel.as<double>(), c.xyz.as<double>());
  T9_LAUNCH_CHECK();
  c.coords_dirty = true;
  c.rpole_valid = false;
  c.bins_valid = false;
}

bool respa_fused(const Ctx& c) { return c.use_valence && c.nmol > 0 && c.respa_path != 0; }

void respa_inner(Ctx& c, double* gfast, double h, int nrep) {
  ValArgs A;
  A.n = c.n;
  A.xyz = c.xyz.as<double>();
  for (int d = 0; d < 3; ++d) A.L[d] = c.periodic ? c.box[d] : 1.0;
  A.periodic = c.periodic ? 1 : 0;
  A.toff = c.vterm_off.as<int>();
  A.term = c.vterm.as<int>();
  A.bij = c.bond_ij.as<int>();
  A.bkr = c.bond_kr.as<double>();
  A.aabc = c.angle_abc.as<int>();
  A.akt = c.angle_kt.as<double>();
  A.uac = c.urey_ac.as<int>();
  A.ukr = c.urey_kr.as<double>();
  A.cb = c.vpoly[0];
  A.qb = c.vpoly[1];
  A.ca = c.vpoly[2];
  A.qa = c.vpoly[3];
  A.pa = c.vpoly[4];
  A.sa = c.vpoly[5];
  A.grad = gfast;
  A.rows = nullptr;
  const int nt = c.nbond + c.nangle + c.nurey;
  c.vterm_grad.alloc(sizeof(double) * 9 * (size_t)std::max(nt, 1));
  KScope ks_(c, KT_VALENCE);
  if (c.mol_small && c.respa_path == 2)
    hipLaunchKernelGGL(k_respa_inner_grp, dim3(ceil_div(c.nmol, kRespaMols)), dim3(256), 0, c.stream, A, c.nmol,
                       c.mol_off.as<int>(), c.mol_atom.as<int>(), c.mol_toff.as<int>(), c.mol_term.as<int>(), c.nbond,
                       c.nangle, gfast, c.mass.as<double>(), h, nrep, c.vel.as<double>(), c.xyz.as<double>());
  else
    hipLaunchKernelGGL(k_respa_inner, dim3(ceil_div(c.nmol, 256)), dim3(256), 0, c.stream, A, c.nmol,
                       c.mol_off.as<int>(), c.mol_atom.as<int>(), c.mol_toff.as<int>(), c.mol_term.as<int>(), c.nbond,
                       c.nangle, c.vterm_grad.as<double>(), gfast, c.mass.as<double>(), h, nrep, c.vel.as<double>(),
                       c.xyz.as<double>());
  T9_LAUNCH_CHECK();
  c.coords_dirty = true;
  c.rpole_valid = false;
  c.bins_valid = false;
}

void md_setup(Ctx& c, const System& s) {
  const int n = c.n;
  hipStream_t st = c.stream;
  c.use_vdw = s.opt.use_vdw;
  c.use_valence = s.opt.use_valence;
  c.vdw_cut = s.opt.vdw_cut;
  c.vdw_taper = s.opt.vdw_taper;
  if (c.use_vdw) {
    upload(c.vparent, s.vdw_parent, st);
    upload(c.vred, s.vdw_red, st);
    upload(c.jvdw_o, s.jvdw, st);
    c.jvdw.alloc(sizeof(int) * n);
    c.njvdw = s.njvdw;
    // tables of (1 / R_min, epsilon), each as hi + lo in the kernel precision.
    // In FP32 the kernel's 1.07 is float(1.07), which scales every pair energy
    // by (float(1.07) / 1.07)^7; epsilon carries the inverse factor.
    const size_t nt = s.vdw_rmin.size();
    const double fix = c.prec == 32 ? std::pow(1.07 / (double)1.07f, 7) : 1.0;
    std::vector<double> irv(2 * nt), eps(2 * nt);
    for (size_t t = 0; t < nt; ++t) {
      const double iv = s.vdw_rmin[t] > 0 ? 1.0 / s.vdw_rmin[t] : 0.0, ep = s.vdw_epsmix[t] * fix;
      if (c.prec == 32) {
        irv[2 * t] = (float)iv;
        irv[2 * t + 1] = (float)(iv - (double)(float)iv);
        eps[2 * t] = (float)ep;
        eps[2 * t + 1] = (float)(ep - (double)(float)ep);
      } else {
        irv[2 * t] = iv;
        eps[2 * t] = ep;
        irv[2 * t + 1] = eps[2 * t + 1] = 0.0;
      }
    }
    if (c.prec == 32) {
      upload(c.vrmin, std::vector<float>(irv.begin(), irv.end()), st);
      upload(c.veps, std::vector<float>(eps.begin(), eps.end()), st);
    } else {
      upload(c.vrmin, irv, st);
      upload(c.veps, eps, st);
    }
    c.vpos.alloc(c.prec == 32 ? sizeof(PosRec<float>) * n : sizeof(PosRec<double>) * n);
    c.vsite_grad.alloc(sizeof(double) * 3 * n);
    // scaled vdW partners: codes over the distinct scale values
    std::vector<double> vals = {1.0};
    std::vector<std::vector<std::pair<int, int>>> ex(n);
    for (size_t q = 0; q < s.vscale_s.size(); ++q) {
      const double v = s.vscale_s[q];
      int code = (int)(std::find(vals.begin(), vals.end(), v) - vals.begin());
      if (code == (int)vals.size()) vals.push_back(v);
      if (code >= kMaxCodes) throw ArgError("too many distinct vdW scale values");
      const int i = s.vscale_ik[2 * q], k = s.vscale_ik[2 * q + 1];
      ex[i].push_back({k, code});
      ex[k].push_back({i, code});
    }
    for (int k = 0; k < kMaxCodes; ++k) c.vscale[k] = k < (int)vals.size() ? vals[k] : 1.0;
    std::vector<int> off(n + 1, 0), ej, ec;
    for (int i = 0; i < n; ++i) {
      off[i + 1] = off[i] + (int)ex[i].size();
      for (auto& pr : ex[i]) {
        ej.push_back(pr.first);
        ec.push_back(pr.second);
      }
    }
    upload(c.vexcl_off, off, st);
    upload(c.vexcl_j, ej, st);
    upload(c.vexcl_code, ec, st);
    // reduced sites hanging on each atom
    std::vector<std::vector<int>> ch(n);
    for (int i = 0; i < n; ++i)
      if (s.vdw_parent[i] != i) ch[s.vdw_parent[i]].push_back(i);
    std::vector<int> coff(n + 1, 0), cl;
    for (int i = 0; i < n; ++i) {
      coff[i + 1] = coff[i] + (int)ch[i].size();
      cl.insert(cl.end(), ch[i].begin(), ch[i].end());
    }
    upload(c.vchild_off, coff, st);
    upload(c.vchild, cl, st);
  }
  if (c.use_valence) {
    upload(c.bond_ij, s.bond_ij, st);
    upload(c.bond_kr, s.bond_kr, st);
    upload(c.angle_abc, s.angle_abc, st);
    upload(c.angle_kt, s.angle_kt, st);
    upload(c.urey_ac, s.urey_ac, st);
    upload(c.urey_kr, s.urey_kr, st);
    c.vpoly[0] = s.ff.bond_cubic;
    c.vpoly[1] = s.ff.bond_quartic;
    c.vpoly[2] = s.ff.angle_cubic;
    c.vpoly[3] = s.ff.angle_quartic;
    c.vpoly[4] = s.ff.angle_pentic;
    c.vpoly[5] = s.ff.angle_sextic;
    std::vector<std::vector<int>> tl(n);
    const int nb = (int)s.bond_kr.size() / 2, na = (int)s.angle_kt.size() / 2, nu = (int)s.urey_kr.size() / 2;
    c.nbond = nb;
    c.nangle = na;
    c.nurey = nu;
    for (int t = 0; t < nb; ++t)
      for (int r = 0; r < 2; ++r) tl[s.bond_ij[2 * t + r]].push_back((0 << 28) | (t << 2) | r);
    for (int t = 0; t < na; ++t)
      for (int r = 0; r < 3; ++r) tl[s.angle_abc[3 * t + r]].push_back((1 << 28) | (t << 2) | r);
    for (int t = 0; t < nu; ++t)
      for (int r = 0; r < 2; ++r) tl[s.urey_ac[2 * t + r]].push_back((2 << 28) | (t << 2) | r);
    std::vector<int> off(n + 1, 0), tt;
    for (int i = 0; i < n; ++i) {
      off[i + 1] = off[i] + (int)tl[i].size();
      tt.insert(tt.end(), tl[i].begin(), tl[i].end());
    }
    upload(c.vterm_off, off, st);
    upload(c.vterm, tt, st);
    // molecules: connected components of the valence terms (union-find);
    // the fused inner loop runs one thread per molecule, so it is used only
    // when no molecule is large
    std::vector<int> par(n);
    for (int i = 0; i < n; ++i) par[i] = i;
    auto root = [&](int a) {
      while (par[a] != a) a = par[a] = par[par[a]];
      return a;
    };
    auto join = [&](int a, int b) {
      a = root(a);
      b = root(b);
      if (a != b) par[std::max(a, b)] = std::min(a, b);
    };
    for (int t = 0; t < nb; ++t) join(s.bond_ij[2 * t], s.bond_ij[2 * t + 1]);
    for (int t = 0; t < na; ++t) {
      join(s.angle_abc[3 * t], s.angle_abc[3 * t + 1]);
      join(s.angle_abc[3 * t + 1], s.angle_abc[3 * t + 2]);
    }
    for (int t = 0; t < nu; ++t) join(s.urey_ac[2 * t], s.urey_ac[2 * t + 1]);
    std::vector<int> mid(n, -1), moff(1, 0), matom, mtoff(1, 0), mterm;
    std::vector<std::vector<int>> atoms_of, terms_of;
    for (int i = 0; i < n; ++i) {  // molecules in the order of their first atom
      const int r = root(i);
      if (mid[r] < 0) {
        mid[r] = (int)atoms_of.size();
        atoms_of.emplace_back();
        terms_of.emplace_back();
      }
      atoms_of[mid[r]].push_back(i);
    }
    for (int t = 0; t < nb + na + nu; ++t) {
      const int a = t < nb ? s.bond_ij[2 * t] : (t < nb + na ? s.angle_abc[3 * (t - nb)] : s.urey_ac[2 * (t - nb - na)]);
      terms_of[mid[root(a)]].push_back(t);
    }
    size_t big = 0, bigt = 0;
    for (size_t m = 0; m < atoms_of.size(); ++m) {
      big = std::max(big, atoms_of[m].size());
      bigt = std::max(bigt, terms_of[m].size());
      matom.insert(matom.end(), atoms_of[m].begin(), atoms_of[m].end());
      moff.push_back((int)matom.size());
      mterm.insert(mterm.end(), terms_of[m].begin(), terms_of[m].end());
      mtoff.push_back((int)mterm.size());
    }
    c.nmol = big <= 64 ? (int)atoms_of.size() : 0;
    c.mol_small = big <= 4 && bigt <= 6;
    // T9HIP_RESPA selects the inner loop: "split" (a kernel per inner step),
    // "thread" (one thread per molecule), otherwise the 8-lane groups where
    // the molecules are small enough
    const char* rp = std::getenv("T9HIP_RESPA");
    c.respa_path = !rp ? 2 : (!std::strcmp(rp, "split") ? 0 : (!std::strcmp(rp, "thread") ? 1 : 2));
    if (c.nmol > 0) {
      upload(c.mol_off, moff, st);
      upload(c.mol_atom, matom, st);
      upload(c.mol_toff, mtoff, st);
      if (mterm.empty()) mterm.push_back(0);
      upload(c.mol_term, mterm, st);
    }
  }
  upload(c.mass, s.mass, st);
  c.mass_h = s.mass;
  c.nfree = std::max(1, 3 * n - 3);
  c.vel.alloc(sizeof(double) * 3 * n);
  c.g_slow.alloc(sizeof(double) * 3 * n);
  c.g_fast.alloc(sizeof(double) * 3 * n);
  T9_HIP(hipMemsetAsync(c.vel.p, 0, sizeof(double) * 3 * n, st));
  T9_HIP(hipMemsetAsync(c.g_fast.p, 0, sizeof(double) * 3 * n, st));
  T9_HIP(hipStreamSynchronize(st));
}

// Maxwell-Boltzmann velocities at `kelvin` from a seeded host stream, net
// momentum removed.
void md_velocities(Ctx& c, double kelvin, uint64_t seed) {
  const int n = c.n;
  const std::vector<double>& mass = c.mass_h;
  std::mt19937_64 gen(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<double> v(3 * n, 0.0);
  double p[3] = {0, 0, 0}, mt = 0;
  for (int i = 0; i < n; ++i) {
    if (mass[i] <= 0) continue;
    const double sd = std::sqrt(kGasR * kelvin * kEkcalD / mass[i]);
    for (int d = 0; d < 3; ++d) {
      v[3 * i + d] = sd * nd(gen);
      p[d] += mass[i] * v[3 * i + d];
    }
    mt += mass[i];
  }
  for (int i = 0; i < n; ++i)
    if (mass[i] > 0)
      for (int d = 0; d < 3; ++d) v[3 * i + d] -= p[d] / mt;
  T9_HIP(hipMemcpyAsync(c.vel.p, v.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
}

double md_kinetic(Ctx& c) {  // kcal/mol
  const int nb = ceil_div(c.n, 256);
  double* rows = c.partials.as<double>();
  hipLaunchKernelGGL(k_ekin, dim3(nb), dim3(256), 0, c.stream, c.n, c.mass.as<double>(), c.vel.as<double>(), rows);
  hipLaunchKernelGGL(k_sum_rows1, dim3(1), dim3(256), 0, c.stream, nb, rows, c.scalars.as<double>() + 40);
  T9_LAUNCH_CHECK();
  T9_HIP(hipMemcpyAsync(&c.h_scalars[40], c.scalars.as<double>() + 40, sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  return 0.5 * c.h_scalars[40] / kEkcalD;
}

void md_kick(Ctx& c, double h1, const double* g1, double h2, const double* g2) {
  hipLaunchKernelGGL(k_kick, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, h1, g1, h2, g2,
                     c.mass.as<double>(), c.vel.as<double>());
  T9_LAUNCH_CHECK();
}

void md_drift(Ctx& c, double h) {
  hipLaunchKernelGGL(k_drift, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, h, c.vel.as<double>(),
                     c.xyz.as<double>());
  T9_LAUNCH_CHECK();
  c.coords_dirty = true;
  c.rpole_valid = false;
  c.bins_valid = false;
}

void md_bussi(Ctx& c, double dt, double tau, double kelvin, double r, double s2) {
  const int nb = ceil_div(c.n, 256);
  double* rows = c.partials.as<double>();
  double* f = c.scalars.as<double>() + 41;
  hipLaunchKernelGGL(k_ekin, dim3(nb), dim3(256), 0, c.stream, c.n, c.mass.as<double>(), c.vel.as<double>(), rows);
  hipLaunchKernelGGL(k_bussi_factor, dim3(1), dim3(256), 0, c.stream, nb, rows, c.nfree, std::exp(-dt / tau), kelvin,
                     r, s2, f);
  hipLaunchKernelGGL(k_scale_v_dev, dim3(nb), dim3(256), 0, c.stream, c.n, f, c.vel.as<double>());
  T9_LAUNCH_CHECK();
}

void md_scale_v(Ctx& c, double a) {
  hipLaunchKernelGGL(k_scale_v, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, a, c.vel.as<double>());
  T9_LAUNCH_CHECK();
}

void md_sorted_to_orig(Ctx& c, const double* sorted, double* orig) {
  hipLaunchKernelGGL(k_sorted_grad_to_orig, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                     c.perm.as<int>(), sorted, orig);
  T9_LAUNCH_CHECK();
}

}  // namespace t9
