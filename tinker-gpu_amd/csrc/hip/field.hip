// Real-space field kernels for gfx950: the direct field of the permanent
// multipoles (dfield, d/p scaling), the mutual field of induced dipoles
// (ufield, u scaling) and the sparse preconditioner sweep.
//
// Pair math follows include/seq/pair_field.h:124-420 and
// src/acc/amoeba/induce.cpp:310-420; exclusions are folded in through the
// scale class of each list entry (Ewald: b_n - (1 - s*lambda) rr_n, which is
// the reference's "full pair + (s-1) non-Ewald correction" in one term).
//
// Execution: one lane per sorted atom, 64 consecutive atoms per wavefront,
// each lane walking its own in-cutoff neighbor list (compacted once per
// coordinate update from the buffered list, see nblist.hip), so no lane
// idles on out-of-range entries and no distance test is repeated.  The
// kPairParts waves of a workgroup share the same 64 atoms and take every
// kPairParts-th list entry; each accumulates in FP64 registers and the parts
// are added in a fixed order, so the result is independent of scheduling
// (no atomics).
//
// Per-pair data layout: pos[j] = (x, y, z, 1/pdamp) so one 16-byte load
// brings coordinates and the Thole radius; when every pair shares one Thole
// value (water, ions) the UTH variants skip the type-table lookup.  Square
// roots and reciprocals use the hardware rsqrt; the Ewald constants
// 2^n a^(2n-1)/sqrt(pi) are precomputed on the host; erfc shares exp(-x^2)
// with the Ewald recursion (t9_erfc_pair).
#include "chgpen.hpp"
#include "ctx.hpp"

namespace t9 {
namespace {

template <class real>
using R4 = PosRec<real>;

template <class real>
struct EwaldK {
  real a;           // Ewald coefficient
  real e1, e2, e3;  // 2a/sqrt(pi), 4a^3/sqrt(pi), 8a^5/sqrt(pi)
};

// Thole damping with the reciprocal radius product rpdik = 1/(pd_i pd_k)
// (0 when either site is unpolarizable: no damping, as the reference's
// damp == 0 branch in include/seq/damp.h:8-16).
template <class real>
__device__ __forceinline__ void thole_r(real r, real rpdik, real pga, real& s3, real& s5, real& s7) {
  const real ratio = r * rpdik;
  const real damp = rpdik == 0 ? (real)-1e16 : -pga * ratio * ratio * ratio;
  const real e = t9_exp_damp(damp);
  s3 = 1 - e;
  s5 = 1 - e * (1 - damp);
  s7 = 1 - e * (1 - damp + (real)0.6 * damp * damp);
}

template <class real, int NB>
__device__ __forceinline__ void ewald_b(real* bn, real r, real rinv, real rr2, const EwaldK<real>& ek) {
  real ex;
  bn[0] = t9_erfc_pair(ek.a * r, ex) * rinv;
  bn[1] = (bn[0] + ek.e1 * ex) * rr2;
  if (NB > 2) bn[2 < NB ? 2 : 0] = (3 * bn[1] + ek.e2 * ex) * rr2;
  if (NB > 3) bn[3 < NB ? 3 : 0] = (5 * bn[2 < NB ? 2 : 0] + ek.e3 * ex) * rr2;
}

// Field of induced dipole pair (ud_k, up_k) at site i, added to o[0..2] (d)
// and o[3..5] (p) in the working precision.  (xr, yr, zr) = x_k - x_i.
template <class real, int EWALD, int UTH>
__device__ __forceinline__ void ufield_pair(uint32_t ent, const R4<real>& pi, const R4<real>& pj, real dx, real dy,
                                            real dz, real px, real py, real pz, int jpi, int jpk,
                                            const real* __restrict__ thlval, real pga_u, const ScaleTable& sc,
                                            const BoxT<real>& box, const EwaldK<real>& ek, real* o) {
  real xr, yr, zr;
  pair_disp(pi, pj, box, xr, yr, zr);
  const real r2 = xr * xr + yr * yr + zr * zr;
  const real rinv = t9_rsqrt(r2);
  const real r = r2 * rinv;
  const real rr2 = rinv * rinv;
  const real rr3 = rinv * rr2;
  const real rr5 = 3 * rr3 * rr2;
  const real pga = UTH ? pga_u : thlval[jpi + jpk];
  real s3, s5, s7;
  thole_r(r, pi.w * pj.w, pga, s3, s5, s7);
  const int code = (int)(ent & (kMaxCodes - 1));
  const real us = code ? (real)sc.u[code] : (real)1;
  real c1, c2;
  if (EWALD) {
    real bn[3];
    ewald_b<real, 3>(bn, r, rinv, rr2, ek);
    c1 = bn[1] - (1 - us * s3) * rr3;
    c2 = bn[2] - (1 - us * s5) * rr5;
  } else {
    c1 = us * s3 * rr3;
    c2 = us * s5 * rr5;
  }
  const real cd = c2 * (dx * xr + dy * yr + dz * zr);
  const real cp = c2 * (px * xr + py * yr + pz * zr);
  o[0] += cd * xr - c1 * dx;
  o[1] += cd * yr - c1 * dy;
  o[2] += cd * zr - c1 * dz;
  o[3] += cp * xr - c1 * px;
  o[4] += cp * yr - c1 * py;
  o[5] += cp * zr - c1 * pz;
}

template <class real, int EWALD, int UTH>
__global__ __launch_bounds__(64 * kPairParts) void k_ufield_real(int n, const int* __restrict__ ncut,
                                                                 const long long* __restrict__ woff,
                                                                 const uint32_t* __restrict__ ecut,
                                                                 const R4<real>* __restrict__ pos,
                                                                 const real* __restrict__ ud,
                                                                 const real* __restrict__ up,
                                                                 const int* __restrict__ jpolar,
                                                                 const real* __restrict__ thlval, int nj, real pga_u,
                                                                 ScaleTable sc, BoxT<real> box, EwaldK<real> ek,
                                                                 real* __restrict__ fd, real* __restrict__ fp,
                                                                 Gate gate) {
  if (gate_closed(gate)) return;
  __shared__ double sh[parts_lds<kPairParts, 6>()];
  const int part = threadIdx.x >> 6;
  const int s = xcd_block() * 64 + (threadIdx.x & 63);
  double ab[6] = {0, 0, 0, 0, 0, 0};
  if (s < n) {
    const int len = ncut[s];
    const uint32_t* e = ecut + woff[s >> 6] + (s & 63);
    const R4<real> pi = pos[s];
    const int jpi = UTH ? 0 : jpolar[s] * nj;
    int k = part;
    constexpr int P = kPairParts;
    // two independent pairs per iteration: their gathers overlap in flight
    for (; k + P < len; k += 2 * P) {
      const uint32_t e0 = e[(long long)k * 64], e1 = e[(long long)(k + P) * 64];
      const int j0 = (int)(e0 >> kCodeBits), j1 = (int)(e1 >> kCodeBits);
      const R4<real> p0 = pos[j0], p1 = pos[j1];
      const real d0x = ud[3 * j0], d0y = ud[3 * j0 + 1], d0z = ud[3 * j0 + 2];
      const real q0x = up[3 * j0], q0y = up[3 * j0 + 1], q0z = up[3 * j0 + 2];
      const real d1x = ud[3 * j1], d1y = ud[3 * j1 + 1], d1z = ud[3 * j1 + 2];
      const real q1x = up[3 * j1], q1y = up[3 * j1 + 1], q1z = up[3 * j1 + 2];
      const int t0 = UTH ? 0 : jpolar[j0], t1 = UTH ? 0 : jpolar[j1];
      // the two pairs are summed in the working precision, then added to the
      // FP64 accumulators (one conversion per component for both)
      real o[6] = {0, 0, 0, 0, 0, 0};
      ufield_pair<real, EWALD, UTH>(e0, pi, p0, d0x, d0y, d0z, q0x, q0y, q0z, jpi, t0, thlval, pga_u, sc, box, ek, o);
      ufield_pair<real, EWALD, UTH>(e1, pi, p1, d1x, d1y, d1z, q1x, q1y, q1z, jpi, t1, thlval, pga_u, sc, box, ek, o);
#pragma unroll
      for (int q = 0; q < 6; ++q) ab[q] += (double)o[q];
    }
    if (k < len) {
      const uint32_t e0 = e[(long long)k * 64];
      const int j0 = (int)(e0 >> kCodeBits);
      const int t0 = UTH ? 0 : jpolar[j0];
      real o[6] = {0, 0, 0, 0, 0, 0};
      ufield_pair<real, EWALD, UTH>(e0, pi, pos[j0], ud[3 * j0], ud[3 * j0 + 1], ud[3 * j0 + 2], up[3 * j0],
                                    up[3 * j0 + 1], up[3 * j0 + 2], jpi, t0, thlval, pga_u, sc, box, ek, o);
#pragma unroll
      for (int q = 0; q < 6; ++q) ab[q] += (double)o[q];
    }
  }
  parts_reduce<kPairParts, 6>(ab, sh);
  if (part == 0 && s < n) {
    fd[3 * s] = (real)ab[0];
    fd[3 * s + 1] = (real)ab[1];
    fd[3 * s + 2] = (real)ab[2];
    fp[3 * s] = (real)ab[3];
    fp[3 * s + 1] = (real)ab[4];
    fp[3 * s + 2] = (real)ab[5];
  }
}

template <class real, int EWALD, int UTH>
__global__ __launch_bounds__(64 * kPairParts) void k_dfield_real(int n, const int* __restrict__ ncut,
                                                     const long long* __restrict__ woff,
                                                     const uint32_t* __restrict__ ecut,
                                                     const R4<real>* __restrict__ pos, const real* __restrict__ rpole,
                                                     const int* __restrict__ jpolar, const real* __restrict__ thlval,
                                                     int nj, real pga_u, ScaleTable sc, BoxT<real> box,
                                                     EwaldK<real> ek, real* __restrict__ fd, real* __restrict__ fp) {
  __shared__ double sh[parts_lds<kPairParts, 6>()];
  const int part = threadIdx.x >> 6;
  const int s = xcd_block() * 64 + (threadIdx.x & 63);
  double ab[6] = {0, 0, 0, 0, 0, 0};
  double &a0 = ab[0], &a1 = ab[1], &a2 = ab[2], &b0 = ab[3], &b1 = ab[4], &b2 = ab[5];
  const int len = s < n ? ncut[s] : 0;
  const uint32_t* e = ecut + (s < n ? woff[s >> 6] + (s & 63) : 0);
  const R4<real> pi = pos[s < n ? s : 0];
  const int jpi = UTH ? 0 : jpolar[s < n ? s : 0] * nj;
  uint32_t nxt = part < len ? e[(long long)part * 64] : 0;
  for (int k = part; k < len; k += kPairParts) {
    const uint32_t ent = nxt;
    if (k + kPairParts < len) nxt = e[(long long)(k + kPairParts) * 64];
    const int j = (int)(ent >> kCodeBits);
    const R4<real> pj = pos[j];
    real xr, yr, zr;
    pair_disp(pi, pj, box, xr, yr, zr);
    const real r2 = xr * xr + yr * yr + zr * zr;
    const real rinv = t9_rsqrt(r2);
    const real r = r2 * rinv;
    const real rr2 = rinv * rinv;
    const real rr3 = rinv * rr2;
    const real rr5 = 3 * rr3 * rr2;
    const real rr7 = 5 * rr5 * rr2;
    const real pga = UTH ? pga_u : thlval[jpi + jpolar[j]];
    real s3, s5, s7;
    thole_r(r, pi.w * pj.w, pga, s3, s5, s7);
    const real* pk = rpole + 10 * (size_t)j;
    const real ck = pk[0], dkx = pk[1], dky = pk[2], dkz = pk[3];
    const real qxx = pk[4], qyy = pk[5], qzz = pk[6], qxy = pk[7], qxz = pk[8], qyz = pk[9];
    const real dkr = dkx * xr + dky * yr + dkz * zr;
    const real qkx = qxx * xr + qxy * yr + qxz * zr;
    const real qky = qxy * xr + qyy * yr + qyz * zr;
    const real qkz = qxz * xr + qyz * yr + qzz * zr;
    const real qkr = qkx * xr + qky * yr + qkz * zr;
    real bn[4] = {0, 0, 0, 0};
    if (EWALD) ewald_b<real, 4>(bn, r, rinv, rr2, ek);
    const int code = (int)(ent & (kMaxCodes - 1));
    const real ds = code ? (real)sc.d[code] : (real)1, ps = code ? (real)sc.p[code] : (real)1;
    real c1, c2, c3;
    if (EWALD) {
      c1 = bn[1] - (1 - ds * s3) * rr3;
      c2 = bn[2] - (1 - ds * s5) * rr5;
      c3 = bn[3] - (1 - ds * s7) * rr7;
    } else {
      c1 = ds * s3 * rr3;
      c2 = ds * s5 * rr5;
      c3 = ds * s7 * rr7;
    }
    real t = -(c1 * ck - c2 * dkr + c3 * qkr);
    real fx = t * xr - c1 * dkx + 2 * c2 * qkx;
    real fy = t * yr - c1 * dky + 2 * c2 * qky;
    real fz = t * zr - c1 * dkz + 2 * c2 * qkz;
    a0 += (double)fx;
    a1 += (double)fy;
    a2 += (double)fz;
    if (ps != ds) {
      if (EWALD) {
        c1 = bn[1] - (1 - ps * s3) * rr3;
        c2 = bn[2] - (1 - ps * s5) * rr5;
        c3 = bn[3] - (1 - ps * s7) * rr7;
      } else {
        c1 = ps * s3 * rr3;
        c2 = ps * s5 * rr5;
        c3 = ps * s7 * rr7;
      }
      t = -(c1 * ck - c2 * dkr + c3 * qkr);
      fx = t * xr - c1 * dkx + 2 * c2 * qkx;
      fy = t * yr - c1 * dky + 2 * c2 * qky;
      fz = t * zr - c1 * dkz + 2 * c2 * qkz;
    }
    b0 += (double)fx;
    b1 += (double)fy;
    b2 += (double)fz;
  }
  parts_reduce<kPairParts, 6>(ab, sh);
  if (part == 0 && s < n) {
    fd[3 * s] = (real)a0;
    fd[3 * s + 1] = (real)a1;
    fd[3 * s + 2] = (real)a2;
    fp[3 * s] = (real)b0;
    fp[3 * s + 1] = (real)b1;
    fp[3 * s + 2] = (real)b2;
  }
}

// The in-cutoff compaction of the buffered list fused with the real-space
// direct field: the same partition and order as the compaction kernel (part
// p of the workgroup takes chunk p of each round of a lane's list; the kept
// entries of a round go out in list order through LDS rows), and every entry
// found within the cutoff also adds its direct-field term, with the pair
// math of k_dfield_real, to the part's FP64 sums, which are added in part
// order at the end.  One pass over the buffered list and its partners'
// positions serves both.
template <class real, int EWALD, int UTH>
__global__ __launch_bounds__(64 * kPairParts) void k_compact_dfield(
    int n, const int* __restrict__ nlist, const long long* __restrict__ woff, const uint32_t* __restrict__ entries,
    const R4<real>* __restrict__ pos, BoxT<real> box, real cut2, real cut2b, int* __restrict__ ncut,
    uint32_t* __restrict__ ecut, int* __restrict__ ncut2, uint32_t* __restrict__ ecut2,
    const real* __restrict__ rpole, const int* __restrict__ jpolar, const real* __restrict__ thlval, int nj,
    real pga_u, ScaleTable sc, EwaldK<real> ek, real* __restrict__ fd, real* __restrict__ fp) {
  constexpr int P = kPairParts;
  __shared__ unsigned short cnt_a[P][64], cnt_b[P][64];
  __shared__ uint32_t rows_a[kCompactRowsA][64], rows_b[kCompactRowsB][64];
  const int part = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int blk = xcd_block();
  const int s = blk * 64 + lane;
  const bool two = ecut2 != nullptr;
  const bool valid = s < n;
  const int len = valid ? nlist[s] : 0;
  const int maxlen = wave_max(len);
  const long long base = woff[blk] + lane;
  const uint32_t* e = entries + base;
  R4<real> pi;
  if (valid) pi = pos[s];
  const int jpi = UTH ? 0 : jpolar[valid ? s : 0] * nj;
  double ab[6] = {0, 0, 0, 0, 0, 0};
  int tot_a = 0, tot_b = 0;  // kept entries of this atom in earlier rounds
  for (int r0 = 0; r0 < maxlen;) {
    const int C = min(64, (maxlen - r0 + P - 1) / P);
    const int kb = r0 + part * C;
    const int ke = min(kb + C, len);
    unsigned long long ma = 0, mb = 0;
    for (int k0 = kb; k0 < ke; k0 += 8) {  // 8 entries, then their 8 partners, in flight
      uint32_t ent[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) ent[u] = k0 + u < ke ? e[(long long)(k0 + u) * 64] : 0u;
      R4<real> pj[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < ke) pj[u] = pos[ent[u] >> kCodeBits];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < ke) {
          real xr, yr, zr;
          pair_disp(pi, pj[u], box, xr, yr, zr);
          const real r2 = xr * xr + yr * yr + zr * zr;
          if (two && r2 <= cut2b) mb |= 1ull << (k0 + u - kb);
          if (r2 <= cut2) {
            ma |= 1ull << (k0 + u - kb);
            const int j = (int)(ent[u] >> kCodeBits);
            const real rinv = t9_rsqrt(r2);
            const real r = r2 * rinv;
            const real rr2 = rinv * rinv;
            const real rr3 = rinv * rr2;
            const real rr5 = 3 * rr3 * rr2;
            const real rr7 = 5 * rr5 * rr2;
            const real pga = UTH ? pga_u : thlval[jpi + jpolar[j]];
            real s3, s5, s7;
            thole_r(r, pi.w * pj[u].w, pga, s3, s5, s7);
            const real* pk = rpole + 10 * (size_t)j;
            const real ck = pk[0], dkx = pk[1], dky = pk[2], dkz = pk[3];
            const real qxx = pk[4], qyy = pk[5], qzz = pk[6], qxy = pk[7], qxz = pk[8], qyz = pk[9];
            const real dkr = dkx * xr + dky * yr + dkz * zr;
            const real qkx = qxx * xr + qxy * yr + qxz * zr;
            const real qky = qxy * xr + qyy * yr + qyz * zr;
            const real qkz = qxz * xr + qyz * yr + qzz * zr;
            const real qkr = qkx * xr + qky * yr + qkz * zr;
            real bn[4] = {0, 0, 0, 0};
            if (EWALD) ewald_b<real, 4>(bn, r, rinv, rr2, ek);
            const int code = (int)(ent[u] & (kMaxCodes - 1));
            const real ds = code ? (real)sc.d[code] : (real)1, ps = code ? (real)sc.p[code] : (real)1;
            real c1, c2, c3;
            if (EWALD) {
              c1 = bn[1] - (1 - ds * s3) * rr3;
              c2 = bn[2] - (1 - ds * s5) * rr5;
              c3 = bn[3] - (1 - ds * s7) * rr7;
            } else {
              c1 = ds * s3 * rr3;
              c2 = ds * s5 * rr5;
              c3 = ds * s7 * rr7;
            }
            real t = -(c1 * ck - c2 * dkr + c3 * qkr);
            real fx = t * xr - c1 * dkx + 2 * c2 * qkx;
            real fy = t * yr - c1 * dky + 2 * c2 * qky;
            real fz = t * zr - c1 * dkz + 2 * c2 * qkz;
            ab[0] += (double)fx;
            ab[1] += (double)fy;
            ab[2] += (double)fz;
            if (ps != ds) {
              if (EWALD) {
                c1 = bn[1] - (1 - ps * s3) * rr3;
                c2 = bn[2] - (1 - ps * s5) * rr5;
                c3 = bn[3] - (1 - ps * s7) * rr7;
              } else {
                c1 = ps * s3 * rr3;
                c2 = ps * s5 * rr5;
                c3 = ps * s7 * rr7;
              }
              t = -(c1 * ck - c2 * dkr + c3 * qkr);
              fx = t * xr - c1 * dkx + 2 * c2 * qkx;
              fy = t * yr - c1 * dky + 2 * c2 * qky;
              fz = t * zr - c1 * dkz + 2 * c2 * qkz;
            }
            ab[3] += (double)fx;
            ab[4] += (double)fy;
            ab[5] += (double)fz;
          }
        }
    }
    cnt_a[part][lane] = (unsigned short)__popcll(ma);
    cnt_b[part][lane] = (unsigned short)__popcll(mb);
    __syncthreads();
    int oa = tot_a, ob = tot_b;
    for (int q = 0; q < P; ++q) {
      if (q < part) {
        oa += cnt_a[q][lane];
        ob += cnt_b[q][lane];
      }
      tot_a += cnt_a[q][lane];
      tot_b += cnt_b[q][lane];
    }
    __syncthreads();
    while (ma | mb) {
      const int ja = ma ? __ffsll((long long)ma) - 1 : 64, jb = mb ? __ffsll((long long)mb) - 1 : 64;
      const int j = min(ja, jb);
      const uint32_t ent = e[(long long)(kb + j) * 64];
      if (j == ja) {
        if (oa < kCompactRowsA) rows_a[oa][lane] = ent;
        else ecut[base + (long long)oa * 64] = ent;
        ++oa;
        ma &= ma - 1;
      }
      if (j == jb) {
        if (ob < kCompactRowsB) rows_b[ob][lane] = ent;
        else ecut2[base + (long long)ob * 64] = ent;
        ++ob;
        mb &= mb - 1;
      }
    }
    r0 += P * C;
  }
  if (valid && part == 0) {
    ncut[s] = tot_a;
    if (two) ncut2[s] = tot_b;
  }
  // whole rows out; slots past an atom's count hold stale values that no
  // kernel reads (it reads positions below ncut)
  const int ra = min(wave_max(tot_a), kCompactRowsA), rb = two ? min(wave_max(tot_b), kCompactRowsB) : 0;
  __syncthreads();
  const long long wb = woff[blk];
  for (int q = threadIdx.x; q < ra * 64; q += 64 * P) ecut[wb + q] = rows_a[q >> 6][q & 63];
  for (int q = threadIdx.x; q < rb * 64; q += 64 * P) ecut2[wb + q] = rows_b[q >> 6][q & 63];
  __syncthreads();
  // the parts' field sums, in part order (the row area is free again)
  parts_reduce<kPairParts, 6>(ab, reinterpret_cast<double*>(&rows_a[0][0]));
  if (part == 0 && valid) {
    fd[3 * s] = (real)ab[0];
    fd[3 * s + 1] = (real)ab[1];
    fd[3 * s + 2] = (real)ab[2];
    fp[3 * s] = (real)ab[3];
    fp[3 * s + 1] = (real)ab[4];
    fp[3 * s + 2] = (real)ab[5];
  }
}

// HIPPO fields (src/cu/hippo/dfieldChgpen_cu1.cc, ufieldChgpen_cu1.cc,
// include/seq/pair_field_chgpen.h): an induced dipole is a point probe, so
// the direct field at i is the field of k's core charge (undamped) plus that
// of k's valence multipole with k's core-valence damping, scaled by the p
// class of the pair (HIPPO's d scale); the mutual field has the
// valence-valence damping and the u class (HIPPO's induce scale).  FP64
// arithmetic whatever the working precision; same layout, parts and order
// as k_dfield_real / k_ufield_real.  cp: caller-order [n][3] core charge,
// valence charge, exponent.
template <class real, int EWALD>
__global__ __launch_bounds__(64 * kPairParts) void k_dfield_chgpen(int n, const int* __restrict__ ncut,
                                                       const long long* __restrict__ woff,
                                                       const uint32_t* __restrict__ ecut,
                                                       const R4<real>* __restrict__ pos, const real* __restrict__ rpole,
                                                       const int* __restrict__ perm, const double* __restrict__ cp,
                                                       ScaleTable sc, BoxT<real> box, EwaldK<double> ek,
                                                       real* __restrict__ fd, real* __restrict__ fp) {
  __shared__ double sh[parts_lds<kPairParts, 6>()];
  const int part = threadIdx.x >> 6;
  const int s = xcd_block() * 64 + (threadIdx.x & 63);
  double ab[6] = {0, 0, 0, 0, 0, 0};
  if (s < n) {
    const int len = ncut[s];
    const uint32_t* e = ecut + woff[s >> 6] + (s & 63);
    const R4<real> pi = pos[s];
    for (int k = part; k < len; k += kPairParts) {
      const uint32_t ent = e[(long long)k * 64];
      const int j = (int)(ent >> kCodeBits);
      const int code = (int)(ent & (kMaxCodes - 1));
      const double ps = code ? sc.p[code] : 1.0;
      double d[3];
      pair_disp64(pi, pos[j], box, d);  // x_k - x_i
      const double xr = d[0], yr = d[1], zr = d[2];
      const double r2 = xr * xr + yr * yr + zr * zr;
      const double r = sqrt(r2), rinv = 1 / r, rr2 = rinv * rinv;
      const double rr3 = rinv * rr2, rr5 = 3 * rr3 * rr2, rr7 = 5 * rr5 * rr2;
      double bn[4] = {0, 0, 0, 0};
      if (EWALD) ewald_b<double, 4>(bn, r, rinv, rr2, ek);
      const int jo = perm[j];
      const double corek = cp[3 * jo], valk = cp[3 * jo + 1];
      double lci[6], lck[6], lik[6];
      chgpen_lambda(r, cp[3 * jo + 2], cp[3 * jo + 2], lci, lck, lik);
      auto kern = [&](double lam, double rr, double b) { return EWALD ? b - (1 - ps * lam) * rr : ps * lam * rr; };
      const double c1 = kern(1.0, rr3, bn[1]);
      const double k1 = kern(lck[1], rr3, bn[1]), k2 = kern(lck[2], rr5, bn[2]), k3 = kern(lck[3], rr7, bn[3]);
      const real* pk = rpole + 10 * (size_t)j;
      const double dkx = pk[1], dky = pk[2], dkz = pk[3];
      const double qxx = pk[4], qyy = pk[5], qzz = pk[6], qxy = pk[7], qxz = pk[8], qyz = pk[9];
      const double dkr = dkx * xr + dky * yr + dkz * zr;
      const double qkx = qxx * xr + qxy * yr + qxz * zr;
      const double qky = qxy * xr + qyy * yr + qyz * zr;
      const double qkz = qxz * xr + qyz * yr + qzz * zr;
      const double qkr = qkx * xr + qky * yr + qkz * zr;
      const double t = -(c1 * corek + k1 * valk - k2 * dkr + k3 * qkr);
      const double fx = t * xr - k1 * dkx + 2 * k2 * qkx;
      const double fy = t * yr - k1 * dky + 2 * k2 * qky;
      const double fz = t * zr - k1 * dkz + 2 * k2 * qkz;
      ab[0] += fx;
      ab[1] += fy;
      ab[2] += fz;
    }
  }
  parts_reduce<kPairParts, 6>(ab, sh);
  if (part == 0 && s < n)
    for (int q = 0; q < 3; ++q) {
      fd[3 * s + q] = (real)ab[q];
      fp[3 * s + q] = (real)ab[q];
    }
}

template <class real, int EWALD>
__global__ __launch_bounds__(64 * kPairParts) void k_ufield_chgpen(int n, const int* __restrict__ ncut,
                                                       const long long* __restrict__ woff,
                                                       const uint32_t* __restrict__ ecut,
                                                       const R4<real>* __restrict__ pos, const real* __restrict__ ud,
                                                       const real* __restrict__ up, const int* __restrict__ perm,
                                                       const double* __restrict__ cp, ScaleTable sc, BoxT<real> box,
                                                       EwaldK<double> ek, real* __restrict__ fd,
                                                       real* __restrict__ fp, Gate gate) {
  if (gate_closed(gate)) return;
  __shared__ double sh[parts_lds<kPairParts, 6>()];
  const int part = threadIdx.x >> 6;
  const int s = xcd_block() * 64 + (threadIdx.x & 63);
  double ab[6] = {0, 0, 0, 0, 0, 0};
  if (s < n) {
    const int len = ncut[s];
    const uint32_t* e = ecut + woff[s >> 6] + (s & 63);
    const R4<real> pi = pos[s];
    const double ai = cp[3 * perm[s] + 2];
    for (int k = part; k < len; k += kPairParts) {
      const uint32_t ent = e[(long long)k * 64];
      const int j = (int)(ent >> kCodeBits);
      const int code = (int)(ent & (kMaxCodes - 1));
      const double us = code ? sc.u[code] : 1.0;
      double d[3];
      pair_disp64(pi, pos[j], box, d);  // x_k - x_i
      const double xr = d[0], yr = d[1], zr = d[2];
      const double r2 = xr * xr + yr * yr + zr * zr;
      const double r = sqrt(r2), rinv = 1 / r, rr2 = rinv * rinv;
      const double rr3 = rinv * rr2, rr5 = 3 * rr3 * rr2;
      double lci[6], lck[6], lik[6];
      chgpen_lambda(r, ai, cp[3 * perm[j] + 2], lci, lck, lik);
      double c1, c2;
      if (EWALD) {
        double bn[3];
        ewald_b<double, 3>(bn, r, rinv, rr2, ek);
        c1 = bn[1] - (1 - us * lik[1]) * rr3;
        c2 = bn[2] - (1 - us * lik[2]) * rr5;
      } else {
        c1 = us * lik[1] * rr3;
        c2 = us * lik[2] * rr5;
      }
      const double dx = ud[3 * j], dy = ud[3 * j + 1], dz = ud[3 * j + 2];
      const double px = up[3 * j], py = up[3 * j + 1], pz = up[3 * j + 2];
      const double cd = c2 * (dx * xr + dy * yr + dz * zr), cq = c2 * (px * xr + py * yr + pz * zr);
      ab[0] += cd * xr - c1 * dx;
      ab[1] += cd * yr - c1 * dy;
      ab[2] += cd * zr - c1 * dz;
      ab[3] += cq * xr - c1 * px;
      ab[4] += cq * yr - c1 * py;
      ab[5] += cq * zr - c1 * pz;
    }
  }
  parts_reduce<kPairParts, 6>(ab, sh);
  if (part == 0 && s < n)
    for (int q = 0; q < 3; ++q) {
      fd[3 * s + q] = (real)ab[q];
      fp[3 * s + q] = (real)ab[3 + q];
    }
}

// One sparse-preconditioner pair: alpha_i alpha_k T_ik r_k added to o[0..2]
// (d) and o[3..5] (p).
template <class real, int UTH>
__device__ __forceinline__ void precond_pair(uint32_t ent, const R4<real>& pi, real poli,
                                             const R4<real>* __restrict__ pos, const real* __restrict__ polarity,
                                             const int* __restrict__ jpolar, const real* __restrict__ thlval, int jpi,
                                             real pga_u, const ScaleTable& sc, const BoxT<real>& box,
                                             const real* __restrict__ rd, const real* __restrict__ rp, real* o) {
  const int j = (int)(ent >> kCodeBits);
  const R4<real> pj = pos[j];
  real xr, yr, zr;
  pair_disp(pi, pj, box, xr, yr, zr);
  const real r2 = xr * xr + yr * yr + zr * zr;
  const real rinv = t9_rsqrt(r2);
  const real r = r2 * rinv;
  const real pga = UTH ? pga_u : thlval[jpi + jpolar[j]];
  real s3, s5, s7;
  thole_r(r, pi.w * pj.w, pga, s3, s5, s7);
  const int code = (int)(ent & (kMaxCodes - 1));
  const real us = code ? (real)sc.u[code] : (real)1;
  const real polik = poli * polarity[j];
  const real ri3 = rinv * rinv * rinv;
  const real rr3 = us * s3 * polik * ri3;
  const real rr5 = 3 * us * s5 * polik * ri3 * rinv * rinv;
  const real m0 = rr5 * xr * xr - rr3, m1 = rr5 * xr * yr, m2 = rr5 * xr * zr;
  const real m3 = rr5 * yr * yr - rr3, m4 = rr5 * yr * zr, m5 = rr5 * zr * zr - rr3;
  const real x0 = rd[3 * j], x1 = rd[3 * j + 1], x2 = rd[3 * j + 2];
  const real y0 = rp[3 * j], y1 = rp[3 * j + 1], y2 = rp[3 * j + 2];
  o[0] += m0 * x0 + m1 * x1 + m2 * x2;
  o[1] += m1 * x0 + m3 * x1 + m4 * x2;
  o[2] += m2 * x0 + m4 * x1 + m5 * x2;
  o[3] += m0 * y0 + m1 * y1 + m2 * y2;
  o[4] += m1 * y0 + m3 * y1 + m4 * y2;
  o[5] += m2 * y0 + m4 * y1 + m5 * y2;
}

// z = udiag * alpha * r + sum_{r_ik <= ucut} alpha_i alpha_k T_ik r_k
// (sparse preconditioner, src/acc/amoeba/induce.cpp:310-420).
template <class real, int UTH>
__global__ __launch_bounds__(64 * kPairParts) void k_precond(int n, int sparse, real udiag, const int* __restrict__ ncut,
                                                 const long long* __restrict__ woff,
                                                 const uint32_t* __restrict__ ecut,
                                                 const R4<real>* __restrict__ pos, const real* __restrict__ polarity,
                                                 const int* __restrict__ jpolar, const real* __restrict__ thlval,
                                                 int nj, real pga_u, ScaleTable sc, BoxT<real> box,
                                                 const real* __restrict__ rd, const real* __restrict__ rp,
                                                 real* __restrict__ zd, real* __restrict__ zp,
                                                 real* __restrict__ z2d, real* __restrict__ z2p,
                                                 double* __restrict__ dots64, Gate gate) {
  if (gate_closed(gate)) return;
  __shared__ double sh[parts_lds<kPairParts, 6>()];
  const int part = threadIdx.x >> 6;
  const int s = xcd_block() * 64 + (threadIdx.x & 63);
  const int si = s < n ? s : 0;
  const real poli = polarity[si];
  const real f = sparse ? udiag * poli : poli;
  double ab[6] = {0, 0, 0, 0, 0, 0};
  double &a0 = ab[0], &a1 = ab[1], &a2 = ab[2], &b0 = ab[3], &b1 = ab[4], &b2 = ab[5];
  if (part == 0 && s < n) {
    a0 = f * rd[3 * s];
    a1 = f * rd[3 * s + 1];
    a2 = f * rd[3 * s + 2];
    b0 = f * rp[3 * s];
    b1 = f * rp[3 * s + 1];
    b2 = f * rp[3 * s + 2];
  }
  if (sparse && s < n) {
    const int len = ncut[s];
    const uint32_t* e = ecut + woff[s >> 6] + (s & 63);
    const R4<real> pi = pos[s];
    const int jpi = UTH ? 0 : jpolar[s] * nj;
    // two entries per iteration, summed in the working precision before the
    // FP64 accumulators
    int k = part;
    constexpr int P = kPairParts;
    for (; k + P < len; k += 2 * P) {
      const uint32_t e0 = e[(long long)k * 64], e1 = e[(long long)(k + P) * 64];
      real o[6] = {0, 0, 0, 0, 0, 0};
      precond_pair<real, UTH>(e0, pi, poli, pos, polarity, jpolar, thlval, jpi, pga_u, sc, box, rd, rp, o);
      precond_pair<real, UTH>(e1, pi, poli, pos, polarity, jpolar, thlval, jpi, pga_u, sc, box, rd, rp, o);
#pragma unroll
      for (int q = 0; q < 6; ++q) ab[q] += (double)o[q];
    }
    if (k < len) {
      real o[6] = {0, 0, 0, 0, 0, 0};
      precond_pair<real, UTH>(e[(long long)k * 64], pi, poli, pos, polarity, jpolar, thlval, jpi, pga_u, sc, box, rd,
                              rp, o);
#pragma unroll
      for (int q = 0; q < 6; ++q) ab[q] += (double)o[q];
    }
  }
  parts_reduce<kPairParts, 6>(ab, sh);
  if (part == 0 && s < n) {
    zd[3 * s] = (real)a0;
    zd[3 * s + 1] = (real)a1;
    zd[3 * s + 2] = (real)a2;
    zp[3 * s] = (real)b0;
    zp[3 * s + 1] = (real)b1;
    zp[3 * s + 2] = (real)b2;
    if (z2d) {  // second copy (the first CG search direction)
      z2d[3 * s] = (real)a0;
      z2d[3 * s + 1] = (real)a1;
      z2d[3 * s + 2] = (real)a2;
      z2p[3 * s] = (real)b0;
      z2p[3 * s + 1] = (real)b1;
      z2p[3 * s + 2] = (real)b2;
    }
  }
  if (dots64 && part == 0) {
    // r.z (d, p) and r.r (d, p) of this workgroup's 64 elements: the same
    // per-element products and wavefront butterfly as the CG dot kernel
    double v[4] = {0, 0, 0, 0};
    if (s < n) {
      v[0] = dot3(rd, (const real*)zd, s);
      v[1] = dot3(rp, (const real*)zp, s);
      v[2] = dot3(rd, rd, s);
      v[3] = dot3(rp, rp, s);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double x = wave_bfly_sum(v[q]);
      if ((threadIdx.x & 63) == 0) dots64[(s >> 6) * 4 + q] = x;
    }
  }
}


template <class real>
EwaldK<real> ewald_k(const Ctx& c) {
  const double a = c.aewald, isp = 0.56418958354775628695;  // 1/sqrt(pi)
  EwaldK<real> k;
  k.a = (real)a;
  k.e1 = (real)(2 * a * isp);
  k.e2 = (real)(4 * a * a * a * isp);
  k.e3 = (real)(8 * a * a * a * a * a * isp);
  return k;
}

void require_cut(const NbList& L) {
  if (!L.cut_valid) throw ArgError("in-cutoff neighbor list is stale (internal ordering error)");
}

template <class real>
void ufield_real_t(Ctx& c, const void* ud, const void* up, void* fd, void* fp, hipStream_t st) {
  const auto& L = c.mlist;
  require_cut(L);
  if (c.chgpen) {
    auto kc = c.ewald ? k_ufield_chgpen<real, 1> : k_ufield_chgpen<real, 0>;
    KScope ks_(c, KT_UFIELD_REAL, 1, st);
    hipLaunchKernelGGL(kc, dim3(c.nwave), dim3(64 * kPairParts), 0, st, c.n, L.ncut.as<int>(),
                       L.off.as<long long>(), L.ecut.as<uint32_t>(), c.pos.as<R4<real>>(), (const real*)ud,
                       (const real*)up, c.perm.as<int>(), c.chgpen_o.as<double>(), c.scale, make_boxT<real>(c),
                       ewald_k<double>(c), (real*)fd, (real*)fp, c.gate);
    T9_LAUNCH_CHECK();
    return;
  }
  auto kern = c.ewald ? (c.uniform_thole ? k_ufield_real<real, 1, 1> : k_ufield_real<real, 1, 0>)
                      : (c.uniform_thole ? k_ufield_real<real, 0, 1> : k_ufield_real<real, 0, 0>);
  KScope ks_(c, KT_UFIELD_REAL, 1, st);
  hipLaunchKernelGGL(kern, dim3(c.nwave), dim3(64 * kPairParts), 0, st, c.n, L.ncut.as<int>(),
                     L.off.as<long long>(), L.ecut.as<uint32_t>(), c.pos.as<R4<real>>(), (const real*)ud,
                     (const real*)up, c.jpolar.as<int>(), c.thlval.as<real>(), c.nj, (real)c.thole_value, c.scale,
                     make_boxT<real>(c), ewald_k<real>(c), (real*)fd, (real*)fp, c.gate);
  T9_LAUNCH_CHECK();
}

// The in-cutoff lists of the current coordinates and the real-space direct
// field in one pass (AMOEBA; the HIPPO path compacts first).
template <class real>
void compact_dfield_real_t(Ctx& c, hipStream_t st) {
  NbList& L = c.mlist;
  if (!nblist_compact_buffers(c)) throw ArgError("no neighbor list to compact (internal ordering error)");
  const bool two = L.cut2nd > 0;
  const double c2 = L.cut * L.cut, c2b = L.cut2nd * L.cut2nd;
  auto kern = c.ewald ? (c.uniform_thole ? k_compact_dfield<real, 1, 1> : k_compact_dfield<real, 1, 0>)
                      : (c.uniform_thole ? k_compact_dfield<real, 0, 1> : k_compact_dfield<real, 0, 0>);
  KScope ks_(c, KT_DFIELD_REAL, 1, st);
  hipLaunchKernelGGL(kern, dim3(c.nwave), dim3(64 * kPairParts), 0, st, c.n, L.nlist.as<int>(),
                     L.off.as<long long>(), L.entries.as<uint32_t>(), c.pos.as<R4<real>>(), make_boxT<real>(c),
                     (real)(c2 > 1e30 ? 1e30 : c2), (real)(two ? c2b : -1.0), L.ncut.as<int>(), L.ecut.as<uint32_t>(),
                     two ? L.ncut2.as<int>() : nullptr, two ? L.ecut2.as<uint32_t>() : nullptr, c.rpole.as<real>(),
                     c.jpolar.as<int>(), c.thlval.as<real>(), c.nj, (real)c.thole_value, c.scale, ewald_k<real>(c),
                     c.field.as<real>(), c.fieldp.as<real>());
  T9_LAUNCH_CHECK();
  L.cut_valid = true;
}

template <class real>
void dfield_real_t(Ctx& c, hipStream_t st) {
  const auto& L = c.mlist;
  if (!L.cut_valid && !c.chgpen && L.cutoff > 0 && L.entries.p) {
    compact_dfield_real_t<real>(c, st);
    return;
  }
  if (!L.cut_valid) nblist_compact(c, st);
  require_cut(L);
  if (c.chgpen) {
    auto kc = c.ewald ? k_dfield_chgpen<real, 1> : k_dfield_chgpen<real, 0>;
    KScope ks_(c, KT_DFIELD_REAL, 1, st);
    hipLaunchKernelGGL(kc, dim3(c.nwave), dim3(64 * kPairParts), 0, st, c.n, L.ncut.as<int>(),
                       L.off.as<long long>(), L.ecut.as<uint32_t>(), c.pos.as<R4<real>>(), c.rpole.as<real>(),
                       c.perm.as<int>(), c.chgpen_o.as<double>(), c.scale, make_boxT<real>(c), ewald_k<double>(c),
                       c.field.as<real>(), c.fieldp.as<real>());
    T9_LAUNCH_CHECK();
    return;
  }
  auto kern = c.ewald ? (c.uniform_thole ? k_dfield_real<real, 1, 1> : k_dfield_real<real, 1, 0>)
                      : (c.uniform_thole ? k_dfield_real<real, 0, 1> : k_dfield_real<real, 0, 0>);
  KScope ks_(c, KT_DFIELD_REAL, 1, st);
  hipLaunchKernelGGL(kern, dim3(c.nwave), dim3(64 * kPairParts), 0, st, c.n, L.ncut.as<int>(),
                     L.off.as<long long>(), L.ecut.as<uint32_t>(), c.pos.as<R4<real>>(), c.rpole.as<real>(),
                     c.jpolar.as<int>(), c.thlval.as<real>(), c.nj, (real)c.thole_value, c.scale, make_boxT<real>(c),
                     ewald_k<real>(c), c.field.as<real>(), c.fieldp.as<real>());
  T9_LAUNCH_CHECK();
}

template <class real>
void precond_t(Ctx& c, const void* rd, const void* rp, void* zd, void* zp, void* z2d, void* z2p, double* dots64,
               int mode) {
  const bool sparse = mode == 0 ? false : (c.opt.pcg_prec && c.opt.usolve_cut > 0 && c.mlist.cut2nd > 0);
  if (mode == 1 && !sparse) throw ArgError("the sparse preconditioner needs pcg-precond and usolve-cutoff > 0");
  const auto& L = c.mlist;
  if (sparse) require_cut(L);
  auto kern = c.uniform_thole ? k_precond<real, 1> : k_precond<real, 0>;
  KScope ks_(c, KT_PRECOND);
  hipLaunchKernelGGL(kern, dim3(c.nwave), dim3(64 * kPairParts), 0, c.stream, c.n, sparse ? 1 : 0,
                     (real)c.opt.udiag, L.ncut2.as<int>(), L.off.as<long long>(), L.ecut2.as<uint32_t>(),
                     c.pos.as<R4<real>>(), c.polarity.as<real>(), c.jpolar.as<int>(), c.thlval.as<real>(), c.nj,
                     (real)c.thole_value, c.scale, make_boxT<real>(c), (const real*)rd, (const real*)rp, (real*)zd,
                     (real*)zp, (real*)z2d, (real*)z2p, dots64, c.gate);
  T9_LAUNCH_CHECK();
}

}  // namespace

void dfield_all(Ctx& c, const void* ua, void* fa) {
  const size_t v3 = (c.prec == 32 ? 4 : 8) * 3 * (size_t)c.n;
  const char* ua_d = (const char*)ua;
  char* fa_d = (char*)fa;
  if (!c.ewald) {
    if (c.prec == 32) dfield_real_t<float>(c, c.stream);
    else dfield_real_t<double>(c, c.stream);
    if (ua) ufield_all(c, (void*)ua_d, (void*)(ua_d + v3), fa_d, fa_d + v3);
    return;
  }
  // real space on the third stream beside the PME spreading and FFT; the
  // PME gather (+=) waits for it.  The mutual field of ua (the ASPC guess
  // of the coming solve) takes the same two routes behind the direct field:
  // real space after it on the third stream, its PME round trip on the main
  // stream while the direct field's real space still runs
  T9_HIP(hipEventRecord(c.ev_ufork, c.stream));
  T9_HIP(hipStreamWaitEvent(c.side2, c.ev_ufork, 0));
  // the in-cutoff lists of the current coordinates, if the caller left them:
  // compacted on the third stream in the same pass as the real-space direct
  // field, beside the PME work of the main stream (whose gathers wait for
  // this stream)
  if (c.prec == 32) dfield_real_t<float>(c, c.side2);
  else dfield_real_t<double>(c, c.side2);
  T9_HIP(hipEventRecord(c.ev_ujoin, c.side2));
  if (!ua) {
    pme_mpole_field(c, c.field.p, c.fieldp.p, c.ev_ujoin);
    return;
  }
  if (c.prec == 32) ufield_real_t<float>(c, ua_d, ua_d + v3, fa_d, fa_d + v3, c.side2);
  else ufield_real_t<double>(c, ua_d, ua_d + v3, fa_d, fa_d + v3, c.side2);
  T9_HIP(hipEventRecord(c.ev_ujoin2, c.side2));
  pme_mpole_dipole_fields(c, c.field.p, c.fieldp.p, c.ev_ujoin, ua_d, ua_d + v3, fa_d, fa_d + v3, c.ev_ujoin2);
}

void ufield_all(Ctx& c, void* ud, void* up, void* fd, void* fp) {
  if (!c.ewald) {
    if (c.prec == 32) ufield_real_t<float>(c, ud, up, fd, fp, c.stream);
    else ufield_real_t<double>(c, ud, up, fd, fp, c.stream);
    return;
  }
  // real space on the third stream (it writes fd, fp) while the PME part
  // spreads and transforms; the gather (fd, fp +=) waits for it
  T9_HIP(hipEventRecord(c.ev_ufork, c.stream));
  T9_HIP(hipStreamWaitEvent(c.side2, c.ev_ufork, 0));
  if (c.prec == 32) ufield_real_t<float>(c, ud, up, fd, fp, c.side2);
  else ufield_real_t<double>(c, ud, up, fd, fp, c.side2);
  T9_HIP(hipEventRecord(c.ev_ujoin, c.side2));
  pme_dipole_field(c, ud, up, fd, fp, c.ev_ujoin);
}

void precond_apply(Ctx& c, const void* rd, const void* rp, void* zd, void* zp, void* z2d, void* z2p,
                   double* dots64, int mode) {
  if (c.prec == 32) precond_t<float>(c, rd, rp, zd, zp, z2d, z2p, dots64, mode);
  else precond_t<double>(c, rd, rp, zd, zp, z2d, z2p, dots64, mode);
}

}  // namespace t9
