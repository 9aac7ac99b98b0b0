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
// each lane walking its own full neighbor list (see nblist.hip).  A lane
// accumulates its atom's field in FP64 registers and stores it once, so the
// result is independent of scheduling (no atomics).
#include "ctx.hpp"

namespace t9 {
namespace {

template <class real>
using R4 = typename RealTraits<real>::real4;

template <class real, int EWALD>
__global__ __launch_bounds__(256) void k_ufield_real(int n, const int* __restrict__ nlist,
                                                     const long long* __restrict__ woff,
                                                     const uint32_t* __restrict__ entries,
                                                     const R4<real>* __restrict__ pos, const real* __restrict__ ud,
                                                     const real* __restrict__ up, const real* __restrict__ pdamp,
                                                     const int* __restrict__ jpolar, const real* __restrict__ thlval,
                                                     int nj, ScaleTable sc, BoxT<real> box, real cut2, real aewald,
                                                     real* __restrict__ fd, real* __restrict__ fp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int len = nlist[s];
  const uint32_t* e = entries + woff[s >> 6] + (s & 63);
  const R4<real> pi = pos[s];
  const real pdi = pdamp[s];
  const real* trow = thlval + jpolar[s] * nj;
  double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
  for (int k = 0; k < len; ++k) {
    const uint32_t ent = e[(long long)k * 64];
    const int j = (int)(ent >> kCodeBits);
    const int code = (int)(ent & (kMaxCodes - 1));
    const R4<real> pj = pos[j];
    real xr = pj.x - pi.x, yr = pj.y - pi.y, zr = pj.z - pi.z;
    if (box.periodic) min_image(xr, yr, zr, box.l[0], box.l[1], box.l[2], box.il[0], box.il[1], box.il[2]);
    const real r2 = xr * xr + yr * yr + zr * zr;
    if (r2 > cut2) continue;
    const real r = t9_sqrt(r2);
    const real rinv = 1 / r;
    const real rr2 = rinv * rinv;
    const real rr3 = rinv * rr2;
    const real rr5 = 3 * rr3 * rr2;
    real s3, s5, s7;
    thole_damp(r, pdi * pdamp[j], trow[jpolar[j]], s3, s5, s7);
    const real us = (real)sc.u[code];
    real c1, c2;
    if (EWALD) {
      real bn[3];
      ewald_bn<3>(bn, r, rinv, rr2, aewald);
      c1 = bn[1] - (1 - us * s3) * rr3;
      c2 = bn[2] - (1 - us * s5) * rr5;
    } else {
      c1 = us * s3 * rr3;
      c2 = us * s5 * rr5;
    }
    const real dx = ud[3 * j], dy = ud[3 * j + 1], dz = ud[3 * j + 2];
    const real px = up[3 * j], py = up[3 * j + 1], pz = up[3 * j + 2];
    const real cd = c2 * (dx * xr + dy * yr + dz * zr);
    const real cp = c2 * (px * xr + py * yr + pz * zr);
    a0 += (double)(cd * xr - c1 * dx);
    a1 += (double)(cd * yr - c1 * dy);
    a2 += (double)(cd * zr - c1 * dz);
    b0 += (double)(cp * xr - c1 * px);
    b1 += (double)(cp * yr - c1 * py);
    b2 += (double)(cp * zr - c1 * pz);
  }
  fd[3 * s] = (real)a0;
  fd[3 * s + 1] = (real)a1;
  fd[3 * s + 2] = (real)a2;
  fp[3 * s] = (real)b0;
  fp[3 * s + 1] = (real)b1;
  fp[3 * s + 2] = (real)b2;
}

template <class real, int EWALD>
__global__ __launch_bounds__(256) void k_dfield_real(int n, const int* __restrict__ nlist,
                                                     const long long* __restrict__ woff,
                                                     const uint32_t* __restrict__ entries,
                                                     const R4<real>* __restrict__ pos, const real* __restrict__ rpole,
                                                     const real* __restrict__ pdamp, const int* __restrict__ jpolar,
                                                     const real* __restrict__ thlval, int nj, ScaleTable sc,
                                                     BoxT<real> box, real cut2, real aewald, real* __restrict__ fd,
                                                     real* __restrict__ fp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int len = nlist[s];
  const uint32_t* e = entries + woff[s >> 6] + (s & 63);
  const R4<real> pi = pos[s];
  const real pdi = pdamp[s];
  const real* trow = thlval + jpolar[s] * nj;
  double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
  for (int k = 0; k < len; ++k) {
    const uint32_t ent = e[(long long)k * 64];
    const int j = (int)(ent >> kCodeBits);
    const int code = (int)(ent & (kMaxCodes - 1));
    const R4<real> pj = pos[j];
    real xr = pj.x - pi.x, yr = pj.y - pi.y, zr = pj.z - pi.z;
    if (box.periodic) min_image(xr, yr, zr, box.l[0], box.l[1], box.l[2], box.il[0], box.il[1], box.il[2]);
    const real r2 = xr * xr + yr * yr + zr * zr;
    if (r2 > cut2) continue;
    const real r = t9_sqrt(r2);
    const real rinv = 1 / r;
    const real rr2 = rinv * rinv;
    const real rr3 = rinv * rr2;
    const real rr5 = 3 * rr3 * rr2;
    const real rr7 = 5 * rr5 * rr2;
    real s3, s5, s7;
    thole_damp(r, pdi * pdamp[j], trow[jpolar[j]], s3, s5, s7);
    const real* pk = rpole + 10 * (size_t)j;
    const real ck = pk[0], dkx = pk[1], dky = pk[2], dkz = pk[3];
    const real qxx = pk[4], qyy = pk[5], qzz = pk[6], qxy = pk[7], qxz = pk[8], qyz = pk[9];
    const real dkr = dkx * xr + dky * yr + dkz * zr;
    const real qkx = qxx * xr + qxy * yr + qxz * zr;
    const real qky = qxy * xr + qyy * yr + qyz * zr;
    const real qkz = qxz * xr + qyz * yr + qzz * zr;
    const real qkr = qkx * xr + qky * yr + qkz * zr;
    real bn[4] = {0, 0, 0, 0};
    if (EWALD) ewald_bn<4>(bn, r, rinv, rr2, aewald);
    const real ds = (real)sc.d[code], ps = (real)sc.p[code];
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
  fd[3 * s] = (real)a0;
  fd[3 * s + 1] = (real)a1;
  fd[3 * s + 2] = (real)a2;
  fp[3 * s] = (real)b0;
  fp[3 * s + 1] = (real)b1;
  fp[3 * s + 2] = (real)b2;
}

// z = udiag * alpha * r + sum_{r_ik <= ucut} alpha_i alpha_k T_ik r_k
// (sparse preconditioner, src/acc/amoeba/induce.cpp:310-420).
template <class real>
__global__ __launch_bounds__(256) void k_precond(int n, int sparse, real udiag, const int* __restrict__ nlist,
                                                 const long long* __restrict__ woff,
                                                 const uint32_t* __restrict__ entries,
                                                 const R4<real>* __restrict__ pos, const real* __restrict__ polarity,
                                                 const real* __restrict__ pdamp, const int* __restrict__ jpolar,
                                                 const real* __restrict__ thlval, int nj, ScaleTable sc,
                                                 BoxT<real> box, real cut2, const real* __restrict__ rd,
                                                 const real* __restrict__ rp, real* __restrict__ zd,
                                                 real* __restrict__ zp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const real poli = polarity[s];
  const real f = sparse ? udiag * poli : poli;
  double a0 = f * rd[3 * s], a1 = f * rd[3 * s + 1], a2 = f * rd[3 * s + 2];
  double b0 = f * rp[3 * s], b1 = f * rp[3 * s + 1], b2 = f * rp[3 * s + 2];
  if (sparse) {
    const int len = nlist[s];
    const uint32_t* e = entries + woff[s >> 6] + (s & 63);
    const R4<real> pi = pos[s];
    const real pdi = pdamp[s];
    const real* trow = thlval + jpolar[s] * nj;
    for (int k = 0; k < len; ++k) {
      const uint32_t ent = e[(long long)k * 64];
      const int j = (int)(ent >> kCodeBits);
      const int code = (int)(ent & (kMaxCodes - 1));
      const R4<real> pj = pos[j];
      real xr = pj.x - pi.x, yr = pj.y - pi.y, zr = pj.z - pi.z;
      if (box.periodic) min_image(xr, yr, zr, box.l[0], box.l[1], box.l[2], box.il[0], box.il[1], box.il[2]);
      const real r2 = xr * xr + yr * yr + zr * zr;
      if (r2 > cut2) continue;
      const real r = t9_sqrt(r2);
      real s3, s5, s7;
      thole_damp(r, pdi * pdamp[j], trow[jpolar[j]], s3, s5, s7);
      const real us = (real)sc.u[code];
      const real polik = poli * polarity[j];
      const real rr3 = us * s3 * polik / (r * r2);
      const real rr5 = 3 * us * s5 * polik / (r * r2 * r2);
      const real m0 = rr5 * xr * xr - rr3, m1 = rr5 * xr * yr, m2 = rr5 * xr * zr;
      const real m3 = rr5 * yr * yr - rr3, m4 = rr5 * yr * zr, m5 = rr5 * zr * zr - rr3;
      const real x0 = rd[3 * j], x1 = rd[3 * j + 1], x2 = rd[3 * j + 2];
      const real y0 = rp[3 * j], y1 = rp[3 * j + 1], y2 = rp[3 * j + 2];
      a0 += (double)(m0 * x0 + m1 * x1 + m2 * x2);
      a1 += (double)(m1 * x0 + m3 * x1 + m4 * x2);
      a2 += (double)(m2 * x0 + m4 * x1 + m5 * x2);
      b0 += (double)(m0 * y0 + m1 * y1 + m2 * y2);
      b1 += (double)(m1 * y0 + m3 * y1 + m4 * y2);
      b2 += (double)(m2 * y0 + m4 * y1 + m5 * y2);
    }
  }
  zd[3 * s] = (real)a0;
  zd[3 * s + 1] = (real)a1;
  zd[3 * s + 2] = (real)a2;
  zp[3 * s] = (real)b0;
  zp[3 * s + 1] = (real)b1;
  zp[3 * s + 2] = (real)b2;
}

template <class real>
BoxT<real> box_of(const Ctx& c) {
  BoxT<real> b;
  for (int d = 0; d < 3; ++d) {
    b.l[d] = (real)(c.periodic ? c.box[d] : 1.0);
    b.il[d] = (real)(c.periodic ? 1.0 / c.box[d] : 1.0);
  }
  b.periodic = c.periodic ? 1 : 0;
  return b;
}

template <class real>
real cut2_of(double cut) {
  double c2 = cut * cut;
  return c2 > 1e30 ? (real)1e30 : (real)c2;
}

template <class real>
void ufield_real_t(Ctx& c, const void* ud, const void* up, void* fd, void* fp) {
  const auto& L = c.mlist;
  auto kern = c.ewald ? k_ufield_real<real, 1> : k_ufield_real<real, 0>;
  KScope ks_(c, KT_UFIELD_REAL);
  hipLaunchKernelGGL(kern, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, L.nlist.as<int>(),
                     L.off.as<long long>(), L.entries.as<uint32_t>(), c.pos.as<R4<real>>(), (const real*)ud,
                     (const real*)up, c.pdamp.as<real>(), c.jpolar.as<int>(), c.thlval.as<real>(), c.nj, c.scale,
                     box_of<real>(c), cut2_of<real>(c.cut), (real)c.aewald, (real*)fd, (real*)fp);
  T9_LAUNCH_CHECK();
}

template <class real>
void dfield_real_t(Ctx& c) {
  const auto& L = c.mlist;
  auto kern = c.ewald ? k_dfield_real<real, 1> : k_dfield_real<real, 0>;
  KScope ks_(c, KT_DFIELD_REAL);
  hipLaunchKernelGGL(kern, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, L.nlist.as<int>(),
                     L.off.as<long long>(), L.entries.as<uint32_t>(), c.pos.as<R4<real>>(), c.rpole.as<real>(),
                     c.pdamp.as<real>(), c.jpolar.as<int>(), c.thlval.as<real>(), c.nj, c.scale, box_of<real>(c),
                     cut2_of<real>(c.cut), (real)c.aewald, c.field.as<real>(), c.fieldp.as<real>());
  T9_LAUNCH_CHECK();
}

template <class real>
void precond_t(Ctx& c, const void* rd, const void* rp, void* zd, void* zp) {
  const bool sparse = c.opt.pcg_prec && c.opt.usolve_cut > 0 && c.ulist.cutoff > 0;
  const auto& L = sparse ? c.ulist : c.mlist;
  KScope ks_(c, KT_PRECOND);
  hipLaunchKernelGGL(k_precond<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, sparse ? 1 : 0,
                     (real)c.opt.udiag, L.nlist.as<int>(), L.off.as<long long>(), L.entries.as<uint32_t>(),
                     c.pos.as<R4<real>>(), c.polarity.as<real>(), c.pdamp.as<real>(), c.jpolar.as<int>(),
                     c.thlval.as<real>(), c.nj, c.scale, box_of<real>(c), cut2_of<real>(c.opt.usolve_cut),
                     (const real*)rd, (const real*)rp, (real*)zd, (real*)zp);
  T9_LAUNCH_CHECK();
}

}  // namespace

void dfield_all(Ctx& c) {
  if (c.prec == 32) dfield_real_t<float>(c);
  else dfield_real_t<double>(c);
  if (c.ewald) pme_mpole_field(c, c.field.p, c.fieldp.p);
}

void ufield_all(Ctx& c, void* ud, void* up, void* fd, void* fp) {
  if (c.prec == 32) ufield_real_t<float>(c, ud, up, fd, fp);
  else ufield_real_t<double>(c, ud, up, fd, fp);
  if (c.ewald) pme_dipole_field(c, ud, up, fd, fp);
}

void precond_apply(Ctx& c, const void* rd, const void* rp, void* zd, void* zp) {
  if (c.prec == 32) precond_t<float>(c, rd, rp, zd, zp);
  else precond_t<double>(c, rd, rp, zd, zp);
}

}  // namespace t9
