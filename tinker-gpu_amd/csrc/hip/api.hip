// C ABI of the device context (t9_ctx_*, t9_nblist_*, t9_mpole_init,
// t9_dfield, t9_ufield, t9_induce): context creation from a host system,
// caller-order <-> spatial-order transfers, and the "prepare" step that keeps
// the neighbor list, sorted coordinates and global multipoles current.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "../../../include/t9hip.h"
#include "ctx.hpp"

const t9::System& t9_system_ref(const t9_system* sys);

struct t9_ctx {
  t9::Ctx c;
};

namespace t9 {

std::vector<double> pme_bsmod(int K, int order);
int pme_tile_edge(int K);
void pme_setup_bins(Ctx& c);

namespace {

template <class real>
__global__ void k_sorted_to_orig3(int n, const int* __restrict__ perm, const real* __restrict__ src,
                                  double* __restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  for (int d = 0; d < 3; ++d) dst[3 * i + d] = (double)src[3 * s + d];
}

template <class real>
__global__ void k_orig_to_sorted3(int n, const int* __restrict__ perm, const double* __restrict__ src,
                                  real* __restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  for (int d = 0; d < 3; ++d) dst[3 * s + d] = (real)src[3 * i + d];
}

template <class real>
__global__ void k_sorted_to_orig10(int n, const int* __restrict__ perm, const real* __restrict__ src,
                                   double* __restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  for (int d = 0; d < 10; ++d) dst[10 * i + d] = (double)src[10 * s + d];
}

template <class T>
void upload(DBuf& b, const std::vector<T>& v, hipStream_t st) {
  b.alloc(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) T9_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

void alloc_vectors(Ctx& c) {
  const size_t vb = rsize(c) * 3 * (size_t)c.n;
  for (DBuf* b : {&c.field, &c.fieldp, &c.udir, &c.udirp, &c.uind, &c.uinp, &c.rsd, &c.rsdp, &c.zrsd, &c.zrsdp,
                  &c.conj, &c.conjp, &c.vec, &c.vecp}) {
    b->alloc(vb);
    T9_HIP(hipMemsetAsync(b->p, 0, vb, c.stream));
  }
  c.partials.alloc(sizeof(double) * 4 * (ceil_div(c.n, 256) + 1));
  c.partials64.alloc(sizeof(double) * 4 * (ceil_div(c.n, 64) + 1));
  c.pcg_trace.alloc(sizeof(double) * kTraceCap);
  c.scalars.alloc(sizeof(double) * 48);
  T9_HIP(hipMemsetAsync(c.scalars.p, 0, sizeof(double) * 48, c.stream));
  T9_HIP(hipHostMalloc((void**)&c.h_scalars, sizeof(double) * 48, hipHostMallocDefault));
}

}  // namespace

void ctx_init(Ctx& c, const System& s, int precision) {
  if (precision != 32 && precision != 64) throw ArgError("precision must be 32 or 64");
  c.prec = precision;
  c.n = s.n;
  c.nwave = ceil_div(s.n, 64);
  c.opt = s.opt;
  c.periodic = s.periodic;
  for (int d = 0; d < 3; ++d) c.box[d] = s.box[d];
  c.ewald = s.opt.ewald;
  c.aewald = s.opt.ewald_alpha;
  c.cut = c.ewald ? s.opt.ewald_cut : s.opt.mpole_cut;
  const bool elec = s.opt.use_mpole || s.opt.use_polar;
  // charge transfer walks the multipole list, whose radius covers the
  // largest cutoff of the terms on it; the in-cutoff compaction keeps the
  // electrostatic cutoff (or the list's, without the electrostatic terms)
  c.list_cut = elec ? c.cut : 0.0;
  if (s.chgtrn) {
    if (!(s.opt.ct_cut > 0) || s.opt.ct_taper > s.opt.ct_cut) throw ArgError("bad chgtrn-cutoff or chgtrn-taper");
    c.list_cut = std::max(c.list_cut, s.opt.ct_cut);
  }
  if (s.repel) {
    if (!(s.opt.rep_cut > 0) || s.opt.rep_taper > s.opt.rep_cut) throw ArgError("bad repuls-cutoff or repuls-taper");
    c.list_cut = std::max(c.list_cut, s.opt.rep_cut);
  }
  if (s.disp) {
    if (!(s.opt.disp_cut > 0) || s.opt.disp_taper > s.opt.disp_cut) throw ArgError("bad disp-cutoff or disp-taper");
    c.list_cut = std::max(c.list_cut, s.opt.disp_cut);
  }
  if (!elec && c.list_cut > 0) c.cut = c.list_cut;
  if (elec) c.list_cut = std::max(c.list_cut, c.cut);
  if (elec && c.periodic && !c.ewald && c.cut > 1e20)
    throw ArgError("periodic systems need ewald or a finite mpole-cutoff");
  if (!c.periodic && c.ewald) throw ArgError("ewald requires a periodic box");
  for (int d = 0; d < 3; ++d) c.grid[d] = s.opt.pme_grid[d];
  T9_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  T9_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
  // fork/join events order kernels of one device: the producer kernel's own
  // release already makes its writes visible to the consumer, so they skip
  // the host-scope fence (the readback event keeps it)
  const unsigned evf = hipEventDisableTiming | hipEventDisableSystemFence;
  T9_HIP(hipEventCreateWithFlags(&c.ev_fork, evf));
  T9_HIP(hipEventCreateWithFlags(&c.ev_join, evf));
  T9_HIP(hipStreamCreateWithFlags(&c.side2, hipStreamNonBlocking));
  T9_HIP(hipEventCreateWithFlags(&c.ev_ufork, evf));
  T9_HIP(hipEventCreateWithFlags(&c.ev_ujoin, evf));
  T9_HIP(hipEventCreateWithFlags(&c.ev_ujoin2, evf));
  T9_HIP(hipStreamCreateWithFlags(&c.side3, hipStreamNonBlocking));
  T9_HIP(hipEventCreateWithFlags(&c.ev_bfork, evf));
  T9_HIP(hipEventCreateWithFlags(&c.ev_bins, evf));
  T9_HIP(hipEventCreateWithFlags(&c.ev_read, hipEventDisableTiming));
  hipStream_t st = c.stream;
  const int n = s.n;

  // ---- static per-atom data, caller order
  upload(c.xyz, s.xyz, st);
  c.xyz_last.alloc(sizeof(double) * 3 * n);
  upload(c.pole, s.pole, st);
  upload(c.frame, s.frame, st);
  upload(c.zaxis, s.zaxis, st);
  upload(c.xaxis, s.xaxis, st);
  upload(c.yaxis, s.yaxis, st);
  upload(c.ysign, s.ysign, st);
  upload(c.polarity_o, s.polarity, st);
  upload(c.pdamp_o, s.pdamp, st);
  upload(c.jpolar_o, s.jpolar, st);
  c.chgpen = s.chgpen;
  if (c.chgpen) {
    std::vector<double> cp(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      cp[3 * i] = s.pcore[i];
      cp[3 * i + 1] = s.pval[i];
      cp[3 * i + 2] = s.palpha[i];
    }
    upload(c.chgpen_o, cp, st);
    T9_HIP(hipStreamSynchronize(st));  // cp is a local host buffer
  }
  c.use_chgtrn = s.chgtrn;
  if (c.use_chgtrn) {
    c.ct_cut = s.opt.ct_cut;
    c.ct_taper = s.opt.ct_taper;
    std::vector<double> ct(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      ct[2 * i] = s.chgct[i];
      ct[2 * i + 1] = s.dmpct[i];
    }
    upload(c.chgtrn_o, ct, st);
    T9_HIP(hipStreamSynchronize(st));  // ct is a local host buffer
  }
  c.use_repel = s.repel;
  if (c.use_repel) {
    c.rep_cut = s.opt.rep_cut;
    c.rep_taper = s.opt.rep_taper;
    std::vector<double> rp(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      rp[3 * i] = s.sizpr[i];
      rp[3 * i + 1] = s.dmppr[i];
      rp[3 * i + 2] = s.elepr[i];
    }
    upload(c.repel_o, rp, st);
    T9_HIP(hipStreamSynchronize(st));  // rp is a local host buffer
  }
  c.use_disp = s.disp;
  if (c.use_disp) {
    c.disp_cut = s.opt.disp_cut;
    c.disp_taper = s.opt.disp_taper;
    std::vector<double> cd(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      cd[2 * i] = s.csix[i];
      cd[2 * i + 1] = s.adisp[i];
    }
    upload(c.disp_o, cd, st);
    T9_HIP(hipStreamSynchronize(st));  // cd is a local host buffer
  }
  c.nj = s.njpolar;
  c.uniform_thole = !s.thlval.empty();
  c.thole_value = s.thlval.empty() ? 0.0 : s.thlval[0];
  for (double v : s.thlval)
    if (v != c.thole_value) c.uniform_thole = false;
  if (c.prec == 32) {
    std::vector<float> t(s.thlval.begin(), s.thlval.end());
    upload(c.thlval, t, st);
  } else {
    upload(c.thlval, s.thlval, st);
  }

  // ---- scale classes and the exclusion CSR (both directions)
  std::map<std::tuple<double, double, double, double, double, double>, int> classes;
  classes[{1.0, 1.0, 1.0, 1.0, 1.0, 1.0}] = 0;
  std::vector<std::vector<std::pair<int, int>>> ex(n);
  for (const auto& p : s.scales) {
    auto key = std::make_tuple(p.m, p.d, p.p, p.u, p.ds, p.rs);
    auto it = classes.find(key);
    int code;
    if (it == classes.end()) {
      code = (int)classes.size();
      if (code >= kMaxCodes) throw ArgError("too many distinct pair scale classes");
      classes[key] = code;
    } else {
      code = it->second;
    }
    ex[p.i].push_back({p.k, code});
    ex[p.k].push_back({p.i, code});
  }
  c.ncodes = (int)classes.size();
  for (int k = 0; k < kMaxCodes; ++k)
    c.scale.m[k] = c.scale.d[k] = c.scale.p[k] = c.scale.u[k] = c.scale.ds[k] = c.scale.rs[k] = 1.0;
  for (auto& kv : classes) {
    int k = kv.second;
    c.scale.m[k] = std::get<0>(kv.first);
    c.scale.d[k] = std::get<1>(kv.first);
    c.scale.p[k] = std::get<2>(kv.first);
    c.scale.u[k] = std::get<3>(kv.first);
    c.scale.ds[k] = std::get<4>(kv.first);
    c.scale.rs[k] = std::get<5>(kv.first);
  }
  std::vector<int> eoff(n + 1, 0), ej, ecode;
  for (int i = 0; i < n; ++i) {
    eoff[i + 1] = eoff[i] + (int)ex[i].size();
    for (auto& pr : ex[i]) {
      ej.push_back(pr.first);
      ecode.push_back(pr.second);
    }
  }
  upload(c.excl_off, eoff, st);
  upload(c.excl_j, ej, st);
  upload(c.excl_code, ecode, st);

  // ---- sorted-order arrays
  const size_t rs = rsize(c);
  // fixed-point frame: the box, or a virtual box four times the extent
  if (c.periodic) {
    for (int d = 0; d < 3; ++d) {
      c.vorg[d] = 0;
      c.vlen[d] = c.box[d];
    }
  } else {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], s.xyz[3 * i + d]);
        hi[d] = std::max(hi[d], s.xyz[3 * i + d]);
      }
    double ext = 0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, n > 0 ? hi[d] - lo[d] : 0.0);
    const double len = 4 * ext + 64;
    for (int d = 0; d < 3; ++d) {
      c.vlen[d] = len;
      c.vorg[d] = (n > 0 ? 0.5 * (lo[d] + hi[d]) : 0.0) - 0.5 * len;
    }
  }
  c.perm.alloc(sizeof(int) * n);
  c.iperm.alloc(sizeof(int) * n);
  c.cellc.alloc(sizeof(int) * n);
  c.pos.alloc(rs == 4 ? sizeof(PosRec<float>) * n : sizeof(PosRec<double>) * n);
  c.rpole.alloc(rs * 10 * n);
  c.polarity.alloc(rs * n);
  c.polinv.alloc(rs * n);
  c.pdamp.alloc(rs * n);
  c.jpolar.alloc(sizeof(int) * n);

  // ---- neighbor-list geometry
  const double buf = s.opt.list_buffer;
  c.mlist.cutoff = c.list_cut > 0 && c.list_cut < 1e20 ? c.list_cut + buf : (c.cut < 1e20 ? c.cut + buf : 1e30);
  const bool sparse = s.opt.pcg_prec && s.opt.usolve_cut > 0;
  // the preconditioner's pairs (usolve cutoff) are a second compaction of the
  // multipole list, which covers them at any time between rebuilds
  c.ulist.cutoff = 0;
  c.mlist.cut = c.cut < 1e20 ? c.cut : 1e30;
  c.mlist.cut2nd = sparse ? std::min(s.opt.usolve_cut, c.mlist.cut) : 0;
  // vdW sites sit off their atoms by (1 - red) times the bond to the parent;
  // the cell search around the atom's cell must cover twice that offset
  double site_slack = 0;
  if (s.opt.use_vdw) {
    c.vlist.cutoff = s.opt.vdw_cut + buf;
    c.vlist.cut = s.opt.vdw_cut;
    double off = 0;
    for (int i = 0; i < n; ++i) {
      const int p = s.vdw_parent[i];
      if (p == i) continue;
      double r2 = 0;
      for (int d = 0; d < 3; ++d) {
        double x = s.xyz[3 * i + d] - s.xyz[3 * p + d];
        if (c.periodic) x -= c.box[d] * std::nearbyint(x / c.box[d]);
        r2 += x * x;
      }
      off = std::max(off, (1.0 - s.vdw_red[i]) * std::sqrt(r2));
    }
    site_slack = 2.0 * (1.5 * off + 0.1);
  }
  c.brute = true;
  c.morton_bits = 0;
  if (c.periodic) {
    const double rl = std::max(c.mlist.cutoff, c.vlist.cutoff + site_slack);
    c.cell_reach = rl;
    for (int d = 0; d < 3; ++d)
      if (std::max(c.mlist.cutoff, c.vlist.cutoff) > 0.5 * c.box[d] + 1e-9)
        throw ArgError("cutoff plus list buffer exceeds half the box edge");
    bool ok = true;
    int mx = 1;
    for (int d = 0; d < 3; ++d) {
      c.ncell[d] = std::min(1024, (int)std::floor(c.box[d] / (0.5 * rl)));
      if (c.ncell[d] < 5) ok = false;
      mx = std::max(mx, c.ncell[d]);
    }
    if (ok) {
      c.brute = false;
      int bits = 0;
      while ((1 << bits) < mx) ++bits;
      c.morton_bits = bits;
    } else {
      c.ncell[0] = c.ncell[1] = c.ncell[2] = 1;
    }
  }

  // ---- PME
  if (c.ewald) {
    std::vector<double> bs;
    std::vector<double> tw;
    for (int d = 0; d < 3; ++d) {
      const int K = c.grid[d];
      if (K < 5) throw ArgError("PME grid too small");
      auto m = pme_bsmod(K, s.opt.pme_order);
      bs.insert(bs.end(), m.begin(), m.end());
      for (int k = 0; k < K; ++k) {
        tw.push_back(std::cos(2 * M_PI * k / K));
        tw.push_back(-std::sin(2 * M_PI * k / K));
      }
      c.tsz[d] = pme_tile_edge(K);
      c.ntile[d] = K / c.tsz[d];
    }
    pme_setup_bins(c);
    pme_check_fft_sizes(c);
    upload(c.bsmod, bs, st);
    if (c.prec == 32) {
      std::vector<float> t(tw.begin(), tw.end());
      upload(c.twiddle, t, st);
    } else {
      upload(c.twiddle, tw, st);
    }
    const size_t ng = (size_t)c.grid[0] * c.grid[1] * c.grid[2];
    c.qgrid.alloc(2 * rs * ng);
    c.pme_w.alloc(rs * 4 * n);
    c.pme_base.alloc(sizeof(int4) * n);
    const int nt = c.nbin[0] * c.nbin[1] * c.nbin[2];
    c.tile_count.alloc(sizeof(int) * nt);
    T9_HIP(hipMemsetAsync(c.tile_count.p, 0, sizeof(int) * nt, st));  // each binning's scan clears it again
    c.tile_start.alloc(sizeof(int) * (nt + 1));
    c.tile_cursor.alloc(sizeof(int) * nt);
    c.tile_atoms.alloc(sizeof(int) * n);
    c.tile_csum.alloc(sizeof(int) * ((nt + 1023) / 1024 + 1));
  }
  alloc_vectors(c);
  if (s.opt.predictor == "ASPC") {
    c.ualt.alloc(rs * 6 * (size_t)n * 16);
    guess_setup(c);
  }
  else if (s.opt.predictor != "NONE") throw ArgError("polar-predict " + s.opt.predictor + " is not supported");
  energy_setup(c, s);
  md_setup(c, s);
  T9_HIP(hipStreamSynchronize(st));
}

void gather_to_host(Ctx& c, const void* dev_sorted, double* host_orig) {
  DBuf tmp;
  tmp.alloc(sizeof(double) * 3 * c.n);
  if (c.prec == 32)
    hipLaunchKernelGGL(k_sorted_to_orig3<float>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                       c.perm.as<int>(), (const float*)dev_sorted, tmp.as<double>());
  else
    hipLaunchKernelGGL(k_sorted_to_orig3<double>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                       c.perm.as<int>(), (const double*)dev_sorted, tmp.as<double>());
  T9_LAUNCH_CHECK();
  T9_HIP(hipMemcpyAsync(host_orig, tmp.p, sizeof(double) * 3 * c.n, hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
}

void scatter_from_host(Ctx& c, const double* host_orig, void* dev_sorted) {
  DBuf tmp;
  tmp.alloc(sizeof(double) * 3 * c.n);
  T9_HIP(hipMemcpyAsync(tmp.p, host_orig, sizeof(double) * 3 * c.n, hipMemcpyHostToDevice, c.stream));
  if (c.prec == 32)
    hipLaunchKernelGGL(k_orig_to_sorted3<float>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                       c.perm.as<int>(), tmp.as<double>(), (float*)dev_sorted);
  else
    hipLaunchKernelGGL(k_orig_to_sorted3<double>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                       c.perm.as<int>(), tmp.as<double>(), (double*)dev_sorted);
  T9_LAUNCH_CHECK();
  T9_HIP(hipStreamSynchronize(c.stream));
}

void mark_readback(Ctx& c) { T9_HIP(hipEventRecord(c.ev_read, c.stream)); }

void wait_marked(Ctx& c) {
  for (;;) {
    const hipError_t e = hipEventQuery(c.ev_read);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) T9_HIP(e);
  }
}

// Make the list, sorted coordinates and global multipoles current.
void prepare(Ctx& c, bool defer_compact) {
  if (!c.list_valid) {
    nblist_build(c);
  } else if (c.coords_dirty) {
    if (nblist_needs_rebuild(c, true)) nblist_build(c);
    else nblist_mark_refreshed(c);
  }
  pme_bin_async(c);  // beside the compaction and the rotation
  if (!c.mlist.cut_valid && !(defer_compact && c.ewald)) nblist_compact(c);
  if (!c.rpole_valid) mpole_init(c);
}

void ktimer_collect(Ctx& c, int id) {
  KTimer& t = c.timers[id];
  T9_HIP(hipStreamSynchronize(c.stream));
  for (size_t k = 0; k + 1 < t.used; k += 2) {
    float ms = 0;
    T9_HIP(hipEventElapsedTime(&ms, t.ev[k], t.ev[k + 1]));
    t.ms += ms;
    t.count += t.nl[k / 2];
  }
  t.used = 0;
}

void ktimer_release(Ctx& c) {
  for (auto& t : c.timers)
    for (auto e : t.ev) (void)hipEventDestroy(e);
}

namespace {

template <class real>
__global__ void k_count_cut(int n, const int* __restrict__ nlist, const long long* __restrict__ woff,
                            const uint32_t* __restrict__ entries, const PosRec<real>* __restrict__ pos,
                            BoxT<real> box, real cut2, unsigned long long* __restrict__ out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const auto pi = pos[s];
  const uint32_t* e = entries + woff[s >> 6] + (s & 63);
  unsigned long long c = 0;
  for (int k = 0; k < nlist[s]; ++k) {
    const int j = (int)(e[(long long)k * 64] >> kCodeBits);
    const auto pj = pos[j];
    real xr, yr, zr;
    pair_disp(pi, pj, box, xr, yr, zr);
    if (xr * xr + yr * yr + zr * zr <= cut2) ++c;
  }
  atomicAdd(out, c);
}

template <class real>
long long count_cut_t(Ctx& c) {
  DBuf cnt;
  cnt.alloc(sizeof(unsigned long long));
  T9_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), c.stream));
  BoxT<real> b = make_boxT<real>(c);
  double c2 = c.cut * c.cut;
  hipLaunchKernelGGL(k_count_cut<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                     c.mlist.nlist.as<int>(), c.mlist.off.as<long long>(), c.mlist.entries.as<uint32_t>(),
                     c.pos.as<PosRec<real>>(), b, (real)(c2 > 1e30 ? 1e30 : c2),
                     cnt.as<unsigned long long>());
  T9_LAUNCH_CHECK();
  unsigned long long h = 0;
  T9_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  return (long long)h;
}

__global__ void k_copy_probe(long long n4, const float4* __restrict__ a, float4* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

}  // namespace

long long count_cutoff_pairs(Ctx& c) { return c.prec == 32 ? count_cut_t<float>(c) : count_cut_t<double>(c); }

void rpole_to_host(Ctx& c, double* out) {
  DBuf tmp;
  tmp.alloc(sizeof(double) * 10 * c.n);
  if (c.prec == 32)
    hipLaunchKernelGGL(k_sorted_to_orig10<float>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                       c.perm.as<int>(), c.rpole.as<float>(), tmp.as<double>());
  else
    hipLaunchKernelGGL(k_sorted_to_orig10<double>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n,
                       c.perm.as<int>(), c.rpole.as<double>(), tmp.as<double>());
  T9_LAUNCH_CHECK();
  T9_HIP(hipMemcpyAsync(out, tmp.p, sizeof(double) * 10 * c.n, hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
}

// One evaluation on the current coordinates: list refresh, induced dipoles,
// then the requested term groups.  Returns the CG iteration count.
int evaluate(Ctx& c, int vers, int terms, double* e16, double* grad, bool grad_on_device = false) {
  int iters = 0;
  if (c.vdw_pending) {  // left over from a call that threw: drain and drop it
    T9_HIP(hipStreamSynchronize(c.side));
    c.vdw_pending = false;
  }
  if (terms & (TERM_ELEC | TERM_VDW | TERM_CHGTRN | TERM_DISP | TERM_REPEL)) {
    c.mlist.cut_valid = c.ulist.cut_valid = false;
    c.bins_valid = false;
    c.rpole_valid = false;
    const bool solve = (terms & TERM_ELEC) && c.opt.use_polar;
    prepare(c, solve);
    if (solve && (terms & TERM_VDW) && c.use_vdw) {
      // the vdW pair kernel does not depend on the dipoles: it runs on the
      // second stream during the solve, into the first energy rows
      c.erows = 0;
      double* rows = energy_rows(c, c.nwave);
      T9_HIP(hipEventRecord(c.ev_fork, c.stream));
      T9_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
      vdw_launch(c, (vers & 4) != 0, c.side, rows);
      T9_HIP(hipEventRecord(c.ev_join, c.side));
      c.vdw_pending = true;
    }
    if (solve) {
      double rms = 0;
      if (induce_pcg(c, &iters, &rms) != 0) throw NotConverged("INDUCE -- induced dipoles are not converged");
    }
  } else if (!c.list_valid) {
    prepare(c);  // the sorted order (perm) the gradient is mapped back through
  }
  energy_run(c, vers, terms, e16, grad, grad_on_device);
  return iters;
}

void fill_energy_out(const double* e, int vers, int iters, t9_energy_out* out) {
  out->em_real = e[EP_MREAL];
  out->em_recip = e[EP_MRECIP];
  out->em_self = e[EP_MSELF];
  out->em = out->em_real + out->em_recip + out->em_self;
  out->ep = e[EP_POLAR];
  out->evdw = e[EP_EVDW];
  out->ebond = e[EP_EBOND];
  out->eangle = e[EP_EANGLE];
  out->eurey = e[EP_EUREY];
  out->ect = e[EP_ECT];
  out->edisp = e[EP_EDISP];
  out->erep = e[EP_EREP];
  out->esum = out->em + out->ep + out->evdw + out->ebond + out->eangle + out->eurey + out->ect + out->edisp +
              out->erep;
  const double v6[6] = {e[EP_VXX], e[EP_VYY], e[EP_VZZ], e[EP_VXY], e[EP_VXZ], e[EP_VYZ]};
  const double v9[9] = {v6[0], v6[3], v6[4], v6[3], v6[1], v6[5], v6[4], v6[5], v6[2]};
  for (int k = 0; k < 9; ++k) out->vir[k] = (vers & T9_CALC_VIRIAL) ? v9[k] : 0.0;
  out->induce_iterations = iters;
}

}  // namespace t9

using t9::guarded;

extern "C" {

int t9_ctx_create(const t9_system* sys, int precision, t9_ctx** out) {
  return guarded([&] {
    if (!sys || !out) throw t9::ArgError("null argument");
    auto* h = new t9_ctx;
    try {
      t9::ctx_init(h->c, t9_system_ref(sys), precision);
    } catch (...) {
      if (h->c.stream) (void)hipStreamDestroy(h->c.stream);
      if (h->c.side) (void)hipStreamDestroy(h->c.side);
      if (h->c.ev_fork) (void)hipEventDestroy(h->c.ev_fork);
      if (h->c.ev_join) (void)hipEventDestroy(h->c.ev_join);
      if (h->c.side2) (void)hipStreamDestroy(h->c.side2);
      if (h->c.ev_ufork) (void)hipEventDestroy(h->c.ev_ufork);
      if (h->c.ev_ujoin) (void)hipEventDestroy(h->c.ev_ujoin);
      if (h->c.ev_ujoin2) (void)hipEventDestroy(h->c.ev_ujoin2);
      if (h->c.side3) (void)hipStreamDestroy(h->c.side3);
      if (h->c.ev_bfork) (void)hipEventDestroy(h->c.ev_bfork);
      if (h->c.ev_bins) (void)hipEventDestroy(h->c.ev_bins);
      if (h->c.ev_read) (void)hipEventDestroy(h->c.ev_read);
      if (h->c.h_scalars) (void)hipHostFree(h->c.h_scalars);
      delete h;
      throw;
    }
    *out = h;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_ctx_destroy(t9_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return T9_OK;
    (void)hipStreamSynchronize(ctx->c.stream);
    (void)hipStreamSynchronize(ctx->c.side);
    (void)hipStreamSynchronize(ctx->c.side2);
    (void)hipStreamSynchronize(ctx->c.side3);
    hipStream_t st = ctx->c.stream, side = ctx->c.side, side2 = ctx->c.side2, side3 = ctx->c.side3;
    hipEvent_t e7 = ctx->c.ev_bfork, e8 = ctx->c.ev_bins;
    hipEvent_t e1 = ctx->c.ev_fork, e2 = ctx->c.ev_join, e3 = ctx->c.ev_ufork, e4 = ctx->c.ev_ujoin;
    hipEvent_t e5 = ctx->c.ev_read, e6 = ctx->c.ev_ujoin2;
    double* hs = ctx->c.h_scalars;
    t9::ktimer_release(ctx->c);
    delete ctx;
    if (hs) (void)hipHostFree(hs);
    if (st) (void)hipStreamDestroy(st);
    if (side) (void)hipStreamDestroy(side);
    if (e1) (void)hipEventDestroy(e1);
    if (e2) (void)hipEventDestroy(e2);
    if (side2) (void)hipStreamDestroy(side2);
    if (e3) (void)hipEventDestroy(e3);
    if (e4) (void)hipEventDestroy(e4);
    if (e5) (void)hipEventDestroy(e5);
    if (e6) (void)hipEventDestroy(e6);
    if (side3) (void)hipStreamDestroy(side3);
    if (e7) (void)hipEventDestroy(e7);
    if (e8) (void)hipEventDestroy(e8);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_set_coords(t9_ctx* ctx, const double* xyz) {
  return guarded([&] {
    auto& c = ctx->c;
    T9_HIP(hipMemcpyAsync(c.xyz.p, xyz, sizeof(double) * 3 * c.n, hipMemcpyHostToDevice, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    c.coords_dirty = true;
    c.rpole_valid = false;
    c.bins_valid = false;
    c.md_ready = false;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_get_coords(t9_ctx* ctx, double* xyz) {
  return guarded([&] {
    auto& c = ctx->c;
    T9_HIP(hipMemcpyAsync(xyz, c.xyz.p, sizeof(double) * 3 * c.n, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_nblist_refresh(t9_ctx* ctx, int* rebuilt) {
  return guarded([&] {
    auto& c = ctx->c;
    bool rb = false;
    if (!c.list_valid || nblist_needs_rebuild(c)) {
      t9::nblist_build(c);
      rb = true;
    } else if (c.coords_dirty) {
      t9::refresh_sorted_coords(c);
    }
    if (rebuilt) *rebuilt = rb ? 1 : 0;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_nblist_build(t9_ctx* ctx) {
  return guarded([&] {
    t9::nblist_build(ctx->c);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_nblist_stats(t9_ctx* ctx, int64_t* out5) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!c.list_valid) t9::nblist_build(c);
    std::vector<int> nl(c.n);
    T9_HIP(hipMemcpyAsync(nl.data(), c.mlist.nlist.p, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    long long pairs = 0;
    for (int v : nl) pairs += v;
    out5[0] = pairs;
    out5[1] = c.n;
    out5[2] = c.nwave;
    out5[3] = c.mlist.total;
    out5[4] = c.builds;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_nblist_pairs(t9_ctx* ctx, int which, int64_t cap, int* pairs, int64_t* count) {
  return guarded([&] {
    auto& c = ctx->c;
    if (which != 0 && which != 1) throw t9::ArgError("which must be 0 (multipole list) or 1 (vdW list)");
    if (cap < 0 || (cap > 0 && !pairs)) throw t9::ArgError("bad pair buffer");
    if (!c.list_valid) t9::nblist_build(c);
    t9::NbList& L = which == 0 ? c.mlist : c.vlist;
    if (L.cutoff <= 0) throw t9::ArgError("that list is not built for this system");
    const int n = c.n, nw = c.nwave;
    std::vector<int> nl(n), perm(n);
    std::vector<long long> off(nw + 1);
    std::vector<uint32_t> ent((size_t)std::max<long long>(L.total, 1));
    T9_HIP(hipMemcpyAsync(nl.data(), L.nlist.p, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipMemcpyAsync(perm.data(), c.perm.p, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipMemcpyAsync(off.data(), L.off.p, sizeof(long long) * (nw + 1), hipMemcpyDeviceToHost, c.stream));
    if (L.total > 0)
      T9_HIP(hipMemcpyAsync(ent.data(), L.entries.p, sizeof(uint32_t) * L.total, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    // each pair appears in both atoms' lists: report it once, caller order i < k
    long long m = 0;
    for (int s = 0; s < n; ++s) {
      const int i = perm[s];
      for (int k = 0; k < nl[s]; ++k) {
        const long long at = off[s >> 6] + 64LL * k + (s & 63);
        if (at < 0 || at >= L.total) throw t9::ArgError("neighbor list entry out of range (internal error)");
        const int j = (int)(ent[at] >> t9::kCodeBits);
        if (j < 0 || j >= n) throw t9::ArgError("neighbor index out of range (internal error)");
        const int kk = perm[j];
        if (i < kk) {
          if (m < cap) {
            pairs[2 * m] = i;
            pairs[2 * m + 1] = kk;
          }
          ++m;
        }
      }
    }
    if (count) *count = m;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_mpole_init(t9_ctx* ctx) {
  return guarded([&] {
    auto& c = ctx->c;
    // a new evaluation: every coordinate-derived cache (in-cutoff lists, PME
    // spline bins, global multipoles) is rebuilt from the current coordinates
    c.mlist.cut_valid = c.ulist.cut_valid = false;
    c.bins_valid = false;
    c.rpole_valid = false;
    t9::prepare(c);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_get_rpole(t9_ctx* ctx, double* rpole) {
  return guarded([&] {
    auto& c = ctx->c;
    t9::prepare(c);
    t9::rpole_to_host(c, rpole);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_dfield(t9_ctx* ctx, double* field, double* fieldp) {
  return guarded([&] {
    auto& c = ctx->c;
    t9::prepare(c);
    t9::dfield_all(c);
    if (field) t9::gather_to_host(c, c.field.p, field);
    if (fieldp) t9::gather_to_host(c, c.fieldp.p, fieldp);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_ufield(t9_ctx* ctx, const double* uind, const double* uinp, double* field, double* fieldp) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!uind || !uinp) throw t9::ArgError("uind and uinp are required");
    t9::prepare(c);
    // use the CG scratch vectors: zrsd/zrsdp hold the inputs, vec/vecp the fields
    t9::scatter_from_host(c, uind, c.zrsd.p);
    t9::scatter_from_host(c, uinp, c.zrsdp.p);
    t9::ufield_all(c, c.zrsd.p, c.zrsdp.p, c.vec.p, c.vecp.p);
    if (field) t9::gather_to_host(c, c.vec.p, field);
    if (fieldp) t9::gather_to_host(c, c.vecp.p, fieldp);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_induce(t9_ctx* ctx, double* uind, double* uinp, t9_induce_stats* stats) {
  return guarded([&] {
    auto& c = ctx->c;
    t9::prepare(c);
    int iters = 0;
    double rms = 0;
    int rc = t9::induce_pcg(c, &iters, &rms);
    if (stats) {
      stats->iterations = iters;
      stats->predictor = c.last_predictor;
      stats->rms_debye = rms;
      stats->guess_rms = c.last_guess_rms;
    }
    if (uind) t9::gather_to_host(c, c.uind.p, uind);
    if (uinp) t9::gather_to_host(c, c.uinp.p, uinp);
    if (rc != 0) throw t9::NotConverged("INDUCE -- induced dipoles are not converged");
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_induce_trace(t9_ctx* ctx, double* eps, int cap, int* count) {
  return guarded([&] {
    auto& c = ctx->c;
    if (cap < 0 || (cap > 0 && !eps)) throw t9::ArgError("t9_induce_trace: bad buffer");
    const int m = std::min(cap, c.pcg_trace_len);
    if (m > 0) {
      T9_HIP(hipMemcpyAsync(eps, c.pcg_trace.p, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
      T9_HIP(hipStreamSynchronize(c.stream));
    }
    if (count) *count = c.pcg_trace_len;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_get_uind(t9_ctx* ctx, double* uind, double* uinp) {
  return guarded([&] {
    auto& c = ctx->c;
    if (uind) t9::gather_to_host(c, c.uind.p, uind);
    if (uinp) t9::gather_to_host(c, c.uinp.p, uinp);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_set_uind(t9_ctx* ctx, const double* uind, const double* uinp) {
  return guarded([&] {
    auto& c = ctx->c;
    t9::prepare(c);
    if (uind) t9::scatter_from_host(c, uind, c.uind.p);
    if (uinp) t9::scatter_from_host(c, uinp, c.uinp.p);
    return T9_OK;
  }, T9_ERR_HIP);
}

namespace {
int energy_common(t9_ctx* ctx, int vers, int terms, t9_energy_out* out, double* grad, bool grad_on_device) {
  auto& c = ctx->c;
  if (vers & ~(T9_CALC_ENERGY | T9_CALC_GRAD | T9_CALC_VIRIAL | T9_CALC_ANALYZE))
    throw t9::ArgError("unknown vers bits");
  if (terms & ~t9::TERM_ALL) throw t9::ArgError("unknown term bits");
  double e[t9::kEnergyRow];
  const int iters = t9::evaluate(c, vers, terms, e, (vers & T9_CALC_GRAD) ? grad : nullptr, grad_on_device);
  long long nem = 0, nep = 0, nev = 0, nct = 0, ndisp = 0, nrep = 0;
  if (vers & T9_CALC_ANALYZE) {
    if (terms & t9::TERM_ELEC) t9::count_interactions(c, &nem, &nep);
    if (terms & t9::TERM_VDW) nev = t9::vdw_count(c);
    if (terms & t9::TERM_CHGTRN) nct = t9::chgtrn_count(c);
    if (terms & t9::TERM_DISP) ndisp = t9::disp_count(c);
    if (terms & t9::TERM_REPEL) nrep = t9::repel_count(c);
  }
  if (out) {
    t9::fill_energy_out(e, vers, iters, out);
    out->nem = nem;
    out->nep = nep;
    out->nev = nev;
    out->nct = nct;
    out->ndisp = ndisp;
    out->nrep = nrep;
  }
  return T9_OK;
}

// Host grid [K3][K2][K1][2] FP64 <-> the context's complex grid.
void grid_from_host(t9::Ctx& c, const double* g) {
  const size_t ng = (size_t)c.grid[0] * c.grid[1] * c.grid[2];
  if (c.prec == 64) {
    T9_HIP(hipMemcpyAsync(c.qgrid.p, g, sizeof(double) * 2 * ng, hipMemcpyHostToDevice, c.stream));
  } else {
    std::vector<float> f(g, g + 2 * ng);
    T9_HIP(hipMemcpyAsync(c.qgrid.p, f.data(), sizeof(float) * 2 * ng, hipMemcpyHostToDevice, c.stream));
  }
  T9_HIP(hipStreamSynchronize(c.stream));
}

void grid_to_host(t9::Ctx& c, double* g) {
  const size_t ng = (size_t)c.grid[0] * c.grid[1] * c.grid[2];
  if (c.prec == 64) {
    T9_HIP(hipMemcpyAsync(g, c.qgrid.p, sizeof(double) * 2 * ng, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
  } else {
    std::vector<float> f(2 * ng);
    T9_HIP(hipMemcpyAsync(f.data(), c.qgrid.p, sizeof(float) * 2 * ng, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    std::copy(f.begin(), f.end(), g);
  }
}

void require_pme(const t9::Ctx& c) {
  if (!c.ewald) throw t9::ArgError("PME sub-steps need ewald");
}

// A PME sub-step on a new evaluation's spline bins: the coordinates may have
// changed since the last call.
void pme_prepare(t9::Ctx& c) {
  require_pme(c);
  c.mlist.cut_valid = c.ulist.cut_valid = false;
  c.bins_valid = false;
  c.rpole_valid = false;
  t9::prepare(c);
}
}  // namespace

int t9_energy(t9_ctx* ctx, int vers, t9_energy_out* out, double* grad) {
  return guarded([&] { return energy_common(ctx, vers, t9::TERM_ALL, out, grad, false); }, T9_ERR_HIP);
}

int t9_energy_terms(t9_ctx* ctx, int vers, int terms, t9_energy_out* out, double* grad) {
  return guarded([&] { return energy_common(ctx, vers, terms, out, grad, false); }, T9_ERR_HIP);
}

int t9_energy_device(t9_ctx* ctx, int vers, int terms, t9_energy_out* out, double* d_grad) {
  return guarded([&] { return energy_common(ctx, vers, terms, out, d_grad, true); }, T9_ERR_HIP);
}

int t9_set_coords_device(t9_ctx* ctx, const double* d_xyz) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!d_xyz) throw t9::ArgError("null coordinates");
    T9_HIP(hipMemcpyAsync(c.xyz.p, d_xyz, sizeof(double) * 3 * c.n, hipMemcpyDeviceToDevice, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    c.coords_dirty = true;
    c.rpole_valid = false;
    c.bins_valid = false;
    c.md_ready = false;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_get_coords_device(t9_ctx* ctx, double* d_xyz) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!d_xyz) throw t9::ArgError("null coordinates");
    T9_HIP(hipMemcpyAsync(d_xyz, c.xyz.p, sizeof(double) * 3 * c.n, hipMemcpyDeviceToDevice, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    return T9_OK;
  }, T9_ERR_HIP);
}

namespace {
int precond_host(t9_ctx* ctx, int mode, const double* rsd, const double* rsdp, double* zrsd, double* zrsdp) {
  auto& c = ctx->c;
  if (!rsd || !rsdp || !zrsd || !zrsdp) throw t9::ArgError("null vector");
  t9::prepare(c);
  // the CG scratch vectors hold the inputs (rsd, rsdp) and outputs (vec, vecp)
  t9::scatter_from_host(c, rsd, c.rsd.p);
  t9::scatter_from_host(c, rsdp, c.rsdp.p);
  t9::precond_apply(c, c.rsd.p, c.rsdp.p, c.vec.p, c.vecp.p, nullptr, nullptr, nullptr, mode);
  t9::gather_to_host(c, c.vec.p, zrsd);
  t9::gather_to_host(c, c.vecp.p, zrsdp);
  return T9_OK;
}
}  // namespace

int t9_diag_precond(t9_ctx* ctx, const double* rsd, const double* rsdp, double* zrsd, double* zrsdp) {
  return guarded([&] { return precond_host(ctx, 0, rsd, rsdp, zrsd, zrsdp); }, T9_ERR_HIP);
}

int t9_sparse_precond_apply(t9_ctx* ctx, const double* rsd, const double* rsdp, double* zrsd, double* zrsdp) {
  return guarded([&] { return precond_host(ctx, 1, rsd, rsdp, zrsd, zrsdp); }, T9_ERR_HIP);
}

int t9_pme_grid_mpole(t9_ctx* ctx, double* grid) {
  return guarded([&] {
    auto& c = ctx->c;
    pme_prepare(c);
    t9::pme_grid_mpole(c);
    if (grid) grid_to_host(c, grid);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_pme_grid_uind(t9_ctx* ctx, const double* uind, const double* uinp, double* grid) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!uind || !uinp) throw t9::ArgError("uind and uinp are required");
    pme_prepare(c);
    t9::scatter_from_host(c, uind, c.zrsd.p);
    t9::scatter_from_host(c, uinp, c.zrsdp.p);
    t9::pme_grid_uind(c, c.zrsd.p, c.zrsdp.p);
    if (grid) grid_to_host(c, grid);
    return T9_OK;
  }, T9_ERR_HIP);
}

namespace {
int pme_grid_op(t9_ctx* ctx, double* grid, int op) {
  auto& c = ctx->c;
  require_pme(c);
  if (grid) grid_from_host(c, grid);
  if (op == 0) t9::pme_conv_grid(c);
  else t9::pme_fft_dir(c, op);
  if (grid) grid_to_host(c, grid);
  else T9_HIP(hipStreamSynchronize(c.stream));
  return T9_OK;
}
}  // namespace

int t9_pme_fft_front(t9_ctx* ctx, double* grid) {
  return guarded([&] { return pme_grid_op(ctx, grid, -1); }, T9_ERR_HIP);
}

int t9_pme_fft_back(t9_ctx* ctx, double* grid) {
  return guarded([&] { return pme_grid_op(ctx, grid, +1); }, T9_ERR_HIP);
}

int t9_pme_conv(t9_ctx* ctx, double* grid) {
  return guarded([&] { return pme_grid_op(ctx, grid, 0); }, T9_ERR_HIP);
}

namespace {
int pme_fphi_host(t9_ctx* ctx, bool mpole, const double* grid, double* fd, double* fp) {
  auto& c = ctx->c;
  require_pme(c);
  t9::prepare(c);
  if (grid) grid_from_host(c, grid);
  t9::pme_fphi(c, mpole, c.vec.p, c.vecp.p);
  if (fd) t9::gather_to_host(c, c.vec.p, fd);
  if (fp) t9::gather_to_host(c, c.vecp.p, fp);
  if (!fd && !fp) T9_HIP(hipStreamSynchronize(c.stream));
  return T9_OK;
}
}  // namespace

int t9_pme_fphi_mpole(t9_ctx* ctx, const double* grid, double* field) {
  return guarded([&] { return pme_fphi_host(ctx, true, grid, field, nullptr); }, T9_ERR_HIP);
}

int t9_pme_fphi_uind(t9_ctx* ctx, const double* grid, double* fd, double* fp) {
  return guarded([&] { return pme_fphi_host(ctx, false, grid, fd, fp); }, T9_ERR_HIP);
}

int t9_fft3d(void* grid, int k1, int k2, int k3, int precision, int sign) {
  return guarded([&] {
    if (!grid) throw t9::ArgError("null grid");
    if (precision != T9_FP32 && precision != T9_FP64) throw t9::ArgError("precision must be T9_FP32 or T9_FP64");
    if (sign != -1 && sign != 1) throw t9::ArgError("sign must be -1 (forward) or +1 (inverse)");
    const int K[3] = {k1, k2, k3};
    for (int d = 0; d < 3; ++d) t9::fft_check_size(K[d], precision);
    t9::DBuf dtw;  // the twiddle table of this call
    dtw.alloc((precision == T9_FP64 ? 16 : 8) * (size_t)(k1 + k2 + k3));
    t9::fft_twiddles_device(dtw.p, precision, K, nullptr);
    t9::fft3d_device(grid, precision, K, sign, dtw.p, nullptr);
    T9_HIP(hipStreamSynchronize(nullptr));
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_fft3d_noalloc(void* grid, int k1, int k2, int k3, int precision, int sign) {
  return guarded([&] {
    if (!grid) throw t9::ArgError("null grid");
    if (precision != T9_FP32 && precision != T9_FP64) throw t9::ArgError("precision must be T9_FP32 or T9_FP64");
    if (sign != -1 && sign != 1) throw t9::ArgError("sign must be -1 (forward) or +1 (inverse)");
    const int K[3] = {k1, k2, k3};
    for (int d = 0; d < 3; ++d) t9::fft_check_size(K[d], precision);
    t9::fft3d_noalloc(grid, precision, K, sign);
    T9_HIP(hipStreamSynchronize(nullptr));
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_md_init(t9_ctx* ctx, double kelvin, uint64_t seed) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!(kelvin >= 0)) throw t9::ArgError("temperature must be >= 0");
    t9::md_velocities(c, kelvin, seed);
    c.md_rng.seed(seed ^ 0x9e3779b97f4a7c15ULL);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_md_set_velocities(t9_ctx* ctx, const double* v) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!v) throw t9::ArgError("null velocities");
    T9_HIP(hipMemcpyAsync(c.vel.p, v, sizeof(double) * 3 * c.n, hipMemcpyHostToDevice, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_md_get_velocities(t9_ctx* ctx, double* v) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!v) throw t9::ArgError("null velocities");
    T9_HIP(hipMemcpyAsync(v, c.vel.p, sizeof(double) * 3 * c.n, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_md_run(t9_ctx* ctx, int nsteps, double dt, int nrespa, double bath_kelvin, double tau, t9_md_stats* stats) {
  return guarded([&] {
    auto& c = ctx->c;
    if (nsteps < 0 || !(dt > 0) || nrespa < 1) throw t9::ArgError("bad nsteps, dt or nrespa");
    if (bath_kelvin > 0 && !(tau > 0)) throw t9::ArgError("the thermostat needs tau > 0");
    double e[t9::kEnergyRow] = {0};
    long long iters = 0;
    const int builds0 = c.builds;
    if (!c.md_ready) {
      iters += t9::evaluate(c, T9_CALC_ENERGY, t9::TERM_ALL, e, nullptr);
      t9::md_sorted_to_orig(c, c.egrad.as<double>(), c.g_slow.as<double>());
      c.md_ready = true;
    }
    const double dta = dt / nrespa;
    double* gs = c.g_slow.as<double>();
    double* gf = c.g_fast.as<double>();
    for (int step = 0; step < nsteps; ++step) {
      // velR1: half kick of the slow forces and the first inner half kick
      t9::md_kick(c, 0.5 * dta, gf, 0.5 * dt, gs);
      // inner steps: drift, then per step the valence gradient, the inner
      // kick (velR0) and the next drift in one pass over the atoms
      if (t9::respa_fused(c)) {
        t9::respa_inner(c, gf, dta, nrespa - 1);  // one launch, one thread per molecule
      } else {
        t9::md_drift(c, dta);
        for (int k = 1; k < nrespa; ++k) {
          if (c.use_valence) {
            t9::valence_inner(c, gf, dta);
          } else {
            t9::md_drift(c, dta);
          }
        }
      }
      // energies come back to the host only for the last step's statistics
      iters += t9::evaluate(c, step == nsteps - 1 ? T9_CALC_ENERGY : 0, t9::TERM_ALL, e, nullptr);
      t9::md_sorted_to_orig(c, c.egrad.as<double>(), gs);
      t9::md_kick(c, 0.5 * dta, gf, 0.5 * dt, gs);  // velR2
      if (bath_kelvin > 0) {
        // Bussi stochastic velocity rescaling over the full step; the two
        // random numbers come from the host stream, the kinetic energy and
        // the factor stay on the device
        std::normal_distribution<double> nd(0.0, 1.0);
        const double r = nd(c.md_rng);
        std::gamma_distribution<double> gd(0.5 * (c.nfree - 1), 2.0);
        const double s2 = c.nfree > 1 ? gd(c.md_rng) : 0.0;
        t9::md_bussi(c, dt, tau, bath_kelvin, r, s2);
      }
    }
    if (stats) {
      t9_energy_out eout;
      t9::fill_energy_out(e, T9_CALC_ENERGY, 0, &eout);
      stats->epot = eout.esum;
      stats->ekin = t9::md_kinetic(c);
      stats->temperature = 2.0 * stats->ekin / (c.nfree * t9::kGasConst);
      stats->induce_iterations = iters;
      stats->list_builds = c.builds - builds0;
      stats->steps = nsteps;
    }
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_md_bussi(t9_ctx* ctx, double dt, double tau, double kelvin, double r, double s2) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!(dt > 0) || !(tau > 0) || !(kelvin > 0)) throw t9::ArgError("bad Bussi arguments");
    if (c.nfree < 2) throw t9::ArgError("the thermostat needs md_init (degrees of freedom)");
    t9::md_bussi(c, dt, tau, kelvin, r, s2);
    T9_HIP(hipStreamSynchronize(c.stream));
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_nparam(void) { return t9::guess_nparam(); }

int t9_guess_set_params(t9_ctx* ctx, const float* params, const float* norm, uint64_t seed) {
  return guarded([&] {
    t9::guess_set_params(ctx->c, params, norm, seed);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_get_params(t9_ctx* ctx, float* params, float* norm) {
  return guarded([&] {
    t9::guess_get_params(ctx->c, params, norm);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_enable(t9_ctx* ctx, int on) {
  return guarded([&] {
    auto& c = ctx->c;
    if (on && !c.guess_ready) throw t9::ArgError("the learned guess has no trained parameters");
    if (on && c.opt.predictor != "ASPC") throw t9::ArgError("the learned guess needs polar-predict ASPC");
    c.guess_on = on != 0;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_capture(t9_ctx* ctx, int stride, int64_t max_samples) {
  return guarded([&] {
    if (ctx->c.opt.predictor != "ASPC") throw t9::ArgError("capture needs polar-predict ASPC");
    t9::guess_capture_start(ctx->c, stride, max_samples);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_stats(t9_ctx* ctx, int64_t* out4) {
  return guarded([&] {
    auto& c = ctx->c;
    out4[0] = c.cap_n;
    out4[1] = c.guess_used;
    out4[2] = c.guess_fallback;
    out4[3] = c.last_predictor;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_train(t9_ctx* ctx, int epochs, int batch, double lr, uint64_t seed, double* loss_first,
                   double* loss_last) {
  return guarded([&] {
    t9::guess_train(ctx->c, epochs, batch, lr, seed, loss_first, loss_last);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_mlp_forward(t9_ctx* ctx, const float* X, int m, float* Y) {
  return guarded([&] {
    auto& c = ctx->c;
    if (m < 1 || !X || !Y) throw t9::ArgError("bad forward arguments");
    if (!c.guess_params) throw t9::ArgError("the learned guess has no parameters");
    t9::DBuf xf, xb, y;
    xf.alloc(sizeof(float) * 32 * (size_t)m);
    xb.alloc(sizeof(uint16_t) * 32 * (size_t)m);
    y.alloc(sizeof(float) * 8 * (size_t)m);
    T9_HIP(hipMemcpyAsync(xf.p, X, sizeof(float) * 32 * (size_t)m, hipMemcpyHostToDevice, c.stream));
    t9::guess_to_bf16(c, 32LL * m, xf.as<float>(), xb.p);
    t9::mlp_forward(c, m, xb.p, y.as<float>());
    std::vector<float> h(8 * (size_t)m);
    T9_HIP(hipMemcpyAsync(h.data(), y.p, sizeof(float) * h.size(), hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    for (int r = 0; r < m; ++r)
      for (int q = 0; q < 6; ++q) Y[6 * (size_t)r + q] = h[8 * (size_t)r + q];
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_mlp_grad(t9_ctx* ctx, const float* X, const float* T, int m, float* grad, double* loss) {
  return guarded([&] {
    auto& c = ctx->c;
    if (m < 1 || !X || !T || !grad) throw t9::ArgError("bad gradient arguments");
    if (!c.guess_params) throw t9::ArgError("the learned guess has no parameters");
    const int np = t9::guess_nparam();
    t9::DBuf xf, xb, tt, g;
    xf.alloc(sizeof(float) * 32 * (size_t)m);
    xb.alloc(sizeof(uint16_t) * 32 * (size_t)m);
    tt.alloc(sizeof(float) * 6 * (size_t)m);
    g.alloc(sizeof(float) * (np + 1));
    T9_HIP(hipMemcpyAsync(xf.p, X, sizeof(float) * 32 * (size_t)m, hipMemcpyHostToDevice, c.stream));
    T9_HIP(hipMemcpyAsync(tt.p, T, sizeof(float) * 6 * (size_t)m, hipMemcpyHostToDevice, c.stream));
    t9::guess_to_bf16(c, 32LL * m, xf.as<float>(), xb.p);
    t9::mlp_grad(c, m, xb.p, tt.as<float>(), g.as<float>());
    std::vector<float> h(np + 1);
    T9_HIP(hipMemcpyAsync(h.data(), g.p, sizeof(float) * h.size(), hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    std::copy(h.begin(), h.begin() + np, grad);
    if (loss) *loss = h[np] / (6.0 * m);
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_adam(t9_ctx* ctx, const float* grad, double lr) {
  return guarded([&] {
    auto& c = ctx->c;
    if (!c.guess_params) throw t9::ArgError("the learned guess has no parameters");
    const int np = t9::guess_nparam();
    t9::DBuf g;
    g.alloc(sizeof(float) * np);
    T9_HIP(hipMemcpyAsync(g.p, grad, sizeof(float) * np, hipMemcpyHostToDevice, c.stream));
    t9::adam_step(c, g.as<float>(), lr);
    T9_HIP(hipStreamSynchronize(c.stream));
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_features(t9_ctx* ctx, float* F) {
  return guarded([&] {
    auto& c = ctx->c;
    if (c.opt.predictor != "ASPC") throw t9::ArgError("features need polar-predict ASPC");
    if (c.nualt < 16 || c.nealt < 16) throw t9::ArgError("features need 16 stored solves");
    c.mlist.cut_valid = c.ulist.cut_valid = false;
    c.bins_valid = false;
    c.rpole_valid = false;
    t9::prepare(c);
    t9::dfield_all(c);
    const auto B = t9::aspc_coefficients();
    double ce[16], cu[2];
    for (int k = 0; k < 16; ++k) ce[k] = B[((c.nealt - 1 - k) % 16 + 16) % 16];
    cu[0] = ((c.nualt - 1) % 16 + 16) % 16;
    cu[1] = ((c.nualt - 2) % 16 + 16) % 16;
    t9::guess_features(c, ce, cu);
    std::vector<float> h(32 * (size_t)c.n);
    std::vector<int> perm(c.n);
    T9_HIP(hipMemcpyAsync(h.data(), c.gfeat.p, sizeof(float) * h.size(), hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipMemcpyAsync(perm.data(), c.perm.p, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
    T9_HIP(hipStreamSynchronize(c.stream));
    for (int s = 0; s < c.n; ++s)
      for (int q = 0; q < 32; ++q) F[32 * (size_t)perm[s] + q] = h[32 * (size_t)s + q];
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_guess_history(t9_ctx* ctx, int kind, int age, double* out) {
  return guarded([&] {
    auto& c = ctx->c;
    if (kind < 0 || kind > 3 || age < 0 || age > 15) throw t9::ArgError("bad history request");
    const int count = kind < 2 ? c.nealt : c.nualt;
    if (age >= std::min(count, 16)) throw t9::ArgError("history entry not stored yet");
    const int slot = ((count - 1 - age) % 16 + 16) % 16;
    const size_t n3 = 3 * (size_t)c.n;
    const t9::DBuf& b = kind < 2 ? c.ealt : c.ualt;
    const size_t off = (size_t)slot * 2 * n3 + ((kind & 1) ? n3 : 0);
    if (c.prec == 32) {
      std::vector<float> h(n3);
      T9_HIP(hipMemcpyAsync(h.data(), b.as<float>() + off, sizeof(float) * n3, hipMemcpyDeviceToHost, c.stream));
      T9_HIP(hipStreamSynchronize(c.stream));
      for (size_t q = 0; q < n3; ++q) out[q] = h[q];
    } else {
      T9_HIP(hipMemcpyAsync(out, b.as<double>() + off, sizeof(double) * n3, hipMemcpyDeviceToHost, c.stream));
      T9_HIP(hipStreamSynchronize(c.stream));
    }
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_sync(t9_ctx* ctx) {
  return guarded([&] {
    T9_HIP(hipStreamSynchronize(ctx->c.stream));
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_kernel_timing(t9_ctx* ctx, int kernel, int mode, double* total_ms, int64_t* launches) {
  return guarded([&] {
    auto& c = ctx->c;
    if (kernel < 0 || kernel >= t9::KT_COUNT) throw t9::ArgError("unknown kernel id");
    auto& t = c.timers[kernel];
    if (mode == 1) {  // start: clear and enable
      t9::ktimer_collect(c, kernel);
      t.ms = 0;
      t.count = 0;
      t.on = true;
    } else if (mode == 0) {  // stop: disable, keep totals
      t9::ktimer_collect(c, kernel);
      t.on = false;
    } else {
      t9::ktimer_collect(c, kernel);
    }
    if (total_ms) *total_ms = t.ms;
    if (launches) *launches = t.count;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_cutoff_pairs(t9_ctx* ctx, int64_t* pairs) {
  return guarded([&] {
    auto& c = ctx->c;
    t9::prepare(c);
    *pairs = t9::count_cutoff_pairs(c) / 2;
    return T9_OK;
  }, T9_ERR_HIP);
}

int t9_copy_probe(int64_t nbytes, int reps, double* ms_per_copy) {
  return guarded([&] {
    if (nbytes < 16 || reps < 1) throw t9::ArgError("bad copy probe size");
    t9::DBuf a, b;
    a.alloc((size_t)nbytes);
    b.alloc((size_t)nbytes);
    T9_HIP(hipMemset(a.p, 0, (size_t)nbytes));
    hipEvent_t e0, e1;
    T9_HIP(hipEventCreate(&e0));
    T9_HIP(hipEventCreate(&e1));
    const long long n4 = nbytes / 16;
    hipLaunchKernelGGL(t9::k_copy_probe, dim3(2048), dim3(256), 0, nullptr, n4, a.as<float4>(), b.as<float4>());
    T9_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(t9::k_copy_probe, dim3(2048), dim3(256), 0, nullptr, n4, a.as<float4>(), b.as<float4>());
    T9_HIP(hipEventRecord(e1, nullptr));
    T9_HIP(hipEventSynchronize(e1));
    float ms = 0;
    T9_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
    return T9_OK;
  }, T9_ERR_HIP);
}

}  // extern "C"
