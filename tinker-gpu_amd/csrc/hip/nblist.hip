// Spatial neighbor list for gfx950.
//
// Replaces the reference's 32-atom Hilbert block list (src/cu/spatial.cu:101-960,
// include/ff/spatial.h) with a wave64 design:
//  1. atoms are binned into cells of edge >= (cutoff + buffer)/2 and sorted
//     by (Morton code of the cell, original index) — a counting sort whose
//     within-cell order is fixed by a per-cell insertion sort, so the order
//     is deterministic;
//  2. every atom gets its own full neighbor list (both directions) of atoms
//     within cutoff + buffer, found in the 5x5x5 surrounding cells;
//  3. lists are stored per wavefront of 64 consecutive sorted atoms,
//     k-major, padded to the longest list of the wave.
// Each entry packs the neighbor's sorted index and its scale class (the
// reference's exclusion bit masks, spatial.h si1..si4).  Pair kernels walk
// their own atom's list and write only their own atom's sums: no atomics,
// fixed summation order, bit-reproducible results.
#include "ctx.hpp"

namespace t9 {
namespace {

constexpr int kScanBlock = 1024;

__device__ __forceinline__ uint32_t spread_bits(uint32_t v) {
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

__device__ __forceinline__ int morton3(int x, int y, int z) {
  return (int)(spread_bits(x) | (spread_bits(y) << 1) | (spread_bits(z) << 2));
}

__global__ void k_cell_assign(int n, const double* __restrict__ xyz, double lx, double ly, double lz, int ncx,
                              int ncy, int ncz, int brute, int* __restrict__ key, int* __restrict__ count,
                              int* __restrict__ cellc_orig) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cx = 0, cy = 0, cz = 0, k = 0;
  if (!brute) {
    double fx = xyz[3 * i] / lx, fy = xyz[3 * i + 1] / ly, fz = xyz[3 * i + 2] / lz;
    fx -= floor(fx);
    fy -= floor(fy);
    fz -= floor(fz);
    cx = min(ncx - 1, (int)(fx * ncx));
    cy = min(ncy - 1, (int)(fy * ncy));
    cz = min(ncz - 1, (int)(fz * ncz));
    k = morton3(cx, cy, cz);
  }
  key[i] = k;
  cellc_orig[i] = cx | (cy << 10) | (cz << 20);
  atomicAdd(&count[k], 1);
}

// Single-block exclusive scan of m ints (chunks of 1024, carried offset).
__global__ __launch_bounds__(kScanBlock) void k_scan_int(const int* __restrict__ in, int* __restrict__ out, int m) {
  __shared__ int sh[kScanBlock];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < m; base += kScanBlock) {
    int idx = base + threadIdx.x;
    int v = idx < m ? in[idx] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < kScanBlock; s <<= 1) {
      int t = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (idx < m) out[idx] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) carry += sh[kScanBlock - 1];
    __syncthreads();
  }
}

// Single-block exclusive scan of m int64 values; out has m + 1 entries.
__global__ __launch_bounds__(kScanBlock) void k_scan_i64(const long long* __restrict__ in, long long* __restrict__ out,
                                                         int m) {
  __shared__ long long sh[kScanBlock];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < m; base += kScanBlock) {
    int idx = base + threadIdx.x;
    long long v = idx < m ? in[idx] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < kScanBlock; s <<= 1) {
      long long t = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (idx < m) out[idx] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) carry += sh[kScanBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = carry;
}

__global__ void k_cell_scatter(int n, const int* __restrict__ key, const int* __restrict__ start,
                               int* __restrict__ cursor, int* __restrict__ unsorted) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = key[i];
  unsorted[start[k] + atomicAdd(&cursor[k], 1)] = i;
}

// Order each cell's atoms by original index (insertion sort per cell).
__global__ void k_cell_sort(int nkeys, const int* __restrict__ start, const int* __restrict__ count,
                            const int* __restrict__ unsorted, int* __restrict__ perm) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  int s0 = start[k], c = count[k];
  for (int a = 0; a < c; ++a) perm[s0 + a] = unsorted[s0 + a];
  for (int a = 1; a < c; ++a) {
    int v = perm[s0 + a];
    int b = a - 1;
    while (b >= 0 && perm[s0 + b] > v) {
      perm[s0 + b + 1] = perm[s0 + b];
      --b;
    }
    perm[s0 + b + 1] = v;
  }
}

__global__ void k_identity(int n, int* __restrict__ perm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) perm[i] = i;
}

template <class real>
__global__ void k_build_sorted(int n, const int* __restrict__ perm, int* __restrict__ iperm,
                               const double* __restrict__ xyz, double* __restrict__ xyz_last, double3 vorg,
                               double3 vlen, int periodic, PosRec<real>* __restrict__ pos,
                               const int* __restrict__ cellc_orig, int* __restrict__ cellc,
                               const double* __restrict__ polarity_o, const double* __restrict__ pdamp_o,
                               const int* __restrict__ jpolar_o, real* __restrict__ polarity,
                               real* __restrict__ polinv, real* __restrict__ pdamp, int* __restrict__ jpolar) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  iperm[i] = s;
  double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  xyz_last[3 * i] = x;
  xyz_last[3 * i + 1] = y;
  xyz_last[3 * i + 2] = z;
  const double org[3] = {vorg.x, vorg.y, vorg.z}, len[3] = {vlen.x, vlen.y, vlen.z};
  pos[s] = make_pos<real>(x, y, z, pdamp_o[i] > 0 ? 1.0 / pdamp_o[i] : 0.0, org, len, periodic);
  cellc[s] = cellc_orig[i];
  double a = polarity_o[i];
  polarity[s] = (real)a;
  polinv[s] = (real)(1.0 / (a > 1e-16 ? a : 1e-16));
  pdamp[s] = (real)pdamp_o[i];
  jpolar[s] = jpolar_o[i];
}

template <class real>
__global__ void k_refresh_pos(int n, const int* __restrict__ perm, const double* __restrict__ xyz, double3 vorg,
                              double3 vlen, int periodic, const double* __restrict__ pdamp_o,
                              PosRec<real>* __restrict__ pos) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  const double org[3] = {vorg.x, vorg.y, vorg.z}, len[3] = {vlen.x, vlen.y, vlen.z};
  pos[s] = make_pos<real>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], pdamp_o[i] > 0 ? 1.0 / pdamp_o[i] : 0.0, org,
                          len, periodic);
}

// Count (fill == 0) or write (fill == 1) each sorted atom's neighbors.
template <class real>
__global__ void k_nb_list(int n, int fill, const PosRec<real>* __restrict__ pos,
                          const int* __restrict__ cellc, int ncx, int ncy, int ncz, int brute,
                          const int* __restrict__ cell_start, const int* __restrict__ cell_count, BoxT<real> box,
                          real rl2, const int* __restrict__ perm, const int* __restrict__ excl_off,
                          const int* __restrict__ excl_j, const int* __restrict__ excl_code, int* __restrict__ nlist,
                          const long long* __restrict__ woff, uint32_t* __restrict__ entries) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const auto pi = pos[s];
  const int oi = perm[s];
  const int e0 = excl_off[oi], e1 = excl_off[oi + 1];
  int cnt = 0;
  uint32_t* out = nullptr;
  int limit = 0x7fffffff;
  if (fill) {
    out = entries + woff[s >> 6] + (s & 63);
    limit = nlist[s];  // the count pass sized this atom's slots
  }
  auto test = [&](int j) {
    if (j == s) return;
    const auto pj = pos[j];
    real xr, yr, zr;
    pair_disp(pi, pj, box, xr, yr, zr);
    real r2 = xr * xr + yr * yr + zr * zr;
    if (r2 < rl2 && cnt < limit) {
      if (fill) {
        int code = 0;
        const int oj = perm[j];
        for (int e = e0; e < e1; ++e)
          if (excl_j[e] == oj) code = excl_code[e];
        out[(long long)cnt * 64] = ((uint32_t)j << kCodeBits) | (uint32_t)code;
      }
      ++cnt;
    }
  };
  if (brute) {
    for (int j = 0; j < n; ++j) test(j);
  } else {
    const int c = cellc[s];
    const int cx = c & 1023, cy = (c >> 10) & 1023, cz = (c >> 20) & 1023;
    for (int dz = -2; dz <= 2; ++dz) {
      int z = cz + dz;
      z += z < 0 ? ncz : 0;
      z -= z >= ncz ? ncz : 0;
      for (int dy = -2; dy <= 2; ++dy) {
        int y = cy + dy;
        y += y < 0 ? ncy : 0;
        y -= y >= ncy ? ncy : 0;
        for (int dx = -2; dx <= 2; ++dx) {
          int x = cx + dx;
          x += x < 0 ? ncx : 0;
          x -= x >= ncx ? ncx : 0;
          int k = morton3(x, y, z);
          int j0 = cell_start[k], j1 = j0 + cell_count[k];
          for (int j = j0; j < j1; ++j) test(j);
        }
      }
    }
  }
  if (!fill) nlist[s] = cnt;
}

__global__ void k_wave_slots(int n, int nwave, const int* __restrict__ nlist, int* __restrict__ maxlen,
                             long long* __restrict__ slots) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwave) return;
  int m = 0;
  for (int l = 0; l < 64; ++l) {
    int s = w * 64 + l;
    if (s < n) m = max(m, nlist[s]);
  }
  maxlen[w] = m;
  slots[w] = 64LL * m;
}

template <class real>
__global__ void k_max_disp(int n, const double* __restrict__ xyz, const double* __restrict__ last, double lx,
                           double ly, double lz, int periodic, double lim2, int* __restrict__ flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double dx = xyz[3 * i] - last[3 * i], dy = xyz[3 * i + 1] - last[3 * i + 1], dz = xyz[3 * i + 2] - last[3 * i + 2];
  if (periodic) {
    dx -= lx * rint(dx / lx);
    dy -= ly * rint(dy / ly);
    dz -= lz * rint(dz / lz);
  }
  if (dx * dx + dy * dy + dz * dz > lim2) atomicOr(flag, 1);
}


template <class real>
void build_list(Ctx& c, NbList& L, const void* posv, const int* eoff, const int* ej, const int* ecode) {
  const int n = c.n, nw = c.nwave;
  const int nb = ceil_div(n, 256);
  L.nlist.alloc(sizeof(int) * n);
  L.maxlen.alloc(sizeof(int) * nw);
  L.off.alloc(sizeof(long long) * (nw + 1));
  DBuf slots;
  slots.alloc(sizeof(long long) * nw);
  BoxT<real> box = make_boxT<real>(c);
  real rl2 = (real)(L.cutoff * L.cutoff);
  const int nkeys = 1 << (3 * c.morton_bits);
  hipLaunchKernelGGL(k_nb_list<real>, dim3(nb), dim3(256), 0, c.stream, n, 0, (const PosRec<real>*)posv,
                     c.cellc.as<int>(), c.ncell[0], c.ncell[1], c.ncell[2], c.brute ? 1 : 0, c.cell_start.as<int>(),
                     c.cell_count.as<int>(), box, rl2, c.perm.as<int>(), eoff, ej, ecode, L.nlist.as<int>(),
                     (const long long*)nullptr, (uint32_t*)nullptr);
  T9_LAUNCH_CHECK();
  (void)nkeys;
  hipLaunchKernelGGL(k_wave_slots, dim3(ceil_div(nw, 256)), dim3(256), 0, c.stream, n, nw, L.nlist.as<int>(),
                     L.maxlen.as<int>(), slots.as<long long>());
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(kScanBlock), 0, c.stream, slots.as<long long>(),
                     L.off.as<long long>(), nw);
  T9_LAUNCH_CHECK();
  long long total = 0;
  T9_HIP(hipMemcpyAsync(&total, L.off.as<long long>() + nw, sizeof(long long), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  if (total < 0 || total > (1LL << 31)) throw ListOverflow("neighbor list exceeds 2^31 slots");
  if ((size_t)total > L.capacity || !L.entries.p) {
    size_t cap = (size_t)(total * 1.2) + 64 * 64;
    L.entries.alloc(sizeof(uint32_t) * cap);
    L.capacity = cap;
  }
  L.total = total;
  hipLaunchKernelGGL(k_nb_list<real>, dim3(nb), dim3(256), 0, c.stream, n, 1, (const PosRec<real>*)posv,
                     c.cellc.as<int>(), c.ncell[0], c.ncell[1], c.ncell[2], c.brute ? 1 : 0, c.cell_start.as<int>(),
                     c.cell_count.as<int>(), box, rl2, c.perm.as<int>(), eoff, ej, ecode, L.nlist.as<int>(),
                     L.off.as<long long>(), L.entries.as<uint32_t>());
  T9_LAUNCH_CHECK();
}

template <class real>
void nblist_build_t(Ctx& c) {
  const int n = c.n;
  const int nb = ceil_div(n, 256);
  const int nkeys = 1 << (3 * c.morton_bits);
  c.cell_key.alloc(sizeof(int) * n);
  c.cell_count.alloc(sizeof(int) * nkeys);
  c.cell_start.alloc(sizeof(int) * nkeys);
  c.cell_cursor.alloc(sizeof(int) * nkeys);
  c.unsorted.alloc(sizeof(int) * n);
  DBuf cellc_orig;
  cellc_orig.alloc(sizeof(int) * n);
  T9_HIP(hipMemsetAsync(c.cell_count.p, 0, sizeof(int) * nkeys, c.stream));
  T9_HIP(hipMemsetAsync(c.cell_cursor.p, 0, sizeof(int) * nkeys, c.stream));
  const double lx = c.periodic ? c.box[0] : 1.0, ly = c.periodic ? c.box[1] : 1.0, lz = c.periodic ? c.box[2] : 1.0;
  hipLaunchKernelGGL(k_cell_assign, dim3(nb), dim3(256), 0, c.stream, n, c.xyz.as<double>(), lx, ly, lz, c.ncell[0],
                     c.ncell[1], c.ncell[2], c.brute ? 1 : 0, c.cell_key.as<int>(), c.cell_count.as<int>(),
                     cellc_orig.as<int>());
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_int, dim3(1), dim3(kScanBlock), 0, c.stream, c.cell_count.as<int>(),
                     c.cell_start.as<int>(), nkeys);
  T9_LAUNCH_CHECK();
  if (c.brute) {
    hipLaunchKernelGGL(k_identity, dim3(nb), dim3(256), 0, c.stream, n, c.perm.as<int>());
    T9_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(256), 0, c.stream, n, c.cell_key.as<int>(),
                       c.cell_start.as<int>(), c.cell_cursor.as<int>(), c.unsorted.as<int>());
    T9_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cell_sort, dim3(ceil_div(nkeys, 256)), dim3(256), 0, c.stream, nkeys,
                       c.cell_start.as<int>(), c.cell_count.as<int>(), c.unsorted.as<int>(), c.perm.as<int>());
    T9_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_build_sorted<real>, dim3(nb), dim3(256), 0, c.stream, n, c.perm.as<int>(), c.iperm.as<int>(),
                     c.xyz.as<double>(), c.xyz_last.as<double>(), make_double3(c.vorg[0], c.vorg[1], c.vorg[2]),
                     make_double3(c.vlen[0], c.vlen[1], c.vlen[2]), c.periodic ? 1 : 0,
                     c.pos.as<PosRec<real>>(), cellc_orig.as<int>(), c.cellc.as<int>(),
                     c.polarity_o.as<double>(), c.pdamp_o.as<double>(), c.jpolar_o.as<int>(), c.polarity.as<real>(),
                     c.polinv.as<real>(), c.pdamp.as<real>(), c.jpolar.as<int>());
  T9_LAUNCH_CHECK();
  build_list<real>(c, c.mlist, c.pos.p, c.excl_off.as<int>(), c.excl_j.as<int>(), c.excl_code.as<int>());
  if (c.ulist.cutoff > 0)
    build_list<real>(c, c.ulist, c.pos.p, c.excl_off.as<int>(), c.excl_j.as<int>(), c.excl_code.as<int>());
  if (c.vlist.cutoff > 0) {
    vdw_sites(c);
    build_list<real>(c, c.vlist, c.vpos.p, c.vexcl_off.as<int>(), c.vexcl_j.as<int>(), c.vexcl_code.as<int>());
  }
  c.builds++;
  c.list_valid = true;
  c.coords_dirty = false;
  c.rpole_valid = false;
  c.bins_valid = false;
  c.mlist.cut_valid = c.ulist.cut_valid = false;
}

// Keep the entries within the interaction cutoff (same FP arithmetic as the
// pair kernels, so those kernels need no distance test of their own).
template <class real>
__global__ void k_compact(int n, const int* __restrict__ nlist, const long long* __restrict__ woff,
                          const uint32_t* __restrict__ entries, const PosRec<real>* __restrict__ pos,
                          BoxT<real> box, real cut2, int* __restrict__ ncut, uint32_t* __restrict__ ecut) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const long long base = woff[s >> 6] + (s & 63);
  const uint32_t* e = entries + base;
  uint32_t* o = ecut + base;
  const auto pi = pos[s];
  int cnt = 0;
  const int len = nlist[s];
  for (int k = 0; k < len; ++k) {
    const uint32_t ent = e[(long long)k * 64];
    const auto pj = pos[ent >> kCodeBits];
    real xr, yr, zr;
    pair_disp(pi, pj, box, xr, yr, zr);
    if (xr * xr + yr * yr + zr * zr <= cut2) o[(long long)(cnt++) * 64] = ent;
  }
  ncut[s] = cnt;
}

template <class real>
void compact_t(Ctx& c) {
  BoxT<real> box = make_boxT<real>(c);
  for (NbList* L : {&c.mlist, &c.ulist}) {
    if (L->cutoff <= 0 || !L->entries.p) continue;
    L->ncut.alloc(sizeof(int) * c.n);
    L->ecut.alloc(sizeof(uint32_t) * L->capacity);
    const double c2 = L->cut * L->cut;
    hipLaunchKernelGGL(k_compact<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, L->nlist.as<int>(),
                       L->off.as<long long>(), L->entries.as<uint32_t>(),
                       c.pos.as<PosRec<real>>(), box, (real)(c2 > 1e30 ? 1e30 : c2),
                       L->ncut.as<int>(), L->ecut.as<uint32_t>());
    T9_LAUNCH_CHECK();
    L->cut_valid = true;
  }
}

template <class real>
bool needs_rebuild_t(Ctx& c) {
  if (!c.list_valid) return true;
  DBuf flag;
  flag.alloc(sizeof(int));
  T9_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), c.stream));
  double half = 0.5 * c.opt.list_buffer;
  hipLaunchKernelGGL(k_max_disp<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.xyz.as<double>(),
                     c.xyz_last.as<double>(), c.box[0], c.box[1], c.box[2], c.periodic ? 1 : 0, half * half,
                     flag.as<int>());
  T9_LAUNCH_CHECK();
  int h = 0;
  T9_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  T9_HIP(hipStreamSynchronize(c.stream));
  return h != 0;
}

template <class real>
void refresh_t(Ctx& c) {
  hipLaunchKernelGGL(k_refresh_pos<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.as<int>(),
                     c.xyz.as<double>(), make_double3(c.vorg[0], c.vorg[1], c.vorg[2]),
                     make_double3(c.vlen[0], c.vlen[1], c.vlen[2]), c.periodic ? 1 : 0, c.pdamp_o.as<double>(),
                     c.pos.as<PosRec<real>>());
  T9_LAUNCH_CHECK();
  if (c.vlist.cutoff > 0) vdw_sites(c);
  c.coords_dirty = false;
  c.rpole_valid = false;
  c.bins_valid = false;
  c.mlist.cut_valid = c.ulist.cut_valid = false;
}

}  // namespace

void nblist_build(Ctx& c) {
  if (c.prec == 32) nblist_build_t<float>(c);
  else nblist_build_t<double>(c);
}

bool nblist_needs_rebuild(Ctx& c) { return c.prec == 32 ? needs_rebuild_t<float>(c) : needs_rebuild_t<double>(c); }

void nblist_compact(Ctx& c) {
  if (c.prec == 32) compact_t<float>(c);
  else compact_t<double>(c);
}

void refresh_sorted_coords(Ctx& c) {
  if (c.prec == 32) refresh_t<float>(c);
  else refresh_t<double>(c);
}

}  // namespace t9
