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
#include <climits>
#include <cstdlib>
#include <cstring>

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

// Single-workgroup exclusive scan of m values into out[0..m-1], and the sum
// into out[m] when `with_total`: each thread sums a contiguous segment, the
// segment sums are scanned across the wavefront by shuffles and across the
// wavefronts through LDS (one barrier).
template <class T>
__global__ __launch_bounds__(kScanBlock) void k_scan(const T* __restrict__ in, T* __restrict__ out, int m,
                                                     int with_total) {
  __shared__ T wsum[kScanBlock / 64];
  const int per = (m + kScanBlock - 1) / kScanBlock;
  const int b = min(m, (int)threadIdx.x * per), e = min(m, b + per);
  T sum = 0;
  for (int t0 = b; t0 < e; t0 += 8) {  // 8 loads in flight
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = t0 + u < e ? in[t0 + u] : (T)0;
#pragma unroll
    for (int u = 0; u < 8; ++u) sum += v[u];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T incl = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  T run = incl - sum;
  for (int u = 0; u < w; ++u) run += wsum[u];
  for (int t0 = b; t0 < e; t0 += 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = t0 + u < e ? in[t0 + u] : (T)0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (t0 + u < e) {
        out[t0 + u] = run;
        run += v[u];
      }
  }
  if (with_total && threadIdx.x == kScanBlock - 1) out[m] = run;
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

// The three lists share one scan over the candidate cells: the multipole
// list (atoms within rm), the preconditioner list (atoms within ru <= rm) and
// the vdW list (sites within rv, sites read from vpos).
template <class real>
struct ListArgs {
  int n, fill, brute;  // fill: 0 count, 1 write at the counted offsets, 2 count and stage in part windows
  const PosRec<real>* pos;
  const PosRec<real>* vpos;  // null: no vdW list
  const int* cellc;
  int nc[3];
  const int* cell_start;
  const int* cell_count;
  BoxT<real> box;
  real rm2, ru2, rv2;        // ru2 <= 0: no preconditioner list
  double reach2;             // squared reach of the cell search (A^2)
  double csize[3];           // cell edge (A)
  double flen[3];            // box edge (FP64 positions are wrapped coordinates)
  const int* perm;
  const int *eoff, *ej, *ecode;     // scaled partners of the atom lists
  const int *veoff, *vej, *vecode;  // scaled partners of the vdW list
  int *nm, *nu, *nv;                // totals (written by k_sum_parts)
  int* parts;                       // [3][kParts][n] counts per part of the cell region
  const long long *om, *ou, *ov;    // slab offsets (fill 1)
  uint32_t *em, *eu, *ev;           // list slabs (fill 1) or part windows (fill 2)
  int wm, wu, wv;                   // part window widths (fill 2)
};

constexpr int kParts = 16;  // waves sharing one wave of atoms (one workgroup)

__device__ __forceinline__ double frac_of(const PosRec<float>& p, int d, const double*) {
  const unsigned v = d == 0 ? (unsigned)p.x : (d == 1 ? (unsigned)p.y : (unsigned)p.z);
  return (double)v * (1.0 / 4294967296.0);
}
__device__ __forceinline__ double frac_of(const PosRec<double>& p, int d, const double* flen) {
  return (d == 0 ? p.x : (d == 1 ? p.y : p.z)) / flen[d];
}

// A lane's scaled partners: the first 8 in registers, the rest read when
// needed (none for water and most of a protein's atoms).
struct Partners {
  int j[8], code[8];
  int e0, e1;
  __device__ __forceinline__ void load(int a, int b, const int* __restrict__ ej, const int* __restrict__ ec) {
    e0 = a;
    e1 = b;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const bool in = a + k < b;
      j[k] = in ? ej[a + k] : -1;
      code[k] = in ? ec[a + k] : 0;
    }
  }
  __device__ __forceinline__ int find(int oj, const int* __restrict__ ej, const int* __restrict__ ec) const {
    int c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) c = j[k] == oj ? code[k] : c;
    for (int e = e0 + 8; e < e1; ++e)
      if (ej[e] == oj) c = ec[e];
    return c;
  }
};

__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ double wave_minf(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_maxf(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ int wrapc(int x, int nc) {
  x += x < 0 ? nc : 0;
  return x - (x >= nc ? nc : 0);
}

// Wave-uniform broadcast of lane t's value (t uniform).
__device__ __forceinline__ int bcast(int v, int t) { return __builtin_amdgcn_readlane(v, t); }
__device__ __forceinline__ float bcast(float v, int t) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), t));
}
__device__ __forceinline__ double bcast(double v, int t) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), t);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), t);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <class real>
__device__ __forceinline__ PosRec<real> bcast(const PosRec<real>& p, int t) {
  PosRec<real> q;
  q.x = bcast(p.x, t);
  q.y = bcast(p.y, t);
  q.z = bcast(p.z, t);
  q.w = bcast(p.w, t);
  return q;
}

// Count (fill == 0) or write (fill == 1) each sorted atom's neighbors.
// The candidates are walked wave-uniformly in chunks of up to 64 atoms of
// one cell: each lane loads one candidate (one load instruction per chunk,
// the next chunk prefetched while the current one is tested), then every
// lane tests every candidate of the chunk, broadcast lane by lane.  The cells
// are those around the wave's bounding box when its atoms span at most 4
// cells per axis (the usual case: 64 consecutive atoms in cell Morton
// order), else the 5^3 stencil of each home cell in turn, tested only by
// that cell's lanes; without cells, all atoms.  The kParts waves of a
// workgroup take the same 64 atoms and split the cells (or the atom range)
// in kParts fixed parts (latency hiding: 16 waves per 64 atoms); each part
// writes its own range of list slots.
template <class real>
__global__ __launch_bounds__(64 * kParts) void k_nb_lists(ListArgs<real> A) {
  const int lane = threadIdx.x & 63, part = threadIdx.x >> 6;
  const int blk = xcd_block();
  const int s0 = blk * 64 + lane;
  const bool active = s0 < A.n;
  const int s = active ? s0 : A.n - 1;  // idle lanes mirror the last atom
  const auto pi = A.pos[s];
  const bool has_u = A.ru2 > 0, has_v = A.vpos != nullptr;
  PosRec<real> vi = pi;
  if (has_v) vi = A.vpos[s];
  const int oi = A.perm[s];
  Partners px, pv;
  if (A.fill) {
    px.load(A.eoff[oi], A.eoff[oi + 1], A.ej, A.ecode);
    if (has_v) pv.load(A.veoff[oi], A.veoff[oi + 1], A.vej, A.vecode);
  }
  int cm = 0, cu = 0, cv = 0;
  int lm = 0x7fffffff, lu = 0x7fffffff, lv = 0x7fffffff;
  uint32_t *om = nullptr, *ou = nullptr, *ov = nullptr;
  const size_t pn = (size_t)kParts * A.n;
  if (A.fill == 2 && active) {
    // a window of wm (wu, wv) slots per atom and part; entries beyond it are
    // counted, not written (the host then writes the lists at the counted offsets)
    const size_t wb = (size_t)blk * kParts + part;
    om = A.em + wb * A.wm * 64 + lane;
    if (has_u) ou = A.eu + wb * A.wu * 64 + lane;
    if (has_v) ov = A.ev + wb * A.wv * 64 + lane;
  }
  if (A.fill == 1 && active) {
    // this part's slots follow those of the parts before it
    int pm = 0, pu = 0, pv = 0;
    for (int q = 0; q < part; ++q) {
      pm += A.parts[(size_t)q * A.n + s];
      pu += A.parts[pn + (size_t)q * A.n + s];
      pv += A.parts[2 * pn + (size_t)q * A.n + s];
    }
    lm = A.parts[(size_t)part * A.n + s];
    lu = A.parts[pn + (size_t)part * A.n + s];
    lv = A.parts[2 * pn + (size_t)part * A.n + s];
    om = A.em + A.om[s >> 6] + lane + 64LL * pm;
    if (has_u) ou = A.eu + A.ou[s >> 6] + lane + 64LL * pu;
    if (has_v) ov = A.ev + A.ov[s >> 6] + lane + 64LL * pv;
  }
  // region of cells: base r0, size rs; pruned against the wave's bounding box
  int c = 0, lo0 = 0, lo1 = 0, lo2 = 0, hi0 = 0, hi1 = 0, hi2 = 0;
  double bl0 = 0, bl1 = 0, bl2 = 0, bh0 = 0, bh1 = 0, bh2 = 0;
  bool compact = true;
  if (!A.brute) {
    c = A.cellc[s];
    const int cx = c & 1023, cy = (c >> 10) & 1023, cz = (c >> 20) & 1023;
    lo0 = wave_min(cx);
    lo1 = wave_min(cy);
    lo2 = wave_min(cz);
    hi0 = wave_max(cx);
    hi1 = wave_max(cy);
    hi2 = wave_max(cz);
    compact = hi0 - lo0 <= 3 && hi1 - lo1 <= 3 && hi2 - lo2 <= 3 && hi0 - lo0 + 5 <= A.nc[0] &&
              hi1 - lo1 + 5 <= A.nc[1] && hi2 - lo2 + 5 <= A.nc[2];
    if (compact) {
      bl0 = wave_minf(frac_of(pi, 0, A.flen) * A.nc[0]);
      bh0 = wave_maxf(frac_of(pi, 0, A.flen) * A.nc[0]);
      bl1 = wave_minf(frac_of(pi, 1, A.flen) * A.nc[1]);
      bh1 = wave_maxf(frac_of(pi, 1, A.flen) * A.nc[1]);
      bl2 = wave_minf(frac_of(pi, 2, A.flen) * A.nc[2]);
      bh2 = wave_maxf(frac_of(pi, 2, A.flen) * A.nc[2]);
    }
  }
  unsigned long long todo = (A.brute || compact) ? 1ull : __ballot(1);
  while (todo) {
    // one walk: the whole box (brute), the wave's region (compact), or the
    // stencil of one home cell
    int r00, r01, r02, rs0, rs1, rs2;
    bool on = active, prune = false;
    if (A.brute || compact) {
      todo = 0;
      r00 = lo0 - 2;
      r01 = lo1 - 2;
      r02 = lo2 - 2;
      rs0 = hi0 - lo0 + 5;
      rs1 = hi1 - lo1 + 5;
      rs2 = hi2 - lo2 + 5;
      prune = !A.brute;
    } else {
      const int lead = __builtin_ctzll(todo);
      const int h = __shfl(c, lead, 64);
      const bool mine = c == h;
      todo &= ~__ballot(mine);
      on = active && mine;
      r00 = (h & 1023) - 2;
      r01 = ((h >> 10) & 1023) - 2;
      r02 = ((h >> 20) & 1023) - 2;
      rs0 = rs1 = rs2 = 5;
    }
    const int ncells = A.brute ? 1 : rs0 * rs1 * rs2;
    const int qa = A.brute ? 0 : part * ncells / kParts, qb = A.brute ? 1 : (part + 1) * ncells / kParts;
    int q = qa - 1, off = 0, total = 0, start = 0;
    // cell table of this walk: lane t holds the atom count and start of cell
    // qa + t (count 0 when pruned), loaded at once; the walk then takes the
    // nonempty cells in order from a ballot, with no memory wait per cell
    const bool table = !A.brute && qb - qa <= 64;
    int tcnt = 0, tstart = 0;
    unsigned long long cmask = 0;
    if (table) {
      const int qq = qa + lane;
      if (qq < qb) {
        const int x = r00 + qq % rs0, y = r01 + (qq / rs0) % rs1, z = r02 + qq / (rs0 * rs1);
        bool keep = true;
        if (prune) {
          const double gx = fmax(0.0, fmax((double)x - bh0, bl0 - (double)(x + 1))) * A.csize[0];
          const double gy = fmax(0.0, fmax((double)y - bh1, bl1 - (double)(y + 1))) * A.csize[1];
          const double gz = fmax(0.0, fmax((double)z - bh2, bl2 - (double)(z + 1))) * A.csize[2];
          keep = gx * gx + gy * gy + gz * gz <= A.reach2;
        }
        if (keep) {
          const int k = morton3(wrapc(x, A.nc[0]), wrapc(y, A.nc[1]), wrapc(z, A.nc[2]));
          tcnt = A.cell_count[k];
          tstart = A.cell_start[k];
        }
      }
      cmask = __ballot(tcnt > 0);
    }
    bool have = true;
    int j0 = 0, cnt = 0;
    PosRec<real> np = pi, nv = vi;
    int no = 0;
    // first chunk, then each iteration prefetches the next one
    for (int stage = 0;; ++stage) {
      // advance (q, off) to the next chunk (j0, cnt)
      if (q >= qa && off + 64 < total) {
        off += 64;
      } else {
        have = false;
        if (table) {
          if (cmask) {
            const int t = __builtin_ctzll(cmask);
            cmask &= cmask - 1;
            q = qa + t;
            total = __builtin_amdgcn_readlane(tcnt, t);
            start = __builtin_amdgcn_readlane(tstart, t);
            have = true;
          } else {
            q = qb;
          }
        } else while (++q < qb) {
          if (A.brute) {  // the atom range is split instead of the cells
            start = part * A.n / kParts;
            total = (part + 1) * A.n / kParts - start;
          } else {
            const int x = r00 + q % rs0, y = r01 + (q / rs0) % rs1, z = r02 + q / (rs0 * rs1);
            if (prune) {
              const double gx = fmax(0.0, fmax((double)x - bh0, bl0 - (double)(x + 1))) * A.csize[0];
              const double gy = fmax(0.0, fmax((double)y - bh1, bl1 - (double)(y + 1))) * A.csize[1];
              const double gz = fmax(0.0, fmax((double)z - bh2, bl2 - (double)(z + 1))) * A.csize[2];
              if (gx * gx + gy * gy + gz * gz > A.reach2) continue;
            }
            const int k = morton3(wrapc(x, A.nc[0]), wrapc(y, A.nc[1]), wrapc(z, A.nc[2]));
            total = __builtin_amdgcn_readfirstlane(A.cell_count[k]);
            start = __builtin_amdgcn_readfirstlane(A.cell_start[k]);
          }
          if (total > 0) {
            have = true;
            break;
          }
        }
        off = 0;
      }
      const PosRec<real> cp = np, cvp = nv;
      const int co = no, cj0 = j0, ccnt = stage == 0 ? 0 : cnt;
      if (have) {
        j0 = start + off;
        cnt = min(64, total - off);
        if (lane < cnt) {  // prefetch: lane t loads candidate j0 + t
          np = A.pos[j0 + lane];
          if (has_v) nv = A.vpos[j0 + lane];
          if (A.fill) no = A.perm[j0 + lane];
        }
      }
      for (int t = 0; t < ccnt; ++t) {
        const int j = cj0 + t;
        const PosRec<real> pj = bcast(cp, t);
        if (!on || j == s) continue;
        real xr, yr, zr;
        pair_disp(pi, pj, A.box, xr, yr, zr);
        const real r2 = xr * xr + yr * yr + zr * zr;
        const int oj = A.fill ? bcast(co, t) : 0;
        if (r2 < A.rm2 && cm < lm) {
          const bool in_u = has_u && r2 < A.ru2 && cu < lu;
          if (A.fill) {
            const int code = px.find(oj, A.ej, A.ecode);
            const uint32_t ent = ((uint32_t)j << kCodeBits) | (uint32_t)code;
            if (cm < A.wm) om[(long long)cm * 64] = ent;
            if (in_u && cu < A.wu) ou[(long long)cu * 64] = ent;
          }
          ++cm;
          cu += in_u ? 1 : 0;
        }
        if (has_v) {
          const PosRec<real> vj = bcast(cvp, t);
          pair_disp(vi, vj, A.box, xr, yr, zr);
          if (xr * xr + yr * yr + zr * zr < A.rv2 && cv < lv) {
            if (A.fill) {
              const int code = pv.find(oj, A.vej, A.vecode);
              if (cv < A.wv) ov[(long long)cv * 64] = ((uint32_t)j << kCodeBits) | (uint32_t)code;
            }
            ++cv;
          }
        }
      }
      if (!have) break;
    }
  }
  if (A.fill != 1 && active) {
    A.parts[(size_t)part * A.n + s] = cm;
    A.parts[pn + (size_t)part * A.n + s] = cu;
    A.parts[2 * pn + (size_t)part * A.n + s] = cv;
  }
}

// Per-atom totals of the part counts, and per wavefront the largest part
// count of each list (wmax[3 * wave + list]; k_max3 reduces them: they size
// the next build's part windows).
__global__ void k_sum_parts(int n, const int* __restrict__ parts, int* __restrict__ nm, int* __restrict__ nu,
                            int* __restrict__ nv, int* __restrict__ wmax) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = s < n;  // every lane takes part in the wavefront maxima
  const size_t pn = (size_t)kParts * n;
  int a = 0, b = 0, c = 0, ma = 0, mb = 0, mc = 0;
  if (on)
    for (int q = 0; q < kParts; ++q) {
      const int x = parts[(size_t)q * n + s], y = parts[pn + (size_t)q * n + s],
                z = parts[2 * pn + (size_t)q * n + s];
      a += x;
      b += y;
      c += z;
      ma = max(ma, x);
      mb = max(mb, y);
      mc = max(mc, z);
    }
  ma = wave_max(ma);
  mb = wave_max(mb);
  mc = wave_max(mc);
  const int wv = s >> 6;
  if ((threadIdx.x & 63) == 0 && wv * 64 < n) {
    wmax[3 * wv] = ma;
    wmax[3 * wv + 1] = mb;
    wmax[3 * wv + 2] = mc;
  }
  if (!on) return;
  nm[s] = a;
  if (nu) nu[s] = b;
  if (nv) nv[s] = c;
}

// Copy the staged part windows of the workgroup's 64 atoms into their list
// slab, parts in order: the entries land where a count-then-write build puts them.
// Rows are assembled kPackRows at a time in LDS (each part drops its entries
// at their final row) and written as whole 256-byte rows; written entry by
// entry, one store of a wavefront would touch 64 different rows.
constexpr int kPackRows = 128;
__global__ __launch_bounds__(64 * kParts) void k_pack_list(int n, const int* __restrict__ parts, int w,
                                                           const uint32_t* __restrict__ stage,
                                                           const long long* __restrict__ off,
                                                           uint32_t* __restrict__ out) {
  __shared__ int cnt[kParts][64];
  __shared__ uint32_t rows[kPackRows][64];
  const int lane = threadIdx.x & 63, part = threadIdx.x >> 6;
  const int blk = xcd_block();
  const int s = blk * 64 + lane;
  const int mine = s < n ? parts[(size_t)part * n + s] : 0;
  cnt[part][lane] = mine;
  __syncthreads();
  int before = 0, len = 0;
  for (int q = 0; q < kParts; ++q) {
    const int v = cnt[q][lane];
    before += q < part ? v : 0;
    len += v;
  }
  const int maxlen = wave_max(len);  // the same in every wavefront of the workgroup
  const uint32_t* src = stage + ((size_t)blk * kParts + part) * w * 64 + lane;
  uint32_t* dst = out + off[blk];
  for (int r0 = 0; r0 < maxlen; r0 += kPackRows) {
    // this lane's window rows [k0, k1) land in the chunk (walking the union
    // of the wavefront's ranges row by row was measured slower: the ranges
    // are far apart)
    const int k0 = max(0, r0 - before), k1 = min(mine, r0 + kPackRows - before);
    for (int k = k0; k < k1; k += 8) {  // 8 loads in flight
      uint32_t v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k + u < k1) v[u] = src[(long long)(k + u) * 64];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k + u < k1) rows[before + k + u - r0][lane] = v[u];
    }
    __syncthreads();
    const int nr = min(kPackRows, maxlen - r0);
    for (int q = threadIdx.x; q < nr * 64; q += 64 * kParts) dst[(long long)(r0 + (q >> 6)) * 64 + (q & 63)] = rows[q >> 6][q & 63];
    __syncthreads();
  }
}

// Largest of the per-wavefront maxima of each list (one workgroup).
__global__ __launch_bounds__(256) void k_max3(int nw, const int* __restrict__ wmax, int* __restrict__ pmax) {
  __shared__ int sm[3][4];
  int m[3] = {0, 0, 0};
  for (int w = threadIdx.x; w < nw; w += 256)
#pragma unroll
    for (int q = 0; q < 3; ++q) m[q] = max(m[q], wmax[3 * w + q]);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    m[q] = wave_max(m[q]);
    if ((threadIdx.x & 63) == 0) sm[q][threadIdx.x >> 6] = m[q];
  }
  __syncthreads();
  if (threadIdx.x < 3) pmax[threadIdx.x] = max(max(sm[threadIdx.x][0], sm[threadIdx.x][1]), max(sm[threadIdx.x][2], sm[threadIdx.x][3]));
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

// Displacement check: any atom that moved more than half the list buffer
// since the last build writes `stamp` (a new value per check) to the flag, so
// the flag needs no clearing between checks.
template <class real>
__global__ void k_max_disp(int n, const double* __restrict__ xyz, const double* __restrict__ last, double lx,
                           double ly, double lz, int periodic, double lim2, int stamp, int* __restrict__ flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double dx = xyz[3 * i] - last[3 * i], dy = xyz[3 * i + 1] - last[3 * i + 1], dz = xyz[3 * i + 2] - last[3 * i + 2];
  if (periodic) {
    dx -= lx * rint(dx / lx);
    dy -= ly * rint(dy / ly);
    dz -= lz * rint(dz / lz);
  }
  if (dx * dx + dy * dy + dz * dz > lim2) *flag = stamp;
}


// Slab offsets and total of one list once its counts are known; the total is
// copied to pinned memory at `h_total` (the host waits once for all lists,
// then list_capacity sizes the slabs).
void list_slots(Ctx& c, NbList& L, long long* h_total) {
  const int n = c.n, nw = c.nwave;
  L.maxlen.alloc(sizeof(int) * nw);
  L.off.alloc(sizeof(long long) * (nw + 1));
  L.slots.alloc(sizeof(long long) * nw);
  hipLaunchKernelGGL(k_wave_slots, dim3(ceil_div(nw, 256)), dim3(256), 0, c.stream, n, nw, L.nlist.as<int>(),
                     L.maxlen.as<int>(), L.slots.as<long long>());
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan<long long>, dim3(1), dim3(kScanBlock), 0, c.stream, L.slots.as<long long>(),
                     L.off.as<long long>(), nw, 1);
  T9_LAUNCH_CHECK();
  T9_HIP(hipMemcpyAsync(h_total, L.off.as<long long>() + nw, sizeof(long long), hipMemcpyDeviceToHost, c.stream));
}

void list_capacity(NbList& L, long long total) {
  if (total < 0 || total > (1LL << 31)) throw ListOverflow("neighbor list exceeds 2^31 slots");
  if ((size_t)total > L.capacity || !L.entries.p) {
    size_t cap = (size_t)(total * 1.2) + 64 * 64;
    L.entries.alloc(sizeof(uint32_t) * cap);
    L.capacity = cap;
  }
  L.total = total;
}

// All three lists from the candidate cells (staged single scan, or count then write).
template <class real>
void build_lists(Ctx& c) {
  const int n = c.n;
  const bool has_u = c.ulist.cutoff > 0, has_v = c.vlist.cutoff > 0;
  ListArgs<real> A;
  A.n = n;
  A.brute = c.brute ? 1 : 0;
  A.pos = c.pos.as<PosRec<real>>();
  A.vpos = has_v ? c.vpos.as<PosRec<real>>() : nullptr;
  A.cellc = c.cellc.as<int>();
  for (int d = 0; d < 3; ++d) {
    A.nc[d] = c.ncell[d];
    A.csize[d] = c.periodic ? c.box[d] / c.ncell[d] : 1.0;
    A.flen[d] = c.periodic ? c.box[d] : 1.0;
  }
  A.cell_start = c.cell_start.as<int>();
  A.cell_count = c.cell_count.as<int>();
  A.box = make_boxT<real>(c);
  A.rm2 = (real)(c.mlist.cutoff * c.mlist.cutoff);
  A.ru2 = has_u ? (real)(c.ulist.cutoff * c.ulist.cutoff) : (real)0;
  A.rv2 = has_v ? (real)(c.vlist.cutoff * c.vlist.cutoff) : (real)0;
  A.reach2 = c.cell_reach * c.cell_reach;
  A.perm = c.perm.as<int>();
  A.eoff = c.excl_off.as<int>();
  A.ej = c.excl_j.as<int>();
  A.ecode = c.excl_code.as<int>();
  A.veoff = has_v ? c.vexcl_off.as<int>() : nullptr;
  A.vej = has_v ? c.vexcl_j.as<int>() : nullptr;
  A.vecode = has_v ? c.vexcl_code.as<int>() : nullptr;
  for (NbList* L : {&c.mlist, &c.ulist, &c.vlist})
    if (L->cutoff > 0) L->nlist.alloc(sizeof(int) * n);
  A.nm = c.mlist.nlist.as<int>();
  A.nu = has_u ? c.ulist.nlist.as<int>() : nullptr;
  A.nv = has_v ? c.vlist.nlist.as<int>() : nullptr;
  // One candidate scan per build: the staged pass counts every list and
  // writes each part's entries into a window sized from the previous build;
  // the lists are then packed in part order.  The first build, a window that
  // overflowed, or T9HIP_LISTS=twopass count first and write at the counted
  // offsets in a second scan (the same lists; T9HIP_LISTS=tiny forces
  // one-entry windows, so the fallback runs).
  const size_t pn = (size_t)kParts * n;
  c.list_parts.alloc(sizeof(int) * 3 * pn);
  A.parts = c.list_parts.as<int>();
  c.list_pmax.alloc(sizeof(int) * (3 * (size_t)c.nwave + 4));  // per-wave maxima, then the 3 list maxima
  const char* mode = std::getenv("T9HIP_LISTS");
  const bool twopass = mode && !std::strcmp(mode, "twopass"), tiny = mode && !std::strcmp(mode, "tiny");
  NbList* lists[3] = {&c.mlist, &c.ulist, &c.vlist};
  DBuf* stage[3] = {&c.stage_m, &c.stage_u, &c.stage_v};
  int w[3];
  bool staged = !twopass;
  for (int q = 0; q < 3; ++q) {
    w[q] = tiny ? 1 : c.list_w[q];
    if (lists[q]->cutoff > 0 && w[q] <= 0) staged = false;
  }
  const int nwg = c.nwave;
  A.fill = 0;
  A.om = A.ou = A.ov = nullptr;
  A.em = A.eu = A.ev = nullptr;
  A.wm = A.wu = A.wv = INT_MAX;
  if (staged) {
    A.fill = 2;
    for (int q = 0; q < 3; ++q)
      if (lists[q]->cutoff > 0) {
        const size_t need = sizeof(uint32_t) * (size_t)w[q] * kParts * 64 * nwg;
        if (need > stage[q]->bytes) stage[q]->alloc(need + need / 2);  // headroom: windows drift up and down
      }
    A.em = c.stage_m.as<uint32_t>();
    A.wm = w[0];
    if (has_u) {
      A.eu = c.stage_u.as<uint32_t>();
      A.wu = w[1];
    }
    if (has_v) {
      A.ev = c.stage_v.as<uint32_t>();
      A.wv = w[2];
    }
  }
  hipLaunchKernelGGL(k_nb_lists<real>, dim3(nwg), dim3(64 * kParts), 0, c.stream, A);
  T9_LAUNCH_CHECK();
  int* wmax = c.list_pmax.as<int>();
  int* pmax = wmax + 3 * (size_t)nwg;
  hipLaunchKernelGGL(k_sum_parts, dim3(ceil_div(n, 256)), dim3(256), 0, c.stream, n, A.parts, A.nm, A.nu, A.nv,
                     wmax);
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_max3, dim3(1), dim3(256), 0, c.stream, nwg, wmax, pmax);
  T9_LAUNCH_CHECK();
  // the totals and the largest part counts come back in one readback
  // (pinned slots 45..47 and 42..43)
  long long* h_tot = reinterpret_cast<long long*>(c.h_scalars + 45);
  int* h_pmax = reinterpret_cast<int*>(c.h_scalars + 42);
  for (int q = 0; q < 3; ++q)
    if (lists[q]->cutoff > 0) list_slots(c, *lists[q], h_tot + q);
  T9_HIP(hipMemcpyAsync(h_pmax, pmax, sizeof(int) * 3, hipMemcpyDeviceToHost, c.stream));
  mark_readback(c);
  wait_marked(c);
  bool overflow = false;
  for (int q = 0; q < 3; ++q)
    if (lists[q]->cutoff > 0) {
      list_capacity(*lists[q], h_tot[q]);
      if (h_pmax[q] > w[q]) overflow = true;
      c.list_w[q] = h_pmax[q] + h_pmax[q] / 4 + 8;  // the next build's window
    }
  if (staged && !overflow) {
    for (int q = 0; q < 3; ++q)
      if (lists[q]->cutoff > 0) {
        hipLaunchKernelGGL(k_pack_list, dim3(nwg), dim3(64 * kParts), 0, c.stream, n, A.parts + q * pn, w[q],
                           stage[q]->as<uint32_t>(), lists[q]->off.as<long long>(), lists[q]->entries.as<uint32_t>());
        T9_LAUNCH_CHECK();
      }
    return;
  }
  A.fill = 1;
  A.wm = A.wu = A.wv = INT_MAX;
  A.om = c.mlist.off.as<long long>();
  A.em = c.mlist.entries.as<uint32_t>();
  if (has_u) {
    A.ou = c.ulist.off.as<long long>();
    A.eu = c.ulist.entries.as<uint32_t>();
  }
  if (has_v) {
    A.ov = c.vlist.off.as<long long>();
    A.ev = c.vlist.entries.as<uint32_t>();
  }
  hipLaunchKernelGGL(k_nb_lists<real>, dim3(nwg), dim3(64 * kParts), 0, c.stream, A);
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
  hipLaunchKernelGGL(k_scan<int>, dim3(1), dim3(kScanBlock), 0, c.stream, c.cell_count.as<int>(),
                     c.cell_start.as<int>(), nkeys, 0);
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
  if (c.vlist.cutoff > 0) vdw_sites(c);
  build_lists<real>(c);
  c.builds++;
  c.list_valid = true;
  c.coords_dirty = false;
  c.rpole_valid = false;
  c.bins_valid = false;
  c.mlist.cut_valid = c.ulist.cut_valid = false;
}

// Keep the entries within the interaction cutoff (same FP arithmetic as the
// pair kernels, so those kernels need no distance test of their own), in
// list order (which keeps the pair kernels' gathers spatially local); with
// cut2b > 0 also the entries within the shorter second radius into
// ecut2 / ncut2, same layout.  kPairParts wavefronts per 64 atoms: in each
// round the wavefronts take consecutive, wave-uniform chunks of C <= 64 list
// positions (so a load instruction reads one contiguous 256-byte row) and
// record which entries fall inside each radius in 64-bit masks; the per-part
// counts are summed in part order through LDS, and each part then places its
// kept entries from its offset on.  The output slab of the 64 atoms is
// assembled in LDS (the first kCompactRowsA / kCompactRowsB rows; later rows,
// rare, go straight to memory) and written as whole 256-byte rows: written
// entry by entry, every store of a wavefront would touch 64 different rows.

template <class real>
__global__ __launch_bounds__(64 * kPairParts) void k_compact_parts(
    int n, const int* __restrict__ nlist, const long long* __restrict__ woff, const uint32_t* __restrict__ entries,
    const PosRec<real>* __restrict__ pos, BoxT<real> box, real cut2, real cut2b, int* __restrict__ ncut,
    uint32_t* __restrict__ ecut, int* __restrict__ ncut2, uint32_t* __restrict__ ecut2) {
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
  PosRec<real> pi;
  if (valid) pi = pos[s];
  int tot_a = 0, tot_b = 0;  // kept entries of this atom in earlier rounds
  for (int k0 = 0; k0 < maxlen;) {
    const int C = min(64, (maxlen - k0 + P - 1) / P);
    const int kb = k0 + part * C;
    const int ke = min(kb + C, len);
    unsigned long long ma = 0, mb = 0;
    for (int k0 = kb; k0 < ke; k0 += 8) {  // 8 entries, then their 8 partners, in flight
      uint32_t ent[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) ent[u] = k0 + u < ke ? e[(long long)(k0 + u) * 64] : 0u;
      PosRec<real> pj[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < ke) pj[u] = pos[ent[u] >> kCodeBits];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < ke) {
          real xr, yr, zr;
          pair_disp(pi, pj[u], box, xr, yr, zr);
          const real r2 = xr * xr + yr * yr + zr * zr;
          if (r2 <= cut2) ma |= 1ull << (k0 + u - kb);
          if (two && r2 <= cut2b) mb |= 1ull << (k0 + u - kb);
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
    k0 += P * C;
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
}

template <class real>
void compact_t(Ctx& c, hipStream_t st) {
  BoxT<real> box = make_boxT<real>(c);
  NbList* L = &c.mlist;
  if (!nblist_compact_buffers(c)) return;
  const bool two = L->cut2nd > 0;
  const double c2 = L->cut * L->cut, c2b = L->cut2nd * L->cut2nd;
  hipLaunchKernelGGL(k_compact_parts<real>, dim3(c.nwave), dim3(64 * kPairParts), 0, st, c.n,
                     L->nlist.as<int>(), L->off.as<long long>(), L->entries.as<uint32_t>(), c.pos.as<PosRec<real>>(),
                     box, (real)(c2 > 1e30 ? 1e30 : c2), (real)(two ? c2b : -1.0), L->ncut.as<int>(),
                     L->ecut.as<uint32_t>(), two ? L->ncut2.as<int>() : nullptr,
                     two ? L->ecut2.as<uint32_t>() : nullptr);
  T9_LAUNCH_CHECK();
  L->cut_valid = true;
}

template <class real>
void refresh_kernels_t(Ctx& c);

template <class real>
bool needs_rebuild_t(Ctx& c, bool refresh) {
  if (!c.list_valid) return true;
  if (!c.disp_flag.p) {  // allocated and cleared once; stamps start at 1
    c.disp_flag.alloc(sizeof(int));
    T9_HIP(hipMemsetAsync(c.disp_flag.p, 0, sizeof(int), c.stream));
  }
  const int stamp = c.disp_stamp = c.disp_stamp == INT_MAX ? 1 : c.disp_stamp + 1;
  double half = 0.5 * c.opt.list_buffer;
  hipLaunchKernelGGL(k_max_disp<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.xyz.as<double>(),
                     c.xyz_last.as<double>(), c.box[0], c.box[1], c.box[2], c.periodic ? 1 : 0, half * half, stamp,
                     c.disp_flag.as<int>());
  T9_LAUNCH_CHECK();
  // read into pinned memory (a spare slot of the scalar readback buffer)
  int* h = reinterpret_cast<int*>(c.h_scalars + kDispSlot);
  T9_HIP(hipMemcpyAsync(h, c.disp_flag.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  mark_readback(c);
  if (refresh) refresh_kernels_t<real>(c);  // runs while the host waits; a rebuild overwrites it
  wait_marked(c);
  return *h == stamp;
}

template <class real>
void refresh_kernels_t(Ctx& c) {
  hipLaunchKernelGGL(k_refresh_pos<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.as<int>(),
                     c.xyz.as<double>(), make_double3(c.vorg[0], c.vorg[1], c.vorg[2]),
                     make_double3(c.vlen[0], c.vlen[1], c.vlen[2]), c.periodic ? 1 : 0, c.pdamp_o.as<double>(),
                     c.pos.as<PosRec<real>>());
  T9_LAUNCH_CHECK();
  if (c.vlist.cutoff > 0) vdw_sites(c);
}

template <class real>
void refresh_t(Ctx& c) {
  refresh_kernels_t<real>(c);
  nblist_mark_refreshed(c);
}

}  // namespace

void nblist_mark_refreshed(Ctx& c) {
  c.coords_dirty = false;
  c.rpole_valid = false;
  c.bins_valid = false;
  c.mlist.cut_valid = c.ulist.cut_valid = false;
}

void nblist_build(Ctx& c) {
  if (c.prec == 32) nblist_build_t<float>(c);
  else nblist_build_t<double>(c);
}

bool nblist_needs_rebuild(Ctx& c, bool refresh) {
  return c.prec == 32 ? needs_rebuild_t<float>(c, refresh) : needs_rebuild_t<double>(c, refresh);
}

bool nblist_compact_buffers(Ctx& c) {
  NbList* L = &c.mlist;
  if (L->cutoff <= 0 || !L->entries.p) return false;
  L->ncut.alloc(sizeof(int) * c.n);
  L->ecut.alloc(sizeof(uint32_t) * L->capacity);
  if (L->cut2nd > 0) {
    L->ncut2.alloc(sizeof(int) * c.n);
    L->ecut2.alloc(sizeof(uint32_t) * L->capacity);
  }
  return true;
}

void nblist_compact(Ctx& c, hipStream_t st) {
  if (!st) st = c.stream;
  if (c.prec == 32) compact_t<float>(c, st);
  else compact_t<double>(c, st);
}

void refresh_sorted_coords(Ctx& c) {
  if (c.prec == 32) refresh_t<float>(c);
  else refresh_t<double>(c);
}

}  // namespace t9
