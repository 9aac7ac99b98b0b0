// Particle-mesh Ewald for gfx950: order-5 B-splines, deterministic
// spreading, an own mixed-radix (2,3,4,5) 3D FFT and the gathers that turn
// the reciprocal potential into fields.
//
// Reference behaviour followed: src/acc/pme.cpp:10-979 and src/cu/pme.cu
// (fractional coordinates w -> (w + 1/2) - floor(w + 1/2), grid base
// floor(K w) - order + 1, x-fastest grid, d dipoles in the real part and p
// dipoles in the imaginary part of one complex grid), src/pme.cpp:84-117 and
// the Fortran dftmod rule (B-spline moduli with the optimal zeta factor),
// pmeConv (src/acc/pme.cpp:220-303, k = 0 zeroed, exp cut at -50), and the
// reciprocal + self fields of src/cu/amoeba/field.cu:12-118.
//
// Spreading ("owner computes"): the grid is cut into tiles whose edge divides
// K; each workgroup owns one tile, takes the atoms whose support base lies in
// its own tile or in the adjacent lower tiles, and each thread sums, in a
// fixed atom order, the contributions to one row of the tile's own points.
// Every grid point is written once by one thread: the grid — and everything
// downstream — is bit-reproducible with no atomics.
//
// FFT: one workgroup transforms 16 (FP32) or 8 (FP64) grid lines in LDS with
// a Stockham auto-sort over radices 4, 2, 3, 5; the influence-function
// multiply is fused between the forward and inverse transforms of the z pass.
#include <cstdlib>
#include <cstring>

#include "bspline.hpp"
#include "ctx.hpp"

namespace t9 {
namespace {

constexpr double kPi = 3.14159265358979323846;

template <class real>
struct C2;
template <>
struct C2<float> {
  using T = float2;
};
template <>
struct C2<double> {
  using T = double2;
};
template <class real>
using cplx_t = typename C2<real>::T;

template <class real>
using R4 = typename RealTraits<real>::real4;

// ------------------------------------------------------------ binning

// Fractional offsets and support bases of every sorted atom, and the count of
// atoms per bin (bin edge B per axis, nb bins per axis) of the support base.
template <class real>
__global__ void k_pme_bin(int n, const int* __restrict__ perm, const double* __restrict__ xyz, double lx, double ly,
                          double lz, int K1, int K2, int K3, int B1, int B2, int B3, int nb1, int nb2,
                          R4<real>* __restrict__ pw, int4* __restrict__ pbase, int* __restrict__ tile_count) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  const double L[3] = {lx, ly, lz};
  const int K[3] = {K1, K2, K3};
  int b[3];
  double w[3];
  for (int d = 0; d < 3; ++d) {
    double f = xyz[3 * i + d] / L[d];
    f = f + 0.5 - floor(f + 0.5);
    double u = K[d] * f;
    int ig = (int)floor(u);
    w[d] = u - ig;
    int bb = ig - 4;
    while (bb < 0) bb += K[d];
    while (bb >= K[d]) bb -= K[d];
    b[d] = bb;
  }
  const int tile = b[0] / B1 + nb1 * (b[1] / B2 + nb2 * (b[2] / B3));
  R4<real> ww;
  ww.x = (real)w[0];
  ww.y = (real)w[1];
  ww.z = (real)w[2];
  ww.w = 0;
  pw[s] = ww;
  pbase[s] = make_int4(b[0], b[1], b[2], tile);
  atomicAdd(&tile_count[tile], 1);
}

__global__ void k_tile_scatter(int n, const int4* __restrict__ pbase, int* __restrict__ cursor,
                               int* __restrict__ atoms) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int t = pbase[s].w;
  atoms[atomicAdd(&cursor[t], 1)] = s;
}

// Each bin's atoms in ascending sorted index (insertion sort, one thread per
// bin; the scatter above fills a bin in arrival order).  This fixes the order
// in which the row-owner spreading adds the atoms of a tile.
__global__ void k_bin_sort(int nbins, const int* __restrict__ start, int* __restrict__ atoms) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  const int lo = start[b], hi = start[b + 1];
  for (int i = lo + 1; i < hi; ++i) {
    const int v = atoms[i];
    int j = i - 1;
    while (j >= lo && atoms[j] > v) {
      atoms[j + 1] = atoms[j];
      --j;
    }
    atoms[j + 1] = v;
  }
}

// Exclusive prefix sums of the per-bin atom counts into start[0..m]
// (start[m] the total), each bin's scatter cursor set to its start, and the
// counts cleared for the next binning.  Two launches over chunks of 1024
// bins, one 1024-thread workgroup per chunk: k_scan_sums writes each chunk's
// total; k_scan_apply adds the totals of the chunks before its own (in chunk
// order), scans its chunk (wavefront shuffles, then the 16 wavefront totals
// through LDS) and writes the starts with coalesced stores.
constexpr int kScanChunk = 1024;

__device__ __forceinline__ int block_scan_1024(int c, int& total) {
  __shared__ int wsum[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int before = 0;
  total = 0;
  for (int u = 0; u < 16; ++u) {
    if (u < w) before += wsum[u];
    total += wsum[u];
  }
  return before + incl - c;  // exclusive
}

__global__ __launch_bounds__(kScanChunk) void k_scan_sums(const int* __restrict__ count, int* __restrict__ csum,
                                                          int m) {
  const int t = blockIdx.x * kScanChunk + (int)threadIdx.x;
  int total;
  (void)block_scan_1024(t < m ? count[t] : 0, total);
  if (threadIdx.x == 0) csum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanChunk) void k_scan_apply(int* __restrict__ count, int* __restrict__ start,
                                                           int* __restrict__ cursor, const int* __restrict__ csum,
                                                           int m) {
  const int t = blockIdx.x * kScanChunk + (int)threadIdx.x;
  int carry = 0;
  for (int b = 0; b < (int)blockIdx.x; ++b) carry += csum[b];
  const int c = t < m ? count[t] : 0;
  int total;
  const int ex = carry + block_scan_1024(c, total);
  if (t < m) {
    start[t] = ex;
    cursor[t] = ex;
    count[t] = 0;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) start[m] = carry + total;
}

// ------------------------------------------------------------ spreading

struct TileGeom {
  int K[3], T[3], nt[3];
};

// MODE 0: induced dipoles d (real part) and p (imaginary part).
// MODE 1: permanent multipoles (real part; imaginary part zero).
// MODE 2: multipoles (real part) and dipoles (imaginary part).
//
// Owner-computes spreading: the workgroup owns one tile.  Phase 1 stages the
// B-spline weights and sources of the atoms whose support base lies in its own
// bin or in one of the seven lower-neighbour bins (chunks of at most
// kSpreadChunk atoms) in LDS.  Phase 2 hands out (atom, grid row) items; each
// adds the row's points that fall inside the tile into an LDS copy of the
// tile as 64-bit fixed-point integers (scale RealTraits::fixed_scale).
// Integer addition is associative, so the tile holds the same bits whatever
// order the items run in; it is converted and written to the grid once.
// Every grid point has one owner: no global atomics, bit-reproducible.
constexpr int kSpreadChunk = 64;
constexpr int kMaxTile = 16;

template <class real, int MODE>
struct SpreadAtom {
  static constexpr int LV = MODE == 0 ? 2 : 3;
  real th[3][LV][5];
  real f[MODE == 0 ? 6 : (MODE == 1 ? 10 : 13)];
  int off[3];   // support base relative to the tile origin
  int mlo[3];   // first support index inside the tile, per axis
  int cnt[3];   // support points inside the tile, per axis (0: none)
};

// Support indices m (0..4) of an atom whose base sits at `off` from the tile
// origin that land inside the tile [0, T), as a contiguous range: the point
// index is off + m, wrapped once by K.
__device__ __forceinline__ void support_range(int off, int T, int K, int& mlo, int& cnt) {
  if (T == K) {  // one tile spans the axis: the whole (wrapped) support is inside
    mlo = 0;
    cnt = 5;
    return;
  }
  // several tiles: T divides K and T >= 4, so K >= T + 4 and the in-tile
  // points form one contiguous run of m
  int a, b;
  if (off >= K - 4) {  // base at the far end of the grid: the support wraps to 0..
    a = K - off;
    b = 4;
  } else if (off >= 0) {
    a = 0;
    b = min(4, T - 1 - off);
  } else {  // base in a lower neighbour tile
    a = -off;
    b = 4;
  }
  mlo = a;
  cnt = b >= a ? b - a + 1 : 0;
}

__device__ __forceinline__ int support_l(int off, int m, int K) {
  int l = off + m;
  if (l >= K) l -= K;
  return l;
}

__device__ __forceinline__ unsigned long long to_fixed(double v, double scale) {
  return (unsigned long long)(long long)rint(v * scale);
}

template <class real, int MODE>
__global__ __launch_bounds__(256) void k_spread(TileGeom g, const int* __restrict__ tile_start,
                                                const int* __restrict__ tile_atoms, const R4<real>* __restrict__ pw,
                                                const int4* __restrict__ pbase, const real* __restrict__ src1,
                                                const real* __restrict__ src2, real a1, real a2, real a3, int chunk,
                                                cplx_t<real>* __restrict__ grid, Gate gate) {
  if (gate_closed(gate)) return;
  using SA = SpreadAtom<real, MODE>;
  constexpr int LV = SA::LV;
  constexpr int NACC = MODE == 1 ? 1 : 2;
  extern __shared__ __align__(16) char smem[];
  const int T1 = g.T[0], T2 = g.T[1], T3 = g.T[2];
  // padded pitches spread the rows of one atom's support over LDS banks
  const int P1 = T1 + 1, P2 = T2 + 1;
  const int nacc = P1 * P2 * T3;
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(smem);  // [NACC][nacc]
  int* pref = reinterpret_cast<int*>(smem + sizeof(unsigned long long) * NACC * nacc);  // [chunk + 1]
  SA* atoms = reinterpret_cast<SA*>(smem + sizeof(unsigned long long) * NACC * nacc +
                                    sizeof(int) * ((chunk + 2) & ~1));
  __shared__ int bin_start[8], bin_count[8], bin_pref[9];
  __shared__ int wsum[4];
  const int tt = blockIdx.x;
  const int t1 = tt % g.nt[0], t2 = (tt / g.nt[0]) % g.nt[1], t3 = tt / (g.nt[0] * g.nt[1]);
  const int lo[3] = {t1 * T1, t2 * T2, t3 * T3};
  if (threadIdx.x < 8) {
    const int c = threadIdx.x;
    const int da = c & 1, db = (c >> 1) & 1, dc = c >> 2;
    int st = 0, ct = 0;
    if (!((da && g.nt[0] == 1) || (db && g.nt[1] == 1) || (dc && g.nt[2] == 1))) {
      int b1 = t1 - da, b2 = t2 - db, b3 = t3 - dc;
      if (b1 < 0) b1 += g.nt[0];
      if (b2 < 0) b2 += g.nt[1];
      if (b3 < 0) b3 += g.nt[2];
      const int bin = b1 + g.nt[0] * (b2 + g.nt[1] * b3);
      st = tile_start[bin];
      ct = tile_start[bin + 1] - st;
    }
    bin_start[c] = st;
    bin_count[c] = ct;
  }
  for (int q = threadIdx.x; q < NACC * nacc; q += blockDim.x) acc[q] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    bin_pref[0] = 0;
    for (int c = 0; c < 8; ++c) bin_pref[c + 1] = bin_pref[c] + bin_count[c];
  }
  __syncthreads();
  const double scale = RealTraits<real>::fixed_scale;
  const int total = bin_pref[8];
  for (int base = 0; base < total; base += chunk) {
    const int nchunk = min(chunk, total - base);
    // ---- phase 1: stage weights and the in-tile support ranges; count the
    // atom's (y, z) rows inside the tile
    int nrow = 0;
    for (int a = threadIdx.x; a < nchunk; a += blockDim.x) {
      const int idx = base + a;
      int c = 0;
      while (idx >= bin_pref[c + 1]) ++c;
      const int s = tile_atoms[bin_start[c] + idx - bin_pref[c]];
      SA& A = atoms[a];
      const int4 b = pbase[s];
      A.off[0] = b.x - lo[0];
      A.off[1] = b.y - lo[1];
      A.off[2] = b.z - lo[2];
      support_range(A.off[0], T1, g.K[0], A.mlo[0], A.cnt[0]);
      support_range(A.off[1], T2, g.K[1], A.mlo[1], A.cnt[1]);
      support_range(A.off[2], T3, g.K[2], A.mlo[2], A.cnt[2]);
      nrow = A.cnt[0] ? A.cnt[1] * A.cnt[2] : 0;
      const R4<real> w = pw[s];
      real th[LV][5];
      bspline5<LV>(w.x, th);
#pragma unroll
      for (int lv = 0; lv < LV; ++lv)
#pragma unroll
        for (int m = 0; m < 5; ++m) A.th[0][lv][m] = th[lv][m];
      bspline5<LV>(w.y, th);
#pragma unroll
      for (int lv = 0; lv < LV; ++lv)
#pragma unroll
        for (int m = 0; m < 5; ++m) A.th[1][lv][m] = th[lv][m];
      bspline5<LV>(w.z, th);
#pragma unroll
      for (int lv = 0; lv < LV; ++lv)
#pragma unroll
        for (int m = 0; m < 5; ++m) A.th[2][lv][m] = th[lv][m];
      if (MODE == 0) {
        A.f[0] = src1[3 * s] * a1;
        A.f[1] = src1[3 * s + 1] * a2;
        A.f[2] = src1[3 * s + 2] * a3;
        A.f[3] = src2[3 * s] * a1;
        A.f[4] = src2[3 * s + 1] * a2;
        A.f[5] = src2[3 * s + 2] * a3;
      } else {
        const real* p = src1 + 10 * (size_t)s;
        A.f[0] = p[0];
        A.f[1] = p[1] * a1;
        A.f[2] = p[2] * a2;
        A.f[3] = p[3] * a3;
        A.f[4] = p[4] * a1 * a1;
        A.f[5] = p[5] * a2 * a2;
        A.f[6] = p[6] * a3 * a3;
        A.f[7] = 2 * p[7] * a1 * a2;
        A.f[8] = 2 * p[8] * a1 * a3;
        A.f[9] = 2 * p[9] * a2 * a3;
        if (MODE == 2) {
          A.f[MODE == 2 ? 10 : 0] = src2[3 * s] * a1;
          A.f[MODE == 2 ? 11 : 0] = src2[3 * s + 1] * a2;
          A.f[MODE == 2 ? 12 : 0] = src2[3 * s + 2] * a3;
        }
      }
    }
    // exclusive scan of the row counts (chunk <= blockDim: one atom per thread)
    {
      int v = nrow;
      const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
      }
      if (lane == 63) wsum[wv] = v;
      __syncthreads();
      int add = 0;
      for (int q = 0; q < wv; ++q) add += wsum[q];
      if ((int)threadIdx.x < nchunk) pref[threadIdx.x + 1] = v + add;
      if (threadIdx.x == 0) pref[0] = 0;
      __syncthreads();
    }
    // ---- phase 2: one item per in-tile row of a staged atom
    const int nitem = pref[nchunk];
    for (int it = threadIdx.x; it < nitem; it += blockDim.x) {
      int lo_a = 0, hi_a = nchunk - 1;  // last atom with pref[a] <= it
      while (lo_a < hi_a) {
        const int mid = (lo_a + hi_a + 1) >> 1;
        if (pref[mid] <= it) lo_a = mid;
        else hi_a = mid - 1;
      }
      const SA& A = atoms[lo_a];
      const int q = it - pref[lo_a];
      const int m2 = A.mlo[1] + q % A.cnt[1], m3 = A.mlo[2] + q / A.cnt[1];
      const int ly = support_l(A.off[1], m2, g.K[1]);
      const int lz = support_l(A.off[2], m3, g.K[2]);
      const real u0 = A.th[1][0][m2], u1 = A.th[1][1][m2];
      const real v0 = A.th[2][0][m3], v1 = A.th[2][1][m3];
      const int rowbase = P1 * (ly + P2 * lz);
      const int m1a = A.mlo[0], m1b = A.mlo[0] + A.cnt[0];
      if (MODE == 0) {
        const real Ad = A.f[0] * u0 * v0, Bd = A.f[1] * u1 * v0 + A.f[2] * u0 * v1;
        const real Ap = A.f[3] * u0 * v0, Bp = A.f[4] * u1 * v0 + A.f[5] * u0 * v1;
        for (int m1 = m1a; m1 < m1b; ++m1) {
          const int lx = support_l(A.off[0], m1, g.K[0]);
          const real t0 = A.th[0][0][m1], t1v = A.th[0][1][m1];
          atomicAdd(&acc[rowbase + lx], to_fixed((double)(Ad * t1v + Bd * t0), scale));
          atomicAdd(&acc[nacc + rowbase + lx], to_fixed((double)(Ap * t1v + Bp * t0), scale));
        }
      } else {
        const real u2 = A.th[1][LV > 2 ? 2 : 0][m2], v2 = A.th[2][LV > 2 ? 2 : 0][m3];
        const real* f = A.f;
        const real R0 = f[0] * u0 * v0 + f[2] * u1 * v0 + f[3] * u0 * v1 + f[5] * u2 * v0 + f[6] * u0 * v2 +
                        f[9] * u1 * v1;
        const real R1 = f[1] * u0 * v0 + f[7] * u1 * v0 + f[8] * u0 * v1;
        const real R2 = f[4] * u0 * v0;
        real Ap = 0, Bp = 0;
        if (MODE == 2) {
          Ap = f[MODE == 2 ? 10 : 0] * u0 * v0;
          Bp = f[MODE == 2 ? 11 : 0] * u1 * v0 + f[MODE == 2 ? 12 : 0] * u0 * v1;
        }
        for (int m1 = m1a; m1 < m1b; ++m1) {
          const int lx = support_l(A.off[0], m1, g.K[0]);
          const real t0 = A.th[0][0][m1], t1v = A.th[0][1][m1], t2v = A.th[0][LV > 2 ? 2 : 0][m1];
          atomicAdd(&acc[rowbase + lx], to_fixed((double)(R0 * t0 + R1 * t1v + R2 * t2v), scale));
          if (MODE == 2) atomicAdd(&acc[(NACC - 1) * nacc + rowbase + lx], to_fixed((double)(Ap * t1v + Bp * t0), scale));
        }
      }
    }
    __syncthreads();
  }
  // ---- convert and write the tile once
  const double inv = 1.0 / scale;
  const int npt = T1 * T2 * T3;
  for (int pnt = threadIdx.x; pnt < npt; pnt += blockDim.x) {
    const int l1 = pnt % T1, r = pnt / T1;
    const int rl2 = r % T2, rl3 = r / T2;
    const int ai = l1 + P1 * (rl2 + P2 * rl3);
    const long long gi = (lo[0] + l1) + (long long)g.K[0] * ((lo[1] + rl2) + (long long)g.K[1] * (lo[2] + rl3));
    cplx_t<real> v;
    v.x = (real)((double)(long long)acc[ai] * inv);
    v.y = NACC == 2 ? (real)((double)(long long)acc[(NACC - 1) * nacc + ai] * inv) : (real)0;
    grid[gi] = v;
  }
}

// Row-owner spreading (grid sizes divisible by 8, at least 16): the grid is
// cut into 8^3 tiles and the support bases into 4^3 bins; every atom whose
// support reaches a tile has its base in one of the 27 bins around the tile's
// own 8 (bin offsets -1, 0, +1 per axis).  One 256-thread workgroup per tile
// stages those atoms in LDS, in bin order and, within a bin, in ascending
// sorted index: their scaled sources and their B-spline weights laid out on
// the tile's 8 points per axis (zero where the support does not reach), so
// the accumulation has no branch and reads each weight row with broadcast
// loads.  Lane (y, z) of each wavefront owns the tile's grid row (y, z): it
// keeps the row's 8 points in registers and adds the atoms one after another;
// wave w takes atoms w, w + 4, ... and the four waves' rows are added in wave
// order at the end.  Every grid point is written once, by one lane, and its
// sum runs in an order fixed by the atoms' positions alone: bit-reproducible,
// with no atomics and no fixed point.
constexpr int kRowTile = 8, kRowBin = 4, kRowChunk = 64;

template <class real, int MODE>
struct RowAtom {
  static constexpr int LV = MODE == 0 ? 2 : 3;
  static constexpr int NF = MODE == 0 ? 6 : (MODE == 1 ? 10 : 13);
  real w[3][LV][kRowTile];  // [axis][derivative][tile point], zero outside the support
  real f[NF];
  int off[3];  // support base relative to the tile origin per axis (-4..7)
};

template <class real, int MODE>
__global__ __launch_bounds__(256) void k_spread_rows(TileGeom g, const int* __restrict__ bin_start,
                                                     const int* __restrict__ bin_atoms,
                                                     const R4<real>* __restrict__ pw, const int4* __restrict__ pbase,
                                                     const real* __restrict__ src1, const real* __restrict__ src2,
                                                     real a1, real a2, real a3, cplx_t<real>* __restrict__ grid,
                                                     Gate gate) {
  if (gate_closed(gate)) return;
  using RA = RowAtom<real, MODE>;
  constexpr int LV = RA::LV;
  constexpr int NACC = MODE == 1 ? 1 : 2;
  // chunks of 64 atoms keep the staging area small enough for more
  // workgroups per CU: measured faster for multipoles and for FP64 dipoles
  // (FP64 step 6 % faster).  FP32 dipoles keep 128: with 64 the spread
  // alone is faster (66 -> 57 us) but the FP32 step time is the same, the
  // earlier spread only overlapping more of the FFT with the mutual field
  constexpr int CH = (MODE != 0 || sizeof(real) == 8) ? kRowChunk : 2 * kRowChunk;
  static_assert(CH % 64 == 0, "the sub-tile lists are built by whole wavefronts");
  constexpr size_t kStage = sizeof(RA) * CH, kPart = sizeof(real) * 2 * 64 * NACC * 4;
  __shared__ __align__(16) char smem[kStage > kPart ? kStage : kPart];
  RA* atoms = reinterpret_cast<RA*>(smem);
  __shared__ int bstart[27], bpref[28];
  __shared__ unsigned char lists[8][CH];
  __shared__ int wcnt[CH / 64][8];
  const int tt = blockIdx.x;  // tiles in launch order (an XCD-contiguous order measured slower)
  const int t1 = tt % g.nt[0], t2 = (tt / g.nt[0]) % g.nt[1], t3 = tt / (g.nt[0] * g.nt[1]);
  const int lo[3] = {t1 * kRowTile, t2 * kRowTile, t3 * kRowTile};
  const int nb[3] = {g.K[0] / kRowBin, g.K[1] / kRowBin, g.K[2] / kRowBin};
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid < 64) {  // the 27 bins in a fixed order, and their exclusive prefix
    int cnt = 0, st = 0;
    if (tid < 27) {
      int b1 = 2 * t1 + tid % 3 - 1, b2 = 2 * t2 + (tid / 3) % 3 - 1, b3 = 2 * t3 + tid / 9 - 1;
      b1 += b1 < 0 ? nb[0] : 0;
      b2 += b2 < 0 ? nb[1] : 0;
      b3 += b3 < 0 ? nb[2] : 0;
      const int bin = b1 + nb[0] * (b2 + nb[1] * b3);
      st = bin_start[bin];
      cnt = bin_start[bin + 1] - st;
      bstart[tid] = st;
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    if (tid < 27) bpref[tid + 1] = incl;
    if (tid == 0) bpref[0] = 0;
  }
  __syncthreads();
  const int total = bpref[27];
  real acc[NACC][4];
#pragma unroll
  for (int q = 0; q < NACC; ++q)
#pragma unroll
    for (int x = 0; x < 4; ++x) acc[q][x] = 0;
  for (int base = 0; base < total; base += CH) {
    const int nchunk = min(CH, total - base);
    // ---- stage: one thread per atom
    if (tid < nchunk) {
      const int idx = base + tid;
      int c = 0;  // bin of idx: last c with bpref[c] <= idx
#pragma unroll
      for (int step = 16; step >= 1; step >>= 1)
        if (c + step < 27 && bpref[c + step] <= idx) c += step;
      const int s = bin_atoms[bstart[c] + idx - bpref[c]];
      RA& A = atoms[tid];
      const int4 b = pbase[s];
      const int bb[3] = {b.x, b.y, b.z};
      const R4<real> w = pw[s];
      const real wf[3] = {w.x, w.y, w.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        int o = bb[d] - lo[d];
        if (o >= kRowTile) o -= g.K[d];  // lower neighbour bin across the periodic boundary
        A.off[d] = o;
        real th[LV][5];
        bspline5<LV>(wf[d], th);
        // the row: zeros, then the support points that fall on the tile
        // (stores at a computed offset instead of selects per point)
#pragma unroll
        for (int lv = 0; lv < LV; ++lv) {
#pragma unroll
          for (int x = 0; x < kRowTile; ++x) A.w[d][lv][x] = 0;
#pragma unroll
          for (int q = 0; q < 5; ++q)
            if ((unsigned)(o + q) < (unsigned)kRowTile) A.w[d][lv][o + q] = th[lv][q];
        }
      }
      if (MODE == 0) {
        A.f[0] = src1[3 * s] * a1;
        A.f[1] = src1[3 * s + 1] * a2;
        A.f[2] = src1[3 * s + 2] * a3;
        A.f[3] = src2[3 * s] * a1;
        A.f[4] = src2[3 * s + 1] * a2;
        A.f[5] = src2[3 * s + 2] * a3;
      } else {
        const real* p = src1 + 10 * (size_t)s;
        A.f[0] = p[0];
        A.f[1] = p[1] * a1;
        A.f[2] = p[2] * a2;
        A.f[3] = p[3] * a3;
        A.f[4] = p[4] * a1 * a1;
        A.f[5] = p[5] * a2 * a2;
        A.f[6] = p[6] * a3 * a3;
        A.f[7] = 2 * p[7] * a1 * a2;
        A.f[8] = 2 * p[8] * a1 * a3;
        A.f[9] = 2 * p[9] * a2 * a3;
        if (MODE == 2) {
          A.f[MODE == 2 ? 10 : 0] = src2[3 * s] * a1;
          A.f[MODE == 2 ? 11 : 0] = src2[3 * s + 1] * a2;
          A.f[MODE == 2 ? 12 : 0] = src2[3 * s + 2] * a3;
        }
      }
    }
    __syncthreads();
    // ---- the staged atoms whose support reaches each 4^3 sub-tile, in
    // staging order (wave ballots, offsets of the earlier staging waves)
    unsigned long long bal[8];
    if (tid < CH) {
      unsigned m = 0;
      if (tid < nchunk) {
        const RA& A = atoms[tid];
#pragma unroll
        for (int st = 0; st < 8; ++st) {
          const int sx = (st & 1) * 4, sy = ((st >> 1) & 1) * 4, sz = (st >> 2) * 4;
          const bool in = A.off[0] <= sx + 3 && A.off[0] + 4 >= sx && A.off[1] <= sy + 3 && A.off[1] + 4 >= sy &&
                          A.off[2] <= sz + 3 && A.off[2] + 4 >= sz;
          m |= in ? 1u << st : 0u;
        }
      }
#pragma unroll
      for (int st = 0; st < 8; ++st) {
        bal[st] = __ballot((m >> st) & 1);
        if (lane == 0) wcnt[wv][st] = __popcll(bal[st]);
      }
    }
    __syncthreads();
    if (tid < CH) {
#pragma unroll
      for (int st = 0; st < 8; ++st)
        if ((bal[st] >> lane) & 1) {
          int o = __popcll(bal[st] & ((1ull << lane) - 1));
          for (int w = 0; w < wv; ++w) o += wcnt[w][st];
          lists[st][o] = (unsigned char)tid;
        }
    }
    __syncthreads();
    // ---- accumulate: 16-lane group gq of wave wv owns the 16 rows of
    // sub-tile (wv & 1) * 4 + gq; waves wv and wv + 2 split its list (even
    // and odd entries)
    {
      const int gq = lane >> 4, r = lane & 15;
      const int st = (wv & 1) * 4 + gq;
      const int sx = (st & 1) * 4, sy = ((st >> 1) & 1) * 4 + (r & 3), sz = (st >> 2) * 4 + (r >> 2);
      int cnt = 0, cmax = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int c = 0;
        for (int w = 0; w < CH / 64; ++w) c += wcnt[w][(wv & 1) * 4 + q];
        cnt = q == gq ? c : cnt;
        cmax = c > cmax ? c : cmax;
      }
      for (int k = wv >> 1; k < cmax; k += 2) {
        if (k >= cnt) continue;
        const RA& A = atoms[lists[st][k]];
        const real u0 = A.w[1][0][sy], u1 = A.w[1][1][sy];
        const real v0 = A.w[2][0][sz], v1 = A.w[2][1][sz];
        const real* f = A.f;
        if (MODE == 0) {
          const real Ad = f[0] * u0 * v0, Bd = f[1] * u1 * v0 + f[2] * u0 * v1;
          const real Ap = f[3] * u0 * v0, Bp = f[4] * u1 * v0 + f[5] * u0 * v1;
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            const real t0 = A.w[0][0][sx + x], t1v = A.w[0][1][sx + x];
            acc[0][x] += Ad * t1v + Bd * t0;
            acc[NACC - 1][x] += Ap * t1v + Bp * t0;
          }
        } else {
          const real u2 = A.w[1][LV > 2 ? 2 : 0][sy], v2 = A.w[2][LV > 2 ? 2 : 0][sz];
          const real R0 = f[0] * u0 * v0 + f[2] * u1 * v0 + f[3] * u0 * v1 + f[5] * u2 * v0 + f[6] * u0 * v2 +
                          f[9] * u1 * v1;
          const real R1 = f[1] * u0 * v0 + f[7] * u1 * v0 + f[8] * u0 * v1;
          const real R2 = f[4] * u0 * v0;
          real Ap = 0, Bp = 0;
          if (MODE == 2) {
            Ap = f[MODE == 2 ? 10 : 0] * u0 * v0;
            Bp = f[MODE == 2 ? 11 : 0] * u1 * v0 + f[MODE == 2 ? 12 : 0] * u0 * v1;
          }
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            const real t0 = A.w[0][0][sx + x], t1v = A.w[0][1][sx + x], t2v = A.w[0][LV > 2 ? 2 : 0][sx + x];
            acc[0][x] += R0 * t0 + R1 * t1v + R2 * t2v;
            if (MODE == 2) acc[NACC - 1][x] += Ap * t1v + Bp * t0;
          }
        }
      }
    }
    __syncthreads();
  }
  // ---- waves 2 and 3 hand their rows to waves 0 and 1, which add them in
  // that order and write every point of their sub-tiles once
  real* part = reinterpret_cast<real*>(smem);  // [2][64][NACC * 4] (the staging area is free)
  if (wv >= 2) {
#pragma unroll
    for (int q = 0; q < NACC; ++q)
#pragma unroll
      for (int x = 0; x < 4; ++x) part[((wv - 2) * 64 + lane) * (NACC * 4) + q * 4 + x] = acc[q][x];
  }
  __syncthreads();
  if (wv < 2) {
#pragma unroll
    for (int q = 0; q < NACC; ++q)
#pragma unroll
      for (int x = 0; x < 4; ++x) acc[q][x] += part[(wv * 64 + lane) * (NACC * 4) + q * 4 + x];
    const int gq = lane >> 4, r = lane & 15;
    const int st = (wv & 1) * 4 + gq;
    const int sx = (st & 1) * 4, sy = ((st >> 1) & 1) * 4 + (r & 3), sz = (st >> 2) * 4 + (r >> 2);
    const long long row = lo[0] + sx + (long long)g.K[0] * ((lo[1] + sy) + (long long)g.K[1] * (lo[2] + sz));
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      cplx_t<real> v;
      v.x = acc[0][x];
      v.y = NACC == 2 ? acc[NACC - 1][x] : (real)0;
      grid[row + x] = v;
    }
  }
}

// ------------------------------------------------------------ gathers

// Field gathers run one thread per atom in 64-thread workgroups: 1,500
// workgroups at 96,000 atoms spread over the 256 CUs more evenly than 375 of
// 256 threads (the dipole gather 24.4 -> 19.9 us in the MD step, the step
// 1.2 % faster), and each XCD's range of sorted atoms stays contiguous.
constexpr int kGatherBlock = 64;

// Field of the dipole grid at each atom: fd -= a * grad phi (real part),
// fp likewise from the imaginary part, plus the self term term * u.
template <class real>
__global__ __launch_bounds__(256) void k_gather_dip(int n, TileGeom g, const R4<real>* __restrict__ pw,
                                                    const int4* __restrict__ pbase,
                                                    const cplx_t<real>* __restrict__ grid, real a1, real a2, real a3,
                                                    real term, const real* __restrict__ ud,
                                                    const real* __restrict__ up, real* __restrict__ fd,
                                                    real* __restrict__ fp, Gate gate) {
  if (gate_closed(gate)) return;
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const R4<real> w = pw[s];
  const int4 b = pbase[s];
  real th1[2][5], th2[2][5], th3[2][5];
  bspline5<2>(w.x, th1);
  bspline5<2>(w.y, th2);
  bspline5<2>(w.z, th3);
  real d1 = 0, d2 = 0, d3 = 0, p1 = 0, p2 = 0, p3 = 0;
  for (int m3 = 0; m3 < 5; ++m3) {
    int z = b.z + m3;
    if (z >= g.K[2]) z -= g.K[2];
    const real v0 = th3[0][m3], v1 = th3[1][m3];
    for (int m2 = 0; m2 < 5; ++m2) {
      int y = b.y + m2;
      if (y >= g.K[1]) y -= g.K[1];
      const real u0 = th2[0][m2], u1 = th2[1][m2];
      const long long row = (long long)g.K[0] * (y + (long long)g.K[1] * z);
      for (int m1 = 0; m1 < 5; ++m1) {
        int x = b.x + m1;
        if (x >= g.K[0]) x -= g.K[0];
        const cplx_t<real> q = grid[row + x];
        const real t0 = th1[0][m1], t1 = th1[1][m1];
        const real w1 = t1 * u0 * v0, w2 = t0 * u1 * v0, w3 = t0 * u0 * v1;
        d1 += q.x * w1;
        d2 += q.x * w2;
        d3 += q.x * w3;
        p1 += q.y * w1;
        p2 += q.y * w2;
        p3 += q.y * w3;
      }
    }
  }
  fd[3 * s] += term * ud[3 * s] - a1 * d1;
  fd[3 * s + 1] += term * ud[3 * s + 1] - a2 * d2;
  fd[3 * s + 2] += term * ud[3 * s + 2] - a3 * d3;
  fp[3 * s] += term * up[3 * s] - a1 * p1;
  fp[3 * s + 1] += term * up[3 * s + 1] - a2 * p2;
  fp[3 * s + 2] += term * up[3 * s + 2] - a3 * p3;
}

// Field of the permanent-multipole grid (real part) plus the self term,
// added to both the d and the p field.
template <class real>
__global__ __launch_bounds__(256) void k_gather_mpole_field(int n, TileGeom g, const R4<real>* __restrict__ pw,
                                                            const int4* __restrict__ pbase,
                                                            const cplx_t<real>* __restrict__ grid, real a1, real a2,
                                                            real a3, real term, const real* __restrict__ rpole,
                                                            real* __restrict__ fd, real* __restrict__ fp) {
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const R4<real> w = pw[s];
  const int4 b = pbase[s];
  real th1[2][5], th2[2][5], th3[2][5];
  bspline5<2>(w.x, th1);
  bspline5<2>(w.y, th2);
  bspline5<2>(w.z, th3);
  real d1 = 0, d2 = 0, d3 = 0;
  for (int m3 = 0; m3 < 5; ++m3) {
    int z = b.z + m3;
    if (z >= g.K[2]) z -= g.K[2];
    const real v0 = th3[0][m3], v1 = th3[1][m3];
    for (int m2 = 0; m2 < 5; ++m2) {
      int y = b.y + m2;
      if (y >= g.K[1]) y -= g.K[1];
      const real u0 = th2[0][m2], u1 = th2[1][m2];
      const long long row = (long long)g.K[0] * (y + (long long)g.K[1] * z);
      for (int m1 = 0; m1 < 5; ++m1) {
        int x = b.x + m1;
        if (x >= g.K[0]) x -= g.K[0];
        const real q = grid[row + x].x;
        const real t0 = th1[0][m1], t1 = th1[1][m1];
        d1 += q * t1 * u0 * v0;
        d2 += q * t0 * u1 * v0;
        d3 += q * t0 * u0 * v1;
      }
    }
  }
  const real* p = rpole + 10 * (size_t)s;
  const real f1 = term * p[1] - a1 * d1, f2 = term * p[2] - a2 * d2, f3 = term * p[3] - a3 * d3;
  fd[3 * s] += f1;
  fd[3 * s + 1] += f2;
  fd[3 * s + 2] += f3;
  fp[3 * s] += f1;
  fp[3 * s + 1] += f2;
  fp[3 * s + 2] += f3;
}

// ------------------------------------------------------------ host side

TileGeom geom(const Ctx& c) {
  TileGeom g;
  for (int d = 0; d < 3; ++d) {
    g.K[d] = c.grid[d];
    g.T[d] = c.tsz[d];
    g.nt[d] = c.ntile[d];
  }
  return g;
}

template <class real>
void pme_bin_t(Ctx& c, hipStream_t st) {
  const int n = c.n;
  const int ntot = c.nbin[0] * c.nbin[1] * c.nbin[2];
  const int nchunk = ceil_div(ntot, kScanChunk);
  // tile_count is zero here: cleared at allocation and by every scan
  hipLaunchKernelGGL(k_pme_bin<real>, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, c.perm.as<int>(),
                     c.xyz.as<double>(), c.box[0], c.box[1], c.box[2], c.grid[0], c.grid[1], c.grid[2], c.bsz[0],
                     c.bsz[1], c.bsz[2], c.nbin[0], c.nbin[1], c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(),
                     c.tile_count.as<int>());
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_sums, dim3(nchunk), dim3(kScanChunk), 0, st, c.tile_count.as<int>(),
                     c.tile_csum.as<int>(), ntot);
  hipLaunchKernelGGL(k_scan_apply, dim3(nchunk), dim3(kScanChunk), 0, st, c.tile_count.as<int>(),
                     c.tile_start.as<int>(), c.tile_cursor.as<int>(), c.tile_csum.as<int>(), ntot);
  T9_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tile_scatter, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, c.pme_base.as<int4>(),
                     c.tile_cursor.as<int>(), c.tile_atoms.as<int>());
  T9_LAUNCH_CHECK();
  // the fixed-point tile sums do not depend on the order within a bin; the
  // row-owner spreading adds atoms in bin order
  if (c.spread_rows) {
    hipLaunchKernelGGL(k_bin_sort, dim3(ceil_div(ntot, 256)), dim3(256), 0, st, ntot, c.tile_start.as<int>(),
                       c.tile_atoms.as<int>());
    T9_LAUNCH_CHECK();
  }
  c.bins_valid = true;
}

// bins of the current coordinates, enqueued on the main stream if no
// binning is valid, and the main stream's wait for an asynchronous one
template <class real>
void ensure_bins(Ctx& c) {
  if (!c.bins_valid) pme_bin_t<real>(c, c.stream);
  if (c.bins_pending) {
    T9_HIP(hipStreamWaitEvent(c.stream, c.ev_bins, 0));
    c.bins_pending = false;
  }
}

template <class real>
void spread_t(Ctx& c, int mode, const void* s1, const void* s2) {
  const int ntot = c.ntile[0] * c.ntile[1] * c.ntile[2];
  const real a1 = (real)(c.grid[0] / c.box[0]), a2 = (real)(c.grid[1] / c.box[1]), a3 = (real)(c.grid[2] / c.box[2]);
  if (c.spread_rows) {
    auto kern = mode == 0 ? k_spread_rows<real, 0> : (mode == 1 ? k_spread_rows<real, 1> : k_spread_rows<real, 2>);
    KScope ks_(c, KT_SPREAD);
    hipLaunchKernelGGL(kern, dim3(ntot), dim3(256), 0, c.stream, geom(c), c.tile_start.as<int>(),
                       c.tile_atoms.as<int>(), c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), (const real*)s1,
                       (const real*)s2, a1, a2, a3, c.qgrid.as<cplx_t<real>>(), c.gate);
    T9_LAUNCH_CHECK();
    return;
  }
  const int rows = c.tsz[1] * c.tsz[2];
  if (rows > 256 || c.tsz[0] > kMaxTile) throw ArgError("spreading tile too large");
  const size_t per = mode == 0 ? sizeof(SpreadAtom<real, 0>)
                               : (mode == 1 ? sizeof(SpreadAtom<real, 1>) : sizeof(SpreadAtom<real, 2>));
  const size_t accb =
      sizeof(unsigned long long) * (mode == 1 ? 1 : 2) * (size_t)(c.tsz[0] + 1) * (c.tsz[1] + 1) * c.tsz[2];
  const size_t budget = 65536 - 1024;  // dynamic LDS next to the kernel's static tables
  if (accb + 32 * (per + sizeof(int)) > budget) throw ArgError("spreading tile too large for LDS");
  const int chunk = (int)std::min<size_t>(kSpreadChunk, (budget - accb - 2 * sizeof(int)) / (per + sizeof(int)));
  const size_t sh = accb + sizeof(int) * ((chunk + 2) & ~1) + per * chunk;
  auto kern = mode == 0 ? k_spread<real, 0> : (mode == 1 ? k_spread<real, 1> : k_spread<real, 2>);
  KScope ks_(c, KT_SPREAD);
  hipLaunchKernelGGL(kern, dim3(ntot), dim3(256), sh, c.stream, geom(c), c.tile_start.as<int>(),
                     c.tile_atoms.as<int>(), c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(),
                     (const real*)s1, (const real*)s2, a1, a2, a3, chunk,
                     c.qgrid.as<cplx_t<real>>(), c.gate);
  T9_LAUNCH_CHECK();
}

template <class real>
void dipole_field_t(Ctx& c, const void* ud, const void* up, void* fd, void* fp, hipEvent_t before_gather) {
  ensure_bins<real>(c);
  spread_t<real>(c, 0, ud, up);
  pme_fft_roundtrip(c);
  if (before_gather) T9_HIP(hipStreamWaitEvent(c.stream, before_gather, 0));
  const real a1 = (real)(c.grid[0] / c.box[0]), a2 = (real)(c.grid[1] / c.box[1]), a3 = (real)(c.grid[2] / c.box[2]);
  const real term = (real)(4.0 / 3.0 * c.aewald * c.aewald * c.aewald / 1.77245385090551602730);
  KScope ks_(c, KT_GATHER);
  hipLaunchKernelGGL(k_gather_dip<real>, dim3(ceil_div(c.n, kGatherBlock)), dim3(kGatherBlock), 0, c.stream, c.n, geom(c),
                     c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), c.qgrid.as<cplx_t<real>>(), a1, a2, a3, term,
                     (const real*)ud, (const real*)up, (real*)fd, (real*)fp, c.gate);
  T9_LAUNCH_CHECK();
}

// direct-field round trip of the multipoles and a dipole round trip on the
// other grid buffer, both transformed before either gather waits for the
// real-space kernels of the third stream (the multipoles' round trip in line
// passes: a plane workgroup needs a whole CU's LDS while the other streams'
// kernels hold every CU)
template <class real>
void mpole_dipole_fields_t(Ctx& c, void* fd, void* fp, hipEvent_t e1, const void* ud, const void* up, void* gd,
                           void* gp, hipEvent_t e2) {
  ensure_bins<real>(c);
  spread_t<real>(c, 1, c.rpole.p, nullptr);
  pme_fft_roundtrip(c, true);
  c.qgrid_m.alloc(c.qgrid.bytes);
  c.qgrid.swap(c.qgrid_m);
  c.qgrid_m_stamp = c.rpole_stamp;
  spread_t<real>(c, 0, ud, up);
  // by now the list and direct-field work of the other streams has mostly
  // drained: plane kernels (about 1 % faster per step than line passes)
  pme_fft_roundtrip(c, false);
  const real a1 = (real)(c.grid[0] / c.box[0]), a2 = (real)(c.grid[1] / c.box[1]), a3 = (real)(c.grid[2] / c.box[2]);
  const real term = (real)(4.0 / 3.0 * c.aewald * c.aewald * c.aewald / 1.77245385090551602730);
  T9_HIP(hipStreamWaitEvent(c.stream, e1, 0));
  {
    KScope ks_(c, KT_GATHER);
    hipLaunchKernelGGL(k_gather_mpole_field<real>, dim3(ceil_div(c.n, kGatherBlock)), dim3(kGatherBlock), 0, c.stream, c.n, geom(c),
                       c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), c.qgrid_m.as<cplx_t<real>>(), a1, a2, a3, term,
                       c.rpole.as<real>(), (real*)fd, (real*)fp);
    T9_LAUNCH_CHECK();
  }
  T9_HIP(hipStreamWaitEvent(c.stream, e2, 0));
  KScope ks_(c, KT_GATHER);
  hipLaunchKernelGGL(k_gather_dip<real>, dim3(ceil_div(c.n, kGatherBlock)), dim3(kGatherBlock), 0, c.stream, c.n, geom(c),
                     c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), c.qgrid.as<cplx_t<real>>(), a1, a2, a3, term,
                     (const real*)ud, (const real*)up, (real*)gd, (real*)gp, Gate{});
  T9_LAUNCH_CHECK();
}

template <class real>
void mpole_field_t(Ctx& c, void* fd, void* fp, hipEvent_t before_gather) {
  ensure_bins<real>(c);
  spread_t<real>(c, 1, c.rpole.p, nullptr);
  // the direct field's round trip runs beside the real-space direct field and
  // the vdW pair kernel of the other streams, which hold every CU: a plane
  // workgroup (a whole CU's LDS) would wait for them to drain, the line
  // passes' small workgroups fit in between
  pme_fft_roundtrip(c, before_gather != nullptr);
  // keep the convolved grid: the reciprocal multipole energy of the same
  // coordinates gathers from it instead of spreading and transforming again.
  // The two grid buffers trade places (no copy); the dipole round trips use
  // the other one.
  c.qgrid_m.alloc(c.qgrid.bytes);
  c.qgrid.swap(c.qgrid_m);
  c.qgrid_m_stamp = c.rpole_stamp;
  const real a1 = (real)(c.grid[0] / c.box[0]), a2 = (real)(c.grid[1] / c.box[1]), a3 = (real)(c.grid[2] / c.box[2]);
  const real term = (real)(4.0 / 3.0 * c.aewald * c.aewald * c.aewald / 1.77245385090551602730);
  if (before_gather) T9_HIP(hipStreamWaitEvent(c.stream, before_gather, 0));
  KScope ks_(c, KT_GATHER);
  hipLaunchKernelGGL(k_gather_mpole_field<real>, dim3(ceil_div(c.n, kGatherBlock)), dim3(kGatherBlock), 0, c.stream, c.n, geom(c),
                     c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), c.qgrid_m.as<cplx_t<real>>(), a1, a2, a3, term,
                     c.rpole.as<real>(), (real*)fd, (real*)fp);
  T9_LAUNCH_CHECK();
}

}  // namespace

void pme_bin(Ctx& c) {
  if (c.prec == 32) ensure_bins<float>(c);
  else ensure_bins<double>(c);
}

// Binning of the current coordinates on side3, beside the list compaction
// and the multipole rotation of the main stream; the first PME kernel of the
// main stream waits for it (ensure_bins).
void pme_bin_async(Ctx& c) {
  if (!c.ewald || c.bins_valid) return;
  T9_HIP(hipEventRecord(c.ev_bfork, c.stream));
  T9_HIP(hipStreamWaitEvent(c.side3, c.ev_bfork, 0));
  if (c.prec == 32) pme_bin_t<float>(c, c.side3);
  else pme_bin_t<double>(c, c.side3);
  T9_HIP(hipEventRecord(c.ev_bins, c.side3));
  c.bins_pending = true;
}

void pme_spread(Ctx& c, int mode, const void* s1, const void* s2) {
  if (c.prec == 32) {
    ensure_bins<float>(c);
    spread_t<float>(c, mode, s1, s2);
  } else {
    ensure_bins<double>(c);
    spread_t<double>(c, mode, s1, s2);
  }
}

void pme_dipole_field(Ctx& c, const void* ud, const void* up, void* fd, void* fp, hipEvent_t before_gather) {
  if (c.prec == 32) dipole_field_t<float>(c, ud, up, fd, fp, before_gather);
  else dipole_field_t<double>(c, ud, up, fd, fp, before_gather);
}

void pme_mpole_dipole_fields(Ctx& c, void* fd, void* fp, hipEvent_t e1, const void* ud, const void* up, void* gd,
                             void* gp, hipEvent_t e2) {
  if (c.prec == 32) mpole_dipole_fields_t<float>(c, fd, fp, e1, ud, up, gd, gp, e2);
  else mpole_dipole_fields_t<double>(c, fd, fp, e1, ud, up, gd, gp, e2);
}

void pme_mpole_field(Ctx& c, void* fd, void* fp, hipEvent_t before_gather) {
  if (c.prec == 32) mpole_field_t<float>(c, fd, fp, before_gather);
  else mpole_field_t<double>(c, fd, fp, before_gather);
}

// PME sub-steps on the context grid (the boundary's gridMpole / gridUind /
// fphiMpole / fphiUind, include/ff/pme.h:64-91).  The gathers return the
// Cartesian field -grad phi of the grid at each sorted atom (real part into
// fd, imaginary part into fp; the multipole grid's real part into both),
// without the self term.
void pme_grid_mpole(Ctx& c) { pme_spread(c, 1, c.rpole.p, nullptr); }

void pme_grid_uind(Ctx& c, const void* ud, const void* up) { pme_spread(c, 0, ud, up); }

namespace {
template <class real>
void fphi_t(Ctx& c, bool mpole, void* fd, void* fp) {
  ensure_bins<real>(c);
  const size_t vb = sizeof(real) * 3 * (size_t)c.n;
  T9_HIP(hipMemsetAsync(fd, 0, vb, c.stream));
  T9_HIP(hipMemsetAsync(fp, 0, vb, c.stream));
  const real a1 = (real)(c.grid[0] / c.box[0]), a2 = (real)(c.grid[1] / c.box[1]), a3 = (real)(c.grid[2] / c.box[2]);
  if (mpole) {
    hipLaunchKernelGGL(k_gather_mpole_field<real>, dim3(ceil_div(c.n, kGatherBlock)), dim3(kGatherBlock), 0, c.stream, c.n, geom(c),
                       c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), c.qgrid.as<cplx_t<real>>(), a1, a2, a3,
                       (real)0, c.rpole.as<real>(), (real*)fd, (real*)fp);
  } else {
    // the self term is off (term 0); its dipole inputs read the zeroed outputs
    T9_HIP(hipMemsetAsync(c.zrsd.p, 0, vb, c.stream));
    hipLaunchKernelGGL(k_gather_dip<real>, dim3(ceil_div(c.n, kGatherBlock)), dim3(kGatherBlock), 0, c.stream, c.n, geom(c),
                       c.pme_w.as<R4<real>>(), c.pme_base.as<int4>(), c.qgrid.as<cplx_t<real>>(), a1, a2, a3,
                       (real)0, c.zrsd.as<real>(), c.zrsd.as<real>(), (real*)fd, (real*)fp, Gate{});
  }
  T9_LAUNCH_CHECK();
}
}  // namespace

void pme_fphi(Ctx& c, bool mpole, void* fd, void* fp) {
  if (c.prec == 32) fphi_t<float>(c, mpole, fd, fp);
  else fphi_t<double>(c, mpole, fd, fp);
}

// B-spline moduli of one axis (the dftmod rule, with the zeta correction),
// FP64 on the host.
std::vector<double> pme_bsmod(int K, int order) {
  auto M = [](int p, double x) {
    // cardinal B-spline by its explicit piecewise sum
    double s = 0, fact = 1;
    for (int k = 1; k < p; ++k) fact *= k;
    for (int k = 0; k <= p; ++k) {
      double t = x - k;
      if (t <= 0) continue;
      double c = 1;
      for (int q = 0; q < k; ++q) c = c * (p - q) / (q + 1);
      s += ((k & 1) ? -c : c) * std::pow(t, p - 1);
    }
    return (x <= 0 || x >= p) ? 0.0 : s / fact;
  };
  std::vector<double> arr(K, 0.0), mod(K, 0.0);
  for (int k = 1; k < order && k < K; ++k) arr[k] = M(order, (double)k);
  for (int m = 0; m < K; ++m) {
    double s1 = 0, s2 = 0;
    for (int j = 0; j < K; ++j) {
      double arg = 2 * kPi * (double)((long long)m * j % K) / K;
      s1 += arr[j] * std::cos(arg);
      s2 += arr[j] * std::sin(arg);
    }
    mod[m] = s1 * s1 + s2 * s2;
  }
  const double eps = 1e-7;
  if (mod[0] < eps) mod[0] = 0.5 * mod[1];
  for (int m = 1; m < K - 1; ++m)
    if (mod[m] < eps) mod[m] = 0.5 * (mod[m - 1] + mod[m + 1]);
  if (mod[K - 1] < eps) mod[K - 1] = 0.5 * mod[K - 2];
  for (int m = 0; m < K; ++m) {
    int k = (m + 1 > K / 2) ? m - K : m;
    double zeta = 1.0;
    if (k != 0) {
      double s1 = 1, s2 = 1, fac = kPi * k / K;
      for (int j = 1; j <= 50; ++j) {
        double a1 = fac / (fac + kPi * j), a2 = fac / (fac - kPi * j);
        s1 += std::pow(a1, order) + std::pow(a2, order);
        s2 += std::pow(a1, 2 * order) + std::pow(a2, 2 * order);
      }
      zeta = s2 / s1;
    }
    mod[m] *= zeta * zeta;
  }
  return mod;
}

// Spreading geometry: row-owner spreading (8^3 tiles over 4^3 bins) when every
// grid size is a multiple of 8 and at least 16 (T9HIP_SPREAD=fixed selects the
// LDS fixed-point kernel), else bins equal to the tiles of the fixed-point
// kernel.
void pme_setup_bins(Ctx& c) {
  const char* e = std::getenv("T9HIP_SPREAD");
  bool rows = !(e && !std::strcmp(e, "fixed"));
  for (int d = 0; d < 3; ++d) rows = rows && c.grid[d] % kRowTile == 0 && c.grid[d] >= 2 * kRowTile;
  c.spread_rows = rows;
  for (int d = 0; d < 3; ++d) {
    if (rows) {
      c.tsz[d] = kRowTile;
      c.ntile[d] = c.grid[d] / kRowTile;
    }
    c.bsz[d] = rows ? kRowBin : c.tsz[d];
    c.nbin[d] = c.grid[d] / c.bsz[d];
  }
}

// Tile edge dividing K, closest to 8, between 4 and 12 (K <= 12: one tile);
// the fixed-point copy of a tile must fit in LDS next to the staged atoms.
int pme_tile_edge(int K) {
  if (K <= 12) return K;
  int best = -1;
  for (int t = 4; t <= 12; ++t)
    if (K % t == 0 && (best < 0 || std::abs(t - 8) < std::abs(best - 8))) best = t;
  if (best < 0) throw ArgError("PME grid size " + std::to_string(K) + " has no divisor between 4 and 12");
  return best;
}

}  // namespace t9
