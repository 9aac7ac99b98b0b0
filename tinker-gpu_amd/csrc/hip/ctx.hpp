// Device context: everything one system needs on one GPU.
//
// Layout in HBM.  Atoms live in two orders:
//  * caller order ("orig"): FP64 coordinates d_xyz[n][3] and the static
//    parameters exactly as the host setup produced them;
//  * spatial order ("sorted"): the order of the last neighbor-list build
//    (cells in Morton order, original index inside a cell).  All per-step
//    working arrays — wrapped coordinates, global multipoles, dipoles,
//    fields, CG vectors — are in this order, [n][3] or [n][10] of `real`,
//    so the pair kernels' lanes read consecutive atoms.
// perm[s] = original index of sorted slot s; iperm is its inverse.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <random>
#include <vector>

#include "common.hpp"
#include "host/system.hpp"

namespace t9 {

// A device allocation that frees itself.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t nbytes) {
    if (nbytes <= bytes && p) return;
    release();
    if (nbytes == 0) nbytes = 16;
    T9_HIP(hipMalloc(&p, nbytes));
    bytes = nbytes;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  void swap(DBuf& o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
};

// Neighbor list in "wave slabs": atoms are grouped 64 to a wavefront in
// sorted order; wave w owns entries [off[w], off[w] + 64*maxlen[w]) laid out
// k-major (entry k of lane l at off[w] + 64*k + l) so a wave's lanes read one
// contiguous 256-byte row per neighbor step.  Entry = (j << 5) | scale class.
struct NbList {
  double cutoff = 0;   // list radius (interaction cutoff + buffer)
  DBuf nlist;          // int [n]     neighbors of each sorted atom
  DBuf maxlen;         // int [nwave]
  DBuf off;            // int64 [nwave + 1]
  DBuf slots;          // int64 [nwave] slab sizes (build scratch)
  DBuf entries;        // uint32 [capacity]
  size_t capacity = 0;
  long long total = 0;  // slots used (including padding)
  long long pairs = 0;  // directed pairs stored
  // Per-step compaction: the entries of each atom that lie within the
  // interaction cutoff on the current coordinates, same slab layout (off).
  double cut = 0;      // interaction cutoff used for the compaction
  DBuf ncut;           // int [n]
  DBuf ecut;           // uint32 [capacity]
  bool cut_valid = false;
  // optional second compaction at a shorter radius (same slab layout): the
  // preconditioner's pairs, taken from the multipole list
  double cut2nd = 0;
  DBuf ncut2, ecut2;
};

// Live kernel timing with HIP events recorded on the context stream around
// selected launches (enabled per kernel id; see t9_kernel_timing).
enum KernelId : int {
  KT_UFIELD_REAL = 0,
  KT_DFIELD_REAL,
  KT_SPREAD,
  KT_FFT,
  KT_GATHER,
  KT_PRECOND,
  KT_PCG_VEC,
  KT_EMPLAR,
  KT_VDW,
  KT_VALENCE,
  KT_COUNT
};

struct KTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;  // pairs (start, stop)
  std::vector<int> nl;         // launches inside each pair
  size_t used = 0;             // events in use
  double ms = 0;
  long long count = 0;
};

struct Ctx {
  int n = 0, nwave = 0;
  KTimer timers[KT_COUNT];
  int prec = 32;
  hipStream_t stream = nullptr;
  // second stream (default priority): the vdW pair kernel runs there while
  // the induced dipoles are solved on `stream`; joined by events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // third stream (default priority): the real-space mutual field runs there
  // while the PME part of the same field spreads and transforms
  hipStream_t side2 = nullptr;
  hipStream_t side3 = nullptr;   // PME binning beside the list compaction
  hipEvent_t ev_bfork = nullptr, ev_bins = nullptr;
  hipEvent_t ev_ufork = nullptr, ev_ujoin = nullptr, ev_ujoin2 = nullptr;
  hipEvent_t ev_read = nullptr;  // marks a readback the host waits for by polling
  bool vdw_pending = false;  // vdW pair kernel enqueued on `side`, not yet joined
  Options opt;
  double box[3] = {0, 0, 0};
  bool periodic = false;
  double vorg[3] = {0, 0, 0};    // origin and edges of the fixed-point frame: the box,
  double vlen[3] = {1, 1, 1};    // or a virtual box around a non-periodic system
  bool ewald = false;
  double aewald = 0;
  double cut = 0;        // real-space cutoff of the multipole/polar pair terms
  int grid[3] = {0, 0, 0};
  int nj = 0;            // polarize types (Thole table is nj x nj)
  bool uniform_thole = false;  // every pair shares one Thole value
  double thole_value = 0;
  ScaleTable scale{};    // m, d, p, u per scale class
  int ncodes = 1;

  // ---- caller order, FP64
  DBuf xyz, xyz_last;            // [n][3]
  DBuf pole;                     // [n][10] local frame
  DBuf frame, zaxis, xaxis, yaxis, ysign;  // int [n]
  DBuf polarity_o, pdamp_o;      // double [n]
  // HIPPO charge penetration (caller order, double [n][3]: core charge,
  // valence charge, valence exponent); chgpen selects the damped multipole
  // pair kernel
  bool chgpen = false;
  DBuf chgpen_o;
  // HIPPO charge transfer: caller order [n][2] (charge, damping exponent),
  // evaluated over the multipole list within ct_cut, tapered from ct_taper
  bool use_chgtrn = false;
  double ct_cut = 0, ct_taper = 0;
  // HIPPO dispersion: caller order [n][2] (C6, damping exponent), over the
  // multipole list within disp_cut, tapered from disp_taper
  bool use_disp = false;
  double disp_cut = 0, disp_taper = 0;
  DBuf disp_o;
  // HIPPO Pauli repulsion: caller order [n][3] (size, damping exponent,
  // valence count), over the multipole list within rep_cut, tapered from
  // rep_taper
  bool use_repel = false;
  double rep_cut = 0, rep_taper = 0;
  DBuf repel_o;
  double list_cut = 0;  // largest cutoff of the terms on the multipole list
  DBuf chgtrn_o;
  DBuf jpolar_o;                 // int [n]
  DBuf excl_off, excl_j, excl_code;  // CSR of scaled partners (both directions)
  DBuf thlval;                   // real [nj*nj]

  // ---- spatial order
  DBuf perm, iperm;              // int [n]
  DBuf cellc;                    // int [n] packed cell coordinates of sorted atoms
  DBuf pos;                      // PosRec<real> [n] (fixed-point FP32 / wrapped FP64) + 1/Thole radius
  DBuf rpole;                    // real [n][10] global multipoles
  DBuf polarity, polinv, pdamp;  // real [n]
  DBuf jpolar;                   // int [n]

  // ---- cell binning
  int ncell[3] = {1, 1, 1};
  int morton_bits = 0;
  bool brute = true;             // all-pairs search (small or non-periodic systems)
  DBuf cell_key, cell_count, cell_start, cell_cursor, unsorted;
  NbList mlist, ulist, vlist;    // multipole/polar (+ preconditioner compaction), unused, vdW sites
  double cell_reach = 0;         // radius the cell search must cover (A)
  int builds = 0;
  bool list_valid = false;

  // ---- CG / field vectors, real [n][3] each, sorted order
  DBuf field, fieldp, udir, udirp, uind, uinp, rsd, rsdp, zrsd, zrsdp, conj, conjp, vec, vecp;
  DBuf partials;                 // double [nblocks * 8]
  Gate gate;                     // set while a CG iteration is enqueued (early exit once converged)
  int pcg_planned = 0;           // iterations the last solve needed: enqueued before the first readback
  DBuf pcg_trace;                // double [kTraceCap]: RMS residual (Debye) of the last solve per iteration
  int pcg_trace_len = 0;         // entries 0..iterations of the last solve (0: initial guess)
  DBuf scalars;                  // double [32] on device
  DBuf hscalars_pinned;          // host-pinned mirror
  double* h_scalars = nullptr;   // pinned, 48 doubles: CG scalars, and the list check's flag at kDispSlot
  DBuf partials64;               // double [4 * ceil(n / 64)]: CG dots per 64 elements (k_precond)
  DBuf disp_flag;                // int: the last displacement check that found an atom beyond half the buffer
  // list build: per-part counts, the largest part count per list, and the
  // part windows of the staged build (widths from the previous build; 0: none yet)
  DBuf list_parts, list_pmax, stage_m, stage_u, stage_v;
  int list_w[3] = {0, 0, 0};
  int disp_stamp = 0;

  // ---- PME
  int tsz[3] = {8, 8, 8};       // spreading tile edge per axis (divides the grid size)
  int ntile[3] = {1, 1, 1};
  bool spread_rows = false;      // row-owner spreading (8^3 tiles over 4^3 bins) instead of LDS fixed point
  int bsz[3] = {8, 8, 8};       // bin edge per axis (4 with row-owner spreading, else the tile edge)
  int nbin[3] = {1, 1, 1};
  DBuf qgrid;                    // real2 [K1*K2*K3]
  DBuf bsmod;                    // real [K1 + K2 + K3]
  DBuf twiddle;                  // real2 [K1 + K2 + K3] forward tables per axis
  DBuf pme_w, pme_base;          // real4 / int4 [n] per sorted atom
  DBuf tile_count, tile_start, tile_cursor, tile_atoms, tile_csum;  // atoms per bin (support base), bin order
  bool bins_valid = false;
  bool bins_pending = false;     // binning enqueued on side3; the main stream has not waited for it yet
  DBuf conv_table;               // double [K1 + K2 + K3] separable factors of the influence function
  DBuf qgrid_m;                  // convolved permanent-multipole grid of the last direct field
  long long qgrid_m_stamp = -1;  // rpole_stamp it belongs to
  long long rpole_stamp = 0;     // incremented by every rotation of the multipoles
  bool conv_table_valid = false; // false after a box change
  int fft_plane = -1;            // 1: x and y transforms per z-plane in LDS (FP32, square planes); -1 undecided

  // ---- energy / gradient / virial (sorted order, FP64)
  DBuf egrad;                    // double [n][3] dE/dx
  DBuf etq;                      // double [n][9] grad phi, grad grad phi (xx yy zz xy xz yz) = dE/dM
  DBuf eframe;                   // double [n][4][3] torque share of the atom and its frame atoms
  DBuf ecphim;                   // double [n][10] reciprocal potential of the multipoles (orders 0-2)
  DBuf escr;                     // real [n][13] packed-grid sources for the reciprocal virial
  DBuf rev_off, rev_ent;         // caller order: frames naming each atom, (owner << 2) | slot
  DBuf epart;                    // double [ecap][kEnergyRow] per-block partial sums
  int erows = 0, ecap = 0;

  // ---- induced-dipole predictor: the last 16 solutions, caller order
  DBuf ualt;                     // real [16][2][n][3]
  int nualt = 0;                 // solutions stored so far (wraps between 16 and 32)
  int last_predictor = 0;        // 0 direct guess, 1 ASPC, 2 learned
  double last_guess_rms = 0;     // RMS residual of the last initial guess (Debye)
  DBuf frot;                     // float [n][9] sorted: rows = local x, y, z axes (written by rotpole)

  // ---- learned initial guess (guess.hip)
  DBuf ealt;                     // real [16][2][n][3] caller order: direct fields of the last 16 solves
  int nealt = 0;
  DBuf uaspc;                    // real [2][n][3] sorted: ASPC guess of the current solve
  DBuf raspc;                    // real [2][n][3] sorted: field, then residual, of the ASPC guess
  DBuf gfeat;                    // float [n][32] sorted raw features of the current solve
  DBuf gde;                      // float [n][6] caller order: field extrapolation errors (d, p)
  DBuf gp, gm, gv, gpb, ggrad;   // FP32 params / Adam moments, bf16 copy, gradient
  DBuf gpack;                    // bf16 weights in MFMA operand order (inference), rebuilt when stale
  bool gpack_valid = false;
  DBuf gpart;                    // float partial gradients [blocks][nparam + 1]
  DBuf gnorm;                    // float [76]: feature mean, 1/std, target mean, std
  std::vector<float> gnorm_h;
  bool guess_params = false;     // parameters set (loaded or initialized)
  bool guess_ready = false;      // parameters and normalization present
  bool guess_on = false;         // use the learned guess when the history allows
  int adam_t = 0;
  long long guess_used = 0, guess_fallback = 0;
  // training capture: raw features and local-frame targets of every stride-th atom
  DBuf cap_x, cap_y;             // float [cap_max][32], [cap_max][6]
  long long cap_n = 0, cap_max = 0;
  int cap_stride = 0;

  // ---- vdW and valence (MD terms)
  bool use_vdw = false, use_valence = false;
  double vdw_cut = 9, vdw_taper = 8.1;
  DBuf vpos;                     // PosRec<real> [n] vdW sites, sorted order
  DBuf vparent, vred;            // caller order: parent atom, reduction factor (double)
  DBuf jvdw_o, jvdw;             // vdW table index, caller / sorted order
  int njvdw = 0;
  DBuf vrmin, veps;              // real [njvdw][njvdw]
  DBuf vexcl_off, vexcl_j, vexcl_code;  // CSR of vdW-scaled partners, caller order
  double vscale[kMaxCodes] = {};
  DBuf vchild_off, vchild;       // caller order: reduced sites hanging on each atom
  DBuf vsite_grad;               // double [n][3] sorted
  DBuf vterm_off, vterm;         // caller order: valence terms touching each atom
  DBuf vterm_grad;               // double [terms][3][3] per-term gradient of its atoms
  int nbond = 0, nangle = 0, nurey = 0;
  // molecules (connected by valence terms) for the fused RESPA inner loop:
  // atoms and term indices of each molecule (nmol = 0: per-step kernels)
  int nmol = 0;
  bool mol_small = false;  // every molecule <= 4 atoms and <= 6 terms: 8 lanes per molecule
  int respa_path = 2;      // inner loop: 0 a kernel per step, 1 a thread per molecule, 2 lane groups
  DBuf mol_off, mol_atom, mol_toff, mol_term;
  DBuf bond_ij, angle_abc, urey_ac;    // int
  DBuf bond_kr, angle_kt, urey_kr;     // double
  double vpoly[6] = {0, 0, 0, 0, 0, 0};  // bond cubic, quartic; angle cubic..sextic
  // molecular dynamics state, caller order, FP64
  DBuf mass, vel, g_slow, g_fast;
  std::vector<double> mass_h;
  bool md_ready = false;        // g_slow / g_fast hold the forces of the current coordinates
  int nfree = 0;                // degrees of freedom of the thermostat
  std::mt19937_64 md_rng;       // thermostat random stream

  // ---- state
  bool coords_dirty = true;      // coordinates changed since the sorted copy
  bool rpole_valid = false;
  bool uind_valid = false;
};

// Per-block partial rows of the energy/virial reductions.
enum {
  EP_MREAL = 0, EP_MRECIP, EP_MSELF, EP_POLAR, EP_VXX, EP_VYY, EP_VZZ, EP_VXY, EP_VXZ, EP_VYZ,
  EP_EVDW, EP_EBOND, EP_EANGLE, EP_EUREY, EP_ECT, EP_EDISP, EP_EREP, EP_N
};
constexpr int kEnergyRow = 24;
static_assert(EP_N <= kEnergyRow, "an energy row holds every EP_* slot");

// ----- entry points implemented across the .hip files
void ctx_init(Ctx& c, const System& s, int precision);
void nblist_build(Ctx& c);
// refresh: enqueue the sorted-coordinate refresh while the host waits for the
// check (a rebuild overwrites it); then call nblist_mark_refreshed if no rebuild
bool nblist_needs_rebuild(Ctx& c, bool refresh = false);
void nblist_mark_refreshed(Ctx& c);
void nblist_compact(Ctx& c, hipStream_t st = nullptr);  // refresh the in-cutoff lists (main stream unless st)
bool nblist_compact_buffers(Ctx& c);  // size the in-cutoff lists; false without a list
void refresh_sorted_coords(Ctx& c);
void mpole_init(Ctx& c);
// c.field/fieldp = full direct field; with ua, also fa = the mutual field of
// the dipoles ua ([2][n][3] sorted: d, then p), computed beside it
void dfield_all(Ctx& c, const void* ua = nullptr, void* fa = nullptr);
void ufield_all(Ctx& c, void* ud, void* up, void* fd, void* fp);  // sorted real vectors
void pme_bin(Ctx& c);
void pme_bin_async(Ctx& c);  // binning of the current coordinates on side3
// forward FFT, influence function, inverse FFT of c.qgrid; `lines`: x and y
// as line passes even where the plane kernel applies (same bits)
void pme_fft_roundtrip(Ctx& c, bool lines = false);
void pme_check_fft_sizes(const Ctx& c);
void pme_dipole_field(Ctx& c, const void* ud, const void* up, void* fd, void* fp, hipEvent_t before_gather = nullptr);  // adds recip+self
void pme_mpole_field(Ctx& c, void* fd, void* fp, hipEvent_t before_gather = nullptr);  // adds recip+self of the permanent multipoles
// the same, then recip+self of the dipoles (ud, up) added to (gd, gp); each
// gather waits for its event
void pme_mpole_dipole_fields(Ctx& c, void* fd, void* fp, hipEvent_t e1, const void* ud, const void* up, void* gd,
                             void* gp, hipEvent_t e2);
// PME sub-steps on c.qgrid: spreading, one-direction transforms (sign -1
// forward, +1 inverse), the influence function alone, and the field of the
// grid at the atoms (fd, fp overwritten, no self term)
void pme_grid_mpole(Ctx& c);
void pme_grid_uind(Ctx& c, const void* ud, const void* up);
void pme_fft_dir(Ctx& c, int sign);
void pme_conv_grid(Ctx& c);
void pme_fphi(Ctx& c, bool mpole, void* fd, void* fp);
// context-free 3D transform of a device grid [K3][K2][K1] (complex FP32 or
// FP64), line passes in place; tw = forward twiddles of the three axes
void fft_check_size(int K, int prec);
void fft_twiddles_device(void* tw, int prec, const int* K, hipStream_t st);  // (K1 + K2 + K3) complex
void fft3d_device(void* grid, int prec, const int* K, int sign, const void* tw, hipStream_t st);
// the same transform with the twiddles in a static per-device table: no allocation
void fft3d_noalloc(void* grid, int prec, const int* K, int sign);
int induce_pcg(Ctx& c, int* iters, double* rms);
std::array<double, 16> aspc_coefficients();
// learned guess (guess.hip): nparam of the network, allocation, per-solve hooks
int guess_nparam();
void guess_setup(Ctx& c);
bool guess_wants_features(const Ctx& c);        // capture on or learned guess usable
void guess_features(Ctx& c, const double* coef_e, const double* coef_u);  // raw features -> c.gfeat
void guess_save_fields(Ctx& c);                 // current direct field -> history
void guess_apply(Ctx& c, void* ud, void* up);   // u += network correction (local -> global)
void guess_capture(Ctx& c);                     // after the solve: features + targets of this step
void guess_set_params(Ctx& c, const float* params, const float* norm, uint64_t seed);
void guess_get_params(Ctx& c, float* params, float* norm);
void guess_to_bf16(Ctx& c, long long count, const float* src_dev, void* dst_dev);
void mlp_forward(Ctx& c, int m, const void* X_bf16_dev, float* Y_dev);  // Y [m][8]
double mlp_grad(Ctx& c, int m, const void* X_bf16_dev, const float* T_dev, float* grad_dev);  // grad [nparam + 1]
void adam_step(Ctx& c, const float* grad_dev, double lr);
void guess_capture_start(Ctx& c, int stride, long long max_samples);
void guess_fit_norm(Ctx& c);
void guess_train(Ctx& c, int epochs, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last);
// dots64: also r.z (d, p) and r.r (d, p) per 64 elements (4 doubles each), or null
// mode -1: the keywords' choice (sparse when pcg-precond and usolve-cutoff
// allow, else diagonal); 0: diagonal z = alpha r; 1: sparse (error if off)
void precond_apply(Ctx& c, const void* rd, const void* rp, void* zd, void* zp, void* z2d = nullptr,
                   void* z2p = nullptr, double* dots64 = nullptr, int mode = -1);  // z (and a second copy of z when given)
// spread mode 0: dipoles (src1 real, src2 imaginary); 1: multipoles (real);
// 2: multipoles src1 (real) + dipoles src2 (imaginary)
void pme_spread(Ctx& c, int mode, const void* s1, const void* s2);
// FFT of c.qgrid: roundtrip != 0 runs forward, influence function, inverse;
// roundtrip == 0 only the forward transforms.  vscale != 0 adds vscale times
// the convolution virial sums of the transformed grid to the energy rows.
void pme_fft_virial(Ctx& c, int roundtrip, double vscale);
double* energy_rows(Ctx& c, int nrows);   // reserve partial rows for one launch
void energy_setup(Ctx& c, const System& s);
void md_setup(Ctx& c, const System& s);   // vdW / valence tables, MD buffers
void vdw_sites(Ctx& c);                   // vdW site positions (sorted) from the coordinates
// vdW (sorted egrad += ..) and valence (caller-order gval) terms of one
// evaluation; partial rows reserved from the energy rows
void vdw_run(Ctx& c, bool virial);
void vdw_launch(Ctx& c, bool virial, hipStream_t st, double* rows);  // pair kernel only
void vdw_finish(Ctx& c);                                             // site gradients onto atoms
void valence_run(Ctx& c, double* gval_caller, bool energy);
// Term groups of one evaluation: the electrostatics (multipole + polarization)
// and vdW are the slow group of the RESPA split, valence the fast group.
enum { TERM_ELEC = 1, TERM_VDW = 2, TERM_VALENCE = 4, TERM_CHGTRN = 8, TERM_DISP = 16, TERM_REPEL = 32, TERM_ALL = 63 };
// vers bit 2: virial; out16[EP_*] (kEnergyRow values) energies / virial; sorted-order dE/dx of
// the electrostatics and vdW in c.egrad, caller-order valence dE/dx in
// c.g_fast; grad_orig (host, caller order, or null) receives their sum
void energy_run(Ctx& c, int vers, int terms, double* out16, double* grad_orig, bool grad_on_device = false);
// readbacks the host waits for: mark after the copy, then wait by polling the
// marker event (a blocking synchronize wakes the thread tens of microseconds late)
void mark_readback(Ctx& c);
void wait_marked(Ctx& c);
// list, sorted coordinates, global multipoles current; defer_compact: the
// in-cutoff lists are left for the induced-dipole solve (dfield_all
// compacts them on the third stream)
void prepare(Ctx& c, bool defer_compact = false);

// molecular dynamics (caller order, FP64)
void md_velocities(Ctx& c, double kelvin, uint64_t seed);  // Maxwell-Boltzmann, net momentum removed
double md_kinetic(Ctx& c);                // kcal/mol
// v -= (h1 g1 + h2 g2) / m, g2 may be null
void md_kick(Ctx& c, double h1, const double* g1, double h2, const double* g2);
void md_drift(Ctx& c, double h);          // x += h v
void count_interactions(Ctx& c, long long* nem, long long* nep);
long long vdw_count(Ctx& c);  // analyze-mode vdW count (0 without the vdW term)
// HIPPO charge transfer: energy rows, virial, sorted egrad += (after the
// electrostatic terms have set egrad), and the analyze-mode count
void chgtrn_run(Ctx& c, bool virial);
long long chgtrn_count(Ctx& c);
// HIPPO dispersion (non-Ewald), the same contract
void disp_run(Ctx& c, bool virial);
long long disp_count(Ctx& c);
// HIPPO repulsion: egrad and etq += (the torque kernel maps etq to forces)
void repel_run(Ctx& c, bool virial);
long long repel_count(Ctx& c);
void respa_inner(Ctx& c, double* gfast, double h, int nrep);  // drift + nrep x (fast forces, kick, drift)
bool respa_fused(const Ctx& c);
void valence_inner(Ctx& c, double* gval, double h);  // valence gradient, kick by h, drift by h
void md_scale_v(Ctx& c, double a);
// Bussi rescaling with host-drawn deviates r (normal) and s2 (chi-square,
// nfree - 1), kinetic energy and factor on the device
void md_bussi(Ctx& c, double dt, double tau, double kelvin, double r, double s2);
void md_sorted_to_orig(Ctx& c, const double* sorted, double* orig);

// sorted <-> caller order transfers of [n][3] real vectors
void gather_to_host(Ctx& c, const void* dev_sorted_real3, double* host_orig);
void scatter_from_host(Ctx& c, const double* host_orig, void* dev_sorted_real3);

inline size_t rsize(const Ctx& c) { return c.prec == 32 ? sizeof(float) : sizeof(double); }

// Box description of the pair kernels (see PosRec).
template <class real>
BoxT<real> make_boxT(const Ctx& c) {
  BoxT<real> b;
  for (int d = 0; d < 3; ++d) {
    b.l[d] = (real)(c.periodic ? c.box[d] : 1.0);
    b.il[d] = (real)(c.periodic ? 1.0 / c.box[d] : 1.0);
    const double q = c.vlen[d] / 4294967296.0;
    b.q[d] = (real)q;
    b.qlo[d] = (real)(q - (double)b.q[d]);
  }
  b.periodic = c.periodic ? 1 : 0;
  return b;
}

// Records start/stop events around `launches` back-to-back launches of one
// kernel family when its timer is on.  The events skip the system-scope
// cache release (device-scope timing only), which would otherwise write L2
// back at every record.
struct KScope {
  KTimer* t;
  hipStream_t st;
  hipEvent_t stop = nullptr;
  KScope(Ctx& c, int id, int launches = 1, hipStream_t on = nullptr)
      : t(c.timers[id].on ? &c.timers[id] : nullptr), st(on ? on : c.stream) {
    if (!t) return;
    if (t->used + 2 > t->ev.size()) {
      hipEvent_t a, b;
      T9_HIP(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
      T9_HIP(hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
      t->ev.push_back(a);
      t->ev.push_back(b);
      t->nl.push_back(0);
    }
    t->nl[t->used / 2] = launches;
    T9_HIP(hipEventRecord(t->ev[t->used], st));
    stop = t->ev[t->used + 1];
    t->used += 2;
  }
  ~KScope() {
    if (t) (void)hipEventRecord(stop, st);
  }
};

// Sum the recorded intervals into ms/count and recycle the events.
void ktimer_collect(Ctx& c, int id);
void ktimer_release(Ctx& c);
long long count_cutoff_pairs(Ctx& c);

}  // namespace t9
