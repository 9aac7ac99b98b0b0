/*
 * t9hip.h — C ABI of the MI355X (gfx950) AMOEBA electrostatics path.
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns T9_OK (0) or a T9_ERR_* code; t9_last_error() describes the last
 * failure on the calling thread.  A context (t9_ctx) owns all device memory
 * for one system on one GPU and is not thread-safe.  Host arrays are laid out
 * [n][3] (xyz, dipoles, fields, gradients) or [n][10] (multipoles) in the
 * caller's atom order, FP64, units as in the reference: Å, e·Å (dipoles),
 * e/Å² (fields), kcal/mol, kcal/mol/Å.
 *
 * Each device entry point names the reference interface it replaces
 * (paths relative to the TinkerTools/tinker-gpu tree).
 */
#ifndef T9HIP_H
#define T9HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
  T9_OK = 0,
  T9_ERR_HIP = 1,            /* a HIP runtime call failed (device missing, fault, OOM) */
  T9_ERR_NOT_CONVERGED = 2,  /* induce: iteration limit reached
                                (src/cu/amoeba/pcg.cu:180-184 throws here) */
  T9_ERR_LIST_OVERFLOW = 3,  /* neighbor list exceeded its capacity
                                (src/cu/spatial.cu:708-716 throws here) */
  T9_ERR_ARG = 4,            /* bad argument or unsupported option */
  T9_ERR_SETUP = 5           /* parameter / coordinate / keyword setup failed */
};

const char* t9_last_error(void);

/* ----------------------------------------------------------- host system */
/* A host-side system: coordinates, force-field parameters, keywords and the
 * derived topology (1-2..1-5 connectivity, polarization groups, local frames,
 * pair scale factors).  Replaces the reference's Fortran setup chain
 * (test/testrt.cpp:184-193: initial, getxyz, mechanic) whose sources are not
 * in the reference tree. */
typedef struct t9_system t9_system;

/* Tinker .xyz + .prm + keyword text (the contents of a .key file).  A
 * parameter file with chgpen records (HIPPO, GORDON1 charge penetration)
 * gives a system whose multipole and polarization terms use the damped
 * core/valence kernels (the reference's empoleChgpen / epolarChgpen /
 * dfieldChgpen / ufieldChgpen, src/cu/hippo/empole.cu, epolar.cu, field.cu);
 * every entry point below then computes the HIPPO terms.  HIPPO's other
 * terms (repulsion, dispersion, charge transfer) and AMOEBA+ are not built. */
int t9_system_from_files(const char* xyz_path, const char* prm_path, const char* key_text,
                         t9_system** out);
/* Seeded AMOEBA water box: nmol rigid waters (O-H 0.9572 Å, 104.52°) on an
 * nside^3 cubic lattice in a cube of edge box_len, random orientation and
 * ±0.1 Å jitter from `seed`, plus n_na Na+ and n_cl Cl- on unused sites at
 * least 3 Å from any oxygen. */
int t9_system_water_box(int nmol, int nside, double box_len, uint64_t seed, int n_na, int n_cl,
                        const char* prm_path, const char* key_text, t9_system** out);
void t9_system_destroy(t9_system* sys);
int t9_system_natoms(const t9_system* sys);
/* Named per-atom arrays.  f64: "xyz" [n][3], "pole" [n][10] (local frame),
 * "polarity", "pdamp", "thole", "mass" [n], "thlval" [nj*nj], "box" [3].
 * i32: "type", "frame", "zaxis", "xaxis", "yaxis", "ysign", "jpolar" [n],
 * "i12" [n][4] (-1 padded). */
int t9_system_get_f64(const t9_system* sys, const char* name, double* out);
int t9_system_get_i32(const t9_system* sys, const char* name, int* out);
/* Scalars: "aewald", "ewald_cut", "ewald", "grid1", "grid2", "grid3",
 * "pme_order", "njpolar", "nscale", "polar_eps", "polar_iter", "pcg_prec",
 * "pcg_guess", "pcg_peek", "usolve_cut", "udiag", "list_buffer", "periodic",
 * "use_mpole", "use_polar", "mpole_cut", "vdw_cut". */
int t9_system_get_scalar(const t9_system* sys, const char* name, double* out);
/* Pairs with any scale factor != 1: ik [nscale][2] (i < k), mdpu [nscale][4]. */
int t9_system_get_scales(const t9_system* sys, int* ik, double* mdpu);
int t9_system_set_xyz(t9_system* sys, const double* xyz);
/* Override a keyword after setup (same names as t9_system_get_scalar). */
int t9_system_set_scalar(t9_system* sys, const char* name, double value);
/* One frame in Tinker's .xyz / dump layout (test/ref/tinkernist.arc, .uind,
 * .udir, .def, .tef; the reference writes them through Tinker's Fortran
 * dump routine, src/mdsave.cpp:41-215, not in the reference tree): header
 * "%6d  title", the box line when periodic, then per atom index, name (3
 * columns), three values in f12.6, type and the 1-2 connected atoms.
 * values [n][3] in caller order: coordinates (.arc, NULL = the system's own),
 * or induced dipoles in Debye (.uind), fields (.def / .tef), ...
 * append != 0 adds the frame to an existing file (an archive). */
int t9_system_write_frame(const t9_system* sys, const char* path, const double* values, int append);

/* ---------------------------------------------------------------- context */
typedef struct t9_ctx t9_ctx;

enum { T9_FP32 = 32, T9_FP64 = 64 };

/* Allocate device state for `sys` on the current HIP device.  precision
 * selects the pair/PME/vector arithmetic (T9_FP32 or T9_FP64); coordinates,
 * dot products, residual norms and energy/virial sums are FP64 in both.
 * Replaces the resource chain initialize() -> epolarData / pmeData /
 * nblistData (src/rcman.cpp:49-80, src/amoeba/epolar.cpp:25-511,
 * src/pme.cpp:119-217, src/nblist.cpp:317-650). */
int t9_ctx_create(const t9_system* sys, int precision, t9_ctx** out);
int t9_ctx_destroy(t9_ctx* ctx);

/* Host -> device coordinates [n][3] (replaces the x/y/z device upload of
 * copyPosToXyz, src/atom.cpp:153-166).  Does not rebuild the list. */
int t9_set_coords(t9_ctx* ctx, const double* xyz);
/* Device -> host coordinates. */
int t9_get_coords(t9_ctx* ctx, double* xyz);

/* Neighbor list refresh: rebuild when any atom moved more than half the list
 * buffer since the last build, else keep.  *rebuilt (may be NULL) reports
 * which.  Replaces nblistRefresh() (src/nblist.cpp:534) and
 * spatialCheck_cu / spatialDataInit_cu (src/cu/spatial.cu:728-960). */
int t9_nblist_refresh(t9_ctx* ctx, int* rebuilt);
/* Force a rebuild (nblistData(RcOp::INIT)). */
int t9_nblist_build(t9_ctx* ctx);
/* List statistics: [0] list pairs (directed), [1] atoms, [2] waves,
 * [3] padded slots, [4] builds so far. */
int t9_nblist_stats(t9_ctx* ctx, int64_t* out5);
/* The pairs stored in a neighbor list (which 0: multipole/polarization list
 * at cutoff + buffer, 1: vdW-site list), each once as caller-order indices
 * (i, k), i < k, into pairs[cap][2]; *count = number of pairs (may exceed
 * cap: then only the first cap are written).  The readback the reference's
 * neighbor-list test takes of nlst / lst (test/nblist.cpp:160-185). */
int t9_nblist_pairs(t9_ctx* ctx, int which, int64_t cap, int* pairs, int64_t* count);

/* Local frame -> global multipoles.  Replaces mpoleInit's chkpole/rotpole
 * (src/amoeba/mpole.cpp:28-58, src/cu/amoeba/rotpole.cu:7-42). */
int t9_mpole_init(t9_ctx* ctx);
int t9_get_rpole(t9_ctx* ctx, double* rpole);

/* Direct field of the permanent multipoles, d and p schemes, [n][3] each.
 * Replaces dfield(field, fieldp) -> dfieldEwald / dfieldNonEwald
 * (include/ff/amoeba/induce.h:8-20, src/amoeba/field.cpp:17-63). */
int t9_dfield(t9_ctx* ctx, double* field, double* fieldp);
/* Mutual field -T·u of two dipole sets.  Replaces ufield(uind, uinp, field,
 * fieldp) -> ufieldEwald / ufieldNonEwald (src/amoeba/field.cpp:76-117). */
int t9_ufield(t9_ctx* ctx, const double* uind, const double* uinp, double* field, double* fieldp);

/* Induced dipole solve statistics. */
typedef struct {
  int iterations;       /* PCG iterations */
  int predictor;        /* 0 direct guess, 1 ASPC, 2 learned */
  double rms_debye;     /* final RMS residual (Debye) */
  double guess_rms;     /* RMS residual of the initial guess (Debye) */
} t9_induce_stats;

/* Mutual induced dipoles by preconditioned CG to polar-eps.  uind/uinp may be
 * NULL: the result then stays on the device (read it with t9_get_uind).
 * Replaces induce(uind, uinp) (include/ff/amoeba/induce.h:34,
 * src/amoeba/induce.cpp:108-113 -> src/cu/amoeba/pcg.cu:14-185). */
int t9_induce(t9_ctx* ctx, double* uind, double* uinp, t9_induce_stats* stats);
/* RMS residual (Debye) of the last solve at each iteration: eps[0] is the
 * initial guess, eps[k] iteration k; *count = iterations + 1 (copies at most
 * cap values, at most 256 are kept).  The values the reference prints per
 * iteration in debug mode (src/cu/amoeba/pcg.cu:150-157). */
int t9_induce_trace(t9_ctx* ctx, double* eps, int cap, int* count);
int t9_get_uind(t9_ctx* ctx, double* uind, double* uinp);
int t9_set_uind(t9_ctx* ctx, const double* uind, const double* uinp);

/* Energy, gradient and virial of the terms enabled by the system's keywords:
 * the multipole and polarization terms and, when enabled, the buffered 14-7
 * vdW and the bond / angle / Urey-Bradley valence terms.  Solves for the
 * induced dipoles first, as the reference's emplar does.  vers is a bit set:
 * T9_CALC_ENERGY, T9_CALC_GRAD, T9_CALC_VIRIAL (the reference's
 * calc::energy/grad/virial, include/tool/rcman.h:105-155).  grad receives
 * dE/dx [n][3] in kcal/mol/Å in caller order when non-NULL.  Replaces
 * energy(vers) (src/energy.cpp:319-450) for these terms: emplar(vers),
 * src/amoeba/emplar.cpp:10-28 and src/cu/amoeba/emplar.cu:546-605, with
 * torque(vers, gx, gy, gz), src/cu/amoeba/torque.cu:9-380; evdw(vers) for
 * AMOEBA (src/evdw.cpp, src/cu/evdw.cu ehal path); ebond / eangle / eurey
 * (src/cu/evalence.cu). */
enum { T9_CALC_ENERGY = 1, T9_CALC_GRAD = 2, T9_CALC_VIRIAL = 4, T9_CALC_ANALYZE = 8 };
typedef struct {
  double esum;          /* sum of the terms below, kcal/mol */
  double em, ep;        /* permanent multipole, polarization */
  double em_real, em_recip, em_self;  /* parts of em (recip/self zero without Ewald) */
  double vir[9];        /* dE/d(strain), kcal/mol, symmetric, all terms */
  int induce_iterations;
  double evdw;          /* buffered 14-7 vdW */
  double ebond, eangle, eurey;  /* valence */
  /* with T9_CALC_ANALYZE (the reference's calc::analyz, energy(calc::v3)):
   * multipole and polarization interaction counts nem / nep, the pairs
   * within the cutoff less the excluded ones, plus one self term per atom
   * with Ewald (src/acc/amoeba/empoleewald.cpp, epolarewald.cpp), and the
   * vdW count nev (src/acc/ehal.cpp); else 0 */
  int64_t nem, nep, nev;
  /* HIPPO charge transfer (echgtrn, src/hippo/echgtrn.cpp:73-100,
   * src/acc/hippo/echgtrn.cpp:13-183) and, with T9_CALC_ANALYZE, its
   * interaction count nct: pairs within chgtrn-cutoff with a nonzero
   * energy, less the excluded (scale 0) pairs */
  double ect;
  int64_t nct;
  /* HIPPO dispersion with a cutoff (edisp, src/hippo/edisp.cpp,
   * src/acc/hippo/edisp.cpp:113-296 in its non-Ewald form) and its count
   * ndisp with T9_CALC_ANALYZE */
  double edisp;
  int64_t ndisp;
  /* HIPPO Pauli repulsion (erepel, src/hippo/erepel.cpp,
   * src/acc/hippo/erepel.cpp:13-215) and its count nrep with
   * T9_CALC_ANALYZE */
  double erep;
  int64_t nrep;
} t9_energy_out;
int t9_energy(t9_ctx* ctx, int vers, t9_energy_out* out, double* grad);
/* Term selection of one energy call, as energy_core selects its terms by the
 * calc flags and the RESPA time-scale group (src/energy.cpp:176-291):
 * T9_TERM_ELEC the multipole + polarization pair (emplar, with the induced
 * dipole solve), T9_TERM_VDW the buffered 14-7 vdW (evdw), T9_TERM_VALENCE
 * bond / angle / Urey-Bradley (ebond, eangle, eurey), T9_TERM_CHGTRN the
 * HIPPO charge transfer (echgtrn), T9_TERM_DISP the HIPPO dispersion
 * (edisp), T9_TERM_REPEL the HIPPO Pauli repulsion (erepel).  A term the system's keywords switch off
 * stays off.  t9_energy is all of them.  A host that keeps its own vdW and
 * valence terms asks for T9_TERM_ELEC only. */
enum { T9_TERM_ELEC = 1, T9_TERM_VDW = 2, T9_TERM_VALENCE = 4, T9_TERM_CHGTRN = 8, T9_TERM_DISP = 16,
       T9_TERM_REPEL = 32, T9_TERM_ALL = 63 };
int t9_energy_terms(t9_ctx* ctx, int vers, int terms, t9_energy_out* out, double* grad);

/* Device-resident coordinates and gradient: d_xyz and d_grad are device
 * pointers to [n][3] FP64 in caller order (the reference keeps x/y/z and the
 * gradient on the device: copyPosToXyz, src/atom.cpp:153-166, and the
 * gx/gy/gz accumulators read by sumGradient, src/energy.cpp:319-448).  No
 * PCIe transfer of per-atom data; d_grad receives dE/dx of the selected
 * terms when vers has T9_CALC_GRAD. */
int t9_set_coords_device(t9_ctx* ctx, const double* d_xyz);
int t9_get_coords_device(t9_ctx* ctx, double* d_xyz);
int t9_energy_device(t9_ctx* ctx, int vers, int terms, t9_energy_out* out, double* d_grad);

/* Preconditioners of the PCG solve on caller-order residuals [n][3]
 * (d and p): z = alpha r (replaces diagPrecond, include/ff/amoeba/induce.h:23,
 * src/cu/diagprecond.cu:5-42), and the sparse form
 * z = udiag alpha r + sum_{r_ik < usolve-cutoff} alpha_i alpha_k T_ik r_k
 * with Thole damping (replaces sparsePrecondApply,
 * include/ff/amoeba/induce.h:25-29, src/cu/amoeba/precond.cu:12-43;
 * T9_ERR_ARG when pcg-precond or usolve-cutoff leave it off). */
int t9_diag_precond(t9_ctx* ctx, const double* rsd, const double* rsdp, double* zrsd, double* zrsdp);
int t9_sparse_precond_apply(t9_ctx* ctx, const double* rsd, const double* rsdp, double* zrsd, double* zrsdp);

/* PME sub-steps on the context's complex grid (Ewald systems only; the
 * reference's PMEUnit calls, include/ff/pme.h:64-91).  grid, when non-NULL,
 * is a host array [K3][K2][K1][2] FP64 (real, imaginary; x fastest): the
 * spreading calls copy the grid out, the transforms and the convolution
 * read it in and write the result back, the gathers read it in; NULL keeps
 * the grid on the device between calls.
 *  - t9_pme_grid_mpole: B-spline spreading of the global multipoles in
 *    fractional form, real part (gridMpole, src/cu/pme.cu:185-255);
 *  - t9_pme_grid_uind: fractional dipoles, d real and p imaginary (gridUind);
 *  - t9_pme_fft_front / t9_pme_fft_back: unnormalised forward (exponent -1)
 *    and inverse (+1) 3D transforms in place (fftfront / fftback,
 *    src/cudart/fft.cpp:16-99);
 *  - t9_pme_conv: the influence function exp(-pi^2 h^2/a^2) /
 *    (pi V h^2 bsmod1 bsmod2 bsmod3), k = 0 and exponents below -50 zeroed
 *    (pmeConv, src/cu/pme.cu:908-1007);
 *  - t9_pme_fphi_mpole / t9_pme_fphi_uind: the Cartesian field -grad phi of
 *    the grid at each atom [n][3], caller order, without the self term (the
 *    field part of fphiMpole / fphiUind + fphiToCphi, src/cu/pme.cu:334-380). */
int t9_pme_grid_mpole(t9_ctx* ctx, double* grid);
int t9_pme_grid_uind(t9_ctx* ctx, const double* uind, const double* uinp, double* grid);
int t9_pme_fft_front(t9_ctx* ctx, double* grid);
int t9_pme_fft_back(t9_ctx* ctx, double* grid);
int t9_pme_conv(t9_ctx* ctx, double* grid);
int t9_pme_fphi_mpole(t9_ctx* ctx, const double* grid, double* field);
int t9_pme_fphi_uind(t9_ctx* ctx, const double* grid, double* fd, double* fp);

/* Context-free unnormalised 3D complex transform in place of a device grid
 * [k3][k2][k1] of interleaved complex T9_FP32 or T9_FP64 values, sign -1
 * forward, +1 inverse, on the current device's null stream (returns when
 * done).  Lengths: products of 2, 3 and 5 up to 255.
 * Replaces the cuFFT plan + cufftExecC2C/Z2Z pair of the reference's FFT
 * unit (src/cudart/fft.cpp:16-99); allocates its twiddle table per call. */
int t9_fft3d(void* grid, int k1, int k2, int k3, int precision, int sign);
/* The same transform without any device allocation: the twiddle table lives
 * in a static per-GPU scratch array of 12,288 bytes (enough for three axes of
 * up to 255 points in T9_FP64), filled by a kernel on every call.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Results are bit-identical to t9_fft3d.  Every call on
 * one GPU shares that GPU's table: calls that overlap on one GPU race.
 * T9_ERR_ARG for lengths beyond the table (or not products of 2, 3, 5). */
int t9_fft3d_noalloc(void* grid, int k1, int k2, int k3, int precision, int sign);

/* ------------------------------------------------------- learned guess */
/* A learned initial guess for the induced dipoles.  No counterpart in the
 * reference: it plugs in where the reference's predictor runs
 * (src/amoeba/induce.cpp:29-69, src/cu/upredict.cu:34-201; polar-predict
 * ASPC is required and stays the fallback).  u0 = u_ASPC + R^T NN(x): R the
 * atom's multipole local frame, x 32 local-frame features (field
 * extrapolation errors of the atom and its frame atoms, current field, last
 * dipole and its change, polarizability, type); NN a 32-128-128-6 BF16 MLP
 * with a linear skip, evaluated and trained with MFMA kernels.  The network
 * only changes the starting point: PCG still converges to polar-eps, and
 * every solve starts from the better of the learned and the ASPC guess by
 * RMS residual (the ASPC guess's residual is formed in the same solve: one
 * extra matrix-vector product, computed beside the direct field).
 * Parameters: flat FP32 [t9_guess_nparam()] in the order W1 [32][128], b1,
 * W2 [128][128], b2, W3 [160][6] (rows 0..127 hidden, 128..159 skip), b3;
 * norm [76]: feature mean [32], 1/std [32], target mean [6], std [6]. */
int t9_guess_nparam(void);
/* params NULL: seeded initialization; norm NULL: fitted by t9_guess_train. */
int t9_guess_set_params(t9_ctx* ctx, const float* params, const float* norm, uint64_t seed);
int t9_guess_get_params(t9_ctx* ctx, float* params, float* norm);
int t9_guess_enable(t9_ctx* ctx, int on);
/* Training data: from the next solve on, every stride-th atom's raw features
 * and converged correction R (u - u_ASPC) are stored on the device, up to
 * max_samples rows (stride 0 stops). */
int t9_guess_capture(t9_ctx* ctx, int stride, int64_t max_samples);
/* [0] captured rows, [1] solves that used the learned guess, [2] fallbacks
 * to ASPC, [3] predictor of the last solve (0 direct, 1 ASPC, 2 learned). */
int t9_guess_stats(t9_ctx* ctx, int64_t* out4);
/* Adam over the captured rows (mean squared error of the normalized
 * targets), seeded pseudo-random order; mean loss of the first/last epoch. */
int t9_guess_train(t9_ctx* ctx, int epochs, int batch, double lr, uint64_t seed, double* loss_first,
                   double* loss_last);
/* Kernel-level entry points (used by the parity tests): network outputs
 * Y [m][6] for normalized features X [m][32] (rounded to BF16 on the device);
 * the mean-squared-error gradient [nparam] against targets T [m][6]; one Adam
 * step with a given gradient; the raw features [n][32] (caller order) that
 * the next solve at the current coordinates would use; history entries
 * (kind 0 E_d, 1 E_p, 2 u_d, 3 u_p; age 0 = newest) [n][3]. */
int t9_guess_mlp_forward(t9_ctx* ctx, const float* X, int m, float* Y);
int t9_guess_mlp_grad(t9_ctx* ctx, const float* X, const float* T, int m, float* grad, double* loss);
int t9_guess_adam(t9_ctx* ctx, const float* grad, double lr);
int t9_guess_features(t9_ctx* ctx, float* F);
int t9_guess_history(t9_ctx* ctx, int kind, int age, double* out);

/* ------------------------------------------------------------ dynamics */
/* Molecular dynamics on the device, caller-order FP64 positions and
 * velocities (Å, Å/ps), masses from the system (g/mol).
 * t9_md_init draws Maxwell-Boltzmann velocities at `kelvin` from `seed`
 * (net momentum removed) and seeds the thermostat stream; it replaces the
 * velocity setup of the reference's Fortran mdinit (not in the reference
 * tree). */
int t9_md_init(t9_ctx* ctx, double kelvin, uint64_t seed);
int t9_md_set_velocities(t9_ctx* ctx, const double* vel);
int t9_md_get_velocities(t9_ctx* ctx, double* vel);
typedef struct {
  double epot;          /* potential energy after the last step, kcal/mol */
  double ekin;          /* kinetic energy after the last step, kcal/mol */
  double temperature;   /* 2 ekin / (nfree R), nfree = 3n - 3 */
  int64_t induce_iterations;  /* CG iterations summed over the run */
  int list_builds;      /* neighbor-list rebuilds during the run */
  int steps;
} t9_md_stats;
/* nsteps RESPA steps of outer length dt (ps) with nrespa inner steps: the
 * slow group (multipole, polarization, vdW) kicks at the outer half steps,
 * the fast group (valence) at the inner steps; nrespa = 1 is velocity
 * Verlet.  bath_kelvin > 0 applies the Bussi stochastic velocity rescaling
 * with relaxation time tau (ps) after every step; bath_kelvin <= 0 is NVE.
 * Replaces BasicIntegrator::dynamic (src/md/integrator.cpp:70-165) with
 * RespaDevice::velR0/velR1/velR2 (src/md/propagator.cpp:174-187) and
 * BussiThermostat::control2 -> bussiThermostat (src/md/thermostat.cpp:77-82,
 * src/mdpt.cpp:41-71). */
int t9_md_run(t9_ctx* ctx, int nsteps, double dt, int nrespa, double bath_kelvin, double tau, t9_md_stats* stats);

/* One Bussi velocity rescaling with given deviates: r standard normal, s2
 * chi-square with nfree - 1 degrees of freedom (t9_md_run draws them itself).
 * The kernel-level step of bussiThermostat (src/mdpt.cpp:41-71): scale =
 * sqrt(c + (s2 + r^2) d + 2 r sqrt(c d)), c = exp(-dt/tau),
 * d = (1 - c) (kelvin / T) / nfree, negative when r + sqrt(c/d) < 0. */
int t9_md_bussi(t9_ctx* ctx, double dt, double tau, double kelvin, double r, double s2);

/* Wait for all work queued on the context's stream. */
int t9_sync(t9_ctx* ctx);

/* ------------------------------------------------------------ measurement */
/* No counterpart in the reference (its only timing is the ns/day printout,
 * src/xdynamic.cpp:175-190).  Live HIP-event timing of one kernel family on
 * the context's stream: mode 1 clears and starts, 0 stops, 2 reads.  Ids:
 * 0 ufield real-space pairs, 1 dfield real-space pairs, 2 PME spread,
 * 3 FFT passes, 4 PME gathers, 5 preconditioner, 6 PCG vector kernels,
 * 7 fused multipole + polarization real-space energy/gradient kernel,
 * 8 vdW pair kernel, 9 valence kernel. */
enum { T9_K_UFIELD = 0, T9_K_DFIELD, T9_K_SPREAD, T9_K_FFT, T9_K_GATHER, T9_K_PRECOND, T9_K_PCG_VEC, T9_K_EMPLAR,
       T9_K_VDW, T9_K_VALENCE };
int t9_kernel_timing(t9_ctx* ctx, int kernel, int mode, double* total_ms, int64_t* launches);
/* Unique atom pairs within the real-space cutoff on the current coordinates
 * (the algorithmic unit of the pair kernels). */
int t9_cutoff_pairs(t9_ctx* ctx, int64_t* pairs);
/* Plain device-to-device copy kernel of nbytes, repeated `reps` times
 * (calibrates memory-traffic counters against a known byte count). */
int t9_copy_probe(int64_t nbytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* T9HIP_H */
