// Host-side system description for the AMOEBA electrostatics path.
//
// Owns what the reference obtains from its Fortran setup (absent from the
// reference tree): keyword values, force-field parameters of the AMOEBA
// subset, connectivity (1-2 .. 1-5), polarization groups, local multipole
// frames and the pair scale factors.  Everything here is FP64 and runs once
// per system; the device never sees these containers, only the flat arrays
// produced from them.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace t9 {

// Physical constants, same values as the reference's
// ext/interface/cpp/tinker/detail/units.hh:16-22.
constexpr double kBohr = 0.529177210544;
constexpr double kCoulomb = 332.0637133;  // electric = coulomb / dielectric
constexpr double kDebye = 4.803206802;
constexpr double kBoltzmann = 0.8314462618;    // g/mol Å^2/ps^2/K
constexpr double kGasConst = 1.9872042586e-3;  // kcal/mol/K
constexpr double kEkcal = 4.184e2;             // (g/mol Å^2/ps^2) per kcal/mol

// Local-frame kinds; numbering is ours, the meaning follows the reference's
// LocalFrame/polaxe (include/ff/amoeba/mpole.h:30-51).
enum Frame : int {
  kFrameNone = 0,
  kFrameZOnly = 1,
  kFrameZThenX = 2,
  kFrameBisector = 3,
  kFrameZBisect = 4,
  kFrame3Fold = 5,
};

// Multipole component order used everywhere in this code base:
// q, dx, dy, dz, qxx, qyy, qzz, qxy, qxz, qyz (the reference's MPL_PME_* order,
// include/ff/amoeba/mpole.h:6-19).
enum Mpl : int { kQ = 0, kDx, kDy, kDz, kQxx, kQyy, kQzz, kQxy, kQxz, kQyz, kMplTotal };

struct MultipoleParam {
  int type = 0;
  int ztyp = 0, xtyp = 0, ytyp = 0;  // signed as in the parameter file
  int frame = kFrameNone;
  double pole[kMplTotal] = {};       // converted to Å-based units
};

struct PolarizeParam {
  double alpha = 0, thole = 0;
  std::vector<int> group_types;
};

struct VdwParam {
  double rad = 0, eps = 0, red = 0;  // radius (as given), epsilon, H reduction
  bool set = false;
};

struct BondParam { double k = 0, r0 = 0; };
struct AngleParam { double k = 0, t0 = 0; };
struct UreyParam { double k = 0, r0 = 0; };

// Force-field parameters of the subset this path needs.
struct ForceField {
  std::string name;
  // scale factors, keyed by the reference's keyword names (amoeba09.prm:22-45)
  double m_scale[6] = {0, 0, 0, 0.4, 0.8, 1};   // index 2..5 = 1-2..1-5
  double p_scale[6] = {0, 0, 0, 1, 1, 1};
  double p_iscale[6] = {0, 0, 0, 0.5, 1, 1};
  double d_scale[5] = {0, 0, 1, 1, 1};          // index 1..4 = 1-1..1-4 groups
  double u_scale[5] = {0, 1, 1, 1, 1};
  double v_scale[6] = {0, 0, 0, 0, 1, 1};
  double w_scale[6] = {1, 1, 1, 1, 1, 1};       // induce-1x-scale (HIPPO mutual field)
  double ds_scale[6] = {0, 0, 0, 0.4, 0.8, 1};  // disp-1x-scale (HIPPO dispersion)
  double rs_scale[6] = {0, 0, 0, 1, 1, 1};      // rep-1x-scale (HIPPO repulsion)
  std::map<int, int> type_class;
  std::map<int, std::string> type_name;
  std::map<int, double> type_mass;
  std::vector<MultipoleParam> mpoles;           // file order
  std::map<int, PolarizeParam> polarize;        // by atom type
  std::map<std::pair<int, int>, double> polpair;  // pair Thole override
  // HIPPO charge penetration by class: core charge, valence damping alpha
  // (the chgpen records of a HIPPO parameter file)
  std::map<int, std::pair<double, double>> chgpen;
  // HIPPO charge transfer by class: charge, damping alpha (chgtrn records)
  std::map<int, std::pair<double, double>> chgtrn;
  // HIPPO dispersion by class: C6 coefficient, damping alpha (dispersion records)
  std::map<int, std::pair<double, double>> disp;
  // HIPPO Pauli repulsion by class: size, damping alpha, valence count
  // (repulsion records)
  std::map<int, std::tuple<double, double, double>> repel;
  std::map<int, VdwParam> vdw;                  // by class
  std::map<std::pair<int, int>, BondParam> bond;  // by class pair (sorted)
  std::map<std::tuple<int, int, int>, AngleParam> angle;
  std::map<std::tuple<int, int, int>, UreyParam> urey;
  double bond_cubic = 0, bond_quartic = 0;
  double angle_cubic = 0, angle_quartic = 0, angle_pentic = 0, angle_sextic = 0;
};

// Keyword values.  Defaults are set explicitly (the reference takes them from
// Fortran that is not in its tree); callers are expected to state every value
// that matters for a run.
struct Options {
  // every term on unless the keywords switch it off, as in Tinker
  bool use_mpole = true, use_polar = true, use_vdw = true, use_valence = true;
  bool use_bond = true, use_angle = true, use_urey = true;  // use_valence: any of the three
  bool use_chgtrn = true;      // HIPPO charge transfer, where the parameters have chgtrn records
  bool use_disp = true;        // HIPPO dispersion, where the parameters have dispersion records
  bool dewald = false;         // Ewald summation of the dispersion (not evaluated here)
  bool use_repel = true;       // HIPPO repulsion, where the parameters have repulsion records
  // valence terms this path does not evaluate (stretch-bend, out-of-plane,
  // torsions, ...): a system that would need them must switch them off
  bool other_valence_off = false;
  bool ewald = false;
  double ewald_cut = 9.0;
  double ewald_alpha = 0;      // 0: derive from the cutoff (ewaldcof rule)
  int pme_grid[3] = {0, 0, 0}; // 0: derive from the box (kewald rule)
  int pme_order = 5;
  double mpole_cut = 1e30;     // non-Ewald real-space cutoff (gas phase: none)
  double polar_eps = 1e-6;
  int polar_iter = 100;
  std::string predictor = "NONE";  // NONE | ASPC | LEARNED
  bool pcg_prec = true, pcg_guess = true;
  double pcg_peek = 1.0;
  double usolve_cut = 4.5;
  double udiag = 2.0;
  double list_buffer = 2.0;
  double vdw_cut = 9.0;
  double vdw_taper = 0;        // 0: 0.9 * vdw_cut (the reference's vdwtaper default)
  double ct_cut = 6.0;         // chgtrn-cutoff
  double ct_taper = 0;         // chgtrn-taper; 0: 0.9 * ct_cut
  double disp_cut = 9.0;       // disp-cutoff
  double disp_taper = 0;       // disp-taper; 0: 0.9 * disp_cut
  double rep_cut = 6.0;        // repuls-cutoff
  double rep_taper = 0;        // repuls-taper; 0: 0.9 * rep_cut
  bool mutual = true;
};

// One pair with at least one scale factor different from 1.  With HIPPO
// charge penetration the four slots hold the multipole scale (m), the
// polar-1x scale with its intra-group variant for the direct field and the
// induced-dipole / permanent-multipole terms (d = p), and the induce-1x scale
// of the mutual field (u), as the reference's mdw exclusion lists.
struct ScalePair {
  int i, k;  // i < k, 0-based original indices
  double m, d, p, u;
  double ds = 1;  // HIPPO dispersion (disp-1x-scale)
  double rs = 1;  // HIPPO repulsion (rep-1x-scale)
};

struct System {
  int n = 0;
  std::string title;
  std::vector<double> xyz;  // [n][3]
  std::vector<int> type, cls;
  std::vector<std::string> name;
  std::vector<std::vector<int>> i12, i13, i14, i15;  // 0-based, sorted
  std::vector<std::vector<int>> ip11, ip12, ip13, ip14;
  std::vector<double> mass;

  double box[3] = {0, 0, 0};
  bool periodic = false;

  // electrostatics, per atom
  std::vector<double> pole;  // [n][10], local frame
  std::vector<int> frame, zaxis, xaxis, yaxis, ysign;  // -1 = unused axis atom
  std::vector<double> polarity, thole, pdamp;
  std::vector<int> jpolar;      // index of the polarize type of each atom
  // HIPPO charge penetration, per atom (empty unless the parameters have
  // chgpen records): the monopole is a core charge pcore plus a damped
  // valence charge pval = q - pcore with damping exponent palpha
  bool chgpen = false;
  std::vector<double> pcore, pval, palpha;
  // HIPPO charge transfer, per atom (empty unless the parameters have chgtrn
  // records and the term is on): charge and damping exponent (a zero
  // exponent stands for 1000, no transfer at any distance)
  bool chgtrn = false;
  std::vector<double> chgct, dmpct;
  // HIPPO dispersion, per atom: C6 coefficient and damping exponent
  bool disp = false;
  std::vector<double> csix, adisp;
  // HIPPO repulsion, per atom: size, damping exponent, valence count
  bool repel = false;
  std::vector<double> sizpr, dmppr, elepr;
  std::vector<double> thlval;   // [njpolar][njpolar] pair Thole values
  int njpolar = 0;

  // vdW, per atom
  std::vector<double> vdw_rad, vdw_eps, vdw_red;
  std::vector<int> vdw_parent;
  // vdW pair parameters by vdW class index (cubic-mean R_min, HHG epsilon)
  std::vector<int> jvdw;          // [n] index into the tables
  int njvdw = 0;
  std::vector<double> vdw_rmin, vdw_epsmix;  // [njvdw][njvdw]
  std::vector<int> vscale_ik;     // [nvscale][2] pairs (i < k) with vdW scale != 1
  std::vector<double> vscale_s;   // [nvscale]

  // valence terms (0-based atoms)
  std::vector<int> bond_ij;       // [nbond][2]
  std::vector<double> bond_kr;    // [nbond][2] force constant, ideal length
  std::vector<int> angle_abc;     // [nangle][3], b the vertex
  std::vector<double> angle_kt;   // [nangle][2] force constant, ideal angle (degrees)
  std::vector<int> urey_ac;       // [nurey][2]
  std::vector<double> urey_kr;    // [nurey][2]

  std::vector<ScalePair> scales;  // sorted by (i, k)
  Options opt;
  ForceField ff;
};

// Parsing and setup (system.cpp).  All throw std::runtime_error on bad input.
void read_prm(const std::string& path, ForceField& ff);
void read_prm_text(const std::string& text, ForceField& ff);  // parameter records of a .key
void read_xyz(const std::string& path, System& s);
void read_key_text(const std::string& text, System& s, std::string* prm_name);
void setup_system(System& s);  // connectivity, groups, frames, scales, vdW

// Seeded rigid-water box (lattice sites, random orientation, jitter) plus
// optional Na+/Cl- ions on unused sites.
void build_water_box(System& s, int nmol, int nsite_side, double box_len, uint64_t seed,
                     int n_na = 0, int n_cl = 0, double ion_min_dist = 3.0);

// Ewald coefficient from the cutoff (erfc(a*rc)/rc = 1e-8 by bisection) and
// the default PME grid size from the box length (1.2 points/Å rounded up to
// a 2,3,5-smooth size).
double ewald_coeff(double cutoff, double eps = 1e-8);
int default_pme_grid(double box_len);

}  // namespace t9
