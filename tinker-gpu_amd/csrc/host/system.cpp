// Host-side setup for the AMOEBA electrostatics path: parameter/coordinate/key
// parsing, connectivity, polarization groups, local frames, pair scales, and
// the seeded water-box generator.
//
// The reference obtains all of this from Fortran Tinker (kmpole, kpolar,
// polargrp, attach, kewald, ewaldcof), which is absent from its tree.  The
// rules below are restated from the behaviour documented by the reference's
// C++ consumers: src/amoeba/epolar.cpp:88-398 (d/p/u scale rules),
// src/elec.cpp:147-503 (m scales), include/seq/rotpole.h:8-224 (frame use),
// src/pme.cpp:84-117 (PME setup).  The golden vectors of test/localframe.cpp
// pin the result.
#include "system.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <functional>
#include <numeric>
#include <set>
#include <sstream>
#include <stdexcept>

namespace t9 {
namespace {

std::vector<std::string> split(const std::string& line) {
  std::vector<std::string> out;
  size_t i = 0, n = line.size();
  while (i < n) {
    while (i < n && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r' || line[i] == ',')) ++i;
    if (i >= n) break;
    if (line[i] == '"') {
      size_t j = line.find('"', i + 1);
      if (j == std::string::npos) j = n;
      out.push_back(line.substr(i + 1, j - i - 1));
      i = j + 1;
    } else {
      size_t j = i;
      while (j < n && line[j] != ' ' && line[j] != '\t' && line[j] != '\r' && line[j] != ',') ++j;
      out.push_back(line.substr(i, j - i));
      i = j;
    }
  }
  return out;
}

std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

bool is_int(const std::string& s) {
  if (s.empty()) return false;
  size_t i = (s[0] == '-' || s[0] == '+') ? 1 : 0;
  if (i >= s.size()) return false;
  for (; i < s.size(); ++i)
    if (!std::isdigit((unsigned char)s[i])) return false;
  return true;
}

double to_d(const std::string& s) {
  std::string t = s;
  for (auto& c : t)
    if (c == 'd' || c == 'D') c = 'e';
  return std::stod(t);
}

void add_unique_sorted(std::vector<int>& v, int x) {
  auto it = std::lower_bound(v.begin(), v.end(), x);
  if (it == v.end() || *it != x) v.insert(it, x);
}

bool contains(const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

// splitmix64: small, well-known seeded generator; the same stream drives every
// seeded construction so CPU and GPU runs start from identical coordinates.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double gauss() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

}  // namespace

// ---------------------------------------------------------------- parsing

namespace {
void parse_prm_lines(const std::vector<std::string>& lines, ForceField& ff);
}

void read_prm(const std::string& path, ForceField& ff) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open parameter file " + path);
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(in, line)) lines.push_back(line);
  parse_prm_lines(lines, ff);
}

// Parameter records in the keyword text (Tinker reads the .key file after
// the parameter file, so its records replace the file's: a multipole record
// of a type replaces every earlier multipole record of that type).
void read_prm_text(const std::string& text, ForceField& ff) {
  std::istringstream in(text);
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(in, line)) lines.push_back(line);
  const size_t first = ff.mpoles.size();
  parse_prm_lines(lines, ff);
  if (ff.mpoles.size() == first) return;
  std::set<int> types;
  for (size_t q = first; q < ff.mpoles.size(); ++q) types.insert(ff.mpoles[q].type);
  std::vector<MultipoleParam> kept;
  for (size_t q = 0; q < ff.mpoles.size(); ++q)
    if (q >= first || !types.count(ff.mpoles[q].type)) kept.push_back(ff.mpoles[q]);
  ff.mpoles.swap(kept);
}

namespace {
void parse_prm_lines(const std::vector<std::string>& lines, ForceField& ff) {
  for (size_t li = 0; li < lines.size(); ++li) {
    auto tok = split(lines[li]);
    if (tok.empty()) continue;
    std::string key = lower(tok[0]);
    // a parameter record starts with its keyword and an integer; other lines
    // that start with a record keyword (the text of a file's header) are not
    // records, and Tinker skips them too
    static const std::set<std::string> kRecords = {"atom", "vdw", "bond", "angle", "ureybrad", "multipole",
                                                    "polarize", "polpair", "chgpen", "chgtrn", "dispersion",
                                                    "repulsion"};
    if (kRecords.count(key) && (tok.size() < 2 || !is_int(tok[1]))) continue;
    auto scale_kw = [&](const char* prefix, double* arr, int lo, int hi, const char* suffix) -> bool {
      for (int j = lo; j <= hi; ++j) {
        std::string kw = std::string(prefix) + std::to_string(j == 1 ? 11 : 10 + j) + suffix;
        if (key == kw && tok.size() >= 2) {
          arr[j] = to_d(tok[1]);
          return true;
        }
      }
      return false;
    };
    if (key == "forcefield" && tok.size() > 1) {
      ff.name = tok[1];
    } else if (scale_kw("mpole-", ff.m_scale, 2, 5, "-scale") ||
               scale_kw("polar-", ff.p_scale, 2, 5, "-scale") ||
               scale_kw("polar-", ff.p_iscale, 2, 5, "-intra") ||
               scale_kw("vdw-", ff.v_scale, 2, 5, "-scale") ||
               scale_kw("induce-", ff.w_scale, 2, 5, "-scale") ||
               scale_kw("disp-", ff.ds_scale, 2, 5, "-scale") ||
               scale_kw("rep-", ff.rs_scale, 2, 5, "-scale")) {
    } else if (key.rfind("direct-1", 0) == 0 && key.size() == 15 && tok.size() >= 2) {
      int j = key[8] - '0';  // direct-1j-scale
      if (j >= 1 && j <= 4) ff.d_scale[j] = to_d(tok[1]);
    } else if (key.rfind("mutual-1", 0) == 0 && key.size() == 15 && tok.size() >= 2) {
      int j = key[8] - '0';
      if (j >= 1 && j <= 4) ff.u_scale[j] = to_d(tok[1]);
    } else if (key == "bond-cubic") {
      ff.bond_cubic = to_d(tok[1]);
    } else if (key == "bond-quartic") {
      ff.bond_quartic = to_d(tok[1]);
    } else if (key == "angle-cubic") {
      ff.angle_cubic = to_d(tok[1]);
    } else if (key == "angle-quartic") {
      ff.angle_quartic = to_d(tok[1]);
    } else if (key == "angle-pentic") {
      ff.angle_pentic = to_d(tok[1]);
    } else if (key == "angle-sextic") {
      ff.angle_sextic = to_d(tok[1]);
    } else if (key == "atom" && tok.size() >= 5) {
      int t = std::stoi(tok[1]);
      ff.type_class[t] = std::stoi(tok[2]);
      ff.type_name[t] = tok[3];
      // atom  type  class  name  "description"  atomic-number  mass  valence
      if (tok.size() >= 7) ff.type_mass[t] = to_d(tok[6]);
    } else if (key == "vdw" && tok.size() >= 4) {
      VdwParam v;
      v.rad = to_d(tok[2]);
      v.eps = to_d(tok[3]);
      v.red = tok.size() >= 5 ? to_d(tok[4]) : 0.0;
      v.set = true;
      ff.vdw[std::stoi(tok[1])] = v;
    } else if (key == "bond" && tok.size() >= 5) {
      int a = std::stoi(tok[1]), b = std::stoi(tok[2]);
      ff.bond[{std::min(a, b), std::max(a, b)}] = {to_d(tok[3]), to_d(tok[4])};
    } else if (key == "angle" && tok.size() >= 6) {
      int a = std::stoi(tok[1]), b = std::stoi(tok[2]), c = std::stoi(tok[3]);
      ff.angle[{std::min(a, c), b, std::max(a, c)}] = {to_d(tok[4]), to_d(tok[5])};
    } else if (key == "ureybrad" && tok.size() >= 6) {
      int a = std::stoi(tok[1]), b = std::stoi(tok[2]), c = std::stoi(tok[3]);
      ff.urey[{std::min(a, c), b, std::max(a, c)}] = {to_d(tok[4]), to_d(tok[5])};
    } else if (key == "multipole" && tok.size() >= 3) {
      MultipoleParam mp;
      // multipole  type  [z  [x  [y]]]  charge ; then 4 continuation lines
      std::vector<int> ax;
      size_t j = 1;
      for (; j + 1 < tok.size() && is_int(tok[j]); ++j) ax.push_back(std::stoi(tok[j]));
      mp.type = ax.at(0);
      mp.ztyp = ax.size() > 1 ? ax[1] : 0;
      mp.xtyp = ax.size() > 2 ? ax[2] : 0;
      mp.ytyp = ax.size() > 3 ? ax[3] : 0;
      double v[10];
      v[0] = to_d(tok.at(j));
      int got = 1;
      size_t lj = li;
      while (got < 10) {
        ++lj;
        if (lj >= lines.size()) throw std::runtime_error("truncated multipole entry");
        for (auto& t : split(lines[lj])) v[got++] = to_d(t);
      }
      li = lj;
      // rows: d(3) ; qxx ; qxy qyy ; qxz qyz qzz
      int z = mp.ztyp, x = mp.xtyp, y = mp.ytyp;
      int fr = kFrameZThenX;
      if (z == 0) fr = kFrameNone;
      if (z != 0 && x == 0) fr = kFrameZOnly;
      if (z < 0 || x < 0) fr = kFrameBisector;
      if (x < 0 && y < 0) fr = kFrameZBisect;
      if (std::max(z, std::max(x, y)) < 0) fr = kFrame3Fold;
      mp.frame = fr;
      const double b = kBohr, b2 = kBohr * kBohr / 3.0;
      mp.pole[kQ] = v[0];
      mp.pole[kDx] = v[1] * b;
      mp.pole[kDy] = v[2] * b;
      mp.pole[kDz] = v[3] * b;
      mp.pole[kQxx] = v[4] * b2;
      mp.pole[kQxy] = v[5] * b2;
      mp.pole[kQyy] = v[6] * b2;
      mp.pole[kQxz] = v[7] * b2;
      mp.pole[kQyz] = v[8] * b2;
      mp.pole[kQzz] = v[9] * b2;
      ff.mpoles.push_back(mp);
    } else if (key == "polarize" && tok.size() >= 3) {
      // AMOEBA: polarize type alpha thole group-types...; HIPPO files have no
      // Thole column (their polarization damping is the charge penetration)
      PolarizeParam p;
      int t = std::stoi(tok[1]);
      p.alpha = to_d(tok[2]);
      const bool hippo = lower(ff.name).rfind("hippo", 0) == 0;
      size_t j = 3;
      if (!hippo && tok.size() >= 4) p.thole = to_d(tok[j++]);
      for (; j < tok.size(); ++j)
        if (is_int(tok[j])) p.group_types.push_back(std::stoi(tok[j]));
      ff.polarize[t] = p;
    } else if (key == "repulsion" && tok.size() >= 5) {
      // repulsion  class  size  alpha  valence
      ff.repel[std::stoi(tok[1])] = std::make_tuple(to_d(tok[2]), to_d(tok[3]), to_d(tok[4]));
    } else if (key == "dispersion" && tok.size() >= 4) {
      // dispersion  class  C6  alpha
      ff.disp[std::stoi(tok[1])] = {to_d(tok[2]), to_d(tok[3])};
    } else if (key == "chgtrn" && tok.size() >= 4) {
      // chgtrn  class  charge  alpha
      ff.chgtrn[std::stoi(tok[1])] = {to_d(tok[2]), to_d(tok[3])};
    } else if (key == "chgpen" && tok.size() >= 4) {
      // chgpen  class  core-charge  alpha
      ff.chgpen[std::stoi(tok[1])] = {to_d(tok[2]), to_d(tok[3])};
    } else if (key == "polpair" && tok.size() >= 4) {
      int a = std::stoi(tok[1]), b = std::stoi(tok[2]);
      ff.polpair[{std::min(a, b), std::max(a, b)}] = to_d(tok[3]);
    }
  }
}
}  // namespace

void read_xyz(const std::string& path, System& s) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open coordinate file " + path);
  std::string line;
  std::getline(in, line);
  auto tok = split(line);
  s.n = std::stoi(tok.at(0));
  {  // the title: the header after the atom count
    const size_t a = line.find_first_not_of(" \t");
    const size_t b = a == std::string::npos ? a : line.find_first_of(" \t", a);
    const size_t c = b == std::string::npos ? b : line.find_first_not_of(" \t", b);
    s.title = c == std::string::npos ? std::string() : line.substr(c);
    while (!s.title.empty() && (s.title.back() == ' ' || s.title.back() == '\r')) s.title.pop_back();
  }
  s.xyz.assign(3 * s.n, 0.0);
  s.type.assign(s.n, 0);
  s.name.assign(s.n, "");
  s.i12.assign(s.n, {});
  int i = 0;
  while (i < s.n && std::getline(in, line)) {
    tok = split(line);
    if (tok.empty()) continue;
    if (i == 0 && !is_int(tok[0])) {  // periodic box line
      s.box[0] = to_d(tok.at(0));
      s.box[1] = to_d(tok.at(1));
      s.box[2] = to_d(tok.at(2));
      s.periodic = true;
      continue;
    }
    s.name[i] = tok.at(1);
    s.xyz[3 * i] = to_d(tok.at(2));
    s.xyz[3 * i + 1] = to_d(tok.at(3));
    s.xyz[3 * i + 2] = to_d(tok.at(4));
    s.type[i] = std::stoi(tok.at(5));
    for (size_t j = 6; j < tok.size(); ++j) add_unique_sorted(s.i12[i], std::stoi(tok[j]) - 1);
    ++i;
  }
  if (i != s.n) throw std::runtime_error("coordinate file has fewer atoms than declared");
}

void read_key_text(const std::string& text, System& s, std::string* prm_name) {
  std::istringstream in(text);
  std::string line;
  Options& o = s.opt;
  bool mpole_only = false, polar_only = false, vdw_only = false, valence_only = false, ct_only = false,
       disp_only = false, rep_only = false;
  int other_off = 0, only_mask = 0;
  while (std::getline(in, line)) {
    auto tok = split(line);
    if (tok.empty()) continue;
    std::string k = lower(tok[0]);
    std::string v1 = tok.size() > 1 ? lower(tok[1]) : "";
    if (k == "parameters") {
      if (prm_name) *prm_name = tok.size() > 1 ? tok[1] : "";
    } else if (k == "a-axis") {
      s.box[0] = s.box[1] = s.box[2] = to_d(tok[1]);
      s.periodic = true;
    } else if (k == "b-axis") {
      s.box[1] = to_d(tok[1]);
    } else if (k == "c-axis") {
      s.box[2] = to_d(tok[1]);
    } else if (k == "ewald") {
      o.ewald = true;
    } else if (k == "cutoff") {
      // one distance for every nonbonded term (Tinker's CUTOFF keyword)
      o.ewald_cut = o.mpole_cut = o.vdw_cut = to_d(tok[1]);
    } else if (k == "ewald-cutoff") {
      o.ewald_cut = to_d(tok[1]);
    } else if (k == "ewald-alpha") {
      o.ewald_alpha = to_d(tok[1]);
    } else if (k == "pme-grid") {
      o.pme_grid[0] = std::stoi(tok[1]);
      o.pme_grid[1] = tok.size() > 2 ? std::stoi(tok[2]) : o.pme_grid[0];
      o.pme_grid[2] = tok.size() > 3 ? std::stoi(tok[3]) : o.pme_grid[0];
    } else if (k == "pme-order") {
      o.pme_order = std::stoi(tok[1]);
    } else if (k == "mpole-cutoff") {
      o.mpole_cut = to_d(tok[1]);
    } else if (k == "vdw-cutoff") {
      o.vdw_cut = to_d(tok[1]);
    } else if (k == "polar-eps") {
      o.polar_eps = to_d(tok[1]);
    } else if (k == "polar-iter") {
      o.polar_iter = std::stoi(tok[1]);
    } else if (k == "polar-predict") {
      o.predictor = tok.size() > 1 ? lower(tok[1]) : "aspc";
      for (auto& c : o.predictor) c = (char)std::toupper((unsigned char)c);
    } else if (k == "pcg-precond") {
      o.pcg_prec = true;
    } else if (k == "pcg-noprecond") {
      o.pcg_prec = false;
    } else if (k == "pcg-guess") {
      o.pcg_guess = true;
    } else if (k == "pcg-noguess") {
      o.pcg_guess = false;
    } else if (k == "pcg-peek") {
      o.pcg_peek = to_d(tok[1]);
    } else if (k == "usolve-cutoff") {
      o.usolve_cut = to_d(tok[1]);
    } else if (k == "usolve-diag") {
      o.udiag = to_d(tok[1]);
    } else if (k == "list-buffer") {
      o.list_buffer = to_d(tok[1]);
    } else if (k == "polarization") {
      o.mutual = (v1 != "direct");
    } else if (k == "multipoleterm") {
      if (v1 == "only") mpole_only = true;
      else o.use_mpole = (v1 != "none");
    } else if (k == "polarizeterm") {
      if (v1 == "only") polar_only = true;
      else o.use_polar = (v1 != "none");
    } else if (k == "repulsionterm") {
      if (v1 == "only") rep_only = true;
      else o.use_repel = (v1 != "none");
    } else if (k == "repuls-cutoff") {
      o.rep_cut = to_d(tok[1]);
    } else if (k == "repuls-taper") {
      o.rep_taper = to_d(tok[1]);
    } else if (k == "dispersionterm") {
      if (v1 == "only") disp_only = true;
      else o.use_disp = (v1 != "none");
    } else if (k == "disp-cutoff") {
      o.disp_cut = to_d(tok[1]);
    } else if (k == "disp-taper") {
      o.disp_taper = to_d(tok[1]);
    } else if (k == "dewald") {
      o.dewald = true;
    } else if (k == "chgtrnterm") {
      if (v1 == "only") ct_only = true;
      else o.use_chgtrn = (v1 != "none");
    } else if (k == "chgtrn-cutoff") {
      o.ct_cut = to_d(tok[1]);
    } else if (k == "chgtrn-taper") {
      o.ct_taper = to_d(tok[1]);
    } else if (k == "vdwterm") {
      if (v1 == "only") vdw_only = true;
      else o.use_vdw = (v1 != "none");
    } else if (k == "bondterm" || k == "angleterm" || k == "ureybradterm") {
      bool& flag = k == "bondterm" ? o.use_bond : (k == "angleterm" ? o.use_angle : o.use_urey);
      if (v1 == "only") {
        valence_only = true;
        only_mask |= k == "bondterm" ? 1 : (k == "angleterm" ? 2 : 4);
      } else {
        flag = (v1 != "none");
      }
    } else if (k == "strbndterm" || k == "opbendterm" || k == "opdistterm" || k == "torsionterm" ||
               k == "pitorsterm" || k == "tortorterm" || k == "angtorterm" || k == "strtorterm" ||
               k == "improperterm" || k == "imptorsterm") {
      if (v1 == "none") ++other_off;
    } else if (k == "vdw-taper") {
      const double v = to_d(tok[1]);
      o.vdw_taper = v <= 1 ? v * o.vdw_cut : v;
    }
  }
  // stretch-bend, out-of-plane bending and the torsion family all named off
  if (other_off >= 7) o.other_valence_off = true;
  if (vdw_only) {
    o.use_mpole = false;
    o.use_polar = false;
    o.use_vdw = true;
    o.use_chgtrn = o.use_disp = o.use_repel = false;
  }
  if (valence_only) {
    o.other_valence_off = true;  // "only" switches every other term off
    o.use_mpole = false;
    o.use_polar = false;
    o.use_vdw = false;
    o.use_chgtrn = o.use_disp = o.use_repel = false;
    o.use_bond = (only_mask & 1) != 0;
    o.use_angle = (only_mask & 2) != 0;
    o.use_urey = (only_mask & 4) != 0;
  }
  if (mpole_only) { o.use_mpole = true; o.use_polar = false; o.use_vdw = false; o.use_bond = o.use_angle = o.use_urey = false; }
  if (polar_only) { o.use_mpole = false; o.use_polar = true; o.use_vdw = false; o.use_bond = o.use_angle = o.use_urey = false; }
  if (mpole_only || polar_only) o.use_chgtrn = o.use_disp = o.use_repel = false;
  if (ct_only || disp_only || rep_only) {
    o.other_valence_off = true;
    o.use_mpole = o.use_polar = o.use_vdw = false;
    o.use_bond = o.use_angle = o.use_urey = false;
    o.use_chgtrn = ct_only;
    o.use_disp = disp_only;
    o.use_repel = rep_only;
  }
  if (vdw_only) o.use_bond = o.use_angle = o.use_urey = false;
  o.use_valence = o.use_bond || o.use_angle || o.use_urey;
}

// ---------------------------------------------------------------- setup

double ewald_coeff(double cutoff, double eps) {
  // Smallest power-of-two bracket, then bisection on erfc(a*rc)/rc = eps.
  double x = 0.5, ratio = eps + 1.0;
  int i = 0;
  while (ratio >= eps) {
    ++i;
    x *= 2.0;
    ratio = std::erfc(x * cutoff) / cutoff;
  }
  double lo = 0.0, hi = x;
  for (int k = 0; k < i + 60; ++k) {
    x = 0.5 * (lo + hi);
    ratio = std::erfc(x * cutoff) / cutoff;
    if (ratio >= eps) lo = x;
    else hi = x;
  }
  return x;
}

int default_pme_grid(double box_len) {
  static const int sizes[] = {2,   4,   6,   8,   10,  12,  16,  18,  20,  24,  30,  32,  36,
                              40,  48,  50,  54,  60,  64,  72,  80,  90,  96,  100, 108, 120,
                              128, 144, 150, 160, 162, 180, 192, 200, 216, 240, 250, 256, 270,
                              288, 300, 320, 324, 360, 384, 400, 432, 450, 480, 486, 500, 512};
  int want = (int)(box_len * 1.2 - 1e-8) + 1;
  for (int s : sizes)
    if (s >= want) return s;
  return sizes[sizeof(sizes) / sizeof(int) - 1];
}

namespace {

void build_couple(System& s) {
  const int n = s.n;
  s.i13.assign(n, {});
  s.i14.assign(n, {});
  s.i15.assign(n, {});
  for (int i = 0; i < n; ++i) {
    for (int j : s.i12[i])
      for (int k : s.i12[j])
        if (k != i && !contains(s.i12[i], k)) add_unique_sorted(s.i13[i], k);
    for (int j : s.i13[i])
      for (int k : s.i12[j])
        if (k != i && !contains(s.i12[i], k) && !contains(s.i13[i], k)) add_unique_sorted(s.i14[i], k);
    for (int j : s.i14[i])
      for (int k : s.i12[j])
        if (k != i && !contains(s.i12[i], k) && !contains(s.i13[i], k) && !contains(s.i14[i], k))
          add_unique_sorted(s.i15[i], k);
  }
}

void build_polgroups(System& s) {
  const int n = s.n;
  std::vector<int> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  std::function<int(int)> find = [&](int a) { return parent[a] == a ? a : parent[a] = find(parent[a]); };
  for (int i = 0; i < n; ++i) {
    auto it = s.ff.polarize.find(s.type[i]);
    if (it == s.ff.polarize.end()) continue;
    for (int j : s.i12[i])
      if (contains(it->second.group_types, s.type[j])) parent[find(i)] = find(j);
  }
  std::vector<std::vector<int>> members(n);
  for (int i = 0; i < n; ++i) members[find(i)].push_back(i);
  s.ip11.assign(n, {});
  s.ip12.assign(n, {});
  s.ip13.assign(n, {});
  s.ip14.assign(n, {});
  for (int i = 0; i < n; ++i) s.ip11[i] = members[find(i)];
  // groups bonded to a set of groups
  auto next_shell = [&](int i, const std::vector<int>& shell, std::vector<int>& out,
                        const std::vector<const std::vector<int>*>& excl) {
    std::set<int> roots;
    for (int a : shell)
      for (int b : s.i12[a]) roots.insert(find(b));
    for (int r : roots)
      for (int m : members[r]) {
        bool skip = false;
        for (auto* e : excl)
          if (contains(*e, m)) skip = true;
        if (!skip) add_unique_sorted(out, m);
      }
    (void)i;
  };
  for (int i = 0; i < n; ++i) {
    next_shell(i, s.ip11[i], s.ip12[i], {&s.ip11[i]});
    next_shell(i, s.ip12[i], s.ip13[i], {&s.ip11[i], &s.ip12[i]});
    next_shell(i, s.ip13[i], s.ip14[i], {&s.ip11[i], &s.ip12[i], &s.ip13[i]});
  }
}

// Multipole parameter assignment: four passes, as the reference's absent
// Fortran kmpole does — (1) z, x, y all from 1-2 atoms, (2) z from 1-2 and
// x, y from 1-3 atoms, (3) z-defining atom only, (4) no connected atoms.
void assign_multipoles(System& s) {
  const int n = s.n;
  s.pole.assign(kMplTotal * n, 0.0);
  s.frame.assign(n, kFrameNone);
  s.zaxis.assign(n, -1);
  s.xaxis.assign(n, -1);
  s.yaxis.assign(n, -1);
  s.ysign.assign(n, 0);
  std::vector<char> done(n, 0);
  auto assign = [&](int i, const MultipoleParam& mp, int z, int x, int y) {
    std::copy(mp.pole, mp.pole + kMplTotal, &s.pole[kMplTotal * i]);
    s.frame[i] = mp.frame;
    s.zaxis[i] = z;
    s.xaxis[i] = x;
    s.yaxis[i] = y;
    s.ysign[i] = (mp.frame == kFrameZThenX && y >= 0) ? (mp.ytyp < 0 ? -1 : 1) : 0;
    done[i] = 1;
  };
  for (int pass = 1; pass <= 4; ++pass) {
    for (int i = 0; i < n; ++i) {
      if (done[i]) continue;
      for (const auto& mp : s.ff.mpoles) {
        if (done[i]) break;
        if (mp.type != s.type[i]) continue;
        int zt = std::abs(mp.ztyp), xt = std::abs(mp.xtyp), yt = std::abs(mp.ytyp);
        if (pass == 4) {
          if (zt == 0) assign(i, mp, -1, -1, -1);
          continue;
        }
        if (zt == 0) continue;
        for (int j : s.i12[i]) {
          if (done[i]) break;
          if (s.type[j] != zt) continue;
          if (pass == 3) {
            if (xt == 0) assign(i, mp, j, -1, -1);
            continue;
          }
          if (xt == 0) {
            if (pass == 1) assign(i, mp, j, -1, -1);
            continue;
          }
          const auto& xl = (pass == 1) ? s.i12[i] : s.i13[i];
          for (int k : xl) {
            if (done[i]) break;
            if (k == j || s.type[k] != xt) continue;
            if (yt == 0) {
              assign(i, mp, j, k, -1);
              continue;
            }
            const auto& yl = (pass == 1) ? s.i12[i] : s.i13[i];
            for (int m : yl) {
              if (m == j || m == k || s.type[m] != yt) continue;
              assign(i, mp, j, k, m);
              break;
            }
            if (!done[i] && pass == 2) {
              // y may also come from the 1-2 shell in the mixed pass
              for (int m : s.i12[i]) {
                if (m == j || m == k || s.type[m] != yt) continue;
                assign(i, mp, j, k, m);
                break;
              }
            }
          }
        }
      }
    }
  }
  // with the electrostatics switched off an atom may lack multipole records
  // (it then carries none)
  if (s.opt.use_mpole || s.opt.use_polar)
    for (int i = 0; i < n; ++i)
      if (!done[i]) throw std::runtime_error("no multipole parameters for atom " + std::to_string(i + 1));
}

void assign_polarize(System& s) {
  const int n = s.n;
  s.polarity.assign(n, 0.0);
  s.thole.assign(n, 0.0);
  s.pdamp.assign(n, 0.0);
  s.jpolar.assign(n, 0);
  std::map<int, int> tmap;
  for (int i = 0; i < n; ++i) {
    auto it = s.ff.polarize.find(s.type[i]);
    if (it != s.ff.polarize.end()) {
      s.polarity[i] = it->second.alpha;
      s.thole[i] = it->second.thole;
    }
    s.pdamp[i] = std::pow(s.polarity[i], 1.0 / 6.0);
    tmap.emplace(s.type[i], 0);
  }
  int idx = 0;
  std::vector<int> types;
  for (auto& kv : tmap) {
    kv.second = idx++;
    types.push_back(kv.first);
  }
  s.njpolar = idx;
  for (int i = 0; i < n; ++i) s.jpolar[i] = tmap[s.type[i]];
  s.thlval.assign(idx * idx, 0.0);
  for (int a = 0; a < idx; ++a)
    for (int b = 0; b < idx; ++b) {
      double ta = 0, tb = 0;
      auto ia = s.ff.polarize.find(types[a]);
      auto ib = s.ff.polarize.find(types[b]);
      if (ia != s.ff.polarize.end()) ta = ia->second.thole;
      if (ib != s.ff.polarize.end()) tb = ib->second.thole;
      double v = std::min(ta, tb);
      auto pp = s.ff.polpair.find({std::min(types[a], types[b]), std::max(types[a], types[b])});
      if (pp != s.ff.polpair.end()) v = pp->second;
      s.thlval[a * idx + b] = v;
    }
}

// HIPPO charge penetration: core charge and alpha of the atom's class, the
// valence charge is the rest of the monopole.
void assign_chgpen(System& s) {
  s.chgpen = !s.ff.chgpen.empty() && (s.opt.use_mpole || s.opt.use_polar);
  s.pcore.clear();
  s.pval.clear();
  s.palpha.clear();
  if (!s.chgpen) return;
  const int n = s.n;
  s.pcore.assign(n, 0.0);
  s.pval.assign(n, 0.0);
  s.palpha.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    s.pval[i] = s.pole[kMplTotal * i + kQ];
    auto it = s.ff.chgpen.find(s.cls[i]);
    if (it == s.ff.chgpen.end()) continue;
    s.pcore[i] = it->second.first;
    s.pval[i] -= it->second.first;
    s.palpha[i] = it->second.second;
  }
}

// HIPPO charge transfer: charge and damping exponent of the atom's class
// (src/hippo/echgtrn.cpp:58-66: a zero exponent becomes 1000).
void assign_chgtrn(System& s) {
  s.chgtrn = !s.ff.chgtrn.empty() && s.opt.use_chgtrn;
  s.chgct.clear();
  s.dmpct.clear();
  if (!s.chgtrn) return;
  s.chgct.assign(s.n, 0.0);
  s.dmpct.assign(s.n, 1000.0);
  for (int i = 0; i < s.n; ++i) {
    auto it = s.ff.chgtrn.find(s.cls[i]);
    if (it == s.ff.chgtrn.end()) continue;
    s.chgct[i] = it->second.first;
    if (it->second.second != 0) s.dmpct[i] = it->second.second;
  }
}

// HIPPO dispersion: C6 and damping exponent of the atom's class.
void assign_disp(System& s) {
  s.disp = !s.ff.disp.empty() && s.opt.use_disp;
  s.csix.clear();
  s.adisp.clear();
  if (!s.disp) return;
  if (s.opt.dewald) throw std::runtime_error("Ewald dispersion (dewald) is not evaluated: use disp-cutoff");
  s.csix.assign(s.n, 0.0);
  s.adisp.assign(s.n, 0.0);
  for (int i = 0; i < s.n; ++i) {
    auto it = s.ff.disp.find(s.cls[i]);
    if (it == s.ff.disp.end()) continue;
    s.csix[i] = it->second.first;
    s.adisp[i] = it->second.second;
  }
}

// HIPPO Pauli repulsion: size, damping exponent and valence count of the
// atom's class.
void assign_repel(System& s) {
  s.repel = !s.ff.repel.empty() && s.opt.use_repel;
  s.sizpr.clear();
  s.dmppr.clear();
  s.elepr.clear();
  if (!s.repel) return;
  s.sizpr.assign(s.n, 0.0);
  s.dmppr.assign(s.n, 0.0);
  s.elepr.assign(s.n, 0.0);
  for (int i = 0; i < s.n; ++i) {
    auto it = s.ff.repel.find(s.cls[i]);
    if (it == s.ff.repel.end()) continue;
    s.sizpr[i] = std::get<0>(it->second);
    s.dmppr[i] = std::get<1>(it->second);
    s.elepr[i] = std::get<2>(it->second);
  }
}

void build_scales(System& s) {
  const int n = s.n;
  const ForceField& ff = s.ff;
  s.scales.clear();
  for (int i = 0; i < n; ++i) {
    std::map<int, ScalePair> row;
    auto get = [&](int k) -> ScalePair& {
      auto it = row.find(k);
      if (it == row.end()) it = row.emplace(k, ScalePair{i, k, 1, 1, 1, 1, 1, 1}).first;
      return it->second;
    };
    const std::vector<int>* cpl[6] = {nullptr, nullptr, &s.i12[i], &s.i13[i], &s.i14[i], &s.i15[i]};
    for (int b = 2; b <= 5; ++b)
      for (int k : *cpl[b]) {
        if (k <= i) continue;
        if (ff.m_scale[b] != 1) get(k).m = ff.m_scale[b];
        if (s.disp && ff.ds_scale[b] != 1) get(k).ds = ff.ds_scale[b];
        if (s.repel && ff.rs_scale[b] != 1) get(k).rs = ff.rs_scale[b];
        double pv = contains(s.ip11[i], k) ? ff.p_iscale[b] : ff.p_scale[b];
        if (ff.p_scale[b] != 1 || ff.p_iscale[b] != 1) get(k).p = pv;
        if (s.chgpen) {  // HIPPO: d = p (bond-based, intra-group variant), u = induce-1x
          if (ff.p_scale[b] != 1 || ff.p_iscale[b] != 1) get(k).d = pv;
          if (ff.w_scale[b] != 1) get(k).u = ff.w_scale[b];
        }
      }
    if (s.chgpen) {
      for (auto& kv : row) {
        const ScalePair& p = kv.second;
        if (p.m != 1 || p.d != 1 || p.p != 1 || p.u != 1 || p.ds != 1 || p.rs != 1) s.scales.push_back(p);
      }
      continue;
    }
    const std::vector<int>* grp[5] = {nullptr, &s.ip11[i], &s.ip12[i], &s.ip13[i], &s.ip14[i]};
    for (int b = 1; b <= 4; ++b)
      for (int k : *grp[b]) {
        if (k <= i) continue;
        if (ff.d_scale[b] != 1) get(k).d = ff.d_scale[b];
        if (ff.u_scale[b] != 1) get(k).u = ff.u_scale[b];
      }
    for (auto& kv : row) {
      const ScalePair& p = kv.second;
      if (p.m != 1 || p.d != 1 || p.p != 1 || p.u != 1 || p.ds != 1 || p.rs != 1) s.scales.push_back(p);
    }
  }
}

void assign_vdw(System& s) {
  const int n = s.n;
  s.vdw_rad.assign(n, 0.0);
  s.vdw_eps.assign(n, 0.0);
  s.vdw_red.assign(n, 0.0);
  s.vdw_parent.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    s.vdw_parent[i] = i;
    auto it = s.ff.vdw.find(s.cls[i]);
    if (it == s.ff.vdw.end()) continue;
    s.vdw_rad[i] = it->second.rad;
    s.vdw_eps[i] = it->second.eps;
    s.vdw_red[i] = it->second.red;
    if (it->second.red != 0 && s.i12[i].size() == 1) s.vdw_parent[i] = s.i12[i][0];
  }
}

// vdW tables: one index per distinct (radius, epsilon) atom class, pair
// R_min by the cubic-mean rule on R-min diameters and epsilon by the HHG rule
// (amoeba09.prm: radiusrule CUBIC-MEAN, radiustype R-MIN, radiussize
// DIAMETER, epsilonrule HHG).  vdW pair scales from 1-2 .. 1-5 connectivity.
void build_vdw_tables(System& s) {
  const int n = s.n;
  std::map<std::pair<double, double>, int> idx;
  s.jvdw.assign(n, 0);
  std::vector<std::pair<double, double>> cls;
  for (int i = 0; i < n; ++i) {
    auto key = std::make_pair(s.vdw_rad[i], s.vdw_eps[i]);
    auto it = idx.find(key);
    if (it == idx.end()) {
      it = idx.emplace(key, (int)cls.size()).first;
      cls.push_back(key);
    }
    s.jvdw[i] = it->second;
  }
  const int m = (int)cls.size();
  s.njvdw = m;
  s.vdw_rmin.assign((size_t)m * m, 0.0);
  s.vdw_epsmix.assign((size_t)m * m, 0.0);
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      const double ra = cls[a].first, rb = cls[b].first, ea = cls[a].second, eb = cls[b].second;
      const double r2 = ra * ra + rb * rb;
      s.vdw_rmin[a * m + b] = r2 > 0 ? (ra * ra * ra + rb * rb * rb) / r2 : 0.0;
      const double se = std::sqrt(ea) + std::sqrt(eb);
      s.vdw_epsmix[a * m + b] = se > 0 ? 4 * ea * eb / (se * se) : 0.0;
    }
  s.vscale_ik.clear();
  s.vscale_s.clear();
  for (int i = 0; i < n; ++i) {
    const std::vector<int>* cpl[6] = {nullptr, nullptr, &s.i12[i], &s.i13[i], &s.i14[i], &s.i15[i]};
    std::map<int, double> row;
    for (int b = 2; b <= 5; ++b)
      for (int k : *cpl[b])
        if (k > i && s.ff.v_scale[b] != 1 && !row.count(k)) row[k] = s.ff.v_scale[b];
    for (auto& kv : row) {
      s.vscale_ik.push_back(i);
      s.vscale_ik.push_back(kv.first);
      s.vscale_s.push_back(kv.second);
    }
  }
}

// Bonds from the 1-2 lists, angles and Urey-Bradley terms from pairs of
// neighbours of each vertex, parameters by atom classes.
void build_valence(System& s) {
  const int n = s.n;
  const ForceField& ff = s.ff;
  s.bond_ij.clear();
  s.bond_kr.clear();
  s.angle_abc.clear();
  s.angle_kt.clear();
  s.urey_ac.clear();
  s.urey_kr.clear();
  for (int i = 0; i < n; ++i)
    for (int j : s.i12[i]) {
      if (j <= i) continue;
      if (!s.opt.use_bond) continue;
      auto it = ff.bond.find({std::min(s.cls[i], s.cls[j]), std::max(s.cls[i], s.cls[j])});
      if (it == ff.bond.end()) throw std::runtime_error("no bond parameters for atoms " + std::to_string(i + 1));
      s.bond_ij.insert(s.bond_ij.end(), {i, j});
      s.bond_kr.insert(s.bond_kr.end(), {it->second.k, it->second.r0});
    }
  for (int b = 0; b < n; ++b) {
    const auto& nb = s.i12[b];
    for (size_t p = 0; p < nb.size(); ++p)
      for (size_t q = p + 1; q < nb.size(); ++q) {
        const int a = nb[p], c = nb[q];
        const auto key = std::make_tuple(std::min(s.cls[a], s.cls[c]), s.cls[b], std::max(s.cls[a], s.cls[c]));
        auto it = ff.angle.find(key);
        if (s.opt.use_angle) {
          if (it == ff.angle.end()) throw std::runtime_error("no angle parameters at atom " + std::to_string(b + 1));
          s.angle_abc.insert(s.angle_abc.end(), {a, b, c});
          s.angle_kt.insert(s.angle_kt.end(), {it->second.k, it->second.t0});
        }
        auto iu = ff.urey.find(key);
        if (s.opt.use_urey && iu != ff.urey.end()) {
          s.urey_ac.insert(s.urey_ac.end(), {a, c});
          s.urey_kr.insert(s.urey_kr.end(), {iu->second.k, iu->second.r0});
        }
      }
  }
}

}  // namespace

void setup_system(System& s) {
  const int n = s.n;
  s.cls.assign(n, 0);
  s.mass.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    auto it = s.ff.type_class.find(s.type[i]);
    s.cls[i] = it == s.ff.type_class.end() ? s.type[i] : it->second;
    auto im = s.ff.type_mass.find(s.type[i]);
    s.mass[i] = im == s.ff.type_mass.end() ? 0.0 : im->second;
  }
  // symmetrise 1-2 lists
  for (int i = 0; i < n; ++i)
    for (int j : std::vector<int>(s.i12[i])) add_unique_sorted(s.i12[j], i);
  build_couple(s);
  build_polgroups(s);
  assign_multipoles(s);
  assign_polarize(s);
  assign_chgpen(s);
  assign_chgtrn(s);
  assign_disp(s);
  assign_repel(s);
  build_scales(s);
  assign_vdw(s);
  build_vdw_tables(s);
  build_valence(s);
  Options& o = s.opt;
  if (o.use_valence && !o.other_valence_off) {
    // bond, angle and Urey-Bradley are the valence terms evaluated here; an
    // atom with three bonds (out-of-plane, stretch-bend) or a 1-4 pair
    // (torsions) would need terms that are not
    for (int i = 0; i < n; ++i)
      if (s.i12[i].size() >= 3 || !s.i14[i].empty())
        throw std::runtime_error(
            "valence terms beyond bond / angle / Urey-Bradley are not evaluated: switch them off "
            "(strbndterm, opbendterm, torsionterm, pitorsterm, tortorterm, angtorterm, strtorterm none) "
            "or switch the valence terms off (bondterm none)");
  }
  if (o.vdw_taper <= 0) o.vdw_taper = 0.9 * o.vdw_cut;
  // HIPPO switching distances: 0.9 of the cutoff by default, a value up to 1
  // a fraction of the cutoff, a larger one a distance
  for (auto tc : {std::make_pair(&o.ct_taper, o.ct_cut), std::make_pair(&o.disp_taper, o.disp_cut),
                  std::make_pair(&o.rep_taper, o.rep_cut)}) {
    double& t = *tc.first;
    if (t <= 0) t = 0.9 * tc.second;
    else if (t <= 1) t *= tc.second;
  }
  if (o.ewald) {
    if (!s.periodic) throw std::runtime_error("ewald requires a periodic box");
    if (o.ewald_alpha <= 0) o.ewald_alpha = ewald_coeff(o.ewald_cut);
    for (int d = 0; d < 3; ++d)
      if (o.pme_grid[d] <= 0) o.pme_grid[d] = default_pme_grid(s.box[d]);
    if (o.pme_order != 5) throw std::runtime_error("pme-order must be 5");
  }
}

void build_water_box(System& s, int nmol, int nside, double L, uint64_t seed, int n_na, int n_cl,
                     double ion_min) {
  const int nsite = nside * nside * nside;
  if (nmol > nsite) throw std::runtime_error("more molecules than lattice sites");
  Rng rng(seed);
  std::vector<int> sites(nsite);
  std::iota(sites.begin(), sites.end(), 0);
  std::vector<int> used(sites);
  if (nmol < nsite) {
    for (int i = nsite - 1; i > 0; --i) {
      int j = (int)(rng.next() % (uint64_t)(i + 1));
      std::swap(used[i], used[j]);
    }
    std::vector<int> unused(used.begin() + nmol, used.end());
    used.resize(nmol);
    std::sort(used.begin(), used.end());
    std::sort(unused.begin(), unused.end());
    sites = unused;
  } else {
    sites.clear();
  }
  const double a = L / nside;
  const double roh = 0.9572, half = 0.5 * 104.52 * M_PI / 180.0;
  const double hb[2][3] = {{roh * std::sin(half), 0, roh * std::cos(half)},
                           {-roh * std::sin(half), 0, roh * std::cos(half)}};
  s.n = 3 * nmol + n_na + n_cl;
  s.xyz.assign(3 * s.n, 0.0);
  s.type.assign(s.n, 0);
  s.name.assign(s.n, "");
  s.i12.assign(s.n, {});
  auto center = [&](int site, double* c) {
    int ix = site % nside, iy = (site / nside) % nside, iz = site / (nside * nside);
    c[0] = (ix + 0.5) * a;
    c[1] = (iy + 0.5) * a;
    c[2] = (iz + 0.5) * a;
  };
  for (int m = 0; m < nmol; ++m) {
    double q[4];
    double qn = 0;
    for (double& v : q) {
      v = rng.gauss();
      qn += v * v;
    }
    qn = 1.0 / std::sqrt(qn);
    for (double& v : q) v *= qn;
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                            {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                            {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
    double c[3];
    center(used[m], c);
    for (int d = 0; d < 3; ++d) c[d] += 0.2 * rng.uniform() - 0.1;
    const int o = 3 * m;
    for (int d = 0; d < 3; ++d) s.xyz[3 * o + d] = c[d];
    for (int h = 0; h < 2; ++h)
      for (int d = 0; d < 3; ++d)
        s.xyz[3 * (o + 1 + h) + d] = c[d] + R[d][0] * hb[h][0] + R[d][1] * hb[h][1] + R[d][2] * hb[h][2];
    s.type[o] = 36;
    s.type[o + 1] = s.type[o + 2] = 37;
    s.name[o] = "O";
    s.name[o + 1] = s.name[o + 2] = "H";
    s.i12[o] = {o + 1, o + 2};
    s.i12[o + 1] = {o};
    s.i12[o + 2] = {o};
  }
  if (n_na + n_cl > 0) {
    for (int i = (int)sites.size() - 1; i > 0; --i) {
      int j = (int)(rng.next() % (uint64_t)(i + 1));
      std::swap(sites[i], sites[j]);
    }
    // Sites at least ion_min from every oxygen are taken in shuffled order;
    // if there are too few of them, the remaining ions go to the free sites
    // farthest from any oxygen (descending distance, ties by shuffled order).
    const int want = n_na + n_cl;
    std::vector<std::pair<double, int>> far;  // (-min distance, shuffled rank)
    std::vector<int> chosen;
    for (size_t r = 0; r < sites.size(); ++r) {
      double c[3];
      center(sites[r], c);
      double md2 = 1e300;
      for (int m = 0; m < nmol; ++m) {
        double r2 = 0;
        for (int d = 0; d < 3; ++d) {
          double dx = s.xyz[9 * m + d] - c[d];
          dx -= L * std::nearbyint(dx / L);
          r2 += dx * dx;
        }
        md2 = std::min(md2, r2);
      }
      if (md2 >= ion_min * ion_min && (int)chosen.size() < want) chosen.push_back((int)r);
      else far.push_back({-md2, (int)r});
    }
    std::sort(far.begin(), far.end());
    for (size_t q = 0; q < far.size() && (int)chosen.size() < want; ++q) chosen.push_back(far[q].second);
    if ((int)chosen.size() != want) throw std::runtime_error("not enough free lattice sites for the ions");
    for (int placed = 0; placed < want; ++placed) {
      double c[3];
      center(sites[chosen[placed]], c);
      const int idx = 3 * nmol + placed;
      for (int d = 0; d < 3; ++d) s.xyz[3 * idx + d] = c[d];
      const bool na = placed < n_na;
      s.type[idx] = na ? 7 : 15;
      s.name[idx] = na ? "Na+" : "Cl-";
    }
  }
  s.box[0] = s.box[1] = s.box[2] = L;
  s.periodic = true;
  s.title = "seeded AMOEBA water box";
}

}  // namespace t9
