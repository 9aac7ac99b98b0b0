// C ABI for the host-side system (t9_system_*): setup from Tinker files or
// the seeded water-box generator, and read access to the derived arrays.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../../include/t9hip.h"
#include "errors.hpp"
#include "system.hpp"

struct t9_system {
  t9::System s;
};

namespace {

std::string dirname_of(const std::string& p) {
  auto k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

int finish_setup(t9_system* out, const char* prm_path, const char* key_text) {
  std::string prm_name;
  if (key_text) t9::read_key_text(key_text, out->s, &prm_name);
  if (!prm_path || !*prm_path) throw std::runtime_error("parameter file path required");
  t9::read_prm(prm_path, out->s.ff);
  if (key_text) t9::read_prm_text(key_text, out->s.ff);
  t9::setup_system(out->s);
  return T9_OK;
}

}  // namespace

extern "C" {

int t9_system_from_files(const char* xyz_path, const char* prm_path, const char* key_text, t9_system** out) {
  return t9::guarded([&] {
    auto* sys = new t9_system;
    try {
      t9::read_xyz(xyz_path, sys->s);
      finish_setup(sys, prm_path, key_text);
    } catch (...) {
      delete sys;
      throw;
    }
    *out = sys;
    return T9_OK;
  }, T9_ERR_SETUP);
}

int t9_system_water_box(int nmol, int nside, double box_len, uint64_t seed, int n_na, int n_cl,
                        const char* prm_path, const char* key_text, t9_system** out) {
  return t9::guarded([&] {
    auto* sys = new t9_system;
    try {
      t9::build_water_box(sys->s, nmol, nside, box_len, seed, n_na, n_cl);
      finish_setup(sys, prm_path, key_text);
    } catch (...) {
      delete sys;
      throw;
    }
    *out = sys;
    return T9_OK;
  }, T9_ERR_SETUP);
}

void t9_system_destroy(t9_system* sys) { delete sys; }

int t9_system_natoms(const t9_system* sys) { return sys ? sys->s.n : 0; }

int t9_system_get_f64(const t9_system* sys, const char* name, double* out) {
  return t9::guarded([&] {
    const t9::System& s = sys->s;
    std::string k = name;
    const std::vector<double>* v = nullptr;
    if (k == "xyz") v = &s.xyz;
    else if (k == "pole") v = &s.pole;
    else if (k == "polarity") v = &s.polarity;
    else if (k == "pdamp") v = &s.pdamp;
    else if (k == "thole") v = &s.thole;
    else if (k == "pcore") v = &s.pcore;
    else if (k == "pval") v = &s.pval;
    else if (k == "palpha") v = &s.palpha;
    else if (k == "chgct") v = &s.chgct;
    else if (k == "dmpct") v = &s.dmpct;
    else if (k == "csix") v = &s.csix;
    else if (k == "adisp") v = &s.adisp;
    else if (k == "sizpr") v = &s.sizpr;
    else if (k == "dmppr") v = &s.dmppr;
    else if (k == "elepr") v = &s.elepr;
    else if (k == "scale_rs") {  // repulsion scale of each pair of t9_system_get_scales
      for (size_t j = 0; j < s.scales.size(); ++j) out[j] = s.scales[j].rs;
      return T9_OK;
    }
    else if (k == "scale_ds") {  // dispersion scale of each pair of t9_system_get_scales
      for (size_t j = 0; j < s.scales.size(); ++j) out[j] = s.scales[j].ds;
      return T9_OK;
    }
    else if (k == "mass") v = &s.mass;
    else if (k == "thlval") v = &s.thlval;
    else if (k == "vdw_rad") v = &s.vdw_rad;
    else if (k == "vdw_eps") v = &s.vdw_eps;
    else if (k == "vdw_red") v = &s.vdw_red;
    else if (k == "vdw_rmin") v = &s.vdw_rmin;
    else if (k == "vdw_epsmix") v = &s.vdw_epsmix;
    else if (k == "vscale_s") v = &s.vscale_s;
    else if (k == "bond_kr") v = &s.bond_kr;
    else if (k == "angle_kt") v = &s.angle_kt;
    else if (k == "urey_kr") v = &s.urey_kr;
    else if (k == "box") {
      std::memcpy(out, s.box, 3 * sizeof(double));
      return T9_OK;
    } else {
      throw t9::ArgError("unknown f64 array " + k);
    }
    std::memcpy(out, v->data(), v->size() * sizeof(double));
    return T9_OK;
  }, T9_ERR_ARG);
}

int t9_system_get_i32(const t9_system* sys, const char* name, int* out) {
  return t9::guarded([&] {
    const t9::System& s = sys->s;
    std::string k = name;
    const std::vector<int>* v = nullptr;
    if (k == "type") v = &s.type;
    else if (k == "frame") v = &s.frame;
    else if (k == "zaxis") v = &s.zaxis;
    else if (k == "xaxis") v = &s.xaxis;
    else if (k == "yaxis") v = &s.yaxis;
    else if (k == "ysign") v = &s.ysign;
    else if (k == "jpolar") v = &s.jpolar;
    else if (k == "vdw_parent") v = &s.vdw_parent;
    else if (k == "jvdw") v = &s.jvdw;
    else if (k == "vscale_ik") v = &s.vscale_ik;
    else if (k == "bond_ij") v = &s.bond_ij;
    else if (k == "angle_abc") v = &s.angle_abc;
    else if (k == "urey_ac") v = &s.urey_ac;
    else if (k == "i12") {
      for (int i = 0; i < s.n; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = j < (int)s.i12[i].size() ? s.i12[i][j] : -1;
      return T9_OK;
    } else {
      throw t9::ArgError("unknown i32 array " + k);
    }
    std::memcpy(out, v->data(), v->size() * sizeof(int));
    return T9_OK;
  }, T9_ERR_ARG);
}

int t9_system_get_scalar(const t9_system* sys, const char* name, double* out) {
  return t9::guarded([&] {
    const t9::System& s = sys->s;
    const t9::Options& o = s.opt;
    std::string k = name;
    double v;
    if (k == "aewald") v = o.ewald_alpha;
    else if (k == "ewald_cut") v = o.ewald_cut;
    else if (k == "ewald") v = o.ewald;
    else if (k == "grid1") v = o.pme_grid[0];
    else if (k == "grid2") v = o.pme_grid[1];
    else if (k == "grid3") v = o.pme_grid[2];
    else if (k == "pme_order") v = o.pme_order;
    else if (k == "njpolar") v = s.njpolar;
    else if (k == "nscale") v = (double)s.scales.size();
    else if (k == "polar_eps") v = o.polar_eps;
    else if (k == "polar_iter") v = o.polar_iter;
    else if (k == "pcg_prec") v = o.pcg_prec;
    else if (k == "pcg_guess") v = o.pcg_guess;
    else if (k == "pcg_peek") v = o.pcg_peek;
    else if (k == "usolve_cut") v = o.usolve_cut;
    else if (k == "udiag") v = o.udiag;
    else if (k == "list_buffer") v = o.list_buffer;
    else if (k == "periodic") v = s.periodic;
    else if (k == "use_mpole") v = o.use_mpole;
    else if (k == "use_polar") v = o.use_polar;
    else if (k == "chgpen") v = s.chgpen;
    else if (k == "chgtrn") v = s.chgtrn;
    else if (k == "ct_cut") v = o.ct_cut;
    else if (k == "ct_taper") v = o.ct_taper;
    else if (k == "disp") v = s.disp;
    else if (k == "disp_cut") v = o.disp_cut;
    else if (k == "disp_taper") v = o.disp_taper;
    else if (k == "repel") v = s.repel;
    else if (k == "rep_cut") v = o.rep_cut;
    else if (k == "rep_taper") v = o.rep_taper;
    else if (k == "mpole_cut") v = o.mpole_cut;
    else if (k == "vdw_cut") v = o.vdw_cut;
    else if (k == "vdw_taper") v = o.vdw_taper;
    else if (k == "use_vdw") v = o.use_vdw;
    else if (k == "use_valence") v = o.use_valence;
    else if (k == "njvdw") v = s.njvdw;
    else if (k == "nvscale") v = (double)s.vscale_s.size();
    else if (k == "nbond") v = (double)s.bond_kr.size() / 2;
    else if (k == "nangle") v = (double)s.angle_kt.size() / 2;
    else if (k == "nurey") v = (double)s.urey_kr.size() / 2;
    else if (k == "bond_cubic") v = s.ff.bond_cubic;
    else if (k == "bond_quartic") v = s.ff.bond_quartic;
    else if (k == "angle_cubic") v = s.ff.angle_cubic;
    else if (k == "angle_quartic") v = s.ff.angle_quartic;
    else if (k == "angle_pentic") v = s.ff.angle_pentic;
    else if (k == "angle_sextic") v = s.ff.angle_sextic;
    else throw t9::ArgError("unknown scalar " + k);
    *out = v;
    return T9_OK;
  }, T9_ERR_ARG);
}

int t9_system_set_scalar(t9_system* sys, const char* name, double value) {
  return t9::guarded([&] {
    t9::Options& o = sys->s.opt;
    std::string k = name;
    if (k == "aewald") o.ewald_alpha = value;
    else if (k == "ewald_cut") o.ewald_cut = value;
    else if (k == "polar_eps") o.polar_eps = value;
    else if (k == "polar_iter") o.polar_iter = (int)value;
    else if (k == "pcg_prec") o.pcg_prec = value != 0;
    else if (k == "pcg_guess") o.pcg_guess = value != 0;
    else if (k == "pcg_peek") o.pcg_peek = value;
    else if (k == "usolve_cut") o.usolve_cut = value;
    else if (k == "udiag") o.udiag = value;
    else if (k == "list_buffer") o.list_buffer = value;
    else if (k == "mpole_cut") o.mpole_cut = value;
    else if (k == "vdw_cut") o.vdw_cut = value;
    else throw t9::ArgError("scalar not settable: " + k);
    return T9_OK;
  }, T9_ERR_ARG);
}

int t9_system_get_scales(const t9_system* sys, int* ik, double* mdpu) {
  return t9::guarded([&] {
    const auto& sc = sys->s.scales;
    for (size_t j = 0; j < sc.size(); ++j) {
      ik[2 * j] = sc[j].i;
      ik[2 * j + 1] = sc[j].k;
      mdpu[4 * j] = sc[j].m;
      mdpu[4 * j + 1] = sc[j].d;
      mdpu[4 * j + 2] = sc[j].p;
      mdpu[4 * j + 3] = sc[j].u;
    }
    return T9_OK;
  }, T9_ERR_ARG);
}

int t9_system_set_xyz(t9_system* sys, const double* xyz) {
  return t9::guarded([&] {
    std::memcpy(sys->s.xyz.data(), xyz, sys->s.xyz.size() * sizeof(double));
    return T9_OK;
  }, T9_ERR_ARG);
}

int t9_system_write_frame(const t9_system* sys, const char* path, const double* values, int append) {
  return t9::guarded([&] {
    if (!sys || !path) throw t9::ArgError("null argument");
    const t9::System& s = sys->s;
    const double* v = values ? values : s.xyz.data();
    std::FILE* f = std::fopen(path, append ? "a" : "w");
    if (!f) throw t9::ArgError(std::string("cannot open ") + path);
    std::fprintf(f, "%6d  %s\n", s.n, s.title.c_str());
    if (s.periodic)
      std::fprintf(f, " %12.6f%12.6f%12.6f%12.6f%12.6f%12.6f\n", s.box[0], s.box[1], s.box[2], 90.0, 90.0, 90.0);
    for (int i = 0; i < s.n; ++i) {
      const std::string name = s.name[i].substr(0, 3);
      std::fprintf(f, "%6d  %-3s%12.6f%12.6f%12.6f%6d", i + 1, name.c_str(), v[3 * i], v[3 * i + 1], v[3 * i + 2],
                   s.type[i]);
      for (int j : s.i12[i]) std::fprintf(f, "%6d", j + 1);
      std::fprintf(f, "\n");
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    if (!ok) throw t9::ArgError(std::string("write failed: ") + path);
    return T9_OK;
  }, T9_ERR_ARG);
}

}  // extern "C"

// Accessor used by the device context (ctx.cpp).
const t9::System& t9_system_ref(const t9_system* sys) { return sys->s; }
