"""Python binding of libt9hip.so (the C ABI declared in include/t9hip.h).

Host-side mirror of the reference's entry points for the AMOEBA electrostatics
path: `System` is the setup (keywords, parameters, topology), `Context` holds
the device state and exposes mpole_init / dfield / ufield / induce /
nblist_refresh with the reference's argument meaning.  Errors from the C ABI
raise `T9Error` (with `.code`), mirroring the reference's TINKER_THROW paths.
This module never falls back to any CPU computation: without the built
library, or without a GPU, device calls fail loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libt9hip.so")
PARAMS = os.path.join(HERE, "params", "amoeba09.prm")

T9_OK, T9_ERR_HIP, T9_ERR_NOT_CONVERGED, T9_ERR_LIST_OVERFLOW, T9_ERR_ARG, T9_ERR_SETUP = range(6)
T9_FP32, T9_FP64 = 32, 64

# Every symbol include/t9hip.h declares (checked by the CPU test suite).
EXPORTS = [
    "t9_last_error", "t9_system_from_files", "t9_system_water_box", "t9_system_destroy", "t9_system_natoms",
    "t9_system_get_f64", "t9_system_get_i32", "t9_system_get_scalar", "t9_system_get_scales",
    "t9_system_set_xyz", "t9_system_set_scalar", "t9_system_write_frame", "t9_ctx_create", "t9_ctx_destroy", "t9_set_coords",
    "t9_get_coords", "t9_nblist_refresh", "t9_nblist_build", "t9_nblist_stats", "t9_mpole_init",
    "t9_get_rpole", "t9_dfield", "t9_ufield", "t9_induce", "t9_induce_trace", "t9_get_uind", "t9_set_uind", "t9_sync",
    "t9_kernel_timing", "t9_cutoff_pairs", "t9_copy_probe", "t9_energy", "t9_md_init", "t9_md_set_velocities",
    "t9_md_get_velocities", "t9_md_run", "t9_guess_nparam", "t9_guess_set_params", "t9_guess_get_params",
    "t9_guess_enable", "t9_guess_capture", "t9_guess_stats", "t9_guess_train", "t9_guess_mlp_forward",
    "t9_guess_mlp_grad", "t9_guess_adam", "t9_guess_features", "t9_guess_history",
    "t9_energy_terms", "t9_set_coords_device", "t9_get_coords_device", "t9_energy_device", "t9_diag_precond",
    "t9_sparse_precond_apply", "t9_pme_grid_mpole", "t9_pme_grid_uind", "t9_pme_fft_front", "t9_pme_fft_back",
    "t9_pme_conv", "t9_pme_fphi_mpole", "t9_pme_fphi_uind", "t9_fft3d", "t9_fft3d_noalloc", "t9_nblist_pairs", "t9_md_bussi",
]
K_UFIELD, K_DFIELD, K_SPREAD, K_FFT, K_GATHER, K_PRECOND, K_PCG_VEC, K_EMPLAR, K_VDW, K_VALENCE = range(10)
CALC_ENERGY, CALC_GRAD, CALC_VIRIAL, CALC_ANALYZE = 1, 2, 4, 8
TERM_ELEC, TERM_VDW, TERM_VALENCE, TERM_CHGTRN, TERM_DISP, TERM_REPEL, TERM_ALL = 1, 2, 4, 8, 16, 32, 63


class T9Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"t9hip error {code}: {msg}")
        self.code = code


class InduceStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("predictor", ctypes.c_int), ("rms_debye", ctypes.c_double),
                ("guess_rms", ctypes.c_double)]


class EnergyOut(ctypes.Structure):
    _fields_ = [("esum", ctypes.c_double), ("em", ctypes.c_double), ("ep", ctypes.c_double),
                ("em_real", ctypes.c_double), ("em_recip", ctypes.c_double), ("em_self", ctypes.c_double),
                ("vir", ctypes.c_double * 9), ("induce_iterations", ctypes.c_int), ("evdw", ctypes.c_double),
                ("ebond", ctypes.c_double), ("eangle", ctypes.c_double), ("eurey", ctypes.c_double),
                ("nem", ctypes.c_int64), ("nep", ctypes.c_int64), ("nev", ctypes.c_int64),
                ("ect", ctypes.c_double), ("nct", ctypes.c_int64), ("edisp", ctypes.c_double),
                ("ndisp", ctypes.c_int64), ("erep", ctypes.c_double), ("nrep", ctypes.c_int64)]


class MdStats(ctypes.Structure):
    _fields_ = [("epot", ctypes.c_double), ("ekin", ctypes.c_double), ("temperature", ctypes.c_double),
                ("induce_iterations", ctypes.c_int64), ("list_builds", ctypes.c_int), ("steps", ctypes.c_int)]


_LIB = None
_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.t9_last_error.restype = ctypes.c_char_p
        L.t9_system_from_files.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_vp)]
        L.t9_system_water_box.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.POINTER(_vp)]
        L.t9_system_destroy.argtypes = [_vp]
        L.t9_system_destroy.restype = None
        L.t9_system_natoms.argtypes = [_vp]
        L.t9_system_get_f64.argtypes = [_vp, ctypes.c_char_p, _d]
        L.t9_system_get_i32.argtypes = [_vp, ctypes.c_char_p, _i]
        L.t9_system_get_scalar.argtypes = [_vp, ctypes.c_char_p, _d]
        L.t9_system_get_scales.argtypes = [_vp, _i, _d]
        L.t9_system_set_xyz.argtypes = [_vp, _d]
        L.t9_system_set_scalar.argtypes = [_vp, ctypes.c_char_p, ctypes.c_double]
        L.t9_system_write_frame.argtypes = [_vp, ctypes.c_char_p, _d, ctypes.c_int]
        for name, args in [
            ("t9_ctx_create", [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
            ("t9_ctx_destroy", [_vp]),
            ("t9_set_coords", [_vp, _d]),
            ("t9_get_coords", [_vp, _d]),
            ("t9_nblist_refresh", [_vp, _i]),
            ("t9_nblist_build", [_vp]),
            ("t9_nblist_stats", [_vp, ctypes.POINTER(ctypes.c_int64)]),
            ("t9_mpole_init", [_vp]),
            ("t9_get_rpole", [_vp, _d]),
            ("t9_dfield", [_vp, _d, _d]),
            ("t9_ufield", [_vp, _d, _d, _d, _d]),
            ("t9_induce", [_vp, _d, _d, ctypes.POINTER(InduceStats)]),
            ("t9_induce_trace", [_vp, _d, ctypes.c_int, _i]),
            ("t9_get_uind", [_vp, _d, _d]),
            ("t9_set_uind", [_vp, _d, _d]),
            ("t9_sync", [_vp]),
            ("t9_kernel_timing", [_vp, ctypes.c_int, ctypes.c_int, _d, ctypes.POINTER(ctypes.c_int64)]),
            ("t9_cutoff_pairs", [_vp, ctypes.POINTER(ctypes.c_int64)]),
            ("t9_copy_probe", [ctypes.c_int64, ctypes.c_int, _d]),
            ("t9_energy", [_vp, ctypes.c_int, ctypes.POINTER(EnergyOut), _d]),
            ("t9_md_init", [_vp, ctypes.c_double, ctypes.c_uint64]),
            ("t9_md_set_velocities", [_vp, _d]),
            ("t9_md_get_velocities", [_vp, _d]),
            ("t9_md_run", [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                           ctypes.POINTER(MdStats)]),
            ("t9_guess_nparam", []),
            ("t9_guess_set_params", [_vp, _f, _f, ctypes.c_uint64]),
            ("t9_guess_get_params", [_vp, _f, _f]),
            ("t9_guess_enable", [_vp, ctypes.c_int]),
            ("t9_guess_capture", [_vp, ctypes.c_int, ctypes.c_int64]),
            ("t9_guess_stats", [_vp, ctypes.POINTER(ctypes.c_int64)]),
            ("t9_guess_train", [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _d, _d]),
            ("t9_guess_mlp_forward", [_vp, _f, ctypes.c_int, _f]),
            ("t9_guess_mlp_grad", [_vp, _f, _f, ctypes.c_int, _f, _d]),
            ("t9_guess_adam", [_vp, _f, ctypes.c_double]),
            ("t9_guess_features", [_vp, _f]),
            ("t9_guess_history", [_vp, ctypes.c_int, ctypes.c_int, _d]),
            ("t9_energy_terms", [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(EnergyOut), _d]),
            ("t9_set_coords_device", [_vp, _vp]),
            ("t9_get_coords_device", [_vp, _vp]),
            ("t9_energy_device", [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(EnergyOut), _vp]),
            ("t9_diag_precond", [_vp, _d, _d, _d, _d]),
            ("t9_sparse_precond_apply", [_vp, _d, _d, _d, _d]),
            ("t9_pme_grid_mpole", [_vp, _d]),
            ("t9_pme_grid_uind", [_vp, _d, _d, _d]),
            ("t9_pme_fft_front", [_vp, _d]),
            ("t9_pme_fft_back", [_vp, _d]),
            ("t9_pme_conv", [_vp, _d]),
            ("t9_pme_fphi_mpole", [_vp, _d, _d]),
            ("t9_pme_fphi_uind", [_vp, _d, _d, _d]),
            ("t9_fft3d", [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            ("t9_fft3d_noalloc", [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            ("t9_nblist_pairs", [_vp, ctypes.c_int, ctypes.c_int64, _i, ctypes.POINTER(ctypes.c_int64)]),
            ("t9_md_bussi", [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double]),
        ]:
            if hasattr(L, name):  # device entry points are absent from a host-only build
                getattr(L, name).argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != T9_OK:
        raise T9Error(rc, lib().t9_last_error().decode())


def _p(a, t=_d):
    return a.ctypes.data_as(t) if a is not None else None


def _pf(a):
    return _p(a, _f)


class System:
    """Host-side system: coordinates, parameters, keywords and topology."""

    def __init__(self, handle):
        self.h = handle
        self.n = lib().t9_system_natoms(handle)

    @classmethod
    def from_files(cls, xyz, prm=PARAMS, key=""):
        h = _vp()
        check(lib().t9_system_from_files(xyz.encode(), prm.encode(), key.encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def water_box(cls, nmol, nside, box_len, seed, n_na=0, n_cl=0, prm=PARAMS, key=""):
        h = _vp()
        check(lib().t9_system_water_box(nmol, nside, box_len, seed, n_na, n_cl, prm.encode(), key.encode(),
                                        ctypes.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.t9_system_destroy(self.h)
            self.h = None

    def f64(self, name, shape):
        out = np.zeros(shape, dtype=np.float64)
        check(lib().t9_system_get_f64(self.h, name.encode(), _p(out)))
        return out

    def i32(self, name, shape):
        out = np.zeros(shape, dtype=np.int32)
        check(lib().t9_system_get_i32(self.h, name.encode(), _p(out, _i)))
        return out

    def scalar(self, name):
        v = ctypes.c_double()
        check(lib().t9_system_get_scalar(self.h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_scalar(self, name, value):
        check(lib().t9_system_set_scalar(self.h, name.encode(), float(value)))

    def write_frame(self, path, values=None, append=False):
        """One frame in Tinker's .xyz / dump layout (values [n,3] in caller
        order: coordinates by default, or dipoles in Debye, fields, ...)."""
        v = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        check(lib().t9_system_write_frame(self.h, str(path).encode(), _p(v) if v is not None else None,
                                          1 if append else 0))

    def set_xyz(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64)
        check(lib().t9_system_set_xyz(self.h, _p(a)))

    def scales(self):
        ns = int(self.scalar("nscale"))
        ik = np.zeros((ns, 2), dtype=np.int32)
        mdpu = np.zeros((ns, 4))
        if ns:
            check(lib().t9_system_get_scales(self.h, _p(ik, _i), _p(mdpu)))
        return ik, mdpu

    def arrays(self):
        """Everything the oracle needs, as numpy arrays and scalars."""
        n = self.n
        nj = int(self.scalar("njpolar"))
        ik, mdpu = self.scales()
        d = dict(n=n, xyz=self.f64("xyz", (n, 3)), pole=self.f64("pole", (n, 10)),
                 polarity=self.f64("polarity", n), pdamp=self.f64("pdamp", n), thole=self.f64("thole", n),
                 thlval=self.f64("thlval", nj * nj), njpolar=nj, frame=self.i32("frame", n),
                 zaxis=self.i32("zaxis", n), xaxis=self.i32("xaxis", n), yaxis=self.i32("yaxis", n),
                 ysign=self.i32("ysign", n), jpolar=self.i32("jpolar", n), type=self.i32("type", n),
                 box=self.f64("box", 3), scale_ik=ik, scale_mdpu=mdpu, nscale=len(ik))
        for k in ("aewald", "ewald_cut", "ewald", "pme_order", "polar_eps", "polar_iter", "pcg_prec", "pcg_guess",
                  "pcg_peek", "usolve_cut", "udiag", "list_buffer", "periodic", "mpole_cut", "use_mpole",
                  "use_polar"):
            d[k] = self.scalar(k)
        d["grid"] = np.array([int(self.scalar("grid%d" % j)) for j in (1, 2, 3)], dtype=np.int32)
        d["chgpen"] = bool(self.scalar("chgpen"))
        if d["chgpen"]:  # HIPPO charge penetration: core charge, valence charge, valence alpha
            d.update(pcore=self.f64("pcore", n), pval=self.f64("pval", n), palpha=self.f64("palpha", n))
        d["chgtrn"] = bool(self.scalar("chgtrn"))
        if d["chgtrn"]:  # HIPPO charge transfer: charge and damping exponent, cutoff and taper
            d.update(chgct=self.f64("chgct", n), dmpct=self.f64("dmpct", n), ct_cut=self.scalar("ct_cut"),
                     ct_taper=self.scalar("ct_taper"))
        d["disp"] = bool(self.scalar("disp"))
        if d["disp"]:  # HIPPO dispersion: C6, damping exponent, cutoff, taper, per-pair scales
            d.update(csix=self.f64("csix", n), adisp=self.f64("adisp", n), disp_cut=self.scalar("disp_cut"),
                     disp_taper=self.scalar("disp_taper"), scale_ds=self.f64("scale_ds", len(d["scale_ik"])))
        d["repel"] = bool(self.scalar("repel"))
        if d["repel"]:  # HIPPO repulsion: size, damping exponent, valence count, cutoff, taper, scales
            d.update(sizpr=self.f64("sizpr", n), dmppr=self.f64("dmppr", n), elepr=self.f64("elepr", n),
                     rep_cut=self.scalar("rep_cut"), rep_taper=self.scalar("rep_taper"),
                     scale_rs=self.f64("scale_rs", len(d["scale_ik"])))
        # MD terms: vdW tables and valence lists
        nj = int(self.scalar("njvdw"))
        nvs, nb, na, nu = (int(self.scalar(k)) for k in ("nvscale", "nbond", "nangle", "nurey"))
        d.update(jvdw=self.i32("jvdw", n), njvdw=nj, vdw_rmin=self.f64("vdw_rmin", nj * nj),
                 vdw_epsmix=self.f64("vdw_epsmix", nj * nj), vdw_red=self.f64("vdw_red", n),
                 vdw_parent=self.i32("vdw_parent", n), mass=self.f64("mass", n),
                 vscale_ik=self.i32("vscale_ik", (nvs, 2)) if nvs else np.zeros((0, 2), np.int32),
                 vscale_s=self.f64("vscale_s", nvs) if nvs else np.zeros(0),
                 bond_ij=self.i32("bond_ij", (nb, 2)) if nb else np.zeros((0, 2), np.int32),
                 bond_kr=self.f64("bond_kr", (nb, 2)) if nb else np.zeros((0, 2)),
                 angle_abc=self.i32("angle_abc", (na, 3)) if na else np.zeros((0, 3), np.int32),
                 angle_kt=self.f64("angle_kt", (na, 2)) if na else np.zeros((0, 2)),
                 urey_ac=self.i32("urey_ac", (nu, 2)) if nu else np.zeros((0, 2), np.int32),
                 urey_kr=self.f64("urey_kr", (nu, 2)) if nu else np.zeros((0, 2)))
        for k in ("vdw_cut", "vdw_taper", "use_vdw", "use_valence", "bond_cubic", "bond_quartic", "angle_cubic",
                  "angle_quartic", "angle_pentic", "angle_sextic"):
            d[k] = self.scalar(k)
        return d


class Context:
    """Device state of one system on the current GPU (t9_ctx)."""

    def __init__(self, system, precision=T9_FP32):
        self.n = system.n
        self.system = system
        h = _vp()
        check(lib().t9_ctx_create(system.h, precision, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            check(lib().t9_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_coords(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64)
        check(lib().t9_set_coords(self.h, _p(a)))

    def coords(self):
        out = np.zeros((self.n, 3))
        check(lib().t9_get_coords(self.h, _p(out)))
        return out

    def nblist_refresh(self):
        r = ctypes.c_int()
        check(lib().t9_nblist_refresh(self.h, ctypes.byref(r)))
        return bool(r.value)

    def nblist_build(self):
        check(lib().t9_nblist_build(self.h))

    def nblist_stats(self):
        out = np.zeros(5, dtype=np.int64)
        check(lib().t9_nblist_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def nblist_pairs(self, which=0):
        """Pairs of a stored list (0 multipole, 1 vdW), caller order, i < k: int array [m, 2]."""
        cnt = ctypes.c_int64(0)
        check(lib().t9_nblist_pairs(self.h, which, 0, None, ctypes.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 2), np.int32)
        check(lib().t9_nblist_pairs(self.h, which, cnt.value, _p(out, _i), ctypes.byref(cnt)))
        return out[:cnt.value]

    def mpole_init(self):
        check(lib().t9_mpole_init(self.h))

    def rpole(self):
        out = np.zeros((self.n, 10))
        check(lib().t9_get_rpole(self.h, _p(out)))
        return out

    def dfield(self):
        fd, fp = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        check(lib().t9_dfield(self.h, _p(fd), _p(fp)))
        return fd, fp

    def ufield(self, ud, up):
        ud = np.ascontiguousarray(ud, dtype=np.float64)
        up = np.ascontiguousarray(up, dtype=np.float64)
        fd, fp = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        check(lib().t9_ufield(self.h, _p(ud), _p(up), _p(fd), _p(fp)))
        return fd, fp

    def induce(self, copy_out=True):
        st = InduceStats()
        if copy_out:
            ud, up = np.zeros((self.n, 3)), np.zeros((self.n, 3))
            check(lib().t9_induce(self.h, _p(ud), _p(up), ctypes.byref(st)))
            return ud, up, st
        check(lib().t9_induce(self.h, None, None, ctypes.byref(st)))
        return None, None, st

    def induce_trace(self):
        """RMS residual (Debye) of the last solve: [initial guess, iteration 1, ...]."""
        eps = np.zeros(256)
        cnt = ctypes.c_int(0)
        check(lib().t9_induce_trace(self.h, _p(eps), 256, ctypes.byref(cnt)))
        return eps[:min(cnt.value, 256)].copy()

    def uind(self):
        ud, up = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        check(lib().t9_get_uind(self.h, _p(ud), _p(up)))
        return ud, up

    def set_uind(self, ud, up):
        ud = np.ascontiguousarray(ud, dtype=np.float64)
        up = np.ascontiguousarray(up, dtype=np.float64)
        check(lib().t9_set_uind(self.h, _p(ud), _p(up)))

    def energy(self, vers=CALC_ENERGY | CALC_GRAD | CALC_VIRIAL, grad=True):
        """Energy of the enabled terms (dict), gradient dE/dx [n,3] (caller
        order, or None) and virial [3,3]."""
        out = EnergyOut()
        g = np.zeros((self.n, 3)) if grad and (vers & CALC_GRAD) else None
        check(lib().t9_energy(self.h, vers, ctypes.byref(out), _p(g) if g is not None else None))
        e = dict(esum=out.esum, em=out.em, ep=out.ep, em_real=out.em_real, em_recip=out.em_recip,
                 em_self=out.em_self, iterations=out.induce_iterations, evdw=out.evdw, ebond=out.ebond,
                 eangle=out.eangle, eurey=out.eurey, ect=out.ect, edisp=out.edisp, erep=out.erep)
        if vers & CALC_ANALYZE:
            e.update(nem=out.nem, nep=out.nep, nev=out.nev, nct=out.nct, ndisp=out.ndisp, nrep=out.nrep)
        return e, g, np.array(out.vir[:]).reshape(3, 3)

    def energy_terms(self, terms, vers=CALC_ENERGY | CALC_GRAD | CALC_VIRIAL, grad=True):
        """As energy(), for the selected term groups (TERM_ELEC | TERM_VDW | TERM_VALENCE | TERM_CHGTRN |
        TERM_DISP | TERM_REPEL)."""
        out = EnergyOut()
        g = np.zeros((self.n, 3)) if grad and (vers & CALC_GRAD) else None
        check(lib().t9_energy_terms(self.h, vers, terms, ctypes.byref(out), _p(g) if g is not None else None))
        e = dict(esum=out.esum, em=out.em, ep=out.ep, em_real=out.em_real, em_recip=out.em_recip,
                 em_self=out.em_self, iterations=out.induce_iterations, evdw=out.evdw, ebond=out.ebond,
                 eangle=out.eangle, eurey=out.eurey, ect=out.ect, edisp=out.edisp, erep=out.erep)
        if vers & CALC_ANALYZE:
            e.update(nem=out.nem, nep=out.nep, nev=out.nev, nct=out.nct, ndisp=out.ndisp, nrep=out.nrep)
        return e, g, np.array(out.vir[:]).reshape(3, 3)

    def set_coords_device(self, d_xyz_ptr):
        """Coordinates from a device pointer ([n][3] FP64, caller order)."""
        check(lib().t9_set_coords_device(self.h, ctypes.c_void_p(int(d_xyz_ptr))))

    def get_coords_device(self, d_xyz_ptr):
        check(lib().t9_get_coords_device(self.h, ctypes.c_void_p(int(d_xyz_ptr))))

    def energy_device(self, terms, d_grad_ptr, vers=CALC_ENERGY | CALC_GRAD | CALC_VIRIAL):
        """Energy of the selected terms; the gradient goes to a device pointer
        ([n][3] FP64, caller order).  Returns the energy dict and the virial."""
        out = EnergyOut()
        check(lib().t9_energy_device(self.h, vers, terms, ctypes.byref(out),
                                     ctypes.c_void_p(int(d_grad_ptr)) if d_grad_ptr else None))
        e = dict(esum=out.esum, em=out.em, ep=out.ep, evdw=out.evdw, ebond=out.ebond, eangle=out.eangle,
                 eurey=out.eurey, ect=out.ect, edisp=out.edisp, erep=out.erep, iterations=out.induce_iterations)
        return e, np.array(out.vir[:]).reshape(3, 3)

    def _precond(self, fn, rd, rp):
        a = np.ascontiguousarray(rd, dtype=np.float64)
        b = np.ascontiguousarray(rp, dtype=np.float64)
        zd, zp = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        check(fn(self.h, _p(a), _p(b), _p(zd), _p(zp)))
        return zd, zp

    def diag_precond(self, rd, rp):
        return self._precond(lib().t9_diag_precond, rd, rp)

    def sparse_precond(self, rd, rp):
        return self._precond(lib().t9_sparse_precond_apply, rd, rp)

    # ---- PME sub-steps; grids are [K3, K2, K1, 2] FP64 (real, imaginary)
    def _grid_shape(self):
        k = [int(self.system.scalar("grid%d" % j)) for j in (1, 2, 3)]
        return (k[2], k[1], k[0], 2)

    def pme_grid_mpole(self):
        g = np.zeros(self._grid_shape())
        check(lib().t9_pme_grid_mpole(self.h, _p(g)))
        return g

    def pme_grid_uind(self, ud, up):
        a = np.ascontiguousarray(ud, dtype=np.float64)
        b = np.ascontiguousarray(up, dtype=np.float64)
        g = np.zeros(self._grid_shape())
        check(lib().t9_pme_grid_uind(self.h, _p(a), _p(b), _p(g)))
        return g

    def _grid_op(self, fn, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64).copy()
        check(fn(self.h, _p(g)))
        return g

    def pme_fft_front(self, grid):
        return self._grid_op(lib().t9_pme_fft_front, grid)

    def pme_fft_back(self, grid):
        return self._grid_op(lib().t9_pme_fft_back, grid)

    def pme_conv(self, grid):
        return self._grid_op(lib().t9_pme_conv, grid)

    def pme_fphi_mpole(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64)
        f = np.zeros((self.n, 3))
        check(lib().t9_pme_fphi_mpole(self.h, _p(g), _p(f)))
        return f

    def pme_fphi_uind(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64)
        fd, fp = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        check(lib().t9_pme_fphi_uind(self.h, _p(g), _p(fd), _p(fp)))
        return fd, fp

    # ---- learned initial guess
    def guess_set_params(self, params=None, norm=None, seed=0):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        q = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        check(lib().t9_guess_set_params(self.h, _pf(p) if p is not None else None,
                                        _pf(q) if q is not None else None, int(seed)))

    def guess_get_params(self):
        p = np.zeros(lib().t9_guess_nparam(), np.float32)
        q = np.zeros(76, np.float32)
        check(lib().t9_guess_get_params(self.h, _pf(p), _pf(q)))
        return p, q

    def guess_enable(self, on=True):
        check(lib().t9_guess_enable(self.h, 1 if on else 0))

    def guess_capture(self, stride, max_samples):
        check(lib().t9_guess_capture(self.h, int(stride), int(max_samples)))

    def guess_stats(self):
        out = np.zeros(4, np.int64)
        check(lib().t9_guess_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return dict(captured=int(out[0]), used=int(out[1]), fallback=int(out[2]), last_predictor=int(out[3]))

    def guess_train(self, epochs, batch=4096, lr=1e-3, seed=0):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().t9_guess_train(self.h, int(epochs), int(batch), float(lr), int(seed), ctypes.byref(a),
                                   ctypes.byref(b)))
        return a.value, b.value

    def guess_mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 6), np.float32)
        check(lib().t9_guess_mlp_forward(self.h, _pf(X), X.shape[0], _pf(Y)))
        return Y

    def guess_mlp_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(lib().t9_guess_nparam(), np.float32)
        loss = ctypes.c_double()
        check(lib().t9_guess_mlp_grad(self.h, _pf(X), _pf(T), X.shape[0], _pf(g), ctypes.byref(loss)))
        return loss.value, g

    def guess_adam(self, grad, lr):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        check(lib().t9_guess_adam(self.h, _pf(g), float(lr)))

    def guess_features(self):
        F = np.zeros((self.n, 32), np.float32)
        check(lib().t9_guess_features(self.h, _pf(F)))
        return F

    def guess_history(self, kind, age):
        out = np.zeros((self.n, 3))
        check(lib().t9_guess_history(self.h, int(kind), int(age), _p(out)))
        return out

    def md_init(self, kelvin, seed):
        check(lib().t9_md_init(self.h, float(kelvin), int(seed)))

    def velocities(self):
        v = np.zeros((self.n, 3))
        check(lib().t9_md_get_velocities(self.h, _p(v)))
        return v

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(lib().t9_md_set_velocities(self.h, _p(v)))

    def md_run(self, nsteps, dt=0.002, nrespa=8, bath_kelvin=0.0, tau=0.1):
        """nsteps RESPA steps (dt ps outer, dt/nrespa inner); bath_kelvin > 0
        applies the Bussi thermostat.  Returns a dict of the run's stats."""
        st = MdStats()
        check(lib().t9_md_run(self.h, int(nsteps), float(dt), int(nrespa), float(bath_kelvin), float(tau),
                              ctypes.byref(st)))
        return dict(epot=st.epot, ekin=st.ekin, temperature=st.temperature, iterations=st.induce_iterations,
                    list_builds=st.list_builds, steps=st.steps)

    def md_bussi(self, dt, tau, kelvin, r, s2):
        """One Bussi rescaling of the velocities with the given deviates."""
        check(lib().t9_md_bussi(self.h, float(dt), float(tau), float(kelvin), float(r), float(s2)))

    def sync(self):
        check(lib().t9_sync(self.h))

    def kernel_timing(self, kernel, mode):
        """mode 1: clear+start, 0: stop, 2: read.  Returns (total_ms, launches)."""
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        check(lib().t9_kernel_timing(self.h, kernel, mode, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def cutoff_pairs(self):
        v = ctypes.c_int64()
        check(lib().t9_cutoff_pairs(self.h, ctypes.byref(v)))
        return v.value


def fft3d(d_grid_ptr, k1, k2, k3, precision, sign):
    """In-place 3D transform of a device grid [k3][k2][k1] complex (sign -1 forward, +1 inverse)."""
    check(lib().t9_fft3d(ctypes.c_void_p(int(d_grid_ptr)), int(k1), int(k2), int(k3), int(precision), int(sign)))


def fft3d_noalloc(d_grid_ptr, k1, k2, k3, precision, sign):
    """The same transform through the allocation-free entry point (static per-GPU twiddle table)."""
    check(lib().t9_fft3d_noalloc(ctypes.c_void_p(int(d_grid_ptr)), int(k1), int(k2), int(k3), int(precision),
                                 int(sign)))


def copy_probe(nbytes, reps=20):
    ms = ctypes.c_double()
    check(lib().t9_copy_probe(int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value
