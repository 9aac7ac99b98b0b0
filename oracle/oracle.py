"""ctypes binding of the FP64 CPU oracle (oracle/oracle.cpp) — TEST INFRASTRUCTURE.

Only tests/, the smoke check and bench.py's CPU-baseline leg use this module;
it is the checker, never the thing measured or shipped.  The oracle takes the
per-atom arrays produced by the host setup (local-frame multipoles, frames,
polarizabilities, Thole table, pair scale list) and recomputes the physics in
double precision.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def lib_path():
    return os.path.join(_HERE, "liboracle.so")


def build():
    """Compile the oracle with g++ (no GPU code)."""
    src = os.path.join(_HERE, "oracle.cpp")
    out = lib_path()
    cmd = f"g++ -std=c++17 -O2 -fopenmp -fPIC -shared -o {out}.tmp {src} && mv {out}.tmp {out}"
    if os.system(cmd) != 0:
        raise RuntimeError("oracle build failed")
    return out


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(lib_path()):
            build()
        L = ctypes.CDLL(lib_path())
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, _d, _d, _i, _i, _i, _i, _i, _d, _d, _i, ctypes.c_int, _d,
                                 ctypes.c_int, _i, _d, _d, ctypes.c_int, ctypes.c_double, ctypes.c_double, _i,
                                 ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_chgpen.argtypes = [ctypes.c_void_p, _d, _d, _d]
        L.orc_damp_chgpen.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _d]
        L.orc_chgtrn.argtypes = [ctypes.c_void_p, _d, ctypes.c_double, ctypes.c_double, _d, _d, _d,
                                 ctypes.POINTER(ctypes.c_longlong)]
        L.orc_disp.argtypes = [ctypes.c_void_p, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int, _i, _d, _d, _d,
                               _d, ctypes.POINTER(ctypes.c_longlong)]
        L.orc_damp_disp.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _d]
        L.orc_repel.argtypes = [ctypes.c_void_p, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int, _i, _d, _d,
                                _d, ctypes.POINTER(ctypes.c_longlong)]
        L.orc_repel.restype = ctypes.c_double
        L.orc_repel_series.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _d]
        L.orc_set_xyz.argtypes = [ctypes.c_void_p, _d]
        L.orc_get_rpole.argtypes = [ctypes.c_void_p, _d]
        L.orc_get_bsmod.argtypes = [ctypes.c_void_p, ctypes.c_int, _d]
        L.orc_dfield.argtypes = [ctypes.c_void_p, _d, _d]
        L.orc_ufield.argtypes = [ctypes.c_void_p, _d, _d, _d, _d]
        L.orc_recip_dipole.argtypes = [ctypes.c_void_p, _d, _d, _d, _d]
        L.orc_recip_mpole.argtypes = [ctypes.c_void_p, _d, _d]
        L.orc_recip_mpole.restype = ctypes.c_double
        L.orc_dft3.argtypes = [ctypes.c_void_p, _d, ctypes.c_int]
        L.orc_induce.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, _d, _d, _i, _d]
        L.orc_induce.restype = ctypes.c_int
        L.orc_induce_from.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, _d, _d, _d, _d, _i, _d,
                                      _d]
        L.orc_induce_from.restype = ctypes.c_int
        L.orc_set_md.argtypes = [ctypes.c_void_p, _i, ctypes.c_int, _d, _d, _d, _i, ctypes.c_int, _i, _d,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int, _i, _d, ctypes.c_int, _i, _d,
                                 ctypes.c_int, _i, _d, _d]
        L.orc_md_energy.argtypes = [ctypes.c_void_p, ctypes.c_int, _d, _d, _d]
        L.orc_md_energy.restype = ctypes.c_int
        L.orc_pme_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, _d, _d, _d]
        L.orc_pme_conv.argtypes = [ctypes.c_void_p, _d]
        L.orc_pme_field.argtypes = [ctypes.c_void_p, _d, ctypes.c_int, _d]
        L.orc_precond.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _d, _d, _d, _d]
        if hasattr(L, "orc_energy"):
            L.orc_energy.argtypes = [ctypes.c_void_p, ctypes.c_int, _d, _d, _d, _d, _d]
            L.orc_energy.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def aspc_coefficients(k=14):
    """Kolafa's always-stable predictor-corrector coefficients B_1..B_{k+2}
    (J. Comput. Chem. 25, 335 (2004)); B[0] weights the most recent solution."""
    from math import comb
    den = comb(2 * k + 2, k + 1)
    return np.array([(-1) ** (j + 1) * j * comb(2 * k + 4, k + 2 - j) / den for j in range(1, k + 3)])


def interaction_counts(a):
    """Analyze-mode interaction counts (nem, nep) of energy(calc::v3):
    src/acc/amoeba/empoleewald.cpp:78-79,132-133,177-180 and
    epolarewald.cpp:128-130,280-282,499-500 (Ewald), empolenonewald.cpp:74-75,
    159-161 and epolarnonewald.cpp:164-165,313-314 (no Ewald).  Pairs within
    the cutoff, less the excluded pairs (m scale 0 for nem, p scale 0 for
    nep); a polarization pair counts only when its energy is nonzero, i.e.
    when one of the two atoms is polarizable; plus one self term per atom
    with Ewald.  Brute force over all pairs (minimum image when periodic);
    a: t9hip.System.arrays()."""
    x = np.asarray(a["xyz"])
    n = len(x)
    ewald = bool(a["ewald"])
    cut = a["ewald_cut"] if ewald else a["mpole_cut"]
    d = x[None, :, :] - x[:, None, :]
    if bool(a["periodic"]):
        box = np.asarray(a["box"])
        d -= box * np.round(d / box)
    iu = np.triu_indices(n, 1)
    within = (d ** 2).sum(-1)[iu] <= cut * cut
    pol = np.asarray(a["polarity"]) != 0
    polpair = (pol[:, None] | pol[None, :])[iu]
    m0 = np.zeros((n, n), bool)
    p0 = np.zeros((n, n), bool)
    for (i, k), sc in zip(a["scale_ik"], a["scale_mdpu"]):
        m0[i, k] = m0[k, i] = sc[0] == 0
        p0[i, k] = p0[k, i] = sc[2] == 0
    self_terms = n if ewald else 0
    nem = int(np.sum(within & ~m0[iu])) + self_terms if a["use_mpole"] else 0
    nep = int(np.sum(within & polpair & ~p0[iu])) + self_terms if a["use_polar"] else 0
    return nem, nep


def damp_chgpen(r, ai, ak):
    """HIPPO damping factors lambda_1..lambda_11 at distance r (Å) for valence
    exponents ai, ak (1/Å): (core-valence with ai, core-valence with ak,
    valence-valence), each an array of 6."""
    out = np.zeros(18)
    _lib().orc_damp_chgpen(float(r), float(ai), float(ak), _p(out))
    return out[:6], out[6:12], out[12:]


def damp_disp(r, ai, ak):
    """HIPPO dispersion damping D(r) and dD/dr for exponents ai, ak (1/Å)."""
    out = np.zeros(2)
    _lib().orc_damp_disp(float(r), float(ai), float(ak), _p(out))
    return out[0], out[1]


def repel_series(r, ai, ak):
    """Radial series B_0..B_5 of the HIPPO repulsion kernel at r (Å) for
    damping exponents ai, ak (1/Å), before the sizes."""
    out = np.zeros(6)
    _lib().orc_repel_series(float(r), float(ai), float(ak), _p(out))
    return out


def vdw_count(a):
    """Analyze-mode vdW count nev (src/acc/ehal.cpp:127,215-216): site pairs
    within the vdW cutoff less the pairs with vdW scale 0.  Sites: reduced
    hydrogens sit at x_p + red (x_i - x_p).  Brute force over all pairs."""
    if not a["use_vdw"]:
        return 0
    x = np.asarray(a["xyz"])
    n = len(x)
    par = np.asarray(a["vdw_parent"])
    red = np.asarray(a["vdw_red"])
    site = x.copy()
    m = par != np.arange(n)
    site[m] = x[par[m]] + red[m, None] * (x[m] - x[par[m]])
    d = site[None, :, :] - site[:, None, :]
    if bool(a["periodic"]):
        box = np.asarray(a["box"])
        d -= box * np.round(d / box)
    iu = np.triu_indices(n, 1)
    within = (d ** 2).sum(-1)[iu] <= a["vdw_cut"] ** 2
    z = np.zeros((n, n), bool)
    for (i, k), sv in zip(a["vscale_ik"], a["vscale_s"]):
        z[i, k] = z[k, i] = sv == 0
    return int(np.sum(within & ~z[iu]))


class Oracle:
    """FP64 reference computation for one system (see module docstring)."""

    def __init__(self, sysd, cut=None):
        """sysd: dict from t9hip.System.arrays() (numpy arrays + scalars)."""
        L = _lib()
        self.n = n = int(sysd["n"])
        self._keep = []

        def f64(x):
            a = np.ascontiguousarray(x, dtype=np.float64)
            self._keep.append(a)
            return a

        def i32(x):
            a = np.ascontiguousarray(x, dtype=np.int32)
            self._keep.append(a)
            return a

        xyz = f64(sysd["xyz"])
        box = f64(sysd["box"]) if sysd["periodic"] else None
        grid = i32(sysd["grid"])
        ik = i32(sysd["scale_ik"]) if len(sysd["scale_ik"]) else i32(np.zeros((1, 2)))
        mdpu = f64(sysd["scale_mdpu"]) if len(sysd["scale_mdpu"]) else f64(np.zeros((1, 4)))
        if cut is None:
            cut = sysd["ewald_cut"] if sysd["ewald"] else sysd["mpole_cut"]
        self.sysd = sysd
        self.h = L.orc_create(
            n, _p(xyz), _p(f64(sysd["pole"])), _p(i32(sysd["frame"]), _i), _p(i32(sysd["zaxis"]), _i),
            _p(i32(sysd["xaxis"]), _i), _p(i32(sysd["yaxis"]), _i), _p(i32(sysd["ysign"]), _i),
            _p(f64(sysd["polarity"])), _p(f64(sysd["pdamp"])), _p(i32(sysd["jpolar"]), _i), int(sysd["njpolar"]),
            _p(f64(sysd["thlval"])), int(sysd["nscale"]), _p(ik, _i), _p(mdpu), _p(box) if box is not None else None,
            int(sysd["ewald"]), float(sysd["aewald"]), float(cut), _p(grid, _i), int(sysd["pme_order"]))
        if sysd.get("chgpen"):
            L.orc_set_chgpen(self.h, _p(f64(sysd["pcore"])), _p(f64(sysd["pval"])), _p(f64(sysd["palpha"])))
        if "jvdw" in sysd:
            nj = int(sysd["njvdw"])
            vs_ik = i32(sysd["vscale_ik"]) if len(sysd["vscale_s"]) else i32(np.zeros((1, 2)))
            vs_s = f64(sysd["vscale_s"]) if len(sysd["vscale_s"]) else f64(np.zeros(1))

            def arr(key, shape_t, dt):
                a = sysd[key]
                return (i32 if dt == "i" else f64)(a if len(a) else np.zeros(shape_t))
            poly = f64([sysd["bond_cubic"], sysd["bond_quartic"], sysd["angle_cubic"], sysd["angle_quartic"],
                        sysd["angle_pentic"], sysd["angle_sextic"]])
            L.orc_set_md(self.h, _p(i32(sysd["jvdw"]), _i), nj, _p(f64(sysd["vdw_rmin"])),
                         _p(f64(sysd["vdw_epsmix"])), _p(f64(sysd["vdw_red"])), _p(i32(sysd["vdw_parent"]), _i),
                         len(sysd["vscale_s"]), _p(vs_ik, _i), _p(vs_s), float(sysd["vdw_cut"]),
                         float(sysd["vdw_taper"]), len(sysd["bond_kr"]), _p(arr("bond_ij", (1, 2), "i"), _i),
                         _p(arr("bond_kr", (1, 2), "d")), len(sysd["angle_kt"]),
                         _p(arr("angle_abc", (1, 3), "i"), _i), _p(arr("angle_kt", (1, 2), "d")),
                         len(sysd["urey_kr"]), _p(arr("urey_ac", (1, 2), "i"), _i), _p(arr("urey_kr", (1, 2), "d")),
                         _p(poly))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_destroy(self.h)
            self.h = None

    def set_xyz(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64)
        _lib().orc_set_xyz(self.h, _p(a))

    def rpole(self):
        out = np.zeros((self.n, 10))
        _lib().orc_get_rpole(self.h, _p(out))
        return out

    def bsmod(self, dim):
        out = np.zeros(int(self.sysd["grid"][dim]))
        _lib().orc_get_bsmod(self.h, dim, _p(out))
        return out

    def dfield(self):
        fd, fp = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        _lib().orc_dfield(self.h, _p(fd), _p(fp))
        return fd, fp

    def ufield(self, ud, up):
        ud = np.ascontiguousarray(ud, dtype=np.float64)
        up = np.ascontiguousarray(up, dtype=np.float64)
        fd, fp = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        _lib().orc_ufield(self.h, _p(ud), _p(up), _p(fd), _p(fp))
        return fd, fp

    def recip_dipole(self, ud, up):
        ud = np.ascontiguousarray(ud, dtype=np.float64)
        up = np.ascontiguousarray(up, dtype=np.float64)
        fd, fp = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        _lib().orc_recip_dipole(self.h, _p(ud), _p(up), _p(fd), _p(fp))
        return fd, fp

    def recip_mpole(self):
        f = np.zeros((self.n, 3))
        fphi = np.zeros((self.n, 20))
        e = _lib().orc_recip_mpole(self.h, _p(f), _p(fphi))
        return e, f, fphi

    def dft3(self, grid_c, sign):
        g = np.ascontiguousarray(np.stack([grid_c.real, grid_c.imag], axis=-1), dtype=np.float64)
        _lib().orc_dft3(self.h, _p(g), sign)
        return g[..., 0] + 1j * g[..., 1]

    def _kgrid(self):
        k1, k2, k3 = (int(v) for v in self.sysd["grid"])
        return (k3, k2, k1, 2)

    def pme_grid_mpole(self):
        """Spread fractional multipoles: grid [K3, K2, K1, 2] (real, imaginary)."""
        g = np.zeros(self._kgrid())
        _lib().orc_pme_grid(self.h, 0, None, None, _p(g))
        return g

    def pme_grid_uind(self, ud, up):
        a = np.ascontiguousarray(ud, dtype=np.float64)
        b = np.ascontiguousarray(up, dtype=np.float64)
        g = np.zeros(self._kgrid())
        _lib().orc_pme_grid(self.h, 1, _p(a), _p(b), _p(g))
        return g

    def pme_conv(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64).copy()
        _lib().orc_pme_conv(self.h, _p(g))
        return g

    def pme_field(self, grid, imag=False):
        g = np.ascontiguousarray(grid, dtype=np.float64)
        f = np.zeros((self.n, 3))
        _lib().orc_pme_field(self.h, _p(g), 1 if imag else 0, _p(f))
        return f

    def precond(self, rd, rp, sparse=True, ucut=None, udiag=None):
        """z = M r for both residual sets (sparse: usolve pair sum)."""
        s = self.sysd
        ucut = s["usolve_cut"] if ucut is None else ucut
        udiag = s["udiag"] if udiag is None else udiag
        a = np.ascontiguousarray(rd, dtype=np.float64)
        b = np.ascontiguousarray(rp, dtype=np.float64)
        zd, zp = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        _lib().orc_precond(self.h, 1 if sparse else 0, float(ucut), float(udiag), _p(a), _p(b), _p(zd), _p(zp))
        return zd, zp

    def md_energy(self, vdw=True, valence=True):
        """vdW and valence energies (dict), gradient [n,3] and virial [3,3]."""
        e4, g, v = np.zeros(4), np.zeros((self.n, 3)), np.zeros((3, 3))
        terms = (1 if vdw else 0) | (2 if valence else 0)
        rc = _lib().orc_md_energy(self.h, terms, _p(e4), _p(g), _p(v))
        assert rc == 0
        return dict(evdw=e4[0], ebond=e4[1], eangle=e4[2], eurey=e4[3]), g, v

    def induce_from(self, u0d, u0p, eps=None, maxiter=None):
        """PCG from the initial guess (u0d, u0p); returns (ud, up, iters, rms,
        guess_rms, rc)."""
        s = self.sysd
        eps = s["polar_eps"] if eps is None else eps
        maxiter = int(s["polar_iter"]) if maxiter is None else maxiter
        a = np.ascontiguousarray(u0d, dtype=np.float64)
        b = np.ascontiguousarray(u0p, dtype=np.float64)
        ud, up = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        it = np.zeros(1, dtype=np.int32)
        rms, g = np.zeros(1), np.zeros(1)
        rc = _lib().orc_induce_from(self.h, eps, maxiter, int(s["pcg_prec"]), int(s["pcg_guess"]), s["pcg_peek"],
                                    s["usolve_cut"], s["udiag"], _p(a), _p(b), _p(ud), _p(up), _p(it, _i), _p(rms),
                                    _p(g))
        return ud, up, int(it[0]), float(rms[0]), float(g[0]), rc

    def energy(self, ud=None, up=None, mpole=True, polar=True):
        """Energies (dict), gradient dE/dx [n,3] and virial [3,3] of the
        multipole and/or polarization terms at the given induced dipoles."""
        n = self.n
        ud = np.zeros((n, 3)) if ud is None else np.ascontiguousarray(ud, dtype=np.float64)
        up = np.zeros((n, 3)) if up is None else np.ascontiguousarray(up, dtype=np.float64)
        e5, g, v = np.zeros(5), np.zeros((n, 3)), np.zeros((3, 3))
        terms = (1 if mpole else 0) | (2 if polar else 0)
        _lib().orc_energy(self.h, terms, _p(ud), _p(up), _p(e5), _p(g), _p(v))
        e = dict(em=e5[0], ep=e5[1], em_real=e5[2], em_recip=e5[3], em_self=e5[4])
        return e, g, v

    def chgtrn(self):
        """HIPPO charge transfer of the system's chgtrn parameters: energy,
        gradient dE/dx [n,3], virial [3,3] and the interaction count."""
        s = self.sysd
        ct = np.ascontiguousarray(np.stack([s["chgct"], s["dmpct"]], axis=1), dtype=np.float64)
        e, g, v = np.zeros(1), np.zeros((self.n, 3)), np.zeros((3, 3))
        cnt = ctypes.c_longlong(0)
        _lib().orc_chgtrn(self.h, _p(ct), float(s["ct_cut"]), float(s["ct_taper"]), _p(e), _p(g), _p(v),
                          ctypes.byref(cnt))
        return float(e[0]), g, v, int(cnt.value)

    def disp(self):
        """HIPPO dispersion with a cutoff of the system's dispersion
        parameters: energy, gradient dE/dx [n,3], virial [3,3] and the
        interaction count."""
        s = self.sysd
        cd = np.ascontiguousarray(np.stack([s["csix"], s["adisp"]], axis=1), dtype=np.float64)
        ik = np.ascontiguousarray(s["scale_ik"], dtype=np.int32).reshape(-1, 2)
        ds = np.ascontiguousarray(s["scale_ds"], dtype=np.float64)
        if len(ds) == 0:
            ik, ds = np.zeros((1, 2), dtype=np.int32), np.zeros(1)
            nds = 0
        else:
            nds = len(ds)
        e, g, v = np.zeros(1), np.zeros((self.n, 3)), np.zeros((3, 3))
        cnt = ctypes.c_longlong(0)
        _lib().orc_disp(self.h, _p(cd), float(s["disp_cut"]), float(s["disp_taper"]), nds, _p(ik, _i), _p(ds),
                        _p(e), _p(g), _p(v), ctypes.byref(cnt))
        return float(e[0]), g, v, int(cnt.value)

    def repel(self):
        """HIPPO Pauli repulsion of the system's repulsion parameters:
        energy, gradient dE/dx [n,3], virial [3,3] and the interaction
        count."""
        s = self.sysd
        rp = np.ascontiguousarray(np.stack([s["sizpr"], s["dmppr"], s["elepr"]], axis=1), dtype=np.float64)
        ik = np.ascontiguousarray(s["scale_ik"], dtype=np.int32).reshape(-1, 2)
        rs = np.ascontiguousarray(s["scale_rs"], dtype=np.float64)
        nrs = len(rs)
        if nrs == 0:
            ik, rs = np.zeros((1, 2), dtype=np.int32), np.zeros(1)
        g, v = np.zeros((self.n, 3)), np.zeros((3, 3))
        cnt = ctypes.c_longlong(0)
        e = _lib().orc_repel(self.h, _p(rp), float(s["rep_cut"]), float(s["rep_taper"]), nrs, _p(ik, _i), _p(rs),
                             _p(g), _p(v), ctypes.byref(cnt))
        return float(e), g, v, int(cnt.value)

    def induce(self, eps=None, maxiter=None, prec=None, guess=None, peek=None, usolve=None, udiag=None):
        s = self.sysd
        eps = s["polar_eps"] if eps is None else eps
        maxiter = int(s["polar_iter"]) if maxiter is None else maxiter
        prec = int(s["pcg_prec"]) if prec is None else int(prec)
        guess = int(s["pcg_guess"]) if guess is None else int(guess)
        peek = s["pcg_peek"] if peek is None else peek
        usolve = s["usolve_cut"] if usolve is None else usolve
        udiag = s["udiag"] if udiag is None else udiag
        ud, up = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        it = np.zeros(1, dtype=np.int32)
        rms = np.zeros(1)
        rc = _lib().orc_induce(self.h, eps, maxiter, prec, guess, peek, usolve, udiag, _p(ud), _p(up),
                               _p(it, _i), _p(rms))
        return ud, up, int(it[0]), float(rms[0]), rc
