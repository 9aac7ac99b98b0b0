"""HIPPO charge-penetration multipoles in the FP64 oracle (CPU only).

* test/emhippo.cpp with test/ref/emhippo.{1,2}.txt: pentane + acetamide (104
  atoms, HIPPO-2019 multipoles and chgpen records) in a 35 x 30 x 25 Å box,
  Ewald at 7 Å and non-Ewald at 9 Å: energy, gradient, virial and the
  interaction count, at the reference test's FP64 tolerances (1e-4, 1e-4,
  1e-3).
* The damping kernels against what the physics fixes: the core-valence
  lambda_1 / r is the potential of a normalized Slater density (Gauss's law,
  integrated numerically here), the valence-valence lambda_1 / r the energy of
  two such densities (a numerical 2D integral), the derivative series
  f_(n+1) = -(1/r) df_n/dr by central differences, and the two-exponent form
  continuous into the equal-exponent form.
"""
import json
import os

import numpy as np
import pytest
from scipy import integrate

import oracle
import t9hip

HIPPO_EWALD_KEY = "neighbor-list\na-axis 35.0\nb-axis 30.0\nc-axis 25.0\nmultipoleterm only\newald-cutoff 7.0\newald\n"
HIPPO_NONEWALD_KEY = "neighbor-list\na-axis 35.0\nb-axis 30.0\nc-axis 25.0\nmultipoleterm only\nmpole-cutoff 9.0\n"


def hippo_system(golden_dir, ewald):
    return t9hip.System.from_files(os.path.join(golden_dir, "c5h12acnh2.xyz"),
                                   prm=os.path.join(golden_dir, "hippo19_subset.prm"),
                                   key=HIPPO_EWALD_KEY if ewald else HIPPO_NONEWALD_KEY)


@pytest.mark.parametrize("ewald", [True, False])
def test_emhippo_goldens(golden_dir, ewald):
    ref = json.load(open(os.path.join(golden_dir, "hippo_ref.json")))["emhippo_ewald" if ewald else "emhippo_nonewald"]
    s = hippo_system(golden_dir, ewald)
    a = s.arrays()
    assert a["chgpen"]
    e, g, v = oracle.Oracle(a).energy(mpole=True, polar=False)
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    print(f"ewald={ewald}: E={e['em']:.6f} ref={ref['energy']} max|dg|={np.abs(g - rg).max():.2e} "
          f"max|dv|={np.abs(v - rv).max():.2e}")
    assert abs(e["em"] - ref["energy"]) < 1e-4
    assert np.abs(g - rg).max() < 1e-4
    assert np.abs(v - rv).max() < 1e-3
    assert oracle.interaction_counts(a)[0] == ref["count"]
    # the multipoles are the HIPPO ones: core + valence = total monopole
    assert np.allclose(a["pcore"] + a["pval"], a["pole"][:, 0])


def _slater_potential(s, al):
    """Potential of the normalized density (al^3 / 8 pi) exp(-al t) at
    distance s: charge inside / s + integral of rho 4 pi t outside."""
    rho = lambda t: al ** 3 / (8 * np.pi) * np.exp(-al * t)
    inner = integrate.quad(lambda t: rho(t) * 4 * np.pi * t * t, 0, s)[0]
    outer = integrate.quad(lambda t: rho(t) * 4 * np.pi * t, s, np.inf)[0]
    return inner / s + outer


@pytest.mark.parametrize("r,al", [(0.7, 4.5439), (1.6, 3.5491), (3.1, 4.9530)])
def test_core_valence_kernel_is_slater_potential(r, al):
    lci, _, _ = oracle.damp_chgpen(r, al, al)
    assert abs(lci[0] / r - _slater_potential(r, al)) < 1e-9


@pytest.mark.parametrize("r,ai,ak", [(1.1, 4.5439, 4.5439), (1.1, 4.3225, 5.9682), (2.3, 3.5491, 4.2097)])
def test_valence_valence_kernel_is_overlap_energy(r, ai, ak):
    """E(r) = int rho_k(t) 2 pi t^2 int_-1^1 phi_i(|R + t u|) dmu dt."""
    phi = lambda s: (1 - (1 + 0.5 * ai * s) * np.exp(-ai * s)) / s  # checked above against Gauss's law
    rho = lambda t: ak ** 3 / (8 * np.pi) * np.exp(-ak * t)
    f = lambda mu, t: rho(t) * 2 * np.pi * t * t * phi(np.sqrt(r * r + t * t + 2 * r * t * mu))
    val = integrate.dblquad(f, 0, 40.0 / ak, -1, 1, epsabs=1e-12, epsrel=1e-10)[0]
    _, _, lik = oracle.damp_chgpen(r, ai, ak)
    assert abs(lik[0] / r - val) < 1e-8 * max(1.0, abs(val))


@pytest.mark.parametrize("ai,ak", [(4.5439, 4.5439), (4.3225, 5.9682), (3.5491, 4.2097)])
def test_damping_derivative_series(ai, ak):
    """lambda_(2n+3) rr_(n+1) = -(1/r) d(lambda_(2n+1) rr_n)/dr for all three kernels."""
    def f(r, which, n):
        lam = oracle.damp_chgpen(r, ai, ak)[which]
        dfac = np.prod(np.arange(2 * n - 1, 0, -2)) if n > 0 else 1.0
        return lam[n] * dfac / r ** (2 * n + 1)
    for r in (0.9, 1.7, 3.2):
        for which in range(3):
            for n in range(5):
                h = 1e-4 * r
                d = (f(r + h, which, n) - f(r - h, which, n)) / (2 * h)
                assert abs(-d / r - f(r, which, n + 1)) < 2e-7 * abs(f(r, which, n + 1)), (r, which, n)


def test_two_exponent_form_meets_equal_form():
    """Just beyond the 1e-3 switch the two-exponent kernel equals the
    equal-exponent kernel at the mean exponent to second order."""
    a = 4.5
    for r in (0.8, 1.5, 2.5):
        _, _, two = oracle.damp_chgpen(r, a - 1.5e-3, a + 1.5e-3)
        _, _, one = oracle.damp_chgpen(r, a - 4e-4, a + 4e-4)
        assert np.abs(two - one).max() < 1e-6


HIPPO_POLAR_EWALD_KEY = "neighbor-list\na-axis 25.0\nb-axis 25.0\nc-axis 30.0\npolarizeterm only\newald-cutoff 7.0\newald\n"
HIPPO_POLAR_NONEWALD_KEY = "neighbor-list\na-axis 25.0\nb-axis 25.0\nc-axis 30.0\npolarizeterm only\nmpole-cutoff 9.0\n"


def hippo_polar_system(golden_dir, ewald):
    return t9hip.System.from_files(os.path.join(golden_dir, "dmso.xyz"),
                                   prm=os.path.join(golden_dir, "hippo19_subset.prm"),
                                   key=HIPPO_POLAR_EWALD_KEY if ewald else HIPPO_POLAR_NONEWALD_KEY)


@pytest.mark.parametrize("ewald", [True, False])
def test_ephippo_goldens(golden_dir, ewald):
    """test/ephippo.cpp: 50 dimethyl sulfoxide molecules (150 atoms,
    25 x 25 x 30 Å): HIPPO mutual polarization (direct field with core /
    valence damping and the polar-1x scales, mutual field with the
    valence-valence damping and the induce-1x scales, one dipole set)."""
    ref = json.load(open(os.path.join(golden_dir, "hippo_ref.json")))["ephippo_ewald" if ewald else "ephippo_nonewald"]
    a = hippo_polar_system(golden_dir, ewald).arrays()
    o = oracle.Oracle(a)
    ud, up, it, rms, rc = o.induce()
    assert rc == 0 and rms < 1e-6
    assert np.abs(ud - up).max() < 1e-14  # the two dipole sets of the solver coincide
    e, g, v = o.energy(ud, up, mpole=False, polar=True)
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    print(f"ewald={ewald}: E={e['ep']:.6f} ref={ref['energy']} it={it} max|dg|={np.abs(g - rg).max():.2e} "
          f"max|dv|={np.abs(v - rv).max():.2e}")
    assert abs(e["ep"] - ref["energy"]) < 1e-4
    assert np.abs(g - rg).max() < 1e-4
    assert np.abs(v - rv).max() < 1e-3
    assert oracle.interaction_counts(a)[1] == ref["count"]


HIPPO_CHGTRN_KEY = "neighbor-list\na-axis 25.0\nb-axis 25.0\nc-axis 30.0\nchgtrnterm only\nchgtrn-cutoff 7.0\n"


def hippo_chgtrn_system(golden_dir):
    return t9hip.System.from_files(os.path.join(golden_dir, "dmso.xyz"),
                                   prm=os.path.join(golden_dir, "hippo19_subset.prm"), key=HIPPO_CHGTRN_KEY)


def test_chgtrn_goldens(golden_dir):
    """test/chgtrn.cpp: HIPPO charge transfer of 50 DMSO molecules (7 Å
    cutoff, taper from 6.3 Å) against test/ref/chgtrn.1.txt at the reference
    test's tolerances (energy and gradient 1e-4, virial 5e-4), and the
    interaction count."""
    ref = json.load(open(os.path.join(golden_dir, "hippo_ref.json")))["chgtrn"]
    a = hippo_chgtrn_system(golden_dir).arrays()
    assert a["chgtrn"] and a["ct_taper"] == pytest.approx(6.3)
    e, g, v, cnt = oracle.Oracle(a).chgtrn()
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    print(f"E={e:.6f} ref={ref['energy']} count={cnt} max|dg|={np.abs(g - rg).max():.2e} "
          f"max|dv|={np.abs(v - rv).max():.2e}")
    assert abs(e - ref["energy"]) < 1e-4
    assert np.abs(g - rg).max() < 1e-4
    assert np.abs(v - rv).max() < 5e-4
    assert cnt == ref["count"]
    # a zero exponent in the parameters stands for 1000 (no transfer)
    assert a["dmpct"].max() == 1000.0


def test_chgtrn_gradient_and_virial_are_derivatives(golden_dir):
    """The oracle's charge-transfer gradient against central differences of
    its energy (atoms whose pairs cross the taper region included), and the
    virial diagonal against the strain derivative (coordinates and box
    scaled together)."""
    s = hippo_chgtrn_system(golden_dir)
    a = s.arrays()
    o = oracle.Oracle(a)
    e0, g, v, _ = o.chgtrn()
    x0 = a["xyz"].copy()
    h = 1e-5
    for i in (0, 17, 101):
        for d in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[i, d] += h
            xm[i, d] -= h
            o.set_xyz(xp)
            ep = o.chgtrn()[0]
            o.set_xyz(xm)
            em = o.chgtrn()[0]
            assert abs((ep - em) / (2 * h) - g[i, d]) < 1e-6 * max(1.0, abs(g[i, d]))
    o.set_xyz(x0)
    for d in range(3):
        es = []
        for eps in (h, -h):
            b = dict(a)
            f = np.ones(3)
            f[d] += eps
            b["xyz"] = x0 * f
            b["box"] = np.asarray(a["box"]) * f
            es.append(oracle.Oracle(b).chgtrn()[0])
        assert abs((es[0] - es[1]) / (2 * h) - v[d, d]) < 1e-5 * max(1.0, abs(v[d, d]))


HIPPO_DISP_KEY = "neighbor-list\na-axis 35.0\nb-axis 30.0\nc-axis 25.0\ndispersionterm only\ndisp-cutoff 5.0\n"


def hippo_disp_system(golden_dir):
    return t9hip.System.from_files(os.path.join(golden_dir, "c5h12acnh2.xyz"),
                                   prm=os.path.join(golden_dir, "hippo19_disp_subset.prm"), key=HIPPO_DISP_KEY)


def test_disp_goldens(golden_dir):
    """test/disp.cpp EDISP-1-NONDEWALD: HIPPO dispersion of pentane +
    acetamide with a 5 Å cutoff (taper from 4.5 Å), against test/ref/disp.1.txt
    at the reference test's tolerances (energy and gradient 1e-4, virial
    1e-3), and the interaction count."""
    ref = json.load(open(os.path.join(golden_dir, "hippo_disp_ref.json")))["disp_cutoff"]
    a = hippo_disp_system(golden_dir).arrays()
    assert a["disp"] and a["disp_taper"] == pytest.approx(4.5)
    e, g, v, cnt = oracle.Oracle(a).disp()
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    print(f"E={e:.6f} ref={ref['energy']} count={cnt} max|dg|={np.abs(g - rg).max():.2e} "
          f"max|dv|={np.abs(v - rv).max():.2e}")
    assert abs(e - ref["energy"]) < 1e-4
    assert np.abs(g - rg).max() < 1e-4
    assert np.abs(v - rv).max() < 1e-3
    assert cnt == ref["count"]


@pytest.mark.parametrize("ai,ak", [(4.5439, 4.5439), (4.3225, 4.5439), (3.1, 4.4)])
def test_disp_damping_limits(ai, ak):
    """The dispersion damping goes to 1 at long range, vanishes as r^3 at
    short range (so C6 D^2 / r^6 stays finite; for equal exponents
    D ~ -(a r)^3 / 96, from e^-d times the polynomial, whose d^3 coefficient
    17/96 exceeds the exponential series' 16/96), and its derivative is dD/dr
    (central differences)."""
    D, _ = oracle.damp_disp(40.0, ai, ak)
    assert abs(D - 1) < 1e-12
    r1, r2 = 2e-3, 4e-3
    q1 = oracle.damp_disp(r1, ai, ak)[0] / r1 ** 3
    q2 = oracle.damp_disp(r2, ai, ak)[0] / r2 ** 3
    assert abs(q1 / q2 - 1) < 1e-2
    if ai == ak:
        assert abs(q1 / (-ai ** 3 / 96) - 1) < 1e-3
    for r in (0.8, 1.9, 3.7):
        h = 1e-5
        d = (oracle.damp_disp(r + h, ai, ak)[0] - oracle.damp_disp(r - h, ai, ak)[0]) / (2 * h)
        assert abs(d - oracle.damp_disp(r, ai, ak)[1]) < 1e-8


def test_disp_two_exponent_form_meets_equal_form():
    """Just beyond the 1e-3 switch the two-exponent damping equals the
    equal-exponent form at the mean exponent to second order."""
    a = 4.3
    for r in (0.9, 2.0, 3.3):
        two = oracle.damp_disp(r, a - 6e-4, a + 6e-4)[0]
        one = oracle.damp_disp(r, a - 4e-4, a + 4e-4)[0]
        assert abs(two - one) < 1e-6


def test_disp_gradient_is_derivative(golden_dir):
    """The oracle's dispersion gradient against central differences of its
    energy (atoms with pairs in the switching region among them)."""
    a = hippo_disp_system(golden_dir).arrays()
    o = oracle.Oracle(a)
    _, g, _, _ = o.disp()
    x0 = a["xyz"].copy()
    h = 1e-4  # the energy's rounding (~1e-12) limits smaller steps
    for i in (0, 9, 60):
        for d in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[i, d] += h
            xm[i, d] -= h
            o.set_xyz(xp)
            ep = o.disp()[0]
            o.set_xyz(xm)
            em = o.disp()[0]
            assert abs((ep - em) / (2 * h) - g[i, d]) < 1e-6 * max(1.0, abs(g[i, d]))
    o.set_xyz(x0)


HIPPO_REP_KEY = "neighbor-list\na-axis 35.0\nb-axis 30.0\nc-axis 25.0\nrepulsionterm only\nrepuls-cutoff 7.0\n"


HIPPO_REP_WATER_KEY = "neighbor-list\nrepulsionterm only\na-axis 25.0\nrepuls-cutoff 7.0\n"
REP_CASES = {"rephippo": ("c5h12acnh2.xyz", "hippo19_subset.prm", HIPPO_REP_KEY, "hippo_ref.json"),
             "rephippo_water": ("h2o10.xyz", "water19_records.prm", HIPPO_REP_WATER_KEY, "hippo_water_ref.json")}


def hippo_rep_system(golden_dir, case="rephippo"):
    xyz, prm, key, _ = REP_CASES[case]
    return t9hip.System.from_files(os.path.join(golden_dir, xyz), prm=os.path.join(golden_dir, prm), key=key)


def rep_ref(golden_dir, case):
    return json.load(open(os.path.join(golden_dir, REP_CASES[case][3])))[case]


@pytest.mark.parametrize("case", sorted(REP_CASES))
def test_repel_goldens(golden_dir, case):
    """test/rephippo.cpp EREPEL-1 (pentane + acetamide, HIPPO-2019) and
    EREPEL-2 (ten waters, the HIPPO water set), 7 Å cutoff: HIPPO Pauli
    repulsion against test/ref/rephippo.{1,2}.txt at the reference test's
    tolerances (energy and gradient 1e-4, virial 1e-3), and the interaction
    count."""
    ref = rep_ref(golden_dir, case)
    a = hippo_rep_system(golden_dir, case).arrays()
    assert a["repel"] and a["rep_taper"] == pytest.approx(6.3)
    e, g, v, cnt = oracle.Oracle(a).repel()
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    print(f"E={e:.6f} ref={ref['energy']} count={cnt} max|dg|={np.abs(g - rg).max():.2e} "
          f"max|dv|={np.abs(v - rv).max():.2e}")
    assert abs(e - ref["energy"]) < 1e-4
    assert np.abs(g - rg).max() < 1e-4
    assert np.abs(v - rv).max() < 1e-3
    assert cnt == ref["count"]


@pytest.mark.parametrize("r,ai,ak", [(1.3, 4.64, 4.64), (2.2, 4.64, 4.2256), (1.8, 3.1, 4.9), (2.6, 4.64, 4.6529)])
def test_repel_kernel_is_orbital_overlap(r, ai, ak):
    """sqrt(D_0) is the overlap of two normalized 1s Slater orbitals of
    exponents ai/2, ak/2 at distance r (a numerical 2D integral), close
    exponents included."""
    za, zb = ai / 2, ak / 2
    phi = lambda z, t: np.sqrt(z ** 3 / np.pi) * np.exp(-z * t)
    f = lambda mu, t: 2 * np.pi * t * t * phi(za, t) * phi(zb, np.sqrt(max(t * t + r * r - 2 * t * r * mu, 0.0)))
    S = integrate.dblquad(f, 0, 40.0 / min(za, zb), -1, 1, epsabs=1e-13, epsrel=1e-11)[0]
    D = oracle.repel_series(r, ai, ak)
    assert abs(np.sqrt(D[0]) - S) < 1e-9 * max(1.0, S)


@pytest.mark.parametrize("ai,ak", [(4.64, 4.64), (4.64, 4.2256), (4.64, 4.6529)])
def test_repel_derivative_series(ai, ak):
    """D_(n+1) = -(1/r) dD_n/dr by central differences."""
    for r in (1.1, 2.4, 4.0):
        h = 1e-5 * r
        Dp, Dm, D = oracle.repel_series(r + h, ai, ak), oracle.repel_series(r - h, ai, ak), oracle.repel_series(r, ai, ak)
        for n in range(5):
            d = (Dp[n] - Dm[n]) / (2 * h)
            assert abs(-d / r - D[n + 1]) < 1e-7 * abs(D[n + 1]), (r, n)
