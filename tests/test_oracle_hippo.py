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
