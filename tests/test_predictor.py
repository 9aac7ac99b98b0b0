"""ASPC predictor and PCG-from-a-guess: closed forms and the reference's
coefficient table (src/cu/upredict.cu:50-65, committed as numbers).  CPU only."""
import json
import os
from fractions import Fraction

import numpy as np

import oracle
import t9hip

WATER_KEY = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid 24 24 24\npme-order 5\npolar-eps 1e-6\n"
             "polarization mutual\nneighbor-list\nlist-buffer 1.5\nusolve-cutoff 4.5\nusolve-diag 2.0\n")


def test_aspc_matches_reference_table(golden_dir):
    ref = np.array(json.load(open(os.path.join(golden_dir, "aspc_reference.json")))["aspc"])
    B = oracle.aspc_coefficients()
    assert len(B) == 16
    # the reference stores rational approximations of the same coefficients
    # (largest relative difference 8.6e-7, e.g. -2329/400 for -5642/969)
    np.testing.assert_allclose(B, ref, rtol=1e-6, atol=0)


def test_aspc_moments():
    """sum B_j = 1 (constants are predicted exactly), odd moments vanish (so
    u(t) = a + b t + c t^3 + ... about the predicted time is exact), and the
    second moment is -16/29 (the predictor is not exact for t^2)."""
    from math import comb
    k = 14
    den = comb(2 * k + 2, k + 1)
    B = [Fraction((-1) ** (j + 1) * j * comb(2 * k + 4, k + 2 - j), den) for j in range(1, k + 3)]
    assert sum(B) == 1
    for p in (1, 3, 5, 7, 9, 11, 13, 15):
        assert sum(b * j ** p for j, b in zip(range(1, k + 3), B)) == 0
    assert sum(b * j ** 2 for j, b in zip(range(1, k + 3), B)) == Fraction(-16, 29)
    Bf = oracle.aspc_coefficients()
    hist = [1.5 - 0.3 * (-j) + 0.01 * (-j) ** 3 for j in range(1, 17)]  # u(-j), predicted u(0) = 1.5
    assert abs(float(np.dot(Bf, hist)) - 1.5) < 1e-9


def test_induce_from_guess_converges_to_same_solution():
    s = t9hip.System.water_box(216, 6, 18.643, 2024, key=WATER_KEY)
    o = oracle.Oracle(s.arrays())
    ud, up, it, rms, rc = o.induce(eps=1e-10)
    rng = np.random.default_rng(3)
    u0d = ud + 0.01 * rng.standard_normal(ud.shape)
    u0p = up + 0.01 * rng.standard_normal(up.shape)
    a, b, it2, rms2, g, rc2 = o.induce_from(u0d, u0p, eps=1e-10)
    assert rc2 == 0 and g > 1e-3
    np.testing.assert_allclose(a, ud, atol=1e-9)
    np.testing.assert_allclose(b, up, atol=1e-9)
    # the exact solution as guess: tiny initial residual, minimum iterations
    a, b, it3, rms3, g3, rc3 = o.induce_from(ud, up, eps=1e-8)
    assert g3 < 1e-8 and it3 == 3
