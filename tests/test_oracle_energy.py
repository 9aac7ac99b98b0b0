"""Pins of the oracle's energy, gradient and virial (multipole + polarization):
the reference's golden values of test/localframe.cpp and
test/localframe3.cpp, central finite differences of the oracle energy (with
the dipoles re-solved), a strain derivative for the virial diagonal, and the
gas-phase identity virial = sym(sum x (x) dE/dx).  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle
import t9hip

K_LF1 = "multipoleterm only\n"
K_LF2 = K_LF1 + "ewald\newald-cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\n"
K_LF3 = "usolve-cutoff 0.01\npolarizeterm only\n"
K_LF4 = K_LF3 + "ewald\newald-cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\n"
# test/localframe3.cpp:9-22 (valence and vdW terms are not part of this path)
K_LF32 = ("cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\nbondterm none\nangleterm none\n"
     "strbndterm none\nureybradterm none\ntorsionterm none\nvdwterm none\newald\n")


def _ref(golden_dir):
    return json.load(open(os.path.join(golden_dir, "localframe_ref.json")))


def _get(ref, case, section, name):
    for k, v in ref.items():
        c, s, nm = k.split("/")
        if c == case and s.startswith(section) and nm == name:
            return np.array(v)
    raise KeyError((case, section, name))


def _solve(o, polar):
    if not polar:
        return None, None
    ud, up, it, rms, rc = o.induce(eps=1e-11)
    assert rc == 0
    return ud, up


def _total(o, polar, mpole, ud, up):
    e, g, v = o.energy(ud, up, mpole=mpole, polar=polar)
    return (e["em"] if mpole else 0.0) + (e["ep"] if polar else 0.0), g, v, e


@pytest.mark.parametrize("case,key,section,mpole,polar", [
    ("Local-Frame-1", K_LF1, "empole", True, False),
    ("Local-Frame-2", K_LF2, "empole", True, False),
    ("Local-Frame-3", K_LF3, "various", False, True),
    ("Local-Frame-4", K_LF4, "various", False, True),
])
def test_localframe_energy_gradient(golden_dir, case, key, section, mpole, polar):
    ref = _ref(golden_dir)
    s = t9hip.System.from_files(os.path.join(golden_dir, "local_frame.xyz"), key=key)
    a = s.arrays()
    o = oracle.Oracle(a)
    ud, up = _solve(o, polar)
    etot, g, v, e = _total(o, polar, mpole, ud, up)
    assert abs(etot - float(_get(ref, case, section, "ref_eng"))) < 1e-4
    # the reference skips the last molecule, an NH3 whose atoms all lie in one
    # plane (test/localframe.cpp:14-22); its 3-fold frame is near-degenerate
    rg = _get(ref, case, section, "ref_grad")
    np.testing.assert_allclose(g[:-4], rg[:-4], atol=1e-4)
    if case == "Local-Frame-2":  # test/localframe.cpp:124-130
        assert abs(e["em_real"] - float(_get(ref, case, section, "ref_ereal"))) < 1e-6
        assert abs(e["em_recip"] - float(_get(ref, case, section, "ref_erecip"))) < 1e-6
        assert abs(e["em_self"] - float(_get(ref, case, section, "ref_eself"))) < 1e-6
    # virial: equal to the reference once the skipped molecule's gradient
    # difference is accounted for (sym(sum x (x) dg) over its 4 atoms)
    x = a["xyz"]
    dg = (rg - g)[-4:]
    corr = x[-4:].T @ dg
    corr = 0.5 * (corr + corr.T)
    np.testing.assert_allclose(v + corr, _get(ref, case, section, "ref_v"), atol=1e-3)
    if not a["periodic"]:
        w = x.T @ g
        np.testing.assert_allclose(v, 0.5 * (w + w.T), atol=1e-8)


def test_localframe3_emplar_ewald(golden_dir):
    """Fused multipole + polarization under PME, every frame type, all atoms
    (test/localframe3.cpp:106-163): energy 1e-4, gradient 1e-4, virial 1e-3."""
    ref = _ref(golden_dir)
    s = t9hip.System.from_files(os.path.join(golden_dir, "local_frame2.xyz"), key=K_LF32)
    o = oracle.Oracle(s.arrays())
    ud, up = _solve(o, True)
    etot, g, v, e = _total(o, True, True, ud, up)
    case, sec = "Local-Frame3-2", "- emplar"
    assert abs(etot - float(_get(ref, case, sec, "ref_eng"))) < 1e-4
    # split given in the reference's comments (test/localframe3.cpp:110-111)
    assert abs(e["em"] - (-92.5088)) < 1e-4
    assert abs(e["ep"] - (-17.3296)) < 1e-4
    np.testing.assert_allclose(g, _get(ref, case, sec, "ref_g"), atol=1e-4)
    np.testing.assert_allclose(v, _get(ref, case, sec, "ref_v"), atol=1e-3)


def _lf32(golden_dir):
    s = t9hip.System.from_files(os.path.join(golden_dir, "local_frame2.xyz"), key=K_LF32)
    return s, s.arrays()


def test_gradient_finite_difference(golden_dir):
    """dE/dx against central differences with the dipoles re-solved, on
    coordinates of atoms with every frame type (PME, scaled pairs)."""
    s, a = _lf32(golden_dir)
    o = oracle.Oracle(a)
    ud, up = _solve(o, True)
    _, g, _, _ = _total(o, True, True, ud, up)
    h = 1e-5
    for atom in (1, 2, 5, 8, 14, 15):
        for d in range(3):
            es = []
            for sgn in (1, -1):
                x = a["xyz"].copy()
                x[atom, d] += sgn * h
                o.set_xyz(x)
                u1, u2 = _solve(o, True)
                es.append(_total(o, True, True, u1, u2)[0])
            fd = (es[0] - es[1]) / (2 * h)
            assert abs(fd - g[atom, d]) < 2e-5 * max(1.0, abs(g[atom, d])), (atom, d, fd, g[atom, d])
    o.set_xyz(a["xyz"])


def test_virial_strain_derivative(golden_dir):
    """Virial diagonal = dE/d(strain) with box and coordinates scaled together."""
    s, a = _lf32(golden_dir)
    o = oracle.Oracle(a)
    ud, up = _solve(o, True)
    _, _, v, _ = _total(o, True, True, ud, up)
    h = 1e-5
    for d in range(3):
        es = []
        for sgn in (1, -1):
            b = dict(a)
            f = np.ones(3)
            f[d] += sgn * h
            b["xyz"] = a["xyz"] * f
            b["box"] = np.asarray(a["box"]) * f
            ob = oracle.Oracle(b)
            u1, u2 = _solve(ob, True)
            es.append(_total(ob, True, True, u1, u2)[0])
        fd = (es[0] - es[1]) / (2 * h)
        assert abs(fd - v[d, d]) < 1e-4 * max(1.0, abs(v[d, d])), (d, fd, v[d, d])


def test_ewald_split_invariance(golden_dir):
    """Energy and gradient do not depend on the Ewald coefficient (fine grid)."""
    s, a = _lf32(golden_dir)
    res = []
    for alpha, K in ((0.40, 64), (0.55, 80)):
        b = dict(a)
        b["aewald"] = alpha
        b["grid"] = np.array([K, K, K], dtype=np.int32)
        b["ewald_cut"] = 9.9
        o = oracle.Oracle(b)
        ud, up = _solve(o, True)
        res.append(_total(o, True, True, ud, up))
    assert abs(res[0][0] - res[1][0]) < 2e-4
    np.testing.assert_allclose(res[0][1], res[1][1], atol=2e-3)
