"""Pins of the FP64 oracle against further reference fixtures and against
what the mathematics fixes (CPU only):

* bond and Urey-Bradley goldens (test/bond.cpp + test/ref/bond.txt on
  trp-cage, test/urey.cpp + test/ref/urey.txt on ten waters): every
  anharmonic coefficient and the Urey-Bradley factor enter these numbers;
* polpair (test/polpair.cpp, test/ref/polpair.{1,2}.txt): a Na+/Cl- pair with
  a pair-specific Thole value, Ewald and non-Ewald;
* NaCl-2 / NaCl-3 multipoles (test/nacl.cpp:207-304): non-Ewald periodic
  cutoff and PME with key-file multipole records;
* Local-Frame3-1 (test/localframe3.cpp:29-104): the fused multipole +
  polarization energy with a non-Ewald periodic cutoff;
* the sparse preconditioner as the first-order Neumann expansion of
  (alpha^-1 - T)^-1 (M = udiag alpha + alpha T alpha inside the usolve
  cutoff), built here from the Thole-damped dipole tensor;
* the angle term's sextic polynomial in the deviation (degrees) and the
  argon-box neighbor sets of test/nblist.hh against a minimum-image brute force.
"""
import json
import os

import numpy as np
import pytest

import oracle
import t9hip

PP_KEY = "neighbor-list\na-axis 20.0\ncutoff 7.0\nvdwterm none\npolpair 7 15 0.05\n"
NACL2_KEY = ("a-axis 18.643\nneighbor-list\nlist-buffer 0.1\nvdw-cutoff 2.6\nmpole-cutoff 5.0\n"
             "multipoleterm only\n")
NACL3_KEY = ("neighbor-list\nlist-buffer 0.1\nvdw-cutoff 2.6\nmpole-cutoff 5.0\na-axis 20.0\npme-grid 40 40 40\n"
             "ewald\newald-cutoff 7.0\nmultipoleterm only\n")
LF3_KEY = ("cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\nbondterm none\nangleterm none\n"
           "strbndterm none\nureybradterm none\ntorsionterm none\nvdwterm none\n")
ARBOX_KEY = "neighbor-list\nlist-buffer 1.0\ncutoff 9.0\nvdwterm only\n"


def more_ref(golden_dir):
    return json.load(open(os.path.join(golden_dir, "more_ref.json")))


def nacl3_key(golden_dir):
    return open(os.path.join(golden_dir, "nacl4_multipoles.key")).read() + NACL3_KEY


def _cmp(e, g, v, ref, eps_e=1e-4, eps_g=1.1e-4, eps_v=1e-3, label=""):
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    print(f"{label}: E={e:.6f} ref={ref['energy']} max|dg|={np.abs(g - rg).max():.2e} "
          f"max|dv|={np.abs(v - rv).max():.2e}")
    assert abs(e - ref["energy"]) < eps_e
    assert np.abs(g - rg).max() < eps_g
    assert np.abs(v - rv).max() < eps_v


@pytest.mark.parametrize("case,xyz,prm,key", [
    ("bond_trpcage", "trpcage.xyz", "amoebapro13_bond_subset.prm", "bondterm only\n"),
    ("urey_h2o10", "h2o10.xyz", "water03_subset.prm", "ureybradterm only\n"),
])
def test_valence_goldens(golden_dir, case, xyz, prm, key):
    ref = more_ref(golden_dir)[case]
    s = t9hip.System.from_files(os.path.join(golden_dir, xyz), prm=os.path.join(golden_dir, prm), key=key)
    n_terms = int(s.scalar("nbond" if case.startswith("bond") else "nurey"))
    assert n_terms == ref["count"]
    e, g, v = oracle.Oracle(s.arrays()).md_energy(vdw=False, valence=True)
    _cmp(e["ebond"] + e["eangle"] + e["eurey"], g, v, ref, label=case)


def _elec(s, mpole=True, polar=True):
    o = oracle.Oracle(s.arrays())
    ud = up = None
    if polar:
        ud, up, it, rms, rc = o.induce()
        assert rc == 0
    e, g, v = o.energy(ud, up, mpole=mpole, polar=polar)
    return (e["em"] if mpole else 0) + (e["ep"] if polar else 0), g, v


@pytest.mark.parametrize("case,extra", [("polpair_ewald", "ewald\n"), ("polpair_nonewald", "")])
def test_polpair(golden_dir, case, extra):
    s = t9hip.System.from_files(os.path.join(golden_dir, "polpair_nacl.xyz"), key=PP_KEY + extra)
    thl = s.f64("thlval", int(s.scalar("njpolar")) ** 2)
    assert np.any(thl == 0.05) and not np.all(thl == thl[0])  # the pair value replaces the type values
    _cmp(*_elec(s), more_ref(golden_dir)[case], label=case)


def test_nacl2_nonewald_pbc(golden_dir):
    s = t9hip.System.from_files(os.path.join(golden_dir, "nacl1.xyz"), key=NACL2_KEY)
    _cmp(*_elec(s, polar=False), more_ref(golden_dir)["NaCl-2"], eps_e=1e-4, eps_g=1e-4, eps_v=1e-3,
         label="NaCl-2")


def test_nacl3_pme_key_multipoles(golden_dir):
    s = t9hip.System.from_files(os.path.join(golden_dir, "nacl4.xyz"), key=nacl3_key(golden_dir))
    pole = s.f64("pole", (2, 10))
    assert abs(pole[0, 3] - 0.12578 * 0.529177210544) < 1e-9  # the key file's record, not the parameter file's
    _cmp(*_elec(s, polar=False), more_ref(golden_dir)["NaCl-3"], eps_e=5e-4, eps_g=5e-4, eps_v=5e-4,
         label="NaCl-3")


def test_localframe3_nonewald_pbc(golden_dir):
    ref = json.load(open(os.path.join(golden_dir, "localframe_ref.json")))
    get = lambda nm: np.array([v for k, v in ref.items() if k.startswith("Local-Frame3-1/") and k.endswith("/" + nm)][0])
    s = t9hip.System.from_files(os.path.join(golden_dir, "local_frame2.xyz"), key=LF3_KEY)
    e, g, v = _elec(s)
    _cmp(e, g, v, {"energy": float(get("ref_eng")), "grad": get("ref_g"), "virial": get("ref_v")},
         label="Local-Frame3-1")


def _thole_tensor(R, ai, ak, a):
    """Thole-damped dipole field tensor T (field at i of a dipole at k is T m),
    u = r / (alpha_i alpha_k)^(1/6), lambda3 = 1 - e^(-a u^3),
    lambda5 = 1 - (1 + a u^3) e^(-a u^3)."""
    r = np.linalg.norm(R)
    u3 = a * (r / (ai * ak) ** (1.0 / 6.0)) ** 3
    ex = np.exp(-u3)
    l3, l5 = 1 - ex, 1 - (1 + u3) * ex
    return l5 * 3 * np.outer(R, R) / r ** 5 - l3 * np.eye(3) / r ** 3


def test_precond_is_neumann_expansion(golden_dir):
    """Two polarizable ions 2.69 Å apart (pair Thole 0.05): z = M r with
    M = [[udiag a_i I, a_i a_k T], [a_k a_i T, udiag a_k I]] inside the
    usolve cutoff, M = diag(a) r (no udiag) for the diagonal form."""
    s = t9hip.System.from_files(os.path.join(golden_dir, "polpair_nacl.xyz"),
                                key=PP_KEY + "ewald\nusolve-cutoff 4.5\nusolve-diag 2.0\n")
    o = oracle.Oracle(s.arrays())
    xyz = s.f64("xyz", (2, 3))
    al = s.f64("polarity", 2)
    T = _thole_tensor(xyz[1] - xyz[0], al[0], al[1], 0.05)
    M = np.block([[2.0 * al[0] * np.eye(3), al[0] * al[1] * T], [al[0] * al[1] * T, 2.0 * al[1] * np.eye(3)]])
    rng = np.random.default_rng(1)
    rd, rp = rng.standard_normal((2, 2, 3))
    zd, zp = o.precond(rd, rp, sparse=True)
    np.testing.assert_allclose(zd.ravel(), M @ rd.ravel(), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(zp.ravel(), M @ rp.ravel(), rtol=1e-12, atol=1e-14)
    zd, zp = o.precond(rd, rp, sparse=False)
    np.testing.assert_allclose(zd, al[:, None] * rd, rtol=1e-14)
    # beyond the usolve cutoff only the diagonal remains
    zd, _ = o.precond(rd, rp, sparse=True, ucut=2.0)
    np.testing.assert_allclose(zd, 2.0 * al[:, None] * rd, rtol=1e-14)


def test_precond_symmetric_water():
    """a . M b = b . M a on the 648-atom box (every pair inside 4.5 Å)."""
    key = ("ewald\newald-cutoff 7.0\npme-grid 24 24 24\nneighbor-list\nlist-buffer 1.5\nusolve-cutoff 4.5\n"
           "usolve-diag 2.0\nvdwterm none\nbondterm none\n")
    s = t9hip.System.water_box(216, 6, 18.643, 2024, key=key)
    o = oracle.Oracle(s.arrays())
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal((2, s.n, 3))
    ma, _ = o.precond(a, a)
    mb, _ = o.precond(b, b)
    lhs, rhs = float(np.sum(a * mb)), float(np.sum(b * ma))
    assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1.0)
    # and the pair part is not empty: it differs from the diagonal form
    d, _ = o.precond(a, a, sparse=False)
    assert np.abs(ma - 2.0 * d).max() > 1e-3


def test_angle_polynomial():
    """The angle energy of one water against E = k dt^2 (1 + c3 dt + c4 dt^2 +
    c5 dt^3 + c6 dt^4) (dt in degrees, k in kcal/mol/rad^2 times (pi/180)^2),
    at a geometry 15 degrees off the ideal angle, so every coefficient counts."""
    s = t9hip.System.water_box(216, 6, 18.643, 2024,
                               key="angleterm only\nneighbor-list\nlist-buffer 1.5\na-axis 18.643\n")
    d = s.arrays()
    a, b, c = d["angle_abc"][0]
    kt = d["angle_kt"][0]
    xyz = d["xyz"].copy()
    # open the angle of molecule 0 to ideal + 15 degrees in the molecular plane
    u, v = xyz[a] - xyz[b], xyz[c] - xyz[b]
    ru = np.linalg.norm(u)
    e1 = u / ru
    w = v - np.dot(v, e1) * e1
    e2 = w / np.linalg.norm(w)
    th = np.radians(kt[1] + 15.0)
    xyz[c] = xyz[b] + np.linalg.norm(v) * (np.cos(th) * e1 + np.sin(th) * e2)
    s.set_xyz(xyz)
    d = s.arrays()
    e, g, _ = oracle.Oracle(d).md_energy(vdw=False, valence=True)
    dt = 15.0
    poly = 1 + d["angle_cubic"] * dt + d["angle_quartic"] * dt ** 2 + d["angle_pentic"] * dt ** 3 + \
        d["angle_sextic"] * dt ** 4
    one = kt[0] * (np.pi / 180) ** 2 * dt * dt * poly
    # the other 215 molecules sit at their construction angle
    others = 0.0
    for q in range(1, len(d["angle_abc"])):
        aa, bb, cc = d["angle_abc"][q]
        x1, x2 = d["xyz"][aa] - d["xyz"][bb], d["xyz"][cc] - d["xyz"][bb]
        t = np.degrees(np.arccos(np.dot(x1, x2) / np.linalg.norm(x1) / np.linalg.norm(x2))) - d["angle_kt"][q][1]
        others += d["angle_kt"][q][0] * (np.pi / 180) ** 2 * t * t * (
            1 + d["angle_cubic"] * t + d["angle_quartic"] * t ** 2 + d["angle_pentic"] * t ** 3 +
            d["angle_sextic"] * t ** 4)
    assert abs(e["eangle"] - (one + others)) < 1e-9 * max(1.0, abs(one + others))
    assert d["angle_sextic"] != 0 and d["angle_cubic"] != 0


def test_arbox_neighbor_sets(golden_dir):
    """test/nblist.hh holds, for 5 frames of 216 argon atoms (21.857 Å box),
    the neighbors k > i within the 9 Å vdW cutoff; the reference's test
    requires its list to contain them after a refresh on each frame
    (test/nblist.cpp:145-193).  Each frame's sets are the minimum-image
    brute force at 9 Å (FP64)."""
    z = np.load(os.path.join(golden_dir, "arbox_nblist.npz"))
    xyz, box, ans = z["xyz"], z["box"], z["ans"]
    L = box[0][0]
    for f in range(len(xyz)):
        d = xyz[f][:, None, :] - xyz[f][None, :, :]
        d -= L * np.round(d / L)
        r = np.sqrt(np.sum(d * d, axis=-1))
        for i in range(216):
            brute = {k for k in range(i + 1, 216) if r[i, k] < 9.0}
            ref = {int(k) for k in ans[f, i] if k >= 0}
            assert ref == brute, (f, i, sorted(ref ^ brute))
