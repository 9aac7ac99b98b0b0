"""Pins of the oracle's MD terms: the reference's vdW golden values for a
Na+/Cl- pair with and without the taper (test/nacl.cpp NaCl-1), central
finite differences of the vdW and valence energies on a water box (reduced
hydrogen sites, every valence term), the strain derivative for the virial
diagonal, and the valence energy at the ideal geometry.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle
import t9hip

NACL_KEY = "a-axis 18.643\nneighbor-list\nlist-buffer 0.1\nvdw-cutoff 2.6\nmpole-cutoff 5.0\nvdwterm only\n"
WATER_KEY = ("ewald\newald-cutoff 4.0\nvdw-cutoff 4.5\npme-grid 16 16 16\nneighbor-list\nlist-buffer 0.5\n"
             "polar-eps 1e-6\n")


@pytest.mark.parametrize("section", ["no switch", "switch, near cut", "switch, near off"])
def test_nacl_vdw_golden(golden_dir, tmp_path, section):
    ref = json.load(open(os.path.join(golden_dir, "nacl_vdw_ref.json")))[section]
    f = tmp_path / "nacl.xyz"
    f.write_text(f"2 NaCl\n1 Na+ 0.0 0.0 0.0 7\n2 Cl- {ref['r']:.6f} 0.0 0.0 15\n")
    s = t9hip.System.from_files(str(f), key=NACL_KEY)
    o = oracle.Oracle(s.arrays(), cut=5.0)
    e, g, v = o.md_energy(valence=False)
    eps = 1e-3  # the reference's margin (test/nacl.cpp:44)
    assert abs(e["evdw"] - ref["eng"]) < eps
    np.testing.assert_allclose(g, np.array(ref["grad"]), atol=eps)
    np.testing.assert_allclose(v, np.array(ref["vir"]), atol=eps)
    assert oracle.vdw_count(s.arrays()) == int(ref["count"])  # test/nacl.cpp ref_count


def _water():
    s = t9hip.System.water_box(27, 3, 9.32, 5, key=WATER_KEY)
    return s, s.arrays()


def test_md_terms_finite_difference():
    s, a = _water()
    o = oracle.Oracle(a)
    e, g, v = o.md_energy()
    assert e["ebond"] > 0 and e["eangle"] > 0  # rigid construction geometry is strained
    h = 1e-6
    x0 = a["xyz"].copy()
    for atom in (0, 1, 2, 13):
        for d in range(3):
            es = []
            for sg in (1, -1):
                x = x0.copy()
                x[atom, d] += sg * h
                o.set_xyz(x)
                ee = o.md_energy()[0]
                es.append(sum(ee.values()))
            fd = (es[0] - es[1]) / (2 * h)
            assert abs(fd - g[atom, d]) < 1e-5 * max(1.0, abs(g[atom, d])), (atom, d, fd, g[atom, d])
    o.set_xyz(x0)


def test_md_terms_virial_strain():
    s, a = _water()
    o = oracle.Oracle(a)
    _, _, v = o.md_energy()
    h = 1e-6
    for d in range(3):
        es = []
        for sg in (1, -1):
            b = dict(a)
            f = np.ones(3)
            f[d] += sg * h
            b["xyz"] = a["xyz"] * f
            b["box"] = np.asarray(a["box"]) * f
            es.append(sum(oracle.Oracle(b).md_energy()[0].values()))
        fd = (es[0] - es[1]) / (2 * h)
        assert abs(fd - v[d, d]) < 1e-5 * max(1.0, abs(v[d, d])), (d, fd, v[d, d])


def test_valence_zero_at_ideal_geometry(tmp_path):
    """One water at the AMOEBA09 ideal geometry (0.9572 Å, 108.50°, H-H at the
    Urey-Bradley length 1.5537 Å only approximately): bond and angle energies
    vanish; the Urey-Bradley term is k (r13 - r0)^2 exactly."""
    th = np.radians(108.50)
    r = 0.9572
    x = np.array([[0, 0, 0], [r, 0, 0], [r * np.cos(th), r * np.sin(th), 0]])
    f = tmp_path / "w.xyz"
    f.write_text("3 water\n" + "".join(f"{i + 1} {n} {x[i, 0]:.12f} {x[i, 1]:.12f} {x[i, 2]:.12f} {t} {b}\n"
                                      for i, (n, t, b) in enumerate([("O", 36, "2 3"), ("H", 37, "1"),
                                                                     ("H", 37, "1")])))
    s = t9hip.System.from_files(str(f), key="multipoleterm none\npolarizeterm none\nvdwterm none\n")
    o = oracle.Oracle(s.arrays())
    e, g, v = o.md_energy(vdw=False)
    assert abs(e["ebond"]) < 1e-18 and abs(e["eangle"]) < 1e-18
    r13 = np.linalg.norm(x[1] - x[2])
    assert abs(e["eurey"] - (-7.60) * (r13 - 1.5337) ** 2) < 1e-12
