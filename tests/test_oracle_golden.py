"""Pins of the FP64 oracle: the reference's own golden vectors, plus checks the
mathematics fixes (Ewald alpha rule, DFT against numpy's FFT, Ewald-splitting
invariance, symmetry of the dipole interaction, the solved linear system).
CPU only."""
import json
import os

import numpy as np
import pytest

import oracle
import t9hip

DEBYE = 4.803206802
LF_KEY3 = "usolve-cutoff 0.01\npolarizeterm only\n"
LF_KEY4 = LF_KEY3 + "ewald\newald-cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\n"


def _ref(golden_dir):
    return json.load(open(os.path.join(golden_dir, "localframe_ref.json")))


def _get(ref, case, section, name):
    for k, v in ref.items():
        c, s, nm = k.split("/")
        if c == case and s.startswith(section) and nm == name:
            return np.array(v)
    raise KeyError((case, section, name))


def _lf_system(golden_dir, key):
    return t9hip.System.from_files(os.path.join(golden_dir, "local_frame.xyz"), key=key)


def _synthetic_dipoles(n):
    # test/localframe.cpp:272-277
    i = np.arange(n)[:, None]
    j = np.arange(3)[None, :]
    return 0.1 * (i + 1) + 0.03 * (j + 1), 0.1 * (i + 1) - 0.03 * (j + 1)


@pytest.mark.parametrize("case,key", [("Local-Frame-3", LF_KEY3), ("Local-Frame-4", LF_KEY4)])
def test_localframe_fields_dipoles_energy(golden_dir, case, key):
    ref = _ref(golden_dir)
    s = _lf_system(golden_dir, key)
    o = oracle.Oracle(s.arrays())
    eps = 1e-4  # the reference's own margin (test/localframe.cpp:250)
    fd, fp = o.dfield()
    np.testing.assert_allclose(fd, _get(ref, case, "dfield", "ref_dir_field_d"), atol=eps)
    np.testing.assert_allclose(fp, _get(ref, case, "dfield", "ref_dir_field_p"), atol=eps)
    ud, up = _synthetic_dipoles(s.n)
    ufd, ufp = o.ufield(ud, up)
    np.testing.assert_allclose(ufd, _get(ref, case, "ufield", "ref_ufield_d"), atol=eps)
    np.testing.assert_allclose(ufp, _get(ref, case, "ufield", "ref_ufield_p"), atol=eps)
    u1, u2, it, rms, rc = o.induce()
    assert rc == 0 and rms < 1e-6
    np.testing.assert_allclose(u1 * DEBYE, _get(ref, case, "default induce", "ref_ud_debye"), atol=eps)
    np.testing.assert_allclose(u2 * DEBYE, _get(ref, case, "default induce", "ref_up_debye"), atol=eps)
    # dot-product polarization energy, -1/2 f sum u_d . E_p (src/cu/amoeba/epolar.cu:15-30)
    e = -0.5 * 332.0637133 * np.sum(u1 * fp)
    assert abs(e - float(_get(ref, case, "epolar via dot", "ref_eng"))) < eps


def test_ewald_alpha_and_grid_rules(golden_dir):
    s = _lf_system(golden_dir, LF_KEY4)
    a = s.scalar("aewald")
    from math import erfc
    assert abs(erfc(a * 7.0) / 7.0 - 1e-8) < 1e-14
    assert [int(s.scalar("grid%d" % j)) for j in (1, 2, 3)] == [24, 24, 24]


def _tinkernist_system(golden_dir, frame, tmp_path):
    z = np.load(os.path.join(golden_dir, "tinkernist.npz"))
    n = len(z["types"])
    xyz = z["arc"][frame]
    lines = [f"{n} tinkernist frame {frame}", "30.0 30.0 30.0 90.0 90.0 90.0"]
    for a in range(n):
        bonds = " ".join(str(b + 1) for b in z["i12"][a] if b >= 0)
        lines.append(f"{a + 1} X {xyz[a, 0]:.6f} {xyz[a, 1]:.6f} {xyz[a, 2]:.6f} {z['types'][a]} {bonds}")
    p = tmp_path / "tn.xyz"
    p.write_text("\n".join(lines) + "\n")
    key = "a-axis 30.0\nneighbor-list\nvdw-cutoff 12.0\newald\newald-cutoff 7.0\npolar-eps 0.00001\n"
    return t9hip.System.from_files(str(p), key=key), z


@pytest.mark.parametrize("frame", [0, 1])
def test_tinkernist_dipoles(golden_dir, tmp_path, frame):
    """All 2,684 atoms of test/ref/tinkernist.{uind,udir} (the reference test
    checks 4 atoms at 1e-3, test/tinkernist.cpp:226-339)."""
    s, z = _tinkernist_system(golden_dir, frame, tmp_path)
    d = s.arrays()
    o = oracle.Oracle(d)
    fd, fp = o.dfield()
    ud, up, it, rms, rc = o.induce()
    assert rc == 0
    np.testing.assert_allclose(d["polarity"][:, None] * fd * DEBYE, z["udir"][frame], atol=2e-5)
    np.testing.assert_allclose(ud * DEBYE, z["uind"][frame], atol=2e-5)
    # .def holds the direct field times the field-unit constant 1439.96455
    np.testing.assert_allclose(fd * 1439.96455, z["def"][frame], atol=2e-3)


def test_dft_matches_numpy_fft(golden_dir):
    s = _lf_system(golden_dir, LF_KEY4)
    o = oracle.Oracle(s.arrays())
    rng = np.random.default_rng(1)
    g = rng.standard_normal((24, 24, 24)) + 1j * rng.standard_normal((24, 24, 24))
    # grid index x + K1*(y + K2*z): numpy axes (z, y, x)
    fwd = o.dft3(g, -1)
    np.testing.assert_allclose(fwd, np.fft.fftn(g), atol=1e-9)
    bwd = o.dft3(g, +1)
    np.testing.assert_allclose(bwd, np.fft.ifftn(g) * g.size, atol=1e-9)


def test_ewald_split_invariance(golden_dir):
    """The total field must not depend on the Ewald coefficient (up to the PME
    discretisation error, which falls as h^4: 1.6e-4 at K=48, 7.6e-6 at K=96)."""
    fields = []
    for alpha in (0.45, 0.60):
        s = _lf_system(golden_dir, LF_KEY4 + "pme-grid 96 96 96\n")
        s.set_scalar("aewald", alpha)
        s.set_scalar("ewald_cut", 9.9)
        o = oracle.Oracle(s.arrays())
        fields.append(o.dfield()[0])
        ud, up = _synthetic_dipoles(s.n)
        fields.append(o.ufield(ud, up)[0])
    np.testing.assert_allclose(fields[0], fields[2], atol=2e-5)
    np.testing.assert_allclose(fields[1], fields[3], atol=2e-5)


@pytest.mark.parametrize("key", [LF_KEY3, LF_KEY4])
def test_dipole_interaction_symmetric(golden_dir, key):
    """-T is symmetric: a . T(b) == b . T(a), for real space and PME alike."""
    s = _lf_system(golden_dir, key)
    o = oracle.Oracle(s.arrays())
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal((2, s.n, 3))
    fa = o.ufield(a, a)[0]
    fb = o.ufield(b, b)[0]
    assert abs(np.sum(b * fa) - np.sum(a * fb)) < 1e-10 * np.abs(fa).max() * s.n


def test_induce_solves_linear_system(golden_dir):
    """u/alpha - ufield(u) == E at convergence (residual recomputed from scratch)."""
    s = _lf_system(golden_dir, LF_KEY4)
    d = s.arrays()
    o = oracle.Oracle(d)
    fd, fp = o.dfield()
    ud, up, it, rms, rc = o.induce(eps=1e-10)
    ufd, ufp = o.ufield(ud, up)
    res = ud / d["polarity"][:, None] - ufd - fd
    assert np.sqrt(np.mean(np.sum((res * d["polarity"][:, None]) ** 2, axis=1))) * DEBYE < 1e-8


# analyze-mode interaction counts (test/localframe.cpp:47,90,131,157,342,519;
# test/localframe3.cpp:114,141-142)
COUNT_CASES = [
    ("Local-Frame-1", "empole -- gas", "local_frame.xyz", "multipoleterm only\n", (201, 0)),
    ("Local-Frame-2", "empole -- pme", "local_frame.xyz",
     "multipoleterm only\newald\newald-cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\n", (222, 0)),
    ("Local-Frame-3", "various epolar", "local_frame.xyz", LF_KEY3, (0, 201)),
    ("Local-Frame-4", "various epolar", "local_frame.xyz", LF_KEY4, (0, 222)),
    ("Local-Frame3-2", "- emplar", "local_frame2.xyz",
     ("cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\nbondterm none\nangleterm none\n"
     "strbndterm none\nureybradterm none\ntorsionterm none\nvdwterm none\newald\n"), (138, 138)),
]


@pytest.mark.parametrize("case,section,xyz,key,expect", COUNT_CASES)
def test_interaction_counts_golden(golden_dir, case, section, xyz, key, expect):
    ref = int(_get(_ref(golden_dir), case, section, "ref_count"))
    assert ref in expect
    s = t9hip.System.from_files(os.path.join(golden_dir, xyz), key=key)
    assert oracle.interaction_counts(s.arrays()) == expect
