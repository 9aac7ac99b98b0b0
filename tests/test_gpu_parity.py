"""GPU parity of the HIP path against the FP64 oracle (and the reference's
golden vectors), through the C ABI.  Tolerances: FP64 kernels 1e-9 relative
to the field scale; FP32 kernels 2e-5 relative (field RMS) and 1e-4 D on
dipoles against the golden values (the reference's own margin)."""
import json
import os

import numpy as np
import pytest

import oracle
import t9hip

pytestmark = pytest.mark.gpu
DEBYE = 4.803206802
LF_KEY3 = "usolve-cutoff 0.01\npolarizeterm only\n"
LF_KEY4 = LF_KEY3 + "ewald\newald-cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\n"
WATER_KEY = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid {K} {K} {K}\npme-order 5\npolar-eps 1e-6\n"
             "polarization mutual\nneighbor-list\nlist-buffer {buf}\nusolve-cutoff 4.5\nusolve-diag 2.0\nvdwterm none\n")


def rel_err(a, b):
    return np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))


def lf(golden_dir, key):
    return t9hip.System.from_files(os.path.join(golden_dir, "local_frame.xyz"), key=key)


def synthetic(n):
    i = np.arange(n)[:, None]
    j = np.arange(3)[None, :]
    return 0.1 * (i + 1) + 0.03 * (j + 1), 0.1 * (i + 1) - 0.03 * (j + 1)


TOL = {t9hip.T9_FP64: 1e-9, t9hip.T9_FP32: 2e-5}


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
@pytest.mark.parametrize("key", [LF_KEY3, LF_KEY4], ids=["gas", "pme"])
def test_localframe_rpole_fields(golden_dir, key, prec):
    print(f"case localframe key={'pme' if 'ewald' in key else 'gas'} prec={prec}", flush=True)
    s = lf(golden_dir, key)
    o = oracle.Oracle(s.arrays())
    ctx = t9hip.Context(s, prec)
    tol = TOL[prec]
    rp = ctx.rpole()
    assert rel_err(rp, o.rpole()) < tol
    fd, fp = ctx.dfield()
    od, op = o.dfield()
    assert rel_err(fd, od) < tol and rel_err(fp, op) < tol
    ud, up = synthetic(s.n)
    ufd, ufp = ctx.ufield(ud, up)
    oud, oup = o.ufield(ud, up)
    assert rel_err(ufd, oud) < tol and rel_err(ufp, oup) < tol


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
@pytest.mark.parametrize("case,key", [("Local-Frame-3", LF_KEY3), ("Local-Frame-4", LF_KEY4)], ids=["lf3", "lf4"])
def test_localframe_induce_golden(golden_dir, case, key, prec):
    print(f"case {case} induce prec={prec}", flush=True)
    ref = json.load(open(os.path.join(golden_dir, "localframe_ref.json")))
    get = lambda sec, nm: np.array([v for k, v in ref.items() if k.split("/")[0] == case
                                    and k.split("/")[1].startswith(sec) and k.split("/")[2] == nm][0])
    s = lf(golden_dir, key)
    ctx = t9hip.Context(s, prec)
    ud, up, st = ctx.induce()
    assert st.rms_debye < 1e-6
    np.testing.assert_allclose(ud * DEBYE, get("default induce", "ref_ud_debye"), atol=1e-4)
    np.testing.assert_allclose(up * DEBYE, get("default induce", "ref_up_debye"), atol=1e-4)
    o = oracle.Oracle(s.arrays())
    oud, oup, it, rms, rc = o.induce()
    tol = 1e-10 if prec == t9hip.T9_FP64 else 2e-5
    assert np.abs(ud - oud).max() * DEBYE < max(tol, 1e-7 if prec == t9hip.T9_FP64 else 2e-5)
    if prec == t9hip.T9_FP64:
        assert st.iterations == it


def water(nmol, nside, L, K, seed=2024, buf=1.0, n_na=0, n_cl=0):
    return t9hip.System.water_box(nmol, nside, L, seed, n_na, n_cl, key=WATER_KEY.format(K=K, buf=buf))


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_water216_fields_induce(prec):
    """216 waters, 18.643 Å box, K = 24 (the CPU-runnable configuration)."""
    print(f"case water216 prec={prec}", flush=True)
    s = water(216, 6, 18.643, 24, buf=1.5)
    o = oracle.Oracle(s.arrays())
    ctx = t9hip.Context(s, prec)
    tol = TOL[prec]
    fd, fp = ctx.dfield()
    od, op = o.dfield()
    assert rel_err(fd, od) < tol and rel_err(fp, op) < tol
    rng = np.random.default_rng(3)
    ud, up = rng.standard_normal((2, s.n, 3)) * 0.05
    ufd, ufp = ctx.ufield(ud, up)
    oud, oup = o.ufield(ud, up)
    assert rel_err(ufd, oud) < tol and rel_err(ufp, oup) < tol
    gud, gup, st = ctx.induce()
    oud, oup, it, rms, rc = o.induce()
    derr = np.sqrt(np.mean((gud - oud) ** 2)) * DEBYE
    assert derr < (1e-9 if prec == t9hip.T9_FP64 else 2e-5), derr
    print(f"  iters gpu={st.iterations} oracle={it} dipole rms err {derr:.3e} D", flush=True)


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_water_cells_ions(prec):
    """4,000 waters + 20 Na+ + 20 Cl- (cell-list path, 48 Å box, K = 48)."""
    print(f"case water4000+ions prec={prec}", flush=True)
    s = water(4000, 16, 49.32, 48, buf=1.0, n_na=20, n_cl=20)
    o = oracle.Oracle(s.arrays())
    ctx = t9hip.Context(s, prec)
    tol = TOL[prec]
    fd, fp = ctx.dfield()
    od, op = o.dfield()
    assert rel_err(fd, od) < tol and rel_err(fp, op) < tol
    rng = np.random.default_rng(5)
    ud, up = rng.standard_normal((2, s.n, 3)) * 0.05
    ufd, ufp = ctx.ufield(ud, up)
    oud, oup = o.ufield(ud, up)
    assert rel_err(ufd, oud) < tol and rel_err(ufp, oup) < tol


def test_tinkernist_induce_golden(golden_dir, tmp_path):
    """2,684-atom reference box: GPU FP32 dipoles vs test/ref/tinkernist.uind."""
    z = np.load(os.path.join(golden_dir, "tinkernist.npz"))
    n = len(z["types"])
    xyz = z["arc"][0]
    lines = [f"{n} tinkernist", "30.0 30.0 30.0 90.0 90.0 90.0"]
    for a in range(n):
        bonds = " ".join(str(b + 1) for b in z["i12"][a] if b >= 0)
        lines.append(f"{a + 1} X {xyz[a, 0]:.6f} {xyz[a, 1]:.6f} {xyz[a, 2]:.6f} {z['types'][a]} {bonds}")
    p = tmp_path / "tn.xyz"
    p.write_text("\n".join(lines) + "\n")
    key = "a-axis 30.0\nneighbor-list\nvdw-cutoff 12.0\newald\newald-cutoff 7.0\npolar-eps 0.00001\n"
    s = t9hip.System.from_files(str(p), key=key)
    for prec in (t9hip.T9_FP64, t9hip.T9_FP32):
        print(f"case tinkernist prec={prec}", flush=True)
        ctx = t9hip.Context(s, prec)
        ud, up, st = ctx.induce()
        np.testing.assert_allclose(ud * DEBYE, z["uind"][0], atol=5e-5)


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_bitwise_reproducible(prec):
    print(f"case determinism prec={prec}", flush=True)
    s = water(216, 6, 18.643, 24, buf=1.5)
    res = []
    for _ in range(2):
        ctx = t9hip.Context(s, prec)
        ud, up, st = ctx.induce()
        res.append((ud, up, st.iterations))
        ctx.close()
    assert res[0][2] == res[1][2]
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_nblist_complete_and_refresh():
    """A rebuild is triggered only after an atom moves more than half the
    buffer (the pair sets themselves: test_gpu_goldens.py)."""
    print("case nblist refresh", flush=True)
    s = water(4000, 16, 49.32, 48, buf=1.0)
    ctx = t9hip.Context(s, t9hip.T9_FP32)
    xyz = s.f64("xyz", (s.n, 3))
    assert ctx.nblist_refresh()  # a new context has no list yet: the first refresh builds it
    assert not ctx.nblist_refresh()
    xyz2 = xyz.copy()
    xyz2[7, 0] += 0.3  # 0.3 Å < half the 1 Å buffer: keep the list
    ctx.set_coords(xyz2)
    assert not ctx.nblist_refresh()
    xyz2[7, 0] += 0.3  # 0.6 Å: rebuild
    ctx.set_coords(xyz2)
    assert ctx.nblist_refresh()

