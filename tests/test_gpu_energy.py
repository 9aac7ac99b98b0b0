"""GPU parity of the multipole + polarization energy, gradient and virial
(t9_energy through the C ABI) against the FP64 oracle and the reference's
golden values (test/localframe.cpp, test/localframe3.cpp).

Tolerances: FP64 kernels 1e-9 relative; FP32 kernels 1e-6 relative on the
total energy and 1e-5 kcal/mol/Å RMS on the gradient against the FP64
oracle (the accuracy bound stated for the mixed-precision form), virial
1e-4 relative to its largest component."""
import json
import os

import numpy as np
import pytest

import oracle
import t9hip

pytestmark = pytest.mark.gpu

K_LF1 = "multipoleterm only\n"
K_LF2 = K_LF1 + "ewald\newald-cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\n"
K_LF3 = "usolve-cutoff 0.01\npolarizeterm only\n"
K_LF4 = K_LF3 + "ewald\newald-cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\n"
# the reference's key for Local-Frame3 (test/localframe3.cpp:9-23): every term but the electrostatics off
K_LF3X = ("cutoff 7.0\nneighbor-list\nlist-buffer 0.1\na-axis 20.0\nbondterm none\nangleterm none\n"
          "strbndterm none\nureybradterm none\ntorsionterm none\nvdwterm none\n")
K_LF32 = K_LF3X + "ewald\n"
WATER_KEY = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid {K} {K} {K}\npme-order 5\npolar-eps 1e-6\n"
             "polarization mutual\nneighbor-list\nlist-buffer {buf}\nusolve-cutoff 4.5\nusolve-diag 2.0\n"
             "vdwterm none\nbondterm none\nangleterm none\nureybradterm none\n")

CASES = {
    "lf1": ("local_frame.xyz", K_LF1, True, False),
    "lf2": ("local_frame.xyz", K_LF2, True, False),
    "lf3": ("local_frame.xyz", K_LF3, False, True),
    "lf4": ("local_frame.xyz", K_LF4, False, True),
    "lf32": ("local_frame2.xyz", K_LF32, True, True),
}


def _oracle_energy(s, mpole, polar):
    o = oracle.Oracle(s.arrays())
    ud = up = None
    if polar:
        ud, up, it, rms, rc = o.induce()
        assert rc == 0
    e, g, v = o.energy(ud, up, mpole=mpole, polar=polar)
    et = (e["em"] if mpole else 0.0) + (e["ep"] if polar else 0.0)
    return et, e, g, v


def _check(prec, eg, gg, vg, eo, go, vo, label):
    et = eg["esum"]
    rel_e = abs(et - eo) / max(1.0, abs(eo))
    g_rms = np.sqrt(np.mean((gg - go) ** 2))
    g_scale = np.sqrt(np.mean(go ** 2))
    v_err = np.abs(vg - vo).max() / max(1.0, np.abs(vo).max())
    print(f"{label} prec={prec} E={et:.10f} oracle={eo:.10f} relE={rel_e:.2e} gradRMS={g_rms:.2e} "
          f"(scale {g_scale:.2f}) vir={v_err:.2e}", flush=True)
    if prec == t9hip.T9_FP64:
        assert rel_e < 1e-9
        assert g_rms < 1e-9 * max(1.0, g_scale)
        assert v_err < 1e-9
    else:
        assert rel_e < 1e-6
        assert g_rms < 1e-5
        assert v_err < 1e-4


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
@pytest.mark.parametrize("case", sorted(CASES))
def test_localframe_energy(golden_dir, case, prec):
    xyz, key, mpole, polar = CASES[case]
    print(f"case energy {case} prec={prec}", flush=True)
    s = t9hip.System.from_files(os.path.join(golden_dir, xyz), key=key)
    ctx = t9hip.Context(s, prec)
    eg, gg, vg = ctx.energy()
    eo, _, go, vo = _oracle_energy(s, mpole, polar)
    _check(prec, eg, gg, vg, eo, go, vo, case)


def test_localframe3_emplar_golden(golden_dir):
    """The reference's own values for the fused Ewald case, every frame type
    (test/localframe3.cpp:106-163): energy 1e-4, gradient 1e-4, virial 1e-3."""
    ref = json.load(open(os.path.join(golden_dir, "localframe_ref.json")))
    pre = "Local-Frame3-2/- emplar -- ewald pbc/"
    s = t9hip.System.from_files(os.path.join(golden_dir, "local_frame2.xyz"), key=K_LF32)
    for prec in (t9hip.T9_FP64, t9hip.T9_FP32):
        ctx = t9hip.Context(s, prec)
        e, g, v = ctx.energy()
        assert abs(e["esum"] - ref[pre + "ref_eng"]) < 1e-4
        np.testing.assert_allclose(g, np.array(ref[pre + "ref_g"]), atol=1e-4)
        np.testing.assert_allclose(v, np.array(ref[pre + "ref_v"]), atol=1e-3)


def water(nmol, nside, L, K, seed=2024, buf=1.0, n_na=0, n_cl=0):
    return t9hip.System.water_box(nmol, nside, L, seed, n_na, n_cl, key=WATER_KEY.format(K=K, buf=buf))


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_water216_energy(prec):
    """216 waters, 18.643 Å box, PME 24^3: the CPU-runnable configuration's
    single-point energy + gradient + virial."""
    print(f"case energy water216 prec={prec}", flush=True)
    s = water(216, 6, 18.643, 24, buf=1.5)
    ctx = t9hip.Context(s, prec)
    eg, gg, vg = ctx.energy()
    eo, _, go, vo = _oracle_energy(s, True, True)
    _check(prec, eg, gg, vg, eo, go, vo, "water216")


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_water_ions_energy(prec):
    """Ions + waters on the cell-list path (1,000 waters, 20 Na+, 20 Cl-)."""
    print(f"case energy water1000+ions prec={prec}", flush=True)
    s = water(1000, 11, 32.6, 40, seed=7, buf=1.0, n_na=20, n_cl=20)
    ctx = t9hip.Context(s, prec)
    eg, gg, vg = ctx.energy()
    eo, _, go, vo = _oracle_energy(s, True, True)
    _check(prec, eg, gg, vg, eo, go, vo, "water1000+ions")


def test_water96k_fp32_vs_fp64():
    """The 96,000-atom box (PME 120^3: whole-plane FP32 FFT kernels, cell
    lists, 8-wave pair kernels) at the same coordinates in FP32 and FP64: the
    FP64 form is pinned to the oracle above, so agreement to the FP32 bounds
    checks the large-system FP32 path end to end."""
    print("case energy water96k fp32 vs fp64", flush=True)
    key = WATER_KEY.format(K=120, buf=2.0).replace("polar-eps 1e-6", "polar-eps 1e-8")
    s = t9hip.System.water_box(32000, 32, 98.65, 2024, 0, 0, key=key)
    e64, g64, v64 = t9hip.Context(s, t9hip.T9_FP64).energy()
    e32, g32, v32 = t9hip.Context(s, t9hip.T9_FP32).energy()
    rel = abs(e32["esum"] - e64["esum"]) / abs(e64["esum"])
    grms = np.sqrt(np.mean((g32 - g64) ** 2))
    vrel = np.abs(v32 - v64).max() / np.abs(v64).max()
    print(f"water96k E64={e64['esum']:.6f} relE={rel:.2e} gradRMS={grms:.2e} vir={vrel:.2e}", flush=True)
    assert rel < 1e-6 and grms < 1e-5 and vrel < 1e-4


@pytest.mark.parametrize("case", ["water216", "water216-k32", "water216-k48", "water216-k64", "water216-k80",
                                  "water216-k96", "water216-k128", "water1000+ions", "water96k"])
def test_fft_plane_matches_lines(case, monkeypatch):
    """FP32 x and y transforms per z-plane in LDS against the line passes
    (T9HIP_FFT=lines): every butterfly is the same arithmetic, so energy,
    gradient and virial are the same bits.  Every plane length with a kernel:
    24 (radices 8, 3), 32 (8, 4), 40 (8, 5), 48 (8, 2, 3), 64 (8, 8), 80
    (8, 2, 5), 96 (8, 4, 3), 120 (8, 3, 5), 128 (8, 8, 2)."""
    print(f"case fft plane vs lines {case}", flush=True)
    if case == "water216":
        s = water(216, 6, 18.643, 24, buf=1.5)
    elif case.startswith("water216-k"):
        s = water(216, 6, 18.643, int(case.split("-k")[1]), buf=1.5)
    elif case == "water1000+ions":
        s = water(1000, 11, 32.6, 40, seed=7, buf=1.0, n_na=20, n_cl=20)
    else:
        s = t9hip.System.water_box(32000, 32, 98.65, 2024, 0, 0, key=WATER_KEY.format(K=120, buf=2.0))
    monkeypatch.delenv("T9HIP_FFT", raising=False)
    ep, gp, vp = t9hip.Context(s, t9hip.T9_FP32).energy()
    monkeypatch.setenv("T9HIP_FFT", "lines")
    el, gl, vl = t9hip.Context(s, t9hip.T9_FP32).energy()
    print(f"{case}: plane E={ep['esum']:.10f} lines E={el['esum']:.10f} max|dg|={np.abs(gp - gl).max():.2e}",
          flush=True)
    assert ep["esum"] == el["esum"] and ep["iterations"] == el["iterations"]
    assert np.array_equal(gp, gl) and np.array_equal(vp, vl)


def test_energy_bitwise_reproducible():
    s = water(216, 6, 18.643, 24, buf=1.5)
    ctx = t9hip.Context(s, t9hip.T9_FP32)
    e1, g1, v1 = ctx.energy()
    e2, g2, v2 = ctx.energy()
    assert e1["esum"] == e2["esum"]
    assert np.array_equal(g1, g2) and np.array_equal(v1, v2)


def test_energy_gradient_finite_difference():
    """FP64 GPU gradient against central differences of the GPU energy."""
    s = water(216, 6, 18.643, 24, buf=1.5)
    ctx = t9hip.Context(s, t9hip.T9_FP64)
    x0 = s.arrays()["xyz"].copy()
    _, g, _ = ctx.energy()
    h = 1e-5
    for atom in (0, 1, 2, 100):
        for d in range(3):
            es = []
            for sg in (1, -1):
                x = x0.copy()
                x[atom, d] += sg * h
                ctx.set_coords(x)
                es.append(ctx.energy(vers=t9hip.CALC_ENERGY)[0]["esum"])
            fd = (es[0] - es[1]) / (2 * h)
            assert abs(fd - g[atom, d]) < 1e-4 * max(1.0, abs(g[atom, d])), (atom, d, fd, g[atom, d])
    ctx.set_coords(x0)


COUNT_GOLDEN = {"lf1": (201, 0), "lf2": (222, 0), "lf3": (0, 201), "lf4": (0, 222), "lf32": (138, 138)}


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
@pytest.mark.parametrize("case", sorted(COUNT_GOLDEN))
def test_interaction_counts(golden_dir, case, prec):
    """energy(calc::v3): the reference's interaction counts nem / nep
    (test/localframe.cpp, test/localframe3.cpp), energy unchanged by the
    analyze bit."""
    xyz, key, mpole, polar = CASES[case]
    s = t9hip.System.from_files(os.path.join(golden_dir, xyz), key=key)
    ctx = t9hip.Context(s, prec)
    e, _, _ = ctx.energy(t9hip.CALC_ENERGY | t9hip.CALC_ANALYZE, grad=False)
    print(f"counts {case} prec={prec}: nem {e['nem']} nep {e['nep']}", flush=True)
    assert (e["nem"], e["nep"]) == COUNT_GOLDEN[case]
    e0, _, _ = ctx.energy(t9hip.CALC_ENERGY, grad=False)
    assert e0["esum"] == e["esum"]


def test_interaction_counts_water_box():
    # FP64: an FP32 distance within rounding of the 7 A cutoff could flip a pair
    s = t9hip.System.water_box(216, 6, 18.643, 2024, key=WATER_KEY.format(K=24, buf=1.0))
    ctx = t9hip.Context(s, t9hip.T9_FP64)
    e, _, _ = ctx.energy(t9hip.CALC_ENERGY | t9hip.CALC_ANALYZE, grad=False)
    assert (e["nem"], e["nep"]) == oracle.interaction_counts(s.arrays())
