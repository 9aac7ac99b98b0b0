"""HIPPO charge-penetration multipoles, polarization, charge transfer,
dispersion and Pauli repulsion on the GPU through the C ABI.

test/emhippo.cpp: pentane + acetamide (104 atoms, HIPPO-2019) in a
35 x 30 x 25 Å box, Ewald at 7 Å (PME) and non-Ewald at 9 Å.  Against the
reference goldens (test/ref/emhippo.{1,2}.txt) with the reference test's
tolerances — FP64: energy 1e-4, gradient 1e-4 (+5e-5, the rounding of the
printed values), virial 1e-3; FP32: energy 1.6e-3 (Ewald) / 5e-3 (non-Ewald),
gradient 3e-4, virial 2e-3 / 2.5e-3 — and against the FP64 oracle (FP64:
1e-10 relative energy, 1e-8 gradient and virial), plus the analyze-mode
interaction count; and test/ephippo.cpp (HIPPO mutual polarization of 50
DMSO molecules, Ewald and non-Ewald) against its goldens and the oracle's
induced dipoles; test/chgtrn.cpp (HIPPO charge transfer of the same DMSO
box) against test/ref/chgtrn.1.txt and the oracle; test/disp.cpp (HIPPO
dispersion with a cutoff) against test/ref/disp.1.txt and the oracle;
test/rephippo.cpp (HIPPO Pauli repulsion of pentane + acetamide) against
test/ref/rephippo.1.txt and the oracle."""
import json
import os

import numpy as np
import pytest

import oracle
import t9hip
from test_oracle_hippo import hippo_system

pytestmark = pytest.mark.gpu

TOL = {(True, t9hip.T9_FP64): (1e-4, 1.5e-4, 1e-3), (False, t9hip.T9_FP64): (1e-4, 1.5e-4, 1e-3),
       (True, t9hip.T9_FP32): (1.6e-3, 3e-4, 2e-3), (False, t9hip.T9_FP32): (5e-3, 3e-4, 2.5e-3)}


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
@pytest.mark.parametrize("ewald", [True, False])
def test_emhippo_gpu(golden_dir, ewald, prec):
    print(f"case emhippo ewald={ewald} prec={prec}", flush=True)
    ref = json.load(open(os.path.join(golden_dir, "hippo_ref.json")))["emhippo_ewald" if ewald else "emhippo_nonewald"]
    s = hippo_system(golden_dir, ewald)
    ctx = t9hip.Context(s, prec)
    e, g, v = ctx.energy_terms(t9hip.TERM_ALL)
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    de, dg, dv = abs(e["esum"] - ref["energy"]), np.abs(g - rg).max(), np.abs(v - rv).max()
    print(f"  E={e['esum']:.6f} ref={ref['energy']} |dE|={de:.2e} max|dg|={dg:.2e} max|dv|={dv:.2e}")
    te, tg, tv = TOL[(ewald, prec)]
    assert de < te and dg < tg and dv < tv
    ea, _, _ = ctx.energy(t9hip.CALC_ENERGY | t9hip.CALC_ANALYZE, grad=False)
    assert ea["nem"] == ref["count"]
    if prec == t9hip.T9_FP64:
        eo, go, vo = oracle.Oracle(s.arrays()).energy(mpole=True, polar=False)
        print(f"  vs oracle: {abs(e['esum'] - eo['em']):.2e} {np.abs(g - go).max():.2e} {np.abs(v - vo).max():.2e}")
        assert abs(e["esum"] - eo["em"]) < 1e-10 * abs(eo["em"])
        assert np.abs(g - go).max() < 1e-8
        assert np.abs(v - vo).max() < 1e-8


TOLP = {t9hip.T9_FP64: (1e-4, 1.5e-4, 1e-3), t9hip.T9_FP32: (1e-4, 1.5e-4, 1e-3)}


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
@pytest.mark.parametrize("ewald", [True, False])
def test_ephippo_gpu(golden_dir, ewald, prec):
    """test/ephippo.cpp (50 DMSO molecules, HIPPO mutual polarization): the
    reference's tolerances (1e-4 energy and gradient, 1e-3 virial, both
    precisions), the interaction count, and the induced dipoles against the
    oracle's (RMS below 1e-5 D: both solves stop at 1e-6 D)."""
    print(f"case ephippo ewald={ewald} prec={prec}", flush=True)
    from test_oracle_hippo import hippo_polar_system
    ref = json.load(open(os.path.join(golden_dir, "hippo_ref.json")))["ephippo_ewald" if ewald else "ephippo_nonewald"]
    s = hippo_polar_system(golden_dir, ewald)
    ctx = t9hip.Context(s, prec)
    ud, up, st = ctx.induce()
    e, g, v = ctx.energy_terms(t9hip.TERM_ALL)
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    de, dg, dv = abs(e["esum"] - ref["energy"]), np.abs(g - rg).max(), np.abs(v - rv).max()
    o = oracle.Oracle(s.arrays())
    oud, oup, it, rms, rc = o.induce()
    du = float(np.sqrt(np.mean((ud - oud) ** 2)) * 4.803206802)
    print(f"  E={e['esum']:.6f} ref={ref['energy']} |dE|={de:.2e} max|dg|={dg:.2e} max|dv|={dv:.2e} "
          f"iters={st.iterations} (oracle {it}) dipole RMS vs oracle {du:.2e} D")
    te, tg, tv = TOLP[prec]
    assert de < te and dg < tg and dv < tv
    assert du < 1e-5
    ea, _, _ = ctx.energy(t9hip.CALC_ENERGY | t9hip.CALC_ANALYZE, grad=False)
    assert ea["nep"] == ref["count"]


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_chgtrn_gpu(golden_dir, prec):
    """test/chgtrn.cpp (HIPPO charge transfer, 50 DMSO molecules, 7 Å): the
    reference's tolerances (energy and gradient 1e-4, virial 5e-4, both
    precisions), the interaction count, and the oracle (FP64: 1e-10 relative
    energy, 1e-9 gradient and virial; FP32 positions are the fixed-point
    records: 1e-7 relative energy, 1e-5 gradient and virial)."""
    print(f"case chgtrn prec={prec}", flush=True)
    from test_oracle_hippo import hippo_chgtrn_system
    ref = json.load(open(os.path.join(golden_dir, "hippo_ref.json")))["chgtrn"]
    s = hippo_chgtrn_system(golden_dir)
    ctx = t9hip.Context(s, prec)
    e, g, v = ctx.energy_terms(t9hip.TERM_ALL)
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    de, dg, dv = abs(e["esum"] - ref["energy"]), np.abs(g - rg).max(), np.abs(v - rv).max()
    eo, go, vo, co = oracle.Oracle(s.arrays()).chgtrn()
    print(f"  E={e['esum']:.6f} ect={e['ect']:.6f} ref={ref['energy']} |dE|={de:.2e} max|dg|={dg:.2e} "
          f"max|dv|={dv:.2e}; vs oracle {abs(e['ect'] - eo):.2e} {np.abs(g - go).max():.2e} {np.abs(v - vo).max():.2e}")
    assert e["ect"] == e["esum"]
    assert de < 1e-4 and dg < 1.5e-4 and dv < 5e-4
    tol = (1e-10, 1e-9) if prec == t9hip.T9_FP64 else (1e-7, 1e-5)
    assert abs(e["ect"] - eo) < tol[0] * abs(eo)
    assert np.abs(g - go).max() < tol[1] and np.abs(v - vo).max() < tol[1] * 10
    ea, _, _ = ctx.energy(t9hip.CALC_ENERGY | t9hip.CALC_ANALYZE, grad=False)
    assert ea["nct"] == ref["count"] == co


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_disp_gpu(golden_dir, prec):
    """test/disp.cpp EDISP-1-NONDEWALD (HIPPO dispersion, pentane +
    acetamide, 5 Å cutoff): the reference's tolerances (energy and gradient
    1e-4, virial 1e-3, both precisions), the interaction count, and the
    oracle (FP64: 1e-10 relative energy, 1e-9 gradient and virial; FP32
    positions are the fixed-point records: 1e-7 relative energy, 1e-5
    gradient and virial)."""
    print(f"case disp prec={prec}", flush=True)
    from test_oracle_hippo import hippo_disp_system
    ref = json.load(open(os.path.join(golden_dir, "hippo_disp_ref.json")))["disp_cutoff"]
    s = hippo_disp_system(golden_dir)
    ctx = t9hip.Context(s, prec)
    e, g, v = ctx.energy_terms(t9hip.TERM_ALL)
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    de, dg, dv = abs(e["esum"] - ref["energy"]), np.abs(g - rg).max(), np.abs(v - rv).max()
    eo, go, vo, co = oracle.Oracle(s.arrays()).disp()
    print(f"  E={e['esum']:.6f} edisp={e['edisp']:.6f} ref={ref['energy']} |dE|={de:.2e} max|dg|={dg:.2e} "
          f"max|dv|={dv:.2e}; vs oracle {abs(e['edisp'] - eo):.2e} {np.abs(g - go).max():.2e} "
          f"{np.abs(v - vo).max():.2e}")
    assert e["edisp"] == e["esum"]
    assert de < 1e-4 and dg < 1.5e-4 and dv < 1e-3
    tol = (1e-10, 1e-9) if prec == t9hip.T9_FP64 else (1e-7, 1e-5)
    assert abs(e["edisp"] - eo) < tol[0] * abs(eo)
    assert np.abs(g - go).max() < tol[1] and np.abs(v - vo).max() < tol[1] * 10
    ea, _, _ = ctx.energy(t9hip.CALC_ENERGY | t9hip.CALC_ANALYZE, grad=False)
    assert ea["ndisp"] == ref["count"] == co


@pytest.mark.parametrize("case", ["rephippo", "rephippo_water"])
@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_repel_gpu(golden_dir, prec, case):
    """test/rephippo.cpp EREPEL-1 (pentane + acetamide) and EREPEL-2 (ten
    waters), HIPPO Pauli repulsion at 7 Å: the reference's tolerances (energy
    and gradient 1e-4, virial 1e-3, both precisions), the interaction count,
    and the oracle (FP64: 1e-10 relative energy, 1e-8 gradient and virial;
    FP32 reads the FP32 rotated multipoles and fixed-point positions: 1e-6
    relative energy, 1e-4 gradient, 1e-3 virial)."""
    print(f"case repel {case} prec={prec}", flush=True)
    from test_oracle_hippo import hippo_rep_system, rep_ref
    ref = rep_ref(golden_dir, case)
    s = hippo_rep_system(golden_dir, case)
    ctx = t9hip.Context(s, prec)
    e, g, v = ctx.energy_terms(t9hip.TERM_ALL)
    rg, rv = np.array(ref["grad"]), np.array(ref["virial"])
    de, dg, dv = abs(e["esum"] - ref["energy"]), np.abs(g - rg).max(), np.abs(v - rv).max()
    eo, go, vo, co = oracle.Oracle(s.arrays()).repel()
    print(f"  E={e['esum']:.6f} erep={e['erep']:.6f} ref={ref['energy']} |dE|={de:.2e} max|dg|={dg:.2e} "
          f"max|dv|={dv:.2e}; vs oracle {abs(e['erep'] - eo):.2e} {np.abs(g - go).max():.2e} "
          f"{np.abs(v - vo).max():.2e}")
    assert e["erep"] == e["esum"]
    assert de < 1e-4 and dg < 1.5e-4 and dv < 1e-3
    tol = (1e-10, 1e-8, 1e-8) if prec == t9hip.T9_FP64 else (1e-6, 1e-4, 1e-3)
    assert abs(e["erep"] - eo) < tol[0] * abs(eo)
    assert np.abs(g - go).max() < tol[1] and np.abs(v - vo).max() < tol[2]
    ea, _, _ = ctx.energy(t9hip.CALC_ENERGY | t9hip.CALC_ANALYZE, grad=False)
    assert ea["nrep"] == ref["count"] == co


HIPPO_ALL_KEY = ("neighbor-list\na-axis 25.0\nb-axis 25.0\nc-axis 30.0\newald\newald-cutoff 7.0\nchgtrn-cutoff 6.0\n"
                 "repuls-cutoff 6.0\ndisp-cutoff 7.0\nvdwterm none\nbondterm none\nangleterm none\n"
                 "ureybradterm none\nstrbndterm none\nopbendterm none\nopdistterm none\ntorsionterm none\n"
                 "pitorsterm none\ntortorterm none\nangtorterm none\nstrtorterm none\nimproperterm none\n"
                 "imptorsterm none\n")


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_hippo_all_terms_gpu(golden_dir, prec):
    """Every HIPPO nonbonded term at once on the 50-DMSO box (Ewald
    multipoles and polarization, Pauli repulsion, dispersion with a cutoff,
    charge transfer): the energy, gradient and virial of one call equal the
    sums of the calls that select one term at a time (the accumulation order
    of the gradient and torque arrays across the term kernels), and the sum
    of the oracle's terms (FP64: 1e-6 relative energy, 2e-5 gradient, 1e-4
    virial, the solve stopping at 1e-6 D; FP32: 1e-5, 1e-4, 1e-3)."""
    print(f"case hippo all terms prec={prec}", flush=True)
    s = t9hip.System.from_files(os.path.join(golden_dir, "dmso.xyz"),
                                prm=os.path.join(golden_dir, "hippo19_subset.prm"), key=HIPPO_ALL_KEY)
    a = s.arrays()
    assert a["chgpen"] and a["repel"] and a["disp"] and a["chgtrn"]
    ctx = t9hip.Context(s, prec)
    e, g, v = ctx.energy_terms(t9hip.TERM_ALL)
    parts = {t: ctx.energy_terms(t) for t in (t9hip.TERM_ELEC, t9hip.TERM_REPEL, t9hip.TERM_DISP,
                                               t9hip.TERM_CHGTRN)}
    es = sum(p[0]["esum"] for p in parts.values())
    gs = sum(p[1] for p in parts.values())
    vs = sum(p[2] for p in parts.values())
    print(f"  E={e['esum']:.6f} (em {e['em']:.4f} ep {e['ep']:.4f} erep {e['erep']:.4f} edisp {e['edisp']:.4f} "
          f"ect {e['ect']:.4f}) sum of parts {es:.6f} max|dg| {np.abs(g - gs).max():.2e} "
          f"max|dv| {np.abs(v - vs).max():.2e}")
    assert abs(e["erep"] - parts[t9hip.TERM_REPEL][0]["erep"]) < 1e-10 * abs(e["erep"])
    assert abs(e["edisp"] - parts[t9hip.TERM_DISP][0]["edisp"]) < 1e-10 * abs(e["edisp"])
    assert abs(e["ect"] - parts[t9hip.TERM_CHGTRN][0]["ect"]) < 1e-10 * abs(e["ect"])
    assert abs(e["esum"] - es) < 1e-7 * abs(es)
    assert np.abs(g - gs).max() < 1e-5 and np.abs(v - vs).max() < 1e-4
    o = oracle.Oracle(a)
    ud, up, _, _, _ = o.induce()
    eo, go, vo = o.energy(ud, up, mpole=True, polar=True)
    er, gr, vr, _ = o.repel()
    ed, gd, vd, _ = o.disp()
    ec, gc, vc, _ = o.chgtrn()
    etot = eo["em"] + eo["ep"] + er + ed + ec
    gtot, vtot = go + gr + gd + gc, vo + vr + vd + vc
    print(f"  vs oracle: {abs(e['esum'] - etot):.2e} {np.abs(g - gtot).max():.2e} {np.abs(v - vtot).max():.2e}")
    tol = (1e-6, 2e-5, 1e-4) if prec == t9hip.T9_FP64 else (1e-5, 1e-4, 1e-3)
    assert abs(e["esum"] - etot) < tol[0] * abs(etot)
    assert np.abs(g - gtot).max() < tol[1] and np.abs(v - vtot).max() < tol[2]


HIPPO_WATER_MD_KEY = ("a-axis 25.0\newald\newald-cutoff 7.0\nrepuls-cutoff 6.0\ndisp-cutoff 7.0\nchgtrn-cutoff 6.0\n"
                      "polar-eps 1e-8\nvdwterm none\n")


def test_hippo_water_nve_gpu(golden_dir):
    """HIPPO water in the MD loop (FP64): ten waters of the HIPPO water set
    with every term on (valence, Ewald multipoles and polarization, Pauli
    repulsion, dispersion, charge transfer), RESPA with two valence steps per
    step, no thermostat, the same 100 fs from the same velocities at 1 fs and
    at 0.5 fs.  The largest deviation of the total energy from its start falls
    with the step as a second-order integrator's does (by more than 2.5x for
    a halved step; forces that were not the gradient of the energy would leave
    a deviation that does not shrink), and stays below 0.2 kcal/mol."""
    print("case hippo water nve", flush=True)
    s = t9hip.System.from_files(os.path.join(golden_dir, "h2o10.xyz"),
                                prm=os.path.join(golden_dir, "water19_records.prm"), key=HIPPO_WATER_MD_KEY)
    a = s.arrays()
    assert a["chgpen"] and a["repel"] and a["disp"] and a["chgtrn"]
    dev = {}
    for dt, nchunk in ((0.001, 10), (0.0005, 20)):
        ctx = t9hip.Context(s, t9hip.T9_FP64)
        ctx.md_init(298.0, 11)
        etot = []
        for _ in range(nchunk):
            st = ctx.md_run(10, dt, 2, 0.0)
            etot.append(st["epot"] + st["ekin"])
        etot = np.array(etot)
        dev[dt] = np.abs(etot - etot[0]).max()
        print(f"  dt={dt * 1000:.1f} fs: E_total first {etot[0]:.5f} last {etot[-1]:.5f} "
              f"max deviation {dev[dt]:.2e} T {st['temperature']:.1f} K")
    assert dev[0.001] < 0.2
    assert dev[0.0005] < dev[0.001] / 2.5
