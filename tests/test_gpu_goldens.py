"""GPU runs of the reference's own fixtures through the C ABI: polpair (a
pair-specific Thole value, so the non-uniform Thole kernels run; Ewald and
non-Ewald), NaCl-2 / NaCl-3 multipoles, Local-Frame3-1 (non-Ewald periodic
cutoff), the Urey-Bradley and bond goldens, the argon-box neighbor sets
(test/nblist.hh) and exact pair sets of the multipole list against an FP64
brute force, and the Bussi rescaling against its closed form.

Tolerances: the reference's own (energy 1e-4, gradient 1e-4 plus the 5e-5
rounding of the printed values, virial 1e-3; NaCl-3 5e-4) in both precisions;
neighbor sets bit-exact (FP64 lists; FP32 lists exact except pairs within
1e-4 Å of the list radius)."""
import json
import os

import numpy as np
import pytest

import t9hip
from test_oracle_pins import (ARBOX_KEY, LF3_KEY, NACL2_KEY, PP_KEY, _cmp, more_ref, nacl3_key)

pytestmark = pytest.mark.gpu
PRECS = [t9hip.T9_FP64, t9hip.T9_FP32]


def _energy(s, prec, terms=t9hip.TERM_ALL):
    ctx = t9hip.Context(s, prec)
    e, g, v = ctx.energy_terms(terms)
    return e, g, v


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("case,extra", [("polpair_ewald", "ewald\n"), ("polpair_nonewald", "")])
def test_polpair_gpu(golden_dir, case, extra, prec):
    print(f"case {case} prec={prec}", flush=True)
    s = t9hip.System.from_files(os.path.join(golden_dir, "polpair_nacl.xyz"), key=PP_KEY + extra)
    e, g, v = _energy(s, prec)
    _cmp(e["esum"], g, v, more_ref(golden_dir)[case], label=f"{case} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_nacl2_gpu(golden_dir, prec):
    print(f"case NaCl-2 prec={prec}", flush=True)
    s = t9hip.System.from_files(os.path.join(golden_dir, "nacl1.xyz"), key=NACL2_KEY)
    e, g, v = _energy(s, prec)
    _cmp(e["esum"], g, v, more_ref(golden_dir)["NaCl-2"], label=f"NaCl-2 prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_nacl3_gpu(golden_dir, prec):
    print(f"case NaCl-3 prec={prec}", flush=True)
    s = t9hip.System.from_files(os.path.join(golden_dir, "nacl4.xyz"), key=nacl3_key(golden_dir))
    e, g, v = _energy(s, prec)
    _cmp(e["esum"], g, v, more_ref(golden_dir)["NaCl-3"], eps_e=5e-4, eps_g=5e-4, eps_v=5e-4,
         label=f"NaCl-3 prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_localframe3_nonewald_gpu(golden_dir, prec):
    print(f"case Local-Frame3-1 prec={prec}", flush=True)
    ref = json.load(open(os.path.join(golden_dir, "localframe_ref.json")))
    get = lambda nm: np.array([v for k, v in ref.items() if k.startswith("Local-Frame3-1/") and k.endswith("/" + nm)][0])
    s = t9hip.System.from_files(os.path.join(golden_dir, "local_frame2.xyz"), key=LF3_KEY)
    e, g, v = _energy(s, prec, t9hip.TERM_ALL)
    _cmp(e["esum"], g, v, {"energy": float(get("ref_eng")), "grad": get("ref_g"), "virial": get("ref_v")},
         label=f"Local-Frame3-1 prec={prec}")
    ctx = t9hip.Context(s, prec)
    ea, _, _ = ctx.energy(t9hip.CALC_ENERGY | t9hip.CALC_ANALYZE, grad=False)
    assert ea["nem"] == 120 and ea["nep"] == 120  # test/localframe3.cpp:35-37


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("case,xyz,prm,key", [
    ("bond_trpcage", "trpcage.xyz", "amoebapro13_bond_subset.prm", "bondterm only\n"),
    ("urey_h2o10", "h2o10.xyz", "water03_subset.prm", "ureybradterm only\n"),
])
def test_valence_goldens_gpu(golden_dir, case, xyz, prm, key, prec):
    print(f"case {case} prec={prec}", flush=True)
    s = t9hip.System.from_files(os.path.join(golden_dir, xyz), prm=os.path.join(golden_dir, prm), key=key)
    e, g, v = _energy(s, prec)
    _cmp(e["esum"], g, v, more_ref(golden_dir)[case], label=f"{case} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_arbox_vdw_list_superset(golden_dir, prec):
    """test/nblist.cpp:145-193: after a refresh on each of five frames the
    vdW list holds every reference neighbor (k > i within 9 Å); on frame 1
    (a fresh build) it is exactly the FP64 brute-force set within the
    10 Å list radius."""
    print(f"case arbox nblist prec={prec}", flush=True)
    z = np.load(os.path.join(golden_dir, "arbox_nblist.npz"))
    xyz, ans = z["xyz"], z["ans"]
    L = float(z["box"][0][0])
    s = t9hip.System.from_files(os.path.join(golden_dir, "arbox_frame1.xyz"), key=ARBOX_KEY)
    ctx = t9hip.Context(s, prec)
    for f in range(len(xyz)):
        ctx.set_coords(xyz[f])
        ctx.nblist_refresh()
        pairs = ctx.nblist_pairs(1)
        have = {(int(i), int(k)) for i, k in pairs}
        need = {(i, int(k)) for i in range(216) for k in ans[f, i] if k >= 0}
        missing = need - have
        assert not missing, (f, sorted(missing)[:10])
        if f == 0:
            d = xyz[0][:, None, :] - xyz[0][None, :, :]
            d -= L * np.round(d / L)
            r = np.sqrt(np.sum(d * d, axis=-1))
            iu = np.triu_indices(216, 1)
            brute = {(int(i), int(k)) for i, k in zip(*iu) if r[i, k] < 10.0}
            diff = brute ^ have
            amb = {p for p in diff if abs(r[p] - 10.0) < 1e-4}
            assert diff == (amb if prec == t9hip.T9_FP32 else set()), sorted(diff)[:10]


def _brute_pairs(xyz, L, rc):
    n = len(xyz)
    out = set()
    near = []
    for a in range(0, n, 1024):
        d = xyz[a:a + 1024, None, :] - xyz[None, :, :]
        d -= L * np.round(d / L)
        r = np.sqrt(np.sum(d * d, axis=-1))
        ii, kk = np.nonzero(r < rc)
        for i, k in zip(ii + a, kk):
            if i < k:
                out.add((int(i), int(k)))
        ii, kk = np.nonzero(np.abs(r - rc) < 1e-4)
        near.extend((int(i + a), int(k)) for i, k in zip(ii, kk) if i + a < k)
    return out, set(near)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("nmol,nside,L,nion,buf", [(4000, 16, 49.32, 20, 1.0), (343, 7, 21.7, 0, 1.0)],
                         ids=["cells_ions", "five_cells"])
def test_multipole_list_exact(prec, nmol, nside, L, nion, buf):
    """The stored multipole list against the FP64 minimum-image brute force
    at the list radius (7 Å + buffer): the same pair set.  4,000 waters +
    ions take the cell search with non-compact per-home-cell walks; the
    21.7 Å box has exactly 5 cells per axis."""
    print(f"case multipole list {nmol} waters prec={prec}", flush=True)
    key = (f"ewald\newald-cutoff 7.0\nvdw-cutoff 7.0\nneighbor-list\nlist-buffer {buf}\nvdwterm none\n"
           "bondterm none\n")
    s = t9hip.System.water_box(nmol, nside, L, 2024, nion, nion, key=key)
    ctx = t9hip.Context(s, prec)
    have = {(int(i), int(k)) for i, k in ctx.nblist_pairs(0)}
    brute, near = _brute_pairs(s.f64("xyz", (s.n, 3)), L, 7.0 + buf)
    diff = have ^ brute
    print(f"  pairs {len(have)} brute {len(brute)} differ {len(diff)} near the radius {len(near)}", flush=True)
    if prec == t9hip.T9_FP64:
        assert not diff, sorted(diff)[:10]
    else:
        assert diff <= near, sorted(diff - near)[:10]


def test_bussi_closed_form():
    """One rescaling with given deviates: the kinetic energy goes from K to
    K' = K [c + (1 - c) (Kb / (nf K)) (r^2 + s2) + 2 r sqrt(c (1 - c) Kb / (nf K))]
    (Bussi, Donadio, Parrinello, J. Chem. Phys. 126, 014101 (2007), eq. A7;
    Kb = nf kT / 2), velocities reversed when r + sqrt(c nf K / ((1 - c) Kb)) < 0;
    tau -> infinity keeps K, tau -> 0 draws K' = Kb (r^2 + s2) / nf."""
    print("case bussi", flush=True)
    s = t9hip.System.water_box(216, 6, 18.643, 3, key="ewald\newald-cutoff 7.0\nvdw-cutoff 7.0\nneighbor-list\nlist-buffer 1.0\n")
    ctx = t9hip.Context(s, t9hip.T9_FP64)
    ctx.md_init(250.0, 9)
    m = s.f64("mass", s.n)
    nf = 3 * s.n - 3
    R = 1.9872042586e-3
    for dt, tau, kelvin, r, s2 in [(0.002, 0.1, 298.0, 0.7, nf - 1 + 3.0), (0.002, 1e9, 298.0, -1.2, nf + 10.0),
                                  (0.002, 1e-6, 310.0, 0.3, nf - 5.0), (0.001, 0.05, 298.0, -2.0, nf - 1.0)]:
        v0 = ctx.velocities()
        k0 = 0.5 * np.sum(m[:, None] * v0 * v0) / 418.4
        ctx.md_bussi(dt, tau, kelvin, r, s2)
        v1 = ctx.velocities()
        k1 = 0.5 * np.sum(m[:, None] * v1 * v1) / 418.4
        c = np.exp(-dt / tau)
        kb = 0.5 * nf * R * kelvin
        expect = k0 * (c + (1 - c) * kb / (nf * k0) * (r * r + s2) + 2 * r * np.sqrt(c * (1 - c) * kb / (nf * k0)))
        print(f"  K {k0:.4f} -> {k1:.4f} expected {expect:.4f}", flush=True)
        assert abs(k1 - expect) < 1e-9 * expect
        sign = np.sign(np.sum(v1 * v0))
        assert sign == (-1 if r + np.sqrt(c * nf * k0 / ((1 - c) * kb)) < 0 else 1)
