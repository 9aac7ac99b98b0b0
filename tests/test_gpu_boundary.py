"""GPU tests of the boundary's sub-step entry points (through the C ABI)
against the FP64 oracle: the PME sub-steps (spreading, one-way transforms,
influence function, gathers), the diagonal and sparse preconditioners, the
energy term selection, device-resident coordinates and gradient, the
context-free 3D FFT against torch.fft (rocFFT), and the non-convergence error
path of the PCG solve.

Tolerances: FP64 kernels 1e-10 relative to the array's RMS; FP32 kernels
2e-5 (grids, fields) and 1e-5 (FFT against rocFFT in FP32)."""
import numpy as np
import pytest

import oracle
import t9hip

pytestmark = pytest.mark.gpu

KEY = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid 24 24 24\npme-order 5\npolar-eps 1e-6\n"
       "polarization mutual\nneighbor-list\nlist-buffer 1.5\nusolve-cutoff 4.5\nusolve-diag 2.0\n")
FIELDS = "vdwterm none\n"
TOL = {t9hip.T9_FP64: 1e-10, t9hip.T9_FP32: 2e-5}


def rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def water216(extra=FIELDS):
    return t9hip.System.water_box(216, 6, 18.643, 2024, key=KEY + extra)


def cgrid(g):
    return g[..., 0] + 1j * g[..., 1]


def rgrid(c):
    return np.stack([c.real, c.imag], axis=-1)


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_pme_substeps(prec):
    print(f"case pme sub-steps prec={prec}", flush=True)
    s = water216()
    o = oracle.Oracle(s.arrays())
    ctx = t9hip.Context(s, prec)
    tol = TOL[prec]
    gm = ctx.pme_grid_mpole()
    om = o.pme_grid_mpole()
    assert rel(gm, om) < tol and np.abs(gm[..., 1]).max() == 0.0
    rng = np.random.default_rng(4)
    ud, up = rng.standard_normal((2, s.n, 3)) * 0.05
    gu = ctx.pme_grid_uind(ud, up)
    ou = o.pme_grid_uind(ud, up)
    assert rel(gu, ou) < tol
    # each step on the oracle's own input
    f = ctx.pme_fft_front(ou)
    of = rgrid(o.dft3(cgrid(ou), -1))
    assert rel(f, of) < tol
    c = ctx.pme_conv(of)
    oc = o.pme_conv(of)
    assert rel(c, oc) < tol
    b = ctx.pme_fft_back(oc)
    ob = rgrid(o.dft3(cgrid(oc), +1))
    assert rel(b, ob) < tol
    fd, fp = ctx.pme_fphi_uind(ob)
    assert rel(fd, o.pme_field(ob, imag=False)) < tol and rel(fp, o.pme_field(ob, imag=True)) < tol
    fm = ctx.pme_fphi_mpole(ob)
    assert rel(fm, o.pme_field(ob, imag=False)) < tol
    # the chain equals the reciprocal-space dipole field of the round trip
    rd, rp = o.recip_dipole(ud, up)
    g = ctx.pme_grid_uind(ud, up)
    g = ctx.pme_fft_back(ctx.pme_conv(ctx.pme_fft_front(g)))
    fd, fp = ctx.pme_fphi_uind(g)
    print(f"  chain rel err d {rel(fd, rd):.2e} p {rel(fp, rp):.2e}", flush=True)
    assert rel(fd, rd) < 5 * tol and rel(fp, rp) < 5 * tol


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_spread_kernels(prec, monkeypatch):
    """Both spreading kernels against the oracle: the row-owner kernel (grid
    sizes divisible by 8: 1,000 waters + 40 ions at K = 40) and the LDS
    fixed-point kernel (the same box with T9HIP_SPREAD=fixed, and at K = 30,
    which the row-owner kernel does not take); multipole and dipole grids."""
    key = KEY.replace("pme-grid 24 24 24", "pme-grid {K} {K} {K}") + FIELDS
    for K, mode in ((40, "rows"), (40, "fixed"), (30, "rows")):
        print(f"case spread K={K} {mode} prec={prec}", flush=True)
        monkeypatch.setenv("T9HIP_SPREAD", mode)
        s = t9hip.System.water_box(1000, 11, 32.6, 7, 20, 20, key=key.format(K=K))
        o = oracle.Oracle(s.arrays())
        ctx = t9hip.Context(s, prec)
        rng = np.random.default_rng(K)
        ud, up = rng.standard_normal((2, s.n, 3)) * 0.05
        em, eu = rel(ctx.pme_grid_mpole(), o.pme_grid_mpole()), rel(ctx.pme_grid_uind(ud, up), o.pme_grid_uind(ud, up))
        print(f"  multipole grid rel err {em:.2e}, dipole grid {eu:.2e}", flush=True)
        assert em < TOL[prec] and eu < TOL[prec]
        ctx.close()


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_preconditioners(prec):
    """z = M r of the sparse preconditioner (usolve 4.5 Å, udiag 2) and of the
    diagonal one against the oracle (the oracle's M is pinned to the
    Neumann expansion of (alpha^-1 - T)^-1 on the CPU)."""
    print(f"case preconditioners prec={prec}", flush=True)
    s = water216()
    o = oracle.Oracle(s.arrays())
    ctx = t9hip.Context(s, prec)
    rng = np.random.default_rng(5)
    rd, rp = rng.standard_normal((2, s.n, 3))
    zd, zp = ctx.sparse_precond(rd, rp)
    od, op = o.precond(rd, rp, sparse=True)
    print(f"  sparse rel err {rel(zd, od):.2e} {rel(zp, op):.2e}", flush=True)
    assert rel(zd, od) < TOL[prec] and rel(zp, op) < TOL[prec]
    zd, zp = ctx.diag_precond(rd, rp)
    od, op = o.precond(rd, rp, sparse=False)
    assert rel(zd, od) < TOL[prec] and rel(zp, op) < TOL[prec]
    s2 = water216(FIELDS + "pcg-noprecond\n")
    with pytest.raises(t9hip.T9Error) as ei:
        t9hip.Context(s2, prec).sparse_precond(rd, rp)
    assert ei.value.code == t9hip.T9_ERR_ARG


def test_energy_terms_no_double_count():
    """With vdW and valence on, the electrostatics-only call returns em + ep
    and the electrostatics gradient, the other groups theirs; the parts add
    up to the all-terms call."""
    print("case energy term selection", flush=True)
    s = water216("vdw-cutoff 7.0\n")
    o = oracle.Oracle(s.arrays())
    ctx = t9hip.Context(s, t9hip.T9_FP64)
    ea, ga, va = ctx.energy()
    ee, ge, ve = ctx.energy_terms(t9hip.TERM_ELEC)
    ev, gv, vv = ctx.energy_terms(t9hip.TERM_VDW)
    ek, gk, vk = ctx.energy_terms(t9hip.TERM_VALENCE)
    ud, up, it, rms, rc = o.induce()
    eo, go, vo = o.energy(ud, up)
    assert abs(ee["esum"] - (eo["em"] + eo["ep"])) < 1e-9 * abs(eo["em"] + eo["ep"])
    assert ee["evdw"] == 0 and ee["ebond"] == 0 and np.abs(ge - go).max() < 1e-8
    mo, gmo, vmo = o.md_energy(vdw=True, valence=False)
    assert abs(ev["esum"] - mo["evdw"]) < 1e-9 * abs(mo["evdw"]) and np.abs(gv - gmo).max() < 1e-8
    ko, gko, vko = o.md_energy(vdw=False, valence=True)
    kt = ko["ebond"] + ko["eangle"] + ko["eurey"]
    assert abs(ek["esum"] - kt) < 1e-9 * abs(kt) and np.abs(gk - gko).max() < 1e-8
    assert abs(ea["esum"] - (ee["esum"] + ev["esum"] + ek["esum"])) < 1e-9 * abs(ea["esum"])
    assert np.abs(ga - (ge + gv + gk)).max() < 1e-8
    assert np.abs(va - (ve + vv + vk)).max() < 1e-8


def test_device_coordinates_and_gradient():
    """t9_set_coords_device / t9_energy_device on torch device tensors give
    the host path's bits."""
    import torch
    print("case device pointers", flush=True)
    s = water216("vdw-cutoff 7.0\n")
    xyz = s.f64("xyz", (s.n, 3))
    xyz[5] += 0.01
    h = t9hip.Context(s, t9hip.T9_FP32)
    h.set_coords(xyz)
    eh, gh, vh = h.energy()
    d = t9hip.Context(s, t9hip.T9_FP32)
    dx = torch.tensor(xyz, dtype=torch.float64, device="cuda")
    dg = torch.zeros((s.n, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    d.set_coords_device(dx.data_ptr())
    ed, vd = d.energy_device(t9hip.TERM_ALL, dg.data_ptr())
    torch.cuda.synchronize()
    assert ed["esum"] == eh["esum"]
    assert np.array_equal(dg.cpu().numpy(), gh)
    back = torch.zeros_like(dx)
    torch.cuda.synchronize()
    d.get_coords_device(back.data_ptr())
    assert np.array_equal(back.cpu().numpy(), xyz)


@pytest.mark.parametrize("prec", [t9hip.T9_FP32, t9hip.T9_FP64])
@pytest.mark.parametrize("dims", [(120, 120, 120), (24, 40, 30), (160, 160, 160)])
def test_fft3d_against_rocfft(prec, dims):
    """The hand-written transform against torch.fft (rocFFT on this device):
    forward and unnormalised inverse."""
    import torch
    k1, k2, k3 = dims
    print(f"case fft3d {dims} prec={prec}", flush=True)
    dt = torch.complex64 if prec == t9hip.T9_FP32 else torch.complex128
    g = torch.Generator().manual_seed(7)
    x = torch.complex(torch.randn((k3, k2, k1), generator=g, dtype=torch.float64),
                      torch.randn((k3, k2, k1), generator=g, dtype=torch.float64)).to(dt).cuda()
    ref_f = torch.fft.fftn(x).cpu().numpy()
    ref_b = (torch.fft.ifftn(x) * (k1 * k2 * k3)).cpu().numpy()
    y = x.clone()
    torch.cuda.synchronize()
    t9hip.fft3d(y.data_ptr(), k1, k2, k3, prec, -1)
    f = y.cpu().numpy()
    y = x.clone()
    torch.cuda.synchronize()
    t9hip.fft3d(y.data_ptr(), k1, k2, k3, prec, +1)
    b = y.cpu().numpy()
    tol = 1e-5 if prec == t9hip.T9_FP32 else 1e-12
    ef = float(np.sqrt(np.mean(np.abs(f - ref_f) ** 2)) / np.sqrt(np.mean(np.abs(ref_f) ** 2)))
    eb = float(np.sqrt(np.mean(np.abs(b - ref_b) ** 2)) / np.sqrt(np.mean(np.abs(ref_b) ** 2)))
    print(f"  rel err forward {ef:.2e} inverse {eb:.2e}", flush=True)
    assert ef < tol and eb < tol


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_not_converged_error(prec):
    """polar-eps 0 cannot be reached (the test is eps < polar-eps; the
    recursively updated residual keeps shrinking geometrically, so any
    positive tolerance is eventually met): after the 100-iteration limit the
    solve reports T9_ERR_NOT_CONVERGED (src/cu/amoeba/pcg.cu:180-184), and
    the context still solves afterwards."""
    print(f"case not converged prec={prec}", flush=True)
    s = water216()
    s.set_scalar("polar_eps", 0.0)
    ctx = t9hip.Context(s, prec)
    with pytest.raises(t9hip.T9Error) as ei:
        ctx.induce()
    assert ei.value.code == t9hip.T9_ERR_NOT_CONVERGED
    assert len(ctx.induce_trace()) == 101
    s.set_scalar("polar_eps", 1e-6)
    ok = t9hip.Context(s, prec)
    ud, up, st = ok.induce()
    assert st.rms_debye < 1e-6
