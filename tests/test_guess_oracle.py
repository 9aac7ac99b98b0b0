"""Pins of the learned-guess oracle (CPU only): BF16 rounding against torch's
bfloat16 conversion, the MLP backward pass against torch autograd and
central finite differences, Adam against torch.optim.Adam, and the local-frame
features against rotation invariance and a hand-evaluated case."""
import numpy as np
import pytest
import torch

import guess_oracle as G
import oracle


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=20000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 20000),
                        np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, -0.0, 65504.0])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(G.bf16(x), ref)
    assert G.bf16(np.float32([1 + 2 ** -8]))[0] == 1.0          # tie to even
    assert G.bf16(np.float32([1 + 3 * 2 ** -8]))[0] == 1 + 2 ** -6


def _torch_loss(p, X, T):
    q = {}
    o = 0
    for name, shp in G.SHAPES:
        k = int(np.prod(shp))
        q[name] = p[o:o + k].reshape(shp)
        o += k
    silu = torch.nn.functional.silu
    H1 = silu(X @ q["w1"] + q["b1"])
    H2 = silu(H1 @ q["w2"] + q["b2"])
    Y = torch.cat([H2, X], 1) @ q["w3"] + q["b3"]
    return ((Y - T) ** 2).mean()


def test_backward_matches_autograd_and_fd():
    rng = np.random.default_rng(2)
    p = G.init_params(3).astype(np.float64)
    p += rng.normal(size=p.shape) * 0.05
    X = rng.normal(size=(11, G.NF))
    T = rng.normal(size=(11, G.NO))
    loss, g = G.loss_and_grad(p, X, T, rounding=False, dtype=np.float64)
    pt = torch.tensor(p, requires_grad=True)
    lt = _torch_loss(pt, torch.tensor(X), torch.tensor(T))
    lt.backward()
    assert abs(loss - lt.item()) < 1e-12
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-9, atol=1e-12)
    # central differences on a few parameters of every block
    h = 1e-6
    o = 0
    for name, shp in G.SHAPES:
        k = int(np.prod(shp))
        for j in (o, o + k // 2, o + k - 1):
            pp, pm = p.copy(), p.copy()
            pp[j] += h
            pm[j] -= h
            lp = G.loss_and_grad(pp, X, T, rounding=False, dtype=np.float64)[0]
            lm = G.loss_and_grad(pm, X, T, rounding=False, dtype=np.float64)[0]
            assert abs((lp - lm) / (2 * h) - g[j]) < 1e-7 * max(1.0, abs(g[j])), (name, j)
        o += k


def test_adam_matches_torch():
    rng = np.random.default_rng(4)
    p = rng.normal(size=50).astype(np.float32)
    m = np.zeros(50, np.float32)
    v = np.zeros(50, np.float32)
    pt = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 6):
        g = rng.normal(size=50).astype(np.float32)
        p, m, v = G.adam(p, g, m, v, t, 1e-3)
        pt.grad = torch.tensor(g)
        opt.step()
    np.testing.assert_allclose(p, pt.detach().numpy(), rtol=1e-6, atol=1e-7)


def _water_frames():
    # two waters: O bisector (frame 3, z = H1, x = H2), H z-then-x (frame 2)
    xyz = np.array([[0, 0, 0], [0.9572, 0, 0], [-0.24, 0.927, 0],
                    [3.0, 0.1, 0.2], [3.5, 0.8, 0.5], [3.3, -0.8, 0.6]], float)
    frame = np.array([3, 2, 2, 3, 2, 2])
    zaxis = np.array([1, 0, 0, 4, 3, 3])
    xaxis = np.array([2, 2, 1, 5, 5, 4])
    return xyz, frame, zaxis, xaxis


def _random_inputs(rng, n):
    hist = lambda: rng.normal(size=(16, n, 3))
    return dict(Ed=rng.normal(size=(n, 3)), Ep=rng.normal(size=(n, 3)), Ed_hist=hist(), Ep_hist=hist(),
                ud_hist=rng.normal(size=(2, n, 3)))


def test_features_rotation_invariant_and_by_hand():
    rng = np.random.default_rng(5)
    xyz, frame, zaxis, xaxis = _water_frames()
    n = len(frame)
    pol = np.array([0.837, 0.496, 0.496] * 2)
    jp = np.array([0, 1, 1] * 2)
    B = oracle.aspc_coefficients()
    inp = _random_inputs(rng, n)
    R = G.frames(xyz, frame, zaxis, xaxis)
    F = G.features(R, zaxis, xaxis, frame, pol, jp, B=B, **inp)
    # a global rotation of coordinates and every vector leaves the features unchanged
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.linalg.det(q))
    rot = lambda a: a @ q.T
    inp2 = {k: rot(v) for k, v in inp.items()}
    F2 = G.features(G.frames(rot(xyz), frame, zaxis, xaxis), zaxis, xaxis, frame, pol, jp, B=B, **inp2)
    np.testing.assert_allclose(F2, F, atol=1e-12)
    # by hand for atom 4 (an H whose z atom is 3 and x atom 5)
    i = 4
    z = xyz[3] - xyz[4]
    z /= np.linalg.norm(z)
    x = xyz[5] - xyz[4]
    x -= np.dot(x, z) * z
    x /= np.linalg.norm(x)
    Ri = np.stack([x, np.cross(z, x), z])
    ext = lambda h, j: sum(B[k] * h[k][j] for k in range(16))
    np.testing.assert_allclose(F[i, 0:3], Ri @ (inp["Ed"][i] - ext(inp["Ed_hist"], i)), atol=1e-12)
    np.testing.assert_allclose(F[i, 9:12], Ri @ (inp["Ep"][3] - ext(inp["Ep_hist"], 3)), atol=1e-12)
    np.testing.assert_allclose(F[i, 12:15], Ri @ (inp["Ed"][5] - ext(inp["Ed_hist"], 5)), atol=1e-12)
    np.testing.assert_allclose(F[i, 21:24], Ri @ (inp["ud_hist"][0][i] - inp["ud_hist"][1][i]), atol=1e-12)
    assert F[i, 27] == pol[i] and F[i, 29] == 1.0 and F[i, 28] == 0.0
    # the bisector frame of an O: z along the sum of the two unit bonds
    b1, b2 = xyz[1] - xyz[0], xyz[2] - xyz[0]
    zb = b1 / np.linalg.norm(b1) + b2 / np.linalg.norm(b2)
    np.testing.assert_allclose(R[0][2], zb / np.linalg.norm(zb), atol=1e-12)


@pytest.mark.parametrize("m", [1, 33, 100])
def test_rounded_forward_close_to_exact(m):
    """BF16 rounding changes outputs by a few 1e-3 relative, not more: the
    rounded oracle is a perturbation of the exact network."""
    rng = np.random.default_rng(6)
    p = G.init_params(7)
    X = rng.normal(size=(m, G.NF)).astype(np.float32)
    Y1, _ = G.forward(p, X, rounding=True)
    Y0, _ = G.forward(p.astype(np.float64), X.astype(np.float64), rounding=False, dtype=np.float64)
    assert np.abs(Y1 - Y0).max() < 2e-2 * max(1.0, np.abs(Y0).max())
