"""CPU restatement of the learned initial-guess predictor (test infrastructure
only: the product runs the HIP kernels in tinker-gpu_amd/csrc/hip/guess.hip).

The reference has no learned predictor; the plug point it replaces is the
ASPC/GEAR/LSQR predictor of src/amoeba/induce.cpp:29-69 and
src/cu/upredict.cu:34-201.  This module pins the GPU kernels:

* features(...): the per-atom inputs in the atom's multipole local frame,
* forward / backward of the 3-layer MLP with a linear skip, in the exact
  arithmetic the MFMA kernels use (inputs and activations rounded to BF16 with
  round-to-nearest-even, products accumulated in FP32),
* adam(...): one Adam update of the FP32 master parameters.

The backward pass is pinned against central finite differences of an FP64
forward without BF16 rounding (tests/test_guess_oracle.py); the features
against a direct evaluation from their definition on a hand-built case.
"""
import numpy as np

NF = 32         # input features
NH = 128        # hidden width
NO = 6          # outputs: correction of u_d, u_p in the local frame
NSKIP = NH + NF  # last layer sees [H2, X]

# flat parameter layout (FP32 master copy)
SHAPES = [("w1", (NF, NH)), ("b1", (NH,)), ("w2", (NH, NH)), ("b2", (NH,)), ("w3", (NSKIP, NO)), ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = p[o:o + k].reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float32).ravel() for name, _ in SHAPES])


def init_params(seed):
    """He-style uniform init from a seeded generator (the device uses the
    same values: they are produced here or by t9_guess_init's own stream)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("b"):
            d[name] = np.zeros(shp, np.float32)
        else:
            lim = np.sqrt(3.0 / shp[0])
            d[name] = rng.uniform(-lim, lim, size=shp).astype(np.float32)
    d["w3"] *= 0.1
    return pack(d)


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as
    float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, rounding=True, dtype=np.float32):
    """X [m, 32] normalized features.  Returns Y [m, 6] and the cache."""
    q = unpack(p.astype(dtype))
    r = bf16 if rounding else (lambda a: a)
    Xb = r(X.astype(dtype)).astype(dtype)
    Z1 = Xb @ r(q["w1"]).astype(dtype) + q["b1"]
    H1 = r(silu(Z1).astype(dtype)).astype(dtype)
    Z2 = H1 @ r(q["w2"]).astype(dtype) + q["b2"]
    H2 = r(silu(Z2).astype(dtype)).astype(dtype)
    C = np.concatenate([H2, Xb], axis=1)
    Y = C @ r(q["w3"]).astype(dtype) + q["b3"]
    return Y, dict(Xb=Xb, Z1=Z1, H1=H1, Z2=Z2, H2=H2, C=C)


def loss_and_grad(p, X, T, rounding=True, dtype=np.float32):
    """Mean squared error over all m*6 outputs and its gradient with respect
    to the flat parameters.  With rounding, the backward pass rounds the
    back-propagated signals that enter a product (dY, dZ2, dZ1) to BF16, as
    the kernels do; weight-gradient products accumulate in FP32."""
    m = X.shape[0]
    Y, c = forward(p, X, rounding, dtype)
    q = unpack(p.astype(dtype))
    r = bf16 if rounding else (lambda a: a)
    D = (Y - T).astype(dtype)
    loss = float(np.mean(D.astype(np.float64) ** 2))
    dY = (2.0 / (m * NO)) * D
    dYb = r(dY.astype(dtype)).astype(dtype)
    g = {}
    g["w3"] = c["C"].T @ dYb
    g["b3"] = dY.sum(0)
    dH2 = dYb @ r(q["w3"]).astype(dtype)[:NH].T
    dZ2 = dH2 * dsilu(c["Z2"])
    dZ2b = r(dZ2.astype(dtype)).astype(dtype)
    g["w2"] = c["H1"].T @ dZ2b
    g["b2"] = dZ2.sum(0)
    dH1 = dZ2b @ r(q["w2"]).astype(dtype).T
    dZ1 = dH1 * dsilu(c["Z1"])
    dZ1b = r(dZ1.astype(dtype)).astype(dtype)
    g["w1"] = c["Xb"].T @ dZ1b
    g["b1"] = dZ1.sum(0)
    return loss, np.concatenate([np.asarray(g[name], dtype=dtype).ravel() for name, _ in SHAPES])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, bias-corrected), all FP32.  t counts from 1."""
    m = (b1 * m + (1 - b1) * g).astype(np.float32)
    v = (b2 * v + (1 - b2) * g * g).astype(np.float32)
    mh = m / np.float32(1 - b1 ** t)
    vh = v / np.float32(1 - b2 ** t)
    p = (p - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))).astype(np.float32)
    return p, m, v


def aspc_extrap(hist, B):
    """sum_k B[k] hist[k]; hist[0] is the most recent entry."""
    return sum(B[k] * hist[k] for k in range(len(B)))


def frames(xyz, frame, zaxis, xaxis, box=None):
    """Rows = local x, y, z axes of each atom's multipole frame (Z-only,
    Z-then-X and Bisector frames; identity for atoms without a frame), as in
    include/seq/rotpole.h:8-224."""
    n = len(frame)
    R = np.tile(np.eye(3), (n, 1, 1))

    def d(j, i):
        v = xyz[j] - xyz[i]
        if box is not None:
            v = v - box * np.round(v / box)
        return v / np.linalg.norm(v)

    for i in range(n):
        fr = frame[i]
        if fr == 0:
            continue
        z = d(zaxis[i], i)
        if fr == 1:
            x = np.array([1.0, 0, 0]) if abs(z[0]) <= 0.866 else np.array([0, 1.0, 0])
        else:
            x = d(xaxis[i], i)
        if fr == 3:
            z = z + x
            z /= np.linalg.norm(z)
        elif fr not in (1, 2):
            raise ValueError("frame type not handled by the feature oracle")
        x = x - np.dot(x, z) * z
        x /= np.linalg.norm(x)
        R[i] = np.stack([x, np.cross(z, x), z])
    return R


def features(R, zaxis, xaxis, frame, polarity, jpolar, Ed, Ep, Ed_hist, Ep_hist, ud_hist, B):
    """Raw (unnormalized) features, caller order, [n, 32]:
    0-5   dE_d, dE_p of the atom: current direct field minus its ASPC
          extrapolation from the previous 16 fields
    6-11  the same for the frame's z-axis atom, 12-17 for its x-axis atom
          (the atom itself when the frame has none)
    18-20 the current direct field E_d
    21-23 u_d(t-1) - u_d(t-2)
    24-26 u_d(t-1)
    27    polarizability
    28-31 one-hot of min(polarize type, 3)
    All vectors in the atom's local frame (rows of R).
    """
    n = len(polarity)
    dEd = Ed - aspc_extrap(Ed_hist, B)
    dEp = Ep - aspc_extrap(Ep_hist, B)
    zi = np.where(frame > 0, zaxis, np.arange(n))
    xi = np.where((frame > 1) & (xaxis >= 0), xaxis, np.arange(n))
    loc = lambda v: np.einsum("nij,nj->ni", R, v)
    F = np.zeros((n, NF))
    F[:, 0:3], F[:, 3:6] = loc(dEd), loc(dEp)
    F[:, 6:9], F[:, 9:12] = loc(dEd[zi]), loc(dEp[zi])
    F[:, 12:15], F[:, 15:18] = loc(dEd[xi]), loc(dEp[xi])
    F[:, 18:21] = loc(Ed)
    F[:, 21:24] = loc(ud_hist[0] - ud_hist[1])
    F[:, 24:27] = loc(ud_hist[0])
    F[:, 27] = polarity
    F[np.arange(n), 28 + np.minimum(jpolar, 3)] = 1.0
    return F
