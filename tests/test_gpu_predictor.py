"""GPU ASPC predictor against the oracle along a smooth synthetic trajectory:
the same initial-guess residual, iteration counts and converged dipoles
(FP64), and the predictor switching on after 16 stored solutions
(src/cu/amoeba/pcg.cu:30-33)."""
import numpy as np
import pytest

import oracle
import t9hip

pytestmark = pytest.mark.gpu
KEY = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid 24 24 24\npme-order 5\npolar-eps 1e-6\n"
       "polarization mutual\nneighbor-list\nlist-buffer 1.5\nusolve-cutoff 4.5\nusolve-diag 2.0\n"
       "polar-predict aspc\nvdwterm none\n")


def frames(x0, nframe, seed=11):
    rng = np.random.default_rng(seed)
    amp = 0.03 * rng.standard_normal(x0.shape)
    w = 0.2 + 0.1 * rng.random(x0.shape)
    ph = 2 * np.pi * rng.random(x0.shape)
    return [x0 + amp * np.sin(w * t + ph) for t in range(nframe)]


@pytest.mark.parametrize("prec", [t9hip.T9_FP64, t9hip.T9_FP32])
def test_aspc_trajectory(prec):
    print(f"case aspc trajectory prec={prec}", flush=True)
    s = t9hip.System.water_box(216, 6, 18.643, 2024, key=KEY)
    x0 = s.arrays()["xyz"].copy()
    ctx = t9hip.Context(s, prec)
    o = oracle.Oracle(s.arrays())
    B = oracle.aspc_coefficients()
    hist = []
    for t, x in enumerate(frames(x0, 22)):
        ctx.set_coords(x)
        ud, up, st = ctx.induce()
        o.set_xyz(x)
        if len(hist) >= 16:
            g_d = sum(B[j] * hist[-1 - j][0] for j in range(16))
            g_p = sum(B[j] * hist[-1 - j][1] for j in range(16))
            od, op, oit, orms, og, rc = o.induce_from(g_d, g_p)
        else:
            od, op, oit, orms, rc = o.induce()
            og = None
        hist.append((od, op))
        assert st.predictor == (1 if t >= 16 else 0)
        # per-iteration residuals: the stopping rule (first iteration >= 3
        # below polar-eps) and the stats agree with the trace
        tr = ctx.induce_trace()
        assert len(tr) == st.iterations + 1
        assert tr[0] == st.guess_rms and tr[-1] == st.rms_debye
        assert tr[-1] < 1e-6 and all(tr[k] >= 1e-6 for k in range(3, st.iterations))
        tol = 1e-9 if prec == t9hip.T9_FP64 else 2e-5
        assert np.abs(ud - od).max() < tol * 10 and np.abs(up - op).max() < tol * 10
        if og is not None:
            print(f"frame {t}: guess rms gpu {st.guess_rms:.3e} oracle {og:.3e}, iters {st.iterations} / {oit}",
                  flush=True)
            assert abs(st.guess_rms - og) < (1e-9 if prec == t9hip.T9_FP64 else 2e-3 * og + 1e-5)
            if prec == t9hip.T9_FP64:
                assert st.iterations == oit
