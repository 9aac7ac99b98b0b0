"""The allocation-free 3D transform (t9_fft3d_noalloc: twiddles in a static
per-GPU table) against the allocating one (t9_fft3d), bit for bit, on two
calls in a row, the second on a smaller grid (the table keeps the first
call's larger contents), in FP32 and FP64, forward and inverse."""
import numpy as np
import pytest

import t9hip

pytestmark = pytest.mark.gpu


def _grid(dims, prec, seed):
    import torch
    k1, k2, k3 = dims
    dt = torch.complex64 if prec == t9hip.T9_FP32 else torch.complex128
    g = torch.Generator().manual_seed(seed)
    return torch.complex(torch.randn((k3, k2, k1), generator=g, dtype=torch.float64),
                         torch.randn((k3, k2, k1), generator=g, dtype=torch.float64)).to(dt).cuda()


@pytest.mark.parametrize("prec", [t9hip.T9_FP32, t9hip.T9_FP64])
def test_fft3d_noalloc_matches_allocating(prec):
    import torch
    for dims, sign, seed in (((120, 120, 120), -1, 11), ((24, 40, 30), +1, 12), ((32, 32, 32), -1, 13)):
        print(f"case fft3d_noalloc {dims} sign={sign} prec={prec}", flush=True)
        x = _grid(dims, prec, seed)
        a, b = x.clone(), x.clone()
        torch.cuda.synchronize()
        t9hip.fft3d(a.data_ptr(), *dims, prec, sign)
        t9hip.fft3d_noalloc(b.data_ptr(), *dims, prec, sign)
        assert torch.equal(a, b)
        assert not torch.equal(a, x)


def test_fft3d_noalloc_rejects_bad_lengths():
    import torch
    x = _grid((8, 8, 8), t9hip.T9_FP32, 1)
    torch.cuda.synchronize()
    for dims in ((7, 8, 8), (8, 8, 256), (0, 8, 8)):
        with pytest.raises(t9hip.T9Error) as e:
            t9hip.fft3d_noalloc(x.data_ptr(), *dims, t9hip.T9_FP32, -1)
        assert e.value.code == t9hip.T9_ERR_ARG
    # the grid was not touched by the rejected calls
    assert torch.equal(x, _grid((8, 8, 8), t9hip.T9_FP32, 1))
    ref = np.fft.fftn(x.cpu().numpy().astype(np.complex128))
    t9hip.fft3d_noalloc(x.data_ptr(), 8, 8, 8, t9hip.T9_FP32, -1)
    assert np.abs(x.cpu().numpy() - ref).max() < 1e-4 * np.abs(ref).max()
