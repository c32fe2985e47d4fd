"""Pins the conv3d oracle (learned preconditioner; no reference counterpart)
against PyTorch's CPU conv3d and against finite differences."""
import numpy as np
import pytest


def test_conv3d_vs_torch(fr):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 5, 6, 7)).astype(np.float32)
    w = rng.standard_normal((4, 3, 3, 3, 3)).astype(np.float32)
    b = rng.standard_normal(4).astype(np.float32)
    y = fr.conv3d_fwd(x, w, b)
    xt = torch.tensor(x[None], requires_grad=True)
    wt = torch.tensor(w, requires_grad=True)
    bt = torch.tensor(b, requires_grad=True)
    yt = torch.nn.functional.conv3d(xt, wt, bt, padding=1)
    assert np.abs(y - yt.detach().numpy()[0]).max() < 1e-4
    dy = rng.standard_normal(y.shape).astype(np.float32)
    yt.backward(torch.tensor(dy[None]))
    dx, dw, db = fr.conv3d_bwd(x, w, dy)
    assert np.abs(dx - xt.grad.numpy()[0]).max() < 1e-4
    assert np.abs(dw - wt.grad.numpy()).max() < 1e-3
    assert np.abs(db - bt.grad.numpy()).max() < 1e-3


def test_conv3d_grad_fd(fr):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 4, 3)).astype(np.float64)
    w = rng.standard_normal((2, 2, 3, 3, 3))
    b = rng.standard_normal(2)
    dy = rng.standard_normal((2, 3, 4, 3))
    L = lambda x_, w_: float((fr.conv3d_fwd(x_, w_, b).astype(np.float64) * dy).sum())
    dx, dw, _ = fr.conv3d_bwd(x, w, dy)
    h = 1e-2
    for idx in [(0, 1, 2, 1), (1, 0, 0, 2)]:
        e = np.zeros_like(x); e[idx] = h
        fd = (L(x + e, w) - L(x - e, w)) / (2 * h)
        assert abs(fd - dx[idx]) < 2e-3 * max(1, abs(fd))
    for idx in [(0, 1, 1, 2, 0), (1, 1, 0, 0, 2)]:
        e = np.zeros_like(w); e[idx] = h
        fd = (L(x, w + e) - L(x, w - e)) / (2 * h)
        assert abs(fd - dw[idx]) < 2e-3 * max(1, abs(fd))
