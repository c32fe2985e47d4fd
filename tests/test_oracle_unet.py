"""Pins the CPU restatement of the U-Net (oracle/unet_ref.py): its backward
pass against central finite differences of its own FP32 forward, and the
pool / unpool / up / sum_children adjoint identities."""
import numpy as np


def _params(seed, scale=0.3):
    import unet_ref as U
    _, n = U.layout()
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(n) * scale).astype(np.float32)
    lay, _ = U.layout()
    for l, L in enumerate(lay):  # zero the K padding as the library does
        W = p[L["woff"]:L["woff"] + L["cout"] * L["kpad"]].reshape(L["cout"], L["kpad"])
        W[:, 27 * L["cin"]:] = 0
    return p


def test_adjoints():
    import unet_ref as U
    rng = np.random.default_rng(0)
    fine = (3, 5, 7, 6)
    a = rng.standard_normal(fine).astype(np.float32)
    c = U.half(fine[1:])
    g = rng.standard_normal((3,) + c).astype(np.float32)
    # <pool(a), g> == <a, unpool(g)> ; <up(g), a> == <g, sum_children(a)>
    assert abs((U.pool(a) * g).sum() - (a * U.unpool(g, fine)).sum()) < 1e-3
    assert abs((U.up(g, fine[1:]) * a).sum() - (g * U.sum_children(a, c)).sum()) < 1e-3


def test_backward_matches_finite_differences():
    import unet_ref as U
    shape = (7, 6, 5)
    V = 7 * 6 * 5
    rng = np.random.default_rng(1)
    x0 = rng.standard_normal(V).astype(np.float32)
    aux = rng.standard_normal(V).astype(np.float32) * 0.5
    # positive biases keep every pre-activation away from the ReLU kink, so the
    # network is smooth in its parameters and central differences are accurate
    p = _params(2, scale=0.01)
    for L in U.layout()[0]:
        p[L["boff"]:L["boff"] + L["cout"]] += 3.0
    dout = rng.standard_normal(V).astype(np.float32)
    out, S = U.forward(p, x0, aux, shape, emulate_bf16=False)
    g = U.backward(p, S, dout)
    L = lambda pp: float(np.dot(U.forward(pp, x0, aux, shape, emulate_bf16=False)[0].astype(np.float64), dout))
    lay, _ = U.layout()
    # with every ReLU active the output is linear in any single parameter, so a
    # large step is exact up to the FP32 noise of the forward pass
    h = 0.1
    checked = 0
    for l, Lr in enumerate(lay):
        for idx in (Lr["woff"] + 5, Lr["woff"] + (Lr["cout"] - 1) * Lr["kpad"] + 3, Lr["boff"]):
            e = np.zeros_like(p); e[idx] = h
            fd = (L(p + e) - L(p - e)) / (2 * h)
            assert abs(fd - g[idx]) <= 2e-2 * max(1.0, abs(fd)), (l, idx, fd, g[idx])
            checked += 1
    assert checked == 18
