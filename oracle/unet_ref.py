"""CPU restatement of the learned preconditioner's U-Net (TEST INFRASTRUCTURE
ONLY; no reference counterpart).  FP32 arithmetic through the oracle's conv3d
(femref.c), numpy pooling / upsampling; optional bf16 rounding at the points
where the GPU stores bf16, so forward outputs compare tightly.

Parameter layout (as include/femgpu.h): per layer W[cout][kpad] with
k = tap * cin + c, tap = kz*9 + ky*3 + kx, then bias[cout].
Layers: 0: 4->16 @D0, 1: 16->32 @D1, 2: 32->64 @D2, 3: 64->32 @D2,
4: 32->16 @D1, 5: 16->1 @D0.
"""
import numpy as np

import femref as fr

CINS = [4, 16, 32, 64, 32, 16]
COUTS = [16, 32, 64, 32, 16, 1]


def kpad(cin, layer):
    return 432 if layer == 5 else (27 * cin + 15) // 16 * 16


def layout():
    out, off = [], 0
    for l in range(6):
        kp = kpad(CINS[l], l)
        woff = off; off += COUTS[l] * kp
        boff = off; off += COUTS[l]
        out.append(dict(cin=CINS[l], cout=COUTS[l], kpad=kp, woff=woff, boff=boff))
    return out, off


def unpack(params):
    """-> list of (w[cout][cin][3][3][3], b[cout])"""
    lay, _ = layout()
    res = []
    for l, L in enumerate(lay):
        W = params[L["woff"]:L["woff"] + L["cout"] * L["kpad"]].reshape(L["cout"], L["kpad"])
        W = W[:, :27 * L["cin"]].reshape(L["cout"], 27, L["cin"])  # [n][tap][c]
        w = np.transpose(W, (0, 2, 1)).reshape(L["cout"], L["cin"], 3, 3, 3)
        res.append((np.ascontiguousarray(w, np.float32),
                    np.ascontiguousarray(params[L["boff"]:L["boff"] + L["cout"]], np.float32)))
    return res


def pack_grad(dws):
    lay, n = layout()
    g = np.zeros(n, np.float32)
    for L, (dw, db) in zip(lay, dws):
        W = np.transpose(dw.reshape(L["cout"], L["cin"], 27), (0, 2, 1)).reshape(L["cout"], -1)
        blk = np.zeros((L["cout"], L["kpad"]), np.float32)
        blk[:, :W.shape[1]] = W
        g[L["woff"]:L["woff"] + blk.size] = blk.ravel()
        g[L["boff"]:L["boff"] + L["cout"]] = db
    return g


def bf16(a):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def half(s):
    return tuple((n + 1) // 2 for n in s)


def pool(a):
    """avg 2x2x2 with ceil on [C][z][y][x]"""
    C, nz, ny, nx = a.shape
    cz, cy, cx = half((nz, ny, nx))
    out = np.zeros((C, cz, cy, cx), np.float32)
    cnt = np.zeros((cz, cy, cx), np.float32)
    for dz in range(2):
        for dy in range(2):
            for dx in range(2):
                sub = a[:, dz::2, dy::2, dx::2]
                out[:, :sub.shape[1], :sub.shape[2], :sub.shape[3]] += sub
                cnt[:sub.shape[1], :sub.shape[2], :sub.shape[3]] += 1
    return out / cnt


def unpool(g, fine):
    """pool backward: distribute coarse gradient / count to the children"""
    C, nz, ny, nx = fine
    cz, cy, cx = g.shape[1:]
    cnt = np.zeros((cz, cy, cx), np.float32)
    for dz in range(2):
        for dy in range(2):
            for dx in range(2):
                cnt[:len(range(dz, nz, 2)), :len(range(dy, ny, 2)), :len(range(dx, nx, 2))] += 1
    return up(g / cnt, fine[1:])


def up(a, shape):
    """nearest upsample [C][cz][cy][cx] -> [C][shape]"""
    nz, ny, nx = shape
    iz, iy, ix = np.arange(nz) // 2, np.arange(ny) // 2, np.arange(nx) // 2
    return a[:, iz][:, :, iy][:, :, :, ix]


def sum_children(g, coarse):
    C = g.shape[0]
    out = np.zeros((C,) + tuple(coarse), np.float32)
    for dz in range(2):
        for dy in range(2):
            for dx in range(2):
                sub = g[:, dz::2, dy::2, dx::2]
                out[:, :sub.shape[1], :sub.shape[2], :sub.shape[3]] += sub
    return out


def forward(params, x0, aux, shape, emulate_bf16=True):
    """x0, aux: flat (x fastest) arrays on the grid `shape` = (nx, ny, nz)."""
    nx, ny, nz = shape
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float32))
    Ws = unpack(params)
    Wq = [(q(w), b) for w, b in Ws]
    D0 = (nz, ny, nx)
    inp = np.zeros((4,) + D0, np.float32)
    inp[0] = np.asarray(x0, np.float32).reshape(D0)
    inp[1] = np.asarray(aux, np.float32).reshape(D0)
    inp = q(inp)
    S = {}
    S["in"] = inp
    S["e0"] = q(np.maximum(fr.conv3d_fwd(inp, *Wq[0]), 0))
    S["p0"] = q(pool(S["e0"]))
    S["e1"] = q(np.maximum(fr.conv3d_fwd(S["p0"], *Wq[1]), 0))
    S["p1"] = q(pool(S["e1"]))
    S["e2"] = q(np.maximum(fr.conv3d_fwd(S["p1"], *Wq[2]), 0))
    S["t1"] = fr.conv3d_fwd(S["e2"], *Wq[3])
    S["d1"] = q(np.maximum(up(S["t1"], S["e1"].shape[1:]) + S["e1"], 0))
    S["t0"] = fr.conv3d_fwd(S["d1"], *Wq[4])
    S["d0"] = q(np.maximum(up(S["t0"], D0) + S["e0"], 0))
    # output layer: f32 weights (the GPU keeps it in f32 on the VALU)
    S["out"] = fr.conv3d_fwd(S["d0"], *Ws[5])[0]
    return S["out"].ravel(), S


def backward(params, S, dout, emulate_bf16=False):
    """FP32 parameter gradient of sum(dout * out) given forward state S.
    emulate_bf16: round to bf16 where the GPU's backward does -- each hidden
    layer's output gradient before its weight gradient, bias gradient and data
    gradient, and that layer's weights in the data gradient; the output
    layer's weight gradient takes bf16(dout), its bias and data gradients the
    f32 dout and weights."""
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float32))
    Ws = unpack(params)
    D0 = S["e0"].shape[1:]
    g_out = np.asarray(dout, np.float32).reshape((1,) + D0)
    grads = [None] * 6
    dx, dw, db = fr.conv3d_bwd(S["d0"], Ws[5][0], g_out)
    if emulate_bf16:
        _, dw, _ = fr.conv3d_bwd(S["d0"], Ws[5][0], q(g_out))
    grads[5] = (dw, db)

    def layer(x, l, g, need_dx=True):
        gq = q(g)
        return fr.conv3d_bwd(x, q(Ws[l][0]), gq, need_dx=need_dx)
    g = dx * (S["d0"] > 0)
    ge0 = g.copy()
    gt0 = sum_children(g, S["t0"].shape[1:])
    dx, dw, db = layer(S["d1"], 4, gt0)
    grads[4] = (dw, db)
    g1 = dx * (S["d1"] > 0)
    ge1 = g1.copy()
    gt1 = sum_children(g1, S["t1"].shape[1:])
    dx, dw, db = layer(S["e2"], 3, gt1)
    grads[3] = (dw, db)
    g2 = dx * (S["e2"] > 0)
    dx, dw, db = layer(S["p1"], 2, g2)
    grads[2] = (dw, db)
    ge1 += unpool(dx, S["e1"].shape)
    g1b = ge1 * (S["e1"] > 0)
    dx, dw, db = layer(S["p0"], 1, g1b)
    grads[1] = (dw, db)
    ge0 += unpool(dx, S["e0"].shape)
    g0 = ge0 * (S["e0"] > 0)
    _, dw, db = layer(S["in"], 0, g0, need_dx=False)
    grads[0] = (dw, db)
    return pack_grad(grads)
