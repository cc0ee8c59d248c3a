"""FP64 numpy restatement of the learned preconditioner's U-Net.

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward/backward kernels
(ouranos_amd/csrc/unet.hip), never the product.  There is no reference
counterpart (the reference solves with PreconditionIdentity,
src/ouranos.cc:161); this file restates the network the library defines:

  in(16: r_hat[8], tau_hat, kappa_hat, 0...) -> enc0a(32) -> enc0b(32)=skip0
  avgpool2 -> enc1a(64) -> enc1b(64)=skip1 -> avgpool2 -> bota(128) -> botb(128)
  upsample2 ++ skip1 (192) -> dec1a(64) -> dec1b(64)
  upsample2 ++ skip0 (96)  -> dec0a(32) -> dec0b(32) -> head 1x1 (8)

3x3x3 convolutions with zero padding, bias and ReLU; tensors channels-last
[sample][z][y][x][c]; parameters flat: per convolution W[27][cout][cin] then
bias[cout] (tap = 9(dz+1) + 3(dy+1) + (dx+1)), head w[8][32] then b[8].
Features per cell: r_hat[8], tau_hat, kappa_hat, g, unused (12).  The
prediction is pred[k] = gain (skip r_hat[k] + (out(+r) - out(-r))[k] / 2), k < 8
(mode bit 0: skip = 1 else 0; bit 1: gain = g else 1) and the loss
mean_b |pred_b - t_b|^2 / |t_b|^2.

With emulate=True every tensor the GPU stores in BF16 is rounded to BF16
(round to nearest even, via FP32) at the same point; the arithmetic between
those points is FP64.  With emulate=False the whole network is FP64, which is
what the finite-difference check of the backward pass uses.
"""
import numpy as np

IN_PAD = 16
IN_REAL = 10
N_OUT = 8
N_FEAT = 12
GAIN = 10
CONVS = [(16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128),
         (192, 64), (64, 64), (96, 32), (32, 32)]


def layout():
    """[(cin, cout, w_off, b_off)], head_w_off, head_b_off, n_params."""
    out, off = [], 0
    for cin, cout in CONVS:
        w = off
        off += 27 * cin * cout
        out.append((cin, cout, w, off))
        off += cout
    return out, off, off + 32 * N_OUT, off + 33 * N_OUT


def bf16(a):
    """Round to BF16 (nearest even) through FP32; returns float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _ident(a):
    return np.asarray(a, dtype=np.float64)


def _shift(xp, tap, D):
    dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
    return xp[:, dz:dz + D, dy:dy + D, dx:dx + D, :]


def conv3(x, W, b):
    """Pre-activation sum_tap x(p + tap) W[tap]^T + b, zero padding."""
    S, D, cin = x.shape[0], x.shape[1], x.shape[-1]
    xp = np.zeros((S, D + 2, D + 2, D + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    y = np.zeros((S, D, D, D, W.shape[1]))
    for tap in range(27):
        y += _shift(xp, tap, D) @ W[tap].T
    return y + b


def conv3_back(x, W, dy):
    """(dW, db, dx) of conv3 for the pre-activation gradient dy."""
    S, D, cin = x.shape[0], x.shape[1], x.shape[-1]
    cout = W.shape[1]
    xp = np.zeros((S, D + 2, D + 2, D + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    dyf = dy.reshape(-1, cout)
    dW = np.stack([dyf.T @ _shift(xp, t, D).reshape(-1, cin) for t in range(27)])
    db = dyf.sum(0)
    dxp = np.zeros_like(xp)
    for tap in range(27):
        dz, dyy, dx = tap // 9, (tap // 3) % 3, tap % 3
        dxp[:, dz:dz + D, dyy:dyy + D, dx:dx + D, :] += dy @ W[tap]
    return dW, db, dxp[:, 1:-1, 1:-1, 1:-1, :]


def pool2(x):
    S, D, C = x.shape[0], x.shape[1], x.shape[-1]
    d = D // 2
    return x.reshape(S, d, 2, d, 2, d, 2, C).mean(axis=(2, 4, 6))


def up2(x):
    return x.repeat(2, 1).repeat(2, 2).repeat(2, 3)


def up2_back(dy):
    return pool2(dy) * 8.0


def _params(params, emulate):
    lay, hw, hb, n = layout()
    p = np.asarray(params, dtype=np.float64)
    if p.size != n:
        raise ValueError(f"expected {n} parameters, got {p.size}")
    q = bf16 if emulate else _ident
    Ws, bs = [], []
    for cin, cout, w, b in lay:
        Ws.append(q(p[w:w + 27 * cin * cout]).reshape(27, cout, cin))
        bs.append(p[b:b + cout])
    return Ws, bs, p[hw:hb].reshape(N_OUT, 32), p[hb:hb + N_OUT]


def forward(params, feat, G, emulate=True, mode=0):
    """feat [B][G^3][12] -> (pred [B][G^3 * 8], cache)."""
    q = bf16 if emulate else _ident
    Ws, bs, hw, hb = _params(params, emulate)
    feat = np.asarray(feat, dtype=np.float64).reshape(-1, G, G, G, N_FEAT)
    B = feat.shape[0]
    x0 = np.zeros((2 * B, G, G, G, IN_PAD))
    x0[:B, ..., 0:8] = feat[..., 0:8]
    x0[B:, ..., 0:8] = -feat[..., 0:8]
    x0[:B, ..., 8:10] = feat[..., 8:10]
    x0[B:, ..., 8:10] = feat[..., 8:10]
    x0 = q(x0)

    def cr(x, i):
        return q(np.maximum(conv3(x, Ws[i], bs[i]), 0.0))

    c = {"x0": x0}
    c["e0a"] = cr(x0, 0)
    c["e0b"] = cr(c["e0a"], 1)
    c["p0"] = q(pool2(c["e0b"]))
    c["e1a"] = cr(c["p0"], 2)
    c["e1b"] = cr(c["e1a"], 3)
    c["p1"] = q(pool2(c["e1b"]))
    c["ba"] = cr(c["p1"], 4)
    c["bb"] = cr(c["ba"], 5)
    c["cat1"] = np.concatenate([up2(c["bb"]), c["e1b"]], axis=-1)
    c["d1a"] = cr(c["cat1"], 6)
    c["d1b"] = cr(c["d1a"], 7)
    c["cat0"] = np.concatenate([up2(c["d1b"]), c["e0b"]], axis=-1)
    c["d0a"] = cr(c["cat0"], 8)
    c["d0b"] = cr(c["d0a"], 9)
    out = c["d0b"] @ hw.T + hb
    if emulate:
        out = out.astype(np.float32).astype(np.float64)
    c["out"] = out
    g = feat[..., GAIN:GAIN + 1] if mode & 2 else np.ones_like(feat[..., GAIN:GAIN + 1])
    skip = feat[..., 0:8] if mode & 1 else 0.0
    pred = g * (skip + 0.5 * (out[:B] - out[B:]))
    c["B"], c["g"] = B, g
    return pred.reshape(B, -1), c


def loss_and_grad(params, feat, target, G, emulate=True, mode=0):
    """Loss mean_b |pred_b - t_b|^2/|t_b|^2 and its gradient (flat, like params)."""
    q = bf16 if emulate else _ident
    lay, hwo, hbo, n = layout()
    Ws, bs, hw, hb = _params(params, emulate)
    pred, c = forward(params, feat, G, emulate, mode)
    B = c["B"]
    t = np.asarray(target, dtype=np.float64).reshape(B, -1)
    num = ((pred - t) ** 2).sum(1)
    den = (t ** 2).sum(1)
    loss = float(np.mean(num / den))
    g = np.broadcast_to(c["g"], (B, G, G, G, N_OUT)).reshape(B, -1)
    d = ((pred - t) * g / (B * den[:, None])).reshape(B, G, G, G, N_OUT)
    dout = np.concatenate([d, -d])
    grad = np.zeros(n)

    def mask(a):
        return (a > 0).astype(np.float64)

    grad[hwo:hbo] = np.einsum("sijkn,sijkc->nc", dout, c["d0b"]).reshape(-1)
    grad[hbo:hbo + N_OUT] = dout.reshape(-1, N_OUT).sum(0)
    dd0b = q(mask(c["d0b"]) * (dout @ hw))

    def back(i, x, dy, m=None):
        dW, db, dx = conv3_back(x, Ws[i], dy)
        cin, cout, w, b = lay[i]
        grad[w:w + 27 * cin * cout] = dW.reshape(-1)
        grad[b:b + cout] = db
        return q(dx if m is None else mask(m) * dx)

    dd0a = back(9, c["d0a"], dd0b, c["d0a"])
    dcat0 = back(8, c["cat0"], dd0a)
    dd1b = q(mask(c["d1b"]) * up2_back(dcat0[..., :64]))
    dd1a = back(7, c["d1a"], dd1b, c["d1a"])
    dcat1 = back(6, c["cat1"], dd1a)
    dbb = q(mask(c["bb"]) * up2_back(dcat1[..., :128]))
    dba = back(5, c["ba"], dbb, c["ba"])
    dp1 = back(4, c["p1"], dba)
    dskip1 = q(mask(c["e1b"]) * (dcat1[..., 128:] + up2(dp1) / 8.0))
    de1a = back(3, c["e1a"], dskip1, c["e1a"])
    dp0 = back(2, c["p0"], de1a)
    dskip0 = q(mask(c["e0b"]) * (dcat0[..., 64:] + up2(dp0) / 8.0))
    de0a = back(1, c["e0a"], dskip0, c["e0a"])
    back(0, c["x0"], de0a)
    return loss, grad


def init_params(seed=0):
    """He-normal weights (padded input channels zero), zero biases; numpy RNG."""
    lay, hw, hb, n = layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for i, (cin, cout, w, b) in enumerate(lay):
        real = IN_REAL if i == 0 else cin
        W = rng.normal(0.0, np.sqrt(2.0 / (27 * real)), (27, cout, cin))
        W[..., real:] = 0.0
        p[w:w + W.size] = W.reshape(-1)
        p[b:b + cout] = rng.normal(0.0, 0.05, cout)
    p[hw:hb] = rng.normal(0.0, 0.3 * np.sqrt(1.0 / 32), hb - hw)
    p[hb:n] = 0.01
    return p.astype(np.float32).astype(np.float64)
