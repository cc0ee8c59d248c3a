"""Pins the U-Net oracle (oracle/unet_oracle.py) on the CPU: its convolution
against a direct loop, the adjoint identities of its backward pieces, BF16
rounding against exact cases, and its full gradient against central finite
differences in FP64."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5,
                     0.0, 3.0e38, 1e-30])
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0, 2.9993e38, 1e-30])
    got = U.bf16(vals)
    # ties go to even; 1+3*2^-9 rounds up; 1+2^-7+2^-8 is a tie between 1+2^-7 (odd) and 1+2^-6
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[2] == 1.0 + 2 ** -7
    assert got[3] == 1.0 + 2 ** -6
    assert got[4] == -2.5 and got[5] == 0.0
    assert abs(got[6] - want[6]) / want[6] < 4e-3
    assert abs(got[7] - 1e-30) / 1e-30 < 4e-3
    # 8 significant bits: every value is within half an ulp
    x = np.random.default_rng(0).normal(size=1000) * 10.0 ** np.random.default_rng(1).integers(
        -5, 5, 1000)
    assert np.all(np.abs(U.bf16(x) - x) <= np.abs(x) * 2 ** -8 * 1.0001)


def test_conv3_matches_direct_loop():
    rng = np.random.default_rng(2)
    S, D, cin, cout = 2, 3, 2, 3
    x = rng.normal(size=(S, D, D, D, cin))
    W = rng.normal(size=(27, cout, cin))
    b = rng.normal(size=cout)
    y = U.conv3(x, W, b)
    ref = np.zeros_like(y)
    for s in range(S):
        for z in range(D):
            for yy in range(D):
                for xx in range(D):
                    acc = b.copy()
                    for dz in (-1, 0, 1):
                        for dy in (-1, 0, 1):
                            for dx in (-1, 0, 1):
                                zz, y2, x2 = z + dz, yy + dy, xx + dx
                                if 0 <= zz < D and 0 <= y2 < D and 0 <= x2 < D:
                                    tap = 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)
                                    acc += W[tap] @ x[s, zz, y2, x2]
                    ref[s, z, yy, xx] = acc
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_conv3_back_is_the_adjoint():
    rng = np.random.default_rng(3)
    S, D, cin, cout = 2, 4, 3, 5
    x = rng.normal(size=(S, D, D, D, cin))
    W = rng.normal(size=(27, cout, cin))
    dy = rng.normal(size=(S, D, D, D, cout))
    dW, db, dx = U.conv3_back(x, W, dy)
    zero = np.zeros(cout)
    # <dy, conv(x)> is bilinear in (x, W): its gradients are dx and dW
    lhs = np.sum(dy * U.conv3(x, W, zero))
    assert np.isclose(lhs, np.sum(dx * x), rtol=1e-12)
    assert np.isclose(lhs, np.sum(dW * W), rtol=1e-12)
    assert np.allclose(db, dy.reshape(-1, cout).sum(0))


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, 4, 4, 4, 3))
    y = rng.normal(size=(2, 2, 2, 2, 3))
    assert np.isclose(np.sum(U.pool2(x) * y), np.sum(x * U.up2(y)) / 8.0, rtol=1e-13)
    z = rng.normal(size=(2, 4, 4, 4, 3))
    assert np.isclose(np.sum(U.up2(y) * z), np.sum(y * U.up2_back(z)), rtol=1e-13)


def test_layout_counts_parameters():
    lay, hw, hb, n = U.layout()
    assert n == sum(27 * a * b + b for a, b in U.CONVS) + 33 * 8
    assert lay[0][:2] == (16, 32) and lay[-1][:2] == (32, 32)
    assert hb == hw + 8 * 32


def features(rng, B, G):
    f = np.zeros((B, G ** 3, U.N_FEAT))
    f[..., 0:8] = rng.normal(size=(B, G ** 3, 8))
    f[..., 8] = rng.uniform(-1, 1, size=(B, G ** 3))
    f[..., 9] = rng.uniform(-1, 0, size=(B, G ** 3))
    f[..., U.GAIN] = rng.uniform(1, 5, size=(B, G ** 3))
    return f


@pytest.mark.parametrize("seed,mode", [(5, 0), (8, 3)])
def test_gradient_matches_central_differences(seed, mode):
    G, B = 4, 2
    rng = np.random.default_rng(seed)
    p = U.init_params(seed)
    feat = features(rng, B, G)
    target = rng.normal(size=(B, G ** 3 * 8))
    loss, grad = U.loss_and_grad(p, feat, target, G, emulate=False, mode=mode)
    lay, hw, hb, n = U.layout()
    picks = []
    for cin, cout, w, b in lay:
        real = U.IN_REAL if cin == 16 else cin
        for _ in range(3):
            t, co, ci = rng.integers(27), rng.integers(cout), rng.integers(real)
            picks.append(w + (t * cout + co) * cin + ci)
        picks.append(b + rng.integers(cout))
    picks += [hw + 3, hw + 117, hw + 250, hb, hb + 5]
    h = 1e-6
    bad = []
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (U.loss_and_grad(pp, feat, target, G, emulate=False, mode=mode)[0] -
              U.loss_and_grad(pm, feat, target, G, emulate=False, mode=mode)[0]) / (2 * h)
        if abs(fd - grad[i]) > 1e-6 * max(1.0, abs(fd)) + 1e-5 * abs(fd):
            bad.append((i, fd, grad[i]))
    assert not bad, bad


def test_emulated_forward_is_close_to_fp64():
    G, B = 4, 1
    rng = np.random.default_rng(6)
    p = U.init_params(6)
    feat = features(rng, B, G)
    a, _ = U.forward(p, feat, G, emulate=False)
    b, _ = U.forward(p, feat, G, emulate=True)
    err = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 0 < err < 3e-2
