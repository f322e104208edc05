"""CPU pins of the FP64 network oracle (oracle/unet.py): the convolution
against a brute-force loop, pool/up against their definitions, and the
analytic gradient against central finite differences."""
import numpy as np

from oracle import unet


def brute_conv(x, w):
    B, nz, ny, nx, cin = x.shape
    y = np.zeros((B, nz, ny, nx, w.shape[0]))
    for b in range(B):
        for z in range(nz):
            for yy in range(ny):
                for xx in range(nx):
                    acc = w[:, 27 * cin].copy()
                    for t, (dz, dy, dx) in enumerate(unet.taps()):
                        q = (z + dz, yy + dy, xx + dx)
                        if 0 <= q[0] < nz and 0 <= q[1] < ny and 0 <= q[2] < nx:
                            acc += w[:, t * cin:(t + 1) * cin] @ x[b, q[0], q[1], q[2]]
                    y[b, z, yy, xx] = acc
    return y


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 3, 4, 5, 8))
    w = rng.normal(size=(3, unet.kpad(8)))
    assert np.allclose(unet.conv3(x, w), brute_conv(x, w), atol=1e-12)
    # tap order: a single tap (dz, dy, dx) = (1, 0, -1) is weight column block t = 2*9 + 1*3 + 0
    w = np.zeros((1, unet.kpad(8)))
    w[0, (2 * 9 + 1 * 3 + 0) * 8 + 5] = 1
    y = unet.conv3(x, w)
    assert np.allclose(y[0, 0, 1, 1, 0], x[0, 1, 1, 0, 5])


def test_conv_backward_is_adjoint():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 3, 3, 4, 8))
    w = rng.normal(size=(4, unet.kpad(8)))
    dy = rng.normal(size=(1, 3, 3, 4, 4))
    dx, dw = unet.conv3_backward(x, w, dy)
    # <conv(x), dy> is linear in x (bias part aside) and in w
    w0 = w.copy(); w0[:, 27 * 8] = 0
    assert np.isclose((unet.conv3(x, w0) * dy).sum(), (x * unet.conv3_backward(x, w0, dy)[0]).sum())
    assert np.isclose((unet.conv3(x, w) * dy).sum(), (w * dw).sum())


def test_pool_and_up():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 3, 5, 4, 2))
    p = unet.pool(x)
    assert p.shape == (1, 2, 3, 2, 2)
    assert np.isclose(p[0, 1, 2, 1, 0], x[0, 2, 4, 2:4, 0].mean())     # 1x1x2 edge block
    assert np.isclose(p[0, 0, 0, 0, 1], x[0, 0:2, 0:2, 0:2, 1].mean())
    # pool_backward and up_backward are the adjoints of pool and up
    g = rng.normal(size=p.shape)
    assert np.isclose((p * g).sum(), (x * unet.pool_backward(g, x.shape)).sum())
    u = unet.up(p, x.shape)
    assert u.shape == x.shape and u[0, 2, 3, 1, 1] == p[0, 1, 1, 0, 1]
    h = rng.normal(size=x.shape)
    assert np.isclose((u * h).sum(), (p * unet.up_backward(h, p.shape)).sum())


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    n = (6, 4, 4)
    B = 2
    params = unet.init_params(5).astype(np.float64)
    params += rng.normal(0, 0.01, params.shape) * (params != 0)
    # give the padding-free bias columns a value too
    for name, w in unet.split_params(params).items():
        ci = w.shape[1]
        w[:, [c for c in range(ci) if c % 27 == 0][:1]] += 0.0
    x = rng.normal(size=(B, n[2], n[1], n[0], 8))
    t = rng.normal(size=(B, n[2], n[1], n[0]))
    mask = rng.random((B, n[2], n[1], n[0])) > 0.3
    loss, g = unet.loss_and_grad(params, x, t, mask)
    live = np.flatnonzero(g != 0)
    pick = rng.choice(live, 40, replace=False)
    eps = 1e-6
    for k in pick:
        pp, pm = params.copy(), params.copy()
        pp[k] += eps; pm[k] -= eps
        lp, _ = unet.loss_and_grad(pp, x, t, mask)
        lm, _ = unet.loss_and_grad(pm, x, t, mask)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0, 1e-40], np.float32)
    r = unet.bf16(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert r[3] == np.float32(-3.140625)
