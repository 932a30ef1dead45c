"""NumPy restatement of the learned-guess U-Net (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker for the GPU kernels in
fvmic2d_amd/csrc/unet.hip, imported by tests/.  It follows the same
architecture and the same bfloat16 rounding points (weights, every stored
activation and activation gradient rounded to bf16; products accumulated in
high precision), so the GPU and this code differ only by the order of the
fp32 accumulations.

Layout: NHWC arrays, weights [tap][Cout][Cin] (tap = ky*K + kx).
"""
import numpy as np


def bf16(x):
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def conv(X, W, b, K, relu):
    """Y[p][co] = act(b + sum_tap X[p + off(tap)] . W[tap][co])."""
    B, H, Wd, Cin = X.shape
    Cout = W.shape[1]
    h = K // 2
    Xp = np.zeros((B, H + 2 * h, Wd + 2 * h, Cin))
    Xp[:, h:h + H, h:h + Wd, :] = X
    Y = np.zeros((B, H, Wd, Cout))
    for tap in range(K * K):
        ky, kx = divmod(tap, K)
        Y += Xp[:, ky:ky + H, kx:kx + Wd, :] @ W[tap].T.astype(np.float64)
    if b is not None:
        Y += b
    if relu:
        Y = np.maximum(Y, 0.0)
    return Y


def conv_dgrad(dY, W, K):
    B, H, Wd, Cout = dY.shape
    Cin = W.shape[2]
    h = K // 2
    dXp = np.zeros((B, H + 2 * h, Wd + 2 * h, Cin))
    for tap in range(K * K):
        ky, kx = divmod(tap, K)
        dXp[:, ky:ky + H, kx:kx + Wd, :] += dY @ W[tap].astype(np.float64)
    return dXp[:, h:h + H, h:h + Wd, :]


def conv_wgrad(X, dY, K):
    B, H, Wd, Cin = X.shape
    Cout = dY.shape[3]
    h = K // 2
    Xp = np.zeros((B, H + 2 * h, Wd + 2 * h, Cin))
    Xp[:, h:h + H, h:h + Wd, :] = X
    dW = np.zeros((K * K, Cout, Cin))
    for tap in range(K * K):
        ky, kx = divmod(tap, K)
        xs = Xp[:, ky:ky + H, kx:kx + Wd, :].reshape(-1, Cin)
        dW[tap] = dY.reshape(-1, Cout).T @ xs
    return dW, dY.reshape(-1, Cout).sum(axis=0)


def pool(X):
    B, H, W, C = X.shape
    return 0.25 * ((X[:, 0::2, 0::2] + X[:, 0::2, 1::2]) + (X[:, 1::2, 0::2] + X[:, 1::2, 1::2]))


def pool_bwd(dY):
    return 0.25 * np.repeat(np.repeat(dY, 2, axis=1), 2, axis=2)


def up(X):
    return np.repeat(np.repeat(X, 2, axis=1), 2, axis=2)


def up_bwd(dY):
    return (dY[:, 0::2, 0::2] + dY[:, 0::2, 1::2]) + (dY[:, 1::2, 0::2] + dY[:, 1::2, 1::2])


def forward_backward(params, arch, X, T, mask, exact=False):
    """Loss and gradients of the U-Net (same naming as fvmic2d_amd.unet.UNet).

    params: dict name -> (W [tap][Cout][Cin], b [Cout]) in float32 (the fp32
    master values; the forward pass uses their bf16 rounding).  arch: (levels,
    base channels, input channels, output channels, real output channels).
    exact=True drops every rounding point (float64 throughout), which makes the
    loss a smooth function of the parameters for a finite-difference check."""
    L, c, cin, cout, creal = arch
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else bf16
    P = {k: (rnd(w).astype(np.float64), b.astype(np.float64)) for k, (w, b) in params.items()}
    r = lambda a: rnd(a).astype(np.float64)
    acts = {}
    x = r(X)
    enc = []
    for i in range(L):
        h1 = r(conv(x, *P["e%d_1" % i], 3, True))
        e = r(conv(h1, *P["e%d_2" % i], 3, True))
        acts["e%d" % i] = (x, h1, e)
        enc.append(e)
        if i < L - 1:
            x = r(pool(e))
    d = enc[L - 1]
    dec = {}
    for i in range(L - 2, -1, -1):
        cat = np.concatenate([up(d), enc[i]], axis=3)
        h1 = r(conv(cat, *P["d%d_1" % i], 3, True))
        d = r(conv(h1, *P["d%d_2" % i], 3, True))
        dec[i] = (cat, h1, d)
    out = conv(d, *P["out"], 1, False)
    if not exact:
        out = out.astype(np.float32).astype(np.float64)
    npix = X.shape[0] * X.shape[1] * X.shape[2]
    diff = (out[..., :creal] - T) * mask
    inv = 1.0 / (npix * creal)
    loss = float((diff * diff).sum() * inv)
    dout = np.zeros_like(out)
    dout[..., :creal] = 2.0 * diff * inv
    dout = r(dout)
    grads = {}
    gW, gb = conv_wgrad(d, dout, 1)
    grads["out"] = (gW, gb)
    dd = r(conv_dgrad(dout, P["out"][0], 1))
    denc = {i: None for i in range(L)}
    for i in range(0, L - 1):
        cat, h1, di = dec[i]
        dz2 = dd * (di > 0)
        gW2, gb2 = conv_wgrad(h1, dz2, 3)
        dh1 = r(conv_dgrad(dz2, P["d%d_2" % i][0], 3))
        dz1 = dh1 * (h1 > 0)
        gW1, gb1 = conv_wgrad(cat, dz1, 3)
        dcat = r(conv_dgrad(dz1, P["d%d_1" % i][0], 3))
        grads["d%d_2" % i] = (gW2, gb2)
        grads["d%d_1" % i] = (gW1, gb1)
        cu = dcat.shape[3] - enc[i].shape[3]
        denc[i] = dcat[..., cu:]
        dd = r(up_bwd(dcat[..., :cu]))
    # dd is now the gradient of enc[L-1] (the bottleneck)
    dnext = dd
    for i in range(L - 1, -1, -1):
        x, h1, e = acts["e%d" % i]
        de = dnext if i == L - 1 else r(denc[i] + dnext)
        dz2 = de * (e > 0)
        gW2, gb2 = conv_wgrad(h1, dz2, 3)
        dh1 = r(conv_dgrad(dz2, P["e%d_2" % i][0], 3))
        dz1 = dh1 * (h1 > 0)
        gW1, gb1 = conv_wgrad(x, dz1, 3)
        grads["e%d_2" % i] = (gW2, gb2)
        grads["e%d_1" % i] = (gW1, gb1)
        if i > 0:
            dx = r(conv_dgrad(dz1, P["e%d_1" % i][0], 3))
            dnext = r(pool_bwd(dx))
    return loss, out, grads
