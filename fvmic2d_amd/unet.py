"""The learned Stokes initial guess: a U-Net run by the hand-written gfx950
kernels of csrc/unet.hip (bf16 activations and weights, fp32 accumulation on
the matrix cores, fp32 Adam master weights).

Architecture (levels L, base channels c): encoder level i has two 3x3
conv+ReLU layers with c*2^i channels and a 2x2 average pool between levels;
decoder level i upsamples (nearest, x2), concatenates the encoder skip and
applies two 3x3 conv+ReLU layers with c*2^i channels; a 1x1 convolution maps
to the output channels.  Tensors are NHWC; channel counts are padded to
multiples of 16 (the input's 5 fields and the output's 3 fields sit in the
first channels, the rest are zero).
"""
import ctypes as C

import numpy as np

from .lib import Context, load

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F = C.c_float

_SIGS = {
    "fvm_nn_mfma_selftest": [P, P, P, P],
    "fvm_nn_conv_fwd": [P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, P, P],
    "fvm_nn_conv_dgrad": [P, P, P, P, I32, I32, I32, I32, I32, I32, P],
    "fvm_nn_wgrad_workspace": [I32, I32, I32, I32, I32, I32],
    "fvm_nn_conv_wgrad": [P, P, P, I32, I32, I32, I32, I32, I32, P, P, P],
    "fvm_nn_relu_bwd": [P, I64, P, P, P],
    "fvm_nn_pool_fwd": [P, P, I32, I32, I32, I32, P],
    "fvm_nn_pool_bwd": [P, P, I32, I32, I32, I32, P],
    "fvm_nn_up_fwd": [P, P, I32, I32, I32, I32, P, I32, I32],
    "fvm_nn_up_bwd": [P, P, I32, I32, I32, I32, I32, I32, P],
    "fvm_nn_copy_channels": [P, P, I64, I32, I32, I32, P, I32, I32],
    "fvm_nn_add_channels": [P, P, I64, I32, I32, I32, P],
    "fvm_nn_mse": [P, P, P, P, I64, I32, I32, F, P, P, P],
    "fvm_nn_adam": [P, I64, P, P, P, P, P, F, F, F, F, I32, F],
    "fvm_nn_f32_to_bf16": [P, I64, P, P],
}


def bind():
    lib = load()
    for n, a in _SIGS.items():
        f = getattr(lib, n)
        f.argtypes = a
        f.restype = C.c_size_t if n == "fvm_nn_wgrad_workspace" else C.c_int
    return lib


def to_bf16_bits(x):
    """float32 -> bfloat16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


class Buf:
    """A device allocation holding a numpy-typed array (bf16 held as uint16)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx, self.shape, self.dtype = ctx, tuple(shape), np.dtype(dtype)
        self.nbytes = max(1, int(np.prod(self.shape)) * self.dtype.itemsize)
        self.ptr = ctx.alloc(self.nbytes)
        self.zero()

    def zero(self):
        self.ctx.call("fvm_memset", P(self.ptr), 0, self.nbytes)

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == int(np.prod(self.shape)), (a.shape, self.shape)
        self.ctx.htod(self.ptr, a)

    def get(self):
        return self.ctx.dtoh(self.ptr, self.shape, self.dtype)

    def __del__(self):
        try:
            self.ctx.free(self.ptr)
        except Exception:
            pass


def _pad16(c):
    return (c + 15) // 16 * 16


class UNet:
    """Parameters, activations and the layer sequence of the learned guess."""

    def __init__(self, ctx=None, levels=4, base=32, cin=5, cout=3, seed=0):
        self.ctx = ctx or Context()
        self.lib = bind()
        self.L, self.c = levels, base
        self.cin_real, self.cout_real = cin, cout
        self.cin, self.cout = _pad16(cin), _pad16(cout)
        self.arch = (levels, base, self.cin, self.cout, cout)
        rng = np.random.default_rng(seed)
        self.layers = []  # (name, K, Cin, Cout, relu)
        ch = [base * 2 ** i for i in range(levels)]
        prev = self.cin
        for i in range(levels):
            self.layers.append(("e%d_1" % i, 3, prev, ch[i], True))
            self.layers.append(("e%d_2" % i, 3, ch[i], ch[i], True))
            prev = ch[i]
        for i in range(levels - 2, -1, -1):
            self.layers.append(("d%d_1" % i, 3, ch[i + 1] + ch[i], ch[i], True))
            self.layers.append(("d%d_2" % i, 3, ch[i], ch[i], True))
        self.layers.append(("out", 1, ch[0], self.cout, False))
        self.ch = ch
        self.p = {}
        for name, K, ci, co, relu in self.layers:
            fan_in = K * K * ci
            W = (rng.standard_normal((K * K, co, ci)) * np.sqrt(2.0 / fan_in)).astype(np.float32)
            if name == "e0_1":
                W[:, :, self.cin_real:] = 0.0
            if name == "out":
                W *= 0.1
                W[:, self.cout_real:, :] = 0.0
            b = np.zeros(co, np.float32)
            ent = {"K": K, "ci": ci, "co": co, "relu": relu,
                   "w": Buf(self.ctx, W.shape, np.float32), "wb": Buf(self.ctx, W.shape, np.uint16),
                   "b": Buf(self.ctx, (co,), np.float32),
                   "gw": Buf(self.ctx, W.shape, np.float32), "gb": Buf(self.ctx, (co,), np.float32),
                   "mw": Buf(self.ctx, W.shape, np.float32), "vw": Buf(self.ctx, W.shape, np.float32),
                   "mb": Buf(self.ctx, (co,), np.float32), "vb": Buf(self.ctx, (co,), np.float32)}
            ent["w"].set(W)
            ent["wb"].set(to_bf16_bits(W))
            ent["b"].set(b)
            self.p[name] = ent
        self.adam_step = 0
        self.shape = None

    # ---- parameters as numpy (tests, checkpoints) ----
    def params_numpy(self):
        return {k: (v["w"].get(), v["b"].get()) for k, v in self.p.items()}

    def grads_numpy(self):
        return {k: (v["gw"].get(), v["gb"].get()) for k, v in self.p.items()}

    def load_params(self, params):
        for k, (W, b) in params.items():
            self.p[k]["w"].set(W)
            self.p[k]["wb"].set(to_bf16_bits(W))
            self.p[k]["b"].set(b)

    # ---- buffers for one batch shape ----
    def _alloc(self, B, H, W):
        if self.shape == (B, H, W):
            return
        assert H % 2 ** (self.L - 1) == 0 and W % 2 ** (self.L - 1) == 0, "H, W must divide by 2^(L-1)"
        ctx = self.ctx
        self.shape = (B, H, W)
        bf = np.uint16
        self.A = {}
        for i in range(self.L):
            h, w = H >> i, W >> i
            cin = self.cin if i == 0 else self.ch[i - 1]
            self.A["x%d" % i] = Buf(ctx, (B, h, w, cin), bf)
            self.A["h%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
            self.A["e%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
            self.A["de%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
            self.A["g1_%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
            self.A["g2_%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
            self.A["dx%d" % i] = Buf(ctx, (B, h, w, cin), bf)
        for i in range(self.L - 1):
            h, w = H >> i, W >> i
            cc = self.ch[i + 1] + self.ch[i]
            self.A["cat%d" % i] = Buf(ctx, (B, h, w, cc), bf)
            self.A["dcat%d" % i] = Buf(ctx, (B, h, w, cc), bf)
            self.A["dh%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
            self.A["d%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
            self.A["dd%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
            self.A["gz%d" % i] = Buf(ctx, (B, h, w, self.ch[i]), bf)
        self.A["out"] = Buf(ctx, (B, H, W, self.cout), np.float32)
        self.A["dout"] = Buf(ctx, (B, H, W, self.cout), bf)
        self.A["ddec_top"] = Buf(ctx, (B, H, W, self.ch[0]), bf)
        wmax = max(K * K * ci * co for _, K, ci, co, _ in self.layers)
        self.A["wflip"] = Buf(ctx, (wmax,), bf)
        ws = max(self.lib.fvm_nn_wgrad_workspace(B, H >> self._lvl(n), W >> self._lvl(n), ci, co, K)
                 for n, K, ci, co, _ in self.layers)
        self.A["work"] = Buf(ctx, (ws // 4 + 1,), np.float32)
        npix = B * H * W
        self.A["lossw"] = Buf(ctx, ((npix * self.cout + 255) // 256 + 1,), np.float32)
        self.A["loss"] = Buf(ctx, (1,), np.float32)

    def _lvl(self, name):
        return 0 if name == "out" else int(name[1])

    def _call(self, fn, *a):
        rc = getattr(self.lib, fn)(self.ctx.h, *a)
        if rc != 0:
            self.ctx._check(rc)

    def _conv(self, name, X, Y, B, H, W, Yf=None):
        e = self.p[name]
        self._call("fvm_nn_conv_fwd", P(X.ptr), P(e["wb"].ptr), P(e["b"].ptr), B, H, W, e["ci"], e["co"], e["K"],
                   1 if e["relu"] else 0, self.cout_real if name == "out" else e["co"],
                   P(Y.ptr) if Y is not None else None, P(Yf.ptr) if Yf is not None else None)

    # ---- forward ----
    def forward(self, X):
        """X: Buf (B, H, W, cin) bf16 -> self.A['out'] (fp32, B H W cout)."""
        B, H, W, _ = X.shape
        self._alloc(B, H, W)
        A = self.A
        self._call("fvm_nn_copy_channels", P(X.ptr), I64(B * H * W), self.cin, self.cin, 0, P(A["x0"].ptr),
                   self.cin, 0)
        for i in range(self.L):
            h, w = H >> i, W >> i
            self._conv("e%d_1" % i, A["x%d" % i], A["h%d" % i], B, h, w)
            self._conv("e%d_2" % i, A["h%d" % i], A["e%d" % i], B, h, w)
            if i < self.L - 1:
                self._call("fvm_nn_pool_fwd", P(A["e%d" % i].ptr), B, h, w, self.ch[i], P(A["x%d" % (i + 1)].ptr))
        d = A["e%d" % (self.L - 1)]
        for i in range(self.L - 2, -1, -1):
            h, w = H >> i, W >> i
            cc = self.ch[i + 1] + self.ch[i]
            self._call("fvm_nn_up_fwd", P(d.ptr), B, h // 2, w // 2, self.ch[i + 1], P(A["cat%d" % i].ptr), cc, 0)
            self._call("fvm_nn_copy_channels", P(A["e%d" % i].ptr), I64(B * h * w), self.ch[i], self.ch[i], 0,
                       P(A["cat%d" % i].ptr), cc, self.ch[i + 1])
            self._conv("d%d_1" % i, A["cat%d" % i], A["dh%d" % i], B, h, w)
            self._conv("d%d_2" % i, A["dh%d" % i], A["d%d" % i], B, h, w)
            d = A["d%d" % i]
        self._conv("out", d, None, B, H, W, Yf=A["out"])
        return A["out"]

    # ---- loss ----
    def loss(self, T, mask=None):
        """Masked MSE over the real output channels against T (Buf fp32, B H W creal);
        leaves d(loss)/d(out) in A['dout'] and returns the loss (synchronises)."""
        B, H, W = self.shape
        npix = B * H * W
        self._call("fvm_nn_mse", P(self.A["out"].ptr), P(T.ptr), P(mask.ptr) if mask is not None else None,
                   I64(npix), self.cout, self.cout_real, F(1.0 / (npix * self.cout_real)), P(self.A["dout"].ptr),
                   P(self.A["lossw"].ptr), P(self.A["loss"].ptr))
        return float(self.A["loss"].get()[0])

    # ---- backward ----
    def _wgrad(self, name, X, dY, B, H, W):
        e = self.p[name]
        self._call("fvm_nn_conv_wgrad", P(X.ptr), P(dY.ptr), B, H, W, e["ci"], e["co"], e["K"],
                   P(self.A["work"].ptr), P(e["gw"].ptr), P(e["gb"].ptr))

    def _dgrad(self, name, dY, dX, B, H, W):
        e = self.p[name]
        self._call("fvm_nn_conv_dgrad", P(dY.ptr), P(e["wb"].ptr), P(self.A["wflip"].ptr), B, H, W, e["ci"],
                   e["co"], e["K"], P(dX.ptr))

    def _relu_bwd(self, y, dy, dz):
        self._call("fvm_nn_relu_bwd", I64(int(np.prod(y.shape))), P(y.ptr), P(dy.ptr), P(dz.ptr))

    def backward(self):
        B, H, W = self.shape
        A, L, ch = self.A, self.L, self.ch
        d0 = A["d0"] if L > 1 else A["e0"]
        self._wgrad("out", d0, A["dout"], B, H, W)
        dd = A["ddec_top"]
        self._dgrad("out", A["dout"], dd, B, H, W)
        for i in range(0, L - 1):
            h, w = H >> i, W >> i
            cc = ch[i + 1] + ch[i]
            self._relu_bwd(A["d%d" % i], dd, A["gz%d" % i])
            self._wgrad("d%d_2" % i, A["dh%d" % i], A["gz%d" % i], B, h, w)
            self._dgrad("d%d_2" % i, A["gz%d" % i], A["dd%d" % i], B, h, w)
            self._relu_bwd(A["dh%d" % i], A["dd%d" % i], A["gz%d" % i])
            self._wgrad("d%d_1" % i, A["cat%d" % i], A["gz%d" % i], B, h, w)
            self._dgrad("d%d_1" % i, A["gz%d" % i], A["dcat%d" % i], B, h, w)
            # skip part -> de_i; upsampled part -> gradient of the level below
            self._call("fvm_nn_copy_channels", P(A["dcat%d" % i].ptr), I64(B * h * w), ch[i], cc, ch[i + 1],
                       P(A["de%d" % i].ptr), ch[i], 0)
            target = A["de%d" % (L - 1)] if i + 1 == L - 1 else A["g1_%d" % (i + 1)]
            self._call("fvm_nn_up_bwd", P(A["dcat%d" % i].ptr), B, h // 2, w // 2, ch[i + 1], cc, 0, P(target.ptr))
            dd = target
        # encoder, bottom up: de_{L-1} holds the bottleneck gradient
        dnext = None
        for i in range(L - 1, -1, -1):
            h, w = H >> i, W >> i
            de = A["de%d" % i]
            if i < L - 1:
                self._call("fvm_nn_add_channels", P(dnext.ptr), I64(B * h * w), ch[i], ch[i], 0, P(de.ptr))
            self._relu_bwd(A["e%d" % i], de, A["g1_%d" % i])
            self._wgrad("e%d_2" % i, A["h%d" % i], A["g1_%d" % i], B, h, w)
            self._dgrad("e%d_2" % i, A["g1_%d" % i], A["g2_%d" % i], B, h, w)
            self._relu_bwd(A["h%d" % i], A["g2_%d" % i], A["g1_%d" % i])
            self._wgrad("e%d_1" % i, A["x%d" % i], A["g1_%d" % i], B, h, w)
            if i > 0:
                self._dgrad("e%d_1" % i, A["g1_%d" % i], A["dx%d" % i], B, h, w)
                self._call("fvm_nn_pool_bwd", P(A["dx%d" % i].ptr), B, h * 2, w * 2, ch[i - 1],
                           P(A["g2_%d" % (i - 1)].ptr))
                dnext = A["g2_%d" % (i - 1)]

    # ---- optimizer ----
    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.adam_step += 1
        for e in self.p.values():
            n = int(np.prod(e["w"].shape))
            self._call("fvm_nn_adam", I64(n), P(e["w"].ptr), P(e["gw"].ptr), P(e["mw"].ptr), P(e["vw"].ptr),
                       P(e["wb"].ptr), F(lr), F(b1), F(b2), F(eps), self.adam_step, F(wd))
            self._call("fvm_nn_adam", I64(e["co"]), P(e["b"].ptr), P(e["gb"].ptr), P(e["mb"].ptr), P(e["vb"].ptr),
                       None, F(lr), F(b1), F(b2), F(eps), self.adam_step, F(0.0))
