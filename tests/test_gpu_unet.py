"""The learned-guess U-Net kernels (csrc/unet.hip) against the NumPy
restatement (oracle/unet_ref.py): the MFMA operand layout, and the loss,
output and every parameter gradient of a small network."""
import numpy as np
import pytest

from oracle import unet_ref as R
from fvmic2d_amd.lib import Context
from fvmic2d_amd.unet import Buf, UNet, bind, from_bf16_bits, to_bf16_bits

pytestmark = pytest.mark.gpu


def test_mfma_layout():
    print("CASE mfma 16x16x32 bf16 layout", flush=True)
    ctx = Context()
    lib = bind()
    rng = np.random.default_rng(0)
    A = R.bf16(rng.standard_normal((16, 32)).astype(np.float32))
    B = R.bf16(rng.standard_normal((32, 16)).astype(np.float32))
    dA, dB, dD = Buf(ctx, (512,), np.uint16), Buf(ctx, (512,), np.uint16), Buf(ctx, (256,), np.float32)
    dA.set(to_bf16_bits(A).ravel())
    dB.set(to_bf16_bits(B).ravel())
    assert lib.fvm_nn_mfma_selftest(ctx.h, dA.ptr, dB.ptr, dD.ptr) == 0
    D = dD.get().reshape(16, 16)
    np.testing.assert_allclose(D, A.astype(np.float64) @ B.astype(np.float64), rtol=1e-5, atol=1e-5)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("L,base,B,H,W", [(3, 16, 2, 16, 16), (2, 32, 1, 8, 24)])
def test_unet_forward_backward_matches_reference(L, base, B, H, W):
    print("CASE unet grad L=%d base=%d %dx%dx%d" % (L, base, B, H, W), flush=True)
    ctx = Context()
    net = UNet(ctx, levels=L, base=base, seed=3)
    rng = np.random.default_rng(5)
    X = np.zeros((B, H, W, net.cin), np.float32)
    X[..., :5] = rng.standard_normal((B, H, W, 5))
    T = rng.standard_normal((B, H, W, 3)).astype(np.float32)
    M = (rng.uniform(size=(B, H, W, 3)) > 0.1).astype(np.float32)
    dX = Buf(ctx, X.shape, np.uint16)
    dX.set(to_bf16_bits(X))
    dT, dM = Buf(ctx, T.shape, np.float32), Buf(ctx, M.shape, np.float32)
    dT.set(T)
    dM.set(M)
    net.forward(dX)
    loss = net.loss(dT, dM)
    net.backward()
    params = net.params_numpy()
    rloss, rout, rgrads = R.forward_backward(params, net.arch, X, T, M)
    assert abs(loss - rloss) <= 2e-3 * abs(rloss)
    assert _rel(net.A["out"].get(), rout) < 5e-3
    g = net.grads_numpy()
    for name, (gw, gb) in rgrads.items():
        ew, eb = _rel(g[name][0], gw), _rel(g[name][1], gb)
        assert ew < 2e-2, (name, "W", ew)
        assert eb < 2e-2 or np.linalg.norm(gb) < 1e-12, (name, "b", eb)


def test_adam_step_decreases_loss():
    print("CASE unet adam", flush=True)
    ctx = Context()
    net = UNet(ctx, levels=2, base=16, seed=1)
    rng = np.random.default_rng(2)
    X = np.zeros((2, 16, 16, 16), np.float32)
    X[..., :5] = rng.standard_normal((2, 16, 16, 5))
    T = (0.5 * X[..., :3]).astype(np.float32)
    dX, dT = Buf(ctx, X.shape, np.uint16), Buf(ctx, T.shape, np.float32)
    dX.set(to_bf16_bits(X))
    dT.set(T)
    losses = []
    for _ in range(30):
        net.forward(dX)
        losses.append(net.loss(dT))
        net.backward()
        net.adam(lr=3e-3)
    assert losses[-1] < 0.5 * losses[0], losses
