"""The GPU U-Net (BF16 on MFMA, FP32 accumulation) against the FP64
restatement oracle/unet_oracle.py on the same parameters, through the C ABI.

Parameters are rounded to BF16 first, so the two differ only by the GPU's
BF16 activations (and BF16 operand rounding of the backward pass: the
output gradient and the activations enter the weight-gradient MFMAs in
BF16). Stated tolerances: prediction relative L2 <= 1e-2; the gradient of
every layer relative L2 <= 1e-1 with cosine >= 0.995 to the FP64 gradient;
the loss within 2e-2 relative. Grid sizes cover odd (2^k + 1) and even sizes (the
clamped up-sampling edge) and every MFMA channel-block count."""
import importlib.util
import os

import numpy as np
import pytest

import incrmt_amd as ia

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("unet_oracle", os.path.join(ROOT, "oracle", "unet_oracle.py"))
uo = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(uo)


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def sample(ny, nx, seed):
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.linspace(0, 1, ny), np.linspace(0, 1, nx), indexing="ij")
    in3 = np.stack([np.sin(3 * x + 2 * y) + 0.3 * rng.standard_normal((ny, nx)),
                    np.tanh(4 * (np.hypot(x - 0.5, y - 0.5) - 0.3)),
                    np.cos(5 * x * y) + 0.3 * rng.standard_normal((ny, nx))], axis=2)
    target = np.sin(2 * x) * np.cos(3 * y)
    return bf16(in3), target.astype(np.float32)


@pytest.mark.parametrize("ny,nx", [(33, 33), (24, 40), (65, 48)])
def test_forward_matches_fp64(ny, nx):
    net = ia.UNet(nx, ny, seed=7)
    p = bf16(net.get_params())
    p[-1] = 0.05   # the output bias, so the prediction is not centred on zero
    net.set_params(p)
    in3, _ = sample(ny, nx, 1)
    got = net.forward(in3)
    want = uo.forward(p.astype(np.float64), in3.astype(np.float64))
    print(f"forward {ny}x{nx}: relative L2 {rel(got, want):.3e}")
    assert rel(got, want) <= 1e-2


@pytest.mark.parametrize("ny,nx", [(33, 33), (24, 40)])
def test_gradients_match_fp64(ny, nx):
    net = ia.UNet(nx, ny, seed=11)
    rng = np.random.default_rng(5)
    p = net.get_params()
    p = bf16(p * (1 + 0.1 * rng.standard_normal(p.size)).astype(np.float32))
    net.set_params(p)
    in3, target = sample(ny, nx, 2)
    loss, g = net.gradient(in3, target)
    lo, go = uo.loss_and_grad(p.astype(np.float64), in3.astype(np.float64), target.astype(np.float64))
    assert abs(loss - lo) <= 2e-2 * lo
    q = 0
    for name, a, b, c, *_ in uo.LAYERS:
        n = c * 9 * (a + b) + c
        gg, oo = g[q:q + n].astype(np.float64), go[q:q + n]
        if name == "E0a":   # the zero padding channel has no gradient
            W = oo[:c * 9 * (a + b)].reshape(c, 9, a + b)
            assert np.all(gg[:c * 9 * (a + b)].reshape(c, 9, a + b)[:, :, 3] == 0)
            assert np.all(W[:, :, 3] == 0)
        r = rel(gg, oo)
        cos = gg @ oo / (np.linalg.norm(gg) * np.linalg.norm(oo))
        print(f"{name}: gradient relative L2 {r:.3e}, cosine {cos:.6f}")
        assert r <= 1e-1 and cos >= 0.995, name
        q += n


def test_training_reduces_the_loss():
    # Adam on one sample: the GPU trainer drives the loss down
    net = ia.UNet(33, 33, seed=3)
    in3, target = sample(33, 33, 4)
    l0 = net.gradient(in3, target)[0]
    for _ in range(150):
        loss = net.train_step(in3, target, 2e-3)
    print(f"loss {l0:.4e} -> {loss:.4e}")
    assert loss < 0.1 * l0
