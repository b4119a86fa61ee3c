"""The FP64 U-Net restatement (oracle/unet_oracle.py) pinned by what the
mathematics fixes: its reverse-mode gradients equal central finite
differences of its loss for parameters of every layer, the up-sampling
interpolates linear data exactly and its adjoint is the transpose, and the
stride-2 convolution reads the kept nodes."""
import importlib.util
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("unet_oracle", os.path.join(ROOT, "oracle", "unet_oracle.py"))
uo = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(uo)


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for _, a, b, c, *_ in uo.LAYERS:
        W = rng.standard_normal((c, 9, a + b)) * np.sqrt(2.0 / (9 * (a + b))) * scale
        parts.append(np.concatenate([W.ravel(), 0.1 * rng.standard_normal(c)]))
    return np.concatenate(parts)


@pytest.mark.parametrize("ny,nx", [(9, 9), (10, 7)])
def test_gradients_match_finite_differences(ny, nx):
    rng = np.random.default_rng(3)
    p = random_params(1)
    x = rng.standard_normal((ny, nx, 3))
    t = rng.standard_normal((ny, nx))
    loss, g = uo.loss_and_grad(p, x, t)
    assert np.isclose(loss, 0.5 * np.mean((uo.forward(p, x) - t) ** 2))
    # parameters from every layer, weights and biases
    idx, q = [], 0
    for _, a, b, c, *_ in uo.LAYERS:
        n = c * 9 * (a + b)
        idx += list(rng.integers(q, q + n, 4)) + [q + n + int(rng.integers(0, c))]
        q += n + c
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (0.5 * np.mean((uo.forward(pp, x) - t) ** 2) - 0.5 * np.mean((uo.forward(pm, x) - t) ** 2)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_upsampling_is_exact_on_linear_data_and_adjoint():
    for nf in (9, 17, 10, 16):
        nc = (nf + 1) // 2
        U = uo.up_matrix(nf, nc)
        xc = np.arange(nc) * 2.0
        xf = U @ xc
        if nf % 2:   # odd: every fine node lies between coarse ones
            assert np.allclose(xf, np.arange(nf))
        else:        # even: the last node copies its clamped neighbour
            assert np.allclose(xf[:-1], np.arange(nf - 1)) and xf[-1] == xc[-1]
        assert np.allclose(U.sum(axis=1), 1.0)
    rng = np.random.default_rng(0)
    C = rng.standard_normal((5, 4, 3))
    G = rng.standard_normal((9, 7, 3))
    assert np.isclose(np.sum(uo.upsample(C, 9, 7) * G), np.sum(C * uo.upsample_bwd(G, 5, 4)))


def test_stride_two_convolution_reads_kept_nodes():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((9, 7, 2))
    W = np.zeros((1, 9, 2))
    W[0, 4, 1] = 1.0          # centre tap, second channel
    out = uo.conv_fwd(X, W, np.zeros(1), 2)
    assert out.shape == (5, 4, 1)
    assert np.array_equal(out[:, :, 0], X[::2, ::2, 1])
    W[0, 4, 1] = 0.0
    W[0, 5, 0] = 1.0          # tap (dy, dx) = (0, +1)
    out = uo.conv_fwd(X, W, np.zeros(1), 2)
    assert np.array_equal(out[:, :3, 0], X[::2, 1::2, 0])
    assert np.all(out[:, 3, 0] == 0)   # beyond the edge: zero padding


def test_parameter_count_matches_the_header():
    assert uo.nparams() == sum(c * 9 * (a + b) + c for _, a, b, c, *_ in uo.LAYERS)
    assert uo.nparams() == 194673
