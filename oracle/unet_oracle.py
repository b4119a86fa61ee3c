"""FP64 restatement of the learned projection guess's U-Net, for checking the
GPU network (TEST INFRASTRUCTURE ONLY: tests/ use it as the checker; the
product never imports it).

The network has no counterpart in the reference. This file states the
architecture declared in include/incrmt.h in numpy float64, forward and
reverse mode by hand, on the same packed parameter vector:

  level l+1 keeps every other node of level l ((n + 1) // 2 nodes);
  conv3x3 with zero padding, stride 1 or 2 (output node X reads input nodes
  s X + (-1, 0, 1)); ReLU after every layer but the last; vertex-centred
  bilinear up-sampling (fine node x at coarse x / 2; an odd node averages
  its two coarse neighbours, the one past an even-sized edge clamped);
  decoder inputs are the up-sampled tensor followed by the skip tensor;
  loss 0.5 mean((out - target)^2).

Its own correctness is pinned by central finite differences of the loss
(tests/test_unet_oracle.py).
"""
import numpy as np

CH = (16, 32, 48, 64)
N_IN = 4
# name, cin1, cin2, cout, stride, level in, level out, relu
LAYERS = [
    ("E0a", N_IN, 0, CH[0], 1, 0, 0, True), ("E0b", CH[0], 0, CH[0], 1, 0, 0, True),
    ("D1", CH[0], 0, CH[1], 2, 0, 1, True), ("E1", CH[1], 0, CH[1], 1, 1, 1, True),
    ("D2", CH[1], 0, CH[2], 2, 1, 2, True), ("E2", CH[2], 0, CH[2], 1, 2, 2, True),
    ("D3", CH[2], 0, CH[3], 2, 2, 3, True), ("B3", CH[3], 0, CH[3], 1, 3, 3, True),
    ("U2", CH[3], CH[2], CH[2], 1, 2, 2, True), ("U1", CH[2], CH[1], CH[1], 1, 1, 1, True),
    ("U0", CH[1], CH[0], CH[0], 1, 0, 0, True), ("Out", CH[0], 0, 1, 1, 0, 0, False),
]


def nparams():
    return sum(c * 9 * (a + b) + c for _, a, b, c, *_ in LAYERS)


def unpack(p):
    """Packed vector -> list of (W [cout, 9, cin], b [cout])."""
    out, q = [], 0
    for _, a, b, c, *_ in LAYERS:
        n = c * 9 * (a + b)
        W = np.asarray(p[q:q + n], dtype=np.float64).reshape(c, 9, a + b)
        q += n
        bias = np.asarray(p[q:q + c], dtype=np.float64)
        q += c
        out.append((W, bias))
    assert q == len(p)
    return out


def pack(params):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in params])


def sizes(n):
    s = [n]
    for _ in range(3):
        s.append((s[-1] + 1) // 2)
    return s


def _taps():
    return [(t // 3 - 1, t % 3 - 1) for t in range(9)]   # (dy, dx) of tap t


def conv_fwd(X, W, b, stride):
    """X [H, W, Cin] -> [Ho, Wo, Cout]; output node Y reads input s Y + dy."""
    H, Wd, _ = X.shape
    Ho, Wo = (H + 1) // 2 if stride == 2 else H, (Wd + 1) // 2 if stride == 2 else Wd
    Xp = np.pad(X, ((1, 1), (1, 1), (0, 0)))
    out = np.zeros((Ho, Wo, W.shape[0])) + b
    for t, (dy, dx) in enumerate(_taps()):
        patch = Xp[1 + dy:1 + dy + stride * (Ho - 1) + 1:stride, 1 + dx:1 + dx + stride * (Wo - 1) + 1:stride, :]
        out += patch @ W[:, t, :].T
    return out


def conv_bwd(X, W, gout, stride):
    """Gradients (dX, dW, db) of conv_fwd for the output gradient gout."""
    H, Wd, _ = X.shape
    Ho, Wo, _ = gout.shape
    Xp = np.pad(X, ((1, 1), (1, 1), (0, 0)))
    dXp = np.zeros_like(Xp)
    dW = np.zeros_like(W)
    g2 = gout.reshape(-1, gout.shape[2])
    for t, (dy, dx) in enumerate(_taps()):
        sl = (slice(1 + dy, 1 + dy + stride * (Ho - 1) + 1, stride), slice(1 + dx, 1 + dx + stride * (Wo - 1) + 1, stride))
        patch = Xp[sl]
        dW[:, t, :] = g2.T @ patch.reshape(-1, patch.shape[2])
        dXp[sl] += gout @ W[:, t, :]
    return dXp[1:-1, 1:-1], dW, g2.sum(axis=0)


def up_matrix(nf, nc):
    """Vertex-centred linear interpolation from nc coarse to nf fine nodes."""
    U = np.zeros((nf, nc))
    for x in range(nf):
        if x % 2 == 0:
            U[x, x // 2] += 1.0
        else:
            U[x, x // 2] += 0.5
            U[x, min(x // 2 + 1, nc - 1)] += 0.5
    return U


def upsample(C, nyf, nxf):
    Uy, Ux = up_matrix(nyf, C.shape[0]), up_matrix(nxf, C.shape[1])
    return np.einsum("yY,YXc,xX->yxc", Uy, C, Ux)


def upsample_bwd(G, nyc, nxc):
    Uy, Ux = up_matrix(G.shape[0], nyc), up_matrix(G.shape[1], nxc)
    return np.einsum("yY,yxc,xX->YXc", Uy, G, Ux)


def forward(p, in3, keep=False):
    """in3 [ny, nx, 3] -> prediction [ny, nx] (and the saved activations)."""
    P = unpack(p)
    ny, nx, _ = in3.shape
    X0 = np.concatenate([np.asarray(in3, dtype=np.float64), np.zeros((ny, nx, 1))], axis=2)
    sy, sx = sizes(ny), sizes(nx)
    act, pre, inp = {}, {}, {}

    def layer(i, x1, x2=None):
        name, _, _, _, stride, _, _, relu = LAYERS[i]
        x = x1 if x2 is None else np.concatenate([x1, x2], axis=2)
        inp[name] = x
        z = conv_fwd(x, P[i][0], P[i][1], stride)
        pre[name] = z
        a = np.maximum(z, 0) if relu else z
        act[name] = a
        return a

    a = layer(0, X0)
    e0 = layer(1, a)
    a = layer(2, e0)
    e1 = layer(3, a)
    a = layer(4, e1)
    e2 = layer(5, a)
    a = layer(6, e2)
    b3 = layer(7, a)
    d2 = layer(8, upsample(b3, sy[2], sx[2]), e2)
    d1 = layer(9, upsample(d2, sy[1], sx[1]), e1)
    d0 = layer(10, upsample(d1, sy[0], sx[0]), e0)
    out = layer(11, d0)[:, :, 0]
    if keep:
        return out, (act, pre, inp, P)
    return out


def loss_and_grad(p, in3, target):
    out, (act, pre, inp, P) = forward(p, in3, keep=True)
    ny, nx = out.shape
    N = nx * ny
    d = out - np.asarray(target, dtype=np.float64)
    loss = 0.5 * np.sum(d * d) / N
    grads = [None] * len(LAYERS)
    gact = {"Out": (d / N)[:, :, None]}

    def back(i):
        name, c1, c2, _, stride, _, _, relu = LAYERS[i]
        g = gact.pop(name)
        if relu:
            g = g * (pre[name] > 0)
        dX, dW, db = conv_bwd(inp[name], P[i][0], g, stride)
        grads[i] = (dW, db)
        return dX[:, :, :c1], dX[:, :, c1:]

    def add(name, g):
        gact[name] = gact[name] + g if name in gact else g

    g, _ = back(11)
    add("U0", g)
    for i, (src, skip) in zip((10, 9, 8), (("U1", "E0b"), ("U2", "E1"), ("B3", "E2"))):
        gu, gs = back(i)
        add(skip, gs)
        add(src, upsample_bwd(gu, act[src].shape[0], act[src].shape[1]))
    for i, prev in ((7, "D3"), (6, "E2"), (5, "D2"), (4, "E1"), (3, "D1"), (2, "E0b"), (1, "E0a")):
        g, _ = back(i)
        add(prev, g)
    back(0)
    return loss, pack(grads)
