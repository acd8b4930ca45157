"""FP64 restatement of the learned warm start's network, for tests only.

This is test infrastructure (the checker), not part of the product: it
restates, in numpy float64, the 3-D convolutional encoder-decoder that the
library runs in BF16 on the matrix cores (am-cfd_amd/csrc/nn.hip,
warmstart.cpp), so that the GPU forward pass, loss and every parameter
gradient can be compared with it.  The network has no counterpart in the
reference (the reference has no learned component); what pins this file:
  - finite differences of its own loss (tests/test_nn_oracle.py),
  - closed forms: an identity tap reproduces its input, nearest upsampling
    followed by its adjoint multiplies by 8.

Layout: tensors (N, D, H, W, C), channels last; weights per layer
[co][tap][ci] with tap = 9 tz + 3 ty + tx over the 3x3x3 stencil, padding 1;
the flat parameter vector holds, per layer, the weights then the biases.
"""
import numpy as np

# (cin, cout, stride, relu) per layer
LAYERS = [(16, 16, 1, 1), (16, 32, 2, 1), (32, 64, 2, 1), (64, 128, 2, 1), (128, 128, 1, 1),
          (192, 64, 1, 1), (96, 32, 1, 1), (48, 16, 1, 1), (16, 16, 1, 0)]
NOUT = 2


def unpack(params):
    out, o = [], 0
    for ci, co, _, _ in LAYERS:
        nw = co * 27 * ci
        w = params[o:o + nw].reshape(co, 27, ci).astype(np.float64)
        o += nw
        b = params[o:o + co].astype(np.float64)
        o += co
        out.append((w, b))
    assert o == params.size
    return out


def _slices(D, H, W, s, tz, ty, tx):
    Do, Ho, Wo = (D - 1) // s + 1, (H - 1) // s + 1, (W - 1) // s + 1
    return (slice(None), slice(tz, tz + s * (Do - 1) + 1, s), slice(ty, ty + s * (Ho - 1) + 1, s),
            slice(tx, tx + s * (Wo - 1) + 1, s)), (Do, Ho, Wo)


def conv(x, w, b, s):
    N, D, H, W, Ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = None
    for tap in range(27):
        sl, (Do, Ho, Wo) = _slices(D, H, W, s, tap // 9, (tap // 3) % 3, tap % 3)
        t = xp[sl] @ w[:, tap, :].T
        y = t if y is None else y + t
    return y + b


def conv_bwd(x, w, s, dy):
    """dW, db, dx of conv(x, w, b, s) for the upstream gradient dy."""
    N, D, H, W, Ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for tap in range(27):
        sl, _ = _slices(D, H, W, s, tap // 9, (tap // 3) % 3, tap % 3)
        dw[:, tap, :] = np.einsum("ndhwo,ndhwc->oc", dy, xp[sl])
        dxp[sl] += dy @ w[:, tap, :]
    return dw, dy.sum(axis=(0, 1, 2, 3)), dxp[:, 1:-1, 1:-1, 1:-1, :]


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def up_bwd(g):
    N, D, H, W, C = g.shape
    return g.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def forward(params, x, emulate=False):
    """Activations of every layer (post-activation) and the head output.
    With emulate, activations are rounded to bfloat16 where the device
    stores them in bfloat16 (every layer output but the head's)."""
    P = unpack(params)
    z, a = [None] * 9, [None] * 9
    inputs = [None] * 9

    def layer(l, inp):
        ci, co, s, relu = LAYERS[l]
        inputs[l] = inp
        z[l] = conv(inp, P[l][0], P[l][1], s)
        a[l] = np.maximum(z[l], 0.0) if relu else z[l]
        if emulate and relu:
            a[l] = bf16(a[l])
        return a[l]

    layer(0, x)
    layer(1, a[0])
    layer(2, a[1])
    layer(3, a[2])
    layer(4, a[3])
    layer(5, np.concatenate([up(a[4]), a[2]], axis=-1))
    layer(6, np.concatenate([up(a[5]), a[1]], axis=-1))
    layer(7, np.concatenate([up(a[6]), a[0]], axis=-1))
    layer(8, a[7])
    return a, z, inputs


def loss(params, x, target, mask):
    a, _, _ = forward(params, x)
    d = a[8][..., :NOUT] - target
    cnt = mask.sum() * NOUT
    return float((mask[..., None] * d * d).sum() / cnt)


def loss_grad(params, x, target, mask, emulate=False):
    """The loss and the flat gradient (same layout as the parameters).  With
    emulate, activations and the gradients entering each layer's backward
    products are rounded to bfloat16 as on the device."""
    P = unpack(params)
    a, z, inputs = forward(params, x, emulate)
    d = a[8][..., :NOUT] - target
    cnt = mask.sum() * NOUT
    L = float((mask[..., None] * d * d).sum() / cnt)
    da = [None] * 9
    da[8] = np.zeros_like(a[8])
    da[8][..., :NOUT] = 2.0 * mask[..., None] * d / cnt
    grads = [None] * 9
    acc = {l: np.zeros_like(a[l]) for l in range(8)}

    def back(l, g):
        ci, co, s, relu = LAYERS[l]
        dz = g * (a[l] > 0) if relu else g
        if emulate:
            dz = bf16(dz)
        dw, db, dx = conv_bwd(inputs[l], P[l][0], s, dz)
        grads[l] = (dw, db)
        return dx

    acc[7] += back(8, da[8])
    dx = back(7, acc[7])
    acc[6] += up_bwd(dx[..., :32]); acc[0] += dx[..., 32:]
    dx = back(6, acc[6])
    acc[5] += up_bwd(dx[..., :64]); acc[1] += dx[..., 64:]
    dx = back(5, acc[5])
    acc[4] += up_bwd(dx[..., :128]); acc[2] += dx[..., 128:]
    acc[3] += back(4, acc[4])
    acc[2] += back(3, acc[3])
    acc[1] += back(2, acc[2])
    acc[0] += back(1, acc[1])
    back(0, acc[0])
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return L, flat


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
