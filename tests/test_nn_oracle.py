"""Pins the FP64 restatement of the warm-start network (oracle/nn_oracle.py)
that the GPU network is checked against: finite differences of its loss for
parameters of every layer, and closed forms of its building blocks."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import nn_oracle  # noqa: E402


def random_case(rng, N=1, D=8, H=8, W=8):
    n = sum(co * 27 * ci + co for ci, co, _, _ in nn_oracle.LAYERS)
    params = np.zeros(n)
    o = 0
    for ci, co, _, _ in nn_oracle.LAYERS:
        nw = co * 27 * ci
        params[o:o + nw] = rng.normal(size=nw) * np.sqrt(2.0 / (27 * ci))
        o += nw
        params[o:o + co] = rng.normal(size=co) * 0.1
        o += co
    x = rng.normal(size=(N, D, H, W, 16))
    tgt = rng.normal(size=(N, D, H, W, 2))
    mask = (rng.uniform(size=(N, D, H, W)) > 0.3).astype(float)
    return params, x, tgt, mask


def test_identity_tap_reproduces_the_input():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 4, 6, 8, 16))
    w = np.zeros((16, 27, 16))
    w[:, 13, :] = np.eye(16)       # the centre tap
    b = rng.normal(size=16)
    assert np.allclose(nn_oracle.conv(x, w, b, 1), x + b, rtol=0, atol=1e-14)
    # stride 2 samples the even positions
    assert np.allclose(nn_oracle.conv(x, w, b, 2), x[:, ::2, ::2, ::2] + b, rtol=0, atol=1e-14)


def test_shift_tap_moves_the_input():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 4, 4, 4, 16))
    w = np.zeros((16, 27, 16))
    w[:, 9 * 1 + 3 * 1 + 2, :] = np.eye(16)   # tx = 2: reads x + 1
    y = nn_oracle.conv(x, w, np.zeros(16), 1)
    assert np.array_equal(y[:, :, :, :-1], x[:, :, :, 1:])
    assert np.all(y[:, :, :, -1] == 0.0)


def test_upsampling_adjoint():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(1, 2, 3, 4, 5))
    g = rng.normal(size=(1, 4, 6, 8, 5))
    # <up(a), g> == <a, up_bwd(g)>; up_bwd(up(a)) == 8 a
    assert np.isclose((nn_oracle.up(a) * g).sum(), (a * nn_oracle.up_bwd(g)).sum(), rtol=1e-12)
    assert np.allclose(nn_oracle.up_bwd(nn_oracle.up(a)), 8 * a)


def test_conv_backward_is_the_adjoint():
    rng = np.random.default_rng(3)
    for s in (1, 2):
        x = rng.normal(size=(1, 4, 4, 4, 16))
        w = rng.normal(size=(32, 27, 16))
        dy = rng.normal(size=nn_oracle.conv(x, w, np.zeros(32), s).shape)
        dw, db, dx = nn_oracle.conv_bwd(x, w, s, dy)
        v = rng.normal(size=x.shape)
        # <conv(v), dy> == <v, dx>
        assert np.isclose((nn_oracle.conv(v, w, np.zeros(32), s) * dy).sum(), (v * dx).sum(), rtol=1e-10)
        u = rng.normal(size=w.shape)
        assert np.isclose((nn_oracle.conv(x, u, np.zeros(32), s) * dy).sum(), (u * dw).sum(), rtol=1e-10)
        assert np.allclose(db, dy.sum(axis=(0, 1, 2, 3)))


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    params, x, tgt, mask = random_case(rng)
    L, g = nn_oracle.loss_grad(params, x, tgt, mask)
    assert np.isclose(L, nn_oracle.loss(params, x, tgt, mask), rtol=1e-14)
    # a few coordinates in every layer's weights and biases
    offs, o = [], 0
    for ci, co, _, _ in nn_oracle.LAYERS:
        nw = co * 27 * ci
        offs += list(o + rng.choice(nw, size=3, replace=False)) + [o + nw + int(rng.integers(co))]
        o += nw + co
    h = 1e-6
    for k in offs:
        p1, p2 = params.copy(), params.copy()
        p1[k] += h
        p2[k] -= h
        fd = (nn_oracle.loss(p1, x, tgt, mask) - nn_oracle.loss(p2, x, tgt, mask)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_bf16_rounding_matches_its_definition():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 1e-30, 65504.0], dtype=np.float32)
    r = nn_oracle.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0   # tie: to even (1.0)
    assert r[1] == 1.0
    assert r[2] == 1.0 + 2 ** -7                  # above the tie: up
    assert abs(r[3] - (-3.140625)) < 1e-12
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)
