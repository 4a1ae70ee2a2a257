"""Pins the oracle's two-headed StatNets (ll2_two_head; twn_PartShare.cpp:93-169,
twn_JustEncoding.cpp) against a numpy composition of the heads as separate networks, written
from the reference's topology (myNetworkLL2.h:9-24):

    PartShare:    encoding -> base net (layers 0-1) -> tail 1 (L) | tail 2 (L^2), 2 layers each
    JustEncoding: encoding -> net 1 (L) | net 2 (L^2), 3 layers each

with 32-wide heads: a hidden layer's neurons 0-31 are head 1, 32-63 head 2; outputs 0-2 head 1,
3-5 head 2. The parameter vector is the one-net layout; the weights between the heads are
ignored by the forward (random values there change nothing) and get no gradient. The backward
is checked against central differences, and a loss on one head's outputs moves no weight of the
other head."""
import numpy as np
import pytest

import nrrs_amd as na
from test_oracle_train import fd_check
from test_oracle_train import setup as train_setup

KIND = {1: ("PartShare", 3), 2: ("JustEncoding", 2)}  # ll2_two_head -> (name, statnet hidden layers)


def _cfg_kw(t):
    return dict(ll2_two_head=t, statnet=dict(n_hidden_layers=KIND[t][1], n_outputs=6))


def _layers(p, g, nl):
    """Split the MLP part of a StatNet parameter vector into its layer matrices."""
    W, pos, fin = [], g, 80
    for layer in range(nl):
        o = 6 if layer == nl - 1 else 64
        W.append(p[pos:pos + o * fin].reshape(o, fin))
        pos += o * fin
        fin = o
    return W


def _heads_numpy(t, W, e):
    """StatNet outputs [n][6] of the two heads, each evaluated as a network of its own."""
    relu = lambda x: np.maximum(x, 0.0)  # noqa: E731
    if t == 1:
        base = relu(W[1] @ relu(W[0] @ e.T))
        y1 = W[3][0:3, 0:32] @ relu(W[2][0:32] @ base)
        y2 = W[3][3:6, 32:64] @ relu(W[2][32:64] @ base)
    else:
        y1 = W[2][0:3, 0:32] @ relu(W[1][0:32, 0:32] @ relu(W[0][0:32] @ e.T))
        y2 = W[2][3:6, 32:64] @ relu(W[1][32:64, 32:64] @ relu(W[0][32:64] @ e.T))
    return np.vstack([y1, y2]).T


def _between_heads(t, nl):
    """[layer] boolean masks of the weights between the heads."""
    m = []
    fin = 80
    first = nl - 2 if t == 1 else 0
    for layer in range(nl):
        o = 6 if layer == nl - 1 else 64
        mk = np.zeros((o, fin), bool)
        if layer > first:
            rh = (np.arange(o) >= (3 if layer == nl - 1 else 32))[:, None]
            ch = (np.arange(fin) >= 32)[None, :]
            mk = rh != ch
        m.append(mk)
        fin = o
    return m


@pytest.mark.parametrize("t", [1, 2])
def test_two_head_forward_is_two_networks(ora, t):
    cfg, hv, film = train_setup(ora, **_cfg_kw(t))
    nl = KIND[t][1] + 1
    p = ora.init_params(cfg, na.STATNET, 3, 0.3)
    g, _ = ora.param_count(cfg, na.STATNET)
    W = _layers(p, g, nl)
    masks = _between_heads(t, nl)
    for Wl, mk in zip(W, masks):
        assert np.all(Wl[mk] == 0.0)  # the initialisation leaves the weights between the heads at zero
        Wl[mk] = np.random.default_rng(1).normal(0, 1, mk.sum())  # and the forward ignores them
    y = ora.net_forward(cfg, na.STATNET, hv, film, None, p)
    e = ora.encode(cfg, na.STATNET, hv, film, None, p)
    want = _heads_numpy(t, W, e)
    assert np.allclose(y, want, rtol=1e-12, atol=1e-12)
    # head 1 (outputs 0-2) does not depend on head 2's weights
    p2 = p.copy()
    W2 = _layers(p2, g, nl)
    W2[-1][3:6] *= 3.0
    if t == 2:
        W2[0][32:64] *= -1.0
    y2 = ora.net_forward(cfg, na.STATNET, hv, film, None, p2)
    assert np.array_equal(y2[:, :3], y[:, :3]) and not np.allclose(y2[:, 3:], y[:, 3:])


@pytest.mark.parametrize("t", [1, 2])
def test_two_head_backward(ora, t):
    cfg, hv, film = train_setup(ora, ll2_step=1, **_cfg_kw(t))
    nl = KIND[t][1] + 1
    params = ora.init_params(cfg, na.STATNET, 7, 0.3)
    g, _ = ora.param_count(cfg, na.STATNET)
    tL = ora.synth_targets(hv.n, 1)
    _, grad, _ = ora.train_grads(cfg, na.STATNET, hv, params, na.LOSS_LL2, film=film, target_L=tL)
    y0 = ora.net_forward(cfg, na.STATNET, hv, film, None, params).T.copy()

    def loss_of(p):
        y = ora.net_forward(cfg, na.STATNET, hv, film, None, p).T
        return ora.loss(cfg, na.LOSS_LL2, hv, y, target_L=tL, y_den=y0)[0]
    assert fd_check(ora, cfg, na.STATNET, hv, params, loss_of, grad) < 1e-4
    for gl, mk in zip(_layers(grad, g, nl), _between_heads(t, nl)):
        assert np.all(gl[mk] == 0.0)
    # a loss on head 2's outputs alone moves no weight of head 1
    dy = np.zeros((6, hv.n))
    dy[3:] = np.random.default_rng(2).normal(0, 1, (3, hv.n))
    _, g2, _ = ora.train_grads(cfg, na.STATNET, hv, params, na.LOSS_EXTERNAL, film=film, dy_ext=dy)
    G = _layers(g2, g, nl)
    assert np.all(G[-1][0:3] == 0.0) and np.any(G[-1][3:6] != 0.0)
    if t == 2:  # JustEncoding: head 1's whole net
        assert np.all(G[0][0:32] == 0.0) and np.all(G[1][0:32] == 0.0)
    else:  # PartShare: head 1's tail; the base net is shared
        assert np.all(G[2][0:32] == 0.0) and np.any(G[0] != 0.0)
