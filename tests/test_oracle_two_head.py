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


# AllSelf (ll2_two_head 3, twn_AllSelf.cpp): two whole nets with their own encodings, each
# 80-64-64-3, NRRS_STATNET for L and STATNET2 for L^2, each stored as the 6-output StatNet whose
# output rows of the other head are not connected.
ALL_SELF = dict(ll2_two_head=3, statnet=dict(n_hidden_layers=2, n_outputs=6))


def _all_self_numpy(W, e, head):
    relu = lambda x: np.maximum(x, 0.0)  # noqa: E731
    rows = slice(3 * head, 3 * head + 3)
    return (W[2][rows] @ relu(W[1] @ relu(W[0] @ e.T))).T


def _all_self_setup(ora, **kw):
    cfg, hv, film = train_setup(ora, **ALL_SELF, **kw)
    p = [ora.init_params(cfg, net, 3 + net, 0.3) for net in (na.STATNET, na.STATNET2)]
    return cfg, hv, film, p


def test_all_self_forward_is_two_networks(ora):
    cfg, hv, film, p = _all_self_setup(ora)
    g, _ = ora.param_count(cfg, na.STATNET)
    assert ora.param_count(cfg, na.STATNET2) == ora.param_count(cfg, na.STATNET)
    for head, (net, pn) in enumerate(zip((na.STATNET, na.STATNET2), p)):
        W = _layers(pn, g, 3)
        other = slice(3 * (1 - head), 3 * (1 - head) + 3)
        assert not W[2][other].any()  # the initialisation leaves the other head's rows at zero
        W[2][other] = np.random.default_rng(4).normal(0, 1, (3, 64))  # and the forward ignores them
        y = ora.net_forward(cfg, net, hv, film, None, pn)
        e = ora.encode(cfg, net, hv, film, None, pn)
        assert np.allclose(y[:, 3 * head:3 * head + 3], _all_self_numpy(W, e, head), rtol=1e-12, atol=1e-12)
        assert not y[:, other].any()
    # each net reads its own table: the encodings differ only in the grid channels
    e0 = ora.encode(cfg, na.STATNET, hv, film, None, p[0])
    e1 = ora.encode(cfg, na.STATNET2, hv, film, None, p[1])
    assert not np.array_equal(e0[:, :32], e1[:, :32]) and np.array_equal(e0[:, 32:], e1[:, 32:])


def test_all_self_backward(ora):
    """Each net's gradient is that of NRRS_LL2 over the six statistics (L from the first net,
    L^2 from the second): central differences of the full loss, and the two nets' loss values
    add up to it ((loss1 + loss2) / 2, twn_AllSelf.cpp:103-108)."""
    cfg, hv, film, p = _all_self_setup(ora, ll2_step=1)
    tL = ora.synth_targets(hv.n, 1)

    def stats(pa, pb):
        ya = ora.net_forward(cfg, na.STATNET, hv, film, None, pa).T
        yb = ora.net_forward(cfg, na.STATNET2, hv, film, None, pb).T
        return np.vstack([ya[:3], yb[3:]])

    y0 = stats(*p)
    full = ora.loss(cfg, na.LOSS_LL2, hv, y0, target_L=tL)[0]
    vals = []
    g, _ = ora.param_count(cfg, na.STATNET)
    for head, net in enumerate((na.STATNET, na.STATNET2)):
        val, grad, _ = ora.train_grads(cfg, net, hv, p[head], na.LOSS_LL2, film=film, target_L=tL)
        vals.append(val)

        def loss_of(q, head=head):
            pa, pb = (q, p[1]) if head == 0 else (p[0], q)
            return ora.loss(cfg, na.LOSS_LL2, hv, stats(pa, pb), target_L=tL, y_den=y0)[0]

        assert fd_check(ora, cfg, net, hv, p[head], loss_of, grad) < 1e-4
        W = _layers(grad, g, 3)
        assert not W[2][3 * (1 - head):3 * (1 - head) + 3].any() and W[2][3 * head:3 * head + 3].any()
    assert vals[0] + vals[1] == pytest.approx(full, rel=1e-12)
    # L only (ll2_step 0): the L^2 net has no loss and no gradient
    cfg0, _, _, _ = _all_self_setup(ora, ll2_step=0)
    val2, grad2, _ = ora.train_grads(cfg0, na.STATNET2, hv, p[1], na.LOSS_LL2, film=film, target_L=tL)
    assert val2 == 0.0 and not grad2.any()
