"""Pins the oracle's shipped nets (net_set 1: the reference's 2net-simple.json:29-75, 210-291):

- worldToLatLong (KiRaRay, absent from the checkout, so its convention is parity unpinned): known
  answers on the axes, and the inverse map latLongToWorld recovers every direction;
- OneBlob (tiny-cuda-nn, absent, parity unpinned): each bin against a numerical quadrature of the
  biweight kernel's density wrapped onto [0, 1) (an evaluation independent of the oracle's CDF
  polynomial), the bins' sum of 1, and the values at a bin centre and on a bin edge;
- the input channels of each net against the reference's input structs (train.h:22-68);
- the MLPs against a numpy evaluation of the 32-wide LeakyReLU nets written from the reference's
  topology (2net-simple.json:69-75, 238-247), with the weights outside the 32-wide nets
  initialised to zero and ignored by the forward;
- the backward against central differences, for both losses, with no gradient outside the
  32-wide nets."""
import numpy as np
import pytest

import nrrs_amd as na
from test_oracle_train import fd_check, rrs_side

N = 96


def _cfg(**kw):
    return na.shipped_config(n_pixels=N, grid=dict(log2_table_size=12), **kw)


def _setup(**kw):
    import oracle_ref as ora

    cfg = _cfg(**kw)
    hv = ora.HostVertices(ora.synth_vertices(N, 5, N))
    film = np.random.default_rng(3).lognormal(-1, 1, (3, N)).astype(np.float32)
    return cfg, hv, film


def _bc(x):
    return 2.0 * (np.sqrt(x) - 1.0)


def test_lat_long_axes_and_inverse(ora):
    for d, uv in (((0, 1, 0), (None, 0.0)), ((0, -1, 0), (None, 1.0)), ((0, 0, -1), (0.5, 0.5)),
                  ((1, 0, 0), (0.75, 0.5)), ((-1, 0, 0), (0.25, 0.5)), ((0, 0, 1), (1.0, 0.5))):
        got = ora.lat_long(d)
        if uv[0] is not None:
            assert got[0] == pytest.approx(uv[0], abs=1e-15)
        assert got[1] == pytest.approx(uv[1], abs=1e-15)
    rng = np.random.default_rng(0)
    for _ in range(200):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        u, v = ora.lat_long(d * 2.5)  # the map normalises its argument
        th, ph = np.pi * v, 2.0 * np.pi * (u - 0.5)
        back = np.array([np.sin(th) * np.sin(ph), np.cos(th), -np.sin(th) * np.cos(ph)])
        assert np.allclose(back, d, atol=1e-12)
        assert 0.0 <= u <= 1.0 and 0.0 <= v <= 1.0


def _one_blob_quadrature(x, n, m=40000):
    """Mass of the wrapped biweight density n * 15/16 (1 - t^2)^2, t = (s - x) n, on each bin, by
    composite Simpson over m intervals per bin (the density's kinks at x +- 1/n cost O(h^2))."""
    out = np.zeros(n)
    for k in range(n):
        s = np.linspace(k / n, (k + 1) / n, m + 1)
        dens = np.zeros_like(s)
        for w in (-1.0, 0.0, 1.0):
            t = (s - x - w) * n
            dens += np.where(np.abs(t) < 1.0, 15.0 / 16.0 * (1.0 - t * t) ** 2 * n, 0.0)
        h = (s[-1] - s[0]) / m
        out[k] = h / 3.0 * (dens[0] + dens[-1] + 4.0 * dens[1:-1:2].sum() + 2.0 * dens[2:-1:2].sum())
    return out


def test_one_blob_against_quadrature(ora):
    rng = np.random.default_rng(1)
    for n in (4, 8):
        for x in list(rng.uniform(0, 1, 12)) + [0.0, 0.5 / n, 1.0 / n, 0.999]:
            ob = ora.one_blob(x, n)
            assert np.all(ob >= -1e-15)
            assert ob.sum() == pytest.approx(1.0, abs=1e-12)
            assert np.allclose(ob, _one_blob_quadrature(x, n), atol=1e-7), (n, x)
    # at a bin centre the kernel (radius one bin) covers the bin and half of each neighbour:
    # C(1/2) - C(-1/2) = 15/16 (1 - 1/12 + 1/80) = 0.79296875, 0.103515625 to each neighbour
    ob = ora.one_blob(2.5 / 8, 8)
    assert ob[2] == pytest.approx(0.79296875, abs=1e-15)
    assert ob[1] == pytest.approx(0.103515625, abs=1e-15) and ob[3] == pytest.approx(0.103515625, abs=1e-15)
    # on the edge x = 0 half the mass wraps into the last bin
    ob = ora.one_blob(0.0, 4)
    assert np.allclose(ob, [0.5, 0.0, 0.0, 0.5], atol=1e-15)


def test_shipped_inputs_follow_the_reference_structs(ora):
    cfg, hv, film = _setup()
    d = hv.d
    rng = np.random.default_rng(2)
    stats = rng.normal(0.5, 1.0, (6, N))
    # StatNet: NRRSNetworkInput0 (pos, lat-long dir, warped roughness) -> HashGrid 8 x 2 +
    # OneBlob(u, 4) + OneBlob(v, 4) + OneBlob(1 - e^-r, 8)
    ps = ora.init_params(cfg, na.STATNET, 1, 0.3)
    e = ora.encode(cfg, na.STATNET, hv, film, None, ps)
    assert np.any(e[:, :16]) and not e[:, 16:32].any() and not e[:, 48:].any()
    for i in range(N):
        u, v = ora.lat_long(d["dir"][:, i])
        wr = 1.0 - np.exp(-np.float64(d["rough"][i]))
        want = np.concatenate([ora.one_blob(u, 4), ora.one_blob(v, 4), ora.one_blob(wr, 8)])
        assert np.allclose(e[i, 32:48], want, atol=1e-15)
    # RRSNet: NRRSNetworkInput1 (train.h:39-56), no grid
    assert ora.param_count(cfg, na.RRSNET)[0] == 0
    pr = ora.init_params(cfg, na.RRSNET, 2, 0.3)
    ref = film[:, d["pixel"]].astype(np.float64).mean(axis=0)
    for training in (False, True):  # the same input without train_sigma
        e = ora.encode(cfg, na.RRSNET, hv, film, stats, pr, training=training)
        assert not e[:, :64].any() and not e[:, 76:].any()
        assert np.allclose(e[:, 64], 1.0 - np.exp(-d["rough"].astype(np.float64)), atol=1e-15)
        assert np.allclose(e[:, 65:68], _bc(d["thp"].astype(np.float64)).T, atol=1e-14)
        assert np.allclose(e[:, 68], _bc(ref), atol=1e-14)
        assert np.allclose(e[:, 69:72], _bc(np.maximum(stats[:3], 0)).T, atol=1e-14)
        assert np.allclose(e[:, 72:75], _bc(np.maximum(stats[3:], 0)).T, atol=1e-14)
        assert np.all(e[:, 75] == 1.0)
    # with train_sigma the training input takes activationSigma of L2 (train.cu:196-200)
    # and the inference input keeps max(L2, 0) (train.cu:41-112)
    cs, _, _ = _setup(train_sigma=1)
    e = ora.encode(cs, na.RRSNET, hv, film, stats, pr, training=True)
    sig = np.vectorize(ora.load().ora_activation_sigma)
    assert np.allclose(e[:, 72:75], _bc(sig(np.maximum(stats[3:], 0))).T, atol=1e-14)
    e = ora.encode(cs, na.RRSNET, hv, film, stats, pr, training=False)
    assert np.allclose(e[:, 72:75], _bc(np.maximum(stats[3:], 0)).T, atol=1e-14)
    # AID RRSNet: NRRSNetworkInput1AID (train.h:58-68) -> HashGrid + OneBlob(u, v; 4) +
    # OneBlob(rough, 4) + Identity(BC(ref mean), BC(thp))
    ca, _, _ = _setup(rrs_aid=1)
    pa = ora.init_params(ca, na.RRSNET, 3, 0.3)
    e = ora.encode(ca, na.RRSNET, hv, film, stats, pa)
    assert np.any(e[:, :16]) and not e[:, 16:32].any() and not e[:, 44:65].any() and not e[:, 69:].any()
    for i in range(N):
        u, v = ora.lat_long(d["dir"][:, i])
        wr = 1.0 - np.exp(-np.float64(d["rough"][i]))
        want = np.concatenate([ora.one_blob(u, 4), ora.one_blob(v, 4), ora.one_blob(wr, 4)])
        assert np.allclose(e[i, 32:44], want, atol=1e-15)
    assert np.allclose(e[:, 65:68], _bc(d["thp"].astype(np.float64)).T, atol=1e-14)
    assert np.allclose(e[:, 68], _bc(ref), atol=1e-14)


def _layers(p, g, nl, n_out):
    W, pos, fin = [], g, 80
    for layer in range(nl):
        o = n_out if layer == nl - 1 else 64
        W.append(p[pos:pos + o * fin].reshape(o, fin))
        pos += o * fin
        fin = o
    return W


def _input_cols(cfg, net):
    if net == na.STATNET:
        return np.r_[0:16, 32:48]
    if not cfg.rrs_aid:
        return np.r_[64:80]
    return np.r_[0:16, 32:44, 65:69]


def _live_masks(cfg, net, nl, n_out):
    """[layer] masks of the weights of the 32-wide net inside the 64-wide storage."""
    m, fin = [], 80
    for layer in range(nl):
        o = n_out if layer == nl - 1 else 64
        mk = np.zeros((o, fin), bool)
        rows = slice(0, o) if layer == nl - 1 else slice(0, 32)
        if layer == 0:
            mk[rows, _input_cols(cfg, net)] = True
        else:
            mk[rows, :32] = True
        m.append(mk)
        fin = o
    return m


def _numpy_net(cfg, net, W, e, slope=0.01):
    """The reference's net: LeakyReLU(0.01) 32-wide hidden layers, linear output, no bias."""
    act = lambda x: np.where(x > 0.0, x, slope * x)  # noqa: E731
    h = e[:, _input_cols(cfg, net)].T
    h = act(W[0][:32, _input_cols(cfg, net)] @ h)
    for Wl in W[1:-1]:
        h = act(Wl[:32, :32] @ h)
    return (W[-1][:, :32] @ h).T


@pytest.mark.parametrize("net,aid", [(na.STATNET, 0), (na.RRSNET, 0), (na.RRSNET, 1)])
def test_shipped_forward_is_the_32_wide_net(ora, net, aid):
    cfg, hv, film = _setup(rrs_aid=aid)
    n_out = 6 if net == na.STATNET else 1
    p = ora.init_params(cfg, net, 4, 0.3)
    g, _ = ora.param_count(cfg, net)
    W = _layers(p, g, 4, n_out)
    stats = np.random.default_rng(6).normal(0.5, 1.0, (6, N)) if net == na.RRSNET else None
    rng = np.random.default_rng(7)
    for Wl, mk in zip(W, _live_masks(cfg, net, 4, n_out)):
        assert np.all(Wl[~mk] == 0.0)  # initialised to zero outside the 32-wide net
        assert np.all(Wl[mk] != 0.0)
        Wl[~mk] = rng.normal(0, 1, (~mk).sum())  # and ignored by the forward
    y = ora.net_forward(cfg, net, hv, film, stats, p)
    e = ora.encode(cfg, net, hv, film, stats, p)
    want = _numpy_net(cfg, net, W, e)
    assert np.allclose(y, want, rtol=1e-12, atol=1e-12)
    # the negative side of the activation matters: the same weights under ReLU give other outputs
    assert not np.allclose(y, _numpy_net(cfg, net, W, e, slope=0.0), rtol=1e-6, atol=1e-9)


def test_shipped_statnet_backward(ora):
    cfg, hv, film = _setup(ll2_step=1)
    p = ora.init_params(cfg, na.STATNET, 7, 0.3)
    g, _ = ora.param_count(cfg, na.STATNET)
    tL = ora.synth_targets(N, 1)
    _, grad, _ = ora.train_grads(cfg, na.STATNET, hv, p, na.LOSS_LL2, film=film, target_L=tL)
    y0 = ora.net_forward(cfg, na.STATNET, hv, film, None, p).T.copy()

    def loss_of(q):
        y = ora.net_forward(cfg, na.STATNET, hv, film, None, q).T
        return ora.loss(cfg, na.LOSS_LL2, hv, y, target_L=tL, y_den=y0)[0]

    assert fd_check(ora, cfg, na.STATNET, hv, p, loss_of, grad) < 1e-4
    for gl, mk in zip(_layers(grad, g, 4, 6), _live_masks(cfg, na.STATNET, 4, 6)):
        assert not gl[~mk].any()
    # the table gradient reaches the 8 levels of the shipped grid only
    assert np.any(grad[:g])


@pytest.mark.parametrize("aid", [0, 1])
def test_shipped_rrsnet_backward(ora, aid):
    cfg, hv, film = _setup(rrs_aid=aid)
    p = ora.init_params(cfg, na.RRSNET, 8, 0.3)
    g, m = ora.param_count(cfg, na.RRSNET)
    stats = np.abs(np.random.default_rng(9).normal(0.5, 1.0, (6, N)))
    rs = ora.RrsSide(**rrs_side(N, 10))
    _, grad, _ = ora.train_grads(cfg, na.RRSNET, hv, p, na.LOSS_RRS, film=film, stats=stats, side=rs)

    def loss_of(q):
        y = ora.net_forward(cfg, na.RRSNET, hv, film, stats, q).T
        return ora.loss(cfg, na.LOSS_RRS, hv, y, film=film, stats=stats, side=rs)[0]

    if g:
        assert fd_check(ora, cfg, na.RRSNET, hv, p, loss_of, grad) < 1e-4
    else:  # no grid: central differences on 256 live MLP weights
        live = np.concatenate([mk.ravel() for mk in _live_masks(cfg, na.RRSNET, 4, 1)])
        idx = np.random.default_rng(11).choice(np.nonzero(live)[0], 256, replace=False)
        scale = np.max(np.abs(grad[idx])) + 1e-30
        worst = 0.0
        for k in idx:
            p1, p2 = p.copy(), p.copy()
            p1[k] += 1e-5
            p2[k] -= 1e-5
            worst = max(worst, abs((loss_of(p1) - loss_of(p2)) / 2e-5 - grad[k]) / scale)
        assert worst < 1e-4
    for gl, mk in zip(_layers(grad, g, 4, 1), _live_masks(cfg, na.RRSNET, 4, 1)):
        assert not gl[~mk].any()
