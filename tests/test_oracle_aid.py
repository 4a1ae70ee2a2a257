"""Pins the oracle's AID RRSNet input (rrs_aid; net1_aid_net2, network.cpp:31, 209-258, 303-343;
NRRSNetworkInput1AID, train.h:58-68; filled by nrrs_generate_inference_data_aid,
train.cu:242-267, and nrrs_generate_training_data_from_0to1_aid, train.cu:114-153).

The AID input holds the position, direction, roughness, reference mean and throughput of a
vertex and nothing else. Against the one-net RRSNet input of the same vertices:

- the grid, direction, roughness, throughput and reference-mean channels are the same values;
- the normal, albedo and statistics channels are zero, and the bias channel is one;
- the input depends on each of the five fields (a change of any moves it) and on no other:
  the normal, the albedos and the StatNet statistics change nothing;
- the NRRS_RRS loss still takes the statistics (the reference's ll2 buffer, network.cpp:431):
  its value for given outputs is that of the one-net configuration, and the backward through
  the AID input agrees with central differences."""
import numpy as np

import nrrs_amd as na
from test_oracle_train import N, fd_check, rrs_side
from test_oracle_train import setup as train_setup

GRID = slice(0, 32)
DIR = slice(32, 48)
NRM = slice(48, 64)
ROUGH, THP, REF = 64, slice(65, 68), 68
STATS = slice(69, 75)
BIAS = 75


def _enc(ora, cfg, hv, film, stats, params, training=False):
    return ora.encode(cfg, na.RRSNET, hv, film, stats, params, training=training)


def _setup(ora, aid):
    cfg, hv, film = train_setup(ora, rrs_aid=aid)
    params = ora.init_params(cfg, na.RRSNET, 8, 0.3)
    stats = np.abs(np.random.default_rng(2).normal(0.5, 0.5, (6, N)))
    return cfg, hv, film, params, stats


def test_aid_channels(ora):
    cfg, hv, film, params, stats = _setup(ora, 1)
    cfg0, _, _, _, _ = _setup(ora, 0)
    for training in (False, True):
        e = _enc(ora, cfg, hv, film, stats, params, training)
        e0 = _enc(ora, cfg0, hv, film, stats, params, training)
        for ch in (GRID, DIR, ROUGH, THP, REF, BIAS):
            assert np.array_equal(e[:, ch], e0[:, ch])
        assert not e[:, NRM].any() and not e[:, STATS].any() and e0[:, NRM].any() and e0[:, STATS].any()
        assert np.all(e[:, BIAS] == 1.0) and not e[:, 76:].any()
    # the same input without statistics (inference) as with them (training)
    assert np.array_equal(_enc(ora, cfg, hv, film, None, params), _enc(ora, cfg, hv, film, stats, params, True))
    # StatNet's input does not change
    sp = ora.init_params(cfg, na.STATNET, 3, 0.3)
    assert np.array_equal(ora.encode(cfg, na.STATNET, hv, film, None, sp), ora.encode(cfg0, na.STATNET, hv, film, None, sp))


def test_aid_input_fields(ora):
    """The input moves with the position, direction, roughness, throughput and film (reference
    mean) of a vertex, and with nothing else."""
    cfg, hv, film, params, stats = _setup(ora, 1)
    base = _enc(ora, cfg, hv, film, stats, params, True)
    d0 = {k: v.copy() for k, v in hv.d.items() if v is not None}

    def moved(field, f):
        d = {k: v.copy() for k, v in d0.items()}
        d[field] = f(d[field])
        h = ora.HostVertices(d)
        return not np.array_equal(_enc(ora, cfg, h, film, stats, params, True), base)

    rot = lambda a: np.roll(a, 1, axis=0)  # noqa: E731  permute the xyz rows (unit vectors stay unit)
    assert moved("pos", lambda a: a + 0.01)
    assert moved("dir", rot)
    assert moved("rough", lambda a: a + 0.1)
    assert moved("thp", lambda a: a * 1.5)
    assert moved("pixel", lambda a: (a + 1) % cfg.n_pixels)
    assert not moved("normal", rot)
    assert not moved("albedo_d", lambda a: a * 0.5)
    assert not moved("albedo_s", lambda a: a * 0.5)
    assert np.array_equal(_enc(ora, cfg, hv, film, stats * 3.0 + 1.0, params, True), base)
    assert not np.array_equal(_enc(ora, cfg, hv, film * 2.0, stats, params, True), base)


def test_aid_loss_takes_the_statistics(ora):
    cfg, hv, film, params, stats = _setup(ora, 1)
    cfg0, _, _, _, _ = _setup(ora, 0)
    side = ora.RrsSide(**rrs_side(N, 4))
    y = np.random.default_rng(5).normal(0, 1, (1, N))
    l1 = ora.loss(cfg, 2, hv, y, film=film, stats=stats, side=side)[0]
    assert l1 == ora.loss(cfg0, 2, hv, y, film=film, stats=stats, side=side)[0]
    assert l1 != ora.loss(cfg, 2, hv, y, film=film, stats=stats * 2.0, side=side)[0]
    val, grad, dy = ora.train_grads(cfg, na.RRSNET, hv, params, 2, film=film, stats=stats, side=side)
    yf = ora.net_forward(cfg, na.RRSNET, hv, film, stats, params).T
    assert val == ora.loss(cfg, 2, hv, yf, film=film, stats=stats, side=side)[0]

    def loss_of(p):
        y = ora.net_forward(cfg, na.RRSNET, hv, film, stats, p).T
        return ora.loss(cfg, 2, hv, y, film=film, stats=stats, side=side)[0]

    assert fd_check(ora, cfg, na.RRSNET, hv, params, loss_of, grad) < 1e-4
    # the first layer's weights on the zero channels get no gradient
    g, _ = ora.param_count(cfg, na.RRSNET)
    W0 = grad[g:g + 64 * 80].reshape(64, 80)
    assert not W0[:, NRM].any() and not W0[:, STATS].any() and W0[:, DIR].any()
