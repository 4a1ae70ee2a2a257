"""Pins the oracle's train_sigma mode (network.h:158; train.cu:196-200): StatNet outputs 3-5
model sigma through activationSigma, so that E[L^2] = sigma^2 + L^2 (the integrator's reading,
integrator.cpp:655-670).

- Training input (reference-visible): the RRSNet input of a training step takes
  BC(activationSigma(max(L2, 0))) in channels 72-74, the inference input keeps BC(max(L2, 0)).
- NRRS_LL2 (restated, nrrs_ll2.h is absent): the sigma output minimises its term where
  sigma^2 + mu^2 = L^2 (mu = the L output, a constant of the gradient): found by bisection on
  the oracle's own dL/dy; the L outputs' gradients are those of the plain mode.
- NRRS_RRS (restated, nrrs_rrs.h is absent): with g1 alone the minimiser is
  q* = sqrt(v / E) - 0.01 with v from the second moments sigma^2 + max(L, 0)^2.
- Both backwards against central differences (StatNet through the frozen-constants surrogate,
  RRSNet through the loss of its training forward).

Plausible mistakes in the sigma code, planted in copies of the oracle source, each fail a check."""
import ctypes as C
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

import nrrs_amd as na
from test_oracle_train import fd_check
from test_oracle_train import setup as train_setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LN2 = math.log(2.0)


def _sigma(y):
    return math.log1p(math.exp(y)) if y < 0 else 0.5 * y + LN2


def _bc(x):
    return 2.0 * (math.sqrt(x) - 1.0)


def _one(ora, P=4):
    cfg0 = na.default_config(n_pixels=P, grid=dict(log2_table_size=12))
    d = ora.synth_vertices(1, 3, P)
    d["thp"][:, 0] = (2.0, 1.0, 0.5)
    d["pixel"][0] = 2
    return ora.HostVertices(d), cfg0


def sigma_checks(ora, lib=None):
    """Names of the failed checks (empty when the sigma mode behaves as documented)."""
    bad = []
    hv, _ = _one(ora)
    P = 4
    film = np.full((3, P), 0.7, np.float32)
    params = ora.init_params(na.default_config(n_pixels=P, grid=dict(log2_table_size=12)), na.RRSNET, 1, 0.1)
    stats = np.array([[0.5], [-1.0], [2.0], [0.0], [2.0], [-3.0]])
    # training input against the inference input
    for ts in (0, 1):
        cfg = na.default_config(n_pixels=P, grid=dict(log2_table_size=12), train_sigma=ts)
        e_inf = ora.encode(cfg, na.RRSNET, hv, film, stats, params)[0]
        e_trn = ora.encode(cfg, na.RRSNET, hv, film, stats, params, training=True)[0]
        want_inf = [_bc(0.0), _bc(2.0), _bc(0.0)]
        want_trn = [_bc(LN2), _bc(1.0 + LN2), _bc(LN2)] if ts else want_inf
        if not np.allclose(e_inf[72:75], want_inf, rtol=0, atol=1e-12):
            bad.append(f"inference input (train_sigma {ts}): {e_inf[72:75]}")
        if not np.allclose(e_trn[72:75], want_trn, rtol=0, atol=1e-12):
            bad.append(f"training input (train_sigma {ts}): {e_trn[72:75]}")
        if not np.array_equal(np.delete(e_inf, [72, 73, 74]), np.delete(e_trn, [72, 73, 74])):
            bad.append("training and inference inputs differ outside the L2 channels")
    # LL2: minimiser of the sigma output, L = (2, 3, 20) clamped at 10, mu = (1, -0.5, 4)
    cfg = na.default_config(n_pixels=P, grid=dict(log2_table_size=12), train_sigma=1, ll2_step=1)
    tL = np.array([[2.0], [3.0], [20.0]], np.float32)
    mu = (1.0, -0.5, 4.0)
    for c in range(3):
        def g(y3, c=c):
            y = np.zeros((6, 1))
            y[:3, 0] = mu
            y[3 + c, 0] = y3
            return ora.loss(cfg, na.LOSS_LL2, hv, y, target_L=tL, lib=lib)[1][3 + c, 0]
        lo, hi = -30.0, 400.0
        if not (g(lo) < 0 < g(hi)):
            bad.append(f"LL2 sigma output {c}: no sign change")
            continue
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if g(mid) < 0 else (lo, mid)
        s = _sigma(0.5 * (lo + hi))
        L = min(float(tL[c, 0]), 10.0)
        if not math.isclose(s * s + mu[c] ** 2, L * L, rel_tol=1e-9):
            bad.append(f"LL2 sigma output {c}: sigma^2 + mu^2 = {s * s + mu[c] ** 2} != {L * L}")
    # the L outputs' gradients do not depend on the mode
    y = np.array([[0.3], [-0.2], [1.5], [0.1], [-0.4], [0.8]])
    d1 = ora.loss(cfg, na.LOSS_LL2, hv, y, target_L=tL, lib=lib)[1]
    cfg0 = na.default_config(n_pixels=P, grid=dict(log2_table_size=12), train_sigma=0, ll2_step=1)
    d0 = ora.loss(cfg0, na.LOSS_LL2, hv, y, target_L=tL, lib=lib)[1]
    if not np.allclose(d1[:3], d0[:3], rtol=1e-12, atol=0):
        bad.append(f"LL2 L-output gradients change with train_sigma: {d1[:3, 0]} vs {d0[:3, 0]}")
    # the sigma outputs' gradients against central differences of the loss, its constants
    # (mu, denominators) held at y; both branches of activationSigma
    for c in range(3, 6):
        h = 1e-6
        yp, ym = y.copy(), y.copy()
        yp[c, 0] += h
        ym[c, 0] -= h
        fd = (ora.loss(cfg, na.LOSS_LL2, hv, yp, target_L=tL, y_den=y, lib=lib)[0] -
              ora.loss(cfg, na.LOSS_LL2, hv, ym, target_L=tL, y_den=y, lib=lib)[0]) / (2 * h)
        if not math.isclose(fd, d1[c, 0], rel_tol=1e-6):
            bad.append(f"LL2 sigma output {c - 3}: dL/dy {d1[c, 0]} != central difference {fd}")
    # RRS g1 minimum with sigma statistics: thp (2, 1, 0.5), L (0.5, -1, 2), sigma outputs (0, 2, -3)
    ref = float(np.float32(0.7))
    m2 = [_sigma(0.0) ** 2 + 0.25, _sigma(2.0) ** 2 + 0.0, _sigma(-3.0) ** 2 + 4.0]
    v = (4.0 * m2[0] + 1.0 * m2[1] + 0.25 * m2[2]) / 3.0 / (ref * ref + 0.01)
    E = 2.0 / P
    cfg = na.default_config(n_pixels=P, grid=dict(log2_table_size=12), train_sigma=1, loss_clamp_on=0,
                            gamma=[1.0, 0.0, 0.0, 0.0])
    side = ora.RrsSide(np.array([0.7], np.float32), np.array([0.3], np.float32), np.array([1.0], np.float32),
                       np.array([0.5], np.float32), 2.0)

    def gr(yq):
        return ora.loss(cfg, na.LOSS_RRS, hv, np.array([[yq]]), film=film, stats=stats, side=side, lib=lib)[1][0, 0]
    lo, hi = -30.0, 4000.0
    if gr(lo) < 0 < gr(hi):
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if gr(mid) < 0 else (lo, mid)
        q = _sigma(0.5 * (lo + hi))
        if not math.isclose(q, math.sqrt(v / E) - 0.01, rel_tol=1e-9):
            bad.append(f"RRS g1 minimum with sigma statistics: {q} != {math.sqrt(v / E) - 0.01}")
    else:
        bad.append("RRS g1: no sign change")
    return bad


def test_train_sigma_pinned(ora):
    bad = sigma_checks(ora)
    assert not bad, bad


def test_ll2_sigma_backward_fd(ora):
    cfg, hv, film = train_setup(ora, train_sigma=1)
    params = ora.init_params(cfg, na.STATNET, 7, 0.3)
    tL = ora.synth_targets(hv.n, 1)
    _, grad, _ = ora.train_grads(cfg, na.STATNET, hv, params, na.LOSS_LL2, film=film, target_L=tL)
    y0 = ora.net_forward(cfg, na.STATNET, hv, film, None, params).T.copy()

    def loss_of(p):  # frozen constants (denominators, mu) of the relative L2
        y = ora.net_forward(cfg, na.STATNET, hv, film, None, p).T
        return ora.loss(cfg, na.LOSS_LL2, hv, y, target_L=tL, y_den=y0)[0]
    assert fd_check(ora, cfg, na.STATNET, hv, params, loss_of, grad) < 1e-4


def test_rrs_sigma_backward_fd(ora):
    cfg, hv, film = train_setup(ora, train_sigma=1)
    n = hv.n
    rng = np.random.default_rng(4)
    stats = np.vstack([rng.lognormal(-1, 1, (3, n)), rng.normal(0, 1.5, (3, n))])
    side = ora.RrsSide(None, rng.uniform(0, 2, n).astype(np.float32), np.ones(n, np.float32),
                       rng.uniform(0.05, 5, n).astype(np.float32), 0.7 * cfg.n_pixels)
    params = ora.init_params(cfg, na.RRSNET, 9, 0.3)
    _, grad, _ = ora.train_grads(cfg, na.RRSNET, hv, params, na.LOSS_RRS, film=film, stats=stats, side=side)

    def loss_of(p):  # the loss of the training forward (its input takes the sigma transform)
        return ora.train_grads(cfg, na.RRSNET, hv, p, na.LOSS_RRS, film=film, stats=stats, side=side)[0]
    assert fd_check(ora, cfg, na.RRSNET, hv, params, loss_of, grad) < 1e-4


# Plausible mistakes in the sigma code, each a (find, replace) on the oracle source.
MUTANTS = {
    "sigma not squared in LL2": ("double m = sg * sg + mu * mu, md", "double m = sg + mu * mu, md"),
    "mu dropped in LL2": ("double m = sg * sg + mu * mu, md", "double m = sg * sg, md"),
    "chain rule of sigma dropped": ("slope = 2.0 * sg * sigma_grad_d(yy);", "slope = 2.0 * sg;"),
    "training input without activationSigma": ("if (training && cfg->train_sigma) l2 = sigma_d(l2);", ""),
    "activationSigma on the inference input": ("if (training && cfg->train_sigma) l2 = sigma_d(l2);",
                                              "if (cfg->train_sigma) l2 = sigma_d(l2);"),
    "mu dropped in RRS": ("l2 = sg * sg + mu * mu;", "l2 = sg * sg;"),
    "sigma not squared in RRS": ("l2 = sg * sg + mu * mu;", "l2 = sg + mu * mu;"),
}


@pytest.fixture(scope="module")
def mutant_dir():
    with tempfile.TemporaryDirectory() as d:
        yield d


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_sigma_mistake_fails_a_check(ora, mutant_dir, name, monkeypatch):
    src = open(os.path.join(ROOT, "oracle", "nrrs_oracle.cpp")).read()
    find, repl = MUTANTS[name]
    assert src.count(find) == 1, f"mutation site not found once: {find}"
    tag = "".join(ch if ch.isalnum() else "_" for ch in name)
    cpp, so = os.path.join(mutant_dir, f"{tag}.cpp"), os.path.join(mutant_dir, f"{tag}.so")
    with open(cpp, "w") as f:
        f.write(src.replace(find, repl))
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
                    f"-I{os.path.join(ROOT, 'oracle')}", "-shared", cpp, "-o", so], check=True)
    lib = ora.bind_loss(C.CDLL(so))
    # the encoders are not part of bind_loss: route the encode calls of the checks to the mutant too
    base = ora.load()
    mut = C.CDLL(so)
    for fn in ("ora_encode", "ora_encode_train"):
        f, ref = getattr(mut, fn), getattr(base, fn)
        f.restype, f.argtypes = ref.restype, ref.argtypes
    enc = _Encoders(mut, base)
    monkeypatch.setattr(ora, "load", lambda: enc)
    bad = sigma_checks(ora, lib)
    print(f"{name}: {len(bad)} failed checks, e.g. {bad[:2]}")
    assert bad, f"the planted mistake '{name}' passes every check"


class _Encoders:
    """The oracle library with its encoders taken from a mutant build."""

    def __init__(self, mut, base):
        self._mut = mut
        self._base = base

    def __getattr__(self, name):
        if name in ("ora_encode", "ora_encode_train"):
            return getattr(self._mut, name)
        return getattr(self._base, name)
