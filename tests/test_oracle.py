"""Pins the CPU oracle before it is trusted as the checker (no GPU needed).

Every check here is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails it: published PCG32
vectors, closed forms and special values, quadrature identities, invariants
and brute force on small cases.
"""
import json
import math
import os

import numpy as np
import pytest

import nrrs_amd as na

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def cfg(**kw):
    # default_config() is host-only code in the HIP library (no device needed)
    return na.default_config(**kw)


# ---------------------------------------------------------------- PCG32
def test_pcg32_known_answers(ora):
    with open(os.path.join(GOLDEN, "pcg32_vectors.json")) as f:
        data = json.load(f)
    for case in data["cases"]:
        got = ora.pcg32(case["seed"], case["stream"], case["skip"], len(case["out"]))
        assert got.tolist() == case["out"], case


def test_pcg32_advance_matches_stepping(ora):
    full = ora.pcg32(7, 3, 0, 300)
    for k in (0, 1, 2, 63, 64, 255, 299):
        assert ora.pcg32(7, 3, k, 1)[0] == full[k]


# ---------------------------------------------------------------- closed forms
def test_activation_sigma_closed_form(ora):
    ln2 = math.log(2.0)
    assert ora.load().ora_activation_sigma(0.0) == pytest.approx(ln2, abs=1e-15)
    assert ora.load().ora_activation_sigma(2.0) == pytest.approx(1.0 + ln2, abs=1e-15)
    assert ora.load().ora_activation_sigma(-30.0) == pytest.approx(math.log1p(math.exp(-30.0)), rel=1e-12)
    assert ora.load().ora_activation_sigma(-30.0) == pytest.approx(9.357622968839e-14, rel=1e-9)
    # C1 at 0: one-sided slopes are both 1/2
    h = 1e-6
    lo = (ora.load().ora_activation_sigma(0.0) - ora.load().ora_activation_sigma(-h)) / h
    hi = (ora.load().ora_activation_sigma(h) - ora.load().ora_activation_sigma(0.0)) / h
    assert lo == pytest.approx(0.5, abs=1e-5) and hi == pytest.approx(0.5, abs=1e-9)


def test_box_cox_known_values(ora):
    bc = ora.load().ora_box_cox
    assert bc(1.0) == 0.0 and bc(0.0) == -2.0 and bc(4.0) == 2.0 and bc(0.25) == -1.0


def test_weight_window(ora):
    ww = ora.load().ora_weight_window
    # size 5: d- = 1/3, d+ = 5/3
    assert ww(1.0 / 6.0, 5.0) == pytest.approx(0.5)
    assert ww(1.0, 5.0) == 1.0
    assert ww(10.0 / 3.0, 5.0) == pytest.approx(2.0)
    assert ww(0.34, 5.0) == 1.0 and ww(1.66, 5.0) == 1.0


def _sh(ora, x, y, z):
    out = np.zeros(16)
    ora.load().ora_sh4(x, y, z, out.ctypes.data)
    return out


def test_sh4_orthonormal(ora):
    # Gauss-Legendre in cos(theta) x uniform in phi integrates products of band<=3
    # harmonics (degree <= 6 polynomials) exactly.
    nt, nphi = 8, 16
    xs, ws = np.polynomial.legendre.leggauss(nt)
    gram = np.zeros((16, 16))
    for z, w in zip(xs, ws):
        r = math.sqrt(1 - z * z)
        for k in range(nphi):
            ph = 2 * math.pi * k / nphi
            y = _sh(ora, r * math.cos(ph), r * math.sin(ph), z)
            gram += w * (2 * math.pi / nphi) * np.outer(y, y)
    assert np.allclose(gram, np.eye(16), atol=1e-12)


def test_sh4_axes(ora):
    c0 = 0.5 / math.sqrt(math.pi)
    c1 = math.sqrt(3 / (4 * math.pi))
    y = _sh(ora, 0.0, 0.0, 1.0)
    assert y[0] == pytest.approx(c0) and y[2] == pytest.approx(c1)
    assert y[1] == 0 and y[3] == 0
    # Y_2^0 at the pole = sqrt(5/(4 pi)), Y_3^0 = sqrt(7/(4 pi))
    assert y[6] == pytest.approx(math.sqrt(5 / (4 * math.pi)))
    assert y[12] == pytest.approx(math.sqrt(7 / (4 * math.pi)))
    y = _sh(ora, 1.0, 0.0, 0.0)
    assert y[3] == pytest.approx(-c1)


# ---------------------------------------------------------------- hash grid
def test_grid_layout_library_matches_oracle(ora):
    c = cfg()
    a = na.grid_layout(c.grid)
    b = ora.grid_layout(c.grid)
    for k in ("res", "offset", "size", "dense", "scale"):
        assert np.array_equal(a[k], b[k]), k
    assert a["n_params"] == b["n_params"]
    # 16 levels, base 16, scale 1.5, T = 2^19: 4 dense levels, then hashed
    assert a["dense"].tolist() == [1] * 4 + [0] * 12
    assert a["res"][0] == 18 and a["res"][1] == 26


def test_grid_indices_in_range_and_dense_injective(ora):
    c = cfg()
    lay = ora.grid_layout(c.grid)
    d = ora.synth_vertices(4096, 11, 4096)
    hv = ora.HostVertices(d)
    idx = ora.grid_indices(c, hv)
    for l in range(c.grid.n_levels):
        assert idx[:, l].max() < lay["size"][l]
    # dense level: corner index is x + r(y + r z) -> decode and compare with the cell
    xn = _normalized(ora, c, hv)
    for l in range(4):
        r = int(lay["res"][l])
        p = xn.astype(np.float64) * float(lay["scale"][l]) + 0.5
        e0 = idx[:, l, 0].astype(np.int64)
        cell = np.stack([e0 % r, (e0 // r) % r, e0 // (r * r)], 1)
        clear = np.abs(p - np.round(p)) > 1e-5  # away from cell boundaries
        assert np.array_equal(cell[clear], np.floor(p)[clear].astype(np.int64))
        for cn in range(8):
            e = idx[:, l, cn].astype(np.int64)
            dec = np.stack([e % r, (e // r) % r, e // (r * r)], 1)
            bits = np.array([cn & 1, (cn >> 1) & 1, (cn >> 2) & 1])
            assert np.array_equal(dec, cell + bits)
    # hashed levels: corners of one cell are (almost always) distinct entries
    distinct = np.array([len(set(idx[i, 8].tolist())) for i in range(256)])
    assert distinct.mean() > 7.9


def _normalized(ora, c, hv):
    xn = np.zeros((hv.n, 3), np.float32)
    mn = np.array(c.aabb_min, np.float32)
    mx = np.array(c.aabb_max, np.float32)
    ora.load().ora_normalize_pos(mn.ctypes.data, mx.ctypes.data, hv.n, hv.d["pos"].ctypes.data, hv.n,
                                 xn.ctypes.data)
    return xn


def test_grid_reproduces_linear_field(ora):
    """A table that stores a linear function of the vertex position is
    interpolated exactly by trilinear weights (dense levels), so features equal
    the function at the continuous position."""
    c = cfg(grid=dict(n_levels=4), n_pixels=512)
    lay = ora.grid_layout(c.grid)
    g, m = ora.param_count(c, na.STATNET)
    params = np.zeros(g + m)
    a0, a1 = np.array([0.3, -1.1, 2.0]), np.array([-0.7, 0.25, 0.5])
    for l in range(4):
        r = int(lay["res"][l])
        e = np.arange(r ** 3)
        x, y, z = e % r, (e // r) % r, e // (r * r)
        base = int(lay["offset"][l]) * 2
        params[base + 2 * e] = a0[0] * x + a0[1] * y + a0[2] * z
        params[base + 2 * e + 1] = a1[0] * x + a1[1] * y + a1[2] * z
    d = ora.synth_vertices(512, 5, 512)
    hv = ora.HostVertices(d)
    film = np.ones((3, 512), np.float32)
    enc = ora.encode(c, na.STATNET, hv, film, None, params)
    xn = _normalized(ora, c, hv)
    for l in range(4):
        p = xn.astype(np.float64) * float(lay["scale"][l]) + 0.5  # continuous cell coordinate
        assert np.allclose(enc[:, 2 * l], p @ a0, atol=1e-4)
        assert np.allclose(enc[:, 2 * l + 1], p @ a1, atol=1e-4)
    # channels of unused levels are zero
    assert np.all(enc[:, 8:32] == 0)


def test_encode_layout(ora):
    c = cfg(n_pixels=64)
    d = ora.synth_vertices(64, 3, 64)
    hv = ora.HostVertices(d)
    film = np.random.default_rng(0).uniform(0, 2, (3, 64)).astype(np.float32)
    g, m = ora.param_count(c, na.STATNET)
    params = np.zeros(g + m)
    e_stat = ora.encode(c, na.STATNET, hv, film, None, params)
    e_rrs = ora.encode(c, na.RRSNET, hv, film, None, params)
    assert np.all(e_stat[:, :32] == 0)
    assert np.allclose(e_stat[:, 64], 1 - np.exp(-d["rough"].astype(np.float64)))
    assert np.allclose(e_stat[:, 65:68], d["albedo_d"].T) and np.all(e_stat[:, 71] == 1)
    assert np.allclose(e_rrs[:, 65:68], 2 * (np.sqrt(d["thp"].T.astype(np.float64)) - 1))
    ref = film[:, d["pixel"]].astype(np.float64).mean(0)
    assert np.allclose(e_rrs[:, 68], 2 * (np.sqrt(ref) - 1))
    assert np.all(e_rrs[:, 75] == 1) and np.all(e_rrs[:, 76:] == 0)


def test_mlp_forward_brute_force(ora):
    """Net forward equals an explicit numpy evaluation of the documented MLP."""
    c = cfg(n_pixels=37)
    d = ora.synth_vertices(37, 9, 37)
    hv = ora.HostVertices(d)
    film = np.ones((3, 37), np.float32)
    for net, nl, no in ((na.STATNET, 5, 6), (na.RRSNET, 4, 1)):
        p = ora.init_params(c, net, 123, 0.5)
        g, m = ora.param_count(c, net)
        enc = ora.encode(c, net, hv, film, None, p)
        w = p[g:]
        h, k, dims = enc, 0, [80] + [64] * (nl - 1) + [no]
        for layer in range(nl):
            W = w[k:k + dims[layer + 1] * dims[layer]].reshape(dims[layer + 1], dims[layer])
            k += W.size
            h = h @ W.T
            if layer < nl - 1:
                h = np.maximum(h, 0)
        assert k == m
        out = ora.net_forward(c, net, hv, film, None, p)
        assert np.allclose(out, h, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- RRS decision
def test_decide_invariants_and_order(ora):
    c = cfg(n_pixels=1000)
    rng = np.random.default_rng(1)
    q = rng.uniform(0.05, 20, 300).astype(np.float32)
    stid = rng.permutation(5000)[:300].astype(np.uint32)
    spec = np.array([4001, 4002, 4003], np.uint32)
    r = ora.rrs_decide(c, q, stid=stid, spec=spec, seed=17, stream=2)
    s = r["s"].astype(np.int64)
    assert r["status"] == 0 and r["overflow"] == 0
    assert r["n_survivors"] == 3 + s.sum()
    expect = np.concatenate([spec, np.repeat(stid, s)])
    assert np.array_equal(r["survivors"], expect)
    assert s.max() <= c.split_cap
    # normalise factor: (P - spec) / sum(ceil q), capped at 1
    tot = np.ceil(q.astype(np.float64)).sum()
    assert r["sum"] == tot
    nf = np.float32(np.float32(1000 - 3) * (np.float32(1.0) / np.float32(tot)))
    assert np.array_equal(r["qn"], (q * min(nf, np.float32(1.0))).astype(np.float32))


def test_decide_expectation_matches_qn(ora):
    """E[s] = q' (roulette / split is unbiased) - checked over many draws of one q'."""
    c = cfg(n_pixels=1 << 30, rrs_no_upscale=1, split_cap=255)
    for qv in (0.05, 0.3, 0.99, 1.0, 2.37, 7.5):
        q = np.full(20000, qv, np.float32)
        r = ora.rrs_decide(c, q, seed=99, stream=int(qv * 100))
        # nf >= 1 here so q' = q; s in {floor, floor+1}
        assert np.all(r["qn"] == np.float32(qv))
        mean = r["s"].astype(np.float64).mean()
        sd = math.sqrt(max(qv - math.floor(qv), 1e-9) * (1 - (qv - math.floor(qv))) / 20000) + 1e-12
        assert abs(mean - qv) < 6 * sd + 1e-9, (qv, mean)
        assert set(np.unique(r["s"])) <= {math.floor(qv), math.floor(qv) + 1}


def test_decide_sum_is_order_independent(ora):
    c = cfg(rrs_use_ceil=0)
    rng = np.random.default_rng(2)
    q = rng.uniform(0.05, 20, 5000).astype(np.float32)
    a = ora.rrs_decide(c, q)["sum"]
    b = ora.rrs_decide(c, q[rng.permutation(5000)])["sum"]
    assert a == b
    assert a == pytest.approx(q.astype(np.float64).sum(), rel=1e-12)


def test_decide_capacity_overflow(ora):
    c = cfg(n_pixels=100)
    q = np.full(50, 4.0, np.float32)
    r = ora.rrs_decide(c, q, capacity=30)
    full = ora.rrs_decide(c, q)
    assert r["status"] == na.ERR_OVERFLOW
    assert r["n_survivors"] == 30
    assert r["overflow"] == full["n_survivors"] - 30
    assert np.array_equal(r["survivors"], full["survivors"][:30])


def test_decide_empty(ora):
    c = cfg()
    r = ora.rrs_decide(c, np.zeros(0, np.float32), spec=np.array([7, 8], np.uint32))
    assert r["n_survivors"] == 2 and r["survivors"].tolist() == [7, 8]


def test_q_modes_closed_form(ora):
    c = cfg(n_pixels=4)
    d = ora.synth_vertices(4, 1, 4)
    hv = ora.HostVertices(d)
    film = np.array([[0.5, 1.0, 0.0, 2.0]] * 3, np.float32)
    raw = np.array([-50.0, 0.0, 3.0, 100.0])
    q = ora.q_f64(c, na.Q_RRSNET, hv, raw, film)
    assert q[0] == pytest.approx(0.05)  # clamp min
    assert q[1] == pytest.approx(math.log(2))
    assert q[2] == pytest.approx(1.5 + math.log(2))
    assert q[3] == pytest.approx(20.0)  # clamp max
    stats = np.ones((6, 4))
    q = ora.q_f64(c, na.Q_STATNET_ADRRS, hv, stats, film)
    thp = d["thp"].astype(np.float64)
    expect = (thp / (film.astype(np.float64) + 0.01)).mean(0)
    assert np.allclose(q, np.clip(expect, 0.05, 20))
