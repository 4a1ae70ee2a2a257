"""Pins the oracle's NRRS_RRS loss (restated: the reference's nrrs_rrs.h lives in the absent
tiny-cuda-nn fork) against its documented definition, per term and through properties that do
not restate it:

    q = activationSigma(y),  v = mean_c(thp_c^2 max(L2_c, 0)) / (ref^2 + 0.01), clamped at
    loss_clamp_max when loss_clamp_on,  E = max(error_sum / pixels, 1e-4),  e = max(error, 1e-4)
    f(q) = g1 (v / (q + 0.01) + q E) + g2 (v / (q + 0.01) + q e) + g3 q + g4 (q - q_t)^2
    loss = (1/n) sum over samples with q_t >= 0 of sample_weight f(q)

- known answers: one sample, each gamma term alone, hand-evaluated;
- minimisers: g1 alone is minimal at q* = sqrt(v / E) - 0.01, g2 alone at sqrt(v / e) - 0.01,
  g4 alone at q_t; found by bisection on the oracle's own dL/dy, so they pin v (the thp^2
  weights, the channel mean, ref^2, the clamp) and the floors of E and e through the loss
  gradient, independently of the value;
- g3 alone has dL/dq = sample_weight / n;
- the reference mean comes from the film pixel when no ref_mean array is given.

Each plausible mistake in the restatement (thp for thp^2, a lost 1/3, ref for ref^2, another
floor, no clamp, swapped efficiency terms, a lost +0.01, an unweighted sample) is planted in a
copy of the oracle source, compiled on the CPU, and must fail at least one of these checks."""
import ctypes as C
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

import nrrs_amd as na

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = 4  # pixels of the film


def _cfg(gamma, **kw):
    return na.default_config(n_pixels=P, gamma=gamma, loss_clamp_on=kw.pop("clamp_on", 1),
                             loss_clamp_max=kw.pop("clamp_max", 10.0), grid=dict(log2_table_size=12), **kw)


def _vertex(ora, thp):
    d = ora.synth_vertices(1, 3, P)
    d["thp"][:, 0] = thp
    d["pixel"][0] = 2
    return ora.HostVertices(d)


def _stats(l2):
    s = np.zeros((6, 1))
    s[3:, 0] = l2
    return s


def _film(rgb_of_pixel2=(0.3, 0.9, 0.9)):
    f = np.full((3, P), 5.0, np.float32)
    f[:, 2] = rgb_of_pixel2
    return f


def _side(ora, ref=0.7, error=0.3, sw=0.5, qt=0.2, error_sum=2.0):
    return ora.RrsSide(None if ref is None else np.array([ref], np.float32), np.array([error], np.float32),
                       np.array([sw], np.float32), np.array([qt], np.float32), error_sum)


BASE = dict(thp=(2.0, 1.0, 0.5), l2=(0.25, 3.0, 4.0))


def _loss(ora, lib, cfg, y, thp=BASE["thp"], l2=BASE["l2"], film=None, **side):
    hv = _vertex(ora, thp)
    return ora.loss(cfg, na.LOSS_RRS, hv, np.array([[y]]), film=_film() if film is None else film, stats=_stats(l2),
                    side=_side(ora, **side), lib=lib)


def _sigma(y):
    return math.log1p(math.exp(y)) if y < 0 else 0.5 * y + math.log(2.0)


def _argmin_q(ora, lib, cfg, **kw):
    """q at the zero of dL/dy (the loss is convex in q on these cases), by bisection in y."""
    lo, hi = -30.0, 4000.0
    g = lambda y: _loss(ora, lib, cfg, y, **kw)[1][0, 0]  # noqa: E731
    if not (g(lo) < 0 < g(hi)):
        return float("nan")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if g(mid) < 0:
            lo = mid
        else:
            hi = mid
    return _sigma(0.5 * (lo + hi))


def rrs_checks(ora, lib=None):
    """Names of the failed checks (empty when the loss behaves as documented)."""
    bad = []
    f32 = lambda x: float(np.float32(x))  # noqa: E731  (side arrays are FP32)
    ref = f32(0.7)
    v = (2.0 ** 2 * 0.25 + 1.0 ** 2 * 3.0 + 0.5 ** 2 * 4.0) / 3.0 / (ref * ref + 0.01)  # 10/3 up to FP32 ref
    q = math.log(2.0)  # y = 0
    E, e, sw, qt = 2.0 / P, f32(0.3), 0.5, f32(0.2)
    want = {0: v / (q + 0.01) + q * E, 1: v / (q + 0.01) + q * e, 2: q, 3: (q - qt) ** 2}
    for k in range(4):
        gam = [0.0] * 4
        gam[k] = 1.0
        got = _loss(ora, lib, _cfg(gam), 0.0)[0]
        if not math.isclose(got, sw * want[k], rel_tol=1e-9):
            bad.append(f"value g{k + 1}: {got} != {sw * want[k]}")
    g1, g2, g4 = [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1]
    cases = [  # (name, gamma, kwargs, expected q*)
        ("g1 minimum", g1, {}, math.sqrt(v / E) - 0.01),
        ("g1 E floor", g1, dict(error_sum=0.0), math.sqrt(v / 1e-4) - 0.01),
        ("g1 thp^2", g1, dict(thp=(4.0, 2.0, 1.0), clamp_on=0), math.sqrt(4 * v / E) - 0.01),
        ("g1 channel mean", g1, dict(thp=(0.0, 0.0, 2.0), l2=(7.0, 7.0, 1.0)),
         math.sqrt(4.0 / 3.0 / (ref * ref + 0.01) / E) - 0.01),
        ("g1 clamp", g1, dict(thp=(20.0, 20.0, 20.0)), math.sqrt(10.0 / E) - 0.01),
        ("g1 no clamp", g1, dict(thp=(20.0, 20.0, 20.0), clamp_on=0),
         math.sqrt(400.0 * (0.25 + 3.0 + 4.0) / 3.0 / (ref * ref + 0.01) / E) - 0.01),
        ("g1 negative L2", g1, dict(l2=(0.25, 3.0, -4.0)),
         math.sqrt((1.0 + 3.0) / 3.0 / (ref * ref + 0.01) / E) - 0.01),
        ("g1 ref from film", g1, dict(ref=None), None),
        ("g2 minimum", g2, dict(error=0.25), math.sqrt(v / 0.25) - 0.01),
        ("g2 e floor", g2, dict(error=0.0), math.sqrt(v / 1e-4) - 0.01),
        ("g4 minimum", g4, dict(qt=1.7), f32(1.7)),
    ]
    for name, gam, kw, q_star in cases:
        kw = dict(kw)
        cfg = _cfg(gam, clamp_on=kw.pop("clamp_on", 1))
        got = _argmin_q(ora, lib, cfg, **kw)
        if q_star is None:  # film pixel 2 holds (0.3, 0.9, 0.9): mean 0.7, as ref_mean above
            q_star = math.sqrt(((2.0 ** 2 * 0.25 + 3.0 + 0.25 * 4.0) / 3.0) /
                               (f32(f32(f32(0.3) + f32(0.9)) + f32(0.9)) ** 2 / 9 + 0.01) / E) - 0.01
            tol = 1e-6
        else:
            tol = 1e-9
        if not math.isclose(got, q_star, rel_tol=tol):
            bad.append(f"{name}: q* {got} != {q_star}")
    # g3 alone: dL/dq = sample_weight / n, so dL/dy = sw * sigma'(y)
    for y, ds in ((-1.0, math.exp(-1.0) / (1.0 + math.exp(-1.0))), (1.0, 0.5)):
        dy = _loss(ora, lib, _cfg([0, 0, 1, 0]), y)[1][0, 0]
        if not math.isclose(dy, sw * ds, rel_tol=1e-9):
            bad.append(f"g3 slope at {y}: {dy} != {sw * ds}")
    return bad


def test_rrs_loss_pinned(ora):
    bad = rrs_checks(ora)
    assert not bad, bad


# Plausible mistakes, each a (find, replace) on the oracle source's loss_rrs.
MUTANTS = {
    "thp instead of thp^2": ("m2 += t * t * l2;", "m2 += t * l2;"),
    "channel sum instead of mean": ("double vv = m2 / 3.0 / (ref * ref + 0.01);",
                                    "double vv = m2 / (ref * ref + 0.01);"),
    "ref instead of ref^2": ("double vv = m2 / 3.0 / (ref * ref + 0.01);", "double vv = m2 / 3.0 / (ref + 0.01);"),
    "E floor 1e-3": ("std::fmax(error_sum / (double)cfg->n_pixels, 1e-4)",
                     "std::fmax(error_sum / (double)cfg->n_pixels, 1e-3)"),
    "no e floor": ("double e = std::fmax((double)err[i], 1e-4);", "double e = (double)err[i];"),
    "no clamp of v": ("if (cfg->loss_clamp_on) vv = std::fmin(vv, (double)cfg->loss_clamp_max);", ""),
    "efficiency terms swapped": ("gam[0] * (vv / qe + q * Ebar) + gam[1] * (vv / qe + q * e)",
                                 "gam[0] * (vv / qe + q * e) + gam[1] * (vv / qe + q * Ebar)"),
    "no +0.01 on q": ("double qe = q + 0.01;", "double qe = q;"),
    "sample weight dropped": ("loss += sw[i] * f / n;", "loss += f / n;"),
    "negative L2 kept": ("double l2 = std::fmax(stats[(size_t)(3 + c) * n + i], 0.0);",
                         "double l2 = stats[(size_t)(3 + c) * n + i];"),
}


@pytest.fixture(scope="module")
def mutant_dir():
    with tempfile.TemporaryDirectory() as d:
        yield d


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_fails_a_check(ora, mutant_dir, name):
    src = open(os.path.join(ROOT, "oracle", "nrrs_oracle.cpp")).read()
    find, repl = MUTANTS[name]
    assert src.count(find) >= 1, f"mutation site not found: {find}"
    mutated = src.replace(find, repl)
    tag = "".join(ch if ch.isalnum() else "_" for ch in name)
    cpp = os.path.join(mutant_dir, f"{tag}.cpp")
    so = os.path.join(mutant_dir, f"{tag}.so")
    with open(cpp, "w") as f:
        f.write(mutated)
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
                    f"-I{os.path.join(ROOT, 'oracle')}", "-shared", cpp, "-o", so], check=True)
    lib = ora.bind_loss(C.CDLL(so))
    bad = rrs_checks(ora, lib)
    print(f"{name}: {len(bad)} failed checks, e.g. {bad[:2]}")
    assert bad, f"the planted mistake '{name}' passes every check"
