"""Oracle of the per-pixel error pipeline against closed forms (train.cu:344-430)."""
import numpy as np


def test_error_closed_forms(ora):
    n = 6
    ref = np.ones((3, n))
    L = np.ones((3, n))
    ref[:, 1] = 0.0                       # black reference: e = mean(L^2) / 0.01
    L[:, 1] = [0.1, 0.2, 0.3]
    L[:, 2] = 3.0                         # (3-1)^2 / 1.01
    ref[:, 3] = 0.0                       # clamped: 100^2 / 0.01 >> 100
    L[:, 3] = 100.0
    L[:, 4] = [2.0, 1.0, 1.0]             # one channel off: (1/1.01) / 3
    err, total = ora.error_compute(ref, L, spp=np.array([1, 4, 9, 16, 25, 36.0]), clamp=True)
    expect = np.array([0.0, (0.01 + 0.04 + 0.09) / 0.01 / 3, 4.0 / 1.01, 100.0, 1.0 / 1.01 / 3, 0.0])
    assert np.allclose(err[0], expect, rtol=1e-14, atol=0)
    assert np.allclose(err[1], expect * np.sqrt([1, 4, 9, 16, 25, 36.0]), rtol=1e-14, atol=0)
    assert abs(total - err[1].sum()) < 1e-12
    err_nc, _ = ora.error_compute(ref, L, spp=None, clamp=False)
    assert err_nc[0, 3] == (100.0 ** 2) / 0.01 and np.array_equal(err_nc[1], err_nc[0])


def test_error_blend_weights(ora):
    rng = np.random.default_rng(2)
    n = 10
    Lf = rng.uniform(0, 5, (3, n)).astype(np.float32)
    spp = rng.integers(1, 9, n).astype(np.uint32)
    w0 = rng.uniform(0, 1, (3, n))
    w = w0.copy()
    ora.error_blend(Lf, spp, 0.0, w)      # blend 0 keeps the running estimate
    assert np.array_equal(w, w0)
    ora.error_blend(Lf, spp, 1.0, w)      # blend 1 takes the frame mean
    assert np.allclose(w, Lf.astype(np.float64) / spp, rtol=1e-15)
    w = w0.copy()
    ora.error_blend(Lf, spp, 0.25, w)
    assert np.allclose(w, 0.75 * w0 + 0.25 * Lf.astype(np.float64) / spp, rtol=1e-14)
