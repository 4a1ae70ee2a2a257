"""Oracle of the path-node buffer and of the training rows built from it, on a hand-built
path tree (workqueue.h:106-189, train.cu:432-592)."""
import numpy as np

import nrrs_amd as na


def _verts(n, pixel, thp, pos_nan=()):
    d = {k: np.zeros((3, n), np.float32) for k in na.VertexBatch.FIELDS3}
    d["pos"][:] = np.arange(3 * n, dtype=np.float32).reshape(3, n) / 10.0
    d["dir"][2] = 1.0
    d["normal"][2] = 1.0
    d["thp"][:] = np.asarray(thp, np.float32).reshape(3, n) if np.ndim(thp) == 2 else thp
    for i in pos_nan:
        d["pos"][0, i] = np.nan
    d["rough"] = np.full(n, 0.5, np.float32)
    d["pixel"] = np.asarray(pixel, np.uint32)
    return d


def test_tree_radiance_and_rows(ora):
    """Tree: roots 0 (pixel 0) and 1 (pixel 1); 0 has children 2 and 3 (a split), 2 has child 4.
    Radiance: 1.0 found at node 4, 0.5 at node 3, 2.0 at node 1."""
    nb = ora.Nodes(8)
    n0 = nb.record(ora.HostVertices(_verts(2, [0, 1], 1.0)), q=[1.5, 0.5], depth=0)
    assert list(n0) == [0, 1]
    n1 = nb.record(ora.HostVertices(_verts(2, [0, 0], 0.5)), parent=[0, 0], q=[2.0, 2.0], depth=1)
    assert list(n1) == [2, 3]
    n2 = nb.record(ora.HostVertices(_verts(1, [0], 0.25)), parent=[2], q=[1.0], depth=2)
    assert list(n2) == [4] and nb.size == 5
    L = np.array([[1.0, 0.5, 2.0]] * 3, np.float32)
    nb.add_radiance(np.array([4, 3, 1], np.int32), L)
    Ln = nb.a["L"][0, :5]
    assert list(Ln) == [1.5, 2.0, 1.0, 0.5, 1.0]     # node 0 gets 1.0 (via 4 -> 2 -> 0) and 0.5 (via 3)
    film = np.ones((3, 2), np.float32) * np.array([[3.0], [6.0], [9.0]], np.float32)
    r = nb.training_data(film, 8)
    assert r["n"] == 5
    # target L = L / thp
    assert np.allclose(r["target_L"][0, :5], [1.5, 2.0, 2.0, 1.0, 4.0])
    # node 3 is the second child of node 0: a duplicate for the RRSNet (q target -1)
    assert list(r["q_target"][:5]) == [1.5, 0.5, 2.0, -1.0, 1.0]
    # pixel 0 has nodes 0, 2, 4 counted (3 is a duplicate), pixel 1 has node 1
    assert list(r["pixel_samples"]) == [3, 1]
    assert np.allclose(r["sample_weight"][:5], [1 / 3, 1.0, 1 / 3, 1 / 3, 1 / 3])
    assert np.allclose(r["ref_mean"][:5], 6.0)
    # copy_rrs_node multiplies thp by q; without duplicate marking node 3 keeps its q
    r2 = nb.training_data(film, 8, copy_rrs=True, mark_dup=False)
    assert np.allclose(r2["v"]["thp"][0, :5], [1.5, 0.5, 1.0, 1.0, 0.25])
    assert list(r2["q_target"][:5]) == [1.5, 0.5, 2.0, 2.0, 1.0]


def test_zero_radiance_nan_capacity(ora):
    nb = ora.Nodes(4)
    thp = np.array([[1, 0, 1, 1, 1], [1, 0, 1, 1, 1], [1, 0, 1, 1, 1]], np.float32)
    out = nb.record(ora.HostVertices(_verts(5, [0, 1, 2, 3, 0], thp, pos_nan=(2,))), parent=[-1, -1, -1, -1, 7])
    # capacity 4: the fifth record keeps its parent and counts as invalid
    assert list(out) == [0, 1, 2, 3, 7] and nb.size == 4 and nb.invalid == 1
    nb.add_radiance(np.array([1, 2], np.int32), np.ones((3, 2), np.float32))
    film = np.ones((3, 4), np.float32)
    r = nb.training_data(film, 8)
    # node 2 has a NaN position; node 1 has zero throughput, so its L is 0 but it is kept
    assert r["n"] == 3 and list(r["v"]["pixel"][:3]) == [0, 1, 3]
    assert np.array_equal(r["target_L"][:, 1], [0.0, 0.0, 0.0])
    r = nb.training_data(film, 8, ignore_zero_L=True)
    assert r["n"] == 0
    r = nb.training_data(film, 2)
    assert r["n"] == 2 and list(r["v"]["pixel"][:2]) == [0, 1]


def test_error_rows_use_scaled_pixel(ora):
    nb = ora.Nodes(4)
    W = 4  # 4 x 2 frame, error image 2 x 1 (scale 2)
    nb.record(ora.HostVertices(_verts(3, [0, 3, 6], 1.0)))
    err = np.array([[0.1, 0.2], [1.1, 1.2]], np.float32)
    r = nb.training_data(np.ones((3, 8), np.float32), 4, err=err, scale=2, width=W)
    # pixel 0 -> (0,0) -> 0; pixel 3 -> (1,0) -> 1; pixel 6 -> (3,1)/2 = (1,0) -> 1
    assert list(r["error"][:3]) == [np.float32(0.1), np.float32(0.2), np.float32(0.2)]
    assert list(r["error_avg"][:3]) == [np.float32(1.1), np.float32(1.2), np.float32(1.2)]
