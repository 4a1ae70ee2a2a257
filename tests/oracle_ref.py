"""ctypes view of the CPU oracle (oracle/_build/libnrrs_oracle.so) for the tests.

The oracle is the checker: tests compare the HIP library against it.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "nrrs-code_amd"))
import nrrs_amd as na  # noqa: E402  (struct definitions only)

ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "libnrrs_oracle.so")

_P = C.c_void_p
_SIGS = {
    "ora_pcg32": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, _P]),
    "ora_grid_layout": (C.c_uint64, [C.POINTER(na.GridConfig), _P, _P, _P, _P, _P]),
    "ora_normalize_pos": (None, [_P, _P, C.c_uint32, _P, C.c_uint32, _P]),
    "ora_grid_indices": (None, [C.POINTER(na.Config), C.c_uint32, C.POINTER(na.Vertices), _P]),
    "ora_activation_sigma": (C.c_double, [C.c_double]),
    "ora_box_cox": (C.c_double, [C.c_double]),
    "ora_weight_window": (C.c_double, [C.c_double, C.c_double]),
    "ora_sh4": (None, [C.c_double, C.c_double, C.c_double, _P]),
    "ora_lat_long": (None, [C.c_double, C.c_double, C.c_double, _P]),
    "ora_one_blob": (None, [C.c_double, C.c_int32, _P]),
    "ora_encode": (None, [C.POINTER(na.Config), C.c_int32, C.c_uint32, C.POINTER(na.Vertices), _P, _P, _P, _P]),
    "ora_encode_train": (None, [C.POINTER(na.Config), C.c_int32, C.c_uint32, C.POINTER(na.Vertices), _P, _P, _P, _P]),
    "ora_net_forward": (None, [C.POINTER(na.Config), C.c_int32, C.c_uint32, C.POINTER(na.Vertices), _P, _P, _P, _P]),
    "ora_param_count": (C.c_uint64, [C.POINTER(na.Config), C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "ora_q": (None, [C.POINTER(na.Config), C.c_int32, C.c_uint32, _P, C.POINTER(na.Vertices), _P, _P]),
    "ora_q_f32": (None, [C.POINTER(na.Config), C.c_int32, C.c_uint32, _P, C.POINTER(na.Vertices), _P, _P]),
    "ora_rrs_decide": (C.c_int32, [C.POINTER(na.Config), C.c_uint32, _P, _P, C.c_uint32, _P, C.c_uint64,
                                   C.c_uint64, _P, _P, _P, C.c_uint32, _P, _P, _P]),
    "ora_synth_vertices": (None, [C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(na.VerticesOut)]),
    "ora_init_params": (None, [C.POINTER(na.Config), C.c_int32, C.c_uint64, C.c_double, _P]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        lib = C.CDLL(ORACLE_LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class HostVertices:
    """Host SoA matching nrrs_vertices (keeps the numpy arrays alive)."""

    def __init__(self, d):
        self.d = {}
        for k in na.VertexBatch.FIELDS3:
            self.d[k] = np.ascontiguousarray(d[k], np.float32)
        self.d["rough"] = np.ascontiguousarray(d["rough"], np.float32)
        self.d["pixel"] = np.ascontiguousarray(d["pixel"], np.uint32)
        self.d["stid"] = None if d.get("stid") is None else np.ascontiguousarray(d["stid"], np.uint32)
        self.n = self.d["rough"].shape[0]
        v = na.Vertices()
        for k in na.VertexBatch.FIELDS3 + ("rough", "pixel"):
            setattr(v, k, self.d[k].ctypes.data)
        v.stid = None if self.d["stid"] is None else self.d["stid"].ctypes.data
        v.stride = self.n
        self.v = v


def pcg32(seed, stream, skip, n):
    out = np.zeros(n, np.uint32)
    load().ora_pcg32(seed, stream, skip, n, out.ctypes.data)
    return out


def grid_layout(grid_cfg):
    L = grid_cfg.n_levels
    res, off, size, dense = (np.zeros(L, np.uint32) for _ in range(4))
    scale = np.zeros(L, np.float32)
    n = load().ora_grid_layout(C.byref(grid_cfg), res.ctypes.data, off.ctypes.data, size.ctypes.data,
                               dense.ctypes.data, scale.ctypes.data)
    return dict(res=res, offset=off, size=size, dense=dense, scale=scale, n_params=int(n))


def synth_vertices(n, seed, n_pixels):
    d = {k: np.zeros((3, n), np.float32) for k in na.VertexBatch.FIELDS3}
    d["rough"] = np.zeros(n, np.float32)
    d["pixel"] = np.zeros(n, np.uint32)
    o = na.VerticesOut()
    for k in na.VertexBatch.FIELDS3 + ("rough", "pixel"):
        setattr(o, k, d[k].ctypes.data)
    o.stride = n
    load().ora_synth_vertices(n, seed, n_pixels, C.byref(o))
    d["stid"] = None
    return d


def lat_long(d):
    uv = np.zeros(2, np.float64)
    load().ora_lat_long(float(d[0]), float(d[1]), float(d[2]), uv.ctypes.data)
    return uv


def one_blob(x, bins):
    out = np.zeros(bins, np.float64)
    load().ora_one_blob(float(x), bins, out.ctypes.data)
    return out


def param_count(cfg, net):
    g, m = C.c_uint64(), C.c_uint64()
    load().ora_param_count(C.byref(cfg), net, C.byref(g), C.byref(m))
    return int(g.value), int(m.value)


def init_params(cfg, net, seed, grid_amp):
    g, m = param_count(cfg, net)
    out = np.zeros(g + m, np.float64)
    load().ora_init_params(C.byref(cfg), net, seed, grid_amp, out.ctypes.data)
    return out


def grid_indices(cfg, hv):
    idx = np.zeros((hv.n, cfg.grid.n_levels, 8), np.uint32)
    load().ora_grid_indices(C.byref(cfg), hv.n, C.byref(hv.v), idx.ctypes.data)
    return idx


def encode(cfg, net, hv, film, stats, params, training=False):
    """Network input [n][80]; `training`: the input of a training step (with train_sigma the
    RRSNet L2 channels take activationSigma(max(L2, 0)) there)."""
    enc = np.zeros((hv.n, na.ENC_WIDTH), np.float64)
    film = None if film is None else np.ascontiguousarray(film, np.float32)
    stats = None if stats is None else np.ascontiguousarray(stats, np.float64)
    params = np.ascontiguousarray(params, np.float64)
    fn = load().ora_encode_train if training else load().ora_encode
    fn(C.byref(cfg), net, hv.n, C.byref(hv.v), _p(film), _p(stats), params.ctypes.data, enc.ctypes.data)
    return enc


def net_forward(cfg, net, hv, film, stats, params):
    n_out = cfg.rrsnet.n_outputs if net == na.RRSNET else cfg.statnet.n_outputs
    out = np.zeros((hv.n, n_out), np.float64)
    film = None if film is None else np.ascontiguousarray(film, np.float32)
    stats = None if stats is None else np.ascontiguousarray(stats, np.float64)
    params = np.ascontiguousarray(params, np.float64)
    load().ora_net_forward(C.byref(cfg), net, hv.n, C.byref(hv.v), _p(film), _p(stats), params.ctypes.data,
                           out.ctypes.data)
    return out


def q_f64(cfg, mode, hv, src, film):
    src = np.ascontiguousarray(src, np.float64)
    q = np.zeros(hv.n, np.float64)
    load().ora_q(C.byref(cfg), mode, hv.n, src.ctypes.data, C.byref(hv.v),
                 np.ascontiguousarray(film, np.float32).ctypes.data, q.ctypes.data)
    return q


def q_f32(cfg, mode, hv, src, film):
    src = np.ascontiguousarray(src, np.float32)
    q = np.zeros(hv.n, np.float32)
    film = np.ascontiguousarray(film, np.float32)
    load().ora_q_f32(C.byref(cfg), mode, hv.n, src.ctypes.data, C.byref(hv.v), film.ctypes.data, q.ctypes.data)
    return q


def rrs_decide(cfg, q, stid=None, spec=None, seed=0, stream=0, capacity=None):
    q = np.ascontiguousarray(q, np.float32)
    n = q.shape[0]
    spec = np.zeros(0, np.uint32) if spec is None else np.ascontiguousarray(spec, np.uint32)
    stid = None if stid is None else np.ascontiguousarray(stid, np.uint32)
    if capacity is None:
        capacity = len(spec) + n * cfg.split_cap
    qn = np.zeros(n, np.float32)
    s = np.zeros(n, np.uint8)
    surv = np.zeros(max(capacity, 1), np.uint32)
    n_surv, ovf = C.c_uint32(), C.c_uint32()
    total = C.c_double()
    st = load().ora_rrs_decide(C.byref(cfg), n, q.ctypes.data, _p(stid), len(spec), spec.ctypes.data, seed, stream,
                               qn.ctypes.data, s.ctypes.data, surv.ctypes.data, capacity, C.byref(n_surv),
                               C.byref(ovf), C.byref(total))
    return dict(status=st, qn=qn, s=s, survivors=surv[: n_surv.value], n_survivors=n_surv.value,
                overflow=ovf.value, sum=total.value)


_SIGS_TRAIN = {
    "ora_loss": (C.c_double, [C.POINTER(na.Config), C.c_int32, C.c_uint32, _P, _P, C.POINTER(na.Vertices), _P, _P,
                              _P, _P, _P, _P, C.c_double, _P, _P]),
    "ora_train_grads": (C.c_double, [C.POINTER(na.Config), C.c_int32, C.c_uint32, C.POINTER(na.Vertices), _P, _P,
                                     _P, C.c_int32, _P, _P, _P, _P, _P, _P, C.c_double, C.c_double, _P, _P]),
    "ora_adam": (None, [C.c_uint64, _P, _P, _P, _P, C.c_double, C.c_double, C.c_double, C.c_double,
                        C.c_double, C.c_uint32]),
    "ora_ema_step": (None, [C.c_uint64, _P, _P, C.c_double, C.c_uint32]),
}


def _load_train():
    lib = load()
    if not getattr(lib, "_train_sigs", False):
        for name, (res, args) in _SIGS_TRAIN.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._train_sigs = True
    return lib


class RrsSide:
    """Host side arrays of the RRS loss (the nrrs_rrs_side fields)."""

    def __init__(self, ref_mean, error, sample_weight, q_target, error_sum):
        self.ref_mean = None if ref_mean is None else np.ascontiguousarray(ref_mean, np.float32)
        self.error = np.ascontiguousarray(error, np.float32)
        self.sample_weight = np.ascontiguousarray(sample_weight, np.float32)
        self.q_target = np.ascontiguousarray(q_target, np.float32)
        self.error_sum = float(error_sum)


def _side_ptrs(side):
    if side is None:
        return None, None, None, None, 0.0
    return _p(side.ref_mean), _p(side.error), _p(side.sample_weight), _p(side.q_target), side.error_sum


def bind_loss(lib):
    """Declares ora_loss on a loaded oracle library (the oracle itself, or a test variant)."""
    name, (res, args) = "ora_loss", _SIGS_TRAIN["ora_loss"]
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
    return lib


def loss(cfg, kind, hv, y, film=None, target_L=None, stats=None, side=None, y_den=None, lib=None):
    """y: [n_out][n] (FP64). Returns (loss, dy). `lib`: another build of the oracle (bind_loss)."""
    y = np.ascontiguousarray(y, np.float64)
    dy = np.zeros_like(y)
    rm, er, sw, qt, es = _side_ptrs(side)
    film = None if film is None else np.ascontiguousarray(film, np.float32)
    target_L = None if target_L is None else np.ascontiguousarray(target_L, np.float32)
    stats = None if stats is None else np.ascontiguousarray(stats, np.float64)
    y_den = None if y_den is None else np.ascontiguousarray(y_den, np.float64)
    lib = lib if lib is not None else _load_train()
    val = lib.ora_loss(C.byref(cfg), kind, hv.n, y.ctypes.data, _p(target_L), C.byref(hv.v), _p(film), _p(stats), rm,
                       er, sw, qt, es, _p(y_den), dy.ctypes.data)
    return val, dy


def train_grads(cfg, net, hv, params, kind, film=None, stats=None, dy_ext=None, target_L=None, side=None,
                loss_scale=1.0):
    """FP64 (loss, grad [n_params], dy [n_out][n])."""
    g, m = param_count(cfg, net)
    n_out = cfg.rrsnet.n_outputs if net == na.RRSNET else cfg.statnet.n_outputs
    grad = np.zeros(g + m, np.float64)
    dy = np.zeros((n_out, hv.n), np.float64)
    params = np.ascontiguousarray(params, np.float64)
    film = None if film is None else np.ascontiguousarray(film, np.float32)
    stats = None if stats is None else np.ascontiguousarray(stats, np.float64)
    dy_ext = None if dy_ext is None else np.ascontiguousarray(dy_ext, np.float64)
    target_L = None if target_L is None else np.ascontiguousarray(target_L, np.float32)
    rm, er, sw, qt, es = _side_ptrs(side)
    val = _load_train().ora_train_grads(C.byref(cfg), net, hv.n, C.byref(hv.v), _p(film), _p(stats),
                                        params.ctypes.data, kind, _p(dy_ext), _p(target_L), rm, er, sw, qt, es,
                                        loss_scale, grad.ctypes.data, dy.ctypes.data)
    return val, grad, dy


def adam(p, m, v, g, lr, b1, b2, eps, l2, t):
    for a in (p, m, v, g):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    _load_train().ora_adam(p.size, p.ctypes.data, m.ctypes.data, v.ctypes.data, g.ctypes.data, lr, b1, b2, eps,
                           l2, t)


def ema_step(ema, p, decay, t):
    """Ema moving average after optimizer step t (in place on a float64 array)."""
    assert ema.dtype == np.float64 and ema.flags.c_contiguous
    p = np.ascontiguousarray(p, np.float64)
    _load_train().ora_ema_step(ema.size, ema.ctypes.data, p.ctypes.data, decay, t)


def synth_targets(n, seed):
    """Radiance targets: log-normal(-1, 1.5) per channel, 5% of samples x100."""
    rng = np.random.default_rng(seed)
    L = rng.lognormal(-1.0, 1.5, (3, n))
    L[:, rng.uniform(size=n) < 0.05] *= 100.0
    return L.astype(np.float32)


def error_blend(L_frame, spp, blend, weighted_L):
    """Running per-pixel radiance blend (in place on a float64 [3][n] array)."""
    assert weighted_L.dtype == np.float64 and weighted_L.flags.c_contiguous
    Lf = np.ascontiguousarray(L_frame, np.float32)
    sp = np.ascontiguousarray(spp, np.uint32)
    lib = load()
    lib.ora_error_blend.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
    lib.ora_error_blend(weighted_L.shape[1], Lf.ctypes.data, sp.ctypes.data, blend, weighted_L.ctypes.data)


def error_compute(ref, weighted_L, spp=None, clamp=True):
    """Per-pixel error [2][n] and the sum of row 1 (FP64)."""
    ref = np.ascontiguousarray(ref, np.float64)
    wl = np.ascontiguousarray(weighted_L, np.float64)
    sp = None if spp is None else np.ascontiguousarray(spp, np.float64)
    n = ref.shape[1]
    err = np.zeros((2, n), np.float64)
    lib = load()
    lib.ora_error_compute.restype = C.c_double
    lib.ora_error_compute.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    total = lib.ora_error_compute(n, ref.ctypes.data, wl.ctypes.data, None if sp is None else sp.ctypes.data,
                                  int(clamp), err.ctypes.data)
    return err, total


class _OraNodes(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("pos", "dir", "normal", "rough", "ad", "as_", "thp", "q", "L", "parent",
                                          "pixel", "depth")] + [("cap", C.c_uint32), ("size", C.c_uint32),
                                                                ("invalid", C.c_uint32)]


class Nodes:
    """Host path-node buffer of the oracle (numpy planes of capacity cap)."""

    def __init__(self, cap):
        self.cap = cap
        f3 = lambda: np.zeros((3, cap), np.float32)  # noqa: E731
        self.a = dict(pos=f3(), dir=f3(), normal=f3(), ad=f3(), as_=f3(), thp=f3(), rough=np.zeros(cap, np.float32),
                      q=np.zeros(cap, np.float32), L=np.zeros((3, cap), np.float64),
                      parent=np.zeros(cap, np.int32), pixel=np.zeros(cap, np.uint32), depth=np.zeros(cap, np.uint32))
        self.s = _OraNodes()
        for k, v in self.a.items():
            setattr(self.s, k, v.ctypes.data)
        self.s.cap = cap

    @property
    def size(self):
        return self.s.size

    @property
    def invalid(self):
        return self.s.invalid

    def record(self, hv, parent=None, q=None, depth=0):
        lib = load()
        lib.ora_nodes_record.argtypes = [C.POINTER(_OraNodes), C.c_uint32, C.POINTER(na.Vertices), C.c_void_p,
                                         C.c_void_p, C.c_uint32, C.c_void_p]
        par = None if parent is None else np.ascontiguousarray(parent, np.int32)
        qq = None if q is None else np.ascontiguousarray(q, np.float32)
        out = np.zeros(hv.n, np.int32)
        lib.ora_nodes_record(C.byref(self.s), hv.n, C.byref(hv.v), _p(par), _p(qq), depth, out.ctypes.data)
        return out

    def add_radiance(self, node, L):
        lib = load()
        lib.ora_nodes_radiance.argtypes = [C.POINTER(_OraNodes), C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32]
        node = np.ascontiguousarray(node, np.int32)
        L = np.ascontiguousarray(L, np.float32)
        lib.ora_nodes_radiance(C.byref(self.s), node.size, node.ctypes.data, L.ctypes.data, L.shape[1])

    def training_data(self, film, capacity, err=None, scale=1, width=1, ignore_zero_L=False, copy_rrs=False,
                      mark_dup=True, pixel_samples=None):
        lib = load()
        lib.ora_training_data.restype = C.c_uint32
        lib.ora_training_data.argtypes = [C.POINTER(_OraNodes), C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                          C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, C.c_uint32,
                                          C.POINTER(na.VerticesOut)] + [C.c_void_p] * 7
        film = np.ascontiguousarray(film, np.float32)
        P = film.shape[1]
        e = None if err is None else np.ascontiguousarray(err, np.float32)
        o = {k: np.zeros((3, capacity), np.float32) for k in na.VertexBatch.FIELDS3}
        o["rough"] = np.zeros(capacity, np.float32)
        o["pixel"] = np.zeros(capacity, np.uint32)
        vo = na.VerticesOut()
        for k in na.VertexBatch.FIELDS3 + ("rough", "pixel"):
            setattr(vo, k, o[k].ctypes.data)
        vo.stride = capacity
        r = dict(target_L=np.zeros((3, capacity)), ref_mean=np.zeros(capacity), error=np.zeros(capacity, np.float32),
                 error_avg=np.zeros(capacity, np.float32), q_target=np.zeros(capacity, np.float32),
                 sample_weight=np.zeros(capacity),
                 pixel_samples=np.zeros(P, np.uint32) if pixel_samples is None else
                 np.ascontiguousarray(pixel_samples, np.uint32))
        n = lib.ora_training_data(C.byref(self.s), film.ctypes.data, P, _p(e), 0 if e is None else e.shape[1], scale,
                                  width, int(ignore_zero_L), int(copy_rrs), int(mark_dup), capacity, C.byref(vo),
                                  *(r[k].ctypes.data for k in ("target_L", "ref_mean", "error", "error_avg",
                                                               "q_target", "sample_weight", "pixel_samples")))
        r["n"] = n
        r["v"] = o
        return r


def consume_survivors(surv, hv_scatter, rrs, node_in=None):
    """Oracle rows of the survivor consumer: dict of planes and node_out."""
    lib = load()
    lib.ora_consume_survivors.argtypes = [C.c_uint32, C.c_void_p, C.POINTER(na.Vertices), C.c_void_p, C.c_void_p,
                                          C.POINTER(na.VerticesOut), C.c_void_p]
    surv = np.ascontiguousarray(surv, np.uint32)
    n = surv.size
    o = {k: np.zeros((3, max(n, 1)), np.float32) for k in na.VertexBatch.FIELDS3}
    o["rough"] = np.zeros(max(n, 1), np.float32)
    o["pixel"] = np.zeros(max(n, 1), np.uint32)
    vo = na.VerticesOut()
    for k in na.VertexBatch.FIELDS3 + ("rough", "pixel"):
        setattr(vo, k, o[k].ctypes.data)
    vo.stride = max(n, 1)
    rrs = np.ascontiguousarray(rrs, np.float32)
    ni = None if node_in is None else np.ascontiguousarray(node_in, np.int32)
    node_out = np.zeros(max(n, 1), np.int32)
    lib.ora_consume_survivors(n, surv.ctypes.data, C.byref(hv_scatter.v), rrs.ctypes.data, _p(ni), C.byref(vo),
                              node_out.ctypes.data)
    return o, node_out


def cpu_step_f32(cfg, hv, film, ps, pr, target_L, err, sw, qt, error_sum, threads, grads=True):
    """FP32 multi-threaded CPU step; returns q, gradients, updated parameters and the times."""
    lib = load()
    lib.ora_cpu_step_f32.restype = C.c_double
    lib.ora_cpu_step_f32.argtypes = [C.POINTER(na.Config), C.c_uint32, C.POINTER(na.Vertices)] + [C.c_void_p] * 7 + \
        [C.c_double, C.c_int32] + [C.c_void_p] * 4
    n = hv.n
    ps32 = np.ascontiguousarray(ps, np.float32).copy()
    pr32 = np.ascontiguousarray(pr, np.float32).copy()
    f = lambda a: np.ascontiguousarray(a, np.float32)  # noqa: E731
    film, tL, er, w, q_t = f(film), f(target_L), f(err), f(sw), f(qt)
    q = np.zeros(n, np.float32)
    gs = np.zeros(ps32.size, np.float32) if grads else None
    gr = np.zeros(pr32.size, np.float32) if grads else None
    times = np.zeros(2)
    total = lib.ora_cpu_step_f32(C.byref(cfg), n, C.byref(hv.v), film.ctypes.data, ps32.ctypes.data, pr32.ctypes.data,
                                 tL.ctypes.data, er.ctypes.data, w.ctypes.data, q_t.ctypes.data, error_sum, threads,
                                 q.ctypes.data, _p(gs), _p(gr), times.ctypes.data)
    return dict(q=q, grad_stat=gs, grad_rrs=gr, params_stat=ps32, params_rrs=pr32, times=times, total=total)
