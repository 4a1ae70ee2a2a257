"""ctypes binding of libnrrs_amd.so (the C ABI declared in include/nrrs_amd.h).

This is the harness-side view of the library: torch tensors provide device
memory and streams, every computation runs in the HIP library. There is no
CPU fallback: loading fails loudly when the shared library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libnrrs_amd.so")

OK, ERR_ARG, ERR_SHAPE, ERR_OOM, ERR_OVERFLOW, ERR_NONFINITE, ERR_HIP, ERR_UNSUPPORTED, ERR_NO_DEVICE = range(9)
FP32, FP16, BF16 = 0, 1, 2
STATNET, RRSNET, STATNET2 = 0, 1, 2  # STATNET2: the second StatNet of AllSelf (ll2_two_head 3)
Q_RRSNET, Q_RRSNET_CHAINED, Q_STATNET_ADRRS = 0, 1, 2
NETS_BENCH, NETS_SHIPPED = 0, 1  # nrrs_config.net_set
ENC_WIDTH = 80


class GridConfig(C.Structure):
    _fields_ = [
        ("n_levels", C.c_uint32),
        ("n_features", C.c_uint32),
        ("log2_table_size", C.c_uint32),
        ("base_resolution", C.c_uint32),
        ("per_level_scale", C.c_float),
    ]


class MlpConfig(C.Structure):
    _fields_ = [("n_hidden_layers", C.c_uint32), ("n_outputs", C.c_uint32)]


class Config(C.Structure):
    _fields_ = [
        ("grid", GridConfig),
        ("statnet", MlpConfig),
        ("rrsnet", MlpConfig),
        ("precision", C.c_int32),
        ("rrs_clamp_min", C.c_float),
        ("rrs_clamp_max", C.c_float),
        ("rrs_normalize_rate", C.c_float),
        ("rrs_use_ceil", C.c_int32),
        ("rrs_no_upscale", C.c_int32),
        ("rrs_weight_window", C.c_int32),
        ("weight_window_size", C.c_float),
        ("split_cap", C.c_uint32),
        ("loss_clamp_on", C.c_int32),
        ("loss_clamp_max", C.c_float),
        ("train_sigma", C.c_int32),
        ("ll2_step", C.c_uint32),
        ("rrs_step", C.c_uint32),
        ("gamma", C.c_float * 4),
        ("lr_statnet", C.c_float),
        ("lr_rrsnet", C.c_float),
        ("adam_beta1", C.c_float),
        ("adam_beta2", C.c_float),
        ("adam_eps", C.c_float),
        ("adam_l2_reg", C.c_float),
        ("loss_scale_init", C.c_float),
        ("loss_scale_growth_interval", C.c_uint32),
        ("aabb_min", C.c_float * 3),
        ("aabb_max", C.c_float * 3),
        ("n_pixels", C.c_uint32),
        ("max_queue", C.c_uint32),
        ("max_train", C.c_uint32),
        ("seed", C.c_uint64),
        ("max_nodes", C.c_uint32),
        ("paired_tables", C.c_uint32),
        ("fp32_level_major", C.c_uint32),
        ("ll2_two_head", C.c_int32),
        ("rrs_aid", C.c_int32),
        ("net_set", C.c_int32),
        ("ema_decay", C.c_float),
    ]


class Vertices(C.Structure):
    _fields_ = [
        ("pos", C.c_void_p),
        ("dir", C.c_void_p),
        ("normal", C.c_void_p),
        ("rough", C.c_void_p),
        ("albedo_d", C.c_void_p),
        ("albedo_s", C.c_void_p),
        ("thp", C.c_void_p),
        ("pixel", C.c_void_p),
        ("stid", C.c_void_p),
        ("stride", C.c_uint32),
    ]


class VerticesOut(C.Structure):
    _fields_ = [
        ("pos", C.c_void_p),
        ("dir", C.c_void_p),
        ("normal", C.c_void_p),
        ("rough", C.c_void_p),
        ("albedo_d", C.c_void_p),
        ("albedo_s", C.c_void_p),
        ("thp", C.c_void_p),
        ("pixel", C.c_void_p),
        ("stride", C.c_uint32),
    ]


class DecideArgs(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("v", Vertices),
        ("film_dev", C.c_void_p),
        ("q_mode", C.c_int32),
        ("spec_stid_dev", C.c_void_p),
        ("n_spec_dev", C.c_void_p),
        ("pcg_seed", C.c_uint64),
        ("pcg_stream", C.c_uint64),
        ("q_dev", C.c_void_p),
        ("qn_dev", C.c_void_p),
        ("split_dev", C.c_void_p),
        ("survivors_dev", C.c_void_p),
        ("capacity", C.c_uint32),
        ("n_survivors_dev", C.c_void_p),
        ("overflow_dev", C.c_void_p),
        ("qn_stid_dev", C.c_void_p),
    ]


class RrsSide(C.Structure):
    _fields_ = [
        ("ref_mean", C.c_void_p),
        ("error", C.c_void_p),
        ("sample_weight", C.c_void_p),
        ("q_target", C.c_void_p),
        ("error_sum", C.c_float),
        ("error_sum_dev", C.c_void_p),
    ]


LOSS_EXTERNAL, LOSS_LL2, LOSS_RRS = 0, 1, 2


class FrameArgs(C.Structure):
    _fields_ = [
        ("n_primary", C.c_uint32),
        ("n_bounces", C.c_uint32),
        ("seed", C.c_uint64),
        ("seed_dev", C.c_void_p),
        ("advance_seed", C.c_int32),
        ("film_dev", C.c_void_p),
        ("queue_dev", C.c_void_p),
        ("capacity", C.c_uint32),
        ("sizes_dev", C.c_void_p),
        ("q_dev", C.c_void_p),
        ("trace_dev", C.c_void_p),
        ("status_dev", C.c_void_p),
    ]


class TileShape(C.Structure):
    _fields_ = [("waves", C.c_uint32), ("m", C.c_uint32)]


class TrainingArgs(C.Structure):
    _fields_ = [
        ("film_dev", C.c_void_p),
        ("error_dev", C.c_void_p),
        ("n_err", C.c_uint32),
        ("error_scale", C.c_uint32),
        ("frame_width", C.c_uint32),
        ("ignore_zero_L", C.c_int32),
        ("copy_rrs_node", C.c_int32),
        ("mark_duplicates", C.c_int32),
        ("capacity", C.c_uint32),
        ("out", VerticesOut),
        ("target_L", C.c_void_p),
        ("ref_mean", C.c_void_p),
        ("error", C.c_void_p),
        ("error_avg", C.c_void_p),
        ("q_target", C.c_void_p),
        ("sample_weight", C.c_void_p),
        ("pixel_samples", C.c_void_p),
        ("n_out_dev", C.c_void_p),
    ]


class ErrorArgs(C.Structure):
    _fields_ = [
        ("ref_dev", C.c_void_p),
        ("weighted_L_dev", C.c_void_p),
        ("spp_dev", C.c_void_p),
        ("clamp", C.c_int32),
        ("error_dev", C.c_void_p),
        ("error_sum_dev", C.c_void_p),
    ]

# name -> (restype, argtypes)
_P = C.c_void_p
_SIGS = {
    "nrrs_status_string": (C.c_char_p, [C.c_int]),
    "nrrs_struct_size": (C.c_size_t, [C.c_int32]),
    "nrrs_default_config": (C.c_int, [C.POINTER(Config)]),
    "nrrs_shipped_config": (C.c_int, [C.POINTER(Config)]),
    "nrrs_create": (C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    "nrrs_destroy": (C.c_int, [_P]),
    "nrrs_grid_layout": (C.c_int, [C.POINTER(GridConfig), _P, _P, _P, _P, _P, C.POINTER(C.c_uint64)]),
    "nrrs_param_count": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "nrrs_params_get": (C.c_int, [_P, C.c_int32, _P]),
    "nrrs_params_set": (C.c_int, [_P, C.c_int32, _P]),
    "nrrs_params_get_inference": (C.c_int, [_P, C.c_int32, _P]),
    "nrrs_encode": (C.c_int, [_P, C.c_int32, C.c_uint32, C.POINTER(Vertices), _P, _P, _P, _P]),
    "nrrs_grid_indices": (C.c_int, [_P, C.c_uint32, C.POINTER(Vertices), _P, _P]),
    "nrrs_statnet_infer": (C.c_int, [_P, C.c_uint32, C.POINTER(Vertices), _P, _P]),
    "nrrs_rrsnet_infer": (C.c_int, [_P, C.c_uint32, C.POINTER(Vertices), _P, _P, _P, _P]),
    "nrrs_rrs_infer_decide": (C.c_int, [_P, C.POINTER(DecideArgs), _P]),
    "nrrs_rrs_decide": (C.c_int, [_P, C.POINTER(DecideArgs), _P]),
    "nrrs_rrs_decide_static": (C.c_int, [C.POINTER(Config), C.POINTER(DecideArgs), _P]),
    "nrrs_synth_vertices": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(VerticesOut), _P]),
    "nrrs_train_step": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(Vertices), _P, _P, C.POINTER(RrsSide),
                                  _P, _P]),
    "nrrs_backward": (C.c_int, [_P, C.c_int32, C.c_uint32, C.POINTER(Vertices), _P, _P, C.c_int32, _P, _P,
                                C.POINTER(RrsSide), C.c_float, _P, _P]),
    "nrrs_grads_zero": (C.c_int, [_P, C.c_int32, _P]),
    "nrrs_grads_get": (C.c_int, [_P, C.c_int32, _P]),
    "nrrs_adam_step": (C.c_int, [_P, C.c_int32, _P]),
    "nrrs_train_state": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_float),
                                   C.POINTER(C.c_uint32)]),
    "nrrs_set_op_timing": (C.c_int, [_P, C.c_int32]),
    "nrrs_op_times": (C.c_int, [_P, C.POINTER(C.c_uint32), _P, _P]),
    "nrrs_frame_run": (C.c_int, [_P, C.POINTER(FrameArgs), _P]),
    "nrrs_frame_set_shape": (C.c_int, [_P, C.POINTER(TileShape)]),
    "nrrs_frame_autotune": (C.c_int, [_P, C.POINTER(FrameArgs), C.POINTER(TileShape), C.POINTER(C.c_float), _P]),
    "nrrs_frame_graph_capture": (C.c_int, [_P, C.POINTER(FrameArgs), _P]),
    "nrrs_frame_graph_launch": (C.c_int, [_P, _P]),
    "nrrs_frame_vertices": (C.c_int, [C.c_uint64, C.c_uint32, _P, C.c_uint32, C.POINTER(VerticesOut), _P]),
    "nrrs_error_blend": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_float, _P, _P]),
    "nrrs_error_compute": (C.c_int, [_P, C.c_uint32, C.POINTER(ErrorArgs), _P]),
    "nrrs_consume_survivors": (C.c_int, [_P, C.c_uint32, _P, _P, C.POINTER(Vertices), _P, _P,
                                         C.POINTER(VerticesOut), _P, _P, _P]),
    "nrrs_nodes_reset": (C.c_int, [_P, _P]),
    "nrrs_nodes_record": (C.c_int, [_P, C.c_uint32, _P, C.POINTER(Vertices), _P, _P, C.c_uint32, _P, _P]),
    "nrrs_nodes_add_radiance": (C.c_int, [_P, C.c_uint32, _P, _P, _P, C.c_uint32, _P]),
    "nrrs_nodes_size": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "nrrs_generate_training_data": (C.c_int, [_P, C.POINTER(TrainingArgs), _P]),
}

_lib = None


def exported_names():
    return list(_SIGS)


def load(path=LIB_PATH):
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libnrrs_amd.so not built ({path}); run nrrs-code_amd/build.py")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for k, cls in enumerate((Config, Vertices, DecideArgs, VerticesOut, RrsSide, FrameArgs, ErrorArgs, TrainingArgs)):
        if lib.nrrs_struct_size(k) != C.sizeof(cls):
            raise RuntimeError(f"ABI mismatch for {cls.__name__}: {lib.nrrs_struct_size(k)} != {C.sizeof(cls)}")
    _lib = lib
    return lib


class NrrsError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {load().nrrs_status_string(code).decode()} ({code})")


def check(code, what=""):
    if code != OK:
        raise NrrsError(code, what)
    return code


def default_config(**overrides):
    cfg = Config()
    check(load().nrrs_default_config(C.byref(cfg)), "default_config")
    return _override(cfg, overrides)


def shipped_config(**overrides):
    """The reference's shipped configuration (nrrs_shipped_config), with overrides."""
    cfg = Config()
    check(load().nrrs_shipped_config(C.byref(cfg)), "shipped_config")
    return _override(cfg, overrides)


def _override(cfg, overrides):
    for k, v in overrides.items():
        if k in ("gamma", "aabb_min", "aabb_max"):
            arr = getattr(cfg, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k in ("grid", "statnet", "rrsnet"):
            sub = getattr(cfg, k)
            for kk, vv in v.items():
                setattr(sub, kk, vv)
        else:
            setattr(cfg, k, v)
    return cfg


def grid_layout(grid_cfg):
    L = grid_cfg.n_levels
    res, off, size, dense = (np.zeros(L, np.uint32) for _ in range(4))
    scale = np.zeros(L, np.float32)
    n = C.c_uint64()
    check(load().nrrs_grid_layout(C.byref(grid_cfg), res.ctypes.data, off.ctypes.data, size.ctypes.data,
                                  dense.ctypes.data, scale.ctypes.data, C.byref(n)), "grid_layout")
    return dict(res=res, offset=off, size=size, dense=dense, scale=scale, n_params=int(n.value))


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class VertexBatch:
    """Device SoA of path vertices (component-major fields of length `stride`)."""

    FIELDS3 = ("pos", "dir", "normal", "albedo_d", "albedo_s", "thp")

    def __init__(self, n, device="cuda", stride=None):
        import torch

        self.n = n
        self.stride = stride or max(n, 1)
        f = dict(dtype=torch.float32, device=device)
        for k in self.FIELDS3:
            setattr(self, k, torch.zeros(3, self.stride, **f))
        self.rough = torch.zeros(self.stride, **f)
        self.pixel = torch.zeros(self.stride, dtype=torch.int32, device=device)
        self.stid = None

    @classmethod
    def from_numpy(cls, d, device="cuda"):
        import torch

        n = d["rough"].shape[0]
        vb = cls(n, device)
        for k in cls.FIELDS3:
            getattr(vb, k).copy_(torch.from_numpy(np.ascontiguousarray(d[k], np.float32)))
        vb.rough.copy_(torch.from_numpy(np.ascontiguousarray(d["rough"], np.float32)))
        vb.pixel.copy_(torch.from_numpy(np.ascontiguousarray(d["pixel"]).astype(np.int32)))
        if d.get("stid") is not None:
            vb.stid = torch.from_numpy(np.ascontiguousarray(d["stid"]).astype(np.int32)).to(device)
        return vb

    def to_numpy(self):
        out = {k: getattr(self, k)[:, : self.n].cpu().numpy() for k in self.FIELDS3}
        out["rough"] = self.rough[: self.n].cpu().numpy()
        out["pixel"] = self.pixel[: self.n].cpu().numpy().astype(np.uint32)
        out["stid"] = None if self.stid is None else self.stid[: self.n].cpu().numpy().astype(np.uint32)
        return out

    def c_struct(self):
        v = Vertices()
        for k in self.FIELDS3 + ("rough", "pixel"):
            setattr(v, k, getattr(self, k).data_ptr())
        v.stid = None if self.stid is None else self.stid.data_ptr()
        v.stride = self.stride
        return v

    def c_out(self):
        o = VerticesOut()
        for k in self.FIELDS3 + ("rough", "pixel"):
            setattr(o, k, getattr(self, k).data_ptr())
        o.stride = self.stride
        return o


def frame_vertices(seed, bounce, queue, n, device="cuda", stream=None):
    """Vertices a frame draws at `bounce` for the path-id queue `queue` (device int32)."""
    vb = VertexBatch(n, device)
    o = vb.c_out()
    check(load().nrrs_frame_vertices(seed, bounce, _ptr(queue), n, C.byref(o), _stream(stream)), "frame_vertices")
    return vb


def synth_vertices(n, seed, n_pixels, device="cuda", stream=None):
    vb = VertexBatch(n, device)
    o = vb.c_out()
    check(load().nrrs_synth_vertices(n, seed, n_pixels, C.byref(o), _stream(stream)), "synth_vertices")
    return vb


class Context:
    """One nrrs_ctx (the NRRSNetwork equivalent)."""

    def __init__(self, cfg=None, **overrides):
        self.lib = load()
        self.cfg = cfg if cfg is not None else default_config(**overrides)
        h = C.c_void_p()
        check(self.lib.nrrs_create(C.byref(self.cfg), C.byref(h)), "nrrs_create")
        self.h = h

    def close(self):
        if self.h:
            check(self.lib.nrrs_destroy(self.h), "nrrs_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def param_count(self, net):
        g, m = C.c_uint64(), C.c_uint64()
        check(self.lib.nrrs_param_count(self.h, net, C.byref(g), C.byref(m)), "param_count")
        return int(g.value), int(m.value)

    def params_get(self, net):
        g, m = self.param_count(net)
        out = np.zeros(g + m, np.float32)
        check(self.lib.nrrs_params_get(self.h, net, out.ctypes.data), "params_get")
        return out

    def params_get_inference(self, net):
        """The parameters inference reads (the Ema moving average when the context keeps one)."""
        g, m = self.param_count(net)
        out = np.zeros(g + m, np.float32)
        check(self.lib.nrrs_params_get_inference(self.h, net, out.ctypes.data), "params_get_inference")
        return out

    def params_set(self, net, arr):
        g, m = self.param_count(net)
        a = np.ascontiguousarray(arr, np.float32)
        assert a.shape == (g + m,)
        check(self.lib.nrrs_params_set(self.h, net, a.ctypes.data), "params_set")

    def encode(self, net, vb, film=None, stats=None, stream=None):
        import torch

        enc = torch.empty(vb.n, ENC_WIDTH, dtype=torch.float32, device=vb.pos.device)
        v = vb.c_struct()
        check(self.lib.nrrs_encode(self.h, net, vb.n, C.byref(v), _ptr(film), _ptr(stats), _ptr(enc),
                                   _stream(stream)), "encode")
        return enc

    def grid_indices(self, vb, stream=None):
        import torch

        idx = torch.empty(vb.n, self.cfg.grid.n_levels, 8, dtype=torch.int32, device=vb.pos.device)
        v = vb.c_struct()
        check(self.lib.nrrs_grid_indices(self.h, vb.n, C.byref(v), _ptr(idx), _stream(stream)), "grid_indices")
        return idx

    def statnet_infer(self, vb, stream=None):
        import torch

        out = torch.empty(6, vb.n, dtype=torch.float32, device=vb.pos.device)
        v = vb.c_struct()
        check(self.lib.nrrs_statnet_infer(self.h, vb.n, C.byref(v), _ptr(out), _stream(stream)), "statnet_infer")
        return out

    def rrsnet_infer(self, vb, film, stats=None, stream=None):
        import torch

        out = torch.empty(vb.n, dtype=torch.float32, device=vb.pos.device)
        v = vb.c_struct()
        check(self.lib.nrrs_rrsnet_infer(self.h, vb.n, C.byref(v), _ptr(film), _ptr(stats), _ptr(out),
                                         _stream(stream)), "rrsnet_infer")
        return out

    def decide_args(self, vb, film, q, survivors, n_surv, overflow=None, q_mode=Q_RRSNET, spec=None,
                    n_spec=None, seed=0, stream_id=0, qn=None, split=None, qn_stid=None):
        a = DecideArgs()
        a.n = vb.n
        a.v = vb.c_struct()
        a.film_dev = None if film is None else film.data_ptr()
        a.q_mode = q_mode
        a.spec_stid_dev = None if spec is None else spec.data_ptr()
        a.n_spec_dev = None if n_spec is None else n_spec.data_ptr()
        a.pcg_seed = seed
        a.pcg_stream = stream_id
        a.q_dev = q.data_ptr()
        a.qn_dev = None if qn is None else qn.data_ptr()
        a.split_dev = None if split is None else split.data_ptr()
        a.survivors_dev = survivors.data_ptr()
        a.capacity = survivors.numel()
        a.n_survivors_dev = n_surv.data_ptr()
        a.overflow_dev = None if overflow is None else overflow.data_ptr()
        a.qn_stid_dev = None if qn_stid is None else qn_stid.data_ptr()
        return a

    def rrs_infer_decide(self, args, stream=None):
        return self.lib.nrrs_rrs_infer_decide(self.h, C.byref(args), _stream(stream))

    def rrs_decide(self, args, stream=None):
        return self.lib.nrrs_rrs_decide(self.h, C.byref(args), _stream(stream))

    # ---- training
    @staticmethod
    def side(ref_mean=None, error=None, sample_weight=None, q_target=None, error_sum=0.0, error_sum_dev=None):
        """nrrs_rrs_side from device tensors (keeps no reference: the caller owns them);
        error_sum_dev, a device scalar, replaces error_sum when given."""
        sd = RrsSide()
        sd.ref_mean = None if ref_mean is None else ref_mean.data_ptr()
        sd.error = error.data_ptr()
        sd.sample_weight = sample_weight.data_ptr()
        sd.q_target = q_target.data_ptr()
        sd.error_sum = float(error_sum)
        sd.error_sum_dev = None if error_sum_dev is None else error_sum_dev.data_ptr()
        return sd

    def consume_survivors(self, survivors, n_survivors, scatter, rrs, capacity, node_in=None, stream=None):
        """Next-bounce rows from the survivor queue: scatter item survivors[k], thp / rrs[sid].
        Returns (VertexBatch, node_out int32, stid_out int32)."""
        import torch

        dev = survivors.device
        vb = VertexBatch(capacity, dev)
        node_out = torch.full((max(capacity, 1),), -2, dtype=torch.int32, device=dev)
        stid_out = torch.full((max(capacity, 1),), -1, dtype=torch.int32, device=dev)
        sv = scatter.c_struct()
        o = vb.c_out()
        check(self.lib.nrrs_consume_survivors(self.h, capacity, _ptr(n_survivors), _ptr(survivors), C.byref(sv),
                                              _ptr(rrs), _ptr(node_in), C.byref(o), _ptr(node_out), _ptr(stid_out),
                                              _stream(stream)), "consume_survivors")
        return vb, node_out, stid_out

    # ---- path nodes and training samples
    def nodes_reset(self, stream=None):
        check(self.lib.nrrs_nodes_reset(self.h, _stream(stream)), "nodes_reset")

    def nodes_record(self, vb, parent=None, q=None, depth=0, n_dev=None, stream=None):
        """Appends one node per vertex; returns the int32 node (or kept parent) per vertex."""
        import torch

        out = torch.empty(max(vb.n, 1), dtype=torch.int32, device=vb.pos.device)
        v = vb.c_struct()
        check(self.lib.nrrs_nodes_record(self.h, vb.n, _ptr(n_dev), C.byref(v), _ptr(parent), _ptr(q), depth,
                                         _ptr(out), _stream(stream)), "nodes_record")
        return out[: vb.n]

    def nodes_add_radiance(self, node, L, stream=None):
        """Adds L [3][n] to node[i] and its ancestors."""
        n = node.shape[0]
        check(self.lib.nrrs_nodes_add_radiance(self.h, n, None, _ptr(node), _ptr(L), L.shape[1], _stream(stream)),
              "nodes_add_radiance")

    def nodes_size(self):
        size, inval = C.c_uint32(), C.c_uint32()
        check(self.lib.nrrs_nodes_size(self.h, C.byref(size), C.byref(inval)), "nodes_size")
        return size.value, inval.value

    def generate_training_data(self, film, capacity, error=None, error_scale=1, frame_width=None,
                               ignore_zero_L=False, copy_rrs_node=False, mark_duplicates=True, pixel_samples=None,
                               stream=None):
        """Training samples from the recorded nodes; returns (VertexBatch, dict of device arrays)."""
        import torch

        dev = film.device
        vb = VertexBatch(capacity, dev)
        f = dict(dtype=torch.float32, device=dev)
        P = film.shape[1]
        o = dict(target_L=torch.zeros(3, capacity, **f), ref_mean=torch.zeros(capacity, **f),
                 error=torch.zeros(capacity, **f), error_avg=torch.zeros(capacity, **f),
                 q_target=torch.zeros(capacity, **f), sample_weight=torch.zeros(capacity, **f),
                 pixel_samples=pixel_samples if pixel_samples is not None else
                 torch.zeros(P, dtype=torch.int32, device=dev),
                 n_out=torch.zeros(1, dtype=torch.int32, device=dev))
        a = TrainingArgs()
        a.film_dev = film.data_ptr()
        a.error_dev = None if error is None else error.data_ptr()
        a.n_err = 0 if error is None else error.shape[1]
        a.error_scale = error_scale
        a.frame_width = frame_width or 1
        a.ignore_zero_L, a.copy_rrs_node, a.mark_duplicates = int(ignore_zero_L), int(copy_rrs_node), int(mark_duplicates)
        a.capacity = capacity
        a.out = vb.c_out()
        for k in ("target_L", "ref_mean", "error", "error_avg", "q_target", "sample_weight", "pixel_samples"):
            setattr(a, k, o[k].data_ptr())
        a.n_out_dev = o["n_out"].data_ptr()
        check(self.lib.nrrs_generate_training_data(self.h, C.byref(a), _stream(stream)), "generate_training_data")
        return vb, o

    def error_blend(self, L_frame, spp, blend, weighted_L, stream=None):
        """weighted_L = weighted_L (1 - blend) + L_frame / spp * blend ([3][P] planes, in place)."""
        check(self.lib.nrrs_error_blend(self.h, weighted_L.shape[1], _ptr(L_frame), _ptr(spp), blend,
                                        _ptr(weighted_L), _stream(stream)), "error_blend")

    def error_compute(self, ref, weighted_L, spp=None, clamp=True, stream=None):
        """Per-pixel error [2][P] (row 1 scaled by sqrt(spp)) and its sum (device scalar)."""
        import torch

        n = ref.shape[1]
        err = torch.empty(2, n, dtype=torch.float32, device=ref.device)
        total = torch.empty(1, dtype=torch.float32, device=ref.device)
        a = ErrorArgs(ref.data_ptr(), weighted_L.data_ptr(), None if spp is None else spp.data_ptr(), int(clamp),
                      err.data_ptr(), total.data_ptr())
        check(self.lib.nrrs_error_compute(self.h, n, C.byref(a), _stream(stream)), "error_compute")
        return err, total

    def backward(self, net, vb, kind, film=None, stats=None, dy=None, target_L=None, side=None, loss_scale=1.0,
                 loss=None, stream=None):
        v = vb.c_struct()
        return self.lib.nrrs_backward(self.h, net, vb.n, C.byref(v), _ptr(film), _ptr(stats), kind, _ptr(dy),
                                      _ptr(target_L), None if side is None else C.byref(side), loss_scale,
                                      _ptr(loss), _stream(stream))

    def grads_zero(self, net, stream=None):
        check(self.lib.nrrs_grads_zero(self.h, net, _stream(stream)), "grads_zero")

    def grads_get(self, net):
        g, m = self.param_count(net)
        out = np.zeros(g + m, np.float32)
        check(self.lib.nrrs_grads_get(self.h, net, out.ctypes.data), "grads_get")
        return out

    def adam_step(self, net, stream=None):
        check(self.lib.nrrs_adam_step(self.h, net, _stream(stream)), "adam_step")

    def train_state(self, net):
        step, sc, sk = C.c_uint32(), C.c_float(), C.c_uint32()
        check(self.lib.nrrs_train_state(self.h, net, C.byref(step), C.byref(sc), C.byref(sk)), "train_state")
        return dict(step=step.value, loss_scale=sc.value, skipped=sk.value)

    # ---- operation timing
    MAX_OP_TIMES = 256

    def set_op_timing(self, on):
        check(self.lib.nrrs_set_op_timing(self.h, int(on)), "set_op_timing")

    def op_times(self):
        """[(name, ms)] of the operations recorded since the previous call (waits for the device)."""
        n = C.c_uint32()
        names = (C.c_char_p * self.MAX_OP_TIMES)()
        ms = (C.c_float * self.MAX_OP_TIMES)()
        check(self.lib.nrrs_op_times(self.h, C.byref(n), names, ms), "op_times")
        return [(names[k].decode(), ms[k]) for k in range(n.value)]

    # ---- batched uneven wavefront (one persistent launch per frame)
    def frame_buffers(self, n_primary, n_bounces, capacity=None, trace=False, record_q=False, device="cuda"):
        """Device buffers of one frame; returns (FrameArgs, dict of tensors)."""
        import torch

        cap = capacity or n_primary
        b = dict(queue=torch.zeros(2 * cap, dtype=torch.int32, device=device),
                 sizes=torch.zeros(n_bounces + 1, dtype=torch.int32, device=device),
                 status=torch.zeros(1, dtype=torch.int32, device=device))
        if trace:
            b["trace"] = torch.zeros((n_bounces + 1, cap), dtype=torch.int32, device=device)
        if record_q:
            b["q"] = torch.zeros((n_bounces, cap), dtype=torch.float32, device=device)
        a = FrameArgs()
        a.n_primary, a.n_bounces, a.capacity = n_primary, n_bounces, cap
        a.queue_dev = b["queue"].data_ptr()
        a.sizes_dev = b["sizes"].data_ptr()
        a.status_dev = b["status"].data_ptr()
        a.trace_dev = b["trace"].data_ptr() if trace else None
        a.q_dev = b["q"].data_ptr() if record_q else None
        return a, b

    def frame_run(self, args, stream=None):
        return self.lib.nrrs_frame_run(self.h, C.byref(args), _stream(stream))

    def frame_set_shape(self, waves, m):
        check(self.lib.nrrs_frame_set_shape(self.h, C.byref(TileShape(waves, m))), "frame_set_shape")

    def frame_autotune(self, args, stream=None):
        best, ms = TileShape(), C.c_float()
        check(self.lib.nrrs_frame_autotune(self.h, C.byref(args), C.byref(best), C.byref(ms), _stream(stream)),
              "frame_autotune")
        return (best.waves, best.m), ms.value

    def frame_graph_capture(self, args, stream):
        check(self.lib.nrrs_frame_graph_capture(self.h, C.byref(args), _stream(stream)), "frame_graph_capture")

    def frame_graph_launch(self, stream=None):
        check(self.lib.nrrs_frame_graph_launch(self.h, _stream(stream)), "frame_graph_launch")

    def train_step(self, vb, net_mask, film=None, target_L=None, side=None, loss=None, stream=None):
        v = vb.c_struct()
        return self.lib.nrrs_train_step(self.h, net_mask, vb.n, C.byref(v), _ptr(film), _ptr(target_L),
                                        None if side is None else C.byref(side), _ptr(loss), _stream(stream))



STATIC_DECIDE_MAX = 1 << 22


def rrs_decide_static(cfg, args, stream=None):
    """nrrs_rrs_decide_static: the decision with the library's fixed device scratch (no context,
    no allocation)."""
    return load().nrrs_rrs_decide_static(C.byref(cfg), C.byref(args), _stream(stream))
