// C ABI of the neural RRS stage: context lifetime, parameters, inference and the
// per-bounce RRS decision. Mirrors the integrator-facing surface of NRRSNetwork
// (src/render/nrrs/net/network.h:22-126) with explicit device pointers instead of
// json-carried raw addresses, and status codes instead of Log(Fatal).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "host_common.hpp"
#include "train_common.hpp"
#include "frame_common.hpp"

using namespace nrrs;

namespace {

nrrs_status hip_status(hipError_t e) {
    if (e == hipSuccess) return NRRS_OK;
    if (e == hipErrorOutOfMemory) return NRRS_ERR_OOM;
    return NRRS_ERR_HIP;
}

#define NRRS_TRY(expr)                         \
    do {                                       \
        nrrs_status st_ = (expr);              \
        if (st_ != NRRS_OK) return st_;        \
    } while (0)
#define NRRS_HIP(expr) NRRS_TRY(hip_status(expr))

// Event pair around one operation when the context's operation timing is on (nrrs_set_op_timing).
struct OpScope {
    nrrs_ctx* c;
    hipStream_t s;
    int k = -1;
    OpScope(nrrs_ctx* ctx, const char* name, hipStream_t stream) : c(ctx), s(stream) {
        if (c->op_timing && c->op_count < NRRS_MAX_OP_TIMES) {
            k = (int)c->op_count++;
            c->op_name[k] = name;
            (void)hipEventRecord(c->op_ev[2 * k], s);
        }
    }
    ~OpScope() {
        if (k >= 0) (void)hipEventRecord(c->op_ev[2 * k + 1], s);
    }
};

struct Pcg32Host {
    uint64_t state, inc;
    Pcg32Host(uint64_t seed, uint64_t stream) {
        inc = (stream << 1u) | 1u;
        state = 0;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * kPcgMult + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    double uniform() { return (double)(next() >> 8) * 0x1p-24; }
};

// Level table of the hash grid. scale_l = base * per_level_scale^l (rounded to
// FP32 once, here); res_l = ceil(scale_l) + 2 so that every corner of a cell of
// a position in [0,1] is inside the level; a level is dense when res^3 <= T.
// Level starts are padded to multiples of 8 entries.
bool make_grid(const nrrs_grid_config& g, GridParams* gp, uint32_t* res, uint32_t* offset,
               uint32_t* size, uint32_t* dense, float* scale, uint64_t* n_params) {
    if (g.n_levels < 1 || g.n_levels > (uint32_t)kMaxLevels || g.n_features != 2 ||
        g.log2_table_size < 4 || g.log2_table_size > 24 || g.base_resolution < 1 ||
        !(g.per_level_scale >= 1.0f))
        return false;
    const uint64_t T = 1ull << g.log2_table_size;
    uint64_t off = 0;
    GridParams p{};
    for (uint32_t l = 0; l < g.n_levels; ++l) {
        double s = (double)g.base_resolution * std::pow((double)g.per_level_scale, (double)l);
        float sf = (float)s;
        if (!(sf < 1.0e6f)) return false;
        uint32_t r = (uint32_t)std::ceil((double)sf) + 2u;
        uint64_t r3 = (uint64_t)r * r * r;
        uint32_t d = r3 <= T ? 1u : 0u;
        uint32_t sz = (uint32_t)(d ? r3 : T);
        p.scale[l] = sf;
        p.res[l] = r;
        p.offset[l] = (uint32_t)off;
        if (d) p.dense_mask |= 1u << l;
        if (res) res[l] = r;
        if (offset) offset[l] = (uint32_t)off;
        if (size) size[l] = sz;
        if (dense) dense[l] = d;
        if (scale) scale[l] = sf;
        off += (sz + 7ull) / 8ull * 8ull;
    }
    if (off > (1ull << 31)) return false;
    p.n_levels = g.n_levels;
    p.hash_mask = (uint32_t)(T - 1);
    p.n_entries = (uint32_t)off;
    if (gp) *gp = p;
    if (n_params) *n_params = off * 2;
    return true;
}

// normalizeSpatialCoord frame (src/render/nrc/nrctrain.h:111-115), computed once
// in FP32 with no contraction (the library is built with -ffp-contract=off on the host).
Frame make_frame(const float* mn, const float* mx) {
    Frame f;
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    float len = std::sqrt(dx * dx + dy * dy + dz * dz);
    float infl = len * 0.005f;
    for (int k = 0; k < 3; ++k) {
        float lo = mn[k] - infl, hi = mx[k] + infl;
        f.c[k] = (lo + hi) * 0.5f;
        f.d[k] = hi - lo;
    }
    return f;
}

// The RRS-decision knobs of a configuration (integrator.h:265, render/test.json:63-67).
DecideParams decide_params(const nrrs_config& cfg) {
    DecideParams d;
    d.clamp_min = cfg.rrs_clamp_min;
    d.clamp_max = cfg.rrs_clamp_max;
    d.rate = cfg.rrs_normalize_rate;
    d.window_size = cfg.weight_window_size;
    d.use_ceil = cfg.rrs_use_ceil;
    d.no_upscale = cfg.rrs_no_upscale;
    d.weight_window = cfg.rrs_weight_window;
    d.split_cap = cfg.split_cap;
    d.n_pixels = cfg.n_pixels;
    return d;
}

uint64_t mlp_count(const nrrs_mlp_config& m) {
    return (uint64_t)kHidden * kEnc + (uint64_t)(m.n_hidden_layers - 1) * kHidden * kHidden +
           (uint64_t)m.n_outputs * kHidden;
}

bool valid_shapes(const nrrs_config& c) {
    // Shapes the fused kernels are compiled for: StatNet 4 hidden -> 6 (one head), 3 hidden
    // (PartShare) or 2 hidden (JustEncoding, and each net of AllSelf) -> 6; RRSNet 3 hidden -> 1.
    // The shipped nets: one-headed StatNet 3 hidden -> 6, RRSNet 3 hidden -> 1.
    static const uint32_t stat_hidden[4] = {4, 3, 2, 2};
    if (c.ll2_two_head < 0 || c.ll2_two_head > 3) return false;
    if (c.net_set == NRRS_NETS_SHIPPED)
        return c.ll2_two_head == 0 && c.statnet.n_hidden_layers == 3 && c.statnet.n_outputs == 6 &&
               c.rrsnet.n_hidden_layers == 3 && c.rrsnet.n_outputs == 1;
    return c.statnet.n_hidden_layers == stat_hidden[c.ll2_two_head] && c.statnet.n_outputs == 6 &&
           c.rrsnet.n_hidden_layers == 3 && c.rrsnet.n_outputs == 1;
}

// Whether a net has a hash grid: every net but the shipped RRSNet without the AID input, which
// reads NRRSNetworkInput1 alone (network.cpp:503-507 builds it without an encoding).
bool has_grid(const nrrs_config& c, int net) {
    return !(net == NRRS_RRSNET && c.net_set == NRRS_NETS_SHIPPED && !c.rrs_aid);
}

// Shipped nets: the live part of the 64-wide, 80-channel storage. Hidden layers are 32 wide;
// layer 0 reads the net's input channels (StatNet: grid and OneBlob channels; RRSNet: the 16
// NRRSNetworkInput1 channels [64,80); AID RRSNet: grid, OneBlob, BC(thp) and BC(ref mean)).
bool shipped_live(const nrrs_config& c, int net, int layer, int n_layers, int row, int col) {
    if (layer < n_layers - 1 && row >= 32) return false;
    if (layer > 0) return col < 32;
    const int grid_ch = has_grid(c, net) ? 2 * (int)c.grid.n_levels : 0;
    if (col < grid_ch) return true;
    if (net != NRRS_RRSNET) return col >= 32 && col < 48;
    if (!c.rrs_aid) return col >= 64;
    return (col >= 32 && col < 44) || (col >= 65 && col < 69);
}

// Layer-0 input chunks (16 channels each, bit c for channels 16c..16c+15) that hold an input of
// the net: all five for the benchmark nets, those with a live column for the shipped ones.
uint32_t l0_chunks(const nrrs_config& c, int net) {
    if (c.net_set != NRRS_NETS_SHIPPED) return 0x1fu;
    const int nl = (int)(net == NRRS_RRSNET ? c.rrsnet : c.statnet).n_hidden_layers + 1;
    uint32_t m = 0;
    for (int col = 0; col < kEnc; ++col)
        if (shipped_live(c, net, 0, nl, 0, col)) m |= 1u << (col / 16);
    return m;
}

// AllSelf (ll2_two_head = 3): the StatNet of L (NRRS_STATNET) and the StatNet of L^2
// (NRRS_STATNET2), each with its own table, optimizer and loss (twn_AllSelf.cpp).
bool all_self(const nrrs_config& c) { return c.ll2_two_head == 3; }
int n_nets(const nrrs_config& c) { return all_self(c) ? 3 : 2; }
bool net_ok(const nrrs_ctx* c, int32_t net) {
    return net == NRRS_STATNET || net == NRRS_RRSNET || (net == NRRS_STATNET2 && all_self(c->cfg));
}
// What a net's kernels compute: NRRS_RRSNET, or NRRS_STATNET for either StatNet.
int kind_of(int net) { return net == NRRS_RRSNET ? NRRS_RRSNET : NRRS_STATNET; }
const nrrs_mlp_config& mlp_of(const nrrs_config& c, int net) { return net == NRRS_RRSNET ? c.rrsnet : c.statnet; }
// AllSelf: the outputs a StatNet predicts and trains (3 head .. 3 head + 2); -1 = all.
int head_of(const nrrs_config& c, int net) {
    return all_self(c) && net != NRRS_RRSNET ? (net == NRRS_STATNET ? 0 : 1) : -1;
}
const char* op_label(int net, const char* statnet, const char* rrsnet, const char* statnet2) {
    return net == NRRS_STATNET ? statnet : (net == NRRS_RRSNET ? rrsnet : statnet2);
}

// Zero the weights between the two heads of a two-headed StatNet (mlp: that net's MLP part); an
// AllSelf StatNet's output rows of the other head.
void zero_between_heads(const nrrs_config& c, int net, float* mlp) {
    if (c.net_set == NRRS_NETS_SHIPPED) {  // the weights outside the 32-wide net
        const nrrs_mlp_config& m = net == NRRS_RRSNET ? c.rrsnet : c.statnet;
        const int nl = (int)m.n_hidden_layers + 1;
        size_t pos = 0;
        int in = kEnc;
        for (int layer = 0; layer < nl; ++layer) {
            const int o = layer == nl - 1 ? (int)m.n_outputs : kHidden;
            for (int r = 0; r < o; ++r)
                for (int k = 0; k < in; ++k)
                    if (!shipped_live(c, net, layer, nl, r, k)) mlp[pos + (size_t)r * in + k] = 0.0f;
            pos += (size_t)o * in;
            in = o;
        }
        return;
    }
    if (net == NRRS_RRSNET || !c.ll2_two_head) return;
    const int nl = (int)c.statnet.n_hidden_layers + 1, first = head_first_layer(c.ll2_two_head, nl);
    if (all_self(c)) {
        const size_t out0 = (size_t)kHidden * kEnc + (size_t)(nl - 2) * kHidden * kHidden;
        for (int r = 0; r < (int)c.statnet.n_outputs; ++r)
            if (r / 3 != head_of(c, net))
                for (int k = 0; k < kHidden; ++k) mlp[out0 + (size_t)r * kHidden + k] = 0.0f;
        return;
    }
    size_t pos = 0;
    int in = kEnc;
    for (int layer = 0; layer < nl; ++layer) {
        const int o = layer == nl - 1 ? (int)c.statnet.n_outputs : kHidden;
        for (int r = 0; r < o; ++r)
            for (int k = 0; k < in; ++k)
                if (head_masked(first, layer, nl, r, k)) mlp[pos + (size_t)r * in + k] = 0.0f;
        pos += (size_t)o * in;
        in = o;
    }
}

template <typename T>
nrrs_status dmalloc(T** p, size_t count) {
    *p = nullptr;
    if (!count) return NRRS_OK;
    return hip_status(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
}

// Parameters of one net: grid U[-1e-4, 1e-4] (the tcnn grid initialisation), then
// every MLP layer Xavier-uniform, from PCG32(seed, 100 + net). A two-headed StatNet takes the
// fan-in and fan-out of each head's own layer (32 wide, 3 outputs) from the split on, and
// zeros between the heads; an AllSelf StatNet has 3 outputs.
void init_params(const nrrs_config& c, int net, uint64_t n_grid, std::vector<float>& out) {
    const nrrs_mlp_config& m = mlp_of(c, net);
    out.resize(n_grid + mlp_count(m));
    Pcg32Host r(c.seed, 100 + (uint64_t)net);
    for (uint64_t k = 0; k < n_grid; ++k) out[k] = (float)(1e-4 * (2.0 * r.uniform() - 1.0));
    size_t pos = n_grid;
    int in = kEnc;
    const int nl = (int)m.n_hidden_layers + 1;
    const int first = net == NRRS_STATNET ? head_first_layer(c.ll2_two_head, nl) : -1;
    for (int layer = 0; layer < nl; ++layer) {
        int o = layer == nl - 1 ? (int)m.n_outputs : kHidden;
        const bool split = first >= 0 && layer >= first;
        int fan_out = split || (head_of(c, net) >= 0 && layer == nl - 1) ? o / 2 : o;
        int fan_in = split && layer > first ? in / 2 : in;
        if (c.net_set == NRRS_NETS_SHIPPED) {  // 32-wide layers; the inputs of the net's own encoding
            fan_out = layer == nl - 1 ? o : 32;
            fan_in = layer > 0 ? 32 : (net == NRRS_RRSNET && !c.rrs_aid ? 16 : 32);
        }
        double lim = std::sqrt(6.0 / (double)(fan_in + fan_out));
        for (int k = 0; k < o * in; ++k) out[pos++] = (float)(lim * (2.0 * r.uniform() - 1.0));
        in = o;
    }
    zero_between_heads(c, net, out.data() + n_grid);
}

EncodeArgs encode_args(nrrs_ctx* ctx, const nrrs_vertices& v, const float* film, const float* stats,
                       uint32_t stats_stride) {
    EncodeArgs a;
    a.v = v;
    a.film = film;
    a.n_pixels = ctx->cfg.n_pixels;
    a.stats = stats;
    a.stats_stride = stats_stride;
    a.grid = ctx->grid;
    a.frame = ctx->frame;
    a.aid = ctx->cfg.rrs_aid;
    a.net_set = ctx->cfg.net_set;
    return a;
}

bool vertices_ok(const nrrs_vertices* v, uint32_t n) {
    return v && v->pos && v->dir && v->normal && v->rough && v->albedo_d && v->albedo_s && v->thp &&
           v->pixel && v->stride >= n;
}

// Paired table from both nets' FP16/BF16 grids; after every change of either.
nrrs_status refresh_pair(nrrs_ctx* c, hipStream_t s) {
    if (!c->grid_pair || !c->net[NRRS_STATNET].grid_lp || !c->net[NRRS_RRSNET].grid_lp) return NRRS_OK;
    OpScope t(c, "pack_pair", s);
    NRRS_HIP(launch_pack_pair(c->net[NRRS_STATNET].grid_lp, c->net[NRRS_RRSNET].grid_lp, c->grid_pair,
                              c->net[NRRS_STATNET].n_grid / 2, s));
    if (!c->grid_pair_ema) return NRRS_OK;
    return hip_status(launch_pack_pair(c->net[NRRS_STATNET].grid_lp_ema, c->net[NRRS_RRSNET].grid_lp_ema,
                                       c->grid_pair_ema, c->net[NRRS_STATNET].n_grid / 2, s));
}

// Low-precision shadows (grid table and packed MFMA tiles) from the FP32 masters, and from their
// moving average when the context keeps one.
nrrs_status refresh_lp(nrrs_ctx* c, int net, hipStream_t s) {
    if (c->cfg.precision == NRRS_FP32) return NRRS_OK;
    NetState& ns = c->net[net];
    NRRS_HIP(launch_pack_lp(c->cfg.precision, ns.grid_f32, ns.n_grid, ns.mlp_f32, ns.n_layers, ns.n_out,
                            ns.grid_lp, ns.mlp_lp, s));
    if (ns.ema)
        NRRS_HIP(launch_pack_lp(c->cfg.precision, ns.ema, ns.n_grid, ns.ema + ns.n_grid, ns.n_layers, ns.n_out,
                                ns.grid_lp_ema, ns.mlp_lp_ema, s));
    return refresh_pair(c, s);
}

// The parameters one net reads: for inference, the moving average of the Ema optimizer when the
// context keeps one (tiny-cuda-nn's Trainer infers with the optimizer's averaged weights,
// 2net-simple.json:16-18), else the masters; for training (`train_side`), the masters.
struct NetView {
    const float* grid_f32;
    const float* mlp_f32;
    const uint16_t* grid_lp;
    const uint16_t* mlp_lp;
    const uint32_t* pair;  // the paired table of the same parameters, or null
};
NetView view_of(const nrrs_ctx* c, int net, bool train_side = false) {
    const NetState& ns = c->net[net];
    if (ns.ema && !train_side)
        return NetView{ns.ema, ns.ema + ns.n_grid, ns.grid_lp_ema, ns.mlp_lp_ema, c->grid_pair_ema};
    return NetView{ns.grid_f32, ns.mlp_f32, ns.grid_lp, ns.mlp_lp, c->grid_pair};
}

// One fused encode + MLP launch of `net` in the context's precision.
hipError_t run_net(nrrs_ctx* c, int net, const EncodeArgs& a_in, int epi, uint32_t n, const EpiArgs& ep,
                   hipStream_t s, bool train_side = false) {
    const NetState& ns = c->net[net];
    const NetView w = view_of(c, net, train_side);
    EncodeArgs a = a_in;
    a.grid = ns.grid;  // the net's own grid (none for the shipped RRSNet)
    a.l0_chunks = l0_chunks(c->cfg, net);
    OpScope t(c, op_label(net, "infer_statnet", "infer_rrsnet", "infer_statnet2"), s);
    const int k = kind_of(net);
    if (c->cfg.precision == NRRS_FP32) {
        if (!c->grid_enc_f32 || n > c->grid_enc_cap)
            return launch_net_fp32(a, k, ns.n_layers, ns.n_out, epi, n, w.grid_f32, w.mlp_f32, ep, s);
        // level-major grid gathers first, then the network kernel reads the encoding
        EncodeArgs ae = a;
        ae.grid_enc_f32 = c->grid_enc_f32;
        ae.grid_enc_stride = c->grid_enc_cap;
        hipError_t e = launch_encode_levels_f32(a, w.grid_f32, n, c->grid_enc_f32, c->grid_enc_cap, s);
        if (e != hipSuccess) return e;
        return launch_net_fp32(ae, k, ns.n_layers, ns.n_out, epi, n, w.grid_f32, w.mlp_f32, ep, s);
    }
    if (!c->grid_enc || n > c->grid_enc_cap)
        return launch_net_mfma(c->cfg.precision, a, k, ns.n_layers, ns.n_out, epi, n, w.grid_lp, w.mlp_lp, ep, s);
    // level-major grid gathers first, then the MLP kernel reads the encoding
    EncodeArgs ae = a;
    ae.grid_enc = c->grid_enc;
    ae.grid_enc_stride = c->grid_enc_cap;
    hipError_t e = launch_encode_levels(c->cfg.precision, ae, w.grid_lp, n, c->grid_enc, c->grid_enc_cap, s);
    if (e != hipSuccess) return e;
    return launch_net_mfma(c->cfg.precision, ae, k, ns.n_layers, ns.n_out, epi, n, w.grid_lp, w.mlp_lp, ep, s);
}

// StatNet statistics [6][stride] of a batch; AllSelf runs both its nets, each writing its head's
// outputs.
hipError_t run_statnets(nrrs_ctx* c, const EncodeArgs& a, float* out, uint32_t stride, uint32_t n, hipStream_t s,
                        bool train_side = false) {
    EpiArgs ep{};
    ep.out = out;
    ep.out_stride = stride;
    if (!all_self(c->cfg)) return run_net(c, NRRS_STATNET, a, EPI_RAW, n, ep, s, train_side);
    ep.head = 0;
    const hipError_t e = run_net(c, NRRS_STATNET, a, EPI_RAW, n, ep, s, train_side);
    if (e != hipSuccess) return e;
    ep.head = 1;
    return run_net(c, NRRS_STATNET2, a, EPI_RAW, n, ep, s, train_side);
}

// Whether a batch of n vertices takes the paired encode pass (both nets in one gather).
bool paired_pass(const nrrs_ctx* c, uint32_t n) {
    return c->grid_pair && c->grid_enc2 && n && n <= c->grid_enc_cap;
}

// Packed training tiles from the FP32 masters (after every parameter change).
nrrs_status refresh_train(nrrs_ctx* c, int net, hipStream_t s) {
    NetState& ns = c->net[net];
    if (!ns.train_ready) return NRRS_OK;
    return hip_status(launch_pack_train(c->cfg.precision, ns.mlp_f32, ns.n_layers, ns.n_out, ns.w_fwd, ns.w_bwd, s));
}

// Training buffers of one net (nrrs_create, when the context trains).
nrrs_status ensure_train(nrrs_ctx* c, int net) {
    NetState& ns = c->net[net];
    if (ns.train_ready) return NRRS_OK;
    const uint64_t np = ns.n_grid + ns.n_mlp;
    const size_t ob = train_op_bytes(c->cfg.precision);
    NRRS_TRY(dmalloc(&ns.grad, np));
    if (c->cfg.precision != NRRS_FP32) {
        NRRS_TRY(dmalloc(&ns.grad_lp, ns.n_grid));
        NRRS_HIP(hipMemset(ns.grad_lp, 0, ns.n_grid * 2));
    }
    NRRS_TRY(dmalloc(&ns.adam_m, np));
    NRRS_TRY(dmalloc(&ns.adam_v, np));
    NRRS_TRY(dmalloc(reinterpret_cast<uint8_t**>(&ns.w_fwd), (size_t)train_tiles_fwd(ns.n_layers) * 64 * ob));
    NRRS_TRY(dmalloc(reinterpret_cast<uint8_t**>(&ns.w_bwd), (size_t)train_tiles_bwd(ns.n_layers) * 64 * ob));
    NRRS_TRY(dmalloc(&ns.dw_partial, (size_t)c->max_train_blocks * ns.n_mlp));
    // the narrow kernels write only the tiles of their 32-wide nets
    NRRS_HIP(hipMemset(ns.dw_partial, 0, (size_t)c->max_train_blocks * ns.n_mlp * sizeof(float)));
    {
        // dense levels come first: their entries end where the first hashed level starts
        const GridParams& gp = ns.grid;
        uint32_t nd = 0;
        for (uint32_t l = 0; l < gp.n_levels; ++l)
            if ((gp.dense_mask >> l) & 1u) nd = (l + 1 < gp.n_levels) ? gp.offset[l + 1] : gp.n_entries;
        ns.n_dense = nd;
        const size_t bytes = (size_t)kDenseGradReplicas * nd * (c->cfg.precision == NRRS_FP32 ? 8 : 4);
        if (bytes) {
            NRRS_TRY(dmalloc(reinterpret_cast<uint8_t**>(&ns.dense_rep), bytes));
            NRRS_HIP(hipMemset(ns.dense_rep, 0, bytes));
        }
    }
    if (c->cfg.precision != NRRS_FP32 && ns.grid.n_levels) {
        // Binned scatter: every level cut into segments of kSegEntries entries, each with room
        // for 1.25x (hashed levels, spread by the hash) or 3x (dense levels, whose segments see
        // the geometry) its share of a full batch's records (4.5 per vertex on average: four
        // x-neighbour pairs, one in eight split in two), at most 8 per vertex; what does not
        // fit is added atomically.
        const uint64_t T = (uint64_t)c->grid.hash_mask + 1u, B8 = 8ull * c->cfg.max_train;
        const uint64_t B45 = (9ull * c->cfg.max_train) / 2;
        uint32_t chunk = kApplyChunkMin;
        while (chunk < kApplyChunkMax && 2ull * chunk <= c->cfg.max_train / 8) chunk *= 2;
        ns.apply_chunk = chunk;
        std::vector<SegInfo> segs;
        std::vector<uint32_t> direct, chunked;
        std::vector<uint2> chunks;
        bool ok = true;
        uint64_t rec_total = 0;
        for (uint32_t l = 0; l < c->grid.n_levels && ok; ++l) {
            const bool dense = (c->grid.dense_mask >> l) & 1u;
            const uint64_t r = c->grid.res[l], size = dense ? r * r * r : T;
            const uint32_t nb = (uint32_t)((size + kSegEntries - 1) / kSegEntries);
            if (nb > kMaxSegBins) {
                ok = false;
                break;
            }
            ns.seg_base[l] = (uint32_t)segs.size();
            ns.seg_nbins[l] = nb;
            for (uint32_t b = 0; b < nb; ++b) {
                const uint64_t ne = std::min<uint64_t>(kSegEntries, size - (uint64_t)b * kSegEntries);
                const uint64_t expect = B45 * ne / size;
                uint64_t cap = dense ? std::min<uint64_t>(B8, 3 * expect + 4096) : expect + expect / 4 + 4096;
                cap = (cap + 63) / 64 * 64;
                const uint32_t idx = (uint32_t)segs.size();
                segs.push_back(SegInfo{(uint32_t)rec_total, (uint32_t)cap, c->grid.offset[l] + b * kSegEntries,
                                       (uint32_t)ne});
                if (dense) {
                    chunked.push_back(idx);
                    for (uint64_t r0 = 0; r0 < cap; r0 += chunk) chunks.push_back(make_uint2(idx, (uint32_t)r0));
                } else {
                    direct.push_back(idx);
                }
                rec_total += cap;
            }
        }
        // slots and slot + list index stay below 2^32
        if (ok && rec_total < 0xFFFF0000ull) {
            NRRS_TRY(dmalloc(&ns.rec_key, rec_total));
            NRRS_TRY(dmalloc(&ns.rec_val, rec_total));
            NRRS_TRY(dmalloc(&ns.gdl, (size_t)c->grid.n_levels * c->cfg.max_train));
            const uint32_t max_runs = (c->cfg.max_train + kRunVert - 1) / kRunVert;
            NRRS_TRY(dmalloc(&ns.bin_runs, (size_t)c->grid.n_levels * kMaxSegBins * max_runs));
            ns.bin_max_tiles = (max_runs + kBinScanRows - 1) / kBinScanRows;
            NRRS_TRY(dmalloc(&ns.bin_tiles, (size_t)c->grid.n_levels * kMaxSegBins * ns.bin_max_tiles));
            NRRS_TRY(dmalloc(&ns.seg_count, segs.size()));
            NRRS_HIP(hipMemset(ns.seg_count, 0, segs.size() * 4));
            ns.n_segs = (uint32_t)segs.size();
            {
                std::vector<uint32_t> end((size_t)c->grid.n_levels * kMaxSegBins, 0u);
                for (uint32_t l = 0; l < c->grid.n_levels; ++l)
                    for (uint32_t b = 0; b < ns.seg_nbins[l]; ++b) {
                        const SegInfo& sg = segs[ns.seg_base[l] + b];
                        end[(size_t)l * kMaxSegBins + b] = sg.rec_off + sg.cap;
                    }
                NRRS_TRY(dmalloc(&ns.seg_end, end.size()));
                NRRS_HIP(hipMemcpy(ns.seg_end, end.data(), end.size() * 4, hipMemcpyHostToDevice));
            }
            NRRS_TRY(dmalloc(reinterpret_cast<SegInfo**>(&ns.segs), segs.size()));
            NRRS_HIP(hipMemcpy(ns.segs, segs.data(), segs.size() * sizeof(SegInfo), hipMemcpyHostToDevice));
            ns.n_direct = (uint32_t)direct.size();
            ns.n_chunked = (uint32_t)chunked.size();
            ns.n_chunks = (uint32_t)chunks.size();
            if (ns.n_direct) {
                NRRS_TRY(dmalloc(&ns.direct, direct.size()));
                NRRS_HIP(hipMemcpy(ns.direct, direct.data(), direct.size() * 4, hipMemcpyHostToDevice));
            }
            if (ns.n_chunks) {
                NRRS_TRY(dmalloc(&ns.chunked, chunked.size()));
                NRRS_HIP(hipMemcpy(ns.chunked, chunked.data(), chunked.size() * 4, hipMemcpyHostToDevice));
                NRRS_TRY(dmalloc(&ns.chunks, chunks.size()));
                NRRS_HIP(hipMemcpy(ns.chunks, chunks.data(), chunks.size() * sizeof(uint2), hipMemcpyHostToDevice));
                NRRS_TRY(dmalloc(&ns.acc, 2 * (size_t)ns.n_dense));
                NRRS_HIP(hipMemset(ns.acc, 0, 16 * (size_t)ns.n_dense));
            }
            NRRS_TRY(dmalloc(&ns.bad, 1));
            NRRS_HIP(hipMemset(ns.bad, 0, 4));
            ns.binned = true;
        }
    }
    NRRS_TRY(dmalloc(&ns.loss_scale, 1));
    NRRS_TRY(dmalloc(&ns.state, 8));
    NRRS_TRY(dmalloc(&ns.loss_partial, c->max_train_blocks));
    NRRS_HIP(hipMemset(ns.grad, 0, np * 4));
    NRRS_HIP(hipMemset(ns.adam_m, 0, np * 4));
    NRRS_HIP(hipMemset(ns.adam_v, 0, np * 4));
    NRRS_HIP(hipMemset(ns.state, 0, 32));
    const float s0 = c->cfg.precision == NRRS_FP32 ? 1.0f : c->cfg.loss_scale_init;
    NRRS_HIP(hipMemcpy(ns.loss_scale, &s0, 4, hipMemcpyHostToDevice));
    if (!c->train_stats) NRRS_TRY(dmalloc(&c->train_stats, (size_t)6 * c->cfg.max_train));
    if (!c->scale_tmp) NRRS_TRY(dmalloc(&c->scale_tmp, 1));
    ns.train_ready = true;
    NRRS_TRY(refresh_train(c, net, nullptr));
    return hip_status(hipDeviceSynchronize());
}

// Binned-scatter arguments of one net (its records, counts and segments).
BinArgs bin_args(nrrs_ctx* c, int net, uint32_t n, const EncodeArgs& a) {
    NetState& ns = c->net[net];
    BinArgs bn{};
    bn.a = a;
    bn.gdl = ns.gdl;
    bn.gdl_stride = c->cfg.max_train;
    bn.n = n;
    bn.rec_key = ns.rec_key;
    bn.rec_val = ns.rec_val;
    bn.seg_count = ns.seg_count;
    bn.runs = ns.bin_runs;
    bn.n_runs = (n + kRunVert - 1) / kRunVert;
    bn.seg_end = ns.seg_end;
    bn.tile_sums = ns.bin_tiles;
    bn.max_tiles = ns.bin_max_tiles;
    bn.segs = static_cast<const SegInfo*>(ns.segs);
    for (int l = 0; l < kMaxLevels; ++l) {
        bn.seg_base[l] = ns.seg_base[l];
        bn.seg_nbins[l] = ns.seg_nbins[l];
    }
    bn.grad = reinterpret_cast<_Float16 __attribute__((ext_vector_type(2)))*>(ns.grad_lp);
    bn.scan_flag = ns.state + 5;
    return bn;
}

// Sums of one net's binned records (keys `key`, its own values) into its table gradient.
// `counts`: the per-segment record counts to read (default: the net's own); `keep`: leave them
// for a second apply of the same records (the joint step's RRSNet reads StatNet's).
hipError_t launch_apply(nrrs_ctx* c, int net, const uint32_t* key, hipStream_t s, uint32_t* counts = nullptr,
                        bool keep = false) {
    NetState& ns = c->net[net];
    ApplyArgs ap{};
    ap.key = key;
    ap.val = ns.rec_val;
    ap.seg_count = counts ? counts : ns.seg_count;
    ap.keep_counts = keep;
    ap.segs = static_cast<const SegInfo*>(ns.segs);
    ap.direct = ns.direct;
    ap.chunks = ns.chunks;
    ap.chunked = ns.chunked;
    ap.n_chunked = ns.n_chunked;
    ap.acc = ns.acc;
    ap.n_acc = ns.n_chunks ? ns.n_dense : 0;
    ap.chunk = ns.apply_chunk;
    ap.bad = ns.bad;
    ap.grad = reinterpret_cast<_Float16 __attribute__((ext_vector_type(2)))*>(ns.grad_lp);
    ap.nonfinite = ns.state + 3;
    return launch_apply_bins(ap, ns.n_direct, ns.n_chunks, s);
}

// Forward + loss + backward of one net into its gradient buffer.
// `pre_enc`: the level-major encoding of this batch and net (and, for a binned net, its record
// counts) already written by a paired encode pass; null = run the net's own encode pass.
// `blocks_out`: stop after the training kernel and report its workgroups (the caller then
// bins, applies and reduces).
nrrs_status run_backward(nrrs_ctx* c, int net, uint32_t n, const nrrs_vertices& v, const float* film,
                         const float* stats, uint32_t stats_stride, int kind, const float* dy,
                         const float* target_L, const nrrs_rrs_side* side, const float* scale, float* y_out,
                         uint32_t out_stride, float* loss_out, hipStream_t s, const uint32_t* pre_enc = nullptr,
                         uint32_t* blocks_out = nullptr) {
    NetState& ns = c->net[net];
    const nrrs_config& cfg = c->cfg;
    if (n > cfg.max_train) return NRRS_ERR_SHAPE;  // per-sample training buffers hold max_train
    TrainArgs t{};
    t.a = encode_args(c, v, film, stats, stats_stride);
    t.a.grid = ns.grid;
    t.a.l0_chunks = l0_chunks(cfg, net);
    // the RRSNet training input takes activationSigma(max(L2, 0)) when StatNet models sigma
    // (train.cu:196-200); inference keeps max(L2, 0)
    t.a.sigma_l2 = cfg.train_sigma && net == NRRS_RRSNET;
    t.train_sigma = cfg.train_sigma;
    t.head_first = net == NRRS_STATNET ? head_first_layer(cfg.ll2_two_head, ns.n_layers) : -1;
    t.head = head_of(cfg, net);
    if (cfg.precision != NRRS_FP32) {
        // level-major grid gathers of the forward, read back by the training kernel
        if (!c->grid_enc || n > c->grid_enc_cap) return NRRS_ERR_SHAPE;
        t.a.grid_enc = pre_enc ? pre_enc : c->grid_enc;
        t.a.grid_enc_stride = c->grid_enc_cap;
        if (!pre_enc) {
            OpScope ot(c, "encode_counts", s);
            NRRS_HIP(launch_encode_levels(cfg.precision, t.a, ns.grid_lp, n, c->grid_enc, c->grid_enc_cap, s,
                                          ns.binned ? ns.bin_runs : nullptr));
        }
    }
    t.net = kind_of(net);
    t.n = n;
    t.table_f32 = cfg.precision == NRRS_FP32 ? ns.grid_f32 : nullptr;
    t.table_lp = reinterpret_cast<const uint32_t*>(ns.grid_lp);
    t.w_fwd = ns.w_fwd;
    t.w_bwd = ns.w_bwd;
    t.bwd_hidden_base = 4;
    t.bwd_last_base = 4 + (ns.n_layers - 2) * 8;
    t.loss_scale = scale;
    t.dy_ext = dy;
    t.target_L = target_L;
    if (side) {
        t.ref_mean = side->ref_mean;
        t.error = side->error;
        t.sample_weight = side->sample_weight;
        t.q_target = side->q_target;
        t.ebar = std::fmax(side->error_sum / (float)cfg.n_pixels, 1e-4f);
        t.error_sum_dev = side->error_sum_dev;
        t.n_pixels_f = (float)cfg.n_pixels;
    }
    for (int k = 0; k < 4; ++k) t.gamma[k] = cfg.gamma[k];
    t.clamp_on = cfg.loss_clamp_on;
    t.clamp_max = cfg.loss_clamp_max;
    t.ll2_step = cfg.ll2_step;
    t.rrs_step = cfg.rrs_step;
    const int active = kind == NRRS_LOSS_LL2 ? (cfg.ll2_step == 0 ? 3 : 6) : 1;
    t.inv_norm = n ? 1.0f / ((float)n * (float)active) : 0.0f;
    t.y_out = y_out;
    t.out_stride = out_stride;
    t.grad_grid = ns.grad;
    t.grad_grid_h2 = reinterpret_cast<_Float16 __attribute__((ext_vector_type(2)))*>(ns.grad_lp);
    t.n_dense = ns.n_dense;
    if (cfg.precision == NRRS_FP32)
        t.dense_f32 = static_cast<float*>(ns.dense_rep);
    else
        t.dense_h2 = static_cast<_Float16 __attribute__((ext_vector_type(2)))*>(ns.dense_rep);
    if (ns.binned) {
        t.gdl = ns.gdl;
        t.gdl_stride = cfg.max_train;
    }
    t.dw_partial = ns.dw_partial;
    t.n_mlp = (uint32_t)ns.n_mlp;
    t.loss_partial = kind == NRRS_LOSS_EXTERNAL ? nullptr : ns.loss_partial;
    const uint32_t vpi = train_vertices_per_iter(cfg.precision, cfg.net_set);
    uint32_t iters = (n + vpi - 1) / vpi;
    uint32_t blocks = iters < c->max_train_blocks ? iters : c->max_train_blocks;
    if (!blocks) return NRRS_OK;
    {
        OpScope ot(c, op_label(net, "train_statnet", "train_rrsnet", "train_statnet2"), s);
        NRRS_HIP(launch_train(cfg.precision, t, ns.n_layers, ns.n_out, kind, blocks, s));
        if (kind != NRRS_LOSS_EXTERNAL)
            NRRS_HIP(launch_loss_finish(ns.loss_partial, blocks, loss_out, ns.state + 4, s));
    }
    if (blocks_out) {  // the caller bins, applies and reduces (joint step)
        *blocks_out = blocks;
        return NRRS_OK;
    }
    if (ns.binned) {
        {
            OpScope ot(c, op_label(net, "bin_statnet", "bin_rrsnet", "bin_statnet2"), s);
            NRRS_HIP(launch_bin_levels(bin_args(c, net, n, t.a), s));
        }
        OpScope ot(c, op_label(net, "apply_statnet", "apply_rrsnet", "apply_statnet2"), s);
        NRRS_HIP(launch_apply(c, net, ns.rec_key, s));
    }
    OpScope ot(c, op_label(net, "reduce_statnet", "reduce_rrsnet", "reduce_statnet2"), s);
    return hip_status(launch_reduce_partials(ns.dw_partial, blocks, (uint32_t)ns.n_mlp, ns.grad + ns.n_grid, s));
}

// Path-node buffer (nrrs_create): 4 words x 3 channels x 7 planes + 4 scalars.
nrrs_status ensure_nodes(nrrs_ctx* c) {
    NodeBuf& nb = c->nodes;
    if (nb.cap) return NRRS_OK;
    const uint64_t cap64 = c->cfg.max_nodes ? c->cfg.max_nodes : 4ull * c->cfg.max_queue;
    if (!cap64 || cap64 > 0x7FFFFFFFull) return NRRS_ERR_SHAPE;
    const size_t cap = (size_t)cap64, n_tiles = cap / kNodeTile + 1;
    const size_t words = cap * (7 * 3 + 2) + cap * 4 + 2 + n_tiles;
    uint32_t* m = nullptr;
    NRRS_TRY(dmalloc(&m, words));
    c->nodes_mem = m;
    float* f = reinterpret_cast<float*>(m);
    nb.pos = f; f += 3 * cap;
    nb.dir = f; f += 3 * cap;
    nb.normal = f; f += 3 * cap;
    nb.ad = f; f += 3 * cap;
    nb.as = f; f += 3 * cap;
    nb.thp = f; f += 3 * cap;
    nb.L = f; f += 3 * cap;
    nb.rough = f; f += cap;
    nb.q = f; f += cap;
    uint32_t* u = reinterpret_cast<uint32_t*>(f);
    nb.parent = reinterpret_cast<int32_t*>(u); u += cap;
    nb.pixel = u; u += cap;
    nb.depth = u; u += cap;
    nb.first = u; u += cap;
    nb.cnt = u; u += 2;
    nb.tiles = u;
    NRRS_HIP(hipMemset(nb.cnt, 0, 8));
    nb.cap = (uint32_t)cap;
    return NRRS_OK;
}

// Folds the dense-level replicas into the table gradient (and clears them).
nrrs_status fold_dense(nrrs_ctx* c, int net, hipStream_t s) {
    NetState& ns = c->net[net];
    if (!ns.dense_rep || ns.binned) return NRRS_OK;  // binned contexts do not use the replicas
    const bool f32 = c->cfg.precision == NRRS_FP32;
    return hip_status(launch_reduce_dense(f32 ? static_cast<float*>(ns.dense_rep) : nullptr, f32 ? nullptr : ns.dense_rep,
                                          ns.n_dense, ns.grad, f32 ? nullptr : ns.grad_lp, s));
}

nrrs_status run_adam(nrrs_ctx* c, int net, hipStream_t s) {
    NetState& ns = c->net[net];
    const nrrs_config& cfg = c->cfg;
    OpScope t(c, op_label(net, "adam_statnet", "adam_rrsnet", "adam_statnet2"), s);
    NRRS_TRY(fold_dense(c, net, s));
    AdamArgs a{};
    a.params = ns.params;
    a.grad = ns.grad;
    a.grad_lp = ns.grad_lp;
    a.m = ns.adam_m;
    a.v = ns.adam_v;
    a.n = ns.n_grid + ns.n_mlp;
    a.n_grid = ns.n_grid;
    a.grid_lp = ns.grid_lp;
    a.lp_precision = cfg.precision;
    a.lr = net == NRRS_RRSNET ? cfg.lr_rrsnet : cfg.lr_statnet;
    if (net != NRRS_RRSNET && all_self(cfg)) a.lr /= 5.0f;  // AllSelf (twn_AllSelf.cpp:9-16)
    a.beta1 = cfg.adam_beta1;
    a.beta2 = cfg.adam_beta2;
    a.eps = cfg.adam_eps;
    a.l2 = cfg.adam_l2_reg;
    a.loss_scale = ns.loss_scale;
    a.state = ns.state;
    a.dynamic_scale = cfg.precision != NRRS_FP32;
    a.growth_interval = cfg.loss_scale_growth_interval ? cfg.loss_scale_growth_interval : 1000;
    a.scan_on_flag = ns.binned;  // the apply kernels report non-finite table-gradient entries
    a.ema = ns.ema;               // Ema: the moving average and its grid shadow follow every taken step
    a.ema_lp = ns.grid_lp_ema;
    a.ema_decay = cfg.ema_decay;
    const float* W_ema = ns.ema ? ns.ema + ns.n_grid : nullptr;
    if (cfg.precision != NRRS_FP32 && ns.train_ready) {
        // the update, then the state and the MLP tile layouts in one launch (the grid shadows are
        // refreshed inside the Adam kernel)
        NRRS_HIP(launch_adam(a, s, false));
        return hip_status(launch_after_adam(cfg.precision, a, ns.mlp_f32, W_ema, ns.n_layers, ns.n_out, ns.mlp_lp,
                                            ns.mlp_lp_ema, ns.w_fwd, ns.w_bwd, s));
    }
    NRRS_HIP(launch_adam(a, s));
    // MLP shadows (the grid shadows are refreshed inside the Adam kernel)
    if (cfg.precision != NRRS_FP32) {
        NRRS_HIP(launch_pack_lp(cfg.precision, nullptr, 0, ns.mlp_f32, ns.n_layers, ns.n_out, nullptr, ns.mlp_lp, s));
        if (W_ema)
            NRRS_HIP(launch_pack_lp(cfg.precision, nullptr, 0, W_ema, ns.n_layers, ns.n_out, nullptr, ns.mlp_lp_ema, s));
    }
    return refresh_train(c, net, s);
}

// Frame scratch for queues of up to `cap` vertices, allocated once by nrrs_create (so a
// captured frame graph never holds pointers that a later call frees).
nrrs_status alloc_frame(nrrs_ctx* c, uint32_t cap) {
    NRRS_TRY(dmalloc(&c->frame_q, cap));
    NRRS_TRY(dmalloc(&c->frame_s, cap));
    NRRS_TRY(dmalloc(&c->frame_tiles, kFrameMaxBlocks));
    NRRS_TRY(dmalloc(&c->frame_x, 3 * (size_t)cap));
    NRRS_TRY(dmalloc(&c->frame_genc, 2 * (size_t)c->grid.n_levels * cap));
    NRRS_TRY(dmalloc(&c->frame_sums, 16));
    NRRS_TRY(dmalloc(&c->frame_bar, kFrameBarWords));
    c->frame_cap = cap;
    return NRRS_OK;
}

nrrs_status frame_params(nrrs_ctx* c, const nrrs_frame_args* a, FrameParams* f) {
    // the frame kernel runs the one-headed StatNet of the benchmark configuration and chains its
    // statistics into RRSNet
    if (c->cfg.ll2_two_head || c->cfg.rrs_aid || c->cfg.net_set != NRRS_NETS_BENCH) return NRRS_ERR_UNSUPPORTED;
    if (!a || !a->film_dev || !a->queue_dev || !a->sizes_dev || !a->status_dev) return NRRS_ERR_ARG;
    if (a->n_bounces < 1 || a->n_bounces > 16 || a->capacity == 0) return NRRS_ERR_ARG;
    if (a->capacity > c->cfg.max_queue) return NRRS_ERR_SHAPE;
    if (c->cfg.precision == NRRS_FP32) return NRRS_ERR_UNSUPPORTED;
    if (a->capacity > c->frame_cap) return NRRS_ERR_SHAPE;
    *f = FrameParams{};
    f->grid = c->grid;
    f->frame = c->frame;
    f->d = c->dec;
    const NetView ws = view_of(c, NRRS_STATNET), wr = view_of(c, NRRS_RRSNET);
    f->table_stat = reinterpret_cast<const uint32_t*>(ws.grid_lp);
    f->table_rrs = reinterpret_cast<const uint32_t*>(wr.grid_lp);
    f->table_pair = reinterpret_cast<const uint2*>(ws.pair);
    f->w_stat = reinterpret_cast<const u32x4_t*>(ws.mlp_lp);
    f->w_rrs = reinterpret_cast<const u32x4_t*>(wr.mlp_lp);
    f->film = a->film_dev;
    f->n_pixels = c->cfg.n_pixels;
    f->n_primary = a->n_primary;
    f->n_bounces = a->n_bounces;
    f->capacity = a->capacity;
    f->cpw = c->frame_m / (32 * c->frame_waves);
    f->seed = a->seed;
    f->seed_dev = a->seed_dev;
    f->advance_seed = a->advance_seed;
    f->queue = a->queue_dev;
    f->sizes = a->sizes_dev;
    f->q_out = a->q_dev;
    f->trace = a->trace_dev;
    f->status = a->status_dev;
    f->q_scr = c->frame_q;
    f->xbuf = c->frame_x;
    f->genc = c->frame_genc;
    f->s_scr = c->frame_s;
    f->tile_sums = c->frame_tiles;
    f->sums = c->frame_sums;
    f->bar = c->frame_bar;
    return NRRS_OK;
}

bool side_ok(const nrrs_rrs_side* side) {
    return side && side->error && side->sample_weight && side->q_target;
}

void free_ctx(nrrs_ctx* c) {
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->aux_stream) hipStreamDestroy(c->aux_stream);
    for (auto& e : c->op_ev)
        if (e) hipEventDestroy(e);
    hipFree(c->err_part);
    hipFree(c->nodes_mem);
    for (auto& ns : c->net) {
        hipFree(ns.params);
        hipFree(ns.grid_lp);
        hipFree(ns.mlp_lp);
        hipFree(ns.ema);
        hipFree(ns.grid_lp_ema);
        hipFree(ns.mlp_lp_ema);
        hipFree(ns.grad);
        hipFree(ns.grad_lp);
        hipFree(ns.dense_rep);
        hipFree(ns.rec_key);
        hipFree(ns.rec_val);
        hipFree(ns.gdl);
        hipFree(ns.bin_runs);
        hipFree(ns.bin_tiles);
        hipFree(ns.seg_end);
        hipFree(ns.seg_count);
        hipFree(ns.segs);
        hipFree(ns.direct);
        hipFree(ns.chunked);
        hipFree(ns.chunks);
        hipFree(ns.acc);
        hipFree(ns.bad);
        hipFree(ns.adam_m);
        hipFree(ns.adam_v);
        hipFree(ns.w_fwd);
        hipFree(ns.w_bwd);
        hipFree(ns.dw_partial);
        hipFree(ns.loss_scale);
        hipFree(ns.state);
        hipFree(ns.loss_partial);
    }
    hipFree(c->train_stats);
    hipFree(c->scale_tmp);
    if (c->frame_exec) hipGraphExecDestroy(c->frame_exec);
    if (c->frame_graph) hipGraphDestroy(c->frame_graph);
    hipFree(c->frame_q);
    hipFree(c->frame_s);
    hipFree(c->frame_tiles);
    hipFree(c->frame_x);
    hipFree(c->frame_genc);
    hipFree(c->frame_sums);
    hipFree(c->frame_bar);
    hipFree(c->sum_dev);
    hipFree(c->tile_sums);
    hipFree(c->tile_offsets);
    hipFree(c->stats_scratch);
    hipFree(c->grid_enc);
    hipFree(c->grid_enc2);
    hipFree(c->grid_enc_f32);
    hipFree(c->grid_pair);
    hipFree(c->grid_pair_ema);
    hipFree(c->q_scratch);
    hipFree(c->s_scratch);
    delete c;
}

}  // namespace

extern "C" {

const char* nrrs_status_string(nrrs_status s) {
    switch (s) {
        case NRRS_OK: return "ok";
        case NRRS_ERR_ARG: return "invalid argument";
        case NRRS_ERR_SHAPE: return "size exceeds the capacity of the context";
        case NRRS_ERR_OOM: return "device out of memory";
        case NRRS_ERR_OVERFLOW: return "survivor queue overflow";
        case NRRS_ERR_NONFINITE: return "non-finite value";
        case NRRS_ERR_HIP: return "HIP runtime error";
        case NRRS_ERR_UNSUPPORTED: return "unsupported configuration";
        case NRRS_ERR_NO_DEVICE: return "no gfx950 device";
    }
    return "unknown status";
}

size_t nrrs_struct_size(int32_t which) {
    switch (which) {
        case 0: return sizeof(nrrs_config);
        case 1: return sizeof(nrrs_vertices);
        case 2: return sizeof(nrrs_decide_args);
        case 3: return sizeof(nrrs_vertices_out);
        case 4: return sizeof(nrrs_rrs_side);
        case 5: return sizeof(nrrs_frame_args);
        case 6: return sizeof(nrrs_error_args);
        case 7: return sizeof(nrrs_training_args);
    }
    return 0;
}

nrrs_status nrrs_default_config(nrrs_config* c) {
    if (!c) return NRRS_ERR_ARG;
    std::memset(c, 0, sizeof(*c));
    c->grid.n_levels = 16;
    c->grid.n_features = 2;
    c->grid.log2_table_size = 19;
    c->grid.base_resolution = 16;
    c->grid.per_level_scale = 1.5f;
    c->statnet.n_hidden_layers = 4;
    c->statnet.n_outputs = 6;
    c->rrsnet.n_hidden_layers = 3;
    c->rrsnet.n_outputs = 1;
    c->precision = NRRS_FP32;
    c->rrs_clamp_min = 0.05f;
    c->rrs_clamp_max = 20.0f;
    c->rrs_normalize_rate = 0.85f;
    c->rrs_use_ceil = 1;
    c->rrs_no_upscale = 1;
    c->rrs_weight_window = 0;
    c->weight_window_size = 5.0f;
    c->split_cap = 8;
    c->loss_clamp_on = 1;
    c->loss_clamp_max = 10.0f;
    c->train_sigma = 0;
    c->ll2_step = 1;
    c->rrs_step = 3;
    c->gamma[0] = 0.01f;
    c->gamma[1] = 0.05f;
    c->gamma[2] = 0.0f;
    c->gamma[3] = 0.01f;
    c->lr_statnet = 1e-2f;
    c->lr_rrsnet = 1e-2f;
    c->adam_beta1 = 0.9f;
    c->adam_beta2 = 0.99f;
    c->adam_eps = 1e-15f;
    c->adam_l2_reg = 0.0f;
    c->loss_scale_init = 32768.0f;
    c->loss_scale_growth_interval = 1000;
    for (int k = 0; k < 3; ++k) {
        c->aabb_min[k] = 0.0f;
        c->aabb_max[k] = 1.0f;
    }
    c->n_pixels = 128 * 128;
    c->max_queue = 1u << 14;
    c->max_train = 1u << 14;
    c->seed = 0x4E525253ull;
    c->paired_tables = 1;
    c->fp32_level_major = 1;
    c->ll2_two_head = 0;
    c->rrs_aid = 0;
    c->net_set = NRRS_NETS_BENCH;
    c->ema_decay = 0.0f;
    return NRRS_OK;
}

nrrs_status nrrs_shipped_config(nrrs_config* c) {
    NRRS_TRY(nrrs_default_config(c));
    c->net_set = NRRS_NETS_SHIPPED;
    c->grid.n_levels = 8;  // 2net-simple.json:46-51
    c->grid.n_features = 2;
    c->grid.log2_table_size = 19;
    c->grid.base_resolution = 16;
    c->grid.per_level_scale = 2.0f;
    c->statnet.n_hidden_layers = 3;  // lines 74, 246
    c->statnet.n_outputs = 6;
    c->rrsnet.n_hidden_layers = 3;
    c->rrsnet.n_outputs = 1;
    c->loss_clamp_on = 1;  // lines 3-5
    c->loss_clamp_max = 10.0f;
    c->train_sigma = 0;
    c->gamma[0] = 0.01f;  // lines 145-148
    c->gamma[1] = 0.05f;
    c->gamma[2] = 0.0f;
    c->gamma[3] = 0.01f;
    // the integrator switches the RRS loss to its on-policy step 3 (NRRSNetwork::setStep,
    // network.cpp:578-590) once the StatNet is trained; the JSON's step 1 is its start value
    c->rrs_step = 3;
    c->lr_statnet = 5e-3f;  // lines 16-28
    c->lr_rrsnet = 5e-4f;   // lines 171-183
    c->adam_beta1 = 0.9f;
    c->adam_beta2 = 0.99f;
    c->adam_eps = 1e-15f;
    c->adam_l2_reg = 3e-6f;
    c->ema_decay = 0.95f;
    return NRRS_OK;
}

nrrs_status nrrs_grid_layout(const nrrs_grid_config* g, uint32_t* res, uint32_t* offset,
                             uint32_t* size, uint32_t* dense, float* scale, uint64_t* n_params) {
    if (!g) return NRRS_ERR_ARG;
    return make_grid(*g, nullptr, res, offset, size, dense, scale, n_params) ? NRRS_OK : NRRS_ERR_ARG;
}

nrrs_status nrrs_create(const nrrs_config* cfg, nrrs_ctx** out) {
    if (!cfg || !out) return NRRS_ERR_ARG;
    *out = nullptr;
    if (cfg->net_set != NRRS_NETS_BENCH && cfg->net_set != NRRS_NETS_SHIPPED) return NRRS_ERR_ARG;
    if (!(cfg->ema_decay >= 0.0f && cfg->ema_decay < 1.0f)) return NRRS_ERR_ARG;
    if (!valid_shapes(*cfg)) return NRRS_ERR_UNSUPPORTED;
    if (cfg->train_sigma != 0 && cfg->train_sigma != 1) return NRRS_ERR_ARG;
    if (cfg->rrs_aid != 0 && cfg->rrs_aid != 1) return NRRS_ERR_ARG;
    if (all_self(*cfg) && cfg->train_sigma) return NRRS_ERR_UNSUPPORTED;
    if (cfg->precision != NRRS_FP32 && cfg->precision != NRRS_FP16 && cfg->precision != NRRS_BF16)
        return NRRS_ERR_UNSUPPORTED;
    if (cfg->max_queue == 0 || cfg->n_pixels == 0 || cfg->split_cap == 0 || cfg->split_cap > 255 ||
        !(cfg->rrs_clamp_min > 0.0f) || !(cfg->rrs_clamp_max >= cfg->rrs_clamp_min))
        return NRRS_ERR_ARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return NRRS_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return NRRS_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return NRRS_ERR_NO_DEVICE;

    nrrs_ctx* c = new (std::nothrow) nrrs_ctx();
    if (!c) return NRRS_ERR_OOM;
    c->cfg = *cfg;
    c->device = dev;
    // the training kernel keeps one workgroup per CU resident (its LDS stage and weight tiles):
    // one persistent workgroup per CU, so the dW partials are one per CU as well; the narrow
    // kernels of the shipped FP16/BF16 nets keep two
    if (prop.multiProcessorCount > 0) c->max_train_blocks = (uint32_t)std::min(prop.multiProcessorCount, 1024);
    if (cfg->net_set == NRRS_NETS_SHIPPED && cfg->precision != NRRS_FP32) c->max_train_blocks *= 2;
    uint64_t n_grid = 0;
    if (!make_grid(cfg->grid, &c->grid, nullptr, nullptr, nullptr, nullptr, nullptr, &n_grid)) {
        delete c;
        return NRRS_ERR_ARG;
    }
    c->frame = make_frame(cfg->aabb_min, cfg->aabb_max);
    c->dec = decide_params(*cfg);
    // default frame tile shape: BF16's split StatNet tiles (110 KB of LDS) leave one workgroup
    // per CU, so it takes the 8-wave shape; FP16 (76 KB) fits two 4-wave workgroups
    if (cfg->precision == NRRS_BF16) {
        c->frame_waves = 8;
        c->frame_m = 256;
    }

    nrrs_status st = NRRS_OK;
    // paired table: one 32-bit word per entry and net, n_grid / 2 entries (both nets with a grid)
    if (cfg->precision != NRRS_FP32 && cfg->paired_tables && has_grid(*cfg, NRRS_RRSNET)) {
        st = dmalloc(&c->grid_pair, n_grid);
        if (st == NRRS_OK && cfg->ema_decay > 0.0f) st = dmalloc(&c->grid_pair_ema, n_grid);
    }
    std::vector<float> host;
    for (int net = 0; net < n_nets(*cfg) && st == NRRS_OK; ++net) {
        const nrrs_mlp_config& m = mlp_of(*cfg, net);
        NetState& ns = c->net[net];
        ns.n_layers = (int)m.n_hidden_layers + 1;
        ns.n_out = (int)m.n_outputs;
        if (has_grid(*cfg, net)) {
            ns.grid = c->grid;
            ns.n_grid = n_grid;
        } else {
            ns.grid = GridParams{};  // n_levels 0: the encoders write no grid channels
            ns.grid.hash_mask = c->grid.hash_mask;
            ns.n_grid = 0;
        }
        ns.n_mlp = mlp_count(m);
        if ((st = dmalloc(&ns.params, ns.n_grid + ns.n_mlp)) != NRRS_OK) break;
        ns.grid_f32 = ns.params;
        ns.mlp_f32 = ns.params + ns.n_grid;
        if (cfg->precision != NRRS_FP32) {
            if ((st = dmalloc(&ns.grid_lp, ns.n_grid)) != NRRS_OK) break;
            if ((st = dmalloc(&ns.mlp_lp, mfma_packed_halves(ns.n_layers))) != NRRS_OK) break;
        }
        if (cfg->ema_decay > 0.0f) {  // the moving average starts at the initial parameters
            if ((st = dmalloc(&ns.ema, ns.n_grid + ns.n_mlp)) != NRRS_OK) break;
            if (cfg->precision != NRRS_FP32) {
                if ((st = dmalloc(&ns.grid_lp_ema, ns.n_grid)) != NRRS_OK) break;
                if ((st = dmalloc(&ns.mlp_lp_ema, mfma_packed_halves(ns.n_layers))) != NRRS_OK) break;
            }
        }
        init_params(*cfg, net, ns.n_grid, host);
        st = hip_status(hipMemcpy(ns.grid_f32, host.data(), ns.n_grid * 4, hipMemcpyHostToDevice));
        if (st == NRRS_OK)
            st = hip_status(hipMemcpy(ns.mlp_f32, host.data() + ns.n_grid, ns.n_mlp * 4, hipMemcpyHostToDevice));
        if (st == NRRS_OK && ns.ema)
            st = hip_status(hipMemcpy(ns.ema, host.data(), (ns.n_grid + ns.n_mlp) * 4, hipMemcpyHostToDevice));
        if (st == NRRS_OK) st = refresh_lp(c, net, nullptr);
        if (st == NRRS_OK) st = hip_status(hipDeviceSynchronize());
    }
    c->max_tiles = (cfg->max_queue + 255) / 256;
    if (st == NRRS_OK) st = dmalloc(&c->sum_dev, 1);
    if (st == NRRS_OK) st = dmalloc(&c->tile_sums, c->max_tiles);
    if (st == NRRS_OK) st = dmalloc(&c->tile_offsets, c->max_tiles);
    if (st == NRRS_OK) st = dmalloc(&c->stats_scratch, (size_t)6 * cfg->max_queue);
    if (st == NRRS_OK && cfg->precision == NRRS_FP32 && cfg->fp32_level_major) {
        c->grid_enc_cap = cfg->max_queue > cfg->max_train ? cfg->max_queue : cfg->max_train;
        st = dmalloc(&c->grid_enc_f32, (size_t)c->grid.n_levels * c->grid_enc_cap);
    }
    if (st == NRRS_OK && cfg->precision != NRRS_FP32) {
        c->grid_enc_cap = cfg->max_queue > cfg->max_train ? cfg->max_queue : cfg->max_train;
        st = dmalloc(&c->grid_enc, (size_t)c->grid.n_levels * c->grid_enc_cap);
        if (st == NRRS_OK) st = dmalloc(&c->grid_enc2, (size_t)c->grid.n_levels * c->grid_enc_cap);
    }
    if (st == NRRS_OK) st = dmalloc(&c->q_scratch, cfg->max_queue);
    if (st == NRRS_OK) st = dmalloc(&c->s_scratch, cfg->max_queue);
    // helper stream at the default priority, and the events of its fork and join
    if (st == NRRS_OK) st = hip_status(hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking));
    if (st == NRRS_OK) st = hip_status(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    if (st == NRRS_OK) st = hip_status(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    // Everything the entry points use is allocated here, as NRRSNetwork::resize allocates its
    // buffers up front (network.cpp:351-421): no later call allocates or synchronises, so a
    // first call can be captured into a graph.
    if (st == NRRS_OK && cfg->max_train) st = ensure_train(c, NRRS_STATNET);
    if (st == NRRS_OK && cfg->max_train) st = ensure_train(c, NRRS_RRSNET);
    if (st == NRRS_OK && cfg->max_train && all_self(*cfg)) st = ensure_train(c, NRRS_STATNET2);
    if (st == NRRS_OK) st = ensure_nodes(c);
    if (st == NRRS_OK) st = dmalloc(&c->err_part, error_partial_count());
    if (st == NRRS_OK && cfg->precision != NRRS_FP32 && !cfg->ll2_two_head && !cfg->rrs_aid &&
        cfg->net_set == NRRS_NETS_BENCH)
        st = alloc_frame(c, cfg->max_queue);
    if (st == NRRS_OK) st = hip_status(hipDeviceSynchronize());
    if (st != NRRS_OK) {
        free_ctx(c);
        return st;
    }
    *out = c;
    return NRRS_OK;
}

nrrs_status nrrs_destroy(nrrs_ctx* ctx) {
    if (!ctx) return NRRS_ERR_ARG;
    hipDeviceSynchronize();
    free_ctx(ctx);
    return NRRS_OK;
}

nrrs_status nrrs_param_count(nrrs_ctx* ctx, int32_t net, uint64_t* n_grid, uint64_t* n_mlp) {
    if (!ctx || !net_ok(ctx, net)) return NRRS_ERR_ARG;
    if (n_grid) *n_grid = ctx->net[net].n_grid;
    if (n_mlp) *n_mlp = ctx->net[net].n_mlp;
    return NRRS_OK;
}

nrrs_status nrrs_params_get(nrrs_ctx* ctx, int32_t net, float* host_dst) {
    if (!ctx || !host_dst || !net_ok(ctx, net)) return NRRS_ERR_ARG;
    const NetState& ns = ctx->net[net];
    NRRS_HIP(hipDeviceSynchronize());
    NRRS_HIP(hipMemcpy(host_dst, ns.grid_f32, ns.n_grid * 4, hipMemcpyDeviceToHost));
    NRRS_HIP(hipMemcpy(host_dst + ns.n_grid, ns.mlp_f32, ns.n_mlp * 4, hipMemcpyDeviceToHost));
    return NRRS_OK;
}

nrrs_status nrrs_params_set(nrrs_ctx* ctx, int32_t net, const float* host_src) {
    if (!ctx || !host_src || !net_ok(ctx, net)) return NRRS_ERR_ARG;
    NetState& ns = ctx->net[net];
    NRRS_HIP(hipDeviceSynchronize());
    NRRS_HIP(hipMemcpy(ns.grid_f32, host_src, ns.n_grid * 4, hipMemcpyHostToDevice));
    // the weights between the heads, or outside the shipped 32-wide nets, stay zero
    if ((net != NRRS_RRSNET && ctx->cfg.ll2_two_head) || ctx->cfg.net_set == NRRS_NETS_SHIPPED) {
        std::vector<float> mlp(host_src + ns.n_grid, host_src + ns.n_grid + ns.n_mlp);
        zero_between_heads(ctx->cfg, net, mlp.data());
        NRRS_HIP(hipMemcpy(ns.mlp_f32, mlp.data(), ns.n_mlp * 4, hipMemcpyHostToDevice));
    } else {
        NRRS_HIP(hipMemcpy(ns.mlp_f32, host_src + ns.n_grid, ns.n_mlp * 4, hipMemcpyHostToDevice));
    }
    // setting the parameters restarts the moving average from them
    if (ns.ema)
        NRRS_HIP(hipMemcpy(ns.ema, ns.params, (ns.n_grid + ns.n_mlp) * 4, hipMemcpyDeviceToDevice));
    NRRS_TRY(refresh_lp(ctx, net, nullptr));
    NRRS_TRY(refresh_train(ctx, net, nullptr));
    return hip_status(hipDeviceSynchronize());
}

nrrs_status nrrs_params_get_inference(nrrs_ctx* ctx, int32_t net, float* host_dst) {
    if (!ctx || !host_dst || !net_ok(ctx, net)) return NRRS_ERR_ARG;
    const NetState& ns = ctx->net[net];
    NRRS_HIP(hipDeviceSynchronize());
    const float* src = ns.ema ? ns.ema : ns.params;
    NRRS_HIP(hipMemcpy(host_dst, src, (ns.n_grid + ns.n_mlp) * 4, hipMemcpyDeviceToHost));
    return NRRS_OK;
}

nrrs_status nrrs_encode(nrrs_ctx* ctx, int32_t net, uint32_t n, const nrrs_vertices* v,
                        const float* film_dev, const float* stats_dev, float* enc_dev, void* stream) {
    if (!ctx || !net_ok(ctx, net) || !vertices_ok(v, n) || !enc_dev) return NRRS_ERR_ARG;
    if (net == NRRS_RRSNET && !film_dev) return NRRS_ERR_ARG;
    EncodeArgs a = encode_args(ctx, *v, film_dev, net == NRRS_RRSNET ? stats_dev : nullptr, n);
    a.grid = ctx->net[net].grid;
    return hip_status(launch_encode(a, kind_of(net), n, view_of(ctx, net).grid_f32, enc_dev, (hipStream_t)stream));
}

nrrs_status nrrs_grid_indices(nrrs_ctx* ctx, uint32_t n, const nrrs_vertices* v, uint32_t* idx_dev,
                              void* stream) {
    if (!ctx || !v || !v->pos || v->stride < n || !idx_dev) return NRRS_ERR_ARG;
    EncodeArgs a = encode_args(ctx, *v, nullptr, nullptr, 0);
    return hip_status(launch_grid_indices(a, n, idx_dev, (hipStream_t)stream));
}

nrrs_status nrrs_statnet_infer(nrrs_ctx* ctx, uint32_t n, const nrrs_vertices* v, float* stats_dev,
                               void* stream) {
    if (!ctx || !vertices_ok(v, n) || !stats_dev) return NRRS_ERR_ARG;
    EncodeArgs a = encode_args(ctx, *v, nullptr, nullptr, 0);
    return hip_status(run_statnets(ctx, a, stats_dev, n, n, (hipStream_t)stream));
}

nrrs_status nrrs_rrsnet_infer(nrrs_ctx* ctx, uint32_t n, const nrrs_vertices* v, const float* film_dev,
                              const float* stats_dev, float* raw_dev, void* stream) {
    if (!ctx || !vertices_ok(v, n) || !film_dev || !raw_dev) return NRRS_ERR_ARG;
    // the shipped RRSNet without the AID input reads the StatNet statistics (NRRSNetworkInput1)
    if (!stats_dev && ctx->cfg.net_set == NRRS_NETS_SHIPPED && !ctx->cfg.rrs_aid) return NRRS_ERR_ARG;
    EncodeArgs a = encode_args(ctx, *v, film_dev, stats_dev, n);
    EpiArgs ep{};
    ep.out = raw_dev;
    ep.out_stride = n;
    return hip_status(run_net(ctx, NRRS_RRSNET, a, EPI_RAW, n, ep, (hipStream_t)stream));
}

static nrrs_status check_decide_args(nrrs_ctx* ctx, const nrrs_decide_args* a) {
    if (!ctx || !a || !a->q_dev || !a->survivors_dev || !a->n_survivors_dev) return NRRS_ERR_ARG;
    if (a->n > ctx->cfg.max_queue) return NRRS_ERR_SHAPE;
    if (a->n_spec_dev && !a->spec_stid_dev) return NRRS_ERR_ARG;
    return NRRS_OK;
}

nrrs_status nrrs_rrs_infer_decide(nrrs_ctx* ctx, const nrrs_decide_args* a, void* stream) {
    NRRS_TRY(check_decide_args(ctx, a));
    if (!vertices_ok(&a->v, a->n) || !a->film_dev) return NRRS_ERR_ARG;
    // the AID input has no statistics channels: RRSNet runs alone (network.cpp:303-343)
    if (a->q_mode == NRRS_Q_RRSNET_CHAINED && ctx->cfg.rrs_aid) return NRRS_ERR_UNSUPPORTED;
    // the shipped RRSNet without the AID input runs only on StatNet statistics
    if (a->q_mode == NRRS_Q_RRSNET && ctx->cfg.net_set == NRRS_NETS_SHIPPED && !ctx->cfg.rrs_aid)
        return NRRS_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    const bool chain = a->q_mode == NRRS_Q_RRSNET_CHAINED && ctx->cfg.precision != NRRS_FP32 &&
                       !ctx->cfg.ll2_two_head && ctx->grid_enc && a->n <= ctx->grid_enc_cap;
    const bool ship_chain = chain && ctx->cfg.net_set == NRRS_NETS_SHIPPED;
    const bool bench_chain = chain && ctx->cfg.net_set == NRRS_NETS_BENCH && ctx->grid_enc2;
    // the chained paths' encode pass clears the sum; the others a launch of its own
    if (!((ship_chain || bench_chain) && a->n)) {
        OpScope t(ctx, "zero_sum", s);
        NRRS_HIP(launch_zero_u64(ctx->sum_dev, s));
    }
    EpiArgs ep{};
    ep.q = a->q_dev;
    ep.sum = ctx->sum_dev;
    ep.d = ctx->dec;
    if (a->q_mode == NRRS_Q_STATNET_ADRRS) {
        EncodeArgs ea = encode_args(ctx, a->v, a->film_dev, nullptr, 0);
        NRRS_HIP(run_net(ctx, NRRS_STATNET, ea, EPI_ADRRS, a->n, ep, s));
    } else if (ship_chain) {
        // shipped nets: StatNet's level-major grid encoding, then StatNet -> RRSNet (no grid) -> q
        // in one launch
        const nrrs_config& cfg = ctx->cfg;
        EncodeArgs ea = encode_args(ctx, a->v, a->film_dev, nullptr, 0);
        ea.grid = ctx->net[NRRS_STATNET].grid;
        ea.l0_chunks = l0_chunks(cfg, NRRS_STATNET);
        ea.grid_enc = ctx->grid_enc;
        ea.grid_enc_stride = ctx->grid_enc_cap;
        ea.zero_u64 = ctx->sum_dev;
        const NetView ws = view_of(ctx, NRRS_STATNET), wr = view_of(ctx, NRRS_RRSNET);
        {
            OpScope t(ctx, "encode_statnet", s);
            NRRS_HIP(launch_encode_levels(cfg.precision, ea, ws.grid_lp, a->n, ctx->grid_enc, ctx->grid_enc_cap, s));
        }
        OpScope t(ctx, "mlp_chained", s);
        NRRS_HIP(launch_net_chained(cfg.precision, ea, nullptr, a->n, ws.mlp_lp, wr.mlp_lp, ep, s));
    } else if (bench_chain) {
        // both nets' level-major grid encodings (one pass over the paired table, else a pass per
        // net), then StatNet -> RRSNet -> q in one launch
        const nrrs_config& cfg = ctx->cfg;
        EncodeArgs ea = encode_args(ctx, a->v, a->film_dev, nullptr, 0);
        ea.grid_enc = ctx->grid_enc;
        ea.grid_enc_stride = ctx->grid_enc_cap;
        ea.zero_u64 = ctx->sum_dev;
        const NetView ws = view_of(ctx, NRRS_STATNET), wr = view_of(ctx, NRRS_RRSNET);
        if (paired_pass(ctx, a->n)) {
            OpScope t(ctx, "encode_pair", s);
            NRRS_HIP(launch_encode_levels_dual(cfg.precision, ea, ws.pair, a->n, ctx->grid_enc, ctx->grid_enc2,
                                               ctx->grid_enc_cap, s));
        } else {
            OpScope t(ctx, "encode_two", s);
            NRRS_HIP(launch_encode_levels(cfg.precision, ea, ws.grid_lp, a->n, ctx->grid_enc, ctx->grid_enc_cap, s));
            NRRS_HIP(launch_encode_levels(cfg.precision, ea, wr.grid_lp, a->n, ctx->grid_enc2, ctx->grid_enc_cap, s));
        }
        OpScope t(ctx, "mlp_chained", s);
        NRRS_HIP(launch_net_chained(cfg.precision, ea, ctx->grid_enc2, a->n, ws.mlp_lp, wr.mlp_lp, ep, s));
    } else if (a->q_mode == NRRS_Q_RRSNET || a->q_mode == NRRS_Q_RRSNET_CHAINED) {
        const float* stats = nullptr;
        if (a->q_mode == NRRS_Q_RRSNET_CHAINED) {  // FP32, two-headed StatNet, or a batch beyond the encoding buffers
            EncodeArgs ea = encode_args(ctx, a->v, nullptr, nullptr, 0);
            NRRS_HIP(run_statnets(ctx, ea, ctx->stats_scratch, ctx->cfg.max_queue, a->n, s));
            stats = ctx->stats_scratch;
        }
        EncodeArgs ea = encode_args(ctx, a->v, a->film_dev, stats, ctx->cfg.max_queue);
        NRRS_HIP(run_net(ctx, NRRS_RRSNET, ea, EPI_RRS, a->n, ep, s));
    } else {
        return NRRS_ERR_ARG;
    }
    OpScope t(ctx, "decide", s);
    return hip_status(launch_decide(ctx, *a, s));
}

nrrs_status nrrs_rrs_decide(nrrs_ctx* ctx, const nrrs_decide_args* a, void* stream) {
    NRRS_TRY(check_decide_args(ctx, a));
    hipStream_t s = (hipStream_t)stream;
    NRRS_HIP(launch_qsum(ctx, a->q_dev, a->n, s));
    return hip_status(launch_decide(ctx, *a, s));
}

nrrs_status nrrs_rrs_decide_static(const nrrs_config* cfg, const nrrs_decide_args* a, void* stream) {
    if (!cfg || !a || !a->q_dev || !a->survivors_dev || !a->n_survivors_dev) return NRRS_ERR_ARG;
    if (a->n_spec_dev && !a->spec_stid_dev) return NRRS_ERR_ARG;
    if (cfg->n_pixels == 0 || cfg->split_cap == 0 || cfg->split_cap > 255) return NRRS_ERR_ARG;
    static_assert(NRRS_STATIC_DECIDE_MAX == kStaticDecideMax, "scratch size of the allocation-free decision");
    if (a->n > NRRS_STATIC_DECIDE_MAX) return NRRS_ERR_SHAPE;
    const DecideParams d = decide_params(*cfg);
    return hip_status(launch_decide_static(d, *a, (hipStream_t)stream));
}

nrrs_status nrrs_synth_vertices(uint32_t n, uint64_t seed, uint32_t n_pixels, const nrrs_vertices_out* o,
                                void* stream) {
    if (!o || !o->pos || !o->dir || !o->normal || !o->rough || !o->albedo_d || !o->albedo_s || !o->thp ||
        !o->pixel || o->stride < n || n_pixels == 0)
        return NRRS_ERR_ARG;
    return hip_status(launch_synth(n, seed, n_pixels, *o, (hipStream_t)stream));
}

nrrs_status nrrs_backward(nrrs_ctx* ctx, int32_t net, uint32_t n, const nrrs_vertices* v, const float* film_dev,
                          const float* stats_dev, int32_t kind, const float* dy_dev, const float* target_L_dev,
                          const nrrs_rrs_side* side, float loss_scale, float* loss_dev, void* stream) {
    if (!ctx || !net_ok(ctx, net) || !vertices_ok(v, n)) return NRRS_ERR_ARG;
    if (n > ctx->cfg.max_train) return NRRS_ERR_SHAPE;
    if (kind == NRRS_LOSS_EXTERNAL) {
        if (!dy_dev) return NRRS_ERR_ARG;
    } else if (kind == NRRS_LOSS_LL2) {
        if (net == NRRS_RRSNET || !target_L_dev) return NRRS_ERR_ARG;
    } else if (kind == NRRS_LOSS_RRS) {
        if (net != NRRS_RRSNET || !side_ok(side) || !stats_dev) return NRRS_ERR_ARG;
    } else {
        return NRRS_ERR_ARG;
    }
    if (net == NRRS_RRSNET && !film_dev) return NRRS_ERR_ARG;
    if (!ctx->net[net].train_ready) return NRRS_ERR_SHAPE;  // max_train 0
    hipStream_t s = (hipStream_t)stream;
    NRRS_HIP(launch_fill_f32(ctx->scale_tmp, loss_scale, 1, s));
    if (loss_dev) NRRS_HIP(launch_fill_f32(loss_dev, 0.0f, 1, s));
    return run_backward(ctx, net, n, *v, film_dev, net == NRRS_RRSNET ? stats_dev : nullptr, n, kind, dy_dev,
                        target_L_dev, side, ctx->scale_tmp, nullptr, 0, loss_dev, s);
}

nrrs_status nrrs_grads_zero(nrrs_ctx* ctx, int32_t net, void* stream) {
    if (!ctx || !net_ok(ctx, net)) return NRRS_ERR_ARG;
    if (!ctx->net[net].train_ready) return NRRS_ERR_SHAPE;  // max_train 0
    NetState& ns = ctx->net[net];
    if (ns.grad_lp)
        NRRS_HIP(launch_fill_f32(reinterpret_cast<float*>(ns.grad_lp), 0.0f, ns.n_grid / 2, (hipStream_t)stream));
    if (ns.dense_rep)
        NRRS_HIP(launch_fill_f32(static_cast<float*>(ns.dense_rep), 0.0f,
                                 (uint64_t)kDenseGradReplicas * ns.n_dense * (ctx->cfg.precision == NRRS_FP32 ? 2 : 1),
                                 (hipStream_t)stream));
    return hip_status(launch_fill_f32(ns.grad, 0.0f, ns.n_grid + ns.n_mlp, (hipStream_t)stream));
}

nrrs_status nrrs_grads_get(nrrs_ctx* ctx, int32_t net, float* host_dst) {
    if (!ctx || !net_ok(ctx, net) || !host_dst) return NRRS_ERR_ARG;
    if (!ctx->net[net].train_ready) return NRRS_ERR_SHAPE;  // max_train 0
    NetState& ns = ctx->net[net];
    NRRS_TRY(fold_dense(ctx, net, nullptr));
    NRRS_HIP(hipDeviceSynchronize());
    NRRS_HIP(hipMemcpy(host_dst, ns.grad, (ns.n_grid + ns.n_mlp) * 4, hipMemcpyDeviceToHost));
    if (ns.grad_lp) {  // FP16 grid gradient of FP16/BF16 contexts
        std::vector<_Float16> h(ns.n_grid);
        NRRS_HIP(hipMemcpy(h.data(), ns.grad_lp, ns.n_grid * 2, hipMemcpyDeviceToHost));
        for (uint64_t k = 0; k < ns.n_grid; ++k) host_dst[k] = (float)h[k];
    }
    return NRRS_OK;
}

nrrs_status nrrs_adam_step(nrrs_ctx* ctx, int32_t net, void* stream) {
    if (!ctx || !net_ok(ctx, net)) return NRRS_ERR_ARG;
    if (!ctx->net[net].train_ready) return NRRS_ERR_SHAPE;  // max_train 0
    NRRS_TRY(run_adam(ctx, net, (hipStream_t)stream));
    return refresh_pair(ctx, (hipStream_t)stream);
}

nrrs_status nrrs_train_state(nrrs_ctx* ctx, int32_t net, uint32_t* step, float* loss_scale, uint32_t* skipped) {
    if (!ctx || !net_ok(ctx, net)) return NRRS_ERR_ARG;
    if (!ctx->net[net].train_ready) return NRRS_ERR_SHAPE;  // max_train 0
    NetState& ns = ctx->net[net];
    uint32_t st[8];
    float sc;
    NRRS_HIP(hipDeviceSynchronize());
    NRRS_HIP(hipMemcpy(st, ns.state, 32, hipMemcpyDeviceToHost));
    NRRS_HIP(hipMemcpy(&sc, ns.loss_scale, 4, hipMemcpyDeviceToHost));
    if (step) *step = st[0];
    if (loss_scale) *loss_scale = sc;
    if (skipped) *skipped = st[2];
    if (st[4]) {
        // a batch loss was NaN/Inf since the previous call (the reference stops on a NaN
        // loss, integrator.cpp:2650-2656): report it once
        NRRS_HIP(hipMemset(ns.state + 4, 0, 4));
        return NRRS_ERR_NONFINITE;
    }
    return NRRS_OK;
}

nrrs_status nrrs_train_step(nrrs_ctx* ctx, uint32_t net_mask, uint32_t n, const nrrs_vertices* v,
                            const float* film_dev, const float* target_L_dev, const nrrs_rrs_side* side,
                            float* loss_dev, void* stream) {
    if (!ctx || !vertices_ok(v, n) || !(net_mask & 3u)) return NRRS_ERR_ARG;
    if (n > ctx->cfg.max_train) return NRRS_ERR_SHAPE;
    if ((net_mask & 1u) && !target_L_dev) return NRRS_ERR_ARG;
    if ((net_mask & 2u) && (!side_ok(side) || !film_dev)) return NRRS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (loss_dev) NRRS_HIP(launch_fill_f32(loss_dev, 0.0f, 2, s));
    const uint32_t ms = ctx->cfg.max_train;
    bool forked = false;
    if (!ctx->net[NRRS_STATNET].train_ready || !ctx->net[NRRS_RRSNET].train_ready) return NRRS_ERR_SHAPE;
    const bool two_stat = all_self(ctx->cfg);  // AllSelf: a second StatNet with its own table
    // joint step: both nets' encodings (and the record counts, the same for both) in one pass
    // over the paired table, before either net's parameters change
    const bool paired =
        (net_mask & 3u) == 3u && ctx->cfg.precision != NRRS_FP32 && !two_stat && paired_pass(ctx, n);
    NetState& s0 = ctx->net[NRRS_STATNET];
    NetState& s1 = ctx->net[NRRS_RRSNET];
    // both nets binned with one segment layout: one record pass writes the records of both
    const bool joint_records = paired && s0.binned && s1.binned && s0.n_segs == s1.n_segs;
    if (paired) {
        EncodeArgs ea = encode_args(ctx, *v, nullptr, nullptr, 0);
        uint32_t* rc = s0.binned ? s0.bin_runs : (s1.binned ? s1.bin_runs : nullptr);
        uint32_t* rc2 = s0.binned && s1.binned && !joint_records ? s1.bin_runs : nullptr;
        OpScope t(ctx, "encode_pair_counts", s);
        NRRS_HIP(launch_encode_levels_dual(ctx->cfg.precision, ea, ctx->grid_pair, n, ctx->grid_enc, ctx->grid_enc2,
                                           ctx->grid_enc_cap, s, rc, rc2));
    }
    if (joint_records) {
        // both training kernels, then the scan and one record pass for both nets, then each
        // net's apply, dW reduction and update (StatNet's on the helper stream)
        uint32_t b0 = 0, b1 = 0;
        NRRS_TRY(run_backward(ctx, NRRS_STATNET, n, *v, film_dev, nullptr, 0, NRRS_LOSS_LL2, nullptr, target_L_dev,
                              nullptr, s0.loss_scale, ctx->train_stats, ms, loss_dev, s, ctx->grid_enc, &b0));
        NRRS_TRY(run_backward(ctx, NRRS_RRSNET, n, *v, film_dev, ctx->train_stats, ms, NRRS_LOSS_RRS, nullptr,
                              nullptr, side, s1.loss_scale, nullptr, 0, loss_dev ? loss_dev + 1 : nullptr, s,
                              ctx->grid_enc2, &b1));
        if (b0) {
            BinArgs bn = bin_args(ctx, NRRS_STATNET, n, encode_args(ctx, *v, nullptr, nullptr, 0));
            {
                OpScope t(ctx, "bin_scan", s);
                NRRS_HIP(launch_bin_scan(bn, s));
            }
            bn.gdl2 = s1.gdl;
            bn.rec_val2 = s1.rec_val;
            bn.scan_flag2 = s1.state + 5;
            bn.grad2 = reinterpret_cast<_Float16 __attribute__((ext_vector_type(2)))*>(s1.grad_lp);
            {
                OpScope t(ctx, "bin_records_pair", s);
                NRRS_HIP(launch_bin_records(bn, s));
            }
            {
                // RRSNet's records take StatNet's slots: its apply reads StatNet's per-segment
                // counts after this one, which leaves them
                OpScope t(ctx, "apply_statnet", s);
                NRRS_HIP(launch_apply(ctx, NRRS_STATNET, s0.rec_key, s, nullptr, true));
            }
            OpScope t(ctx, "reduce_statnet", s);
            NRRS_HIP(launch_reduce_partials(s0.dw_partial, b0, (uint32_t)s0.n_mlp, s0.grad + s0.n_grid, s));
        }
        // the StatNet update touches only StatNet's buffers: helper stream, beside RRSNet's apply
        if (ctx->aux_stream) {
            NRRS_HIP(hipEventRecord(ctx->ev_fork, s));
            NRRS_HIP(hipStreamWaitEvent(ctx->aux_stream, ctx->ev_fork, 0));
            const nrrs_status st = run_adam(ctx, NRRS_STATNET, ctx->aux_stream);
            NRRS_HIP(hipEventRecord(ctx->ev_join, ctx->aux_stream));
            forked = true;
            if (st != NRRS_OK) {
                NRRS_HIP(hipStreamWaitEvent(s, ctx->ev_join, 0));
                return st;
            }
        } else {
            NRRS_TRY(run_adam(ctx, NRRS_STATNET, s));
        }
        nrrs_status st = NRRS_OK;
        if (b1) {
            hipError_t e;
            {
                OpScope t(ctx, "apply_rrsnet", s);
                e = launch_apply(ctx, NRRS_RRSNET, s0.rec_key, s, s0.seg_count);
            }
            if (e == hipSuccess) {
                OpScope t(ctx, "reduce_rrsnet", s);
                e = launch_reduce_partials(s1.dw_partial, b1, (uint32_t)s1.n_mlp, s1.grad + s1.n_grid, s);
            }
            st = hip_status(e);
        }
        if (st == NRRS_OK) st = run_adam(ctx, NRRS_RRSNET, s);
        if (forked) {
            const nrrs_status sj = hip_status(hipStreamWaitEvent(s, ctx->ev_join, 0));
            if (st == NRRS_OK) st = sj;
        }
        if (st != NRRS_OK) return st;
        return refresh_pair(ctx, s);
    }
    if (net_mask & 1u) {
        NetState& ns = s0;
        // StatNet step (myNetworkLL2.cpp:41-89); its forward outputs feed RRSNet
        NRRS_TRY(run_backward(ctx, NRRS_STATNET, n, *v, film_dev, nullptr, 0, NRRS_LOSS_LL2, nullptr, target_L_dev,
                              nullptr, ns.loss_scale, ctx->train_stats, ms, loss_dev, s,
                              paired ? ctx->grid_enc : nullptr));
        // AllSelf: the StatNet of L^2 (outputs 3-5 of the statistics; its loss adds to loss_dev[0])
        if (two_stat)
            NRRS_TRY(run_backward(ctx, NRRS_STATNET2, n, *v, film_dev, nullptr, 0, NRRS_LOSS_LL2, nullptr,
                                  target_L_dev, nullptr, ctx->net[NRRS_STATNET2].loss_scale, ctx->train_stats, ms,
                                  loss_dev, s));
        if ((net_mask & 2u) && ctx->aux_stream) {
            // The StatNet update touches only StatNet's buffers: it runs on the helper stream
            // beside the RRSNet backward.
            NRRS_HIP(hipEventRecord(ctx->ev_fork, s));
            NRRS_HIP(hipStreamWaitEvent(ctx->aux_stream, ctx->ev_fork, 0));
            nrrs_status st = run_adam(ctx, NRRS_STATNET, ctx->aux_stream);
            if (st == NRRS_OK && two_stat) st = run_adam(ctx, NRRS_STATNET2, ctx->aux_stream);
            NRRS_HIP(hipEventRecord(ctx->ev_join, ctx->aux_stream));
            forked = true;
            if (st != NRRS_OK) {
                NRRS_HIP(hipStreamWaitEvent(s, ctx->ev_join, 0));
                return st;
            }
        } else {
            NRRS_TRY(run_adam(ctx, NRRS_STATNET, s));
            if (two_stat) NRRS_TRY(run_adam(ctx, NRRS_STATNET2, s));
        }
    }
    // the helper stream's work joins the caller's stream on every return from here
    struct Join {
        nrrs_ctx* c;
        hipStream_t s;
        bool on;
        ~Join() {
            if (on) (void)hipStreamWaitEvent(s, c->ev_join, 0);
        }
    } join{ctx, s, forked};
    if (net_mask & 2u) {
        if (!(net_mask & 1u)) {  // the statistics of the StatNet being trained, as its training forward gives
            EncodeArgs ea = encode_args(ctx, *v, nullptr, nullptr, 0);
            NRRS_HIP(run_statnets(ctx, ea, ctx->train_stats, ms, n, s, true));
        }
        NetState& ns = ctx->net[NRRS_RRSNET];
        NRRS_TRY(run_backward(ctx, NRRS_RRSNET, n, *v, film_dev, ctx->train_stats, ms, NRRS_LOSS_RRS, nullptr,
                              nullptr, side, ns.loss_scale, nullptr, 0, loss_dev ? loss_dev + 1 : nullptr, s,
                              paired ? ctx->grid_enc2 : nullptr));
        NRRS_TRY(run_adam(ctx, NRRS_RRSNET, s));
    }
    // the paired table follows both nets' updated grids once the StatNet update has joined
    if (join.on) {
        NRRS_HIP(hipStreamWaitEvent(s, ctx->ev_join, 0));
        join.on = false;
    }
    return refresh_pair(ctx, s);
}

nrrs_status nrrs_set_op_timing(nrrs_ctx* ctx, int32_t on) {
    if (!ctx) return NRRS_ERR_ARG;
    if (on && !ctx->op_ev[0]) {
        for (auto& e : ctx->op_ev) NRRS_HIP(hipEventCreate(&e));
    }
    ctx->op_timing = on != 0;
    ctx->op_count = 0;
    return NRRS_OK;
}

nrrs_status nrrs_op_times(nrrs_ctx* ctx, uint32_t* count, const char** names, float* ms) {
    if (!ctx || !count) return NRRS_ERR_ARG;
    const uint32_t n = ctx->op_count;
    NRRS_HIP(hipDeviceSynchronize());  // operations of every stream (the helper stream as well)
    for (uint32_t k = 0; k < n; ++k) {
        if (names) names[k] = ctx->op_name[k];
        if (ms) NRRS_HIP(hipEventElapsedTime(ms + k, ctx->op_ev[2 * k], ctx->op_ev[2 * k + 1]));
    }
    *count = n;
    ctx->op_count = 0;
    return NRRS_OK;
}

nrrs_status nrrs_frame_set_shape(nrrs_ctx* ctx, const nrrs_tile_shape* shape) {
    if (!ctx || !shape) return NRRS_ERR_ARG;
    const uint32_t w = shape->waves, m = shape->m;
    if ((w != 1 && w != 2 && w != 4 && w != 8) || (m != 64 && m != 128 && m != 256) || m % (32 * w)) return NRRS_ERR_ARG;
    ctx->frame_waves = w;
    ctx->frame_m = m;
    return NRRS_OK;
}

nrrs_status nrrs_frame_run(nrrs_ctx* ctx, const nrrs_frame_args* a, void* stream) {
    if (!ctx) return NRRS_ERR_ARG;
    FrameParams f;
    NRRS_TRY(frame_params(ctx, a, &f));
    return hip_status(launch_frame(ctx->cfg.precision, f, (int)ctx->frame_waves, 0, (hipStream_t)stream));
}

nrrs_status nrrs_frame_autotune(nrrs_ctx* ctx, const nrrs_frame_args* a, nrrs_tile_shape* best, float* best_ms,
                                void* stream) {
    if (!ctx) return NRRS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const nrrs_tile_shape shapes[] = {{1, 64}, {1, 128}, {1, 256}, {2, 64}, {2, 128}, {2, 256}, {4, 128}, {4, 256},
                                      {8, 256}};
    hipEvent_t e0, e1;
    NRRS_HIP(hipEventCreate(&e0));
    NRRS_HIP(hipEventCreate(&e1));
    float best_t = 1e30f;
    nrrs_tile_shape best_s = {ctx->frame_waves, ctx->frame_m};
    nrrs_status st = NRRS_OK;
    for (const nrrs_tile_shape& sh : shapes) {
        if ((st = nrrs_frame_set_shape(ctx, &sh)) != NRRS_OK) break;
        FrameParams f;
        if ((st = frame_params(ctx, a, &f)) != NRRS_OK) break;
        float t = 0.0f;
        for (int rep = 0; rep < 2 && st == NRRS_OK; ++rep) {  // first launch warms up
            st = hip_status(hipEventRecord(e0, s));
            if (st == NRRS_OK) st = hip_status(launch_frame(ctx->cfg.precision, f, (int)sh.waves, 0, s));
            if (st == NRRS_OK) st = hip_status(hipEventRecord(e1, s));
            if (st == NRRS_OK) st = hip_status(hipEventSynchronize(e1));
            if (st == NRRS_OK) st = hip_status(hipEventElapsedTime(&t, e0, e1));
        }
        if (st != NRRS_OK) break;
        uint32_t status = 0;
        if ((st = hip_status(hipMemcpy(&status, a->status_dev, 4, hipMemcpyDeviceToHost))) != NRRS_OK) break;
        if (status & 1u) {
            st = NRRS_ERR_HIP;
            break;
        }
        if (t < best_t) {
            best_t = t;
            best_s = sh;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    NRRS_TRY(st);
    NRRS_TRY(nrrs_frame_set_shape(ctx, &best_s));
    if (best) *best = best_s;
    if (best_ms) *best_ms = best_t;
    return NRRS_OK;
}

nrrs_status nrrs_frame_graph_capture(nrrs_ctx* ctx, const nrrs_frame_args* a, void* stream) {
    if (!ctx) return NRRS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (!s) return NRRS_ERR_ARG;  // the legacy default stream cannot be captured
    FrameParams f;
    NRRS_TRY(frame_params(ctx, a, &f));  // allocates outside the capture
    if (ctx->frame_exec) {
        hipGraphExecDestroy(ctx->frame_exec);
        ctx->frame_exec = nullptr;
    }
    if (ctx->frame_graph) {
        hipGraphDestroy(ctx->frame_graph);
        ctx->frame_graph = nullptr;
    }
    NRRS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_frame(ctx->cfg.precision, f, (int)ctx->frame_waves, 0, s);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return NRRS_ERR_HIP;
    }
    ctx->frame_graph = g;
    return hip_status(hipGraphInstantiate(&ctx->frame_exec, g, nullptr, nullptr, 0));
}

nrrs_status nrrs_frame_graph_launch(nrrs_ctx* ctx, void* stream) {
    if (!ctx || !ctx->frame_exec) return NRRS_ERR_ARG;
    return hip_status(hipGraphLaunch(ctx->frame_exec, (hipStream_t)stream));
}

nrrs_status nrrs_frame_vertices(uint64_t seed, uint32_t bounce, const uint32_t* queue_dev, uint32_t n,
                                const nrrs_vertices_out* o, void* stream) {
    if (!o || !queue_dev || !o->pos || !o->dir || !o->normal || !o->rough || !o->albedo_d || !o->albedo_s ||
        !o->thp || !o->pixel || o->stride < n)
        return NRRS_ERR_ARG;
    return hip_status(launch_bounce_vertices(seed, bounce, queue_dev, n, *o, (hipStream_t)stream));
}

nrrs_status nrrs_error_blend(nrrs_ctx* ctx, uint32_t n_pixels, const float* L_frame_dev, const uint32_t* spp_dev,
                             float blend, float* weighted_L_dev, void* stream) {
    if (!ctx || (n_pixels && (!L_frame_dev || !spp_dev || !weighted_L_dev))) return NRRS_ERR_ARG;
    return hip_status(launch_error_blend(n_pixels, L_frame_dev, spp_dev, blend, weighted_L_dev, (hipStream_t)stream));
}

nrrs_status nrrs_error_compute(nrrs_ctx* ctx, uint32_t n_pixels, const nrrs_error_args* a, void* stream) {
    if (!ctx || !a || (n_pixels && (!a->ref_dev || !a->weighted_L_dev || !a->error_dev))) return NRRS_ERR_ARG;
    return hip_status(launch_error_compute(n_pixels, a->ref_dev, a->weighted_L_dev, a->spp_dev, a->clamp, a->error_dev,
                                           ctx->err_part, a->error_sum_dev, (hipStream_t)stream));
}

nrrs_status nrrs_nodes_reset(nrrs_ctx* ctx, void* stream) {
    if (!ctx) return NRRS_ERR_ARG;
    if (!ctx->nodes.cap) return NRRS_ERR_SHAPE;
    return hip_status(launch_nodes_reset(ctx->nodes, (hipStream_t)stream));
}

nrrs_status nrrs_nodes_record(nrrs_ctx* ctx, uint32_t n, const uint32_t* n_dev, const nrrs_vertices* v,
                              const int32_t* parent_dev, const float* q_dev, uint32_t depth, int32_t* node_out_dev,
                              void* stream) {
    if (!ctx || !vertices_ok(v, n) || (n && !node_out_dev)) return NRRS_ERR_ARG;
    if (!ctx->nodes.cap) return NRRS_ERR_SHAPE;
    return hip_status(
        launch_nodes_record(ctx->nodes, n, n_dev, *v, parent_dev, q_dev, depth, node_out_dev, (hipStream_t)stream));
}

nrrs_status nrrs_nodes_add_radiance(nrrs_ctx* ctx, uint32_t n, const uint32_t* n_dev, const int32_t* node_dev,
                                    const float* L_dev, uint32_t L_stride, void* stream) {
    if (!ctx || (n && (!node_dev || !L_dev || L_stride < n))) return NRRS_ERR_ARG;
    if (!ctx->nodes.cap) return NRRS_ERR_SHAPE;
    return hip_status(launch_nodes_radiance(ctx->nodes, n, n_dev, node_dev, L_dev, L_stride, (hipStream_t)stream));
}

nrrs_status nrrs_nodes_size(nrrs_ctx* ctx, uint32_t* size, uint32_t* invalid) {
    if (!ctx) return NRRS_ERR_ARG;
    if (!ctx->nodes.cap) return NRRS_ERR_SHAPE;
    uint32_t h[2];
    NRRS_HIP(hipDeviceSynchronize());
    NRRS_HIP(hipMemcpy(h, ctx->nodes.cnt, 8, hipMemcpyDeviceToHost));
    if (size) *size = h[0];
    if (invalid) *invalid = h[1];
    return NRRS_OK;
}

nrrs_status nrrs_generate_training_data(nrrs_ctx* ctx, const nrrs_training_args* a, void* stream) {
    if (!ctx || !a) return NRRS_ERR_ARG;
    const nrrs_vertices_out& o = a->out;
    if (!a->film_dev || !a->target_L || !a->ref_mean || !a->error || !a->error_avg || !a->q_target ||
        !a->sample_weight || !a->pixel_samples || !a->n_out_dev || !o.pos || !o.dir || !o.normal || !o.rough ||
        !o.albedo_d || !o.albedo_s || !o.thp || !o.pixel || o.stride < a->capacity)
        return NRRS_ERR_ARG;
    if (a->error_dev && (!a->error_scale || !a->frame_width || a->frame_width % a->error_scale)) return NRRS_ERR_ARG;
    if (!ctx->nodes.cap) return NRRS_ERR_SHAPE;
    return hip_status(launch_training_data(ctx->nodes, *a, ctx->cfg.n_pixels, (hipStream_t)stream));
}

nrrs_status nrrs_consume_survivors(nrrs_ctx* ctx, uint32_t capacity, const uint32_t* n_survivors_dev,
                                   const uint32_t* survivors_dev, const nrrs_vertices* scatter, const float* rrs_dev,
                                   const int32_t* node_in_dev, const nrrs_vertices_out* out, int32_t* node_out_dev,
                                   uint32_t* stid_out_dev, void* stream) {
    if (!ctx || !out || !scatter || !vertices_ok(scatter, 0)) return NRRS_ERR_ARG;
    if (capacity && (!n_survivors_dev || !survivors_dev || !rrs_dev || !out->pos || !out->dir || !out->normal ||
                     !out->rough || !out->albedo_d || !out->albedo_s || !out->thp || !out->pixel ||
                     out->stride < capacity))
        return NRRS_ERR_ARG;
    return hip_status(launch_consume(capacity, n_survivors_dev, survivors_dev, *scatter, rrs_dev, node_in_dev, *out,
                                     node_out_dev, stid_out_dev, (hipStream_t)stream));
}

}  // extern "C"
