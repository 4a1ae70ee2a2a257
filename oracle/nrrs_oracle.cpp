// CPU oracle of the neural RRS stage - TEST INFRASTRUCTURE ONLY (the checker).
// FP64 restatement of the reference algorithms; see nrrs_oracle.h for what pins
// each function. Built with -ffp-contract=off so FP32 expressions round exactly
// as written.
#include "nrrs_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr uint64_t kPcgMult = 6364136223846793005ull;

struct Pcg32 {
    uint64_t state, inc;
    // pcg32_srandom_r: O'Neill, "PCG: A Family of Simple Fast Space-Efficient
    // Statistically Good Algorithms for Random Number Generation" (2014).
    Pcg32(uint64_t seed, uint64_t stream) {
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
    // LCG jump-ahead (F. Brown, "Random number generation with arbitrary
    // strides", 1994): state <- A^delta state + C(delta).
    void advance(uint64_t delta) {
        uint64_t acc_mult = 1, acc_plus = 0, cur_mult = kPcgMult, cur_plus = inc;
        while (delta) {
            if (delta & 1u) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta >>= 1u;
        }
        state = acc_mult * state + acc_plus;
    }
};

uint32_t pcg_at(uint64_t seed, uint64_t stream, uint64_t pos) {
    Pcg32 r(seed, stream);
    r.advance(pos);
    return r.next();
}

float u24(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

struct Level {
    uint32_t res, offset, size, dense;
    float scale;
};

std::vector<Level> levels_of(const nrrs_grid_config& g, uint64_t* n_entries) {
    std::vector<Level> L(g.n_levels);
    const uint64_t T = 1ull << g.log2_table_size;
    uint64_t off = 0;
    for (uint32_t l = 0; l < g.n_levels; ++l) {
        double s = (double)g.base_resolution * std::pow((double)g.per_level_scale, (double)l);
        L[l].scale = (float)s;
        L[l].res = (uint32_t)std::ceil((double)L[l].scale) + 2u;
        uint64_t r3 = (uint64_t)L[l].res * L[l].res * L[l].res;
        L[l].dense = r3 <= T ? 1u : 0u;
        L[l].size = (uint32_t)(L[l].dense ? r3 : T);
        L[l].offset = (uint32_t)off;
        off += (L[l].size + 7ull) / 8ull * 8ull;
    }
    if (n_entries) *n_entries = off;
    return L;
}

// normalizeSpatialCoord (src/render/nrc/nrctrain.h:111-115): the AABB is
// inflated by 0.5% of its diagonal length, then p -> 0.5 + (p - c) / d.
void aabb_frame(const float* mn, const float* mx, float* c, float* d) {
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    float len = std::sqrt(dx * dx + dy * dy + dz * dz);
    float infl = len * 0.005f;
    for (int k = 0; k < 3; ++k) {
        float lo = mn[k] - infl, hi = mx[k] + infl;
        c[k] = (lo + hi) * 0.5f;
        d[k] = hi - lo;
    }
}

// fmin/fmax semantics: a NaN coordinate maps to 0.
float norm_coord(float p, float c, float d) {
    float x = 0.5f + (p - c) / d;
    return std::fmin(std::fmax(x, 0.0f), 1.0f);
}

uint32_t corner_index(const Level& L, uint32_t log2T, uint32_t cx, uint32_t cy, uint32_t cz) {
    if (L.dense) return cx + L.res * (cy + L.res * cz);
    // instant-NGP spatial hash (Mueller et al. 2022, eq. 4)
    uint32_t h = cx ^ (cy * 2654435761u) ^ (cz * 805459861u);
    return h & ((1u << log2T) - 1u);
}

double sigma_d(double x) { return x < 0 ? std::log1p(std::exp(x)) : 0.5 * x + 0.6931471805599453; }
double sigma_grad_d(double x) { return x < 0 ? std::exp(x) / (1.0 + std::exp(x)) : 0.5; }

float sigma_f(float x) {
    // activationSigma (src/render/nrrs/net/networkcommon.h:61-68) in FP32
    return (x < 0) ? std::log(1.0f + std::exp(x)) : (0.5f * x + 0.6931471805599453f);
}

double bc_d(double x) { return 2.0 * (std::sqrt(x) - 1.0); }


// The shipped RRSNet without the AID input has no grid: network.cpp:503-507 builds it with
// create_network (no encoding) on the 16 channels of NRRSNetworkInput1 (train.h:39-56).
bool net_has_grid(const nrrs_config* cfg, int net) {
    return !(net == NRRS_RRSNET && cfg->net_set == NRRS_NETS_SHIPPED && !cfg->rrs_aid);
}

// Hidden activation: ReLU, or tiny-cuda-nn's LeakyReLU (slope 0.01) for the shipped nets
// (2net-simple.json:72, 240).
double act_d(const nrrs_config* cfg, double x) {
    if (cfg->net_set == NRRS_NETS_SHIPPED) return x > 0.0 ? x : 0.01 * x;
    return x > 0.0 ? x : 0.0;
}
// Its slope, read from the activation's output (which has the sign of its input).
double act_slope_d(const nrrs_config* cfg, double h) {
    if (h > 0.0) return 1.0;
    return cfg->net_set == NRRS_NETS_SHIPPED ? 0.01 : 0.0;
}

// worldToLatLong (KiRaRay; called at src/render/nrrs/device.cu:102): azimuth about +y measured
// from -z, u = atan2(x, -z) / 2pi + 1/2, and polar angle from +y, v = acos(y) / pi.
void lat_long_d(double x, double y, double z, double* u, double* v) {
    const double r = std::sqrt(x * x + y * y + z * z);
    x /= r, y /= r, z /= r;
    *u = std::atan2(x, -z) / (2.0 * M_PI) + 0.5;
    *v = std::acos(std::min(1.0, std::max(-1.0, y))) / M_PI;
}

// OneBlob (tiny-cuda-nn "OneBlob", Mueller et al. 2019): bin k of n holds the mass that a
// biweight kernel 15/16 (1 - t^2)^2, t = (s - x) n (radius one bin), wrapped periodically onto
// [0, 1), puts on [k/n, (k+1)/n). F is the kernel's CDF in t, clamped outside [-1, 1].
double biweight_cdf(double t) {
    t = std::min(1.0, std::max(-1.0, t));
    return 0.5 + 15.0 / 16.0 * (t - 2.0 * t * t * t / 3.0 + t * t * t * t * t / 5.0);
}
void one_blob_d(double x, int n, double* out) {
    for (int k = 0; k < n; ++k) {
        double m = 0;
        for (int w = -1; w <= 1; ++w) {
            const double c = x + w;  // a periodic copy of the centre
            m += biweight_cdf(((double)(k + 1) / n - c) * n) - biweight_cdf(((double)k / n - c) * n);
        }
        out[k] = m;
    }
}

void mlp_dims(const nrrs_config* cfg, int net, int* n_layers, int* n_out) {
    const nrrs_mlp_config& m = net == NRRS_RRSNET ? cfg->rrsnet : cfg->statnet;
    *n_layers = (int)m.n_hidden_layers + 1;
    *n_out = (int)m.n_outputs;
}

// Which inputs each neuron reads, per layer ([layer][row * in + col], 1 = connected).
// Two-headed StatNets restated (twn_PartShare.cpp:93-169, twn_JustEncoding.cpp; the heads are
// 32 wide here): head 1 predicts L (outputs 0-2), head 2 L^2 (outputs 3-5). PartShare: a base
// net (all layers but the last two) on the encoding, then per head a tail of two layers on the
// base net's output; JustEncoding: per head a whole net on the encoding. A head's first layer
// reads the shared input whole; its later layers read only the same head's neurons
// (neurons 0-31 of a hidden layer are head 1, 32-63 head 2). AllSelf (twn_AllSelf.cpp): two
// whole nets with their own encodings, net NRRS_STATNET for L and net 2 for L^2; each is the
// one-head StatNet whose output rows of the other head are not connected.
std::vector<std::vector<char>> connections(const nrrs_config* cfg, int net) {
    int nl, no;
    mlp_dims(cfg, net, &nl, &no);
    std::vector<std::vector<char>> conn(nl);
    const bool all_self = net != NRRS_RRSNET && cfg->ll2_two_head == 3;
    const bool two = net == NRRS_STATNET && (cfg->ll2_two_head == 1 || cfg->ll2_two_head == 2);
    const int head_start = cfg->ll2_two_head == 1 ? nl - 2 : 0;  // first layer of the heads
    int in = ORA_ENC_WIDTH;
    for (int layer = 0; layer < nl; ++layer) {
        const int o = layer == nl - 1 ? no : ORA_HIDDEN;
        conn[layer].assign((size_t)o * in, 1);
        if (two && layer > head_start)
            for (int r = 0; r < o; ++r) {
                const int head = layer == nl - 1 ? r / 3 : r / 32;
                for (int k = 0; k < in; ++k) conn[layer][(size_t)r * in + k] = (k / 32 == head);
            }
        if (all_self && layer == nl - 1)
            for (int r = 0; r < o; ++r)
                if (r / 3 != (net == NRRS_STATNET ? 0 : 1))
                    for (int k = 0; k < in; ++k) conn[layer][(size_t)r * in + k] = 0;
        if (cfg->net_set == NRRS_NETS_SHIPPED) {
            // 32-wide hidden layers; layer 0 reads the channels of the net's shipped input
            // (the encoding layout of encode_impl)
            for (int r = 0; r < o; ++r)
                for (int k = 0; k < in; ++k) {
                    bool live = (layer == nl - 1 || r < 32);
                    if (layer > 0) {
                        live = live && k < 32;
                    } else {
                        const int gch = net_has_grid(cfg, net) ? 2 * (int)cfg->grid.n_levels : 0;
                        bool in_ch;
                        if (k < gch)
                            in_ch = true;
                        else if (net != NRRS_RRSNET)
                            in_ch = k >= 32 && k < 48;  // OneBlob(u, v) 8 + OneBlob(rough) 8
                        else if (!cfg->rrs_aid)
                            in_ch = k >= 64;  // NRRSNetworkInput1, 16 channels
                        else
                            in_ch = (k >= 32 && k < 44) || (k >= 65 && k < 69);
                        live = live && in_ch;
                    }
                    if (!live) conn[layer][(size_t)r * in + k] = 0;
                }
        }
        in = o;
    }
    return conn;
}

}  // namespace

extern "C" {

void ora_pcg32(uint64_t seed, uint64_t stream, uint64_t skip, uint32_t n, uint32_t* out) {
    Pcg32 r(seed, stream);
    r.advance(skip);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next();
}

uint64_t ora_grid_layout(const nrrs_grid_config* g, uint32_t* res, uint32_t* offset,
                         uint32_t* size, uint32_t* dense, float* scale) {
    uint64_t n = 0;
    std::vector<Level> L = levels_of(*g, &n);
    for (uint32_t l = 0; l < g->n_levels; ++l) {
        if (res) res[l] = L[l].res;
        if (offset) offset[l] = L[l].offset;
        if (size) size[l] = L[l].size;
        if (dense) dense[l] = L[l].dense;
        if (scale) scale[l] = L[l].scale;
    }
    return n * g->n_features;
}

void ora_normalize_pos(const float* aabb_min, const float* aabb_max, uint32_t n,
                       const float* pos, uint32_t stride, float* xn) {
    float c[3], d[3];
    aabb_frame(aabb_min, aabb_max, c, d);
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) xn[i * 3 + k] = norm_coord(pos[k * stride + i], c[k], d[k]);
}

void ora_grid_indices(const nrrs_config* cfg, uint32_t n, const nrrs_vertices* v, uint32_t* idx) {
    const nrrs_grid_config& g = cfg->grid;
    std::vector<Level> L = levels_of(g, nullptr);
    float c[3], d[3];
    aabb_frame(cfg->aabb_min, cfg->aabb_max, c, d);
    for (uint32_t i = 0; i < n; ++i) {
        float x[3];
        for (int k = 0; k < 3; ++k) x[k] = norm_coord(v->pos[k * v->stride + i], c[k], d[k]);
        for (uint32_t l = 0; l < g.n_levels; ++l) {
            uint32_t cell[3];
            for (int k = 0; k < 3; ++k) {
                float p = std::fma(x[k], L[l].scale, 0.5f);
                cell[k] = (uint32_t)std::floor(p);
            }
            for (uint32_t cn = 0; cn < 8; ++cn)
                idx[((size_t)i * g.n_levels + l) * 8 + cn] =
                    corner_index(L[l], g.log2_table_size, cell[0] + (cn & 1u),
                                 cell[1] + ((cn >> 1) & 1u), cell[2] + ((cn >> 2) & 1u));
        }
    }
}

double ora_activation_sigma(double x) { return sigma_d(x); }
void ora_lat_long(double x, double y, double z, double* uv) { lat_long_d(x, y, z, uv, uv + 1); }
void ora_one_blob(double x, int32_t bins, double* out) { one_blob_d(x, bins, out); }
double ora_box_cox(double x) { return bc_d(x); }

double ora_weight_window(double q, double size) {
    // RRSWeightWindow (src/render/common/commoncudautils.h:14-28)
    double dminus = 2.0 / (1.0 + size), dplus = dminus * size;
    if (q < dminus) return q / dminus;
    if (q > dplus) return q / dplus;
    return 1.0;
}

void ora_sh4(double x, double y, double z, double* o) {
    // Real spherical harmonics, bands 0..3 (Sloan, "Stupid Spherical Harmonics Tricks", 2008)
    double xx = x * x, yy = y * y, zz = z * z;
    o[0] = 0.28209479177387814;
    o[1] = -0.48860251190291987 * y;
    o[2] = 0.48860251190291987 * z;
    o[3] = -0.48860251190291987 * x;
    o[4] = 1.0925484305920792 * x * y;
    o[5] = -1.0925484305920792 * y * z;
    o[6] = 0.94617469575755997 * zz - 0.31539156525251999;
    o[7] = -1.0925484305920792 * x * z;
    o[8] = 0.54627421529603959 * (xx - yy);
    o[9] = 0.59004358992664352 * y * (-3.0 * xx + yy);
    o[10] = 2.8906114426405538 * x * y * z;
    o[11] = 0.45704579946446572 * y * (1.0 - 5.0 * zz);
    o[12] = 0.3731763325901154 * z * (5.0 * zz - 3.0);
    o[13] = 0.45704579946446572 * x * (1.0 - 5.0 * zz);
    o[14] = 1.4453057213202769 * z * (xx - yy);
    o[15] = 0.59004358992664352 * x * (-xx + 3.0 * yy);
}

uint64_t ora_param_count(const nrrs_config* cfg, int32_t net, uint64_t* n_grid, uint64_t* n_mlp) {
    uint64_t ng = net_has_grid(cfg, net) ? ora_grid_layout(&cfg->grid, nullptr, nullptr, nullptr, nullptr, nullptr) : 0;
    int nl, no;
    mlp_dims(cfg, net, &nl, &no);
    uint64_t nm = (uint64_t)ORA_HIDDEN * ORA_ENC_WIDTH +
                  (uint64_t)(nl - 2) * ORA_HIDDEN * ORA_HIDDEN + (uint64_t)no * ORA_HIDDEN;
    if (n_grid) *n_grid = ng;
    if (n_mlp) *n_mlp = nm;
    return ng + nm;
}

}  // extern "C"

namespace {

// Encoding of n vertices [n][80]. `training`: the RRSNet input of a training step, which with
// train_sigma takes activationSigma(max(L2, 0)) (nrrs_generate_training_data_from_0to1,
// src/render/nrrs/train.cu:196-200) where the inference input takes max(L2, 0)
// (nrrs_generate_inference_data_from_0to1, train.cu:41-112).
void encode_impl(const nrrs_config* cfg, int32_t net, uint32_t n, const nrrs_vertices* v, const float* film,
                 const double* stats, const double* params, double* enc, bool training) {
    const nrrs_grid_config& g = cfg->grid;
    std::vector<Level> L = levels_of(g, nullptr);
    float c[3], d[3];
    aabb_frame(cfg->aabb_min, cfg->aabb_max, c, d);
    const uint32_t S = v->stride;
    for (uint32_t i = 0; i < n; ++i) {
        double* e = enc + (size_t)i * ORA_ENC_WIDTH;
        std::fill(e, e + ORA_ENC_WIDTH, 0.0);
        float x[3];
        for (int k = 0; k < 3; ++k) x[k] = norm_coord(v->pos[k * S + i], c[k], d[k]);
        const uint32_t n_levels = net_has_grid(cfg, net) ? g.n_levels : 0;
        for (uint32_t l = 0; l < n_levels; ++l) {
            uint32_t cell[3];
            double fr[3];
            for (int k = 0; k < 3; ++k) {
                float p = std::fma(x[k], L[l].scale, 0.5f);
                float fl = std::floor(p);
                cell[k] = (uint32_t)fl;
                fr[k] = (double)(p - fl);
            }
            double f0 = 0, f1 = 0;
            for (uint32_t cn = 0; cn < 8; ++cn) {
                double w = 1.0;
                for (int k = 0; k < 3; ++k) w *= ((cn >> k) & 1u) ? fr[k] : 1.0 - fr[k];
                uint32_t id = corner_index(L[l], g.log2_table_size, cell[0] + (cn & 1u),
                                           cell[1] + ((cn >> 1) & 1u), cell[2] + ((cn >> 2) & 1u));
                const double* t = params + ((size_t)L[l].offset + id) * 2;
                f0 += w * t[0];
                f1 += w * t[1];
            }
            e[2 * l] = f0;
            e[2 * l + 1] = f1;
        }
        // AID RRSNet input (NRRSNetworkInput1AID, src/render/nrrs/train.h:58-68; built by
        // nrrs_generate_inference_data_aid, train.cu:242-267, and for training by
        // nrrs_generate_training_data_from_0to1_aid, train.cu:114-153): position, direction, roughness,
        // reference mean and throughput only; the normal, albedo and statistics channels stay zero
        const bool aid = net == NRRS_RRSNET && cfg->rrs_aid != 0;
        if (cfg->net_set == NRRS_NETS_SHIPPED) {
            // shipped inputs (2net-simple.json:29-68, 210-291)
            const double wr = 1.0 - std::exp(-(double)v->rough[i]);  // warp_roughness, train.h:95
            uint32_t px = v->pixel[i];
            if (net == NRRS_RRSNET && !aid) {
                // NRRSNetworkInput1 (train.h:39-56): BC(max(L,0)), BC(max(L2,0)) (with train_sigma
                // the training input takes activationSigma of it, train.cu:196-200), BC(thp),
                // roughness, BC(ref mean), then the padding {1, 0, 0, 0, 0}
                const double ref = ((double)film[px] + (double)film[cfg->n_pixels + px] +
                                    (double)film[2 * (size_t)cfg->n_pixels + px]) / 3.0;
                for (int k = 0; k < 3; ++k) {
                    const double L = stats ? std::max(stats[(size_t)k * n + i], 0.0) : 0.0;
                    const double l2 = stats ? std::max(stats[(size_t)(3 + k) * n + i], 0.0) : 0.0;
                    e[69 + k] = bc_d(L);
                    e[72 + k] = bc_d(cfg->train_sigma && training ? sigma_d(l2) : l2);
                    e[65 + k] = bc_d((double)v->thp[k * S + i]);
                }
                e[64] = wr;
                e[68] = bc_d(ref);
                e[75] = 1.0;
                continue;
            }
            // NRRSNetworkInput0 (train.h:22-29) / NRRSNetworkInput1AID (train.h:58-68): the
            // lat-long direction through OneBlob (4 bins per coordinate), the roughness through
            // OneBlob (StatNet 8 bins, AID 4), the AID input's Box-Cox reference mean and thp
            double uu, vv;
            lat_long_d(v->dir[i], v->dir[S + i], v->dir[2 * S + i], &uu, &vv);
            one_blob_d(uu, 4, e + 32);
            one_blob_d(vv, 4, e + 36);
            one_blob_d(wr, aid ? 4 : 8, e + 40);
            if (aid) {
                const double ref = ((double)film[px] + (double)film[cfg->n_pixels + px] +
                                    (double)film[2 * (size_t)cfg->n_pixels + px]) / 3.0;
                for (int k = 0; k < 3; ++k) e[65 + k] = bc_d((double)v->thp[k * S + i]);
                e[68] = bc_d(ref);
            }
            continue;
        }
        ora_sh4(v->dir[i], v->dir[S + i], v->dir[2 * S + i], e + 32);
        if (!aid) ora_sh4(v->normal[i], v->normal[S + i], v->normal[2 * S + i], e + 48);
        // warp_roughness (src/render/nrrs/train.h:95)
        e[64] = 1.0 - std::exp(-(double)v->rough[i]);
        if (net != NRRS_RRSNET) {  // StatNet (either net of AllSelf)
            for (int k = 0; k < 3; ++k) {
                e[65 + k] = v->albedo_d[k * S + i];
                e[68 + k] = v->albedo_s[k * S + i];
            }
            e[71] = 1.0;
        } else {
            // RRSNet input: Box-Cox of thp, ref mean, StatNet stats
            // (nrrs_generate_inference_data_from_0to1, src/render/nrrs/train.cu:41-112)
            uint32_t px = v->pixel[i];
            double ref = ((double)film[px] + (double)film[cfg->n_pixels + px] +
                          (double)film[2 * (size_t)cfg->n_pixels + px]) / 3.0;
            for (int k = 0; k < 3; ++k) e[65 + k] = bc_d((double)v->thp[k * S + i]);
            e[68] = bc_d(ref);
            for (int k = 0; k < 3 && !aid; ++k) {
                if (stats) {
                    e[69 + k] = bc_d(std::max(stats[(size_t)k * n + i], 0.0));
                    double l2 = std::max(stats[(size_t)(3 + k) * n + i], 0.0);
                    if (training && cfg->train_sigma) l2 = sigma_d(l2);
                    e[72 + k] = bc_d(l2);
                } else {
                    e[69 + k] = v->albedo_d[k * S + i];
                    e[72 + k] = v->albedo_s[k * S + i];
                }
            }
            e[75] = 1.0;
        }
    }
}

}  // namespace

extern "C" {

void ora_encode(const nrrs_config* cfg, int32_t net, uint32_t n, const nrrs_vertices* v,
                const float* film, const double* stats, const double* params, double* enc) {
    encode_impl(cfg, net, n, v, film, stats, params, enc, false);
}

void ora_encode_train(const nrrs_config* cfg, int32_t net, uint32_t n, const nrrs_vertices* v,
                      const float* film, const double* stats, const double* params, double* enc) {
    encode_impl(cfg, net, n, v, film, stats, params, enc, true);
}

void ora_net_forward(const nrrs_config* cfg, int32_t net, uint32_t n, const nrrs_vertices* v,
                     const float* film, const double* stats, const double* params, double* out) {
    uint64_t ng = 0;
    ora_param_count(cfg, net, &ng, nullptr);
    int nl, no;
    mlp_dims(cfg, net, &nl, &no);
    std::vector<double> enc((size_t)n * ORA_ENC_WIDTH);
    ora_encode(cfg, net, n, v, film, stats, params, enc.data());
    const double* W = params + ng;
    const std::vector<std::vector<char>> conn = connections(cfg, net);
    std::vector<double> h(ORA_ENC_WIDTH), t(ORA_HIDDEN);
    for (uint32_t i = 0; i < n; ++i) {
        int in = ORA_ENC_WIDTH;
        h.assign(enc.begin() + (size_t)i * ORA_ENC_WIDTH, enc.begin() + (size_t)(i + 1) * ORA_ENC_WIDTH);
        const double* w = W;
        for (int layer = 0; layer < nl; ++layer) {
            int o = layer == nl - 1 ? no : ORA_HIDDEN;
            t.assign(o, 0.0);
            for (int r = 0; r < o; ++r) {
                double acc = 0;
                for (int k = 0; k < in; ++k)
                    if (conn[layer][(size_t)r * in + k]) acc += w[(size_t)r * in + k] * h[k];
                t[r] = layer == nl - 1 ? acc : act_d(cfg, acc);
            }
            w += (size_t)o * in;
            h.assign(t.begin(), t.begin() + o);
            in = o;
        }
        for (int r = 0; r < no; ++r) out[(size_t)i * no + r] = h[r];
    }
}

void ora_q(const nrrs_config* cfg, int32_t mode, uint32_t n, const double* src,
           const nrrs_vertices* v, const float* film, double* q) {
    const uint32_t S = v->stride, P = cfg->n_pixels;
    for (uint32_t i = 0; i < n; ++i) {
        double r;
        if (mode == NRRS_Q_STATNET_ADRRS) {
            // integrator.cpp:1246-1255
            uint32_t px = v->pixel[i];
            r = 0;
            for (int k = 0; k < 3; ++k) {
                double L = std::fmin(std::fmax(src[(size_t)k * n + i], 1e-4), 1e4);
                r += L * (double)v->thp[k * S + i] / ((double)film[(size_t)k * P + px] + 0.01);
            }
            r /= 3.0;
        } else {
            r = sigma_d(src[i]);  // integrator.cpp:1203-1219
        }
        if (cfg->rrs_weight_window) r = ora_weight_window(r, cfg->weight_window_size);
        // NaN q becomes the clamp minimum (fmin/fmax semantics)
        q[i] = std::fmin(std::fmax(r, (double)cfg->rrs_clamp_min), (double)cfg->rrs_clamp_max);
    }
}

void ora_q_f32(const nrrs_config* cfg, int32_t mode, uint32_t n, const float* src,
               const nrrs_vertices* v, const float* film, float* q) {
    const uint32_t S = v->stride, P = cfg->n_pixels;
    for (uint32_t i = 0; i < n; ++i) {
        float r;
        if (mode == NRRS_Q_STATNET_ADRRS) {
            uint32_t px = v->pixel[i];
            float acc = 0.f;
            for (int k = 0; k < 3; ++k) {
                float L = std::fmin(std::fmax(src[(size_t)k * n + i], 1e-4f), 1e4f);
                acc += L * v->thp[k * S + i] / (film[(size_t)k * P + px] + 0.01f);
            }
            r = acc / 3.0f;
        } else {
            r = sigma_f(src[i]);
        }
        if (cfg->rrs_weight_window) {
            float dm = 2.0f / (1.0f + cfg->weight_window_size), dp = dm * cfg->weight_window_size;
            r = r < dm ? r / dm : (r > dp ? r / dp : 1.0f);
        }
        q[i] = std::fmin(std::fmax(r, cfg->rrs_clamp_min), cfg->rrs_clamp_max);
    }
}

int32_t ora_rrs_decide(const nrrs_config* cfg, uint32_t n, const float* q, const uint32_t* stid,
                       uint32_t n_spec, const uint32_t* spec_stid, uint64_t seed, uint64_t stream,
                       float* qn, uint8_t* s, uint32_t* survivors, uint32_t capacity,
                       uint32_t* n_survivors, uint32_t* overflow, double* sum_out) {
    // Global sum of q (or ceil q) - calcSum2PassAsync (commoncudautilshost.cpp:6-22) -
    // restated as an exact fixed-point sum with 2^-32 resolution, so the result is
    // independent of summation order.
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        float t = cfg->rrs_use_ceil ? std::ceil(q[i]) : q[i];
        acc += (uint64_t)((double)t * 4294967296.0);
    }
    double sum = (double)acc * 0x1p-32;
    if (sum_out) *sum_out = sum;
    float sum_f = (float)sum;
    float rcp = sum_f > 0.0f ? 1.0f / sum_f : 0.0f;
    // normalizeFactor (integrator.cpp:1376-1377)
    float nf = (float)(cfg->n_pixels - n_spec) * rcp;
    nf = nf * (cfg->rrs_use_ceil ? 1.0f : cfg->rrs_normalize_rate);
    uint32_t written = 0;
    uint64_t total = 0;
    for (uint32_t k = 0; k < n_spec; ++k) {
        if (written < capacity) survivors[written++] = spec_stid[k];
        ++total;
    }
    Pcg32 rng(seed, stream);
    for (uint32_t i = 0; i < n; ++i) {
        // integrator.cpp:1391-1397
        float r = q[i];
        r = cfg->rrs_no_upscale ? r * (nf < 1.0f ? nf : 1.0f) : r * nf;
        if (qn) qn[i] = r;
        float u = u24(rng.next());  // vertex i draws output number i of the stream
        // integrator.cpp:1415-1416, 1454: s = floor(q') + [u <= frac(q')]
        int si = (int)r;
        float left = r - (float)si;
        si += (u <= left) ? 1 : 0;
        if ((uint32_t)si > cfg->split_cap) si = (int)cfg->split_cap;
        if (s) s[i] = (uint8_t)si;
        uint32_t id = stid ? stid[i] : i;
        for (int c = 0; c < si; ++c) {
            if (written < capacity) survivors[written++] = id;
            ++total;
        }
    }
    if (n_survivors) *n_survivors = written;
    uint32_t dropped = (uint32_t)(total - written);
    if (overflow) *overflow = dropped;
    return dropped ? NRRS_ERR_OVERFLOW : 0;
}

void ora_synth_vertices(uint32_t n, uint64_t seed, uint32_t n_pixels, const nrrs_vertices_out* o) {
    // 64 triangles, corners uniform in [0,1]^3 (stream 0, outputs 0..575)
    float tri[64 * 9];
    {
        Pcg32 r(seed, 0);
        for (int k = 0; k < 64 * 9; ++k) tri[k] = u24(r.next());
    }
    const uint32_t S = o->stride;
    const float two_pi = 6.283185307179586f;
    for (uint32_t i = 0; i < n; ++i) {
        // position: stream 1, 3 outputs per vertex
        uint32_t a = pcg_at(seed, 1, 3ull * i), b = pcg_at(seed, 1, 3ull * i + 1),
                 cc = pcg_at(seed, 1, 3ull * i + 2);
        const float* T = tri + (a & 63u) * 9;
        float su = std::sqrt(u24(b));
        float b0 = 1.0f - su, b1 = u24(cc) * su, b2 = 1.0f - b0 - b1;
        for (int k = 0; k < 3; ++k) o->pos[k * S + i] = b0 * T[k] + b1 * T[3 + k] + b2 * T[6 + k];
        // direction (stream 2) and normal (stream 3): uniform on the sphere
        for (int f = 0; f < 2; ++f) {
            float* dst = f ? o->normal : o->dir;
            float u1 = u24(pcg_at(seed, 2 + f, 2ull * i)), u2 = u24(pcg_at(seed, 2 + f, 2ull * i + 1));
            float z = 1.0f - 2.0f * u1;
            float rr = std::sqrt(std::max(0.0f, 1.0f - z * z));
            float ph = two_pi * u2;
            dst[i] = rr * std::cos(ph);
            dst[S + i] = rr * std::sin(ph);
            dst[2 * S + i] = z;
        }
        o->rough[i] = 0.05f + 0.95f * u24(pcg_at(seed, 4, i));
        for (int k = 0; k < 3; ++k) {
            o->albedo_d[k * S + i] = u24(pcg_at(seed, 5, 3ull * i + k));
            o->albedo_s[k * S + i] = u24(pcg_at(seed, 6, 3ull * i + k));
            // log-normal(mu=-1, sigma=1): Box-Muller, stream 7, 2 outputs per channel
            float ua = u24(pcg_at(seed, 7, 6ull * i + 2 * k)), ub = u24(pcg_at(seed, 7, 6ull * i + 2 * k + 1));
            float nrm = std::sqrt(-2.0f * std::log(1.0f - ua)) * std::cos(two_pi * ub);
            o->thp[k * S + i] = std::exp(-1.0f + nrm);
        }
        o->pixel[i] = i % n_pixels;
    }
}

void ora_init_params(const nrrs_config* cfg, int32_t net, uint64_t seed, double grid_amp,
                     double* params) {
    uint64_t ng = 0, nm = 0;
    ora_param_count(cfg, net, &ng, &nm);
    Pcg32 r(seed, 100 + (uint64_t)net);
    for (uint64_t k = 0; k < ng; ++k) params[k] = grid_amp * (2.0 * u24(r.next()) - 1.0);
    int nl, no;
    mlp_dims(cfg, net, &nl, &no);
    double* w = params + ng;
    int in = ORA_ENC_WIDTH;
    const std::vector<std::vector<char>> conn = connections(cfg, net);
    for (int layer = 0; layer < nl; ++layer) {
        int o = layer == nl - 1 ? no : ORA_HIDDEN;
        double lim = std::sqrt(6.0 / (double)(in + o));
        for (int k = 0; k < o * in; ++k) w[k] = conn[layer][k] ? lim * (2.0 * u24(r.next()) - 1.0) : 0.0;
        w += (size_t)o * in;
        in = o;
    }
}

}  // extern "C"

// ---------------------------------------------------------------- training
namespace {

// NRRS_LL2 restated (the loss source, nrrs_ll2.h, is not in the reference checkout):
// relative L2 of the 3 (step 0) or 6 (step >= 1) outputs against t = (L, L^2), L
// clamped to loss_clamp_max when loss_clamp_on; the denominator y^2 + 0.01 is a
// constant of the gradient (relative-L2 convention). Loss = mean over samples and
// active channels. With train_sigma, output 3 + c is a sigma: its prediction of L^2 is
// m = activationSigma(y_{3+c})^2 + y_c^2 (as the integrator reads the outputs,
// integrator.cpp:655-670), scored as (m - L^2)^2 / (m^2 + 0.01) with y_c and the
// denominator constants of the gradient. y_den supplies those constants (null: y itself).
// head >= 0: the terms of outputs 3 head .. 3 head + 2 only (the loss of one AllSelf net).
double loss_ll2(const nrrs_config* cfg, uint32_t n, const double* y, const float* tL, const double* y_den,
                double* dy, int head = -1) {
    const int active = cfg->ll2_step == 0 ? 3 : 6;
    const double norm = 1.0 / ((double)n * active);
    double loss = 0;
    for (uint32_t i = 0; i < n; ++i) {
        for (int c = 0; c < 6; ++c) {
            double t = std::fmax((double)tL[(size_t)(c % 3) * n + i], 0.0);
            if (cfg->loss_clamp_on) t = std::fmin(t, (double)cfg->loss_clamp_max);
            if (c >= 3) t = t * t;
            double yy = y[(size_t)c * n + i];
            const double* yc = y_den ? y_den : y;  // constants of the gradient
            double yd = yc[(size_t)c * n + i];
            double den = yd * yd + 0.01;
            double d = yy - t;
            double slope = 1.0;  // d(prediction)/dy
            if (c >= 3 && cfg->train_sigma) {
                double mu = yc[(size_t)(c - 3) * n + i];
                double sg = sigma_d(yy), sgd = sigma_d(yd);
                double m = sg * sg + mu * mu, md = sgd * sgd + mu * mu;
                den = md * md + 0.01;
                d = m - t;
                slope = 2.0 * sg * sigma_grad_d(yy);
            }
            if (c < active && (head < 0 || c / 3 == head)) {
                loss += d * d / den * norm;
                if (dy) dy[(size_t)c * n + i] = 2.0 * d / den * slope * norm;
            } else if (dy) {
                dy[(size_t)c * n + i] = 0.0;
            }
        }
    }
    return loss;
}

// NRRS_RRS restated (nrrs_rrs.h is not in the reference checkout). With q = sigma(y),
// v = mean_c(thp_c^2 max(L2_c,0)) / (ref^2 + 0.01) (clamped at loss_clamp_max when
// loss_clamp_on; with train_sigma the second moment is activationSigma(s_{3+c})^2 +
// max(s_c, 0)^2 of the StatNet outputs s), E = max(error_sum / pixels, 1e-4),
// e = max(error, 1e-4):
//   f(q) = g1 (v/(q+0.01) + q E) + g2 (v/(q+0.01) + q e) + g3 q + g4 (q - q_t)^2
// loss = (1/n) sum over samples with q_t >= 0 of sample_weight * f(q); zero for step < 3.
double loss_rrs(const nrrs_config* cfg, uint32_t n, const double* y, const nrrs_vertices* v,
                const float* film, const double* stats, const float* ref_mean, const float* err,
                const float* sw, const float* qt, double error_sum, double* dy) {
    double loss = 0;
    const double Ebar = std::fmax(error_sum / (double)cfg->n_pixels, 1e-4);
    const double* g = nullptr;
    double gam[4] = {cfg->gamma[0], cfg->gamma[1], cfg->gamma[2], cfg->gamma[3]};
    (void)g;
    const uint32_t S = v->stride, P = cfg->n_pixels;
    for (uint32_t i = 0; i < n; ++i) {
        if (dy) dy[i] = 0.0;
        if (cfg->rrs_step < 3 || qt[i] < 0.0f) continue;
        double ref;
        if (ref_mean) {
            ref = ref_mean[i];
        } else {
            uint32_t px = v->pixel[i];
            ref = ((double)film[px] + (double)film[P + px] + (double)film[2 * (size_t)P + px]) / 3.0;
        }
        double m2 = 0;
        for (int c = 0; c < 3; ++c) {
            double t = v->thp[c * S + i];
            double l2 = std::fmax(stats[(size_t)(3 + c) * n + i], 0.0);
            if (cfg->train_sigma) {
                double sg = sigma_d(stats[(size_t)(3 + c) * n + i]), mu = std::fmax(stats[(size_t)c * n + i], 0.0);
                l2 = sg * sg + mu * mu;
            }
            m2 += t * t * l2;
        }
        double vv = m2 / 3.0 / (ref * ref + 0.01);
        if (cfg->loss_clamp_on) vv = std::fmin(vv, (double)cfg->loss_clamp_max);
        double e = std::fmax((double)err[i], 1e-4);
        double q = sigma_d(y[i]);
        double qe = q + 0.01;
        double f = gam[0] * (vv / qe + q * Ebar) + gam[1] * (vv / qe + q * e) + gam[2] * q +
                   gam[3] * (q - qt[i]) * (q - qt[i]);
        double fp = -(gam[0] + gam[1]) * vv / (qe * qe) + gam[0] * Ebar + gam[1] * e + gam[2] +
                    2.0 * gam[3] * (q - qt[i]);
        loss += sw[i] * f / n;
        if (dy) dy[i] = sw[i] * fp * sigma_grad_d(y[i]) / n;
    }
    return loss;
}

}  // namespace

extern "C" {

double ora_loss(const nrrs_config* cfg, int32_t kind, uint32_t n, const double* y, const float* target_L,
                const nrrs_vertices* v, const float* film, const double* stats, const float* ref_mean,
                const float* err, const float* sw, const float* qt, double error_sum, const double* y_den,
                double* dy) {
    if (kind == NRRS_LOSS_LL2) return loss_ll2(cfg, n, y, target_L, y_den, dy);
    if (kind == NRRS_LOSS_RRS)
        return loss_rrs(cfg, n, y, v, film, stats, ref_mean, err, sw, qt, error_sum, dy);
    return 0.0;
}

// Forward (activations kept), loss, backward of one net in FP64. grad: [grid | mlp],
// overwritten with d(loss_scale * loss)/dparams. dy_ext: [n_out][n] for kind EXTERNAL.
double ora_train_grads(const nrrs_config* cfg, int32_t net, uint32_t n, const nrrs_vertices* v,
                       const float* film, const double* stats, const double* params, int32_t kind,
                       const double* dy_ext, const float* target_L, const float* ref_mean, const float* err,
                       const float* sw, const float* qt, double error_sum, double loss_scale, double* grad,
                       double* dy_out) {
    uint64_t ng = 0, nm = 0;
    ora_param_count(cfg, net, &ng, &nm);
    int nl, no;
    mlp_dims(cfg, net, &nl, &no);
    std::fill(grad, grad + ng + nm, 0.0);
    std::vector<double> enc((size_t)n * ORA_ENC_WIDTH);
    encode_impl(cfg, net, n, v, film, stats, params, enc.data(), true);
    const double* W = params + ng;
    double* dW = grad + ng;
    const std::vector<std::vector<char>> conn = connections(cfg, net);
    // forward, keeping every layer input
    std::vector<std::vector<double>> act(nl);  // act[l] = input of layer l, [n][in_l]
    act[0] = enc;
    std::vector<double> y((size_t)no * n);
    {
        size_t woff = 0;
        int in = ORA_ENC_WIDTH;
        for (int layer = 0; layer < nl; ++layer) {
            int o = layer == nl - 1 ? no : ORA_HIDDEN;
            const double* w = W + woff;
            std::vector<double> out((size_t)n * o);
            for (uint32_t i = 0; i < n; ++i)
                for (int r = 0; r < o; ++r) {
                    double acc = 0;
                    for (int k = 0; k < in; ++k)
                        if (conn[layer][(size_t)r * in + k])
                            acc += w[(size_t)r * in + k] * act[layer][(size_t)i * in + k];
                    out[(size_t)i * o + r] = layer == nl - 1 ? acc : act_d(cfg, acc);
                }
            if (layer < nl - 1) {
                act[layer + 1] = out;
            } else {
                for (uint32_t i = 0; i < n; ++i)
                    for (int r = 0; r < o; ++r) y[(size_t)r * n + i] = out[(size_t)i * o + r];
            }
            woff += (size_t)o * in;
            in = o;
        }
    }
    std::vector<double> dy((size_t)no * n, 0.0);
    double loss = 0;
    if (kind == NRRS_LOSS_EXTERNAL) {
        std::copy(dy_ext, dy_ext + (size_t)no * n, dy.begin());
    } else {
        // AllSelf: each net trains on its own head's loss, half of the NRRS_LL2 mean over 6 outputs
        // ((loss1 + loss2) / 2, twn_AllSelf.cpp:103-108)
        const int head = kind == NRRS_LOSS_LL2 && net != NRRS_RRSNET && cfg->ll2_two_head == 3
                             ? (net == NRRS_STATNET ? 0 : 1)
                             : -1;
        if (head >= 0)
            loss = loss_ll2(cfg, n, y.data(), target_L, nullptr, dy.data(), head);
        else
            loss = ora_loss(cfg, kind, n, y.data(), target_L, v, film, stats, ref_mean, err, sw, qt, error_sum,
                            nullptr, dy.data());
    }
    for (auto& d : dy) d *= loss_scale;
    if (dy_out) std::copy(dy.begin(), dy.end(), dy_out);
    // backward
    std::vector<size_t> woffs(nl);
    {
        size_t woff = 0;
        int in = ORA_ENC_WIDTH;
        for (int layer = 0; layer < nl; ++layer) {
            int o = layer == nl - 1 ? no : ORA_HIDDEN;
            woffs[layer] = woff;
            woff += (size_t)o * in;
            in = o;
        }
    }
    std::vector<double> delta((size_t)n * no), prev;
    for (uint32_t i = 0; i < n; ++i)
        for (int r = 0; r < no; ++r) delta[(size_t)i * no + r] = dy[(size_t)r * n + i];
    int o = no;
    for (int layer = nl - 1; layer >= 0; --layer) {
        int in = layer == 0 ? ORA_ENC_WIDTH : ORA_HIDDEN;
        const double* w = W + woffs[layer];
        double* dw = dW + woffs[layer];
        const std::vector<double>& a = act[layer];
        prev.assign((size_t)n * in, 0.0);
        for (uint32_t i = 0; i < n; ++i) {
            for (int r = 0; r < o; ++r) {
                double dr = delta[(size_t)i * o + r];
                if (dr == 0.0) continue;
                for (int k = 0; k < in; ++k) {
                    if (!conn[layer][(size_t)r * in + k]) continue;
                    dw[(size_t)r * in + k] += dr * a[(size_t)i * in + k];
                    prev[(size_t)i * in + k] += w[(size_t)r * in + k] * dr;
                }
            }
            if (layer > 0)
                for (int k = 0; k < in; ++k)
                {
                    const double sl = act_slope_d(cfg, a[(size_t)i * in + k]);
                    if (sl == 0.0)
                        prev[(size_t)i * in + k] = 0.0;
                    else
                        prev[(size_t)i * in + k] *= sl;
                }
        }
        delta.swap(prev);
        o = in;
    }
    // delta = dL/denc; scatter the grid channels into the table gradient
    {
        const nrrs_grid_config& g = cfg->grid;
        std::vector<Level> L = levels_of(g, nullptr);
        float c[3], d[3];
        aabb_frame(cfg->aabb_min, cfg->aabb_max, c, d);
        for (uint32_t i = 0; i < n; ++i) {
            float x[3];
            for (int k = 0; k < 3; ++k) x[k] = norm_coord(v->pos[k * v->stride + i], c[k], d[k]);
            for (uint32_t l = 0; l < (net_has_grid(cfg, net) ? g.n_levels : 0u); ++l) {
                uint32_t cell[3];
                double fr[3];
                for (int k = 0; k < 3; ++k) {
                    float p = std::fma(x[k], L[l].scale, 0.5f);
                    float fl = std::floor(p);
                    cell[k] = (uint32_t)fl;
                    fr[k] = (double)(p - fl);
                }
                for (uint32_t cn = 0; cn < 8; ++cn) {
                    double w = 1.0;
                    for (int k = 0; k < 3; ++k) w *= ((cn >> k) & 1u) ? fr[k] : 1.0 - fr[k];
                    uint32_t id = corner_index(L[l], g.log2_table_size, cell[0] + (cn & 1u),
                                               cell[1] + ((cn >> 1) & 1u), cell[2] + ((cn >> 2) & 1u));
                    double* t = grad + ((size_t)L[l].offset + id) * 2;
                    t[0] += w * delta[(size_t)i * ORA_ENC_WIDTH + 2 * l];
                    t[1] += w * delta[(size_t)i * ORA_ENC_WIDTH + 2 * l + 1];
                }
            }
        }
    }
    return loss;
}

// Ema optimizer (2net-simple.json:16-18; tiny-cuda-nn's "Ema", absent from the checkout):
// after optimizer step t >= 1, the debiased exponential moving average of the parameters,
// ema_t = sum_k (1 - d) d^(t-k) p_k / (1 - d^t), kept as the recursion
// ema_t = (d (1 - d^(t-1)) ema_{t-1} + (1 - d) p_t) / (1 - d^t).
void ora_ema_step(uint64_t n, double* ema, const double* p, double decay, uint32_t t) {
    const double keep = decay * (1.0 - std::pow(decay, (double)(t - 1))), norm = 1.0 - std::pow(decay, (double)t);
    for (uint64_t i = 0; i < n; ++i) ema[i] = (keep * ema[i] + (1.0 - decay) * p[i]) / norm;
}

// Adam (Kingma & Ba 2015) with L2 regularisation added to the gradient; step t >= 1.
void ora_adam(uint64_t n, double* p, double* m, double* v, const double* g, double lr, double b1, double b2,
              double eps, double l2, uint32_t t) {
    double c1 = 1.0 - std::pow(b1, (double)t), c2 = 1.0 - std::pow(b2, (double)t);
    for (uint64_t i = 0; i < n; ++i) {
        double gg = g[i] + l2 * p[i];
        m[i] = b1 * m[i] + (1.0 - b1) * gg;
        v[i] = b2 * v[i] + (1.0 - b2) * gg * gg;
        p[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
    }
}

// handleIntersections (integrator.cpp:772-794): copy of scatter item sid, thp /= q'(sid).
void ora_consume_survivors(uint32_t n, const uint32_t* surv, const nrrs_vertices* sc, const float* rrs,
                           const int32_t* node_in, const nrrs_vertices_out* out, int32_t* node_out) {
    const uint32_t S = sc->stride, So = out->stride;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t sid = surv[k];
        if (sid >= S) continue;
        for (int c = 0; c < 3; ++c) {
            out->pos[c * So + k] = sc->pos[c * S + sid];
            out->dir[c * So + k] = sc->dir[c * S + sid];
            out->normal[c * So + k] = sc->normal[c * S + sid];
            out->albedo_d[c * So + k] = sc->albedo_d[c * S + sid];
            out->albedo_s[c * So + k] = sc->albedo_s[c * S + sid];
            out->thp[c * So + k] = sc->thp[c * S + sid] / rrs[sid];
        }
        out->rough[k] = sc->rough[sid];
        out->pixel[k] = sc->pixel[sid];
        node_out[k] = node_in ? node_in[sid] : -1;
    }
}

// NRRSPathNodesBuffer::recordNode (workqueue.h:142-168), one vertex after the other.
void ora_nodes_record(ora_nodes* nb, uint32_t n, const nrrs_vertices* v, const int32_t* parent, const float* q,
                      uint32_t depth, int32_t* out) {
    const uint32_t C = nb->cap, S = v->stride;
    for (uint32_t i = 0; i < n; ++i) {
        const int32_t par = parent ? parent[i] : -1;
        if (nb->size >= C) {
            nb->invalid += 1;
            out[i] = par;
            continue;
        }
        const uint32_t j = nb->size++;
        for (int c = 0; c < 3; ++c) {
            nb->pos[c * C + j] = v->pos[c * S + i];
            nb->dir[c * C + j] = v->dir[c * S + i];
            nb->normal[c * C + j] = v->normal[c * S + i];
            nb->ad[c * C + j] = v->albedo_d[c * S + i];
            nb->as[c * C + j] = v->albedo_s[c * S + i];
            nb->thp[c * C + j] = v->thp[c * S + i];
            nb->L[c * C + j] = 0.0;
        }
        nb->rough[j] = v->rough[i];
        nb->pixel[j] = v->pixel[i];
        nb->parent[j] = par;
        nb->q[j] = q ? q[i] : 1.0f;
        nb->depth[j] = depth;
        out[i] = (int32_t)j;
    }
}

// NRRSPathNodesBuffer::recordRadiance (workqueue.h:122-134): the node and every ancestor.
void ora_nodes_radiance(ora_nodes* nb, uint32_t n, const int32_t* node, const float* L, uint32_t Ls) {
    const uint32_t C = nb->cap;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t j = node[i];
        if (j < 0 || (uint32_t)j >= nb->size) continue;
        while (j != -1) {
            for (int c = 0; c < 3; ++c) nb->L[c * C + j] += (double)L[c * Ls + i];
            j = nb->parent[j];
        }
    }
}

// nrrs_generate_training_data (train.cu:432-570) with rows in node order, the duplicate
// rule "the sibling with the lowest node index keeps its q target", then
// nrrs_calculate_sample_weight (train.cu:572-592).
uint32_t ora_training_data(const ora_nodes* nb, const float* film, uint32_t n_pixels, const float* err, uint32_t n_err,
                           uint32_t scale, uint32_t width, int32_t ignore_zero_L, int32_t copy_rrs,
                           int32_t mark_dup, uint32_t capacity, const nrrs_vertices_out* out, double* target_L,
                           double* ref_mean, float* error, float* error_avg, float* q_target, double* sample_weight,
                           uint32_t* pixel_samples) {
    const uint32_t C = nb->cap, S = out->stride, size = nb->size;
    std::vector<uint32_t> first(size, 0xFFFFFFFFu);
    for (uint32_t j = 0; j < size; ++j) {
        const int32_t p = nb->parent[j];
        if (p >= 0 && (uint32_t)p < size && j < first[p]) first[p] = j;
    }
    uint32_t rows = 0;
    for (uint32_t j = 0; j < size; ++j) {
        double L[3];
        bool nan = false;
        for (int c = 0; c < 3; ++c) {
            const double th = nb->thp[c * C + j];
            double l = th > 1e-5 ? nb->L[c * C + j] / th : 0.0;
            l = l < 0.0 ? 0.0 : l;
            L[c] = l;
            nan = nan || std::isnan(l) || std::isnan(nb->pos[c * C + j]) || std::isnan(nb->dir[c * C + j]);
        }
        if (nan) continue;
        if (ignore_zero_L && !((L[0] + L[1] + L[2]) / 3.0 > 0.0)) continue;
        if (rows >= capacity) {
            ++rows;
            continue;
        }
        const uint32_t r = rows++;
        const float q = nb->q[j];
        for (int c = 0; c < 3; ++c) {
            out->pos[c * S + r] = nb->pos[c * C + j];
            out->dir[c * S + r] = nb->dir[c * C + j];
            out->normal[c * S + r] = nb->normal[c * C + j];
            out->albedo_d[c * S + r] = nb->ad[c * C + j];
            out->albedo_s[c * S + r] = nb->as[c * C + j];
            out->thp[c * S + r] = copy_rrs ? nb->thp[c * C + j] * q : nb->thp[c * C + j];
            target_L[(size_t)c * capacity + r] = L[c];
        }
        out->rough[r] = nb->rough[j];
        const uint32_t px = nb->pixel[j];
        out->pixel[r] = px;
        ref_mean[r] = px < n_pixels
                          ? ((double)film[px] + (double)film[n_pixels + px] + (double)film[2 * n_pixels + px]) / 3.0
                          : 0.0;
        float e0 = 0.0f, e1 = 0.0f;
        if (err) {
            const uint32_t off = ((px % width) / scale) + ((px / width) / scale) * (width / scale);
            if (off < n_err) {
                e0 = err[off];
                e1 = err[(size_t)n_err + off];
            }
        }
        error[r] = e0;
        error_avg[r] = e1;
        const int32_t p = nb->parent[j];
        const bool dup = mark_dup && p >= 0 && (uint32_t)p < size && first[p] != j;
        q_target[r] = dup ? -1.0f : q;
        if (!dup && px < n_pixels) pixel_samples[px] += 1;
    }
    const uint32_t n_out = rows < capacity ? rows : capacity;
    for (uint32_t r = 0; r < n_out; ++r) {
        const uint32_t px = out->pixel[r];
        const uint32_t k = px < n_pixels ? pixel_samples[px] : 0u;
        sample_weight[r] = k ? 1.0 / (double)k : 0.0;
    }
    return n_out;
}

// nrrs_update_weighted_L (train.cu:344-363): L = L (1 - b) + (L_frame / spp) b per channel.
void ora_error_blend(uint32_t n, const float* L_frame, const uint32_t* spp, double blend, double* weighted_L) {
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const size_t k = (size_t)c * n + i;
            weighted_L[k] = weighted_L[k] * (1.0 - blend) + ((double)L_frame[k] / (double)spp[i]) * blend;
        }
}

// nrrs_calculate_error (train.cu:365-430): e = mean_c (L - ref)^2 / (ref^2 + 0.01), clamp at
// 100, row 1 = e * sqrt(spp); the sum of row 1 as calcSum2PassAsync<false> forms it
// (integrator.cpp:2466-2471).
double ora_error_compute(uint32_t n, const double* ref, const double* weighted_L, const double* spp, int32_t clamp,
                         double* err) {
    double total = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double e = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double r = ref[(size_t)c * n + i], d = weighted_L[(size_t)c * n + i] - r;
            e += d * d / (r * r + 1e-2);
        }
        e /= 3.0;
        if (clamp) e = std::fmin(e, 100.0);
        err[i] = e;
        const double e2 = spp ? e * std::sqrt(spp[i]) : e;
        err[(size_t)n + i] = e2;
        total += e2;
    }
    return total;
}

}  // extern "C"
