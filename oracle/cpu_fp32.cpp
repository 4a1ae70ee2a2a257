// FP32 multi-threaded CPU path of the benchmark step - TEST INFRASTRUCTURE ONLY (the CPU
// baseline of the measurement; tests check it against the FP64 oracle).
//
// One step on n vertices: StatNet -> RRSNet chained inference, q, the exact normalisation
// sum, the split and the stable survivor compaction; then one training step per net
// (forward, loss, backward, Adam), the RRSNet on the StatNet statistics of the same
// vertices. The formulas are those of nrrs_oracle.cpp in FP32. Vertices are split into
// contiguous ranges, one per thread; the table-gradient scatter is partitioned by table
// range: each thread files its corner records per owner range, then every owner adds the
// records of its range (no atomics, no per-thread copy of the 52 MB table gradient).
#include "nrrs_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kEnc = ORA_ENC_WIDTH, kHid = ORA_HIDDEN;
constexpr uint64_t kMult = 6364136223846793005ull;

template <typename F>
void parallel_for(int threads, uint32_t n, F f) {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        const uint32_t b = (uint32_t)((uint64_t)n * t / threads), e = (uint32_t)((uint64_t)n * (t + 1) / threads);
        pool.emplace_back([=] { f(t, b, e); });
    }
    for (auto& th : pool) th.join();
}

// PCG32 output number `pos` of (seed, stream): jump-ahead (Brown 1994) then one step.
uint32_t pcg_at(uint64_t seed, uint64_t stream, uint64_t pos) {
    const uint64_t inc = (stream << 1u) | 1u;
    uint64_t st = 0;
    st = st * kMult + inc;
    st += seed;
    st = st * kMult + inc;
    uint64_t am = 1, ap = 0, cm = kMult, cp = inc;
    for (uint64_t d = pos; d; d >>= 1) {
        if (d & 1u) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm *= cm;
    }
    st = am * st + ap;
    const uint32_t xs = (uint32_t)(((st >> 18u) ^ st) >> 27u), rot = (uint32_t)(st >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

struct Grid {
    uint32_t L = 0, log2T = 0;
    uint32_t res[16], off[16], size[16], dense[16];
    float scale[16];
    uint64_t n_entries = 0;
};

struct Cell {
    uint32_t id[8];
    float w[8];
};

inline float norm_coord(float p, float c, float d) {
    float x = 0.5f + (p - c) / d;
    return std::fmin(std::fmax(x, 0.0f), 1.0f);
}

inline void level_cell(const Grid& g, uint32_t l, const float* x, Cell& cl) {
    uint32_t c[3];
    float fr[3];
    for (int k = 0; k < 3; ++k) {
        float p = std::fma(x[k], g.scale[l], 0.5f);
        float fl = std::floor(p);
        c[k] = (uint32_t)fl;
        fr[k] = p - fl;
    }
    for (uint32_t cn = 0; cn < 8; ++cn) {
        const uint32_t cx = c[0] + (cn & 1u), cy = c[1] + ((cn >> 1) & 1u), cz = c[2] + ((cn >> 2) & 1u);
        uint32_t id;
        if (g.dense[l]) {
            id = cx + g.res[l] * (cy + g.res[l] * cz);
        } else {
            id = (cx ^ (cy * 2654435761u) ^ (cz * 805459861u)) & ((1u << g.log2T) - 1u);
        }
        cl.id[cn] = g.off[l] + id;
        float w = 1.0f;
        for (int k = 0; k < 3; ++k) w *= ((cn >> k) & 1u) ? fr[k] : 1.0f - fr[k];
        cl.w[cn] = w;
    }
}

inline void sh4(float x, float y, float z, float* o) {
    const float xx = x * x, yy = y * y, zz = z * z;
    o[0] = 0.28209479177387814f;
    o[1] = -0.48860251190291987f * y;
    o[2] = 0.48860251190291987f * z;
    o[3] = -0.48860251190291987f * x;
    o[4] = 1.0925484305920792f * x * y;
    o[5] = -1.0925484305920792f * y * z;
    o[6] = 0.94617469575755997f * zz - 0.31539156525251999f;
    o[7] = -1.0925484305920792f * x * z;
    o[8] = 0.54627421529603959f * (xx - yy);
    o[9] = 0.59004358992664352f * y * (-3.0f * xx + yy);
    o[10] = 2.8906114426405538f * x * y * z;
    o[11] = 0.45704579946446572f * y * (1.0f - 5.0f * zz);
    o[12] = 0.3731763325901154f * z * (5.0f * zz - 3.0f);
    o[13] = 0.45704579946446572f * x * (1.0f - 5.0f * zz);
    o[14] = 1.4453057213202769f * z * (xx - yy);
    o[15] = 0.59004358992664352f * x * (-xx + 3.0f * yy);
}

inline float bc(float x) { return 2.0f * (std::sqrt(x) - 1.0f); }
inline float sigma(float x) { return x < 0 ? std::log1p(std::exp(x)) : 0.5f * x + 0.6931471805599453f; }
inline float sigma_grad(float x) { return x < 0 ? std::exp(x) / (1.0f + std::exp(x)) : 0.5f; }

struct Ctx {
    const nrrs_config* cfg;
    Grid g;
    float c[3], d[3];
    const nrrs_vertices* v;
    const float* film;
    uint32_t P;
};

// Encoding of vertex i (natural channel order, as the oracle); the grid corner cells of
// each level go to cells (for the backward).
void encode(const Ctx& C, int net, uint32_t i, const float* params, const float* st, float* e, Cell* cells) {
    const nrrs_vertices& v = *C.v;
    const uint32_t S = v.stride;
    std::fill(e, e + kEnc, 0.0f);
    float x[3];
    for (int k = 0; k < 3; ++k) x[k] = norm_coord(v.pos[k * S + i], C.c[k], C.d[k]);
    for (uint32_t l = 0; l < C.g.L; ++l) {
        Cell& cl = cells[l];
        level_cell(C.g, l, x, cl);
        float f0 = 0, f1 = 0;
        for (int cn = 0; cn < 8; ++cn) {
            const float* t = params + (size_t)cl.id[cn] * 2;
            f0 += cl.w[cn] * t[0];
            f1 += cl.w[cn] * t[1];
        }
        e[2 * l] = f0;
        e[2 * l + 1] = f1;
    }
    sh4(v.dir[i], v.dir[S + i], v.dir[2 * S + i], e + 32);
    sh4(v.normal[i], v.normal[S + i], v.normal[2 * S + i], e + 48);
    e[64] = 1.0f - std::exp(-v.rough[i]);
    if (net == NRRS_STATNET) {
        for (int k = 0; k < 3; ++k) {
            e[65 + k] = v.albedo_d[k * S + i];
            e[68 + k] = v.albedo_s[k * S + i];
        }
        e[71] = 1.0f;
    } else {
        const uint32_t px = v.pixel[i];
        const float ref = (C.film[px] + C.film[C.P + px] + C.film[2 * C.P + px]) / 3.0f;
        for (int k = 0; k < 3; ++k) e[65 + k] = bc(v.thp[k * S + i]);
        e[68] = bc(ref);
        for (int k = 0; k < 3; ++k) {
            e[69 + k] = bc(std::max(st[k], 0.0f));
            e[72 + k] = bc(std::max(st[3 + k], 0.0f));
        }
        e[75] = 1.0f;
    }
}

// Forward of one vertex; acts[l] = input of layer l ([nl][kEnc] scratch), y = outputs.
void forward(const float* W, int nl, int no, const float* e, float* acts, float* y) {
    std::memcpy(acts, e, kEnc * sizeof(float));
    int in = kEnc;
    for (int layer = 0; layer < nl; ++layer) {
        const int o = layer == nl - 1 ? no : kHid;
        const float* h = acts + (size_t)layer * kEnc;
        float* out = layer == nl - 1 ? y : acts + (size_t)(layer + 1) * kEnc;
        for (int r = 0; r < o; ++r) {
            float acc = 0;
            const float* w = W + (size_t)r * in;
            for (int k = 0; k < in; ++k) acc += w[k] * h[k];
            out[r] = layer == nl - 1 ? acc : std::max(acc, 0.0f);
        }
        W += (size_t)o * in;
        in = o;
    }
}

struct Rec {
    uint32_t id;
    float g0, g1;
};

struct NetRun {
    const float* params;
    float* grad;  // [grid | mlp]
    uint64_t ng, nm;
    int nl, no;
};

// One training step of a net: per thread forward / loss / backward over its vertex range,
// dW per thread, corner records per owner range; then the owners add their records.
void train_net(const Ctx& C, int net, NetRun& R, uint32_t n, const float* stats, const float* tL, const float* err,
               const float* sw, const float* qt, double error_sum, int threads) {
    std::fill(R.grad, R.grad + R.ng + R.nm, 0.0f);
    std::vector<std::vector<float>> dW(threads, std::vector<float>(R.nm, 0.0f));
    // recs[t][owner]
    std::vector<std::vector<std::vector<Rec>>> recs(threads, std::vector<std::vector<Rec>>(threads));
    const uint64_t ne = R.ng / 2;
    const int active = C.cfg->ll2_step == 0 ? 3 : 6;
    const float ll2_norm = 1.0f / ((float)n * active);
    const float Ebar = (float)std::fmax(error_sum / (double)C.P, 1e-4);
    const float* gam = C.cfg->gamma;
    parallel_for(threads, n, [&](int t, uint32_t b, uint32_t e_) {
        for (auto& r : recs[t]) r.reserve((size_t)(e_ - b) * 8 * C.g.L / threads * 5 / 4 + 64);
        std::vector<float> enc(kEnc), acts((size_t)R.nl * kEnc), y(8), delta(kEnc), prev(kEnc);
        std::vector<Cell> cells(C.g.L);
        float* dw = dW[t].data();
        const float* W = R.params + R.ng;
        const uint32_t S = C.v->stride;
        for (uint32_t i = b; i < e_; ++i) {
            float st[6];
            if (net == NRRS_RRSNET)
                for (int k = 0; k < 6; ++k) st[k] = stats[(size_t)k * n + i];
            encode(C, net, i, R.params, st, enc.data(), cells.data());
            forward(W, R.nl, R.no, enc.data(), acts.data(), y.data());
            // dL/dy
            float dy[8] = {0};
            if (net == NRRS_STATNET) {
                for (int c = 0; c < active; ++c) {
                    float tg = std::fmax(tL[(size_t)(c % 3) * n + i], 0.0f);
                    if (C.cfg->loss_clamp_on) tg = std::fmin(tg, C.cfg->loss_clamp_max);
                    if (c >= 3) tg = tg * tg;
                    const float den = y[c] * y[c] + 0.01f;
                    dy[c] = 2.0f * (y[c] - tg) / den * ll2_norm;
                }
            } else if (C.cfg->rrs_step >= 3 && qt[i] >= 0.0f) {
                const uint32_t px = C.v->pixel[i];
                const float ref = (C.film[px] + C.film[C.P + px] + C.film[2 * C.P + px]) / 3.0f;
                float m2 = 0;
                for (int c = 0; c < 3; ++c) {
                    const float th = C.v->thp[c * S + i];
                    m2 += th * th * std::fmax(st[3 + c], 0.0f);
                }
                float vv = m2 / 3.0f / (ref * ref + 0.01f);
                if (C.cfg->loss_clamp_on) vv = std::fmin(vv, C.cfg->loss_clamp_max);
                const float ee = std::fmax(err[i], 1e-4f), q = sigma(y[0]), qe = q + 0.01f;
                const float fp = -(gam[0] + gam[1]) * vv / (qe * qe) + gam[0] * Ebar + gam[1] * ee + gam[2] +
                                 2.0f * gam[3] * (q - qt[i]);
                dy[0] = sw[i] * fp * sigma_grad(y[0]) / (float)n;
            }
            // backward
            std::fill(delta.begin(), delta.end(), 0.0f);
            for (int r = 0; r < R.no; ++r) delta[r] = dy[r];
            size_t woffs[8];
            {
                size_t w = 0;
                int in = kEnc;
                for (int layer = 0; layer < R.nl; ++layer) {
                    const int o = layer == R.nl - 1 ? R.no : kHid;
                    woffs[layer] = w;
                    w += (size_t)o * in;
                    in = o;
                }
            }
            int o = R.no;
            for (int layer = R.nl - 1; layer >= 0; --layer) {
                const int in = layer == 0 ? kEnc : kHid;
                const float* w = W + woffs[layer];
                float* g = dw + woffs[layer];
                const float* a = acts.data() + (size_t)layer * kEnc;
                std::fill(prev.begin(), prev.begin() + in, 0.0f);
                for (int r = 0; r < o; ++r) {
                    const float dr = delta[r];
                    if (dr == 0.0f) continue;
                    for (int k = 0; k < in; ++k) {
                        g[(size_t)r * in + k] += dr * a[k];
                        prev[k] += w[(size_t)r * in + k] * dr;
                    }
                }
                if (layer > 0)
                    for (int k = 0; k < in; ++k)
                        if (!(a[k] > 0.0f)) prev[k] = 0.0f;
                std::swap(delta, prev);
                o = in;
            }
            for (uint32_t l = 0; l < C.g.L; ++l) {
                const float g0 = delta[2 * l], g1 = delta[2 * l + 1];
                for (int cn = 0; cn < 8; ++cn) {
                    const uint32_t id = cells[l].id[cn];
                    const int owner = (int)((uint64_t)id * threads / ne);
                    recs[t][owner].push_back(Rec{id, cells[l].w[cn] * g0, cells[l].w[cn] * g1});
                }
            }
        }
    });
    parallel_for(threads, (uint32_t)threads, [&](int, uint32_t b, uint32_t e_) {
        for (uint32_t owner = b; owner < e_; ++owner)
            for (int t = 0; t < threads; ++t)
                for (const Rec& r : recs[t][owner]) {
                    R.grad[(size_t)r.id * 2] += r.g0;
                    R.grad[(size_t)r.id * 2 + 1] += r.g1;
                }
    });
    for (int t = 0; t < threads; ++t)
        for (uint64_t k = 0; k < R.nm; ++k) R.grad[R.ng + k] += dW[t][k];
}

}  // namespace

extern "C" {

double ora_cpu_step_f32(const nrrs_config* cfg, uint32_t n, const nrrs_vertices* v, const float* film,
                        float* params_stat, float* params_rrs, const float* target_L, const float* err,
                        const float* sw, const float* qt, double error_sum, int32_t threads, float* q_out,
                        float* grad_stat, float* grad_rrs, double* times) {
    using clk = std::chrono::steady_clock;
    Ctx C{};
    C.cfg = cfg;
    C.v = v;
    C.film = film;
    C.P = cfg->n_pixels;
    {
        uint32_t res[16], off[16], size[16], dense[16];
        float scale[16];
        const uint64_t np = ora_grid_layout(&cfg->grid, res, off, size, dense, scale);
        C.g.L = cfg->grid.n_levels;
        C.g.log2T = cfg->grid.log2_table_size;
        C.g.n_entries = np / 2;
        for (uint32_t l = 0; l < C.g.L; ++l) {
            C.g.res[l] = res[l];
            C.g.off[l] = off[l];
            C.g.size[l] = size[l];
            C.g.dense[l] = dense[l];
            C.g.scale[l] = scale[l];
        }
    }
    {
        const float* mn = cfg->aabb_min;
        const float* mx = cfg->aabb_max;
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        float infl = std::sqrt(dx * dx + dy * dy + dz * dz) * 0.005f;
        for (int k = 0; k < 3; ++k) {
            float lo = mn[k] - infl, hi = mx[k] + infl;
            C.c[k] = (lo + hi) * 0.5f;
            C.d[k] = hi - lo;
        }
    }
    if (threads < 1) threads = 1;
    NetRun Rs{}, Rr{};
    uint64_t ng, nm;
    ora_param_count(cfg, NRRS_STATNET, &ng, &nm);
    Rs.params = params_stat;
    Rs.ng = ng;
    Rs.nm = nm;
    Rs.nl = (int)cfg->statnet.n_hidden_layers + 1;
    Rs.no = (int)cfg->statnet.n_outputs;
    ora_param_count(cfg, NRRS_RRSNET, &ng, &nm);
    Rr.params = params_rrs;
    Rr.ng = ng;
    Rr.nm = nm;
    Rr.nl = (int)cfg->rrsnet.n_hidden_layers + 1;
    Rr.no = (int)cfg->rrsnet.n_outputs;
    std::vector<float> gs(Rs.ng + Rs.nm), gr(Rr.ng + Rr.nm);
    Rs.grad = grad_stat ? grad_stat : gs.data();
    Rr.grad = grad_rrs ? grad_rrs : gr.data();

    const auto t0 = clk::now();
    // ---- inference: StatNet -> RRSNet -> q
    std::vector<float> stats((size_t)6 * n), q(n);
    parallel_for(threads, n, [&](int, uint32_t b, uint32_t e_) {
        std::vector<float> enc(kEnc), acts((size_t)8 * kEnc), y(8);
        std::vector<Cell> cells(C.g.L);
        for (uint32_t i = b; i < e_; ++i) {
            encode(C, NRRS_STATNET, i, Rs.params, nullptr, enc.data(), cells.data());
            forward(Rs.params + Rs.ng, Rs.nl, Rs.no, enc.data(), acts.data(), y.data());
            float st[6];
            for (int k = 0; k < 6; ++k) stats[(size_t)k * n + i] = st[k] = y[k];
            encode(C, NRRS_RRSNET, i, Rr.params, st, enc.data(), cells.data());
            forward(Rr.params + Rr.ng, Rr.nl, Rr.no, enc.data(), acts.data(), y.data());
            float r = sigma(y[0]);
            if (cfg->rrs_weight_window) {
                const float dm = 2.0f / (1.0f + cfg->weight_window_size), dp = dm * cfg->weight_window_size;
                r = r < dm ? r / dm : (r > dp ? r / dp : 1.0f);
            }
            q[i] = std::fmin(std::fmax(r, cfg->rrs_clamp_min), cfg->rrs_clamp_max);
        }
    });
    // ---- decision: exact sum, split, stable compaction
    std::vector<uint64_t> part(threads, 0);
    std::vector<uint32_t> cnt(threads, 0);
    std::vector<uint8_t> s(n);
    parallel_for(threads, n, [&](int t, uint32_t b, uint32_t e_) {
        uint64_t a = 0;
        for (uint32_t i = b; i < e_; ++i) {
            const float tq = cfg->rrs_use_ceil ? std::ceil(q[i]) : q[i];
            a += (uint64_t)((double)tq * 4294967296.0);
        }
        part[t] = a;
    });
    uint64_t acc = 0;
    for (uint64_t a : part) acc += a;
    const float sum_f = (float)((double)acc * 0x1p-32);
    const float rcp = sum_f > 0.0f ? 1.0f / sum_f : 0.0f;
    float nf = (float)cfg->n_pixels * rcp;
    nf = nf * (cfg->rrs_use_ceil ? 1.0f : cfg->rrs_normalize_rate);
    parallel_for(threads, n, [&](int t, uint32_t b, uint32_t e_) {
        uint32_t c = 0;
        for (uint32_t i = b; i < e_; ++i) {
            float r = q[i];
            r = cfg->rrs_no_upscale ? r * (nf < 1.0f ? nf : 1.0f) : r * nf;
            const float u = (float)(pcg_at(1, 0, i) >> 8) * 0x1p-24f;
            int si = (int)r;
            const float left = r - (float)si;
            si += (u <= left) ? 1 : 0;
            if ((uint32_t)si > cfg->split_cap) si = (int)cfg->split_cap;
            s[i] = (uint8_t)si;
            c += (uint32_t)si;
        }
        cnt[t] = c;
    });
    uint32_t total = 0;
    std::vector<uint32_t> base(threads);
    for (int t = 0; t < threads; ++t) {
        base[t] = total;
        total += cnt[t];
    }
    std::vector<uint32_t> surv(total);
    parallel_for(threads, n, [&](int t, uint32_t b, uint32_t e_) {
        uint32_t w = base[t];
        for (uint32_t i = b; i < e_; ++i)
            for (int c = 0; c < s[i]; ++c) surv[w++] = i;
    });
    const auto t1 = clk::now();
    // ---- training: StatNet then RRSNet (on the StatNet statistics), Adam after each
    train_net(C, NRRS_STATNET, Rs, n, nullptr, target_L, err, sw, qt, error_sum, threads);
    train_net(C, NRRS_RRSNET, Rr, n, stats.data(), target_L, err, sw, qt, error_sum, threads);
    for (NetRun* R : {&Rs, &Rr}) {
        float* p = const_cast<float*>(R->params);
        const float* g = R->grad;
        const uint64_t np = R->ng + R->nm;
        const float lr = R == &Rs ? cfg->lr_statnet : cfg->lr_rrsnet, b1 = cfg->adam_beta1, b2 = cfg->adam_beta2;
        const float c1 = 1.0f - b1, c2 = 1.0f - b2;  // bias corrections of step 1
        parallel_for(threads, (uint32_t)np, [&](int, uint32_t b, uint32_t e_) {
            for (uint32_t k = b; k < e_; ++k) {
                const float gg = g[k] + cfg->adam_l2_reg * p[k];
                const float m = (1.0f - b1) * gg, vv = (1.0f - b2) * gg * gg;
                p[k] -= lr * (m / c1) / (std::sqrt(vv / c2) + cfg->adam_eps);
            }
        });
    }
    const auto t2 = clk::now();
    if (q_out) std::copy(q.begin(), q.end(), q_out);
    if (times) {
        times[0] = std::chrono::duration<double>(t1 - t0).count();
        times[1] = std::chrono::duration<double>(t2 - t1).count();
    }
    return std::chrono::duration<double>(t2 - t0).count();
}

}  // extern "C"
