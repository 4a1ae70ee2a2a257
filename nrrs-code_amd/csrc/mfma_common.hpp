#pragma once
// MFMA building blocks of the mixed-precision inference path (v_mfma_f32_32x32x16_{bf16,f16},
// FP32 accumulation), shared by the per-net kernels and the persistent frame kernel.
//
// Layout. A wavefront processes 32 vertices at a time; lane l owns vertex
// l % 32 and half h = l / 32 of every 16-wide K block, which is exactly the B
// operand layout of the 32x32x16 MFMA (B[k][n]: n = l % 32, k = 8h + j). The
// encoding is computed directly in that layout (each half gathers 8 of the 16
// grid levels). The weights are the A operand (A[m][k]: m = l % 32, k = 8h + j),
// pre-packed per (layer, 32-row tile, K block) into 1 KiB LDS tiles.
//
// The accumulator of a layer (D[m][n]: n = l % 32, m = 8(r/4) + 4h + r%4 for
// register r) is fed to the next layer without any shuffle: registers 8g..8g+7
// of output tile t become K block 2t+g, and the next layer's weights are packed
// with their input columns permuted to match, p(kb,h,j) = 16kb + 8(j/4) + 4h + j%4.
#include "device_common.hpp"

namespace nrrs {
namespace mfma {


typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct BF16 {
    typedef bf16x8 vec8;
    // StatNet runs in split precision (see mlp_forward): with 8-bit mantissas its statistics,
    // which reach q through a square root near zero, would carry q errors of about 1e-2
    static constexpr bool kSplitStat = true;
    __device__ static float2 pair(uint32_t u) {
        return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
    }
    __device__ static vec8 pack(const float* x) {
        vec8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (__bf16)x[j];
        return v;
    }
    __device__ static f32x16 mma(u32x4 a, vec8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vec8, a), b, c, 0, 0, 0);
    }
};

struct F16 {
    typedef f16x8 vec8;
    static constexpr bool kSplitStat = false;
    __device__ static float2 pair(uint32_t u) {
        _Float16 lo = __builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu));
        _Float16 hi = __builtin_bit_cast(_Float16, (uint16_t)(u >> 16));
        return make_float2((float)lo, (float)hi);
    }
    __device__ static vec8 pack(const float* x) {
        vec8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (_Float16)f32_rounded(x[j]);
        return v;
    }
    __device__ static f32x16 mma(u32x4 a, vec8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(vec8, a), b, c, 0, 0, 0);
    }
};

template <typename P>
__device__ __forceinline__ uint16_t lp_bits(float f);
template <>
__device__ __forceinline__ uint16_t lp_bits<BF16>(float f) {
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}
template <>
__device__ __forceinline__ uint16_t lp_bits<F16>(float f) {
    return __builtin_bit_cast(uint16_t, (_Float16)f32_rounded(f));
}

constexpr int kTileHalves = 64 * 8;  // one packed A tile: 64 lanes x 8 values

// Packed A tiles of a net, 1 KiB each. Plain: layer 0 (2 row tiles x 5 K chunks), each hidden
// layer (2 x 4), the output layer (4 K chunks) and its residual tiles (4: lp(W - lp(W))).
// Split (every layer in two-term precision but the grid chunks 0-1 of layer 0): layer 0, its
// residual tiles of K chunks 2-4 (2 x 3), then every hidden layer's 8 tiles followed by their 8
// residual tiles, then the output layer as above.
__host__ __device__ constexpr int tiles_of(int n_layers, bool split = false) {
    return 10 + (split ? 6 : 0) + (n_layers - 2) * 8 * (split ? 2 : 1) + 8;
}
// Whether a net runs in split precision: the StatNet (6 outputs, one or two heads) in
// precisions with kSplitStat.
template <typename P>
__host__ __device__ constexpr bool split_of(int n_out) {
    return P::kSplitStat && n_out == 6;
}
// First tile of hidden layer l (1..n_layers-2) and of the output layer.
__host__ __device__ constexpr int hidden_base(int layer, bool split) {
    return 10 + (split ? 6 : 0) + (layer - 1) * 8 * (split ? 2 : 1);
}
__host__ __device__ constexpr int out_base(int n_layers, bool split) { return hidden_base(n_layers - 1, split); }

// Input column of the packed K element (layer, kb, h, j).
__device__ __forceinline__ int packed_col(int layer, int kb, int h, int j) {
    return layer == 0 ? kb * 16 + 8 * h + j : 16 * kb + 8 * (j >> 2) + 4 * h + (j & 3);
}

// Tile t, lane `lane` (0..63) of the packed A tiles of an MLP (layout of tiles_of): 8 weights of
// the A operand (residual tiles hold lp(W - lp(W)); the difference is exact in FP32).
template <typename P>
__device__ __forceinline__ void pack_mlp_tile(const float* __restrict__ W, int n_layers, int n_out, int split,
                                              uint16_t* __restrict__ out, int t, int lane) {
    int layer, ot, kb, base_w = 0, in = kEnc, rows = kHidden;
    bool residual = false;
    const int l0_tiles = split ? 16 : 10, per_hidden = split ? 16 : 8;
    if (t < 10) {
        layer = 0;
        ot = t / 5;
        kb = t % 5;
    } else if (t < l0_tiles) {  // residuals of layer 0's K chunks 2-4
        layer = 0;
        ot = (t - 10) / 3;
        kb = 2 + (t - 10) % 3;
        residual = true;
    } else {
        int u = t - l0_tiles;
        layer = 1 + u / per_hidden;
        if (layer <= n_layers - 2) {
            const int v = u % per_hidden;
            residual = v >= 8;
            ot = (v % 8) / 4;
            kb = v % 4;
        } else {
            layer = n_layers - 1;
            ot = 0;
            kb = u - (n_layers - 2) * per_hidden;  // 0..3: W, 4..7: residuals
            residual = kb >= 4;
            kb &= 3;
        }
        base_w = kHidden * kEnc + (layer - 1) * kHidden * kHidden;
        in = kHidden;
        rows = layer == n_layers - 1 ? n_out : kHidden;
    }
    const int h = lane >> 5, row = ot * 32 + (lane & 31);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float w = row < rows ? W[base_w + row * in + packed_col(layer, kb, h, j)] : 0.0f;
        if (residual) w -= P::pair(lp_bits<P>(w)).x;
        out[((size_t)t * 64 + lane) * 8 + j] = lp_bits<P>(w);
    }
}

// The 40 channels of a vertex that half h of the wavefront feeds to layer 0:
// channel kb*16 + 8h + j goes to element j of K block kb. Chunks 0-1 are the grid
// (levels 4h..4h+3 and 8+4h..8+4h+3), chunks 2-4 the SH and per-vertex channels.
template <typename P>
__device__ __forceinline__ void encode_grid_half(const VertexRec& v, const GridParams& grid,
                                                 const uint32_t* __restrict__ table, int h, typename P::vec8* B) {
    float e[8];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int ll = 0; ll < 4; ++ll) {
            const int l = 8 * kb + 4 * h + ll;
            float2 f = make_float2(0.0f, 0.0f);
            if (l < (int)grid.n_levels) f = grid_level_pairs<PairLoadLp<P>>(table, grid, l, v.x);
            e[2 * ll] = f.x;
            e[2 * ll + 1] = f.y;
        }
        B[kb] = P::pack(e);
    }
}

// Two-term operand of 8 FP32 values: hi = lp(x), lo = lp(x - hi) (the difference is exact in
// FP32), so that hi + lo holds x to about 16 bits.
template <typename P>
__device__ __forceinline__ void pack_split(const float* x, typename P::vec8& hi, typename P::vec8& lo) {
    float r[8];
    hi = P::pack(x);
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = x[j] - (float)hi[j];
    lo = P::pack(r);
}

// K chunks 2-4 (SH-4 of the direction and the normal, per-vertex channels) of the lane's half;
// with Bl, also their residual terms (Bl[0..2] for chunks 2..4).
template <typename P>
__device__ __forceinline__ void encode_tail_half(const VertexRec& v, int net, int h, float ref, const float* st,
                                                 typename P::vec8* B, typename P::vec8* Bl = nullptr,
                                                 bool aid = false, int set = NRRS_NETS_BENCH) {
    float t[kEnc], e[8];
    encode_tail_rec(v, net, ref, st, t, false, aid, set);
#pragma unroll
    for (int kb = 2; kb < 5; ++kb) {
#pragma unroll
        for (int j = 0; j < 8; ++j) e[j] = h ? t[kb * 16 + 8 + j] : t[kb * 16 + j];
        if (Bl)
            pack_split<P>(e, B[kb], Bl[kb - 2]);
        else
            B[kb] = P::pack(e);
    }
}

template <typename P>
__device__ __forceinline__ void encode_half_rec(const VertexRec& v, const GridParams& grid, int net,
                                                const uint32_t* __restrict__ table, int h, float ref,
                                                const float* st, typename P::vec8* B, typename P::vec8* Bl,
                                                bool aid = false, int set = NRRS_NETS_BENCH) {
    encode_grid_half<P>(v, grid, table, h, B);
    encode_tail_half<P>(v, net, h, ref, st, B, Bl, aid, set);
}

template <typename P>
__device__ __forceinline__ void encode_half(const EncodeArgs& a, int net, const uint32_t* __restrict__ table,
                                            uint32_t i, int h, typename P::vec8* B, typename P::vec8* Bl = nullptr) {
    VertexRec v = vertex_from_soa(a, i);
    const float ref = net == NRRS_RRSNET ? film_mean(a.film, a.n_pixels, v.pixel) : 0.0f;
    if (net == NRRS_RRSNET && a.stats && !a.aid) {  // two calls rather than a selected pointer (st stays in registers)
        float st[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) st[k] = a.stats[(size_t)k * a.stats_stride + i];
        encode_half_rec<P>(v, a.grid, net, table, h, ref, st, B, Bl, false, a.net_set);
    } else {
        encode_half_rec<P>(v, a.grid, net, table, h, ref, nullptr, B, Bl, a.aid != 0, a.net_set);
    }
}

// Grid part of the B operands from the level-major encoding: chunk kb of lane half h holds
// levels 8kb + 4h + {0..3}, each a packed feature pair (feature 0 in the low half).
template <typename P>
__device__ __forceinline__ void grid_half_from_levels(const uint32_t* __restrict__ enc, uint32_t stride,
                                                      uint32_t n_levels, uint32_t i, int h, typename P::vec8* B) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        u32x4 u;
#pragma unroll
        for (int ll = 0; ll < 4; ++ll) {
            const uint32_t l = 8 * kb + 4 * h + ll;
            u[ll] = l < n_levels ? enc[(size_t)l * stride + i] : 0u;
        }
        B[kb] = __builtin_bit_cast(typename P::vec8, u);
    }
}

// encode_half with the grid channels read from the level-major encoding. SET: the network set
// when the caller knows it at compile time, else -1 (a.net_set).
template <typename P, int SET = -1>
__device__ __forceinline__ void encode_half_pre(const EncodeArgs& a, int net, uint32_t i, int h, typename P::vec8* B,
                                                typename P::vec8* Bl = nullptr) {
    const int set = SET >= 0 ? SET : a.net_set;
    VertexRec v = vertex_from_soa(a, i);
    const float ref = net == NRRS_RRSNET ? film_mean(a.film, a.n_pixels, v.pixel) : 0.0f;
    grid_half_from_levels<P>(a.grid_enc, a.grid_enc_stride, a.grid.n_levels, i, h, B);
    if (net == NRRS_RRSNET && a.stats && !a.aid) {  // two calls rather than a selected pointer (st stays in registers)
        float st[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) st[k] = a.stats[(size_t)k * a.stats_stride + i];
        encode_tail_half<P>(v, net, h, ref, st, B, Bl, false, set);
    } else {
        encode_tail_half<P>(v, net, h, ref, nullptr, B, Bl, a.aid != 0, set);
    }
}

__device__ __forceinline__ u32x4 lds_tile(const u32x4* __restrict__ sW, int tile, int lane) {
    return sW[tile * 64 + lane];
}

// Hidden-layer activation of an FP32 accumulator value (Activation).
template <int ACT>
__device__ __forceinline__ float act_fwd(float x) {
    if constexpr (ACT == ACT_LEAKY)
        return x > 0.0f ? x : x * kLeakySlope;
    else
        return fmaxf(x, 0.0f);
}

// Hidden width W: 64, or 32 for the shipped nets, whose 64-wide storage holds zeros outside the
// 32-wide net: their kernels compute output tile 0 and K chunks 0-1 of the hidden layers only,
// and of layer 0 the K chunks set in a mask (the chunks that hold the net's inputs).
template <int W>
constexpr int kOutTiles = W / 32;
template <int W>
constexpr int kHidChunks = W / 16;

template <typename P, int ACT = ACT_RELU, int W = 64>
__device__ __forceinline__ void relu_to_operands(const f32x16* D, typename P::vec8* H) {
#pragma unroll
    for (int ot = 0; ot < kOutTiles<W>; ++ot) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = act_fwd<ACT>(D[ot][8 * g + j]);
            H[2 * ot + g] = P::pack(x);
        }
    }
}

// ReLU of the last hidden layer as a two-term operand: H = lp(x) and L = lp(x - H) (the
// difference is exact in FP32), so that H + L holds x to about 16 bits.
template <typename P, int ACT = ACT_RELU, int W = 64>
__device__ __forceinline__ void relu_to_operands_split(const f32x16* D, typename P::vec8* H, typename P::vec8* L) {
#pragma unroll
    for (int ot = 0; ot < kOutTiles<W>; ++ot) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = act_fwd<ACT>(D[ot][8 * g + j]);
            pack_split<P>(x, H[2 * ot + g], L[2 * ot + g]);
        }
    }
}

// Fused MLP of one net from the layer-0 operands B (5 K chunks); weights are the
// packed A tiles at sW (layout of tiles_of). Returns the output accumulator: output c of the
// lane's vertex sits in lane half c / 4, register c % 4.
// The output layer runs in three products, W_lo h_hi + W_hi h_lo + W_hi h_hi, with the weights
// and the last hidden activations each split into two low-precision terms, so that its result
// is about as accurate as an FP32 product of the FP32 activations (the outputs reach q through
// a square root near zero, where an 8-bit-mantissa product is too coarse). SPLIT runs every
// hidden layer the same way, and layer 0 too on its SH and per-vertex chunks (Bl: their residual
// terms), the grid chunks staying single-term. W, l0: the hidden width and the live layer-0
// K chunks (above).
template <typename P, int NL, bool SPLIT = false, int ACT = ACT_RELU, int W = 64>
__device__ __forceinline__ f32x16 mlp_forward(const u32x4* __restrict__ sW, const typename P::vec8* B,
                                              const typename P::vec8* Bl, int lane, uint32_t l0 = 0x1fu) {
    constexpr int kOT = kOutTiles<W>, kKB = kHidChunks<W>;
    f32x16 D[2];
#pragma unroll
    for (int ot = 0; ot < kOT; ++ot) {
        D[ot] = f32x16{};
        if constexpr (SPLIT) {
#pragma unroll
            for (int kb = 2; kb < 5; ++kb) {
                if (W == 32 && !((l0 >> kb) & 1u)) continue;
                D[ot] = P::mma(lds_tile(sW, 10 + ot * 3 + (kb - 2), lane), B[kb], D[ot]);
                D[ot] = P::mma(lds_tile(sW, ot * 5 + kb, lane), Bl[kb - 2], D[ot]);
            }
        }
#pragma unroll
        for (int kb = 0; kb < 5; ++kb)
            if (W == 64 || ((l0 >> kb) & 1u)) D[ot] = P::mma(lds_tile(sW, ot * 5 + kb, lane), B[kb], D[ot]);
    }
    typename P::vec8 H[4], L[4];
#pragma unroll
    for (int layer = 1; layer < NL - 1; ++layer) {
        const int base = hidden_base(layer, SPLIT);
        if constexpr (SPLIT)
            relu_to_operands_split<P, ACT, W>(D, H, L);
        else
            relu_to_operands<P, ACT, W>(D, H);
#pragma unroll
        for (int ot = 0; ot < kOT; ++ot) {
            D[ot] = f32x16{};
            if constexpr (SPLIT) {
#pragma unroll
                for (int kb = 0; kb < kKB; ++kb) {
                    D[ot] = P::mma(lds_tile(sW, base + 8 + ot * 4 + kb, lane), H[kb], D[ot]);
                    D[ot] = P::mma(lds_tile(sW, base + ot * 4 + kb, lane), L[kb], D[ot]);
                }
            }
#pragma unroll
            for (int kb = 0; kb < kKB; ++kb) D[ot] = P::mma(lds_tile(sW, base + ot * 4 + kb, lane), H[kb], D[ot]);
        }
    }
    relu_to_operands_split<P, ACT, W>(D, H, L);
    f32x16 Y = f32x16{};
    const int base = out_base(NL, SPLIT);
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb) Y = P::mma(lds_tile(sW, base + 4 + kb, lane), H[kb], Y);
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb) Y = P::mma(lds_tile(sW, base + kb, lane), L[kb], Y);
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb) Y = P::mma(lds_tile(sW, base + kb, lane), H[kb], Y);
    return Y;
}

}  // namespace mfma
}  // namespace nrrs
