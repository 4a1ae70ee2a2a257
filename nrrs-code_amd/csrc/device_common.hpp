// Device-side building blocks of the neural RRS stage (gfx950).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "host_common.hpp"

// No implicit FMA contraction in the device code that includes this header: the
// encoding and the epilogues then round identically in every kernel that inlines
// them (the per-net kernels, the persistent frame kernel, the training kernel), so
// their results agree bit for bit. Fused multiply-adds that are wanted are written
// out (__fmaf_rn); the MLPs run on MFMA.
#pragma clang fp contract(off)

namespace nrrs {

// ---------------------------------------------------------------- PCG32
// O'Neill (2014) pcg32, XSH-RR output, settable stream. Vertex i of a stream
// draws output number i, found by LCG jump-ahead (Brown 1994) so that every
// thread is independent of every other.
__device__ __forceinline__ uint32_t pcg_output(uint64_t old) {
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

// Jump-ahead table. The LCG step x -> a x + c with c = 2 stream + 1 (mod 2^64), taken k times,
// is x -> A(k) x + c G(k) with A(k) = a^k and G(k) = 1 + a + ... + a^(k-1), both independent of
// the stream. Entry [n][v] holds {A, G} for k = v 16^n, so a jump of `delta` steps composes one
// entry per hexadecimal digit of delta (Brown 1994's jump-ahead, taken a digit rather than a
// bit at a time; the maps are powers of one map, so the order of composition is free).
struct PcgAffine {
    uint64_t A, G;
};
struct PcgJumpTable {
    PcgAffine e[16][16];
};
constexpr PcgJumpTable make_pcg_jump_table() {
    PcgJumpTable t{};
    uint64_t bA = kPcgMult, bG = 1;  // one step
    for (int n = 0; n < 16; ++n) {
        uint64_t A = 1, G = 0;  // zero steps
        for (int v = 0; v < 16; ++v) {
            t.e[n][v] = PcgAffine{A, G};
            G = bA * G + bG;  // one more base step: 16^n steps
            A = bA * A;
        }
        bA = A;  // 16^(n+1) steps
        bG = G;
    }
    return t;
}
__constant__ const PcgJumpTable kPcgJump = make_pcg_jump_table();

// State after seeding (pcg32_srandom_r), jumped `delta` steps ahead.
__device__ __forceinline__ uint64_t pcg_state_at(uint64_t seed, uint64_t stream, uint64_t delta) {
    const uint64_t inc = (stream << 1u) | 1u;
    uint64_t state = inc;                     // state = 0; next()
    state = (state + seed) * kPcgMult + inc;  // state += seed; next()
    uint64_t A = 1, G = 0;
    for (int n = 0; delta; ++n, delta >>= 4u) {
        const PcgAffine e = kPcgJump.e[n][delta & 15u];
        G = e.A * G + e.G;
        A = e.A * A;
    }
    return A * state + inc * G;
}

__device__ __forceinline__ uint32_t pcg_next(uint64_t& state, uint64_t stream) {
    uint64_t old = state;
    state = old * kPcgMult + ((stream << 1u) | 1u);
    return pcg_output(old);
}

// 24-bit dyadic uniform in [0,1)
__device__ __forceinline__ float u24(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// ---------------------------------------------------------------- synthetic vertices
// Recipe of the benchmark workloads: fields on separate PCG32 streams of one seed;
// vertex i draws fixed output numbers of each stream, so vertices are independent.
//   stream 0: 64 triangles, 9 corner coordinates each, uniform in [0,1]
//   stream 1: (triangle = x0 & 63, barycentrics from u1, u2) at outputs 3i..3i+2
//   stream 2/3: direction / normal uniform on the sphere, outputs 2i, 2i+1
//   stream 4: roughness U[0.05, 1]
//   stream 5/6: diffuse / specular albedo U[0,1]^3, outputs 3i+k
//   stream 7: throughput log-normal(-1, 1) per channel, Box-Muller from 6i+2k, 6i+2k+1
struct SynthVertex {
    float pos[3], dir[3], nrm[3], rough, ad[3], as[3], thp[3];
};

// Position of synthetic vertex i (streams 0 and 1 of the recipe).
__device__ __forceinline__ void synth_position(uint64_t seed, uint64_t i, float* pos) {
    uint64_t st = pcg_state_at(seed, 1, 3ull * i);
    uint32_t a = pcg_next(st, 1), b = pcg_next(st, 1), c = pcg_next(st, 1);
    uint64_t tst = pcg_state_at(seed, 0, 9ull * (a & 63u));
    float T[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) T[k] = u24(pcg_next(tst, 0));
    float su = sqrtf(u24(b));
    float b0 = 1.0f - su, b1 = u24(c) * su, b2 = 1.0f - b0 - b1;
#pragma unroll
    for (int k = 0; k < 3; ++k) pos[k] = b0 * T[k] + b1 * T[3 + k] + b2 * T[6 + k];
}

// Unit vector of stream 2 (direction) or 3 (normal), outputs 2i, 2i+1.
__device__ __forceinline__ void synth_sphere(uint64_t seed, uint64_t stream, uint64_t i, float* dst) {
    const float two_pi = 6.283185307179586f;
    uint64_t st = pcg_state_at(seed, stream, 2ull * i);
    float u1 = u24(pcg_next(st, stream)), u2 = u24(pcg_next(st, stream));
    float z = 1.0f - 2.0f * u1;
    float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float ph = two_pi * u2;
    dst[0] = rr * cosf(ph);
    dst[1] = rr * sinf(ph);
    dst[2] = z;
}

// Albedo of stream 5 (diffuse) or 6 (specular), outputs 3i+k.
__device__ __forceinline__ void synth_albedo(uint64_t seed, uint64_t stream, uint64_t i, float* dst) {
    uint64_t st = pcg_state_at(seed, stream, 3ull * i);
#pragma unroll
    for (int k = 0; k < 3; ++k) dst[k] = u24(pcg_next(st, stream));
}

// Throughput channel k (stream 7, outputs 6i+2k and 6i+2k+1).
__device__ __forceinline__ float synth_thp(uint64_t seed, uint64_t i, uint32_t k) {
    const float two_pi = 6.283185307179586f;
    uint64_t st = pcg_state_at(seed, 7, 6ull * i + 2ull * k);
    float ua = u24(pcg_next(st, 7)), ub = u24(pcg_next(st, 7));
    float nrm = sqrtf(-2.0f * logf(1.0f - ua)) * cosf(two_pi * ub);
    return expf(-1.0f + nrm);
}

__device__ __forceinline__ float synth_rough(uint64_t seed, uint64_t i) {
    return 0.05f + 0.95f * u24(pcg_output(pcg_state_at(seed, 4, i)));
}

__device__ __forceinline__ SynthVertex synth_vertex(uint64_t seed, uint64_t i) {
    SynthVertex v;
    synth_position(seed, i, v.pos);
    synth_sphere(seed, 2, i, v.dir);
    synth_sphere(seed, 3, i, v.nrm);
    v.rough = synth_rough(seed, i);
    synth_albedo(seed, 5, i, v.ad);
    synth_albedo(seed, 6, i, v.as);
#pragma unroll
    for (int k = 0; k < 3; ++k) v.thp[k] = synth_thp(seed, i, k);
    return v;
}

// The non-position fields of synthetic vertex i drawn by the two halves of a wavefront that
// hold the same vertex in lanes l and l ^ 32 (h = lane half): half h draws stream 2 + h (direction
// / normal), 5 + h (diffuse / specular albedo) and throughput channels h and h + 2 (channel 3
// is drawn and dropped), so that both halves run the same code on their own streams; the
// other half's share comes over by lane exchange. Bitwise the fields of synth_vertex. Every
// lane of the wavefront must take part.
__device__ __forceinline__ void synth_fields_pair(uint64_t seed, uint64_t i, int h, float* dir, float* nrm,
                                                  float* ad, float* as, float* thp, float* rough) {
    float sph[3], alb[3];
    synth_sphere(seed, 2 + h, i, sph);
    synth_albedo(seed, 5 + h, i, alb);
    const float t0 = synth_thp(seed, i, (uint32_t)h), t1 = synth_thp(seed, i, (uint32_t)h + 2u);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float os = __shfl_xor(sph[k], 32, 64), oa = __shfl_xor(alb[k], 32, 64);
        dir[k] = h ? os : sph[k];
        nrm[k] = h ? sph[k] : os;
        ad[k] = h ? oa : alb[k];
        as[k] = h ? alb[k] : oa;
    }
    const float o0 = __shfl_xor(t0, 32, 64), o1 = __shfl_xor(t1, 32, 64);
    thp[0] = h ? o0 : t0;
    thp[1] = h ? t0 : o0;
    thp[2] = h ? o1 : t1;
    *rough = synth_rough(seed, i);
}

// The FP32 value x, rounded to FP32 before whatever conversion follows. Without it, the FP16
// conversion of a multiply-add result can be selected as one mixed-precision FMA that rounds
// once, straight to FP16 (v_fma_mixlo_f16), where another kernel inlining the same code gets an
// FP32 FMA and a conversion (two roundings, which differ in rare ties). Operands converted
// through this helper round the same way in every kernel.
__device__ __forceinline__ float f32_rounded(float x) {
    asm("" : "+v"(x));
    return x;
}

// ---------------------------------------------------------------- scalar maps
// activationSigma (src/render/nrrs/net/networkcommon.h:61-68): softplus for x<0,
// 0.5x + ln 2 above; C1 at 0.
__device__ __forceinline__ float activation_sigma(float x) {
    return (x < 0.0f) ? logf(1.0f + expf(x)) : (0.5f * x + 0.6931471805599453f);
}
__device__ __forceinline__ float activation_sigma_grad(float x) {
    if (x < 0.0f) {
        float e = expf(x);
        return e / (1.0f + e);
    }
    return 0.5f;
}

// box_cox_transform with lambda = 0.5 (src/render/nrrs/train.cu:19-39)
__device__ __forceinline__ float box_cox(float x) { return 2.0f * (sqrtf(x) - 1.0f); }

// warp_roughness (src/render/nrrs/train.h:95)
__device__ __forceinline__ float warp_roughness(float r) { return 1.0f - expf(-r); }

// normalizeSpatialCoord (src/render/nrc/nrctrain.h:111-115), clamped to [0,1]
__device__ __forceinline__ float norm_coord(float p, float c, float d) {
    float x = 0.5f + (p - c) / d;
    return fminf(fmaxf(x, 0.0f), 1.0f);
}

// ---------------------------------------------------------------- shipped-net encodings
// worldToLatLong (KiRaRay, called at device.cu:102 and workqueue.h:160; the library is not in
// the checkout): u = atan2(x, -z) / (2 pi) + 1/2, v = acos(y) / pi of the normalised direction.
__device__ __forceinline__ void lat_long(const float* d, float* uv) {
    const float inv = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const float x = d[0] * inv, y = d[1] * inv, z = d[2] * inv;
    uv[0] = atan2f(x, -z) * 0.15915494309189535f + 0.5f;
    uv[1] = acosf(fminf(fmaxf(y, -1.0f), 1.0f)) * 0.3183098861837907f;
}

// OneBlob encoding of x in [0,1] over `bins` bins (tiny-cuda-nn's "OneBlob", 2net-simple.json:
// 53-66): bin k holds the mass that a biweight kernel K(u) = 15/16 (1 - u^2)^2 of radius one bin
// width, centred at x and wrapped periodically onto [0,1), puts on [k / bins, (k+1) / bins).
// With C the kernel's CDF, C(u) = 1/2 + 15/16 u (1 - 2u^2/3 + u^4/5) on [-1,1] (0 below, 1
// above), bin k = sum over s in {-1,0,1} of C(k + 1 - t - s bins) - C(k - t - s bins), t = x bins.
// The kernel covers [t - 1, t + 1], so with m = floor(t) and f = t - m only three bins (mod bins)
// receive mass: m - 1 takes C(-f), m takes C(1 - f) - C(-f), m + 1 takes 1 - C(1 - f) (two
// evaluations of C instead of 3 (bins + 1)). The bins sum to 1.
__device__ __forceinline__ float quartic_cdf(float u) {
    u = fminf(fmaxf(u, -1.0f), 1.0f);
    const float u2 = u * u;
    return 0.5f + 0.9375f * u * (1.0f - u2 * (2.0f / 3.0f) + u2 * u2 * 0.2f);
}
template <int BINS>
__device__ __forceinline__ void one_blob(float x, float* o) {
    static_assert(BINS >= 4 && (BINS & (BINS - 1)) == 0, "three distinct bins, index wrap by mask");
    const float t = x * (float)BINS, fm = floorf(t), f = t - fm;
    const int m = (int)fm & (BINS - 1), ml = (m + BINS - 1) & (BINS - 1), mr = (m + 1) & (BINS - 1);
    const float c0 = quartic_cdf(-f), c1 = quartic_cdf(1.0f - f);
    const float w_l = c0, w_m = c1 - c0, w_r = 1.0f - c1;
#pragma unroll
    for (int k = 0; k < BINS; ++k) o[k] = k == ml ? w_l : (k == m ? w_m : (k == mr ? w_r : 0.0f));
}

// ---------------------------------------------------------------- SH degree 4
// Real spherical harmonics, bands 0..3 (Sloan 2008), written to o[0..15].
__device__ __forceinline__ void sh4(float x, float y, float z, float* o) {
    float xx = x * x, yy = y * y, zz = z * z;
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

// ---------------------------------------------------------------- hash grid
// Cell position of level l: fma(x, scale_l, 0.5), corner c = floor + bit.
struct LevelCell {
    uint32_t c[3];
    float f[3];
};

__device__ __forceinline__ LevelCell level_cell(const GridParams& g, int l, const float* x) {
    LevelCell lc;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float p = __fmaf_rn(x[k], g.scale[l], 0.5f);
        float fl = floorf(p);
        lc.c[k] = (uint32_t)fl;
        lc.f[k] = p - fl;
    }
    return lc;
}

// Index of corner `cn` (bit0 = x, bit1 = y, bit2 = z) within level l.
__device__ __forceinline__ uint32_t corner_index(const GridParams& g, int l, const LevelCell& lc,
                                                 int cn) {
    uint32_t cx = lc.c[0] + (cn & 1), cy = lc.c[1] + ((cn >> 1) & 1), cz = lc.c[2] + ((cn >> 2) & 1);
    if ((g.dense_mask >> l) & 1u) return cx + g.res[l] * (cy + g.res[l] * cz);
    // instant-NGP spatial hash (Mueller et al. 2022, eq. 4)
    return (cx ^ (cy * 2654435761u) ^ (cz * 805459861u)) & g.hash_mask;
}

__device__ __forceinline__ float corner_weight(const LevelCell& lc, int cn) {
    float wx = (cn & 1) ? lc.f[0] : 1.0f - lc.f[0];
    float wy = (cn & 2) ? lc.f[1] : 1.0f - lc.f[1];
    float wz = (cn & 4) ? lc.f[2] : 1.0f - lc.f[2];
    return (wx * wy) * wz;
}

// Record code of the x-adjacent corners e0 (x bit 0) and e1 (x bit 1) of one (y, z) of a cell,
// for the binned table-gradient scatter: 0..2 when e1 = e0 ^ (2^(code+1) - 1), so that one
// record {e0, code, both values} carries the pair (both entries then lie in one 8192-entry
// segment); 3 when the corners need a record each. Dense levels have e1 = e0 + 1, hashed levels
// e1 = e0 ^ (cx ^ (cx + 1)) (the x prime of the hash is 1): code 3 is left for about one pair
// in eight.
__device__ __forceinline__ uint32_t pair_code(uint32_t e0, uint32_t e1) {
    const uint32_t m = e0 ^ e1;
    return m == 1u ? 0u : m == 3u ? 1u : m == 7u ? 2u : 3u;
}

// Corner values of a level fetched as x-adjacent pairs. In a dense level the
// corners (x, y, z) and (x+1, y, z) are consecutive entries; in a hashed level
// with even x they are the aligned pair {h & ~1, h | 1} (the x prime of the
// spatial hash is 1). One 2-entry load then replaces two, which cuts the cache
// lines a vertex touches from 8 to 4..8 per level. Values are summed in corner
// order 0..7, exactly as the per-corner form.
template <typename TT>
struct PairLoad;
template <>
struct PairLoad<float> {  // FP32 table: entry = 2 floats
    typedef float T;
    static constexpr bool kQuad = false;
    __device__ __forceinline__ static float2 one(const float* t, uint32_t e) {
        return reinterpret_cast<const float2*>(t)[e];
    }
    __device__ __forceinline__ static void two(const float* t, uint32_t e, float2& a, float2& b) {
        float4 q;
        __builtin_memcpy(&q, t + 2 * (size_t)e, sizeof(q));
        a = make_float2(q.x, q.y);
        b = make_float2(q.z, q.w);
    }
};
template <typename Dec>
struct PairLoadLp {  // FP16/BF16 table: entry = one 32-bit word, decoded by Dec::pair
    typedef uint32_t T;
    static constexpr bool kQuad = true;  // hashed x-neighbours via aligned 4-entry loads
    __device__ __forceinline__ static float2 one(const uint32_t* t, uint32_t e) { return Dec::pair(t[e]); }
    __device__ __forceinline__ static void two(const uint32_t* t, uint32_t e, float2& a, float2& b) {
        uint2 q;
        __builtin_memcpy(&q, t + e, sizeof(q));
        a = Dec::pair(q.x);
        b = Dec::pair(q.y);
    }
    // Entries h0 and h1 of a level whose first entry is 16-byte aligned: one 16-byte load of
    // the aligned quad of h0, which also holds h1 when h0 ^ h1 < 4; else a second load.
    __device__ __forceinline__ static void x_pair(const uint32_t* t, uint32_t h0, uint32_t h1, float2& a, float2& b) {
        const uint4 q = *reinterpret_cast<const uint4*>(t + (h0 & ~3u));
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
        a = Dec::pair(w[h0 & 3u]);
        b = ((h0 ^ h1) < 4u) ? Dec::pair(w[h1 & 3u]) : Dec::pair(t[h1]);
    }
};

template <typename L>
__device__ __forceinline__ float2 grid_level_pairs(const typename L::T* __restrict__ table, const GridParams& g,
                                                   int l, const float* x) {
    LevelCell lc = level_cell(g, l, x);
    const uint32_t off = g.offset[l];
    float2 v[8];
    if ((g.dense_mask >> l) & 1u) {
        const uint32_t r = g.res[l];
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            const uint32_t cy = lc.c[1] + (yz & 1), cz = lc.c[2] + (yz >> 1);
            L::two(table, off + lc.c[0] + r * (cy + r * cz), v[2 * yz], v[2 * yz + 1]);
        }
    } else if constexpr (L::kQuad) {
        // x-neighbours h0, h1 = h0 ^ (cx ^ (cx+1)) share an aligned quad of entries for 3 cells
        // in 4 (the x prime of the hash is 1); level starts are 8-entry aligned
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            const uint32_t cy = lc.c[1] + (yz & 1), cz = lc.c[2] + (yz >> 1);
            const uint32_t yzh = (cy * 2654435761u) ^ (cz * 805459861u);
            const uint32_t h0 = (lc.c[0] ^ yzh) & g.hash_mask, h1 = ((lc.c[0] + 1) ^ yzh) & g.hash_mask;
            L::x_pair(table + off, h0, h1, v[2 * yz], v[2 * yz + 1]);
        }
    } else {
        const bool odd = lc.c[0] & 1u;
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            const uint32_t cy = lc.c[1] + (yz & 1), cz = lc.c[2] + (yz >> 1);
            const uint32_t yzh = (cy * 2654435761u) ^ (cz * 805459861u);
            const uint32_t h0 = (lc.c[0] ^ yzh) & g.hash_mask;
            float2 lo, hi;
            L::two(table, off + (h0 & ~1u), lo, hi);
            v[2 * yz] = (h0 & 1u) ? hi : lo;
            if (odd) {
                v[2 * yz + 1] = L::one(table, off + (((lc.c[0] + 1) ^ yzh) & g.hash_mask));
            } else {
                v[2 * yz + 1] = (h0 & 1u) ? lo : hi;
            }
        }
    }
    // corners with x bit 0 and with x bit 1 summed apart, in (y, z) order, then added: the order
    // of the two-lane gathers (grid_level_half), which give the same bits
    float2 a0 = make_float2(0.0f, 0.0f), a1 = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int yz = 0; yz < 4; ++yz) {
        const float w0 = corner_weight(lc, yz << 1), w1 = corner_weight(lc, (yz << 1) | 1);
        a0.x = __fmaf_rn(w0, v[2 * yz].x, a0.x);
        a0.y = __fmaf_rn(w0, v[2 * yz].y, a0.y);
        a1.x = __fmaf_rn(w1, v[2 * yz + 1].x, a1.x);
        a1.y = __fmaf_rn(w1, v[2 * yz + 1].y, a1.y);
    }
    return make_float2(a0.x + a1.x, a0.y + a1.y);
}

// Both nets' FP16/BF16 tables interleaved per entry (paired table): entry e is the 8 bytes
// {StatNet word, RRSNet word}, so one gather serves both nets. The x-adjacent entries of a dense
// level, and of a hashed level with even x (e1 = e0 ^ 1), come in one 16-byte load; the other
// hashed x-pairs add an 8-byte load for the second entry (half the lanes). That is 6 loads per
// hashed level for both nets against 2 x 5 with the per-net tables. Returns {StatNet f0, f1,
// RRSNet f0, f1}; each net's sum runs in corner order 0..7 with the same roundings as
// grid_level_pairs, so the values are bitwise those of the per-net tables.
template <typename Dec>
__device__ __forceinline__ float4 grid_level_pairs_dual(const uint2* __restrict__ table, const GridParams& g, int l,
                                                        const float* x) {
    LevelCell lc = level_cell(g, l, x);
    const uint2* t = table + g.offset[l];
    uint2 v[8];
    if ((g.dense_mask >> l) & 1u) {
        const uint32_t r = g.res[l];
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            const uint32_t cy = lc.c[1] + (yz & 1), cz = lc.c[2] + (yz >> 1);
            uint4 q;
            __builtin_memcpy(&q, t + lc.c[0] + r * (cy + r * cz), sizeof(q));  // 8-byte aligned
            v[2 * yz] = make_uint2(q.x, q.y);
            v[2 * yz + 1] = make_uint2(q.z, q.w);
        }
    } else {
        // level starts are 8-entry aligned, so the pair {h0 & ~1, h0 | 1} is 16-byte aligned
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            const uint32_t cy = lc.c[1] + (yz & 1), cz = lc.c[2] + (yz >> 1);
            const uint32_t yzh = (cy * 2654435761u) ^ (cz * 805459861u);
            const uint32_t h0 = (lc.c[0] ^ yzh) & g.hash_mask, h1 = ((lc.c[0] + 1) ^ yzh) & g.hash_mask;
            const uint4 q = *reinterpret_cast<const uint4*>(t + (h0 & ~1u));
            const bool odd = h0 & 1u;
            v[2 * yz] = make_uint2(odd ? q.z : q.x, odd ? q.w : q.y);
            uint2 other = make_uint2(odd ? q.x : q.z, odd ? q.y : q.w);  // entry h0 ^ 1
            if ((h0 ^ h1) != 1u) {  // global address space stated, else this load becomes a flat one
                typedef const __attribute__((address_space(1))) unsigned long long GlobalWords;
                const unsigned long long w = *(GlobalWords*)(t + h1);
                other = make_uint2((uint32_t)w, (uint32_t)(w >> 32));
            }
            v[2 * yz + 1] = other;
        }
    }
    // the corner order of grid_level_pairs
    float4 a0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), a1 = a0;
#pragma unroll
    for (int yz = 0; yz < 4; ++yz) {
        const float w0 = corner_weight(lc, yz << 1), w1 = corner_weight(lc, (yz << 1) | 1);
        const float2 s0 = Dec::pair(v[2 * yz].x), r0 = Dec::pair(v[2 * yz].y);
        const float2 s1 = Dec::pair(v[2 * yz + 1].x), r1 = Dec::pair(v[2 * yz + 1].y);
        a0.x = __fmaf_rn(w0, s0.x, a0.x);
        a0.y = __fmaf_rn(w0, s0.y, a0.y);
        a0.z = __fmaf_rn(w0, r0.x, a0.z);
        a0.w = __fmaf_rn(w0, r0.y, a0.w);
        a1.x = __fmaf_rn(w1, s1.x, a1.x);
        a1.y = __fmaf_rn(w1, s1.y, a1.y);
        a1.z = __fmaf_rn(w1, r1.x, a1.z);
        a1.w = __fmaf_rn(w1, r1.y, a1.w);
    }
    return make_float4(a0.x + a1.x, a0.y + a1.y, a0.z + a1.z, a0.w + a1.w);
}

// Two-lane form of the gathers: lanes 2k and 2k + 1 hold the same vertex and gather its corners
// with x bit xb = lane & 1, one 8-byte entry (both nets) or one 4-byte entry (one net) per (y, z),
// so that the two entries of an x-neighbour pair, which lie in one cache line in all but about
// one cell in 16 of a hashed level (their entries differ in the low bits only), come in one
// request of one load instruction. Returns the lane's partial sums over its four corners in
// (y, z) order; partner lanes add them as a0 + a1 (the order of grid_level_pairs). `table` is a
// paired table (DUAL) or a one-net table of 32-bit FP16/BF16 pairs.
template <typename Dec, bool DUAL>
__device__ __forceinline__ float4 grid_level_half(const uint32_t* __restrict__ table, const GridParams& g, int l,
                                                  const LevelCell& lc, uint32_t xb) {
    const bool dense = (g.dense_mask >> l) & 1u;
    const uint32_t cx = lc.c[0] + xb, r = g.res[l], off = g.offset[l];
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int yz = 0; yz < 4; ++yz) {
        const uint32_t cy = lc.c[1] + (yz & 1), cz = lc.c[2] + (yz >> 1);
        const uint32_t e = dense ? cx + r * (cy + r * cz) : (cx ^ (cy * 2654435761u) ^ (cz * 805459861u)) & g.hash_mask;
        const float w = corner_weight(lc, (int)xb | (yz << 1));
        if constexpr (DUAL) {
            const uint2 v = reinterpret_cast<const uint2*>(table)[off + e];
            const float2 s = Dec::pair(v.x), q = Dec::pair(v.y);
            acc.x = __fmaf_rn(w, s.x, acc.x);
            acc.y = __fmaf_rn(w, s.y, acc.y);
            acc.z = __fmaf_rn(w, q.x, acc.z);
            acc.w = __fmaf_rn(w, q.y, acc.w);
        } else {
            const float2 s = Dec::pair(table[off + e]);
            acc.x = __fmaf_rn(w, s.x, acc.x);
            acc.y = __fmaf_rn(w, s.y, acc.y);
        }
    }
    return acc;
}

// ---------------------------------------------------------------- encoding
// Network input layout (80 channels) of the benchmark nets (net_set 0):
//   [0,32)  hash grid, feature f of level l at 2l+f
//   [32,48) SH4(direction)      [48,64) SH4(normal)
//   StatNet [64] warp(rough) [65,68) albedo_d [68,71) albedo_s [71] 1 [72,80) 0
//   RRSNet  [64] warp(rough) [65,68) BC(thp) [68] BC(ref mean)
//           [69,72) BC(max(L,0)) or albedo_d   [72,75) BC(max(L2,0)) or albedo_s
//           [75] 1 [76,80) 0
// The constant channel stands in for the bias the fused MLP does not have
// (the reference injects 1 as its padding channel, train.cu:94-96).
// Shipped nets (net_set 1; 2net-simple.json:29-68, 210-291), the other channels zero:
//   StatNet [0, 2 levels) hash grid, [32,36) OneBlob(u), [36,40) OneBlob(v) of the lat-long
//           direction, [40,48) OneBlob(warp(rough), 8 bins)                    (32 inputs)
//   RRSNet  [64] warp(rough) [65,68) BC(thp) [68] BC(ref mean) [69,72) BC(max(L,0))
//           [72,75) BC(max(L2,0)) [75] 1 [76,80) 0 (NRRSNetworkInput1, no grid; 16 inputs)
//   RRSNet with rrs_aid: [0, 2 levels) hash grid, [32,40) OneBlob(u, v), [40,44)
//           OneBlob(warp(rough), 4 bins), [65,68) BC(thp), [68] BC(ref mean)   (32 inputs)
// The OneBlob sums give the StatNet and the AID RRSNet their constant.
__device__ __forceinline__ void load_xn(const EncodeArgs& a, uint32_t i, float* x) {
    const uint32_t S = a.v.stride;
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k] = norm_coord(a.v.pos[k * S + i], a.frame.c[k], a.frame.d[k]);
}

// One path vertex in registers (position already in the normalised frame).
struct VertexRec {
    float x[3], dir[3], nrm[3], rough, ad[3], as[3], thp[3];
    uint32_t pixel;
};

__device__ __forceinline__ VertexRec vertex_from_soa(const EncodeArgs& a, uint32_t i) {
    VertexRec v;
    const uint32_t S = a.v.stride;
    load_xn(a, i, v.x);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v.dir[k] = a.v.dir[k * S + i];
        v.nrm[k] = a.v.normal[k * S + i];
        v.ad[k] = a.v.albedo_d[k * S + i];
        v.as[k] = a.v.albedo_s[k * S + i];
        v.thp[k] = a.v.thp[k * S + i];
    }
    v.rough = a.v.rough[i];
    v.pixel = a.v.pixel[i];
    return v;
}

// Mean of the pixel estimate (film) of a pixel.
__device__ __forceinline__ float film_mean(const float* film, uint32_t n_pixels, uint32_t px) {
    return (film[px] + film[n_pixels + px] + film[2 * n_pixels + px]) * (1.0f / 3.0f);
}

// Non-grid channels [32,80) of a vertex. RRSNet: ref = pixel mean, st = StatNet
// statistics (L_rgb, L2_rgb) or null; sigma_l2: the L2 channels are sigma outputs and enter
// as activationSigma(max(L2, 0)) (the training input with train_sigma, train.cu:196-200).
// aid: RRSNet's AID input (NRRSNetworkInput1AID, train.h:58-68; train.cu:114-153, 242-267):
// position, direction, roughness, reference mean and throughput; the normal, albedo and
// statistics channels are zero.
__device__ __forceinline__ void encode_tail_rec(const VertexRec& v, int net, float ref, const float* st, float* e,
                                                bool sigma_l2 = false, bool aid = false, int set = NRRS_NETS_BENCH) {
    aid = aid && net == NRRS_RRSNET;
    if (set == NRRS_NETS_SHIPPED) {
#pragma unroll
        for (int k = 32; k < kEnc; ++k) e[k] = 0.0f;
        const float wr = warp_roughness(v.rough);
        if (net != NRRS_RRSNET || aid) {
            // NRRSNetworkInput0 (train.h:22-29, nrrs_generate_inference_data, train.cu:269-287) and
            // NRRSNetworkInput1AID (train.h:58-68, train.cu:114-153, 242-267)
            float uv[2];
            lat_long(v.dir, uv);
            one_blob<4>(uv[0], e + 32);
            one_blob<4>(uv[1], e + 36);
            if (aid) {
                one_blob<4>(wr, e + 40);
#pragma unroll
                for (int k = 0; k < 3; ++k) e[65 + k] = box_cox(v.thp[k]);
                e[68] = box_cox(ref);
            } else {
                one_blob<8>(wr, e + 40);
            }
        } else {
            // NRRSNetworkInput1 (train.h:39-56, nrrs_generate_inference_data_from_0to1,
            // train.cu:41-112; training: train.cu:155-240); without statistics they count as zero
            e[64] = wr;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e[65 + k] = box_cox(v.thp[k]);
                const float l = st ? fmaxf(st[k], 0.0f) : 0.0f, l2 = st ? fmaxf(st[3 + k], 0.0f) : 0.0f;
                e[69 + k] = box_cox(l);
                e[72 + k] = box_cox(sigma_l2 ? activation_sigma(l2) : l2);
            }
            e[68] = box_cox(ref);
            e[75] = 1.0f;
        }
        return;
    }
    sh4(v.dir[0], v.dir[1], v.dir[2], e + 32);
    if (aid) {
#pragma unroll
        for (int k = 48; k < 64; ++k) e[k] = 0.0f;
    } else {
        sh4(v.nrm[0], v.nrm[1], v.nrm[2], e + 48);
    }
#pragma unroll
    for (int k = 64; k < kEnc; ++k) e[k] = 0.0f;
    e[64] = warp_roughness(v.rough);
    if (net == NRRS_STATNET) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            e[65 + k] = v.ad[k];
            e[68 + k] = v.as[k];
        }
        e[71] = 1.0f;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) e[65 + k] = box_cox(v.thp[k]);
        e[68] = box_cox(ref);
        if (aid) {
            // channels 69-74 stay zero
        } else if (st) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e[69 + k] = box_cox(fmaxf(st[k], 0.0f));
                const float l2 = fmaxf(st[3 + k], 0.0f);
                e[72 + k] = box_cox(sigma_l2 ? activation_sigma(l2) : l2);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e[69 + k] = v.ad[k];
                e[72 + k] = v.as[k];
            }
        }
        e[75] = 1.0f;
    }
}

// Non-grid channels of vertex i of a SoA batch. SET: the network set when the caller knows it at
// compile time (the loads and channels of the other set are then dropped), else -1 (a.net_set).
template <int SET = -1>
__device__ __forceinline__ void encode_tail(const EncodeArgs& a, int net, uint32_t i, float* e) {
    const int set = SET >= 0 ? SET : a.net_set;
    VertexRec v = vertex_from_soa(a, i);
    const float ref = net == NRRS_RRSNET ? film_mean(a.film, a.n_pixels, v.pixel) : 0.0f;
    // two calls rather than a selected pointer, which would put st in memory
    if (net == NRRS_RRSNET && a.stats && !a.aid) {
        float st[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) st[k] = a.stats[(size_t)k * a.stats_stride + i];
        encode_tail_rec(v, net, ref, st, e, a.sigma_l2 != 0, false, set);
    } else {
        encode_tail_rec(v, net, ref, nullptr, e, false, a.aid != 0, set);
    }
}

template <typename TT>
__device__ __forceinline__ void encode_full(const EncodeArgs& a, int net, const TT* __restrict__ table,
                                            uint32_t i, float* e) {
    float x[3];
    load_xn(a, i, x);
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) {
        if (l < (int)a.grid.n_levels) {
            float2 f = grid_level_pairs<PairLoad<TT>>(table, a.grid, l, x);
            e[2 * l] = f.x;
            e[2 * l + 1] = f.y;
        } else {
            e[2 * l] = 0.0f;
            e[2 * l + 1] = 0.0f;
        }
    }
    encode_tail(a, net, i, e);
}

// ---------------------------------------------------------------- RRS epilogue
// q from the RRSNet raw output (integrator.cpp:1203-1219) or from StatNet L
// (integrator.cpp:1246-1255); optional weight window; clamp (integrator.cpp:1277).
__device__ __forceinline__ float rrs_window_clamp(const DecideParams& d, float r) {
    if (d.weight_window) {
        float dm = 2.0f / (1.0f + d.window_size), dp = dm * d.window_size;
        r = r < dm ? r / dm : (r > dp ? r / dp : 1.0f);
    }
    return fminf(fmaxf(r, d.clamp_min), d.clamp_max);
}

// Exact fixed-point term (2^-32 resolution) of the normalisation sum.
__device__ __forceinline__ uint64_t rrs_sum_term(const DecideParams& d, float q) {
    float t = d.use_ceil ? ceilf(q) : q;
    return (uint64_t)((double)t * 4294967296.0);
}

// Block-wide sum of a u64 (blockDim.x multiple of 64), one atomic per block.
__device__ __forceinline__ void block_atomic_add_u64(uint64_t v, unsigned long long* dst) {
    __shared__ unsigned long long part[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += part[w];
        if (s) atomicAdd(dst, s);
    }
}

// ---------------------------------------------------------------- scans
// Inclusive prefix sum over the 64 lanes of a wavefront.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(v, off, 64);
        if (lane >= off) v += y;
    }
    return v;
}

// Increments LDS counter ctr[b] once for every active lane and returns a distinct value in
// [old count, new count) to each lane (its rank among the lanes of that counter). When every
// active lane hits one counter, a single atomic adds them all: per-lane atomics on one LDS
// address are serialised.
__device__ __forceinline__ uint32_t lds_count(uint32_t* ctr, uint32_t b) {
    const uint64_t act = __ballot(1);
    const uint32_t b0 = __builtin_amdgcn_readfirstlane(b);
    if (__ballot(b == b0) == act) {
        const uint32_t lt = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
        uint32_t old = 0u;
        if (lt == 0u) old = atomicAdd(&ctr[b0], (uint32_t)__popcll(act));
        return __builtin_amdgcn_readfirstlane(old) + lt;
    }
    return atomicAdd(&ctr[b], 1u);
}

// Exclusive prefix over the block; *total receives the block sum. Contains a barrier.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) {
        uint32_t x = wsum[w];
        base += (w < wave) ? x : 0u;
        tot += x;
    }
    *total = tot;
    return base + inc - v;
}

}  // namespace nrrs
