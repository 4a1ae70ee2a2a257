// Training: fused forward, loss, backward and weight-gradient kernel on MFMA, for
// FP32 (v_mfma_f32_32x32x2_f32), FP16 and BF16 (v_mfma_f32_32x32x16_*) operands
// with FP32 accumulation.
//
// One workgroup = kWaves wavefronts = 32 kWaves vertices per iteration (persistent loop):
// 8 waves and 256 vertices for FP16/BF16, 4 waves and 128 vertices for FP32.
// Operand layout is the one of the inference kernel: lane l owns vertex l % 32
// and element j of every 16-wide K chunk with k = 8(l/32) + j; layer outputs are
// chained without shuffles through the permuted column order
// p(kb,h,j) = 16kb + 8(j/4) + 4h + j%4.
//
// Backward sweep, layer l = L-1 .. 0:
//   1. recompute the layer input h_l from the encoding, or from the input of layer 2 kept
//      from the forward (only the encoding, that checkpoint and one layer of activations
//      are ever live: activations are recomputed, not stored);
//   2. stage delta_l and h_l of the workgroup's vertices in LDS as [neuron][vertex];
//   3. dW_l += delta_l h_l^T on MFMA with the vertices as the contraction index,
//      each wave owning a fixed subset of the 32x32 dW tiles in its registers for
//      the whole kernel (written once per workgroup, reduced deterministically);
//   4. delta_{l-1} = (W_l^T delta_l) * [h_l > 0] on MFMA with pre-packed W^T tiles;
//      at l = 0 the 32 grid channels of dL/denc are scatter-added into the table (FP32)
//      or written out for the binned scatter k_bin_levels (FP16/BF16).
#include "device_common.hpp"
#include "mfma_common.hpp"
#include "train_common.hpp"

namespace nrrs {
namespace {

using mfma::kHidChunks;
using mfma::kOutTiles;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// 16-bit operands: ReLU on the rounded values (a negative FP16/BF16 value has its sign bit set,
// so as an int16 it is negative; rounding keeps the sign, so this is round(max(x, 0)) for every
// finite x), and delta * [h > 0] for ReLU outputs h (h > 0 iff its bits are nonzero; a positive
// input that rounds to zero, below 2^-25 in FP16, counts as inactive).
template <typename OP>
__device__ __forceinline__ OP relu16(OP v) {
    return __builtin_bit_cast(OP, __builtin_elementwise_max(__builtin_bit_cast(i16x8, v), i16x8{}));
}
// 0xffff in each 16-bit half of h that is nonzero: min(half, 1), then 0 - that. Written as
// packed 16-bit VALU operations in asm, since the compiler otherwise rewrites the mask into
// per-element FP32 compares and selects.
__device__ __forceinline__ uint32_t nonzero16_mask(uint32_t h) {
    uint32_t m;
    asm("v_pk_min_u16 %0, %1, %2\n\tv_pk_sub_u16 %0, 0, %0" : "=&v"(m) : "v"(h), "s"(0x00010001u));
    return m;
}
template <typename OP>
__device__ __forceinline__ OP mask16(OP d, OP h) {
    u32x4 dd = __builtin_bit_cast(u32x4, d);
    const u32x4 hh = __builtin_bit_cast(u32x4, h);
#pragma unroll
    for (int k = 0; k < 4; ++k) dd[k] &= nonzero16_mask(hh[k]);
    return __builtin_bit_cast(OP, dd);
}

// Operand policies: 8 values per lane per 16-deep K chunk.
// kWaves: waves per workgroup. FP16/BF16 use 8 (one workgroup of 8 waves per CU, two
// waves per SIMD, 256 VGPRs per lane); FP32 uses 4, since its FP32 staging of 8 waves
// would not fit in LDS.
struct TBF16 {
    static constexpr bool kF32Table = false;
    static constexpr int kWaves = 8;
    typedef bf16x8 op;
    typedef __bf16 elem;
    __device__ static op pack(const float* x) {
        op v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (__bf16)x[j];
        return v;
    }
    __device__ static op relu_pack(const float* x) { return relu16(pack(x)); }
    __device__ static op mask_by(op d, op h) { return mask16(d, h); }
    __device__ static float get(const op& v, int j) { return (float)v[j]; }
    __device__ static f32x16 mma16(op a, op b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static float2 pair(uint32_t u) {
        return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
    }
};
struct TF16 {
    static constexpr bool kF32Table = false;
    static constexpr int kWaves = 8;
    typedef f16x8 op;
    typedef _Float16 elem;
    __device__ static op pack(const float* x) {
        op v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (_Float16)f32_rounded(x[j]);
        return v;
    }
    __device__ static op relu_pack(const float* x) { return relu16(pack(x)); }
    __device__ static op mask_by(op d, op h) { return mask16(d, h); }
    __device__ static float get(const op& v, int j) { return (float)v[j]; }
    __device__ static f32x16 mma16(op a, op b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    __device__ static float2 pair(uint32_t u) {
        _Float16 lo = __builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu));
        _Float16 hi = __builtin_bit_cast(_Float16, (uint16_t)(u >> 16));
        return make_float2((float)lo, (float)hi);
    }
};
struct TF32 {
    static constexpr bool kF32Table = true;
    static constexpr int kWaves = 4;
    typedef f32x8 op;
    typedef float elem;
    __device__ static op pack(const float* x) {
        op v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = x[j];
        return v;
    }
    __device__ static op relu_pack(const float* x) {
        op v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = fmaxf(x[j], 0.0f);
        return v;
    }
    __device__ static op mask_by(op d, op h) {
#pragma unroll
        for (int j = 0; j < 8; ++j) d[j] = h[j] > 0.0f ? d[j] : 0.0f;
        return d;
    }
    __device__ static float get(const op& v, int j) { return v[j]; }
    // A 16-deep chunk as 8 K=2 MFMAs: step s pairs k = s (lanes 0-31) with k = 8 + s
    // (lanes 32-63), so every k of the chunk is contracted once.
    __device__ static f32x16 mma16(op a, op b, f32x16 c) {
#pragma unroll
        for (int s = 0; s < 8; ++s) c = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], b[s], c, 0, 0, 0);
        return c;
    }
};

constexpr uint32_t kDenseReplicas = kDenseGradReplicas;
constexpr int kStageRowsD = 64;        // delta rows
constexpr int kStageRowsH = 96;        // input rows (80 padded to 3 x 32)

// FP16/BF16 operands stage vertex-major (below); the kernels of 32-wide nets on them are the
// narrow kernels.
template <typename P>
constexpr bool kTrStage = sizeof(typename P::elem) == 2;
template <typename P, int W>
constexpr bool kNarrow = W == 32 && kTrStage<P>;

// Workgroup shape. The narrow kernels run workgroups of 4 waves, two per CU: each needs 35 KB of
// LDS (their weight tiles are read from global memory, where they stay in L1), and the two
// workgroups' independent barriers let one hide the latency of the other's global loads.
template <typename P, int W = 64>
struct Geo {
    static constexpr int kWaves = kNarrow<P, W> ? 4 : P::kWaves;
    static constexpr int kThreads = 64 * kWaves;
    static constexpr int kVert = 32 * kWaves;   // vertices per workgroup iteration
    static constexpr int kStageLd = kVert + 8;  // LDS row stride (elements), breaks bank aliasing
};

// FP16/BF16 operands stage vertex-major ([vertex][neuron], neurons contiguous), so that a lane
// writes the neurons of its vertex with 8- and 16-byte stores, and the dW MFMAs read their
// operands with the transposing LDS load ds_read_b64_tr_b16 (stage_tr). FP32 operands stage
// neuron-major ([neuron][vertex]) and read 8 consecutive vertices of a row (stage_op).
// Rows of the vertex-major stages: 64 neurons (channels 0-63 of the layer-0 input) padded to 68
// elements; channels 64-95 of the layer-0 input (80-95 zero) in a second stage padded to 36. With
// rows of 34 and 18 dwords, the 8-byte stores of 32 consecutive rows (one per lane) start on 32
// different even banks of a 64-bank LDS, and rows 8 apart start 16 banks apart, so the four rows
// that one group of 16 lanes reads with a transposing load (stage_tr) are taken 8 vertices apart:
// k-step s of the dW contraction covers vertices 32 (s / 2) + 4 (s % 2) + 8j + 2h + r (lane half
// h, load r, element j).
constexpr int kTrRow = 68, kTrRow2 = 36;  // elements per row of the main and the second stage
__host__ __device__ constexpr int tr_kstep_row(int s) { return 32 * (s >> 1) + 4 * (s & 1); }

template <typename P, bool TR = kTrStage<P>>
struct Stage {
    typename P::elem d[kStageRowsD * Geo<P>::kStageLd];
    typename P::elem h[kStageRowsH * Geo<P>::kStageLd];
};
template <typename P>
struct Stage<P, true> {
    typename P::elem d[Geo<P>::kVert * kTrRow];
    typename P::elem h[Geo<P>::kVert * kTrRow];
    typename P::elem h2[Geo<P>::kVert * kTrRow2];
};
// Narrow kernels: the delta stage holds 32 neurons in rows of kTrRow2 elements.
template <typename P>
struct Stage32 {
    typename P::elem d[Geo<P, 32>::kVert * kTrRow2];
    typename P::elem h[Geo<P, 32>::kVert * kTrRow];
    typename P::elem h2[Geo<P, 32>::kVert * kTrRow2];
};
template <typename P, int W>
using StageOf = std::conditional_t<kNarrow<P, W>, Stage32<P>, Stage<P>>;
// weight tiles in LDS: the 8-wave FP16/BF16 kernels of the 64-wide nets
template <typename P, int W>
constexpr bool kTilesInLds = P::kWaves == 8 && W == 64;
// dW tiles of the narrow kernels: tile u = ti for input tile ti of layer 0 (0-2), u = l + 2 for
// the one tile of layer l > 0; tile u on wave u % waves, in register slot u / waves.
template <int NL>
constexpr int kNarrowTiles = NL + 2;

// Number of 32x32 dW tiles of layer l (out rows x in cols).
__host__ __device__ constexpr int dw_tiles(int l, int nl) {
    return (l == 0 ? 2 * 3 : (l == nl - 1 ? 1 * 2 : 2 * 2));
}
// dW tiles are numbered across layers (layer 0 first); wave w of W owns the tiles g
// with g % W == w and accumulates tile g in register slot g / W.
__host__ __device__ constexpr int dw_tile_base(int l, int nl) {
    int s = 0;
    for (int k = 0; k < l; ++k) s += dw_tiles(k, nl);
    return s;
}
__host__ __device__ constexpr int dw_slots(int nl, int w) { return (dw_tile_base(nl, nl) + w - 1) / w; }
// First tile at or after `base` owned by wave w of W.
__device__ __forceinline__ int dw_first(int base, int w, int W) { return base + ((w - base) % W + W) % W; }

template <typename P>
__device__ __forceinline__ typename P::op load_op(const typename P::op* __restrict__ W, int tile, int lane) {
    return W[tile * 64 + lane];
}

// LDS operand: 8 consecutive vertices of row `row`, starting at vertex v0.
template <typename P>
__device__ __forceinline__ typename P::op stage_op(const typename P::elem* s, int row, int v0) {
    const typename P::elem* p = s + row * Geo<P>::kStageLd + v0;
    typename P::op v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = p[j];
    return v;
}

// LDS operand of the vertex-major stages: lane l gets neuron n0 + l % 32 of the 8 vertices of
// its lane half h = l / 32 in one k-step (the A/B operand layout of v_mfma_f32_32x32x16 with the
// k-step's vertices as k), in two ds_read_b64_tr_b16. That load treats each group of 16 lanes as
// a unit: lane i of the group receives element i % 4 of the 4-element (8-byte) loads of lanes
// 4j + i / 4, j = 0..3 (checked on the device with an exact-integer probe). So lane p of the group
// loads the four neurons n0 + 16 (group % 2) + 4 (p % 4) .. +3 of the vertex of element j = p / 4,
// and lane i ends with neuron n0 + 16 (group % 2) + i of the vertices of elements 0..3 (load r
// gives operand elements 4r..4r+3). `s` points at the k-step's first row and the tile's first
// neuron; off[r] is the lane's offset (tr_lane_offsets).
typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) i16x4 lds_i16x4;
template <typename P>
__device__ __forceinline__ typename P::op stage_tr(const typename P::elem* s, int off0, int off1) {
    const i16x4 r0 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_i16x4*)(s + off0));
    const i16x4 r1 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_i16x4*)(s + off1));
    typedef short i16x8 __attribute__((ext_vector_type(8)));
    const i16x8 r = {r0[0], r0[1], r0[2], r0[3], r1[0], r1[1], r1[2], r1[3]};
    return __builtin_bit_cast(typename P::op, r);
}
// The lane's offsets for loads r = 0, 1 in a stage of `len` elements per row: the row of element
// j = (l % 16) / 4 of load r in lane half h (8j + 2h + r after the k-step's first row) and the
// lane's four neurons 16 (group % 2) + 4 (l % 4).
__device__ __forceinline__ int tr_lane_offset(int lane, int r, int len) {
    return (8 * ((lane & 15) >> 2) + 2 * (lane >> 5) + r) * len + 16 * ((lane >> 4) & 1) + 4 * (lane & 3);
}

// Stores of one operand of 8 16-bit values into row `row` of a vertex-major stage (len elements
// per row): elements 0-3 at column c0, 4-7 at column c1 (8-byte stores).
template <typename P>
__device__ __forceinline__ void stage_put(typename P::elem* s, int len, int row, const typename P::op& v, int c0,
                                          int c1) {
    const u32x4 w = __builtin_bit_cast(u32x4, v);
    typename P::elem* r = s + row * len;
    *reinterpret_cast<uint2*>(r + c0) = make_uint2(w[0], w[1]);
    *reinterpret_cast<uint2*>(r + c1) = make_uint2(w[2], w[3]);
}

// Forward of hidden layer `layer` (1..NL-2) or layer 0 from operands X (5 chunks
// for layer 0, 4 otherwise) into fp32 accumulators D[2] (D[0] only for hidden width W = 32,
// which takes of layer 0 the K chunks set in l0).
template <typename P, int NL, int W = 64>
__device__ __forceinline__ void fwd_layer(const typename P::op* __restrict__ Wf, int layer, const typename P::op* X,
                                          f32x16* D, int lane, uint32_t l0 = 0x1fu) {
#pragma unroll
    for (int ot = 0; ot < kOutTiles<W>; ++ot) {
        D[ot] = f32x16{};
        if (layer == 0) {
#pragma unroll
            for (int kb = 0; kb < 5; ++kb)
                if (W == 64 || ((l0 >> kb) & 1u)) D[ot] = P::mma16(load_op<P>(Wf, ot * 5 + kb, lane), X[kb], D[ot]);
        } else {
            const int base = 10 + (layer - 1) * 8;
#pragma unroll
            for (int kb = 0; kb < kHidChunks<W>; ++kb)
                D[ot] = P::mma16(load_op<P>(Wf, base + ot * 4 + kb, lane), X[kb], D[ot]);
        }
    }
}

// Activation of accumulators into operands. ReLU: its outputs double as the mask [h > 0] of
// the backward (P::mask_by). LeakyReLU (slope kLeakySlope): x or slope * x, rounded once.
template <typename P, int ACT, int W = 64>
__device__ __forceinline__ void relu_ops(const f32x16* D, typename P::op* H) {
#pragma unroll
    for (int ot = 0; ot < kOutTiles<W>; ++ot) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = D[ot][8 * g + j];
            if constexpr (ACT == ACT_LEAKY) {
#pragma unroll
                for (int j = 0; j < 8; ++j) x[j] = x[j] > 0.0f ? x[j] : x[j] * kLeakySlope;
                H[2 * ot + g] = P::pack(x);
            } else {
                H[2 * ot + g] = P::relu_pack(x);
            }
        }
    }
}

// delta of a hidden layer from its FP32 accumulators x8 and the layer's activations H:
// x * [h > 0] (ReLU), x * (h > 0 ? 1 : slope) (LeakyReLU, whose output has the sign of its input).
template <typename P, int ACT>
__device__ __forceinline__ typename P::op act_mask(float* x8, const typename P::op& H) {
    if constexpr (ACT == ACT_LEAKY) {
#pragma unroll
        for (int j = 0; j < 8; ++j) x8[j] = P::get(H, j) > 0.0f ? x8[j] : x8[j] * kLeakySlope;
        return P::pack(x8);
    } else {
        return P::mask_by(P::pack(x8), H);
    }
}

// Input of layer `upto` (> from): forward through layers from..upto-1, starting from X, the
// input of layer `from` (the encoding when from = 0).
template <typename P, int NL, int ACT, int W = 64>
__device__ __forceinline__ void forward_to(const typename P::op* __restrict__ Wf, int from, const typename P::op* X,
                                           int upto, typename P::op* H, int lane, uint32_t l0 = 0x1fu) {
    f32x16 D[2];
    fwd_layer<P, NL, W>(Wf, from, X, D, lane, l0);
    for (int layer = from + 1; layer < upto; ++layer) {
        relu_ops<P, ACT, W>(D, H);
        fwd_layer<P, NL, W>(Wf, layer, H, D, lane, l0);
    }
    relu_ops<P, ACT, W>(D, H);
}

// The input of this layer is kept from the forward pass (a checkpoint) for the backward sweep:
// the inputs of the layers above it are recomputed from it, those below from the encoding. Of
// the recompute the StatNet (5 layers) then runs 2 layers per iteration instead of 6, the
// RRSNet (4 layers) 1 instead of 3. 0: no checkpoint.
template <int NL>
constexpr int kCkpt = NL >= 4 ? 2 : 0;

// BIN: dL/denc of the grid goes to t.gdl for the binned scatter (FP16/BF16); otherwise it is
// added to the table with atomics. Separate instances keep the unused path's registers out.
// W: hidden width (64, or 32 for the shipped nets: only the live tiles are computed; their
// encoding is that of the shipped network set).
template <typename P, int NL, int NOUT, int LOSS, bool BIN, int ACT, int W>
__global__ __launch_bounds__((64 * Geo<P, W>::kWaves)) __attribute__((amdgpu_waves_per_eu(kNarrow<P, W> ? 2 : 1)))
void k_train(TrainArgs t) {
    static_assert(!(BIN && P::kF32Table), "the binned scatter is for FP16/BF16 tables");
    static_assert(W == 64 || W == 32, "hidden width");
    const uint32_t l0 = t.a.l0_chunks;  // layer-0 input chunks that hold the net's inputs
    typedef typename P::op op;
    constexpr int kWaves = Geo<P, W>::kWaves, kThreads = Geo<P, W>::kThreads, kVert = Geo<P, W>::kVert;
    constexpr int kStageLd = Geo<P>::kStageLd;
    constexpr bool kN = kNarrow<P, W>;
    constexpr int kDRow = kN ? kTrRow2 : kTrRow;  // elements per row of the FP16/BF16 delta stage
    constexpr int kNet = NOUT == 1 ? NRRS_RRSNET : NRRS_STATNET;
    constexpr int kSet = W == 32 ? NRRS_NETS_SHIPPED : NRRS_NETS_BENCH;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typedef StageOf<P, W> St;
    St& st = *reinterpret_cast<St*>(smem_raw);
    __shared__ double loss_part[kWaves];  // loss per wave (FP64, summed in a fixed order)

    // the wave index in a scalar register: tile ownership and slot choices below are then
    // scalar branches
    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    // FP16/BF16: the packed forward and W^T tiles live in LDS behind the stage (they are read
    // by every layer of every iteration); FP32: from global memory (cache)
    constexpr int kTilesF = 10 + (NL - 2) * 8 + 4, kTilesB = 4 + (NL - 2) * 8 + 2;
    op* sW = reinterpret_cast<op*>(smem_raw + sizeof(St));
    if constexpr (kTilesInLds<P, W>) {
        const op* gF = reinterpret_cast<const op*>(t.w_fwd);
        const op* gB = reinterpret_cast<const op*>(t.w_bwd);
        for (int k = threadIdx.x; k < (kTilesF + kTilesB) * 64; k += kThreads)
            sW[k] = k < kTilesF * 64 ? gF[k] : gB[k - kTilesF * 64];
    }
    const op* Wf0 = kTilesInLds<P, W> ? sW : reinterpret_cast<const op*>(t.w_fwd);
    const op* Wb0 = kTilesInLds<P, W> ? sW + kTilesF * 64 : reinterpret_cast<const op*>(t.w_bwd);
    const float S = *t.loss_scale;
    const float ebar = t.error_sum_dev ? fmaxf(*t.error_sum_dev / t.n_pixels_f, 1e-4f) : t.ebar;
    constexpr int kSlots = kN ? (kNarrowTiles<NL> + kWaves - 1) / kWaves : dw_slots(NL, kWaves);
    f32x16 acc[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) acc[s] = f32x16{};
    if (lane == 0) loss_part[wave] = 0.0;  // by the wave that accumulates into it

    // zero the padded input channels once (layer 0 has 80 of 96)
    if constexpr (kTrStage<P>) {
        for (int k = threadIdx.x; k < kTrRow2 * kVert; k += kThreads) st.h2[k] = (typename P::elem)0.0f;
    } else {
        for (int k = threadIdx.x; k < 16 * kStageLd; k += kThreads) st.h[80 * kStageLd + k] = (typename P::elem)0.0f;
    }
    __syncthreads();  // weight tiles in LDS before the first forward

    const uint32_t n_iter = (t.n + kVert - 1) / kVert;
    for (uint32_t it = blockIdx.x; it < n_iter; it += gridDim.x) {
        const op* Wf = Wf0;
        const op* Wb = Wb0;
        const uint32_t i = it * kVert + wave * 32 + col;
        const bool live = i < t.n;
        const int vcol = wave * 32 + col;  // vertex column in the stage

        // ---- the iteration's inputs. Every load is issued here, before any is used: the film
        // gather through the pixel index is the only dependent one (loaded where the encoding and
        // the loss use them, behind their branches and arithmetic, the groups of loads had been
        // round trips of their own).
        VertexRec v{};
        float st6[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};  // RRSNet: StatNet statistics
        float lin[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // dL/dy (EXTERNAL) or target L (LL2) of outputs 4h + r
        float rq_qt = -1.0f, rq_e = 0.0f, rq_w = 0.0f, rq_refm = 0.0f;  // NRRS_RRS
        op E[5];
        if (live) {
            if constexpr (!P::kF32Table) {
                // FP16/BF16 grid channels: the level-major encoding computed before this kernel;
                // lane half h holds channels 16kb + 8h + j, levels 8kb + 4h + {0..3}
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) {
                    u32x4 u;
#pragma unroll
                    for (int ll = 0; ll < 4; ++ll) {
                        const uint32_t l = 8 * kb + 4 * h + ll;
                        u[ll] = l < t.a.grid.n_levels ? t.a.grid_enc[(size_t)l * t.a.grid_enc_stride + i] : 0u;
                    }
                    E[kb] = __builtin_bit_cast(op, u);
                }
            }
            v = vertex_from_soa(t.a, i);
            if constexpr (kNet == NRRS_RRSNET) {
                if (t.a.stats) {
#pragma unroll
                    for (int k = 0; k < 6; ++k) st6[k] = t.a.stats[(size_t)k * t.a.stats_stride + i];
                }
            }
            if constexpr (LOSS == NRRS_LOSS_EXTERNAL) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (4 * h + r < NOUT) lin[r] = t.dy_ext[(size_t)(4 * h + r) * t.n + i];
            } else if constexpr (LOSS == NRRS_LOSS_LL2) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (4 * h + r < NOUT) lin[r] = t.target_L[(size_t)((4 * h + r) % 3) * t.a.v.stride + i];
            } else if constexpr (LOSS == NRRS_LOSS_RRS) {
                if (t.rrs_step >= 3) {
                    rq_qt = t.q_target[i];
                    rq_e = t.error[i];
                    rq_w = t.sample_weight[i];
                    if (t.ref_mean) rq_refm = t.ref_mean[i];
                }
            }
        }
        // mean of the pixel estimate: RRSNet input channel and the NRRS_RRS loss's reference
        const float pix_ref = live && kNet == NRRS_RRSNET ? film_mean(t.a.film, t.a.n_pixels, v.pixel) : 0.0f;
        const float rq_ref = t.ref_mean ? rq_refm : pix_ref;

        // ---- encoding (B layout, natural channel order): lane half h holds channels 16kb + 8h + j
        {
            float e8[8];
            if (live) {
                if constexpr (P::kF32Table) {
                    // FP32: the grid gathered here, each lane half only its 8 levels
                    float x[3];
                    load_xn(t.a, i, x);
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
                        for (int ll = 0; ll < 4; ++ll) {
                            const int l = 8 * kb + 4 * h + ll;
                            float2 f = make_float2(0.0f, 0.0f);
                            if (l < (int)t.a.grid.n_levels)
                                f = grid_level_pairs<PairLoad<float>>(t.table_f32, t.a.grid, l, x);
                            e8[2 * ll] = f.x;
                            e8[2 * ll + 1] = f.y;
                        }
                        E[kb] = P::pack(e8);
                    }
                }
                // the non-grid channels (as encode_tail, from the values loaded above)
                float e[kEnc];
                if (kNet == NRRS_RRSNET && t.a.stats && !t.a.aid)
                    encode_tail_rec(v, kNet, pix_ref, st6, e, t.a.sigma_l2 != 0, false, kSet);
                else
                    encode_tail_rec(v, kNet, pix_ref, nullptr, e, false, t.a.aid != 0, kSet);
#pragma unroll
                for (int kb = 2; kb < 5; ++kb) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) e8[j] = h ? e[kb * 16 + 8 + j] : e[kb * 16 + j];
                    E[kb] = P::pack(e8);
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) e8[j] = 0.0f;
#pragma unroll
                for (int kb = 0; kb < 5; ++kb) E[kb] = P::pack(e8);
            }
        }

        // ---- forward to the output, loss, dL/dy
        op H[4];
        constexpr int C = kCkpt<NL>;
        op Hc[4];  // input of layer C (the checkpoint)
        if constexpr (C > 0) {
            forward_to<P, NL, ACT, W>(Wf, 0, E, C, Hc, lane, l0);
            if constexpr (NL - 1 > C) {
                forward_to<P, NL, ACT, W>(Wf, C, Hc, NL - 1, H, lane, l0);
            } else {
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) H[kb] = Hc[kb];
            }
        } else {
            forward_to<P, NL, ACT, W>(Wf, 0, E, NL - 1, H, lane, l0);
        }
        f32x16 Y = f32x16{};
        {
            const int base = 10 + (NL - 2) * 8;
#pragma unroll
            for (int kb = 0; kb < kHidChunks<W>; ++kb) Y = P::mma16(load_op<P>(Wf, base + kb, lane), H[kb], Y);
        }
        // output c sits in lane half c/4, register c%4
        float dyr[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float lterm = 0.0f;  // this lane's loss terms of the iteration
        // With sigma outputs, output c >= 3 needs L output c - 3: c = 3 sits in this lane (half 0),
        // c = 4, 5 in lane half 1 need outputs 1, 2 of the other half
        float y_mu[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (LOSS == NRRS_LOSS_LL2 && NOUT == 6) {
#pragma unroll
            for (int r = 0; r < 3; ++r) y_mu[r] = __shfl_xor(Y[r], 32);
        }
        if (live) {
            if (LOSS == NRRS_LOSS_EXTERNAL) {
#pragma unroll
                for (int r = 0; r < 4; ++r) dyr[r] = lin[r];
            } else if (LOSS == NRRS_LOSS_LL2) {
                const int active = t.ll2_step == 0 ? 3 : 6;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = 4 * h + r;
                    if (c < NOUT && c < active && (t.head < 0 || c / 3 == t.head)) {
                        float tg = fmaxf(lin[r], 0.0f);
                        if (t.clamp_on) tg = fminf(tg, t.clamp_max);
                        if (c >= 3) tg = tg * tg;
                        const float y = Y[r];
                        if (c >= 3 && t.train_sigma) {
                            // second moment sigma^2 + mu^2 against L^2, mu (output c - 3) held
                            // constant: the gradient reaches only the sigma output
                            // (indices static in r: a lane-dependent index would put y_mu in memory)
                            const float mu = h == 0 ? Y[0] : (r == 0 ? y_mu[1] : y_mu[2]);
                            const float sg = activation_sigma(y);
                            const float m2 = sg * sg + mu * mu;
                            const float rden = 1.0f / (m2 * m2 + 0.01f);  // one division for both terms
                            const float d = m2 - tg;
                            lterm += d * d * rden * t.inv_norm;
                            dyr[r] = 2.0f * d * rden * 2.0f * sg * activation_sigma_grad(y) * t.inv_norm;
                        } else {
                            const float rden = 1.0f / (y * y + 0.01f);
                            const float d = y - tg;
                            lterm += d * d * rden * t.inv_norm;
                            dyr[r] = 2.0f * d * rden * t.inv_norm;
                        }
                    }
                }
            } else if (LOSS == NRRS_LOSS_RRS) {
                if (h == 0 && t.rrs_step >= 3) {
                    const float qt = rq_qt;
                    if (qt >= 0.0f) {
                        const float ref = rq_ref;
                        float m2 = 0.0f;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const float th = v.thp[c];
                            const float s2 = st6[3 + c];
                            float l2;
                            if (t.train_sigma) {  // E[L^2] = sigma^2 + max(L, 0)^2
                                const float sg = activation_sigma(s2);
                                const float mu = fmaxf(st6[c], 0.0f);
                                l2 = sg * sg + mu * mu;
                            } else {
                                l2 = fmaxf(s2, 0.0f);
                            }
                            m2 += th * th * l2;
                        }
                        float vv = m2 * (1.0f / 3.0f) / (ref * ref + 0.01f);
                        if (t.clamp_on) vv = fminf(vv, t.clamp_max);
                        const float e = fmaxf(rq_e, 1e-4f);
                        const float y = Y[0];
                        const float q = activation_sigma(y);
                        const float qe = q + 0.01f;
                        const float w = rq_w;
                        const float rq = 1.0f / qe, vq = vv * rq;  // v / (q + 0.01), one division
                        const float f = t.gamma[0] * (vq + q * ebar) + t.gamma[1] * (vq + q * e) +
                                        t.gamma[2] * q + t.gamma[3] * (q - qt) * (q - qt);
                        const float fp = -(t.gamma[0] + t.gamma[1]) * vq * rq + t.gamma[0] * ebar +
                                         t.gamma[1] * e + t.gamma[2] + 2.0f * t.gamma[3] * (q - qt);
                        lterm += w * f * t.inv_norm;
                        dyr[0] = w * fp * activation_sigma_grad(y) * t.inv_norm;
                    }
                }
            }
            if (t.y_out) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = 4 * h + r;
                    if (c < NOUT && (t.head < 0 || c / 3 == t.head)) t.y_out[(size_t)c * t.out_stride + i] = Y[r];
                }
            }
        }
        if constexpr (LOSS != NRRS_LOSS_EXTERNAL) {
            // FP64 butterfly over the wave (every lane ends with the same total), added by lane 0
            // to the wave's running sum: the batch loss does not depend on scheduling
            double ls = (double)lterm;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) ls += __shfl_xor(ls, o, 64);
            if (lane == 0) loss_part[wave] += ls;
        }
        // dy operand (k = c = 8h + j): half 0 holds outputs 0..7, fetching 4..7 from half 1
        op DY;
        {
            float x8[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float other = __shfl_xor(dyr[r], 32, 64);
                x8[r] = h ? 0.0f : dyr[r] * S;
                x8[4 + r] = h ? 0.0f : other * S;
            }
            DY = P::pack(x8);
        }

        // ---- backward sweep
        op Dl[4];  // delta of the current layer as operands (k-chunks of its output neurons)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) Dl[kb] = DY;  // only Dl[0] is used for the output layer
#pragma unroll
        for (int layer = NL - 1; layer >= 0; --layer) {
            const bool last = layer == NL - 1;
            // (1) layer input h_l and its ReLU mask
            if (!last && layer > 0) {
                if (C > 0 && layer > C) {
                    forward_to<P, NL, ACT, W>(Wf, C, Hc, layer, H, lane, l0);
                } else if (C > 0 && layer == C) {
#pragma unroll
                    for (int kb = 0; kb < 4; ++kb) H[kb] = Hc[kb];
                } else {
                    forward_to<P, NL, ACT, W>(Wf, 0, E, layer, H, lane, l0);
                }
            }
            // (2) stage delta_l rows [out neuron][vertex] and h_l rows [in neuron][vertex]
            __syncthreads();  // previous layer's dW reads are done
            if constexpr (kTrStage<P>) {
                // vertex-major: operand element j of chunk kb is neuron 16kb + 8(j/4) + 4h + j%4
                // (columns 16kb + 4h and 16kb + 8 + 4h), of the encoding channel 16kb + 8h + j
                // (columns 16kb + 8h and 16kb + 8h + 4; chunk 4 in the second stage)
                if (last) {
                    if (h == 0) stage_put<P>(st.d, kDRow, vcol, Dl[0], 0, 4);  // outputs 0-7
                } else {
#pragma unroll
                    for (int kb = 0; kb < kHidChunks<W>; ++kb)
                        stage_put<P>(st.d, kDRow, vcol, Dl[kb], 16 * kb + 4 * h, 16 * kb + 8 + 4 * h);
                }
                if (layer == 0) {
#pragma unroll
                    for (int kb = 0; kb < 4; ++kb)
                        stage_put<P>(st.h, kTrRow, vcol, E[kb], 16 * kb + 8 * h, 16 * kb + 8 * h + 4);
                    stage_put<P>(st.h2, kTrRow2, vcol, E[4], 8 * h, 8 * h + 4);
                } else {
#pragma unroll
                    for (int kb = 0; kb < kHidChunks<W>; ++kb)
                        stage_put<P>(st.h, kTrRow, vcol, H[kb], 16 * kb + 4 * h, 16 * kb + 8 + 4 * h);
                }
            } else {
                if (last) {
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        if (h == 0 && j < NOUT) st.d[j * kStageLd + vcol] = (typename P::elem)P::get(Dl[0], j);
                } else {
#pragma unroll
                    for (int kb = 0; kb < kHidChunks<W>; ++kb)
#pragma unroll
                        for (int j = 0; j < 8; ++j)
                            st.d[(16 * kb + 8 * (j >> 2) + 4 * h + (j & 3)) * kStageLd + vcol] =
                                (typename P::elem)P::get(Dl[kb], j);
                }
                if (layer == 0) {
#pragma unroll
                    for (int kb = 0; kb < 5; ++kb)
#pragma unroll
                        for (int j = 0; j < 8; ++j)
                            st.h[(kb * 16 + 8 * h + j) * kStageLd + vcol] = (typename P::elem)P::get(E[kb], j);
                } else {
#pragma unroll
                    for (int kb = 0; kb < kHidChunks<W>; ++kb)
#pragma unroll
                        for (int j = 0; j < 8; ++j)
                            st.h[(16 * kb + 8 * (j >> 2) + 4 * h + (j & 3)) * kStageLd + vcol] =
                                (typename P::elem)P::get(H[kb], j);
                }
            }
            __syncthreads();
            // (3) dW tiles of this layer owned by this wave
            if constexpr (kN) {
                // narrow: this layer's tiles u owned by this wave (layer 0: those of input
                // chunks that hold inputs)
                for (int u = layer == 0 ? 0 : layer + 2; u < (layer == 0 ? 3 : layer + 3); ++u) {
                    const int ti = layer == 0 ? u : 0;
                    if (u % kWaves != wave || (layer == 0 && ((l0 >> (2 * ti)) & 3u) == 0)) continue;
                    int ln = lane;
                    asm volatile("" : "+v"(ln));
                    const int tro[2] = {tr_lane_offset(ln, 0, kTrRow), tr_lane_offset(ln, 1, kTrRow)};
                    const int tro2[2] = {tr_lane_offset(ln, 0, kTrRow2), tr_lane_offset(ln, 1, kTrRow2)};
#pragma unroll
                    for (int s2 = 0; s2 < kSlots; ++s2) {
                        if (s2 == u / kWaves) {
#pragma unroll
                            for (int s = 0; s < kVert / 16; ++s) {
                                const int row = tr_kstep_row(s);
                                const op A = stage_tr<P>(st.d + row * kTrRow2, tro2[0], tro2[1]);
                                const op B = ti < 2 ? stage_tr<P>(st.h + row * kTrRow + 32 * ti, tro[0], tro[1])
                                                    : stage_tr<P>(st.h2 + row * kTrRow2, tro2[0], tro2[1]);
                                acc[s2] = P::mma16(A, B, acc[s2]);
                            }
                        }
                    }
                }
            } else {
                const int n_ti = layer == 0 ? 3 : 2;
                const int base = dw_tile_base(layer, NL), n_tiles = dw_tiles(layer, NL);
                for (int g = dw_first(base, wave, kWaves); g < base + n_tiles; g += kWaves) {
                    const int tt = g - base, slot = g / kWaves;
                    const int to = tt / n_ti, ti = tt % n_ti;
                    // W = 32: the tiles outside the net (and of layer 0 those of input chunks
                    // without inputs) are zero and stay so
                    if (W == 32 && !(to == 0 && (layer == 0 ? ((l0 >> (2 * ti)) & 3u) != 0 : ti == 0))) continue;
                    // offsets of the transposing loads (FP16/BF16 stages), recomputed per tile
                    // (cheaper than keeping four registers through the whole kernel)
                    int ln = lane;
                    asm volatile("" : "+v"(ln));
                    [[maybe_unused]] const int tro[2] = {tr_lane_offset(ln, 0, kTrRow), tr_lane_offset(ln, 1, kTrRow)};
                    [[maybe_unused]] const int tro2[2] = {tr_lane_offset(ln, 0, kTrRow2),
                                                          tr_lane_offset(ln, 1, kTrRow2)};
                    // the MFMAs accumulate straight into the tile's register slot; the slot is
                    // wave-uniform, so its choice is a scalar branch
#pragma unroll
                    for (int s2 = 0; s2 < kSlots; ++s2) {
                        if (s2 == slot) {
#pragma unroll
                            for (int s = 0; s < kVert / 16; ++s) {
                                op A, B;
                                if constexpr (kTrStage<P>) {
                                    const int row = tr_kstep_row(s);
                                    A = stage_tr<P>(st.d + row * kDRow + 32 * to, tro[0], tro[1]);
                                    B = ti < 2 ? stage_tr<P>(st.h + row * kTrRow + 32 * ti, tro[0], tro[1])
                                               : stage_tr<P>(st.h2 + row * kTrRow2, tro2[0], tro2[1]);
                                } else {
                                    A = stage_op<P>(st.d, to * 32 + col, 16 * s + 8 * h);
                                    B = stage_op<P>(st.h, ti * 32 + col, 16 * s + 8 * h);
                                }
                                acc[s2] = P::mma16(A, B, acc[s2]);
                            }
                        }
                    }
                }
            }
            // (4) propagate delta to the layer input
            if (layer > 0) {
                f32x16 D[2];
#pragma unroll
                for (int ot = 0; ot < kOutTiles<W>; ++ot) {
                    D[ot] = f32x16{};
                    if (last) {
                        D[ot] = P::mma16(load_op<P>(Wb, t.bwd_last_base + ot, lane), Dl[0], D[ot]);
                    } else {
                        const int base = t.bwd_hidden_base + (layer - 1) * 8;
#pragma unroll
                        for (int kb = 0; kb < kHidChunks<W>; ++kb)
                            D[ot] = P::mma16(load_op<P>(Wb, base + ot * 4 + kb, lane), Dl[kb], D[ot]);
                    }
                }
                // delta_{l-1} = D * [h_l > 0]
#pragma unroll
                for (int ot = 0; ot < kOutTiles<W>; ++ot)
#pragma unroll
                    for (int g = 0; g < 2; ++g) {
                        float x8[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) x8[j] = D[ot][8 * g + j];
                        Dl[2 * ot + g] = act_mask<P, ACT>(x8, H[2 * ot + g]);
                    }
            } else if (t.a.grid.n_levels) {  // (a net without a grid has no table gradient)
                // dL/denc for the 32 grid channels: rows c = 8(r/4) + 4h + r%4
                f32x16 G = f32x16{};
#pragma unroll
                for (int kb = 0; kb < kHidChunks<W>; ++kb) G = P::mma16(load_op<P>(Wb, kb, lane), Dl[kb], G);
                if constexpr (BIN) {
                    // Binned scatter (FP16/BF16): the level pairs of dL/denc go to HBM as
                    // [level][vertex] FP16 pairs (the precision of the records they become) and
                    // k_bin_levels turns them into records. Lane (h, col) holds channels
                    // 8(r/4) + 4h + r%4, so registers r and r + 1 (r even) are features 0 and 1 of
                    // level 4(r/4) + 2h + (r%4)/2.
                    if (live) {
#pragma unroll
                        for (int r = 0; r < 16; r += 2) {
                            const uint32_t l = 4 * (r >> 2) + 2 * h + ((r & 3) >> 1);
                            if (l < t.a.grid.n_levels)
                                t.gdl[(size_t)l * t.gdl_stride + i] =
                                    __builtin_bit_cast(uint32_t, h2v{(_Float16)G[r], (_Float16)G[r + 1]});
                        }
                    }
                } else {
                    // Direct scatter to the table. The level gradients are staged in LDS as
                    // [grid channel][vertex] (the delta rows are free once every wave is past its
                    // dW tiles). Lanes 2k and 2k+1 then take vertex k of this wave's 32, so that the
                    // two lanes of a pair update neighbouring table words in the same atomic
                    // instruction, which the memory pipeline merges into one request:
                    //  - FP16 table: lane parity = x bit of the corner; the x-adjacent corners are
                    //    consecutive entries in a dense level and the aligned pair {h & ~1, h | 1} of
                    //    a hashed level when the cell's x is even;
                    //  - FP32 table: lane parity = feature, both features of a corner are one 8-byte
                    //    entry.
                    __syncthreads();
                    // [32][kVert] in the delta stage, or in the input stage of the narrow kernels
                    static_assert((kN ? sizeof(st.h) : sizeof(st.d)) >= 32 * kVert * sizeof(float),
                                  "grid-gradient stage");
                    float* gst = reinterpret_cast<float*>(kN ? (void*)st.h : (void*)st.d);
#pragma unroll
                    for (int r = 0; r < 16; ++r) gst[(8 * (r >> 2) + 4 * h + (r & 3)) * kVert + vcol] = G[r];
                    // the values a lane reads were written by its own wave
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const int pv = lane >> 1, pb = lane & 1;
                    const uint32_t iv = it * kVert + wave * 32 + pv;
                    const bool live_v = iv < t.n;
                    float x[3] = {0.0f, 0.0f, 0.0f};
                    if (live_v) load_xn(t.a, iv, x);
#pragma unroll 1
                    for (int l = 0; l < (int)t.a.grid.n_levels; ++l) {
                        const float g0 = gst[(2 * l) * kVert + wave * 32 + pv];
                        const float g1 = gst[(2 * l + 1) * kVert + wave * 32 + pv];
                        const bool dense = (t.a.grid.dense_mask >> l) & 1u;
                        if (!live_v) continue;
                        LevelCell lc = level_cell(t.a.grid, l, x);
                        // Dense (coarse) levels are small and hit by every vertex: their updates go
                        // to one of kDenseReplicas private copies (chosen by workgroup), summed before
                        // the optimiser step, which spreads the contention on each entry.
                        const size_t rep = dense ? (size_t)(blockIdx.x % kDenseReplicas) * t.n_dense : 0;
                        if (t.grad_grid_h2) {
#pragma unroll
                            for (int yz = 0; yz < 4; ++yz) {
                                const int cn = pb | (yz << 1);
                                const uint32_t ent = t.a.grid.offset[l] + corner_index(t.a.grid, l, lc, cn);
                                const float w = corner_weight(lc, cn);
                                h2v gv = {(_Float16)(w * g0), (_Float16)(w * g1)};  // one packed FP16 atomic
                                h2v* dst = dense ? t.dense_h2 + rep + ent : t.grad_grid_h2 + ent;
                                __builtin_amdgcn_global_atomic_fadd_v2f16(dst, gv);
                            }
                        } else {
                            const float g = pb ? g1 : g0;
#pragma unroll
                            for (int cn = 0; cn < 8; ++cn) {
                                const uint32_t ent = t.a.grid.offset[l] + corner_index(t.a.grid, l, lc, cn);
                                const float w = corner_weight(lc, cn);
                                float* dst = dense ? t.dense_f32 + 2 * (rep + ent) : t.grad_grid + 2 * (size_t)ent;
                                atomicAdd(dst + pb, w * g);
                            }
                        }
                    }
                }
            }
        }
    }

    // ---- per-workgroup dW partials (rows < out, cols < in) and the loss
    float* part = t.dw_partial + (size_t)blockIdx.x * t.n_mlp;
    if constexpr (kN) {
        // the wave's narrow tiles; partials outside them stay zero (cleared once at allocation)
        for (int u = wave; u < kNarrowTiles<NL>; u += kWaves) {
            const int layer = u < 3 ? 0 : u - 2, ti = u < 3 ? u : 0;
            const int in = layer == 0 ? kEnc : kHidden;
            const int out = layer == NL - 1 ? NOUT : kHidden;
            const int woff = layer == 0 ? 0 : kHidden * kEnc + (layer - 1) * kHidden * kHidden;
            f32x16 a = f32x16{};
#pragma unroll
            for (int s2 = 0; s2 < kSlots; ++s2)
                if (s2 == u / kWaves) a = acc[s2];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * h + (r & 3);
                const int cc = ti * 32 + col;
                if (row < out && cc < in)
                    part[woff + row * in + cc] = head_masked(t.head_first, layer, NL, row, cc) ? 0.0f : a[r];
            }
        }
    }
#pragma unroll
    for (int layer = 0; layer < (kN ? 0 : NL); ++layer) {
        const int n_ti = layer == 0 ? 3 : 2;
        const int n_tiles = dw_tiles(layer, NL);
        const int in = layer == 0 ? kEnc : kHidden;
        const int out = layer == NL - 1 ? NOUT : kHidden;
        const int woff = layer == 0 ? 0 : kHidden * kEnc + (layer - 1) * kHidden * kHidden;
        const int base = dw_tile_base(layer, NL);
        for (int g = dw_first(base, wave, kWaves); g < base + n_tiles; g += kWaves) {
            const int tt = g - base, slot = g / kWaves;
            const int to = tt / n_ti, ti = tt % n_ti;
            f32x16 a = f32x16{};
#pragma unroll
            for (int s2 = 0; s2 < kSlots; ++s2)
                if (s2 == slot) a = acc[s2];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = to * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                const int cc = ti * 32 + col;
                if (row < out && cc < in)
                    part[woff + row * in + cc] = head_masked(t.head_first, layer, NL, row, cc) ? 0.0f : a[r];
            }
        }
    }
    __syncthreads();
    if (LOSS != NRRS_LOSS_EXTERNAL && threadIdx.x == 0 && t.loss_partial) {
        double s = 0.0;
        for (int w = 0; w < kWaves; ++w) s += loss_part[w];
        t.loss_partial[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------- packing
// Forward tiles: identical to the inference packing (A operand of W_l).
// Backward tiles (A operand of W_l^T): [0,4) layer 0 -> 32 grid channels (rows)
// x 64 outputs (K, permuted); then hidden layers 1..L-2: 2 row tiles x 4 K chunks
// (rows = input neuron, K = output neuron permuted); then the output layer:
// 2 row tiles x 1 K chunk (K = output c = 8h + j, natural).
__device__ __forceinline__ int pcol(int kb, int h, int j) { return 16 * kb + 8 * (j >> 2) + 4 * h + (j & 3); }

template <typename E>
__device__ __forceinline__ void pack_train_tile(const float* __restrict__ W, int nl, int n_out, E* __restrict__ fwd,
                                                E* __restrict__ bwd, int n_fwd, int n_bwd, int tile, int lane) {
    const int h = lane >> 5, r32 = lane & 31;
    auto woff = [&](int layer) { return layer == 0 ? 0 : kHidden * kEnc + (layer - 1) * kHidden * kHidden; };
    if (tile < n_fwd) {
        int layer, ot, kb;
        if (tile < 10) {
            layer = 0;
            ot = tile / 5;
            kb = tile % 5;
        } else {
            const int u = tile - 10;
            layer = 1 + u / 8;
            if (layer <= nl - 2) {
                ot = (u % 8) / 4;
                kb = u % 4;
            } else {
                layer = nl - 1;
                ot = 0;
                kb = u - (nl - 2) * 8;
            }
        }
        const int in = layer == 0 ? kEnc : kHidden;
        const int rows = layer == nl - 1 ? n_out : kHidden;
        const int row = ot * 32 + r32;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = layer == 0 ? kb * 16 + 8 * h + j : pcol(kb, h, j);
            const float w = row < rows ? W[woff(layer) + row * in + c] : 0.0f;
            fwd[((size_t)tile * 64 + lane) * 8 + j] = (E)w;
        }
        return;
    }
    const int bt = tile - n_fwd;
    if (bt >= n_bwd) return;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float w = 0.0f;
        if (bt < 4) {  // layer 0: row = grid channel r32, K = output neuron pcol(bt, h, j)
            w = W[woff(0) + pcol(bt, h, j) * kEnc + r32];
        } else if (bt < 4 + (nl - 2) * 8) {
            const int u = bt - 4, layer = 1 + u / 8, ot = (u % 8) / 4, kb = u % 4;
            w = W[woff(layer) + pcol(kb, h, j) * kHidden + ot * 32 + r32];
        } else {
            const int ot = bt - 4 - (nl - 2) * 8;
            const int c = 8 * h + j;
            w = c < n_out ? W[woff(nl - 1) + c * kHidden + ot * 32 + r32] : 0.0f;
        }
        bwd[((size_t)bt * 64 + lane) * 8 + j] = (E)w;
    }
}

// One wave per tile.
template <typename E>
__global__ void k_pack_train(const float* __restrict__ W, int nl, int n_out, E* __restrict__ fwd, E* __restrict__ bwd,
                             int n_fwd, int n_bwd) {
    pack_train_tile<E>(W, nl, n_out, fwd, bwd, n_fwd, n_bwd, blockIdx.x, threadIdx.x);
}

// After an optimizer step of an FP16/BF16 net, one wave per workgroup: workgroup 0 updates the
// step and loss-scale state, the next n_inf the inference tiles of the MLP (P's layout, with
// the split-precision residual tiles where the net has them), then, with a moving average
// W_ema, n_inf more of its inference tiles, the rest the training tiles (E) of the masters W.
// Three dependent small launches become one.
template <typename P, typename E>
__global__ void k_after_adam(AdamArgs a, const float* __restrict__ W, const float* __restrict__ W_ema, int nl,
                             int n_out, int split, uint16_t* __restrict__ inf, uint16_t* __restrict__ inf_ema,
                             int n_inf, E* __restrict__ fwd, E* __restrict__ bwd, int n_fwd, int n_bwd) {
    int b = (int)blockIdx.x - 1;
    if (b < 0) {
        if (threadIdx.x == 0) adam_finish_state(a);
        return;
    }
    if (b < n_inf) {
        mfma::pack_mlp_tile<P>(W, nl, n_out, split, inf, b, threadIdx.x);
        return;
    }
    b -= n_inf;
    if (W_ema) {
        if (b < n_inf) {
            mfma::pack_mlp_tile<P>(W_ema, nl, n_out, split, inf_ema, b, threadIdx.x);
            return;
        }
        b -= n_inf;
    }
    pack_train_tile<E>(W, nl, n_out, fwd, bwd, n_fwd, n_bwd, b, threadIdx.x);
}

template <typename P, int NL, int NOUT, int LOSS, bool BIN, int ACT, int W>
hipError_t launch_train_one_bin(const TrainArgs& t, uint32_t blocks, hipStream_t s) {
    constexpr size_t smem =
        sizeof(StageOf<P, W>) + (kTilesInLds<P, W> ? (size_t)(10 + (NL - 2) * 8 + 4 + 4 + (NL - 2) * 8 + 2) * 64 *
                                                         sizeof(typename P::op)
                                                   : 0);
    static std::atomic<bool> attr[kMaxDevices];
    hipError_t e =
        set_dyn_lds_once(attr, reinterpret_cast<const void*>(&k_train<P, NL, NOUT, LOSS, BIN, ACT, W>), smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_train<P, NL, NOUT, LOSS, BIN, ACT, W>), dim3(blocks), dim3(Geo<P, W>::kThreads), smem, s,
                       t);
    return hipGetLastError();
}

template <typename P, int NL, int NOUT, int LOSS, int ACT = ACT_RELU, int W = 64>
hipError_t launch_train_one(const TrainArgs& t, uint32_t blocks, hipStream_t s) {
    if constexpr (P::kF32Table) {
        if (t.gdl) return hipErrorInvalidValue;
        return launch_train_one_bin<P, NL, NOUT, LOSS, false, ACT, W>(t, blocks, s);
    } else {
        return t.gdl ? launch_train_one_bin<P, NL, NOUT, LOSS, true, ACT, W>(t, blocks, s)
                     : launch_train_one_bin<P, NL, NOUT, LOSS, false, ACT, W>(t, blocks, s);
    }
}

}  // namespace

uint32_t train_vertices_per_iter(int precision, int net_set) {
    if (precision == NRRS_FP32) return Geo<TF32>::kVert;
    return net_set == NRRS_NETS_SHIPPED ? Geo<TBF16, 32>::kVert : Geo<TBF16>::kVert;
}

namespace {

template <typename P>
hipError_t dispatch_train(const TrainArgs& t, int nl, int n_out, int loss, uint32_t blocks, hipStream_t s) {
    if (t.a.net_set == NRRS_NETS_SHIPPED) {
        // the shipped nets: 3 hidden 32-wide LeakyReLU layers, StatNet -> 6, RRSNet -> 1
        if (nl == 4 && n_out == 6) {
            if (loss == NRRS_LOSS_EXTERNAL)
                return launch_train_one<P, 4, 6, NRRS_LOSS_EXTERNAL, ACT_LEAKY, 32>(t, blocks, s);
            if (loss == NRRS_LOSS_LL2) return launch_train_one<P, 4, 6, NRRS_LOSS_LL2, ACT_LEAKY, 32>(t, blocks, s);
        }
        if (nl == 4 && n_out == 1) {
            if (loss == NRRS_LOSS_EXTERNAL)
                return launch_train_one<P, 4, 1, NRRS_LOSS_EXTERNAL, ACT_LEAKY, 32>(t, blocks, s);
            if (loss == NRRS_LOSS_RRS) return launch_train_one<P, 4, 1, NRRS_LOSS_RRS, ACT_LEAKY, 32>(t, blocks, s);
        }
        return hipErrorInvalidValue;
    }
    if (nl == 5 && n_out == 6) {
        if (loss == NRRS_LOSS_EXTERNAL) return launch_train_one<P, 5, 6, NRRS_LOSS_EXTERNAL>(t, blocks, s);
        if (loss == NRRS_LOSS_LL2) return launch_train_one<P, 5, 6, NRRS_LOSS_LL2>(t, blocks, s);
    }
    // two-headed StatNets: PartShare (4 layers), JustEncoding (3 layers)
    if (nl == 4 && n_out == 6) {
        if (loss == NRRS_LOSS_EXTERNAL) return launch_train_one<P, 4, 6, NRRS_LOSS_EXTERNAL>(t, blocks, s);
        if (loss == NRRS_LOSS_LL2) return launch_train_one<P, 4, 6, NRRS_LOSS_LL2>(t, blocks, s);
    }
    if (nl == 3 && n_out == 6) {
        if (loss == NRRS_LOSS_EXTERNAL) return launch_train_one<P, 3, 6, NRRS_LOSS_EXTERNAL>(t, blocks, s);
        if (loss == NRRS_LOSS_LL2) return launch_train_one<P, 3, 6, NRRS_LOSS_LL2>(t, blocks, s);
    }
    if (nl == 4 && n_out == 1) {
        if (loss == NRRS_LOSS_EXTERNAL) return launch_train_one<P, 4, 1, NRRS_LOSS_EXTERNAL>(t, blocks, s);
        if (loss == NRRS_LOSS_RRS) return launch_train_one<P, 4, 1, NRRS_LOSS_RRS>(t, blocks, s);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------- binned scatter
// The encode pass of the backward counted the corners of every (level, kRunVert vertices,
// segment), one 256-byte row of counts per chunk. Two kernels turn these counts into the first
// record slot of every chunk's run (chunk order within each segment) and each segment's total,
// over tiles of kScanRows chunks (wave g of a workgroup takes rows g * kScanRowsPerWave.., lane
// s segment s): k_bin_tiles sums each tile; k_bin_scan adds the sums of the level's earlier
// tiles, scans its own rows and writes the running slots. Slots are clamped to the segment's
// end, so that a full segment (clustered input) sends the rest of its records to direct
// atomics.
constexpr int kScanWaves = 16, kScanRowsPerWave = 16, kScanRows = kScanWaves * kScanRowsPerWave;
static_assert(kScanRows == (int)kBinScanRows, "tile size");

__global__ __launch_bounds__(64 * kScanWaves) void k_bin_tiles(BinArgs b, uint32_t n_tiles) {
    __shared__ uint32_t part[kScanWaves][kMaxSegBins];
    const uint32_t l = blockIdx.x / n_tiles, tile = blockIdx.x % n_tiles, s = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t r0 = tile * kScanRows + w * kScanRowsPerWave;
    const uint32_t* r = b.runs + ((size_t)l * b.n_runs) * kMaxSegBins + s;
    uint32_t sum = 0u;
#pragma unroll
    for (int k = 0; k < kScanRowsPerWave; ++k)
        if (r0 + k < b.n_runs) sum += r[(size_t)(r0 + k) * kMaxSegBins];
    part[w][s] = sum;
    __syncthreads();
    if (w == 0) {
        uint32_t tot = 0u;
#pragma unroll
        for (int k = 0; k < kScanWaves; ++k) tot += part[k][s];
        b.tile_sums[((size_t)l * n_tiles + tile) * kMaxSegBins + s] = tot;
    }
}

__global__ __launch_bounds__(64 * kScanWaves) void k_bin_scan(BinArgs b, uint32_t n_tiles) {
    static_assert(kMaxSegBins == 64, "one lane per segment");
    __shared__ uint32_t part[kScanWaves][kMaxSegBins];
    const uint32_t l = blockIdx.x / n_tiles, tile = blockIdx.x % n_tiles, s = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t r0 = tile * kScanRows + w * kScanRowsPerWave;
    uint32_t* r = b.runs + ((size_t)l * b.n_runs) * kMaxSegBins + s;
    uint32_t v[kScanRowsPerWave], sum = 0u;
#pragma unroll
    for (int k = 0; k < kScanRowsPerWave; ++k) {
        v[k] = r0 + k < b.n_runs ? r[(size_t)(r0 + k) * kMaxSegBins] : 0u;
        sum += v[k];
    }
    uint64_t base = 0u;  // records of this segment in the level's earlier tiles
    for (uint32_t t = 0; t < tile; ++t) base += b.tile_sums[((size_t)l * n_tiles + t) * kMaxSegBins + s];
    part[w][s] = sum;
    __syncthreads();
    uint64_t before = base, total = base;
#pragma unroll
    for (int k = 0; k < kScanWaves; ++k) {
        before += k < (int)w ? part[k][s] : 0u;
        total += part[k][s];
    }
    if (s >= b.seg_nbins[l]) return;  // rows of segments the level does not have stay zero
    const uint32_t seg = b.seg_base[l] + s;
    const SegInfo sg = b.segs[seg];
    const uint64_t end = (uint64_t)sg.rec_off + sg.cap;
    uint64_t run = sg.rec_off + before;
#pragma unroll
    for (int k = 0; k < kScanRowsPerWave; ++k) {
        if (r0 + k < b.n_runs) r[(size_t)(r0 + k) * kMaxSegBins] = (uint32_t)(run < end ? run : end);
        run += v[k];
    }
    if (w == 0 && tile == n_tiles - 1) b.seg_count[seg] = (uint32_t)total;
}

// k_bin_levels: workgroup b takes level b / chunks and kBinVert vertices (one run of the scan),
// so the GPU walks the levels one after the other. Each thread takes one vertex: per (y, z) of its
// cell, one record for the x-adjacent corner pair or, for pair code 3, one per corner:
//  1. every record gets its rank within its segment from LDS atomics;
//  2. wave 0, lane s = segment s: an exclusive scan of the counts places the segment's records
//     in the workgroup's segment-ordered list; the workgroup's slots of segment s start at the
//     first slot of its first run;
//  3. each record {entry in level, code, vertex and corner} is put into the LDS list at
//     (segment start + rank);
//  4. the list is copied out with consecutive threads on consecutive records, so the stores of
//     a wave form a few contiguous runs rather than one cache line per lane; each record's value
//     (FP16 w * dL/denc of its corners) is formed there from its vertex's cell fractions and
//     dL/denc, staged in LDS (a 2-byte index per record instead of the 8-byte value keeps the
//     list, and so the workgroup's LDS, small enough for more resident workgroups).
// Records beyond a full segment's slots are added directly to the table gradient with packed
// FP16 atomics.
// 1024 threads: 71 KB of LDS (79 KB for both nets), two workgroups per CU; half the workgroups
// of 512-vertex ones, so half the per-workgroup segment scans and copy-out tails (3% faster).
// One run per workgroup: its slots of each segment start at the run's first slot.
constexpr int kBinVert = kRunVert;
static_assert(kBinVert == (int)kRunVert, "the first-slot read takes one run per workgroup");
constexpr int kBinMaxRec = 8 * kBinVert;  // every pair split

// Value of a record: the FP16 pairs w * dL/denc of the corners (x = 0, 1) of (y, z) = yz, both in
// one u64 for a paired record, corner xb alone for code 3.
// Value of a record: for a pair (codes 0-2), the FP16 pair of w_y w_z g, the (y, z) weight of the
// pair times dL/denc, which the apply kernels split between the two x corners with the x
// fraction carried in the key; for a single corner (code 3), the FP16 pair of its full weight
// times dL/denc.
__device__ __forceinline__ uint32_t record_value(const LevelCell& lc, int yz, float2 g, uint32_t code, uint32_t xb) {
    float w;
    if (code != 3u) {
        const float wy = (yz & 1) ? lc.f[1] : 1.0f - lc.f[1];
        const float wz = (yz & 2) ? lc.f[2] : 1.0f - lc.f[2];
        w = wy * wz;
    } else {
        w = corner_weight(lc, (yz << 1) | (int)xb);
    }
    const h2v gv = {(_Float16)(w * g.x), (_Float16)(w * g.y)};
    return __builtin_bit_cast(uint32_t, gv);
}

// dL/denc pair of the training kernel (FP16 pair) as FP32 (exact)
__device__ __forceinline__ float2 unpack_h2(uint32_t u) {
    const h2v v = __builtin_bit_cast(h2v, u);
    return make_float2((float)v.x, (float)v.y);
}

// x fraction of a cell in 2^-16 units (truncated: 0..65535), the key's upper half
__device__ __forceinline__ uint32_t x_fraction16(float fx) { return (uint32_t)(fx * 65536.0f); }

// The two corner values of a pair record, for the direct FP16 atomics of a full segment.
__device__ __forceinline__ void pair_corners(const LevelCell& lc, int yz, float2 g, h2v& v0, h2v& v1) {
    const float w0 = corner_weight(lc, yz << 1), w1 = corner_weight(lc, (yz << 1) | 1);
    v0 = h2v{(_Float16)(w0 * g.x), (_Float16)(w0 * g.y)};
    v1 = h2v{(_Float16)(w1 * g.x), (_Float16)(w1 * g.y)};
}

// DUAL: the records of both nets of a joint step in one pass. Their corners, and so their keys and
// slots, are the same (one grid layout, one set of record counts); the values come from each
// net's dL/denc (gdl, gdl2) and go to rec_val and rec_val2, the keys once to rec_key.
template <bool DUAL>
__global__ __launch_bounds__(kBinVert) void k_bin_levels(BinArgs b, uint32_t n_chunks) {
    static_assert(kMaxSegBins == 64, "one lane per segment");
    __shared__ uint32_t cnt[kMaxSegBins], start[kMaxSegBins], dst[kMaxSegBins], lim[kMaxSegBins], tot;
    __shared__ uint32_t list_key[kBinMaxRec];  // entry in level | code << 24
    __shared__ uint16_t list_src[kBinMaxRec];  // vertex in the workgroup << 3 | yz << 1 | x bit
    __shared__ float vf[3][kBinVert];          // cell fractions and dL/denc of the workgroup's vertices:
    __shared__ float2 vg[kBinVert];            // the values are formed at copy-out
    __shared__ float2 vg2[DUAL ? kBinVert : 1];
    // Workgroups are dispatched to the 8 XCDs in turn (workgroup w to XCD w % 8): slot w / 8 of
    // XCD w % 8 takes level (w / 8) % L of chunk 8 ((w / 8) / L) + w % 8, so the L levels of a
    // chunk run together on one XCD and read its positions from HBM once, then from that XCD's
    // L2. Every mapping is correct (the slots come from the scan); this one saves traffic.
    const uint32_t L = b.a.grid.n_levels, slot = blockIdx.x >> 3;
    const uint32_t l = slot % L, chunk = (slot / L) * 8u + (blockIdx.x & 7u), i = chunk * kBinVert + threadIdx.x;
    if (chunk >= n_chunks) return;  // whole workgroup
    uint32_t first = 0u, end = 0u;  // wave 0: this workgroup's first slot and the segment end
    if (threadIdx.x < kMaxSegBins) {
        cnt[threadIdx.x] = 0u;
        first = b.runs[((size_t)l * b.n_runs + chunk * (kBinVert / kRunVert)) * kMaxSegBins + threadIdx.x];
        end = b.seg_end[l * kMaxSegBins + threadIdx.x];
    }
    __syncthreads();
    const bool live = i < b.n;
    // dense (coarse) levels have few segments, often one per wave: aggregated LDS atomics
    const bool dense = (b.a.grid.dense_mask >> l) & 1u;
    uint32_t e[8], rk[8], code[4];  // e[2yz], e[2yz+1]: corners x = 0, 1 of (y, z) = yz
    if (live) {
        float x[3];
        load_xn(b.a, i, x);
        const LevelCell lc = level_cell(b.a.grid, (int)l, x);
#pragma unroll
        for (int k = 0; k < 3; ++k) vf[k][threadIdx.x] = lc.f[k];
        vg[threadIdx.x] = unpack_h2(b.gdl[(size_t)l * b.gdl_stride + i]);
        if (DUAL) vg2[threadIdx.x] = unpack_h2(b.gdl2[(size_t)l * b.gdl_stride + i]);
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            e[2 * yz] = corner_index(b.a.grid, (int)l, lc, yz << 1);  // < level size
            e[2 * yz + 1] = corner_index(b.a.grid, (int)l, lc, (yz << 1) | 1);
            code[yz] = pair_code(e[2 * yz], e[2 * yz + 1]);
            rk[2 * yz] = dense ? lds_count(cnt, e[2 * yz] >> kSegShift) : atomicAdd(&cnt[e[2 * yz] >> kSegShift], 1u);
            if (code[yz] == 3u)
                rk[2 * yz + 1] = dense ? lds_count(cnt, e[2 * yz + 1] >> kSegShift)
                                       : atomicAdd(&cnt[e[2 * yz + 1] >> kSegShift], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < kMaxSegBins) {
        const uint32_t sb = threadIdx.x, c = cnt[sb];
        const uint32_t inc = wave_inclusive_scan(c);
        start[sb] = inc - c;
        dst[sb] = first - (inc - c);  // list entry k of this segment goes to slot dst + k (mod 2^32)
        lim[sb] = end;
        if (sb == kMaxSegBins - 1) tot = inc;
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            const uint32_t e0 = e[2 * yz], e1 = e[2 * yz + 1];
            const uint32_t src = (threadIdx.x << 3) | (uint32_t)(yz << 1);
            const uint32_t k0 = start[e0 >> kSegShift] + rk[2 * yz];
            list_key[k0] = e0 | (code[yz] << 24);
            list_src[k0] = (uint16_t)src;
            if (code[yz] == 3u) {
                const uint32_t k1 = start[e1 >> kSegShift] + rk[2 * yz + 1];
                list_key[k1] = e1 | (3u << 24);
                list_src[k1] = (uint16_t)(src | 1u);
            }
        }
    }
    __syncthreads();
    const uint32_t n_rec = tot;
    for (uint32_t k = threadIdx.x; k < n_rec; k += kBinVert) {
        const uint32_t key = list_key[k], ent = key & 0xFFFFFFu, c = key >> 24, sb = ent >> kSegShift;
        const uint32_t src = list_src[k], vi = src >> 3;
        LevelCell lc;
#pragma unroll
        for (int q = 0; q < 3; ++q) lc.f[q] = vf[q][vi];
        const int yz = (int)((src >> 1) & 3u);
        const uint32_t d = dst[sb] + k;
        if (d < lim[sb]) {
            const uint32_t u = c != 3u ? x_fraction16(lc.f[0]) : 0u;
            b.rec_key[d] = (ent & (kSegEntries - 1u)) | (c << kSegShift) | (u << 16);
            b.rec_val[d] = record_value(lc, yz, vg[vi], c, src & 1u);
            if (DUAL) b.rec_val2[d] = record_value(lc, yz, vg2[vi], c, src & 1u);
        } else {  // segment full: add directly (the optimizer then scans the table gradient)
            if (b.scan_flag) atomicOr(b.scan_flag, 1u);
            if (DUAL && b.scan_flag2) atomicOr(b.scan_flag2, 1u);
            const uint32_t partner = ent ^ ((2u << (c & 3u)) - 1u);
#pragma unroll
            for (int net2 = 0; net2 < (DUAL ? 2 : 1); ++net2) {
                h2v* gl = (net2 ? b.grad2 : b.grad) + b.a.grid.offset[l];
                const float2 g = net2 ? vg2[vi] : vg[vi];
                if (c != 3u) {
                    h2v v0, v1;
                    pair_corners(lc, yz, g, v0, v1);
                    __builtin_amdgcn_global_atomic_fadd_v2f16(gl + ent, v0);
                    __builtin_amdgcn_global_atomic_fadd_v2f16(gl + partner, v1);
                } else {
                    __builtin_amdgcn_global_atomic_fadd_v2f16(
                        gl + ent, __builtin_bit_cast(h2v, record_value(lc, yz, g, c, src & 1u)));
                }
            }
        }
    }
}

// Sums of the binned records. Every FP16 value is an integer multiple of 2^-24 below 2^16, so
// the records are summed exactly as 64-bit fixed point with integer LDS atomics (LDS float
// atomics are an order of magnitude slower); the sums do not depend on the record order.
// k_apply: one workgroup per segment of a hashed level adds its sum to the table gradient
// (apply_segment); segments of the dense levels, which receive many records per entry, are cut
// into chunks whose sums are added to a 64-bit accumulator with coalesced integer atomics
// (apply_chunk); k_apply_finish adds the accumulator to the table gradient. A non-finite
// record turns a table-gradient entry into NaN, so the finite check of the optimiser step
// sees it. Counts and accumulators are cleared for the next launch.
constexpr int kApplyThreads = 1024;
constexpr int kApplyBatch = 2;

// One feature of a record into the LDS sums of plane `acc` (feature f of entry e at
// f * kSegEntries + e: consecutive entries are 8 bytes apart, so the entries of a wave's records
// spread over all LDS banks). The FP16 value is an exact count of 2^-24 (its magnitude is below
// 2^41); a pair record's value is split exactly between its two corners: the x = 1 corner (the
// partner) takes v u / 2^16, truncated toward zero, the x = 0 corner the rest, so the two parts
// sum to v and the sums do not depend on the record order. A single-corner record has u = 0, so
// its partner receives zero. Branch-free: zero adds cost less than divergent lanes; a
// non-finite value adds nothing and sets `bad`.
__device__ __forceinline__ void apply_feature(unsigned long long* acc, uint32_t e, uint32_t partner, uint32_t u,
                                              uint32_t h, uint32_t& bad) {
    const uint32_t ex = (h >> 10) & 31u, m = h & 1023u;
    const uint32_t norm = ex != 0u ? 1u : 0u, finite = ex != 31u ? 1u : 0u;
    bad |= finite ^ 1u;
    // written as masks and shifts: the compiler turns nested selects here into branches
    const unsigned long long mag =
        ((unsigned long long)(m | (norm << 10)) << (ex - norm)) & (0ull - (unsigned long long)finite);
    const unsigned long long t = (mag * u) >> 16, x0 = mag - t;  // mag * u < 2^57
    const bool neg = (h & 0x8000u) != 0u;
    // zero parts (the partner of a single-corner record, zero values) skip their atomic: the
    // lanes drop out of the LDS instruction, whose cost grows with its active lanes
    if (x0) atomicAdd(&acc[e], neg ? 0ull - x0 : x0);
    if (t) atomicAdd(&acc[partner], neg ? 0ull - t : t);
}

// One record (key, value) into the LDS sums of its segment. Key: entry within the segment
// (bits 0-12) | pair code << 13 | x fraction in 2^-16 units << 16; value: FP16 pair (features 0, 1).
__device__ __forceinline__ void apply_record(unsigned long long* acc, uint32_t key, uint32_t v, uint32_t& bad) {
    const uint32_t e = key & (kSegEntries - 1u), c = (key >> kSegShift) & 3u, u = key >> 16;
    const uint32_t partner = e ^ ((2u << c) - 1u);  // c = 3: e ^ 15, in the segment, receives zero
    apply_feature(acc, e, partner, u, v & 0xffffu, bad);
    apply_feature(acc + kSegEntries, e, partner, u, v >> 16, bad);
}

// Records [r0, r1) of a segment into the LDS sums, four per thread and load.
__device__ __forceinline__ void apply_range(unsigned long long* acc_s, const uint32_t* key, const uint32_t* val,
                                            uint32_t r0, uint32_t r1, uint32_t& bad) {
    // r0 is a multiple of 64 (segment slots and chunks start there): 16-byte loads
    const uint4* k4 = static_cast<const uint4*>(__builtin_assume_aligned(key + r0, 16));
    const uint4* v4 = static_cast<const uint4*>(__builtin_assume_aligned(val + r0, 16));
    const uint32_t n4 = (r1 - r0) / 4;
    constexpr uint32_t kStep = kApplyThreads * kApplyBatch;
    uint32_t j0 = 0;
    for (; j0 + kStep <= n4; j0 += kStep) {  // full batches: every load in bounds, all issued first
        uint4 kk[kApplyBatch], q[kApplyBatch];
#pragma unroll
        for (int b = 0; b < kApplyBatch; ++b) {
            kk[b] = k4[j0 + b * kApplyThreads + threadIdx.x];
            q[b] = v4[j0 + b * kApplyThreads + threadIdx.x];
        }
#pragma unroll
        for (int b = 0; b < kApplyBatch; ++b) {
            apply_record(acc_s, kk[b].x, q[b].x, bad);
            apply_record(acc_s, kk[b].y, q[b].y, bad);
            apply_record(acc_s, kk[b].z, q[b].z, bad);
            apply_record(acc_s, kk[b].w, q[b].w, bad);
        }
    }
    for (uint32_t j = j0 + threadIdx.x; j < n4; j += kApplyThreads) {
        const uint4 kk = k4[j], q = v4[j];
        apply_record(acc_s, kk.x, q.x, bad);
        apply_record(acc_s, kk.y, q.y, bad);
        apply_record(acc_s, kk.z, q.z, bad);
        apply_record(acc_s, kk.w, q.w, bad);
    }
    for (uint32_t r = r0 + 4 * n4 + threadIdx.x; r < r1; r += kApplyThreads) apply_record(acc_s, key[r], val[r], bad);
}

__device__ __forceinline__ h2v add_fixed(h2v o, long long s0, long long s1) {
    return h2v{(_Float16)((float)o.x + (float)((double)s0 * 0x1p-24)),
               (_Float16)((float)o.y + (float)((double)s1 * 0x1p-24))};
}

// One hashed-level segment (`direct` entry b) by one workgroup. acc_s: [2][kSegEntries], units
// of 2^-24.
__device__ __forceinline__ void apply_segment(const ApplyArgs& a, uint32_t b, unsigned long long* acc_s,
                                              uint32_t& bad_s) {
    const uint32_t seg = a.direct[b];
    const SegInfo sg = a.segs[seg];
    for (uint32_t k = threadIdx.x; k < 2 * kSegEntries; k += kApplyThreads) acc_s[k] = 0ull;
    if (threadIdx.x == 0) bad_s = 0u;
    __syncthreads();
    uint32_t bad = 0u;
    apply_range(acc_s, a.key + sg.rec_off, a.val + sg.rec_off, 0, min(a.seg_count[seg], sg.cap), bad);
    if (bad) bad_s = 1u;
    __syncthreads();
    h2v* g = a.grad + sg.entry0;
    bool inf = false;  // a written entry is not finite (an FP16 overflow of the sum)
    for (uint32_t e = threadIdx.x; e < sg.n_entries; e += kApplyThreads) {
        const long long s0 = (long long)acc_s[e], s1 = (long long)acc_s[kSegEntries + e];
        if (s0 != 0 || s1 != 0) {
            const h2v v = add_fixed(g[e], s0, s1);
            g[e] = v;
            inf |= !isfinite((float)v.x) || !isfinite((float)v.y);
        }
    }
    if (a.nonfinite && __any(inf) && (threadIdx.x & 63) == 0) atomicOr(a.nonfinite, 1u);
    if (threadIdx.x == 0) {
        if (bad_s) {
            g[0] = h2v{(_Float16)__builtin_nanf(""), (_Float16)__builtin_nanf("")};
            if (a.nonfinite) atomicOr(a.nonfinite, 1u);
        }
        if (!a.keep_counts) a.seg_count[seg] = 0;
    }
}

// One chunk of a dense-level segment (`chunks` entry b) by one workgroup.
__device__ __forceinline__ void apply_chunk(const ApplyArgs& a, uint32_t b, unsigned long long* acc_s) {
    const uint2 ch = a.chunks[b];  // {segment, first record}
    const SegInfo sg = a.segs[ch.x];
    const uint32_t n = min(a.seg_count[ch.x], sg.cap);
    if (ch.y >= n) return;  // whole workgroup
    for (uint32_t k = threadIdx.x; k < 2 * sg.n_entries; k += kApplyThreads)
        acc_s[(k & 1u) * kSegEntries + (k >> 1)] = 0ull;
    __syncthreads();
    uint32_t bad = 0u;
    apply_range(acc_s, a.key + sg.rec_off, a.val + sg.rec_off, ch.y, min(ch.y + a.chunk, n), bad);
    if (bad) atomicOr(a.bad, 1u);
    __syncthreads();
    unsigned long long* acc = a.acc + 2 * (size_t)sg.entry0;
    for (uint32_t k = threadIdx.x; k < 2 * sg.n_entries; k += kApplyThreads) {  // global: [entry][2]
        const unsigned long long v = acc_s[(k & 1u) * kSegEntries + (k >> 1)];
        if (v) atomicAdd(acc + k, v);
    }
}

// Both kinds in one launch: workgroups [0, n_direct) take the hashed segments, the rest the
// dense-level chunks, so that the chunks fill the CUs the last segments leave idle (and the
// launch gap between two kernels goes away).
__global__ __launch_bounds__(kApplyThreads) void k_apply(ApplyArgs a, uint32_t n_direct) {
    extern __shared__ unsigned long long acc_s[];
    __shared__ uint32_t bad_s;
    if (blockIdx.x < n_direct)
        apply_segment(a, blockIdx.x, acc_s, bad_s);
    else
        apply_chunk(a, blockIdx.x - n_direct, acc_s);
}

__global__ void k_apply_finish(ApplyArgs a) {
    bool inf = false;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < a.n_acc; e += gridDim.x * blockDim.x) {
        const long long s0 = (long long)a.acc[2 * e], s1 = (long long)a.acc[2 * e + 1];
        if (s0 != 0 || s1 != 0) {
            const h2v v = add_fixed(a.grad[e], s0, s1);
            a.grad[e] = v;
            inf |= !isfinite((float)v.x) || !isfinite((float)v.y);
            a.acc[2 * e] = 0ull;
            a.acc[2 * e + 1] = 0ull;
        }
    }
    if (a.nonfinite && __any(inf) && (threadIdx.x & 63) == 0) atomicOr(a.nonfinite, 1u);
    if (blockIdx.x == 0) {
        if (!a.keep_counts)
            for (uint32_t k = threadIdx.x; k < a.n_chunked; k += blockDim.x) a.seg_count[a.chunked[k]] = 0;
        if (threadIdx.x == 0 && *a.bad) {
            a.grad[0] = h2v{(_Float16)__builtin_nanf(""), (_Float16)__builtin_nanf("")};
            if (a.nonfinite) atomicOr(a.nonfinite, 1u);
            *a.bad = 0u;
        }
    }
}

// Sum of the dense-level gradient replicas into the table gradient; the replicas are cleared.
__global__ void k_reduce_dense(float* __restrict__ rep_f32, _Float16* __restrict__ rep_h, uint32_t n_dense,
                               float* __restrict__ grad, _Float16* __restrict__ grad_h) {
    // element k of [n_dense][2]
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < 2 * n_dense; k += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (uint32_t r = 0; r < kDenseReplicas; ++r) {
            const size_t o = (size_t)r * 2 * n_dense + k;
            if (rep_h) {
                s += (float)rep_h[o];
                rep_h[o] = (_Float16)0.0f;
            } else {
                s += rep_f32[o];
                rep_f32[o] = 0.0f;
            }
        }
        if (grad_h) {
            grad_h[k] = (_Float16)((float)grad_h[k] + s);
        } else {
            grad[k] += s;
        }
    }
}

// Sum over workgroups of the dW partials, in a fixed order: a block takes 64 parameters,
// its 16 thread groups take every 16th partial row (8 loads in flight), then the groups
// are summed in group order.
constexpr int kRedParams = 64, kRedGroups = 16, kLossThreads = 256;

// Batch loss from the per-workgroup FP64 partials of the training kernel, summed in a fixed
// order; a NaN/Inf total counts one non-finite step in *nonfinite.
__global__ __launch_bounds__(kLossThreads) void k_loss_finish(const double* __restrict__ part, uint32_t n_blocks,
                                                              float* loss_out, uint32_t* nonfinite) {
    __shared__ double red[kLossThreads];
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < n_blocks; b += kLossThreads) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
#pragma unroll
    for (int o = kLossThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double v = red[0];
        if (loss_out) *loss_out += (float)v;  // the caller clears it (AllSelf's two nets add)
        if (nonfinite && !isfinite(v)) *nonfinite += 1u;
    }
}

__global__ __launch_bounds__(kRedParams* kRedGroups) void k_reduce_partials(const float* __restrict__ part,
                                                                           uint32_t n_blocks, uint32_t n_mlp,
                                                                           float* __restrict__ grad) {
    __shared__ float red[kRedGroups][kRedParams];
    const uint32_t pl = threadIdx.x % kRedParams, g = threadIdx.x / kRedParams;
    const uint32_t k = blockIdx.x * kRedParams + pl;
    float s = 0.0f;
    if (k < n_mlp) {
        uint32_t b = g;
        for (; b + 7 * kRedGroups < n_blocks; b += 8 * kRedGroups) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[(size_t)(b + u * kRedGroups) * n_mlp + k];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; b < n_blocks; b += kRedGroups) s += part[(size_t)b * n_mlp + k];
    }
    red[g][pl] = s;
    __syncthreads();
    if (g == 0 && k < n_mlp) {
        float t = 0.0f;
#pragma unroll
        for (int q = 0; q < kRedGroups; ++q) t += red[q][pl];
        grad[k] += t;
    }
}

}  // namespace

int train_tiles_fwd(int nl) { return 10 + (nl - 2) * 8 + 4; }
int train_tiles_bwd(int nl) { return 4 + (nl - 2) * 8 + 2; }
size_t train_op_bytes(int precision) { return precision == NRRS_FP32 ? 32 : 16; }

hipError_t launch_pack_train(int precision, const float* W, int nl, int n_out, void* fwd, void* bwd,
                             hipStream_t s) {
    const int nf = train_tiles_fwd(nl), nb = train_tiles_bwd(nl);
    if (precision == NRRS_FP32)
        hipLaunchKernelGGL(k_pack_train<float>, dim3(nf + nb), dim3(64), 0, s, W, nl, n_out,
                           static_cast<float*>(fwd), static_cast<float*>(bwd), nf, nb);
    else if (precision == NRRS_BF16)
        hipLaunchKernelGGL(k_pack_train<__bf16>, dim3(nf + nb), dim3(64), 0, s, W, nl, n_out,
                           static_cast<__bf16*>(fwd), static_cast<__bf16*>(bwd), nf, nb);
    else
        hipLaunchKernelGGL(k_pack_train<_Float16>, dim3(nf + nb), dim3(64), 0, s, W, nl, n_out,
                           static_cast<_Float16*>(fwd), static_cast<_Float16*>(bwd), nf, nb);
    return hipGetLastError();
}

hipError_t launch_after_adam(int precision, const AdamArgs& a, const float* W, const float* W_ema, int nl, int n_out,
                             uint16_t* mlp_lp, uint16_t* mlp_lp_ema, void* fwd, void* bwd, hipStream_t s) {
    const int nf = train_tiles_fwd(nl), nb = train_tiles_bwd(nl);
    if (precision == NRRS_BF16) {
        const int split = mfma::split_of<mfma::BF16>(n_out), ni = mfma::tiles_of(nl, split);
        const dim3 grid(1 + ni * (W_ema ? 2 : 1) + nf + nb);
        hipLaunchKernelGGL((k_after_adam<mfma::BF16, __bf16>), grid, dim3(64), 0, s, a, W, W_ema, nl, n_out, split,
                           mlp_lp, mlp_lp_ema, ni, static_cast<__bf16*>(fwd), static_cast<__bf16*>(bwd), nf, nb);
    } else if (precision == NRRS_FP16) {
        const int split = mfma::split_of<mfma::F16>(n_out), ni = mfma::tiles_of(nl, split);
        const dim3 grid(1 + ni * (W_ema ? 2 : 1) + nf + nb);
        hipLaunchKernelGGL((k_after_adam<mfma::F16, _Float16>), grid, dim3(64), 0, s, a, W, W_ema, nl, n_out, split,
                           mlp_lp, mlp_lp_ema, ni, static_cast<_Float16*>(fwd), static_cast<_Float16*>(bwd), nf, nb);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_train(int precision, const TrainArgs& t, int nl, int n_out, int loss, uint32_t blocks,
                        hipStream_t s) {
    if (precision == NRRS_FP32) return dispatch_train<TF32>(t, nl, n_out, loss, blocks, s);
    if (precision == NRRS_BF16) return dispatch_train<TBF16>(t, nl, n_out, loss, blocks, s);
    return dispatch_train<TF16>(t, nl, n_out, loss, blocks, s);
}

hipError_t launch_apply_bins(const ApplyArgs& a, uint32_t n_direct, uint32_t n_chunks, hipStream_t s) {
    constexpr size_t smem = 2 * kSegEntries * sizeof(unsigned long long);
    static std::atomic<bool> attr[kMaxDevices];
    hipError_t e = set_dyn_lds_once(attr, reinterpret_cast<const void*>(&k_apply), smem);
    if (e != hipSuccess) return e;
    if (n_direct + n_chunks)
        hipLaunchKernelGGL(k_apply, dim3(n_direct + n_chunks), dim3(kApplyThreads), smem, s, a, n_direct);
    if (n_chunks) {
        const uint32_t blocks = std::min<uint32_t>((a.n_acc + 255) / 256, 2048);
        hipLaunchKernelGGL(k_apply_finish, dim3(blocks), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_bin_scan(const BinArgs& b, hipStream_t s) {
    if (!b.n) return hipSuccess;
    if (b.n_runs != (b.n + kRunVert - 1) / kRunVert) return hipErrorInvalidValue;
    const uint32_t n_tiles = (b.n_runs + kScanRows - 1) / kScanRows;
    if (n_tiles > b.max_tiles) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bin_tiles, dim3(b.a.grid.n_levels * n_tiles), dim3(64 * kScanWaves), 0, s, b, n_tiles);
    hipLaunchKernelGGL(k_bin_scan, dim3(b.a.grid.n_levels * n_tiles), dim3(64 * kScanWaves), 0, s, b, n_tiles);
    return hipGetLastError();
}

hipError_t launch_bin_records(const BinArgs& b, hipStream_t s) {
    if (!b.n) return hipSuccess;
    if (b.n_runs != (b.n + kRunVert - 1) / kRunVert) return hipErrorInvalidValue;
    const bool dual = b.gdl2 != nullptr;
    if (dual && (!b.rec_val2 || !b.grad2)) return hipErrorInvalidValue;
    const uint32_t n_chunks = (b.n + kBinVert - 1) / kBinVert;
    const uint64_t blocks = 8ull * ((n_chunks + 7) / 8) * b.a.grid.n_levels;  // XCD-major order, see k_bin_levels
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    if (dual)
        hipLaunchKernelGGL(k_bin_levels<true>, dim3((uint32_t)blocks), dim3(kBinVert), 0, s, b, n_chunks);
    else
        hipLaunchKernelGGL(k_bin_levels<false>, dim3((uint32_t)blocks), dim3(kBinVert), 0, s, b, n_chunks);
    return hipGetLastError();
}

hipError_t launch_bin_levels(const BinArgs& b, hipStream_t s) {
    const hipError_t e = launch_bin_scan(b, s);
    if (e != hipSuccess) return e;
    return launch_bin_records(b, s);
}

hipError_t launch_reduce_dense(float* rep_f32, void* rep_h, uint32_t n_dense, float* grad, void* grad_h,
                               hipStream_t s) {
    if (!n_dense) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_dense, dim3(512), dim3(256), 0, s, rep_f32, static_cast<_Float16*>(rep_h), n_dense,
                       grad, static_cast<_Float16*>(grad_h));
    return hipGetLastError();
}

hipError_t launch_loss_finish(const double* part, uint32_t n_blocks, float* loss_out, uint32_t* nonfinite,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_loss_finish, dim3(1), dim3(kLossThreads), 0, s, part, n_blocks, loss_out, nonfinite);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const float* part, uint32_t n_blocks, uint32_t n_mlp, float* grad, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3((n_mlp + kRedParams - 1) / kRedParams), dim3(kRedParams * kRedGroups), 0,
                       s, part, n_blocks, n_mlp, grad);
    return hipGetLastError();
}

}  // namespace nrrs
