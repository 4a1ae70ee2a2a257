// Mixed-precision path: hash-grid + SH encoding fused with the MLP on MFMA
// (v_mfma_f32_32x32x16_{bf16,f16}, FP32 accumulation).
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
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct BF16 {
    typedef bf16x8 vec8;
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
    __device__ static float2 pair(uint32_t u) {
        _Float16 lo = __builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu));
        _Float16 hi = __builtin_bit_cast(_Float16, (uint16_t)(u >> 16));
        return make_float2((float)lo, (float)hi);
    }
    __device__ static vec8 pack(const float* x) {
        vec8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (_Float16)x[j];
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
    return __builtin_bit_cast(uint16_t, (_Float16)f);
}

constexpr int kWaves = 4;
constexpr int kBlockT = 64 * kWaves;
constexpr int kTileHalves = 64 * 8;  // one packed A tile: 64 lanes x 8 values

__host__ __device__ constexpr int tiles_of(int n_layers) { return 10 + (n_layers - 2) * 8 + 4; }

// Input column of the packed K element (layer, kb, h, j).
__device__ __forceinline__ int packed_col(int layer, int kb, int h, int j) {
    return layer == 0 ? kb * 16 + 8 * h + j : 16 * kb + 8 * (j >> 2) + 4 * h + (j & 3);
}

// One thread per (tile, lane): 8 weights of the A operand.
template <typename P>
__global__ void k_pack_mlp(const float* __restrict__ W, int n_layers, int n_out, uint16_t* __restrict__ out) {
    const int t = blockIdx.x;      // tile
    const int lane = threadIdx.x;  // 0..63
    int layer, ot, kb, base_w = 0, in = kEnc, rows = kHidden;
    if (t < 10) {
        layer = 0;
        ot = t / 5;
        kb = t % 5;
    } else {
        int u = t - 10;
        layer = 1 + u / 8;
        if (layer <= n_layers - 2) {
            ot = (u % 8) / 4;
            kb = u % 4;
        } else {
            layer = n_layers - 1;
            ot = 0;
            kb = u - (n_layers - 2) * 8;
        }
        base_w = kHidden * kEnc + (layer - 1) * kHidden * kHidden;
        in = kHidden;
        rows = layer == n_layers - 1 ? n_out : kHidden;
    }
    const int h = lane >> 5, row = ot * 32 + (lane & 31);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float w = row < rows ? W[base_w + row * in + packed_col(layer, kb, h, j)] : 0.0f;
        out[((size_t)t * 64 + lane) * 8 + j] = lp_bits<P>(w);
    }
}

template <typename P>
__global__ void k_to_lp(const float* __restrict__ src, uint16_t* __restrict__ dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = lp_bits<P>(src[i]);
}

// The 40 channels of vertex i that half h of the wavefront feeds to layer 0:
// channel kb*16 + 8h + j goes to element j of K block kb.
template <typename P>
__device__ __forceinline__ void encode_half(const EncodeArgs& a, int net, const uint32_t* __restrict__ table,
                                            uint32_t i, int h, typename P::vec8* B) {
    float e[8];
    float x[3];
    load_xn(a, i, x);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int ll = 0; ll < 4; ++ll) {
            const int l = 8 * kb + 4 * h + ll;
            float2 f = make_float2(0.0f, 0.0f);
            if (l < (int)a.grid.n_levels) f = grid_level_pairs<PairLoadLp<P>>(table, a.grid, l, x);
            e[2 * ll] = f.x;
            e[2 * ll + 1] = f.y;
        }
        B[kb] = P::pack(e);
    }
    float t[kEnc];
    encode_tail(a, net, i, t);
#pragma unroll
    for (int kb = 2; kb < 5; ++kb) {
#pragma unroll
        for (int j = 0; j < 8; ++j) e[j] = h ? t[kb * 16 + 8 + j] : t[kb * 16 + j];
        B[kb] = P::pack(e);
    }
}

__device__ __forceinline__ u32x4 lds_tile(const u32x4* __restrict__ sW, int tile, int lane) {
    return sW[tile * 64 + lane];
}

template <typename P>
__device__ __forceinline__ void relu_to_operands(const f32x16* D, typename P::vec8* H) {
#pragma unroll
    for (int ot = 0; ot < 2; ++ot) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = fmaxf(D[ot][8 * g + j], 0.0f);
            H[2 * ot + g] = P::pack(x);
        }
    }
}

template <typename P, int NL, int NOUT, int EPI>
__global__ __launch_bounds__(kBlockT) void k_net_mfma(EncodeArgs a, int net, uint32_t n,
                                                      const uint32_t* __restrict__ table,
                                                      const u32x4* __restrict__ Wp, EpiArgs ep) {
    extern __shared__ u32x4 sW[];
    constexpr int kTiles = tiles_of(NL);
    for (int k = threadIdx.x; k < kTiles * 64; k += kBlockT) sW[k] = Wp[k];
    __syncthreads();

    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const int wave = threadIdx.x >> 6;
    const uint32_t n_cb = (n + 31) / 32;
    uint64_t term = 0;
    for (uint32_t cb = blockIdx.x * kWaves + wave; cb < n_cb; cb += gridDim.x * kWaves) {
        const uint32_t i = cb * 32 + col;
        const bool live = i < n;
        typename P::vec8 B[5];
        if (live) {
            encode_half<P>(a, net, table, i, h, B);
        } else {
#pragma unroll
            for (int kb = 0; kb < 5; ++kb) B[kb] = typename P::vec8{};
        }
        f32x16 D[2];
#pragma unroll
        for (int ot = 0; ot < 2; ++ot) {
            D[ot] = f32x16{};
#pragma unroll
            for (int kb = 0; kb < 5; ++kb) D[ot] = P::mma(lds_tile(sW, ot * 5 + kb, lane), B[kb], D[ot]);
        }
        typename P::vec8 H[4];
#pragma unroll
        for (int layer = 1; layer < NL - 1; ++layer) {
            relu_to_operands<P>(D, H);
            const int base = 10 + (layer - 1) * 8;
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) {
                D[ot] = f32x16{};
#pragma unroll
                for (int kb = 0; kb < 4; ++kb)
                    D[ot] = P::mma(lds_tile(sW, base + ot * 4 + kb, lane), H[kb], D[ot]);
            }
        }
        relu_to_operands<P>(D, H);
        f32x16 Y = f32x16{};
        const int base = 10 + (NL - 2) * 8;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) Y = P::mma(lds_tile(sW, base + kb, lane), H[kb], Y);

        // output c of vertex i sits in lane half c / 4, register c % 4
        if (EPI == EPI_RAW) {
            if (live) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = 4 * h + r;
                    if (c < NOUT) ep.out[(size_t)c * ep.out_stride + i] = Y[r];
                }
            }
        } else if (live && h == 0) {
            float r;
            if (EPI == EPI_RRS) {
                r = activation_sigma(Y[0]);
            } else {
                const uint32_t px = a.v.pixel[i], S = a.v.stride, Pn = a.n_pixels;
                float acc = 0.0f;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    float L = fminf(fmaxf(Y[k], 1e-4f), 1e4f);
                    acc += L * a.v.thp[k * S + i] / (a.film[k * Pn + px] + 0.01f);
                }
                r = acc / 3.0f;
            }
            float q = rrs_window_clamp(ep.d, r);
            ep.q[i] = q;
            term += rrs_sum_term(ep.d, q);
        }
        if (EPI == EPI_ADRRS && ep.out && live) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int c = 4 * h + r;
                if (c < NOUT) ep.out[(size_t)c * ep.out_stride + i] = Y[r];
            }
        }
    }
    if (EPI != EPI_RAW) block_atomic_add_u64(term, ep.sum);
}

template <typename P, int NL, int NOUT, int EPI>
hipError_t launch_one(const EncodeArgs& a, int net, uint32_t n, const uint16_t* table, const uint16_t* Wp,
                      const EpiArgs& ep, hipStream_t s) {
    constexpr size_t smem = (size_t)tiles_of(NL) * kTileHalves * 2;
    const uint32_t n_cb = (n + 31) / 32;
    uint32_t blocks = (n_cb + kWaves - 1) / kWaves;
    blocks = blocks > 4096 ? 4096 : blocks;
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL((k_net_mfma<P, NL, NOUT, EPI>), dim3(blocks), dim3(kBlockT), smem, s, a, net, n,
                       reinterpret_cast<const uint32_t*>(table), reinterpret_cast<const u32x4*>(Wp), ep);
    return hipGetLastError();
}

template <typename P>
hipError_t dispatch(const EncodeArgs& a, int net, int n_layers, int n_out, int epi, uint32_t n,
                    const uint16_t* table, const uint16_t* Wp, const EpiArgs& ep, hipStream_t s) {
    if (n_layers == 5 && n_out == 6) {
        if (epi == EPI_RAW) return launch_one<P, 5, 6, EPI_RAW>(a, net, n, table, Wp, ep, s);
        if (epi == EPI_ADRRS) return launch_one<P, 5, 6, EPI_ADRRS>(a, net, n, table, Wp, ep, s);
    }
    if (n_layers == 4 && n_out == 1) {
        if (epi == EPI_RAW) return launch_one<P, 4, 1, EPI_RAW>(a, net, n, table, Wp, ep, s);
        if (epi == EPI_RRS) return launch_one<P, 4, 1, EPI_RRS>(a, net, n, table, Wp, ep, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

size_t mfma_packed_halves(int n_layers) { return (size_t)tiles_of(n_layers) * kTileHalves; }

hipError_t launch_pack_lp(int precision, const float* grid, uint64_t n_grid, const float* W, int n_layers,
                          int n_out, uint16_t* grid_lp, uint16_t* W_lp, hipStream_t s) {
    const int tiles = tiles_of(n_layers);
    if (precision == NRRS_BF16) {
        hipLaunchKernelGGL(k_pack_mlp<BF16>, dim3(tiles), dim3(64), 0, s, W, n_layers, n_out, W_lp);
        if (grid) hipLaunchKernelGGL(k_to_lp<BF16>, dim3(2048), dim3(256), 0, s, grid, grid_lp, n_grid);
    } else if (precision == NRRS_FP16) {
        hipLaunchKernelGGL(k_pack_mlp<F16>, dim3(tiles), dim3(64), 0, s, W, n_layers, n_out, W_lp);
        if (grid) hipLaunchKernelGGL(k_to_lp<F16>, dim3(2048), dim3(256), 0, s, grid, grid_lp, n_grid);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_net_mfma(int precision, const EncodeArgs& a, int net, int n_layers, int n_out, int epi,
                           uint32_t n, const uint16_t* table, const uint16_t* Wp, const EpiArgs& ep,
                           hipStream_t s) {
    if (precision == NRRS_BF16) return dispatch<BF16>(a, net, n_layers, n_out, epi, n, table, Wp, ep, s);
    if (precision == NRRS_FP16) return dispatch<F16>(a, net, n_layers, n_out, epi, n, table, Wp, ep, s);
    return hipErrorInvalidValue;
}

}  // namespace nrrs
