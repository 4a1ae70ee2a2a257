// Plain FP32 path: hash-grid + SH encoding fused with the MLP on the vector ALUs.
// One thread owns one vertex; the MLP weights sit in LDS and are read as
// wave-uniform float4 broadcasts, the activations stay in VGPRs.
#include "device_common.hpp"

namespace nrrs {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_encode(EncodeArgs a, int net, uint32_t n,
                                                   const float* __restrict__ table,
                                                   float* __restrict__ enc) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float e[kEnc];
    encode_full<float>(a, net, table, i, e);
#pragma unroll
    for (int k = 0; k < kEnc; ++k) enc[(size_t)i * kEnc + k] = e[k];
}

__global__ __launch_bounds__(kBlock) void k_grid_indices(EncodeArgs a, uint32_t n,
                                                         uint32_t* __restrict__ idx) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float x[3];
    load_xn(a, i, x);
    for (int l = 0; l < (int)a.grid.n_levels; ++l) {
        LevelCell lc = level_cell(a.grid, l, x);
#pragma unroll
        for (int cn = 0; cn < 8; ++cn)
            idx[((size_t)i * a.grid.n_levels + l) * 8 + cn] = corner_index(a.grid, l, lc, cn);
    }
}

// out = W h (W row-major [OUT][IN] in LDS), optional hidden-layer activation (ACT_RELU /
// ACT_LEAKY; the output layer is linear).
template <int IN, int OUT, bool RELU, int ACT = ACT_RELU>
__device__ __forceinline__ void dense_fp32(const float* __restrict__ sW, const float* hin, float* hout) {
#pragma unroll
    for (int o = 0; o < OUT; ++o) {
        float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
        for (int k = 0; k < IN; k += 8) {
            float4 w0 = *reinterpret_cast<const float4*>(sW + o * IN + k);
            float4 w1 = *reinterpret_cast<const float4*>(sW + o * IN + k + 4);
            a0 = __fmaf_rn(w0.x, hin[k + 0], a0);
            a1 = __fmaf_rn(w0.y, hin[k + 1], a1);
            a0 = __fmaf_rn(w0.z, hin[k + 2], a0);
            a1 = __fmaf_rn(w0.w, hin[k + 3], a1);
            a0 = __fmaf_rn(w1.x, hin[k + 4], a0);
            a1 = __fmaf_rn(w1.y, hin[k + 5], a1);
            a0 = __fmaf_rn(w1.z, hin[k + 6], a0);
            a1 = __fmaf_rn(w1.w, hin[k + 7], a1);
        }
        float acc = a0 + a1;
        if (!RELU)
            hout[o] = acc;
        else if (ACT == ACT_LEAKY)
            hout[o] = acc > 0.0f ? acc : acc * kLeakySlope;
        else
            hout[o] = fmaxf(acc, 0.0f);
    }
}

// Level-major FP32 grid encoding: workgroup b takes level b / n_chunks and kChunk vertices, so
// that the GPU walks the levels one after the other and each level's table (4 MB at T = 2^19)
// stays in L2 while it is gathered.
constexpr uint32_t kChunk = 1024;

__global__ __launch_bounds__(kChunk) void k_encode_levels_f32(EncodeArgs a, const float* __restrict__ table, uint32_t n,
                                                              uint32_t n_chunks, float2* __restrict__ out,
                                                              uint32_t stride) {
    const uint32_t l = blockIdx.x / n_chunks, i = (blockIdx.x % n_chunks) * kChunk + threadIdx.x;
    if (l >= a.grid.n_levels || i >= n) return;
    float x[3];
    load_xn(a, i, x);
    out[(size_t)l * stride + i] = grid_level_pairs<PairLoad<float>>(table, a.grid, (int)l, x);
}

template <int NL, int NOUT>
constexpr int mlp_floats() {
    return kHidden * kEnc + (NL - 2) * kHidden * kHidden + NOUT * kHidden;
}

template <int NL, int NOUT, int EPI, int ACT>
__global__ __launch_bounds__(kBlock) void k_net_fp32(EncodeArgs a, int net, uint32_t n,
                                                     const float* __restrict__ table,
                                                     const float* __restrict__ W, EpiArgs ep) {
    extern __shared__ float4 smem4[];
    float* sW = reinterpret_cast<float*>(smem4);
    constexpr int kW = mlp_floats<NL, NOUT>();
    for (int k = threadIdx.x; k < kW / 4; k += kBlock)
        smem4[k] = reinterpret_cast<const float4*>(W)[k];
    __syncthreads();

    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < n;
    float y[NOUT];
    {
        float h[kEnc];
        float t[kHidden];
        if (live && a.grid_enc_f32) {
            // grid channels from the level-major encoding, the rest computed here
#pragma unroll
            for (int l = 0; l < kMaxLevels; ++l) {
                const float2 f = l < (int)a.grid.n_levels ? a.grid_enc_f32[(size_t)l * a.grid_enc_stride + i]
                                                           : make_float2(0.0f, 0.0f);
                h[2 * l] = f.x;
                h[2 * l + 1] = f.y;
            }
            encode_tail(a, net, i, h);
        } else if (live) {
            encode_full<float>(a, net, table, i, h);
        } else {
#pragma unroll
            for (int k = 0; k < kEnc; ++k) h[k] = 0.0f;
        }
        dense_fp32<kEnc, kHidden, true, ACT>(sW, h, t);
        const float* w = sW + kHidden * kEnc;
#pragma unroll
        for (int layer = 1; layer < NL - 1; ++layer) {
            dense_fp32<kHidden, kHidden, true, ACT>(w, t, h);
#pragma unroll
            for (int k = 0; k < kHidden; ++k) t[k] = h[k];
            w += kHidden * kHidden;
        }
        dense_fp32<kHidden, NOUT, false>(w, t, y);
    }

    if (EPI == EPI_RAW) {
        if (live) {
#pragma unroll
            for (int j = 0; j < NOUT; ++j)
                if (ep.head < 0 || j / 3 == ep.head) ep.out[(size_t)j * ep.out_stride + i] = y[j];
        }
        return;
    }
    uint64_t term = 0;
    if (live) {
        float r;
        if (EPI == EPI_RRS) {
            r = activation_sigma(y[0]);
        } else {
            // StatNet ADRRS form (integrator.cpp:1246-1255)
            const uint32_t px = a.v.pixel[i], S = a.v.stride, P = a.n_pixels;
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float L = fminf(fmaxf(y[k], 1e-4f), 1e4f);
                acc += L * a.v.thp[k * S + i] / (a.film[k * P + px] + 0.01f);
            }
            r = acc / 3.0f;
            if (ep.out) {
#pragma unroll
                for (int j = 0; j < NOUT; ++j) ep.out[(size_t)j * ep.out_stride + i] = y[j];
            }
        }
        float q = rrs_window_clamp(ep.d, r);
        ep.q[i] = q;
        term = rrs_sum_term(ep.d, q);
    }
    block_atomic_add_u64(term, ep.sum);
}

template <int NL, int NOUT, int EPI, int ACT = ACT_RELU>
hipError_t launch_one(const EncodeArgs& a, int net, uint32_t n, const float* table, const float* W,
                      const EpiArgs& ep, hipStream_t s) {
    constexpr size_t smem = sizeof(float) * mlp_floats<NL, NOUT>();
    static std::atomic<bool> attr[kMaxDevices];
    hipError_t e = set_dyn_lds_once(attr, reinterpret_cast<const void*>(&k_net_fp32<NL, NOUT, EPI, ACT>), smem);
    if (e != hipSuccess) return e;
    dim3 grid((n + kBlock - 1) / kBlock);
    if (grid.x == 0) return hipSuccess;
    hipLaunchKernelGGL((k_net_fp32<NL, NOUT, EPI, ACT>), grid, dim3(kBlock), smem, s, a, net, n, table, W, ep);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_encode(const EncodeArgs& a, int net, uint32_t n, const float* table, float* enc,
                         hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_encode, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, net, n, table, enc);
    return hipGetLastError();
}

hipError_t launch_encode_levels_f32(const EncodeArgs& a, const float* table, uint32_t n, float2* out, uint32_t stride,
                                    hipStream_t s) {
    if (!n || !a.grid.n_levels) return hipSuccess;  // no grid: nothing to gather
    const uint32_t n_chunks = (n + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_encode_levels_f32, dim3(n_chunks * a.grid.n_levels), dim3(kChunk), 0, s, a, table, n,
                       n_chunks, out, stride);
    return hipGetLastError();
}

hipError_t launch_grid_indices(const EncodeArgs& a, uint32_t n, uint32_t* idx, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_grid_indices, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, n, idx);
    return hipGetLastError();
}

// Supported shapes: StatNet 4 hidden -> 6, RRSNet 3 hidden -> 1 (the benchmark
// configurations), the two-headed StatNets 3 and 2 hidden -> 6; other shapes report
// hipErrorInvalidValue.
hipError_t launch_net_fp32(const EncodeArgs& a, int net, int n_layers, int n_out, int epi, uint32_t n,
                           const float* table, const float* W, const EpiArgs& ep, hipStream_t s) {
    if (a.net_set == NRRS_NETS_SHIPPED) {
        // the shipped nets: 3 hidden LeakyReLU layers, StatNet -> 6, RRSNet -> 1
        if (n_layers == 4 && n_out == 6) {
            if (epi == EPI_RAW) return launch_one<4, 6, EPI_RAW, ACT_LEAKY>(a, net, n, table, W, ep, s);
            if (epi == EPI_ADRRS) return launch_one<4, 6, EPI_ADRRS, ACT_LEAKY>(a, net, n, table, W, ep, s);
        }
        if (n_layers == 4 && n_out == 1) {
            if (epi == EPI_RAW) return launch_one<4, 1, EPI_RAW, ACT_LEAKY>(a, net, n, table, W, ep, s);
            if (epi == EPI_RRS) return launch_one<4, 1, EPI_RRS, ACT_LEAKY>(a, net, n, table, W, ep, s);
        }
        return hipErrorInvalidValue;
    }
    if (n_layers == 5 && n_out == 6) {
        if (epi == EPI_RAW) return launch_one<5, 6, EPI_RAW>(a, net, n, table, W, ep, s);
        if (epi == EPI_ADRRS) return launch_one<5, 6, EPI_ADRRS>(a, net, n, table, W, ep, s);
    }
    // two-headed StatNets (PartShare, JustEncoding): the weights between the heads are zeros
    if (n_layers == 4 && n_out == 6) {
        if (epi == EPI_RAW) return launch_one<4, 6, EPI_RAW>(a, net, n, table, W, ep, s);
        if (epi == EPI_ADRRS) return launch_one<4, 6, EPI_ADRRS>(a, net, n, table, W, ep, s);
    }
    if (n_layers == 3 && n_out == 6) {
        if (epi == EPI_RAW) return launch_one<3, 6, EPI_RAW>(a, net, n, table, W, ep, s);
        if (epi == EPI_ADRRS) return launch_one<3, 6, EPI_ADRRS>(a, net, n, table, W, ep, s);
    }
    if (n_layers == 4 && n_out == 1) {
        if (epi == EPI_RAW) return launch_one<4, 1, EPI_RAW>(a, net, n, table, W, ep, s);
        if (epi == EPI_RRS) return launch_one<4, 1, EPI_RRS>(a, net, n, table, W, ep, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace nrrs
