// Adam on the FP32 master parameters (Kingma & Ba 2015; L2 regularisation added
// to the gradient), with dynamic loss scaling for FP16/BF16 contexts: a gradient
// with a NaN/Inf skips the step and halves the scale; `growth_interval` finite
// steps in a row double it. Everything stays on the device (no host sync).
#include "device_common.hpp"
#include "train_common.hpp"

namespace nrrs {
namespace {

// FP16 infinity / NaN: all exponent bits set (each half of a 32-bit word)
__device__ __forceinline__ bool half_pair_bad(uint32_t w) {
    return (w & 0x7c00u) == 0x7c00u || (w & 0x7c000000u) == 0x7c000000u;
}

// scan_flag: null, or the flag that says whether the FP16 table gradient needs the scan (its
// producers report non-finite entries themselves otherwise).
__global__ void k_check_finite(const float* __restrict__ g, uint64_t n, const uint16_t* __restrict__ glp,
                               uint64_t n_lp, uint32_t* flag, const uint32_t* scan_flag) {
    bool bad = false;
    const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = i0; i < n; i += step) bad |= !isfinite(g[i]);
    if (glp && (!scan_flag || *scan_flag)) {
        // 8 halves per 16-byte load (the buffer is 16-byte aligned), four loads in flight per
        // thread, then the tail
        const uint4* g8 = reinterpret_cast<const uint4*>(glp);
        const uint64_t n8 = n_lp / 8;
        auto quad_bad = [](const uint4& q) {
            return half_pair_bad(q.x) | half_pair_bad(q.y) | half_pair_bad(q.z) | half_pair_bad(q.w);
        };
        uint64_t k = i0;
        for (; k + 3 * step < n8; k += 4 * step) {
            const uint4 q0 = g8[k], q1 = g8[k + step], q2 = g8[k + 2 * step], q3 = g8[k + 3 * step];
            bad |= quad_bad(q0) | quad_bad(q1) | quad_bad(q2) | quad_bad(q3);
        }
        for (; k < n8; k += step) bad |= quad_bad(g8[k]);
        for (uint64_t i = 8 * n8 + i0; i < n_lp; i += step) bad |= (glp[i] & 0x7c00u) == 0x7c00u;
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

struct AdamStep {
    float inv_s, c1, c2;
    float e_keep, e_take;  // Ema weights of the old average and of the new parameters
};

__device__ __forceinline__ AdamStep adam_step(const AdamArgs& a) {
    const uint32_t t = a.state[0] + 1;
    AdamStep k{1.0f / *a.loss_scale, 1.0f - powf(a.beta1, (float)t), 1.0f - powf(a.beta2, (float)t), 0.0f, 1.0f};
    if (a.ema) {
        // debiased moving average (the bias correction of Adam's moments applied to the weights)
        const float d_old = 1.0f - powf(a.ema_decay, (float)(t - 1)), d_new = 1.0f - powf(a.ema_decay, (float)t);
        k.e_keep = a.ema_decay * d_old / d_new;
        k.e_take = (1.0f - a.ema_decay) / d_new;
    }
    return k;
}

__device__ __forceinline__ float ema_one(const AdamStep& k, float e, float p) { return k.e_keep * e + k.e_take * p; }

// One parameter: gradient g (loss-scaled), master p and moments m, v updated in place.
__device__ __forceinline__ void adam_one(const AdamArgs& a, const AdamStep& k, float g, float& p, float& m, float& v) {
    g = g * k.inv_s + a.l2 * p;
    m = a.beta1 * m + (1.0f - a.beta1) * g;
    v = a.beta2 * v + (1.0f - a.beta2) * g * g;
    p -= a.lr * (m / k.c1) / (sqrtf(v / k.c2) + a.eps);
}

__device__ __forceinline__ uint16_t lp_bits(const AdamArgs& a, float p) {
    return a.lp_precision == NRRS_BF16 ? __builtin_bit_cast(uint16_t, (__bf16)p) : __builtin_bit_cast(uint16_t, (_Float16)p);
}

// One parameter i (gradient read and cleared, update, shadow refreshed).
__device__ __forceinline__ void adam_scalar(const AdamArgs& a, const AdamStep& k, uint32_t skip, uint64_t i) {
    float g;
    if (a.grad_lp && i < a.n_grid) {
        g = (float)__builtin_bit_cast(_Float16, a.grad_lp[i]);
        a.grad_lp[i] = 0;
    } else {
        g = a.grad[i];
        a.grad[i] = 0.0f;
    }
    if (skip) return;
    float p = a.params[i], m = a.m[i], v = a.v[i];
    adam_one(a, k, g, p, m, v);
    a.m[i] = m;
    a.v[i] = v;
    a.params[i] = p;
    if (a.grid_lp && i < a.n_grid) a.grid_lp[i] = lp_bits(a, p);
    if (a.ema) {
        const float e = ema_one(k, a.ema[i], p);
        a.ema[i] = e;
        if (a.ema_lp && i < a.n_grid) a.ema_lp[i] = lp_bits(a, e);
    }
}

// Any n_grid: one parameter per thread and iteration. The 16-level grids with base resolutions
// 8-23 and T = 2^17 or 2^19 all have even entry counts, so n_grid is a multiple of 4 and k_adam
// runs there; this kernel keeps any other layout correct.
__global__ void k_adam_scalar(AdamArgs a) {
    const uint32_t skip = a.state[3];
    const AdamStep k = adam_step(a);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * blockDim.x)
        adam_scalar(a, k, skip, i);
}

// Four consecutive parameters per thread and iteration: 16-byte loads and stores of the FP32
// arrays, 8-byte ones of the FP16 gradient and the low-precision shadow. Needs n_grid % 4 == 0
// (a group is all grid or all MLP); the last n % 4 parameters are taken one by one.
__global__ void k_adam(AdamArgs a) {
    const uint32_t skip = a.state[3];
    const AdamStep k = adam_step(a);
    const uint64_t n4 = a.n / 4, stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n4; c += stride) {
        const uint64_t i = 4 * c;
        const bool lp = a.grad_lp && i < a.n_grid;
        float g[4];
        if (lp) {
            const uint2 w = reinterpret_cast<const uint2*>(a.grad_lp)[c];
            const uint32_t h[4] = {w.x & 0xffffu, w.x >> 16, w.y & 0xffffu, w.y >> 16};
#pragma unroll
            for (int u = 0; u < 4; ++u) g[u] = (float)__builtin_bit_cast(_Float16, (uint16_t)h[u]);
            reinterpret_cast<uint2*>(a.grad_lp)[c] = make_uint2(0u, 0u);
        } else {
            const float4 w = reinterpret_cast<const float4*>(a.grad)[c];
            g[0] = w.x, g[1] = w.y, g[2] = w.z, g[3] = w.w;
            reinterpret_cast<float4*>(a.grad)[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        if (skip) continue;
        float4 p = reinterpret_cast<const float4*>(a.params)[c], m = reinterpret_cast<const float4*>(a.m)[c],
               v = reinterpret_cast<const float4*>(a.v)[c];
        adam_one(a, k, g[0], p.x, m.x, v.x);
        adam_one(a, k, g[1], p.y, m.y, v.y);
        adam_one(a, k, g[2], p.z, m.z, v.z);
        adam_one(a, k, g[3], p.w, m.w, v.w);
        reinterpret_cast<float4*>(a.m)[c] = m;
        reinterpret_cast<float4*>(a.v)[c] = v;
        reinterpret_cast<float4*>(a.params)[c] = p;
        if (a.grid_lp && i < a.n_grid)
            reinterpret_cast<uint2*>(a.grid_lp)[c] = make_uint2(lp_bits(a, p.x) | (uint32_t)lp_bits(a, p.y) << 16,
                                                                lp_bits(a, p.z) | (uint32_t)lp_bits(a, p.w) << 16);
        if (a.ema) {
            float4 e = reinterpret_cast<const float4*>(a.ema)[c];
            e = make_float4(ema_one(k, e.x, p.x), ema_one(k, e.y, p.y), ema_one(k, e.z, p.z), ema_one(k, e.w, p.w));
            reinterpret_cast<float4*>(a.ema)[c] = e;
            if (a.ema_lp && i < a.n_grid)
                reinterpret_cast<uint2*>(a.ema_lp)[c] = make_uint2(lp_bits(a, e.x) | (uint32_t)lp_bits(a, e.y) << 16,
                                                                   lp_bits(a, e.z) | (uint32_t)lp_bits(a, e.w) << 16);
        }
    }
    for (uint64_t i = 4 * n4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride)
        adam_scalar(a, k, skip, i);
}

__global__ void k_adam_finish(AdamArgs a) { adam_finish_state(a); }

// Paired table from the two nets' FP16/BF16 grids: entry e = {StatNet word e, RRSNet word e}.
// Two entries per thread: 8-byte reads, one 16-byte write (full lines, no partial writes).
__global__ void k_pack_pair(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint4* __restrict__ pair,
                            uint64_t n_entries) {
    const uint64_t n2 = n_entries / 2;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint2 x = reinterpret_cast<const uint2*>(a)[k], y = reinterpret_cast<const uint2*>(b)[k];
        pair[k] = make_uint4(x.x, y.x, x.y, y.y);
    }
    if ((n_entries & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t* p = reinterpret_cast<uint32_t*>(pair) + 2 * (n_entries - 1);
        p[0] = a[n_entries - 1];
        p[1] = b[n_entries - 1];
    }
}

__global__ void k_fill(float* p, float v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

hipError_t launch_adam(const AdamArgs& a, hipStream_t s, bool finish) {
    // with an FP16 grid gradient only the MLP part of `grad` is live
    const uint64_t off = a.grad_lp ? a.n_grid : 0;
    // with producer-side reporting the table scan is the rare case (a full segment): fewer blocks
    hipLaunchKernelGGL(k_check_finite, dim3(a.scan_on_flag ? 256 : 2048), dim3(256), 0, s, a.grad + off, a.n - off,
                       a.grad_lp, a.grad_lp ? a.n_grid : 0, a.state + 3, a.scan_on_flag ? a.state + 5 : nullptr);
    if (a.n_grid % 4 == 0)
        hipLaunchKernelGGL(k_adam, dim3(2048), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_adam_scalar, dim3(2048), dim3(256), 0, s, a);
    if (finish) hipLaunchKernelGGL(k_adam_finish, dim3(1), dim3(1), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pack_pair(const uint16_t* stat_lp, const uint16_t* rrs_lp, uint32_t* pair, uint64_t n_entries,
                            hipStream_t s) {
    if (n_entries)
        hipLaunchKernelGGL(k_pack_pair, dim3(2048), dim3(256), 0, s, reinterpret_cast<const uint32_t*>(stat_lp),
                           reinterpret_cast<const uint32_t*>(rrs_lp), reinterpret_cast<uint4*>(pair), n_entries);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float v, uint64_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

}  // namespace nrrs
