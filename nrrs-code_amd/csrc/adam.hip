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

__global__ void k_check_finite(const float* __restrict__ g, uint64_t n, const uint16_t* __restrict__ glp,
                               uint64_t n_lp, uint32_t* flag) {
    bool bad = false;
    const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = i0; i < n; i += step) bad |= !isfinite(g[i]);
    if (glp) {
        // 8 halves per 16-byte load (the buffer is 16-byte aligned), then the tail
        const uint4* g8 = reinterpret_cast<const uint4*>(glp);
        const uint64_t n8 = n_lp / 8;
        for (uint64_t k = i0; k < n8; k += step) {
            const uint4 q = g8[k];
            bad |= half_pair_bad(q.x) | half_pair_bad(q.y) | half_pair_bad(q.z) | half_pair_bad(q.w);
        }
        for (uint64_t i = 8 * n8 + i0; i < n_lp; i += step) bad |= (glp[i] & 0x7c00u) == 0x7c00u;
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

__global__ void k_adam(AdamArgs a) {
    const uint32_t skip = a.state[3];
    const float inv_s = 1.0f / *a.loss_scale;
    const uint32_t t = a.state[0] + 1;
    const float c1 = 1.0f - powf(a.beta1, (float)t), c2 = 1.0f - powf(a.beta2, (float)t);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * blockDim.x) {
        float g;
        if (a.grad_lp && i < a.n_grid) {
            g = (float)__builtin_bit_cast(_Float16, a.grad_lp[i]);
            a.grad_lp[i] = 0;
        } else {
            g = a.grad[i];
            a.grad[i] = 0.0f;
        }
        if (skip) continue;
        float p = a.params[i];
        g = g * inv_s + a.l2 * p;
        float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * g;
        float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * g * g;
        a.m[i] = m;
        a.v[i] = v;
        p -= a.lr * (m / c1) / (sqrtf(v / c2) + a.eps);
        a.params[i] = p;
        if (a.grid_lp && i < a.n_grid) {
            const uint16_t b = a.lp_precision == NRRS_BF16 ? __builtin_bit_cast(uint16_t, (__bf16)p)
                                                            : __builtin_bit_cast(uint16_t, (_Float16)p);
            a.grid_lp[i] = b;
        }
    }
}

__global__ void k_adam_finish(AdamArgs a) {
    uint32_t* st = a.state;
    if (st[3]) {
        st[2] += 1;
        st[1] = 0;
        if (a.dynamic_scale) *a.loss_scale = fmaxf(*a.loss_scale * 0.5f, 1.0f);
    } else {
        st[0] += 1;
        st[1] += 1;
        if (a.dynamic_scale && st[1] >= a.growth_interval) {
            *a.loss_scale = fminf(*a.loss_scale * 2.0f, 65536.0f * 65536.0f);
            st[1] = 0;
        }
    }
    st[3] = 0;
}

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

hipError_t launch_adam(const AdamArgs& a, hipStream_t s) {
    // with an FP16 grid gradient only the MLP part of `grad` is live
    const uint64_t off = a.grad_lp ? a.n_grid : 0;
    hipLaunchKernelGGL(k_check_finite, dim3(1024), dim3(256), 0, s, a.grad + off, a.n - off, a.grad_lp,
                       a.grad_lp ? a.n_grid : 0, a.state + 3);
    hipLaunchKernelGGL(k_adam, dim3(2048), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_adam_finish, dim3(1), dim3(1), 0, s, a);
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
