// Per-pixel error that feeds the RRS loss: running weighted radiance (train.cu:344-363),
// relative squared error with clamp and sqrt(spp) scaling (train.cu:365-430) and its sum
// (integrator.cpp:2466-2471). Elementwise over [3][pixels] planes; the sum is a fixed-order
// FP64 reduction (per-block partials, then one block), so it is reproducible.
#include "host_common.hpp"

namespace nrrs {
namespace {

constexpr int kErrThreads = 256;
constexpr int kErrBlocks = 256;

__global__ void k_error_blend(uint32_t n, const float* __restrict__ Lf, const uint32_t* __restrict__ spp, float blend,
                              float* __restrict__ wL) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float inv = 1.0f / (float)spp[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t k = (size_t)c * n + i;
            wL[k] = wL[k] * (1.0f - blend) + (Lf[k] * inv) * blend;
        }
    }
}

__device__ __forceinline__ double block_sum_f64(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
    return t;
}

__global__ __launch_bounds__(kErrThreads) void k_error_compute(uint32_t n, const float* __restrict__ ref,
                                                               const float* __restrict__ wL,
                                                               const float* __restrict__ spp, int clamp,
                                                               float* __restrict__ err, double* __restrict__ part) {
    __shared__ double red[kErrThreads / 64];
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float e = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float r = ref[(size_t)c * n + i], d = wL[(size_t)c * n + i] - r;
            e += d * d / (r * r + 1e-2f);
        }
        e = e / 3.0f;
        if (clamp) e = fminf(e, 100.0f);
        err[i] = e;
        const float e2 = spp ? e * sqrtf(spp[i]) : e;
        err[(size_t)n + i] = e2;
        acc += (double)e2;
    }
    const double t = block_sum_f64(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kErrThreads) void k_error_sum(const double* __restrict__ part, int n_part,
                                                           float* __restrict__ out) {
    __shared__ double red[kErrThreads / 64];
    double acc = 0.0;
    for (int k = threadIdx.x; k < n_part; k += blockDim.x) acc += part[k];
    const double t = block_sum_f64(acc, red);
    if (threadIdx.x == 0) *out = (float)t;
}

}  // namespace

hipError_t launch_error_blend(uint32_t n, const float* Lf, const uint32_t* spp, float blend, float* wL, hipStream_t s) {
    if (!n) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((n + kErrThreads - 1) / kErrThreads, 4096);
    hipLaunchKernelGGL(k_error_blend, dim3(blocks), dim3(kErrThreads), 0, s, n, Lf, spp, blend, wL);
    return hipGetLastError();
}

size_t error_partial_count() { return kErrBlocks; }

hipError_t launch_error_compute(uint32_t n, const float* ref, const float* wL, const float* spp, int clamp, float* err,
                                double* part, float* sum, hipStream_t s) {
    hipLaunchKernelGGL(k_error_compute, dim3(kErrBlocks), dim3(kErrThreads), 0, s, n, ref, wL, spp, clamp, err, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !sum) return e;
    hipLaunchKernelGGL(k_error_sum, dim3(1), dim3(kErrThreads), 0, s, part, kErrBlocks, sum);
    return hipGetLastError();
}

}  // namespace nrrs
