// Synthetic path vertices of the benchmark workloads, drawn on the device.
// Recipe (fields on separate PCG32 streams of one seed; vertex i draws fixed
// output numbers of each stream, so every vertex is independent):
//   stream 0: 64 triangles, 9 corner coordinates each, uniform in [0,1]
//   stream 1: (triangle = x0 & 63, barycentrics from u1, u2) at outputs 3i..3i+2
//   stream 2/3: direction / normal uniform on the sphere, outputs 2i, 2i+1
//   stream 4: roughness U[0.05, 1]
//   stream 5/6: diffuse / specular albedo U[0,1]^3, outputs 3i+k
//   stream 7: throughput log-normal(-1, 1) per channel, Box-Muller from 6i+2k, 6i+2k+1
//   pixel = i mod n_pixels
#include "device_common.hpp"

namespace nrrs {
namespace {

__device__ __forceinline__ uint32_t draw(uint64_t seed, uint64_t stream, uint64_t pos) {
    return pcg_output(pcg_state_at(seed, stream, pos));
}

__global__ __launch_bounds__(256) void k_synth(uint32_t n, uint64_t seed, uint32_t n_pixels,
                                               nrrs_vertices_out o) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t S = o.stride;
    const float two_pi = 6.283185307179586f;
    {
        uint64_t st = pcg_state_at(seed, 1, 3ull * i);
        uint32_t a = pcg_next(st, 1), b = pcg_next(st, 1), c = pcg_next(st, 1);
        uint64_t tst = pcg_state_at(seed, 0, 9ull * (a & 63u));
        float T[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) T[k] = u24(pcg_next(tst, 0));
        float su = sqrtf(u24(b));
        float b0 = 1.0f - su, b1 = u24(c) * su, b2 = 1.0f - b0 - b1;
#pragma unroll
        for (int k = 0; k < 3; ++k) o.pos[k * S + i] = b0 * T[k] + b1 * T[3 + k] + b2 * T[6 + k];
    }
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        float* dst = f ? o.normal : o.dir;
        uint64_t st = pcg_state_at(seed, 2 + f, 2ull * i);
        float u1 = u24(pcg_next(st, 2 + f)), u2 = u24(pcg_next(st, 2 + f));
        float z = 1.0f - 2.0f * u1;
        float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float ph = two_pi * u2;
        dst[i] = rr * cosf(ph);
        dst[S + i] = rr * sinf(ph);
        dst[2 * S + i] = z;
    }
    o.rough[i] = 0.05f + 0.95f * u24(draw(seed, 4, i));
    {
        uint64_t sa = pcg_state_at(seed, 5, 3ull * i), ss = pcg_state_at(seed, 6, 3ull * i),
                 st = pcg_state_at(seed, 7, 6ull * i);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            o.albedo_d[k * S + i] = u24(pcg_next(sa, 5));
            o.albedo_s[k * S + i] = u24(pcg_next(ss, 6));
            float ua = u24(pcg_next(st, 7)), ub = u24(pcg_next(st, 7));
            float nrm = sqrtf(-2.0f * logf(1.0f - ua)) * cosf(two_pi * ub);
            o.thp[k * S + i] = expf(-1.0f + nrm);
        }
    }
    o.pixel[i] = i % n_pixels;
}

}  // namespace

hipError_t launch_synth(uint32_t n, uint64_t seed, uint32_t n_pixels, const nrrs_vertices_out& o,
                        hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_synth, dim3((n + 255) / 256), dim3(256), 0, s, n, seed, n_pixels, o);
    return hipGetLastError();
}

}  // namespace nrrs
