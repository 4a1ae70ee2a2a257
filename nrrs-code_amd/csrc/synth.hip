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

__global__ __launch_bounds__(256) void k_synth(uint32_t n, uint64_t seed, uint32_t n_pixels,
                                               nrrs_vertices_out o) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t S = o.stride;
    const SynthVertex v = synth_vertex(seed, i);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o.pos[k * S + i] = v.pos[k];
        o.dir[k * S + i] = v.dir[k];
        o.normal[k * S + i] = v.nrm[k];
        o.albedo_d[k * S + i] = v.ad[k];
        o.albedo_s[k * S + i] = v.as[k];
        o.thp[k * S + i] = v.thp[k];
    }
    o.rough[i] = v.rough;
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
