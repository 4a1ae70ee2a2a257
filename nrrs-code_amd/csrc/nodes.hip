// Path-node buffer and the training samples built from it
// (NRRSPathNodesBuffer, src/render/nrrs/workqueue.h:106-189; nrrs_generate_training_data
// and nrrs_calculate_sample_weight, src/render/nrrs/train.cu:432-592).
//
// Records are appended in vertex order (node = size + i) instead of one atomic per vertex,
// and the samples are a stable compaction of the kept nodes (tile counts, a scan of the tile
// totals, a scatter that re-scans each tile), so the rows come out in node order. Duplicate
// siblings are decided by the lowest node index (atomicMin), not by atomic arrival order.
#include "device_common.hpp"

namespace nrrs {
namespace {

constexpr int kNodeThreads = 256;
constexpr int kTileThreads = kNodeTile;  // one thread per node of a tile
constexpr float kThpEps = 1e-5f;          // throughput below which a channel's L is 0

__global__ void k_nodes_reset(uint32_t* cnt) {
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0u;
}

__device__ __forceinline__ uint32_t count_of(uint32_t n, const uint32_t* n_dev) {
    return n_dev ? min(*n_dev, n) : n;
}

__global__ void k_nodes_record(NodeBuf nb, uint32_t n, const uint32_t* n_dev, nrrs_vertices v,
                               const int32_t* __restrict__ parent, const float* __restrict__ q, uint32_t depth,
                               int32_t* __restrict__ out) {
    const uint32_t m = count_of(n, n_dev), base = nb.cnt[0], C = nb.cap, S = v.stride;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const int32_t par = parent ? parent[i] : -1;
        const uint64_t j64 = (uint64_t)base + i;
        if (j64 >= C) {
            out[i] = par;
            continue;
        }
        const uint32_t j = (uint32_t)j64;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            nb.pos[c * C + j] = v.pos[c * S + i];
            nb.dir[c * C + j] = v.dir[c * S + i];
            nb.normal[c * C + j] = v.normal[c * S + i];
            nb.ad[c * C + j] = v.albedo_d[c * S + i];
            nb.as[c * C + j] = v.albedo_s[c * S + i];
            nb.thp[c * C + j] = v.thp[c * S + i];
            nb.L[c * C + j] = 0.0f;
        }
        nb.rough[j] = v.rough[i];
        nb.pixel[j] = v.pixel[i];
        nb.parent[j] = par;
        nb.q[j] = q ? q[i] : 1.0f;
        nb.depth[j] = depth;
        out[i] = (int32_t)j;
    }
}

// Runs after k_nodes_record on the same stream: every record of that launch read the same base.
__global__ void k_nodes_commit(NodeBuf nb, uint32_t n, const uint32_t* n_dev) {
    const uint64_t end = (uint64_t)nb.cnt[0] + count_of(n, n_dev);
    if (end > nb.cap) {
        nb.cnt[1] += (uint32_t)(end - nb.cap);
        nb.cnt[0] = nb.cap;
    } else {
        nb.cnt[0] = (uint32_t)end;
    }
}

// Radiance to a node and all its ancestors. A parent always has a lower index than its
// child, so the walk ends; a link that does not decrease stops it.
__global__ void k_nodes_radiance(NodeBuf nb, uint32_t n, const uint32_t* n_dev, const int32_t* __restrict__ node,
                                 const float* __restrict__ L, uint32_t Ls) {
    const uint32_t m = count_of(n, n_dev), size = nb.cnt[0], C = nb.cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        int32_t j = node[i];
        if (j < 0 || (uint32_t)j >= size) continue;
        const float l0 = L[i], l1 = L[Ls + i], l2 = L[2 * Ls + i];
        while (j >= 0) {
            atomicAdd(nb.L + j, l0);
            atomicAdd(nb.L + C + j, l1);
            atomicAdd(nb.L + 2 * C + j, l2);
            const int32_t p = nb.parent[j];
            if (p >= j) break;
            j = p;
        }
    }
}

__global__ void k_td_first_init(NodeBuf nb) {
    const uint32_t size = nb.cnt[0];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < size; j += gridDim.x * blockDim.x)
        nb.first[j] = 0xFFFFFFFFu;
}

__global__ void k_td_first(NodeBuf nb) {
    const uint32_t size = nb.cnt[0];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < size; j += gridDim.x * blockDim.x) {
        const int32_t p = nb.parent[j];
        if (p >= 0 && (uint32_t)p < size) atomicMin(nb.first + p, j);
    }
}

// Target radiance of node j and whether it becomes a sample (train.cu:470-519).
__device__ __forceinline__ bool node_sample(const NodeBuf& nb, uint32_t j, int ignore_zero_L, float* L) {
    const uint32_t C = nb.cap;
    bool nan = false;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float th = nb.thp[c * C + j];
        float l = th > kThpEps ? nb.L[c * C + j] / th : 0.0f;
        l = l < 0.0f ? 0.0f : l;  // max(L, 0) keeps a NaN
        L[c] = l;
        nan |= __builtin_isnan(l) || __builtin_isnan(nb.pos[c * C + j]) || __builtin_isnan(nb.dir[c * C + j]);
    }
    if (nan) return false;
    return !ignore_zero_L || (L[0] + L[1] + L[2]) / 3.0f > 0.0f;
}

__global__ __launch_bounds__(kTileThreads) void k_td_count(NodeBuf nb, int ignore_zero_L) {
    const uint32_t size = nb.cnt[0], j = blockIdx.x * kTileThreads + threadIdx.x;
    if (blockIdx.x * kTileThreads >= size) return;  // whole block: uniform
    float L[3];
    const uint32_t keep = (j < size && node_sample(nb, j, ignore_zero_L, L)) ? 1u : 0u;
    uint32_t total;
    block_exclusive_scan<kTileThreads>(keep, &total);
    if (threadIdx.x == 0) nb.tiles[blockIdx.x] = total;
}

// Exclusive scan of the tile counts in one block; *n_out = min(total, capacity).
__global__ __launch_bounds__(kTileThreads) void k_td_scan(NodeBuf nb, uint32_t capacity, uint32_t* n_out) {
    const uint32_t n_tiles = (nb.cnt[0] + kTileThreads - 1) / kTileThreads;
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < n_tiles; t0 += kTileThreads) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t v = t < n_tiles ? nb.tiles[t] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<kTileThreads>(v, &total);
        if (t < n_tiles) nb.tiles[t] = carry + ex;
        carry += total;
        __syncthreads();  // wsum of the scan is reused by the next chunk
    }
    if (threadIdx.x == 0) *n_out = min(carry, capacity);
}

struct TdArgs {
    nrrs_training_args a;
    uint32_t n_pixels;
};

__global__ __launch_bounds__(kTileThreads) void k_td_scatter(NodeBuf nb, TdArgs ta) {
    const nrrs_training_args& a = ta.a;
    const uint32_t size = nb.cnt[0], j = blockIdx.x * kTileThreads + threadIdx.x, C = nb.cap;
    if (blockIdx.x * kTileThreads >= size) return;
    float L[3];
    const bool keep = j < size && node_sample(nb, j, a.ignore_zero_L, L);
    uint32_t total;
    const uint32_t row = nb.tiles[blockIdx.x] + block_exclusive_scan<kTileThreads>(keep ? 1u : 0u, &total);
    if (!keep || row >= a.capacity) return;
    const uint32_t S = a.out.stride, px = nb.pixel[j];
    const int32_t p = nb.parent[j];
    const bool dup = a.mark_duplicates && p >= 0 && (uint32_t)p < size && nb.first[p] != j;
    const float q = nb.q[j];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.out.pos[c * S + row] = nb.pos[c * C + j];
        a.out.dir[c * S + row] = nb.dir[c * C + j];
        a.out.normal[c * S + row] = nb.normal[c * C + j];
        a.out.albedo_d[c * S + row] = nb.ad[c * C + j];
        a.out.albedo_s[c * S + row] = nb.as[c * C + j];
        const float th = nb.thp[c * C + j];
        a.out.thp[c * S + row] = a.copy_rrs_node ? th * q : th;
        a.target_L[(size_t)c * a.capacity + row] = L[c];
    }
    a.out.rough[row] = nb.rough[j];
    a.out.pixel[row] = px;
    const uint32_t P = ta.n_pixels;
    float ref = 0.0f;
    if (px < P) ref = (a.film_dev[px] + a.film_dev[P + px] + a.film_dev[2 * P + px]) / 3.0f;
    a.ref_mean[row] = ref;
    float e0 = 0.0f, e1 = 0.0f;
    if (a.error_dev) {
        // offsetFrame2Scaled (myFilm.h:6-15)
        const uint32_t W = a.frame_width, sc = a.error_scale;
        const uint32_t off = ((px % W) / sc) + ((px / W) / sc) * (W / sc);
        if (off < a.n_err) {
            e0 = a.error_dev[off];
            e1 = a.error_dev[(size_t)a.n_err + off];
        }
    }
    a.error[row] = e0;
    a.error_avg[row] = e1;
    a.q_target[row] = dup ? -1.0f : q;
    if (!dup && px < P) atomicAdd(a.pixel_samples + px, 1u);
}

__global__ void k_td_weights(nrrs_training_args a, uint32_t n_pixels) {
    const uint32_t m = *a.n_out_dev;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
        const uint32_t px = a.out.pixel[r];
        const uint32_t k = px < n_pixels ? a.pixel_samples[px] : 0u;
        a.sample_weight[r] = k ? 1.0f / (float)k : 0.0f;
    }
}

uint32_t grid_for(uint32_t n) { return std::max<uint32_t>(1u, std::min<uint32_t>((n + kNodeThreads - 1) / kNodeThreads, 8192)); }

}  // namespace

hipError_t launch_nodes_reset(const NodeBuf& nb, hipStream_t s) {
    hipLaunchKernelGGL(k_nodes_reset, dim3(1), dim3(64), 0, s, nb.cnt);
    return hipGetLastError();
}

hipError_t launch_nodes_record(const NodeBuf& nb, uint32_t n, const uint32_t* n_dev, const nrrs_vertices& v,
                               const int32_t* parent, const float* q, uint32_t depth, int32_t* out, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_nodes_record, dim3(grid_for(n)), dim3(kNodeThreads), 0, s, nb, n, n_dev, v, parent, q, depth,
                       out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_nodes_commit, dim3(1), dim3(1), 0, s, nb, n, n_dev);
    return hipGetLastError();
}

hipError_t launch_nodes_radiance(const NodeBuf& nb, uint32_t n, const uint32_t* n_dev, const int32_t* node,
                                 const float* L, uint32_t L_stride, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_nodes_radiance, dim3(grid_for(n)), dim3(kNodeThreads), 0, s, nb, n, n_dev, node, L, L_stride);
    return hipGetLastError();
}

hipError_t launch_training_data(const NodeBuf& nb, const nrrs_training_args& a, uint32_t n_pixels, hipStream_t s) {
    const uint32_t tiles = (nb.cap + kNodeTile - 1) / kNodeTile;
    const uint32_t g = grid_for(nb.cap);
    if (a.mark_duplicates) {
        hipLaunchKernelGGL(k_td_first_init, dim3(g), dim3(kNodeThreads), 0, s, nb);
        hipLaunchKernelGGL(k_td_first, dim3(g), dim3(kNodeThreads), 0, s, nb);
    }
    hipLaunchKernelGGL(k_td_count, dim3(tiles), dim3(kTileThreads), 0, s, nb, a.ignore_zero_L);
    hipLaunchKernelGGL(k_td_scan, dim3(1), dim3(kTileThreads), 0, s, nb, a.capacity, a.n_out_dev);
    TdArgs ta{a, n_pixels};
    hipLaunchKernelGGL(k_td_scatter, dim3(tiles), dim3(kTileThreads), 0, s, nb, ta);
    hipLaunchKernelGGL(k_td_weights, dim3(grid_for(a.capacity)), dim3(kNodeThreads), 0, s, a, n_pixels);
    return hipGetLastError();
}

}  // namespace nrrs
