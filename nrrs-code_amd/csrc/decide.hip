// RRS normalisation, split decision and stable survivor compaction
// (src/render/nrrs/integrator.cpp:1301-1460, src/render/common/commonworkqueue.h:19-25).
//
// The reference appends survivors with one global fetch_add per vertex, so its
// queue order is nondeterministic. Here the queue is a stable compaction:
// specular stids first (in the order given), then for each inference index i in
// ascending order, s_i copies of stid_i. It is built by a tile reduce, a scan of
// the tile totals and a scatter that re-scans each tile with wavefront prefix sums.
#include <mutex>

#include "device_common.hpp"

namespace nrrs {
namespace {

constexpr int kTile = 256;
constexpr int kScanThreads = 1024;

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

struct SplitArgs {
    const float* q;
    uint32_t n;
    const unsigned long long* sum;
    const uint32_t* n_spec;
    uint64_t seed, stream;
    float* qn;
    float* qn_stid;
    const uint32_t* stid;
    uint8_t* s;
    uint32_t* tile_sums;
    DecideParams d;
};

__global__ __launch_bounds__(kTile) void k_split(SplitArgs a) {
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    uint32_t si = 0;
    if (i < a.n) {
        // normalizeFactor (integrator.cpp:1376-1377) from the exact sum
        double sum = (double)(*a.sum) * 0x1p-32;
        float sum_f = (float)sum;
        float rcp = sum_f > 0.0f ? 1.0f / sum_f : 0.0f;
        uint32_t nspec = a.n_spec ? *a.n_spec : 0u;
        float nf = (float)(a.d.n_pixels - nspec) * rcp;
        nf = nf * (a.d.use_ceil ? 1.0f : a.d.rate);
        // integrator.cpp:1391-1397
        float r = a.q[i];
        r = a.d.no_upscale ? r * (nf < 1.0f ? nf : 1.0f) : r * nf;
        if (a.qn) a.qn[i] = r;
        if (a.qn_stid) a.qn_stid[a.stid ? a.stid[i] : i] = r;
        float u = u24(pcg_output(pcg_state_at(a.seed, a.stream, i)));
        // integrator.cpp:1415-1416, 1454
        int s = (int)r;
        float left = r - (float)s;
        s += (u <= left) ? 1 : 0;
        si = (uint32_t)min(s, (int)a.d.split_cap);
        a.s[i] = (uint8_t)si;
    }
    uint32_t total;
    block_exclusive_scan<kTile>(si, &total);
    if (threadIdx.x == 0) a.tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const uint32_t* __restrict__ tile_sums,
                                                             uint32_t n_tiles,
                                                             const uint32_t* __restrict__ n_spec,
                                                             uint32_t capacity,
                                                             uint32_t* __restrict__ tile_offsets,
                                                             uint32_t* __restrict__ n_survivors,
                                                             uint32_t* __restrict__ overflow) {
    const uint32_t per = (n_tiles + kScanThreads - 1) / kScanThreads;
    const uint32_t t0 = threadIdx.x * per;
    uint32_t local = 0;
    for (uint32_t t = t0; t < t0 + per && t < n_tiles; ++t) local += tile_sums[t];
    uint32_t total;
    uint32_t run = block_exclusive_scan<kScanThreads>(local, &total);
    const uint32_t nspec = n_spec ? *n_spec : 0u;
    run += nspec;
    for (uint32_t t = t0; t < t0 + per && t < n_tiles; ++t) {
        tile_offsets[t] = run;
        run += tile_sums[t];
    }
    if (threadIdx.x == 0) {
        uint64_t all = (uint64_t)total + nspec;
        uint32_t kept = all < capacity ? (uint32_t)all : capacity;
        *n_survivors = kept;
        if (overflow) *overflow = (uint32_t)(all - kept);
    }
}

__global__ __launch_bounds__(kTile) void k_scatter(const uint8_t* __restrict__ s, uint32_t n,
                                                   const uint32_t* __restrict__ stid,
                                                   const uint32_t* __restrict__ tile_offsets,
                                                   uint32_t capacity, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    const uint32_t si = i < n ? s[i] : 0u;
    uint32_t total;
    uint32_t pos = tile_offsets[blockIdx.x] + block_exclusive_scan<kTile>(si, &total);
    if (si) {
        const uint32_t id = stid ? stid[i] : i;
        for (uint32_t c = 0; c < si; ++c)
            if (pos + c < capacity) out[pos + c] = id;
    }
}

__global__ void k_spec_copy(const uint32_t* __restrict__ spec, const uint32_t* __restrict__ n_spec,
                            uint32_t capacity, uint32_t* __restrict__ out) {
    const uint32_t ns = *n_spec;
    const uint32_t m = ns < capacity ? ns : capacity;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
        out[k] = spec[k];
}

// Normalisation sum of a given clamped q (decision-only entry point).
__global__ __launch_bounds__(kTile) void k_qsum(const float* __restrict__ q, uint32_t n, DecideParams d,
                                                unsigned long long* sum) {
    const uint32_t i = blockIdx.x * kTile + threadIdx.x;
    uint64_t term = i < n ? rrs_sum_term(d, q[i]) : 0ull;
    block_atomic_add_u64(term, sum);
}

// handleIntersections' consumer loop (integrator.cpp:772-794) on SoA rows.
__global__ void k_consume(uint32_t capacity, const uint32_t* __restrict__ n_dev, const uint32_t* __restrict__ surv,
                          nrrs_vertices sc, const float* __restrict__ rrs, const int32_t* __restrict__ node_in,
                          nrrs_vertices_out o, int32_t* __restrict__ node_out, uint32_t* __restrict__ stid_out) {
    const uint32_t n = min(*n_dev, capacity), S = sc.stride, So = o.stride;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t sid = surv[k];
        if (sid >= S) continue;
        const float r = rrs[sid];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            o.pos[c * So + k] = sc.pos[c * S + sid];
            o.dir[c * So + k] = sc.dir[c * S + sid];
            o.normal[c * So + k] = sc.normal[c * S + sid];
            o.albedo_d[c * So + k] = sc.albedo_d[c * S + sid];
            o.albedo_s[c * So + k] = sc.albedo_s[c * S + sid];
            o.thp[c * So + k] = sc.thp[c * S + sid] / r;
        }
        o.rough[k] = sc.rough[sid];
        o.pixel[k] = sc.pixel[sid];
        if (node_out) node_out[k] = node_in ? node_in[sid] : -1;
        if (stid_out) stid_out[k] = sid;
    }
}

}  // namespace

hipError_t launch_consume(uint32_t capacity, const uint32_t* n_dev, const uint32_t* surv, const nrrs_vertices& sc,
                          const float* rrs, const int32_t* node_in, const nrrs_vertices_out& o, int32_t* node_out,
                          uint32_t* stid_out, hipStream_t s) {
    if (!capacity) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((capacity + 255) / 256, 8192);
    hipLaunchKernelGGL(k_consume, dim3(blocks), dim3(256), 0, s, capacity, n_dev, surv, sc, rrs, node_in, o, node_out,
                       stid_out);
    return hipGetLastError();
}

namespace {

// Scratch of one decision: the exact sum, the per-tile survivor totals and first slots, and
// the split counts when the caller keeps none.
struct DecideScratch {
    unsigned long long* sum;
    uint32_t* tile_sums;
    uint32_t* tile_offsets;
    uint8_t* s;
    uint32_t max_tiles;
};

hipError_t qsum_into(const DecideParams& d, const float* q, uint32_t n, unsigned long long* sum, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, sum);
    if (n) hipLaunchKernelGGL(k_qsum, dim3((n + kTile - 1) / kTile), dim3(kTile), 0, s, q, n, d, sum);
    return hipGetLastError();
}

hipError_t decide_with(const DecideParams& d, const DecideScratch& sc, const nrrs_decide_args& a, hipStream_t s) {
    const uint32_t tiles = (a.n + kTile - 1) / kTile;
    if (tiles > sc.max_tiles) return hipErrorInvalidValue;
    uint8_t* split = a.split_dev ? a.split_dev : sc.s;
    if (tiles) {
        SplitArgs sa;
        sa.q = a.q_dev;
        sa.n = a.n;
        sa.sum = sc.sum;
        sa.n_spec = a.n_spec_dev;
        sa.seed = a.pcg_seed;
        sa.stream = a.pcg_stream;
        sa.qn = a.qn_dev;
        sa.qn_stid = a.qn_stid_dev;
        sa.stid = a.v.stid;
        sa.s = split;
        sa.tile_sums = sc.tile_sums;
        sa.d = d;
        hipLaunchKernelGGL(k_split, dim3(tiles), dim3(kTile), 0, s, sa);
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s, sc.tile_sums, tiles, a.n_spec_dev, a.capacity,
                       sc.tile_offsets, a.n_survivors_dev, a.overflow_dev);
    if (a.n_spec_dev && a.spec_stid_dev)
        hipLaunchKernelGGL(k_spec_copy, dim3(64), dim3(256), 0, s, a.spec_stid_dev, a.n_spec_dev, a.capacity,
                           a.survivors_dev);
    if (tiles)
        hipLaunchKernelGGL(k_scatter, dim3(tiles), dim3(kTile), 0, s, split, a.n, a.v.stid, sc.tile_offsets,
                           a.capacity, a.survivors_dev);
    return hipGetLastError();
}

DecideScratch ctx_scratch(nrrs_ctx* ctx) {
    return DecideScratch{ctx->sum_dev, ctx->tile_sums, ctx->tile_offsets, ctx->s_scratch, ctx->max_tiles};
}

// Scratch of the allocation-free decision (nrrs_rrs_decide_static): one array per device, in
// the code object, [0, 256) the sum, then the tile totals, the tile first slots and the split
// counts of kStaticDecideMax vertices (about 4.3 MB). Host code uses the address that
// hipGetSymbolAddress returns on each device, never the array's name.
constexpr uint32_t kStaticTiles = kStaticDecideMax / kTile;
constexpr size_t kStaticTileOff = 256, kStaticOffOff = kStaticTileOff + 4 * (size_t)kStaticTiles,
                 kStaticSplitOff = kStaticOffOff + 4 * (size_t)kStaticTiles,
                 kStaticScratchBytes = kStaticSplitOff + kStaticDecideMax;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

std::once_flag g_scratch_once[kMaxDevices];
uint8_t* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

// The device address of `scratch` on the current device (looked up once per device).
hipError_t static_scratch(uint8_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::call_once(g_scratch_once[dev], [dev] {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? static_cast<uint8_t*>(p) : nullptr;
    });
    if (g_scratch_err[dev] != hipSuccess) return g_scratch_err[dev];
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

hipError_t launch_qsum(nrrs_ctx* ctx, const float* q, uint32_t n, hipStream_t s) {
    return qsum_into(ctx->dec, q, n, ctx->sum_dev, s);
}

hipError_t launch_zero_u64(unsigned long long* p, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_decide(nrrs_ctx* ctx, const nrrs_decide_args& a, hipStream_t s) {
    return decide_with(ctx->dec, ctx_scratch(ctx), a, s);
}

hipError_t launch_decide_static(const DecideParams& d, const nrrs_decide_args& a, hipStream_t s) {
    if (a.n > kStaticDecideMax) return hipErrorInvalidValue;
    uint8_t* base = nullptr;
    hipError_t e = static_scratch(&base);
    if (e != hipSuccess) return e;
    DecideScratch sc;
    sc.sum = reinterpret_cast<unsigned long long*>(base);
    sc.tile_sums = reinterpret_cast<uint32_t*>(base + kStaticTileOff);
    sc.tile_offsets = reinterpret_cast<uint32_t*>(base + kStaticOffOff);
    sc.s = base + kStaticSplitOff;
    sc.max_tiles = kStaticTiles;
    if ((e = qsum_into(d, a.q_dev, a.n, sc.sum, s)) != hipSuccess) return e;
    return decide_with(d, sc, a, s);
}

}  // namespace nrrs
