// Batched uneven wavefront: all bounces of a frame in one persistent launch.
//
// Bounce b processes the queue of path ids left by bounce b-1 (bounce 0: every
// primary path). For queue slot k the path vertex is drawn from the synthetic
// recipe with seed_b = seed + b * 0x9E3779B97F4A7C15 (it stands in for the
// closest-hit program that produces the inference queue, device.cu:97-108),
// then: level-major grid encodings of both nets (every workgroup walks the (net, level,
// vertex chunk) items in the same order, so the GPU gathers from one level's table at a
// time and it stays in L2) -> StatNet -> RRSNet (statistics passed in registers) -> q -> exact
// normalisation sum -> split decision -> stable compaction of the survivors into
// the next queue. Queue sizes never leave the device, so the host neither waits
// nor reads anything back between bounces (the reference reads the queue size back
// every bounce, integrator.cpp:1533-1555).
//
// Phases are separated by a software grid barrier. The grid never exceeds the
// number of workgroups the device keeps resident, and every wait is bounded: a
// barrier that does not complete within its spin limit sets status bit 0 and every
// workgroup leaves at that barrier or the next one it reaches.
#include "frame_common.hpp"
#include "mfma_common.hpp"

namespace nrrs {
namespace {

using namespace mfma;

constexpr uint32_t kSpinLimit = 1u << 22;
constexpr uint64_t kBounceSeedStep = 0x9E3779B97F4A7C15ull;
constexpr uint32_t kBarGroups = 8, kBarStride = 32;  // barrier groups; u32 words per counter line

// Bounded sense-by-generation barrier over all workgroups of the launch, in two levels:
// workgroup w arrives at the counter of group w % kBarGroups (each counter on a line of its
// own), the last arrival of a group at the top counter, and the last of those advances the
// generation. Arrivals at one counter are serialised by the memory system, so two levels cut
// the serial chain from nblocks to nblocks / kBarGroups + kBarGroups. One agent-scope release
// fence before arriving and one acquire fence after leaving; the spin itself uses relaxed
// loads (an acquire load per spin iteration would invalidate the L2 of every spinning
// workgroup's XCD on each poll). bar: [0] top count, [1] generation, [kBarStride * (1 + g)]
// arrivals of group g; zeroed by k_bar_reset before every launch.
// A workgroup whose wait exceeds the spin limit sets status bit 0 and the abort bit of the
// generation word. The generation only ever grows by atomic adds, so the bit stays: every
// workgroup that is waiting at that barrier, arrives at it later or enters a later barrier sees
// it and leaves. The decision is taken by one thread and shared through LDS, so a workgroup
// never splits.
constexpr uint32_t kBarAbort = 0x80000000u;

__device__ __forceinline__ bool grid_sync(uint32_t* bar, uint32_t nblocks, uint32_t* status) {
    __shared__ uint32_t pass;
    __syncthreads();
    if (threadIdx.x == 0) {
        bool ok = true;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gen & kBarAbort) {
            ok = false;  // an earlier wait timed out: leave without arriving
        } else {
            const uint32_t g = blockIdx.x % kBarGroups, groups = nblocks < kBarGroups ? nblocks : kBarGroups;
            const uint32_t in_group = nblocks / kBarGroups + (g < nblocks % kBarGroups ? 1u : 0u);
            uint32_t* gc = bar + kBarStride * (1u + g);
            bool last = false;
            if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == in_group - 1u) {
                // the reset completes before the top arrival (release): no workgroup can pass this
                // barrier, and reach the group counter again, while the reset is in flight
                __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                last = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == groups - 1u;
            }
            if (last) {
                __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t old = __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                ok = (old & kBarAbort) == 0u;
            } else {
                uint32_t it = 0, cur;
                while ((cur = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == gen) {
                    if (++it > kSpinLimit) {
                        atomicOr(status, 1u);
                        __hip_atomic_fetch_or(bar + 1, kBarAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        cur = kBarAbort;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                ok = (cur & kBarAbort) == 0u;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        pass = ok ? 1u : 0u;
    }
    __syncthreads();
    return pass != 0u;
}

// Clears the barrier words before a frame launch (a kernel, so that a captured frame graph
// holds it as a kernel node).
__global__ void k_bar_reset(uint32_t* bar) {
    for (uint32_t k = threadIdx.x; k < kFrameBarWords; k += blockDim.x) bar[k] = 0u;
}

template <int BLOCK>
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* total, uint32_t* wsum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    __syncthreads();
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) {
        const uint32_t x = wsum[w];
        base += (w < wave) ? x : 0u;
        tot += x;
    }
    *total = tot;
    return base + inc - v;
}

__device__ __forceinline__ void store_position(const FrameParams& f, uint64_t seed_b, uint32_t k) {
    float p[3];
    synth_position(seed_b, k, p);
#pragma unroll
    for (int c = 0; c < 3; ++c) f.xbuf[(size_t)c * f.capacity + k] = norm_coord(p[c], f.frame.c[c], f.frame.d[c]);
}

template <typename P, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_frame(FrameParams f) {
    constexpr int BLOCK = 64 * WAVES;
    constexpr bool kSplit = split_of<P>(6);
    constexpr int kTilesS = tiles_of(5, kSplit), kTilesR = tiles_of(4);
    extern __shared__ u32x4 smem[];
    u32x4* sWs = smem;
    u32x4* sWr = smem + kTilesS * 64;
    __shared__ uint32_t wsum[WAVES];
    for (int k = threadIdx.x; k < kTilesS * 64; k += BLOCK) sWs[k] = reinterpret_cast<const u32x4*>(f.w_stat)[k];
    for (int k = threadIdx.x; k < kTilesR * 64; k += BLOCK) sWr[k] = reinterpret_cast<const u32x4*>(f.w_rrs)[k];

    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31, wave = threadIdx.x >> 6;
    const uint32_t nb = gridDim.x;
    const uint32_t cap = f.capacity;
    const uint64_t seed = f.seed_dev ? *f.seed_dev : f.seed;

    // frame set-up: queue 0 = all primary paths, per-bounce sums cleared
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        f.sizes[0] = f.n_primary < cap ? f.n_primary : cap;
        for (uint32_t b = 0; b < f.n_bounces; ++b) f.sums[b] = 0ull;
        if (f.n_primary > cap) atomicOr(f.status, 2u);
    }
    for (uint32_t k = blockIdx.x * BLOCK + threadIdx.x; k < f.n_primary && k < cap; k += nb * BLOCK) {
        f.queue[k] = k;
        store_position(f, seed, k);
    }
    if (!grid_sync(f.bar, nb, f.status)) return;

    for (uint32_t b = 0; b < f.n_bounces; ++b) {
        const uint32_t n = __hip_atomic_load(f.sizes + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t* qin = f.queue + (size_t)(b & 1u) * cap;
        uint32_t* qout = f.queue + (size_t)((b + 1) & 1u) * cap;
        const uint64_t seed_b = seed + (uint64_t)b * kBounceSeedStep;
        if (f.trace)
            for (uint32_t k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += nb * BLOCK)
                f.trace[(size_t)b * cap + k] = qin[k];

        // ---- phase A0: level-major grid encodings of both nets. Items are (net, level, BLOCK
        // vertices), walked by all workgroups in the same order; lanes 2k and 2k + 1 gather the x bit 0
        // and x bit 1 corners of vertices k and BLOCK/2 + k of an item (grid_level_half), as the
        // stand-alone encode pass does. With the paired table an item covers both nets of a level.
        {
            const uint32_t L = f.grid.n_levels, n_ch = (n + BLOCK - 1) / BLOCK;
            const bool paired = f.table_pair != nullptr;
            const uint32_t items = (paired ? 1u : 2u) * L * n_ch;
            const uint32_t xb = threadIdx.x & 1u, kk = threadIdx.x >> 1;
            for (uint32_t it = blockIdx.x; it < items; it += nb) {
                const uint32_t nl = it / n_ch, base = (it % n_ch) * BLOCK;  // nl = net * L + level
                const int lev = (int)(nl % L);
                const uint32_t* table = paired ? reinterpret_cast<const uint32_t*>(f.table_pair)
                                               : (nl < L ? f.table_stat : f.table_rrs);
#pragma unroll
                for (uint32_t half = 0; half < 2; ++half) {
                    const uint32_t i = base + half * (BLOCK / 2) + kk;
                    const uint32_t ic = i < n ? i : 0u;  // out-of-range lanes gather vertex 0, not stored
                    const float x[3] = {f.xbuf[ic], f.xbuf[cap + ic], f.xbuf[2 * cap + ic]};
                    const LevelCell lc = level_cell(f.grid, lev, x);
                    const float4 acc = paired ? grid_level_half<P, true>(table, f.grid, lev, lc, xb)
                                              : grid_level_half<P, false>(table, f.grid, lev, lc, xb);
                    float4 oth;
                    oth.x = __shfl_xor(acc.x, 1, 64);
                    oth.y = __shfl_xor(acc.y, 1, 64);
                    oth.z = __shfl_xor(acc.z, 1, 64);
                    oth.w = __shfl_xor(acc.w, 1, 64);
                    const float4 a0 = xb ? oth : acc, a1 = xb ? acc : oth;
                    if (i < n) {
                        if (paired && xb)
                            f.genc[(size_t)(L + nl) * cap + i] =
                                (uint32_t)lp_bits<P>(a0.z + a1.z) | ((uint32_t)lp_bits<P>(a0.w + a1.w) << 16);
                        else if (!xb)
                            f.genc[(size_t)nl * cap + i] =
                                (uint32_t)lp_bits<P>(a0.x + a1.x) | ((uint32_t)lp_bits<P>(a0.y + a1.y) << 16);
                    }
                }
            }
        }
        if (!grid_sync(f.bar, nb, f.status)) return;

        // ---- phase A: StatNet -> RRSNet -> q, exact normalisation sum
        {
            const uint32_t n_cb = (n + 31) / 32;
            const uint32_t total_waves = nb * WAVES;
            const uint32_t gw = blockIdx.x * WAVES + wave;
            uint64_t term = 0;
            for (uint32_t chunk = gw; chunk * f.cpw < n_cb; chunk += total_waves) {
                for (uint32_t j = 0; j < f.cpw; ++j) {
                    const uint32_t cb = chunk * f.cpw + j;
                    if (cb >= n_cb) break;
                    const uint32_t i = cb * 32 + col;
                    const bool live = i < n;
                    typename P::vec8 B[5], Br[2], Bl[3];
                    VertexRec v;
                    float ref = 0.0f;
                    // the vertex's fields, drawn by both lane halves (all lanes take part; the
                    // grid channels come from the encodings, so the position is not needed here)
                    synth_fields_pair(seed_b, live ? i : 0u, h, v.dir, v.nrm, v.ad, v.as, v.thp, &v.rough);
                    if (live) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) v.x[k] = 0.0f;
                        v.pixel = qin[i];
                        ref = film_mean(f.film, f.n_pixels, v.pixel);
                        const uint32_t L = f.grid.n_levels;
                        grid_half_from_levels<P>(f.genc, cap, L, i, h, B);
                        grid_half_from_levels<P>(f.genc + (size_t)L * cap, cap, L, i, h, Br);
                        encode_tail_half<P>(v, NRRS_STATNET, h, 0.0f, nullptr, B, kSplit ? Bl : nullptr);
                    } else {
#pragma unroll
                        for (int kb = 0; kb < 5; ++kb) B[kb] = typename P::vec8{};
#pragma unroll
                        for (int kb = 0; kb < 3; ++kb) Bl[kb] = typename P::vec8{};
                        Br[0] = Br[1] = typename P::vec8{};
                    }
                    // K chunks 2 and 3 (SH-4 of the direction and the normal) are the same in
                    // both nets' inputs
                    const typename P::vec8 sh_dir = B[2], sh_nrm = B[3];
                    const f32x16 Ys = mlp_forward<P, 5, kSplit>(sWs, B, Bl, lane);
                    // statistics of the lane's vertex: outputs 0..3 in half 0, 4..5 in half 1
                    float st[6];
                    {
                        float a[4], o[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            a[r] = Ys[r];
                            o[r] = __shfl_xor(a[r], 32, 64);
                        }
                        if (h == 0) {
                            st[0] = a[0]; st[1] = a[1]; st[2] = a[2]; st[3] = a[3]; st[4] = o[0]; st[5] = o[1];
                        } else {
                            st[0] = o[0]; st[1] = o[1]; st[2] = o[2]; st[3] = o[3]; st[4] = a[0]; st[5] = a[1];
                        }
                    }
                    if (live) {
                        B[0] = Br[0];
                        B[1] = Br[1];
                        encode_tail_half<P>(v, NRRS_RRSNET, h, ref, st, B);
                        B[2] = sh_dir;  // (the recomputed chunks are dead and not evaluated)
                        B[3] = sh_nrm;
                    }
                    const f32x16 Yr = mlp_forward<P, 4>(sWr, B, nullptr, lane);
                    if (live && h == 0) {
                        const float q = rrs_window_clamp(f.d, activation_sigma(Yr[0]));
                        f.q_scr[i] = q;
                        if (f.q_out) f.q_out[(size_t)b * cap + i] = q;
                        term += rrs_sum_term(f.d, q);
                    }
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) term += __shfl_xor(term, o, 64);
            if (lane == 0 && term) atomicAdd(f.sums + b, (unsigned long long)term);
        }
        if (!grid_sync(f.bar, nb, f.status)) return;

        // ---- phase B: normalise, draw, split counts. Workgroup w takes the contiguous tiles
        // [w * tpw, (w + 1) * tpw), so that the queue is filled in tile order by the
        // workgroups one after the other and only per-workgroup totals cross the barrier.
        const uint32_t n_tiles = (n + BLOCK - 1) / BLOCK, tpw = (n_tiles + nb - 1) / nb;
        const uint32_t t_begin = min(blockIdx.x * tpw, n_tiles), t_end = min(t_begin + tpw, n_tiles);
        {
            const double sum = (double)__hip_atomic_load(f.sums + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) *
                               0x1p-32;
            const float sum_f = (float)sum;
            const float rcp = sum_f > 0.0f ? 1.0f / sum_f : 0.0f;
            float nf = (float)f.d.n_pixels * rcp;
            nf = nf * (f.d.use_ceil ? 1.0f : f.d.rate);
            uint32_t cnt = 0;
            for (uint32_t t = t_begin; t < t_end; ++t) {
                const uint32_t i = t * BLOCK + threadIdx.x;
                if (i < n) {
                    float r = f.q_scr[i];
                    r = f.d.no_upscale ? r * (nf < 1.0f ? nf : 1.0f) : r * nf;
                    const float u = u24(pcg_output(pcg_state_at(seed, 1000ull + b, i)));
                    int s = (int)r;
                    const float left = r - (float)s;
                    s += (u <= left) ? 1 : 0;
                    const uint32_t si = (uint32_t)min(s, (int)f.d.split_cap);
                    f.s_scr[i] = (uint8_t)si;
                    cnt += si;
                }
            }
            uint32_t tot;
            block_scan_excl<BLOCK>(cnt, &tot, wsum);
            if (threadIdx.x == 0) f.tile_sums[blockIdx.x] = tot;  // survivors of this workgroup
        }
        if (!grid_sync(f.bar, nb, f.status)) return;

        // ---- phase C: stable scatter of the survivors into the next queue. Every workgroup
        // sums the totals of the workgroups before it (its first slot) and of all (the next
        // queue size), then walks its tiles in order.
        uint32_t n1;
        {
            uint32_t before = 0, all = 0;
            for (uint32_t w = threadIdx.x; w < nb; w += BLOCK) {
                const uint32_t v = f.tile_sums[w];
                all += v;
                before += w < blockIdx.x ? v : 0u;
            }
            uint32_t run, total;
            block_scan_excl<BLOCK>(before, &run, wsum);
            block_scan_excl<BLOCK>(all, &total, wsum);
            n1 = total < cap ? total : cap;
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                __hip_atomic_store(f.sizes + b + 1, n1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (total > cap) atomicOr(f.status, 2u);
            }
            for (uint32_t t = t_begin; t < t_end; ++t) {
                const uint32_t i = t * BLOCK + threadIdx.x;
                const uint32_t si = i < n ? f.s_scr[i] : 0u;
                uint32_t tot;
                const uint32_t pos = run + block_scan_excl<BLOCK>(si, &tot, wsum);
                if (si) {
                    const uint32_t id = qin[i];
                    for (uint32_t c = 0; c < si; ++c)
                        if (pos + c < cap) qout[pos + c] = id;
                }
                run += tot;
            }
        }
        // positions of the next bounce (slot k draws vertex k of that bounce's seed); the
        // current positions were last read in phase A0
        if (b + 1 < f.n_bounces) {
            for (uint32_t k = blockIdx.x * BLOCK + threadIdx.x; k < n1; k += nb * BLOCK)
                store_position(f, seed + (uint64_t)(b + 1) * kBounceSeedStep, k);
        }
        if (!grid_sync(f.bar, nb, f.status)) return;
    }
    if (f.trace) {
        const uint32_t n = __hip_atomic_load(f.sizes + f.n_bounces, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t* qf = f.queue + (size_t)(f.n_bounces & 1u) * cap;
        for (uint32_t k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += nb * BLOCK)
            f.trace[(size_t)f.n_bounces * cap + k] = qf[k];
    }
    if (f.seed_dev && f.advance_seed && blockIdx.x == 0 && threadIdx.x == 0) {
        // the next replay of a captured frame draws a fresh frame; every workgroup has
        // read the seed before the first barrier
        *f.seed_dev = seed + 0xD1B54A32D192ED03ull;
    }
}

__global__ void k_bounce_vertices(uint64_t seed, uint32_t bounce, const uint32_t* __restrict__ queue, uint32_t n,
                                  nrrs_vertices_out o) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const SynthVertex v = synth_vertex(seed + (uint64_t)bounce * kBounceSeedStep, i);
    const uint32_t S = o.stride;
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
    o.pixel[i] = queue[i];
}

template <typename P, int WAVES>
hipError_t launch_frame_t(const FrameParams& f, uint32_t max_blocks, hipStream_t s) {
    const size_t smem = (size_t)(tiles_of(5, split_of<P>(6)) + tiles_of(4)) * 64 * 16;
    auto kern = k_frame<P, WAVES>;
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern),
                                                                64 * WAVES, smem);
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    uint32_t blocks = (uint32_t)per_cu * (uint32_t)cus;  // all resident at once
    if (max_blocks && blocks > max_blocks) blocks = max_blocks;
    if (blocks > kFrameMaxBlocks) blocks = kFrameMaxBlocks;  // the per-workgroup totals' scratch
    hipLaunchKernelGGL(k_bar_reset, dim3(1), dim3(256), 0, s, f.bar);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * WAVES), smem, s, f);
    return hipGetLastError();
}

template <typename P>
hipError_t launch_frame_p(const FrameParams& f, int waves, uint32_t max_blocks, hipStream_t s) {
    if (waves == 1) return launch_frame_t<P, 1>(f, max_blocks, s);
    if (waves == 2) return launch_frame_t<P, 2>(f, max_blocks, s);
    if (waves == 4) return launch_frame_t<P, 4>(f, max_blocks, s);
    if (waves == 8) return launch_frame_t<P, 8>(f, max_blocks, s);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_frame(int precision, const FrameParams& f, int waves, uint32_t max_blocks, hipStream_t s) {
    if (precision == NRRS_BF16) return launch_frame_p<BF16>(f, waves, max_blocks, s);
    if (precision == NRRS_FP16) return launch_frame_p<F16>(f, waves, max_blocks, s);
    return hipErrorInvalidValue;
}

hipError_t launch_bounce_vertices(uint64_t seed, uint32_t bounce, const uint32_t* queue, uint32_t n,
                                  const nrrs_vertices_out& o, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_bounce_vertices, dim3((n + 255) / 256), dim3(256), 0, s, seed, bounce, queue, n, o);
    return hipGetLastError();
}

}  // namespace nrrs
