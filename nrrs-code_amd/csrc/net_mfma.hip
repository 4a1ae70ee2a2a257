// Mixed-precision path: hash-grid + SH encoding fused with the MLP on MFMA
// (v_mfma_f32_32x32x16_{bf16,f16}, FP32 accumulation). The operand layouts are
// described in mfma_common.hpp.
#include "mfma_common.hpp"
#include "train_common.hpp"

namespace nrrs {
namespace {

using namespace mfma;

constexpr int kWaves = 4;
constexpr int kBlockT = 64 * kWaves;

// One thread per (tile, lane): 8 weights of the A operand, in the tile layout of tiles_of
// (residual tiles hold lp(W - lp(W)); the difference is exact in FP32).
template <typename P>
__global__ void k_pack_mlp(const float* __restrict__ W, int n_layers, int n_out, int split, uint16_t* __restrict__ out) {
    pack_mlp_tile<P>(W, n_layers, n_out, split, out, blockIdx.x, threadIdx.x);
}

template <typename P>
__global__ void k_to_lp(const float* __restrict__ src, uint16_t* __restrict__ dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = lp_bits<P>(src[i]);
}

// Level-major grid encoding: workgroup b takes level b / n_chunks and the kRunVert vertices of
// chunk b % n_chunks, so consecutive workgroups share one level's table and the GPU walks the
// levels one after the other (each level's table stays in L2 while it is gathered). The pass is
// bound by the L2 request rate of its random gathers, so each vertex is taken by two lanes
// (grid_level_half): lane 2k + xb gathers the corners with x bit xb, and the two entries of an
// x-neighbour pair come in one request. Lanes 2k and 2k + 1 take vertices k and kRunVert/2 + k of
// the chunk in turn.
// DUAL: `table` is the paired table of both nets; StatNet's encoding goes to `out` (written by
// the even lane), RRSNet's to `out2` (odd lane), and the record counts (the same for both nets,
// whose tables share one layout) to `run_counts` and, when given, `run_counts2`.
template <typename P, bool DUAL>
__global__ __launch_bounds__(kRunVert) void k_encode_levels(EncodeArgs a, const uint32_t* __restrict__ table, uint32_t n,
                                                            uint32_t n_chunks, uint32_t* __restrict__ out,
                                                            uint32_t* __restrict__ out2, uint32_t stride,
                                                            uint32_t* __restrict__ run_counts,
                                                            uint32_t* __restrict__ run_counts2) {
    if (a.zero_u64 && blockIdx.x == 0 && threadIdx.x == 0) *a.zero_u64 = 0ull;
    const uint32_t l = blockIdx.x / n_chunks, chunk = blockIdx.x % n_chunks;
    if (l >= a.grid.n_levels) return;  // whole workgroup
    const uint32_t xb = threadIdx.x & 1u, k = threadIdx.x >> 1;
    __shared__ uint32_t hist[kMaxSegBins];
    if (run_counts) {
        if (threadIdx.x < kMaxSegBins) hist[threadIdx.x] = 0u;
        __syncthreads();
    }
    const bool dense = (a.grid.dense_mask >> l) & 1u;
#pragma unroll
    for (uint32_t half = 0; half < 2; ++half) {
        const uint32_t i = chunk * kRunVert + half * (kRunVert / 2) + k;
        const bool live = i < n;
        float x[3] = {0.0f, 0.0f, 0.0f};  // lanes past the batch gather cell 0 (in bounds), store nothing
        if (live) load_xn(a, i, x);
        const LevelCell lc = level_cell(a.grid, (int)l, x);
        const float4 acc = grid_level_half<P, DUAL>(table, a.grid, (int)l, lc, xb);
        float4 oth;
        oth.x = __shfl_xor(acc.x, 1, 64);
        oth.y = __shfl_xor(acc.y, 1, 64);
        if constexpr (DUAL) {
            oth.z = __shfl_xor(acc.z, 1, 64);
            oth.w = __shfl_xor(acc.w, 1, 64);
        }
        const float4 a0 = xb ? oth : acc, a1 = xb ? acc : oth;
        if (live) {
            const size_t o = (size_t)l * stride + i;
            if (DUAL && xb)
                out2[o] = (uint32_t)lp_bits<P>(a0.z + a1.z) | ((uint32_t)lp_bits<P>(a0.w + a1.w) << 16);
            else if (!xb)
                out[o] = (uint32_t)lp_bits<P>(a0.x + a1.x) | ((uint32_t)lp_bits<P>(a0.y + a1.y) << 16);
        }
        if (run_counts && live) {  // records of this vertex's (y, z) rows 2xb and 2xb + 1
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int yz = 2 * (int)xb + j;
                const uint32_t e0 = corner_index(a.grid, (int)l, lc, yz << 1);
                const uint32_t e1 = corner_index(a.grid, (int)l, lc, (yz << 1) | 1);
                const bool two = pair_code(e0, e1) == 3u;  // a record per corner
                if (dense) {  // few segments: aggregated atomics
                    lds_count(hist, e0 >> kSegShift);
                    if (two) lds_count(hist, e1 >> kSegShift);
                } else {
                    atomicAdd(&hist[e0 >> kSegShift], 1u);
                    if (two) atomicAdd(&hist[e1 >> kSegShift], 1u);
                }
            }
        }
    }
    if (run_counts) {  // uniform: table-gradient records per segment of this level and chunk
        __syncthreads();
        if (threadIdx.x < kMaxSegBins) {
            const size_t o = ((size_t)l * n_chunks + chunk) * kMaxSegBins + threadIdx.x;
            run_counts[o] = hist[threadIdx.x];
            if (DUAL && run_counts2) run_counts2[o] = hist[threadIdx.x];
        }
    }
}

// Four waves per SIMD (at most 128 registers): the weight tiles are read from LDS in every
// iteration of the persistent loop rather than kept in registers, which would take 120-152
// of them and leave one or two waves per SIMD to hide the latency of the loads.
template <typename P, int NL, int NOUT, int EPI, int ACT, int W>
__global__ __launch_bounds__(kBlockT, 4) void k_net_mfma(EncodeArgs a, int net, uint32_t n,
                                                      const uint32_t* __restrict__ table,
                                                      const u32x4* __restrict__ Wp, EpiArgs ep) {
    extern __shared__ u32x4 sW[];
    constexpr bool kSplit = split_of<P>(NOUT);
    constexpr int kTiles = tiles_of(NL, kSplit);
    for (int k = threadIdx.x; k < kTiles * 64; k += kBlockT) sW[k] = Wp[k];
    __syncthreads();

    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const int wave = threadIdx.x >> 6;
    const uint32_t n_cb = (n + 31) / 32;
    uint64_t term = 0;
    for (uint32_t cb = blockIdx.x * kWaves + wave; cb < n_cb; cb += gridDim.x * kWaves) {
        asm volatile("" ::: "memory");  // keeps the weight-tile loads from LDS inside the loop
        const uint32_t i = cb * 32 + col;
        const bool live = i < n;
        typename P::vec8 B[5], Bl[3];
        if (live) {
            if (a.grid_enc)
                encode_half_pre<P>(a, net, i, h, B, kSplit ? Bl : nullptr);
            else
                encode_half<P>(a, net, table, i, h, B, kSplit ? Bl : nullptr);
        } else {
#pragma unroll
            for (int kb = 0; kb < 5; ++kb) B[kb] = typename P::vec8{};
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) Bl[kb] = typename P::vec8{};
        }
        f32x16 Y = mlp_forward<P, NL, kSplit, ACT, W>(sW, B, Bl, lane, a.l0_chunks);

        // output c of vertex i sits in lane half c / 4, register c % 4
        if (EPI == EPI_RAW) {
            if (live) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = 4 * h + r;
                    if (c < NOUT && (ep.head < 0 || c / 3 == ep.head)) ep.out[(size_t)c * ep.out_stride + i] = Y[r];
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

// Chained StatNet -> RRSNet inference in one launch: the StatNet statistics of a vertex stay in
// registers (exchanged between the two lane halves) and feed the RRSNet input directly, which
// saves writing and re-reading them and a second pass over the vertex records. The grid
// channels come from level-major encodings (a.grid_enc: StatNet, enc_rrs: RRSNet; the shipped
// RRSNet has no grid). Both nets' weight tiles sit in LDS: FP16 76 KB, 8-wave workgroups, two
// per CU; BF16 (split StatNet) 106 KB, one 16-wave workgroup per CU. Four waves per SIMD either
// way (at most 128 registers). Epilogue as k_net_mfma's EPI_RRS: q and its exact sum term.
// SET: the network set (NRRS_NETS_BENCH: StatNet 80-64x4-6, RRSNet 80-64x3-1, ReLU;
// NRRS_NETS_SHIPPED: 32-wide LeakyReLU nets, StatNet 3 hidden layers, RRSNet on
// NRRSNetworkInput1 without a grid).
template <typename P, int SET = NRRS_NETS_BENCH>
struct Chain {
    static constexpr bool kShipped = SET == NRRS_NETS_SHIPPED;
    static constexpr int kLayersS = kShipped ? 4 : 5;
    static constexpr int kWidth = kShipped ? 32 : 64;
    static constexpr int kAct = kShipped ? ACT_LEAKY : ACT_RELU;
    static constexpr uint32_t kL0Rrs = kShipped ? 0x10u : 0x1fu;  // RRSNet's layer-0 input chunks
    static constexpr int kWaves = P::kSplitStat ? 16 : 8;
    static constexpr int kThreads = 64 * kWaves;
    static constexpr int kPerCu = P::kSplitStat ? 1 : 2;
    static constexpr int kTilesS = tiles_of(kLayersS, split_of<P>(6)), kTilesR = tiles_of(4);
};

template <typename P, int SET>
__global__ __launch_bounds__((Chain<P, SET>::kThreads)) void k_net_chained(EncodeArgs a,
                                                                        const uint32_t* __restrict__ enc_rrs,
                                                                        uint32_t n, const u32x4* __restrict__ Ws,
                                                                        const u32x4* __restrict__ Wr, EpiArgs ep) {
    typedef Chain<P, SET> Ch;
    extern __shared__ u32x4 sW[];
    constexpr int kChainWaves = Ch::kWaves, kChainT = Ch::kThreads;
    constexpr int kTilesS = Ch::kTilesS, kTilesR = Ch::kTilesR;
    constexpr bool kSplit = split_of<P>(6);
    u32x4* sWs = sW;
    u32x4* sWr = sW + kTilesS * 64;
    for (int k = threadIdx.x; k < kTilesS * 64; k += kChainT) sWs[k] = Ws[k];
    for (int k = threadIdx.x; k < kTilesR * 64; k += kChainT) sWr[k] = Wr[k];
    __syncthreads();

    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const int wave = threadIdx.x >> 6;
    const uint32_t n_cb = (n + 31) / 32;
    uint64_t term = 0;
    for (uint32_t cb = blockIdx.x * kChainWaves + wave; cb < n_cb; cb += gridDim.x * kChainWaves) {
        asm volatile("" ::: "memory");  // keeps the weight-tile loads from LDS inside the loop
        const uint32_t i = cb * 32 + col;
        const bool live = i < n;
        typename P::vec8 B[5], Bl[3];
        if (live) {
            encode_half_pre<P, SET>(a, NRRS_STATNET, i, h, B, kSplit ? Bl : nullptr);
        } else {
#pragma unroll
            for (int kb = 0; kb < 5; ++kb) B[kb] = typename P::vec8{};
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) Bl[kb] = typename P::vec8{};
        }
        // benchmark nets: K chunks 2 and 3 (SH-4 of the direction and the normal) are the same in
        // both nets' inputs
        const typename P::vec8 sh_dir = B[2], sh_nrm = B[3];
        const f32x16 Ys = mlp_forward<P, Ch::kLayersS, kSplit, Ch::kAct, Ch::kWidth>(sWs, B, Bl, lane, a.l0_chunks);
        // statistics of the lane's vertex: outputs 0..3 in half 0, 4..5 in half 1
        float st[6];
        {
            float o[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) o[r] = __shfl_xor(Ys[r], 32, 64);
            if (h == 0) {
                st[0] = Ys[0]; st[1] = Ys[1]; st[2] = Ys[2]; st[3] = Ys[3]; st[4] = o[0]; st[5] = o[1];
            } else {
                st[0] = o[0]; st[1] = o[1]; st[2] = o[2]; st[3] = o[3]; st[4] = Ys[0]; st[5] = Ys[1];
            }
        }
        if (live) {
            const VertexRec v = vertex_from_soa(a, i);
            const float ref = film_mean(a.film, a.n_pixels, v.pixel);
            if constexpr (Ch::kShipped) {
                B[0] = B[1] = typename P::vec8{};
            } else {
                grid_half_from_levels<P>(enc_rrs, a.grid_enc_stride, a.grid.n_levels, i, h, B);
            }
            encode_tail_half<P>(v, NRRS_RRSNET, h, ref, st, B, nullptr, false, SET);
            if constexpr (!Ch::kShipped) {
                B[2] = sh_dir;  // (the recomputed chunks are dead and not evaluated)
                B[3] = sh_nrm;
            }
        } else {
#pragma unroll
            for (int kb = 0; kb < 5; ++kb) B[kb] = typename P::vec8{};
        }
        const f32x16 Yr = mlp_forward<P, 4, false, Ch::kAct, Ch::kWidth>(sWr, B, nullptr, lane, Ch::kL0Rrs);
        if (live && h == 0) {
            const float q = rrs_window_clamp(ep.d, activation_sigma(Yr[0]));
            ep.q[i] = q;
            term += rrs_sum_term(ep.d, q);
        }
    }
    block_atomic_add_u64(term, ep.sum);
}

template <typename P, int SET>
hipError_t launch_chained(const EncodeArgs& a, const uint32_t* enc_rrs, uint32_t n, const uint16_t* Ws,
                          const uint16_t* Wr, const EpiArgs& ep, hipStream_t s) {
    constexpr size_t smem = (size_t)(Chain<P, SET>::kTilesS + Chain<P, SET>::kTilesR) * kTileHalves * 2;
    constexpr int kChainWaves = Chain<P, SET>::kWaves;
    static std::atomic<bool> attr[kMaxDevices];
    hipError_t e = set_dyn_lds_once(attr, reinterpret_cast<const void*>(&k_net_chained<P, SET>), smem);
    if (e != hipSuccess) return e;
    // persistent: the resident workgroups only, each loading both nets' tiles once
    static std::atomic<uint32_t> resident_of[kMaxDevices];
    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    uint32_t resident = dev >= 0 && dev < kMaxDevices ? resident_of[dev].load(std::memory_order_relaxed) : 0u;
    if (!resident) {
        int cus = 0;
        if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
        resident = (uint32_t)Chain<P, SET>::kPerCu * (uint32_t)(cus > 0 ? cus : 1);
        if (dev >= 0 && dev < kMaxDevices) resident_of[dev].store(resident, std::memory_order_relaxed);
    }
    const uint32_t n_cb = (n + 31) / 32;
    uint32_t blocks = (n_cb + kChainWaves - 1) / kChainWaves;
    blocks = blocks > resident ? resident : blocks;
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL((k_net_chained<P, SET>), dim3(blocks), dim3(Chain<P, SET>::kThreads), smem, s, a, enc_rrs, n,
                       reinterpret_cast<const u32x4*>(Ws), reinterpret_cast<const u32x4*>(Wr), ep);
    return hipGetLastError();
}

template <typename P, int NL, int NOUT, int EPI, int ACT = ACT_RELU, int W = 64>
hipError_t launch_one(const EncodeArgs& a, int net, uint32_t n, const uint16_t* table, const uint16_t* Wp,
                      const EpiArgs& ep, hipStream_t s) {
    constexpr size_t smem = (size_t)tiles_of(NL, split_of<P>(NOUT)) * kTileHalves * 2;
    static std::atomic<bool> attr[kMaxDevices];
    hipError_t e =
        set_dyn_lds_once(attr, reinterpret_cast<const void*>(&k_net_mfma<P, NL, NOUT, EPI, ACT, W>), smem);
    if (e != hipSuccess) return e;
    const uint32_t n_cb = (n + 31) / 32;
    uint32_t blocks = (n_cb + kWaves - 1) / kWaves;
    blocks = blocks > 4096 ? 4096 : blocks;
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL((k_net_mfma<P, NL, NOUT, EPI, ACT, W>), dim3(blocks), dim3(kBlockT), smem, s, a, net, n,
                       reinterpret_cast<const uint32_t*>(table), reinterpret_cast<const u32x4*>(Wp), ep);
    return hipGetLastError();
}

template <typename P>
hipError_t dispatch(const EncodeArgs& a, int net, int n_layers, int n_out, int epi, uint32_t n,
                    const uint16_t* table, const uint16_t* Wp, const EpiArgs& ep, hipStream_t s) {
    if (a.net_set == NRRS_NETS_SHIPPED) {
        // the shipped nets: 3 hidden 32-wide LeakyReLU layers, StatNet -> 6, RRSNet -> 1
        if (n_layers == 4 && n_out == 6) {
            if (epi == EPI_RAW) return launch_one<P, 4, 6, EPI_RAW, ACT_LEAKY, 32>(a, net, n, table, Wp, ep, s);
            if (epi == EPI_ADRRS) return launch_one<P, 4, 6, EPI_ADRRS, ACT_LEAKY, 32>(a, net, n, table, Wp, ep, s);
        }
        if (n_layers == 4 && n_out == 1) {
            if (epi == EPI_RAW) return launch_one<P, 4, 1, EPI_RAW, ACT_LEAKY, 32>(a, net, n, table, Wp, ep, s);
            if (epi == EPI_RRS) return launch_one<P, 4, 1, EPI_RRS, ACT_LEAKY, 32>(a, net, n, table, Wp, ep, s);
        }
        return hipErrorInvalidValue;
    }
    if (n_layers == 5 && n_out == 6) {
        if (epi == EPI_RAW) return launch_one<P, 5, 6, EPI_RAW>(a, net, n, table, Wp, ep, s);
        if (epi == EPI_ADRRS) return launch_one<P, 5, 6, EPI_ADRRS>(a, net, n, table, Wp, ep, s);
    }
    if (n_layers == 4 && n_out == 1) {
        if (epi == EPI_RAW) return launch_one<P, 4, 1, EPI_RAW>(a, net, n, table, Wp, ep, s);
        if (epi == EPI_RRS) return launch_one<P, 4, 1, EPI_RRS>(a, net, n, table, Wp, ep, s);
    }
    // two-headed StatNets (PartShare, JustEncoding): the weights between the heads are zeros
    if (n_layers == 4 && n_out == 6) {
        if (epi == EPI_RAW) return launch_one<P, 4, 6, EPI_RAW>(a, net, n, table, Wp, ep, s);
        if (epi == EPI_ADRRS) return launch_one<P, 4, 6, EPI_ADRRS>(a, net, n, table, Wp, ep, s);
    }
    if (n_layers == 3 && n_out == 6) {
        if (epi == EPI_RAW) return launch_one<P, 3, 6, EPI_RAW>(a, net, n, table, Wp, ep, s);
        if (epi == EPI_ADRRS) return launch_one<P, 3, 6, EPI_ADRRS>(a, net, n, table, Wp, ep, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_encode_levels(int precision, const EncodeArgs& a, const uint16_t* table, uint32_t n, uint32_t* out,
                                uint32_t stride, hipStream_t s, uint32_t* run_counts) {
    if (!n || !a.grid.n_levels) return hipSuccess;  // no grid: nothing to gather
    const uint32_t n_chunks = (n + kRunVert - 1) / kRunVert;
    const dim3 grid(n_chunks * a.grid.n_levels);
    const uint32_t* t = reinterpret_cast<const uint32_t*>(table);
    if (precision == NRRS_BF16)
        hipLaunchKernelGGL((k_encode_levels<BF16, false>), grid, dim3(kRunVert), 0, s, a, t, n, n_chunks, out,
                           nullptr, stride, run_counts, nullptr);
    else if (precision == NRRS_FP16)
        hipLaunchKernelGGL((k_encode_levels<F16, false>), grid, dim3(kRunVert), 0, s, a, t, n, n_chunks, out,
                           nullptr, stride, run_counts, nullptr);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_encode_levels_dual(int precision, const EncodeArgs& a, const uint32_t* pair_table, uint32_t n,
                                     uint32_t* out_stat, uint32_t* out_rrs, uint32_t stride, hipStream_t s,
                                     uint32_t* run_counts, uint32_t* run_counts2) {
    if (!n) return hipSuccess;
    const uint32_t n_chunks = (n + kRunVert - 1) / kRunVert;
    const dim3 grid(n_chunks * a.grid.n_levels);
    if (precision == NRRS_BF16)
        hipLaunchKernelGGL((k_encode_levels<BF16, true>), grid, dim3(kRunVert), 0, s, a, pair_table, n, n_chunks,
                           out_stat, out_rrs, stride, run_counts, run_counts2);
    else if (precision == NRRS_FP16)
        hipLaunchKernelGGL((k_encode_levels<F16, true>), grid, dim3(kRunVert), 0, s, a, pair_table, n, n_chunks,
                           out_stat, out_rrs, stride, run_counts, run_counts2);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Room for the larger (split) layout.
size_t mfma_packed_halves(int n_layers) { return (size_t)tiles_of(n_layers, true) * kTileHalves; }

hipError_t launch_pack_lp(int precision, const float* grid, uint64_t n_grid, const float* W, int n_layers,
                          int n_out, uint16_t* grid_lp, uint16_t* W_lp, hipStream_t s) {
    if (precision == NRRS_BF16) {
        const int split = split_of<BF16>(n_out);
        if (W)
            hipLaunchKernelGGL(k_pack_mlp<BF16>, dim3(tiles_of(n_layers, split)), dim3(64), 0, s, W, n_layers, n_out,
                               split, W_lp);
        if (grid)
            hipLaunchKernelGGL(k_to_lp<BF16>, dim3(2048), dim3(256), 0, s, grid, grid_lp, n_grid);
    } else if (precision == NRRS_FP16) {
        const int split = split_of<F16>(n_out);
        if (W)
            hipLaunchKernelGGL(k_pack_mlp<F16>, dim3(tiles_of(n_layers, split)), dim3(64), 0, s, W, n_layers, n_out,
                               split, W_lp);
        if (grid)
            hipLaunchKernelGGL(k_to_lp<F16>, dim3(2048), dim3(256), 0, s, grid, grid_lp, n_grid);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_net_chained(int precision, const EncodeArgs& a, const uint32_t* enc_rrs, uint32_t n,
                              const uint16_t* Ws, const uint16_t* Wr, const EpiArgs& ep, hipStream_t s) {
    if (a.stats || a.aid) return hipErrorInvalidValue;
    if (a.net_set == NRRS_NETS_SHIPPED) {  // the shipped RRSNet has no grid
        if (!a.grid_enc && a.grid.n_levels) return hipErrorInvalidValue;
        if (precision == NRRS_BF16) return launch_chained<BF16, NRRS_NETS_SHIPPED>(a, nullptr, n, Ws, Wr, ep, s);
        if (precision == NRRS_FP16) return launch_chained<F16, NRRS_NETS_SHIPPED>(a, nullptr, n, Ws, Wr, ep, s);
        return hipErrorInvalidValue;
    }
    if (!a.grid_enc || !enc_rrs) return hipErrorInvalidValue;
    if (precision == NRRS_BF16) return launch_chained<BF16, NRRS_NETS_BENCH>(a, enc_rrs, n, Ws, Wr, ep, s);
    if (precision == NRRS_FP16) return launch_chained<F16, NRRS_NETS_BENCH>(a, enc_rrs, n, Ws, Wr, ep, s);
    return hipErrorInvalidValue;
}

hipError_t launch_net_mfma(int precision, const EncodeArgs& a, int net, int n_layers, int n_out, int epi,
                           uint32_t n, const uint16_t* table, const uint16_t* Wp, const EpiArgs& ep,
                           hipStream_t s) {
    if (precision == NRRS_BF16) return dispatch<BF16>(a, net, n_layers, n_out, epi, n, table, Wp, ep, s);
    if (precision == NRRS_FP16) return dispatch<F16>(a, net, n_layers, n_out, epi, n, table, Wp, ep, s);
    return hipErrorInvalidValue;
}

}  // namespace nrrs
