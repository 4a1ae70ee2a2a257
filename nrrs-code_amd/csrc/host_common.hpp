// Shared host/device definitions of the neural RRS stage.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "../../include/nrrs_amd.h"

namespace nrrs {

constexpr int kEnc = 80;       // network input width (64 encoded + 16 per-vertex channels)
constexpr int kHidden = 64;    // MLP width
constexpr int kMaxLevels = 16;
constexpr int kMaxLayers = 7;  // n_hidden_layers <= 6
constexpr uint64_t kPcgMult = 6364136223846793005ull;
// Hidden-layer activations: ReLU (the benchmark nets) and LeakyReLU with slope 0.01 (the
// shipped nets, 2net-simple.json:72, 240; tiny-cuda-nn's LeakyReLU slope)
enum Activation { ACT_RELU = 0, ACT_LEAKY = 1 };
constexpr float kLeakySlope = 0.01f;

// Two-headed StatNet (nrrs_config.ll2_two_head): the first layer split into the two 32-wide
// heads, or -1 for one head. PartShare: the tails' two layers (hidden, output); JustEncoding:
// every layer.
__host__ __device__ inline int head_first_layer(int two_head, int n_layers) {
    return two_head == 1 ? n_layers - 2 : (two_head == 2 ? 0 : -1);
}
// Whether W_layer[row][col] joins the two heads (held at zero): from the layer after the split
// on, output rows belong to head row >= 32 (the output layer: row >= 3, L then L^2) and input
// columns to head col >= 32; the split layer itself reads the shared input whole.
__host__ __device__ inline bool head_masked(int first, int layer, int n_layers, int row, int col) {
    if (first < 0 || layer <= first) return false;
    const bool rh = row >= (layer == n_layers - 1 ? 3 : 32), ch = col >= 32;
    return rh != ch;
}

// Per-level hash-grid constants, passed to kernels by value.
struct GridParams {
    float scale[kMaxLevels];
    uint32_t res[kMaxLevels];
    uint32_t offset[kMaxLevels];  // first entry of the level
    uint32_t dense_mask;          // bit l set: level l is a dense (res^3 <= T) grid
    uint32_t n_levels;
    uint32_t hash_mask;           // T - 1
    uint32_t n_entries;
};

// normalizeSpatialCoord frame: x = 0.5 + (p - c) / d.
struct Frame {
    float c[3];
    float d[3];
};

struct EncodeArgs {
    nrrs_vertices v;
    const float* film;        // [3][n_pixels]
    uint32_t n_pixels;
    const float* stats;       // [6][stats_stride] StatNet outputs, or null
    uint32_t stats_stride;
    GridParams grid;
    Frame frame;
    // Level-major grid encoding [level][grid_enc_stride] of packed FP16/BF16 feature pairs
    // (k_encode_levels); when set, the MFMA kernels read it instead of gathering.
    const uint32_t* grid_enc = nullptr;
    uint32_t grid_enc_stride = 0;
    // FP32 contexts: level-major grid encoding [level][grid_enc_stride] of FP32 feature pairs
    // (k_encode_levels_f32); when set, the FP32 network kernel reads it instead of gathering.
    const float2* grid_enc_f32 = nullptr;
    // RRSNet training input with sigma outputs: channels 72-74 take activationSigma(max(L2, 0))
    int sigma_l2 = 0;
    // RRSNet's AID input: no normal, albedo or statistics channels (the statistics, when set,
    // still feed the NRRS_RRS loss)
    int aid = 0;
    // network set of the non-grid channels (NRRS_NETS_BENCH / NRRS_NETS_SHIPPED)
    int net_set = 0;
    // layer-0 K chunks (input channels 16c..16c+15, bit c) that hold an input of the net; the
    // 32-wide shipped nets' kernels multiply only these
    uint32_t l0_chunks = 0x1fu;
    // k_encode_levels only: one thread sets this word to 0 (the RRS stage's sum, which the net
    // kernel launched next accumulates), in place of a launch of its own
    unsigned long long* zero_u64 = nullptr;
};

struct DecideParams {
    float clamp_min, clamp_max, rate, window_size;
    int32_t use_ceil, no_upscale, weight_window;
    uint32_t split_cap;
    uint32_t n_pixels;
};

// Outputs of the fused network kernels.
struct EpiArgs {
    float* out;                  // EPI_RAW: [n_out][out_stride]
    uint32_t out_stride;
    int head = -1;               // EPI_RAW: >= 0 writes outputs 3 head .. 3 head + 2 only (AllSelf)
    float* q;                    // EPI_RRS/EPI_ADRRS: [n]
    unsigned long long* sum;     // fixed-point normalisation sum
    DecideParams d;
};

enum Epilogue { EPI_RAW = 0, EPI_RRS = 1, EPI_ADRRS = 2 };

// Path-node buffer (SoA, capacity cap) and the scratch of the training-data pass.
struct NodeBuf {
    float* pos = nullptr;       // [3][cap]
    float* dir = nullptr;       // [3][cap]
    float* normal = nullptr;    // [3][cap]
    float* rough = nullptr;     // [cap]
    float* ad = nullptr;        // [3][cap]
    float* as = nullptr;        // [3][cap]
    float* thp = nullptr;       // [3][cap]
    float* q = nullptr;         // [cap]
    float* L = nullptr;         // [3][cap] radiance found after the node
    int32_t* parent = nullptr;  // [cap]
    uint32_t* pixel = nullptr;  // [cap]
    uint32_t* depth = nullptr;  // [cap]
    uint32_t* cnt = nullptr;    // [0] nodes, [1] invalid records
    uint32_t* first = nullptr;  // [cap] lowest child of each node
    uint32_t* tiles = nullptr;  // [cap / kNodeTile + 1] kept rows per tile, then row offsets
    uint32_t cap = 0;
};
constexpr uint32_t kNodeTile = 1024;

struct NetState {
    int n_layers = 0, n_out = 0;
    GridParams grid{};               // the net's grid (n_levels = 0: a net without one)
    uint64_t n_grid = 0, n_mlp = 0;  // FP32 parameter counts
    float* params = nullptr;         // FP32 masters [grid | mlp]
    float* grid_f32 = nullptr;       // = params: grid [entries][2]
    float* mlp_f32 = nullptr;        // = params + n_grid: MLP weights, row-major W_l[out][in]
    uint16_t* grid_lp = nullptr;     // FP16/BF16 shadow of the grid
    uint16_t* mlp_lp = nullptr;      // FP16/BF16 shadow of the MLP (MFMA tile order)
    // Ema optimizer (nrrs_config.ema_decay > 0): debiased moving average of the masters [grid | mlp],
    // the parameters every inference entry point reads, and its FP16/BF16 shadows
    float* ema = nullptr;
    uint16_t* grid_lp_ema = nullptr;
    uint16_t* mlp_lp_ema = nullptr;
    // training state (allocated by nrrs_create when max_train > 0)
    bool train_ready = false;
    float* grad = nullptr;           // [grid | mlp]
    uint16_t* grad_lp = nullptr;     // FP16 grid gradient of FP16/BF16 contexts [n_grid]
    void* dense_rep = nullptr;       // dense-level gradient replicas (FP32 pairs or FP16 pairs)
    uint32_t n_dense = 0;            // table entries of the dense levels
    // binned table-gradient scatter (FP16/BF16 contexts); see train_common.hpp
    bool binned = false;
    uint32_t* rec_key = nullptr;     // gradient records: keys
    uint32_t* rec_val = nullptr;     // and values
    uint32_t* gdl = nullptr;         // [level][max_train] dL/denc FP16 pairs of the training kernel
    uint32_t* bin_runs = nullptr;    // [level][chunk][segment] run counts / first slots
    uint32_t* bin_tiles = nullptr;   // [level][tile][segment] scan scratch
    uint32_t bin_max_tiles = 0;
    uint32_t* seg_end = nullptr;     // [level][kMaxSegBins] end slot of every segment (0: none)
    uint32_t* seg_count = nullptr;   // records per segment
    uint32_t n_segs = 0;
    void* segs = nullptr;            // SegInfo per segment
    uint32_t seg_base[kMaxLevels] = {}, seg_nbins[kMaxLevels] = {};
    uint32_t* direct = nullptr;      // segments applied by one workgroup each
    uint32_t n_direct = 0;
    uint32_t* chunked = nullptr;     // segments applied in chunks
    uint32_t n_chunked = 0;
    uint2* chunks = nullptr;         // {segment, first record}
    uint32_t n_chunks = 0;
    uint32_t apply_chunk = 0;        // records per chunk of the chunked apply
    unsigned long long* acc = nullptr;  // fixed-point sums of the chunked segments
    uint32_t* bad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    void* w_fwd = nullptr;           // packed training tiles (forward / transposed)
    void* w_bwd = nullptr;
    float* dw_partial = nullptr;     // [max_train_blocks][n_mlp]
    float* loss_scale = nullptr;     // device scalar
    uint32_t* state = nullptr;       // device [8]: step, good steps, skipped, non-finite gradient flag,
                                     // non-finite batch losses
    double* loss_partial = nullptr;  // [max_train_blocks] FP64 batch-loss partials of the training kernel
};

}  // namespace nrrs

struct nrrs_ctx {
    nrrs_config cfg;
    int device = 0;
    nrrs::GridParams grid;
    nrrs::Frame frame;
    nrrs::DecideParams dec;
    nrrs::NetState net[3];  // StatNet, RRSNet, and AllSelf's second StatNet (ll2_two_head = 3)
    // RRS-stage scratch
    unsigned long long* sum_dev = nullptr;  // [1]
    double* err_part = nullptr;             // per-block partial sums of the pixel error
    uint32_t* grid_enc = nullptr;           // [levels][grid_enc_cap] level-major encoding (FP16/BF16)
    uint32_t* grid_enc2 = nullptr;          // RRSNet's encoding of a paired pass, same shape
    uint32_t grid_enc_cap = 0;
    uint32_t* grid_pair = nullptr;          // paired table: [entries][2 nets] FP16/BF16 words, or null
    uint32_t* grid_pair_ema = nullptr;      // the same of the moving averages (Ema contexts), or null
    float2* grid_enc_f32 = nullptr;         // FP32 level-major encoding [levels][grid_enc_cap], or null
    nrrs::NodeBuf nodes;                    // path nodes of the training paths
    void* nodes_mem = nullptr;
    uint32_t* tile_sums = nullptr;          // [max tiles]
    uint32_t* tile_offsets = nullptr;       // [max tiles]
    float* stats_scratch = nullptr;         // [6][max_queue]
    float* q_scratch = nullptr;             // [max_queue]
    uint8_t* s_scratch = nullptr;           // [max_queue]
    uint32_t max_tiles = 0;
    // frame (persistent multi-bounce) state
    uint32_t frame_cap = 0;
    float* frame_q = nullptr;
    uint8_t* frame_s = nullptr;
    uint32_t* frame_tiles = nullptr;        // per-workgroup survivor totals of the frame kernel
    float* frame_x = nullptr;               // [3][cap] positions of the current bounce
    uint32_t* frame_genc = nullptr;         // [2][levels][cap] level-major encodings
    unsigned long long* frame_sums = nullptr;  // [16]
    uint32_t* frame_bar = nullptr;          // frame kernel's grid barrier
    uint32_t frame_waves = 4, frame_m = 128;
    hipGraph_t frame_graph = nullptr;
    hipGraphExec_t frame_exec = nullptr;
    // training scratch
    float* train_stats = nullptr;           // [6][max_train] StatNet outputs of the batch
    float* scale_tmp = nullptr;             // device scalar for nrrs_backward's loss scale
    uint32_t max_train_blocks = 256;  // training workgroups: one per CU, two for the narrow kernels (set at create)
    // helper stream (default priority) on which a joint training step runs the StatNet update
    // beside the RRSNet backward; joined to the caller's stream before the step returns
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // operation timing (nrrs_set_op_timing): an event pair per recorded operation
    bool op_timing = false;
    uint32_t op_count = 0;
    hipEvent_t op_ev[2 * NRRS_MAX_OP_TIMES] = {};
    const char* op_name[NRRS_MAX_OP_TIMES] = {};
};

namespace nrrs {

// ------------------------------------------------------------ launchers
hipError_t launch_encode(const EncodeArgs& a, int net, uint32_t n, const float* table, float* enc,
                         hipStream_t s);
hipError_t launch_grid_indices(const EncodeArgs& a, uint32_t n, uint32_t* idx, hipStream_t s);
hipError_t launch_net_fp32(const EncodeArgs& a, int net, int n_layers, int n_out, int epi, uint32_t n,
                           const float* table, const float* W, const EpiArgs& ep, hipStream_t s);
// Level-major FP32 grid encoding of one net's table: out[l * stride + i] = the feature pair of
// level l of vertex i, one level after the other over the batch.
hipError_t launch_encode_levels_f32(const EncodeArgs& a, const float* table, uint32_t n, float2* out, uint32_t stride,
                                    hipStream_t s);

// Launch set-up that is done once per device: `done` is the caller's static table, one flag per
// device ordinal (devices beyond kMaxDevices repeat the set-up on every launch).
constexpr int kMaxDevices = 64;
inline hipError_t set_dyn_lds_once(std::atomic<bool>* done, const void* kernel, size_t bytes) {
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return e;
    const bool cached = d >= 0 && d < kMaxDevices;
    if (cached && done[d].load(std::memory_order_acquire)) return hipSuccess;
    e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess && cached) done[d].store(true, std::memory_order_release);
    return e;
}

size_t mfma_packed_halves(int n_layers);
// FP16/BF16 shadows of one net: packed MLP tiles and, with `grid`, the grid table.
hipError_t launch_pack_lp(int precision, const float* grid, uint64_t n_grid, const float* W, int n_layers,
                          int n_out, uint16_t* grid_lp, uint16_t* W_lp, hipStream_t s);
// Level-major grid encoding: one workgroup per (level, kRunVert = 1024 vertices), levels in
// order, so the whole GPU gathers from one level's table at a time (it stays in L2). With
// run_counts, also counts the corners of every (level, table segment, 1024 vertices) for the
// binned scatter: run_counts[(level * ceil(n / 1024) + chunk) * kMaxSegBins + segment].
hipError_t launch_encode_levels(int precision, const EncodeArgs& a, const uint16_t* table, uint32_t n, uint32_t* out,
                                uint32_t stride, hipStream_t s, uint32_t* run_counts = nullptr);
// The same pass for both nets from the paired table (entry e = {StatNet word, RRSNet word}):
// StatNet's encoding to out_stat, RRSNet's to out_rrs, the record counts (one set serves both
// nets) to run_counts and run_counts2 where given.
hipError_t launch_encode_levels_dual(int precision, const EncodeArgs& a, const uint32_t* pair_table, uint32_t n,
                                     uint32_t* out_stat, uint32_t* out_rrs, uint32_t stride, hipStream_t s,
                                     uint32_t* run_counts = nullptr, uint32_t* run_counts2 = nullptr);
// Chained StatNet (4 hidden, 6 outputs) -> RRSNet (3 hidden, 1 output) inference with the RRS
// epilogue (q, exact sum) in one launch; a.grid_enc holds StatNet's level-major encoding,
// enc_rrs RRSNet's (same stride), a.stats must be null. Ws / Wr: packed MFMA tiles.
hipError_t launch_net_chained(int precision, const EncodeArgs& a, const uint32_t* enc_rrs, uint32_t n,
                              const uint16_t* Ws, const uint16_t* Wr, const EpiArgs& ep, hipStream_t s);
hipError_t launch_net_mfma(int precision, const EncodeArgs& a, int net, int n_layers, int n_out, int epi,
                           uint32_t n, const uint16_t* table, const uint16_t* Wp, const EpiArgs& ep,
                           hipStream_t s);

hipError_t launch_zero_u64(unsigned long long* p, hipStream_t s);
hipError_t launch_nodes_reset(const NodeBuf& nb, hipStream_t s);
hipError_t launch_nodes_record(const NodeBuf& nb, uint32_t n, const uint32_t* n_dev, const nrrs_vertices& v,
                               const int32_t* parent, const float* q, uint32_t depth, int32_t* out, hipStream_t s);
hipError_t launch_nodes_radiance(const NodeBuf& nb, uint32_t n, const uint32_t* n_dev, const int32_t* node,
                                 const float* L, uint32_t L_stride, hipStream_t s);
hipError_t launch_training_data(const NodeBuf& nb, const nrrs_training_args& a, uint32_t n_pixels, hipStream_t s);
hipError_t launch_error_blend(uint32_t n, const float* Lf, const uint32_t* spp, float blend, float* wL, hipStream_t s);
size_t error_partial_count();
hipError_t launch_error_compute(uint32_t n, const float* ref, const float* wL, const float* spp, int clamp, float* err,
                                double* part, float* sum, hipStream_t s);
hipError_t launch_qsum(nrrs_ctx* ctx, const float* q, uint32_t n, hipStream_t s);
hipError_t launch_decide(nrrs_ctx* ctx, const nrrs_decide_args& a, hipStream_t s);
// Vertices nrrs_rrs_decide_static takes at most (its scratch is a fixed device array).
constexpr uint32_t kStaticDecideMax = 1u << 22;
// Normalisation sum + decision + compaction with the fixed device scratch instead of a
// context's (allocates nothing; calls that overlap on one device race on that scratch).
hipError_t launch_decide_static(const DecideParams& d, const nrrs_decide_args& a, hipStream_t s);
hipError_t launch_consume(uint32_t capacity, const uint32_t* n_dev, const uint32_t* surv, const nrrs_vertices& sc,
                          const float* rrs, const int32_t* node_in, const nrrs_vertices_out& o, int32_t* node_out,
                          uint32_t* stid_out, hipStream_t s);

hipError_t launch_synth(uint32_t n, uint64_t seed, uint32_t n_pixels, const nrrs_vertices_out& o,
                        hipStream_t s);

}  // namespace nrrs
