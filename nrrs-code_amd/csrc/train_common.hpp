// Arguments and launchers of the training kernels.
#pragma once

#include "host_common.hpp"

namespace nrrs {

// Private copies of the dense-level table gradient (see the scatter in k_train).
constexpr uint32_t kDenseGradReplicas = 16;
// Segments of the binned table-gradient scatter: every level is cut into segments of
// kSegEntries entries (a level has at most kMaxSegBins of them).
constexpr uint32_t kSegShift = 13;
constexpr uint32_t kSegEntries = 1u << kSegShift;
constexpr uint32_t kMaxSegBins = 64;
// The encode pass counts the records of every (level, segment, kRunVert vertices); a scan of
// these counts gives each run its first record slot.
constexpr uint32_t kRunVert = 1024;
// Chunks per tile of the scan of the run counts.
constexpr uint32_t kBinScanRows = 256;
// Records one workgroup of the chunked apply sums (segments of the dense levels, where one
// segment can receive several records per vertex): max_train / 8 rounded down to a power of
// two in [kApplyChunkMin, kApplyChunkMax], so that a full batch gives a similar number of
// chunks at every batch size.
constexpr uint32_t kApplyChunkMin = 1u << 12, kApplyChunkMax = 1u << 17;

struct SegInfo {
    uint32_t rec_off;    // first record slot in the record buffer
    uint32_t cap;        // record slots; records beyond are added atomically by k_train
    uint32_t entry0;     // first table entry of the segment
    uint32_t n_entries;  // entries of the segment (the last one of a level can be short)
};

// Table-gradient records: key (u32) = entry within the segment (bits 0-12) | pair code << 13
// (see pair_code) | x fraction of the cell in 2^-16 units << 16 (pair codes 0-2); value (u32) =
// FP16 pair of w_y w_z dL/denc for a pair of the key's entry (x bit 0) and its x-neighbour
// entry ^ (2^(code+1) - 1), split between the two by the x fraction when summed; for code 3
// the FP16 pair of the key's entry alone.
struct ApplyArgs {
    const uint32_t* key;             // record keys, indexed like the slots of the segments
    const uint32_t* val;             // record values
    uint32_t* seg_count;             // records appended per segment, cleared after use
    const SegInfo* segs;
    const uint32_t* direct;          // segments summed by one workgroup each (hashed levels)
    const uint2* chunks;             // {segment, first record} of the chunked apply (dense levels)
    const uint32_t* chunked;         // the segments of the chunked apply
    uint32_t n_chunked;
    uint32_t chunk;                  // records per chunk (a power of two)
    unsigned long long* acc;         // [n_acc][2] fixed-point sums of the chunked segments
    uint32_t n_acc;                  // entries covered by acc (the dense levels come first)
    uint32_t* bad;                   // non-finite record seen by the chunked apply
    _Float16 __attribute__((ext_vector_type(2)))* grad;  // FP16 pair table gradient
    uint32_t* nonfinite;             // the optimizer's non-finite-gradient flag: set when a sum or a
                                     // gradient entry written is not finite
    bool keep_counts = false;        // leave seg_count for a second apply of the same records
};

struct TrainArgs {
    EncodeArgs a;                  // vertices, film, RRSNet statistics input, grid
    int net;
    uint32_t n;
    const float* table_f32;        // FP32 grid (FP32 contexts) or null
    const uint32_t* table_lp;      // FP16/BF16 grid (two features per 32-bit word)
    const void* w_fwd;             // packed forward A tiles
    const void* w_bwd;             // packed W^T A tiles
    int bwd_hidden_base, bwd_last_base;
    const float* loss_scale;       // device scalar
    // loss inputs
    const float* dy_ext;           // [n_out][n] (external loss)
    const float* target_L;         // [3][n] (StatNet)
    const float* ref_mean;         // RRS side arrays [n]
    const float* error;
    const float* sample_weight;
    const float* q_target;
    float ebar;                    // max(error_sum / pixels, 1e-4)
    const float* error_sum_dev;    // device error sum replacing ebar's error_sum, or null
    float n_pixels_f;
    float gamma[4];
    int clamp_on;
    float clamp_max;
    int train_sigma;               // StatNet channels 3-5 model sigma (1) or E[L^2] (0)
    uint32_t ll2_step, rrs_step;
    float inv_norm;                // loss normalisation
    // outputs
    float* y_out;                  // [n_out][out_stride] forward outputs, or null
    uint32_t out_stride;
    float* grad_grid;              // [entries][2], atomically accumulated (FP32 contexts)
    _Float16 __attribute__((ext_vector_type(2)))* grad_grid_h2;  // [entries] FP16 pairs (FP16/BF16)
    float* dense_f32;              // [replicas][n_dense][2] (FP32 contexts)
    _Float16 __attribute__((ext_vector_type(2)))* dense_h2;  // [replicas][n_dense] (FP16/BF16)
    uint32_t n_dense;              // entries of the dense levels (they come first in the table)
    // Binned scatter (FP16/BF16 contexts): the kernel writes dL/denc of the grid as
    // [level][vertex] FP16 pairs (row stride gdl_stride) for launch_bin_levels instead of
    // adding it to the table. Null selects direct atomics (and dense-level replicas).
    uint32_t* gdl;
    uint32_t gdl_stride;
    float* dw_partial;             // [gridDim.x][n_mlp]
    int head = -1;                 // AllSelf StatNet: NRRS_LL2 terms and y_out of outputs
                                   // 3 head .. 3 head + 2 only (-1: every output)
    int head_first;                // two-headed StatNet: first split layer (-1: one head);
                                   // weights between the heads get a zero gradient
    uint32_t n_mlp;
    double* loss_partial;          // [gridDim.x] FP64 batch-loss partials (LL2 / RRS losses), or null
};

int train_tiles_fwd(int nl);
int train_tiles_bwd(int nl);
size_t train_op_bytes(int precision);
hipError_t launch_pack_train(int precision, const float* W, int nl, int n_out, void* fwd, void* bwd, hipStream_t s);
hipError_t launch_train(int precision, const TrainArgs& t, int nl, int n_out, int loss, uint32_t blocks,
                        hipStream_t s);
struct BinArgs {
    EncodeArgs a;                  // vertices and grid (positions of the records' corners)
    const uint32_t* gdl;           // [level][gdl_stride] dL/denc FP16 pairs written by the training kernel
    uint32_t gdl_stride;
    uint32_t n;
    uint32_t* rec_key;             // record keys and values (see ApplyArgs)
    uint32_t* rec_val;
    uint32_t* seg_count;           // records per segment (written by the scan)
    uint32_t* runs;                // [level][n_runs][kMaxSegBins]: record counts of the kRunVert-vertex
                                   // chunks from the encode pass, replaced by their first slots
    uint32_t n_runs;               // chunks: ceil(n / kRunVert)
    uint32_t* tile_sums;           // [level][max_tiles][kMaxSegBins] scan scratch
    uint32_t max_tiles;            // tiles of kBinScanRows chunks tile_sums holds per level
    const uint32_t* seg_end;       // [level][kMaxSegBins] one past the last slot of each segment
    const SegInfo* segs;
    uint32_t seg_base[kMaxLevels];   // first segment of each level
    uint32_t seg_nbins[kMaxLevels];  // segments of each level
    _Float16 __attribute__((ext_vector_type(2)))* grad;  // FP16 pair table gradient (full segments)
    // set when a full segment's records are added to the table gradient directly (their sums
    // are then not checked by the apply kernels; the optimizer scans the table), or null
    uint32_t* scan_flag;
    uint32_t* scan_flag2;
    // second net of a joint step (same grid, same record slots), or null: its dL/denc, its
    // record values (the keys are shared) and its table gradient
    const uint32_t* gdl2;
    uint32_t* rec_val2;
    _Float16 __attribute__((ext_vector_type(2)))* grad2;
};
// Scans the run counts into first slots and segment totals, then writes the records of every
// (vertex, level) (FP16 w * dL/denc of its corners, x-neighbours paired) into the slots of
// their segments, walking the levels one after the other.
hipError_t launch_bin_levels(const BinArgs& b, hipStream_t s);
// The two halves of launch_bin_levels: the scan, then the record pass (with gdl2 set, the
// records of both nets of a joint step in one pass).
hipError_t launch_bin_scan(const BinArgs& b, hipStream_t s);
hipError_t launch_bin_records(const BinArgs& b, hipStream_t s);
// Sums the binned records into the table gradient: n_direct segments with one workgroup
// each, n_chunks chunks of the other segments, then their fixed-point sums.
hipError_t launch_apply_bins(const ApplyArgs& a, uint32_t n_direct, uint32_t n_chunks, hipStream_t s);
// Vertices one training workgroup processes per iteration.
uint32_t train_vertices_per_iter(int precision, int net_set);
hipError_t launch_reduce_dense(float* rep_f32, void* rep_h, uint32_t n_dense, float* grad, void* grad_h,
                               hipStream_t s);
hipError_t launch_reduce_partials(const float* part, uint32_t n_blocks, uint32_t n_mlp, float* grad, hipStream_t s);
// Batch loss = FP64 sum of the training kernel's per-workgroup partials (fixed order), written
// to loss_out (or not, when null); a non-finite sum increments *nonfinite.
hipError_t launch_loss_finish(const double* part, uint32_t n_blocks, float* loss_out, uint32_t* nonfinite,
                              hipStream_t s);

struct AdamArgs {
    float* params;        // [n] FP32 masters
    float* grad;          // [n], zeroed after use (its grid part is unused with grad_lp)
    uint16_t* grad_lp;    // [n_grid] FP16 grid gradient (FP16/BF16 contexts) or null
    float* m;
    float* v;
    uint64_t n;
    uint64_t n_grid;      // leading grid parameters (their low-precision shadow is refreshed)
    uint16_t* grid_lp;    // or null
    int lp_precision;
    float lr, beta1, beta2, eps, l2;
    float* loss_scale;    // device
    uint32_t* state;      // device: [0] step, [1] good steps, [2] skipped, [3] non-finite gradient flag,
                          // [4] non-finite batch losses (read and cleared by nrrs_train_state),
                          // [5] the table gradient was written by direct atomics (scan it)
    // Binned contexts, whose apply kernels report non-finite table-gradient entries themselves:
    // the finite check scans the FP16 table gradient only when state[5] is set. Else false: it
    // always scans.
    bool scan_on_flag;
    int dynamic_scale;
    uint32_t growth_interval;
    // Ema: after the update of step t, ema = (decay (1 - decay^(t-1)) ema + (1 - decay) p) /
    // (1 - decay^t) (a skipped step leaves it), its grid part also written to ema_lp. Null: off.
    float* ema = nullptr;
    uint16_t* ema_lp = nullptr;
    float ema_decay = 0.0f;
};
// Loss-scale and step state after an optimizer step (one thread): a step whose gradient had a
// NaN/Inf is counted as skipped and halves the scale; growth_interval finite steps in a row
// double it.
__device__ inline void adam_finish_state(const AdamArgs& a) {
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
    st[5] = 0;
}
// The finite check and the update; with `finish` also the state update (else the caller runs
// adam_finish_state, as launch_after_adam does).
hipError_t launch_adam(const AdamArgs& a, hipStream_t s, bool finish = true);
// After launch_adam(..., false) of an FP16/BF16 net: the state update, the inference tiles
// (mlp_lp) and the training tiles (fwd, bwd) of the updated MLP W, and, with a moving average
// W_ema (or null), its inference tiles (mlp_lp_ema), in one launch.
hipError_t launch_after_adam(int precision, const AdamArgs& a, const float* W, const float* W_ema, int nl, int n_out,
                             uint16_t* mlp_lp, uint16_t* mlp_lp_ema, void* fwd, void* bwd, hipStream_t s);
hipError_t launch_fill_f32(float* p, float v, uint64_t n, hipStream_t s);
// Paired table (entry e = {StatNet word e, RRSNet word e}) from the nets' FP16/BF16 grids,
// both 16-byte aligned with n_entries words each.
hipError_t launch_pack_pair(const uint16_t* stat_lp, const uint16_t* rrs_lp, uint32_t* pair, uint64_t n_entries,
                            hipStream_t s);

}  // namespace nrrs
