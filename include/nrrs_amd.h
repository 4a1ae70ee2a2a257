/*
 * nrrs_amd.h - C ABI of the MI355X (gfx950) neural Russian-roulette-and-splitting stage.
 *
 * One context owns everything the reference's NRRSNetwork owns
 * (src/render/nrrs/net/network.h:22-214): the StatNet and RRSNet parameters
 * (hash grid + MLP, FP32 master copies and FP16/BF16 shadows), their Adam state,
 * and the scratch buffers of the RRS decision. The caller owns the queues and the
 * film and passes device pointers, as the integrator does
 * (src/render/nrrs/integrator.cpp:1057-1483).
 *
 * Conventions
 *  - Every function returns an nrrs_status (0 = ok) and never throws.
 *  - Pointers named *_dev are device pointers; everything else is host memory.
 *  - Work is enqueued on the hipStream_t passed as `stream` (NULL = default stream)
 *    and is asynchronous unless the comment says otherwise.
 *  - Per-vertex arrays are component-major: field c of vertex i is at [c * stride + i].
 */
#ifndef NRRS_AMD_H
#define NRRS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NRRS_OK = 0,
    NRRS_ERR_ARG = 1,         /* null pointer / invalid value */
    NRRS_ERR_SHAPE = 2,       /* size beyond the capacity fixed at create time */
    NRRS_ERR_OOM = 3,         /* device allocation failed */
    NRRS_ERR_OVERFLOW = 4,    /* survivor queue capacity exceeded (queue truncated) */
    NRRS_ERR_NONFINITE = 5,   /* NaN / Inf met where a finite value is required */
    NRRS_ERR_HIP = 6,         /* HIP runtime error */
    NRRS_ERR_UNSUPPORTED = 7, /* configuration this build does not implement */
    NRRS_ERR_NO_DEVICE = 8    /* no usable gfx950 device */
} nrrs_status;

typedef enum { NRRS_FP32 = 0, NRRS_FP16 = 1, NRRS_BF16 = 2 } nrrs_precision;

/* Opaque context: one per device, used from one host thread at a time. */
typedef struct nrrs_ctx nrrs_ctx;

/* Network ids, as NRRSNetwork::LL2_NET / RRS_NET (src/render/nrrs/net/network.h:211-212). */
/* NRRS_STATNET2: the second StatNet of AllSelf (ll2_two_head = 3), the L^2 head. */
typedef enum { NRRS_STATNET = 0, NRRS_RRSNET = 1, NRRS_STATNET2 = 2 } nrrs_net;

/* Where the RRS factor q comes from (src/render/nrrs/integrator.cpp:1184-1256). */
typedef enum {
    NRRS_Q_RRSNET = 0,         /* q = activationSigma(RRSNet(encode(v)))                 */
    NRRS_Q_RRSNET_CHAINED = 1, /* StatNet stats feed the RRSNet input, then as above     */
    NRRS_Q_STATNET_ADRRS = 2   /* q = mean(clamp(L,1e-4,1e4) * thp / (I_px + 0.01))      */
} nrrs_q_mode;

/* Multiresolution hash grid (tcnn "HashGrid", 2net-simple.json:46-51). */
typedef struct {
    uint32_t n_levels;        /* 1..16                                   */
    uint32_t n_features;      /* must be 2                               */
    uint32_t log2_table_size; /* T = 2^log2_table_size entries per level  */
    uint32_t base_resolution;
    float per_level_scale;
} nrrs_grid_config;

typedef struct {
    uint32_t n_hidden_layers; /* hidden layers of width 64 (1..6)      */
    uint32_t n_outputs;       /* StatNet 6, RRSNet 1                      */
} nrrs_mlp_config;

typedef struct {
    nrrs_grid_config grid;   /* shared by both nets (each net has its own table) */
    nrrs_mlp_config statnet; /* 4 hidden, 6 outputs                              */
    nrrs_mlp_config rrsnet;  /* 3 hidden, 1 output                               */
    int32_t precision;       /* nrrs_precision of the MLP contractions           */

    /* RRS decision: integrator.h:265, render/test.json:63-67 */
    float rrs_clamp_min;        /* 0.05 */
    float rrs_clamp_max;        /* 20   */
    float rrs_normalize_rate;   /* 0.85, used when !rrs_use_ceil */
    int32_t rrs_use_ceil;       /* 1 */
    int32_t rrs_no_upscale;     /* 1 = donnot_normalize_when_less_than_one */
    int32_t rrs_weight_window;  /* 0 = off (RRSWeightWindow, commoncudautils.h:14-28) */
    float weight_window_size;   /* 5 */
    uint32_t split_cap;         /* largest split count, 8 */

    /* Losses (2net-simple.json:3-4, 145-148) */
    int32_t loss_clamp_on;  /* 1  */
    float loss_clamp_max;   /* 10 */
    /* train_sigma (network.h:158, train.cu:196-200): 0 = StatNet channels 3-5 model E[L^2];
     * 1 = they model sigma through activationSigma, E[L^2] = sigma^2 + L^2. With 1 the
     * RRSNet training input takes activationSigma(max(L2, 0)) where inference takes
     * max(L2, 0), as the reference's two input builders do. */
    int32_t train_sigma;    /* 0 */
    uint32_t ll2_step;      /* 0: L only, >=1: L and L^2 */
    uint32_t rrs_step;      /* 3: on-policy RRSNet loss; <3: RRSNet gradient is zero */
    float gamma[4];         /* 0.01, 0.05, 0.0, 0.01 */

    /* Adam (2net-simple.json:16-28), FP32 master weights */
    float lr_statnet, lr_rrsnet;
    float adam_beta1, adam_beta2, adam_eps, adam_l2_reg;
    float loss_scale_init;          /* dynamic loss scaling start (FP16/BF16), 2^15 */
    uint32_t loss_scale_growth_interval;

    /* Scene and frame */
    float aabb_min[3], aabb_max[3];
    uint32_t n_pixels;   /* pixels of the film (resolution) */
    uint32_t max_queue;  /* capacity of one inference batch (vertices) */
    uint32_t max_train;  /* capacity of one training batch */
    uint64_t seed;       /* parameter initialisation seed */
    uint32_t max_nodes;  /* capacity of the path-node buffer (0: 4 x max_queue) */
    /* FP16/BF16 contexts: 1 (default) keeps, beside each net's table, a paired table with both
     * nets' entries interleaved (8 bytes per entry, 4 x the grid parameters in bytes), so that
     * the chained inference, the joint training step and the frame kernel gather both nets'
     * grids in one pass; 0 gathers each net's table in a pass of its own. Results are bitwise
     * the same either way. No reference counterpart. */
    uint32_t paired_tables;
    /* FP32 contexts: 1 (default) gathers the grid level by level over a batch into a level-major
     * encoding that the network kernel reads (each level's table stays in L2 meanwhile); 0
     * gathers all levels of a vertex inside the network kernel. Results are bitwise the same.
     * No reference counterpart. */
    uint32_t fp32_level_major;
    /* Two-headed StatNet (ll2_use_two_head / ll2_two_head_type, myNetworkLL2.cpp:21-29,
     * 296-322): 0 = one head (default); 1 = PartShare (twn_PartShare.cpp: shared encoding and
     * base net, a tail per head), with statnet.n_hidden_layers = 3 (base: 2 layers, tails: 2);
     * 2 = JustEncoding (twn_JustEncoding.cpp: shared encoding, a net per head), with
     * statnet.n_hidden_layers = 2. Each head is 32 wide; the two heads of a layer form one
     * 64-wide layer whose weights between the heads are held at zero (set to zero by
     * nrrs_params_set, never trained). Outputs 0-2 (L) come from head 1, 3-5 (L^2) from head 2;
     * NRRS_LL2 over the 6 outputs is the mean of the two heads' losses. 3 = AllSelf
     * (twn_AllSelf.cpp: two nets, each with its own encoding, optimizer and loss), with
     * statnet.n_hidden_layers = 2: net NRRS_STATNET predicts L (outputs 0-2), net NRRS_STATNET2
     * L^2 (outputs 3-5); each is a 64-wide StatNet whose output rows of the other head are
     * held at zero, trained on its head's half of NRRS_LL2 with Adam at lr_statnet / 5
     * (twn_AllSelf.cpp:9-16); train_sigma = 1 is unsupported with it. The frame kernel
     * (nrrs_frame_*) returns NRRS_ERR_UNSUPPORTED for every two-headed StatNet. */
    int32_t ll2_two_head;
    /* AID RRSNet input (net1_aid_net2, network.cpp:31, 209-219, 242-258, 303-308, 340-343;
     * NRRSNetworkInput1AID, train.h:58-68): 0 = RRSNet reads the StatNet statistics (default);
     * 1 = RRSNet reads position, direction, roughness, reference mean and throughput only
     * (its normal, albedo and statistics channels are zero) and runs without StatNet at
     * inference (NRRS_Q_RRSNET). Training still runs StatNet: the NRRS_RRS loss takes its
     * statistics of the batch (the reference's ll2 buffer). NRRS_Q_RRSNET_CHAINED and the frame
     * kernel return NRRS_ERR_UNSUPPORTED with 1. */
    int32_t rrs_aid;
    /* Network set. 0 = the benchmark nets (default): 64-wide ReLU MLPs on the 80-channel input
     * (grid, SH-4 of direction and normal, per-vertex channels). 1 = the reference's shipped
     * nets (2net-simple.json:29-75, 210-247, 250-291): 32-wide hidden layers with LeakyReLU
     * (slope 0.01), no bias. StatNet reads NRRSNetworkInput0 (train.h:22-29: position, lat-long
     * direction, warped roughness) encoded as HashGrid (cfg.grid) + OneBlob of the two
     * direction coordinates (4 bins each) + OneBlob of the roughness (8 bins). RRSNet reads
     * NRRSNetworkInput1 (train.h:39-56: Box-Cox of the StatNet statistics, of the throughput
     * and of the reference mean, the roughness and a constant 1) and has no grid; with
     * rrs_aid it reads NRRSNetworkInput1AID instead (train.h:58-68): HashGrid + OneBlob of the
     * direction (4 bins each) + OneBlob of the roughness (4 bins) + Box-Cox of the reference mean
     * and the throughput. The layers keep the 64-wide, 80-channel storage of set 0: the weights
     * outside the 32-wide net are held at zero (nrrs_params_set zeroes them; they receive zero
     * gradients). Requires ll2_two_head = 0 and statnet / rrsnet of 3 hidden layers; without
     * rrs_aid, RRSNet always reads statistics (NRRS_Q_RRSNET is refused); the frame kernel
     * returns NRRS_ERR_UNSUPPORTED. nrrs_shipped_config fills the whole shipped configuration. */
    int32_t net_set;
    /* Ema optimizer (Ema{Adam}, 2net-simple.json:16-18, 171-173): 0 = off (default); in (0, 1):
     * every Adam step also updates the debiased exponential moving average of each net's
     * parameters with this decay, and the inference entry points (nrrs_statnet_infer,
     * nrrs_rrsnet_infer, nrrs_rrs_infer_decide, nrrs_frame_*) read the averaged parameters;
     * training reads the trained ones. */
    float ema_decay;
} nrrs_config;

/* Network sets of nrrs_config.net_set. */
#define NRRS_NETS_BENCH 0
#define NRRS_NETS_SHIPPED 1

/* SoA path vertices of one bounce (replaces NRRSInferenceWorkItem, workitem.h:9-15). */
typedef struct {
    const float* pos;       /* [3][stride] world-space position          */
    const float* dir;       /* [3][stride] unit direction                 */
    const float* normal;    /* [3][stride] unit shading normal            */
    const float* rough;     /* [stride] roughness                         */
    const float* albedo_d;  /* [3][stride]                                */
    const float* albedo_s;  /* [3][stride]                                */
    const float* thp;       /* [3][stride] throughput before RRS          */
    const uint32_t* pixel;  /* [stride] film pixel of the path            */
    const uint32_t* stid;   /* [stride] scatter-queue index; NULL = i     */
    uint32_t stride;
} nrrs_vertices;

/* Arguments of the per-bounce RRS stage. */
typedef struct {
    uint32_t n;                   /* inference queue size                               */
    nrrs_vertices v;              /* device SoA                                         */
    const float* film_dev;        /* [3][n_pixels] per-pixel estimate I (RGB)           */
    int32_t q_mode;               /* nrrs_q_mode                                        */
    const uint32_t* spec_stid_dev;/* [n_spec] specular stids, already in the queue       */
    const uint32_t* n_spec_dev;   /* device count of specular stids (may be NULL = 0)    */
    uint64_t pcg_seed;            /* PCG32 (seed, stream); vertex i draws output number i */
    uint64_t pcg_stream;
    /* outputs (device) */
    float* q_dev;                 /* [n] clamped q before normalisation (mRRSArray)       */
    float* qn_dev;                /* [n] normalised q' (mScatterRayQueue->mRRS), or NULL  */
    uint8_t* split_dev;           /* [n] split count s, or NULL                           */
    uint32_t* survivors_dev;      /* [capacity] compacted survivor stids                  */
    uint32_t capacity;
    uint32_t* n_survivors_dev;    /* device: number of survivors (incl. specular)         */
    uint32_t* overflow_dev;       /* device: survivors dropped for lack of capacity, or NULL */
    float* qn_stid_dev;           /* q' stored at its scatter item, qn_stid[stid_i] (the
                                     reference's mScatterRayQueue->mRRS[stid],
                                     integrator.cpp:1408), or NULL                        */
} nrrs_decide_args;

/* ---------------- configuration and lifetime ---------------- */

/* Fills the synthetic-benchmark defaults (16-level grid, 64-wide MLPs, render/test.json RRS knobs). */
nrrs_status nrrs_default_config(nrrs_config* cfg);

/* Fills the reference's shipped configuration (common/configs/nrrs/nrrs/2net-simple.json with
 * the render/test.json RRS knobs): net_set NRRS_NETS_SHIPPED; grid 8 levels x 2 features,
 * T = 2^19, base 16, scale 2.0 (lines 46-51); StatNet and RRSNet 3 hidden layers
 * (lines 74, 246); NRRS_LL2 with clamp 10 (lines 3-4); NRRS_RRS gammas 0.01, 0.05, 0, 0.01
 * (lines 145-148); Adam lr 5e-3 / 5e-4, beta (0.9, 0.99), eps 1e-15, l2 3e-6 inside
 * Ema(decay 0.95) (lines 16-28, 171-183). Scene, capacities and precision as
 * nrrs_default_config. */
nrrs_status nrrs_shipped_config(nrrs_config* cfg);

/* sizeof of the ABI structs, so a binding can check its own layout (no reference counterpart). */
size_t nrrs_struct_size(int32_t which); /* 0 config, 1 vertices, 2 decide_args, 3 vertices_out */

/* Replaces NRRSNetwork(json&) + resize/reset (network.cpp:28-35, 351-576). Allocates every
 * device buffer and initialises parameters (grid U[-1e-4,1e-4], MLP Xavier-uniform) from cfg->seed.
 * Synchronous. */
nrrs_status nrrs_create(const nrrs_config* cfg, nrrs_ctx** out);
nrrs_status nrrs_destroy(nrrs_ctx* ctx);
const char* nrrs_status_string(nrrs_status s);

/* Hash-grid level table: res/offset/size in entries, dense flag, FP32 scale per level.
 * Host-only. Returns the total number of grid parameters in *n_params. */
nrrs_status nrrs_grid_layout(const nrrs_grid_config* g, uint32_t* res, uint32_t* offset,
                             uint32_t* size, uint32_t* dense, float* scale, uint64_t* n_params);

/* ---------------- parameters (fixtures / finite differences) ---------------- */

/* Number of FP32 parameters of one net: grid entries * 2, then MLP weights. */
nrrs_status nrrs_param_count(nrrs_ctx* ctx, int32_t net, uint64_t* n_grid, uint64_t* n_mlp);
/* Host copy of [grid | W0 | W1 | ... ] (row-major W_l[out][in]). Synchronous. */
nrrs_status nrrs_params_get(nrrs_ctx* ctx, int32_t net, float* host_dst);
/* Upload, refresh low-precision shadows; with Ema (ema_decay > 0) the moving average restarts
 * at these parameters. Synchronous. */
nrrs_status nrrs_params_set(nrrs_ctx* ctx, int32_t net, const float* host_src);
/* Host copy of the parameters the inference entry points read, in the layout of
 * nrrs_params_get: the Ema moving average when ema_decay > 0, else the trained parameters
 * (tiny-cuda-nn's Trainer infers with its optimizer's averaged weights; 2net-simple.json:16-18).
 * Synchronous. */
nrrs_status nrrs_params_get_inference(nrrs_ctx* ctx, int32_t net, float* host_dst);

/* ---------------- inference (network.cpp:297-349) ---------------- */

/* Encoded network input [n][80] of `net` (debug / parity). stats_dev: [6][n] StatNet outputs
 * feeding a chained RRSNet, or NULL. */
nrrs_status nrrs_encode(nrrs_ctx* ctx, int32_t net, uint32_t n, const nrrs_vertices* v,
                        const float* film_dev, const float* stats_dev, float* enc_dev, void* stream);

/* Hash-grid table indices [n][levels][8] (corner bit0=x, bit1=y, bit2=z), within the level. */
nrrs_status nrrs_grid_indices(nrrs_ctx* ctx, uint32_t n, const nrrs_vertices* v,
                              uint32_t* idx_dev, void* stream);

/* StatNet inference: stats_dev [6][n] = (L_rgb, L2_rgb). */
nrrs_status nrrs_statnet_infer(nrrs_ctx* ctx, uint32_t n, const nrrs_vertices* v,
                               float* stats_dev, void* stream);

/* RRSNet inference: raw_dev [n] = network output before activationSigma. */
nrrs_status nrrs_rrsnet_infer(nrrs_ctx* ctx, uint32_t n, const nrrs_vertices* v,
                              const float* film_dev, const float* stats_dev, float* raw_dev,
                              void* stream);

/* ---------------- the RRS stage (integrator.cpp:1057-1483) ---------------- */

/* Inference + q + clamp/ceil + exact global sum + normalise + split + stable compaction. */
nrrs_status nrrs_rrs_infer_decide(nrrs_ctx* ctx, const nrrs_decide_args* a, void* stream);

/* Decision only, from a given clamped q (a->q_dev is an input here). */
nrrs_status nrrs_rrs_decide(nrrs_ctx* ctx, const nrrs_decide_args* a, void* stream);

/* Allocation-free form of nrrs_rrs_decide. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc) and may hold no context. The decision knobs and n_pixels come
 * from `cfg`; a->q_dev is an input, as for nrrs_rrs_decide, and the results are the same. Its
 * scratch (exact sum, tile totals, split counts) is a fixed array in the library's code object,
 * one per device, holding up to NRRS_STATIC_DECIDE_MAX vertices: a->n above that returns
 * NRRS_ERR_SHAPE. Every call on a device shares that device's scratch, so calls that overlap
 * on one device (several streams or host threads) race; serialise them. */
#define NRRS_STATIC_DECIDE_MAX (1u << 22)
nrrs_status nrrs_rrs_decide_static(const nrrs_config* cfg, const nrrs_decide_args* a, void* stream);

/* Synthetic vertices of the benchmark workloads, generated on the device from PCG32
 * (seed, field stream). out: buffers of a nrrs_vertices with stride >= n (written). */
typedef struct {
    float* pos; float* dir; float* normal; float* rough; float* albedo_d; float* albedo_s;
    float* thp; uint32_t* pixel; uint32_t stride;
} nrrs_vertices_out;
nrrs_status nrrs_synth_vertices(uint32_t n, uint64_t seed, uint32_t n_pixels,
                                const nrrs_vertices_out* out, void* stream);

/* Survivor consumer (handleIntersections, integrator.cpp:772-794): for k < min(*n_survivors_dev,
 * capacity), row k of out is scatter item sid = survivors_dev[k] of the scatter-queue SoA
 * `scatter` (indexed by stid; scatter->stid unused) with its throughput divided by rrs_dev[sid],
 * the q' stored per scatter item (1 for specular items). node_in_dev[sid] is carried to
 * node_out_dev[k] and sid to stid_out_dev[k] when those are given. Rows whose sid is not
 * below scatter->stride are left unwritten. */
nrrs_status nrrs_consume_survivors(nrrs_ctx* ctx, uint32_t capacity, const uint32_t* n_survivors_dev,
                                   const uint32_t* survivors_dev, const nrrs_vertices* scatter, const float* rrs_dev,
                                   const int32_t* node_in_dev, const nrrs_vertices_out* out, int32_t* node_out_dev,
                                   uint32_t* stid_out_dev, void* stream);

/* ---------------- training (network.cpp:196-295, myNetworkLL2.cpp:41-89) ---------------- */

/* Per-sample side inputs of the RRSNet loss (the arrays NRRSNetwork::getLossHelper hands
 * out, network.h:49-61), device pointers of length n. */
typedef struct {
    const float* ref_mean;      /* mean of the sample's pixel estimate; NULL = from the film */
    const float* error;         /* relative error of the sample's pixel                     */
    const float* sample_weight; /* 1 / (training samples of the pixel)                      */
    const float* q_target;      /* q used when the path was traced; -1 = duplicate node      */
    float error_sum;            /* sum of the per-pixel error over the film                 */
    const float* error_sum_dev; /* device scalar replacing error_sum when not NULL (the sum
                                   nrrs_error_compute writes, integrator.cpp:2466-2471)      */
} nrrs_rrs_side;

/* One training step: StatNet (net_mask bit 0) on the NRRS_LL2 loss against target_L
 * [3][v->stride] (the vertices' row stride; entries 0..n-1 used), then RRSNet (bit 1) on the NRRS_RRS loss, its StatNet-statistics inputs taken
 * from the StatNet forward of the same batch; backward; Adam on each trained net.
 * loss_dev (optional, device, 2 floats) receives the two batch losses. */
nrrs_status nrrs_train_step(nrrs_ctx* ctx, uint32_t net_mask, uint32_t n, const nrrs_vertices* v,
                            const float* film_dev, const float* target_L_dev, const nrrs_rrs_side* side,
                            float* loss_dev, void* stream);

/* Loss kind of nrrs_backward: dL/dy given, or one of the two losses. */
typedef enum { NRRS_LOSS_EXTERNAL = 0, NRRS_LOSS_LL2 = 1, NRRS_LOSS_RRS = 2 } nrrs_loss_kind;

/* Forward + backward of one net, accumulating dL/dparams (scaled by loss_scale) into the
 * context's gradient buffer. kind EXTERNAL: dy_dev [n_out][n] is dL/dy. LL2: target_L_dev
 * [3][v->stride].
 * RRS: side + stats_dev (StatNet statistics [6][n], also the RRSNet input). loss_dev: 1 float. */
nrrs_status nrrs_backward(nrrs_ctx* ctx, int32_t net, uint32_t n, const nrrs_vertices* v,
                          const float* film_dev, const float* stats_dev, int32_t kind,
                          const float* dy_dev, const float* target_L_dev, const nrrs_rrs_side* side,
                          float loss_scale, float* loss_dev, void* stream);

/* Gradient buffer of one net: zero it / copy it to the host ([grid | mlp], synchronous). */
nrrs_status nrrs_grads_zero(nrrs_ctx* ctx, int32_t net, void* stream);
nrrs_status nrrs_grads_get(nrrs_ctx* ctx, int32_t net, float* host_dst);

/* Adam on the FP32 masters from the gradient buffer (divided by the current loss scale);
 * refreshes the FP16/BF16 shadows and zeroes the gradient buffer. With dynamic loss scaling
 * (FP16/BF16 contexts) a non-finite gradient skips the update and halves the scale. */
nrrs_status nrrs_adam_step(nrrs_ctx* ctx, int32_t net, void* stream);

/* Training state of one net: Adam step count, current loss scale, skipped steps. Synchronous. */
nrrs_status nrrs_train_state(nrrs_ctx* ctx, int32_t net, uint32_t* step, float* loss_scale,
                             uint32_t* skipped);

/* ---------------- operation timing (measurement support, no reference counterpart) ----------------
 * With timing on, the library records a pair of HIP events on the launching stream around each
 * operation of the RRS stage and of training (one launcher: one kernel or a few, e.g.
 * "encode_pair", "mlp_chained", "decide", "train_statnet", "bin_records", "apply_rrsnet",
 * "adam_statnet"), up to NRRS_MAX_OP_TIMES per collection. nrrs_op_times waits for the device,
 * returns the names (static strings) and durations in ms of the operations recorded since the
 * previous call, in launch order, and starts a new collection. Turning timing on creates the
 * events (synchronous). */
#define NRRS_MAX_OP_TIMES 256
nrrs_status nrrs_set_op_timing(nrrs_ctx* ctx, int32_t on);
nrrs_status nrrs_op_times(nrrs_ctx* ctx, uint32_t* count, const char** names, float* ms);

/* ---------------- batched uneven wavefront: one persistent launch per frame ---------------- */

/* One frame of `n_bounces` RRS stages. Queue 0 holds the primary paths [0, n_primary)
 * (path id = pixel). Bounce b draws vertex k of its queue from the synthetic recipe with
 * seed + b * 0x9E3779B97F4A7C15 (standing in for the closest-hit program that fills the
 * inference queue, device.cu:97-108), runs StatNet -> RRSNet -> q -> normalise -> split
 * (PCG32 stream (seed, 1000 + b), vertex k draws output k) and compacts the survivors'
 * path ids into the next queue. Sizes stay on the device. FP16/BF16 contexts only. */
typedef struct {
    uint32_t n_primary;
    uint32_t n_bounces;       /* 1..16 */
    uint64_t seed;
    uint64_t* seed_dev;       /* optional device seed (read instead of `seed`), for graph replays */
    int32_t advance_seed;     /* with seed_dev: advance it at the end of the frame */
    const float* film_dev;    /* [3][n_pixels] */
    uint32_t* queue_dev;      /* [2][capacity] ping-pong path-id queues */
    uint32_t capacity;        /* <= max_queue of the context */
    uint32_t* sizes_dev;      /* [n_bounces + 1] size of the queue entering each bounce, then the final one */
    float* q_dev;             /* optional [n_bounces][capacity] clamped q per bounce, or NULL */
    uint32_t* trace_dev;      /* optional [n_bounces + 1][capacity] every queue, or NULL */
    uint32_t* status_dev;     /* [1] bit 0 barrier timeout, bit 1 queue overflow (must start at 0) */
} nrrs_frame_args;

/* Tile shape of the frame kernel: wavefronts per workgroup (1, 2, 4, 8) and vertices a
 * workgroup takes per step M (64, 128, 256; a multiple of 32 * waves). */
typedef struct {
    uint32_t waves;
    uint32_t m;
} nrrs_tile_shape;

nrrs_status nrrs_frame_run(nrrs_ctx* ctx, const nrrs_frame_args* a, void* stream);
nrrs_status nrrs_frame_set_shape(nrrs_ctx* ctx, const nrrs_tile_shape* shape);
/* Times every valid shape on this frame (synchronous) and keeps the fastest. */
nrrs_status nrrs_frame_autotune(nrrs_ctx* ctx, const nrrs_frame_args* a, nrrs_tile_shape* best, float* best_ms,
                                void* stream);
/* Captures one frame (with the current shape) into a hipGraph owned by the context. */
nrrs_status nrrs_frame_graph_capture(nrrs_ctx* ctx, const nrrs_frame_args* a, void* stream);
nrrs_status nrrs_frame_graph_launch(nrrs_ctx* ctx, void* stream);
/* The vertices bounce `bounce` of a frame draws for a queue (for tests and tools). */
nrrs_status nrrs_frame_vertices(uint64_t seed, uint32_t bounce, const uint32_t* queue_dev, uint32_t n,
                                const nrrs_vertices_out* out, void* stream);

/* ---------------- per-pixel error feeding NRRS_RRS (train.cu:344-430) ---------------- */

/* Running per-pixel radiance used by the error (nrrs_update_weighted_L, train.cu:344-363):
 * weighted_L = weighted_L (1 - blend) + (L_frame / spp) blend per channel, all [3][n_pixels];
 * spp_dev: samples of the pixel in L_frame. */
nrrs_status nrrs_error_blend(nrrs_ctx* ctx, uint32_t n_pixels, const float* L_frame_dev, const uint32_t* spp_dev,
                             float blend, float* weighted_L_dev, void* stream);

/* Relative squared error per pixel (nrrs_calculate_error, train.cu:365-430):
 * e = mean_c (L_c - ref_c)^2 / (ref_c^2 + 0.01), min(e, 100) when clamp; row 0 of error_dev
 * holds e, row 1 e * sqrt(spp) (spp_dev, samples per pixel this frame; NULL: row 1 = e).
 * error_sum_dev receives the sum of row 1 (integrator.cpp:2466-2471), summed in FP64 in a
 * fixed order. */
typedef struct {
    const float* ref_dev;        /* [3][n] reference estimate (accumulated film)   */
    const float* weighted_L_dev; /* [3][n] running estimate                        */
    const float* spp_dev;        /* [n] or NULL                                    */
    int32_t clamp;               /* clamp e at 100                                 */
    float* error_dev;            /* [2][n]                                         */
    float* error_sum_dev;        /* device scalar, or NULL                         */
} nrrs_error_args;
nrrs_status nrrs_error_compute(nrrs_ctx* ctx, uint32_t n_pixels, const nrrs_error_args* a, void* stream);

/* ---------------- training data from path nodes (workqueue.h:106-189, train.cu:432-592) ----------------
 * The context holds a buffer of path nodes: every non-specular vertex of a training path is
 * recorded with a link to the node of the previous vertex of its path; radiance found later
 * along the path is added to the node and all its ancestors; at the end of the frame the
 * nodes become training samples. Node indices are int32, -1 = none. */

/* Empties the node buffer (NRRSPathNodesBuffer::reset). */
nrrs_status nrrs_nodes_reset(nrrs_ctx* ctx, void* stream);

/* Appends one node per vertex (NRRSPathNodesBuffer::recordNode, workqueue.h:142-168):
 * vertex i -> node size + i, in vertex order. parent_dev[i]: node of the path's previous
 * vertex (NULL: all -1); q_dev[i]: the q' used at this vertex (NULL: 1); depth: bounce.
 * n_dev (optional, device): the vertex count, n then being its upper bound.
 * node_out_dev[i] receives the new node, or parent_dev[i] when the buffer is full (the
 * vertex then counts as invalid), as the reference's callers keep the previous node. */
nrrs_status nrrs_nodes_record(nrrs_ctx* ctx, uint32_t n, const uint32_t* n_dev, const nrrs_vertices* v,
                              const int32_t* parent_dev, const float* q_dev, uint32_t depth, int32_t* node_out_dev,
                              void* stream);

/* Adds radiance L_dev [3][L_stride] (entry i) to node node_dev[i] and every ancestor
 * (NRRSPathNodesBuffer::recordRadiance, workqueue.h:122-134); nodes < 0 or not yet
 * recorded are skipped. FP32 atomic adds: the sums are order-dependent in their last bits. */
nrrs_status nrrs_nodes_add_radiance(nrrs_ctx* ctx, uint32_t n, const uint32_t* n_dev, const int32_t* node_dev,
                                    const float* L_dev, uint32_t L_stride, void* stream);

/* Node count and invalid (dropped) records, read back to the host (synchronises). */
nrrs_status nrrs_nodes_size(nrrs_ctx* ctx, uint32_t* size, uint32_t* invalid);

/* Training samples from the recorded nodes (nrrs_generate_training_data, train.cu:432-570,
 * and nrrs_calculate_sample_weight, train.cu:572-592). Node j gives target
 * L = max(L_j / thp_j per channel where thp_j > 1e-5 else 0, 0); it is dropped when L, its
 * position or its direction has a NaN, or when ignore_zero_L and mean(L) == 0. Kept nodes
 * become rows in node order (the reference's atomic push order is not reproducible).
 * With mark_duplicates, a node whose parent has an earlier child (lower node index) gets
 * q_target -1 and is not counted in pixel_samples; other rows get q_target = the node's q.
 * copy_rrs_node multiplies the row's thp by the node's q. ref_mean = mean RGB of the film
 * pixel; error / error_avg = rows 0 / 1 of error_dev at the pixel mapped to the error image
 * (offsetFrame2Scaled, myFilm.h:6-15); sample_weight = 1 / pixel_samples[pixel] (0 if 0)
 * after all rows are counted. pixel_samples accumulates (zero it per frame). Rows beyond
 * capacity are dropped; n_out_dev receives the row count. */
typedef struct {
    const float* film_dev;     /* [3][n_pixels] */
    const float* error_dev;    /* [2][n_err] per-pixel error, or NULL (errors 0) */
    uint32_t n_err;            /* pixels of the error image */
    uint32_t error_scale;      /* frame / error image, per axis (>= 1) */
    uint32_t frame_width;
    int32_t ignore_zero_L;
    int32_t copy_rrs_node;
    int32_t mark_duplicates;
    uint32_t capacity;         /* rows of the outputs */
    nrrs_vertices_out out;     /* training vertices (stride >= capacity) */
    float* target_L;           /* [3][capacity] */
    float* ref_mean;           /* [capacity] */
    float* error;
    float* error_avg;
    float* q_target;
    float* sample_weight;
    uint32_t* pixel_samples;   /* [n_pixels] */
    uint32_t* n_out_dev;       /* device scalar */
} nrrs_training_args;
nrrs_status nrrs_generate_training_data(nrrs_ctx* ctx, const nrrs_training_args* a, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NRRS_AMD_H */
