// Arguments of the persistent frame kernel.
#pragma once

#include "host_common.hpp"

namespace nrrs {

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

struct FrameParams {
    GridParams grid;
    Frame frame;
    DecideParams d;
    const uint32_t* table_stat;   // FP16/BF16 grids (two features per word)
    const uint32_t* table_rrs;
    const uint2* table_pair;      // both nets' grids interleaved per entry, or null
    const u32x4_t* w_stat;        // packed MFMA tiles
    const u32x4_t* w_rrs;
    const float* film;
    uint32_t n_pixels;
    uint32_t n_primary, n_bounces, capacity, cpw;
    uint64_t seed;
    uint64_t* seed_dev;           // when set, the seed is read here (graph replays)
    int advance_seed;
    uint32_t* queue;              // [2][capacity]
    uint32_t* sizes;              // [n_bounces + 1]
    float* q_out;                 // [n_bounces][capacity] or null
    uint32_t* trace;              // [n_bounces + 1][capacity] or null
    uint32_t* status;
    // scratch
    float* q_scr;
    uint8_t* s_scr;
    uint32_t* tile_sums;          // per-workgroup survivor totals (at least kFrameMaxBlocks)
    unsigned long long* sums;     // [n_bounces]
    float* xbuf;                  // [3][capacity] normalised positions of the current bounce
    uint32_t* genc;               // [2][levels][capacity] level-major grid encodings (StatNet, RRSNet)
    uint32_t* bar;                // [kFrameBarWords] grid barrier
};

// Workgroups of a frame launch at most; the tile-sum scratch holds at least this many totals.
constexpr uint32_t kFrameMaxBlocks = 4096;
// u32 words of the grid barrier: top count, generation and 8 group counters on 128-byte lines.
constexpr uint32_t kFrameBarWords = 32 * 9;

hipError_t launch_frame(int precision, const FrameParams& f, int waves, uint32_t max_blocks, hipStream_t s);
hipError_t launch_bounce_vertices(uint64_t seed, uint32_t bounce, const uint32_t* queue, uint32_t n,
                                  const nrrs_vertices_out& o, hipStream_t s);

}  // namespace nrrs
