// Internal state of one simulation handle and the kernel launchers behind the
// C ABI.  Device layout (HBM), per handle:
//   particle state  pos float2[Na] x2 (ping-pong), vel float2[Na] x2, mass f32[Na]
//                   (direct sum: one mass copy, in the caller's order; Barnes-Hut:
//                   two mass copies and perm u32[Na] x2, the state of each set in
//                   that step's sorted order, perm[k][slot] = caller's index)
//   direct sum      one-sided: partial float2[S*Na] per-slice partial accelerations;
//                   symmetric: row partials float2[tiles*16*64] then column
//                   partials float2[(sum over groups of B - 16g)*64]
//   Barnes-Hut      cells uint2[Na] x2 (the sort's input {cell, slot} by particle
//                   index), keys/idx u32[Na] x2, radix histograms, sdep i8[Na+1],
//                   cpos u32[Na], prod float4[Na], leaf records,
//                   tile results,
//                   breaker u32[cap], node4 float4[cap] {x, y, mass, radius}
//                   (a particle node's radius word holds minus its run length),
//                   ncount u32[cap],
//                   sorted_src u32[Na], device counters
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "gsim.h"

namespace gsim {

// Fixed domain of the reference: [-500, 500)^2, a 1024^2 leaf grid, depth 10
// (BarnesHutSimulation.cpp:579,581).
constexpr float kSimSize = 500.0f;
constexpr uint32_t kTreeSize = 1024;
constexpr uint32_t kTreeDepth = 10;
// GSIM_TUNE_BH_ORDER = 2 (auto): every kOrderPeriod steps, kOrderTrial
// consecutive W = 32 steps alternate heaviest-first and index order
constexpr uint32_t kOrderTrial = 4;
constexpr uint64_t kOrderPeriod = 1024;
constexpr uint64_t kOrderReplan = 8;  // steps between plans of the order
constexpr uint32_t kKilledKey = 1u << 20;  // sorts after every 20-bit cell
// 15-bit leaves (opt-in, beyond the reference): a 32768^2 leaf grid, depth 15,
// 30-bit keys sorted in four 8-bit passes.
constexpr uint32_t kDeepBits = 15;
constexpr uint32_t kKilledKeyDeep = 1u << 30;  // sorts after every 30-bit cell
constexpr uint32_t kTileDepth = 5;          // tile kernel: one depth-5 cell per workgroup
constexpr uint32_t kTiles = 1u << (2 * kTileDepth);

struct BhCounters {
    uint32_t n_in;      // particles inside the domain (padding included)
    uint32_t n_cells;   // cell nodes (cells holding >= 2 particles)
    uint32_t n_nodes;   // n_in + n_cells = counts[0]
    uint32_t n_leaves;  // occupied leaf cells (depth 10, or 15 with 15-bit leaves)
    uint32_t next_group;         // work counter of the persistent traversal
    uint32_t n_gran;             // 15-bit leaves: occupied depth-10 cells
    uint32_t sort_stall;         // look-back waits that hit their bound (never expected)
    uint32_t n_heavy;            // static traversal: groups on the heavy list this step
    uint32_t tiles_busy;         // occupied depth-5 tiles
    uint32_t tiles_done;         // occupied tiles reduced (the last one reduces the top)
};

// Radix sort passes (single and batched builds): tiles of kSortTile keys,
// kSortRounds per thread of a kSortThreads workgroup.  Look-back status
// buffers hold 256 words per tile.
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;
// Tiles per group of the radix passes' two-level look-back; each pass's
// status words hold its tile words, then one group word set per kLookGroup
// tiles.
constexpr uint32_t kLookGroup = 16;
// The single-system build counts every pass's digits into kDigitCopies
// copies of the totals (workgroup b into copy b mod kDigitCopies), so that
// fewer workgroups add into each word; a pass sums the copies.
constexpr uint32_t kDigitCopies = 8;

// Particle runs a walking wave queues in its LDS before summing them (the
// staged walk's run stack; the batch may use fewer, BatchTune::run_queue).
constexpr uint32_t kRunQueue = 64;

// Static traversal: at most this many groups per step go to the heavy list
// (walked by four waves each beside the main launch).
constexpr uint32_t kMaxHeavy = 512;
// ... and only when the launch has more groups than the GPU holds at once
// (256 CUs x 4 SIMDs x 8 waves of the static traversal).
constexpr uint32_t kHeavyMinGroups = 8192;

struct TileCell {  // one cell of the dense tile/top reduction
    uint32_t n, y, f, pad;
    float sx, sy, sm, pad2;
};

struct BhBuffers {
    size_t na = 0, cap = 0, sort_blocks = 0, prep_blocks = 0;
    uint32_t bits = 10;              // leaf grid bits per axis the buffers are sized for
    // cells: the sort's input {cell, slot} by particle (caller's) index,
    // alternating between steps (the walk writes the next build's into the
    // one this build does not read); keys/idx: the radix passes' buffers
    uint2 *cells[2] = {};
    uint32_t *keys[2] = {}, *idx[2] = {};
    int kin = 0;                     // the next build's input (cells[kin])
    int kcells = 0;                  // the last build's input
    bool walk_keys = false;          // cells[kin] was written by the last walk
    uint32_t *status = nullptr;      // [4 * sort_blocks * 256] look-back words per pass
    uint32_t *digit_total = nullptr; // [2][kDigitCopies][4 * 256] per-pass digit totals, two halves
    uint32_t parity = 0;             // the half the next build counts into
    uint32_t *breaker = nullptr;     // [cap] sorted position where cell node #o starts
    uint32_t *block_cells = nullptr; // [prep_blocks] then scanned
    uint32_t *block_leaves = nullptr;
    int8_t *sdep = nullptr;          // [Na] shared depth with the previous sorted key
    uint32_t *cpos = nullptr;        // [Na] cell nodes starting before sorted position k
    float4 *prod = nullptr;          // [Na] (m*x, m*y, m, -) in sorted order
    uint32_t *leaf_first = nullptr;  // [Na+1]
    uint32_t *leaf_key = nullptr;    // [Na]
    float4 *leaf_sum = nullptr;      // [Na] (sx, sy, sm, n as bits)
    uint32_t *tile_range = nullptr;  // [2*kTiles] first/end leaf of each depth-5 tile
    TileCell *top = nullptr;         // [kTiles]
    float4 *node4 = nullptr;         // [cap]
    uint32_t *ncount = nullptr;      // [cap]
    float4 *nsum = nullptr;          // [cap] 15-bit leaves only: a cell node's
                                     // (sx, sy, sm, particle count as bits)
    // 15-bit leaves only: the occupied depth-10 cells ("granules") that feed
    // the dense tile reduction (first sorted position, 20-bit cell, subtree
    // size; their sums go to leaf_sum), and the per-block granule counts
    uint32_t *gran_first = nullptr;  // [Na+1]
    uint32_t *gran_key = nullptr;    // [Na]
    uint32_t *gran_y = nullptr;      // [Na]
    uint32_t *block_gran = nullptr;  // [prep_blocks]
    uint32_t *sorted_src = nullptr;  // [Na]
    BhCounters *ctr = nullptr;
    uint32_t sorted_slot = 0;        // which keys/idx buffer holds the sorted result
    // balanced traversal: each group's walk cost from the previous step and
    // the exclusive prefix of (cost + 1) that cuts the groups into segments
    uint32_t *gcost = nullptr;       // [Na/64], zero before the first walk
    uint32_t *gchunk = nullptr;      // [Na/64 + 1] chunks each group summed last step (issue priority)
    uint32_t *plan = nullptr;        // [kMaxHeavy] the heavy list
    uint8_t *heavy = nullptr;        // [Na/64 + 1] 1: on the heavy list this step
    uint32_t *order = nullptr;       // [Na/128] static walk's launch order of the workgroups
    bool order_valid = false;        // order holds a plan (a permutation of the units)
    uint32_t *gpre = nullptr;        // [Na/64 + 1]
    // each group's real-lane node visits and acceptances in the last walk,
    // written by the group's wave and summed on a dump (one same-address
    // atomic per wave cost ~5% of the walk)
    unsigned long long *gvis = nullptr;  // [Na/32 + 1] one per group (W = 32 or 64)
    unsigned long long *gacc = nullptr;  // [Na/32 + 1]
    // adaptive bounding box (opt-in): {cx, cy, h, r0} and per-block extremes
    float *box = nullptr;            // [4]
    float4 *box_part = nullptr;      // [256] (xmin, xmax, ymin, ymax)
};

// Events recorded around one step: e[0] start, e[1] between the two phases,
// e[2] end.  Barnes-Hut: (e0,e1) build, (e1,e2) traversal; direct sum:
// (e0,e1) force, (e1,e2) combine + integrate.
struct PhaseEvents {
    hipEvent_t e[3];
    bool bh;
};

// Symmetric direct sum (ds_kernels.hip): row groups of kDsSymRows 64-particle
// blocks; a workgroup tile is a row group times `width` = 4 x cols column
// blocks (width divides kDsSymRows, so the tile count of each group has a
// closed form).
constexpr uint32_t kDsSymRows = 16;
constexpr size_t kDsSymMinParticles = 16384;   // default range of the symmetric kernel
constexpr size_t kDsSymMaxParticles = 262144;  // (its partial sums grow as N^2 / 64)
constexpr size_t kDsSymMinWaves = 32768;       // auto: 2 columns per wave from this many waves

struct DsSymPlan {
    uint32_t B = 0, G = 0;      // 64-particle blocks, row groups
    uint32_t cols = 0;          // column blocks per wave (1 or 2)
    uint32_t width = 0;         // column blocks per workgroup tile (4 x cols)
    uint32_t ntiles = 0;        // workgroups of the force launch
    size_t ipart = 0, jpart = 0;  // float2 entries of the two partial arrays
};

// First tile of row group g: sum over h < g of ceil((B - R h) / width)
// = g * cb - step * g (g - 1) / 2 with cb = ceil(B / width), step = R / width.
__host__ __device__ inline size_t ds_sym_tile_base(uint32_t g, uint32_t cb, uint32_t step) {
    return (size_t)g * cb - (size_t)step * g * (g - (g ? 1u : 0u)) / 2;
}
// First column-partial row of group g: sum over h < g of (B - R h).
__host__ __device__ inline size_t ds_sym_jbase(uint32_t g, uint32_t B) {
    return (size_t)g * B - (size_t)kDsSymRows * g * (g - (g ? 1u : 0u)) / 2;
}

}  // namespace gsim

struct gsim_sim {
    gsim_params params{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;        // Barnes-Hut: the heavy list's walks, beside the main one
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    size_t n = 0, na = 0;
    // FP32 constants squared/multiplied on the host as the reference does
    // (DirectSimulation.cpp:179, BarnesHutSimulation.cpp:575-578).
    float eps2 = 0, theta2 = 0, dt = 0, kv = 0, kp = 0;
    uint32_t W = 64;
    uint32_t trav_waves_per_simd = 8;  // persistent traversal width
    uint32_t trav_mode = 2;            // bit 0 persistent, bit 1 LDS-staged runs,
                                       // bit 2 cost-balanced segments (staged);
                                       // default 2: one wave per group, staged
    uint32_t trav_segments = 512;      // balanced traversal: workgroups (2 per CU)
    uint32_t trav_split = 2;           // static: heavy list above this x mean chunk count;
                                       // balanced: split above this x mean cost (0: never)
    uint32_t bh_box = 0;               // adaptive bounding box instead of the fixed domain
    // Static walk's launch order: order_mode 0 index, 1 heaviest first (by last
    // step's visits), 2 auto: a trial of kOrderTrial steps alternating the
    // two, timed by events, keeps the faster once the host sees the trial's
    // last event complete (polled at each later step; never waited for)
    uint32_t order_mode = 2;
    uint32_t trav_order = 0;           // the order the next walk uses (0/1)
    bool order_fresh = false;          // this step is a trial step: plan afresh
    hipEvent_t order_ev[2 * gsim::kOrderTrial] = {};
    uint32_t order_rec = 0;            // trial steps recorded
    bool order_wait = false;           // the trial's events are outstanding
    uint64_t order_next = 1;           // first step of the next trial
    uint32_t xcd_order = 1;            // static traversals: each XCD takes runs of this
                                       // many consecutive workgroups (xcd_block; <= 1:
                                       // round-robin)
    uint32_t ds_rcp = 0;               // direct sum: 1/r2 from the transcendental unit
    // state
    float2 *pos[2] = {}, *vel[2] = {};
    float *mass = nullptr;
    int cur = 0;  // index of the input set
    // Barnes-Hut keeps the state in the last step's sorted order: perm[k][slot]
    // is the caller's index of the particle in slot of set k, and mass_next
    // receives the next step's masses (swapped with mass after each step)
    uint32_t *perm[2] = {};
    float *mass_next = nullptr;
    // direct sum
    float2 *partial = nullptr;
    size_t partial_cap = 0;            // float2 entries allocated
    uint32_t ds_R = 1, ds_S = 1, ds_S_used = 1;
    uint32_t ds_sym = 0, ds_sym_cols = 0;  // symmetric kernel; column blocks per tile (0 auto)
    gsim::DsSymPlan ds_plan{};
    // Barnes-Hut
    gsim::BhBuffers bh;
    // debug / timing
    float2 *accel = nullptr;
    uint32_t debug_flags = 0;
    bool timing = false;
    std::vector<gsim::PhaseEvents> pending;
    std::vector<hipEvent_t> event_pool;
    gsim_timing acc{};
    uint64_t steps = 0;
    bool has_tree = false;
};

namespace gsim {

// Direct sum (ds_kernels.hip)
void ds_pick_shape(size_t na, uint32_t *R, uint32_t *S);
bool ds_sym_default(size_t na);
DsSymPlan ds_sym_plan(size_t na, uint32_t cols);
hipError_t ds_launch_partial(gsim_sim *s, const float2 *pos, const float *mass);
hipError_t ds_launch_combine(gsim_sim *s, const float2 *pos, const float2 *vel,
                             float2 *pos_out, float2 *vel_out);

// Barnes-Hut (bh_kernels.hip)
size_t bh_capacity(size_t na, uint32_t bits);
hipError_t bh_alloc(BhBuffers *b, size_t na, uint32_t bits);
void bh_free(BhBuffers *b);
hipError_t bh_launch_build(gsim_sim *s, const float2 *pos, const float *mass,
                           const uint32_t *perm_in, float2 *pos_out, const float2 *vel,
                           float2 *vel_out, float *mass_out, uint32_t *perm_out);
hipError_t bh_launch_traverse(gsim_sim *s, const float2 *pos, const float2 *vel,
                              float2 *pos_out, float2 *vel_out, const uint32_t *perm_out);
hipError_t bh_launch_morton30(gsim_sim *s, const float2 *pos, uint32_t *out);
// Whether the handle's walk is a static launch of four-group workgroups,
// which GSIM_TUNE_BH_ORDER can reorder (W = 32, or W = 64 neither persistent
// nor balanced).
inline bool order_launch(const gsim_sim *s) { return s->W == 32 || (s->trav_mode & 5u) == 0; }

}  // namespace gsim
