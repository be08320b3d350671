// Batched multi-system stepping: many independent systems of uneven size in
// one device arena, all stepped by the same kernels (one launch per stage for
// every system), replayed from a hipGraph.  Sizes and membership live in
// device memory, so the captured graph stays valid while systems come and go
// within the arena's capacity.
#pragma once

#include <cstddef>
#include <map>
#include <vector>

#include "internal.hpp"

namespace gsim {

// One system's slot in the arena and its constants (host mirror + device copy).
struct SysDesc {
    uint32_t off, na, n, algo;       // arena offset (multiple of 64), padded/real count
    float eps2, theta2, kv, kp;      // softening^2, theta^2, velocity and position factors
    float dt, pad0;
    uint32_t bh_rank, sorted_base;   // Barnes-Hut: rank in the global sort, first sorted position
    uint32_t ds_part_off, ds_jlen;   // direct sum: partial-sum offset (float2) and slice length
    uint32_t ds_slices;
    uint32_t w;                      // Barnes-Hut vote width: 64 or 32
};

// Per Barnes-Hut rank, written by the emit kernel every step.
struct RankStats {
    uint32_t root;      // first node slot of the system's tree
    uint32_t end_slot;  // one past its last node slot
    uint32_t n_in;      // particles inside the domain
    uint32_t pad;
};

// The leaf kernel lists the (system, depth-5 tile) pairs holding leaves in
// kLiveLists lists picked by a hash of the pair, each with its own counter:
// one counter for all of them serialises thousands of same-address atomics.
constexpr uint32_t kLiveLists = 64;

// The force launch's work list is dealt to kForceQueues interleaved queues
// (item i to queue i mod kForceQueues); the waves of queue q take its items
// one at a time from a counter, so a wave that drew light items takes more.
constexpr uint32_t kForceQueues = 64;

struct BatchCounts {
    uint32_t n_bh;            // Barnes-Hut systems (ranks)
    uint32_t n_bh_particles;  // sum of their padded counts = sorted positions
    uint32_t n_ds;            // direct-sum systems
    uint32_t n_ds_particles;  // sum of their padded counts
    uint32_t n_items;         // work items of the force launch
    uint32_t n_cells;         // cell nodes of all trees (emit)
    uint32_t n_leaves;        // occupied leaves of all trees (emit)
    uint32_t n_sys;           // all systems (direct sum and Barnes-Hut)
    uint32_t pad;
    uint32_t sort_stall;      // radix look-back waits that hit their bound (never expected)
    uint32_t sort_tile[4];    // tile counter of each radix pass
    uint32_t n_live[kLiveLists];  // (system, depth-5 tile) pairs holding leaves, per list
    uint32_t qnext[kForceQueues];  // force launch: items taken from each queue (zeroed each step)
    unsigned long long accepts;  // Barnes-Hut real-lane acceptances, all steps
    unsigned long long visits;   // Barnes-Hut real-lane node visits, all steps
};
// the two totals are copied together
static_assert(offsetof(BatchCounts, visits) == offsetof(BatchCounts, accepts) + 8,
              "accepts and visits must be adjacent");

constexpr uint32_t kUnlimitedSteps = 0xffffffffu;

// Radix tiles of the batched sort: sort_rounds keys per thread (8, 12 or 16:
// tiles of 2,048, 3,072 or 4,096 keys; 12 by default, the autotuner picks).
// The look-back status words are sized for the smallest tile.
constexpr int kBatchSortRoundsDefault = 12;
constexpr int kBatchSortTileMin = 8 * kSortThreads;

// Launch parameters of a batched step that never change results (the
// autotuner's knobs): persistent force-launch width, radix tile size, leaf
// runs a wave queues before summing them (its LDS run stack), and the walk's
// issue-priority step and carry divisor (0: off).
struct BatchTune {
    // 6 of the force kernel's 7 resident waves per SIMD: on the 256-system
    // mix 2.36 ms per step against 2.91 with 7 and 2.78 with 8 (a second
    // round of waves)
    uint32_t waves_per_simd = 6;
    uint32_t sort_rounds = kBatchSortRoundsDefault;
    uint32_t run_queue = 64;   // <= kRunQueue
    uint32_t prio_step = 32;   // summed chunks per priority level
    uint32_t prio_carry = 16;  // last step's chunks / this = the starting level
};

constexpr uint32_t kBatchRankShift = 21;                   // key = rank << 21 | cell (bit 20 = killed)
constexpr uint32_t kBatchMaxSystems = 1u << (32 - kBatchRankShift);  // 2048
constexpr uint32_t kBatchSentinel = 0xffffffffu;
constexpr uint32_t kMaxForceWaves = 1024u * 8u;  // persistent force launch: 8 waves per SIMD
// Largest arena: node slots (3 per particle) and partial sums stay within
// 32-bit indices.
constexpr size_t kBatchMaxParticles = (0x7fffffffull / 8) / 64 * 64;

struct BatchDevice {
    // arena
    float2 *pos[2] = {}, *vel[2] = {};
    float *mass = nullptr;
    float2 *partial = nullptr;
    // membership
    SysDesc *desc = nullptr;
    uint32_t *bh_sys = nullptr, *bh_prefix = nullptr;  // [max_systems (+1)]
    uint32_t *ds_sys = nullptr, *ds_prefix = nullptr;
    uint2 *items = nullptr;
    uint32_t *remaining = nullptr;  // steps each system may still take (kUnlimitedSteps = no limit)
    BatchCounts *counts = nullptr;
    RankStats *stats = nullptr;
    // Barnes-Hut scratch over all sorted positions
    uint32_t *keys[2] = {}, *idx[2] = {};
    uint32_t *status = nullptr;       // [2][tiles_cap * 256] radix look-back words, by pass parity
    uint32_t *digit_total = nullptr;  // [4 * 256] per-pass digit totals
    uint32_t *ichunk = nullptr;       // [item_capacity] chunks each work item summed last step (issue priority)
    uint32_t *block_cells = nullptr, *block_leaves = nullptr;
    int8_t *sdep = nullptr;
    uint32_t *cpos = nullptr, *sorted_src = nullptr;
    float4 *prod = nullptr, *leaf_sum = nullptr;
    uint32_t *leaf_first = nullptr, *leaf_key = nullptr;
    uint32_t *tile_range = nullptr;
    uint32_t *live_tiles = nullptr;  // [kLiveLists][live_per_list] tiles with leaves, any order
    TileCell *top = nullptr;
    // nodes of all trees
    float4 *node4 = nullptr;
    uint32_t *ncount = nullptr, *breaker = nullptr;
    // each force wave's Barnes-Hut acceptances, then its node visits, in the
    // last step, summed into counts->accepts / visits by the step's last
    // kernel (a same-address atomic per wave costs more than the sum)
    unsigned long long *wave_acc = nullptr;  // [2 * kMaxForceWaves]
};

struct BatchSystem {
    gsim_params params;
    uint32_t off = 0, na = 0, n = 0;
    size_t node_reserve = 0;
    uint64_t steps = 0;          // steps taken (exact after a synchronisation)
    uint64_t added_at = 0;       // batch step count when the system joined
    uint64_t target = ~0ull;     // total steps to take (~0 = unlimited)
};

}  // namespace gsim

struct gsim_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t capacity = 0;        // particles (padded) in the arena
    size_t node_capacity = 0;   // node slots
    size_t partial_capacity = 0;
    size_t item_capacity = 0;
    uint32_t max_systems = 0;
    size_t tiles_cap = 0, blocks_cap = 0;
    gsim::BatchDevice d;
    std::map<uint32_t, gsim::BatchSystem> systems;  // id -> system
    std::map<uint32_t, uint32_t> free_ranges;       // arena offset -> length (units of particles)
    uint32_t next_id = 1;
    size_t nodes_reserved = 0, partial_used = 0;
    bool dirty = true;
    int cur = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    gsim::BatchTune tune;
    uint64_t steps = 0;
    // host mirror of the membership arrays
    std::vector<gsim::SysDesc> h_desc;
    gsim::BatchCounts h_counts{};
};

namespace gsim {

hipError_t batch_launch_step(gsim_batch *b, int in, int out);

}  // namespace gsim
