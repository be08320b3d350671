// One-sweep stable LSD radix sort pieces shared by the single-system build
// (bh_kernels.hip) and the batched build (batch_kernels.hip).
//
// A pass sorts tiles of kSortTile (key, value) pairs on one 8-bit digit.  A
// workgroup takes the next tile from a counter, so every lower tile belongs to
// a workgroup that is already running.  Within the tile, key r * 256 + t
// (round r, thread t) is ranked among the equal digits of its wave by
// ballots; the per-(round, wave, digit) counts go to LDS and thread d scans
// digit d's counts in (round, wave) order, which is input order, so the sort
// is stable and the tile's digit counts fall out of the same scan.  With
// kStage the tile is staged in LDS in digit order and written out from there,
// so consecutive lanes store consecutive positions of a digit's run (pays for
// many tiles of spatially clustered keys; without it each key is stored
// straight to its position).  The tile publishes its counts in one status word per (tile, digit) -- 0 (not
// yet counted), kStA | the tile's count, kStP | the inclusive prefix over
// tiles 0..tile -- and sums its predecessors' words back to the first kStP
// (decoupled look-back).  A tile publishes kStA before it reads any
// predecessor, so every wait ends; waits are bounded all the same
// (kSpinLimit re-reads, counted in *stall).
#pragma once
#include "bh_device.hpp"
#include "internal.hpp"

namespace gsim {

constexpr uint32_t kStA = 1u << 30, kStP = 2u << 30, kStVal = kStA - 1u;
constexpr int kLookWindow = 4;             // predecessors read at once per digit
constexpr uint32_t kSpinLimit = 1u << 20;  // bound on re-reads of one status word

// kR: rounds of kSortThreads keys per tile (kSortRounds unless the launch
// picks another tile size).
template <bool kStage, int kR = kSortRounds>
struct SortShared {
    static constexpr int kRows = kR * (kSortThreads / 64);
    static constexpr int kTile = kR * kSortThreads;
    uint32_t slot[kRows * 256];  // per-(round, wave) digit counts, then offsets
    uint32_t stage_key[kStage ? kTile : 1];  // the tile in digit order
    uint32_t stage_val[kStage ? kTile : 1];
    uint32_t first[256];             // tile position of the first key of each digit
    uint32_t adj[256];               // output position - tile position, per digit
    uint32_t wsum[4];
    uint32_t tile;
};

__device__ __forceinline__ uint32_t status_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void status_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lanes of this wave whose digit equals this lane's (valid lanes only).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
        const bool set = (d >> bit) & 1u;
        const uint64_t bb = __ballot(set);
        m &= set ? bb : ~bb;
    }
    return m;
}

// Adds the digits of kPasses passes of `key` to the LDS histograms h, one
// atomic per lane; with kUniform, one atomic for a digit that every valid
// lane of the wave shares (the upper digits of keys grouped by system).
template <int kPasses, bool kUniform>
__device__ __forceinline__ void hist_add(uint32_t (*h)[256], uint32_t key, bool valid) {
    if (!kUniform) {
#pragma unroll
        for (int p = 0; p < kPasses; ++p)
            if (valid) atomicAdd(&h[p][(key >> (8 * p)) & 255u], 1u);
        return;
    }
    const uint64_t live = __ballot(valid);
    if (!live) return;
    const int first = __builtin_ctzll(live);
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int p = 0; p < kPasses; ++p) {
        const uint32_t d = (key >> (8 * p)) & 255u;
        const uint32_t d0 = __shfl(d, first, 64);
        if (__ballot(valid && d != d0) == 0) {
            if (lane == (uint32_t)first) atomicAdd(&h[p][d], (uint32_t)__popcll(live));
        } else if (valid) {
            atomicAdd(&h[p][d], 1u);
        }
    }
}

// Adds predecessors j, j-1, ..., j-kW+1 of this digit's chain to *excl until
// one holds an inclusive prefix (sets *done); predecessors below tile 0 count
// as an empty prefix.  The window is read at once; while a word it needs is
// not yet published, the whole window is read again (a word only goes from 0
// to a count to a prefix, and any published form of it is usable), so
// several late words cost one wait, not one each.
template <int kW>
__device__ __forceinline__ void look_back(const uint32_t *st, int j, uint32_t *excl, bool *done,
                                          uint32_t *stall) {
    uint32_t v[kW];
#pragma unroll 1
    for (uint32_t spin = 0;; ++spin) {
#pragma unroll
        for (int q = 0; q < kW; ++q)
            v[q] = j - q >= 0 ? status_load(st + (size_t)(j - q) * 256) : kStP;
        uint32_t add = 0;
        bool found = false, blocked = false;
#pragma unroll
        for (int q = 0; q < kW; ++q) {
            if (!found && !blocked) {
                const uint32_t x = v[q];
                if ((x >> 30) == 0) {
                    blocked = true;
                } else {
                    add += x & kStVal;
                    found = (x & kStP) != 0;
                }
            }
        }
        if (!blocked) {
            *excl += add;
            *done = found;
            return;
        }
        if (spin >= kSpinLimit) {  // never expected: counted, and the prefix taken as is
            atomicAdd(stall, 1u);
            *excl += add;
            *done = true;
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Exclusive prefix of this pass's digit totals for digit threadIdx.x
// (synchronises the workgroup).
template <bool kStage, int kR>
__device__ __forceinline__ uint32_t digit_base(const uint32_t *__restrict__ digit_total,
                                               SortShared<kStage, kR> &sh) {
    const uint32_t total = digit_total[threadIdx.x];
    uint32_t all = 0;
    return block_scan256(total, &all, sh.wsum) - total;
}

// Clears this thread's column of sh.slot (sort_tile's precondition, ordered
// by a barrier of the caller).
template <bool kStage, int kR>
__device__ __forceinline__ void sort_tile_clear(SortShared<kStage, kR> &sh) {
#pragma unroll
    for (int k = 0; k < SortShared<kStage, kR>::kRows; ++k) sh.slot[k * 256 + threadIdx.x] = 0;
}

// Two-level look-back (kLookGroup tiles to a group): the exclusive prefix of
// this thread's digit before `tile` (c: the tile's own count) is the sum of
// the counts of the tiles before it in its group, read at once, plus the
// exclusive prefix of its group, found by decoupled look-back over the
// groups' words.  Tile words (st[j * 256]) only ever hold kStA | count; the
// last tile of each group publishes the group's sum and then its inclusive
// prefix (gst[g * 256]); the groups are looked back over kLookGroup at a time.
// Every wait is for a lower tile, so it ends; waits are bounded as in
// look_back.  The chain of look-backs is the number of groups long instead
// of the number of tiles.  (kLookGroup: internal.hpp.)
__device__ __forceinline__ uint32_t grouped_prefix(uint32_t *st, uint32_t *gst, uint32_t tile,
                                                   uint32_t tiles, uint32_t c,
                                                   uint32_t *__restrict__ stall) {
    const uint32_t g = tile / kLookGroup, first = g * kLookGroup;
    status_store(st + (size_t)tile * 256, kStA | c);
    // the counts of the tiles before this one in its group, read together
    // (again together while any is unpublished)
    uint32_t inner = 0;
#pragma unroll 1
    for (uint32_t spin = 0;; ++spin) {
        uint32_t v[kLookGroup - 1];
#pragma unroll
        for (uint32_t q = 0; q + 1 < kLookGroup; ++q)
            v[q] = first + q < tile ? status_load(st + (size_t)(first + q) * 256) : kStA;
        uint32_t sum = 0;
        bool ready = true;
#pragma unroll
        for (uint32_t q = 0; q + 1 < kLookGroup; ++q) {
            ready = ready && (v[q] >> 30) != 0;
            sum += v[q] & kStVal;
        }
        if (ready || spin >= kSpinLimit) {
            if (!ready) atomicAdd(stall, 1u);  // never expected
            inner = sum;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    const bool last = tile + 1 == first + kLookGroup || tile + 1 == tiles;
    if (last) status_store(gst + (size_t)g * 256, (g ? kStA : kStP) | (inner + c));
    uint32_t gex = 0;
    if (g) {
        // groups' sums are published at about the same time, and their
        // prefixes one after another: read a wide window of them at once
        int j = (int)g - 1;
        bool done = false;
        while (!done) {
            look_back<kLookGroup>(gst, j, &gex, &done, stall);
            j -= (int)kLookGroup;
        }
        if (last) status_store(gst + (size_t)g * 256, kStP | (gex + inner + c));
    }
    return gex + inner;
}

// Sorts tile `tile` of keys_in/idx_in on digit (key >> shift) & 255 into
// keys_out/idx_out.  gbase: digit_base() of this thread's digit; status: this
// pass's [tiles][256] words; sh.slot cleared by sort_tile_clear() before a
// barrier.  Ends with sh in use: a barrier must precede its next use.
// kGrouped: grouped_prefix instead of the tile-by-tile look-back (tiles: the
// pass's tile count; gstatus: its [groups][256] words, zeroed); with totals
// (this pass's digit totals, split over `copies` copies `copy_stride` words
// apart) gbase is not given but computed here, the totals loaded ahead of the
// keys and scanned once the tile is ranked.  packed_in: the input as {key,
// index} pairs instead of keys_in/idx_in.
template <bool kStage, int kR, bool kGrouped = false>
__device__ __forceinline__ void sort_tile(const uint32_t *__restrict__ keys_in,
                                          const uint32_t *__restrict__ idx_in,
                                          uint32_t *__restrict__ keys_out,
                                          uint32_t *__restrict__ idx_out, uint32_t n,
                                          uint32_t shift, uint32_t tile, uint32_t gbase,
                                          uint32_t *__restrict__ status,
                                          uint32_t *__restrict__ stall,
                                          SortShared<kStage, kR> &sh, uint32_t tiles = 0,
                                          uint32_t *__restrict__ gstatus = nullptr,
                                          const uint32_t *__restrict__ totals = nullptr,
                                          uint32_t copies = 0, uint32_t copy_stride = 0,
                                          const uint2 *__restrict__ packed_in = nullptr) {
    constexpr int kRows = SortShared<kStage, kR>::kRows;
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const uint32_t base = tile * SortShared<kStage, kR>::kTile;
    uint32_t total = 0;
    if (kGrouped && totals)
        for (uint32_t c = 0; c < copies; ++c) total += totals[c * copy_stride + t];
    uint32_t key[kR], val[kR], rank[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        const uint32_t i = base + r * kSortThreads + t;
        if (packed_in) {
            const uint2 kv = i < n ? packed_in[i] : make_uint2(0u, 0u);
            key[r] = kv.x;
            val[r] = kv.y;
        } else {
            key[r] = i < n ? keys_in[i] : 0u;
            val[r] = i < n ? idx_in[i] : 0u;
        }
    }
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        const bool valid = base + r * kSortThreads + t < n;
        const uint32_t d = (key[r] >> shift) & 255u;
        const uint64_t m = digit_peers(d, valid);
        rank[r] = (uint32_t)__popcll(m & below);
        if (valid && rank[r] == 0) sh.slot[(r * (kSortThreads / 64) + wave) * 256 + d] = (uint32_t)__popcll(m);
    }
    lds_barrier();
    uint32_t c = 0;  // keys of digit t before row k, then the tile's count
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const uint32_t x = sh.slot[k * 256 + t];
        sh.slot[k * 256 + t] = c;
        c += x;
    }
    if (kGrouped && totals) {
        uint32_t all = 0;
        gbase = block_scan256(total, &all, sh.wsum) - total;  // synchronises
    }
    uint32_t tile_n = 0, first = 0;
    if (kStage) {
        first = block_scan256(c, &tile_n, sh.wsum) - c;  // synchronises
        sh.first[t] = first;
    }
    uint32_t *const st = status + t;  // this digit's word of tile j: st[j * 256]
    uint32_t excl = 0;
    if (kGrouped) {
        excl = grouped_prefix(st, gstatus + t, tile, tiles, c, stall);
    } else {
        status_store(st + (size_t)tile * 256, (tile ? kStA : kStP) | c);
        if (tile) {
            int j = (int)tile - 1;
            bool done = false;
            while (!done) {
                look_back<kLookWindow>(st, j, &excl, &done, stall);
                j -= kLookWindow;
            }
            status_store(st + (size_t)tile * 256, kStP | (excl + c));
        }
    }
    sh.adj[t] = gbase + excl - first;
    lds_barrier();
    if (!kStage) {
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            if (base + r * kSortThreads + t < n) {
                const uint32_t d = (key[r] >> shift) & 255u;
                const uint32_t off =
                    sh.adj[d] + sh.slot[(r * (kSortThreads / 64) + wave) * 256 + d] + rank[r];
                if (off < n) {
                    keys_out[off] = key[r];
                    idx_out[off] = val[r];
                }
            }
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        if (base + r * kSortThreads + t < n) {
            const uint32_t d = (key[r] >> shift) & 255u;
            const uint32_t at =
                sh.first[d] + sh.slot[(r * (kSortThreads / 64) + wave) * 256 + d] + rank[r];
            sh.stage_key[at] = key[r];
            sh.stage_val[at] = val[r];
        }
    }
    lds_barrier();
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        const uint32_t at = r * kSortThreads + t;
        if (at < tile_n) {
            const uint32_t k = sh.stage_key[at];
            const uint32_t off = sh.adj[(k >> shift) & 255u] + at;
            if (off < n) {
                keys_out[off] = k;
                idx_out[off] = sh.stage_val[at];
            }
        }
    }
}

}  // namespace gsim
