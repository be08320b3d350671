// Batched stepping kernels: every kernel covers all systems of a batch and
// reads its sizes from device memory (BatchCounts), so a step captured once
// in a hipGraph serves any membership that fits the arena.
//
// Barnes-Hut systems share one stable radix sort on (rank << 21 | cell):
// sorted positions group by system, and because node slots are
// position + scanned cell-node count, each system's pre-order tree lands
// contiguously in one global node array ([root, end_slot) per rank, recorded
// by the emit kernel).  The per-system semantics are those of the
// single-system kernels in bh_kernels.hip (same cells, same canonical order,
// same reduction association, same vote), so a system's result in a batch is
// bit-identical to its result alone.  Direct-sum systems use the same slice
// split as the single-system path for their size, so they match bit for bit
// as well.
#include "batch.hpp"
#include "bh_device.hpp"
#include "radix_device.hpp"

#include <algorithm>

namespace gsim {
namespace {

constexpr uint32_t kCellMask = (1u << 20) - 1u;
constexpr uint32_t kKilledBit = 1u << 20;

// Largest r in [0, n) with prefix[r] <= f (prefix ascending, prefix[0] = 0).
__device__ __forceinline__ uint32_t rank_of(const uint32_t *__restrict__ prefix, uint32_t n,
                                            uint32_t f) {
    uint32_t lo = 0, hi = n;  // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= f) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---- keys ------------------------------------------------------------------

__global__ __launch_bounds__(256) void bb_keys_kernel(
    const SysDesc *__restrict__ desc, const uint32_t *__restrict__ bh_sys,
    const uint32_t *__restrict__ bh_prefix, const uint32_t *__restrict__ remaining,
    BatchCounts *__restrict__ counts,
    float2 *__restrict__ pos, float *__restrict__ mass, float2 *__restrict__ pos_out,
    const float2 *__restrict__ vel, float2 *__restrict__ vel_out, uint32_t *__restrict__ keys,
    uint32_t *__restrict__ idx, uint32_t *__restrict__ tile_range,
    RankStats *__restrict__ stats, uint32_t *__restrict__ digit_total,
    uint32_t *__restrict__ status, uint32_t tile_keys) {
    const uint32_t nb = counts->n_bh, n = counts->n_bh_particles;
    const uint32_t tid = blockIdx.x * 256 + threadIdx.x, stride = gridDim.x * 256;
    for (uint32_t t = tid; t < nb * 2 * kTiles; t += stride) tile_range[t] = 0;
    for (uint32_t t = tid; t < nb; t += stride) stats[t] = RankStats{0u, 0u, 0u, 0u};
    for (uint32_t t = tid; t < 4 * 256; t += stride) digit_total[t] = 0;
    for (uint32_t t = tid; t < kLiveLists; t += stride) counts->n_live[t] = 0;
    // the first radix pass's look-back words (each pass clears the next one's)
    const uint32_t status_words = (n + tile_keys - 1) / tile_keys * 256;
    for (uint32_t t = tid; t < status_words; t += stride) status[t] = 0;
    if (tid == 0) {
        counts->n_cells = 0;
        counts->n_leaves = 0;
        counts->sort_stall = 0;
    }
    if (tid < 4) counts->sort_tile[tid] = 0;
    for (uint32_t f = tid; f < n; f += stride) {
        const uint32_t r = rank_of(bh_prefix, nb, f);
        const uint32_t di = bh_sys[r];
        const SysDesc &d = desc[di];
        const uint32_t i = d.off + (f - bh_prefix[r]);
        const float2 p = pos[i];
        if (remaining[di] == 0u) {
            // Budget spent: the system keeps its state exactly (no kill of
            // particles that left the domain on its last step) and builds an
            // empty tree, so the force launch skips it.
            pos_out[i] = p;
            vel_out[i] = vel[i];
            keys[f] = (r << kBatchRankShift) | kKilledBit;
            idx[f] = i;
            continue;
        }
        uint32_t c = cell20(p.x, p.y);
        if (c == 0xffffffffu) {  // InitShader.comp:110-117
            const float k = -kSimSize * 2.0f;
            pos[i] = make_float2(k, k);
            pos_out[i] = make_float2(k, k);
            mass[i] = 0.0f;
            vel_out[i] = vel[i];
            c = kKilledBit;
        }
        keys[f] = (r << kBatchRankShift) | c;
        idx[f] = i;
    }
}

// ---- stable LSD radix sort over all sorted positions (4 x 8 bits) -----------
// One-sweep passes (radix_device.hpp) with persistent workgroups: the live
// tile count comes from device memory.

// digit_total[pass * 256 + d] += keys of digit d in pass `pass`; the per-wave
// aggregation matters here: a wave's keys usually share one system, so the
// upper digits repeat across the wave.
template <int kR>
__global__ __launch_bounds__(kSortThreads) void bb_hist_kernel(
    const uint32_t *__restrict__ keys, const BatchCounts *__restrict__ counts,
    uint32_t *__restrict__ digit_total) {
    constexpr uint32_t kTile = kR * kSortThreads;
    __shared__ uint32_t h[4][256];
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int p = 0; p < 4; ++p) h[p][t] = 0;
    __syncthreads();
    const uint32_t n = counts->n_bh_particles;
    for (uint32_t base = blockIdx.x * kTile; base < n; base += gridDim.x * kTile) {
        uint32_t k[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            const uint32_t i = base + r * kSortThreads + t;
            k[r] = i < n ? keys[i] : 0u;
        }
#pragma unroll
        for (int r = 0; r < kR; ++r)
            hist_add<4, true>(h, k[r], base + r * kSortThreads + t < n);
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t c = h[p][t];
        if (c) atomicAdd(&digit_total[p * 256 + t], c);
    }
}

// Pass `pass` (digit (key >> 8 pass) & 255): workgroups take tiles from
// counts->sort_tile[pass] until the live tiles run out.  status: this pass's
// [tiles][256] look-back words, zeroed; status_next: the next pass's, zeroed
// here tile by tile.
template <int kR>
__global__ __launch_bounds__(kSortThreads) void bb_onesweep_kernel(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ idx_in,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ idx_out,
    BatchCounts *__restrict__ counts, uint32_t pass, const uint32_t *__restrict__ digit_total,
    uint32_t *__restrict__ status, uint32_t *__restrict__ status_next) {
    __shared__ SortShared<true, kR> sh;
    constexpr uint32_t kTile = kR * kSortThreads;
    const uint32_t n = counts->n_bh_particles;
    const uint32_t ntiles = (n + kTile - 1) / kTile;
    const uint32_t gbase = digit_base(digit_total + 256 * pass, sh);  // synchronises
    for (;;) {
        if (threadIdx.x == 0) sh.tile = atomicAdd(&counts->sort_tile[pass], 1u);
        sort_tile_clear(sh);
        __syncthreads();
        const uint32_t tile = sh.tile;
        if (tile >= ntiles) break;
        status_next[(size_t)tile * 256 + threadIdx.x] = 0;
        sort_tile<true>(keys_in, idx_in, keys_out, idx_out, n, 8u * pass, tile, gbase, status,
                  &counts->sort_stall, sh);
        __syncthreads();
    }
}

// ---- slot scan and emission ----------------------------------------------------

struct BPos {
    int sk, sk1;
    uint32_t ns, lf, rank, key;
    bool inside;
};

// s(k) and s(k+1) never cross a system boundary: the first position of a
// system behaves like position 0 of a single system, its last in-domain
// position like the last one.
__device__ __forceinline__ BPos bpos_info(const uint32_t *__restrict__ keys, uint32_t n,
                                          uint32_t k) {
    BPos q{-1, -1, 0u, 0u, 0u, 0u, false};
    if (k >= n) return q;
    q.key = keys[k];
    q.rank = q.key >> kBatchRankShift;
    if (q.key & kKilledBit) return q;
    q.inside = true;
    if (k > 0) {
        const uint32_t pk = keys[k - 1];
        if ((pk >> kBatchRankShift) == q.rank) q.sk = shared_depth(pk & kCellMask, q.key & kCellMask);
    }
    if (k + 1 < n) {
        const uint32_t nk = keys[k + 1];
        if ((nk >> kBatchRankShift) == q.rank && !(nk & kKilledBit))
            q.sk1 = shared_depth(q.key & kCellMask, nk & kCellMask);
    }
    q.ns = q.sk1 > q.sk ? (uint32_t)(q.sk1 - q.sk) : 0u;
    q.lf = q.sk < 10 ? 1u : 0u;
    return q;
}

__global__ __launch_bounds__(256) void bb_prep_kernel(const uint32_t *__restrict__ keys,
                                                      const BatchCounts *__restrict__ counts,
                                                      uint32_t *__restrict__ block_cells,
                                                      uint32_t *__restrict__ block_leaves) {
    __shared__ uint32_t wa[4], wb[4];
    const uint32_t n = counts->n_bh_particles;
    const uint32_t nblocks = (n + 255) / 256;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        const BPos q = bpos_info(keys, n, blk * 256 + threadIdx.x);
        uint32_t a = q.ns, b = q.lf;
        wave_scan2(a, b);
        if (lane == 63) {
            wa[wave] = a;
            wb[wave] = b;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            block_cells[blk] = wa[0] + wa[1] + wa[2] + wa[3];
            block_leaves[blk] = wb[0] + wb[1] + wb[2] + wb[3];
        }
        __syncthreads();
    }
}

// Exclusive scan of the per-block sums (blockIdx.x 0: cells, 1: leaves).
__global__ __launch_bounds__(1024) void bb_block_scan_kernel(uint32_t *__restrict__ a,
                                                             uint32_t *__restrict__ b,
                                                             const BatchCounts *__restrict__ counts) {
    __shared__ uint32_t wsum[16];
    scan_inplace_1024(blockIdx.x == 0 ? a : b, (counts->n_bh_particles + 255) / 256, wsum);
}

__global__ __launch_bounds__(256) void bb_emit_kernel(
    const uint32_t *__restrict__ keys, const uint32_t *__restrict__ sidx,
    const float2 *__restrict__ pos, const float *__restrict__ mass,
    BatchCounts *__restrict__ counts, const uint32_t *__restrict__ bh_prefix,
    const uint32_t *__restrict__ block_cells, const uint32_t *__restrict__ block_leaves,
    int8_t *__restrict__ sdep, uint32_t *__restrict__ cpos, float4 *__restrict__ prod,
    uint32_t *__restrict__ leaf_first, uint32_t *__restrict__ leaf_key,
    float4 *__restrict__ node4, uint32_t *__restrict__ ncount, uint32_t *__restrict__ sorted_src,
    uint32_t *__restrict__ breaker, RankStats *__restrict__ stats) {
    __shared__ uint32_t wa[4], wb[4];
    const uint32_t n = counts->n_bh_particles;
    const uint32_t nblocks = (n + 255) / 256;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        const uint32_t k = blk * 256 + threadIdx.x;
        // the particle's gather does not depend on the scan below: start it first
        const uint32_t i = k < n ? sidx[k] : 0u;
        const float2 p = k < n ? pos[i] : make_float2(0.0f, 0.0f);
        const float m = k < n ? mass[i] : 0.0f;
        const BPos q = bpos_info(keys, n, k);
        uint32_t a = q.ns, b = q.lf;
        wave_scan2(a, b);
        if (lane == 63) {
            wa[wave] = a;
            wb[wave] = b;
        }
        __syncthreads();
        uint32_t C = a - q.ns + block_cells[blk];
        uint32_t L = b - q.lf + block_leaves[blk];
        for (uint32_t w = 0; w < wave; ++w) {
            C += wa[w];
            L += wb[w];
        }
        __syncthreads();
        if (k >= n) continue;
        if (k + 1 == n) {  // totals over all systems, and the end of the last leaf
            counts->n_cells = C + q.ns;
            counts->n_leaves = L + q.lf;
            leaf_first[L + q.lf] = q.inside ? k + 1 : n;
        }
        if (!q.inside) {
            sorted_src[k] = kBatchSentinel;
            continue;
        }
        const uint32_t slot = k + C + q.ns;
        node4[slot] = make_float4(p.x, p.y, m, 0.0f);
        ncount[slot] = 1u;
        sorted_src[k] = i;
        sdep[k] = (int8_t)q.sk;
        cpos[k] = C;
        prod[k] = make_float4(p.x * m, p.y * m, m, 0.0f);
        if (q.lf) {
            leaf_first[L] = k;
            leaf_key[L] = q.key;
        }
        for (uint32_t j = 0; j < q.ns; ++j) breaker[C + j] = k;
        if (q.sk < 0) stats[q.rank].root = k + C;
        if (q.sk1 < 0) {
            stats[q.rank].end_slot = k + 1 + C + q.ns;
            stats[q.rank].n_in = k + 1 - bh_prefix[q.rank];
        }
    }
}

// ---- leaf sums and particle runs ------------------------------------------------

// The live list of (system r, depth-5 tile c), T = r kTiles + c: (c + 23 r)
// mod 64.  Each system gives each list exactly kTiles / 64 of its tiles, so a
// list never holds more than max_systems kTiles / 64 (its capacity), and the
// odd stride spreads the systems' few central tiles over the lists.
__device__ __forceinline__ uint32_t live_list(uint32_t T) {
    return ((T & (kTiles - 1u)) + 23u * (T / kTiles)) & (kLiveLists - 1u);
}

__device__ __forceinline__ uint32_t tile_of(uint32_t key) {
    return (key >> kBatchRankShift) * kTiles + ((key >> 10) & (kTiles - 1));
}

__global__ __launch_bounds__(256) void bb_leaf_kernel(
    const uint32_t *__restrict__ leaf_first, const uint32_t *__restrict__ leaf_key,
    const float4 *__restrict__ prod, const uint32_t *__restrict__ keys,
    const uint32_t *__restrict__ cpos, const uint32_t *__restrict__ breaker,
    BatchCounts *counts, const uint32_t *__restrict__ bh_prefix,
    const RankStats *__restrict__ stats, float4 *__restrict__ leaf_sum,
    uint32_t *__restrict__ tile_range, uint32_t *__restrict__ live_tiles, uint32_t live_per_list,
    float4 *__restrict__ node4) {
    __shared__ float4 stage[4][64];
    const uint32_t n = counts->n_bh_particles, nl = counts->n_leaves, ncells = counts->n_cells;
    const uint32_t waves = gridDim.x * 4;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + w);
    // small leaves: each wave takes 64 consecutive leaves, one per lane
    for (uint32_t base = wave_id * 64; base < nl; base += waves * 64) {
        const uint32_t L = base + lane;
        const bool valid = L < nl;
        uint32_t f = 0, e = 0, key = 0;
        if (valid) {
            key = leaf_key[L];
            const uint32_t r = key >> kBatchRankShift;
            const uint32_t sys_end = bh_prefix[r] + stats[r].n_in;
            e = min(leaf_first[L + 1], sys_end);
            f = min(leaf_first[L], e);
        }
        const bool small = valid && e - f <= kSmallLeaf;
        if (small) {
            const float4 r = lane_leaf_sum(prod, f, e);
            leaf_sum[L] = make_float4(r.x, r.y, r.z, __uint_as_float(e - f));
        }
        if (valid) {
            const uint32_t T = tile_of(key);
            if (L == 0 || tile_of(leaf_key[L - 1]) != T) {
                tile_range[2 * T] = L;
                const uint32_t q = live_list(T);
                live_tiles[q * live_per_list + atomicAdd(&counts->n_live[q], 1u)] = T;
            }
            if (L + 1 == nl || tile_of(leaf_key[L + 1]) != T) tile_range[2 * T + 1] = L + 1;
        }
    }
    // larger leaves: leaf L goes to wave L mod waves (neighbouring dense
    // leaves to different waves), streamed by the whole wave
    for (uint32_t i0 = 0; wave_id + i0 * waves < nl; i0 += 64) {
        const uint32_t L = wave_id + (i0 + lane) * waves;
        uint32_t f = 0, e = 0;
        if (L < nl) {
            const uint32_t r = leaf_key[L] >> kBatchRankShift;
            e = min(leaf_first[L + 1], bh_prefix[r] + stats[r].n_in);
            f = min(leaf_first[L], e);
        }
        unsigned long long big = __ballot(L < nl && e - f > kSmallLeaf);
        while (big) {
            const uint32_t j = (uint32_t)__builtin_ctzll(big);
            big &= big - 1;
            const uint32_t bf = __builtin_amdgcn_readfirstlane(__shfl(f, j, 64));
            const uint32_t be = __builtin_amdgcn_readfirstlane(__shfl(e, j, 64));
            const float4 r = wave_leaf_sum(stage[w], prod, bf, be, lane);
            if (lane == 0)
                leaf_sum[wave_id + (i0 + j) * waves] = make_float4(r.x, r.y, r.z, __uint_as_float(be - bf));
        }
    }
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        const BPos q = bpos_info(keys, n, k);
        if (!q.inside) continue;
        const uint32_t c1 = cpos[k] + q.ns;  // C(k+1)
        uint32_t run_end = c1 < ncells ? breaker[c1] : n;
        run_end = min(run_end, bh_prefix[q.rank] + stats[q.rank].n_in);
        set_run_length(node4, k + c1, run_end > k ? run_end - k : 1u);
    }
}

// ---- dense reduction per (system, depth-5 tile), then per system --------------

__device__ __forceinline__ int tile_level_offset(int d) {
    return d == 9 ? 0 : d == 8 ? 256 : d == 7 ? 320 : d == 6 ? 336 : 340;
}

__global__ __launch_bounds__(256) void bb_tile_kernel(
    const uint32_t *__restrict__ tile_range, const uint32_t *__restrict__ leaf_first,
    const uint32_t *__restrict__ leaf_key, const float4 *__restrict__ leaf_sum,
    const int8_t *__restrict__ sdep, const uint32_t *__restrict__ cpos,
    const BatchCounts *__restrict__ counts, const uint32_t *__restrict__ live_tiles,
    uint32_t live_per_list, float4 *__restrict__ node4, uint32_t *__restrict__ ncount,
    TileCell *__restrict__ top) {
    __shared__ TileCell l10[1024];
    __shared__ TileCell up[341];
    // gridDim.x / kLiveLists workgroups per list (gridDim.x a multiple of kLiveLists)
    const uint32_t q = blockIdx.x % kLiveLists, per = gridDim.x / kLiveLists;
    const uint32_t nlive = counts->n_live[q];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = blockIdx.x / kLiveLists; i < nlive; i += per) {  // only tiles holding leaves
        const uint32_t T = live_tiles[q * live_per_list + i];
        const uint32_t lb = tile_range[2 * T], le = tile_range[2 * T + 1];
        for (uint32_t c = t; c < 1024; c += 256)
            l10[c] = TileCell{0u, 0u, 0xffffffffu, 0u, 0.0f, 0.0f, 0.0f, 0.0f};
        __syncthreads();
        for (uint32_t L = lb + t; L < le; L += 256) {
            const uint32_t c = leaf_key[L] & 1023u;
            const float4 s = leaf_sum[L];
            const uint32_t n = __float_as_uint(s.w);
            TileCell cell{n, n + (n >= 2u ? 1u : 0u), leaf_first[L], 0u, s.x, s.y, s.z, 0.0f};
            l10[c] = cell;
            emit_cell(cell, 10, sdep, cpos, node4, ncount);
        }
        __syncthreads();
        up[t] = reduce4(l10[4 * t], l10[4 * t + 1], l10[4 * t + 2], l10[4 * t + 3]);
        __syncthreads();
        for (int d = 8; d >= 5; --d) {
            const uint32_t cnt = 1u << (2 * (d - 5));
            const int o = tile_level_offset(d), oc = tile_level_offset(d + 1);
            if (t < cnt)
                up[o + t] = reduce4(up[oc + 4 * t], up[oc + 4 * t + 1], up[oc + 4 * t + 2],
                                    up[oc + 4 * t + 3]);
            __syncthreads();
        }
        for (uint32_t c = t; c < 341; c += 256) {
            const int d = c < 256 ? 9 : c < 320 ? 8 : c < 336 ? 7 : c < 340 ? 6 : 5;
            emit_cell(up[c], d, sdep, cpos, node4, ncount);
        }
        if (t == 0) top[T] = up[340];
        __syncthreads();
    }
}

__device__ __forceinline__ int top_level_offset(int d) {
    return d == 5 ? 0 : d == 4 ? 1024 : d == 3 ? 1280 : d == 2 ? 1344 : d == 1 ? 1360 : 1364;
}

__global__ __launch_bounds__(1024) void bb_top_kernel(const TileCell *__restrict__ top,
                                                      const uint32_t *__restrict__ tile_range,
                                                      const int8_t *__restrict__ sdep,
                                                      const uint32_t *__restrict__ cpos,
                                                      const BatchCounts *__restrict__ counts,
                                                      float4 *__restrict__ node4,
                                                      uint32_t *__restrict__ ncount) {
    __shared__ TileCell lv[1365];
    const uint32_t t = threadIdx.x;
    for (uint32_t r = blockIdx.x; r < counts->n_bh; r += gridDim.x) {
        const uint32_t T = r * kTiles + t;
        lv[t] = tile_range[2 * T] != tile_range[2 * T + 1]
                    ? top[T]
                    : TileCell{0u, 0u, 0xffffffffu, 0u, 0.0f, 0.0f, 0.0f, 0.0f};
        __syncthreads();
        for (int d = 4; d >= 0; --d) {
            const uint32_t cnt = 1u << (2 * d);
            const int o = top_level_offset(d), oc = top_level_offset(d + 1);
            if (t < cnt)
                lv[o + t] = reduce4(lv[oc + 4 * t], lv[oc + 4 * t + 1], lv[oc + 4 * t + 2],
                                    lv[oc + 4 * t + 3]);
            __syncthreads();
        }
        if (t < 341) {
            const uint32_t c = 1024 + t;
            const int d = c < 1280 ? 4 : c < 1344 ? 3 : c < 1360 ? 2 : c < 1364 ? 1 : 0;
            emit_cell(lv[c], d, sdep, cpos, node4, ncount);
        }
        __syncthreads();
    }
}

// ---- one persistent force launch for every system ------------------------------

// Barnes-Hut group g of one system: the walk of bh_kernels.hip over the
// system's node range [root, end_slot) with its own constants.
// kHalf: a group of the W = 32 vote (the published runs' subgroup width), one
// 32-particle group per wave with sorted position g * 32 + l held by lanes l
// and l + 32, walked exactly as bh_traverse32_kernel walks it (both halves
// vote alike; each half sums its part of every chunk (sum_run_queue), and the
// halves' partials are added before the acceleration), so the result is that
// kernel's bit for bit.
template <bool kHalf>
__device__ void bb_walk_group(uint32_t g, const SysDesc &d, const RankStats &st, float4 *stage,
                              uint2 *queue, const uint32_t *__restrict__ sorted_src,
                              const float2 *__restrict__ pos, const float2 *__restrict__ vel,
                              const float4 *__restrict__ node4,
                              const uint32_t *__restrict__ ncount, float2 *__restrict__ pos_out,
                              float2 *__restrict__ vel_out, uint32_t node_cap,
                              unsigned long long &vis, unsigned long long &acc,
                              uint32_t *__restrict__ chunks_io, const BatchTune &tune) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t src = kHalf ? sorted_src[d.sorted_base + g * 32 + (lane & 31u)]
                               : sorted_src[d.sorted_base + g * 64 + lane];
    const bool real = src != kBatchSentinel;
    const uint32_t nreal = kHalf ? (uint32_t)__popcll(__ballot(real) & 0xffffffffull)
                                 : (uint32_t)__popcll(__ballot(real));
    float px = 0.0f, py = 0.0f;
    if (real) {
        const float2 p = pos[src];
        px = p.x;
        py = p.y;
    }
    const uint32_t end = __builtin_amdgcn_readfirstlane(min(st.end_slot, node_cap));
    const uint32_t root = __builtin_amdgcn_readfirstlane(st.root);
    Partials P;
    // start at the issue priority this work item reached in the last step
    const uint32_t prio0 = tune.prio_carry ? *chunks_io / tune.prio_carry : 0u;
    uint32_t chunks = 0;
    (void)walk_partials<true, kHalf>(root, end, stage, queue, node4, ncount, px, py, real, nreal,
                                     d.eps2, d.theta2, 0xfu, P, vis, acc, tune.prio_step, prio0,
                                     &chunks, tune.run_queue);
    if (lane == 0) *chunks_io = chunks;
    if (kHalf) {
        P.x0 += __shfl_xor(P.x0, 32, 64);
        P.y0 += __shfl_xor(P.y0, 32, 64);
        P.x1 += __shfl_xor(P.x1, 32, 64);
        P.y1 += __shfl_xor(P.y1, 32, 64);
        P.x2 += __shfl_xor(P.x2, 32, 64);
        P.y2 += __shfl_xor(P.y2, 32, 64);
        P.x3 += __shfl_xor(P.x3, 32, 64);
        P.y3 += __shfl_xor(P.y3, 32, 64);
    }
    if (real && (!kHalf || lane < 32)) {
        float ax, ay;
        partial_sum(P, ax, ay);
        const float2 v = vel[src];
        const float nvx = v.x + ax * d.kv, nvy = v.y + ay * d.kv;
        pos_out[src] = make_float2(px + (v.x + nvx) * d.kp, py + (v.y + nvy) * d.kp);
        vel_out[src] = make_float2(nvx, nvy);
    }
}

// Direct-sum tile: 128 targets (two per lane) x one source slice; the same
// arithmetic as ds_partial_kernel, so the per-slice partials are identical.
__device__ void bb_ds_tile(uint32_t sub, const SysDesc &d, const float2 *__restrict__ pos,
                           const float *__restrict__ mass, float2 *__restrict__ partial) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tb = sub >> 16, slice = sub & 0xffffu;
    const uint32_t t0 = tb * 128 + lane, t1 = t0 + 64;
    const float2 p0 = t0 < d.na ? pos[d.off + t0] : make_float2(0.0f, 0.0f);
    const float2 p1 = t1 < d.na ? pos[d.off + t1] : make_float2(0.0f, 0.0f);
    float ax0 = 0.0f, ay0 = 0.0f, ax1 = 0.0f, ay1 = 0.0f;
    const uint32_t j0 = d.off + slice * d.ds_jlen;
    const uint32_t j1 = min(d.off + d.na, j0 + d.ds_jlen);
    const float eps2 = d.eps2;
    for (uint32_t j = j0; j < j1; j += 8) {  // slices are whole multiples of 64
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float2 s = pos[j + u];
            const float m = mass[j + u];
            {
                const float dx = s.x - p0.x, dy = s.y - p0.y;
                const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, eps2));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float tt = (m * inv) * (inv * inv);
                ax0 = __builtin_fmaf(dx, tt, ax0);
                ay0 = __builtin_fmaf(dy, tt, ay0);
            }
            {
                const float dx = s.x - p1.x, dy = s.y - p1.y;
                const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, eps2));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float tt = (m * inv) * (inv * inv);
                ax1 = __builtin_fmaf(dx, tt, ax1);
                ay1 = __builtin_fmaf(dy, tt, ay1);
            }
        }
    }
    float2 *dst = partial + d.ds_part_off + (size_t)slice * d.na;
    if (t0 < d.na) dst[t0] = make_float2(ax0, ay0);
    if (t1 < d.na) dst[t1] = make_float2(ax1, ay1);
}

__global__ __launch_bounds__(256) void bb_force_kernel(
    const uint2 *__restrict__ items, const SysDesc *__restrict__ desc,
    const uint32_t *__restrict__ remaining, const RankStats *__restrict__ stats,
    BatchCounts *__restrict__ counts,
    const uint32_t *__restrict__ sorted_src, const float2 *__restrict__ pos,
    const float2 *__restrict__ vel, const float *__restrict__ mass,
    const float4 *__restrict__ node4, const uint32_t *__restrict__ ncount,
    float2 *__restrict__ pos_out, float2 *__restrict__ vel_out, float2 *__restrict__ partial,
    uint32_t node_cap, unsigned long long *__restrict__ wave_acc, uint32_t *__restrict__ ichunk,
    BatchTune tune) {
    __shared__ float4 stage[4][64];
    __shared__ uint2 runq[4][kRunQueue];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t nitems = counts->n_items;
    // this wave's queue (kForceQueues queues, each shared by the waves of
    // one residue): items are taken one at a time, the next one's ticket
    // drawn before the current item is walked
    const uint32_t q = (blockIdx.x * 4 + w) % kForceQueues;
    uint32_t *const qctr = &counts->qnext[q];
    auto draw = [&]() -> uint32_t {
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(qctr, 1u);
        return __builtin_amdgcn_readfirstlane(__shfl(j, 0, 64));
    };
    unsigned long long acc = 0, vis = 0;
    uint32_t next = draw();
    while (true) {
        const uint32_t it = q + kForceQueues * next;
        if (it >= nitems) break;  // the queue is drained (every later ticket is beyond too)
        next = draw();
        const uint2 item = items[it];
        const SysDesc d = desc[item.x];
        const bool live = remaining[item.x] != 0u;
        if (d.algo == GSIM_BARNES_HUT) {
            const RankStats st = stats[d.bh_rank];
            // no particle inside the domain, or a spent budget (the keys
            // kernel has already written the system's output)
            if (!live || st.n_in == 0) continue;
            if (d.w == 32)
                bb_walk_group<true>(item.y, d, st, stage[w], runq[w], sorted_src, pos, vel, node4,
                                    ncount, pos_out, vel_out, node_cap, vis, acc, &ichunk[it],
                                    tune);
            else
                bb_walk_group<false>(item.y, d, st, stage[w], runq[w], sorted_src, pos, vel,
                                     node4, ncount, pos_out, vel_out, node_cap, vis, acc,
                                     &ichunk[it], tune);
        } else if (live) {
            bb_ds_tile(item.y, d, pos, mass, partial);
        }
    }
    if (lane == 0) {
        wave_acc[blockIdx.x * 4 + w] = acc;
        wave_acc[kMaxForceWaves + blockIdx.x * 4 + w] = vis;
    }
}

// End of a step: spend one step of every limited budget, and add the force
// waves' acceptances and visits to the running totals (one atomic each per
// block).
__global__ __launch_bounds__(256) void bb_finish_kernel(uint32_t *__restrict__ remaining,
                                                        BatchCounts *__restrict__ counts,
                                                        const unsigned long long *__restrict__ wave_acc,
                                                        uint32_t nwaves) {
    __shared__ unsigned long long part[2][4];
    // the force launch's queue counters, for the next step
    if (blockIdx.x == 0 && threadIdx.x < kForceQueues) counts->qnext[threadIdx.x] = 0u;
    for (uint32_t s = blockIdx.x * 256 + threadIdx.x; s < counts->n_sys; s += gridDim.x * 256) {
        const uint32_t r = remaining[s];
        if (r != 0u && r != kUnlimitedSteps) remaining[s] = r - 1u;
    }
    unsigned long long a = 0, v = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < nwaves; i += gridDim.x * 256) {
        a += wave_acc[i];
        v += wave_acc[kMaxForceWaves + i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        v += __shfl_xor(v, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = a;
        part[1][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long ta = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        const unsigned long long tv = part[1][0] + part[1][1] + part[1][2] + part[1][3];
        if (ta) atomicAdd(&counts->accepts, ta);
        if (tv) atomicAdd(&counts->visits, tv);
    }
}

// Direct-sum combine + integrate (SimShader.comp:68-69) for every DS particle.
__global__ __launch_bounds__(256) void bb_ds_combine_kernel(
    const SysDesc *__restrict__ desc, const uint32_t *__restrict__ remaining,
    const uint32_t *__restrict__ ds_sys,
    const uint32_t *__restrict__ ds_prefix, const BatchCounts *__restrict__ counts,
    const float2 *__restrict__ partial, const float2 *__restrict__ pos,
    const float2 *__restrict__ vel, float2 *__restrict__ pos_out, float2 *__restrict__ vel_out) {
    const uint32_t nd = counts->n_ds, n = counts->n_ds_particles;
    for (uint32_t f = blockIdx.x * 256 + threadIdx.x; f < n; f += gridDim.x * 256) {
        const uint32_t r = rank_of(ds_prefix, nd, f);
        const uint32_t di = ds_sys[r];
        const SysDesc d = desc[di];
        const uint32_t li = f - ds_prefix[r];
        if (remaining[di] == 0u) {  // budget spent: hold the state
            pos_out[d.off + li] = pos[d.off + li];
            vel_out[d.off + li] = vel[d.off + li];
            continue;
        }
        float ax = 0.0f, ay = 0.0f;
        for (uint32_t s = 0; s < d.ds_slices; ++s) {
            const float2 a = partial[d.ds_part_off + (size_t)s * d.na + li];
            ax = ax + a.x;
            ay = ay + a.y;
        }
        const uint32_t i = d.off + li;
        const float2 p = pos[i], v = vel[i];
        const float nvx = v.x + ax * d.kv, nvy = v.y + ay * d.kv;
        pos_out[i] = make_float2(p.x + ((v.x + nvx) * 0.5f) * d.dt, p.y + ((v.y + nvy) * 0.5f) * d.dt);
        vel_out[i] = make_float2(nvx, nvy);
    }
}

}  // namespace

// One batched step: Barnes-Hut build for all BH systems, one persistent force
// launch for every system, the direct-sum combine.  Launched onto b->stream
// (captured into a graph by the caller).
hipError_t batch_launch_step(gsim_batch *b, int in, int out) {
    BatchDevice &d = b->d;
    hipStream_t st = b->stream;
    const uint32_t g_part = (uint32_t)std::min<size_t>((b->capacity + 255) / 256, 4096);
    const uint32_t live_per_list = (uint32_t)(((size_t)b->max_systems * kTiles + kLiveLists - 1) / kLiveLists);
    const BatchTune &tune = b->tune;
    const uint32_t rounds = tune.sort_rounds;
    if (rounds != 8 && rounds != 12 && rounds != 16) return hipErrorInvalidValue;
    const uint32_t tile_keys = rounds * kSortThreads;
    const size_t tiles = (b->capacity + tile_keys - 1) / tile_keys;  // <= tiles_cap
    const uint32_t g_tiles = (uint32_t)std::min<size_t>(tiles, 2048);
    const uint32_t g_hist = (uint32_t)std::min<size_t>(tiles, 1024);
    bb_keys_kernel<<<g_part, 256, 0, st>>>(d.desc, d.bh_sys, d.bh_prefix, d.remaining, d.counts, d.pos[in],
                                           d.mass, d.pos[out], d.vel[in], d.vel[out], d.keys[0],
                                           d.idx[0], d.tile_range, d.stats, d.digit_total,
                                           d.status, tile_keys);
#define GSIM_BB_SORT(R)                                                                          \
    do {                                                                                         \
        bb_hist_kernel<R><<<g_hist, kSortThreads, 0, st>>>(d.keys[0], d.counts, d.digit_total); \
        for (int pass = 0; pass < 4; ++pass) {                                                   \
            const int i = pass & 1, o = i ^ 1;                                                   \
            bb_onesweep_kernel<R><<<g_tiles, kSortThreads, 0, st>>>(                              \
                d.keys[i], d.idx[i], d.keys[o], d.idx[o], d.counts, (uint32_t)pass,              \
                d.digit_total, d.status + half * (pass & 1), d.status + half * ((pass + 1) & 1)); \
        }                                                                                        \
    } while (0)
    // 4 passes: keys[0] -> keys[1] -> keys[0] -> keys[1] -> keys[0]; look-back
    // words alternate between the two status halves
    const size_t half = 256 * b->tiles_cap;
    if (rounds == 8) GSIM_BB_SORT(8);
    else if (rounds == 12) GSIM_BB_SORT(12);
    else GSIM_BB_SORT(16);
#undef GSIM_BB_SORT
    const uint32_t *sk = d.keys[0], *si = d.idx[0];
    const uint32_t g_blk = (uint32_t)std::min<size_t>(b->blocks_cap, 4096);
    bb_prep_kernel<<<g_blk, 256, 0, st>>>(sk, d.counts, d.block_cells, d.block_leaves);
    bb_block_scan_kernel<<<2, 1024, 0, st>>>(d.block_cells, d.block_leaves, d.counts);
    bb_emit_kernel<<<g_blk, 256, 0, st>>>(sk, si, d.pos[in], d.mass, d.counts, d.bh_prefix,
                                          d.block_cells, d.block_leaves, d.sdep, d.cpos, d.prod,
                                          d.leaf_first, d.leaf_key, d.node4, d.ncount,
                                          d.sorted_src, d.breaker, d.stats);
    bb_leaf_kernel<<<4096, 256, 0, st>>>(d.leaf_first, d.leaf_key, d.prod, sk, d.cpos, d.breaker,
                                         d.counts, d.bh_prefix, d.stats, d.leaf_sum,
                                         d.tile_range, d.live_tiles, live_per_list, d.node4);
    const uint32_t g_tile = 4096;  // 64 workgroups per live list
    bb_tile_kernel<<<g_tile, 256, 0, st>>>(d.tile_range, d.leaf_first, d.leaf_key, d.leaf_sum,
                                           d.sdep, d.cpos, d.counts, d.live_tiles, live_per_list,
                                           d.node4, d.ncount, d.top);
    bb_top_kernel<<<std::min<uint32_t>(b->max_systems, 256), 1024, 0, st>>>(
        d.top, d.tile_range, d.sdep, d.cpos, d.counts, d.node4, d.ncount);
    const uint32_t waves = 1024u * b->tune.waves_per_simd;
    bb_force_kernel<<<waves / 4, 256, 0, st>>>(d.items, d.desc, d.remaining, d.stats, d.counts, d.sorted_src,
                                               d.pos[in], d.vel[in], d.mass, d.node4, d.ncount,
                                               d.pos[out], d.vel[out], d.partial,
                                               (uint32_t)b->node_capacity, d.wave_acc, d.ichunk,
                                               tune);
    bb_ds_combine_kernel<<<g_part, 256, 0, st>>>(d.desc, d.remaining, d.ds_sys, d.ds_prefix,
                                                 d.counts, d.partial, d.pos[in], d.vel[in],
                                                 d.pos[out], d.vel[out]);
    bb_finish_kernel<<<8, 256, 0, st>>>(d.remaining, d.counts, d.wave_acc, waves);
    return hipGetLastError();
}

}  // namespace gsim
