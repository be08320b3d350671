// Barnes-Hut build and traversal for gfx950.
//
// Tree semantics follow the reference's fixed-domain quadtree
// (BarnesHutSimulation.cpp:885-918 and its seven shaders): domain [-500,500)^2,
// 1024^2 leaf grid, 11 levels, a node for every particle and for every cell
// holding >= 2 particles, flattened in pre-order with subtree sizes.  The
// build here is sparse instead of the reference's dense 1.4M-cell pyramid:
//
//   1. keys      20-bit Morton cell per particle (InitShader.comp:80-98); the
//                out-of-domain kill of InitShader.comp:110-117.
//   2. sort      stable LSD radix sort of (cell, index), 3 passes of 8 bits,
//                so a leaf's particles are in ascending index order: one
//                histogram launch, then one launch per pass (look-back).
//   3. prep      s(k) = depth shared by sorted keys k-1 and k; the cell nodes
//                that start at k are the depths in (s(k), s(k+1)], so an
//                exclusive scan C(k) of their number gives every pre-order
//                slot: particle k -> k + C(k+1); the depth-d cell starting at
//                k -> k + C(k) + (d - s(k) - 1).  Per-block counts here.
//   4. emit      the scan (each block adds up the counts before it), particle
//                nodes, sorted_src, m*p products, leaf records.
//   5. leafsum   leaf sums in ascending particle index (the canonical order of
//                the reference's float atomics, InitShader.comp:107-109).
//   6. tile      one workgroup per depth-5 cell: the 32x32 leaf tile is
//                reduced in LDS level by level with ((c0+c1)+c2)+c3 exactly as
//                TreeInitShader.comp:78-83, counts y = sum + [sum > 1]; cells
//                with y > 1 are written as nodes (TreeMoveShader.comp:60-77).
//   7. top       the last tile workgroup to finish reduces the 1024 depth-5
//                cells to the root.
//   8. traverse  groups of W sorted slots walk one path; a node is accepted
//                iff all W lanes accept it (ForceShader.comp:111-141), then
//                integrate (ForceShader.comp:145-150).
#include "bh_device.hpp"
#include "radix_device.hpp"

namespace gsim {
namespace {


// ---- 1. keys ----------------------------------------------------------------

// ---- 0. adaptive bounding box (opt-in; beyond the reference) -------------------
// Extremes of the finite positions: per-block partials, then one block that
// also applies the box formula {cx, cy, h, r0} (order-independent, so exact).

__device__ __forceinline__ void minmax4_wave(float &a, float &b, float &c, float &d) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a = fminf(a, __shfl_xor(a, off, 64));
        b = fmaxf(b, __shfl_xor(b, off, 64));
        c = fminf(c, __shfl_xor(c, off, 64));
        d = fmaxf(d, __shfl_xor(d, off, 64));
    }
}

__global__ __launch_bounds__(256) void bh_box_partial_kernel(const float2 *__restrict__ pos,
                                                             uint32_t na,
                                                             float4 *__restrict__ part) {
    __shared__ float4 wv[4];
    float x0 = __builtin_inff(), x1 = -__builtin_inff(), y0 = __builtin_inff(),
          y1 = -__builtin_inff();
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < na; i += gridDim.x * 256) {
        const float2 p = pos[i];
        if (!__builtin_isfinite(p.x) || !__builtin_isfinite(p.y)) continue;
        x0 = fminf(x0, p.x);
        x1 = fmaxf(x1, p.x);
        y0 = fminf(y0, p.y);
        y1 = fmaxf(y1, p.y);
    }
    minmax4_wave(x0, x1, y0, y1);
    if ((threadIdx.x & 63) == 0) wv[threadIdx.x >> 6] = make_float4(x0, x1, y0, y1);
    lds_barrier();
    if (threadIdx.x == 0) {
        float4 r = wv[0];
        for (int w = 1; w < 4; ++w) {
            r.x = fminf(r.x, wv[w].x);
            r.y = fmaxf(r.y, wv[w].y);
            r.z = fminf(r.z, wv[w].z);
            r.w = fmaxf(r.w, wv[w].w);
        }
        part[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void bh_box_final_kernel(const float4 *__restrict__ part,
                                                           uint32_t nparts,
                                                           float *__restrict__ box) {
    __shared__ float4 wv[4];
    float x0 = __builtin_inff(), x1 = -__builtin_inff(), y0 = __builtin_inff(),
          y1 = -__builtin_inff();
    for (uint32_t i = threadIdx.x; i < nparts; i += 256) {
        const float4 r = part[i];
        x0 = fminf(x0, r.x);
        x1 = fmaxf(x1, r.y);
        y0 = fminf(y0, r.z);
        y1 = fmaxf(y1, r.w);
    }
    minmax4_wave(x0, x1, y0, y1);
    if ((threadIdx.x & 63) == 0) wv[threadIdx.x >> 6] = make_float4(x0, x1, y0, y1);
    lds_barrier();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < 4; ++w) {
        wv[0].x = fminf(wv[0].x, wv[w].x);
        wv[0].y = fmaxf(wv[0].y, wv[w].y);
        wv[0].z = fminf(wv[0].z, wv[w].z);
        wv[0].w = fmaxf(wv[0].w, wv[w].w);
    }
    float xmin = wv[0].x, xmax = wv[0].y, ymin = wv[0].z, ymax = wv[0].w;
    if (!(xmin <= xmax)) xmin = xmax = ymin = ymax = 0.0f;  // no finite particle
    const float cx = (xmin + xmax) * 0.5f, cy = (ymin + ymax) * 0.5f;
    const float w = fmaxf(xmax - xmin, ymax - ymin);
    float h = w * 0.501953125f;  // half the extent plus 2^-9: the farthest maps below 1024
    if (!(h >= 1e-3f)) h = 1e-3f;
    box[0] = cx;
    box[1] = cy;
    box[2] = h;
    box[3] = (h * h) * 2.0f;
}

// kDeep: 30-bit keys of the 32768^2 grid (15-bit leaves) instead of 20-bit cells.
// One block of kR x kSortThreads particles (kR per thread) per workgroup.
// The digits of every radix pass are counted into digit_total, this build's
// half of the totals (kDigitCopies copies of 1024 words, workgroup b adding
// into copy b mod kDigitCopies; zeroed by the previous build's keys kernel,
// or at allocation), while digit_next, the other half, is zeroed for the next
// build; the radix passes' look-back words are zeroed as well.
// The particles are in the last step's sorted order: slot i holds the
// particle of original index perm[i].  Cells and slots are written at the
// original index (cells[perm[i]] = {cell, i}), so the sort runs over original
// indices, whose order breaks ties in a leaf, and carries each particle's
// slot.  An out-of-domain particle gets the killed key; its
// kill (InitShader.comp:110-117) is written by the emit kernel at the slot it
// sorts to.  kWalk: the last walk (write_next_key) and emit already wrote the
// cells; only the digits are counted.
template <bool kBox, bool kDeep, int kR, bool kWalk>
__global__ __launch_bounds__(kSortThreads) void bh_keys_kernel(
    const float2 *__restrict__ pos, const uint32_t *__restrict__ perm, uint32_t na,
    uint2 *__restrict__ cells, uint32_t *__restrict__ tile_range,
    uint32_t *__restrict__ digit_total, uint32_t *__restrict__ digit_next,
    uint32_t *__restrict__ status, uint32_t status_words, BhCounters *__restrict__ ctr,
    const float *__restrict__ box) {
    constexpr int kPasses = kDeep ? 4 : 3;
    __shared__ uint32_t h[kPasses][256];
    const uint32_t t = threadIdx.x;
    const uint32_t g = blockIdx.x * kSortThreads + t, stride = gridDim.x * kSortThreads;
    for (uint32_t j = g; j < 2 * kTiles; j += stride) tile_range[j] = 0;
    for (uint32_t j = g; j < kDigitCopies * 1024; j += stride) digit_next[j] = 0;
    for (uint32_t j = g; j < status_words; j += stride) status[j] = 0;
    if (g == 0) {
        ctr->next_group = 0;
        ctr->sort_stall = 0;
        ctr->tiles_busy = 0;
        ctr->tiles_done = 0;
    }
#pragma unroll
    for (int p = 0; p < kPasses; ++p) h[p][t] = 0;
    lds_barrier();
    const uint32_t base = blockIdx.x * (kR * kSortThreads);
    if constexpr (kWalk) {
        uint32_t c[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            const uint32_t i = base + r * kSortThreads + t;
            c[r] = i < na ? cells[i].x : 0u;
        }
#pragma unroll
        for (int r = 0; r < kR; ++r)
            if (base + r * kSortThreads + t < na) hist_add<kPasses, false>(h, c[r], true);
    } else {
    float2 pr[kR];
    uint32_t orig[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        const uint32_t i = base + r * kSortThreads + t;
        pr[r] = i < na ? pos[i] : make_float2(0.0f, 0.0f);
        orig[r] = i < na ? perm[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        const uint32_t i = base + r * kSortThreads + t;
        if (i >= na) continue;
        const float2 p = pr[r];
        uint32_t c;
        if (kDeep) c = kBox ? cell30_box(p.x, p.y, box[0], box[1], box[2]) : cell30(p.x, p.y);
        else c = kBox ? cell20_box(p.x, p.y, box[0], box[1], box[2]) : cell20(p.x, p.y);
        if (c == 0xffffffffu) c = kDeep ? kKilledKeyDeep : kKilledKey;
        cells[orig[r]] = make_uint2(c, i);
        hist_add<kPasses, false>(h, c, true);
    }
    }
    lds_barrier();
    uint32_t *const copy = digit_total + (blockIdx.x % kDigitCopies) * 1024;
#pragma unroll
    for (int p = 0; p < kPasses; ++p) {
        const uint32_t n = h[p][t];
        if (n) atomicAdd(&copy[p * 256 + t], n);
    }
}

// ---- 2. stable LSD radix sort, one sweep per pass (radix_device.hpp) -----------

// One pass, one tile per workgroup: digit (key >> shift) & 255, with the
// two-level look-back (grouped_prefix).  digit_total: this pass's 256 totals,
// split over kDigitCopies copies 1024 words apart; status: this pass's
// [tiles][256] tile words followed by its [groups][256] group words, zeroed;
// Tiles of kR x kSortThreads keys.
template <int kR>
__global__ __launch_bounds__(kSortThreads) void radix_onesweep_kernel(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ idx_in,
    const uint2 *__restrict__ cells_in, uint32_t *__restrict__ keys_out,
    uint32_t *__restrict__ idx_out, uint32_t n, uint32_t shift,
    const uint32_t *__restrict__ digit_total, uint32_t *__restrict__ status,
    uint32_t *__restrict__ stall) {
    __shared__ SortShared<false, kR> sh;
    sort_tile_clear(sh);  // the tile's first barrier follows its ranking
    // Tile = workgroup index.  A tile only waits for lower tiles, which the
    // dispatcher starts first, and a pass's workgroups all fit on the GPU at
    // once, so every wait ends (and each is bounded besides).
    sort_tile<false, kR, true>(keys_in, idx_in, keys_out, idx_out, n, shift, blockIdx.x, 0u,
                               status, stall, sh, gridDim.x, status + (size_t)gridDim.x * 256,
                               digit_total, kDigitCopies, 1024u, cells_in);
}

// ---- 3./4. slot scan and emission --------------------------------------------

// Per sorted position k: s(k), s(k+1) and the two scanned quantities.
struct PosInfo {
    int sk, sk1;
    uint32_t ns, lf;
    bool inside;
    uint32_t key;
    bool g10;  // first particle of an occupied depth-10 cell (15-bit leaves)
};

// bits: 10 (the reference's grid) or 15 (15-bit leaves); keys at or above
// 1 << 2*bits are killed particles.
__device__ __forceinline__ PosInfo pos_info(const uint32_t *__restrict__ keys, uint32_t na,
                                            uint32_t k, int bits) {
    PosInfo q{-1, -1, 0u, 0u, false, 0u, false};
    if (k >= na) return q;
    const uint32_t killed = 1u << (2 * bits);
    q.key = keys[k];
    if (q.key >= killed) return q;
    q.inside = true;
    q.sk = k == 0 ? -1 : shared_depth_bits(keys[k - 1], q.key, bits);
    if (k + 1 < na) {
        const uint32_t nk = keys[k + 1];
        if (nk < killed) q.sk1 = shared_depth_bits(q.key, nk, bits);
    }
    q.ns = q.sk1 > q.sk ? (uint32_t)(q.sk1 - q.sk) : 0u;
    q.lf = q.sk < bits ? 1u : 0u;
    q.g10 = q.sk < 10;
    return q;
}

// block_gran (15-bit leaves only, else nullptr): per-block counts of the
// occupied depth-10 cells as well.
__global__ __launch_bounds__(256) void bh_prep_kernel(const uint32_t *__restrict__ keys,
                                                      uint32_t na, int bits,
                                                      uint32_t *__restrict__ block_cells,
                                                      uint32_t *__restrict__ block_leaves,
                                                      uint32_t *__restrict__ block_gran) {
    __shared__ uint32_t wa[4], wb[4], wc[4];
    const PosInfo q = pos_info(keys, na, blockIdx.x * 256 + threadIdx.x, bits);
    uint32_t a = q.ns, b = q.lf;
    wave_scan2(a, b);
    const uint32_t c = (uint32_t)__popcll(__ballot(q.g10));
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 63) {
        wa[wave] = a;
        wb[wave] = b;
        wc[wave] = c;
    }
    lds_barrier();
    if (threadIdx.x == 0) {
        block_cells[blockIdx.x] = wa[0] + wa[1] + wa[2] + wa[3];
        block_leaves[blockIdx.x] = wb[0] + wb[1] + wb[2] + wb[3];
        if (block_gran) block_gran[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
    }
}

// block_cells, block_leaves (and block_gran): bh_prep_kernel's per-block
// counts; each block adds up those of the blocks before it (integer sums, a
// few loads per thread) instead of a separate scan launch.
__global__ __launch_bounds__(256) void bh_emit_kernel(
    const uint32_t *__restrict__ keys, const uint32_t *__restrict__ sidx,
    const float2 *__restrict__ pos, const float *__restrict__ mass,
    const float2 *__restrict__ vel, const uint32_t *__restrict__ perm_in,
    float2 *__restrict__ pos_out, float2 *__restrict__ vel_out, float *__restrict__ mass_out,
    uint32_t *__restrict__ perm_out, uint2 *__restrict__ ncells, uint32_t na, int bits,
    const uint32_t *__restrict__ block_cells, const uint32_t *__restrict__ block_leaves,
    int8_t *__restrict__ sdep, uint32_t *__restrict__ cpos, float4 *__restrict__ prod,
    uint32_t *__restrict__ leaf_first, uint32_t *__restrict__ leaf_key,
    float4 *__restrict__ node4, uint32_t *__restrict__ ncount,
    uint32_t *__restrict__ sorted_src, uint32_t *__restrict__ breaker,
    BhCounters *__restrict__ ctr, const uint32_t *__restrict__ block_gran,
    uint32_t *__restrict__ gran_first, uint32_t *__restrict__ gran_key) {
    __shared__ uint32_t wa[4], wb[4], wc[4], pa[4], pb[4], pc[4];
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    // the particle now at sorted position k, from its slot i of the last
    // step's order (nearly the same order, so nearly sequential reads); k is
    // its slot in the next order
    const uint32_t i = k < na ? sidx[k] : 0u;
    const float2 p = k < na ? pos[i] : make_float2(0.0f, 0.0f);
    const float m = k < na ? mass[i] : 0.0f;
    const uint32_t orig = k < na ? perm_in[i] : 0u;
    // the counts of the blocks before this one, eight loads per thread at once
    uint32_t xa = 0, xb = 0, xc = 0;
    for (uint32_t j0 = threadIdx.x; j0 < blockIdx.x; j0 += 8 * 256) {
        uint32_t va[8], vb[8], vc[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            const uint32_t j = j0 + 256 * u;
            const bool in = j < blockIdx.x;
            va[u] = in ? block_cells[j] : 0u;
            vb[u] = in ? block_leaves[j] : 0u;
            vc[u] = in && block_gran ? block_gran[j] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            xa += va[u];
            xb += vb[u];
            xc += vc[u];
        }
    }
    const PosInfo q = pos_info(keys, na, k, bits);
    uint32_t a = q.ns, b = q.lf;
    wave_scan2(a, b);
    const unsigned long long gb = __ballot(q.g10);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        xa += __shfl_xor(xa, off, 64);
        xb += __shfl_xor(xb, off, 64);
        xc += __shfl_xor(xc, off, 64);
    }
    if (lane == 63) {
        wa[wave] = a;
        wb[wave] = b;
        wc[wave] = (uint32_t)__popcll(gb);
        pa[wave] = xa;
        pb[wave] = xb;
        pc[wave] = xc;
    }
    lds_barrier();
    uint32_t ea = a - q.ns + ((pa[0] + pa[1]) + (pa[2] + pa[3]));
    uint32_t eb = b - q.lf + ((pb[0] + pb[1]) + (pb[2] + pb[3]));
    uint32_t ec = (uint32_t)__popcll(gb & ((1ull << lane) - 1ull)) +
                  ((pc[0] + pc[1]) + (pc[2] + pc[3]));
    for (uint32_t w = 0; w < wave; ++w) {
        ea += wa[w];
        eb += wb[w];
        ec += wc[w];
    }
    if (k >= na) return;
    perm_out[k] = orig;
    if (!q.inside) {
        // killed (InitShader.comp:110-117): not walked; the reference leaves
        // vel_out stale, here it keeps the input velocity
        const float kk = -kSimSize * 2.0f;
        pos_out[k] = make_float2(kk, kk);
        vel_out[k] = vel[i];
        mass_out[k] = 0.0f;
        if (ncells)  // the next build's cell, where the walk writes the others
            ncells[orig] = make_uint2(bits == (int)kDeepBits ? kKilledKeyDeep : kKilledKey, k);
        sorted_src[k] = na;  // ClearShader.comp:81-82 sentinel
        if (k == 0) {        // nothing inside the domain
            ctr->n_in = 0;
            ctr->n_cells = 0;
            ctr->n_nodes = 0;
            ctr->n_leaves = 0;
            ctr->n_gran = 0;
            leaf_first[0] = 0;
            if (gran_first) gran_first[0] = 0;
        }
        return;
    }
    mass_out[k] = m;
    const uint32_t C = ea, L = eb;
    const uint32_t slot = k + C + q.ns;
    // ParticleSortShader.comp:99-107 (canonical order)
    node4[slot] = make_float4(p.x, p.y, m, 0.0f);
    ncount[slot] = 1u;
    sorted_src[k] = i;
    sdep[k] = (int8_t)q.sk;
    cpos[k] = C;
    prod[k] = make_float4(p.x * m, p.y * m, m, 0.0f);  // pos *= mass, InitShader.comp:102
    if (q.lf) {
        leaf_first[L] = k;
        leaf_key[L] = q.key;
    }
    for (uint32_t j = 0; j < q.ns; ++j) breaker[C + j] = k;
    if (gran_first && q.g10) {
        gran_first[ec] = k;
        gran_key[ec] = q.key >> 10;  // the 20-bit cell
    }
    if (q.sk1 < 0) {  // last in-domain position
        ctr->n_in = k + 1;
        ctr->n_cells = C + q.ns;
        ctr->n_nodes = k + 1 + C + q.ns;
        ctr->n_leaves = L + q.lf;
        leaf_first[L + q.lf] = k + 1;
        ctr->n_gran = ec + (q.g10 ? 1u : 0u);
        if (gran_first) gran_first[ec + (q.g10 ? 1u : 0u)] = k + 1;
    }
}

// ---- 5. leaf sums and particle runs ------------------------------------------

constexpr uint32_t kLeafBlocks = 4096;
constexpr uint32_t kBoxParts = 256;  // blocks of the box reduction

// Phase A, one wave per leaf: the sum runs over the leaf's particles in
// sorted (= ascending index) order, the canonical order of the reference's
// float atomics; the wave stages 64 products at a time in LDS with one
// coalesced load and adds them in order.  Phase B, one thread per sorted position: the
// number of consecutive particle nodes from that particle's slot, i.e. up to
// the next position where a cell node starts (breaker[C(k+1)]).
// With 15-bit leaves (nsum != nullptr) a leaf of two or more particles is
// written as its depth-15 cell node at once, with its sums in nsum for the
// level reduction; there are no tiles.
__device__ __forceinline__ void emit_deep_leaf(const float4 &r, uint32_t f, uint32_t n,
                                               const int8_t *__restrict__ sdep,
                                               const uint32_t *__restrict__ cpos,
                                               float4 *__restrict__ node4,
                                               uint32_t *__restrict__ ncount,
                                               float4 *__restrict__ nsum, float r0) {
    if (n < 2u) return;
    const TileCell c{n, n + 1u, f, 0u, r.x, r.y, r.z, 0.0f};
    emit_cell(c, (int)kDeepBits, sdep, cpos, node4, ncount, r0);
    const uint32_t slot = f + cpos[f] + (uint32_t)((int)kDeepBits - (int)sdep[f] - 1);
    nsum[slot] = make_float4(r.x, r.y, r.z, __uint_as_float(n));
}

__global__ __launch_bounds__(256) void bh_leaf_kernel(
    const uint32_t *__restrict__ leaf_first, const uint32_t *__restrict__ leaf_key,
    const float4 *__restrict__ prod, const uint32_t *__restrict__ keys, uint32_t na, int bits,
    const uint32_t *__restrict__ cpos, const uint32_t *__restrict__ breaker,
    const BhCounters *__restrict__ ctr, float4 *__restrict__ leaf_sum,
    uint32_t *__restrict__ tile_range, const int8_t *__restrict__ sdep,
    float4 *__restrict__ node4, uint32_t *__restrict__ ncount,
    float4 *__restrict__ nsum, const float *__restrict__ box, uint32_t *__restrict__ tiles_busy) {
    __shared__ float4 stage[4][64];
    const uint32_t n_in = ctr->n_in, nl = ctr->n_leaves, ncells = ctr->n_cells;
    const bool deep = nsum != nullptr;
    const float r0 = box ? box[3] : kFixedR0;
    const uint32_t waves = gridDim.x * 4;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + w);
    // the first particle's run length: its loads go out first, its dependent
    // load (breaker) once the leaves are summed
    const uint32_t k0 = blockIdx.x * 256 + threadIdx.x;
    const bool has0 = k0 < n_in;
    const uint32_t kk0 = has0 ? k0 : 0u;
    const uint32_t key0 = keys[kk0], cpos0 = cpos[kk0];
    const uint32_t keyp = keys[kk0 ? kk0 - 1 : 0], keyn = keys[kk0 + 1 < na ? kk0 + 1 : kk0];
    // larger leaves: leaf L goes to wave (waves - 1 - L) mod waves, so the
    // largest ones (dense cores, low L when few) avoid the first waves, which
    // hold the small leaves; neighbouring dense leaves go to different waves;
    // each is streamed by the whole wave; the 10-bit leaf's cell node is
    // written here (15-bit leaves: its depth-15 cell)
    const uint32_t rw = waves - 1u - wave_id;
    for (uint32_t i0 = 0; rw + i0 * waves < nl; i0 += 64) {
        const uint32_t L = rw + (i0 + lane) * waves;
        uint32_t f = 0, e = 0;
        if (L < nl) {
            e = min(leaf_first[L + 1], n_in);
            f = min(leaf_first[L], e);
        }
        unsigned long long big = __ballot(L < nl && e - f > kSmallLeaf);
        while (big) {
            const uint32_t j = (uint32_t)__builtin_ctzll(big);
            big &= big - 1;
            const uint32_t bf = __builtin_amdgcn_readfirstlane(__shfl(f, j, 64));
            const uint32_t be = __builtin_amdgcn_readfirstlane(__shfl(e, j, 64));
            const uint32_t slot = deep ? 0u : cell_slot(bf, (int)kTreeDepth, sdep, cpos);
            const float4 r = wave_leaf_sum(stage[w], prod, bf, be, lane);
            if (lane == 0) {
                const uint32_t nb = be - bf;
                if (deep) {
                    emit_deep_leaf(r, bf, nb, sdep, cpos, node4, ncount, nsum, r0);
                } else {
                    leaf_sum[rw + (i0 + j) * waves] = make_float4(r.x, r.y, r.z, __uint_as_float(nb));
                    emit_cell_at(TileCell{nb, nb + 1u, bf, 0u, r.x, r.y, r.z, 0.0f},
                                 (int)kTreeDepth, slot, node4, ncount, r0);
                }
            }
        }
    }
    // small leaves: each wave takes 64 consecutive leaves, one per lane
    for (uint32_t base = wave_id * 64; base < nl; base += waves * 64) {
        const uint32_t L = base + lane;
        const bool valid = L < nl;
        uint32_t f = 0, e = 0;
        if (valid) {
            e = min(leaf_first[L + 1], n_in);
            f = min(leaf_first[L], e);
        }
        const bool small = valid && e - f <= kSmallLeaf;
        if (small) {
            const float4 r = lane_leaf_sum(prod, f, e);
            const uint32_t nb = e - f;
            if (deep) {
                emit_deep_leaf(r, f, nb, sdep, cpos, node4, ncount, nsum, r0);
            } else {
                leaf_sum[L] = make_float4(r.x, r.y, r.z, __uint_as_float(nb));
                emit_cell(TileCell{nb, nb + (nb >= 2u ? 1u : 0u), f, 0u, r.x, r.y, r.z, 0.0f},
                          (int)kTreeDepth, sdep, cpos, node4, ncount, r0);
            }
        }
        if (valid && !deep) {
            const uint32_t T = leaf_key[L] >> 10;
            if (L == 0 || (leaf_key[L - 1] >> 10) != T) {
                tile_range[2 * T] = L;
                atomicAdd(tiles_busy, 1u);
            }
            if (L + 1 == nl || (leaf_key[L + 1] >> 10) != T) tile_range[2 * T + 1] = L + 1;
        }
    }
    uint32_t slot0 = 0, end0 = 0;
    if (has0) {
        // pos_info(keys, na, k0, bits) from the loads above
        const uint32_t killed = 1u << (2 * bits);
        const int sk = k0 == 0 ? -1 : shared_depth_bits(keyp, key0, bits);
        const int sk1 = k0 + 1 < na && keyn < killed ? shared_depth_bits(key0, keyn, bits) : -1;
        const uint32_t c1 = cpos0 + (sk1 > sk ? (uint32_t)(sk1 - sk) : 0u);  // C(k+1)
        slot0 = k0 + c1;
        end0 = c1 < ncells ? breaker[c1] : n_in;
    }
    if (k0 < n_in) set_run_length(node4, slot0, end0 > k0 ? end0 - k0 : 1u);
    for (uint32_t k = k0 + gridDim.x * 256; k < n_in; k += gridDim.x * 256) {
        const PosInfo q = pos_info(keys, na, k, bits);
        const uint32_t c1 = cpos[k] + q.ns;  // C(k+1)
        const uint32_t run_end = c1 < ncells ? breaker[c1] : n_in;
        set_run_length(node4, k + c1, run_end > k ? run_end - k : 1u);
    }
}

// ---- 6'. sparse level reduction (15-bit leaves) ---------------------------------

// The sums of the depth-d cell (>= 2 particles) that starts at sorted position
// k, from its children in pre-order at slot + 1 on: a particle node (count 1)
// contributes 0 + m*p, a cell node its sums in nsum; each child's quadrant
// comes from the key bits of its first particle.  Sums are ((c0 + c1) + c2) + c3
// with an empty quadrant +0: the dense tile reduction's rule, evaluated only
// where particles are.  Returns (sx, sy, sm) and the particle count n and
// subtree size y.
struct CellSums {
    float sx, sy, sm;
    uint32_t n, y;
};

__device__ __forceinline__ CellSums child_sums(const uint32_t *__restrict__ keys,
                                               const float4 *__restrict__ prod,
                                               const uint32_t *__restrict__ ncount,
                                               const float4 *__restrict__ nsum, uint32_t k,
                                               uint32_t slot, int d, uint32_t n_in,
                                               uint32_t n_nodes) {
    const uint32_t shift = 2u * (kDeepBits - 1u - (uint32_t)d);  // child quadrant bits
    const uint32_t prefix = keys[k] >> (shift + 2u);
    float4 v0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v1 = v0, v2 = v0, v3 = v0;
    uint32_t c = slot + 1u, f = k, n = 0u, y = 1u;
    // at most four children (distinct quadrants)
    for (int it = 0; it < 4 && f < n_in && c < n_nodes; ++it) {
        // one round of loads per child: both candidate sums are read before
        // the count says which one the child is
        const uint32_t key = keys[f];
        const uint32_t cy = ncount[c];
        const float4 p = prod[f];
        const float4 s = nsum[c];
        if ((key >> (shift + 2u)) != prefix) break;
        float4 v;
        uint32_t cn;
        if (cy == 1u) {  // a particle node
            v = make_float4(0.0f + p.x, 0.0f + p.y, 0.0f + p.z, 0.0f);
            cn = 1u;
        } else {
            v = s;
            cn = __float_as_uint(v.w);
        }
        const uint32_t q = (key >> shift) & 3u;
        if (q == 0u) v0 = v;
        else if (q == 1u) v1 = v;
        else if (q == 2u) v2 = v;
        else v3 = v;
        n += cn;
        y += cy;
        f += cn;
        c += cy;
    }
    return CellSums{((v0.x + v1.x) + v2.x) + v3.x, ((v0.y + v1.y) + v2.y) + v3.y,
                    ((v0.z + v1.z) + v2.z) + v3.z, n, y};
}

// One launch per depth d = 14..11: every cell node of depth d (a sorted
// position k with s(k) < d <= s(k+1) starts one) gets its node and its sums.
__global__ __launch_bounds__(256) void bh_level_kernel(
    const uint32_t *__restrict__ keys, const int8_t *__restrict__ sdep,
    const uint32_t *__restrict__ cpos, const float4 *__restrict__ prod,
    const BhCounters *__restrict__ ctr, float4 *__restrict__ node4,
    uint32_t *__restrict__ ncount, float4 *__restrict__ nsum, int d,
    const float *__restrict__ box) {
    const uint32_t n_in = ctr->n_in, n_nodes = ctr->n_nodes;
    const float r0 = box ? box[3] : kFixedR0;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n_in; k += gridDim.x * 256) {
        const int sk = sdep[k];
        const int sk1 = k + 1 < n_in ? (int)sdep[k + 1] : -1;
        if (!(sk < d && d <= sk1)) continue;
        const uint32_t slot = k + cpos[k] + (uint32_t)(d - sk - 1);
        const CellSums c = child_sums(keys, prod, ncount, nsum, k, slot, d, n_in, n_nodes);
        float px = c.sx, py = c.sy;
        if (c.sm != 0.0f) {
            px = px / c.sm;
            py = py / c.sm;
        }
        ncount[slot] = c.y;
        node4[slot] = make_float4(px, py, c.sm, cell_radius(d, r0));
        nsum[slot] = make_float4(c.sx, c.sy, c.sm, __uint_as_float(c.n));
    }
}

// Depth 10 and up with 15-bit leaves: the cells of depth <= 10 are those of a
// 10-bit tree over the same particles, so the dense tile and top kernels
// reduce them once each occupied depth-10 cell (granule) has its record: the
// sums of its depth-10 cell (child_sums) when it holds >= 2 particles, else
// 0 + m*p of its one particle; particle count; subtree size.  Also marks each
// depth-5 tile's granule range for the tile kernel.
__global__ __launch_bounds__(256) void bh_gran_kernel(
    const uint32_t *__restrict__ keys, const int8_t *__restrict__ sdep,
    const uint32_t *__restrict__ cpos, const float4 *__restrict__ prod,
    const BhCounters *__restrict__ ctr, const uint32_t *__restrict__ ncount,
    const float4 *__restrict__ nsum, const uint32_t *__restrict__ gran_first,
    const uint32_t *__restrict__ gran_key, float4 *__restrict__ gran_sum,
    uint32_t *__restrict__ gran_y, uint32_t *__restrict__ tile_range,
    uint32_t *__restrict__ tiles_busy) {
    const uint32_t n_in = ctr->n_in, n_nodes = ctr->n_nodes, ng = ctr->n_gran;
    for (uint32_t L = blockIdx.x * 256 + threadIdx.x; L < ng; L += gridDim.x * 256) {
        const uint32_t k = gran_first[L];
        const int sk = sdep[k];
        const int sk1 = k + 1 < n_in ? (int)sdep[k + 1] : -1;
        if (sk1 >= 10) {
            const uint32_t slot = k + cpos[k] + (uint32_t)(10 - sk - 1);
            const CellSums c = child_sums(keys, prod, ncount, nsum, k, slot, 10, n_in, n_nodes);
            gran_sum[L] = make_float4(c.sx, c.sy, c.sm, __uint_as_float(c.n));
            gran_y[L] = c.y;
        } else {
            const float4 p = prod[k];
            gran_sum[L] = make_float4(0.0f + p.x, 0.0f + p.y, 0.0f + p.z, __uint_as_float(1u));
            gran_y[L] = 1u;
        }
        const uint32_t T = gran_key[L] >> 10;
        if (L == 0 || (gran_key[L - 1] >> 10) != T) {
            tile_range[2 * T] = L;
            atomicAdd(tiles_busy, 1u);
        }
        if (L + 1 == ng || (gran_key[L + 1] >> 10) != T) tile_range[2 * T + 1] = L + 1;
    }
}

// ---- 6./7. dense reduction per tile, then the top ------------------------------

// Level offsets inside `up`: depth 9 at 0 (256 cells), 8 at 256 (64),
// 7 at 320 (16), 6 at 336 (4), 5 at 340 (1).
__device__ __forceinline__ int tile_level_offset(int d) {
    return d == 9 ? 0 : d == 8 ? 256 : d == 7 ? 320 : d == 6 ? 336 : 340;
}

// The last occupied tile's workgroup to finish (a device-scope counter,
// zeroed by the keys kernel, against the number of occupied tiles the leaf
// pass counted) also reduces the 1024 depth-5 cells to the root and writes
// depths 4..0, so the top needs no launch of its own.  Only occupied tiles
// write their record to `top` (fenced before they count); empty tiles return
// at once, and the last workgroup takes their empty record from tile_range.
__global__ __launch_bounds__(256) void bh_tile_kernel(
    const uint32_t *__restrict__ tile_range, const uint32_t *__restrict__ leaf_first,
    const uint32_t *__restrict__ leaf_key, const float4 *__restrict__ leaf_sum,
    const int8_t *__restrict__ sdep, const uint32_t *__restrict__ cpos,
    float4 *__restrict__ node4, uint32_t *__restrict__ ncount, TileCell *__restrict__ top,
    const float *__restrict__ box, const uint32_t *__restrict__ leaf_y,
    BhCounters *__restrict__ ctr) {
    // leaf_y (15-bit leaves): the depth-10 records are granules whose subtree
    // size is given; otherwise leaves of n particles, y = n + [n >= 2]
    __shared__ TileCell l10[1024];
    __shared__ TileCell up[341];
    __shared__ uint32_t last;
    const uint32_t T = blockIdx.x, t = threadIdx.x;
    const uint32_t lb = tile_range[2 * T], le = tile_range[2 * T + 1];
    if (lb == le) return;  // the whole workgroup: an empty tile
    const float r0 = box ? box[3] : kFixedR0;
    for (uint32_t c = t; c < 1024; c += 256)
        l10[c] = TileCell{0u, 0u, 0xffffffffu, 0u, 0.0f, 0.0f, 0.0f, 0.0f};
    lds_barrier();
    // a tile holds at most 1024 leaves (granules): four per thread, loaded
    // together; a 10-bit leaf's own cell node was written with its sum
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t L = lb + t + 256 * u;
        if (L >= le) break;
        const uint32_t c = leaf_key[L] & 1023u;
        const float4 s = leaf_sum[L];
        const uint32_t n = __float_as_uint(s.w);
        const uint32_t y = leaf_y ? leaf_y[L] : n + (n >= 2u ? 1u : 0u);
        TileCell cell{n, y, leaf_first[L], 0u, s.x, s.y, s.z, 0.0f};
        l10[c] = cell;
        if (leaf_y) emit_cell(cell, 10, sdep, cpos, node4, ncount, r0);
    }
    lds_barrier();
    {
        const uint32_t c = t;  // 256 depth-9 cells
        up[c] = reduce4(l10[4 * c], l10[4 * c + 1], l10[4 * c + 2], l10[4 * c + 3]);
    }
    lds_barrier();
    for (int d = 8; d >= 5; --d) {
        const uint32_t cnt = 1u << (2 * (d - 5));
        const int o = tile_level_offset(d), oc = tile_level_offset(d + 1);
        if (t < cnt)
            up[o + t] = reduce4(up[oc + 4 * t], up[oc + 4 * t + 1], up[oc + 4 * t + 2],
                                up[oc + 4 * t + 3]);
        lds_barrier();
    }
    for (uint32_t c = t; c < 341; c += 256) {
        const int d = c < 256 ? 9 : c < 320 ? 8 : c < 336 ? 7 : c < 340 ? 6 : 5;
        emit_cell(up[c], d, sdep, cpos, node4, ncount, r0);
    }
    if (t == 0) {
        top[T] = up[340];
        __threadfence();  // top[T] before the count (release)
        last = atomicAdd(&ctr->tiles_done, 1u) + 1u == ctr->tiles_busy ? 1u : 0u;
    }
    // a full barrier: thread 0's acquire of the other tiles' top[] is handed
    // on to the whole workgroup
    __syncthreads();
    if (!last) return;
    __threadfence();  // every tile's top[] after the count (acquire)
    // depths 4..0 in `up` at the offsets of depths 9..5
    TileCell c4[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t q = 4 * t + u;
        c4[u] = tile_range[2 * q] != tile_range[2 * q + 1]
                    ? top[q]
                    : TileCell{0u, 0u, 0xffffffffu, 0u, 0.0f, 0.0f, 0.0f, 0.0f};
    }
    up[t] = reduce4(c4[0], c4[1], c4[2], c4[3]);
    lds_barrier();
    for (int d = 3; d >= 0; --d) {
        const uint32_t cnt = 1u << (2 * d);
        const int o = tile_level_offset(d + 5), oc = tile_level_offset(d + 6);
        if (t < cnt)
            up[o + t] = reduce4(up[oc + 4 * t], up[oc + 4 * t + 1], up[oc + 4 * t + 2],
                                up[oc + 4 * t + 3]);
        lds_barrier();
    }
    for (uint32_t c = t; c < 341; c += 256) {
        const int d = c < 256 ? 4 : c < 320 ? 3 : c < 336 ? 2 : c < 340 ? 1 : 0;
        emit_cell(up[c], d, sdep, cpos, node4, ncount, r0);
    }
}

// ---- 8. traversal + integrate ----------------------------------------------------

// W = 64: one wave is one group, so the node index is wave-uniform and node
// records arrive through scalar loads.  A particle node (radius 0) is accepted
// by every lane, so a run of consecutive particle nodes (a leaf bucket) is
// streamed as a plain direct sum without votes; cell nodes are voted on with
// r2 = (dx*dx + dy*dy) + eps2 rounded exactly as the reference's shader, so
// the visited and accepted node sets are the reference's.
// The next build's sort input, written by the walk on the fixed domain: the
// cell of the particle's new position (InitShader.comp:73-119), at the
// particle's caller index o = perm[slot] (cells[o] = {cell, slot}), so the
// next build reads cells instead of positions.  cells == nullptr: the next
// build computes the cells itself (adaptive box, first step).
struct NextKeys {
    const uint32_t *perm;  // this step's output order: slot -> caller's index
    uint2 *cells;
    uint32_t deep;  // 30-bit cells
};

__device__ __forceinline__ void write_next_key(const NextKeys &nk, uint32_t slot, float2 p) {
    uint32_t c = nk.deep ? cell30(p.x, p.y) : cell20(p.x, p.y);
    if (c == 0xffffffffu) c = nk.deep ? kKilledKeyDeep : kKilledKey;
    nk.cells[nk.perm[slot]] = make_uint2(c, slot);
}

// One group's walk (all four partials) and integrate.
template <bool kDebugAccel, bool kStaged>
__device__ __forceinline__ void walk_group(
    uint32_t g, uint32_t total, float4 *stage, uint2 *queue, const uint32_t *__restrict__ sorted_src,
    const float2 *__restrict__ pos, const float2 *__restrict__ vel,
    const float4 *__restrict__ node4, const uint32_t *__restrict__ ncount,
    float2 *__restrict__ pos_out,
    float2 *__restrict__ vel_out, float2 *__restrict__ acc_dbg, uint32_t na, float eps2,
    float theta2, float kv, float kp, unsigned long long *__restrict__ gvis,
    unsigned long long *__restrict__ gacc, const NextKeys &nk, uint32_t *__restrict__ gcost = nullptr,
    uint32_t *__restrict__ gchunk = nullptr, const uint8_t *__restrict__ heavy = nullptr) {
    if (heavy && heavy[g]) return;  // walked by bh_traverse_heavy_kernel
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t dst = g * 64 + lane;  // this particle's slot in the next order
    const uint32_t src = sorted_src[dst];
    const bool real = src < na;  // ForceShader.comp:80-92
    float px = 0.0f, py = 0.0f;
    if (real) {
        const float2 p = pos[src];
        px = p.x;
        py = p.y;
    }
    const uint32_t nreal = (uint32_t)__popcll(__ballot(real));
    Partials P;
    unsigned long long vis = 0, acc = 0;
    // start at the issue priority this group reached in the last step
    const uint32_t prio0 = gchunk ? gchunk[g] / kPrioCarry : 0u;
    uint32_t chunks = 0;
    const uint32_t cost = walk_partials<kStaged>(0u, total, stage, queue, node4, ncount, px, py,
                                                 real, nreal, eps2, theta2, 0xfu, P, vis, acc,
                                                 kPrioStepSingle, prio0, &chunks);
    if (gchunk && lane == 0) gchunk[g] = chunks;
    if (lane == 0) {  // this group's visits and acceptances, summed on a dump
        gvis[g] = vis;
        gacc[g] = acc;
    }
    if (real) {
        float ax, ay;
        partial_sum(P, ax, ay);
        const float2 v = vel[src];
        const float nvx = v.x + ax * kv, nvy = v.y + ay * kv;
        const float2 np = make_float2(px + (v.x + nvx) * kp, py + (v.y + nvy) * kp);
        pos_out[dst] = np;
        vel_out[dst] = make_float2(nvx, nvy);
        if (nk.cells) write_next_key(nk, dst, np);
        if (kDebugAccel) acc_dbg[dst] = make_float2(ax, ay);
    }
    if (gcost && lane == 0) gcost[g] = cost;
}

// The static traversal's heavy list from the last step's chunk counts
// (gchunk): with q = floor(4 x chunks / mean) - 4 x split_x, the groups with
// q >= qt go to plan[0, n_heavy) and are flagged in heavy[] (every other flag
// is cleared), qt >= 0 the smallest threshold that lists at most `cap` groups
// (the heaviest, whatever their index); q saturates at 63, and if more than
// `cap` groups reach it the first `cap` of them by index are listed.  With no
// counts yet (the first step) the list is empty.  One workgroup.
__global__ __launch_bounds__(1024) void bh_heavy_plan_kernel(const uint32_t *__restrict__ gchunk,
                                                             uint32_t groups, uint32_t split_x,
                                                             uint32_t cap,
                                                             uint32_t *__restrict__ plan,
                                                             uint8_t *__restrict__ heavy,
                                                             BhCounters *__restrict__ ctr) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t hist[64];
    __shared__ uint32_t qt_sh;
    __shared__ unsigned long long total_sh;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    unsigned long long part = 0;
    for (uint32_t g = t; g < groups; g += 1024) part += gchunk[g];
    if (t == 0) total_sh = 0;
    if (t < 64) hist[t] = 0;
    lds_barrier();
    atomicAdd(&total_sh, part);
    lds_barrier();
    const unsigned long long total = total_sh;
    const bool any = total > 0 && split_x > 0 && cap > 0;
    auto quarter = [&](uint32_t g) -> uint32_t {  // q as above, 64 below the multiple
        const unsigned long long q = 4ull * gchunk[g] * groups / total;
        const unsigned long long lo = 4ull * split_x;
        return q < lo ? 64u : (q - lo > 63ull ? 63u : (uint32_t)(q - lo));
    };
    if (any)
        for (uint32_t g = t; g < groups; g += 1024) {
            const uint32_t q = quarter(g);
            if (q < 64) atomicAdd(&hist[q], 1u);
        }
    lds_barrier();
    if (t == 0) {
        uint32_t qt = 64;  // 64: nothing listed
        if (any) {
            uint32_t above = 0;  // groups with q >= k
            for (int k = 63; k >= 0; --k) {
                above += hist[k];
                if (above > cap && k < 63) break;
                qt = (uint32_t)k;
            }
        }
        qt_sh = qt;
    }
    lds_barrier();
    const uint32_t qt = qt_sh;
    uint32_t carry = 0;  // heavy groups before this chunk of 1024
    for (uint32_t base = 0; base < groups; base += 1024) {
        const uint32_t g = base + t;
        uint32_t q = 64;
        if (g < groups && qt < 64) q = quarter(g);
        const bool h = q < 64 && q >= qt;
        uint32_t x = h ? 1u : 0u;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        lds_barrier();
        uint32_t before = carry + (h ? x - 1u : x);
        uint32_t chunk_total = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            if (w < wave) before += wsum[w];
            chunk_total += wsum[w];
        }
        const bool listed = h && before < cap;
        if (listed) plan[before] = g;
        if (g < groups) heavy[g] = listed ? 1u : 0u;
        carry += chunk_total;
        lds_barrier();  // wsum is rewritten by the next chunk
    }
    if (t == 0) ctr->n_heavy = min(carry, cap);
}

// The W = 32 launch order, by workgroup (four consecutive groups, the launch
// unit): by last step's walk cost (the four groups' real-lane visits, gvis),
// heaviest first, in buckets of a quarter of the mean (the last bucket holds
// everything from 7.75x the mean up) and by index within a bucket, so the
// dense-core groups start in the launch's first round and the light ones
// fill its end.  No costs yet (first step): index order.  Which wave walks a
// group never changes a result.  One workgroup: it is bound by its own
// instruction count, which ordering units of four groups keeps small.
constexpr uint32_t kOrderBuckets = 32;
constexpr uint32_t kOrderAhead = 4;

// The lanes of the wave whose 5-bit bucket equals this lane's (five ballots,
// whatever the number of distinct buckets); q = kOrderBuckets: no unit.
__device__ __forceinline__ unsigned long long bucket_peers(uint32_t q) {
    unsigned long long m = __ballot(q < kOrderBuckets);
#pragma unroll
    for (int bit = 0; bit < 5; ++bit) {
        const unsigned long long b = __ballot((q >> bit) & 1u);
        m &= ((q >> bit) & 1u) ? b : ~b;
    }
    return m;
}

__global__ __launch_bounds__(1024) void bh_order_plan_kernel(
    const unsigned long long *__restrict__ gvis, uint32_t groups, uint32_t *__restrict__ order) {
    __shared__ unsigned long long wsum[16];
    __shared__ uint32_t base[2][kOrderBuckets];
    __shared__ uint32_t wcnt[2][16][kOrderBuckets];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t units = (groups + 3) / 4;
    const unsigned long long below = (1ull << lane) - 1ull;
    // a unit's cost, clamped to 32 bits (a float conversion of one word)
    auto unit_cost = [&](uint32_t u, unsigned long long &sum) -> uint32_t {
        unsigned long long v[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t g = 4 * u + k;
            v[k] = gvis[g < groups ? g : groups - 1];  // unconditional: loads stay together
        }
        unsigned long long c = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) c += 4 * u + k < groups ? v[k] : 0ull;
        sum += u < units ? c : 0ull;
        return c > 0xffffffffull ? 0xffffffffu : (uint32_t)c;
    };
    uint32_t cost[kOrderAhead];  // this thread's units t + 1024 i, i < kOrderAhead
    unsigned long long part = 0;
#pragma unroll
    for (uint32_t i = 0; i < kOrderAhead; ++i) cost[i] = unit_cost(t + 1024 * i, part);
    for (uint32_t u = t + 1024 * kOrderAhead; u < units; u += 1024) (void)unit_cost(u, part);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if (lane == 0) wsum[wave] = part;
    if (t < kOrderBuckets) base[0][t] = 0;
    lds_barrier();
    unsigned long long total = 0;
#pragma unroll
    for (uint32_t w = 0; w < 16; ++w) total += wsum[w];
    // bucket = cost in quarters of the mean unit cost (float: any order is
    // a correct one)
    const float scale = total ? 4.0f * (float)units / (float)total : 0.0f;
    auto bucket_of = [&](uint32_t u, uint32_t c) -> uint32_t {
        if (u >= units) return kOrderBuckets;  // no unit
        const float q = (float)c * scale;
        return q >= (float)(kOrderBuckets - 1) ? kOrderBuckets - 1 : (uint32_t)q;
    };
    auto bucket = [&](uint32_t u, uint32_t i) -> uint32_t {
        if (i < kOrderAhead) return bucket_of(u, cost[i]);
        unsigned long long dummy = 0;
        return bucket_of(u, unit_cost(u, dummy));
    };
    // histogram: the lowest lane of each bucket in the wave adds its lanes
    for (uint32_t i = 0; 1024 * i < units; ++i) {
        const uint32_t q = bucket(t + 1024 * i, i);
        const unsigned long long m = bucket_peers(q);
        if (q < kOrderBuckets && (m & below) == 0ull) atomicAdd(&base[0][q], (uint32_t)__popcll(m));
    }
    lds_barrier();
    if (t == 0) {  // bucket starts, heaviest bucket first
        uint32_t run = 0;
        for (int k = (int)kOrderBuckets - 1; k >= 0; --k) {
            const uint32_t c = base[0][k];
            base[0][k] = run;
            run += c;
        }
    }
    lds_barrier();
    // placement, 1024 units a round; wcnt and base alternate by round, so
    // two barriers a round suffice
    for (uint32_t i = 0; 1024 * i < units; ++i) {
        const uint32_t p = i & 1u;
        const uint32_t u = t + 1024 * i;
        const uint32_t q = bucket(u, i);
        if (lane < kOrderBuckets) wcnt[p][wave][lane] = 0;
        const unsigned long long m = bucket_peers(q);
        const uint32_t rank = (uint32_t)__popcll(m & below);
        // the lowest lane of each bucket writes the wave's count (after the
        // zeroing above: one wave's LDS writes complete in order)
        if (q < kOrderBuckets && rank == 0) wcnt[p][wave][q] = (uint32_t)__popcll(m);
        lds_barrier();
        if (q < kOrderBuckets) {
            uint32_t c[16];
#pragma unroll
            for (uint32_t w = 0; w < 16; ++w) c[w] = wcnt[p][w][q];  // all at once
            uint32_t pos = base[p][q] + rank;
#pragma unroll
            for (uint32_t w = 0; w < 16; ++w) pos += w < wave ? c[w] : 0u;
            order[pos] = u;
        }
        if (t < kOrderBuckets) {
            uint32_t c = base[p][t];
#pragma unroll
            for (uint32_t w = 0; w < 16; ++w) c += wcnt[p][w][t];
            base[p ^ 1u][t] = c;
        }
        lds_barrier();
    }
}

// The heavy list, one workgroup of four waves per group, beside the main
// traversal (which skips these groups): wave 0 decides -- the walk of
// walk_partials, same votes and counts -- and queues the particle runs in the
// workgroup's LDS, then all four waves sum the queued chunks, wave k those of
// partial k, wave 0 also holding the accepted cells; at the end the shares
// (split_share) are added in order by wave 0, which integrates.  The same
// bits as one wave walking the group, with the summing spread over four SIMDs.
template <bool kDebugAccel>
__global__ __launch_bounds__(256) void bh_traverse_heavy_kernel(
    const uint32_t *__restrict__ plan, const BhCounters *__restrict__ ctr,
    const uint32_t *__restrict__ sorted_src, const float2 *__restrict__ pos,
    const float2 *__restrict__ vel, const float4 *__restrict__ node4,
    const uint32_t *__restrict__ ncount, float2 *__restrict__ pos_out,
    float2 *__restrict__ vel_out, float2 *__restrict__ acc_dbg, uint32_t na, uint32_t cap,
    float eps2, float theta2, float kv, float kp, unsigned long long *__restrict__ gvis,
    unsigned long long *__restrict__ gacc, uint32_t *__restrict__ gchunk, NextKeys nk) {
    __shared__ float4 stage[4][64];
    __shared__ uint2 runq[kRunQueue];
    __shared__ uint32_t qs[3];  // queued runs, their first chunk number, walk finished
    __shared__ float2 share[4][64];
    if (blockIdx.x >= ctr->n_heavy) return;  // the whole workgroup
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t g = __builtin_amdgcn_readfirstlane(plan[blockIdx.x]);
    // node count is bounded by the node arrays; every step advances ind by at
    // least one, so the walk ends
    const uint32_t end = __builtin_amdgcn_readfirstlane(min(ctr->n_nodes, cap));
    const uint32_t dst = g * 64 + lane;  // this particle's slot in the next order
    const uint32_t src = sorted_src[dst];
    const bool real = src < na;  // ForceShader.comp:80-92
    float px = 0.0f, py = 0.0f;
    if (real) {
        const float2 p = pos[src];
        px = p.x;
        py = p.y;
    }
    const uint32_t nreal = (uint32_t)__popcll(__ballot(real));
    __builtin_amdgcn_s_setprio(3);  // these groups are the launch's tail
    Partials P;
    unsigned long long vis = 0, acc = 0;
    uint32_t ind = 0, chunk = 0;  // wave 0's walk
    while (true) {
        if (w == 0) {
            const uint32_t qch = chunk;
            uint32_t qn = 0;
            while (ind < end && qn < kRunQueue) {
                const float4 nd = node4[ind];
                const uint32_t cnt = ncount[ind];
                // one round trip per step, as in walk_partials
                asm volatile("" ::"s"(nd.x), "s"(nd.y), "s"(nd.z), "s"(nd.w), "s"(cnt));
                if (nd.w <= 0.0f) {  // a particle node: w = -(particle nodes left in its run)
                    const uint32_t pr = (uint32_t)(-nd.w);
                    const uint32_t n = __builtin_amdgcn_readfirstlane(min(max(pr, 1u), end - ind));
                    vis += (unsigned long long)n * nreal;
                    acc += (unsigned long long)n * nreal;
                    if (lane == 0) runq[qn] = make_uint2(ind, n);
                    ++qn;
                    chunk += (n + 63u) >> 6;
                    ind += n;
                    continue;
                }
                const float dx = nd.x - px;
                const float dy = nd.y - py;
                const float r2 = (dx * dx + dy * dy) + eps2;  // ForceShader.comp:124-125
                const bool ok = (nd.w <= r2 * theta2) || !real;
                vis += nreal;
                if (__ballot(ok) == ~0ull) {
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    const float sc = (nd.z * inv) * (inv * inv);
                    P.xc = __builtin_fmaf(dx, sc, P.xc);
                    P.yc = __builtin_fmaf(dy, sc, P.yc);
                    acc += nreal;
                    ind = __builtin_amdgcn_readfirstlane(ind + max(cnt, 1u));
                } else {
                    ind = __builtin_amdgcn_readfirstlane(ind + 1u);
                }
            }
            if (lane == 0) {
                qs[0] = qn;
                qs[1] = qch;
                qs[2] = ind >= end ? 1u : 0u;
            }
        }
        lds_barrier();
        const uint32_t qn = __builtin_amdgcn_readfirstlane(qs[0]);
        const uint32_t qch = __builtin_amdgcn_readfirstlane(qs[1]);
        const uint32_t done = __builtin_amdgcn_readfirstlane(qs[2]);
        if (qn) sum_run_queue(runq, qn, qch, stage[w], node4, px, py, eps2, 1u << w, P);
        lds_barrier();  // runq and qs are rewritten by the next batch
        if (done) break;
    }
    float ax, ay;
    split_share(P, w, ax, ay);
    share[w][lane] = make_float2(ax, ay);
    lds_barrier();
    __builtin_amdgcn_s_setprio(0);
    if (w != 0) return;
    if (lane == 0) {
        gvis[g] = vis;
        gacc[g] = acc;
        gchunk[g] = chunk;
    }
    if (real) {
        ax = ((share[0][lane].x + share[1][lane].x) + share[2][lane].x) + share[3][lane].x;
        ay = ((share[0][lane].y + share[1][lane].y) + share[2][lane].y) + share[3][lane].y;
        const float2 v = vel[src];
        const float nvx = v.x + ax * kv, nvy = v.y + ay * kv;
        const float2 np = make_float2(px + (v.x + nvx) * kp, py + (v.y + nvy) * kp);
        pos_out[dst] = np;
        vel_out[dst] = make_float2(nvx, nvy);
        if (nk.cells) write_next_key(nk, dst, np);
        if (kDebugAccel) acc_dbg[dst] = make_float2(ax, ay);
    }
}

// Exclusive prefix of (cost + 1) over the groups, and the total at [groups]
// (one workgroup).  The +1 makes the first step, with no costs yet, split the
// groups evenly by count.
__global__ __launch_bounds__(1024) void bh_cost_scan_kernel(const uint32_t *__restrict__ gcost,
                                                            uint32_t groups,
                                                            uint32_t *__restrict__ gpre) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (groups + 1023) / 1024;
    const uint32_t lo = min(groups, t * per), hi = min(groups, lo + per);
    uint32_t sum = 0;
    for (uint32_t g = lo; g < hi; ++g) sum += gcost[g] + 1u;
    part[t] = sum;
    lds_barrier();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        lds_barrier();
        part[t] += v;
        lds_barrier();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (uint32_t g = lo; g < hi; ++g) {
        gpre[g] = run;
        run += gcost[g] + 1u;
    }
    if (t == 1023) gpre[groups] = part[1023];
}

// Cost-balanced traversal: workgroup k of P walks the contiguous groups whose
// cost prefix starts in [k T / P, (k+1) T / P) (T = total cost of the last
// walk), so every workgroup gets the same estimated work while neighbouring
// groups -- which share their near field in the caches -- stay together.
// Dense-core groups (last cost above `split_x` times the mean) would still
// bound the launch alone on one wave (one walks ~5,400 particle nodes at the
// 500k workload), so up to kSplitSlots of them per segment are split into
// four units, one per partial of the canonical summation (walk_partials);
// the four waves leave their partials in LDS and the last to finish adds them
// in order and integrates -- the same bits as one wave walking the group.
// The 16 waves take units from an LDS counter: split parts first, then the
// segment's other groups.  Which wave walks what never changes a result.
// Measured at 500k (two disks, theta 1), traversal ms: static one wave per
// group 0.336, balanced 0.320, balanced with groups above 5x the mean split
// 0.311-0.316 (per-CU work max/mean 1.45 -> 1.13).
constexpr uint32_t kSplitSlots = 16;
constexpr uint32_t kSegList = 1024;

// kSplit = false: every unit is a whole group (one specialised walk in the
// kernel, all four partials on one wave).
template <bool kDebugAccel, bool kSplit>
__global__ __launch_bounds__(1024) void bh_traverse_bal_kernel(
    const uint32_t *__restrict__ sorted_src, const float2 *__restrict__ pos,
    const float2 *__restrict__ vel, const float4 *__restrict__ node4,
    const uint32_t *__restrict__ ncount,
    BhCounters *__restrict__ ctr, float2 *__restrict__ pos_out, float2 *__restrict__ vel_out,
    float2 *__restrict__ acc_dbg, uint32_t na, uint32_t cap, float eps2, float theta2,
    float kv, float kp, const uint32_t *__restrict__ gpre, uint32_t *__restrict__ gcost,
    uint32_t split_x, unsigned long long *__restrict__ gvis,
    unsigned long long *__restrict__ gacc, NextKeys nk) {
    __shared__ float4 stage[16][64];
    __shared__ uint2 runq[16][kRunQueue];
    __shared__ uint32_t seg[4];            // first group, end, next unit, list fill
    __shared__ uint32_t list[kSegList];    // split groups first, then the others
    __shared__ float2 hpart[kSplitSlots][4][64];
    __shared__ uint32_t hdone[kSplitSlots];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t groups = na / 64;
    const unsigned long long total_cost = gpre[groups];
    if (threadIdx.x == 0) {
        const uint32_t tl = (uint32_t)(total_cost * blockIdx.x / gridDim.x);
        const uint32_t th = (uint32_t)(total_cost * (blockIdx.x + 1) / gridDim.x);
        auto first_at_least = [&](uint32_t v) -> uint32_t {  // first g with gpre[g] >= v
            uint32_t a = 0, b = groups;
            while (a < b) {
                const uint32_t m = (a + b) >> 1;
                if (gpre[m] < v) a = m + 1;
                else b = m;
            }
            return a;
        };
        seg[0] = first_at_least(tl);
        seg[1] = blockIdx.x + 1 == gridDim.x ? groups : first_at_least(th);
        seg[2] = 0;
        seg[3] = 0;
    }
    if (threadIdx.x < kSplitSlots) hdone[threadIdx.x] = 0;
    lds_barrier();
    const uint32_t lo = seg[0], hi = seg[1];
    const uint32_t count = hi - lo, listed = min(count, kSegList);
    const uint32_t i = threadIdx.x;  // one listed group per thread (kSegList == blockDim)
    bool split = false;
    if (kSplit && split_x && i < listed &&
        (unsigned long long)(gcost[lo + i] + 1u) * groups > split_x * total_cost) {
        const uint32_t slot = atomicAdd(&seg[3], 1u);
        if (slot < kSplitSlots) {
            list[slot] = lo + i;
            split = true;
        }
    }
    lds_barrier();
    const uint32_t nsplit = min(seg[3], kSplitSlots);
    lds_barrier();
    if (threadIdx.x == 0) seg[3] = nsplit;
    lds_barrier();
    if (i < listed && !split) list[atomicAdd(&seg[3], 1u)] = lo + i;
    lds_barrier();
    const uint32_t total_nodes = __builtin_amdgcn_readfirstlane(min(ctr->n_nodes, cap));
    const uint32_t units = 4 * nsplit + (count - nsplit);
    while (true) {
        uint32_t u = 0;
        if (lane == 0) u = atomicAdd(&seg[2], 1u);
        u = __builtin_amdgcn_readfirstlane(__shfl(u, 0, 64));
        if (u >= units) break;
        // a whole group (all partials), or one partial of a split group; one
        // walk call for both keeps a single copy of the walk in the kernel
        const bool whole = !kSplit || u >= 4 * nsplit;
        const uint32_t v = u - 4 * nsplit + nsplit;
        const uint32_t h = u >> 2, part = u & 3u;
        uint32_t g;
        if (whole) g = v < listed ? list[v] : lo + v;
        else g = list[h];  // h < nsplit <= kSplitSlots
        g = __builtin_amdgcn_readfirstlane(g);
        const uint32_t parts = !kSplit ? 0xfu : whole ? 0xfu : 1u << part;
        const uint32_t dst = g * 64 + lane;  // this particle's slot in the next order
        const uint32_t src = sorted_src[dst];
        const bool real = src < na;  // ForceShader.comp:80-92
        float px = 0.0f, py = 0.0f;
        if (real) {
            const float2 p = pos[src];
            px = p.x;
            py = p.y;
        }
        const uint32_t nreal = (uint32_t)__popcll(__ballot(real));
        Partials P;
        unsigned long long vis = 0, acc = 0;
        const uint32_t cost = walk_partials<true>(0u, total_nodes, stage[w], runq[w], node4, ncount,
                                                  px, py, real, nreal, eps2, theta2, parts, P, vis,
                                                  acc, kPrioStepSingle);
        if (lane == 0 && (parts & 1u)) {  // the counting wave of the group
            gvis[g] = vis;
            gacc[g] = acc;
        }
        float ax = 0.0f, ay = 0.0f;
        if (whole) {
            partial_sum(P, ax, ay);
            if (lane == 0) gcost[g] = cost;
        } else {
            split_share(P, part, ax, ay);
            hpart[h][part][lane] = make_float2(ax, ay);
            if (part == 0 && lane == 0) gcost[g] = cost;
            __threadfence_block();
            uint32_t done = 0;
            if (lane == 0) done = atomicAdd(&hdone[h], 1u);
            done = __builtin_amdgcn_readfirstlane(__shfl(done, 0, 64));
            if (done != 3u) continue;
            __threadfence_block();  // the last of the four adds the shares in order
            ax = ((hpart[h][0][lane].x + hpart[h][1][lane].x) + hpart[h][2][lane].x) +
                 hpart[h][3][lane].x;
            ay = ((hpart[h][0][lane].y + hpart[h][1][lane].y) + hpart[h][2][lane].y) +
                 hpart[h][3][lane].y;
        }
        if (real) {
            const float2 vv = vel[src];
            const float nvx = vv.x + ax * kv, nvy = vv.y + ay * kv;
            const float2 np = make_float2(px + (vv.x + nvx) * kp, py + (vv.y + nvy) * kp);
            pos_out[dst] = np;
            vel_out[dst] = make_float2(nvx, nvy);
            if (nk.cells) write_next_key(nk, dst, np);
            if (kDebugAccel) acc_dbg[dst] = make_float2(ax, ay);
        }
    }
}

// kPersistent: waves take groups from a device counter until all na/64 are
// done (every wave leaves once the counter passes the last group); otherwise
// wave i walks group i.
template <bool kDebugAccel, bool kStaged, bool kPersistent>
__global__ __launch_bounds__(256) void bh_traverse64_kernel(
    const uint32_t *__restrict__ sorted_src, const float2 *__restrict__ pos,
    const float2 *__restrict__ vel, const float4 *__restrict__ node4,
    const uint32_t *__restrict__ ncount,
    BhCounters *__restrict__ ctr, float2 *__restrict__ pos_out, float2 *__restrict__ vel_out,
    float2 *__restrict__ acc_dbg, uint32_t na, uint32_t cap, float eps2, float theta2,
    float kv, float kp, unsigned long long *__restrict__ gvis,
    unsigned long long *__restrict__ gacc, uint32_t *__restrict__ gchunk,
    const uint8_t *__restrict__ heavy, uint32_t xcd, NextKeys nk,
    const uint32_t *__restrict__ order) {
    __shared__ float4 stage[kStaged ? 4 : 1][64];
    __shared__ uint2 runq[kStaged ? 4 : 1][kStaged ? kRunQueue : 1];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t groups = na / 64;
    // node count is bounded by the node arrays; every iteration advances ind by
    // at least one, so each walk ends
    const uint32_t total = __builtin_amdgcn_readfirstlane(min(ctr->n_nodes, cap));
    float4 *st = stage[kStaged ? w : 0];
    uint2 *rq = runq[kStaged ? w : 0];
    if (kPersistent) {
        while (true) {
            uint32_t g = 0;
            if (lane == 0) g = atomicAdd(&ctr->next_group, 1u);
            g = __builtin_amdgcn_readfirstlane(__shfl(g, 0, 64));
            if (g >= groups) break;
            walk_group<kDebugAccel, kStaged>(g, total, st, rq, sorted_src, pos, vel, node4,
                                             ncount, pos_out, vel_out, acc_dbg, na, eps2,
                                             theta2, kv, kp, gvis, gacc, nk, nullptr, gchunk);
        }
    } else {
        const uint32_t b0 = xcd_block(blockIdx.x, gridDim.x, xcd);
        // order: a permutation of the gridDim.x workgroups' units
        const uint32_t b = order ? __builtin_amdgcn_readfirstlane(order[b0]) : b0;
        const uint32_t g = b * 4 + w;
        if (g >= groups) return;  // whole wave beyond the last group
        walk_group<kDebugAccel, kStaged>(g, total, st, rq, sorted_src, pos, vel, node4, ncount,
                                         pos_out, vel_out, acc_dbg, na, eps2, theta2, kv,
                                         kp, gvis, gacc, nk, nullptr, gchunk, heavy);
    }
}

// W = 32 (the published runs' subgroup width): one 32-particle group per
// wave, particle slot g * 32 + l held by lanes l and l + 32, walked by the
// staged W = 64 walk (both halves vote alike, so the 64-lane vote is the
// group's: the same visits and acceptances); each half sums its part of every
// chunk (sum_run_queue).  Per partial k the halves' sums are added, then the
// acceleration is (((c + p0) + p1) + p2) + p3 as at W = 64.
template <bool kDebugAccel>
__global__ __launch_bounds__(256) void bh_traverse32_kernel(
    const uint32_t *__restrict__ sorted_src, const float2 *__restrict__ pos,
    const float2 *__restrict__ vel, const float4 *__restrict__ node4,
    const uint32_t *__restrict__ ncount, BhCounters *__restrict__ ctr,
    float2 *__restrict__ pos_out, float2 *__restrict__ vel_out, float2 *__restrict__ acc_dbg,
    uint32_t na, uint32_t cap, float eps2, float theta2, float kv, float kp,
    unsigned long long *__restrict__ gvis, unsigned long long *__restrict__ gacc, uint32_t xcd,
    NextKeys nk, const uint32_t *__restrict__ order) {
    __shared__ float4 stage[4][64];
    __shared__ uint2 runq[4][kRunQueue];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t b0 = xcd_block(blockIdx.x, gridDim.x, xcd);
    // order: a permutation of the gridDim.x workgroups' units
    const uint32_t b = order ? __builtin_amdgcn_readfirstlane(order[b0]) : b0;
    const uint32_t g = b * 4 + w;
    if (g >= na / 32) return;  // whole wave beyond the last group
    // node count is bounded by the node arrays; every step advances the walk
    const uint32_t total = __builtin_amdgcn_readfirstlane(min(ctr->n_nodes, cap));
    const uint32_t dst = g * 32 + (lane & 31u);  // this particle's slot in the next order
    const uint32_t src = sorted_src[dst];
    const bool real = src < na;  // ForceShader.comp:80-92
    float px = 0.0f, py = 0.0f;
    if (real) {
        const float2 p = pos[src];
        px = p.x;
        py = p.y;
    }
    const uint32_t nreal = (uint32_t)__popcll(__ballot(real) & 0xffffffffull);
    Partials P;
    unsigned long long vis = 0, acc = 0;
    (void)walk_partials<true, true>(0u, total, stage[w], runq[w], node4, ncount, px, py, real,
                                    nreal, eps2, theta2, 0xfu, P, vis, acc, kPrioStepSingle);
    P.x0 += __shfl_xor(P.x0, 32, 64);
    P.y0 += __shfl_xor(P.y0, 32, 64);
    P.x1 += __shfl_xor(P.x1, 32, 64);
    P.y1 += __shfl_xor(P.y1, 32, 64);
    P.x2 += __shfl_xor(P.x2, 32, 64);
    P.y2 += __shfl_xor(P.y2, 32, 64);
    P.x3 += __shfl_xor(P.x3, 32, 64);
    P.y3 += __shfl_xor(P.y3, 32, 64);
    if (lane == 0) {  // this group's visits and acceptances, summed on a dump
        gvis[g] = vis;
        gacc[g] = acc;
    }
    if (real && lane < 32) {
        float ax, ay;
        partial_sum(P, ax, ay);
        const float2 v = vel[src];
        const float nvx = v.x + ax * kv, nvy = v.y + ay * kv;
        const float2 np = make_float2(px + (v.x + nvx) * kp, py + (v.y + nvy) * kp);
        pos_out[dst] = np;
        vel_out[dst] = make_float2(nvx, nvy);
        if (nk.cells) write_next_key(nk, dst, np);
        if (kDebugAccel) acc_dbg[dst] = make_float2(ax, ay);
    }
}

// 30-bit Morton codes (15 bits per axis) of the same normalised coordinate;
// code >> 10 equals the 20-bit cell because scaling by 32 is exact.
__global__ __launch_bounds__(256) void bh_morton30_kernel(const float2 *__restrict__ pos,
                                                          uint32_t na, uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= na) return;
    const float2 p = pos[i];
    const float qx = unit_coord(p.x) * 32768.0f, qy = unit_coord(p.y) * 32768.0f;
    uint32_t c = 0xffffffffu;
    if (qx >= 0.0f && qx < 32768.0f && qy >= 0.0f && qy < 32768.0f)
        c = part1by1((uint32_t)qx) | (part1by1((uint32_t)qy) << 1);
    out[i] = c;
}

}  // namespace

size_t bh_capacity(size_t na, uint32_t bits) {
    // every particle plus at most min(4^d, na/2) cells with >= 2 particles per level
    size_t cells = 0;
    for (int d = 0; d <= (int)bits; ++d) {
        const size_t lvl = (size_t)1 << (2 * d);
        cells += lvl < na / 2 ? lvl : na / 2;
    }
    return na + cells + 64;
}

hipError_t bh_alloc(BhBuffers *b, size_t na, uint32_t bits) {
    b->na = na;
    b->bits = bits;
    b->cap = bh_capacity(na, bits);
    b->sort_blocks = (na + kSortTile - 1) / kSortTile;
    b->prep_blocks = (na + 255) / 256;
    hipError_t e = hipSuccess;
#define GSIM_ALLOC(ptr, count)                                                  \
    do {                                                                        \
        if (e == hipSuccess) e = hipMalloc((void **)&(ptr), sizeof(*(ptr)) * (count)); \
    } while (0)
    for (int k = 0; k < 2; ++k) {
        GSIM_ALLOC(b->keys[k], na);
        GSIM_ALLOC(b->idx[k], na);
        GSIM_ALLOC(b->cells[k], na);
    }
    // four passes of tile words and group words
    GSIM_ALLOC(b->status, 4 * 256 * (b->sort_blocks + (b->sort_blocks + kLookGroup - 1) / kLookGroup));
    GSIM_ALLOC(b->digit_total, 2 * kDigitCopies * 1024);
    GSIM_ALLOC(b->breaker, b->cap);
    GSIM_ALLOC(b->block_cells, b->prep_blocks);
    GSIM_ALLOC(b->block_leaves, b->prep_blocks);
    GSIM_ALLOC(b->sdep, na + 1);
    GSIM_ALLOC(b->cpos, na);
    GSIM_ALLOC(b->prod, na);
    GSIM_ALLOC(b->leaf_first, na + 1);
    GSIM_ALLOC(b->leaf_key, na);
    GSIM_ALLOC(b->leaf_sum, na);
    GSIM_ALLOC(b->tile_range, 2 * kTiles);
    GSIM_ALLOC(b->top, kTiles);
    GSIM_ALLOC(b->node4, b->cap);
    GSIM_ALLOC(b->ncount, b->cap);
    if (bits == kDeepBits) {
        GSIM_ALLOC(b->nsum, b->cap);
        GSIM_ALLOC(b->gran_first, na + 1);
        GSIM_ALLOC(b->gran_key, na);
        GSIM_ALLOC(b->gran_y, na);
        GSIM_ALLOC(b->block_gran, b->prep_blocks);
    }
    GSIM_ALLOC(b->sorted_src, na);
    GSIM_ALLOC(b->ctr, 1);
    GSIM_ALLOC(b->box, 4);
    GSIM_ALLOC(b->box_part, kBoxParts);
    GSIM_ALLOC(b->gcost, na / 64);
    GSIM_ALLOC(b->gchunk, na / 64 + 1);
    GSIM_ALLOC(b->plan, kMaxHeavy);
    GSIM_ALLOC(b->order, na / 128 + 1);
    GSIM_ALLOC(b->heavy, na / 64 + 1);
    GSIM_ALLOC(b->gvis, na / 32 + 1);  // one slot per group of W = 32 or 64
    GSIM_ALLOC(b->gacc, na / 32 + 1);
    GSIM_ALLOC(b->gpre, na / 64 + 1);
#undef GSIM_ALLOC
    if (e == hipSuccess) e = hipMemset(b->ctr, 0, sizeof(BhCounters));
    if (e == hipSuccess) e = hipMemset(b->digit_total, 0, sizeof(uint32_t) * 2 * kDigitCopies * 1024);
    if (e == hipSuccess) e = hipMemset(b->gcost, 0, sizeof(uint32_t) * (na / 64));
    if (e == hipSuccess) e = hipMemset(b->gchunk, 0, sizeof(uint32_t) * (na / 64 + 1));
    if (e == hipSuccess) e = hipMemset(b->gvis, 0, sizeof(unsigned long long) * (na / 32 + 1));
    if (e == hipSuccess) e = hipMemset(b->gacc, 0, sizeof(unsigned long long) * (na / 32 + 1));
    if (e == hipSuccess) e = hipMemset(b->ncount, 0, sizeof(uint32_t) * b->cap);
    if (e == hipSuccess) e = hipMemset(b->node4, 0, sizeof(float4) * b->cap);
    return e;
}

void bh_free(BhBuffers *b) {
    for (int k = 0; k < 2; ++k) {
        (void)hipFree(b->keys[k]);
        (void)hipFree(b->idx[k]);
        (void)hipFree(b->cells[k]);
    }
    void *ptrs[] = {b->status, b->digit_total, b->breaker, b->block_cells, b->block_leaves, b->sdep, b->cpos, b->prod,
                    b->leaf_first, b->leaf_key, b->leaf_sum, b->tile_range, b->top,
                    b->node4, b->ncount, b->nsum, b->gran_first, b->gran_key, b->gran_y,
                    b->block_gran, b->sorted_src, b->ctr, b->gcost, b->gchunk, b->plan, b->order, b->heavy, b->gpre, b->gvis, b->gacc, b->box,
                    b->box_part};
    for (void *p : ptrs) (void)hipFree(p);
    *b = BhBuffers{};
}

// Where this step's walk writes the next build's keys (the buffer the current
// build does not read), or none when the next build must compute them.
static NextKeys next_keys(gsim_sim *s, const uint32_t *perm_out) {
    BhBuffers &b = s->bh;
    if (s->bh_box || !perm_out) return NextKeys{nullptr, nullptr, 0u};
    return NextKeys{perm_out, b.cells[b.kin ^ 1], b.bits == kDeepBits ? 1u : 0u};
}

hipError_t bh_launch_build(gsim_sim *s, const float2 *pos, const float *mass,
                           const uint32_t *perm_in, float2 *pos_out, const float2 *vel,
                           float2 *vel_out, float *mass_out, uint32_t *perm_out) {
    BhBuffers &b = s->bh;
    const uint32_t na = (uint32_t)s->na;
    const uint32_t g256 = (na + 255) / 256;
    const bool deep = b.bits == kDeepBits;
    const int bits = (int)b.bits;
    const float *box = nullptr;
    if (s->bh_box) {
        const uint32_t nparts = std::min<uint32_t>(g256, kBoxParts);
        bh_box_partial_kernel<<<nparts, 256, 0, s->stream>>>(pos, na, b.box_part);
        bh_box_final_kernel<<<1, 256, 0, s->stream>>>(b.box_part, nparts, b.box);
        box = b.box;
    }
    // one tile of kSortTile keys per workgroup, in the keys kernel and in
    // every pass; each pass's status holds its tile words, then its group words
    const uint32_t nblk = (na + kSortTile - 1) / kSortTile;
    const uint32_t kblk = nblk;
    const uint32_t pass_words = (nblk + (nblk + kLookGroup - 1) / kLookGroup) * 256u;
    const int passes = deep ? 4 : 3;
    const uint32_t status_words = (uint32_t)passes * pass_words;
    // digit totals alternate between two halves by build parity
    uint32_t *const totals = b.digit_total + kDigitCopies * 1024 * b.parity;
    uint32_t *const totals_next = b.digit_total + kDigitCopies * 1024 * (b.parity ^ 1u);
    b.parity ^= 1u;
    // the sort's input: written by the last walk (fixed domain), else here
    const int kin = b.kin;
    const bool walk = b.walk_keys && !box;
    b.kcells = kin;
#define GSIM_KEYS(B, D, W)                                                                    \
    bh_keys_kernel<B, D, kSortRounds, W><<<kblk, kSortThreads, 0, s->stream>>>(                \
        pos, perm_in, na, b.cells[kin], b.tile_range, totals, totals_next, b.status,          \
        status_words, b.ctr, box)
    if (box) { if (deep) GSIM_KEYS(true, true, false); else GSIM_KEYS(true, false, false); }
    else if (walk) { if (deep) GSIM_KEYS(false, true, true); else GSIM_KEYS(false, false, true); }
    else { if (deep) GSIM_KEYS(false, true, false); else GSIM_KEYS(false, false, false); }
#undef GSIM_KEYS
    // 20-bit keys: cells[kin] -> keys[0] -> keys[1] -> keys[0]; 30-bit keys take
    // a fourth pass -> keys[1].  cells[kin] keeps the per-particle cells.
    const int in_of[4] = {-1, 0, 1, 0}, out_of[4] = {0, 1, 0, 1};
    for (int pass = 0; pass < passes; ++pass) {
        const int i = in_of[pass], o = out_of[pass];
        uint32_t *const st = b.status + (size_t)pass * pass_words;
        radix_onesweep_kernel<kSortRounds><<<nblk, kSortThreads, 0, s->stream>>>(
            i < 0 ? nullptr : b.keys[i], i < 0 ? nullptr : b.idx[i], i < 0 ? b.cells[kin] : nullptr,
            b.keys[o], b.idx[o], na, 8u * pass, totals + 256 * pass, st,
            &b.ctr->sort_stall);
    }
    b.sorted_slot = deep ? 1 : 0;
    const uint32_t *sk = b.keys[b.sorted_slot], *si = b.idx[b.sorted_slot];
    const uint32_t pb = (uint32_t)b.prep_blocks;
    bh_prep_kernel<<<pb, 256, 0, s->stream>>>(sk, na, bits, b.block_cells, b.block_leaves,
                                              deep ? b.block_gran : nullptr);
    const NextKeys nk = next_keys(s, perm_out);
    bh_emit_kernel<<<pb, 256, 0, s->stream>>>(sk, si, pos, mass, vel, perm_in, pos_out, vel_out,
                                              mass_out, perm_out, nk.cells, na, bits,
                                              b.block_cells,
                                              b.block_leaves, b.sdep, b.cpos, b.prod,
                                              b.leaf_first, b.leaf_key, b.node4, b.ncount,
                                              b.sorted_src, b.breaker, b.ctr,
                                              deep ? b.block_gran : nullptr,
                                              deep ? b.gran_first : nullptr, b.gran_key);
    const uint32_t lb = pb < kLeafBlocks ? pb : kLeafBlocks;
    bh_leaf_kernel<<<lb, 256, 0, s->stream>>>(b.leaf_first, b.leaf_key, b.prod, sk, na, bits,
                                              b.cpos, b.breaker, b.ctr, b.leaf_sum, b.tile_range,
                                              b.sdep, b.node4, b.ncount,
                                              deep ? b.nsum : nullptr, box, &b.ctr->tiles_busy);
    if (deep) {
        // sparse levels below depth 10, then depth-10 records for the dense path
        for (int d = (int)kDeepBits - 1; d > (int)kTreeDepth; --d)
            bh_level_kernel<<<pb, 256, 0, s->stream>>>(sk, b.sdep, b.cpos, b.prod, b.ctr, b.node4,
                                                       b.ncount, b.nsum, d, box);
        bh_gran_kernel<<<pb, 256, 0, s->stream>>>(sk, b.sdep, b.cpos, b.prod, b.ctr, b.ncount,
                                                  b.nsum, b.gran_first, b.gran_key, b.leaf_sum,
                                                  b.gran_y, b.tile_range, &b.ctr->tiles_busy);
        bh_tile_kernel<<<kTiles, 256, 0, s->stream>>>(b.tile_range, b.gran_first, b.gran_key,
                                                      b.leaf_sum, b.sdep, b.cpos, b.node4,
                                                      b.ncount, b.top, box, b.gran_y, b.ctr);
    } else {
        bh_tile_kernel<<<kTiles, 256, 0, s->stream>>>(b.tile_range, b.leaf_first, b.leaf_key,
                                                      b.leaf_sum, b.sdep, b.cpos, b.node4,
                                                      b.ncount, b.top, box, nullptr, b.ctr);
    }
    return hipGetLastError();
}

// The static walk's workgroup order (null: index order), re-planned from the
// last walk's costs every kOrderReplan steps (costs change slowly) and on
// every trial step, on the walk's own stream: a plan on the side stream
// beside the build measured slower (the cross-stream fork and join cost more
// than the plan).
static const uint32_t *launch_order(gsim_sim *s, uint32_t groups) {
    BhBuffers &b = s->bh;
    if (!s->trav_order) return nullptr;
    if (!b.order_valid || s->order_fresh || s->steps % kOrderReplan == 0) {
        bh_order_plan_kernel<<<1, 1024, 0, s->stream>>>(b.gvis, groups, b.order);
        b.order_valid = s->steps > 0;  // the first step's plan has no costs yet
    }
    return b.order;
}

hipError_t bh_launch_traverse(gsim_sim *s, const float2 *pos, const float2 *vel,
                              float2 *pos_out, float2 *vel_out, const uint32_t *perm_out) {
    BhBuffers &b = s->bh;
    const uint32_t na = (uint32_t)s->na;
    const uint32_t grid = (na + 255) / 256;
    const NextKeys nk = next_keys(s, perm_out);
    // the next build reads the cells this walk writes
    b.walk_keys = nk.cells != nullptr;
    if (b.walk_keys) b.kin ^= 1;
    const bool dbg = (s->debug_flags & GSIM_DEBUG_ACCEL) != 0;
    float2 *acc = dbg ? s->accel : nullptr;
    if (s->W == 64) {
        const uint32_t groups = na / 64;
        if (s->trav_mode & 4u) {
            bh_cost_scan_kernel<<<1, 1024, 0, s->stream>>>(b.gcost, groups, b.gpre);
            const uint32_t P = s->trav_segments;
#define GSIM_BAL(D, S)                                                                          \
    bh_traverse_bal_kernel<D, S><<<P, 1024, 0, s->stream>>>(                                   \
        b.sorted_src, pos, vel, b.node4, b.ncount, b.ctr, pos_out, vel_out, acc, na,          \
        (uint32_t)b.cap, s->eps2, s->theta2, s->kv, s->kp, b.gpre, b.gcost, s->trav_split,   \
        b.gvis, b.gacc, nk)
            if (dbg) { if (s->trav_split) GSIM_BAL(true, true); else GSIM_BAL(true, false); }
            else { if (s->trav_split) GSIM_BAL(false, true); else GSIM_BAL(false, false); }
#undef GSIM_BAL
            return hipGetLastError();
        }
        const bool persistent = (s->trav_mode & 1u) != 0, staged = (s->trav_mode & 2u) != 0;
        uint32_t g = grid;
        // static staged launch with more groups than the GPU holds at once:
        // last step's heaviest groups go to four-wave workgroups on the side
        // stream (forked after the plan, joined back below), beside the main
        // launch, whose later rounds leave them room.  When every group is
        // resident from the start the side workgroups only find room at the
        // end and lengthen the launch (two disks at 500k: +5%), so they stay
        // off there.
        // The heavy list and the heaviest-first launch order (below) answer
        // the same tail; with the order active the list only lengthened the
        // walk (1M, W = 64: two disks 0.91 -> 0.95, galaxy collision 1.21 ->
        // 1.23 ms per step), so it runs with index order only.
        const uint8_t *heavy = nullptr;
        if (!persistent && staged && s->trav_split && s->side && groups > kHeavyMinGroups &&
            !s->trav_order) {
            bh_heavy_plan_kernel<<<1, 1024, 0, s->stream>>>(b.gchunk, groups, s->trav_split,
                                                            kMaxHeavy, b.plan, b.heavy, b.ctr);
            hipError_t e = hipEventRecord(s->ev_fork, s->stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(s->side, s->ev_fork, 0);
            if (e != hipSuccess) return e;
#define GSIM_HEAVY(D)                                                                        \
    bh_traverse_heavy_kernel<D><<<kMaxHeavy, 256, 0, s->side>>>(                              \
        b.plan, b.ctr, b.sorted_src, pos, vel, b.node4, b.ncount, pos_out, vel_out, acc, na, \
        (uint32_t)b.cap, s->eps2, s->theta2, s->kv, s->kp, b.gvis, b.gacc, b.gchunk, nk)
            if (dbg) GSIM_HEAVY(true);
            else GSIM_HEAVY(false);
#undef GSIM_HEAVY
            heavy = b.heavy;
        }
        const uint32_t *order = persistent ? nullptr : launch_order(s, groups);
        if (persistent) {
            // waves_per_simd waves on each of the 1024 SIMDs, at most one per group
            uint32_t waves = 1024u * s->trav_waves_per_simd;
            if (waves > groups) waves = groups;
            g = (waves + 3) / 4;
        }
#define GSIM_TRAV(D, S, P)                                                                 \
    bh_traverse64_kernel<D, S, P><<<g, 256, 0, s->stream>>>(                               \
        b.sorted_src, pos, vel, b.node4, b.ncount, b.ctr, pos_out, vel_out, acc, na, \
        (uint32_t)b.cap, s->eps2, s->theta2, s->kv, s->kp, b.gvis, b.gacc, b.gchunk, heavy, \
        s->xcd_order, nk, order)
        if (dbg) {
            if (persistent) { if (staged) GSIM_TRAV(true, true, true); else GSIM_TRAV(true, false, true); }
            else { if (staged) GSIM_TRAV(true, true, false); else GSIM_TRAV(true, false, false); }
        } else {
            if (persistent) { if (staged) GSIM_TRAV(false, true, true); else GSIM_TRAV(false, false, true); }
            else { if (staged) GSIM_TRAV(false, true, false); else GSIM_TRAV(false, false, false); }
        }
#undef GSIM_TRAV
        if (heavy) {  // the next launch on the stream waits for the heavy list too
            hipError_t e = hipEventRecord(s->ev_join, s->side);
            if (e == hipSuccess) e = hipStreamWaitEvent(s->stream, s->ev_join, 0);
            if (e != hipSuccess) return e;
        }
    } else {
        const uint32_t g32 = (na / 32 + 3) / 4;  // four 32-particle groups per workgroup
        const uint32_t *order = launch_order(s, na / 32);
        if (dbg)
            bh_traverse32_kernel<true><<<g32, 256, 0, s->stream>>>(
                b.sorted_src, pos, vel, b.node4, b.ncount, b.ctr, pos_out, vel_out, acc, na,
                (uint32_t)b.cap, s->eps2, s->theta2, s->kv, s->kp, b.gvis, b.gacc, s->xcd_order,
                nk, order);
        else
            bh_traverse32_kernel<false><<<g32, 256, 0, s->stream>>>(
                b.sorted_src, pos, vel, b.node4, b.ncount, b.ctr, pos_out, vel_out, acc, na,
                (uint32_t)b.cap, s->eps2, s->theta2, s->kv, s->kp, b.gvis, b.gacc, s->xcd_order,
                nk, order);
    }
    return hipGetLastError();
}

hipError_t bh_launch_morton30(gsim_sim *s, const float2 *pos, uint32_t *out) {
    const uint32_t na = (uint32_t)s->na;
    bh_morton30_kernel<<<(na + 255) / 256, 256, 0, s->stream>>>(pos, na, out);
    return hipGetLastError();
}

}  // namespace gsim
