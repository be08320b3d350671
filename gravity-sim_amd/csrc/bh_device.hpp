// Device helpers shared by the single-system and batched Barnes-Hut kernels.
#pragma once

#include "internal.hpp"

namespace gsim {

__device__ __forceinline__ uint32_t part1by1(uint32_t v) {
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// (p / S + 1) * 0.5, the normalised coordinate of InitShader.comp:80.
// Division is IEEE (no reciprocal contraction); the file is built with
// -ffp-contract=off, so each operation rounds as written.
__device__ __forceinline__ float unit_coord(float p) { return ((p / kSimSize) + 1.0f) * 0.5f; }

// A workgroup barrier that orders LDS only: __syncthreads() also waits for
// the wave's outstanding global stores, which costs a store's latency at
// every barrier that follows one.  For barriers that guard LDS data alone.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint32_t cell20(float x, float y) {
    const float qx = unit_coord(x) * (float)kTreeSize;
    const float qy = unit_coord(y) * (float)kTreeSize;
    if (qx >= 0.0f && qx < (float)kTreeSize && qy >= 0.0f && qy < (float)kTreeSize)
        return part1by1((uint32_t)qx) | (part1by1((uint32_t)qy) << 1);
    return 0xffffffffu;
}

// Depth of the deepest cell shared by two in-domain cells (0..10).
__device__ __forceinline__ int shared_depth(uint32_t a, uint32_t b) {
    const uint32_t x = a ^ b;
    if (x == 0) return 10;
    return 9 - ((31 - __clz(x)) >> 1);
}

// The same for keys of `bits` bits per axis (0..bits).
__device__ __forceinline__ int shared_depth_bits(uint32_t a, uint32_t b, int bits) {
    const uint32_t x = a ^ b;
    if (x == 0) return bits;
    return bits - 1 - ((31 - __clz(x)) >> 1);
}

// 15-bit cells (30-bit keys): the normalised coordinate of cell20 scaled by
// 32768 instead of 1024.  Scaling by a power of two is exact, so key >> 10 is
// the 20-bit cell.
__device__ __forceinline__ uint32_t cell30(float x, float y) {
    const float qx = unit_coord(x) * 32768.0f, qy = unit_coord(y) * 32768.0f;
    if (qx >= 0.0f && qx < 32768.0f && qy >= 0.0f && qy < 32768.0f)
        return part1by1((uint32_t)qx) | (part1by1((uint32_t)qy) << 1);
    return 0xffffffffu;
}

// Workgroups of a launch are dispatched to the 8 XCDs round-robin (workgroup
// b to XCD b mod 8), and each XCD has its own L2.  xcd_block renumbers them so
// that XCD x runs runs of c consecutive logical workgroups (logical blocks
// [8ck + cx, 8ck + cx + c) for k = 0, 1, ...), spatially adjacent groups,
// which walk nearly the same cells and leaf runs, then sharing an L2; the
// last nb mod 8c workgroups keep their number.  From 8c >= nb on, each XCD
// takes one contiguous range (q or q + 1 workgroups for nb = 8q + r).
// c <= 1 is the plain round-robin order.  A bijection on [0, nb).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb, uint32_t c) {
    if (c <= 1u) return b;
    const uint32_t x = b & 7u, i = b >> 3;
    if (8u * c >= nb) {
        const uint32_t q = nb >> 3, r = nb & 7u;
        return x * q + min(x, r) + i;
    }
    const uint32_t span = 8u * c, whole = nb / span * span;
    if (b >= whole) return b;
    return (i / c) * span + x * c + i % c;
}

// Radius base of the fixed domain: S*S*2 (TreeMoveShader.comp:70-71).
constexpr float kFixedR0 = kSimSize * kSimSize * 2.0f;

// r0 / 4^d: S*S*2 / 4^d for the fixed domain, (h*h)*2 / 4^d for the adaptive
// box; exact in FP32 (division by a power of two).
__device__ __forceinline__ float cell_radius(int d, float r0) {
    return r0 / (float)(1u << (2 * d));
}

// Adaptive box {cx, cy, h, r0}: ((p - c) / h + 1) * 0.5 * 1024, clamped to
// [0, 1023]; only non-finite positions are outside.
__device__ __forceinline__ uint32_t cell20_box(float x, float y, float cx, float cy, float h) {
    if (!__builtin_isfinite(x) || !__builtin_isfinite(y)) return 0xffffffffu;
    const float qx = ((((x - cx) / h) + 1.0f) * 0.5f) * (float)kTreeSize;
    const float qy = ((((y - cy) / h) + 1.0f) * 0.5f) * (float)kTreeSize;
    const uint32_t ix = min(qx > 0.0f ? (uint32_t)qx : 0u, kTreeSize - 1);
    const uint32_t iy = min(qy > 0.0f ? (uint32_t)qy : 0u, kTreeSize - 1);
    return part1by1(ix) | (part1by1(iy) << 1);
}

// cell20_box on the 32768^2 grid (clamped to [0, 32767]).
__device__ __forceinline__ uint32_t cell30_box(float x, float y, float cx, float cy, float h) {
    if (!__builtin_isfinite(x) || !__builtin_isfinite(y)) return 0xffffffffu;
    const float qx = ((((x - cx) / h) + 1.0f) * 0.5f) * 32768.0f;
    const float qy = ((((y - cy) / h) + 1.0f) * 0.5f) * 32768.0f;
    const uint32_t ix = min(qx > 0.0f ? (uint32_t)qx : 0u, 32767u);
    const uint32_t iy = min(qy > 0.0f ? (uint32_t)qy : 0u, 32767u);
    return part1by1(ix) | (part1by1(iy) << 1);
}

__device__ __forceinline__ void accumulate(float nx, float ny, float m, float px, float py,
                                           float eps2, float &ax, float &ay) {
    const float dx = nx - px, dy = ny - py;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, eps2));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float s = (m * inv) * (inv * inv);
    ax = __builtin_fmaf(dx, s, ax);
    ay = __builtin_fmaf(dy, s, ay);
}

__device__ __forceinline__ TileCell reduce4(const TileCell &a, const TileCell &b,
                                            const TileCell &c, const TileCell &d) {
    TileCell r;
    r.n = a.n + b.n + c.n + d.n;
    r.y = a.y + b.y + c.y + d.y;
    r.y += r.y > 1u ? 1u : 0u;  // TreeInitShader.comp:83
    r.f = min(min(a.f, b.f), min(c.f, d.f));
    r.pad = 0;
    r.sx = ((a.sx + b.sx) + c.sx) + d.sx;  // TreeInitShader.comp:79-80
    r.sy = ((a.sy + b.sy) + c.sy) + d.sy;
    r.sm = ((a.sm + b.sm) + c.sm) + d.sm;
    r.pad2 = 0.0f;
    return r;
}

// The depth-d cell node c (first particle c.f) at its pre-order slot, given.
__device__ __forceinline__ void emit_cell_at(const TileCell &c, int d, uint32_t slot,
                                             float4 *__restrict__ node4,
                                             uint32_t *__restrict__ ncount,
                                             float r0 = kFixedR0) {
    if (c.y <= 1u) return;
    float px = c.sx, py = c.sy;
    if (c.sm != 0.0f) {  // TreeMoveShader.comp:66-67
        px = px / c.sm;
        py = py / c.sm;
    }
    ncount[slot] = c.y;
    node4[slot] = make_float4(px, py, c.sm, cell_radius(d, r0));
}

// Depth d of the cell starting at sorted position f: slot f + C(f) + (d - s(f) - 1).
__device__ __forceinline__ uint32_t cell_slot(uint32_t f, int d, const int8_t *__restrict__ sdep,
                                              const uint32_t *__restrict__ cpos) {
    return f + cpos[f] + (uint32_t)(d - (int)sdep[f] - 1);
}

__device__ __forceinline__ void emit_cell(const TileCell &c, int d,
                                          const int8_t *__restrict__ sdep,
                                          const uint32_t *__restrict__ cpos,
                                          float4 *__restrict__ node4,
                                          uint32_t *__restrict__ ncount,
                                          float r0 = kFixedR0) {
    if (c.y <= 1u) return;
    emit_cell_at(c, d, cell_slot(c.f, d, sdep, cpos), node4, ncount, r0);
}

// A particle node's fourth word (the reference's radius 0): minus the number
// of particle nodes from it to the next cell node, capped at 2^24 so it is
// exact in FP32 (a longer run continues from the next slot, which holds its
// own count).  Every reader takes w <= 0 as a particle node.
__device__ __forceinline__ void set_run_length(float4 *__restrict__ node4, uint32_t slot,
                                               uint32_t n) {
    reinterpret_cast<float *>(node4)[4 * (size_t)slot + 3] = -(float)min(n, 1u << 24);
}

// Exclusive scan of data[0, count) in place by one workgroup of 1024 threads:
// chunks of 4096 values, four consecutive values per thread (coalesced), wave
// shuffles, the 16 wave sums scanned by wave 0, a carry between chunks.
// wsum: 16 words of LDS.
__device__ __forceinline__ void scan_inplace_1024(uint32_t *__restrict__ data, uint32_t count,
                                                  uint32_t *wsum) {
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < count; base += 4096) {
        const uint32_t i0 = base + 4 * t;
        uint32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = i0 + u < count ? data[i0 + u] : 0u;
        const uint32_t s = (v[0] + v[1]) + (v[2] + v[3]);
        uint32_t x = s;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        lds_barrier();
        if (wave == 0) {
            uint32_t w = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const uint32_t y = __shfl_up(w, off, 64);
                if (lane >= (uint32_t)off) w += y;
            }
            if (lane < 16) wsum[lane] = w;
        }
        lds_barrier();
        uint32_t run = carry + (wave ? wsum[wave - 1] : 0u) + (x - s);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (i0 + u < count) data[i0 + u] = run;
            run += v[u];
        }
        carry += wsum[15];
        lds_barrier();  // wsum is rewritten by the next chunk
    }
}

// Inclusive scan of one value across a 256-thread workgroup; returns the
// inclusive prefix and writes the workgroup total.
__device__ __forceinline__ uint32_t block_scan256(uint32_t x, uint32_t *total,
                                                  uint32_t *wsum /* [4] LDS */) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += u;
    }
    if (lane == 63) wsum[wave] = x;
    lds_barrier();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += wsum[w];
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    lds_barrier();
    return x + before;
}

// Inclusive wave scan of two counters.
__device__ __forceinline__ void wave_scan2(uint32_t &a, uint32_t &b) {
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ua = __shfl_up(a, off, 64);
        const uint32_t ub = __shfl_up(b, off, 64);
        if (lane >= (uint32_t)off) {
            a += ua;
            b += ub;
        }
    }
}

// ---- leaf sums: m*x, m*y, m of a leaf's particles in ascending sorted order --
// Both forms add the products one after another in the same order, so a leaf
// gets the same bits whichever sums it.

// Leaves of at most kSmallLeaf particles are summed by one lane.
constexpr uint32_t kSmallLeaf = 24;

// Loads go out eight at a time (one round trip per eight products); the
// additions keep the sequential order.
__device__ __forceinline__ float4 lane_leaf_sum(const float4 *__restrict__ prod, uint32_t f,
                                                uint32_t e) {
    typedef float v4 __attribute__((ext_vector_type(4)));  // stays in registers
    const v4 *vp = reinterpret_cast<const v4 *>(prod);
    const v4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    float sx = 0.0f, sy = 0.0f, sm = 0.0f;
    for (uint32_t k = f; k < e; k += 8) {
        v4 q[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) q[u] = k + u < e ? vp[k + u] : zero;
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            const bool in = k + u < e;
            sx = in ? sx + q[u].x : sx;
            sy = in ? sy + q[u].y : sy;
            sm = in ? sm + q[u].z : sm;
        }
    }
    return make_float4(sx, sy, sm, 0.0f);
}

// A larger leaf [f, e) (wave-uniform bounds) by the whole wave: 64 products
// per coalesced load, four chunks in flight ahead of the one being summed,
// staged in the wave's LDS component by component (64 m*x, 64 m*y, 64 m).
// Lanes 0, 1 and 2 each sum one component, reading it 16 bytes at a time, so
// the wave issues one add per product; each component's order is the leaf's
// particle order, from +0.  The dependent chain of adds is what remains.
__device__ __forceinline__ float4 wave_leaf_sum(float4 *stage, const float4 *__restrict__ prod,
                                                uint32_t f, uint32_t e, uint32_t lane) {
    typedef float v4 __attribute__((ext_vector_type(4)));  // stays in registers
    const v4 *vp = reinterpret_cast<const v4 *>(prod);
    float *st = reinterpret_cast<float *>(stage);
    const uint32_t comp = lane < 3u ? lane : 0u;  // the component this lane sums
    const v4 *sc = reinterpret_cast<const v4 *>(st + 64 * comp);
    float s = 0.0f;
    // loads past the leaf read its last record instead (never summed): no
    // lane-dependent branch, so each wait is for the one chunk it needs
    v4 q[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) q[u] = vp[min(f + 64 * u + lane, e - 1)];
    for (uint32_t base = f; base < e; base += 256) {
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t cb = base + 64 * u;
            if (cb >= e) break;
            st[lane] = q[u].x;
            st[64 + lane] = q[u].y;
            st[128 + lane] = q[u].z;
            __builtin_amdgcn_wave_barrier();
            q[u] = vp[min(cb + 256 + lane, e - 1)];
            const uint32_t cnt = min(64u, e - cb);
            if (cnt == 64u) {
#pragma unroll
                for (int h = 0; h < 16; h += 4) {
                    v4 r[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) r[j] = sc[h + j];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        s = s + r[j].x;
                        s = s + r[j].y;
                        s = s + r[j].z;
                        s = s + r[j].w;
                    }
                }
            } else {
                for (uint32_t k = 0; k < cnt; ++k) s = s + st[64 * comp + k];
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    return make_float4(__shfl(s, 0, 64), __shfl(s, 1, 64), __shfl(s, 2, 64), 0.0f);
}

// ---- the walk of one W = 64 group --------------------------------------------
// Summation order, the same in every W = 64 launch, single-system or batched:
// accepted cells are added in walk order into c; the particle runs (leaf
// buckets) are cut into 64-node chunks numbered in walk order across runs;
// chunk j is summed node by node from zero and its sum added to partial
// j mod 4; the acceleration is (((c + p0) + p1) + p2) + p3.  One wave can
// compute all of them, or four waves one partial each (a dense-core group
// split across waves; the wave of p0 also adds the cells and hands over
// c + p0), with the same bits.
// Native 4-float vector: conditional copies stay in registers (HIP's float4
// is a union-based struct whose selects the compiler may route through
// private memory).
typedef float f4v __attribute__((ext_vector_type(4)));

// The (x, y, mass) words of node i, or zeros: a run sum uses no more, and a
// 16-byte load whose fourth word is dead lets the compiler reuse that
// register while the load is still in flight, which forces a wait for the
// load at the reuse (a prefetch issued before the previous chunk's sum would
// then stall at that sum's first instruction).
typedef float f3v __attribute__((ext_vector_type(3)));
__device__ __forceinline__ f3v load_node3(const float4 *__restrict__ node4, uint32_t i, bool ok) {
    f3v r = {0.0f, 0.0f, 0.0f};
    if (ok) {
        const float *p = reinterpret_cast<const float *>(node4 + i);
        r.x = p[0];
        r.y = p[1];
        r.z = p[2];
    }
    return r;
}

struct Partials {
    float xc = 0.0f, yc = 0.0f;  // accepted cells
    float x0 = 0.0f, y0 = 0.0f, x1 = 0.0f, y1 = 0.0f, x2 = 0.0f, y2 = 0.0f, x3 = 0.0f, y3 = 0.0f;
};

// Adds a chunk's sum to partial k (wave-uniform).  Selects rather than a
// switch: a switch over the fields lets the compiler index Partials as an
// array in private memory.
__device__ __forceinline__ void add_partial(Partials &P, uint32_t k, float cx, float cy) {
    const float ax0 = P.x0 + cx, ay0 = P.y0 + cy, ax1 = P.x1 + cx, ay1 = P.y1 + cy;
    const float ax2 = P.x2 + cx, ay2 = P.y2 + cy, ax3 = P.x3 + cx, ay3 = P.y3 + cy;
    P.x0 = k == 0u ? ax0 : P.x0;
    P.y0 = k == 0u ? ay0 : P.y0;
    P.x1 = k == 1u ? ax1 : P.x1;
    P.y1 = k == 1u ? ay1 : P.y1;
    P.x2 = k == 2u ? ax2 : P.x2;
    P.y2 = k == 2u ? ay2 : P.y2;
    P.x3 = k == 3u ? ax3 : P.x3;
    P.y3 = k == 3u ? ay3 : P.y3;
}

__device__ __forceinline__ void partial_sum(const Partials &P, float &ax, float &ay) {
    ax = (((P.xc + P.x0) + P.x1) + P.x2) + P.x3;
    ay = (((P.yc + P.y0) + P.y1) + P.y2) + P.y3;
}

// What the wave computing partial k of a split group hands over: c + p0 for
// k = 0, else pk; added in order they give partial_sum's bits.
__device__ __forceinline__ void split_share(const Partials &P, uint32_t k, float &ax, float &ay) {
    switch (k) {
        case 0: ax = P.xc + P.x0; ay = P.yc + P.y0; break;
        case 1: ax = P.x1; ay = P.y1; break;
        case 2: ax = P.x2; ay = P.y2; break;
        default: ax = P.x3; ay = P.y3; break;
    }
}


// The run chunks [queue[0], queue[qn]) of one queue batch, chunk numbers from
// ch0: each chunk this wave owns ((parts >> (chunk & 3)) & 1) is loaded with
// one coalesced load while the previous one is summed (across run
// boundaries; a chunk is staged in LDS before the next one's load is issued),
// staged in the wave's LDS as packed 12-byte (x, y, m) records,
// summed node by node from zero and added to its partial.  kHalf: the wave
// holds one 32-particle group twice (lane l and l + 32 the same particle) and
// each half sums part of every chunk: lanes 0-31 its first s nodes, lanes
// 32-63 the rest, s = the chunk's node count halved and rounded up to a
// multiple of 4 (32 for a full chunk), so a partial chunk costs the wave
// about half its nodes rather than up to all of them.
template <bool kHalf = false>
__device__ __forceinline__ void sum_run_queue(const uint2 *queue, uint32_t qn, uint32_t ch0,
                                              float4 *stage, const float4 *__restrict__ node4,
                                              float px, float py, float eps2, uint32_t parts,
                                              Partials &P) {
    const uint32_t lane = threadIdx.x & 63;
    float *st = reinterpret_cast<float *>(stage);
    const f4v *st4 = reinterpret_cast<const f4v *>(stage);
    // cursor: run e, chunk start base, run end rend, chunk number ch
    uint32_t e = 0, base = 0, rend = 0, ch = ch0;
    auto run_at = [&](uint32_t i, uint32_t &b, uint32_t &r) {
        const uint2 q = queue[i];
        b = __builtin_amdgcn_readfirstlane(q.x);
        r = b + __builtin_amdgcn_readfirstlane(q.y);
    };
    auto owned = [&](uint32_t c) { return ((parts >> (c & 3u)) & 1u) != 0u; };
    auto step = [&](uint32_t &ee, uint32_t &bb, uint32_t &rr, uint32_t &cc) {
        bb += 64u;
        ++cc;
        if (bb >= rr) {
            ++ee;
            if (ee < qn) run_at(ee, bb, rr);
        }
    };
    if (qn == 0) return;
    run_at(0, base, rend);
    while (e < qn && !owned(ch)) step(e, base, rend, ch);
    f3v cur = load_node3(node4, base + lane, e < qn && lane < min(64u, rend - base));
    while (e < qn) {
        const uint32_t cnt = min(64u, rend - base), k = ch & 3u;
        // stage this chunk first: its load (the previous iteration's
        // prefetch) is the only one in flight, so the wait for it cannot
        // also wait for the next chunk's load, issued right after
        st[3 * lane] = cur.x;
        st[3 * lane + 1] = cur.y;
        st[3 * lane + 2] = cur.z;
        uint32_t ne = e, nb = base, nr = rend, nch = ch;
        do {
            step(ne, nb, nr, nch);
        } while (ne < qn && !owned(nch));
        const f3v next = load_node3(node4, nb + lane, ne < qn && lane < min(64u, nr - nb));
        __builtin_amdgcn_wave_barrier();
        float cx = 0.0f, cy = 0.0f;
        uint32_t j = 0, jend = cnt;
        if (kHalf) {
            const uint32_t split = ((cnt + 7u) >> 3) << 2;  // a multiple of 4: float4 reads
            j = lane < 32u ? 0u : split;
            jend = lane < 32u ? min(split, cnt) : cnt;
        }
        for (; j + 8 <= jend; j += 8) {
            f4v r[6];
#pragma unroll
            for (int u = 0; u < 6; ++u) r[u] = st4[3 * (j / 4) + u];
            const float *q = reinterpret_cast<const float *>(r);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                accumulate(q[3 * u], q[3 * u + 1], q[3 * u + 2], px, py, eps2, cx, cy);
        }
        for (; j < jend; ++j)
            accumulate(st[3 * j], st[3 * j + 1], st[3 * j + 2], px, py, eps2, cx, cy);
        add_partial(P, k, cx, cy);
        __builtin_amdgcn_wave_barrier();
        cur = next;
        e = ne;
        base = nb;
        rend = nr;
        ch = nch;
    }
}

// Walks the pre-order nodes [ind, end) for the group's lanes (ForceShader.comp
// :111-141): a cell is accepted iff every real lane has radius <= (dx*dx +
// dy*dy + eps2) * theta2, rounded as the shader writes it, then ind += count,
// else ind++.  Particle nodes are always accepted, so a run of them is a plain
// direct sum; a particle node's fourth word (the reference's radius 0) holds
// minus the number of particle nodes from it to the next cell node, so one
// 16-byte record tells a cell from a run and gives the run's length.  `parts` (bit k) selects the partials this wave computes; the
// wave computing partial 0 adds the cells and counts the real-lane visits
// and acceptances.  kStaged: the walk decides first and queues the particle
// runs (up to qmax <= kRunQueue, `queue` is the wave's LDS; where the queue
// is cut changes no sum, as chunks are numbered across runs) -- no decision
// depends on
// a force -- then sums the queued runs with the next chunk always in flight
// (sum_run_queue), raising the wave's issue priority one level per prio_step
// chunks summed (0: never); the node record, its count and its run length are loaded
// together, so each step of the walk is one round trip.  Otherwise runs are
// summed as they come, through scalar loads.  Returns the walk's VALU weight
// (a cell visit ~12 instructions, a particle ~10) in units of 4, capped at
// 65535.
// Issue-priority steps of the staged walk, in summed chunks of 64 source
// particles per priority level (s_setprio 1..3).  Dense-core groups are the
// tail of a launch whose groups are all resident from the start; letting them
// take issue slots ahead of lighter waves shortens it.  Measured best: 56-72
// for the single-system 500k two-disk step (traversal -4.5%), 32 for the
// batched workload (+1.8%; the batch's step and carry are autotuner knobs,
// BatchTune); 0 turns it off.
constexpr uint32_t kPrioStepSingle = 64;
// A single-system group also starts at level (chunks it summed last step) /
// kPrioCarry: the heavy groups lead from the first instruction.  Measured on
// the same workload: 24 -> traversal a further -4% (16: none, 32: -2%, 64: +1%).
// The batched force launch does the same per work item.
constexpr uint32_t kPrioCarry = 24;

__device__ __forceinline__ void set_issue_priority(uint32_t level) {
    if (level >= 3) __builtin_amdgcn_s_setprio(3);
    else if (level == 2) __builtin_amdgcn_s_setprio(2);
    else if (level == 1) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// kHalf (staged only): one 32-particle group held twice by the wave, see
// sum_run_queue; the votes are the group's, as both halves vote alike.
template <bool kStaged, bool kHalf = false>
__device__ __forceinline__ uint32_t walk_partials(
    uint32_t ind, uint32_t end, float4 *stage, uint2 *queue, const float4 *__restrict__ node4,
    const uint32_t *__restrict__ ncount, float px, float py, bool real, uint32_t nreal,
    float eps2, float theta2, uint32_t parts, Partials &P, unsigned long long &vis,
    unsigned long long &acc, uint32_t prio_step, uint32_t prio0 = 0,
    uint32_t *chunks_out = nullptr, uint32_t qmax = kRunQueue) {
    static_assert(kStaged || !kHalf, "half-wave groups need the staged walk");
    const uint32_t lane = threadIdx.x & 63;
    const bool first = (parts & 1u) != 0;
    uint32_t chunk = 0, cells = 0, runs = 0, qn = 0, qch = 0;
    uint32_t level = 0;
    if (kStaged && prio_step && prio0) {
        level = min(prio0, 3u);
        set_issue_priority(level);
    }
    while (true) {
      while (ind < end && qn < qmax) {
        const float4 nd = node4[ind];
        const uint32_t cnt = ncount[ind];
        // one round trip per step: both loads are issued here and waited for
        // together (otherwise the compiler sinks the x, y half of the record
        // and the count into the branches that use them, three dependent
        // round trips for an accepted cell)
        asm volatile("" ::"s"(nd.x), "s"(nd.y), "s"(nd.z), "s"(nd.w), "s"(cnt));
        if (nd.w <= 0.0f) {  // a particle node: w = -(particle nodes left in its run)
            const uint32_t pr = (uint32_t)(-nd.w);
            const uint32_t n = __builtin_amdgcn_readfirstlane(min(max(pr, 1u), end - ind));
            if (first) {
                vis += (unsigned long long)n * nreal;
                acc += (unsigned long long)n * nreal;
            }
            runs += n;
            if (kStaged) {
                if (lane == 0) queue[qn] = make_uint2(ind, n);
                ++qn;
                chunk += (n + 63u) >> 6;
                ind += n;
                continue;
            }
            const uint32_t rend = ind + n;
            for (uint32_t base = ind; base < rend; base += 64, ++chunk) {
                const uint32_t k = chunk & 3u;
                if (!((parts >> k) & 1u)) continue;
                const uint32_t ce = base + min(64u, rend - base);
                float cx = 0.0f, cy = 0.0f;
                uint32_t j = base;
                for (; j + 4 <= ce; j += 4) {
                    float4 q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) q[u] = node4[j + u];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        accumulate(q[u].x, q[u].y, q[u].z, px, py, eps2, cx, cy);
                }
                for (; j < ce; ++j) {
                    const float4 q = node4[j];
                    accumulate(q.x, q.y, q.z, px, py, eps2, cx, cy);
                }
                add_partial(P, k, cx, cy);
            }
            ind = rend;
            continue;
        }
        ++cells;
        const float dx = nd.x - px;
        const float dy = nd.y - py;
        const float r2 = (dx * dx + dy * dy) + eps2;  // ForceShader.comp:124-125
        const bool ok = (nd.w <= r2 * theta2) || !real;
        if (first) vis += nreal;
        if (__ballot(ok) == ~0ull) {
            if (first) {
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float sc = (nd.z * inv) * (inv * inv);
                P.xc = __builtin_fmaf(dx, sc, P.xc);
                P.yc = __builtin_fmaf(dy, sc, P.yc);
                acc += nreal;
            }
            ind = __builtin_amdgcn_readfirstlane(ind + max(cnt, 1u));
        } else {
            ind = __builtin_amdgcn_readfirstlane(ind + 1u);
        }
      }
      if (kStaged && qn) {  // the one place queued runs are summed
          // a wave that has summed many chunks is a dense-core group, the
          // tail of the launch: it takes issue priority over lighter waves
          if (prio_step) {
              const uint32_t want = min(chunk / prio_step, 3u);
              if (want > level) {
                  level = want;
                  set_issue_priority(level);
              }
          }
          __builtin_amdgcn_wave_barrier();
          sum_run_queue<kHalf>(queue, qn, qch, stage, node4, px, py, eps2, parts, P);
          qn = 0;
          qch = chunk;
      }
      if (ind >= end) break;
    }
    if (kStaged && level) __builtin_amdgcn_s_setprio(0);  // a persistent wave goes on to other work
    if (chunks_out) *chunks_out = chunk;
    return min((cells * 12u + runs * 10u) >> 2, 65535u);
}

}  // namespace gsim
