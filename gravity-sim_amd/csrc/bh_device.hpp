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

__device__ __forceinline__ void emit_cell(const TileCell &c, int d,
                                          const int8_t *__restrict__ sdep,
                                          const uint32_t *__restrict__ cpos,
                                          float4 *__restrict__ node4,
                                          uint32_t *__restrict__ ncount,
                                          float r0 = kFixedR0) {
    if (c.y <= 1u) return;
    const uint32_t f = c.f;
    const uint32_t slot = f + cpos[f] + (uint32_t)(d - (int)sdep[f] - 1);
    float px = c.sx, py = c.sy;
    if (c.sm != 0.0f) {  // TreeMoveShader.comp:66-67
        px = px / c.sm;
        py = py / c.sm;
    }
    ncount[slot] = c.y;
    node4[slot] = make_float4(px, py, c.sm, cell_radius(d, r0));
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
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += wsum[w];
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
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

__device__ __forceinline__ float4 lane_leaf_sum(const float4 *__restrict__ prod, uint32_t f,
                                                uint32_t e) {
    float sx = 0.0f, sy = 0.0f, sm = 0.0f;
    for (uint32_t k = f; k < e; ++k) {
        const float4 q = prod[k];
        sx = sx + q.x;
        sy = sy + q.y;
        sm = sm + q.z;
    }
    return make_float4(sx, sy, sm, 0.0f);
}

// A larger leaf [f, e) (wave-uniform bounds) by the whole wave: 64 products
// per coalesced load into the wave's LDS stage, the next chunk in flight while
// the current one is summed (every lane computes the same running sum).
__device__ __forceinline__ float4 wave_leaf_sum(float4 *stage, const float4 *__restrict__ prod,
                                                uint32_t f, uint32_t e, uint32_t lane) {
    typedef float v4 __attribute__((ext_vector_type(4)));  // stays in registers
    const v4 *vp = reinterpret_cast<const v4 *>(prod);
    const v4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    float sx = 0.0f, sy = 0.0f, sm = 0.0f;
    v4 next = f + lane < e ? vp[f + lane] : zero;
    for (uint32_t base = f; base < e; base += 64) {
        const uint32_t cnt = min(64u, e - base);
        reinterpret_cast<v4 *>(stage)[lane] = next;
        __builtin_amdgcn_wave_barrier();
        const uint32_t nb = base + 64 + lane;
        next = nb < e ? vp[nb] : zero;
        uint32_t k = 0;
        for (; k + 8 <= cnt; k += 8) {
            float4 q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) q[u] = stage[k + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                sx = sx + q[u].x;
                sy = sy + q[u].y;
                sm = sm + q[u].z;
            }
        }
        for (; k < cnt; ++k) {
            const float4 q = stage[k];
            sx = sx + q.x;
            sy = sy + q.y;
            sm = sm + q.z;
        }
        __builtin_amdgcn_wave_barrier();
    }
    return make_float4(sx, sy, sm, 0.0f);
}

// ---- the walk of one W = 64 group --------------------------------------------
// Summation order, the same in every W = 64 launch, single-system or batched:
// the particle runs (leaf buckets) are cut into 64-node chunks numbered in
// walk order across runs; chunk c is accumulated node by node into partial
// c mod 4, every accepted cell into partial 0, and the acceleration is
// ((p0 + p1) + p2) + p3.  One wave can compute all four partials, or four
// waves one each (a dense-core group split across waves), with the same bits.
// Native 4-float vector: conditional copies stay in registers (HIP's float4
// is a union-based struct whose selects the compiler may route through
// private memory).
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v load_node(const float4 *__restrict__ node4, uint32_t i, bool ok) {
    return ok ? reinterpret_cast<const f4v *>(node4)[i] : f4v{0.0f, 0.0f, 0.0f, 0.0f};
}

struct Partials {
    float x0 = 0.0f, y0 = 0.0f, x1 = 0.0f, y1 = 0.0f, x2 = 0.0f, y2 = 0.0f, x3 = 0.0f, y3 = 0.0f;
};

// Runs f(ax, ay) on partial k (wave-uniform), each case on named registers so
// the partials never leave them.
template <class F>
__device__ __forceinline__ void with_partial(Partials &P, uint32_t k, F &&f) {
    switch (k) {
        case 0: f(P.x0, P.y0); break;
        case 1: f(P.x1, P.y1); break;
        case 2: f(P.x2, P.y2); break;
        default: f(P.x3, P.y3); break;
    }
}

__device__ __forceinline__ void partial_sum(const Partials &P, float &ax, float &ay) {
    ax = ((P.x0 + P.x1) + P.x2) + P.x3;
    ay = ((P.y0 + P.y1) + P.y2) + P.y3;
}

// Walks the pre-order nodes [ind, end) for the group's lanes (ForceShader.comp
// :111-141): a cell is accepted iff every real lane has radius <= (dx*dx +
// dy*dy + eps2) * theta2, rounded as the shader writes it, then ind += count,
// else ind++.  Particle nodes are always accepted, so a run of them is a plain
// direct sum.  `parts` (bit k) selects the partials this wave computes; the
// wave computing partial 0 adds the cells and counts the real-lane visits
// and acceptances.  kStaged: run chunks come through the wave's LDS stage (one
// coalesced load per chunk, broadcast reads; walking all four partials, the
// next chunk of a run is loaded while the current one is summed); otherwise
// through scalar loads.  Returns
// the walk's VALU weight (a cell visit ~12 instructions, a particle ~10) in
// units of 4, capped at 65535.
template <bool kStaged>
__device__ __forceinline__ uint32_t walk_partials(
    uint32_t ind, uint32_t end, float4 *stage, const float4 *__restrict__ node4,
    const uint32_t *__restrict__ ncount, const uint32_t *__restrict__ prun, float px, float py,
    bool real, uint32_t nreal, float eps2, float theta2, uint32_t parts, Partials &P,
    unsigned long long &vis, unsigned long long &acc) {
    const uint32_t lane = threadIdx.x & 63;
    const bool first = (parts & 1u) != 0;
    uint32_t chunk = 0, cells = 0, runs = 0;
    while (ind < end) {
        const float4 nd = node4[ind];
        if (nd.w == 0.0f) {
            uint32_t n = prun[ind];
            n = __builtin_amdgcn_readfirstlane(min(max(n, 1u), end - ind));
            const uint32_t rend = ind + n;
            if (kStaged) {
                // all partials: every chunk of the run is this wave's, so the
                // next one is loaded while the current one is summed
                // the stage holds (x, y, m) records packed at 12 bytes, so a
                // broadcast read of 48 bytes brings four particles
                const bool all = parts == 0xfu;
                f4v next = load_node(node4, ind + lane, all && lane < min(64u, n));
                float *st = reinterpret_cast<float *>(stage);
                const f4v *st4 = reinterpret_cast<const f4v *>(stage);
                for (uint32_t base = ind; base < rend; base += 64, ++chunk) {
                    const uint32_t k = chunk & 3u;
                    if (!((parts >> k) & 1u)) continue;
                    const uint32_t cnt = min(64u, rend - base);
                    f4v cur;
                    if (all) {
                        cur = next;
                        const uint32_t nb = base + 64;
                        const uint32_t ncnt = nb < rend ? min(64u, rend - nb) : 0u;
                        next = load_node(node4, nb + lane, lane < ncnt);
                    } else {
                        cur = load_node(node4, base + lane, lane < cnt);
                    }
                    st[3 * lane] = cur.x;
                    st[3 * lane + 1] = cur.y;
                    st[3 * lane + 2] = cur.z;
                    __builtin_amdgcn_wave_barrier();
                    with_partial(P, k, [&](float &ax, float &ay) {
                        uint32_t j = 0;
                        for (; j + 8 <= cnt; j += 8) {
                            f4v r[6];
#pragma unroll
                            for (int u = 0; u < 6; ++u) r[u] = st4[3 * (j / 4) + u];
                            const float *q = reinterpret_cast<const float *>(r);
#pragma unroll
                            for (int u = 0; u < 8; ++u)
                                accumulate(q[3 * u], q[3 * u + 1], q[3 * u + 2], px, py, eps2, ax, ay);
                        }
                        for (; j < cnt; ++j)
                            accumulate(st[3 * j], st[3 * j + 1], st[3 * j + 2], px, py, eps2, ax, ay);
                    });
                    __builtin_amdgcn_wave_barrier();
                }
            } else {
                for (uint32_t base = ind; base < rend; base += 64, ++chunk) {
                    const uint32_t k = chunk & 3u;
                    if (!((parts >> k) & 1u)) continue;
                    const uint32_t ce = base + min(64u, rend - base);
                    with_partial(P, k, [&](float &ax, float &ay) {
                        uint32_t j = base;
                        for (; j + 4 <= ce; j += 4) {
                            float4 q[4];
#pragma unroll
                            for (int u = 0; u < 4; ++u) q[u] = node4[j + u];
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                accumulate(q[u].x, q[u].y, q[u].z, px, py, eps2, ax, ay);
                        }
                        for (; j < ce; ++j) {
                            const float4 q = node4[j];
                            accumulate(q.x, q.y, q.z, px, py, eps2, ax, ay);
                        }
                    });
                }
            }
            if (first) {
                vis += (unsigned long long)n * nreal;
                acc += (unsigned long long)n * nreal;
            }
            runs += n;
            ind = rend;
            continue;
        }
        ++cells;
        const uint32_t cnt = ncount[ind];
        const float dx = nd.x - px;
        const float dy = nd.y - py;
        const float r2 = (dx * dx + dy * dy) + eps2;  // ForceShader.comp:124-125
        const bool ok = (nd.w <= r2 * theta2) || !real;
        if (first) vis += nreal;
        if (__ballot(ok) == ~0ull) {
            if (first) {
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float sc = (nd.z * inv) * (inv * inv);
                P.x0 = __builtin_fmaf(dx, sc, P.x0);
                P.y0 = __builtin_fmaf(dy, sc, P.y0);
                acc += nreal;
            }
            ind = __builtin_amdgcn_readfirstlane(ind + max(cnt, 1u));
        } else {
            ind = __builtin_amdgcn_readfirstlane(ind + 1u);
        }
    }
    return min((cells * 12u + runs * 10u) >> 2, 65535u);
}

}  // namespace gsim
