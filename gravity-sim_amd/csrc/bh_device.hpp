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

__device__ __forceinline__ float cell_radius(int d) {
    // TreeMoveShader.comp:70-71: S*S*2 / 4^d, exact in FP32
    return (kSimSize * kSimSize * 2.0f) / (float)(1u << (2 * d));
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
                                          uint32_t *__restrict__ ncount) {
    if (c.y <= 1u) return;
    const uint32_t f = c.f;
    const uint32_t slot = f + cpos[f] + (uint32_t)(d - (int)sdep[f] - 1);
    float px = c.sx, py = c.sy;
    if (c.sm != 0.0f) {  // TreeMoveShader.comp:66-67
        px = px / c.sm;
        py = py / c.sm;
    }
    ncount[slot] = c.y;
    node4[slot] = make_float4(px, py, c.sm, cell_radius(d));
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
    const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float sx = 0.0f, sy = 0.0f, sm = 0.0f;
    float4 next = f + lane < e ? prod[f + lane] : zero4;
    for (uint32_t base = f; base < e; base += 64) {
        const uint32_t cnt = min(64u, e - base);
        stage[lane] = next;
        __builtin_amdgcn_wave_barrier();
        const uint32_t nb = base + 64 + lane;
        next = nb < e ? prod[nb] : zero4;
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

}  // namespace gsim
