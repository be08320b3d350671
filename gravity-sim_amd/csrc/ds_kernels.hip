// Direct-sum force and integrate for gfx950.
//
// Follows the all-pairs step of SimShader.comp:37-74: for every target i,
// a_i = sum_j d * (m_j r^3) with d = p_j - p_i, r = rsqrt(d.d + eps2), then
// v' = v + a*(dt*G), p' = p + ((v+v')*0.5)*dt.
//
// Work split: a target block of 256*R targets (R per lane) times a source
// slice of `jlen` particles is one workgroup.  Source particles are read with
// wave-uniform indices, so they arrive through scalar loads (s_load) and sit
// in SGPRs: no LDS traffic and no VALU cycles spent on broadcast.  Each slice
// writes a partial acceleration; the combine kernel adds the S partials in
// slice order (deterministic) and integrates.
#include "internal.hpp"

namespace gsim {
namespace {

constexpr int kDsThreads = 256;

// kRcp: 1/r2 comes from v_rcp_f32 (the transcendental unit, which issues
// beside the VALU on gfx950) instead of inv*inv, one VALU multiply less per
// interaction.
template <int R, bool kRcp>
__global__ __launch_bounds__(kDsThreads) void ds_partial_kernel(
    const float2 *__restrict__ pos, const float *__restrict__ mass,
    float2 *__restrict__ partial, uint32_t na, uint32_t jlen, float eps2) {
    const uint32_t slice = blockIdx.y;
    const uint32_t base = blockIdx.x * (kDsThreads * R) + threadIdx.x;
    float px[R], py[R], ax[R], ay[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = base + r * kDsThreads;
        const float2 p = i < na ? pos[i] : make_float2(0.0f, 0.0f);
        px[r] = p.x; py[r] = p.y;
        ax[r] = 0.0f; ay[r] = 0.0f;
    }
    const uint32_t j0 = slice * jlen;
    const uint32_t j1 = min(na, j0 + jlen);
    // na and jlen are multiples of 64, so the source range is whole 8-groups.
    for (uint32_t j = j0; j < j1; j += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float2 s = pos[j + u];
            const float m = mass[j + u];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dx = s.x - px[r];
                const float dy = s.y - py[r];
                const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, eps2));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float inv2 = kRcp ? __builtin_amdgcn_rcpf(r2) : inv * inv;
                const float t = (m * inv) * inv2;
                ax[r] = __builtin_fmaf(dx, t, ax[r]);
                ay[r] = __builtin_fmaf(dy, t, ay[r]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = base + r * kDsThreads;
        if (i < na) partial[(size_t)slice * na + i] = make_float2(ax[r], ay[r]);
    }
}

__global__ __launch_bounds__(256) void ds_combine_kernel(
    const float2 *__restrict__ partial, uint32_t slices, const float2 *__restrict__ pos,
    const float2 *__restrict__ vel, float2 *__restrict__ pos_out,
    float2 *__restrict__ vel_out, float2 *__restrict__ acc_dbg, uint32_t na, float kv,
    float dt) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= na) return;
    float ax = 0.0f, ay = 0.0f;
    for (uint32_t s = 0; s < slices; ++s) {
        const float2 a = partial[(size_t)s * na + i];
        ax = ax + a.x;
        ay = ay + a.y;
    }
    const float2 p = pos[i], v = vel[i];
    // SimShader.comp:68-69
    const float nvx = v.x + ax * kv, nvy = v.y + ay * kv;
    pos_out[i] = make_float2(p.x + ((v.x + nvx) * 0.5f) * dt, p.y + ((v.y + nvy) * 0.5f) * dt);
    vel_out[i] = make_float2(nvx, nvy);
    if (acc_dbg) acc_dbg[i] = make_float2(ax, ay);
}

}  // namespace

// Launch shape by size: two targets per lane (each source load then feeds two
// interactions), and enough source slices for ~6k workgroups so 256 CUs stay
// full to the end; each slice keeps >= 512 sources.  Measured at 50,048
// particles: 0.546 ms/step for 2 x 64 against 0.721 ms for 1 x 16.
void ds_pick_shape(size_t na, uint32_t *R, uint32_t *S) {
    const uint32_t r = 2;
    const size_t tblocks = (na + kDsThreads * r - 1) / (kDsThreads * r);
    size_t s = (6144 + tblocks - 1) / tblocks;
    const size_t max_s = na / 512 ? na / 512 : 1;
    if (s > max_s) s = max_s;
    if (s < 1) s = 1;
    if (s > 64) s = 64;
    *R = r;
    *S = (uint32_t)s;
}

hipError_t ds_launch_partial(gsim_sim *s, const float2 *pos, const float *mass) {
    const uint32_t na = (uint32_t)s->na;
    const uint32_t R = s->ds_R, S = s->ds_S;
    // slice length: multiple of 64 covering na in S pieces
    uint32_t jlen = (uint32_t)(((na + S - 1) / S + 63) / 64 * 64);
    const uint32_t slices = (na + jlen - 1) / jlen;
    dim3 grid((na + kDsThreads * R - 1) / (kDsThreads * R), slices);
#define GSIM_DS(RR, RCP) \
    ds_partial_kernel<RR, RCP><<<grid, kDsThreads, 0, s->stream>>>(pos, mass, s->partial, na, jlen, s->eps2)
    const bool rcp = s->ds_rcp != 0;
    switch (R) {
        case 1: if (rcp) GSIM_DS(1, true); else GSIM_DS(1, false); break;
        case 2: if (rcp) GSIM_DS(2, true); else GSIM_DS(2, false); break;
        case 4: if (rcp) GSIM_DS(4, true); else GSIM_DS(4, false); break;
        default: return hipErrorInvalidValue;
    }
#undef GSIM_DS
    s->ds_S_used = slices;
    return hipGetLastError();
}

hipError_t ds_launch_combine(gsim_sim *s, const float2 *pos, const float2 *vel,
                             float2 *pos_out, float2 *vel_out) {
    const uint32_t na = (uint32_t)s->na;
    float2 *acc = (s->debug_flags & GSIM_DEBUG_ACCEL) ? s->accel : nullptr;
    ds_combine_kernel<<<(na + 255) / 256, 256, 0, s->stream>>>(
        s->partial, s->ds_S_used, pos, vel, pos_out, vel_out, acc, na, s->kv, s->dt);
    return hipGetLastError();
}

}  // namespace gsim
