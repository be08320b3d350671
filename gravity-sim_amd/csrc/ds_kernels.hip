// Direct-sum force and integrate for gfx950.
//
// Follows the all-pairs step of SimShader.comp:37-74: for every target i,
// a_i = sum_j d * (m_j r^3) with d = p_j - p_i, r = rsqrt(d.d + eps2), then
// v' = v + a*(dt*G), p' = p + ((v+v')*0.5)*dt.
//
// One-sided form -- work split: a target block of 256*R targets (R per lane) times a source
// slice of `jlen` particles is one workgroup.  Source particles are read with
// wave-uniform indices, so they arrive through scalar loads (s_load) and sit
// in SGPRs: no LDS traffic and no VALU cycles spent on broadcast.  Each slice
// writes a partial acceleration; the combine kernel adds the S partials in
// slice order (deterministic) and integrates.
#include <cmath>
#include <cstring>
#include <mutex>

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

// ---- symmetric direct sum: every unordered pair evaluated once ----------------
//
// The 64-particle blocks of the padded set form a B x B block matrix of
// interactions; only its upper triangle is evaluated, each pair (i, j) once,
// adding  m_j d inv^3  to a_i and  -m_i d inv^3  to a_j (d = p_j - p_i; the
// pair's r2 and inv are the same numbers from either side).  Row groups of
// kDsSymRows blocks sit in one wave (each lane holds one particle of every row
// block); a workgroup tile is one row group times 4 * `cols` column blocks,
// `cols` per wave.  For a column
// block the wave holds one source particle per lane and rotates the sources,
// with their accumulated reactions, one lane per step (DPP wave_ror), so after
// 64 steps every lane has met every source once and each reaction is back in
// its owner's lane.  Per pair: 12 VALU + 1 v_rsq_f32 for two interactions,
// against 9 + 1 for one in the one-sided kernel; the 5 rotates per step are
// shared by the group's 16 rows (16 rows beat 8 by 3-5%: 0.420 against 0.431
// ms/step at 50k, 1.537 against 1.609 at 100k).
//
// Partial sums, both in fixed order: per tile, each row particle's sum over
// the tile's columns (ipart; the four waves' sums added in wave order through
// LDS); per row group and column block, each column
// particle's reactions from the group's rows (jpart).  The combine kernel adds
// a particle's row partials in column order, then its column partials in
// group order, and integrates.  Diagonal blocks are evaluated one-sided (so a
// particle meets itself, as in the reference sum: with zero softening that
// term is NaN there too).

// Wave rotate right by one lane (DPP wave_ror:1, 0x13C): lane l receives lane
// l-1.  Every source lane is valid, so no "old" value is needed.
__device__ __forceinline__ float ror1(float v) {
    return __builtin_bit_cast(float,
                              __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x13C, 0xF, 0xF, true));
}

// One column block against the wave's row blocks.  kInGroup: the column lies
// inside the row group, so rows below it are skipped, the diagonal row is
// one-sided and only rows above it react.
template <bool kInGroup>
__device__ __forceinline__ void ds_sym_column(const float (&ix)[kDsSymRows],
                                              const float (&iy)[kDsSymRows],
                                              const float (&im)[kDsSymRows],
                                              float (&aix)[kDsSymRows], float (&aiy)[kDsSymRows],
                                              float jx, float jy, float jm, float &ajx_out,
                                              float &ajy_out, float eps2, uint32_t I0, uint32_t J) {
    float ajx = 0.0f, ajy = 0.0f;
#pragma unroll 2
    for (int k = 0; k < 64; ++k) {
#pragma unroll
        for (int r = 0; r < kDsSymRows; ++r) {
            if (kInGroup && I0 + r > J) continue;
            const float dx = jx - ix[r], dy = jy - iy[r];
            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, eps2));
            const float inv = __builtin_amdgcn_rsqf(r2);
            const float inv3 = (inv * inv) * inv;
            const float ti = jm * inv3;
            aix[r] = __builtin_fmaf(dx, ti, aix[r]);
            aiy[r] = __builtin_fmaf(dy, ti, aiy[r]);
            if (!kInGroup || I0 + r < J) {
                const float tj = im[r] * inv3;
                ajx = __builtin_fmaf(-dx, tj, ajx);
                ajy = __builtin_fmaf(-dy, tj, ajy);
            }
        }
        jx = ror1(jx);
        jy = ror1(jy);
        jm = ror1(jm);
        ajx = ror1(ajx);
        ajy = ror1(ajy);
    }
    ajx_out = ajx;
    ajy_out = ajy;
}

// Row group of tile t: the largest g with tile_base(g) <= t.
__device__ __forceinline__ uint32_t ds_sym_group(uint32_t t, uint32_t G, uint32_t cb,
                                                 uint32_t step) {
    uint32_t lo = 0, hi = G;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ds_sym_tile_base(mid, cb, step) <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void ds_sym_kernel(const float2 *__restrict__ pos,
                                                     const float *__restrict__ mass,
                                                     float2 *__restrict__ ipart,
                                                     float2 *__restrict__ jpart, uint32_t B,
                                                     uint32_t G, uint32_t cols, float eps2) {
    constexpr int R = kDsSymRows;
    __shared__ float2 red[3][R][64];  // row sums of waves 1..3
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t T = blockIdx.x;  // one tile per workgroup
    const uint32_t width = 4 * cols;
    const uint32_t cb = (B + width - 1) / width, step = R / width;
    const uint32_t g = ds_sym_group(T, G, cb, step);
    const uint32_t c = T - ds_sym_tile_base(g, cb, step);
    const uint32_t I0 = g * R;
    const uint32_t Jt = I0 + c * width;
    const uint32_t J0 = min(B, Jt + w * cols), J1 = min(B, Jt + (w + 1) * cols);
    float ix[R], iy[R], im[R], aix[R], aiy[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = (I0 + r) * 64 + lane;
        const bool ok = I0 + r < B;
        const float2 p = ok ? pos[i] : make_float2(0.0f, 0.0f);
        ix[r] = p.x;
        iy[r] = p.y;
        im[r] = ok ? mass[i] : 0.0f;
        aix[r] = 0.0f;
        aiy[r] = 0.0f;
    }
    const size_t jrow = ds_sym_jbase(g, B);
    for (uint32_t J = J0; J < J1; ++J) {
        const uint32_t j = J * 64 + lane;
        const float2 q = pos[j];
        const float qm = mass[j];
        float ajx, ajy;
        if (J >= I0 + R)
            ds_sym_column<false>(ix, iy, im, aix, aiy, q.x, q.y, qm, ajx, ajy, eps2, I0, J);
        else
            ds_sym_column<true>(ix, iy, im, aix, aiy, q.x, q.y, qm, ajx, ajy, eps2, I0, J);
        jpart[(jrow + (J - I0)) * 64 + lane] = make_float2(ajx, ajy);
    }
    if (w > 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) red[w - 1][r][lane] = make_float2(aix[r], aiy[r]);
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (I0 + r >= B) continue;
        float sx = aix[r], sy = aiy[r];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const float2 v = red[u][r][lane];
            sx = sx + v.x;
            sy = sy + v.y;
        }
        ipart[((size_t)T * R + r) * 64 + lane] = make_float2(sx, sy);
    }
}

// One workgroup per 64-particle block: its four waves each add a quarter of
// the block's partial list in order, then the quarters are added in order.
__global__ __launch_bounds__(256) void ds_sym_combine_kernel(
    const float2 *__restrict__ ipart, const float2 *__restrict__ jpart, uint32_t B, uint32_t width,
    const float2 *__restrict__ pos, const float2 *__restrict__ vel, float2 *__restrict__ pos_out,
    float2 *__restrict__ vel_out, float2 *__restrict__ acc_dbg, float kv, float dt) {
    constexpr uint32_t R = kDsSymRows;
    __shared__ float2 q[4][64];
    const uint32_t P = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t g = P / R, r = P % R;
    const uint32_t cb = (B + width - 1) / width, step = R / width;
    const uint32_t ni = cb - step * g;  // row partials: the group's tiles
    const uint32_t nl = ni + g + 1;     // then one column partial per group 0..g
    const uint32_t per = (nl + 3) / 4;
    const uint32_t lo = min(nl, w * per), hi = min(nl, lo + per);
    const size_t tb = ds_sym_tile_base(g, cb, step);
    auto addr = [&](uint32_t e) -> size_t {
        if (e < ni) return ((tb + e) * R + r) * 64 + lane;
        const uint32_t h = e - ni;
        return (ds_sym_jbase(h, B) + (P - h * R)) * 64 + lane;
    };
    float ax = 0.0f, ay = 0.0f;
    uint32_t e = lo;
    for (; e + 4 <= hi; e += 4) {
        float2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const size_t a = addr(e + u);
            v[u] = e + u < ni ? ipart[a] : jpart[a];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ax = ax + v[u].x;
            ay = ay + v[u].y;
        }
    }
    for (; e < hi; ++e) {
        const size_t a = addr(e);
        const float2 v = e < ni ? ipart[a] : jpart[a];
        ax = ax + v.x;
        ay = ay + v.y;
    }
    q[w][lane] = make_float2(ax, ay);
    __syncthreads();
    if (w != 0) return;
    ax = ((q[0][lane].x + q[1][lane].x) + q[2][lane].x) + q[3][lane].x;
    ay = ((q[0][lane].y + q[1][lane].y) + q[2][lane].y) + q[3][lane].y;
    const uint32_t i = P * 64 + lane;
    const float2 p = pos[i], v = vel[i];
    // SimShader.comp:68-69
    const float nvx = v.x + ax * kv, nvy = v.y + ay * kv;
    pos_out[i] = make_float2(p.x + ((v.x + nvx) * 0.5f) * dt, p.y + ((v.y + nvy) * 0.5f) * dt);
    vel_out[i] = make_float2(nvx, nvy);
    if (acc_dbg) acc_dbg[i] = make_float2(ax, ay);
}

}  // namespace

DsSymPlan ds_sym_plan(size_t na, uint32_t cols) {
    DsSymPlan p{};
    p.B = (uint32_t)(na / 64);
    p.G = (p.B + kDsSymRows - 1) / kDsSymRows;
    auto tiles = [&](uint32_t width) -> size_t {
        const uint32_t cb = (p.B + width - 1) / width, step = kDsSymRows / width;
        return ds_sym_tile_base(p.G, cb, step);
    };
    // Two columns per wave once that still gives ~32k waves.  Measured
    // (ms/step, 1 / 2 columns): 20k 0.101 / 0.118, 50k 0.424 / 0.432, 100k
    // 1.639 / 1.607.
    if (cols == 0) cols = tiles(8) * 4 >= kDsSymMinWaves ? 2 : 1;
    p.cols = cols;
    p.width = 4 * cols;
    p.ntiles = (uint32_t)tiles(p.width);
    p.ipart = (size_t)p.ntiles * kDsSymRows * 64;
    p.jpart = ds_sym_jbase(p.G, p.B) * 64;
    return p;
}

bool ds_sym_default(size_t na) { return na >= kDsSymMinParticles && na <= kDsSymMaxParticles; }

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
    if (s->ds_sym) {
        const DsSymPlan &p = s->ds_plan;
        ds_sym_kernel<<<p.ntiles, 256, 0, s->stream>>>(pos, mass, s->partial,
                                                       s->partial + p.ipart, p.B, p.G, p.cols,
                                                       s->eps2);
        return hipGetLastError();
    }
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
    if (s->ds_sym) {
        const DsSymPlan &p = s->ds_plan;
        ds_sym_combine_kernel<<<p.B, 256, 0, s->stream>>>(s->partial, s->partial + p.ipart, p.B,
                                                          p.width, pos, vel, pos_out, vel_out, acc,
                                                          s->kv, s->dt);
        return hipGetLastError();
    }
    ds_combine_kernel<<<(na + 255) / 256, 256, 0, s->stream>>>(
        s->partial, s->ds_S_used, pos, vel, pos_out, vel_out, acc, na, s->kv, s->dt);
    return hipGetLastError();
}

}  // namespace gsim

// ---- allocation-free direct-sum step ---------------------------------------------
//
// gsim_ds_step_noalloc runs one one-sided direct-sum step on caller-owned
// device arrays.  Its only scratch, the per-slice partial accelerations, lives
// in `scratch`, a file-scope device array that every GPU holds for the life of
// the process, so the call allocates nothing and can run where hipMalloc
// cannot (under stream capture, inside a serving loop that forbids it).  In
// host code `scratch` names a host-side stand-in, never device memory: the
// device address comes from hipGetSymbolAddress, looked up once per device
// with that device current, and only that address reaches a kernel.  Every
// call on one GPU shares its array; the partial kernel writes every entry the
// combine reads (slices x Na), so nothing needs clearing between calls, and
// calls that overlap on one GPU race on it.

namespace gsim {
namespace {

constexpr size_t kNoallocScratchEntries = GSIM_NOALLOC_SCRATCH_BYTES / sizeof(float2);
static __device__ __align__(256) float2 scratch[kNoallocScratchEntries];

constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
float2 *g_noalloc_ptr[kNoallocMaxDevices] = {};

// Looks up this device's copy of `scratch` (and checks it is a gfx950) on the
// first call per device; later calls read the cached address.  The caller has
// made `device` current.
gsim_status noalloc_setup(int device, float2 **out) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) {
        *out = g_noalloc_ptr[device];
        return GSIM_OK;
    }
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess)
        return fail(GSIM_ERROR_DEVICE, "hipGetDeviceProperties failed: %s", hipGetErrorString(e));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(GSIM_ERROR_NO_DEVICE, "device %d is %s; libgsim is built for gfx950 only",
                    device, prop.gcnArchName);
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(GSIM_ERROR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                    hipGetErrorString(e));
    g_noalloc_ptr[device] = static_cast<float2 *>(ptr);
    *out = g_noalloc_ptr[device];
    return GSIM_OK;
}

// Launches the partial and combine kernels; the caller has made the device
// current and checked every argument.
gsim_status noalloc_launch(float2 *part, const gsim_params *p, const float2 *pos,
                           const float2 *vel, const float *mass, float2 *pos_out,
                           float2 *vel_out, uint32_t na, hipStream_t stream) {
    uint32_t R = 2, S = 1;
    ds_pick_shape(na, &R, &S);  // the shape gsim_step picks for a one-sided handle
    const size_t fit = kNoallocScratchEntries / na;
    if (S > fit) S = (uint32_t)fit;
    const uint32_t jlen = ((na + S - 1) / S + 63) / 64 * 64;
    const uint32_t slices = (na + jlen - 1) / jlen;
    // constants as gsim_create forms them (DirectSimulation.cpp:179, SimShader.comp:68)
    const float eps2 = p->softening_len * p->softening_len;
    const float dt = p->simulation_time;
    const float kv = dt * p->gravitational_const;
    const dim3 grid((na + kDsThreads * 2 - 1) / (kDsThreads * 2), slices);
    ds_partial_kernel<2, false><<<grid, kDsThreads, 0, stream>>>(pos, mass, part, na, jlen, eps2);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        ds_combine_kernel<<<(na + 255) / 256, 256, 0, stream>>>(part, slices, pos, vel, pos_out,
                                                                vel_out, nullptr, na, kv, dt);
        e = hipGetLastError();
    }
    if (e != hipSuccess)
        return fail(GSIM_ERROR_DEVICE, "direct-sum launch failed: %s", hipGetErrorString(e));
    return GSIM_OK;
}

}  // namespace
}  // namespace gsim

extern "C" gsim_status gsim_ds_step_noalloc(const gsim_params *params, const float *pos_xy,
                                            const float *vel_xy, const float *mass,
                                            float *pos_out_xy, float *vel_out_xy, size_t n,
                                            void *stream) {
    using gsim::fail;
    gsim::clear_error();
    if (!params || !pos_xy || !vel_xy || !mass || !pos_out_xy || !vel_out_xy)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_ds_step_noalloc: NULL argument");
    if (params->algorithm != GSIM_DIRECT_SUM)
        return fail(GSIM_ERROR_INVALID_ARGUMENT,
                    "gsim_ds_step_noalloc: only the direct-sum algorithm runs without allocation");
    if (n == 0 || n % 64 != 0)
        return fail(GSIM_ERROR_INVALID_ARGUMENT,
                    "gsim_ds_step_noalloc: n = %zu must be a nonzero multiple of 64 (pad with "
                    "zero-mass particles)", n);
    if (n > gsim::kNoallocScratchEntries)
        return fail(GSIM_ERROR_CAPACITY,
                    "gsim_ds_step_noalloc: %zu particles exceed the %zu the %u-byte static "
                    "scratch holds", n, gsim::kNoallocScratchEntries,
                    (unsigned)GSIM_NOALLOC_SCRATCH_BYTES);
    if (pos_out_xy == pos_xy || vel_out_xy == vel_xy || pos_out_xy == vel_xy ||
        vel_out_xy == pos_xy)
        return fail(GSIM_ERROR_INVALID_ARGUMENT,
                    "gsim_ds_step_noalloc: output arrays must not alias the inputs");
    const float vals[] = {params->gravitational_const, params->simulation_time,
                          params->softening_len};
    for (float v : vals)
        if (!std::isfinite(v))
            return fail(GSIM_ERROR_INVALID_ARGUMENT, "simulation parameters must be finite");
    const int device = params->device;
    if (device < 0 || device >= gsim::kNoallocMaxDevices)
        return fail(GSIM_ERROR_NO_DEVICE, "gsim_ds_step_noalloc: device %d out of range", device);
    int prev = -1;
    hipError_t e = hipGetDevice(&prev);
    if (e != hipSuccess)
        return fail(GSIM_ERROR_NO_DEVICE, "hipGetDevice failed: %s", hipGetErrorString(e));
    if (prev != device && (e = hipSetDevice(device)) != hipSuccess)
        return fail(GSIM_ERROR_NO_DEVICE, "hipSetDevice(%d) failed: %s", device,
                    hipGetErrorString(e));
    float2 *part = nullptr;
    gsim_status st = gsim::noalloc_setup(device, &part);
    if (st == GSIM_OK)
        st = gsim::noalloc_launch(part, params, reinterpret_cast<const float2 *>(pos_xy),
                                  reinterpret_cast<const float2 *>(vel_xy), mass,
                                  reinterpret_cast<float2 *>(pos_out_xy),
                                  reinterpret_cast<float2 *>(vel_out_xy), (uint32_t)n,
                                  static_cast<hipStream_t>(stream));
    if (prev != device) (void)hipSetDevice(prev);
    return st;
}
