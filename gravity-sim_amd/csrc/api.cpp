// C-ABI entry points of libgsim: handle lifetime, stepping, readback, debug
// dumps and HIP-event timing.  HIP kernels live in ds_kernels.hip and
// bh_kernels.hip; this file is host code built by hipcc.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.hpp"

namespace gsim {

namespace {
thread_local char g_error[512] = {0};
}

gsim_status fail(gsim_status st, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return st;
}

void clear_error() { g_error[0] = 0; }

}  // namespace gsim

using gsim::fail;

namespace {

gsim_status hip_fail(hipError_t e, const char *what) {
    return fail(GSIM_ERROR_DEVICE, "%s failed: %s", what, hipGetErrorString(e));
}

#define GSIM_HIP(call)                                       \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(e_, #call);    \
    } while (0)

size_t align_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

// out[perm[slot]] = in[slot] for the first n caller indices, w floats each
void unpermute(const std::vector<uint32_t> &perm, const float *in, float *out, size_t w, size_t n) {
    for (size_t slot = 0; slot < perm.size(); ++slot) {
        const size_t i = perm[slot];
        if (i >= n) continue;
        for (size_t c = 0; c < w; ++c) out[i * w + c] = in[slot * w + c];
    }
}

void free_sim(gsim_sim *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (int k = 0; k < 2; ++k) {
        (void)hipFree(s->pos[k]);
        (void)hipFree(s->vel[k]);
        (void)hipFree(s->perm[k]);
    }
    (void)hipFree(s->mass);
    (void)hipFree(s->mass_next);
    (void)hipFree(s->partial);
    (void)hipFree(s->accel);
    gsim::bh_free(&s->bh);
    for (auto &p : s->pending)
        for (hipEvent_t e : p.e) s->event_pool.push_back(e);
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->order_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

hipEvent_t take_event(gsim_sim *s) {
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Converts recorded event pairs into accumulated milliseconds.
gsim_status drain_timing(gsim_sim *s) {
    if (s->pending.empty()) return GSIM_OK;
    GSIM_HIP(hipStreamSynchronize(s->stream));
    for (auto &p : s->pending) {
        float first = 0.0f, second = 0.0f, whole = 0.0f;
        GSIM_HIP(hipEventElapsedTime(&first, p.e[0], p.e[1]));
        GSIM_HIP(hipEventElapsedTime(&second, p.e[1], p.e[2]));
        GSIM_HIP(hipEventElapsedTime(&whole, p.e[0], p.e[2]));
        if (p.bh) {
            s->acc.build_ms += first;
            s->acc.force_ms += second;
        } else {
            s->acc.force_ms += first;
            s->acc.integrate_ms += second;
        }
        s->acc.step_ms += whole;
        s->acc.steps += 1;
        s->acc.force_launches += 1;
        for (hipEvent_t e : p.e) s->event_pool.push_back(e);
    }
    s->pending.clear();
    return GSIM_OK;
}

// Grows the partial-sum buffer to hold `entries` float2 (the caller has
// synchronised the stream).
gsim_status ds_reserve_partial(gsim_sim *s, size_t entries) {
    if (entries <= s->partial_cap) return GSIM_OK;
    (void)hipFree(s->partial);
    s->partial = nullptr;
    s->partial_cap = 0;
    hipError_t e = hipMalloc((void **)&s->partial, sizeof(float2) * entries);
    if (e != hipSuccess)
        return fail(GSIM_ERROR_OUT_OF_MEMORY, "direct-sum partial sums (%zu bytes): %s",
                    sizeof(float2) * entries, hipGetErrorString(e));
    s->partial_cap = entries;
    return GSIM_OK;
}

// Selects the symmetric or the one-sided kernel and sizes its partial sums.
gsim_status ds_configure(gsim_sim *s, bool sym) {
    s->ds_sym = sym ? 1u : 0u;
    if (sym) {
        s->ds_plan = gsim::ds_sym_plan(s->na, s->ds_sym_cols);
        return ds_reserve_partial(s, s->ds_plan.ipart + s->ds_plan.jpart);
    }
    return ds_reserve_partial(s, s->na * s->ds_S);
}

// GSIM_TUNE_BH_ORDER = 2: whether this step's walk is a trial walk (and with
// which order); otherwise the walk keeps the order picked last.
bool order_trial_step(gsim_sim *s) {
    s->order_fresh = false;
    if (s->order_mode != 2 || !gsim::order_launch(s)) return false;
    if (s->order_wait || s->steps < s->order_next || s->order_rec >= gsim::kOrderTrial) return false;
    if (!s->order_ev[0])
        for (hipEvent_t &e : s->order_ev)
            if (hipEventCreate(&e) != hipSuccess) {
                (void)hipGetLastError();
                s->order_mode = 0;  // no events: index order
                s->trav_order = 0;
                return false;
            }
    s->trav_order = (s->order_rec & 1u) ? 0u : 1u;  // heaviest first, index, ...
    s->order_fresh = true;
    return true;
}

// Once the trial's last step has completed (never waiting for it), keeps the
// faster order.  The trial's heaviest-first steps plan afresh each time,
// where later steps plan every kOrderReplan steps: the comparison leans
// toward index order by most of a plan's cost.
void order_poll(gsim_sim *s) {
    if (!s->order_wait) return;
    const hipEvent_t last = s->order_ev[2 * gsim::kOrderTrial - 1];
    const hipError_t q = hipEventQuery(last);
    (void)hipGetLastError();  // not ready is no error of the step
    if (q == hipErrorNotReady) return;
    float ms[2] = {0.0f, 0.0f};
    bool ok = q == hipSuccess;
    for (uint32_t i = 0; ok && i < gsim::kOrderTrial; ++i) {
        float t = 0.0f;
        ok = hipEventElapsedTime(&t, s->order_ev[2 * i], s->order_ev[2 * i + 1]) == hipSuccess;
        ms[(i & 1u) ? 0 : 1] += t;
    }
    (void)hipGetLastError();
    s->trav_order = ok && ms[1] < ms[0] ? 1u : 0u;
    s->order_wait = false;
    s->order_rec = 0;
    s->order_next = s->steps + gsim::kOrderPeriod;
}

gsim_status run_steps(gsim_sim *s, uint32_t count) {
    const bool bh = s->params.algorithm == GSIM_BARNES_HUT;
    for (uint32_t k = 0; k < count; ++k) {
        const int in = s->cur, out = s->cur ^ 1;
        hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
        if (s->timing) {
            for (auto &e : ev) {
                e = take_event(s);
                if (!e) return fail(GSIM_ERROR_DEVICE, "hipEventCreate failed");
            }
            GSIM_HIP(hipEventRecord(ev[0], s->stream));
        }
        if (bh) {
            // the walk's workgroup order is fixed for the whole step; a trial
            // step is timed whole
            const bool trial = order_trial_step(s);
            if (trial) GSIM_HIP(hipEventRecord(s->order_ev[2 * s->order_rec], s->stream));
            // the build writes the next order's masses, permutation and killed
            // particles; the walk writes the rest of pos[out] and vel[out]
            GSIM_HIP(gsim::bh_launch_build(s, s->pos[in], s->mass, s->perm[in], s->pos[out],
                                           s->vel[in], s->vel[out], s->mass_next,
                                           s->perm[out]));
            if (s->timing) GSIM_HIP(hipEventRecord(ev[1], s->stream));
            GSIM_HIP(gsim::bh_launch_traverse(s, s->pos[in], s->vel[in], s->pos[out],
                                              s->vel[out], s->perm[out]));
            if (trial) {
                GSIM_HIP(hipEventRecord(s->order_ev[2 * s->order_rec + 1], s->stream));
                if (++s->order_rec == gsim::kOrderTrial) s->order_wait = true;
            } else {
                order_poll(s);
            }
            std::swap(s->mass, s->mass_next);
            s->has_tree = true;
        } else {
            GSIM_HIP(gsim::ds_launch_partial(s, s->pos[in], s->mass));
            if (s->timing) GSIM_HIP(hipEventRecord(ev[1], s->stream));
            GSIM_HIP(gsim::ds_launch_combine(s, s->pos[in], s->vel[in], s->pos[out],
                                             s->vel[out]));
        }
        if (s->timing) {
            GSIM_HIP(hipEventRecord(ev[2], s->stream));
            s->pending.push_back({{ev[0], ev[1], ev[2]}, bh});
            if (s->pending.size() > 100000) {
                gsim_status st = drain_timing(s);
                if (st != GSIM_OK) return st;
            }
        }
        s->cur = out;
        ++s->steps;
    }
    return GSIM_OK;
}

}  // namespace

extern "C" {

const char *gsim_last_error(void) { return gsim::g_error; }

const char *gsim_version(void) { return "gsim-mi355x 0.1.0 (gfx950)"; }

void gsim_default_params(gsim_params *p) {
    if (!p) return;
    // ProgramInfo defaults, Main.cpp:59-63
    p->gravitational_const = 1.0f;
    p->simulation_time = 0.001f;
    p->simulation_speed = 1.0f;
    p->softening_len = 0.2f;
    p->accuracy_parameter = 1.0f;
    p->algorithm = GSIM_DIRECT_SUM;
    p->group_width = 64;
    p->device = 0;
}

size_t gsim_required_alignment(int32_t algorithm) {
    // DirectSimulation.cpp:26-28 (WORKGROUP_SIZE) and BarnesHutSimulation.cpp:834-836
    return (algorithm == GSIM_DIRECT_SUM || algorithm == GSIM_BARNES_HUT) ? 64 : 0;
}

gsim_status gsim_create(const gsim_params *params, const float *pos_xy, const float *vel_xy,
                        const float *mass, size_t n, gsim_sim **out) {
    gsim::clear_error();
    if (!out) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_create: out is NULL");
    *out = nullptr;
    if (!params || !pos_xy || !vel_xy || !mass)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_create: NULL argument");
    if (n == 0)
        return fail(GSIM_ERROR_INVALID_ARGUMENT,
                    "The simulation must contain at least one particle!");
    if (params->algorithm != GSIM_DIRECT_SUM && params->algorithm != GSIM_BARNES_HUT)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "A valid simulation algorithm must be given!");
    const uint32_t W = params->group_width ? params->group_width : 64;
    if (W != 32 && W != 64)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "group_width must be 32 or 64 (got %u)", W);
    const float vals[] = {params->gravitational_const, params->simulation_time,
                          params->simulation_speed, params->softening_len,
                          params->accuracy_parameter};
    for (float v : vals)
        if (!std::isfinite(v))
            return fail(GSIM_ERROR_INVALID_ARGUMENT, "simulation parameters must be finite");
    const size_t na = align_up(n, 64);
    if (na > 0x7fffffffull / 4)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "particle count %zu too large", n);

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(GSIM_ERROR_NO_DEVICE, "no HIP device visible");
    if (params->device < 0 || params->device >= ndev)
        return fail(GSIM_ERROR_NO_DEVICE, "device %d not present (%d visible)", params->device, ndev);
    hipDeviceProp_t prop;
    GSIM_HIP(hipGetDeviceProperties(&prop, params->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(GSIM_ERROR_NO_DEVICE, "device %d is %s; libgsim is built for gfx950 only",
                    params->device, prop.gcnArchName);
    GSIM_HIP(hipSetDevice(params->device));

    gsim_sim *s = new (std::nothrow) gsim_sim();
    if (!s) return fail(GSIM_ERROR_OUT_OF_MEMORY, "host allocation failed");
    s->params = *params;
    s->params.group_width = W;
    s->device = params->device;
    s->n = n;
    s->na = na;
    s->W = W;
    const float G = params->gravitational_const;
    const float soft = params->softening_len, acc = params->accuracy_parameter;
    s->eps2 = soft * soft;  // DirectSimulation.cpp:179, BarnesHutSimulation.cpp:577
    s->theta2 = acc * acc;  // BarnesHutSimulation.cpp:578
    if (params->algorithm == GSIM_BARNES_HUT) {
        s->dt = params->simulation_time * params->simulation_speed;  // :575
        s->kv = G * s->dt;   // GRAVITATIONAL_CONST * SIMULATION_TIME, ForceShader :145
        s->kp = 0.5f * s->dt;
    } else {
        s->dt = params->simulation_time;  // SimShader ignores --simulation-speed
        s->kv = s->dt * G;                // SIMULATION_TIME * GRAVITATIONAL_CONST, :68
        s->kp = 0.5f * s->dt;
    }

    auto bail = [&](gsim_status st) {
        free_sim(s);
        return st;
    };
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreateWithFlags"));
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipMalloc((void **)&s->pos[k], sizeof(float2) * na);
        if (e == hipSuccess) e = hipMalloc((void **)&s->vel[k], sizeof(float2) * na);
    }
    if (e == hipSuccess) e = hipMalloc((void **)&s->mass, sizeof(float) * na);
    if (e == hipSuccess) e = hipMalloc((void **)&s->accel, sizeof(float2) * na);
    if (params->algorithm == GSIM_BARNES_HUT) {
        for (int k = 0; k < 2 && e == hipSuccess; ++k)
            e = hipMalloc((void **)&s->perm[k], sizeof(uint32_t) * na);
        if (e == hipSuccess) e = hipMalloc((void **)&s->mass_next, sizeof(float) * na);
    }
    if (e != hipSuccess) return bail(fail(GSIM_ERROR_OUT_OF_MEMORY, "device allocation failed: %s",
                                          hipGetErrorString(e)));

    // Padding particles are {0,0,0,0,0} (ParticleSystem.cpp:487-491).
    std::vector<float> hp(2 * na, 0.0f), hv(2 * na, 0.0f), hm(na, 0.0f);
    std::memcpy(hp.data(), pos_xy, sizeof(float) * 2 * n);
    std::memcpy(hv.data(), vel_xy, sizeof(float) * 2 * n);
    std::memcpy(hm.data(), mass, sizeof(float) * n);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipMemcpy(s->pos[k], hp.data(), sizeof(float) * 2 * na, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(s->vel[k], hv.data(), sizeof(float) * 2 * na, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(s->mass, hm.data(), sizeof(float) * na, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->accel, 0, sizeof(float2) * na);
    if (e == hipSuccess && s->perm[0]) {
        // the caller's order to start with
        std::vector<uint32_t> id(na);
        for (size_t i = 0; i < na; ++i) id[i] = (uint32_t)i;
        for (int k = 0; k < 2 && e == hipSuccess; ++k)
            e = hipMemcpy(s->perm[k], id.data(), sizeof(uint32_t) * na, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return bail(hip_fail(e, "upload"));

    if (params->algorithm == GSIM_DIRECT_SUM) {
        gsim::ds_pick_shape(na, &s->ds_R, &s->ds_S);
        gsim_status st = ds_configure(s, gsim::ds_sym_default(na));
        if (st != GSIM_OK) return bail(st);
    } else {
        e = gsim::bh_alloc(&s->bh, na, gsim::kTreeDepth);
        if (e != hipSuccess)
            return bail(fail(GSIM_ERROR_OUT_OF_MEMORY, "Barnes-Hut buffers: %s", hipGetErrorString(e)));
        // the heavy list's walks run on a second stream (default priority),
        // forked from and joined back to s->stream by events
        e = hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming);
        if (e != hipSuccess) return bail(hip_fail(e, "side stream"));
    }
    *out = s;
    return GSIM_OK;
}

gsim_status gsim_step(gsim_sim *s, uint32_t count) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_step: NULL handle");
    if (count == 0) return GSIM_OK;  // BarnesHutSimulation.cpp:851-852
    GSIM_HIP(hipSetDevice(s->device));
    return run_steps(s, count);
}

gsim_status gsim_synchronize(gsim_sim *s) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_synchronize: NULL handle");
    GSIM_HIP(hipStreamSynchronize(s->stream));
    return GSIM_OK;
}

gsim_status gsim_read(gsim_sim *s, float *pos_xy, float *vel_xy, float *mass) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_read: NULL handle");
    GSIM_HIP(hipSetDevice(s->device));
    GSIM_HIP(hipStreamSynchronize(s->stream));
    if (!s->perm[0]) {  // direct sum: the caller's order
        if (pos_xy) GSIM_HIP(hipMemcpy(pos_xy, s->pos[s->cur], sizeof(float2) * s->n, hipMemcpyDeviceToHost));
        if (vel_xy) GSIM_HIP(hipMemcpy(vel_xy, s->vel[s->cur], sizeof(float2) * s->n, hipMemcpyDeviceToHost));
        if (mass) GSIM_HIP(hipMemcpy(mass, s->mass, sizeof(float) * s->n, hipMemcpyDeviceToHost));
        return GSIM_OK;
    }
    // Barnes-Hut: the state is in the last step's sorted order
    const size_t na = s->na;
    std::vector<uint32_t> perm(na);
    GSIM_HIP(hipMemcpy(perm.data(), s->perm[s->cur], sizeof(uint32_t) * na, hipMemcpyDeviceToHost));
    std::vector<float> tmp(2 * na);
    const float *src[3] = {(const float *)s->pos[s->cur], (const float *)s->vel[s->cur], s->mass};
    float *dst[3] = {pos_xy, vel_xy, mass};
    for (int a = 0; a < 3; ++a) {
        if (!dst[a]) continue;
        const size_t w = a < 2 ? 2 : 1;
        GSIM_HIP(hipMemcpy(tmp.data(), src[a], sizeof(float) * w * na, hipMemcpyDeviceToHost));
        unpermute(perm, tmp.data(), dst[a], w, s->n);
    }
    return GSIM_OK;
}

gsim_status gsim_destroy(gsim_sim *s) {
    gsim::clear_error();
    if (!s) return GSIM_OK;
    (void)hipSetDevice(s->device);
    free_sim(s);
    return GSIM_OK;
}

size_t gsim_particle_count(const gsim_sim *s) { return s ? s->n : 0; }
size_t gsim_aligned_particle_count(const gsim_sim *s) { return s ? s->na : 0; }
uint64_t gsim_steps_done(const gsim_sim *s) { return s ? s->steps : 0; }

gsim_status gsim_set_debug(gsim_sim *s, uint32_t flags) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_set_debug: NULL handle");
    s->debug_flags = flags;
    return GSIM_OK;
}

gsim_status gsim_set_ds_shape(gsim_sim *s, uint32_t R, uint32_t S) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_set_ds_shape: NULL handle");
    if (s->params.algorithm != GSIM_DIRECT_SUM)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_set_ds_shape: not a direct-sum handle");
    if (R != 1 && R != 2 && R != 4)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "targets per lane must be 1, 2 or 4");
    if (S == 0) {
        uint32_t r;
        gsim::ds_pick_shape(s->na, &r, &S);
    }
    if (S > 256 || (size_t)S * 64 > s->na)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "slices must be <= 256 and <= Na/64");
    GSIM_HIP(hipSetDevice(s->device));
    GSIM_HIP(hipStreamSynchronize(s->stream));
    s->ds_R = R;
    s->ds_S = S;
    return ds_configure(s, false);
}

gsim_status gsim_set_tuning(gsim_sim *s, int32_t knob, uint32_t value) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_set_tuning: NULL handle");
    // A W = 32 handle has one traversal launch (bh_traverse32_kernel): the
    // W = 64 launch-shape knobs would be silently ignored, so they are refused.
    if (s->params.algorithm == GSIM_BARNES_HUT && s->W == 32 &&
        (knob == GSIM_TUNE_BH_WAVES_PER_SIMD || knob == GSIM_TUNE_BH_TRAVERSAL ||
         knob == GSIM_TUNE_BH_SPLIT || knob == GSIM_TUNE_BH_SEGMENTS))
        return fail(GSIM_ERROR_INVALID_ARGUMENT,
                    "tuning knob %d selects W = 64 traversal launches; a group width 32 "
                    "handle has one launch shape", knob);
    switch (knob) {
        case GSIM_TUNE_DS_RCP:
            s->ds_rcp = value ? 1u : 0u;
            return GSIM_OK;
        case GSIM_TUNE_BH_WAVES_PER_SIMD:
            if (value < 1 || value > 8)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "waves per SIMD must be in [1, 8]");
            s->trav_waves_per_simd = value;
            return GSIM_OK;
        case GSIM_TUNE_BH_TRAVERSAL:
            if (value > 3 && value != 6)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "traversal mode must be 0..3 or 6");
            s->trav_mode = value;
            return GSIM_OK;
        case GSIM_TUNE_BH_ADAPTIVE_BOX:
            if (s->params.algorithm != GSIM_BARNES_HUT)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "not a Barnes-Hut handle");
            GSIM_HIP(hipSetDevice(s->device));
            GSIM_HIP(hipStreamSynchronize(s->stream));
            s->bh_box = value ? 1u : 0u;
            return GSIM_OK;
        case GSIM_TUNE_BH_LEAF_BITS: {
            if (s->params.algorithm != GSIM_BARNES_HUT)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "not a Barnes-Hut handle");
            if (value != gsim::kTreeDepth && value != gsim::kDeepBits)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "leaf bits must be 10 or 15");
            if (value == s->bh.bits) return GSIM_OK;
            GSIM_HIP(hipSetDevice(s->device));
            GSIM_HIP(hipStreamSynchronize(s->stream));
            // the node arrays are sized for the depth: reallocate them
            gsim::bh_free(&s->bh);
            s->has_tree = false;
            const hipError_t e = gsim::bh_alloc(&s->bh, s->na, value);
            if (e != hipSuccess) {
                gsim::bh_free(&s->bh);
                (void)gsim::bh_alloc(&s->bh, s->na, gsim::kTreeDepth);
                return fail(GSIM_ERROR_OUT_OF_MEMORY, "Barnes-Hut buffers for %u-bit leaves: %s",
                            value, hipGetErrorString(e));
            }
            return GSIM_OK;
        }
        case GSIM_TUNE_BH_SPLIT:
            if (value > 1000) return fail(GSIM_ERROR_INVALID_ARGUMENT, "split threshold must be in [0, 1000]");
            s->trav_split = value;
            return GSIM_OK;
        case GSIM_TUNE_BH_ORDER:
            if (value > 2) return fail(GSIM_ERROR_INVALID_ARGUMENT, "group order must be 0, 1 or 2");
            s->order_mode = value;
            s->trav_order = value == 1 ? 1u : 0u;
            s->order_rec = 0;
            s->order_wait = false;
            s->order_next = s->steps + 1;
            return GSIM_OK;
        case GSIM_TUNE_BH_XCD_ORDER:
            if (value > 4096)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "XCD run length must be in [0, 4096]");
            s->xcd_order = value;
            return GSIM_OK;
        case GSIM_TUNE_BH_SEGMENTS:
            if (value < 1 || value > 65536)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "segments must be in [1, 65536]");
            s->trav_segments = value;
            return GSIM_OK;
        case GSIM_TUNE_DS_SYMMETRIC:
        case GSIM_TUNE_DS_SYM_COLS: {
            if (s->params.algorithm != GSIM_DIRECT_SUM)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "not a direct-sum handle");
            if (knob == GSIM_TUNE_DS_SYM_COLS && value > 2)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "column blocks per wave must be 0, 1 or 2");
            GSIM_HIP(hipSetDevice(s->device));
            GSIM_HIP(hipStreamSynchronize(s->stream));
            bool sym = s->ds_sym != 0;
            if (knob == GSIM_TUNE_DS_SYMMETRIC) sym = value != 0;
            else s->ds_sym_cols = value;
            return ds_configure(s, sym);
        }
        default:
            return fail(GSIM_ERROR_INVALID_ARGUMENT, "unknown tuning knob %d", knob);
    }
}

gsim_status gsim_set_timing(gsim_sim *s, int32_t enable) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_set_timing: NULL handle");
    s->timing = enable != 0;
    return GSIM_OK;
}

gsim_status gsim_get_timing(gsim_sim *s, gsim_timing *out, int32_t reset) {
    gsim::clear_error();
    if (!s || !out) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_get_timing: NULL argument");
    GSIM_HIP(hipSetDevice(s->device));
    gsim_status st = drain_timing(s);
    if (st != GSIM_OK) return st;
    *out = s->acc;
    if (reset) s->acc = gsim_timing{};
    return GSIM_OK;
}

gsim_status gsim_debug_dump(gsim_sim *s, int32_t what, void *buf, size_t bytes, size_t *written) {
    gsim::clear_error();
    if (!s || !written) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_debug_dump: NULL argument");
    GSIM_HIP(hipSetDevice(s->device));
    GSIM_HIP(hipStreamSynchronize(s->stream));
    const size_t na = s->na;
    if (what == GSIM_DUMP_ACCEL) {
        *written = sizeof(float2) * na;
        if (!buf || bytes < *written) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
        GSIM_HIP(hipMemcpy(buf, s->accel, *written, hipMemcpyDeviceToHost));
        if (s->perm[0]) {  // Barnes-Hut writes it in the last step's output order
            std::vector<uint32_t> perm(na);
            GSIM_HIP(hipMemcpy(perm.data(), s->perm[s->cur], sizeof(uint32_t) * na,
                               hipMemcpyDeviceToHost));
            std::vector<float> tmp((const float *)buf, (const float *)buf + 2 * na);
            unpermute(perm, tmp.data(), (float *)buf, 2, na);
        }
        return GSIM_OK;
    }
    if (s->params.algorithm != GSIM_BARNES_HUT)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "dump %d needs a Barnes-Hut handle", what);
    gsim::BhCounters ctr{};
    GSIM_HIP(hipMemcpy(&ctr, s->bh.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
    if (ctr.sort_stall)
        return fail(GSIM_ERROR_DEVICE, "radix sort look-back gave up on %u status words: the tree is invalid",
                    ctr.sort_stall);
    const size_t nodes = s->has_tree ? ctr.n_nodes : 0;
    switch (what) {
        case GSIM_DUMP_CELLS: {
            *written = sizeof(uint32_t) * na;
            if (!buf || bytes < *written) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
            std::vector<uint32_t> pairs(2 * na);  // {cell, slot} by particle index
            GSIM_HIP(hipMemcpy(pairs.data(), s->bh.cells[s->bh.kcells], sizeof(uint32_t) * 2 * na,
                               hipMemcpyDeviceToHost));
            uint32_t *u = (uint32_t *)buf;
            for (size_t i = 0; i < na; ++i) u[i] = pairs[2 * i];
            for (size_t i = 0; i < na; ++i)
                if (u[i] >= (1u << (2 * s->bh.bits))) u[i] = 0xffffffffu;
            return GSIM_OK;
        }
        case GSIM_DUMP_BOX:
            *written = 4 * sizeof(float);
            if (!buf || bytes < *written) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
            if (!s->bh_box) return fail(GSIM_ERROR_INVALID_ARGUMENT, "the adaptive box is off");
            GSIM_HIP(hipMemcpy(buf, s->bh.box, *written, hipMemcpyDeviceToHost));
            return GSIM_OK;
        case GSIM_DUMP_MORTON30: {
            if (s->bh_box)
                return fail(GSIM_ERROR_INVALID_ARGUMENT, "30-bit codes are of the fixed domain only");
            *written = sizeof(uint32_t) * na;
            if (!buf || bytes < *written) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
            // positions the last step read (its input set)
            const int in = s->steps ? (s->cur ^ 1) : s->cur;
            GSIM_HIP(gsim::bh_launch_morton30(s, s->pos[in], s->bh.keys[1]));
            GSIM_HIP(hipStreamSynchronize(s->stream));
            std::vector<uint32_t> perm(na), code(na);
            GSIM_HIP(hipMemcpy(code.data(), s->bh.keys[1], *written, hipMemcpyDeviceToHost));
            GSIM_HIP(hipMemcpy(perm.data(), s->perm[in], sizeof(uint32_t) * na, hipMemcpyDeviceToHost));
            uint32_t *u = (uint32_t *)buf;
            for (size_t slot = 0; slot < na; ++slot) u[perm[slot]] = code[slot];
            return GSIM_OK;
        }
        case GSIM_DUMP_SORTED_SRC:
            *written = sizeof(uint32_t) * na;
            if (!buf || bytes < *written) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
            GSIM_HIP(hipMemcpy(buf, s->bh.sorted_src, *written, hipMemcpyDeviceToHost));
            {
                // slots of the last step's input order -> the caller's indices
                std::vector<uint32_t> perm(na);
                const int in = s->steps ? (s->cur ^ 1) : s->cur;
                GSIM_HIP(hipMemcpy(perm.data(), s->perm[in], sizeof(uint32_t) * na,
                                   hipMemcpyDeviceToHost));
                uint32_t *u = (uint32_t *)buf;
                for (size_t k = 0; k < na; ++k)
                    if (u[k] < na) u[k] = perm[u[k]];
            }
            return GSIM_OK;
        case GSIM_DUMP_NODE_TOTAL:
            *written = sizeof(uint32_t);
            if (!buf || bytes < *written) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
            *(uint32_t *)buf = (uint32_t)nodes;
            return GSIM_OK;
        case GSIM_DUMP_COUNTERS:
            *written = 2 * sizeof(uint64_t);
            if (!buf || bytes < *written) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
            {
                const size_t groups = na / s->W;
                std::vector<unsigned long long> v(groups), a(groups);
                if (groups) {
                    GSIM_HIP(hipMemcpy(v.data(), s->bh.gvis, groups * 8, hipMemcpyDeviceToHost));
                    GSIM_HIP(hipMemcpy(a.data(), s->bh.gacc, groups * 8, hipMemcpyDeviceToHost));
                }
                uint64_t sv = 0, sa = 0;
                for (size_t g = 0; g < groups; ++g) {
                    sv += v[g];
                    sa += a[g];
                }
                ((uint64_t *)buf)[0] = sv;
                ((uint64_t *)buf)[1] = sa;
            }
            return GSIM_OK;
        case GSIM_DUMP_COUNTS:
            *written = sizeof(uint32_t) * nodes;
            if (bytes < *written || (!buf && nodes)) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
            if (nodes) GSIM_HIP(hipMemcpy(buf, s->bh.ncount, *written, hipMemcpyDeviceToHost));
            return GSIM_OK;
        case GSIM_DUMP_RADII:
        case GSIM_DUMP_NODE_POS:
        case GSIM_DUMP_NODE_MASS: {
            const size_t per = what == GSIM_DUMP_NODE_POS ? 2 : 1;
            *written = sizeof(float) * per * nodes;
            if (bytes < *written || (!buf && nodes)) return fail(GSIM_ERROR_CAPACITY, "buffer too small");
            if (!nodes) return GSIM_OK;
            std::vector<float4> tmp(nodes);
            GSIM_HIP(hipMemcpy(tmp.data(), s->bh.node4, sizeof(float4) * nodes, hipMemcpyDeviceToHost));
            float *f = (float *)buf;
            for (size_t k = 0; k < nodes; ++k) {
                if (what == GSIM_DUMP_RADII) f[k] = tmp[k].w > 0.0f ? tmp[k].w : 0.0f;  // run length -> 0
                else if (what == GSIM_DUMP_NODE_MASS) f[k] = tmp[k].z;
                else { f[2 * k] = tmp[k].x; f[2 * k + 1] = tmp[k].y; }
            }
            return GSIM_OK;
        }
        default:
            return fail(GSIM_ERROR_INVALID_ARGUMENT, "unknown dump selector %d", what);
    }
}

}  // extern "C"
