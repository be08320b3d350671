// C-ABI entry points of the batched form: an arena of many independent
// systems stepped together, each step a replay of a captured hipGraph.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "batch.hpp"

using gsim::fail;

namespace {

gsim_status hip_fail(hipError_t e, const char *what) {
    return fail(GSIM_ERROR_DEVICE, "%s failed: %s", what, hipGetErrorString(e));
}

#define GSIM_HIP(call)                                      \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #call);   \
    } while (0)

size_t align64(size_t n) { return (n + 63) / 64 * 64; }

// Cell nodes a system of na particles can produce (at most min(4^d, na/2) per level).
size_t cell_bound(size_t na) {
    size_t cells = 0;
    for (int d = 0; d <= 10; ++d) cells += std::min<size_t>((size_t)1 << (2 * d), na / 2);
    return cells;
}

void destroy_graphs(gsim_batch *b) {
    for (auto &g : b->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void free_batch(gsim_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    destroy_graphs(b);
    gsim::BatchDevice &d = b->d;
    void *ptrs[] = {d.pos[0], d.pos[1], d.vel[0], d.vel[1], d.mass, d.partial, d.desc, d.remaining,
                    d.bh_sys, d.bh_prefix, d.ds_sys, d.ds_prefix, d.items, d.counts, d.stats,
                    d.keys[0], d.keys[1], d.idx[0], d.idx[1], d.status, d.digit_total,
                    d.block_cells, d.block_leaves, d.sdep, d.cpos, d.sorted_src, d.prod,
                    d.leaf_sum, d.leaf_first, d.leaf_key, d.tile_range, d.live_tiles, d.top, d.node4,
                    d.ncount, d.breaker, d.wave_acc, d.ichunk};
    for (void *p : ptrs) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

bool take_range(gsim_batch *b, uint32_t len, uint32_t *off) {
    for (auto it = b->free_ranges.begin(); it != b->free_ranges.end(); ++it) {
        if (it->second >= len) {
            *off = it->first;
            const uint32_t rest = it->second - len, start = it->first + len;
            b->free_ranges.erase(it);
            if (rest) b->free_ranges[start] = rest;
            return true;
        }
    }
    return false;
}

void give_range(gsim_batch *b, uint32_t off, uint32_t len) {
    auto it = b->free_ranges.emplace(off, len).first;
    auto next = std::next(it);
    if (next != b->free_ranges.end() && it->first + it->second == next->first) {
        it->second += next->second;
        b->free_ranges.erase(next);
    }
    if (it != b->free_ranges.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            b->free_ranges.erase(it);
        }
    }
}

// Rebuilds and uploads the membership arrays (descriptors, ranks, work items).
gsim_status upload_membership(gsim_batch *b) {
    std::vector<gsim::SysDesc> desc;
    std::vector<uint32_t> bh_sys, bh_prefix{0}, ds_sys, ds_prefix{0};
    std::vector<uint2> items;
    std::vector<uint32_t> remaining;
    uint32_t part_off = 0;
    GSIM_HIP(hipStreamSynchronize(b->stream));  // step counts below are exact from here
    for (auto &kv : b->systems) {
        gsim::BatchSystem &s = kv.second;
        const uint64_t taken = b->steps - s.added_at;
        s.steps = s.target == ~0ull ? taken : std::min<uint64_t>(taken, s.target);
        remaining.push_back(s.target == ~0ull ? gsim::kUnlimitedSteps
                                              : (uint32_t)std::min<uint64_t>(s.target - s.steps,
                                                                             gsim::kUnlimitedSteps - 1));
        gsim::SysDesc d{};
        d.off = s.off;
        d.na = s.na;
        d.n = s.n;
        d.algo = (uint32_t)s.params.algorithm;
        const float soft = s.params.softening_len, acc = s.params.accuracy_parameter;
        const float G = s.params.gravitational_const;
        d.eps2 = soft * soft;
        d.theta2 = acc * acc;
        if (d.algo == GSIM_BARNES_HUT) {
            d.dt = s.params.simulation_time * s.params.simulation_speed;
            d.kv = G * d.dt;
        } else {
            d.dt = s.params.simulation_time;
            d.kv = d.dt * G;
        }
        d.kp = 0.5f * d.dt;
        const uint32_t di = (uint32_t)desc.size();
        if (d.algo == GSIM_BARNES_HUT) {
            d.bh_rank = (uint32_t)bh_sys.size();
            d.sorted_base = bh_prefix.back();
            bh_sys.push_back(di);
            bh_prefix.push_back(bh_prefix.back() + s.na);
            // one work item per vote group: 64 sorted positions, or 32 at W = 32
            d.w = s.params.group_width;
            for (uint32_t g = 0; g < s.na / d.w; ++g) items.push_back(make_uint2(di, g));
        } else {
            uint32_t R = 0, S = 0;
            gsim::ds_pick_shape(s.na, &R, &S);
            const uint32_t jlen = (uint32_t)(((s.na + S - 1) / S + 63) / 64 * 64);
            const uint32_t slices = (s.na + jlen - 1) / jlen;
            d.ds_jlen = jlen;
            d.ds_slices = slices;
            d.ds_part_off = part_off;
            part_off += slices * s.na;
            ds_sys.push_back(di);
            ds_prefix.push_back(ds_prefix.back() + s.na);
            for (uint32_t tb = 0; tb < (s.na + 127) / 128; ++tb)
                for (uint32_t sl = 0; sl < slices; ++sl) items.push_back(make_uint2(di, (tb << 16) | sl));
        }
        desc.push_back(d);
    }
    if (items.size() > b->item_capacity || part_off > b->partial_capacity)
        return fail(GSIM_ERROR_CAPACITY, "batch work list exceeds its capacity");
    gsim::BatchCounts c{};
    c.n_sys = (uint32_t)desc.size();
    c.n_bh = (uint32_t)bh_sys.size();
    c.n_bh_particles = bh_prefix.back();
    c.n_ds = (uint32_t)ds_sys.size();
    c.n_ds_particles = ds_prefix.back();
    c.n_items = (uint32_t)items.size();
    // the acceptance and visit totals carry over
    GSIM_HIP(hipMemcpy(&c.accepts, &b->d.counts->accepts, sizeof(c.accepts), hipMemcpyDeviceToHost));
    GSIM_HIP(hipMemcpy(&c.visits, &b->d.counts->visits, sizeof(c.visits), hipMemcpyDeviceToHost));
    auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    GSIM_HIP(up(b->d.desc, desc.data(), desc.size() * sizeof(gsim::SysDesc)));
    GSIM_HIP(up(b->d.bh_sys, bh_sys.data(), bh_sys.size() * 4));
    GSIM_HIP(up(b->d.bh_prefix, bh_prefix.data(), bh_prefix.size() * 4));
    GSIM_HIP(up(b->d.ds_sys, ds_sys.data(), ds_sys.size() * 4));
    GSIM_HIP(up(b->d.ds_prefix, ds_prefix.data(), ds_prefix.size() * 4));
    GSIM_HIP(up(b->d.items, items.data(), items.size() * sizeof(uint2)));
    // the work items changed: their last-step chunk counts no longer apply
    GSIM_HIP(hipMemsetAsync(b->d.ichunk, 0, sizeof(uint32_t) * b->item_capacity, b->stream));
    GSIM_HIP(up(b->d.remaining, remaining.data(), remaining.size() * 4));
    GSIM_HIP(up(b->d.counts, &c, sizeof(c)));
    b->h_desc = desc;
    b->h_counts = c;
    b->dirty = false;
    return GSIM_OK;
}

gsim_status launch_steps(gsim_batch *b, uint32_t count) {
    for (uint32_t k = 0; k < count; ++k) {
        const int in = b->cur;
        if (!b->graph[in]) {
            hipGraph_t g = nullptr;
            GSIM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = gsim::batch_launch_step(b, in, in ^ 1);
            hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hip_fail(e, "batched step launch");
            }
            GSIM_HIP(e2);
            e = hipGraphInstantiate(&b->graph[in], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            GSIM_HIP(e);
        }
        GSIM_HIP(hipGraphLaunch(b->graph[in], b->stream));
        b->cur ^= 1;
        ++b->steps;
    }
    return GSIM_OK;
}

}  // namespace

extern "C" {

gsim_status gsim_batch_create(int32_t device, size_t capacity_particles, uint32_t max_systems,
                              gsim_batch **out) {
    gsim::clear_error();
    if (!out) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_create: out is NULL");
    *out = nullptr;
    if (capacity_particles == 0 || max_systems == 0 || max_systems > gsim::kBatchMaxSystems)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "capacity must be > 0 and max_systems in [1, %u]",
                    gsim::kBatchMaxSystems);
    const size_t cap = align64(capacity_particles);
    if (cap > gsim::kBatchMaxParticles)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "capacity %zu exceeds the arena maximum of %zu particles",
                    capacity_particles, gsim::kBatchMaxParticles);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(GSIM_ERROR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(GSIM_ERROR_NO_DEVICE, "device %d not present", device);
    hipDeviceProp_t prop;
    GSIM_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(GSIM_ERROR_NO_DEVICE, "device %d is %s; libgsim is built for gfx950 only", device,
                    prop.gcnArchName);
    GSIM_HIP(hipSetDevice(device));
    gsim_batch *b = new (std::nothrow) gsim_batch();
    if (!b) return fail(GSIM_ERROR_OUT_OF_MEMORY, "host allocation failed");
    b->device = device;
    b->capacity = cap;
    b->max_systems = max_systems;
    b->node_capacity = 3 * cap + 64;
    b->partial_capacity = 4 * cap;
    b->item_capacity = cap / 8 + cap / 64 + 16 * (size_t)max_systems + 64;
    b->tiles_cap = (cap + gsim::kBatchSortTileMin - 1) / gsim::kBatchSortTileMin;
    b->blocks_cap = (cap + 255) / 256;
    b->free_ranges[0] = (uint32_t)cap;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    gsim::BatchDevice &d = b->d;
#define GSIM_BALLOC(ptr, count)                                                          \
    do {                                                                                 \
        if (e == hipSuccess) e = hipMalloc((void **)&(ptr), sizeof(*(ptr)) * (count));   \
    } while (0)
    for (int k = 0; k < 2; ++k) {
        GSIM_BALLOC(d.pos[k], cap);
        GSIM_BALLOC(d.vel[k], cap);
        GSIM_BALLOC(d.keys[k], cap);
        GSIM_BALLOC(d.idx[k], cap);
    }
    GSIM_BALLOC(d.mass, cap);
    GSIM_BALLOC(d.partial, b->partial_capacity);
    GSIM_BALLOC(d.desc, max_systems);
    GSIM_BALLOC(d.bh_sys, max_systems);
    GSIM_BALLOC(d.bh_prefix, max_systems + 1);
    GSIM_BALLOC(d.ds_sys, max_systems);
    GSIM_BALLOC(d.ds_prefix, max_systems + 1);
    GSIM_BALLOC(d.items, b->item_capacity);
    GSIM_BALLOC(d.remaining, max_systems);
    GSIM_BALLOC(d.counts, 1);
    GSIM_BALLOC(d.stats, max_systems);
    GSIM_BALLOC(d.status, 2 * 256 * b->tiles_cap);
    GSIM_BALLOC(d.digit_total, 4 * 256);
    GSIM_BALLOC(d.block_cells, b->blocks_cap);
    GSIM_BALLOC(d.block_leaves, b->blocks_cap);
    GSIM_BALLOC(d.sdep, cap);
    GSIM_BALLOC(d.cpos, cap);
    GSIM_BALLOC(d.sorted_src, cap);
    GSIM_BALLOC(d.prod, cap);
    GSIM_BALLOC(d.leaf_sum, cap);
    GSIM_BALLOC(d.leaf_first, cap + 1);
    GSIM_BALLOC(d.leaf_key, cap);
    GSIM_BALLOC(d.tile_range, 2 * (size_t)gsim::kTiles * max_systems);
    GSIM_BALLOC(d.live_tiles, (size_t)gsim::kTiles * max_systems);
    GSIM_BALLOC(d.top, (size_t)gsim::kTiles * max_systems);
    GSIM_BALLOC(d.node4, b->node_capacity);
    GSIM_BALLOC(d.ncount, b->node_capacity);
    GSIM_BALLOC(d.breaker, b->node_capacity);
    GSIM_BALLOC(d.wave_acc, 2 * gsim::kMaxForceWaves);
    GSIM_BALLOC(d.ichunk, b->item_capacity);
#undef GSIM_BALLOC
    if (e == hipSuccess) e = hipMemset(d.counts, 0, sizeof(gsim::BatchCounts));
    if (e == hipSuccess) e = hipMemset(d.wave_acc, 0, sizeof(unsigned long long) * 2 * gsim::kMaxForceWaves);
    if (e == hipSuccess) e = hipMemset(d.ichunk, 0, sizeof(uint32_t) * b->item_capacity);
    if (e == hipSuccess) e = hipMemset(d.ncount, 0, sizeof(uint32_t) * b->node_capacity);
    if (e == hipSuccess) e = hipMemset(d.node4, 0, sizeof(float4) * b->node_capacity);
    if (e != hipSuccess) {
        gsim_status st = fail(GSIM_ERROR_OUT_OF_MEMORY, "batch allocation failed: %s", hipGetErrorString(e));
        free_batch(b);
        return st;
    }
    *out = b;
    return GSIM_OK;
}

gsim_status gsim_batch_add(gsim_batch *b, const gsim_params *p, const float *pos_xy,
                           const float *vel_xy, const float *mass, size_t n, uint32_t *system_id) {
    gsim::clear_error();
    if (!b || !p || !pos_xy || !vel_xy || !mass || !system_id)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_add: NULL argument");
    if (n == 0) return fail(GSIM_ERROR_INVALID_ARGUMENT, "The simulation must contain at least one particle!");
    if (p->algorithm != GSIM_DIRECT_SUM && p->algorithm != GSIM_BARNES_HUT)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "A valid simulation algorithm must be given!");
    const uint32_t W = p->group_width ? p->group_width : 64;
    if (W != 32 && W != 64)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "group_width must be 32 or 64 (got %u)", W);
    const float vals[] = {p->gravitational_const, p->simulation_time, p->simulation_speed,
                          p->softening_len, p->accuracy_parameter};
    for (float v : vals)
        if (!std::isfinite(v)) return fail(GSIM_ERROR_INVALID_ARGUMENT, "simulation parameters must be finite");
    if (b->systems.size() >= b->max_systems)
        return fail(GSIM_ERROR_CAPACITY, "batch holds %u systems already", b->max_systems);
    const size_t na = align64(n);
    // Large direct-sum systems fill the GPU alone and are stepped alone with
    // the symmetric kernel; the batch keeps the one-sided kernel for the rest.
    if (p->algorithm == GSIM_DIRECT_SUM && na >= gsim::kDsSymMinParticles)
        return fail(GSIM_ERROR_INVALID_ARGUMENT,
                    "direct-sum systems of %zu or more padded particles are stepped alone (gsim_create)",
                    gsim::kDsSymMinParticles);
    const size_t nodes = p->algorithm == GSIM_BARNES_HUT ? na + cell_bound(na) : 0;
    if (b->nodes_reserved + nodes > b->node_capacity)
        return fail(GSIM_ERROR_CAPACITY, "batch node arrays are full");
    size_t part = 0;
    if (p->algorithm == GSIM_DIRECT_SUM) {
        uint32_t R = 0, S = 0;
        gsim::ds_pick_shape(na, &R, &S);
        const size_t jlen = ((na + S - 1) / S + 63) / 64 * 64;
        part = ((na + jlen - 1) / jlen) * na;
        if (b->partial_used + part > b->partial_capacity)
            return fail(GSIM_ERROR_CAPACITY, "batch partial-sum space is full");
    }
    uint32_t off = 0;
    if (!take_range(b, (uint32_t)na, &off))
        return fail(GSIM_ERROR_CAPACITY, "no free arena range of %zu particles", na);
    GSIM_HIP(hipSetDevice(b->device));
    GSIM_HIP(hipStreamSynchronize(b->stream));
    std::vector<float> hp(2 * na, 0.0f), hv(2 * na, 0.0f), hm(na, 0.0f);
    std::memcpy(hp.data(), pos_xy, sizeof(float) * 2 * n);
    std::memcpy(hv.data(), vel_xy, sizeof(float) * 2 * n);
    std::memcpy(hm.data(), mass, sizeof(float) * n);
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipMemcpy(b->d.pos[k] + off, hp.data(), sizeof(float2) * na, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d.vel[k] + off, hv.data(), sizeof(float2) * na, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(b->d.mass + off, hm.data(), sizeof(float) * na, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        give_range(b, off, (uint32_t)na);
        return hip_fail(e, "batch upload");
    }
    gsim::BatchSystem s;
    s.params = *p;
    s.params.group_width = W;
    s.off = off;
    s.na = (uint32_t)na;
    s.n = (uint32_t)n;
    s.node_reserve = nodes;
    s.added_at = b->steps;
    const uint32_t id = b->next_id++;
    b->systems[id] = s;
    b->nodes_reserved += nodes;
    b->partial_used += part;
    b->dirty = true;
    *system_id = id;
    return GSIM_OK;
}

gsim_status gsim_batch_remove(gsim_batch *b, uint32_t id) {
    gsim::clear_error();
    if (!b) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_remove: NULL handle");
    auto it = b->systems.find(id);
    if (it == b->systems.end()) return fail(GSIM_ERROR_INVALID_ARGUMENT, "no system %u in the batch", id);
    GSIM_HIP(hipSetDevice(b->device));
    GSIM_HIP(hipStreamSynchronize(b->stream));
    const gsim::BatchSystem &s = it->second;
    give_range(b, s.off, s.na);
    b->nodes_reserved -= s.node_reserve;
    if (s.params.algorithm == GSIM_DIRECT_SUM) {
        uint32_t R = 0, S = 0;
        gsim::ds_pick_shape(s.na, &R, &S);
        const size_t jlen = ((s.na + S - 1) / S + 63) / 64 * 64;
        b->partial_used -= ((s.na + jlen - 1) / jlen) * s.na;
    }
    b->systems.erase(it);
    b->dirty = true;
    return GSIM_OK;
}

gsim_status gsim_batch_step(gsim_batch *b, uint32_t count) {
    gsim::clear_error();
    if (!b) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_step: NULL handle");
    if (count == 0) return GSIM_OK;
    GSIM_HIP(hipSetDevice(b->device));
    if (b->dirty) {
        gsim_status st = upload_membership(b);
        if (st != GSIM_OK) return st;
    }
    return launch_steps(b, count);
}

gsim_status gsim_batch_set_steps(gsim_batch *b, uint32_t id, uint64_t more_steps) {
    gsim::clear_error();
    if (!b) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_set_steps: NULL handle");
    auto it = b->systems.find(id);
    if (it == b->systems.end()) return fail(GSIM_ERROR_INVALID_ARGUMENT, "no system %u in the batch", id);
    GSIM_HIP(hipSetDevice(b->device));
    GSIM_HIP(hipStreamSynchronize(b->stream));
    gsim::BatchSystem &s = it->second;
    const uint64_t taken = b->steps - s.added_at;
    const uint64_t done = s.target == ~0ull ? taken : std::min<uint64_t>(taken, s.target);
    // re-anchor so that `done` steps stay counted and `more_steps` remain
    s.added_at = b->steps - done;
    s.target = more_steps == ~0ull ? ~0ull : done + more_steps;
    b->dirty = true;
    return GSIM_OK;
}

gsim_status gsim_batch_steps_done(gsim_batch *b, uint32_t id, uint64_t *steps) {
    gsim::clear_error();
    if (!b || !steps) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_steps_done: NULL argument");
    auto it = b->systems.find(id);
    if (it == b->systems.end()) return fail(GSIM_ERROR_INVALID_ARGUMENT, "no system %u in the batch", id);
    const gsim::BatchSystem &s = it->second;
    const uint64_t taken = b->steps - s.added_at;
    *steps = s.target == ~0ull ? taken : std::min<uint64_t>(taken, s.target);
    return GSIM_OK;
}

gsim_status gsim_batch_accepts(gsim_batch *b, uint64_t *accepts) {
    gsim::clear_error();
    if (!b || !accepts) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_accepts: NULL argument");
    GSIM_HIP(hipSetDevice(b->device));
    GSIM_HIP(hipStreamSynchronize(b->stream));
    unsigned long long a = 0;
    GSIM_HIP(hipMemcpy(&a, &b->d.counts->accepts, sizeof(a), hipMemcpyDeviceToHost));
    *accepts = a;
    return GSIM_OK;
}

gsim_status gsim_batch_visits(gsim_batch *b, uint64_t *visits) {
    gsim::clear_error();
    if (!b || !visits) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_visits: NULL argument");
    GSIM_HIP(hipSetDevice(b->device));
    GSIM_HIP(hipStreamSynchronize(b->stream));
    unsigned long long v = 0;
    GSIM_HIP(hipMemcpy(&v, &b->d.counts->visits, sizeof(v), hipMemcpyDeviceToHost));
    *visits = v;
    return GSIM_OK;
}

gsim_status gsim_batch_synchronize(gsim_batch *b) {
    gsim::clear_error();
    if (!b) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_synchronize: NULL handle");
    GSIM_HIP(hipStreamSynchronize(b->stream));
    return GSIM_OK;
}

gsim_status gsim_batch_read(gsim_batch *b, uint32_t id, float *pos_xy, float *vel_xy, float *mass) {
    gsim::clear_error();
    if (!b) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_read: NULL handle");
    auto it = b->systems.find(id);
    if (it == b->systems.end()) return fail(GSIM_ERROR_INVALID_ARGUMENT, "no system %u in the batch", id);
    GSIM_HIP(hipSetDevice(b->device));
    GSIM_HIP(hipStreamSynchronize(b->stream));
    uint32_t stall = 0;
    GSIM_HIP(hipMemcpy(&stall, &b->d.counts->sort_stall, sizeof(stall), hipMemcpyDeviceToHost));
    if (stall)
        return fail(GSIM_ERROR_DEVICE, "radix sort look-back gave up on %u status words: the batch state is invalid",
                    stall);
    const gsim::BatchSystem &s = it->second;
    if (pos_xy) GSIM_HIP(hipMemcpy(pos_xy, b->d.pos[b->cur] + s.off, sizeof(float2) * s.n, hipMemcpyDeviceToHost));
    if (vel_xy) GSIM_HIP(hipMemcpy(vel_xy, b->d.vel[b->cur] + s.off, sizeof(float2) * s.n, hipMemcpyDeviceToHost));
    if (mass) GSIM_HIP(hipMemcpy(mass, b->d.mass + s.off, sizeof(float) * s.n, hipMemcpyDeviceToHost));
    return GSIM_OK;
}

gsim_status gsim_batch_query(const gsim_batch *b, gsim_batch_info *out) {
    gsim::clear_error();
    if (!b || !out) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_query: NULL argument");
    size_t used = 0, largest = 0;
    for (auto &kv : b->free_ranges) largest = std::max<size_t>(largest, kv.second);
    for (auto &kv : b->systems) used += kv.second.na;
    out->systems = (uint32_t)b->systems.size();
    out->max_systems = b->max_systems;
    out->capacity_particles = b->capacity;
    out->used_particles = used;
    out->largest_free_range = largest;
    out->node_capacity = b->node_capacity;
    out->nodes_reserved = b->nodes_reserved;
    out->steps = b->steps;
    out->waves_per_simd = b->tune.waves_per_simd;
    return GSIM_OK;
}

size_t gsim_batch_bytes_per_particle(void) {
    // state 36 B, Barnes-Hut scratch 67 B, 3 node slots x 28 B, partials 32 B
    return 36 + 67 + 3 * 28 + 32;
}

gsim_status gsim_batch_set_waves(gsim_batch *b, uint32_t waves_per_simd) {
    gsim::clear_error();
    if (!b) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_set_waves: NULL handle");
    gsim_batch_tuning t;
    (void)gsim_batch_get_tuning(b, &t);
    t.waves_per_simd = waves_per_simd;
    return gsim_batch_set_tuning(b, &t);
}

gsim_status gsim_batch_get_tuning(const gsim_batch *b, gsim_batch_tuning *out) {
    gsim::clear_error();
    if (!b || !out) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_get_tuning: NULL argument");
    out->waves_per_simd = b->tune.waves_per_simd;
    out->sort_tile = b->tune.sort_rounds * (uint32_t)gsim::kSortThreads;
    out->run_queue = b->tune.run_queue;
    out->prio_step = b->tune.prio_step;
    out->prio_carry = b->tune.prio_carry;
    return GSIM_OK;
}

gsim_status gsim_batch_set_tuning(gsim_batch *b, const gsim_batch_tuning *t) {
    gsim::clear_error();
    if (!b || !t) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_set_tuning: NULL argument");
    if (t->waves_per_simd < 1 || t->waves_per_simd > 8)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "waves per SIMD must be in [1, 8]");
    if (t->sort_tile != 2048 && t->sort_tile != 3072 && t->sort_tile != 4096)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "sort tile must be 2048, 3072 or 4096 keys");
    if (t->run_queue < 1 || t->run_queue > gsim::kRunQueue)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "run queue must be in [1, %u]", gsim::kRunQueue);
    if (t->prio_step > 4096 || t->prio_carry > 4096)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "priority step and carry must be in [0, 4096]");
    gsim::BatchTune n;
    n.waves_per_simd = t->waves_per_simd;
    n.sort_rounds = t->sort_tile / (uint32_t)gsim::kSortThreads;
    n.run_queue = t->run_queue;
    n.prio_step = t->prio_step;
    n.prio_carry = t->prio_carry;
    if (std::memcmp(&n, &b->tune, sizeof(n)) != 0) {
        GSIM_HIP(hipSetDevice(b->device));
        GSIM_HIP(hipStreamSynchronize(b->stream));
        destroy_graphs(b);  // the launch parameters are baked into the graphs
        b->tune = n;
    }
    return GSIM_OK;
}

gsim_status gsim_batch_autotune(gsim_batch *b, uint32_t steps_per_trial, uint32_t *chosen) {
    gsim::clear_error();
    if (!b) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_batch_autotune: NULL handle");
    if (steps_per_trial == 0) steps_per_trial = 5;
    GSIM_HIP(hipSetDevice(b->device));
    if (b->dirty) {
        gsim_status st = upload_membership(b);
        if (st != GSIM_OK) return st;
    }
    // Trials run on the live arena, so snapshot the state and restore it after
    // each: the batch continues from exactly where it was.  The snapshot copies
    // run on the null stream, so wait for the batch's own stream first.
    GSIM_HIP(hipStreamSynchronize(b->stream));
    const size_t cap = b->capacity;
    float2 *snap_pos = nullptr, *snap_vel = nullptr;
    float *snap_mass = nullptr;
    GSIM_HIP(hipMalloc((void **)&snap_pos, sizeof(float2) * cap));
    hipError_t e = hipMalloc((void **)&snap_vel, sizeof(float2) * cap);
    if (e == hipSuccess) e = hipMalloc((void **)&snap_mass, sizeof(float) * cap);
    const int cur0 = b->cur;
    const uint64_t steps0 = b->steps;
    // trial steps also spend device step budgets: restore them with the state
    std::vector<uint32_t> rem_snapshot(b->h_counts.n_sys);
    if (!rem_snapshot.empty())
        GSIM_HIP(hipMemcpy(rem_snapshot.data(), b->d.remaining, 4 * rem_snapshot.size(), hipMemcpyDeviceToHost));
    // ... and add to the acceptance and visit totals: restore those too
    unsigned long long work0[2] = {0, 0};
    GSIM_HIP(hipMemcpy(work0, &b->d.counts->accepts, sizeof(work0), hipMemcpyDeviceToHost));
    auto restore = [&]() -> hipError_t {
        hipError_t r = hipStreamSynchronize(b->stream);
        if (r == hipSuccess) r = hipMemcpy(b->d.pos[cur0], snap_pos, sizeof(float2) * cap, hipMemcpyDeviceToDevice);
        if (r == hipSuccess) r = hipMemcpy(b->d.vel[cur0], snap_vel, sizeof(float2) * cap, hipMemcpyDeviceToDevice);
        if (r == hipSuccess) r = hipMemcpy(b->d.mass, snap_mass, sizeof(float) * cap, hipMemcpyDeviceToDevice);
        if (r == hipSuccess && !rem_snapshot.empty())
            r = hipMemcpy(b->d.remaining, rem_snapshot.data(), 4 * rem_snapshot.size(), hipMemcpyHostToDevice);
        if (r == hipSuccess)
            r = hipMemcpy(&b->d.counts->accepts, work0, sizeof(work0), hipMemcpyHostToDevice);
        b->cur = cur0;
        b->steps = steps0;
        return r;
    };
    if (e == hipSuccess) e = hipMemcpy(snap_pos, b->d.pos[cur0], sizeof(float2) * cap, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(snap_vel, b->d.vel[cur0], sizeof(float2) * cap, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(snap_mass, b->d.mass, sizeof(float) * cap, hipMemcpyDeviceToDevice);
    gsim_status st = GSIM_OK;
    // One trial: the candidate's graphs are captured and warmed by one step,
    // then `steps_per_trial` steps are timed, twice, each from the snapshot;
    // the faster of the two counts.
    auto trial = [&](const gsim_batch_tuning &t, double *ms_out) -> bool {
        *ms_out = 1e300;
        if (e != hipSuccess || st != GSIM_OK) return false;
        st = gsim_batch_set_tuning(b, &t);
        for (int rep = 0; rep < 2 && st == GSIM_OK && e == hipSuccess; ++rep) {
            st = launch_steps(b, 1);
            if (st != GSIM_OK) break;
            e = hipStreamSynchronize(b->stream);
            const auto t0 = std::chrono::steady_clock::now();
            st = launch_steps(b, steps_per_trial);
            if (st != GSIM_OK) break;
            e = hipStreamSynchronize(b->stream);
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            *ms_out = std::min(*ms_out, ms);
            if (e == hipSuccess) e = restore();
        }
        return e == hipSuccess && st == GSIM_OK;
    };
    // Coordinate descent over the knobs, each over its candidates from the
    // current best point; a candidate must beat it by 0.5% to replace it.
    gsim_batch_tuning best;
    (void)gsim_batch_get_tuning(b, &best);
    double best_ms = 1e300;
    bool ok = trial(best, &best_ms);
    struct Knob {
        uint32_t gsim_batch_tuning::*field;
        std::vector<uint32_t> values;
    };
    const Knob knobs[] = {
        {&gsim_batch_tuning::waves_per_simd, {2, 4, 6, 8}},
        {&gsim_batch_tuning::sort_tile, {2048, 3072, 4096}},
        {&gsim_batch_tuning::run_queue, {16, 32, 64}},
        {&gsim_batch_tuning::prio_step, {0, 16, 32, 64}},
        {&gsim_batch_tuning::prio_carry, {0, 8, 16, 32}},
    };
    for (const Knob &k : knobs) {
        for (uint32_t v : k.values) {
            if (!ok) break;
            if (best.*(k.field) == v) continue;
            gsim_batch_tuning cand = best;
            cand.*(k.field) = v;
            double ms = 1e300;
            ok = trial(cand, &ms);
            if (ok && ms < best_ms * 0.995) {
                best = cand;
                best_ms = ms;
            }
        }
    }
    if (e == hipSuccess && st == GSIM_OK) st = gsim_batch_set_tuning(b, &best);
    (void)hipFree(snap_pos);
    (void)hipFree(snap_vel);
    (void)hipFree(snap_mass);
    if (st != GSIM_OK) return st;
    if (e != hipSuccess) return hip_fail(e, "autotune");
    if (chosen) *chosen = best.waves_per_simd;
    return GSIM_OK;
}

gsim_status gsim_batch_destroy(gsim_batch *b) {
    gsim::clear_error();
    if (!b) return GSIM_OK;
    (void)hipSetDevice(b->device);
    free_batch(b);
    return GSIM_OK;
}

}  // extern "C"
