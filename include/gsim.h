/*
 * gsim.h — C ABI of the MI355X force-and-integrate library (libgsim.so).
 *
 * Plain pointers and sizes only.  One gsim_sim owns its device state and one
 * HIP stream; calls on one handle must be serialised by the caller (the
 * reference is single-threaded too).  Every function returns a gsim_status;
 * on failure gsim_last_error() holds a thread-local message, the analogue of
 * the reference's gsim::Exception text (src/Debug/Exception.hpp:10-72).
 *
 * Reference surface replaced (file:line in victordobra/Gravity-Sim):
 *   gsim_required_alignment  DirectSimulation::GetRequiredParticleAlignment
 *                            (DirectSimulation.cpp:26-28) and the Barnes-Hut one
 *                            (BarnesHutSimulation.cpp:834-836)
 *   gsim_create              ParticleSystem ctor upload (ParticleSystem.cpp:438-559,
 *                            150-407) + DirectSimulation / BarnesHutSimulation ctor
 *                            (DirectSimulation.hpp:24, BarnesHutSimulation.hpp:24)
 *   gsim_step                RunSimulations(uint32_t) (DirectSimulation.cpp:271-353,
 *                            BarnesHutSimulation.cpp:849-952): asynchronous,
 *                            stream-ordered
 *   gsim_synchronize         vkDeviceWaitIdle before readback (Main.cpp:286)
 *   gsim_read                ParticleSystem::GetParticles (ParticleSystem.cpp:561-726)
 *   gsim_destroy             ~DirectSimulation / ~BarnesHutSimulation / ~ParticleSystem
 *   gsim_generate            the four generation presets (ParticleSystem.cpp:15-149)
 *   gsim_load_particles      file constructor parsing (ParticleSystem.cpp:451-484)
 *   gsim_save_particles      SaveParticles (ParticleSystem.cpp:727-750)
 */
#ifndef GSIM_H
#define GSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSIM_ABI_VERSION 1

typedef enum gsim_status {
    GSIM_OK = 0,
    GSIM_ERROR_INVALID_ARGUMENT = 1,
    GSIM_ERROR_OUT_OF_MEMORY = 2,
    GSIM_ERROR_DEVICE = 3,       /* a HIP call failed */
    GSIM_ERROR_NO_DEVICE = 4,    /* no gfx950 device visible */
    GSIM_ERROR_IO = 5,           /* particle file could not be read/written */
    GSIM_ERROR_CAPACITY = 6      /* a fixed-size buffer would overflow */
} gsim_status;

/* --simulation-algorithm (Main.cpp:197-202) */
typedef enum gsim_algorithm {
    GSIM_DIRECT_SUM = 0,
    GSIM_BARNES_HUT = 1
} gsim_algorithm;

/* --generate-type (Main.cpp:176-185, ParticleSystem.hpp:12-18) */
typedef enum gsim_generate_type {
    GSIM_GENERATE_RANDOM = 0,
    GSIM_GENERATE_GALAXY = 1,
    GSIM_GENERATE_GALAXY_COLLISION = 2,
    GSIM_GENERATE_SYMMETRICAL_GALAXY_COLLISION = 3,
    /* Synthetic benchmark inputs (not reference presets): */
    GSIM_GENERATE_EXPONENTIAL_DISK = 16,      /* one disk, radius = size, scale = size/5 */
    GSIM_GENERATE_TWO_EXPONENTIAL_DISKS = 17  /* two approaching disks at +-40 */
} gsim_generate_type;

/* Simulation parameters; the reference's ProgramInfo defaults (Main.cpp:59-63)
 * come from gsim_default_params. */
typedef struct gsim_params {
    float gravitational_const; /* --gravitational-const, default 1            */
    float simulation_time;     /* --simulation-time (dt), default 1e-3        */
    float simulation_speed;    /* --simulation-speed, default 1; Barnes-Hut
                                  only, as in the reference (SimShader.comp:68) */
    float softening_len;       /* --softening-len, default 0.2                */
    float accuracy_parameter;  /* --accuracy-parameter (theta), default 1     */
    int32_t algorithm;         /* gsim_algorithm                              */
    uint32_t group_width;      /* Barnes-Hut vote width W: 64 (one wave64) or
                                  32 (the subgroup of the published runs);
                                  0 selects 64                                */
    int32_t device;            /* HIP device ordinal, default 0               */
} gsim_params;

typedef struct gsim_sim gsim_sim;

/* Debug dump selectors for gsim_debug_dump (state of the most recent step). */
typedef enum gsim_dump {
    GSIM_DUMP_CELLS = 1,       /* u32[Na]  20-bit cell per particle (30-bit with 15-bit
                                  leaves), 0xffffffff = killed                          */
    GSIM_DUMP_MORTON30 = 2,    /* u32[Na]  30-bit Morton code, 0xffffffff = killed     */
    GSIM_DUMP_SORTED_SRC = 3,  /* u32[Na]  particle index per sorted slot (Na = empty)  */
    GSIM_DUMP_COUNTS = 4,      /* u32[nodes] subtree sizes                              */
    GSIM_DUMP_RADII = 5,       /* f32[nodes]                                            */
    GSIM_DUMP_NODE_POS = 6,    /* f32[2*nodes]                                          */
    GSIM_DUMP_NODE_MASS = 7,   /* f32[nodes]                                            */
    GSIM_DUMP_NODE_TOTAL = 8,  /* u32[1]   node count of the last tree (counts[0])      */
    GSIM_DUMP_ACCEL = 9,       /* f32[2*Na] accelerations of the last step, by particle
                                  index (needs gsim_set_debug(sim, GSIM_DEBUG_ACCEL))  */
    GSIM_DUMP_COUNTERS = 10,   /* u64[2]   real-lane node visits and acceptances of
                                  the last Barnes-Hut step                              */
    GSIM_DUMP_BOX = 11         /* f32[4]   adaptive box {cx, cy, h, r0} of the last
                                  step (GSIM_TUNE_BH_ADAPTIVE_BOX only)                 */
} gsim_dump;

#define GSIM_DEBUG_ACCEL 1u

/* Timing phases returned by gsim_get_timing (milliseconds, summed). */
typedef struct gsim_timing {
    double force_ms;      /* direct-sum force kernel / Barnes-Hut traversal   */
    double build_ms;      /* Barnes-Hut: keys, sort and tree build            */
    double integrate_ms;  /* direct-sum partial-sum combine and integrate     */
    double step_ms;       /* whole steps                                       */
    uint64_t steps;       /* steps covered                                     */
    uint64_t force_launches;
} gsim_timing;

void gsim_default_params(gsim_params *params);
const char *gsim_last_error(void);
const char *gsim_version(void);

/* 64 for both algorithms: N is padded to a multiple with {0,0,0,0,0}
 * particles (ParticleSystem.cpp:487-491). */
size_t gsim_required_alignment(int32_t algorithm);

/* Copies n particles in (host arrays: pos_xy[2n], vel_xy[2n], mass[n]) and pads
 * them to the alignment.  The library owns the device copies. */
gsim_status gsim_create(const gsim_params *params, const float *pos_xy,
                        const float *vel_xy, const float *mass, size_t n,
                        gsim_sim **out);

/* Enqueues count steps on the handle's stream and returns (no host wait). */
gsim_status gsim_step(gsim_sim *sim, uint32_t count);

/* Waits for every enqueued step. */
gsim_status gsim_synchronize(gsim_sim *sim);

/* Synchronises, then copies the n unpadded particles out in original index
 * order.  Any pointer may be NULL to skip that field. */
gsim_status gsim_read(gsim_sim *sim, float *pos_xy, float *vel_xy, float *mass);

gsim_status gsim_destroy(gsim_sim *sim);

size_t gsim_particle_count(const gsim_sim *sim);
size_t gsim_aligned_particle_count(const gsim_sim *sim);
uint64_t gsim_steps_done(const gsim_sim *sim);

/* Debug intermediates for parity tests; synchronises.  *written receives the
 * number of bytes copied (or required, when bytes is too small). */
gsim_status gsim_debug_dump(gsim_sim *sim, int32_t what, void *buf,
                            size_t bytes, size_t *written);
gsim_status gsim_set_debug(gsim_sim *sim, uint32_t flags);

/* HIP-event timing of each kernel phase on the handle's stream. */
gsim_status gsim_set_timing(gsim_sim *sim, int32_t enable);
gsim_status gsim_get_timing(gsim_sim *sim, gsim_timing *out, int32_t reset);

/* Selects the one-sided direct-sum kernel with this launch shape: targets per
 * lane (1, 2 or 4) and source slices (0 = pick by size).  Results stay within the documented tolerance for every
 * shape; the partial sums are combined in slice order, so a shape is
 * deterministic. */
gsim_status gsim_set_ds_shape(gsim_sim *sim, uint32_t targets_per_lane,
                              uint32_t slices);

/* Launch tuning that never changes results beyond the documented tolerance
 * (visit/accept counts and trees are unaffected by every knob), and two
 * opt-in options beyond the reference, GSIM_TUNE_BH_ADAPTIVE_BOX and
 * GSIM_TUNE_BH_LEAF_BITS.  The traversal launch knobs (WAVES_PER_SIMD,
 * TRAVERSAL, SEGMENTS, SPLIT) shape the W = 64 launches only: on a Barnes-Hut
 * handle of group width 32 they return GSIM_ERROR_INVALID_ARGUMENT. */
typedef enum gsim_tuning {
    GSIM_TUNE_DS_RCP = 1,             /* 0/1: 1/r^2 from v_rcp instead of a multiply  */
    GSIM_TUNE_BH_WAVES_PER_SIMD = 2,  /* 1..8: persistent traversal width             */
    GSIM_TUNE_BH_TRAVERSAL = 3,       /* bit 0 persistent groups, bit 1 LDS-staged runs
                                         (2, one wave per group with staged runs, is the
                                         default); 6 = cost-balanced contiguous
                                         segments, with dense-core groups split          */
    GSIM_TUNE_DS_SYMMETRIC = 4,       /* 0/1: each pair once (Newton's third law); the
                                         default for 16,384 <= Na <= 262,144           */
    GSIM_TUNE_DS_SYM_COLS = 5,        /* symmetric kernel: 1 or 2 column blocks per
                                         wave, 0 = by size                              */
    GSIM_TUNE_BH_SEGMENTS = 6,        /* balanced traversal: workgroups of 16 waves
                                         (default 512, two per CU)                      */
    GSIM_TUNE_BH_SPLIT = 7,           /* groups whose last walk reaches this multiple
                                         of the mean are walked by 4 waves: in the
                                         default traversal in index order (chunk
                                         count, the heaviest 512 at most, beside the
                                         main launch) and the balanced one (cost);
                                         0 = never; default 2                          */
    GSIM_TUNE_BH_ADAPTIVE_BOX = 8,    /* 0/1, beyond the reference and off by default:
                                         each step's tree spans the bounding square of
                                         the particles (centre c, half-size h = extent
                                         x (0.5 + 2^-9)) instead of [-500, 500), so
                                         only non-finite positions are killed; cell
                                         radius (h*h)*2 / 4^d.  Changes the topology,
                                         so results differ from the reference's.        */
    GSIM_TUNE_BH_LEAF_BITS = 9,       /* 10 (the reference's 1024^2 leaf grid, the
                                         default) or 15, beyond the reference: leaves
                                         on a 32768^2 grid at depth 15, so a dense knot
                                         is no longer one bucket that every group
                                         streams whole.  The sums follow the same
                                         quadrant rule; cells and leaves below depth 10
                                         change the topology, so results differ from
                                         the reference's.  Combines with the adaptive
                                         box.  Reallocates the node arrays (synchronous;
                                         they grow by up to 2.5 x Na entries).          */
    GSIM_TUNE_BH_XCD_ORDER = 10,      /* 0..4096: the static traversals hand each of
                                         the 8 XCDs (each with its own L2) runs of this
                                         many consecutive workgroups instead of every
                                         eighth one, so spatially adjacent groups share
                                         an L2 (one contiguous range per XCD once the
                                         runs reach an eighth of the launch); 0 or 1 =
                                         round-robin, the default: measured at 500k,
                                         runs of 2-32 change nothing and contiguous
                                         ranges are 16% slower.  W = 32 and 64.         */
    GSIM_TUNE_BH_ORDER = 11           /* static walk's launch order (W = 32, and W =
                                         64 neither persistent nor balanced), by
                                         workgroup of four groups: 0 index order, 1
                                         heaviest first (by the last step's visits), 2
                                         auto (the default): every 1,024 steps, four
                                         steps alternate the two and the faster is kept
                                         (decided once the host sees them complete;
                                         never waited for)                              */
} gsim_tuning;
gsim_status gsim_set_tuning(gsim_sim *sim, int32_t knob, uint32_t value);

/* ---- allocation-free direct-sum step ---------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP stream capture, serving loops that forbid hipMalloc):
 * one direct-sum step, the one-sided kernels of gsim_step with the slice count
 * gsim_set_ds_shape(sim, 2, 0) picks (capped so slices x n fits the scratch),
 * on caller-owned DEVICE arrays of n particles (n a nonzero multiple of 64;
 * pad with {0,0,0,0,0} particles as gsim_create does).  Inputs are read and
 * the outputs written on `stream` (a hipStream_t; NULL = the null stream) of
 * params->device, asynchronously; outputs must not alias inputs.  G,
 * simulation_time and softening_len mean what they mean for gsim_create;
 * simulation_speed, accuracy_parameter and group_width are ignored.
 * The partial sums live in a static device array of
 * GSIM_NOALLOC_SCRATCH_BYTES that each GPU the call runs on holds until the
 * process ends, so n is at most GSIM_NOALLOC_SCRATCH_BYTES / 8 (4,194,304);
 * larger n returns GSIM_ERROR_CAPACITY.  The call allocates nothing.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
#define GSIM_NOALLOC_SCRATCH_BYTES (32u << 20)
gsim_status gsim_ds_step_noalloc(const gsim_params *params, const float *pos_xy,
                                 const float *vel_xy, const float *mass,
                                 float *pos_out_xy, float *vel_out_xy, size_t n,
                                 void *stream);

/* ---- batched form: many independent systems stepped together ---------------
 * A batch is one device arena holding many systems of uneven size (direct sum
 * or Barnes-Hut, each with its own parameters).  Every stage of a step runs
 * as one launch over all systems (Barnes-Hut systems share one radix sort and
 * one node array; one persistent launch computes every system's forces), and
 * the whole step is replayed from a captured hipGraph.  Membership can change
 * between steps without re-capturing.  A system's results are bit-identical
 * to stepping it alone with gsim_step (the same group width, 32 or 64, and
 * the default launch shape).
 * Direct-sum systems of 16,384 or more padded particles are refused: they
 * fill the GPU alone and are stepped alone with the symmetric kernel.
 * No counterpart in the reference, which runs one system per process. */
typedef struct gsim_batch gsim_batch;

typedef struct gsim_batch_info {
    uint32_t systems, max_systems;
    uint64_t capacity_particles;   /* padded particles the arena holds       */
    uint64_t used_particles;
    uint64_t largest_free_range;   /* largest system (padded) that fits now   */
    uint64_t node_capacity, nodes_reserved;
    uint64_t steps;
    uint32_t waves_per_simd;
} gsim_batch_info;

gsim_status gsim_batch_create(int32_t device, size_t capacity_particles,
                              uint32_t max_systems, gsim_batch **out);
/* Uploads one system (host arrays as for gsim_create) into a free arena
 * range; GSIM_ERROR_CAPACITY when it does not fit now. */
gsim_status gsim_batch_add(gsim_batch *batch, const gsim_params *params,
                           const float *pos_xy, const float *vel_xy,
                           const float *mass, size_t n, uint32_t *system_id);
gsim_status gsim_batch_remove(gsim_batch *batch, uint32_t system_id);
/* Enqueues count steps of every system in the batch (asynchronous). */
gsim_status gsim_batch_step(gsim_batch *batch, uint32_t count);
/* Limits a system to `more_steps` further steps (UINT64_MAX: unlimited); once
 * spent, the system holds its state while the others keep stepping. */
gsim_status gsim_batch_set_steps(gsim_batch *batch, uint32_t system_id, uint64_t more_steps);
gsim_status gsim_batch_steps_done(gsim_batch *batch, uint32_t system_id, uint64_t *steps);
/* Barnes-Hut real-lane acceptances (interactions) over all steps so far. */
gsim_status gsim_batch_accepts(gsim_batch *batch, uint64_t *accepts);
/* Barnes-Hut real-lane node visits (vote evaluations) over all steps so far. */
gsim_status gsim_batch_visits(gsim_batch *batch, uint64_t *visits);
gsim_status gsim_batch_synchronize(gsim_batch *batch);
gsim_status gsim_batch_read(gsim_batch *batch, uint32_t system_id,
                            float *pos_xy, float *vel_xy, float *mass);
gsim_status gsim_batch_query(const gsim_batch *batch, gsim_batch_info *out);
/* Device bytes the arena needs per padded particle (for HBM budgeting). */
size_t gsim_batch_bytes_per_particle(void);
/* Launch parameters of a batched step.  None changes any result: systems
 * stay bit-identical to their alone runs under every setting. */
typedef struct gsim_batch_tuning {
    uint32_t waves_per_simd;  /* persistent force-launch width, 1..8 (default 6)      */
    uint32_t sort_tile;       /* radix sort tile: 2048, 3072 (default) or 4096 keys  */
    uint32_t run_queue;       /* leaf runs a walking wave queues in its LDS stack
                                 before summing them, 1..64 (default 64)            */
    uint32_t prio_step;       /* summed 64-node chunks per issue-priority level of a
                                 walking wave, 0 = off (default 32)                  */
    uint32_t prio_carry;      /* a work item starts at level (chunks it summed last
                                 step) / prio_carry, 0 = off (default 16)            */
} gsim_batch_tuning;
gsim_status gsim_batch_get_tuning(const gsim_batch *batch, gsim_batch_tuning *out);
gsim_status gsim_batch_set_tuning(gsim_batch *batch, const gsim_batch_tuning *tuning);
/* Persistent force-launch width only (waves per SIMD, 1..8). */
gsim_status gsim_batch_set_waves(gsim_batch *batch, uint32_t waves_per_simd);
/* Autotuner: coordinate descent over the gsim_batch_tuning knobs on the live
 * batch (waves per SIMD, sort tile, run queue, priority step, priority
 * carry, in that order), each candidate timed twice over steps_per_trial
 * steps (0: 5) from a snapshot of the state that is restored after every
 * trial, the faster run counting; keeps the fastest setting (read it back
 * with gsim_batch_get_tuning).  *chosen_waves_per_simd may be NULL. */
gsim_status gsim_batch_autotune(gsim_batch *batch, uint32_t steps_per_trial,
                                uint32_t *chosen_waves_per_simd);
gsim_status gsim_batch_destroy(gsim_batch *batch);

/* ---- serving: continuous batching of independent requests ------------------
 * A server owns a batch sized by an HBM budget and a worker thread.  Clients
 * submit systems (any thread); the worker admits queued requests first-come
 * first-served while the arena (and the optional active-particle cap) has
 * room, steps every admitted system together, and retires each system after
 * exactly its requested steps, freeing its space for the next request.  Each
 * result is bit-identical to running the system alone.  No counterpart in the
 * reference, which runs one system per process. */
typedef struct gsim_server gsim_server;

typedef struct gsim_request_info {
    uint64_t n, steps;
    double queue_ms;    /* submit -> admitted into the batch */
    double latency_ms;  /* submit -> result on the host      */
} gsim_request_info;

typedef struct gsim_server_info {
    uint64_t completed, failed, queued, running;
    uint64_t batch_steps;          /* batched steps the worker has enqueued   */
    double p50_ms, p99_ms, max_ms; /* submit -> result latency percentiles    */
    double span_s;                 /* first submit -> last completion         */
    uint64_t ds_interactions;      /* direct sum: padded N^2 per step, completed requests */
    uint64_t bh_interactions;      /* Barnes-Hut accepted interactions, all steps so far  */
    uint64_t capacity_particles;   /* arena capacity the budget bought, all arenas */
    uint64_t bh_visits;            /* Barnes-Hut real-lane node visits, all steps so far  */
    uint32_t arenas;               /* arenas the budget was split into          */
} gsim_server_info;

/* hbm_budget_bytes buys gsim_batch_bytes_per_particle per padded particle,
 * split into arenas (batches) of at most 268,435,392 particles each (32-bit
 * node indices inside an arena): a 200 GB budget is 4 arenas, 913M
 * particles.  Each arena has its own stream, so their steps overlap; a
 * request goes to the first arena with room, and must fit in one.
 * max_systems applies per arena (at most 2048); max_active_particles caps
 * the particles stepping at once (0: the arenas' capacity is the only
 * limit). */
gsim_status gsim_server_create(int32_t device, size_t hbm_budget_bytes, uint32_t max_systems,
                               size_t max_active_particles, gsim_server **out);
/* The same with each arena capped at arena_particles (0: the 268,435,392
 * maximum); at most 64 arenas. */
gsim_status gsim_server_create_arenas(int32_t device, size_t hbm_budget_bytes,
                                      uint32_t max_systems, size_t max_active_particles,
                                      size_t arena_particles, gsim_server **out);
/* Queues one system for `steps` steps; inputs are copied.  GSIM_ERROR_CAPACITY
 * when the system alone exceeds the budget. */
gsim_status gsim_server_submit(gsim_server *server, const gsim_params *params,
                               const float *pos_xy, const float *vel_xy, const float *mass,
                               size_t n, uint32_t steps, uint64_t *request_id);
/* Waits up to timeout_s seconds; *done = 1 when the request has completed. */
gsim_status gsim_server_wait(gsim_server *server, uint64_t request_id, double timeout_s,
                             int32_t *done);
/* Copies a completed request's final state (any pointer may be NULL). */
gsim_status gsim_server_result(gsim_server *server, uint64_t request_id, float *pos_xy,
                               float *vel_xy, float *mass, gsim_request_info *info);
/* Forgets a completed request and its host copy. */
gsim_status gsim_server_release(gsim_server *server, uint64_t request_id);
/* Waits until no request is queued or running; *idle = 1 if so. */
gsim_status gsim_server_drain(gsim_server *server, double timeout_s, int32_t *idle);
gsim_status gsim_server_query(gsim_server *server, gsim_server_info *out);
gsim_status gsim_server_destroy(gsim_server *server);

/* ---- particle sets: presets and the text format --------------------------- */

/* Fills n particles (pos_xy[2n], vel_xy[2n], mass[n]) from a preset with an
 * explicit seed.  For GSIM_GENERATE_SYMMETRICAL_GALAXY_COLLISION n must be even
 * (the reference rounds it down, ParticleSystem.cpp:522-525). */
gsim_status gsim_generate(int32_t type, size_t n, float generate_size,
                          float min_mass, float max_mass,
                          float gravitational_const, uint64_t seed,
                          float *pos_xy, float *vel_xy, float *mass);

/* Counts the particles of a text file ("%f%f%f%f%f" records until the first
 * failed match), then loads up to capacity of them. */
gsim_status gsim_count_particles(const char *path, size_t *n);
gsim_status gsim_load_particles(const char *path, float *pos_xy, float *vel_xy,
                                float *mass, size_t capacity, size_t *n);

/* Writes "%.7f %.7f %.7f %.7f %.7f\n" per particle. */
gsim_status gsim_save_particles(const char *path, const float *pos_xy,
                                const float *vel_xy, const float *mass,
                                size_t n);

#ifdef __cplusplus
}
#endif
#endif
