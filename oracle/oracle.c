/*
 * CPU oracle for the force-and-integrate step — TEST INFRASTRUCTURE ONLY
 * (the checker for tests/, the smoke check and bench.py's CPU baseline; the
 * product never loads it).  Compiled with -ffp-contract=off so every FP32
 * operation rounds exactly as written; no -ffast-math.
 *
 * Parity status of each function (how it is pinned is in tests/):
 *   oracle_cell20          pinned: exhaustive 2^20 bit-loop Morton check,
 *                          boundary known-answer cases, 30->20 bit prefix.
 *   oracle_ds_accel_f32    pinned: two-body and ring closed forms, Newton III,
 *                          FP64 sum, bit-exact vs BH traversal at theta = 0.
 *   oracle_ds_accel_f64    pinned: closed forms (two-body, ring centre),
 *                          Newton III on 512 random bodies.
 *   oracle_ds_accel_f64_at pinned: equal, row for row, to oracle_ds_accel_f64
 *                          on the listed targets (any order, repeats).
 *   oracle_ds_step         pinned: free particle x + v*dt; one two-body
 *                          step (G != 1) within FP32 rounding of the float64
 *                          physics; five steps of 128 bodies within FP32
 *                          rounding of oracle_ds_step_f64.
 *   oracle_ds_step_f64     pinned: free particle exact over 1 and 7 steps;
 *                          two bodies from rest and two moving bodies over
 *                          two steps against Newton's law in numpy (unequal
 *                          masses, G != 1, softening comparable to the
 *                          separation); momentum and uniform centre-of-mass
 *                          motion over 1,000 steps of 64 bodies; a softened
 *                          circular orbit closing after one period with
 *                          first-order convergence in dt.
 *   oracle_bh_build        pinned: recursive pointer-quadtree brute force on
 *                          small N, y = n + #cells invariants, dyadic COM.
 *   oracle_bh_build_sparse pinned: bit-identical to oracle_bh_build /
 *                          oracle_bh_build_box at 10 bits; recursive
 *                          pointer-quadtree brute force at 15 bits; the
 *                          15-bit code >> 10 is the 20-bit cell.
 *   oracle_bh_traverse     pinned: recursive group-vote brute force for
 *                          W in {1,32,64}, theta = 0 == sorted direct sum,
 *                          theta -> inf == pull of the global COM.
 *   oracle_bh_integrate    pinned: kill semantics (killed particles keep
 *                          their velocity); a theta = 0 step with G != 1
 *                          within FP32 rounding of oracle_ds_step_f64.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- tiny parallel-for ------------------------------------------------------ */

typedef void (*range_fn)(size_t lo, size_t hi, void *ctx);
typedef struct {
    range_fn fn;
    void *ctx;
    size_t lo, hi;
} range_job;

static void *range_thread(void *p) {
    range_job *j = (range_job *)p;
    j->fn(j->lo, j->hi, j->ctx);
    return NULL;
}

/* Splits [0, n) into nthreads contiguous chunks.  Results never depend on the
 * thread count: every output element is computed by exactly one thread with
 * the same operation order. */
static void parallel_for(size_t n, int nthreads, range_fn fn, void *ctx) {
    if (nthreads <= 1 || n < 2) {
        fn(0, n, ctx);
        return;
    }
    if ((size_t)nthreads > n) nthreads = (int)n;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    range_job *jobs = (range_job *)malloc(sizeof(range_job) * (size_t)nthreads);
    size_t chunk = (n + (size_t)nthreads - 1) / (size_t)nthreads;
    int started = 0;
    for (int t = 0; t < nthreads; ++t) {
        size_t lo = (size_t)t * chunk, hi = lo + chunk;
        if (lo >= n) break;
        if (hi > n) hi = n;
        jobs[t].fn = fn; jobs[t].ctx = ctx; jobs[t].lo = lo; jobs[t].hi = hi;
        if (pthread_create(&th[t], NULL, range_thread, &jobs[t]) != 0) {
            fn(lo, hi, ctx); /* fall back to the calling thread */
            jobs[t].fn = NULL;
        }
        started = t + 1;
    }
    for (int t = 0; t < started; ++t)
        if (jobs[t].fn) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

/* ---- direct sum ------------------------------------------------------------ */

typedef struct {
    size_t na;
    const float *pos, *mass;
    float eps2;
    float *acc;
} ds32_ctx;

/* SimShader.comp:39-65: j ascending over [0, na); the 64-wide LDS tiling of
 * the reference does not change the order. */
static void ds32_range(size_t lo, size_t hi, void *p) {
    ds32_ctx *c = (ds32_ctx *)p;
    for (size_t i = lo; i < hi; ++i) {
        const float px = c->pos[2 * i], py = c->pos[2 * i + 1];
        float ax = 0.0f, ay = 0.0f;
        for (size_t j = 0; j < c->na; ++j) {
            const float dx = c->pos[2 * j] - px;             /* :55 */
            const float dy = c->pos[2 * j + 1] - py;
            const float r2 = (dx * dx + dy * dy) + c->eps2;  /* :56 dot + eps2 */
            const float r = 1.0f / sqrtf(r2);                /* inversesqrt */
            const float s = ((c->mass[j] * r) * r) * r;      /* :59 */
            ax = ax + dx * s;
            ay = ay + dy * s;
        }
        c->acc[2 * i] = ax;
        c->acc[2 * i + 1] = ay;
    }
}

void oracle_ds_accel_f32(size_t na, const float *pos, const float *mass,
                         float eps2, float *acc, int nthreads) {
    ds32_ctx c = {na, pos, mass, eps2, acc};
    parallel_for(na, nthreads, ds32_range, &c);
}

typedef struct {
    size_t na;
    const float *pos, *mass;
    double eps2;
    double *acc;
} ds64_ctx;

static void ds64_range(size_t lo, size_t hi, void *p) {
    ds64_ctx *c = (ds64_ctx *)p;
    for (size_t i = lo; i < hi; ++i) {
        const double px = c->pos[2 * i], py = c->pos[2 * i + 1];
        double ax = 0.0, ay = 0.0;
        for (size_t j = 0; j < c->na; ++j) {
            const double dx = (double)c->pos[2 * j] - px;
            const double dy = (double)c->pos[2 * j + 1] - py;
            const double r2 = dx * dx + dy * dy + c->eps2;
            const double r = 1.0 / sqrt(r2);
            const double s = (double)c->mass[j] * r * r * r;
            ax += dx * s;
            ay += dy * s;
        }
        c->acc[2 * i] = ax;
        c->acc[2 * i + 1] = ay;
    }
}

void oracle_ds_accel_f64(size_t na, const float *pos, const float *mass,
                         double eps2, double *acc, int nthreads) {
    ds64_ctx c = {na, pos, mass, eps2, acc};
    parallel_for(na, nthreads, ds64_range, &c);
}

typedef struct {
    ds64_ctx base;
    const uint32_t *targets;
} ds64t_ctx;

static void ds64t_range(size_t lo, size_t hi, void *p) {
    ds64t_ctx *c = (ds64t_ctx *)p;
    for (size_t k = lo; k < hi; ++k) {
        const size_t i = c->targets[k];
        const double px = c->base.pos[2 * i], py = c->base.pos[2 * i + 1];
        double ax = 0.0, ay = 0.0;
        for (size_t j = 0; j < c->base.na; ++j) {
            const double dx = (double)c->base.pos[2 * j] - px;
            const double dy = (double)c->base.pos[2 * j + 1] - py;
            const double r2 = dx * dx + dy * dy + c->base.eps2;
            const double r = 1.0 / sqrt(r2);
            const double s = (double)c->base.mass[j] * r * r * r;
            ax += dx * s;
            ay += dy * s;
        }
        c->base.acc[2 * k] = ax;
        c->base.acc[2 * k + 1] = ay;
    }
}

void oracle_ds_accel_f64_at(size_t na, const float *pos, const float *mass, double eps2,
                            const uint32_t *targets, size_t ntargets, double *acc,
                            int nthreads) {
    ds64t_ctx c = {{na, pos, mass, eps2, acc}, targets};
    parallel_for(ntargets, nthreads, ds64t_range, &c);
}

void oracle_ds_step(size_t na, const float *pos_in, const float *vel_in,
                    const float *mass_in, float *pos_out, float *vel_out,
                    float G, float dt, float eps2, int nthreads) {
    float *acc = (float *)malloc(sizeof(float) * 2 * (na ? na : 1));
    oracle_ds_accel_f32(na, pos_in, mass_in, eps2, acc, nthreads);
    const float kv = dt * G; /* SIMULATION_TIME * GRAVITATIONAL_CONST, :68 */
    for (size_t i = 0; i < na; ++i) {
        for (int k = 0; k < 2; ++k) {
            const float v = vel_in[2 * i + k];
            const float nv = v + acc[2 * i + k] * kv;               /* :68 */
            const float np = pos_in[2 * i + k] + ((v + nv) * 0.5f) * dt; /* :69 */
            vel_out[2 * i + k] = nv;
            pos_out[2 * i + k] = np;
        }
    }
    free(acc);
}

typedef struct {
    size_t na;
    const double *pos, *mass;
    double eps2;
    double *acc;
} ds64s_ctx;

static void ds64s_range(size_t lo, size_t hi, void *p) {
    ds64s_ctx *c = (ds64s_ctx *)p;
    for (size_t i = lo; i < hi; ++i) {
        double ax = 0.0, ay = 0.0;
        for (size_t j = 0; j < c->na; ++j) {
            const double dx = c->pos[2 * j] - c->pos[2 * i];
            const double dy = c->pos[2 * j + 1] - c->pos[2 * i + 1];
            const double r2 = dx * dx + dy * dy + c->eps2;
            const double r = 1.0 / sqrt(r2);
            const double s = c->mass[j] * r * r * r;
            ax += dx * s;
            ay += dy * s;
        }
        c->acc[2 * i] = ax;
        c->acc[2 * i + 1] = ay;
    }
}

void oracle_ds_step_f64(size_t na, double *pos, double *vel, const double *mass,
                        double G, double dt, double eps2, int nthreads) {
    double *acc = (double *)malloc(sizeof(double) * 2 * (na ? na : 1));
    ds64s_ctx c = {na, pos, mass, eps2, acc};
    parallel_for(na, nthreads, ds64s_range, &c);
    for (size_t i = 0; i < 2 * na; ++i) {
        const double v = vel[i];
        const double nv = v + acc[i] * (dt * G);
        pos[i] = pos[i] + (v + nv) * 0.5 * dt;
        vel[i] = nv;
    }
    free(acc);
}

/* ---- Barnes-Hut: cells ------------------------------------------------------ */

/* InitShader.comp:84-98: spread the 10 low bits of v to the even bits. */
static uint32_t part1by1(uint32_t v) {
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

uint32_t oracle_cell20(float x, float y) {
    /* InitShader.comp:80-83: (pos / S + 1) * 0.5, then *= TREE_SIZE */
    float qx = ((x / ORACLE_SIM_SIZE) + 1.0f) * 0.5f;
    float qy = ((y / ORACLE_SIM_SIZE) + 1.0f) * 0.5f;
    qx = qx * (float)ORACLE_TREE_SIZE;
    qy = qy * (float)ORACLE_TREE_SIZE;
    if (qx >= 0.0f && qx < (float)ORACLE_TREE_SIZE && qy >= 0.0f &&
        qy < (float)ORACLE_TREE_SIZE) {
        const uint32_t ix = (uint32_t)qx, iy = (uint32_t)qy;
        return part1by1(ix) | (part1by1(iy) << 1);
    }
    return 0xffffffffu;
}

/* ---- adaptive bounding box (beyond the reference; opt-in) ------------------- */

/* The square that holds every particle with finite coordinates: centre at the
 * midpoint of the extremes, half-size h = max extent x (0.5 + 2^-9) (so the
 * farthest particle maps below cell 1024), at least 1e-3; r0 = (h*h)*2 plays
 * the role of S*S*2 in the cell radius r0 / 4^d.  FP32, operations as
 * written.  No finite particle: the box is (0, 0, 1e-3, 2e-6). */
void oracle_bh_box(size_t na, const float *pos, float box[4]) {
    float xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (size_t i = 0; i < na; ++i) {
        const float x = pos[2 * i], y = pos[2 * i + 1];
        if (!isfinite(x) || !isfinite(y)) continue;
        xmin = fminf(xmin, x); xmax = fmaxf(xmax, x);
        ymin = fminf(ymin, y); ymax = fmaxf(ymax, y);
    }
    if (!(xmin <= xmax)) { xmin = xmax = ymin = ymax = 0.0f; }
    const float cx = (xmin + xmax) * 0.5f, cy = (ymin + ymax) * 0.5f;
    const float w = fmaxf(xmax - xmin, ymax - ymin);
    float h = w * 0.501953125f;
    if (!(h >= 1e-3f)) h = 1e-3f;
    box[0] = cx; box[1] = cy; box[2] = h; box[3] = (h * h) * 2.0f;
}

/* Cell of a particle in the box: ((p - c) / h + 1) * 0.5 * 1024, clamped to
 * [0, 1023]; only non-finite positions are outside. */
uint32_t oracle_cell20_box(float x, float y, const float box[4]) {
    if (!isfinite(x) || !isfinite(y)) return 0xffffffffu;
    float qx = (((x - box[0]) / box[2]) + 1.0f) * 0.5f;
    float qy = (((y - box[1]) / box[2]) + 1.0f) * 0.5f;
    qx = qx * (float)ORACLE_TREE_SIZE;
    qy = qy * (float)ORACLE_TREE_SIZE;
    uint32_t ix = qx > 0.0f ? (uint32_t)qx : 0u, iy = qy > 0.0f ? (uint32_t)qy : 0u;
    if (ix > ORACLE_TREE_SIZE - 1) ix = ORACLE_TREE_SIZE - 1;
    if (iy > ORACLE_TREE_SIZE - 1) iy = ORACLE_TREE_SIZE - 1;
    return part1by1(ix) | (part1by1(iy) << 1);
}

static uint32_t cell_of(float x, float y, const float *box) {
    return box ? oracle_cell20_box(x, y, box) : oracle_cell20(x, y);
}

size_t oracle_bh_capacity(size_t na) { return na + ORACLE_PYRAMID_CELLS; }

/* Offsets of level d inside one dense pyramid array: (4^d - 1)/3. */
static size_t level_offset(unsigned d) { return ((size_t)1 << (2 * d)) / 3; }

static long build_impl(size_t na, float *pos, float *mass, float *pos_out,
                       float *mass_out, const float *box, oracle_tree *tree,
                       size_t *n_inside) {
    const size_t cells = ORACLE_PYRAMID_CELLS;
    /* dense per-level arrays (BarnesHutSimulation.cpp:143-234) */
    uint32_t *cn = (uint32_t *)calloc(cells, sizeof(uint32_t));
    uint32_t *cy = (uint32_t *)calloc(cells, sizeof(uint32_t));
    uint32_t *stx = (uint32_t *)calloc(cells, sizeof(uint32_t));
    uint32_t *sty = (uint32_t *)calloc(cells, sizeof(uint32_t));
    float *sx = (float *)calloc(cells, sizeof(float));
    float *sy = (float *)calloc(cells, sizeof(float));
    float *sm = (float *)calloc(cells, sizeof(float));
    if (!cn || !cy || !stx || !sty || !sx || !sy || !sm) {
        free(cn); free(cy); free(stx); free(sty); free(sx); free(sy); free(sm);
        return -1;
    }
    const size_t L10 = level_offset(10);
    size_t inside = 0;
    /* each particle's cell from Init, reused by ParticleSort (a killed
     * particle's rewritten position must not be mapped again) */
    uint32_t *cell = (uint32_t *)malloc(sizeof(uint32_t) * (na ? na : 1));
    if (!cell) {
        free(cn); free(cy); free(stx); free(sty); free(sx); free(sy); free(sm);
        return -1;
    }

    /* ClearShader.comp:81-82: every sorted slot starts as the sentinel. */
    for (size_t i = 0; i < na; ++i) tree->sorted_src[i] = (uint32_t)na;

    /* InitShader.comp:73-119, particles in ascending index (canonical order of
     * the float atomics). */
    for (size_t i = 0; i < na; ++i) {
        const float x = pos[2 * i], y = pos[2 * i + 1], m = mass[i];
        const uint32_t c = cell_of(x, y, box);
        cell[i] = c;
        if (c != 0xffffffffu) {
            const size_t k = L10 + c;
            cn[k] += 1u;
            if (cy[k] == 1u) cy[k] = 2u; /* race-free form of :105-106 */
            cy[k] += 1u;
            sx[k] = sx[k] + x * m; /* pos *= mass, :102 */
            sy[k] = sy[k] + y * m;
            sm[k] = sm[k] + m;
            ++inside;
        } else {
            const float kill = -ORACLE_SIM_SIZE * 2.0f; /* :113 */
            pos[2 * i] = kill; pos[2 * i + 1] = kill; mass[i] = 0.0f;
            if (pos_out) { pos_out[2 * i] = kill; pos_out[2 * i + 1] = kill; }
            if (mass_out) mass_out[i] = 0.0f;
        }
    }
    if (n_inside) *n_inside = inside;

    /* TreeInitShader.comp:58-90, d = 9..0 */
    for (int d = 9; d >= 0; --d) {
        const size_t Ld = level_offset((unsigned)d), Lc = level_offset((unsigned)d + 1);
        const size_t nc = (size_t)1 << (2 * d);
        for (size_t c = 0; c < nc; ++c) {
            const size_t ch = Lc + 4 * c;
            uint32_t n = cn[ch] + cn[ch + 1] + cn[ch + 2] + cn[ch + 3];
            uint32_t y = cy[ch] + cy[ch + 1] + cy[ch + 2] + cy[ch + 3];
            y += (y > 1u) ? 1u : 0u;                               /* :83 */
            cn[Ld + c] = n;
            cy[Ld + c] = y;
            sx[Ld + c] = ((sx[ch] + sx[ch + 1]) + sx[ch + 2]) + sx[ch + 3]; /* :79 */
            sy[Ld + c] = ((sy[ch] + sy[ch + 1]) + sy[ch + 2]) + sy[ch + 3];
            sm[Ld + c] = ((sm[ch] + sm[ch + 1]) + sm[ch + 2]) + sm[ch + 3]; /* :80 */
        }
    }

    /* TreeSortShader.comp:57-94, d = 0..9; the root start is (0, 0). */
    stx[0] = 0u; sty[0] = 0u;
    for (unsigned d = 0; d <= 9; ++d) {
        const size_t Ld = level_offset(d), Lc = level_offset(d + 1);
        const size_t nc = (size_t)1 << (2 * d);
        for (size_t c = 0; c < nc; ++c) {
            uint32_t ox = stx[Ld + c], oy = sty[Ld + c];
            if (cy[Ld + c] > 1u) oy += 1u;                         /* :76-77 */
            for (size_t k = 0; k < 4; ++k) {
                const size_t ch = Lc + 4 * c + k;
                stx[ch] = ox; sty[ch] = oy;
                ox += cn[ch]; oy += cy[ch];                        /* :80-85 */
            }
        }
    }

    /* TreeMoveShader.comp:53-84, d = 0..10 */
    long status = 0;
    for (unsigned d = 0; d <= 10; ++d) {
        const size_t Ld = level_offset(d);
        const size_t nc = (size_t)1 << (2 * d);
        const float r0 = box ? box[3] : ORACLE_SIM_SIZE * ORACLE_SIM_SIZE * 2.0f;
        const float radius = r0 / (float)(1u << (2 * d));          /* :70-71 */
        for (size_t c = 0; c < nc; ++c) {
            const size_t k = Ld + c;
            if (cy[k] > 1u) {
                float px = sx[k], py = sy[k];
                const float m = sm[k];
                if (m != 0.0f) { px = px / m; py = py / m; }      /* :66-67 */
                const uint32_t slot = sty[k];
                if ((size_t)slot >= tree->cap) { status = -1; continue; }
                tree->counts[slot] = cy[k];
                tree->radii[slot] = radius;
                tree->node_pos[2 * (size_t)slot] = px;
                tree->node_pos[2 * (size_t)slot + 1] = py;
                tree->node_mass[slot] = m;
                if (d == 10) sty[k] += 1u;                         /* :80-81 */
            }
        }
    }

    /* ParticleSortShader.comp:71-110, ascending index (canonical). */
    for (size_t i = 0; i < na && status == 0; ++i) {
        const float x = pos[2 * i], y = pos[2 * i + 1];
        const uint32_t c = cell[i];
        if (c == 0xffffffffu) continue;
        const size_t k = L10 + c;
        const uint32_t dsrc = stx[k]++;
        tree->sorted_src[dsrc] = (uint32_t)i;
        const uint32_t slot = sty[k]++;
        if ((size_t)slot >= tree->cap) { status = -1; break; }
        tree->counts[slot] = 1u;
        tree->radii[slot] = 0.0f;
        tree->node_pos[2 * (size_t)slot] = x;
        tree->node_pos[2 * (size_t)slot + 1] = y;
        tree->node_mass[slot] = mass[i];
    }

    free(cn); free(cy); free(stx); free(sty); free(sx); free(sy); free(sm); free(cell);
    if (status != 0) return -1;
    return inside ? (long)tree->counts[0] : 0;
}

long oracle_bh_build(size_t na, float *pos, float *mass, float *pos_out,
                     float *mass_out, oracle_tree *tree, size_t *n_inside) {
    return build_impl(na, pos, mass, pos_out, mass_out, NULL, tree, n_inside);
}

long oracle_bh_build_box(size_t na, float *pos, float *mass, float *pos_out,
                         float *mass_out, const float box[4], oracle_tree *tree,
                         size_t *n_inside) {
    return build_impl(na, pos, mass, pos_out, mass_out, box, tree, n_inside);
}

/* ---- Barnes-Hut: sparse build at 10 or 15 bits per axis (opt-in) ----------- */

/* Cell of a position on a 2^bits grid per axis, from the same normalised
 * coordinate as oracle_cell20 / oracle_cell20_box scaled by 2^bits instead of
 * 1024.  Scaling by a power of two is exact and truncation commutes with it,
 * so the 15-bit code shifted right by 10 is the 20-bit cell. */
uint32_t oracle_cell_bits(float x, float y, const float *box, unsigned bits) {
    const uint32_t T = 1u << bits;
    float qx, qy;
    if (box) {
        if (!isfinite(x) || !isfinite(y)) return 0xffffffffu;
        qx = (((x - box[0]) / box[2]) + 1.0f) * 0.5f;
        qy = (((y - box[1]) / box[2]) + 1.0f) * 0.5f;
        qx = qx * (float)T;
        qy = qy * (float)T;
        uint32_t ix = qx > 0.0f ? (uint32_t)qx : 0u, iy = qy > 0.0f ? (uint32_t)qy : 0u;
        if (ix > T - 1) ix = T - 1;
        if (iy > T - 1) iy = T - 1;
        return part1by1(ix) | (part1by1(iy) << 1);
    }
    qx = ((x / ORACLE_SIM_SIZE) + 1.0f) * 0.5f;
    qy = ((y / ORACLE_SIM_SIZE) + 1.0f) * 0.5f;
    qx = qx * (float)T;
    qy = qy * (float)T;
    if (qx >= 0.0f && qx < (float)T && qy >= 0.0f && qy < (float)T)
        return part1by1((uint32_t)qx) | (part1by1((uint32_t)qy) << 1);
    return 0xffffffffu;
}

size_t oracle_bh_capacity_sparse(size_t na, unsigned bits) {
    /* every particle plus, per depth, at most min(4^d, na/2) cells holding
     * two or more particles */
    size_t cells = 0;
    for (unsigned d = 0; d <= bits; ++d) {
        const size_t lvl = (size_t)1 << (2 * d);
        cells += lvl < na / 2 ? lvl : na / 2;
    }
    return na + cells + 1;
}

typedef struct {
    uint32_t key, idx;
} kv_pair;

static int kv_cmp(const void *a, const void *b) {
    const kv_pair *x = (const kv_pair *)a, *y = (const kv_pair *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0);
}

typedef struct {
    const kv_pair *kv;
    const float *pos, *mass;
    unsigned bits;
    float r0;
    oracle_tree *tree;
    size_t next; /* next free pre-order slot */
    int overflow;
} sparse_ctx;

typedef struct {
    uint32_t n, y;
    float sx, sy, sm;
} sparse_sum;

/* The cell of depth d holding sorted positions [lo, hi): the dense pyramid's
 * rule evaluated only where particles are.  A depth-`bits` leaf adds its
 * particles' m*p in ascending index from zero; a shallower cell is
 * ((c0 + c1) + c2) + c3 over its quadrants, an empty quadrant being +0 (the
 * pyramid's untouched cell); y = sum of the children's y, + 1 when above 1.
 * A cell of two or more particles takes the pre-order slot before its
 * subtree; particles are nodes of their own. */
static sparse_sum sparse_rec(sparse_ctx *c, unsigned d, size_t lo, size_t hi) {
    sparse_sum r = {0u, 0u, 0.0f, 0.0f, 0.0f};
    const size_t n = hi - lo;
    if (n == 0) return r;
    oracle_tree *t = c->tree;
    const size_t slot = c->next;
    if (n >= 2) c->next += 1;
    if (d == c->bits) {
        for (size_t k = lo; k < hi; ++k) {
            const uint32_t i = c->kv[k].idx;
            const float x = c->pos[2 * (size_t)i], y = c->pos[2 * (size_t)i + 1];
            const float m = c->mass[i];
            r.sx = r.sx + x * m;
            r.sy = r.sy + y * m;
            r.sm = r.sm + m;
            const size_t ps = c->next++;
            if (ps >= t->cap) { c->overflow = 1; continue; }
            t->sorted_src[k] = i;
            t->counts[ps] = 1u;
            t->radii[ps] = 0.0f;
            t->node_pos[2 * ps] = x;
            t->node_pos[2 * ps + 1] = y;
            t->node_mass[ps] = m;
        }
        r.n = (uint32_t)n;
        r.y = (uint32_t)n + (n >= 2 ? 1u : 0u);
    } else {
        const unsigned shift = 2u * (c->bits - 1u - d);
        sparse_sum v[4];
        size_t a = lo;
        for (uint32_t q = 0; q < 4; ++q) {
            size_t b = a;
            while (b < hi && ((c->kv[b].key >> shift) & 3u) == q) ++b;
            v[q] = sparse_rec(c, d + 1, a, b);
            a = b;
        }
        r.n = v[0].n + v[1].n + v[2].n + v[3].n;
        r.y = v[0].y + v[1].y + v[2].y + v[3].y;
        r.y += r.y > 1u ? 1u : 0u;
        r.sx = ((v[0].sx + v[1].sx) + v[2].sx) + v[3].sx;
        r.sy = ((v[0].sy + v[1].sy) + v[2].sy) + v[3].sy;
        r.sm = ((v[0].sm + v[1].sm) + v[2].sm) + v[3].sm;
    }
    if (n >= 2) {
        if (slot >= t->cap) { c->overflow = 1; return r; }
        float px = r.sx, py = r.sy;
        if (r.sm != 0.0f) { px = px / r.sm; py = py / r.sm; }
        t->counts[slot] = r.y;
        t->radii[slot] = c->r0 / (float)(1u << (2 * d));
        t->node_pos[2 * slot] = px;
        t->node_pos[2 * slot + 1] = py;
        t->node_mass[slot] = r.sm;
    }
    return r;
}

long oracle_bh_build_sparse(size_t na, float *pos, float *mass, float *pos_out,
                            float *mass_out, const float *box, unsigned bits,
                            oracle_tree *tree, size_t *n_inside) {
    if (bits != 10 && bits != 15) return -1;
    kv_pair *kv = (kv_pair *)malloc(sizeof(kv_pair) * (na ? na : 1));
    if (!kv) return -1;
    size_t inside = 0;
    for (size_t i = 0; i < na; ++i) {
        tree->sorted_src[i] = (uint32_t)na;
        const uint32_t c = oracle_cell_bits(pos[2 * i], pos[2 * i + 1], box, bits);
        if (c != 0xffffffffu) {
            kv[inside].key = c;
            kv[inside].idx = (uint32_t)i;
            ++inside;
        } else { /* the kill of oracle_bh_build */
            const float kill = -ORACLE_SIM_SIZE * 2.0f;
            pos[2 * i] = kill; pos[2 * i + 1] = kill; mass[i] = 0.0f;
            if (pos_out) { pos_out[2 * i] = kill; pos_out[2 * i + 1] = kill; }
            if (mass_out) mass_out[i] = 0.0f;
        }
    }
    if (n_inside) *n_inside = inside;
    qsort(kv, inside, sizeof(kv_pair), kv_cmp);
    sparse_ctx c = {kv, pos, mass, bits,
                    box ? box[3] : ORACLE_SIM_SIZE * ORACLE_SIM_SIZE * 2.0f, tree, 0, 0};
    (void)sparse_rec(&c, 0, 0, inside);
    free(kv);
    if (c.overflow) return -1;
    return (long)c.next;
}

/* ---- Barnes-Hut: traversal ------------------------------------------------- */

typedef struct {
    size_t na;
    uint32_t W;
    const float *pos, *mass;
    const oracle_tree *tree;
    size_t n_nodes;
    float eps2, theta2;
    float *acc;
    double *acc64;              /* may be NULL */
    uint64_t *visits, *accepts; /* per group */
    uint8_t *ambiguous;         /* per group, may be NULL */
} trav_ctx;

static void trav_range(size_t glo, size_t ghi, void *p) {
    trav_ctx *c = (trav_ctx *)p;
    const uint32_t W = c->W;
    float *lx = (float *)malloc(sizeof(float) * W * 6);
    float *ly = lx + W, *ax = lx + 2 * W, *ay = lx + 3 * W, *ddx = lx + 4 * W,
          *ddy = lx + 5 * W;
    double *bx = (double *)malloc(sizeof(double) * W * 2), *by = bx + W;
    uint8_t *real = (uint8_t *)malloc(W);
    uint32_t *src = (uint32_t *)malloc(sizeof(uint32_t) * W);
    const oracle_tree *t = c->tree;
    for (size_t g = glo; g < ghi; ++g) {
        uint32_t nreal = 0;
        for (uint32_t l = 0; l < W; ++l) {
            const size_t slot = g * W + l;
            src[l] = t->sorted_src[slot];                /* ForceShader :80 */
            real[l] = src[l] != (uint32_t)c->na;
            lx[l] = real[l] ? c->pos[2 * (size_t)src[l]] : 0.0f;
            ly[l] = real[l] ? c->pos[2 * (size_t)src[l] + 1] : 0.0f;
            ax[l] = 0.0f; ay[l] = 0.0f;
            bx[l] = 0.0; by[l] = 0.0;
            nreal += real[l];
        }
        uint64_t vis = 0, acc_n = 0;
        uint8_t amb = 0;
        size_t ind = 0;
        const size_t total = c->n_nodes; /* counts[0], :106 */
        while (ind < total) {
            const float nx = t->node_pos[2 * ind], ny = t->node_pos[2 * ind + 1];
            const float rad = t->radii[ind];
            int all = 1;
            for (uint32_t l = 0; l < W; ++l) {
                const float dx = nx - lx[l], dy = ny - ly[l];
                const float r2 = (dx * dx + dy * dy) + c->eps2;
                ddx[l] = dx; ddy[l] = dy;
                const float lim = r2 * c->theta2;
                const int ok = (rad <= lim) || !real[l];
                all &= ok;
                if (real[l] && rad > 0.0f) {
                    const float gap = fabsf(rad - lim);
                    if (gap <= 4.0f * (nextafterf(rad, INFINITY) - rad)) amb = 1;
                }
            }
            vis += nreal;
            if (all) {
                const float m = t->node_mass[ind];
                for (uint32_t l = 0; l < W; ++l) {
                    const float dx = ddx[l], dy = ddy[l];
                    const float r2 = (dx * dx + dy * dy) + c->eps2;
                    const float r = 1.0f / sqrtf(r2);
                    const float s = ((m * r) * r) * r;     /* :134 */
                    ax[l] = ax[l] + dx * s;
                    ay[l] = ay[l] + dy * s;
                    if (c->acc64) {
                        /* the same accepted node in double precision */
                        const double ex = (double)nx - (double)lx[l];
                        const double ey = (double)ny - (double)ly[l];
                        const double q = 1.0 / sqrt(ex * ex + ey * ey + (double)c->eps2);
                        const double w = (double)m * q * q * q;
                        bx[l] += ex * w;
                        by[l] += ey * w;
                    }
                }
                acc_n += nreal;
                ind += t->counts[ind];                     /* :136 */
            } else {
                ++ind;                                     /* :139 */
            }
        }
        for (uint32_t l = 0; l < W; ++l) {
            if (!real[l]) continue;
            c->acc[2 * (size_t)src[l]] = ax[l];
            c->acc[2 * (size_t)src[l] + 1] = ay[l];
            if (c->acc64) {
                c->acc64[2 * (size_t)src[l]] = bx[l];
                c->acc64[2 * (size_t)src[l] + 1] = by[l];
            }
        }
        c->visits[g] = vis;
        c->accepts[g] = acc_n;
        if (c->ambiguous) c->ambiguous[g] = amb;
    }
    free(lx); free(bx); free(real); free(src);
}

void oracle_bh_traverse(size_t na, uint32_t W, const float *pos,
                        const float *mass, const oracle_tree *tree,
                        size_t n_nodes, float eps2, float theta2, float *acc,
                        double *acc64, uint64_t *visits, uint64_t *accepts,
                        uint64_t *ambiguous, uint8_t *group_ambiguous,
                        int nthreads) {
    if (W == 0 || na % W != 0) return;
    const size_t groups = na / W;
    uint64_t *gv = (uint64_t *)calloc(groups ? groups : 1, sizeof(uint64_t));
    uint64_t *ga = (uint64_t *)calloc(groups ? groups : 1, sizeof(uint64_t));
    uint8_t *gamb = (uint8_t *)calloc(groups ? groups : 1, 1);
    trav_ctx c = {na, W, pos, mass, tree, n_nodes, eps2, theta2, acc, acc64, gv, ga, gamb};
    if (n_nodes > 0) parallel_for(groups, nthreads, trav_range, &c);
    uint64_t V = 0, A = 0, AM = 0;
    for (size_t g = 0; g < groups; ++g) { V += gv[g]; A += ga[g]; AM += gamb[g]; }
    if (group_ambiguous) memcpy(group_ambiguous, gamb, groups);
    if (visits) *visits = V;
    if (accepts) *accepts = A;
    if (ambiguous) *ambiguous = AM;
    free(gv); free(ga); free(gamb);
}

void oracle_bh_integrate(size_t na, const float *pos, const float *vel,
                         const float *mass, const float *acc, float *pos_out,
                         float *vel_out, float *mass_out, float G, float dt) {
    const float kv = G * dt;    /* GRAVITATIONAL_CONST * SIMULATION_TIME, :145 */
    const float kp = 0.5f * dt; /* 0.5 * SIMULATION_TIME, :146 */
    for (size_t i = 0; i < na; ++i) {
        if (oracle_cell20(pos[2 * i], pos[2 * i + 1]) == 0xffffffffu) {
            vel_out[2 * i] = vel[2 * i];
            vel_out[2 * i + 1] = vel[2 * i + 1];
            continue; /* killed: position and mass were written by the build */
        }
        for (int k = 0; k < 2; ++k) {
            const float v = vel[2 * i + k];
            const float nv = v + acc[2 * i + k] * kv;
            vel_out[2 * i + k] = nv;
            pos_out[2 * i + k] = pos[2 * i + k] + (v + nv) * kp;
        }
        if (mass_out) mass_out[i] = mass[i];
    }
}

int oracle_bh_step(size_t na, float *pos_in, float *vel_in, float *mass_in,
                   float *pos_out, float *vel_out, float *mass_out, float G,
                   float dt, float eps2, float theta2, uint32_t W,
                   uint64_t *visits, uint64_t *accepts, int nthreads) {
    const size_t cap = oracle_bh_capacity(na);
    oracle_tree t;
    t.cap = cap;
    t.counts = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    t.radii = (float *)malloc(sizeof(float) * cap);
    t.node_pos = (float *)malloc(sizeof(float) * 2 * cap);
    t.node_mass = (float *)malloc(sizeof(float) * cap);
    t.sorted_src = (uint32_t *)malloc(sizeof(uint32_t) * (na ? na : 1));
    float *acc = (float *)calloc(2 * (na ? na : 1), sizeof(float));
    int rc = 0;
    if (!t.counts || !t.radii || !t.node_pos || !t.node_mass || !t.sorted_src || !acc) {
        rc = -2;
    } else {
        const long nn = oracle_bh_build(na, pos_in, mass_in, pos_out, mass_out, &t, NULL);
        if (nn < 0) {
            rc = -1;
        } else {
            oracle_bh_traverse(na, W, pos_in, mass_in, &t, (size_t)nn, eps2, theta2,
                               acc, NULL, visits, accepts, NULL, NULL, nthreads);
            oracle_bh_integrate(na, pos_in, vel_in, mass_in, acc, pos_out, vel_out,
                                mass_out, G, dt);
        }
    }
    free(t.counts); free(t.radii); free(t.node_pos); free(t.node_mass);
    free(t.sorted_src); free(acc);
    return rc;
}
