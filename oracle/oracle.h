/*
 * CPU oracle for the force-and-integrate step — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it.  The product path
 * (libgsim) does not link, load or fall back to it.
 *
 * It restates, in plain C compiled with -ffp-contract=off (no FMA
 * contraction, IEEE division and square root), the FP32 arithmetic of the
 * reference's compute shaders in their source operation order:
 *   - direct sum:  src/Simulation/Direct/Shaders/SimShader.comp:37-74
 *   - Barnes-Hut:  src/Simulation/BarnesHut/Shaders/{Clear,Init,TreeInit,
 *                  TreeSort,TreeMove,ParticleSort,Force}Shader.comp, in the
 *                  dispatch order of BarnesHutSimulation.cpp:885-918.
 * The Barnes-Hut build is restated with the reference's dense 11-level
 * pyramid (1,398,101 cells), canonicalised where the reference is racy:
 * leaf sums and within-leaf order follow ascending particle index, the leaf
 * node count uses the race-free y = n + [n >= 2], and the root start is 0.
 *
 * Each exported function names the reference lines it follows.  Pinning of
 * each function (golden vectors, closed forms, brute force) lives in tests/.
 */
#ifndef GSIM_ORACLE_H
#define GSIM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Domain constants of the reference (BarnesHutSimulation.cpp:579,581). */
#define ORACLE_SIM_SIZE 500.0f
#define ORACLE_TREE_SIZE 1024u
#define ORACLE_TREE_DEPTH 10u
#define ORACLE_PYRAMID_CELLS 1398101u

/* ---- direct sum ---------------------------------------------------------- */

/* FP32 accelerations a_i = sum_j d * (((m_j*r)*r)*r), r = 1/sqrt(d.d + eps2),
 * j ascending over all na particles (SimShader.comp:43-65).  acc is [2*na]. */
void oracle_ds_accel_f32(size_t na, const float *pos, const float *mass,
                         float eps2, float *acc, int nthreads);

/* The same sum in FP64 (inputs widened, eps2 given in double), the error
 * yardstick for every FP32 result. */
void oracle_ds_accel_f64(size_t na, const float *pos, const float *mass,
                         double eps2, double *acc, int nthreads);
/* FP64 direct sum for the listed targets only: acc[2k..2k+1] is the
 * acceleration of particle targets[k] (the error yardstick on a sample). */
void oracle_ds_accel_f64_at(size_t na, const float *pos, const float *mass, double eps2,
                            const uint32_t *targets, size_t ntargets, double *acc,
                            int nthreads);

/* One direct-sum step (SimShader.comp:37-74 with the host constants of
 * DirectSimulation.cpp:175-181): v' = v + a*(dt*G); p' = p + ((v+v')*0.5)*dt.
 * Mass is not written. */
void oracle_ds_step(size_t na, const float *pos_in, const float *vel_in,
                    const float *mass_in, float *pos_out, float *vel_out,
                    float G, float dt, float eps2, int nthreads);

/* The same step carried entirely in FP64 (state and arithmetic), the
 * trajectory yardstick.  pos/vel are [2*na] doubles, updated in place. */
void oracle_ds_step_f64(size_t na, double *pos, double *vel, const double *mass,
                        double G, double dt, double eps2, int nthreads);

/* ---- Barnes-Hut ------------------------------------------------------------ */

/* 20-bit cell of one position (InitShader.comp:80-98): q = ((p/S + 1)*0.5)*T,
 * in bounds iff 0 <= q < T on both axes.  Returns 0xffffffff when out of
 * bounds (the particle is killed by the reference). */
uint32_t oracle_cell20(float x, float y);

/* Node arrays of the flattened pre-order tree (BarnesHutSimulation.cpp:57-66). */
typedef struct {
    uint32_t *counts;    /* [cap] subtree size, 1 for a particle node        */
    float *radii;        /* [cap] 500000/4^d for cells, 0 for particles      */
    float *node_pos;     /* [2*cap] centre of mass (or particle position)    */
    float *node_mass;    /* [cap]                                            */
    uint32_t *sorted_src;/* [na]  particle index per sorted slot, na = empty */
    size_t cap;
} oracle_tree;

/* Clear + Init + TreeInit + TreeSort + TreeMove + ParticleSort on one input
 * set.  pos/mass are the input set and are modified for killed particles
 * (pos := (-1000,-1000), mass := 0, InitShader.comp:110-117); pos_out and
 * mass_out receive the same kill writes when non-NULL.  Returns the number of
 * nodes written (counts[0] when at least one particle is in bounds, else 0),
 * or -1 when cap is too small. */
long oracle_bh_build(size_t na, float *pos, float *mass, float *pos_out,
                     float *mass_out, oracle_tree *tree, size_t *n_inside);

/* Adaptive bounding box (beyond the reference, opt-in): box = {cx, cy, h, r0}
 * from the extremes of the finite positions (see oracle.c), the cell of a
 * particle in it, and the build of oracle_bh_build over that box instead of
 * the fixed [-500, 500) domain: only non-finite positions are killed, and the
 * cell radius is r0 / 4^d. */
void oracle_bh_box(size_t na, const float *pos, float box[4]);
uint32_t oracle_cell20_box(float x, float y, const float box[4]);
long oracle_bh_build_box(size_t na, float *pos, float *mass, float *pos_out,
                         float *mass_out, const float box[4], oracle_tree *tree,
                         size_t *n_inside);

/* Sparse build at `bits` = 10 or 15 bits per axis (15: beyond the reference,
 * opt-in).  The dense pyramid's rule evaluated recursively over the occupied
 * cells only: a depth-`bits` leaf adds its particles' m*p in ascending index
 * from zero, a shallower cell is ((c0+c1)+c2)+c3 over its quadrants with an
 * empty quadrant +0, y = sum + [sum > 1], pre-order slots.  At 10 bits it is
 * bit-identical to oracle_bh_build (box == NULL) / oracle_bh_build_box; at 15
 * bits the cells are those of oracle_cell_bits and the leaves sit at depth
 * 15 (radius r0 / 4^15).  box == NULL means the fixed [-500, 500) domain with
 * its kill.  Returns the node count, or -1 when cap is too small or bits is
 * neither 10 nor 15. */
uint32_t oracle_cell_bits(float x, float y, const float *box, unsigned bits);
size_t oracle_bh_capacity_sparse(size_t na, unsigned bits);
long oracle_bh_build_sparse(size_t na, float *pos, float *mass, float *pos_out,
                            float *mass_out, const float *box, unsigned bits,
                            oracle_tree *tree, size_t *n_inside);
/* Group traversal (ForceShader.comp:78-141): groups of W consecutive sorted
 * slots walk one path; a node is accepted iff every non-sentinel lane has
 * radius <= (d.d + eps2)*theta2.  Writes acc[2*src] for non-sentinel lanes
 * only.  visits/accepts (may be NULL) receive the real-lane node visits and
 * acceptances; ambiguous (may be NULL) counts the groups that had a vote with
 * a real lane within 4 ulp of the threshold, and group_ambiguous (may be
 * NULL, [na/W]) flags them.  acc64 (may be NULL) receives the same accepted
 * node set summed in double precision: the error yardstick for a GPU walk
 * that adds the same terms in another order. */
void oracle_bh_traverse(size_t na, uint32_t W, const float *pos,
                        const float *mass, const oracle_tree *tree,
                        size_t n_nodes, float eps2, float theta2, float *acc,
                        double *acc64,
                        uint64_t *visits, uint64_t *accepts,
                        uint64_t *ambiguous, uint8_t *group_ambiguous,
                        int nthreads);

/* Integrate of ForceShader.comp:145-150 for every in-bounds particle:
 * v' = v + a*(G*dt); p' = p + (v+v')*(0.5*dt).  Killed particles keep
 * vel_out = vel_in (the reference leaves it stale). */
void oracle_bh_integrate(size_t na, const float *pos, const float *vel,
                         const float *mass, const float *acc,
                         float *pos_out, float *vel_out, float *mass_out,
                         float G, float dt);

/* Full Barnes-Hut step (BarnesHutSimulation.cpp:885-918).  dt is
 * simulation_time*simulation_speed.  in_* are modified by kills. */
int oracle_bh_step(size_t na, float *pos_in, float *vel_in, float *mass_in,
                   float *pos_out, float *vel_out, float *mass_out, float G,
                   float dt, float eps2, float theta2, uint32_t W,
                   uint64_t *visits, uint64_t *accepts, int nthreads);

/* Node capacity the reference allocates: na + 1,398,101. */
size_t oracle_bh_capacity(size_t na);

#ifdef __cplusplus
}
#endif
#endif
