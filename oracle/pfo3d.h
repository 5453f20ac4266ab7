/*
 * TEST INFRASTRUCTURE ONLY. The CPU oracle of the 3D per-cell phase-field
 * step: the checker for the gfx950 kernels. Only tests, the smoke check and
 * the CPU baseline leg of the benchmark load it; no product code links it.
 *
 * Restated from the reference (file:line of the behaviour each part follows):
 *   CellState3D layout                include/pf3d/kernels.cuh:98-137
 *   flags, fatal mask                 include/pf3d/kernels.cuh:55-91
 *   sphere seeding                    src/pf3d/kernels.cu:1071-1122
 *   S scatter                         src/pf3d/kernels.cu:1148-1186
 *   measurement + commit              src/pf3d/kernels.cu:427-496,1359-1409,1591-1737
 *   motion / tumble                   src/pf3d/kernels.cu:1739-1788,
 *                                     include/pf3d/rng.cuh:19-150
 *   moment-free update                src/pf3d/kernels.cu:498-793,
 *                                     src/pf3d/kernels_periodic_slab_walk.cu:54-318
 *   repair / finalize / restore /
 *   advance                           src/pf3d/kernels.cu:2128-2263
 *   step sequence                     src/pf3d/sim.cu:1460-1651
 *   seed-radius bisection             include/pf3d/reference.hpp:712-745
 *
 * Periodic XYZ geometry with cubic brick storage only (the reference's
 * default geometry).
 */
#ifndef PFO3D_H
#define PFO3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PFO3D_FLAG_S_OVERFLOW = 0,
    PFO3D_FLAG_Q_CLAMP = 1,
    PFO3D_FLAG_S_NEGATIVE = 2,
    PFO3D_FLAG_NONFINITE = 3,
    PFO3D_FLAG_V_NONPOSITIVE = 4,
    PFO3D_FLAG_SUPPORT_EXHAUSTED = 5,
    PFO3D_FLAG_SUPPORT_EDGE = 6,
    PFO3D_FLAG_DESTINATION_CLIP = 7,
    PFO3D_FLAG_INVALID_GEOMETRY = 8,
    PFO3D_FLAG_SLAB_TOP_CONTACT = 9,
    PFO3D_FLAG_COUNT = 10
};
#define PFO3D_FATAL_MASK 0x3BFu  /* every flag except SUPPORT_EDGE */

/* Reduction / operation orders the oracle reproduces. */
#define PFO3D_ORDER_MI355X 0    /* this build's kernels, bit for bit */
#define PFO3D_ORDER_REFERENCE 1 /* the reference's tiles, warps and stencil order */

typedef struct pfo3d_cell {
    int64_t global_id;
    int64_t origin_x, origin_y, origin_z;
    float polarity_x, polarity_y, polarity_z;
    float velocity_x, velocity_y, velocity_z;
    float gamma, v_A, R_tgt;
    float phi_max;
    uint32_t shift_ctr;
    uint32_t tumble_ctr;
    uint32_t flags;
    uint32_t pad0;
    int32_t bb_lo_x, bb_hi_x, bb_lo_y, bb_hi_y, bb_lo_z, bb_hi_z;
    int32_t pending_shift_x, pending_shift_y, pending_shift_z;
    uint32_t pad1;
    double V, Cx, Cy, Cz, surface, Ix, Iy, Iz;
    uint32_t storage_edge;
    uint32_t pad2;
    double wall_overlap;
    double outside_slit_volume;
    uint32_t reserved[10];
} __attribute__((aligned(64))) pfo3d_cell;

_Static_assert(sizeof(pfo3d_cell) == 256, "CellState3D is 256 bytes");
_Static_assert(offsetof(pfo3d_cell, storage_edge) == 192, "storage_edge offset");
_Static_assert(offsetof(pfo3d_cell, V) == 128, "V offset");

/* One 3D step. Field/aggregate buffers swap every begun step; counters[0]
 * counts accepted steps (the Philox step), counters[1] begun steps (the
 * full-moment cadence). */
typedef struct pfo3d_step_args {
    const float* phi_in;
    float* phi_out;
    const uint32_t* S_in;
    uint32_t* S_out;
    pfo3d_cell* cells;
    pfo3d_cell* accepted;
    uint32_t* flags;
    uint64_t* counters;
    void* partials;      /* device scratch; unused by the oracle */
    uint32_t* arrivals;  /* device scratch; unused by the oracle */
    int N, B, L, P;
    int shards;
    int max_shift;
    int support_margin;
    int full_moment_every;
    float dt;
    float bulk_scale;
    float interaction_coeff;
    float motility_coeff;
    double V0;
    double volume_scale;
    double p_tumble;
    uint64_t polarity_stream;
    uint64_t aging_steps;
} pfo3d_step_args;

_Static_assert(sizeof(pfo3d_step_args) == 168, "step args layout");

/* k_initialize_spheres with this project's tanh. centres: 3*N doubles. */
void pfo3d_init_spheres(float* phi_even, float* phi_odd, pfo3d_cell* cells,
                        const double* centres, int N, int B, float lambda,
                        float seed_radius);
/* k_scatter_current: S += q(phi) (caller clears S). */
void pfo3d_scatter(const float* phi, pfo3d_cell* cells, uint32_t* S,
                   uint32_t* flags, int N, int B, int L, int P);
/* Measurement of every cell + commit (+ motion when apply_motion). */
void pfo3d_measure(const pfo3d_step_args* a, int order, int apply_tumble,
                   int compute_surface, int apply_motion, int threads);
/* The whole begin_step + finish_step sequence (binary32). */
int pfo3d_step(const pfo3d_step_args* a, int order, int threads);
/* FP64 variant: phi in double, velocities kept in vel (3N doubles), and
 * coef = {dt, bulk_scale, interaction, motility} in double. */
int pfo3d_step_f64(const pfo3d_step_args* a, const double* phi_in,
                   double* phi_out, double* vel, const double* coef, int threads);
void pfo3d_measure_f64(const pfo3d_step_args* a, const double* phi,
                       double* vel, const double* coef, int apply_tumble,
                       int compute_surface, int apply_motion, int threads);
void pfo3d_scatter_f64(const double* phi, pfo3d_cell* cells, uint32_t* S,
                       uint32_t* flags, int N, int B, int L, int P);

/* Helpers exposed for the primitive tests. */
double pfo3d_calibrate_seed_radius(double radius, double lambda, int B,
                                   double tol, int max_iter, int* iters);
void pfo3d_initial_polarity(int64_t gid, uint64_t stream, double out[3]);
int pfo3d_tumble_sample(uint64_t step, int64_t gid, uint64_t stream,
                        double p, const double cur[3], double out[3]);
float pfo3d_lap27(const float* f27, int order);
int pfo3d_brick_edge(double radius, double lambda);

#ifdef __cplusplus
}
#endif
#endif /* PFO3D_H */
