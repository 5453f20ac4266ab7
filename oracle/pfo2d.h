/*
 * TEST INFRASTRUCTURE ONLY: CPU oracle of the 2D per-cell phase-field update.
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py may load
 * the library built from this file (oracle/_build/libpfo.so). It is the
 * checker of the HIP kernels, never a fallback for them.
 *
 * What it restates (reference file:line):
 *   process_cell (classes 0-4)        src/kernels.cu:53-592
 *   process_cell_fallback (class 5)   src/kernels.cu:604-937
 *   step_body clear-ahead, queue      src/kernels.cu:948-1024
 *   k_init_tiles / k_init_moments /
 *   k_scatter_all                     src/kernels.cu:1050-1190
 *   k_verify_cells / k_verify_S       src/kernels.cu:1193-1240
 *   k_pack_traj                       src/kernels.cu:1242-1265
 *   grid+jitter placement, ic_theta,
 *   ic_v_A                            src/sim.cu:108-144, include/kernels.cuh:153-172
 *   CellState / StepArgs layouts      include/kernels.cuh:27-116
 *
 * Arithmetic: binary32 per-pixel arithmetic in source order with no
 * multiply-add contraction (the semantics the reference's own host fixture
 * tests pin, tests/mobility_model_cpu_test.cpp:77-150), fp64 reductions.
 *
 * Reduction order is a parameter, because the reference fixes it by its CUDA
 * launch shape (768 threads, 32-lane warps):
 *   PFO_ORDER_MI355X    the traversal of this project's gfx950 kernel
 *                       (1024 threads, 64-lane wavefronts) -- the oracle then
 *                       reproduces the kernel bit for bit;
 *   PFO_ORDER_REFERENCE the reference kernel's traversal.
 * Each order is: per-thread sequential fp64 partial sums in raster order,
 * a lane tree (offsets W/2 .. 1) inside each wave, then waves in ascending
 * index added to 0.0.
 */
#ifndef PFO2D_H
#define PFO2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PFO_ORDER_MI355X = 0, PFO_ORDER_REFERENCE = 1 };

enum {
    PFO_FLAG_S_OVERFLOW = 0, PFO_FLAG_Q_CLAMP = 1, PFO_FLAG_SUPPORT_CLIP = 2,
    PFO_FLAG_CLASS_EXHAUSTED = 3, PFO_FLAG_S_NEGATIVE = 4,
    PFO_FLAG_NONFINITE = 5, PFO_FLAG_V_NONPOS = 6,
    PFO_FLAG_CLASS_UNSUPPORTED = 7, PFO_FLAG_COUNT = 8
};

#define PFO_TILE_PITCH 288
#define PFO_TILE_AREA (288 * 288)

/* 192-byte mirror of CellState (include/kernels.cuh:27-57). */
typedef struct __attribute__((aligned(64))) pfo_cell {
    int32_t gx0, gy0;
    uint8_t cls, cls_written[2], pad0;
    int32_t global_id;
    float gamma, v_A, R_tgt, theta;
    double V, Cx, Cy, perim;
    int32_t bb_lo_x, bb_hi_x, bb_lo_y, bb_hi_y;
    float vx, vy;
    double Ix, Iy;
    uint32_t promote_ctr, shift_ctr, tumble_ctr;
    float phi_max;
    uint32_t reserved[2];
    double adv_q[6];
    uint64_t adv_count;
    uint32_t reserved_unused;
    float M_pf;
} pfo_cell;

/* 240-byte mirror of MomentAccum (common/velocity_moments.h:30-40). */
typedef struct pfo_moment {
    double q[12];
    uint64_t count;
    double held[6];
    uint64_t spatial_count, last_sample, last_interval;
    double last_totals[4];
    double correction[4];
} pfo_moment;

typedef struct pfo_traj {
    int32_t global_id, cls;
    float cx, cy, vx, vy, theta, volume, perim, gamma, v_A, phi_max;
} pfo_traj;

/* Arguments of one step; host pointers. Mirrors StepArgs
 * (include/kernels.cuh:74-116) with the device-resident step counter and
 * cursors replaced by the step value. phi_in/phi_out are float* for
 * pfo2d_step and double* for pfo2d_step_f64. */
typedef struct pfo2d_step_args {
    const void* phi_in;
    void* phi_out;
    const uint32_t* S_rd;
    uint32_t* S_sc;
    uint32_t* S_cl;
    pfo_cell* cell;
    uint8_t* cell_cls;
    const uint32_t* perm;
    uint32_t* flags;
    double* vel_f64;          /* FP64 oracle only: 2N carried velocities */
    uint64_t step;
    int N, L, P, parity_out;
    float dt;
    double A0, vol_scale;
    float bulk_scale, rep_coeff, mot_coeff;
    double dt_d, bulk_d, rep_d, mot_d;   /* FP64 oracle coefficients */
    uint64_t polarity_seed;
    double p_tumble;
    int full_moment_every;
    uint64_t clear_ahead_words;
    /* velocity observation (VelocityStep, include/kernels.cuh:216): 0 none,
     * 1 advection, 2 spatial (moments: N accumulators) */
    int observation;
    pfo_moment* moments;
    double physical_dt;
} pfo2d_step_args;

/* One step over all cells (cells in perm order; each cell independent).
 * Returns 0, or -1 on invalid arguments. threads <= 0 uses OpenMP's default. */
int pfo2d_step(const pfo2d_step_args* a, int order, int threads);
int pfo2d_step_f64(const pfo2d_step_args* a, int threads);

/* Initial condition. */
void pfo2d_init_tiles(float* phi_a, float* phi_b, pfo_cell* cell,
                      const uint8_t* cls, int N, int L, const float* cx,
                      const float* cy, float radius_eff, float kappa_iface);
void pfo2d_init_tiles_f64(double* phi_a, double* phi_b, pfo_cell* cell,
                          const uint8_t* cls, int N, int L, const float* cx,
                          const float* cy, double radius_eff,
                          double kappa_iface);
void pfo2d_init_moments(pfo_cell* cell, const uint8_t* cls, const float* phi,
                        int N, int order);
void pfo2d_init_moments_f64(pfo_cell* cell, const uint8_t* cls,
                            const double* phi, int N);
void pfo2d_scatter_all(const float* phi, const pfo_cell* cell,
                       const uint8_t* cls, uint32_t* S, int N, int L, int P,
                       uint32_t* flags);
void pfo2d_scatter_all_f64(const double* phi, const pfo_cell* cell,
                           const uint8_t* cls, uint32_t* S, int N, int L,
                           int P, uint32_t* flags);
void pfo2d_verify_cells(const float* phi, const pfo_cell* cell,
                        const uint8_t* cls, int N, double* out_V,
                        float* out_outside_max);
uint32_t pfo2d_verify_S(const uint32_t* S, size_t n);
void pfo2d_morton_perm(const pfo_cell* cell, int N, int L, uint32_t* perm);
void pfo2d_velocity_moments(const pfo2d_step_args* A, pfo_moment* out, double pdt);
void pfo2d_pack_traj(const pfo_cell* cell, const uint8_t* cls, pfo_traj* out,
                     int N, int L);

/* Scalars. */
int pfo2d_tumble(uint64_t step, int32_t gid, uint64_t stream, double p_tumble,
                 float* theta);
void pfo_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double pfo_u53(uint32_t a, uint32_t b);
uint32_t pfo_qof(float phi);
float pfo_ic_theta(int32_t gid, uint64_t stream);
double pfo_ic_v_A(int32_t gid, uint64_t seed, double v_A, double sigma);
void pfo_sincosf(float theta, float* s, float* c);
double pfo_tanh(double y);
float pfo_rhs(float lap, float phi, float So, float gx, float gy, float gam,
              float dwC, float volC, float repC, float vx, float vy,
              float mrel);
float pfo_lap9(const float* p, int pitch);
int pfo_class_containing(int ex, int ey, int slack);
int pfo_domain_side_for(int n, double radius, double rho);
int pfo_s_pitch_for(int side);

/* Default grid+jitter placement (src/sim.cu:108-129). */
void pfo2d_grid_jitter_centres(int N, int L, uint64_t seed, float* cx,
                               float* cy);
/* Seeded jittered-grid harness placement used by the benchmark workloads:
 * cell i at ((i%g)+0.5)*L/g, ((i/g)+0.5)*L/g plus uniform jitter in
 * [-amp, +amp] from Philox counter (i, 0x4A495454, 0, 0), key seed. */
void pfo2d_harness_centres(int N, int L, int g, double amp, uint64_t seed,
                           float* cx, float* cy);

#ifdef __cplusplus
}
#endif
#endif /* PFO2D_H */
