/*
 * pfh.h -- C ABI of the MI355X (gfx950) phase-field hot path.
 *
 * This is the drop-in boundary for the reference's 2D launch/operator API
 * (include/kernels.cuh:176-244). Every entry point takes plain device
 * pointers and sizes, never allocates or frees caller memory, returns an int
 * status (0 = success, negative = contract or HIP error; nothing throws
 * across the ABI) and is safe to record into a HIP graph unless stated
 * otherwise. `stream` is a hipStream_t passed as void* (NULL = default
 * stream) so FFI callers need no HIP headers.
 *
 * POD records keep the reference's byte layout so checkpoints, host code and
 * tests read them unchanged; the static asserts below pin every offset the
 * reference pins.
 */
#ifndef PFH_H
#define PFH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define PFH_OK 0
#define PFH_ERR_INVALID (-1)     /* contract violation (null, size, range) */
#define PFH_ERR_HIP (-2)         /* a HIP runtime call failed */
#define PFH_ERR_UNSUPPORTED (-3) /* feature outside this build's scope */
#define PFH_ERR_NO_DEVICE (-4)   /* no gfx950 device visible */

/* ---- sticky integrity flags: include/params.cuh:434-445 ---- */
enum {
    PFH_FLAG_S_OVERFLOW = 0,
    PFH_FLAG_Q_CLAMP = 1,
    PFH_FLAG_SUPPORT_CLIP = 2,
    PFH_FLAG_CLASS_EXHAUSTED = 3,
    PFH_FLAG_S_NEGATIVE = 4,
    PFH_FLAG_NONFINITE = 5,
    PFH_FLAG_V_NONPOS = 6,
    PFH_FLAG_CLASS_UNSUPPORTED = 7,
    PFH_FLAG_COUNT = 8
};

/* ---- 2D geometry: include/params.cuh:121-147 ---- */
#define PFH_TILE_PITCH 288
#define PFH_TILE_AREA (288 * 288)
#define PFH_NUM_CLASSES 6

/* Runtime-only observer accumulator (common/velocity_moments.h:26-30). */
typedef struct pfh_advection_accum {
    double q[6];
    unsigned long long count;
} pfh_advection_accum;

/* Per-cell state: include/kernels.cuh:27-57. 192 bytes, 64-byte aligned. */
typedef struct pfh_cell_state {
    int32_t gx0, gy0;
    uint8_t cls;
    uint8_t cls_written[2];
    uint8_t pad0;
    int32_t global_id;
    float gamma, v_A, R_tgt;
    float theta;
    double V;
    double Cx, Cy;
    double perim;
    int32_t bb_lo_x, bb_hi_x, bb_lo_y, bb_hi_y;
    float vx, vy;
    double Ix, Iy;
    uint32_t promote_ctr;
    uint32_t shift_ctr;
    uint32_t tumble_ctr;
    float phi_max;
    uint32_t reserved[2];
    pfh_advection_accum advection;
    uint32_t reserved_unused;
    float M_pf;
} __attribute__((aligned(64))) pfh_cell_state;

/* Trajectory record: include/kernels.cuh:60-70. */
typedef struct pfh_traj_packed_cell {
    int32_t global_id;
    int32_t cls;
    float cx, cy;
    float vx, vy;
    float theta;
    float volume;
    float perim;
    float gamma, v_A;
    float phi_max;
} pfh_traj_packed_cell;

/* Immutable arguments of one 2D step: include/kernels.cuh:74-116.
 * Step counters and the queue cursors stay device-resident so one set of
 * arguments can be replayed from a graph (parity slot = step % 2, S rotation
 * slot = step % 3, as in src/sim.cu:578-618). */
typedef struct pfh2d_step_args {
    const float* phi_in;
    float* phi_out;
    const uint32_t* S_rd;
    uint32_t* S_sc;
    uint32_t* S_cl;
    pfh_cell_state* cell;
    uint8_t* cell_cls;
    const uint32_t* perm;
    unsigned long long* cursor_use;
    unsigned long long* cursor_clear;
    const unsigned long long* step_rd;
    unsigned long long* step_wr;
    uint32_t* flags;
    void* moments;            /* observer accumulators: must be NULL */
    double physical_dt;
    int N;
    int L;
    int P;
    int parity_out;
    float dt;
    double A0;
    double vol_scale;
    float bulk_scale;
    float rep_coeff;
    float mot_coeff;
    unsigned long long seed;
    unsigned long long polarity_seed;
    double p_tumble;
    int full_moment_every;
    unsigned long long clear_ahead_words;
} pfh2d_step_args;

/* Kernel resource report: include/kernels.cuh:230-244 (KernelStats). */
typedef struct pfh_kernel_stats {
    int regs;             /* architected VGPRs per thread */
    int agprs;            /* accumulation VGPRs per thread */
    int sgprs;
    size_t local_bytes;   /* private segment (spills + stack) per thread */
    int static_smem;      /* LDS bytes */
    int dynamic_smem;
    int ctas_per_sm;      /* workgroups resident per CU */
    int warps_per_sm;     /* wavefronts resident per CU */
    double occupancy;     /* resident threads / max threads per CU */
} pfh_kernel_stats;

#ifndef __cplusplus
_Static_assert(sizeof(pfh_cell_state) == 192, "pfh_cell_state must be 192 B");
_Static_assert(_Alignof(pfh_cell_state) == 64, "pfh_cell_state alignment");
_Static_assert(offsetof(pfh_cell_state, advection) == 128, "advection moved");
_Static_assert(offsetof(pfh_cell_state, reserved) == 120, "reserved moved");
_Static_assert(offsetof(pfh_cell_state, M_pf) == 188, "M_pf moved");
_Static_assert(sizeof(pfh_traj_packed_cell) == 48, "traj record size");
#else
static_assert(sizeof(pfh_cell_state) == 192, "pfh_cell_state must be 192 B");
static_assert(alignof(pfh_cell_state) == 64, "pfh_cell_state alignment");
static_assert(offsetof(pfh_cell_state, advection) == 128, "advection moved");
static_assert(offsetof(pfh_cell_state, reserved) == 120, "reserved moved");
static_assert(offsetof(pfh_cell_state, M_pf) == 188, "M_pf moved");
static_assert(sizeof(pfh_traj_packed_cell) == 48, "traj record size");
#endif

/* ---- library ---- */
/* ABI version (major*10000 + minor*100 + patch). */
int pfh_abi_version(void);
/* Last HIP error text recorded by this thread's failing call ("" if none). */
const char* pfh_last_error(void);

/* Plain 16-byte-vector device copy of `bytes` (a multiple of 16, 16 B
 * aligned). No counterpart in the reference: it is the known-size kernel that
 * calibrates hardware traffic counters in the measurement tooling. */
int pfh_stream_copy(void* dst, const void* src, size_t bytes, void* stream);

/* ---- 2D launch API ---- */

/* Replaces configure_k_step_smem() (include/kernels.cuh:219): sets the
 * kernel's LDS attribute on the current device. Idempotent. */
int pfh2d_configure_step(void);
/* Replaces configure_morton_smem(int) (include/kernels.cuh:220). */
int pfh2d_configure_morton(int smem_bytes);
/* Replaces k_step_grid(device) (include/kernels.cuh:221): the persistent
 * grid, one 1024-thread workgroup per compute unit. Returns the grid size,
 * or a negative status. */
int pfh2d_step_grid(int device);

/* Replaces launch_step (include/kernels.cuh:223-225, src/kernels.cu:1329-1363).
 * One persistent work-queue launch advances every cell by one step.
 * l2_base/l2_bytes/l2_hit_ratio are accepted and ignored: gfx950 has no
 * L2 access-policy window. observation must be 0 (Disabled); the velocity
 * observers are outside this build (returns PFH_ERR_UNSUPPORTED). */
int pfh2d_launch_step(const pfh2d_step_args* args, int grid, void* stream,
                      const void* l2_base, size_t l2_bytes,
                      float l2_hit_ratio, int observation);

/* Replaces query_kernel_stats (include/kernels.cuh:243-244) for k_step. */
int pfh2d_query_step_stats(int device, pfh_kernel_stats* out);

/* Replaces k_init_tiles (src/kernels.cu:1050-1094): zero both tiles of every
 * cell and draw phi = 0.5 (1 - tanh(kappa_iface (r - radius_eff))) on the
 * active window of class cls[n] centred on (seed_cx[n], seed_cy[n]); sets
 * gx0/gy0/cls/cls_written. */
int pfh2d_init_tiles(float* phi_a, float* phi_b, pfh_cell_state* cell,
                     const uint8_t* cls, int N, int L, const float* seed_cx,
                     const float* seed_cy, float radius_eff,
                     float kappa_iface, void* stream);
/* Replaces k_init_moments (src/kernels.cu:1098-1165): V, Cx, Cy, support box
 * and phi_max of the stored field in the update's reduction order. */
int pfh2d_init_moments(pfh_cell_state* cell, const uint8_t* cls,
                       const float* phi, int N, void* stream);
/* Replaces k_scatter_all (src/kernels.cu:1167-1190): S += sum_n q(phi_n).
 * S must be zeroed by the caller. */
int pfh2d_scatter_all(const float* phi, const pfh_cell_state* cell,
                      const uint8_t* cls, uint32_t* S, int N, int L, int P,
                      uint32_t* flags, void* stream);
/* Replaces k_zero_u32 (src/kernels.cu:1044-1048). */
int pfh2d_zero_u32(uint32_t* p, size_t n, void* stream);
/* Replaces k_verify_cells (src/kernels.cu:1193-1232). */
int pfh2d_verify_cells(const float* phi, const pfh_cell_state* cell,
                       const uint8_t* cls, int N, double* out_V,
                       float* out_outside_max, void* stream);
/* Replaces k_verify_S (src/kernels.cu:1234-1240): atomic max into *out_max
 * (caller zeroes it). */
int pfh2d_verify_S(const uint32_t* S, size_t n, uint32_t* out_max,
                   void* stream);
/* Replaces k_pack_traj (src/kernels.cu:1242-1265). out may be device memory
 * or mapped pinned host memory. */
int pfh2d_pack_traj(const pfh_cell_state* cell, const uint8_t* cls,
                    pfh_traj_packed_cell* out, int N, int L, void* stream);
/* Replaces k_morton_sort (src/kernels.cu:1267-1302): M is a power of two
 * >= N, M <= 16384. */
int pfh2d_morton_sort(const pfh_cell_state* cell, uint32_t* perm, int N,
                      int M, int L, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PFH_H */
