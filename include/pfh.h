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

/* Runtime-only observer accumulator (common/velocity_moments.h:26-30):
 * q = integrated active (x, y), interaction (x, y) and stored-velocity
 * rounding (x, y) velocity components; count = recorded steps. */
typedef struct pfh_advection_accum {
    double q[6];
    unsigned long long count;
} pfh_advection_accum;

/* Velocity observation modes of pfh2d_launch_step (VelocityStep,
 * include/kernels.cuh:216). */
#define PFH_OBSERVE_DISABLED 0
#define PFH_OBSERVE_ADVECTION 1
#define PFH_OBSERVE_SPATIAL 2

/* Per-cell velocity-moment accumulator (MomentAccum,
 * common/velocity_moments.h:30-40), 240 bytes. Column pairs (x, y) of q:
 * 0 active, 2 interaction, 4 interfacial, 6 overlap, 8 area, 10 advection
 * correction. held: the interfacial and overlap rates (G, K) of the last
 * spatial sample, then the normalised same-cell neighbour overlaps
 * (sum phi phi_east / V, sum phi phi_north / V). */
typedef struct pfh_moment_accum {
    double q[12];
    unsigned long long count;
    double held[6];
    unsigned long long spatial_count;
    unsigned long long last_sample;
    unsigned long long last_interval;
    double last_totals[4];
    double correction[4];
} pfh_moment_accum;

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
    pfh_moment_accum* moments; /* spatial observation: N accumulators */
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

/* One replica of a batched 2D step (no reference counterpart; the batched
 * ensemble generalises the persistent queue of src/kernels.cu:948-1042 to
 * tickets over every replica's cells): its periodic side L, S row pitch P,
 * the word offset of its aggregate inside each S slot arena, and its Philox
 * streams (seed, and the polarity stream of its tumbles). 32 bytes. */
typedef struct pfh2d_replica {
    unsigned long long S_offset;
    int32_t L;
    int32_t P;
    unsigned long long seed;
    unsigned long long polarity_seed;
} pfh2d_replica;

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
_Static_assert(sizeof(pfh2d_replica) == 32, "replica record size");
_Static_assert(sizeof(pfh_moment_accum) == 240, "MomentAccum is 240 B");
_Static_assert(offsetof(pfh_moment_accum, held) == 104, "MomentAccum.held moved");
#else
static_assert(sizeof(pfh_cell_state) == 192, "pfh_cell_state must be 192 B");
static_assert(alignof(pfh_cell_state) == 64, "pfh_cell_state alignment");
static_assert(offsetof(pfh_cell_state, advection) == 128, "advection moved");
static_assert(offsetof(pfh_cell_state, reserved) == 120, "reserved moved");
static_assert(offsetof(pfh_cell_state, M_pf) == 188, "M_pf moved");
static_assert(sizeof(pfh_traj_packed_cell) == 48, "traj record size");
static_assert(sizeof(pfh2d_replica) == 32, "replica record size");
static_assert(sizeof(pfh_moment_accum) == 240, "MomentAccum is 240 B");
static_assert(offsetof(pfh_moment_accum, held) == 104, "MomentAccum.held moved");
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
/* Replaces configure_morton_smem(int) (include/kernels.cuh:220). The sort's
 * LDS is static (sized for 16384 cells), so this only validates the size. */
int pfh2d_configure_morton(int smem_bytes);
/* Replaces k_step_grid(device) (include/kernels.cuh:221): the persistent
 * grid, one 1024-thread workgroup per compute unit. Returns the grid size,
 * or a negative status. */
int pfh2d_step_grid(int device);

/* Replaces launch_step (include/kernels.cuh:223-225, src/kernels.cu:1329-1363).
 * One persistent work-queue launch advances every cell by one step.
 * l2_base/l2_bytes/l2_hit_ratio are accepted and ignored: gfx950 has no
 * L2 access-policy window. observation: PFH_OBSERVE_DISABLED (the plain
 * step), PFH_OBSERVE_ADVECTION (each cell's advection accumulator records
 * the step's velocity components, k_step_advection) or PFH_OBSERVE_SPATIAL
 * (also the interfacial/overlap centroid rates into args->moments,
 * k_step_moments, include/fused_moments.cuh:23-127). Observation changes no
 * field, aggregate or cell state other than the accumulators. S_sc may be
 * NULL (no reference counterpart): the step then scatters nothing, for a
 * caller that needs only the new field and cell states (the surrogate's check
 * step). */
int pfh2d_launch_step(const pfh2d_step_args* args, int grid, void* stream,
                      const void* l2_base, size_t l2_bytes,
                      float l2_hit_ratio, int observation);
/* Replaces launch_velocity_moments (include/velocity_moments.cuh:7-8,
 * src/velocity_moments.cu:43-128): the independent dense observer. For each
 * cell, from the step's input field, aggregate and cell state (the step it
 * observes has not run yet): out[n].q[0..3] += physical_dt x (active x, y;
 * interaction x, y velocity), q[4..9] += physical_dt x the interfacial,
 * overlap and area centroid rates, q[10..11] += physical_dt x (advection
 * rate - active - interaction), ++count. fp64 sums. */
int pfh2d_velocity_moments(const pfh2d_step_args* args, pfh_moment_accum* out,
                           double physical_dt, void* stream);
/* pfh2d_launch_step gated on a device word: the launch does nothing unless
 * *run_if != 0 (run_if may be NULL: always runs). No reference counterpart:
 * lets a captured surrogate block carry steps whose execution depends on a
 * device-side decision (the check step runs only while the block is live). */
int pfh2d_launch_step_if(const pfh2d_step_args* args, const int* run_if, int grid,
                         void* stream);
/* pfh2d_launch_step_if for a surrogate check step (no reference
 * counterpart): the same new field in phi_out, and each cell's new window
 * position and class (gx0, gy0, cls, cls_written, cell_cls); no other
 * CellState field, flag or aggregate is written. S_sc must be NULL. */
int pfh2d_launch_check_step_if(const pfh2d_step_args* args, const int* run_if, int grid,
                               void* stream);

/* Batched step: one persistent launch advances every cell of many
 * independent replicas by one step (generalises launch_step,
 * src/kernels.cu:1329-1363, and the ticket queue of :948-1042 to an
 * ensemble). `a` describes the ensemble as one population: phi_in/phi_out
 * hold every replica's tiles, cell/cell_cls/perm every replica's cells (perm
 * is the queue order over all N = sum N_r cells; each cell's global_id is
 * its id within its replica), S_rd/S_sc/S_cl are the bases of the three S
 * slot arenas and clear_ahead_words is the arena size; L, P, seed and
 * polarity_seed of `a` are not read. replicas (device, num_replicas records)
 * and cell_replica (device, N entries: the replica of each cell) give each
 * cell's domain. max_side / max_pitch bound every replica's L and P; sides
 * from 256 are accepted: a replica narrower than 287 pixels cannot use the
 * 286-wide fallback class (a cell that would need it raises
 * CLASS_EXHAUSTED). Capture-safe; allocates nothing. */
int pfh2d_launch_batch_step(const pfh2d_step_args* a, const pfh2d_replica* replicas,
                            const uint16_t* cell_replica, int num_replicas, int max_side,
                            int max_pitch, int grid, void* stream);
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
/* Replaces k_zero_u32 (src/kernels.cu:1044-1048): n words set to 0, any
 * 4 B alignment. */
int pfh2d_zero_u32(uint32_t* p, size_t n, void* stream);
/* Replaces k_verify_cells (src/kernels.cu:1193-1232): per cell, out_V[n] =
 * sum of phi^2 over the active window of class cls[n] (the update's
 * reduction order, so it is the V the moments pass would store) and
 * out_outside_max[n] = max |phi| over the rest of the tile. phi 16 B
 * aligned. cell is not read (kept for the reference's signature). */
int pfh2d_verify_cells(const float* phi, const pfh_cell_state* cell,
                       const uint8_t* cls, int N, double* out_V,
                       float* out_outside_max, void* stream);
/* Replaces k_verify_S (src/kernels.cu:1234-1240): atomic max into *out_max
 * (caller zeroes it). */
int pfh2d_verify_S(const uint32_t* S, size_t n, uint32_t* out_max,
                   void* stream);
/* Replaces k_pack_traj (src/kernels.cu:1242-1265). out may be device memory
 * or mapped pinned host memory; cell and out 16 B aligned. */
int pfh2d_pack_traj(const pfh_cell_state* cell, const uint8_t* cls,
                    pfh_traj_packed_cell* out, int N, int L, void* stream);
/* Replaces k_morton_sort (src/kernels.cu:1267-1302): perm = the cells sorted
 * by (Morton code of the truncated centroid, index). M is a power of two
 * >= N, M <= 16384 (validated as in the reference; the sort itself does not
 * pad to M). */
int pfh2d_morton_sort(const pfh_cell_state* cell, uint32_t* perm, int N,
                      int M, int L, void* stream);

/* ---- 2D host-side set-up (no GPU work) ---- */
/* Philox4x32-10 (common/philox.cuh:32-59): out = the ten-round block of
 * counter ctr under key. The kernels' own implementation, for host callers. */
int pfh_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* philox_uniform53 (common/philox.cuh:61-67): ((a << 32 | b) >> 11) 2^-53. */
double pfh_uniform53(uint32_t a, uint32_t b);
/* domain_side_for (include/params.cuh:562-569): ceil(sqrt(N pi R^2 / rho)). */
int pfh2d_domain_side(int num_cells, double radius, double rho);
/* s_pitch_for (include/params.cuh:573-580): 32 ceil(side / 32). */
int pfh2d_s_pitch(int side);
/* ic_theta (include/kernels.cuh:153-158): initial polarity angle of a cell. */
float pfh2d_ic_theta(int32_t gid, uint64_t polarity_stream);
/* ic_v_A (include/kernels.cuh:161-172): lognormal per-cell speed, median v_A. */
double pfh2d_ic_v_A(int32_t gid, uint64_t seed, double v_A, double sigma);
/* The default grid + jitter placement of Sim::seed_positions
 * (src/sim.cu:108-129): binary32 centres of N cells in an L x L domain. */
int pfh2d_grid_jitter_centres(int N, int L, uint64_t seed, float* cx, float* cy);
/* palmieri_centre_table_fnv1a64 (include/palmieri_initializer.hpp:105-122):
 * FNV-1a over the little-endian words x0, y0, x1, y1, ... */
uint64_t pfh_centre_table_fnv1a64(const float* x, const float* y, int n);

/* ====================================================================== */
/* ---- 3D: include/pf3d/kernels.cuh (periodic XYZ, cubic bricks) ---- */

/* Sticky integrity flags: include/pf3d/kernels.cuh:55-91. CellState3D.flags
 * holds them as bits; the global array holds one count per flag. */
enum {
    PFH3D_FLAG_S_OVERFLOW = 0,
    PFH3D_FLAG_Q_CLAMP = 1,
    PFH3D_FLAG_S_NEGATIVE = 2,
    PFH3D_FLAG_NONFINITE = 3,
    PFH3D_FLAG_V_NONPOSITIVE = 4,
    PFH3D_FLAG_SUPPORT_EXHAUSTED = 5,
    PFH3D_FLAG_SUPPORT_EDGE = 6,
    PFH3D_FLAG_DESTINATION_CLIP = 7,
    PFH3D_FLAG_INVALID_GEOMETRY = 8,
    PFH3D_FLAG_SLAB_TOP_CONTACT = 9,
    PFH3D_FLAG_COUNT = 10
};
#define PFH3D_FATAL_MASK 0x3BFu  /* all but SUPPORT_EDGE (kernels.cuh:74-83) */
#define PFH3D_MAX_SHIFT 4        /* kMaximumShiftPerStep, src/pf3d/sim.cu:27 */
#define PFH3D_MAX_EDGE 64        /* bricks up to 64^3 (one x group per lane) */

/* Per-cell state: CellState3D, include/pf3d/kernels.cuh:98-137. 256 B. */
typedef struct pfh3d_cell_state {
    int64_t global_id;
    int64_t origin_x, origin_y, origin_z;   /* unwrapped brick origin */
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
    uint32_t storage_edge;                  /* must be 0: base brick */
    uint32_t pad2;
    double wall_overlap;
    double outside_slit_volume;
    uint32_t reserved[10];
} __attribute__((aligned(64))) pfh3d_cell_state;

/* One shard's moment reduction: MomentPartial3D (kernels.cuh:201-211). */
typedef struct pfh3d_moment_partial {
    double V, Cx, Cy, Cz, surface, Ix, Iy, Iz;
    int32_t lo_x, hi_x, lo_y, hi_y, lo_z, hi_z;
    uint32_t phi_max_bits, padding;
} __attribute__((aligned(16))) pfh3d_moment_partial;

/* Trajectory record: TrajPackedCell3D (kernels.cuh:318-328). */
typedef struct pfh3d_traj_cell {
    int64_t global_id;
    double x, y, z;
    double volume, surface, wall_overlap, outside_slit_volume;
    float velocity_x, velocity_y, velocity_z;
    float polarity_x, polarity_y, polarity_z;
    float gamma, v_A, phi_max;
} pfh3d_traj_cell;

/* Arguments of one 3D step: the union of MeasureArgs3D and UpdateArgs3D
 * (include/pf3d/kernels.cuh:232-316) for the periodic cube of side L with S
 * row pitch P (= 32 ceil(L/32), src/pf3d/sim.cu:276). Fields and aggregates
 * swap every begun step (src/pf3d/sim.cu:1632-1640). counters[0] is the
 * accepted-step count (the Philox step, d_step_), counters[1] the begun-step
 * count (the full-moment cadence, steps_done_). partials holds
 * N*shards pfh3d_moment_partial and arrivals N uint32 words that are zero
 * between launches (the library leaves them zero). */
typedef struct pfh3d_step_args {
    const float* phi_in;
    float* phi_out;
    const uint32_t* S_in;
    uint32_t* S_out;
    pfh3d_cell_state* cells;
    pfh3d_cell_state* accepted;   /* accepted-state snapshot (d_accepted_cells_) */
    uint32_t* flags;              /* PFH3D_FLAG_COUNT global counters */
    uint64_t* counters;           /* [0] accepted steps, [1] begun steps */
    pfh3d_moment_partial* partials;
    uint32_t* arrivals;
    int N, B, L, P;
    int shards;                   /* measurement CTAs per cell, 1..B */
    int max_shift;                /* 1..PFH3D_MAX_SHIFT */
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
} pfh3d_step_args;

#ifndef __cplusplus
_Static_assert(sizeof(pfh3d_cell_state) == 256, "CellState3D is 256 B");
_Static_assert(offsetof(pfh3d_cell_state, storage_edge) == 192, "storage_edge moved");
_Static_assert(sizeof(pfh3d_moment_partial) == 96, "MomentPartial3D is 96 B");
_Static_assert(sizeof(pfh3d_traj_cell) == 104, "TrajPackedCell3D is 104 B");
_Static_assert(sizeof(pfh3d_step_args) == 168, "step args layout");
#else
static_assert(sizeof(pfh3d_cell_state) == 256, "CellState3D is 256 B");
static_assert(offsetof(pfh3d_cell_state, storage_edge) == 192, "storage_edge moved");
static_assert(sizeof(pfh3d_moment_partial) == 96, "MomentPartial3D is 96 B");
static_assert(sizeof(pfh3d_traj_cell) == 104, "TrajPackedCell3D is 104 B");
static_assert(sizeof(pfh3d_step_args) == 168, "step args layout");
#endif

/* ---- the reference's 3D launch records (include/pf3d/kernels.cuh:153-338,
 * include/pf3d/rng.cuh:15-17, tile_partition.cuh:9-18, phase_storage.hpp:24)
 * with the same member types and order, hence the same layout. ---- */
typedef struct pfh3d_layout {          /* SLayout3D */
    int nx, ny, nz;
    int pitch_x;
    uint32_t boundary_flags;           /* 7: periodic XYZ (the only one here) */
} pfh3d_layout;
typedef struct pfh3d_wall_field {      /* WeightedWallField3D */
    const float* values;
    int planes;
} pfh3d_wall_field;
typedef struct pfh3d_storage {         /* CellFieldStorage3D */
    int z_cap;                         /* 0: cubic storage */
} pfh3d_storage;
typedef struct pfh3d_selection {       /* CellSelection3D: ids NULL = all */
    const int* ids;
    int count;
} pfh3d_selection;
typedef struct pfh3d_exchange_region { /* TileExchangeRegion3D */
    int y_lo[4];
    int y_hi[4];
    int y_count;
    int z_lo, z_hi;
} pfh3d_exchange_region;
typedef struct pfh3d_vec3 {            /* Vec3 */
    double x, y, z;
} pfh3d_vec3;
typedef struct pfh3d_init_args {       /* InitArgs3D */
    float* phi_even;
    float* phi_odd;
    pfh3d_cell_state* cells;
    const pfh3d_vec3* centres;
    int N;
    int B;
    pfh3d_layout layout;
    float lambda;
    float seed_radius;
    pfh3d_storage storage;
} pfh3d_init_args;
typedef struct pfh3d_scatter_args {    /* ScatterArgs3D */
    const float* phi;
    pfh3d_cell_state* cells;
    uint32_t* S;
    uint32_t* global_flags;
    pfh3d_layout layout;
    int N;
    int B;
    pfh3d_selection selection;
    pfh3d_storage storage;
} pfh3d_scatter_args;
typedef struct pfh3d_measure_args {    /* MeasureArgs3D */
    const float* phi;
    const uint32_t* S;
    pfh3d_wall_field wall;
    pfh3d_cell_state* cells;
    uint32_t* global_flags;
    const uint64_t* step;
    pfh3d_layout layout;
    int N;
    int B;
    pfh3d_moment_partial* partials;
    int shards;
    uint64_t polarity_stream;
    uint64_t aging_steps;
    double p_tumble;
    float motility_coeff;
    int max_shift;
    int support_margin;
    uint8_t apply_tumble;              /* C++ bool */
    uint8_t compute_surface;           /* C++ bool */
    pfh3d_selection selection;
    pfh3d_storage storage;
} pfh3d_measure_args;
typedef struct pfh3d_update_args {     /* UpdateArgs3D */
    float* phi_in;
    float* phi_out;
    const uint32_t* S_in;
    uint32_t* S_out;
    pfh3d_wall_field wall;
    pfh3d_cell_state* cells;
    uint32_t* global_flags;
    unsigned long long* work_cursor;
    pfh3d_layout layout;
    int N;
    int B;
    float dt;
    double V0;
    double volume_scale;
    float bulk_scale;
    float interaction_coeff;
    uint8_t compute_surface;           /* C++ bool */
    pfh3d_selection selection;
    int32_t tile_pass;                 /* UpdateTilePass3D: 0 = All */
    pfh3d_exchange_region exchange_region;
    pfh3d_storage storage;
} pfh3d_update_args;
typedef struct pfh3d_verify_cell {     /* VerifyCell3D */
    double measured_V;
    double state_V;
    float max_abs_phi;
    uint32_t nonfinite_count;
    uint32_t support_edge_count;
    uint32_t verified;
} pfh3d_verify_cell;

#ifndef __cplusplus
#define PFH_SA _Static_assert
#else
#define PFH_SA static_assert
#endif
PFH_SA(sizeof(pfh3d_layout) == 20, "SLayout3D");
PFH_SA(sizeof(pfh3d_init_args) == 72 && offsetof(pfh3d_init_args, layout) == 40 &&
       offsetof(pfh3d_init_args, storage) == 68, "InitArgs3D");
PFH_SA(sizeof(pfh3d_scatter_args) == 88 && offsetof(pfh3d_scatter_args, layout) == 32 &&
       offsetof(pfh3d_scatter_args, selection) == 64, "ScatterArgs3D");
PFH_SA(sizeof(pfh3d_measure_args) == 168 && offsetof(pfh3d_measure_args, layout) == 56 &&
       offsetof(pfh3d_measure_args, partials) == 88 &&
       offsetof(pfh3d_measure_args, polarity_stream) == 104 &&
       offsetof(pfh3d_measure_args, apply_tumble) == 140 &&
       offsetof(pfh3d_measure_args, selection) == 144, "MeasureArgs3D");
PFH_SA(offsetof(pfh3d_update_args, layout) == 72 && offsetof(pfh3d_update_args, V0) == 104 &&
       offsetof(pfh3d_update_args, compute_surface) == 128 &&
       offsetof(pfh3d_update_args, selection) == 136 &&
       offsetof(pfh3d_update_args, tile_pass) == 152 &&
       offsetof(pfh3d_update_args, exchange_region) == 156 &&
       sizeof(pfh3d_update_args) == 208, "UpdateArgs3D");
PFH_SA(sizeof(pfh3d_verify_cell) == 32, "VerifyCell3D");
#undef PFH_SA

/* Replaces launch_initialize_spheres(const InitArgs3D&) (kernels.cuh:417). */
int pfh3d_launch_initialize_spheres(const pfh3d_init_args* a, void* stream);
/* Replaces launch_scatter_current(const ScatterArgs3D&) (kernels.cuh:420-421). */
int pfh3d_launch_scatter_current(const pfh3d_scatter_args* a, void* stream);
/* Replaces launch_measure_cells(const MeasureArgs3D&) and, with apply_motion,
 * the following launch_apply_cell_motion (kernels.cuh:436-447). The
 * reference finalizes the shard partials in a second kernel; here the last
 * shard of each cell to finish does it, counting arrivals in `arrivals`
 * (N words, zero between launches; may be NULL when shards == 1). dt is the
 * step's binary32 dt (UpdateArgs3D::dt), read by the motion. */
int pfh3d_launch_measure(const pfh3d_measure_args* a, uint32_t* arrivals, int apply_motion,
                         float dt, void* stream);
/* Replaces launch_update_tiled_fast / _sharded_fast / periodic_slab_walk_fast
 * (const UpdateArgs3D&; kernels.cuh:506-533). S_out (zeroed by the caller)
 * is required; slabs = z slabs per cell (0: automatic). */
int pfh3d_launch_update(const pfh3d_update_args* a, int slabs, void* stream);
/* Replaces launch_verify_cells (kernels.cuh:478-481): per cell, sum of
 * phi^2 (fp64), max |phi|, non-finite voxels and voxels above 1e-5 on a
 * face of the cube, and the stored V. */
int pfh3d_verify_cells(const float* phi, const pfh3d_cell_state* cells, pfh3d_verify_cell* out,
                       int N, int B, const pfh3d_layout* layout, void* stream);
/* Replaces launch_verify_S (kernels.cuh:488-489): atomic max of the
 * aggregate's words into *out_max (caller zeroes it). */
int pfh3d_verify_S(const uint32_t* S, const pfh3d_layout* layout, uint32_t* out_max,
                   void* stream);

/* Parameter record of a PF3D schema-1 checkpoint: ParamsRecord3D
 * (common/checkpoint_format_3d.h:176-228), packed, 288 B, every offset as the
 * reference pins it. This build writes and reads the periodic-XYZ cube
 * (boundary_flags 7, Nx = Ny = Nz) with base bricks only. */
#pragma pack(push, 1)
typedef struct pfh3d_ckpt_params {
    int64_t Nx, Ny, Nz;
    double dx, dy, dz;
    uint32_t brick_edge;
    uint32_t boundary_flags;
    double dt, t_end, lambda, gamma_normal, gamma_soft, soft_fraction, kappa;
    double target_radius, mu, v_A, v_A_sigma, xi, tau, volume_fraction;
    uint64_t seed, polarity_seed;
    int64_t print_interval, full_moment_every, verify_every;
    uint32_t integrator_id;        /* 1: explicit Euler */
    uint32_t stencil_id;           /* 2: 27-point */
    uint32_t rng_id;               /* 1: Philox4x32-10 */
    uint32_t trajectory_interval;
    uint64_t initialization_hash;
    double aging_time;
    uint64_t promoted_measure_policy;
    uint64_t promoted_measure_auto_wave_ctas;
    int64_t channel_height, channel_padding;
    double wall_kappa, wall_width;
} pfh3d_ckpt_params;
#pragma pack(pop)

/* CRC-64/ECMA-182 (common/checkpoint_format_3d.h:78-115): polynomial
 * 0x42F0E1EBA9EA3693, MSB first, no reflection, init and final xor 0;
 * continues `crc` over `bytes` bytes. Host only. */
uint64_t pfh_crc64_ecma(uint64_t crc, const void* data, size_t bytes);
/* Replaces checkpoint_write_3d (src/pf3d/checkpoint.cu:1009-1235): writes a
 * PF3D schema-1 file (header, params, N x {record, B^3 cube}) through
 * <path>.tmp and a rename, with the file CRC. cells: N host records; phi:
 * the N device cubes of the current field. Host function; synchronises
 * `stream` per cube. */
int pfh3d_checkpoint_write(const char* path, int64_t step, double time,
                           const pfh3d_ckpt_params* params, const pfh3d_cell_state* cells,
                           int N, int B, const float* phi, uint64_t base_measure_shards,
                           void* stream);
/* Replaces checkpoint_probe_3d / checkpoint_load_3d
 * (src/pf3d/checkpoint.cu:819-1007): validates header, parameters, every
 * record and the file CRC; fills params, step, time, shards, N and B. With
 * cells and phi NULL it only probes; otherwise cells (N host records) and phi
 * (N device cubes) receive the state. Errors name the reason in
 * pfh_last_error. */
int pfh3d_checkpoint_read(const char* path, pfh3d_ckpt_params* params_out, int64_t* step,
                          double* time, uint64_t* base_measure_shards, int* N, int* B,
                          pfh3d_cell_state* cells, float* phi, void* stream);

/* Replaces launch_initialize_spheres (kernels.cuh:417;
 * src/pf3d/kernels.cu:1071-1122): origin = rint(centre - (B-1)/2) and
 * phi = 0.5 (1 - tanh(k (r - seed_radius))), k = sqrt(7.5)/lambda; resets the
 * per-cell dynamics fields. centres: 3*N doubles (x, y, z). phi_odd may be
 * NULL; otherwise it is zeroed. */
int pfh3d_init_spheres(float* phi_even, float* phi_odd, pfh3d_cell_state* cells,
                       const double* centres, int N, int B, int L,
                       float lambda, float seed_radius, void* stream);
/* Replaces launch_clear_S (kernels.cuh:418-419). */
int pfh3d_clear_S(uint32_t* S, int L, int P, void* stream);
/* Replaces launch_scatter_current (kernels.cuh:420-421;
 * src/pf3d/kernels.cu:1148-1186): S += q(phi) of every cell. */
int pfh3d_scatter_current(const float* phi, pfh3d_cell_state* cells,
                          uint32_t* S, uint32_t* flags, int N, int B, int L,
                          int P, void* stream);
/* Replaces launch_measure_cells / launch_measure_cells_only followed by
 * launch_apply_cell_motion (kernels.cuh:436-447; src/pf3d/kernels.cu:
 * 1591-1788): measures phi_in against S_in, commits V, centroid, support box,
 * phi_max, interaction integral, surface (compute_surface) and the planned
 * shift, then (apply_motion) the tumble and velocity. Reads counters[0] when
 * apply_tumble. Uses phi_in, S_in, cells, flags, counters, partials, arrivals. */
int pfh3d_measure_cells(const pfh3d_step_args* a, int apply_tumble,
                        int compute_surface, int apply_motion, void* stream);
/* Replaces launch_update_tiled_fast, launch_update_tiled_sharded_fast and
 * launch_update_periodic_slab_walk_fast (kernels.cuh:506-533): the
 * moment-free update phi_in -> phi_out with the pending recentring shift,
 * scattering q(phi_out) into S_out (which the caller zeroed). Every
 * destination voxel is written. slabs = z slabs per cell (0: automatic). */
int pfh3d_update(const pfh3d_step_args* a, int slabs, void* stream);
/* Replaces launch_repair_after_fatal, launch_finalize_origins,
 * launch_restore_cells_after_fatal and launch_advance_step
 * (kernels.cuh:470-481; src/pf3d/kernels.cu:2128-2263) as one launch. */
int pfh3d_finish_step(const pfh3d_step_args* a, void* stream);
/* Replaces Sim3D::begin_step + finish_step (src/pf3d/sim.cu:1460-1651) for
 * the periodic base-brick path: snapshot, measurement (surface every
 * full_moment_every begun steps), motion, S_out clear, update, repair /
 * finalize / restore / advance. Three launches, no host synchronisation, no
 * allocation: safe to capture into a HIP graph. */
int pfh3d_step(const pfh3d_step_args* a, int slabs, void* stream);
/* pfh3d_step that allocates nothing and takes no scratch from the caller:
 * no counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). a->partials and a->arrivals are ignored; the shard
 * partials and arrival words live in a library-owned array of 800 KiB per
 * device (held until the process ends), so N <= 8192 and
 * N * shards <= 8192, else PFH_ERR_INVALID. Every call on one device shares
 * that array: calls that overlap on one device (concurrent streams) race. */
int pfh3d_step_noalloc(const pfh3d_step_args* a, int slabs, void* stream);
/* Replaces launch_pack_trajectory (kernels.cuh:482-485). */
int pfh3d_pack_trajectory(const pfh3d_cell_state* cells, pfh3d_traj_cell* out,
                          int N, void* stream);
/* z slabs per cell that pfh3d_update uses when slabs == 0. */
int pfh3d_update_slabs(int N, int B, int device);
/* Replaces Sim3D::resize_cells (src/pf3d/adaptive_storage.cu:143-266) for a
 * whole population: every cell's cube of edge B_old in phi_old is embedded,
 * centred, in a cube of edge B_new in phi_new (offset (B_new - B_old) / 2 on
 * each axis, zeros around it), and its state moved with it as
 * translate_state does (origin - offset, Cx/Cy/Cz + offset V, support box +
 * offset, pending shifts 0, the support flags cleared); the population's
 * SUPPORT_EXHAUSTED and SUPPORT_EDGE counters are cleared. Values keep their
 * world positions, so S is unchanged. 8 <= B_old < B_new <= 64, multiples
 * of 8. */
int pfh3d_grow_bricks(const float* phi_old, float* phi_new, pfh3d_cell_state* cells,
                      uint32_t* flags, int N, int B_old, int B_new, void* stream);

/* ---- 3D host-side set-up (no GPU work) ---- */
/* next_brick_edge (include/pf3d/brick_classes.hpp:23-32) for domain side L:
 * current + 32 rounded up to 16, at most the largest multiple of 8 below L
 * and 64 (this build's largest brick); 0 when no larger brick fits. */
int pfh3d_next_brick_edge(int current, int L);
/* automatic_brick_edge (include/pf3d/params.cuh:122-152). */
int pfh3d_brick_edge(double radius, double lambda);
/* domain_side_for (include/pf3d/params.cuh:224-233). */
int pfh3d_domain_side(int num_cells, double radius, double rho);
/* calibrate_seed_radius (include/pf3d/reference.hpp:712-745): bisection of
 * the phi^2 volume of a centred sphere to 4 pi R^3 / 3 (relative tolerance
 * 1e-10, at most 80 halvings). Returns the iteration count, or a negative
 * status. */
int pfh3d_calibrate_seed_radius(double radius, double lambda, int B,
                                double* seed_radius);
/* palmieri3d_sequential_centres (include/pf3d/palmieri_initializer.hpp:72-123):
 * cell 0 at the box centre, then mt19937_64 proposals (x, y, z, high 53 bits)
 * accepted at periodic distance >= radius, stored as binary32. Returns the
 * number of candidates drawn, or a negative status. */
int pfh3d_sequential_centres(int N, double side, double radius, uint64_t seed,
                             float* x, float* y, float* z);
/* initial_polarity (include/pf3d/rng.cuh:41-52), in double. */
int pfh3d_initial_polarity(int64_t global_id, uint64_t stream, double out[3]);

/* ====================================================================== */
/* ---- learned surrogate of the 2D step ----
 * No counterpart in the reference: a BF16 CNN (4 -> 32 -> 32 -> 32 -> 1
 * channels, 3x3 convolutions, ReLU, residual output) that advances a cell's
 * 96x96 crop by K plain steps, with its training kernels and the
 * accept/fallback kernels. Activations are NHWC bfloat16 (uint16_t bit
 * patterns): the input has 8 channels (phi, S_other, cos theta, sin theta,
 * four zero channels), hidden layers 32. MFMA-layer weights are [32][9][cin]
 * (tap = 3 ky + kx), the output layer's [9][32] fp32. Crops are centred on
 * the rounded centroid of the cell's stored field (include/kernels.cuh:27-57
 * CellState, tile layout include/params.cuh:121-147). */
#define PFH_SUR_CROP 96

/* Network input X [N][96][96][8], exact phi crop phic [N][96][96] and the
 * crop origins origin [N][2] (world pixel of crop (0,0)). S_other =
 * (S - q(phi)) 2^-27 is exact. */
int pfh_sur_crop(const float* phi, const pfh_cell_state* cells, const uint8_t* cls,
                 const uint32_t* S, int N, int L, int P, uint16_t* X, float* phic,
                 int* origin, void* stream);
/* phi of each cell on the recorded crop windows (after the cells moved). */
int pfh_sur_recrop(const float* phi, const pfh_cell_state* cells, const uint8_t* cls,
                   const int* origin, int N, int L, float* out, void* stream);
/* Y = relu?(conv3x3(X, W) + bias) [* (mask > 0)] on MFMA; cin = 8 or 32,
 * bias and mask may be NULL. */
int pfh_sur_conv(int cin, const uint16_t* X, const uint16_t* W, const float* bias,
                 const uint16_t* mask, uint16_t* Y, int ncrops, int relu, void* stream);
/* Output layer: dphi = b4 + conv3x3(H, W4); out = phic + dphi (either output
 * may be NULL). */
int pfh_sur_out(const uint16_t* H, const float* W4, float b4, const float* phic, float* dphi,
                float* out, int ncrops, void* stream);
/* Whole-network inference in one kernel (activations stay in LDS): X as for
 * pfh_sur_conv (8 channels, the first 4 used), W1 [32][9][8], W2 and W3
 * [32][9][32] bf16, biases fp32, W4 [9][32] fp32 (applied as bf16 hi + lo).
 * pred = phic + dphi; dphi may be NULL. X, W2 and W3 16-byte aligned. Does
 * nothing when go is non-NULL and *go == 0. */
int pfh_sur_infer(const uint16_t* X, const uint16_t* W1, const float* b1, const uint16_t* W2,
                  const float* b2, const uint16_t* W3, const float* b3, const float* W4, float b4,
                  const float* phic, float* dphi, float* pred, int ncrops, const int* go,
                  void* stream);
/* pfh_sur_infer with the inputs built in the kernel from the field, the
 * aggregate and the cell states (the crops of pfh_sur_crop, same values):
 * writes the crop origins and pred = phi crop + dphi; dphi may be NULL. */
int pfh_sur_infer_cells(const float* phi, const pfh_cell_state* cells, const uint8_t* cls,
                        const uint32_t* S, int N, int L, int P, const uint16_t* W1,
                        const float* b1, const uint16_t* W2, const float* b2, const uint16_t* W3,
                        const float* b3, const float* W4, float b4, int* origin, float* dphi,
                        float* pred, const int* go, void* stream);
/* y = bfloat16(x) with the conversion pfh_sur_infer applies to hidden
 * activations (n even, y 4-byte aligned). */
int pfh_sur_round_bf16(const float* x, uint16_t* y, int n, void* stream);
/* Squared-error loss per crop (fp64) and g = scale (dphi - target). */
int pfh_sur_loss(const float* dphi, const float* target, float* g, double* crop_loss,
                 int ncrops, float scale, void* stream);
/* Output-layer backward: dH (bf16, masked by H > 0) and per-crop partials of
 * dW4 (288) then db4 (1). */
int pfh_sur_out_bwd(const float* g, const uint16_t* H, const float* W4, uint16_t* dH,
                    float* dW4_part, int ncrops, void* stream);
/* Per-crop partials of dW [32][9][cin] then db [32] of an MFMA layer. */
int pfh_sur_wgrad(int cin, const uint16_t* dY, const uint16_t* X, float* part, int ncrops,
                  void* stream);
/* out[e] = sum over crops n (ascending) of part[n][e]. */
int pfh_sur_reduce(const float* part, int ncrops, int width, float* out, void* stream);
/* Backward-data weights of a 32 -> 32 layer: Wt[ci][t][co] = W[co][8-t][ci]. */
int pfh_sur_flip(const uint16_t* W, uint16_t* Wt, void* stream);
/* Adam step on fp32 master weights; wb (may be NULL) receives bf16 copies.
 * c1, c2: bias corrections 1 - beta^t. */
int pfh_sur_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, int n, float lr,
                 float b1, float b2, float eps, float c1, float c2, void* stream);
/* Per-crop relative L2 error of a against b. */
int pfh_sur_rel_l2(const float* a, const float* b, float* out, int ncrops, void* stream);
/* Write the predicted crops of accepted cells into their tiles' active
 * windows (window pixels outside the crop become 0). */
int pfh_sur_writeback(float* phi, const pfh_cell_state* cells, const uint8_t* cls,
                      const int* origin, const float* pred, const int* accept, int N, int L,
                      void* stream);
/* Acceptance test: err[n] = relative L2 of pred against the check step's
 * field on the recorded crop windows; rejected[0] += cells with err > tol;
 * err_max (may be NULL) = running maximum of err as float bits. Does nothing
 * when go is non-NULL and *go == 0. */
int pfh_sur_check(const float* phi_chk, const pfh_cell_state* cells, const uint8_t* cls,
                  const int* origin, int N, int L, const float* pred, float tol, float* err,
                  uint32_t* rejected, uint32_t* err_max, const int* go, void* stream);
/* Per-cell acceptance: for each cell, err[n] = relative L2 of the committed
 * field (the prediction on the crop, zero elsewhere) against the check
 * step's field over the crop and the check window; accept[n] = 1 when
 * err <= tol, the error relative to the block's own increment (check field
 * minus phi_n on the crop) <= tol_inc, and acc[n] + err <= budget, in which
 * case acc[n] += err (the cell's accumulated committed error); *rejected
 * counts the others; err_max (may be NULL) = running maximum of err. */
int pfh_sur_check_cells(const float* phi_chk, const pfh_cell_state* cells_chk,
                        const uint8_t* cls_chk, const float* phi_now,
                        const pfh_cell_state* cells_now, const uint8_t* cls_now,
                        const int* origin, int N, int L, const float* pred, float tol,
                        float tol_inc, float budget, float* acc, float* err, int* accept,
                        uint32_t* rejected, uint32_t* err_max, void* stream);
/* Commit with per-cell fallback: accepted cells take the prediction in
 * their window, the others the check step's field with its window position
 * and class (cls_written[parity] and cls updated); q(phi) scattered into the
 * zeroed aggregate S and every cell's moments recomputed. */
int pfh_sur_commit_cells(float* phi, pfh_cell_state* cells, uint8_t* cls, const int* origin,
                         const float* pred, const int* accept, const float* phi_chk,
                         const pfh_cell_state* cells_chk, const uint8_t* cls_chk, int parity,
                         int N, int L, int P, uint32_t* S, uint32_t* flags, void* stream);
/* Accepted block: prediction into every cell's window, q(phi) scattered into
 * the zeroed aggregate S, moments as pfh2d_init_moments computes them.
 * Does nothing when skip_if is non-NULL and *skip_if != 0. */
int pfh_sur_commit(float* phi, pfh_cell_state* cells, const uint8_t* cls, const int* origin,
                   const float* pred, int N, int L, int P, uint32_t* S, uint32_t* flags,
                   const int* skip_if, void* stream);
/* a[0..words) = b[0..words) = 0 unless *skip_if != 0 (16-byte aligned). */
int pfh_sur_zero2(uint32_t* a, uint32_t* b, unsigned long long words, const int* skip_if,
                  void* stream);
/* End of a block: when *rejected == 0 the step counters advance by K and the
 * queue cursors restart; otherwise, if *go != 0, *go = 0 (every later gated
 * launch does nothing), *abort_step = the block's first step (stepc[pin])
 * and *rejected_blocks (may be NULL) += 1. */
int pfh_sur_advance(long long* stepc, unsigned long long* cursor, int K, int pin, int* go,
                    const int* rejected, long long* abort_step, uint32_t* rejected_blocks,
                    void* stream);
/* Start of a block (nothing when *go == 0): copies of the cell states and
 * classes for the check step, S_chk[0..words) zeroed (words may be 0 and
 * S_chk NULL), cursors zeroed, both step slots = *step_now, *rejected = 0. */
int pfh_sur_prep(const int* go, int* rejected, const pfh_cell_state* cells,
                 pfh_cell_state* cells_chk, const uint8_t* cls, uint8_t* cls_chk, int N,
                 uint32_t* S_chk, unsigned long long words, unsigned long long* cur_chk,
                 long long* step_chk, const long long* step_now, void* stream);
/* The polarity draws of K skipped plain steps, steps *step0 .. *step0+K-1
 * (src/kernels.cu:98-114): theta and tumble_ctr of every cell. step0 is a
 * device pointer (the accepted-step counter), so the launch can be captured
 * in a graph. Does nothing when skip_if is non-NULL and *skip_if != 0. */
int pfh_sur_tumble(pfh_cell_state* cells, int N, const long long* step0, int K,
                   unsigned long long polarity_seed, double p_tumble, const int* skip_if,
                   void* stream);
/* The end of a block in one launch (no reference counterpart): pfh_sur_tumble
 * with skip_if = rejected, taking the block's first step from step_first (the
 * check step's private counter, step_chk[0] of pfh_sur_prep), and
 * pfh_sur_advance. One launch fewer per block than the two calls; the same
 * effect, since the draws do not read stepc. */
int pfh_sur_finish(pfh_cell_state* cells, int N, const long long* step_first, int K,
                   unsigned long long polarity_seed, double p_tumble, long long* stepc,
                   unsigned long long* cursor, int pin, int* go, const int* rejected,
                   long long* abort_step, uint32_t* rejected_blocks, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PFH_H */
