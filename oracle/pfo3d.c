/*
 * TEST INFRASTRUCTURE ONLY: CPU oracle of the 3D per-cell step (see
 * pfo3d.h for the reference file:line of every part). It is the checker of
 * the gfx950 kernels and the CPU baseline of the benchmark; no product code
 * links it.
 *
 * Two operation orders are reproduced:
 *   PFO3D_ORDER_MI355X     this build's kernels bit for bit: measurement
 *                          shards are contiguous z ranges, one thread per
 *                          four-voxel x group of a plane, 64-lane wave trees;
 *                          the 27-point Laplacian is summed per plane
 *                          (centre, edges, corners) and then across planes.
 *   PFO3D_ORDER_REFERENCE  the reference's 32x16x8 tiles dealt round-robin
 *                          to shards, 256 threads in 32-lane warps, and its
 *                          z/y/x single-accumulator Laplacian.
 * Binary32 arithmetic follows source order without contraction.
 */
#include "pfo3d.h"
#include "pfo_math.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SUPPORT_EPS 1.0e-5f
#define PFO_PI_D 3.141592653589793238462643383279502884
#define QINV_D 7.450580596923828125e-9

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float ubits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static inline void raise_flag(pfo3d_cell* c, uint32_t* flags, int f) {
    __atomic_fetch_or(&c->flags, 1u << f, __ATOMIC_RELAXED);
    if (flags) __atomic_fetch_add(&flags[f], 1u, __ATOMIC_RELAXED);
}

static inline int global_fatal(const uint32_t* flags) {
    if (!flags) return 0;
    for (int f = 0; f < PFO3D_FLAG_COUNT; ++f)
        if (((PFO3D_FATAL_MASK >> f) & 1u) &&
            __atomic_load_n(&flags[f], __ATOMIC_RELAXED) != 0u)
            return 1;
    return 0;
}

static inline int fatal_present(const pfo3d_cell* c, const uint32_t* flags) {
    return (__atomic_load_n(&c->flags, __ATOMIC_RELAXED) & PFO3D_FATAL_MASK) != 0u ||
           global_fatal(flags);
}

/* src/pf3d/device_common.cuh:38-54 */
static inline int wrap_origin(int64_t o, int n) {
    int64_t v = o % (int64_t)n;
    if (v < 0) v += n;
    return (int)v;
}
static inline int wrap_off(int o, int off, int n) {
    int v = o + off;
    if (v >= n) v -= n;
    return v;
}
static inline size_t s_index(int L, int P, int x, int y, int z) {
    return ((size_t)z * (size_t)L + (size_t)y) * (size_t)P + (size_t)x;
}
static inline size_t lidx(int x, int y, int z, int B) {
    return ((size_t)z * (size_t)B + (size_t)y) * (size_t)B + (size_t)x;
}

typedef struct { double V, Cx, Cy, Cz, surf, Ix, Iy, Iz; } acc8;
typedef struct { int lo[3], hi[3]; uint32_t pmax; } box3;

static inline void box_empty(box3* b, int B) {
    for (int i = 0; i < 3; ++i) { b->lo[i] = B; b->hi[i] = -1; }
    b->pmax = 0u;
}
static inline void box_merge(box3* t, const box3* p) {
    for (int i = 0; i < 3; ++i) {
        if (p->lo[i] < t->lo[i]) t->lo[i] = p->lo[i];
        if (p->hi[i] > t->hi[i]) t->hi[i] = p->hi[i];
    }
    if (p->pmax > t->pmax) t->pmax = p->pmax;
}
static inline void acc_add(acc8* t, const acc8* p) {
    t->V += p->V; t->Cx += p->Cx; t->Cy += p->Cy; t->Cz += p->Cz;
    t->surf += p->surf; t->Ix += p->Ix; t->Iy += p->Iy; t->Iz += p->Iz;
}

/* Lane tree (shuffle-down by ws/2 .. 1) inside each wave of ws lanes, then
 * waves in ascending order into a zero accumulator. */
static acc8 tree_reduce(const acc8* part, int nt, int ws) {
    acc8 tmp[64];
    acc8 total;
    memset(&total, 0, sizeof total);
    for (int w = 0; w < nt / ws; ++w) {
        for (int i = 0; i < ws; ++i) tmp[i] = part[w * ws + i];
        for (int d = ws / 2; d > 0; d >>= 1)
            for (int i = 0; i < d; ++i) acc_add(&tmp[i], &tmp[i + d]);
        acc_add(&total, &tmp[0]);
    }
    return total;
}

/* ---- commit_measurement, src/pf3d/kernels.cu:427-496 ---- */
static inline int clamp_shift(int proposed, int limit, int lo, int hi, int B) {
    if (proposed > limit) proposed = limit;
    if (proposed < -limit) proposed = -limit;
    if (hi >= lo) {
        if (proposed > lo) proposed = lo;
        if (proposed < hi - (B - 1)) proposed = hi - (B - 1);
    }
    return proposed;
}

static void commit(const pfo3d_step_args* a, int n, const acc8* t,
                   const box3* b, int surface) {
    pfo3d_cell* s = &a->cells[n];
    const int B = a->B;
    s->V = t->V; s->Cx = t->Cx; s->Cy = t->Cy; s->Cz = t->Cz;
    if (surface) s->surface = t->surf;
    s->Ix = t->Ix; s->Iy = t->Iy; s->Iz = t->Iz;
    s->bb_lo_x = b->lo[0]; s->bb_hi_x = b->hi[0];
    s->bb_lo_y = b->lo[1]; s->bb_hi_y = b->hi[1];
    s->bb_lo_z = b->lo[2]; s->bb_hi_z = b->hi[2];
    s->phi_max = ubits(b->pmax);
    int sh[3] = {0, 0, 0};
    if (t->V > 0.0 && isfinite(t->V)) {
        const double middle = 0.5 * (double)(B - 1);
        const double c[3] = {t->Cx, t->Cy, t->Cz};
        for (int i = 0; i < 3; ++i) {
            sh[i] = (int)rint(c[i] / t->V - middle);
            sh[i] = clamp_shift(sh[i], a->max_shift, b->lo[i], b->hi[i], B);
        }
    } else {
        raise_flag(s, a->flags, PFO3D_FLAG_V_NONPOSITIVE);
    }
    s->pending_shift_x = sh[0];
    s->pending_shift_y = sh[1];
    s->pending_shift_z = sh[2];
    if (b->hi[0] >= b->lo[0]) {
        const int lm = a->support_margin / 2;
        const int um = a->support_margin - lm;
        int bad = 0;
        for (int i = 0; i < 3; ++i) {
            const int ext = b->hi[i] - b->lo[i] + 1;
            if (ext + a->support_margin > B || b->lo[i] - sh[i] < lm ||
                b->hi[i] - sh[i] >= B - um)
                bad = 1;
        }
        if (bad) raise_flag(s, a->flags, PFO3D_FLAG_SUPPORT_EXHAUSTED);
    }
}

/* ---- polarity sampling, include/pf3d/rng.cuh:19-139 ---- */
static void isotropic_unit(double axial_u, double azimuth_u, double out[3]) {
    const double z = 2.0 * axial_u - 1.0;
    const double rsq = 1.0 - z * z;
    const double radial = sqrt(rsq > 0.0 ? rsq : 0.0);
    const double az = 2.0 * PFO_PI_D * azimuth_u;
    double s, c;
    pfo_sincos_d(az, &s, &c);
    out[0] = radial * c;
    out[1] = radial * s;
    out[2] = z;
}

void pfo3d_initial_polarity(int64_t gid, uint64_t stream, double out[3]) {
    const uint64_t g = (uint64_t)gid;
    const pfo_philox4 r = pfo_philox4x32_10((uint32_t)g, (uint32_t)(g >> 32),
                                            0x504F4C49u, 0u, (uint32_t)stream,
                                            (uint32_t)(stream >> 32));
    isotropic_unit(pfo_uniform53(r.v[0], r.v[1]), pfo_uniform53(r.v[2], r.v[3]), out);
}

int pfo3d_tumble_sample(uint64_t step, int64_t gid, uint64_t stream, double p,
                        const double cur[3], double out[3]) {
    const uint64_t g = (uint64_t)gid;
    const uint32_t k0 = (uint32_t)stream, k1 = (uint32_t)(stream >> 32);
    const pfo_philox4 ev = pfo_philox4x32_10((uint32_t)step, (uint32_t)(step >> 32),
                                             (uint32_t)g,
                                             (uint32_t)(g >> 32) ^ 0x54554D45u, k0, k1);
    if (!(pfo_uniform53(ev.v[0], ev.v[1]) < p)) {
        out[0] = cur[0]; out[1] = cur[1]; out[2] = cur[2];
        return 0;
    }
    const pfo_philox4 ve = pfo_philox4x32_10((uint32_t)step, (uint32_t)(step >> 32),
                                             (uint32_t)g,
                                             (uint32_t)(g >> 32) ^ 0x54554D56u, k0, k1);
    isotropic_unit(pfo_uniform53(ev.v[2], ev.v[3]), pfo_uniform53(ve.v[0], ve.v[1]), out);
    return 1;
}

/* k_apply_cell_motion, src/pf3d/kernels.cu:1739-1788. */
static void motion_cell(const pfo3d_step_args* a, int n, int apply_tumble) {
    pfo3d_cell* s = &a->cells[n];
    if (fatal_present(s, a->flags)) return;
    const uint64_t step = a->counters ? a->counters[0] : 0u;
    const int active = step >= a->aging_steps;
    double pol[3] = {s->polarity_x, s->polarity_y, s->polarity_z};
    int tumbled = 0;
    if (apply_tumble && active) {
        double out[3];
        tumbled = pfo3d_tumble_sample(step - a->aging_steps, s->global_id,
                                      a->polarity_stream, a->p_tumble, pol, out);
        pol[0] = out[0]; pol[1] = out[1]; pol[2] = out[2];
    }
    const float px = (float)pol[0], py = (float)pol[1], pz = (float)pol[2];
    const float speed = active ? s->v_A : 0.0f;
    const float vx = speed * px + a->motility_coeff * (float)s->Ix;
    const float vy = speed * py + a->motility_coeff * (float)s->Iy;
    const float vz = speed * pz + a->motility_coeff * (float)s->Iz;
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz) || !isfinite(vx) ||
        !isfinite(vy) || !isfinite(vz)) {
        raise_flag(s, a->flags, PFO3D_FLAG_NONFINITE);
        return;
    }
    s->polarity_x = px; s->polarity_y = py; s->polarity_z = pz;
    s->velocity_x = vx; s->velocity_y = vy; s->velocity_z = vz;
    if (tumbled) ++s->tumble_ctr;
}

/* FP64 motion: the same decision, v = v_A p + mot I in double. */
static void motion_cell_f64(const pfo3d_step_args* a, int n, int apply_tumble,
                            double mot_d, double* vel) {
    pfo3d_cell* s = &a->cells[n];
    if (fatal_present(s, a->flags)) return;
    const uint64_t step = a->counters ? a->counters[0] : 0u;
    const int active = step >= a->aging_steps;
    double pol[3] = {s->polarity_x, s->polarity_y, s->polarity_z};
    int tumbled = 0;
    if (apply_tumble && active) {
        double out[3];
        tumbled = pfo3d_tumble_sample(step - a->aging_steps, s->global_id,
                                      a->polarity_stream, a->p_tumble, pol, out);
        pol[0] = out[0]; pol[1] = out[1]; pol[2] = out[2];
    }
    const float px = (float)pol[0], py = (float)pol[1], pz = (float)pol[2];
    const double speed = active ? (double)s->v_A : 0.0;
    vel[3 * n + 0] = speed * px + mot_d * s->Ix;
    vel[3 * n + 1] = speed * py + mot_d * s->Iy;
    vel[3 * n + 2] = speed * pz + mot_d * s->Iz;
    s->polarity_x = px; s->polarity_y = py; s->polarity_z = pz;
    s->velocity_x = (float)vel[3 * n + 0];
    s->velocity_y = (float)vel[3 * n + 1];
    s->velocity_z = (float)vel[3 * n + 2];
    if (tumbled) ++s->tumble_ctr;
}

/* 27-point Laplacian of a 3x3x3 neighbourhood n[(dz+1)*9+(dy+1)*3+(dx+1)].
 * Weights 14/3/1/-128 over 30 (include/pf3d/params.cuh:73-90). */
static inline float lap27_f32(const float* n, int order) {
    if (order == PFO3D_ORDER_REFERENCE) {
        /* src/pf3d/device_common.cuh:218-242: z, y, x into one accumulator */
        float w = -128.0f * n[13];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int d = (dx != 0) + (dy != 0) + (dz != 0);
                    if (d == 0) continue;
                    const float wt = d == 1 ? 14.0f : d == 2 ? 3.0f : 1.0f;
                    w += wt * n[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
                }
        return w * (float)(1.0 / 30.0);
    }
    /* Per plane: e = (W + E) + (N + S), c = (NW + NE) + (SW + SE);
     * off-centre planes weigh (14 A + 3 e) + c, the centre plane
     * (-128 A + 14 e) + 3 c; then (below + centre) + above. */
    float q[3];
    for (int p = 0; p < 3; ++p) {
        const float* m = n + 9 * p;
        const float A = m[4];
        const float e = (m[3] + m[5]) + (m[1] + m[7]);
        const float c = (m[0] + m[2]) + (m[6] + m[8]);
        q[p] = (p == 1) ? (-128.0f * A + 14.0f * e) + 3.0f * c
                        : (14.0f * A + 3.0f * e) + c;
    }
    return ((q[0] + q[1]) + q[2]) * (float)(1.0 / 30.0);
}

static inline double lap27_f64(const double* n) {
    double w = 0.0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int d = (dx != 0) + (dy != 0) + (dz != 0);
                const double wt = d == 0 ? -128.0 : d == 1 ? 14.0 : d == 2 ? 3.0 : 1.0;
                w += wt * n[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
            }
    return w / 30.0;
}

float pfo3d_lap27(const float* f27, int order) { return lap27_f32(f27, order); }

/* ---- binary32 and FP64 instantiations ---- */
#define REAL float
#define SFX _f32
#define IS_F64 0
#include "pfo3d_step.inc"
#undef REAL
#undef SFX
#undef IS_F64

#define REAL double
#define SFX _f64
#define IS_F64 1
#include "pfo3d_step.inc"
#undef REAL
#undef SFX
#undef IS_F64

/* ---- k_initialize_spheres, src/pf3d/kernels.cu:1071-1122 ---- */
void pfo3d_init_spheres(float* phi_even, float* phi_odd, pfo3d_cell* cells,
                        const double* centres, int N, int B, float lambda,
                        float seed_radius) {
    const size_t words = (size_t)B * B * B;
    const float k = 2.7386127875258306f / lambda;
#pragma omp parallel for schedule(dynamic, 1)
    for (int n = 0; n < N; ++n) {
        pfo3d_cell* s = &cells[n];
        const double cx = centres[3 * n], cy = centres[3 * n + 1], cz = centres[3 * n + 2];
        s->origin_x = llrint(cx - 0.5 * (double)(B - 1));
        s->origin_y = llrint(cy - 0.5 * (double)(B - 1));
        s->origin_z = llrint(cz - 0.5 * (double)(B - 1));
        s->shift_ctr = s->tumble_ctr = 0u;
        s->flags = 0u;
        s->pad0 = s->pad1 = s->pad2 = 0u;
        s->pending_shift_x = s->pending_shift_y = s->pending_shift_z = 0;
        s->V = s->Cx = s->Cy = s->Cz = 0.0;
        s->surface = s->Ix = s->Iy = s->Iz = 0.0;
        s->wall_overlap = s->outside_slit_volume = 0.0;
        s->phi_max = 0.0f;
        memset(s->reserved, 0, sizeof s->reserved);
        float* even = phi_even + (size_t)n * words;
        float* odd = phi_odd ? phi_odd + (size_t)n * words : NULL;
        for (size_t q = 0; q < words; ++q) {
            const int x = (int)(q % (size_t)B);
            const int y = (int)((q / (size_t)B) % (size_t)B);
            const int z = (int)(q / ((size_t)B * B));
            const double dx = (double)(s->origin_x + x) - cx;
            const double dy = (double)(s->origin_y + y) - cy;
            const double dz = (double)(s->origin_z + z) - cz;
            const float r = (float)sqrt(dx * dx + dy * dy + dz * dz);
            const float t = (float)pfo_tanh_d((double)(k * (r - seed_radius)));
            even[q] = 0.5f * (1.0f - t);
            if (odd) odd[q] = 0.0f;
        }
    }
}

void pfo3d_scatter(const float* phi, pfo3d_cell* cells, uint32_t* S,
                   uint32_t* flags, int N, int B, int L, int P) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int n = 0; n < N; ++n) scatter_cell_f32(phi, cells, S, flags, n, B, L, P);
}

void pfo3d_scatter_f64(const double* phi, pfo3d_cell* cells, uint32_t* S,
                       uint32_t* flags, int N, int B, int L, int P) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int n = 0; n < N; ++n) scatter_cell_f64(phi, cells, S, flags, n, B, L, P);
}

/* measure (+commit) of every cell, then the motion launch */
void pfo3d_measure(const pfo3d_step_args* a, int order, int apply_tumble,
                   int compute_surface, int apply_motion, int threads) {
    const int nt = threads > 0 ? threads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int n = 0; n < a->N; ++n) measure_commit_f32(a, a->phi_in, n, order, compute_surface);
    if (apply_motion)
        for (int n = 0; n < a->N; ++n) motion_cell(a, n, apply_tumble);
}

void pfo3d_measure_f64(const pfo3d_step_args* a, const double* phi, double* vel,
                       const double* coef, int apply_tumble, int compute_surface,
                       int apply_motion, int threads) {
    const int nt = threads > 0 ? threads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int n = 0; n < a->N; ++n)
        measure_commit_f64(a, phi, n, PFO3D_ORDER_REFERENCE, compute_surface);
    if (apply_motion)
        for (int n = 0; n < a->N; ++n)
            motion_cell_f64(a, n, apply_tumble, coef[3], vel);
}

/* Repair, finalize origins, restore and advance: src/pf3d/kernels.cu:
 * 2128-2263. The repair rebuilds S_out as a copy of S_in: after the repair
 * every field equals its step input, and S_in is exactly the aggregate of
 * those fields, so the rebuilt aggregate is the same bit pattern. */
static void finish_step(const pfo3d_step_args* a, int fatal, size_t elem,
                        const void* phi_in, void* phi_out) {
    const size_t words = (size_t)a->B * a->B * a->B;
    if (fatal) {
        memcpy(phi_out, phi_in, elem * words * (size_t)a->N);
        memcpy(a->S_out, a->S_in, sizeof(uint32_t) * (size_t)a->L * a->L * a->P);
        for (int n = 0; n < a->N; ++n) {
            a->cells[n].pending_shift_x = 0;
            a->cells[n].pending_shift_y = 0;
            a->cells[n].pending_shift_z = 0;
        }
    }
    for (int n = 0; n < a->N; ++n) {
        pfo3d_cell* s = &a->cells[n];
        if ((s->flags & PFO3D_FATAL_MASK) == 0u) {
            const int sx = s->pending_shift_x, sy = s->pending_shift_y,
                      sz = s->pending_shift_z;
            s->origin_x += sx;
            s->origin_y += sy;
            s->origin_z += sz;
            if (sx != 0 || sy != 0 || sz != 0) ++s->shift_ctr;
        }
        s->pending_shift_x = s->pending_shift_y = s->pending_shift_z = 0;
    }
    if (fatal) memcpy(a->cells, a->accepted, sizeof(pfo3d_cell) * (size_t)a->N);
    if (a->counters) {
        if (!fatal) a->counters[0] += 1u;
        a->counters[1] += 1u;
    }
}

int pfo3d_step(const pfo3d_step_args* a, int order, int threads) {
    if (!a || a->N <= 0 || a->B <= 0 || a->B % 8 != 0 || a->L <= a->B ||
        a->P < a->L || a->shards < 1 || a->shards > a->B)
        return -1;
    const int nt = threads > 0 ? threads : 1;
    memcpy(a->accepted, a->cells, sizeof(pfo3d_cell) * (size_t)a->N);
    const int surface = a->full_moment_every > 0 &&
                        a->counters[1] % (uint64_t)a->full_moment_every == 0u;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int n = 0; n < a->N; ++n) measure_commit_f32(a, a->phi_in, n, order, surface);
    for (int n = 0; n < a->N; ++n) motion_cell(a, n, 1);
    memset(a->S_out, 0, sizeof(uint32_t) * (size_t)a->L * a->L * a->P);
    const float coef[4] = {a->dt, a->bulk_scale, a->interaction_coeff, 0.0f};
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int n = 0; n < a->N; ++n)
        update_cell_f32(a, a->phi_in, a->phi_out, coef, NULL, n, order);
    finish_step(a, global_fatal(a->flags), sizeof(float), a->phi_in, a->phi_out);
    return 0;
}

int pfo3d_step_f64(const pfo3d_step_args* a, const double* phi_in,
                   double* phi_out, double* vel, const double* coef, int threads) {
    if (!a || a->N <= 0 || a->B <= 0 || a->L <= a->B) return -1;
    const int nt = threads > 0 ? threads : 1;
    memcpy(a->accepted, a->cells, sizeof(pfo3d_cell) * (size_t)a->N);
    const int surface = a->full_moment_every > 0 &&
                        a->counters[1] % (uint64_t)a->full_moment_every == 0u;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int n = 0; n < a->N; ++n)
        measure_commit_f64(a, phi_in, n, PFO3D_ORDER_REFERENCE, surface);
    for (int n = 0; n < a->N; ++n) motion_cell_f64(a, n, 1, coef[3], vel);
    memset(a->S_out, 0, sizeof(uint32_t) * (size_t)a->L * a->L * a->P);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int n = 0; n < a->N; ++n)
        update_cell_f64(a, phi_in, phi_out, coef, vel, n, PFO3D_ORDER_REFERENCE);
    finish_step(a, global_fatal(a->flags), sizeof(double), phi_in, phi_out);
    return 0;
}

/* ---- include/pf3d/reference.hpp:655-745 (bisection on the phi^2 volume) */
static double seeded_volume(int edge, double seed_radius, double lambda) {
    const double centre = 0.5 * (double)(edge - 1);
    const double k = sqrt(7.5) / lambda;
    double volume = 0.0;
    for (int z = 0; z < edge; ++z) {
        const double dz = (double)z - centre;
        for (int y = 0; y < edge; ++y) {
            const double dy = (double)y - centre;
            for (int x = 0; x < edge; ++x) {
                const double dx = (double)x - centre;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                const double phi = 0.5 * (1.0 - tanh(k * (r - seed_radius)));
                volume += phi * phi;
            }
        }
    }
    return volume;
}

double pfo3d_calibrate_seed_radius(double radius, double lambda, int B,
                                   double tol, int max_iter, int* iters) {
    const double target = 4.0 * PFO_PI_D * radius * radius * radius / 3.0;
    double lo = 0.0, hi = (double)B, mid = 0.0;
    if (seeded_volume(B, lo, lambda) > target || seeded_volume(B, hi, lambda) < target)
        return -1.0;
    int it;
    for (it = 1; it <= max_iter; ++it) {
        mid = 0.5 * (lo + hi);
        const double v = seeded_volume(B, mid, lambda);
        if (fabs(v - target) / target <= tol) break;
        if (v < target) lo = mid; else hi = mid;
    }
    if (iters) *iters = it > max_iter ? max_iter : it;
    return mid;
}

/* automatic_brick_edge, include/pf3d/params.cuh:122-150 */
int pfo3d_brick_edge(double radius, double lambda) {
    const double k = 2.7386127875258306 / lambda;
    const double eff = radius + 0.5 / k;
    const double tail = log(1.0 / 1.0e-5 - 1.0) / (2.0 * k);
    const double req = 2.0 * (eff + tail) + 8.0;
    const int c = (int)ceil(req);
    return ((c + 7) / 8) * 8;
}
