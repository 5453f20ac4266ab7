/*
 * TEST INFRASTRUCTURE ONLY: CPU oracle of the 2D per-cell update (see
 * pfo2d.h for scope and citations). Built by oracle/Makefile into
 * oracle/_build/libpfo.so with -ffp-contract=off so that the compiler fuses
 * nothing the reference's host semantics do not fuse.
 */
#include "pfo2d.h"
#include "pfo_math.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define PFO_PI 3.14159265358979323846
#define PFO_NUM_CLASSES 6
#define PFO_CLASS_FALLBACK 5
#define PFO_CLASS_STAGED_MAX 3
#define PFO_CONTROL_LANES 16   /* MI355X order: threads 0-15 own no pixels */
#define PFO_PROMOTE_SLACK 8
#define PFO_DEMOTE_SLACK 20
#define PFO_DEMOTE_DWELL 200
#define PFO_MAX_SHIFT 4
#define PFO_SUPPORT_EPS 1e-5f
#define PFO_NT_MAX 1024
#define PFO_SRC_MAX (286 * 286)

typedef struct { int wx, wy, tx0, ty0; } pfo_shape;

/* Shape classes, include/params.cuh:138-147. */
static const pfo_shape kCls[PFO_NUM_CLASSES] = {
    {144, 144, 64, 64}, {176, 144, 32, 64}, {144, 176, 64, 32},
    {160, 160, 32, 32}, {224, 224, 32, 32}, {286, 286, 1, 1},
};

static inline int pfo_wrapi(int v, int L) {
    v %= L;
    return v < 0 ? v + L : v;
}
static inline int pfo_clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

int pfo_class_containing(int ex, int ey, int slack) {
    int best = -1, best_area = 0;
    for (int c = 0; c < PFO_NUM_CLASSES; ++c) {
        const int area = kCls[c].wx * kCls[c].wy;
        if (ex + slack <= kCls[c].wx && ey + slack <= kCls[c].wy &&
            (best < 0 || area < best_area)) {
            best = c;
            best_area = area;
        }
    }
    return best;
}

/* ---- reduction layouts ---- */
typedef struct { int order, nt, ws; } pfo_layout;

static pfo_layout pfo_layout_of(int order) {
    pfo_layout l;
    l.order = order;
    if (order == PFO_ORDER_REFERENCE) { l.nt = 768; l.ws = 32; }
    else { l.nt = 1024; l.ws = 64; }
    return l;
}

/* Thread owning source pixel (x,y) in the interaction and stencil passes.
 * MI355X staged classes: bands of rows, one column per pixel thread; the
 * pixel threads are threads 16..1023 (threads 0-15 fetch the next cell):
 * NB = 1008/WX bands of RB = ceil(WY/NB) rows, thread 16 + band*WX + x;
 * non-staged classes: flat pixel index mod 1024. Reference: strips of 16
 * rows with a 768 stride (src/kernels.cu:265) for classes 0-4; warp rows and
 * lane columns for the fallback (:721-724). */
static inline int pfo_thr_src(const pfo_layout* l, int CLS, int WX, int WY,
                              int x, int y) {
    if (l->order == PFO_ORDER_REFERENCE) {
        if (CLS <= 4) return ((y % 16) * WX + x) % 768;
        return (y % 24) * 32 + (x % 32);
    }
    if (CLS <= PFO_CLASS_STAGED_MAX) {
        const int nb = (1024 - PFO_CONTROL_LANES) / WX;
        const int rb = (WY + nb - 1) / nb;
        return PFO_CONTROL_LANES + (y / rb) * WX + x;
    }
    return (y * WX + x) % 1024;
}

/* Thread owning the moment terms of destination pixel (a,b) whose source is
 * (i,j); -1 when the pixel is not covered and the kernel does not visit it
 * (its contribution is an exact zero). */
static inline int pfo_thr_mom(const pfo_layout* l, int CLS, int WX, int WY,
                              int dwx, int a, int b, int i, int j) {
    if (l->order == PFO_ORDER_REFERENCE) return (b % 24) * 32 + (a % 32);
    if (CLS <= PFO_CLASS_STAGED_MAX) {
        if ((unsigned)i >= (unsigned)WX || (unsigned)j >= (unsigned)WY) return -1;
        return pfo_thr_src(l, CLS, WX, WY, i, j);
    }
    return (b * dwx + a) % 1024;
}

static inline int pfo_thr_per(const pfo_layout* l, int dwx, int a, int b) {
    if (l->order == PFO_ORDER_REFERENCE) return (b % 24) * 32 + (a % 32);
    return (b * dwx + a) % 1024;
}

/* Lane tree inside each wave, then waves ascending into 0.0. */
static double pfo_tree(const double* part, const pfo_layout* l) {
    double tmp[64];
    double total = 0.0;
    for (int w = 0; w < l->nt / l->ws; ++w) {
        for (int i = 0; i < l->ws; ++i) tmp[i] = part[w * l->ws + i];
        for (int d = l->ws / 2; d > 0; d >>= 1)
            for (int i = 0; i < d; ++i) tmp[i] += tmp[i + d];
        total += tmp[0];
    }
    return total;
}

/* Run-and-tumble draw of one cell-step (src/kernels.cu:98-112): counter
 * (step_lo, step_hi, gid, 0xA5A5A5A5), key = polarity stream; a tumble when
 * u53(v0, v1) < p, then theta = (float)(2 pi u53(v2, v3)). Returns 1 on a
 * tumble and writes the new angle; otherwise leaves *theta. */
int pfo2d_tumble(uint64_t step, int32_t gid, uint64_t stream, double p_tumble,
                 float* theta) {
    const pfo_philox4 r = pfo_philox4x32_10(
        (uint32_t)(step & 0xFFFFFFFFull), (uint32_t)(step >> 32),
        (uint32_t)gid, 0xA5A5A5A5u,
        (uint32_t)(stream & 0xFFFFFFFFull), (uint32_t)(stream >> 32));
    if (pfo_uniform53(r.v[0], r.v[1]) < p_tumble) {
        *theta = (float)(2.0 * PFO_PI * pfo_uniform53(r.v[2], r.v[3]));
        return 1;
    }
    return 0;
}

/* ---- velocity observers (include/fused_moments.cuh:23-127,
 * common/velocity_moments.h:42-100), in the kernels' order ---- */
#define PFO_OBS_SUMS 6     /* gx, gy, kx, ky, bx, by partial arrays */

/* One pixel of the spatial observer: weight 2 phi / V times the passive
 * interfacial and overlap sources in fp64, times the offset from the input
 * centroid; and the same-cell neighbour overlaps phi phi_east, phi phi_north. */
static inline void pfo_obs_pixel(double* po, int t, int x, int y, double v, double e,
                                 double w, double nn, double ss, double ne, double nw,
                                 double se, double sw, double other, double gam,
                                 double dwC, double repC, double g2iv, double cx,
                                 double cy) {
    const double axial = ((nn + ss) + e) + w;
    const double lap = (((((4.0 * axial + ne) + nw) + se) + sw) - 20.0 * v) / 6.0;
    const double weight = g2iv * v;
    const double well = ((dwC * v) * (1.0 - v)) * (1.0 - 2.0 * v);
    const double interfacial = weight * (gam * lap - well);
    const double dx = (double)x - cx, dy = (double)y - cy;
    po[0 * PFO_NT_MAX + t] += dx * interfacial;
    po[1 * PFO_NT_MAX + t] += dy * interfacial;
    const double overlap = (((-weight) * repC) * v) * other;
    po[2 * PFO_NT_MAX + t] += dx * overlap;
    po[3 * PFO_NT_MAX + t] += dy * overlap;
    po[4 * PFO_NT_MAX + t] += v * e;
    po[5 * PFO_NT_MAX + t] += v * nn;
}

static inline double pfo_linear_correction(double dt, uint64_t steps, double first,
                                           double last) {
    return steps ? dt * 0.5 * (double)(steps - 1ull) * (last - first) : 0.0;
}

/* The record after the velocity: record_advection (fused_moments.cuh:96-127)
 * and, on a spatial step, begin_spatial_sample (velocity_moments.h:59-77) and
 * the new held rates (fused_moments.cuh:55-88). sums = gx, gy, kx, ky, bx, by
 * over the cell. vxf, vyf: the stored binary32 velocity. */
static void pfo_obs_record(const pfo2d_step_args* A, int n, double ax, double ay,
                           double sIx, double sIy, double vxf, double vyf,
                           const double* sums) {
    pfo_cell* cs = &A->cell[n];
    const double ix = (double)A->mot_coeff * sIx;
    const double iy = (double)A->mot_coeff * sIy;
    const double dt = A->physical_dt;
    pfo_moment* m = A->observation == 2 ? &A->moments[n] : NULL;
    if (m) {
        const double tot[4] = {cs->adv_q[0] + cs->adv_q[2], cs->adv_q[1] + cs->adv_q[3],
                               cs->adv_q[4], cs->adv_q[5]};
        const uint64_t interval = cs->adv_count - m->last_sample;
        if (m->spatial_count) {
            for (int j = 0; j < 4; ++j) m->q[4 + j] += (dt * (double)interval) * m->held[j];
            for (int xy = 0; xy < 2; ++xy) {
                const double beta = m->held[4 + xy];
                m->q[10 + xy] += (beta - 1.0) * (tot[xy] - m->last_totals[xy]) +
                                 beta * (tot[2 + xy] - m->last_totals[2 + xy]);
            }
        }
        for (int j = 0; j < 4; ++j) m->last_totals[j] = tot[j];
        m->last_interval = interval;
        m->last_sample = cs->adv_count;
        ++m->spatial_count;
        m->held[4] = sums[4] / cs->V;
        m->held[5] = sums[5] / cs->V;
    }
    cs->adv_q[0] += dt * ax;
    cs->adv_q[1] += dt * ay;
    cs->adv_q[2] += dt * ix;
    cs->adv_q[3] += dt * iy;
    cs->adv_q[4] += dt * ((vxf - ax) - ix);
    cs->adv_q[5] += dt * ((vyf - ay) - iy);
    ++cs->adv_count;
    if (m) {
        const double mrel = (double)cs->M_pf / 0.5;
        for (int j = 0; j < 4; ++j) {
            double rate = sums[j];
            if (mrel != 1.0) rate *= mrel;
            if (m->spatial_count > 1)
                m->correction[j] +=
                    pfo_linear_correction(dt, m->last_interval, m->held[j], rate);
            m->held[j] = rate;
        }
    }
}

/* ---- binary32 and FP64 instantiations of the step ---- */
#define REAL float
#define SFX _f32
#define IS_F64 0
#include "pfo2d_step.inc"
#undef REAL
#undef SFX
#undef IS_F64

#define REAL double
#define SFX _f64
#define IS_F64 1
#include "pfo2d_step.inc"
#undef REAL
#undef SFX
#undef IS_F64

int pfo2d_step(const pfo2d_step_args* a, int order, int threads) {
    return step_all_f32(a, order, threads);
}

int pfo2d_step_f64(const pfo2d_step_args* a, int threads) {
    return step_all_f64(a, PFO_ORDER_MI355X, threads);
}

/* k_init_tiles, src/kernels.cu:1050-1094, with tanh from pfo_tanh_d. */
void pfo2d_init_tiles(float* phi_a, float* phi_b, pfo_cell* cell,
                      const uint8_t* clsv, int N, int L, const float* seed_cx,
                      const float* seed_cy, float radius_eff,
                      float kappa_iface) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int n = 0; n < N; ++n) {
        const int c = (int)clsv[n];
        const int wx = kCls[c].wx, wy = kCls[c].wy;
        const int tx0 = kCls[c].tx0, ty0 = kCls[c].ty0;
        float* ta = phi_a + (size_t)n * PFO_TILE_AREA;
        float* tb = phi_b + (size_t)n * PFO_TILE_AREA;
        memset(ta, 0, sizeof(float) * PFO_TILE_AREA);
        memset(tb, 0, sizeof(float) * PFO_TILE_AREA);
        const float cx = seed_cx[n], cy = seed_cy[n];
        const int gx0 = pfo_wrapi((int)lrintf(cx - 0.5f * (float)(wx - 1)), L);
        const int gy0 = pfo_wrapi((int)lrintf(cy - 0.5f * (float)(wy - 1)), L);
        for (int b = 0; b < wy; ++b)
            for (int a = 0; a < wx; ++a) {
                float dx = (float)(gx0 + a) - cx;
                float dy = (float)(gy0 + b) - cy;
                if (dx > 0.5f * (float)L) dx -= (float)L;
                if (dx < -0.5f * (float)L) dx += (float)L;
                if (dy > 0.5f * (float)L) dy -= (float)L;
                if (dy < -0.5f * (float)L) dy += (float)L;
                const float r = sqrtf(dx * dx + dy * dy);
                const float t = (float)pfo_tanh_d((double)(kappa_iface * (r - radius_eff)));
                ta[(size_t)(ty0 + b) * PFO_TILE_PITCH + tx0 + a] = 0.5f * (1.0f - t);
            }
        cell[n].gx0 = gx0;
        cell[n].gy0 = gy0;
        cell[n].cls = (uint8_t)c;
        cell[n].cls_written[0] = (uint8_t)c;
        cell[n].cls_written[1] = (uint8_t)c;
    }
}

void pfo2d_init_tiles_f64(double* phi_a, double* phi_b, pfo_cell* cell,
                          const uint8_t* clsv, int N, int L,
                          const float* seed_cx, const float* seed_cy,
                          double radius_eff, double kappa_iface) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int n = 0; n < N; ++n) {
        const int c = (int)clsv[n];
        const int wx = kCls[c].wx, wy = kCls[c].wy;
        const int tx0 = kCls[c].tx0, ty0 = kCls[c].ty0;
        double* ta = phi_a + (size_t)n * PFO_TILE_AREA;
        double* tb = phi_b + (size_t)n * PFO_TILE_AREA;
        memset(ta, 0, sizeof(double) * PFO_TILE_AREA);
        memset(tb, 0, sizeof(double) * PFO_TILE_AREA);
        const float cx = seed_cx[n], cy = seed_cy[n];
        const int gx0 = pfo_wrapi((int)lrintf(cx - 0.5f * (float)(wx - 1)), L);
        const int gy0 = pfo_wrapi((int)lrintf(cy - 0.5f * (float)(wy - 1)), L);
        for (int b = 0; b < wy; ++b)
            for (int a = 0; a < wx; ++a) {
                double dx = (double)(gx0 + a) - (double)cx;
                double dy = (double)(gy0 + b) - (double)cy;
                if (dx > 0.5 * L) dx -= L;
                if (dx < -0.5 * L) dx += L;
                if (dy > 0.5 * L) dy -= L;
                if (dy < -0.5 * L) dy += L;
                const double r = sqrt(dx * dx + dy * dy);
                ta[(size_t)(ty0 + b) * PFO_TILE_PITCH + tx0 + a] =
                    0.5 * (1.0 - pfo_tanh_d(kappa_iface * (r - radius_eff)));
            }
        cell[n].gx0 = gx0;
        cell[n].gy0 = gy0;
        cell[n].cls = (uint8_t)c;
        cell[n].cls_written[0] = (uint8_t)c;
        cell[n].cls_written[1] = (uint8_t)c;
    }
}

void pfo2d_init_moments(pfo_cell* cell, const uint8_t* cls, const float* phi,
                        int N, int order) {
    init_moments_f32(cell, cls, phi, N, order);
}
void pfo2d_init_moments_f64(pfo_cell* cell, const uint8_t* cls,
                            const double* phi, int N) {
    init_moments_f64(cell, cls, phi, N, PFO_ORDER_MI355X);
}
void pfo2d_scatter_all(const float* phi, const pfo_cell* cell,
                       const uint8_t* cls, uint32_t* S, int N, int L, int P,
                       uint32_t* flags) {
    scatter_all_f32(phi, cell, cls, S, N, L, P, flags);
}
void pfo2d_scatter_all_f64(const double* phi, const pfo_cell* cell,
                           const uint8_t* cls, uint32_t* S, int N, int L,
                           int P, uint32_t* flags) {
    scatter_all_f64(phi, cell, cls, S, N, L, P, flags);
}

/* k_verify_cells, src/kernels.cu:1193-1232: the window's sum of phi^2 (here
 * in the update's reduction order, i.e. the V the moments pass stores) and
 * max |phi| over the tile outside the window. */
void pfo2d_verify_cells(const float* phi, const pfo_cell* cell,
                        const uint8_t* clsv, int N, double* out_V,
                        float* out_outside_max) {
    pfo_cell* tmp = (pfo_cell*)calloc((size_t)(N > 0 ? N : 1), sizeof(pfo_cell));
    if (N > 0) memcpy(tmp, cell, sizeof(pfo_cell) * (size_t)N);
    init_moments_f32(tmp, clsv, phi, N, PFO_ORDER_MI355X);
    for (int n = 0; n < N; ++n) {
        const int c = (int)clsv[n];
        const int wx = kCls[c].wx, wy = kCls[c].wy;
        const int tx0 = kCls[c].tx0, ty0 = kCls[c].ty0;
        const float* t = phi + (size_t)n * PFO_TILE_AREA;
        float omax = 0.0f;
        for (int i = 0; i < PFO_TILE_AREA; ++i) {
            const int x = i % PFO_TILE_PITCH, y = i / PFO_TILE_PITCH;
            const int inside = x >= tx0 && x < tx0 + wx && y >= ty0 && y < ty0 + wy;
            if (!inside) omax = fmaxf(omax, fabsf(t[i]));
        }
        out_V[n] = tmp[n].V;
        out_outside_max[n] = omax;
    }
    free(tmp);
}

/* Morton queue order (k_morton_sort, src/kernels.cu:1267-1302): cells sorted
 * by the 64-bit key (Morton code of the truncated, wrapped centroid << 32 |
 * index); the Morton code interleaves x into the even and y into the odd
 * bits of 16-bit coordinates. */
static uint32_t pfo_part1by1(uint32_t v) {
    uint32_t r = 0u;
    for (int i = 0; i < 16; ++i) r |= ((v >> i) & 1u) << (2 * i);
    return r;
}
static int pfo_cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}
void pfo2d_morton_perm(const pfo_cell* cell, int N, int L, uint32_t* perm) {
    uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(N > 0 ? N : 1));
    for (int i = 0; i < N; ++i) {
        const double V = cell[i].V > 0.0 ? cell[i].V : 1.0;
        const uint32_t x = (uint32_t)pfo_wrapi(cell[i].gx0 + (int)(cell[i].Cx / V), L);
        const uint32_t y = (uint32_t)pfo_wrapi(cell[i].gy0 + (int)(cell[i].Cy / V), L);
        const uint32_t m = pfo_part1by1(x & 0xFFFFu) | (pfo_part1by1(y & 0xFFFFu) << 1);
        k[i] = ((uint64_t)m << 32) | (uint32_t)i;
    }
    qsort(k, (size_t)N, sizeof(uint64_t), pfo_cmp_u64);
    for (int i = 0; i < N; ++i) perm[i] = (uint32_t)(k[i] & 0xFFFFFFFFull);
    free(k);
}

uint32_t pfo2d_verify_S(const uint32_t* S, size_t n) {
    uint32_t m = 0u;
    for (size_t i = 0; i < n; ++i) m = S[i] > m ? S[i] : m;
    return m;
}

/* The independent dense velocity observer (launch_velocity_moments,
 * src/velocity_moments.cu:43-128) on one step's input: force, tumble-drawn
 * polarity, then the centroid rates of the four velocity sources, every sum
 * in pixel order with one fp64 accumulator per quantity. out[n] accumulates
 * pdt x the rates as the GPU's pfh2d_velocity_moments does. */
void pfo2d_velocity_moments(const pfo2d_step_args* A, pfo_moment* out, double pdt) {
    for (int n = 0; n < A->N; ++n) {
        const pfo_cell* c = &A->cell[n];
        const pfo_shape sc = kCls[c->cls];
        const float* tile = (const float*)A->phi_in + (size_t)n * PFO_TILE_AREA;
#define AT(xx, yy)                                                                  \
        (((unsigned)(xx) < (unsigned)sc.wx && (unsigned)(yy) < (unsigned)sc.wy)     \
             ? tile[(size_t)(sc.ty0 + (yy)) * PFO_TILE_PITCH + sc.tx0 + (xx)] : 0.0f)
        double f0 = 0.0, f1 = 0.0;
        for (int y = 0; y < sc.wy; ++y)
            for (int x = 0; x < sc.wx; ++x) {
                const float p = AT(x, y);
                const float gx = 0.5f * (AT(x + 1, y) - AT(x - 1, y));
                const float gy = 0.5f * (AT(x, y + 1) - AT(x, y - 1));
                const uint32_t all = A->S_rd[(size_t)pfo_wrapi(c->gy0 + y, A->L) * A->P +
                                             pfo_wrapi(c->gx0 + x, A->L)];
                const uint32_t self = pfo_q_of(p);
                const float so = all >= self ? (float)(all - self) * PFO_QINV : NAN;
                f0 += (double)(p * gx * so);
                f1 += (double)(p * gy * so);
            }
        float theta = c->theta;
        pfo2d_tumble(A->step, c->global_id, A->polarity_seed, A->p_tumble, &theta);
        float sn, co;
        pfo_sincosf_theta(theta, &sn, &co);
        const double vel[4] = {(double)c->v_A * (double)co, (double)c->v_A * (double)sn,
                               (double)A->mot_coeff * f0, (double)A->mot_coeff * f1};
        const double cx = c->Cx / c->V, cy = c->Cy / c->V;
        const double dw = (double)(A->bulk_scale * c->gamma);
        const double vol = (double)(float)(A->vol_scale * (A->A0 - c->V));
        const double mob = (double)(c->M_pf / 0.5f);
        const double vx = (double)(float)(vel[0] + vel[2]);
        const double vy = (double)(float)(vel[1] + vel[3]);
        double m[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int y = 0; y < sc.wy; ++y)
            for (int x = 0; x < sc.wx; ++x) {
                const double p = AT(x, y);
                const double e = AT(x + 1, y), w = AT(x - 1, y);
                const double nn = AT(x, y + 1), ss = AT(x, y - 1);
                const double ne = AT(x + 1, y + 1), nw = AT(x - 1, y + 1);
                const double se = AT(x + 1, y - 1), sw = AT(x - 1, y - 1);
                const double lap =
                    (((((4.0 * (((nn + ss) + e) + w) + ne) + nw) + se) + sw) - 20.0 * p) / 6.0;
                const uint32_t all = A->S_rd[(size_t)pfo_wrapi(c->gy0 + y, A->L) * A->P +
                                             pfo_wrapi(c->gx0 + x, A->L)];
                const uint32_t self = pfo_q_of((float)p);
                const double so = all >= self ? (double)((float)(all - self) * PFO_QINV) : NAN;
                double src[4] = {(double)c->gamma * lap - ((dw * p) * (1.0 - p)) * (1.0 - 2.0 * p),
                                 -((double)A->rep_coeff * p) * so, vol * p,
                                 -(vx * (0.5 * (e - w)) + vy * (0.5 * (nn - ss)))};
                if (mob != 1.0) { src[0] *= mob; src[1] *= mob; src[2] *= mob; }
                const double wxp = ((2.0 * ((double)x - cx)) * p) / c->V;
                const double wyp = ((2.0 * ((double)y - cy)) * p) / c->V;
                for (int j = 0; j < 4; ++j) {
                    m[2 * j] += wxp * src[j];
                    m[2 * j + 1] += wyp * src[j];
                }
            }
#undef AT
        pfo_moment* z = &out[n];
        for (int j = 0; j < 4; ++j) z->q[j] += pdt * vel[j];
        for (int j = 0; j < 6; ++j) z->q[4 + j] += pdt * m[j];
        z->q[10] += pdt * ((m[6] - vel[0]) - vel[2]);
        z->q[11] += pdt * ((m[7] - vel[1]) - vel[3]);
        ++z->count;
    }
}

/* k_pack_traj, src/kernels.cu:1242-1265. */
void pfo2d_pack_traj(const pfo_cell* cell, const uint8_t* clsv, pfo_traj* out,
                     int N, int L) {
    for (int n = 0; n < N; ++n) {
        const pfo_cell* cs = &cell[n];
        const double V = cs->V > 0.0 ? cs->V : 1.0;
        const double cx = (double)cs->gx0 + cs->Cx / V;
        const double cy = (double)cs->gy0 + cs->Cy / V;
        pfo_traj t;
        t.global_id = cs->global_id;
        t.cls = (int)clsv[n];
        t.cx = (float)(cx - floor(cx / (double)L) * (double)L);
        t.cy = (float)(cy - floor(cy / (double)L) * (double)L);
        t.vx = cs->vx;
        t.vy = cs->vy;
        t.theta = cs->theta;
        t.volume = (float)cs->V;
        t.perim = (float)cs->perim;
        t.gamma = cs->gamma;
        t.v_A = cs->v_A;
        t.phi_max = cs->phi_max;
        out[n] = t;
    }
}

/* ---- scalar entry points for tests ---- */
void pfo_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const pfo_philox4 r = pfo_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3],
                                            key[0], key[1]);
    memcpy(out, r.v, sizeof(r.v));
}
double pfo_u53(uint32_t a, uint32_t b) { return pfo_uniform53(a, b); }
uint32_t pfo_qof(float phi) { return pfo_q_of(phi); }

/* ic_theta, include/kernels.cuh:153-158: counter (gid, 2, 0, 0). */
float pfo_ic_theta(int32_t gid, uint64_t stream) {
    const pfo_philox4 r = pfo_philox4x32_10((uint32_t)gid, 2u, 0u, 0u,
                                            (uint32_t)(stream & 0xFFFFFFFFull),
                                            (uint32_t)(stream >> 32));
    return (float)(2.0 * PFO_PI * pfo_uniform53(r.v[0], r.v[1]));
}

/* ic_v_A, include/kernels.cuh:161-172 (host libm, lognormal median v_A). */
double pfo_ic_v_A(int32_t gid, uint64_t seed, double v_A, double sigma) {
    if (!(sigma > 0.0)) return v_A;
    const pfo_philox4 r = pfo_philox4x32_10((uint32_t)gid, 3u, 0u, 0u,
                                            (uint32_t)(seed & 0xFFFFFFFFull),
                                            (uint32_t)(seed >> 32));
    double u1 = pfo_uniform53(r.v[0], r.v[1]);
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = pfo_uniform53(r.v[2], r.v[3]);
    const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * PFO_PI * u2);
    return v_A * exp(sigma * z);
}

void pfo_sincosf(float theta, float* s, float* c) { pfo_sincosf_theta(theta, s, c); }
double pfo_tanh(double y) { return pfo_tanh_d(y); }

/* phase_field_rhs, include/kernels.cuh:121-138. */
float pfo_rhs(float lap, float phi, float So, float gx, float gy, float gam,
              float dwC, float volC, float repC, float vx, float vy,
              float mrel) {
    const float passive = gam * lap
        - dwC * (phi * (1.0f - phi) * (1.0f - 2.0f * phi))
        + volC * phi
        - repC * phi * So;
    const float adv = vx * gx + vy * gy;
    return (mrel == 1.0f) ? passive - adv : mrel * passive - adv;
}

/* 9-point Laplacian at p with row pitch `pitch` (include/params.cuh:75-92). */
float pfo_lap9(const float* p, int pitch) {
    const float nC = p[pitch], sC = p[-pitch], cE = p[1], cW = p[-1];
    const float nE = p[pitch + 1], nW = p[pitch - 1];
    const float sE = p[-pitch + 1], sW = p[-pitch - 1];
    return (4.0f * (nC + sC + cE + cW) + 1.0f * (nE + nW + sE + sW)
            + -20.0f * p[0]) * (float)(1.0 / 6.0);
}

int pfo_domain_side_for(int n, double radius, double rho) {
    const double a = (double)n * (PFO_PI * radius * radius) / rho;
    const double side = ceil(sqrt(a));
    return (isfinite(side) && side >= 1.0 && side <= 2147483647.0) ? (int)side : -1;
}

int pfo_s_pitch_for(int side) {
    if (side <= 0) return -1;
    const long long pitch = 32LL * (((long long)side + 31LL) / 32LL);
    return pitch <= 2147483647LL ? (int)pitch : -1;
}

/* Default grid+jitter placement, src/sim.cu:108-129. */
void pfo2d_grid_jitter_centres(int N, int L, uint64_t seed, float* cx,
                               float* cy) {
    const int nx = (int)ceil(sqrt((double)N));
    const double sp = (double)L / (double)nx;
    for (int i = 0; i < N; ++i) {
        const int gxi = i % nx, gyi = i / nx;
        const pfo_philox4 r = pfo_philox4x32_10((uint32_t)i, 1u, 0u, 0u,
                                                (uint32_t)(seed & 0xFFFFFFFFull),
                                                (uint32_t)(seed >> 32));
        const double jx = (pfo_uniform53(r.v[0], r.v[1]) - 0.5) * 0.10 * sp;
        const double jy = (pfo_uniform53(r.v[2], r.v[3]) - 0.5) * 0.10 * sp;
        double x = ((double)gxi + 0.5 + ((gyi & 1) ? 0.5 : 0.0)) * sp + jx;
        double y = ((double)gyi + 0.5) * sp + jy;
        x -= floor(x / (double)L) * (double)L;
        y -= floor(y / (double)L) * (double)L;
        cx[i] = (float)x;
        cy[i] = (float)y;
    }
}

void pfo2d_harness_centres(int N, int L, int g, double amp, uint64_t seed,
                           float* cx, float* cy) {
    const double sp = (double)L / (double)g;
    for (int i = 0; i < N; ++i) {
        const pfo_philox4 r = pfo_philox4x32_10((uint32_t)i, 0x4A495454u, 0u, 0u,
                                                (uint32_t)(seed & 0xFFFFFFFFull),
                                                (uint32_t)(seed >> 32));
        const double jx = (2.0 * pfo_uniform53(r.v[0], r.v[1]) - 1.0) * amp;
        const double jy = (2.0 * pfo_uniform53(r.v[2], r.v[3]) - 1.0) * amp;
        double x = ((double)(i % g) + 0.5) * sp + jx;
        double y = ((double)(i / g) + 0.5) * sp + jy;
        x -= floor(x / (double)L) * (double)L;
        y -= floor(y / (double)L) * (double)L;
        cx[i] = (float)x;
        cy[i] = (float)y;
    }
}
