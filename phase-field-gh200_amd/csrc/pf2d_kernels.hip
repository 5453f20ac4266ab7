// gfx950 kernels of the 2D per-cell phase-field update and their C ABI.
//
// The step is one persistent work-queue launch: one 1024-thread workgroup per
// compute unit takes cell tickets from a device cursor (reference step_body,
// src/kernels.cu:948-1024) and advances one cell at a time through:
//   A  stage the cell's active window of phi in LDS (16 B vector loads) and
//      hold its S window in VGPRs; a column-walk 3x3 register stencil gives
//      the interaction integral terms and, for every pixel, the passive
//      right-hand side and both gradients, which stay in VGPRs;
//   -  fixed-order fp64 reduction (64-lane tree, 16 wavefronts ascending) of
//      the interaction integral and the cell velocity;
//   B  the explicit Euler update from the held terms plus advection, the
//      shifted store, the Q5.27 scatter into the next S with integer atomics,
//      and the moments/support/phi_max of the stored field;
//   -  a second fixed-order reduction and the CellState commit.
// Classes 0-3 take the LDS path; classes 4 (224x224, larger than gfx950's
// 160 KB LDS once fp32 phi is staged) and 5 (fallback) read the tile from
// global memory.
//
// Arithmetic follows the reference's host semantics exactly: binary32
// operations in source order with no multiply-add contraction (the build
// passes -ffp-contract=off), fp64 reductions, Q5.27 integer scatter.
// Reference: process_cell src/kernels.cu:53-592, process_cell_fallback
// :604-937, init/verify/pack/morton :1044-1302.

#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include "../../include/pfh.h"
#include "pfh_device.h"
#include "pfh_host.h"

namespace pfh {

// Control block written by thread 0 and read by the workgroup.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct Ctl {
    int sx, sy, gx0, gy0, gx0n, gy0n, dcls, zero_tile, fm, pctr, tumbled;
    int dwx, dwy, dtx0, dty0, no_margin;
    float gam, dwC, volC, vA, cosp, sinp, theta, mrel, vx, vy;
    // spatial velocity observer: 2 / V and the centroid of the step's input
    // (include/fused_moments.cuh:16-21)
    double og2iv, ocx, ocy;
};

// The cell a workgroup processes next: filled by the control wavefront
// during the previous cell, double-buffered.
struct Slot {
    int n, cls;
    Ctl ctl;
    // the cell's domain in a batched step: side, S pitch and the word
    // offset of its replica's S in each slot arena
    int L, P, s_off_lo, s_off_hi;
};

struct Smem {
    pfh2d_step_args args;        // LDS copy for the out-of-line class paths
    double red[kWaves * 8];
    int ired[kWaves * 8];
    Slot slot[2];
    uint32_t csraw[2][48];       // CellState (192 B) of the next cell
    const pfh2d_replica* reps;   // batched step only: replica table
    const uint16_t* cell_rep;    // batched step only: replica of each cell
    int fm;                      // full-moment step (one in full_moment_every)
};

// The staged window, an LDS object of its own: the global-to-LDS window
// copies write only here, so the compiler does not hold other LDS reads
// (slots, partials) until those copies land.
static __shared__ __attribute__((aligned(16))) float s_window[kPhiLdsFloats];

// Control lanes: the first 16 lanes of the workgroup (kControlLanes). No
// LDS-path class gives them pixels (7x144, 5x176, 7x144 and 6x160 pixel
// threads are at most 1008), so they fetch the next cell while the other
// lanes compute. They sit in wave 0, which the SIMD issues first and which
// therefore finishes its pixels first: the fetch's serial latency (the next
// cell's scalars) then stays off the barrier's critical path, where it lay
// in the last wave.
constexpr int kLeaderLane = 0;                        // lane 0 of wave 0
constexpr int kScalarsThread = kWave;                 // lane 0 of wave 1

// threadIdx.x made opaque inside the persistent loop: without this, LICM
// hoists every per-thread index and address of the cell body out of the loop
// and keeps them all live (and spilled) across cells.
__device__ __forceinline__ int opaque_tid() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

__device__ __forceinline__ void flag_add(uint32_t* flags, int i, uint32_t v) {
    atomicAdd(&flags[i], v);
}
__device__ __forceinline__ void flag_or(uint32_t* flags, int i) {
    atomicOr(&flags[i], 1u);
}

// Barriers on the staged path order LDS only (lds_barrier): every value one
// thread passes to another there goes through LDS, and each thread's global
// stores and scatters are read back, if at all, by the same thread.

// Global (address space 1) views of the hot buffers: pointers read out of the
// argument struct are otherwise generic, and flat instructions would tie every
// LDS wait to outstanding memory operations.
#define PFH_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ PFH_GLOBAL T* gptr(T* p) {
    return (PFH_GLOBAL T*)p;
}
__device__ __forceinline__ uint32_t atomic_add_g(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(gptr(p), v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}

// The step arguments cell_scalars reads, passed by value so that the kernel
// argument block never has its address taken (which copies it to scratch).
struct ScalarArgs {
    const pfh_cell_state* cell;
    uint32_t* flags;
    unsigned long long polarity_seed;
    double p_tumble, vol_scale, A0;
    float bulk_scale;
    int L, parity_out;
    int fm;              // this launch's step is a full-moment step
};

__device__ __forceinline__ ScalarArgs scalar_args(const pfh2d_step_args& A, int fm) {
    ScalarArgs s;
    s.cell = A.cell;
    s.flags = A.flags;
    s.polarity_seed = A.polarity_seed;
    s.p_tumble = A.p_tumble;
    s.vol_scale = A.vol_scale;
    s.A0 = A.A0;
    s.bulk_scale = A.bulk_scale;
    s.L = A.L;
    s.parity_out = A.parity_out;
    s.fm = fm;
    return s;
}

// Per-cell scalars: tumble, polarity, coefficients, class hysteresis and the
// recentring shift (src/kernels.cu:93-191 and :624-711), from a copy of the
// cell's state. Out of line: its double-precision polynomial constants must
// not be hoisted out of the persistent loop into the update's registers.
__device__ __noinline__ void cell_scalars(const ScalarArgs A, const pfh_cell_state* csp,
                                         int CLS, unsigned long long step, Ctl* ctl) {
    const int WX = shape_of(CLS).wx, WY = shape_of(CLS).wy;
    const pfh_cell_state& cs = *csp;
    const Philox4 r = philox4x32_10(
        (uint32_t)(step & 0xFFFFFFFFull), (uint32_t)(step >> 32),
        (uint32_t)cs.global_id, 0xA5A5A5A5u,
        (uint32_t)(A.polarity_seed & 0xFFFFFFFFull),
        (uint32_t)(A.polarity_seed >> 32));
    float theta = cs.theta;
    int tumbled = 0;
    if (uniform53(r.v[0], r.v[1]) < A.p_tumble) {
        theta = (float)(2.0 * kPi * uniform53(r.v[2], r.v[3]));
        tumbled = 1;
    }
    double sd, cd;
    sincos_d((double)theta, &sd, &cd);
    const float gam = cs.gamma;
    const float dwC = A.bulk_scale * gam;
    const float volC = (float)(A.vol_scale * (A.A0 - cs.V));
    const float mrel = cs.M_pf * (float)(1.0 / 0.5);
    // the centroid (also the observers' reference point)
    const double cxr = cs.Cx / cs.V;
    const double cyr = cs.Cy / cs.V;
    ctl->og2iv = 2.0 / cs.V;
    ctl->ocx = cxr;
    ctl->ocy = cyr;

    const int ex = cs.bb_hi_x - cs.bb_lo_x + 1;
    const int ey = cs.bb_hi_y - cs.bb_lo_y + 1;
    int dcls = CLS, no_margin = 0;
    unsigned pctr = cs.promote_ctr;
    if (ex > 0) {
        if (ex + kPromoteSlack > WX || ey + kPromoteSlack > WY) {
            int grow = class_containing(ex, ey, kPromoteSlack);
            // a window must lie strictly inside the periodic domain: classes
            // as wide as the side are unavailable (the batched step's
            // 256-pixel replicas run without the 286-pixel fallback class)
            if (grow >= 0 && max(shape_of(grow).wx, shape_of(grow).wy) >= A.L) grow = -1;
            if (grow >= 0) {
                dcls = grow;
            } else if (CLS == kClassFallback && ex <= WX && ey <= WY) {
                no_margin = 1;
            } else {
                flag_add(A.flags, PFH_FLAG_CLASS_EXHAUSTED, 1u);
            }
            pctr = 0u;
        } else {
            const int small = class_containing(ex, ey, kDemoteSlack);
            if (small >= 0 && small != CLS &&
                shape_of(small).wx * shape_of(small).wy < WX * WY) {
                if (++pctr >= (unsigned)kDemoteDwell) { dcls = small; pctr = 0u; }
            } else {
                pctr = 0u;
            }
        }
    }
    const Shape dc = shape_of(dcls);
    int sx = 0, sy = 0;
    if (cs.V > 0.0) {
        sx = __double2int_rn(cxr - 0.5 * (double)(dc.wx - 1));
        sy = __double2int_rn(cyr - 0.5 * (double)(dc.wy - 1));
        if (dcls == CLS) {
            sx = max(-kMaxShiftPerStep, min(kMaxShiftPerStep, sx));
            sy = max(-kMaxShiftPerStep, min(kMaxShiftPerStep, sy));
        } else {
            sx = max(-WX, min(WX, sx));
            sy = max(-WY, min(WY, sy));
        }
    } else {
        flag_or(A.flags, PFH_FLAG_V_NONPOS);
    }
    ctl->sx = sx;
    ctl->sy = sy;
    ctl->gx0 = cs.gx0;
    ctl->gy0 = cs.gy0;
    ctl->gx0n = wrapi(cs.gx0 + sx, A.L);
    ctl->gy0n = wrapi(cs.gy0 + sy, A.L);
    ctl->dcls = dcls;
    ctl->zero_tile = (dcls != (int)cs.cls_written[A.parity_out]);
    ctl->fm = A.fm;
    ctl->pctr = (int)pctr;
    ctl->tumbled = tumbled;
    ctl->dwx = dc.wx;
    ctl->dwy = dc.wy;
    ctl->dtx0 = dc.tx0;
    ctl->dty0 = dc.ty0;
    ctl->no_margin = no_margin;
    ctl->gam = gam;
    ctl->dwC = dwC;
    ctl->volC = volC;
    ctl->vA = cs.v_A;
    ctl->cosp = (float)cd;
    ctl->sinp = (float)sd;
    ctl->theta = theta;
    ctl->mrel = mrel;
}

// Zero every tile pixel outside the destination window (class change). The
// window itself is fully rewritten by the update, so no ordering is needed.
__device__ void zero_outside_window(float* tile, int dwx, int dwy, int dtx0,
                                    int dty0) {
    for (int i = opaque_tid(); i < kTileArea; i += kNT) {
        const int x = i % kTilePitch, y = i / kTilePitch;
        if (!((unsigned)(x - dtx0) < (unsigned)dwx &&
              (unsigned)(y - dty0) < (unsigned)dwy))
            tile[i] = 0.0f;
    }
}

// ---- velocity observers (include/fused_moments.cuh:23-127,
// common/velocity_moments.h:42-100). Observation mode kObs: 0 none,
// 1 advection, 2 spatial. Nothing here changes the step's physics. ----

// Per-lane sums of the spatial observer: the interfacial (gx, gy) and
// overlap (kx, ky) centroid rates and the same-cell neighbour overlaps
// sum phi phi_east, sum phi phi_north (bx, by).
struct ObsSums {
    double gx = 0.0, gy = 0.0, kx = 0.0, ky = 0.0, bx = 0.0, by = 0.0;
};

// One pixel of the spatial observer (RelaxMoments::pixel,
// include/fused_moments.cuh:36-53, and the beta sums of
// src/kernels.cu:287-291): the pixel's weight 2 phi / V times its passive
// interfacial and overlap sources, in fp64 from the binary32 stencil, times
// its offset from the input centroid (g2iv = 2 / V, cx, cy).
__device__ __forceinline__ void observe_pixel(ObsSums& o, int x, int y, float c, float e,
                                              float w, float nn, float ss, float ne, float nw,
                                              float se, float sw, float other, float gam,
                                              float dwC, float repC, double g2iv, double cx,
                                              double cy) {
    const double v = (double)c;
    const double axial = (((double)nn + (double)ss) + (double)e) + (double)w;
    const double lap =
        (((((4.0 * axial + (double)ne) + (double)nw) + (double)se) + (double)sw) - 20.0 * v) /
        6.0;
    const double weight = g2iv * v;
    const double well = (((double)dwC * v) * (1.0 - v)) * (1.0 - 2.0 * v);
    const double interfacial = weight * ((double)gam * lap - well);
    const double dx = (double)x - cx, dy = (double)y - cy;
    o.gx += dx * interfacial;
    o.gy += dy * interfacial;
    const double overlap = (((-weight) * (double)repC) * v) * (double)other;
    o.kx += dx * overlap;
    o.ky += dy * overlap;
    o.bx += v * (double)e;
    o.by += v * (double)nn;
}

// Discrete linear-interpolation correction of a held rate
// (common/velocity_moments.h:80-85).
__device__ __forceinline__ double linear_correction(double dt, unsigned long long steps,
                                                    double first, double last) {
    return steps ? dt * 0.5 * (double)(steps - 1ull) * (last - first) : 0.0;
}

// Thread 0 after the velocity: the advection record of the step
// (record_advection, include/fused_moments.cuh:96-127) and, on a spatial
// step, the sample's bookkeeping (begin_spatial_sample,
// common/velocity_moments.h:59-77) and the new held rates
// (RelaxMoments::finish, include/fused_moments.cuh:55-88). red holds the
// wave sums: slots 2, 3 the neighbour overlaps, 4-7 gx, gy, kx, ky.
template <int kObs>
__device__ void record_observers(const pfh2d_step_args& A, int n, const Ctl& c, double sIx,
                                 double sIy, const double* red) {
    pfh_cell_state* cs = &A.cell[n];
    pfh_advection_accum& a = cs->advection;
    const double ax = (double)c.vA * (double)c.cosp;
    const double ay = (double)c.vA * (double)c.sinp;
    const double ix = (double)A.mot_coeff * sIx;
    const double iy = (double)A.mot_coeff * sIy;
    const double dt = A.physical_dt;
    if constexpr (kObs == 2) {
        pfh_moment_accum& m = A.moments[n];
        const double tot[4] = {a.q[0] + a.q[2], a.q[1] + a.q[3], a.q[4], a.q[5]};
        const unsigned long long interval = a.count - m.last_sample;
        if (m.spatial_count) {
            for (int j = 0; j < 4; ++j) m.q[4 + j] += (dt * (double)interval) * m.held[j];
            for (int xy = 0; xy < 2; ++xy) {
                const double beta = m.held[4 + xy];
                m.q[10 + xy] += (beta - 1.0) * (tot[xy] - m.last_totals[xy]) +
                                beta * (tot[2 + xy] - m.last_totals[2 + xy]);
            }
        }
        for (int j = 0; j < 4; ++j) m.last_totals[j] = tot[j];
        m.last_interval = interval;
        m.last_sample = a.count;
        ++m.spatial_count;
        double bx = 0.0, by = 0.0;
        for (int w = 0; w < kWaves; ++w) {
            bx += red[w * 8 + 2];
            by += red[w * 8 + 3];
        }
        m.held[4] = bx / cs->V;
        m.held[5] = by / cs->V;
    }
    a.q[0] += dt * ax;
    a.q[1] += dt * ay;
    a.q[2] += dt * ix;
    a.q[3] += dt * iy;
    a.q[4] += dt * (((double)c.vx - ax) - ix);
    a.q[5] += dt * (((double)c.vy - ay) - iy);
    ++a.count;
    if constexpr (kObs == 2) {
        pfh_moment_accum& m = A.moments[n];
        const double mrel = (double)cs->M_pf / 0.5;
        for (int j = 0; j < 4; ++j) {
            double rate = 0.0;
            for (int w = 0; w < kWaves; ++w) rate += red[w * 8 + 4 + j];
            // only the passive rates are held here; the area source has an
            // exactly zero centroid moment
            if (mrel != 1.0) rate *= mrel;
            if (m.spatial_count > 1)
                m.correction[j] += linear_correction(dt, m.last_interval, m.held[j], rate);
            m.held[j] = rate;
        }
    }
}

// Velocity from the interaction integral (src/kernels.cu:300-342): wave
// trees, then thread 0 adds the 16 wave sums in order and commits; with an
// observer, the observer's sums ride in the same reduction.
template <int kObs = 0>
__device__ void reduce_velocity(const pfh2d_step_args& A, int n, Smem& sm,
                                Ctl& c, double aIx, double aIy, ObsSums ob = ObsSums()) {
    const int tid = opaque_tid();
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    aIx = wave_sum(aIx);
    aIy = wave_sum(aIy);
    if constexpr (kObs == 2) {
        ob.bx = wave_sum(ob.bx);
        ob.by = wave_sum(ob.by);
        ob.gx = wave_sum(ob.gx);
        ob.gy = wave_sum(ob.gy);
        ob.kx = wave_sum(ob.kx);
        ob.ky = wave_sum(ob.ky);
    }
    if (lane == 0) {
        sm.red[wave * 8 + 0] = aIx;
        sm.red[wave * 8 + 1] = aIy;
        if constexpr (kObs == 2) {
            double* r = sm.red + wave * 8;
            r[2] = ob.bx; r[3] = ob.by;
            r[4] = ob.gx; r[5] = ob.gy; r[6] = ob.kx; r[7] = ob.ky;
        }
    }
    lds_barrier();
    if (threadIdx.x == 0) {
        double sIx = 0.0, sIy = 0.0;
        for (int w = 0; w < kWaves; ++w) {
            sIx += sm.red[w * 8 + 0];
            sIy += sm.red[w * 8 + 1];
        }
        const double vxd = (double)c.vA * (double)c.cosp
                         + (double)A.mot_coeff * sIx;
        const double vyd = (double)c.vA * (double)c.sinp
                         + (double)A.mot_coeff * sIy;
        c.vx = (float)vxd;
        c.vy = (float)vyd;
        pfh_cell_state* cs = &A.cell[n];
        cs->vx = (float)vxd;
        cs->vy = (float)vyd;
        cs->Ix = sIx;
        cs->Iy = sIy;
        if constexpr (kObs != 0) record_observers<kObs>(A, n, c, sIx, sIy, sm.red);
    }
    lds_barrier();
}

// Moments/support reduction and CellState commit (src/kernels.cu:542-591).
struct NoHook {
    __device__ void operator()() const {}
};

// after_partials runs on every thread once the wave partials are published
// (every wave has finished the cell's pixel phases).
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// old_raw: the cell's CellState as fetched at the start of its turn (LDS),
// for the counters that the commit advances.
template <int CLS, class Hook = NoHook>
__device__ void reduce_commit(const pfh2d_step_args& A, int n, Smem& sm,
                              const Ctl& c, const uint32_t* old_raw,
                              double aV, double aCx, double aCy, double aPer,
                              int blox, int bhix, int bloy, int bhiy,
                              int pmaxb, Hook after_partials = Hook()) {
    const int tid = opaque_tid();
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    aV = wave_sum(aV);
    aCx = wave_sum(aCx);
    aCy = wave_sum(aCy);
    aPer = wave_sum(aPer);
    blox = wave_min(blox);
    bhix = wave_max(bhix);
    bloy = wave_min(bloy);
    bhiy = wave_max(bhiy);
    pmaxb = wave_max(pmaxb);
    if (lane == 0) {
        double* rw = sm.red + wave * 8;
        rw[0] = aV; rw[1] = aCx; rw[2] = aCy; rw[3] = aPer;
        int* iw = sm.ired + wave * 8;
        iw[0] = blox; iw[1] = bhix; iw[2] = bloy; iw[3] = bhiy; iw[4] = pmaxb;
    }
    lds_barrier();
    after_partials();
    if (__builtin_amdgcn_readfirstlane(wave) == 0) {
        // Lanes 0-3 add the 16 wave partials of V, Cx, Cy and the perimeter,
        // lanes 4-8 reduce the support box and phi_max, each in wave order
        // from the same initial value: the sums of one thread doing all nine.
        double sl = 0.0;
        int ml = 0;
        if (lane < 4) {
            for (int w = 0; w < kWaves; ++w) sl += sm.red[w * 8 + lane];
        } else if (lane < 9) {
            const int j = lane - 4;                  // 0 lo_x, 1 hi_x, 2 lo_y, 3 hi_y, 4 pmax
            const bool is_max = (j & 1) != 0 || j == 4;
            ml = j == 0 ? c.dwx : j == 2 ? c.dwy : j == 4 ? 0 : -1;
            for (int w = 0; w < kWaves; ++w) {
                const int v = sm.ired[w * 8 + j];
                ml = is_max ? max(ml, v) : min(ml, v);
            }
        }
        const double sV = readlane_f64(sl, 0), sCx = readlane_f64(sl, 1);
        const double sCy = readlane_f64(sl, 2), sPer = readlane_f64(sl, 3);
        const int Blox = __builtin_amdgcn_readlane(ml, 4);
        const int Bhix = __builtin_amdgcn_readlane(ml, 5);
        const int Bloy = __builtin_amdgcn_readlane(ml, 6);
        const int Bhiy = __builtin_amdgcn_readlane(ml, 7);
        const int Pmax = __builtin_amdgcn_readlane(ml, 8);
        const pfh_cell_state* old = reinterpret_cast<const pfh_cell_state*>(old_raw);
        if (lane == 0) {
        pfh_cell_state* cs = &A.cell[n];
        cs->gx0 = c.gx0n;
        cs->gy0 = c.gy0n;
        cs->cls = (uint8_t)c.dcls;
        cs->cls_written[A.parity_out] = (uint8_t)c.dcls;
        cs->theta = c.theta;
        cs->V = sV;
        cs->Cx = sCx;
        cs->Cy = sCy;
        if (c.fm) cs->perim = sPer;
        cs->bb_lo_x = Blox;
        cs->bb_hi_x = Bhix;
        cs->bb_lo_y = Bloy;
        cs->bb_hi_y = Bhiy;
        cs->promote_ctr = (uint32_t)c.pctr;
        cs->shift_ctr = old->shift_ctr + ((c.sx | c.sy) ? 1u : 0u);
        cs->tumble_ctr = old->tumble_ctr + (uint32_t)c.tumbled;
        cs->phi_max = __uint_as_float((uint32_t)Pmax);
        A.cell_cls[n] = (uint8_t)c.dcls;
        if (CLS == kClassFallback) {
            cs->reserved[0] = 1u;
            const bool edge = Bhix >= 0 && (Blox == 0 || Bhix == c.dwx - 1 ||
                                            Bloy == 0 || Bhiy == c.dwy - 1);
            if (c.no_margin || (c.dcls == kClassFallback && edge))
                cs->reserved[1] = old->reserved[1] + 1u;
        }
        }
    }
    lds_barrier();
}

// The check step's commit: the cell's new window position and class (all its
// consumer reads), between the same two barriers as reduce_commit.
template <class Hook>
__device__ void commit_position(const pfh2d_step_args& A, int n, Smem& sm, const Ctl& c,
                                Hook after_partials) {
    lds_barrier();
    after_partials();
    if (threadIdx.x == 0) {
        pfh_cell_state* cs = &A.cell[n];
        cs->gx0 = c.gx0n;
        cs->gy0 = c.gy0n;
        cs->cls = (uint8_t)c.dcls;
        cs->cls_written[A.parity_out] = (uint8_t)c.dcls;
        A.cell_cls[n] = (uint8_t)c.dcls;
    }
    lds_barrier();
}

// -------------------------------------------------------------------------
// Next-cell fetch by the control lanes, in four stages so that each stage
// only waits on a load issued one phase earlier:
//   1 the queue ticket (atomic on the cursor);
//   2 the cell id perm[ticket];
//   3 its class and its 192-byte CellState (3 dwords per control lane);
//   4 the CellState copy into LDS, the slot, and the per-cell scalars.
// An unsupported class id is flagged here (reference dispatch default,
// src/kernels.cu:1017-1021) and skipped by the dispatch.
// -------------------------------------------------------------------------
struct Fetch {
    unsigned long long tk;
    int nn, cls;
    uint32_t w0, w1, w2;
};

__device__ __forceinline__ void fetch1(const pfh2d_step_args& A, Fetch& f) {
    if (((int)threadIdx.x & (kWave - 1)) == kLeaderLane)
        f.tk = atomicAdd(A.cursor_use, 1ull);
}

__device__ __forceinline__ void fetch2(const pfh2d_step_args& A, Fetch& f) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)f.tk, kLeaderLane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(f.tk >> 32), kLeaderLane);
    const unsigned long long k = ((unsigned long long)hi << 32) | lo;
    f.nn = k < (unsigned long long)A.N ? (int)A.perm[k] : -1;
}

__device__ __forceinline__ void fetch3(const pfh2d_step_args& A, Fetch& f) {
    if (f.nn >= 0) {
        const int j = ((int)threadIdx.x & (kWave - 1)) - kLeaderLane;
        const uint32_t* w = reinterpret_cast<const uint32_t*>(A.cell + f.nn);
        f.cls = (int)A.cell_cls[f.nn];
        f.w0 = w[j];
        f.w1 = w[j + 16];
        f.w2 = w[j + 32];
    }
}

// 4a (control lanes): the CellState copy, the id and the class into LDS.
__device__ __forceinline__ void fetch4a(const pfh2d_step_args& A, Smem& sm, Fetch& f,
                                        int which) {
    const int j = ((int)threadIdx.x & (kWave - 1)) - kLeaderLane;
    int cls = -1;
    if (f.nn >= 0) {
        sm.csraw[which][j] = f.w0;
        sm.csraw[which][j + 16] = f.w1;
        sm.csraw[which][j + 32] = f.w2;
        cls = f.cls < kNumClasses ? f.cls : kNumClasses;
    }
    if (j == 0) {
        if (cls == kNumClasses) atomicAdd(&A.flags[PFH_FLAG_CLASS_UNSUPPORTED], 1u);
        sm.slot[which].n = f.nn;
        sm.slot[which].cls = cls;
    }
}

// 4b (one thread, after 4a is visible to it): the per-cell scalars.
template <bool kBatch>
__device__ __forceinline__ void fetch4b(const pfh2d_step_args& A, Smem& sm, int which,
                                        unsigned long long step) {
    const int cls = sm.slot[which].cls;
    if (cls >= 0 && cls < kNumClasses) {
        ScalarArgs s = scalar_args(A, sm.fm);
        if constexpr (kBatch) {
            // batched step: the cell's replica sets its domain and stream
            Slot& sl = sm.slot[which];
            const pfh2d_replica& rp = sm.reps[sm.cell_rep[sl.n]];
            s.L = rp.L;
            s.polarity_seed = rp.polarity_seed;
            sl.L = rp.L;
            sl.P = rp.P;
            sl.s_off_lo = (int)(uint32_t)rp.S_offset;
            sl.s_off_hi = (int)(uint32_t)(rp.S_offset >> 32);
        }
        cell_scalars(s, reinterpret_cast<const pfh_cell_state*>(sm.csraw[which]), cls, step,
                     &sm.slot[which].ctl);
    }
}

template <bool kBatch>
__device__ __forceinline__ void fetch4(const pfh2d_step_args& A, Smem& sm, Fetch& f,
                                       int which, unsigned long long step) {
    fetch4a(A, sm, f, which);
    if (((int)threadIdx.x & (kWave - 1)) == kLeaderLane) fetch4b<kBatch>(A, sm, which, step);
}

template <bool kBatch = false>
__device__ __forceinline__ void fetch_all(const pfh2d_step_args& A, Smem& sm, int which,
                                          unsigned long long step) {
    Fetch f;
    fetch1(A, f);
    fetch2(A, f);
    fetch3(A, f);
    fetch4<kBatch>(A, sm, f, which, step);
}

// -------------------------------------------------------------------------
// LDS path: classes 0-3.
//
// LDS holds the window with a one-pixel zero ring: row r (= y + 1, y in
// [-1, WY]) starts at r*PX with four pad floats (the last one is column -1)
// and column WX falls on the next row's first pad, so the stencil never needs
// an edge test. Thread (band, x) walks rows y0..y0+RB-1 of column x with a
// rolling 3x3 neighbourhood: three LDS reads per row.
//
// Exact shortcuts (phi never takes the value -0, so adding or subtracting a
// signed zero changes no sum or right-hand side):
//   - a wave-row where no lane overlaps another cell (S_other == 0) skips the
//     repulsion term and the fp64 interaction-integral terms, all zero;
//   - a wave-row where every lane has c == passive == gx == gy == 0 has
//     phi_next == +0 exactly and only stores the zeros.
// -------------------------------------------------------------------------
struct Roll {
    float sW, sC, sE, cW, cC, cE, nW, nC, nE;
};

__device__ __forceinline__ void load_row(const float* p, float& w, float& c, float& e) {
    w = p[-1];
    c = p[0];
    e = p[1];
}

__device__ __forceinline__ bool wave_any(bool v) {
    return __builtin_amdgcn_ballot_w64(v) != 0ull;
}

// First part of phase_field_rhs (include/kernels.cuh:127-136): every passive
// term but the repulsion, the reference's terms in its order with each
// product-and-add fused (one rounding where the reference's source has a
// multiply and an add): the 9-point Laplacian (edge sum x 4 + diagonal sum,
// then - 20 x centre, over 6), then gam x lap - dwC x phi(1 - phi)(1 - 2 phi),
// then + volC x phi.
__device__ __forceinline__ float passive_head_of(float nC, float sC, float cE, float cW,
                                                 float nE, float nW, float sE, float sW,
                                                 float cC, float gam, float dwC, float volC) {
    const float edge = nC + sC + cE + cW;
    const float diag = nE + nW + sE + sW;
    const float lap = fmaf(-20.0f, cC, fmaf(4.0f, edge, diag)) * (float)(1.0 / 6.0);
    const float dw = (cC * (1.0f - cC)) * fmaf(-2.0f, cC, 1.0f);
    return fmaf(volC, cC, fmaf(gam, lap, -(dwC * dw)));
}

__device__ __forceinline__ float passive_head(const Roll& q, float gam, float dwC,
                                              float volC) {
    return passive_head_of(q.nC, q.sC, q.cE, q.cW, q.nE, q.nW, q.sE, q.sW, q.cC, gam, dwC,
                           volC);
}

// The repulsion term: passive = head - repC x phi x S_other, fused.
__device__ __forceinline__ float with_repulsion(float head, float repC, float cC, float So) {
    return fmaf(-(repC * cC), So, head);
}

// The Euler update phi + dt (mrel x passive - (vx gx + vy gy)), fused. The
// reference takes passive - adv when mrel == 1 (include/kernels.cuh:118-138):
// 1 x passive is passive exactly, so one expression gives both.
__device__ __forceinline__ float euler_update(float cC, float gxv, float gyv, float pas,
                                              float vx, float vy, float mrel, float dt) {
    const float adv = fmaf(vy, gyv, vx * gxv);
    return fmaf(dt, fmaf(mrel, pas, -adv), cC);
}

__device__ __forceinline__ float s_other_of(uint32_t qS, float c, uint32_t& negc) {
    const uint32_t qc = q_of(c);
    const bool neg = qS < qc;
    negc += neg ? 2u : 0u;          // the reference evaluates S_other once per pass
    return neg ? 0.0f : (float)(qS - qc) * kQInv;
}

// The same, counting this lane's negative differences once each (the
// subtraction's borrow is added in with carry).
__device__ __forceinline__ float s_other_lane(uint32_t qS, float c, uint32_t& neg1) {
    const uint32_t qc = q_of(c);
    const bool neg = qS < qc;
    neg1 += (uint32_t)neg;
    return neg ? 0.0f : (float)(qS - qc) * kQInv;
}

// Class-0 window prefetch: rows [r0, r1) of cell nn's class-0 window,
// copied from HBM straight into LDS (global-to-LDS loads, no registers), row
// r to dst0 + (r - r0) * PX. Waves W0..W0+NW-1 take rows r0 + (w - W0) + NW j;
// one instruction moves one 576 B row with 36 lanes of 16 B. Wave 0 (thread
// 0 commits the cell, the control lanes fetch) is never among them, so that
// none of its own memory waits has to wait for the copies. A wave's copies
// are complete when its vmcnt has passed them; LDS release fences (and so
// every barrier here) wait for them too.
template <int W0 = 1, int NW = kWaves - 1>
__device__ __forceinline__ void copy_rows0(const pfh2d_step_args& A, int nn, float* dst0,
                                           int r0, int r1) {
    using G = Staged<0>;
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
    const PFH_GLOBAL float* src = gptr(A.phi_in) + (size_t)nn * kTileArea
                                + (size_t)G::TY0 * kTilePitch + G::TX0 + 4 * lane;
    if (wave >= W0 && wave < W0 + NW && lane < G::NCH) {
        for (int r = r0 + wave - W0; r < r1; r += NW)
            __builtin_amdgcn_global_load_lds(
                (const PFH_GLOBAL void*)(src + (size_t)r * kTilePitch),
                (__attribute__((address_space(3))) void*)(dst0 + (r - r0) * G::PX),
                16, 0, 0);
    }
}

// kCheck: a check step (the surrogate's acceptance field), which needs the
// new field and the cell's window position only: no scatter, moments,
// support, flags or perimeter, and a commit of the position alone.
// The cell's domain: the launch's own population or, in a batched step,
// its replica's side, S pitch and S offset, made wave-uniform.
struct Dom {
    int L, P;
    unsigned long long off;
};
template <bool kBatch>
__device__ __forceinline__ Dom cell_dom(const pfh2d_step_args& A, const Slot& c) {
    if constexpr (kBatch) {
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)c.s_off_lo);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)c.s_off_hi);
        return {__builtin_amdgcn_readfirstlane(c.L), __builtin_amdgcn_readfirstlane(c.P),
                ((unsigned long long)hi << 32) | lo};
    } else {
        return {A.L, A.P, 0ull};
    }
}

template <int CLS, bool kCheck = false, bool kBatch = false, int kObs = 0>
__device__ void cell_staged(const pfh2d_step_args& A, int n,
                            unsigned long long step, Smem& sm, Ctl& c, int nxt,
                            int& pre, int& pslot, int* ring = nullptr) {
    using G = Staged<CLS>;
    constexpr int WX = G::WX, WY = G::WY, TX0 = G::TX0, TY0 = G::TY0;
    constexpr int PX = G::PX;
    constexpr int RB = G::RB, NCH = G::NCH, NCHUNK = G::NCHUNK, KCH = G::KCH;
    constexpr bool kExact = G::NB * RB == WY;   // no partial band
    const int tid = opaque_tid();
    const int pix = tid - kControlLanes;                 // pixel thread index
    const bool control = pix < 0;
    const int x = control ? 0 : pix % WX;
    const int band = control ? 0 : pix / WX;
    const bool active = !control && band < G::NB;
    const int y0 = band * RB;
    Fetch nf;
    if (control) fetch1(A, nf);
    const Dom dm = cell_dom<kBatch>(A, sm.slot[nxt ^ 1]);
    const int L = dm.L, P = dm.P;
    // the origin from the LDS copy of the CellState: a global load here would
    // make the S loads wait for every load issued before it
    const int gx0 = c.gx0;
    const int gy0 = c.gy0;
    float* phi_s = s_window;
    // The window in LDS: class 0 in its two pieces (piece 1 in slot pslot),
    // the other classes whole. rowp(r) is the start of row r's data; cbw is
    // this thread's column at row y0 - 1 (a band's rows are all in one piece).
    float* base1 = phi_s;
    float* base2 = phi_s;
    if constexpr (CLS == 0) {
        base1 = phi_s + pslot * Split0::P1_FLOATS;
        base2 = phi_s + 2 * Split0::P1_FLOATS;
    }
    auto rowp = [&](int r) -> float* {
        if constexpr (CLS == 0)
            return r <= Split0::P2_R0 ? base1 + (r + 1) * PX + 4
                                      : base2 + (r - Split0::P2_R0) * PX + 4;
        else
            return phi_s + (r + 1) * PX + 4;
    };
    float* cbw = base1 + y0 * PX + 4 + x;
    if constexpr (CLS == 0) {
        if (band >= Split0::KB) cbw = base2 + (y0 - 1 - Split0::P2_R0) * PX + 4 + x;
    }
    const int gx = (gx0 + x < L) ? gx0 + x : gx0 + x - L;
    // k is a compile-time row index: rows below RFULL exist in every band, so
    // only the last band's tail rows are tested at run time.
    auto row_ok = [&](int k) { return kExact || k < G::RFULL || y0 + k < WY; };

    // Window rows of phi in flight (16 B each) ...
    const PFH_GLOBAL float* tin = gptr(A.phi_in) + (size_t)n * kTileArea;
    const PFH_GLOBAL uint32_t* S_rd = gptr(A.S_rd + dm.off);
    const bool have = CLS == 0 && pre != 0;     // window already copied into LDS
    f32x4 ph[KCH];
    if (!have) {
#pragma unroll
        for (int k = 0; k < KCH; ++k) {
            const int u = tid + k * kNT;
            if (u < NCHUNK) {
                const int r = u / NCH, c4 = u - (u / NCH) * NCH;
                ph[k] = *reinterpret_cast<const PFH_GLOBAL f32x4*>(
                    tin + (size_t)(TY0 + r) * kTilePitch + TX0 + 4 * c4);
            }
        }
    }
    // ... and this thread's column of the S window. S is addressed by a
    // 32-bit byte offset (4 P L < 2^32, checked at launch) that advances one
    // row pitch per row and wraps once at the domain edge.
    const uint32_t s_pitch = 4u * (uint32_t)P, s_wrap = s_pitch * (uint32_t)L;
    uint32_t s_off0;
    {
        int gy = gy0 + y0;
        if (gy >= L) gy -= L;
        s_off0 = (uint32_t)gy * s_pitch + 4u * (uint32_t)gx;
    }
    // Every lane issues all RB loads (control lanes and rows past the window
    // read in-bounds S words that no row uses): with no branch around a load,
    // each row of phase A waits only for its own word.
    auto load_qs = [&](uint32_t (&qS)[RB]) {
        uint32_t off = s_off0;
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            qS[k] = *reinterpret_cast<const PFH_GLOBAL uint32_t*>(
                reinterpret_cast<const PFH_GLOBAL char*>(S_rd) + off);
            off += s_pitch;
            if (off >= s_wrap) off -= s_wrap;
            asm volatile("" : "+v"(off));   // one live offset, not one per row
        }
    };
    uint32_t qS[RB];
    load_qs(qS);
    if (have) {
        // This wave's global-to-LDS copies, issued during the previous cell,
        // are older than the RB loads just issued: retire all but those.
        static_assert(RB <= 63, "vmcnt is 6 bits");
        asm volatile("s_waitcnt vmcnt(%0)" :: "n"(RB) : "memory");
    } else {
#pragma unroll
        for (int k = 0; k < KCH; ++k) {
            const int u = tid + k * kNT;
            if (u < NCHUNK) {
                const int r = u / NCH, c4 = u - (u / NCH) * NCH;
                if constexpr (CLS == 0) {             // rows P2_R0, R1 go to both pieces
                    if (r <= Split0::R1)
                        *reinterpret_cast<f32x4*>(base1 + (r + 1) * PX + 4 + 4 * c4) = ph[k];
                    if (r >= Split0::P2_R0)
                        *reinterpret_cast<f32x4*>(base2 + (r - Split0::P2_R0) * PX + 4 + 4 * c4) = ph[k];
                } else {
                    *reinterpret_cast<f32x4*>(phi_s + (r + 1) * PX + 4 + 4 * c4) = ph[k];
                }
            }
        }
    }
    // zero ring: the pads of every row (column -1 and column WX) and rows -1, WY
    if constexpr (CLS == 0) {
        // The class-0 layout's ring is written by nothing but this set-up
        // (copies, staging and the perimeter pass write data columns only):
        // both piece-1 slots and piece 2 are set up once, and again after a
        // cell of another class has used the LDS.
        if (*ring == 0) {
            constexpr int n1 = (Split0::P1_ROWS + 1) * 4, n2 = (Split0::P2_ROWS + 1) * 4;
            float* const b2 = phi_s + 2 * Split0::P1_FLOATS;
            for (int i = tid; i < 2 * n1 + n2 + 3 * WX; i += kNT) {
                if (i < 2 * n1) {
                    const int sl = i >= n1 ? 1 : 0, j = i - sl * n1;
                    phi_s[sl * Split0::P1_FLOATS + (j >> 2) * PX + (j & 3)] = 0.0f;
                } else if (i < 2 * n1 + n2) {
                    const int j = i - 2 * n1;
                    b2[(j >> 2) * PX + (j & 3)] = 0.0f;
                } else {
                    const int j = i - 2 * n1 - n2;
                    if (j < 2 * WX) {
                        const int sl = j >= WX ? 1 : 0;
                        phi_s[sl * Split0::P1_FLOATS + 4 + (j - sl * WX)] = 0.0f;
                    } else {
                        b2[(WY - Split0::P2_R0) * PX + 4 + (j - 2 * WX)] = 0.0f;
                    }
                }
            }
            *ring = 1;
        }
    } else {
        for (int i = tid; i < (WY + 3) * 4 + 2 * WX; i += kNT) {
            if (i < (WY + 3) * 4) phi_s[(i >> 2) * PX + (i & 3)] = 0.0f;
            else {
                const int j = i - (WY + 3) * 4;
                phi_s[(j < WX ? 0 : (WY + 1) * PX - WX) + 4 + j] = 0.0f;
            }
        }
    }
    lds_barrier();
    // The id and then the CellState of the next cell: wave 0 waits for them
    // here, while the other waves of its SIMD start phase A.
    if (control) {
        fetch2(A, nf);
        fetch3(A, nf);
    }

    const float gam = c.gam, dwC = c.dwC, volC = c.volC;
    const float repC = A.rep_coeff;
    if (c.zero_tile)
        zero_outside_window(A.phi_out + (size_t)n * kTileArea, c.dwx, c.dwy,
                            c.dtx0, c.dty0);

    // ---- phase A: interaction integral and the held passive RHS ----
    float pas[RB];
    double aIx = 0.0, aIy = 0.0;
    ObsSums ob;
    if (active) {
        uint32_t neg1 = 0u;
        Roll q;
        load_row(cbw, q.sW, q.sC, q.sE);
        load_row(cbw + PX, q.cW, q.cC, q.cE);
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            pas[k] = 0.0f;
            if (row_ok(k)) {
                load_row(cbw + (k + 2) * PX, q.nW, q.nC, q.nE);
                const float head = passive_head(q, gam, dwC, volC);
                const float So = s_other_lane(qS[k], q.cC, neg1);
                if constexpr (kObs == 2)
                    observe_pixel(ob, x, y0 + k, q.cC, q.cE, q.cW, q.nC, q.sC, q.nE, q.nW,
                                  q.sE, q.sW, So, gam, dwC, repC, c.og2iv, c.ocx, c.ocy);
                if (wave_any(So != 0.0f)) {
                    pas[k] = with_repulsion(head, repC, q.cC, So);
                    const float gxv = 0.5f * (q.cE - q.cW);
                    const float gyv = 0.5f * (q.nC - q.sC);
                    aIx += (double)(q.cC * gxv * So);
                    aIy += (double)(q.cC * gyv * So);
                } else {
                    pas[k] = head;
                }
                q.sW = q.cW; q.sC = q.cC; q.sE = q.cE;
                q.cW = q.nW; q.cC = q.nC; q.cE = q.nE;
            }
            // One row per scheduling region keeps the walk's register
            // footprint at one row instead of the whole column.
            __builtin_amdgcn_sched_barrier(0);
        }
        // two events per negative pixel (the reference evaluates S_other once
        // per pass)
        if (neg1) flag_add(A.flags, PFH_FLAG_S_NEGATIVE, 2u * neg1);
    }

    if (control) fetch4a(A, sm, nf, nxt);       // visible after the next barrier
    reduce_velocity<kObs>(A, n, sm, c, aIx, aIy, ob);
    // A class-0 cell next in the queue gets its window copied in ahead:
    // piece 1 into the other slot (its last reader was the previous cell) as
    // soon as the first waves finish phase B, piece 2 once every wave has.
    int next_nn = -1;
    if constexpr (CLS == 0) {
        const int ncls = __builtin_amdgcn_readfirstlane(sm.slot[nxt].cls);
        const int nn = __builtin_amdgcn_readfirstlane(sm.slot[nxt].n);
        if (ncls == 0 && nn >= 0) next_nn = nn;
    }

    // ---- phase B: Euler update, shifted store, scatter, moments ----
    const float vx = c.vx, vy = c.vy, mrel = c.mrel, dt = A.dt;
    const int sx = c.sx, sy = c.sy, dwx = c.dwx, dwy = c.dwy;
    const int dtx0 = c.dtx0, dty0 = c.dty0;
    PFH_GLOBAL float* tout = gptr(A.phi_out) + (size_t)n * kTileArea;
    double aV = 0.0, aCx = 0.0, aCy = 0.0, aPer = 0.0;
    int blox = dwx, bhix = -1, bloy = dwy, bhiy = -1, pmaxb = 0;
    bool nonfin = false, clampf = false, ovf = false;
    auto update = [&](float cC, float gxv, float gyv, float pas_k) -> float {
        return euler_update(cC, gxv, gyv, pas_k, vx, vy, mrel, dt);
    };
    const int a = x - sx;
    if (active) {
        const bool ain = (unsigned)a < (unsigned)dwx;
        uint32_t olds[RB];
        int rlo = dwy, rhi = -1;
        const float* cb = cbw;
        float sC = cb[0];
        float cW, cC, cE;
        load_row(cb + PX, cW, cC, cE);
        PFH_GLOBAL float* orow = tout + (ptrdiff_t)(dty0 + y0 - sy) * kTilePitch + dtx0 + a;
        // the aggregate's base in a VGPR pair: the per-row atomic then needs
        // neither the argument block nor a scalar wait
        const bool scatter = A.S_sc != nullptr;   // NULL: a step that scatters nothing
        PFH_GLOBAL char* S_sc = reinterpret_cast<PFH_GLOBAL char*>(
            gptr(scatter ? A.S_sc + dm.off : A.S_sc));
        asm volatile("" : "+v"(S_sc));
        uint32_t s_off = s_off0;
        const double ad = (double)a;
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            const int y = y0 + k;
            olds[k] = 0u;
            if (row_ok(k)) {
                float nW, nC, nE;
                load_row(cb + (k + 2) * PX, nW, nC, nE);
                const float gxv = 0.5f * (cE - cW);
                const float gyv = 0.5f * (nC - sC);
                const int b = y - sy;
                const bool valid = ain && (unsigned)b < (unsigned)dwy;
                const float pn = update(cC, gxv, gyv, pas[k]);
                if (valid) *orow = pn;
                if constexpr (!kCheck) {
                // A lane whose destination is outside the window contributes
                // an exact zero to every sum below.
                const float pv = valid ? pn : 0.0f;
                const float sq = pv * pv;
                nonfin |= !isfinite(pv);
                clampf |= (sq > kQClampPhiSq);
                pmaxb = max(pmaxb, (int)__float_as_uint(fabsf(pv)));
                const uint32_t qv = __float2uint_rn(fminf(sq, kQClampPhiSq) * kQScale);
                if (qv && scatter)
                    olds[k] = __hip_atomic_fetch_add(
                        reinterpret_cast<PFH_GLOBAL uint32_t*>(S_sc + s_off), qv,
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const double d = (double)pv * (double)pv;
                aV += d;
                aCy += d * (double)b;
                if (pv > kSupportEps) { rlo = min(rlo, b); rhi = max(rhi, b); }
                }
                sC = cC;
                cW = nW; cC = nC; cE = nE;
                orow += kTilePitch;
                asm volatile("" : "+v"(orow));   // a pointer increment, not a multiply
                s_off += s_pitch;
                if (s_off >= s_wrap) s_off -= s_wrap;
                asm volatile("" : "+v"(s_off));
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        aCx = aV * ad;                 // every pixel of this thread has column a
        if (rhi >= 0) { blox = a; bhix = a; bloy = rlo; bhiy = rhi; }
        // Exact S_overflow test with one register per pixel: every addend is
        // at most 2^29, so a wrap needs old >= 2^32 - 2^29; only such
        // candidates recompute q from this thread's own stored value.
#pragma unroll
        for (int k = 0; k < (kCheck ? 0 : RB); ++k) {
            if (olds[k] >= 0xE0000000u) {
                const int b = y0 + k - sy;
                const uint32_t qv = q_of(tout[(size_t)(dty0 + b) * kTilePitch + dtx0 + a]);
                ovf |= olds[k] > 0xFFFFFFFFu - qv;
            }
        }
    }
    if (nonfin) flag_or(A.flags, PFH_FLAG_NONFINITE);
    if (clampf) flag_or(A.flags, PFH_FLAG_Q_CLAMP);
    if (ovf) flag_or(A.flags, PFH_FLAG_S_OVERFLOW);

    // Destination pixels whose source lies outside the window are zero.
    if (c.dcls == CLS) {
        const int ny = sy > 0 ? sy : -sy, nx = sx > 0 ? sx : -sx;
        const int rows = ny * dwx, total = rows + nx * dwy;
        for (int i = tid; i < total; i += kNT) {
            int aa, b;
            if (i < rows) {
                const int rr = i / dwx;
                aa = i - rr * dwx;
                b = sy > 0 ? dwy - sy + rr : rr;
            } else {
                const int j = i - rows, cc = j / dwy;
                b = j - cc * dwy;
                aa = sx > 0 ? dwx - sx + cc : cc;
            }
            tout[(size_t)(dty0 + b) * kTilePitch + dtx0 + aa] = 0.0f;
        }
    } else {
        for (int i = tid; i < dwx * dwy; i += kNT) {
            const int b = i / dwx, aa = i - (i / dwx) * dwx;
            const int ii = aa + sx, jj = b + sy;
            if (!((unsigned)ii < (unsigned)WX && (unsigned)jj < (unsigned)WY))
                tout[(size_t)(dty0 + b) * kTilePitch + dtx0 + aa] = 0.0f;
        }
    }

    // Perimeter on full-moment steps (one step in full_moment_every) from the
    // new field in the source layout with source-window zeros (the staged
    // semantics, src/kernels.cu:512-538). phi_next is recomputed here, bit for
    // bit as above, instead of being held through the update.
    if (!kCheck && c.fm) {
        float pns[RB];
        if (active) {
            uint32_t qS2[RB];
            load_qs(qS2);
            uint32_t negc2 = 0u;
            Roll q;
            load_row(cbw, q.sW, q.sC, q.sE);
            load_row(cbw + PX, q.cW, q.cC, q.cE);
#pragma unroll
            for (int k = 0; k < RB; ++k) {
                pns[k] = 0.0f;
                if (row_ok(k)) {
                    load_row(cbw + (k + 2) * PX, q.nW, q.nC, q.nE);
                    const float So = s_other_of(qS2[k], q.cC, negc2);
                    const float pk = with_repulsion(passive_head(q, gam, dwC, volC), repC,
                                                    q.cC, So);
                    pns[k] = update(q.cC, 0.5f * (q.cE - q.cW), 0.5f * (q.nC - q.sC), pk);
                    q.sW = q.cW; q.sC = q.cC; q.sE = q.cE;
                    q.cW = q.nW; q.cC = q.nC; q.cE = q.nE;
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        lds_barrier();                         // all reads of phi_s done
        if (active) {
#pragma unroll
            for (int k = 0; k < RB; ++k)
                if (row_ok(k)) cbw[(k + 1) * PX] = pns[k];
        }
        lds_barrier();
        for (int i = tid; i < dwx * dwy; i += kNT) {
            const int b = i / dwx, aa = i - (i / dwx) * dwx;
            const int ii = aa + sx, jj = b + sy;
            const bool iin = (unsigned)ii < (unsigned)WX;
            const bool jin = (unsigned)jj < (unsigned)WY;
            float e = 0.0f, w = 0.0f, nn = 0.0f, ss = 0.0f;
            if (jin) {
                const float* pr = rowp(jj) + ii;
                if ((unsigned)(ii + 1) < (unsigned)WX) e = pr[1];
                if ((unsigned)(ii - 1) < (unsigned)WX) w = pr[-1];
            }
            if (iin) {
                if ((unsigned)(jj + 1) < (unsigned)WY) nn = rowp(jj + 1)[ii];
                if ((unsigned)(jj - 1) < (unsigned)WY) ss = rowp(jj - 1)[ii];
            }
            const float pgx = 0.5f * (e - w), pgy = 0.5f * (nn - ss);
            aPer += (double)sqrtf(pgx * pgx + pgy * pgy);
        }
    }
    // The next cell's scalars, on a thread of wave 1: the first waves of each
    // SIMD finish phase B early, wave 0 (control lanes, commit) the latest.
    if (tid == kScalarsThread) fetch4b<kBatch>(A, sm, nxt, step);
    if constexpr (CLS == 0) {
        // piece 1 of the next window: each wave issues its rows as it
        // finishes phase B
        if (next_nn >= 0)
            copy_rows0(A, next_nn, phi_s + (pslot ^ 1) * Split0::P1_FLOATS + PX + 4, 0,
                       Split0::R1 + 1);
    }
    if constexpr (kCheck) {
        commit_position(A, n, sm, c, [&]() {
            if (next_nn >= 0) copy_rows0(A, next_nn, base2 + 4, Split0::P2_R0, WY);
        });
        pre = next_nn >= 0 ? 1 : 0;
        if (pre) pslot ^= 1;
    } else if constexpr (CLS == 0) {
        reduce_commit<CLS>(A, n, sm, c, sm.csraw[nxt ^ 1], aV, aCx, aCy, aPer,
                           blox, bhix, bloy, bhiy, pmaxb, [&]() {
            if (next_nn >= 0) copy_rows0(A, next_nn, base2 + 4, Split0::P2_R0, WY);
        });
        pre = next_nn >= 0 ? 1 : 0;
        if (pre) pslot ^= 1;
    } else {
        reduce_commit<CLS>(A, n, sm, c, sm.csraw[nxt ^ 1], aV, aCx, aCy, aPer,
                           blox, bhix, bloy, bhiy, pmaxb);
    }
}

// -------------------------------------------------------------------------
// Global-memory path: class 4 (window zeros synthesized) and the fallback
// class 5 (the tile's own one-pixel zero ring), src/kernels.cu:604-937.
// -------------------------------------------------------------------------
template <int CLS, bool kBatch = false, int kObs = 0>
__device__ void cell_global(const pfh2d_step_args& A, int n,
                            unsigned long long step, Smem& sm, Ctl& c, int nxt) {
    constexpr int WX = kShapes[CLS].wx, WY = kShapes[CLS].wy;
    constexpr int TX0 = kShapes[CLS].tx0, TY0 = kShapes[CLS].ty0;
    const int tid = opaque_tid();
    const Dom dm = cell_dom<kBatch>(A, sm.slot[nxt ^ 1]);
    const int L = dm.L, P = dm.P;
    const uint32_t* S_rd = A.S_rd + dm.off;
    uint32_t* S_sc = A.S_sc ? A.S_sc + dm.off : nullptr;
    if (tid < kControlLanes) fetch_all<kBatch>(A, sm, nxt, step);
    const int gx0 = c.gx0, gy0 = c.gy0;
    const float* tin = A.phi_in + (size_t)n * kTileArea;
    float* tout = A.phi_out + (size_t)n * kTileArea;
    auto pin = [&](int xx, int yy) -> float {
        if (CLS != kClassFallback &&
            !((unsigned)xx < (unsigned)WX && (unsigned)yy < (unsigned)WY))
            return 0.0f;
        return tin[(size_t)(TY0 + yy) * kTilePitch + TX0 + xx];
    };
    if (c.zero_tile) zero_outside_window(tout, c.dwx, c.dwy, c.dtx0, c.dty0);

    double aIx = 0.0, aIy = 0.0;
    uint32_t negc = 0u;
    ObsSums ob;
    for (int p = tid; p < WX * WY; p += kNT) {
        const int y = p / WX, x = p - (p / WX) * WX;
        const float cc = pin(x, y);
        const float gxv = 0.5f * (pin(x + 1, y) - pin(x - 1, y));
        const float gyv = 0.5f * (pin(x, y + 1) - pin(x, y - 1));
        const uint32_t qS = S_rd[(size_t)wrapi(gy0 + y, L) * P + wrapi(gx0 + x, L)];
        const uint32_t qc = q_of(cc);
        float So = 0.0f;
        if (qS >= qc) So = (float)(qS - qc) * kQInv;
        else negc += 1u;
        aIx += (double)(cc * gxv * So);
        aIy += (double)(cc * gyv * So);
        if constexpr (kObs == 2)
            observe_pixel(ob, x, y, cc, pin(x + 1, y), pin(x - 1, y), pin(x, y + 1),
                          pin(x, y - 1), pin(x + 1, y + 1), pin(x - 1, y + 1),
                          pin(x + 1, y - 1), pin(x - 1, y - 1), So, c.gam, c.dwC,
                          A.rep_coeff, c.og2iv, c.ocx, c.ocy);
    }
    reduce_velocity<kObs>(A, n, sm, c, aIx, aIy, ob);

    const float gam = c.gam, dwC = c.dwC, volC = c.volC, repC = A.rep_coeff;
    const float vx = c.vx, vy = c.vy, mrel = c.mrel, dt = A.dt;
    const int sx = c.sx, sy = c.sy, dwx = c.dwx, dwy = c.dwy;
    const int dtx0 = c.dtx0, dty0 = c.dty0;
    for (int p = tid; p < WX * WY; p += kNT) {
        const int y = p / WX, x = p - (p / WX) * WX;
        const float cC = pin(x, y);
        const float cE = pin(x + 1, y), cW = pin(x - 1, y);
        const float nC = pin(x, y + 1), sC = pin(x, y - 1);
        const float nE = pin(x + 1, y + 1), nW = pin(x - 1, y + 1);
        const float sE = pin(x + 1, y - 1), sW = pin(x - 1, y - 1);
        const float gxv = 0.5f * (cE - cW);
        const float gyv = 0.5f * (nC - sC);
        const uint32_t qS = S_rd[(size_t)wrapi(gy0 + y, L) * P + wrapi(gx0 + x, L)];
        const uint32_t qc = q_of(cC);
        float So = 0.0f;
        if (qS >= qc) So = (float)(qS - qc) * kQInv;
        else negc += 1u;
        const float passive = with_repulsion(
            passive_head_of(nC, sC, cE, cW, nE, nW, sE, sW, cC, gam, dwC, volC), repC, cC, So);
        const float pn = euler_update(cC, gxv, gyv, passive, vx, vy, mrel, dt);
        const int a = x - sx, b = y - sy;
        if ((unsigned)a < (unsigned)dwx && (unsigned)b < (unsigned)dwy)
            tout[(size_t)(dty0 + b) * kTilePitch + dtx0 + a] = pn;
    }
    if (negc) flag_add(A.flags, PFH_FLAG_S_NEGATIVE, negc);
    for (int i = tid; i < dwx * dwy; i += kNT) {
        const int b = i / dwx, a = i - (i / dwx) * dwx;
        const int ii = a + sx, jj = b + sy;
        if (!((unsigned)ii < (unsigned)WX && (unsigned)jj < (unsigned)WY))
            tout[(size_t)(dty0 + b) * kTilePitch + dtx0 + a] = 0.0f;
    }
    __threadfence();      // stores reach L2; L1 lines are invalidated
    __syncthreads();

    const int gx0n = c.gx0n, gy0n = c.gy0n, fm = c.fm;
    double aV = 0.0, aCx = 0.0, aCy = 0.0, aPer = 0.0;
    int blox = dwx, bhix = -1, bloy = dwy, bhiy = -1, pmaxb = 0;
    bool nonfin = false, clampf = false, ovf = false;
    for (int i = tid; i < dwx * dwy; i += kNT) {
        const int b = i / dwx, a = i - (i / dwx) * dwx;
        const size_t row = (size_t)(dty0 + b) * kTilePitch + dtx0;
        const float pn = tout[row + a];
        nonfin |= !isfinite(pn);
        clampf |= (pn * pn > kQClampPhiSq);
        pmaxb = max(pmaxb, (int)__float_as_uint(fabsf(pn)));
        const uint32_t q = q_of(pn);
        if (q && S_sc) {
            const uint32_t old = atomicAdd(
                &S_sc[(size_t)wrapi(gy0n + b, L) * P + wrapi(gx0n + a, L)], q);
            ovf |= old > 0xFFFFFFFFu - q;
        }
        const double d = (double)pn * (double)pn;
        aV += d;
        aCx += d * (double)a;
        aCy += d * (double)b;
        if (pn > kSupportEps) {
            blox = min(blox, a); bhix = max(bhix, a);
            bloy = min(bloy, b); bhiy = max(bhiy, b);
        }
        if (fm) {
            const float e = a + 1 < dwx ? tout[row + a + 1] : 0.0f;
            const float w = a > 0 ? tout[row + a - 1] : 0.0f;
            const float nn = b + 1 < dwy ? tout[row + kTilePitch + a] : 0.0f;
            const float ss = b > 0 ? tout[row - kTilePitch + a] : 0.0f;
            const float pgx = 0.5f * (e - w), pgy = 0.5f * (nn - ss);
            aPer += (double)sqrtf(pgx * pgx + pgy * pgy);
        }
    }
    if (nonfin) flag_or(A.flags, PFH_FLAG_NONFINITE);
    if (clampf) flag_or(A.flags, PFH_FLAG_Q_CLAMP);
    if (ovf) flag_or(A.flags, PFH_FLAG_S_OVERFLOW);
    reduce_commit<CLS>(A, n, sm, c, sm.csraw[nxt ^ 1], aV, aCx, aCy, aPer,
                       blox, bhix, bloy, bhiy, pmaxb);
}

// Classes other than the round class 0 run out of line, with their own
// register allocation, so their longer per-thread columns (up to 29 rows) do
// not raise the pressure of the hot class-0 path.
template <int CLS, bool kBatch = false, int kObs = 0>
__device__ __noinline__ void cell_cold(const pfh2d_step_args* A, int n,
                                       unsigned long long step, Smem* sm, int cur) {
    if constexpr (CLS <= kClassStagedMax) {
        int none = 0, slot0 = 0;
        cell_staged<CLS, false, kBatch, kObs>(*A, n, step, *sm, sm->slot[cur].ctl, cur ^ 1,
                                              none, slot0);
    }
    else
        cell_global<CLS, kBatch, kObs>(*A, n, step, *sm, sm->slot[cur].ctl, cur ^ 1);
}

// -------------------------------------------------------------------------
// The persistent step kernel (src/kernels.cu:948-1030).
// -------------------------------------------------------------------------
// Batched form (kBatch): A describes the whole ensemble as one population
// (tiles, cell states, classes and the queue order over every replica's
// cells; each S slot one arena holding every replica's aggregate; the
// clear-ahead sweeps the whole arena), and each cell's side, S pitch, S
// offset and polarity stream come from its replica's record.
template <bool kCheck, bool kBatch = false, int kObs = 0>
__global__ __launch_bounds__(kNT, 1) void k_step(const pfh2d_step_args A, const int* run_if,
                                                 const pfh2d_replica* reps,
                                                 const uint16_t* cell_rep) {
    __shared__ Smem sm;
    const int tid = (int)threadIdx.x;
    if (run_if != nullptr && *run_if == 0) return;   // gated-off launch: no effect

    // Clear S[(step+2)%3] ahead: this slot is neither read nor scattered
    // into by this launch. Non-temporal stores keep it out of the caches.
    if (A.clear_ahead_words > 0ull) {
        u32x4* dst = reinterpret_cast<u32x4*>(A.S_cl);
        const u32x4 zero4 = {0u, 0u, 0u, 0u};
        const size_t nvec = (size_t)(A.clear_ahead_words >> 2);
        const size_t stride = (size_t)gridDim.x * kNT;
        for (size_t i = (size_t)blockIdx.x * kNT + tid; i < nvec; i += stride)
            __builtin_nontemporal_store(zero4, &dst[i]);
        for (size_t i = (nvec << 2) + (size_t)blockIdx.x * kNT + tid;
             i < (size_t)A.clear_ahead_words; i += stride)
            A.S_cl[i] = 0u;
    }
    // Cursor and step slots alternate with the phi parity; the kernel
    // boundary orders the reset and publication for the next launch.
    if (blockIdx.x == 0 && tid == 0) {
        *A.cursor_clear = 0ull;
        *A.step_wr = *A.step_rd + 1ull;
    }
    const unsigned long long step = *A.step_rd;
    if (tid == 0) {
        sm.args = A;
        sm.fm = (A.full_moment_every > 0 &&
                 ((step + 1ull) % (unsigned long long)A.full_moment_every) == 0ull) ? 1 : 0;
        if constexpr (kBatch) {
            sm.reps = reps;
            sm.cell_rep = cell_rep;
        }
    }

    // The control wavefront fetches each next cell one cell ahead (ticket,
    // id, class, CellState, scalars) while the other wavefronts compute, so
    // the queue's global round trips stay off the critical path.
    if (tid < kControlLanes) fetch_all<kBatch>(A, sm, 0, step);
    lds_barrier();
    int cur = 0;
    int pre = 0;                       // the current class-0 window is already in LDS
    int pslot = 0;                     // slot of the class-0 window's piece 1
    int ring = 0;                      // the class-0 layout's zero ring is set up
    for (;;) {
        const int cls = __builtin_amdgcn_readfirstlane(sm.slot[cur].cls);
        const int n = __builtin_amdgcn_readfirstlane(sm.slot[cur].n);
        if (cls < 0) break;
        switch (cls) {
            case 0:
                cell_staged<0, kCheck, kBatch, kObs>(A, n, step, sm, sm.slot[cur].ctl, cur ^ 1,
                                                     pre, pslot, &ring);
                break;
            case 1: cell_cold<1, kBatch, kObs>(&sm.args, n, step, &sm, cur); break;
            case 2: cell_cold<2, kBatch, kObs>(&sm.args, n, step, &sm, cur); break;
            case 3: cell_cold<3, kBatch, kObs>(&sm.args, n, step, &sm, cur); break;
            case 4: cell_cold<4, kBatch, kObs>(&sm.args, n, step, &sm, cur); break;
            case 5: cell_cold<5, kBatch, kObs>(&sm.args, n, step, &sm, cur); break;
            default:                       // unsupported id, flagged at fetch
                if (tid < kControlLanes) fetch_all<kBatch>(A, sm, cur ^ 1, step);
                lds_barrier();
                break;
        }
        if (cls != 0) {
            pre = 0;
            ring = 0;
            // Retire the out-of-line paths' register reloads here, so that the
            // loop head does not wait on them (with everything else in flight).
            __builtin_amdgcn_s_waitcnt(0);
        }
        cur ^= 1;
    }
}

// -------------------------------------------------------------------------
// Initial condition, verification, trajectory, ordering.
//
// Layout shared by these kernels: a tile row is 72 groups of 16 B; work is
// spread over (cell, band of rows) workgroups so that a small population
// still fills the chip, every global access is a 16 B vector or a lane per
// word of one row, and flag words take at most one atomic per wave.
// -------------------------------------------------------------------------
constexpr int kRowGroups = kTilePitch / 4;          // 16 B groups per tile row
constexpr int kInitBandRows = 32;                   // 288 = 9 bands
constexpr int kInitBands = kTilePitch / kInitBandRows;
constexpr int kScatterBandRows = 16;
constexpr int kScatterBands = (286 + kScatterBandRows - 1) / kScatterBandRows;

__device__ __forceinline__ float f4_get(const float4& v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}

// Fill n words with v: 16 B stores on the aligned body, words at both ends.
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t head = (size_t)((0u - (uint32_t)(uintptr_t)p) & 15u) >> 2;
    if (head > n) head = n;
    if (tid < head) p[tid] = v;
    u32x4* body = reinterpret_cast<u32x4*>(p + head);
    const size_t nv = (n - head) >> 2;
    const u32x4 v4 = {v, v, v, v};
    for (size_t i = tid; i < nv; i += stride) body[i] = v4;
    const size_t done = head + 4 * nv;
    if (tid < n - done) p[done + tid] = v;
}

// Smoothed disk of one window pixel (k_init_tiles' profile,
// src/kernels.cu:1072-1085): binary32 distance with minimum-image wrap,
// 0.5 (1 - tanh(k (r - R_eff))).
__device__ __forceinline__ float disk_value(int gx0, int gy0, int a, int b, float cx,
                                            float cy, int L, float radius_eff,
                                            float kappa_iface) {
    const float half = 0.5f * (float)L;
    float dx = (float)(gx0 + a) - cx;
    float dy = (float)(gy0 + b) - cy;
    if (dx > half) dx -= (float)L;
    if (dx < -half) dx += (float)L;
    if (dy > half) dy -= (float)L;
    if (dy < -half) dy += (float)L;
    const float r = sqrtf(dx * dx + dy * dy);
    const float t = (float)tanh_d((double)(kappa_iface * (r - radius_eff)));
    return 0.5f * (1.0f - t);
}

// grid (N, 9): workgroup (n, band) writes tile rows band*32 .. band*32+31 of
// both tiles of cell n in one pass of 16 B stores: the disk inside the class
// window, zeros elsewhere (phi_b all zeros). Band 0 records the origin.
__global__ __launch_bounds__(256) void k_init_tiles(float* phi_a, float* phi_b,
                                                    pfh_cell_state* cell,
                                                    const uint8_t* clsv, int L,
                                                    const float* seed_cx,
                                                    const float* seed_cy,
                                                    float radius_eff, float kappa_iface) {
    const int n = (int)blockIdx.x;
    const int cls = (int)clsv[n];
    const Shape sc = shape_of(cls);
    const float cx = seed_cx[n], cy = seed_cy[n];
    const int gx0 = wrapi((int)rintf(cx - 0.5f * (float)(sc.wx - 1)), L);
    const int gy0 = wrapi((int)rintf(cy - 0.5f * (float)(sc.wy - 1)), L);
    float4* ta = reinterpret_cast<float4*>(phi_a + (size_t)n * kTileArea);
    float4* tb = reinterpret_cast<float4*>(phi_b + (size_t)n * kTileArea);
    const int row0 = (int)blockIdx.y * kInitBandRows;
    const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int g = (int)threadIdx.x; g < kInitBandRows * kRowGroups; g += (int)blockDim.x) {
        const int ty = row0 + g / kRowGroups, gx = g % kRowGroups;
        const int b = ty - sc.ty0;
        float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if ((unsigned)b < (unsigned)sc.wy) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int a = 4 * gx + k - sc.tx0;
                if ((unsigned)a < (unsigned)sc.wx)
                    v[k] = disk_value(gx0, gy0, a, b, cx, cy, L, radius_eff, kappa_iface);
            }
        }
        const size_t o = (size_t)ty * kRowGroups + gx;
        ta[o] = make_float4(v[0], v[1], v[2], v[3]);
        tb[o] = zero4;
    }
    if (blockIdx.y == 0 && threadIdx.x == 0) {
        pfh_cell_state* cs = &cell[n];
        cs->gx0 = gx0;
        cs->gy0 = gy0;
        cs->cls = (uint8_t)cls;
        cs->cls_written[0] = (uint8_t)cls;
        cs->cls_written[1] = (uint8_t)cls;
    }
}

// Moments of the stored field with the update's thread mapping (column bands
// for the LDS classes, flat pixel index for classes 4-5).
__global__ __launch_bounds__(kNT) void k_init_moments(pfh_cell_state* cell,
                                                      const uint8_t* clsv,
                                                      const float* phi, int N) {
    const int n = blockIdx.x;
    if (n >= N) return;
    const int cls = (int)clsv[n];
    const Shape sc = shape_of(cls);
    const float* t = phi + (size_t)n * kTileArea + (size_t)sc.ty0 * kTilePitch + sc.tx0;
    const WindowMoments m =
        window_moments(cls, [&](int a, int b) { return t[(size_t)b * kTilePitch + a]; });
    if (threadIdx.x == 0) store_moments(&cell[n], m);
}

// grid (N, 18): workgroup (n, band) adds q(phi) of window rows band*16 ..
// band*16+15 of cell n into S; wave w takes rows w, w+4, ..., its lanes one
// word of the row each (coalesced loads, atomics only where q != 0). Flag
// conditions are collected per lane and raised by one atomic per wave.
__global__ __launch_bounds__(256) void k_scatter_all(const float* phi,
                                                     const pfh_cell_state* cell,
                                                     const uint8_t* clsv, uint32_t* S,
                                                     int L, int P, uint32_t* flags) {
    const int n = (int)blockIdx.x;
    const Shape sc = shape_of((int)clsv[n]);
    const int r0 = (int)blockIdx.y * kScatterBandRows;
    if (r0 >= sc.wy) return;
    const int r1 = min(r0 + kScatterBandRows, sc.wy);
    const int gx0 = cell[n].gx0, gy0 = cell[n].gy0;
    const int lane = (int)threadIdx.x & (kWave - 1), wave = (int)threadIdx.x / kWave;
    const float* win = phi + (size_t)n * kTileArea + (size_t)sc.ty0 * kTilePitch + sc.tx0;
    uint32_t overflow = 0u, clamp = 0u;
    for (int b = r0 + wave; b < r1; b += (int)blockDim.x / kWave) {
        uint32_t* srow = S + (size_t)wrapi(gy0 + b, L) * (size_t)P;
        const float* prow = win + (size_t)b * kTilePitch;
        for (int a = lane; a < sc.wx; a += kWave) {
            const float pn = prow[a];
            const uint32_t q = q_of(pn);
            if (q) {
                const uint32_t old = __hip_atomic_fetch_add(&srow[wrapi(gx0 + a, L)], q,
                                                            __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT);
                overflow |= old > 0xFFFFFFFFu - q;
                clamp |= pn * pn > kQClampPhiSq;
            }
        }
    }
    const bool any_overflow = __ballot(overflow) != 0ull;
    const bool any_clamp = __ballot(clamp) != 0ull;
    if (lane == 0) {
        if (any_overflow) atomicOr(&flags[PFH_FLAG_S_OVERFLOW], 1u);
        if (any_clamp) atomicOr(&flags[PFH_FLAG_Q_CLAMP], 1u);
    }
}

// Per cell: V of the stored window in the update's reduction order (the
// moments pass), and max |phi| over the tile outside the window (a 16 B sweep
// of the whole tile). Sim::verify compares V with CellState.V.
__global__ __launch_bounds__(kNT) void k_verify_cells(const float* phi,
                                                      const uint8_t* clsv,
                                                      double* out_V,
                                                      float* out_outside_max) {
    __shared__ float wmax[kWaves];
    const int n = (int)blockIdx.x;
    const int cls = (int)clsv[n];
    const Shape sc = shape_of(cls);
    const float* tile = phi + (size_t)n * kTileArea;
    const float4* t4 = reinterpret_cast<const float4*>(tile);
    float omax = 0.0f;
    for (int g = (int)threadIdx.x; g < kTileArea / 4; g += kNT) {
        const int y = g / kRowGroups, x0 = 4 * (g % kRowGroups);
        const float4 v = t4[g];
        const bool row_in = (unsigned)(y - sc.ty0) < (unsigned)sc.wy;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (!(row_in && (unsigned)(x0 + k - sc.tx0) < (unsigned)sc.wx))
                omax = fmaxf(omax, fabsf(f4_get(v, k)));
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) omax = fmaxf(omax, __shfl_xor(omax, d, kWave));
    if (((int)threadIdx.x & (kWave - 1)) == 0) wmax[threadIdx.x / kWave] = omax;
    const float* win = tile + (size_t)sc.ty0 * kTilePitch + sc.tx0;
    const WindowMoments m =
        window_moments(cls, [&](int a, int b) { return win[(size_t)b * kTilePitch + a]; });
    if (threadIdx.x == 0) {
        float o = 0.0f;
        for (int w = 0; w < kWaves; ++w) o = fmaxf(o, wmax[w]);
        out_V[n] = m.V;
        out_outside_max[n] = o;
    }
}

// Plain streaming copy: the known-size reference kernel used to calibrate
// hardware byte counters against bytes actually moved.
__global__ __launch_bounds__(256) void k_stream_copy(u32x4* dst, const u32x4* src,
                                                     size_t nvec) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Max word of S into *out_max: 16 B loads on the aligned body, a per-wave
// maximum, then one atomic per wave that saw a nonzero word.
__global__ __launch_bounds__(256) void k_verify_S(const uint32_t* S, size_t n,
                                                  uint32_t* out_max) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t head = (size_t)((0u - (uint32_t)(uintptr_t)S) & 15u) >> 2;
    if (head > n) head = n;
    uint32_t m = tid < head ? S[tid] : 0u;
    const u32x4* body = reinterpret_cast<const u32x4*>(S + head);
    const size_t nv = (n - head) >> 2;
    for (size_t i = tid; i < nv; i += stride) {
        const u32x4 v = body[i];
        m = max(m, max(max(v.x, v.y), max(v.z, v.w)));
    }
    const size_t done = head + 4 * nv;
    if (tid < n - done) m = max(m, S[done + tid]);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d, kWave));
    if (((int)threadIdx.x & (kWave - 1)) == 0 && m) atomicMax(out_max, m);
}

// Trajectory records (k_pack_traj's fields, src/kernels.cu:1242-1265) of 64
// cells per one-wave workgroup: the 64 CellStates (12 KiB) arrive in LDS by
// 16 B loads, lane c packs cell c into LDS, and the 3 KiB of records leave by
// 16 B stores (write-combined when out is mapped host memory).
constexpr int kCellVecs = (int)(sizeof(pfh_cell_state) / 16);
constexpr int kTrajVecs = (int)(sizeof(pfh_traj_packed_cell) / 16);
static_assert(sizeof(pfh_cell_state) % 16 == 0 && sizeof(pfh_traj_packed_cell) % 16 == 0,
              "16 B record vectors");

__global__ __launch_bounds__(kWave) void k_pack_traj(const pfh_cell_state* cell,
                                                     const uint8_t* clsv,
                                                     pfh_traj_packed_cell* out, int N,
                                                     int L) {
    __shared__ u32x4 in[kWave * kCellVecs];
    __shared__ u32x4 rec[kWave * kTrajVecs];
    const int lane = (int)threadIdx.x;
    const int base = (int)blockIdx.x * kWave;
    const int cnt = min(kWave, N - base);
    const u32x4* src = reinterpret_cast<const u32x4*>(cell + base);
    for (int i = lane; i < cnt * kCellVecs; i += kWave) in[i] = src[i];
    __syncthreads();
    if (lane < cnt) {
        const pfh_cell_state& cs = *reinterpret_cast<const pfh_cell_state*>(&in[lane * kCellVecs]);
        pfh_traj_packed_cell& t = *reinterpret_cast<pfh_traj_packed_cell*>(&rec[lane * kTrajVecs]);
        const double V = cs.V > 0.0 ? cs.V : 1.0;
        const double Ld = (double)L;
        const double cx = (double)cs.gx0 + cs.Cx / V;
        const double cy = (double)cs.gy0 + cs.Cy / V;
        t.global_id = cs.global_id;
        t.cls = (int)clsv[base + lane];
        t.cx = (float)(cx - floor(cx / Ld) * Ld);
        t.cy = (float)(cy - floor(cy / Ld) * Ld);
        t.vx = cs.vx;
        t.vy = cs.vy;
        t.theta = cs.theta;
        t.volume = (float)cs.V;
        t.perim = (float)cs.perim;
        t.gamma = cs.gamma;
        t.v_A = cs.v_A;
        t.phi_max = cs.phi_max;
    }
    __syncthreads();
    u32x4* dst = reinterpret_cast<u32x4*>(out + base);
    for (int i = lane; i < cnt * kTrajVecs; i += kWave) dst[i] = rec[i];
}

// ---- Morton ordering of the cell queue ----
// Key of cell i: 2D Morton code (x in the even bits, y in the odd bits, 16
// bits each) of its truncated centroid, wrapped into the domain
// (src/kernels.cu:1271-1279). The queue is the cells sorted by (key, i).
//
// One workgroup of 1024 threads: a stable least-significant-digit radix sort
// of the indices on 4-bit digits, entirely in LDS. Per pass, each wave ranks
// its 64 elements in their digit bucket with 16 ballots, the per-(digit,
// round, wave) counts are scanned in that order (stable across waves and
// rounds of 1024 elements), and the indices move to their bucket positions.
// Only as many passes run as the domain's coordinates have bits.
constexpr int kMortonMax = 16384;
constexpr int kMortonRounds = kMortonMax / kNT;
constexpr int kDigitBits = 4;
constexpr int kBuckets = 1 << kDigitBits;
constexpr int kCountSlots = kBuckets * kMortonRounds * kWaves;   // 4096

__device__ __forceinline__ uint32_t spread_bits16(uint32_t v) {
    // bit i of v -> bit 2i of the result (v < 2^16)
    uint32_t r = 0u;
#pragma unroll
    for (int i = 0; i < 16; ++i) r |= ((v >> i) & 1u) << (2 * i);
    return r;
}

__global__ __launch_bounds__(kNT) void k_morton_sort(const pfh_cell_state* cell,
                                                     uint32_t* perm, int N, int L,
                                                     int passes) {
    __shared__ uint32_t key[kMortonMax];
    __shared__ uint16_t idx[2][kMortonMax];
    __shared__ uint32_t cnt[kCountSlots];
    __shared__ uint32_t wsum[kWaves];
    const int tid = (int)threadIdx.x;
    const int lane = tid & (kWave - 1), wave = tid / kWave;
    const int rounds = (N + kNT - 1) / kNT;
    for (int i = tid; i < N; i += kNT) {
        const pfh_cell_state& cs = cell[i];
        const double V = cs.V > 0.0 ? cs.V : 1.0;
        const uint32_t x = (uint32_t)wrapi(cs.gx0 + (int)(cs.Cx / V), L);
        const uint32_t y = (uint32_t)wrapi(cs.gy0 + (int)(cs.Cy / V), L);
        key[i] = spread_bits16(x & 0xFFFFu) | (spread_bits16(y & 0xFFFFu) << 1);
        idx[0][i] = (uint16_t)i;
    }
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    int cur = 0;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = pass * kDigitBits;
        uint32_t digit[kMortonRounds], rank[kMortonRounds];
        // 1. per-wave digit counts and ranks within the wave
#pragma unroll
        for (int r = 0; r < kMortonRounds; ++r) {
            if (r >= rounds) break;
            const int i = r * kNT + tid;
            const uint32_t d = i < N ? (key[idx[cur][i]] >> shift) & (kBuckets - 1)
                                     : (uint32_t)kBuckets;
            uint64_t mine = 0ull;
            uint32_t c = 0u;
#pragma unroll
            for (int k = 0; k < kBuckets; ++k) {
                const uint64_t m = __ballot(d == (uint32_t)k);
                if (d == (uint32_t)k) mine = m;
                if (lane == k) c = (uint32_t)__popcll(m);
            }
            digit[r] = d;
            rank[r] = (uint32_t)__popcll(mine & lt_mask);
            if (lane < kBuckets) cnt[(lane * rounds + r) * kWaves + wave] = c;
        }
        __syncthreads();
        // 2. exclusive scan of the counts in (digit, round, wave) order
        const int slots = kBuckets * rounds * kWaves;
        const int per = (slots + kNT - 1) / kNT;
        uint32_t local = 0u;
        for (int k = 0; k < per; ++k) {
            const int s = tid * per + k;
            if (s < slots) local += cnt[s];
        }
        uint32_t incl = local;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)incl, d, kWave);
            if (lane >= d) incl += o;
        }
        if (lane == kWave - 1) wsum[wave] = incl;
        __syncthreads();
        uint32_t run = incl - local;
        for (int w = 0; w < wave; ++w) run += wsum[w];
        for (int k = 0; k < per; ++k) {
            const int s = tid * per + k;
            if (s < slots) {
                const uint32_t c = cnt[s];
                cnt[s] = run;
                run += c;
            }
        }
        __syncthreads();
        // 3. stable moves into the other index buffer
#pragma unroll
        for (int r = 0; r < kMortonRounds; ++r) {
            if (r >= rounds) break;
            const int i = r * kNT + tid;
            if (i < N)
                idx[cur ^ 1][cnt[(digit[r] * rounds + r) * kWaves + wave] + rank[r]] =
                    idx[cur][i];
        }
        __syncthreads();
        cur ^= 1;
    }
    for (int i = tid; i < N; i += kNT) perm[i] = idx[cur][i];
}

// -------------------------------------------------------------------------
// Independent dense velocity observer (src/velocity_moments.cu:43-128): one
// 256-thread workgroup per cell, straight from the step's input field and
// aggregate in global memory (no window staging, no shift, no class
// hysteresis: none of the step's code paths), two passes over the window:
// the interaction force, then the centroid rates of the four velocity
// sources (passive interfacial, overlap, area; advection) in fp64.
// Lane-strided sums, 64-lane trees, waves in ascending order.
// -------------------------------------------------------------------------
constexpr int kObsNT = 256, kObsWaves = kObsNT / kWave;

template <int K>
__device__ void obs_block_sum(double (&v)[K], double* sh) {
    const int lane = (int)threadIdx.x & (kWave - 1), wave = (int)threadIdx.x / kWave;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        v[j] = wave_sum(v[j]);
        if (lane == 0) sh[j * kObsWaves + wave] = v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double t = 0.0;
            for (int w = 0; w < kObsWaves; ++w) t += sh[j * kObsWaves + w];
            sh[j * kObsWaves] = t;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kObsNT) void k_velocity_observe(const pfh2d_step_args A,
                                                             pfh_moment_accum* out,
                                                             double pdt) {
    __shared__ double sh[8 * kObsWaves];
    __shared__ double vel[4];
    const int n = (int)blockIdx.x;
    const pfh_cell_state& c = A.cell[n];
    const Shape sc = shape_of((int)c.cls);
    const int L = A.L, P = A.P, gx0 = c.gx0, gy0 = c.gy0;
    const float* tile = A.phi_in + (size_t)n * kTileArea;
    auto at = [&](int x, int y) -> float {
        return ((unsigned)x < (unsigned)sc.wx && (unsigned)y < (unsigned)sc.wy)
            ? tile[(size_t)(sc.ty0 + y) * kTilePitch + sc.tx0 + x] : 0.0f;
    };
    // S_other at a window pixel; an invalid aggregate (S below the cell's own
    // share) gives NaN, so the sample is refused downstream
    auto other = [&](int x, int y, float p) -> float {
        const uint32_t all = A.S_rd[(size_t)wrap_once(gy0 + y, L) * P + wrap_once(gx0 + x, L)];
        const uint32_t self = q_of(p);
        return all >= self ? (float)(all - self) * kQInv : __int_as_float(0x7fc00000);
    };
    const int W = sc.wx * sc.wy;
    double f[2] = {0.0, 0.0};
    for (int k = (int)threadIdx.x; k < W; k += kObsNT) {
        const int y = k / sc.wx, x = k - y * sc.wx;
        const float p = at(x, y);
        const float gx = 0.5f * (at(x + 1, y) - at(x - 1, y));
        const float gy = 0.5f * (at(x, y + 1) - at(x, y - 1));
        const float so = other(x, y, p);
        // the same binary32 product as the step's interaction integrand
        f[0] += (double)(p * gx * so);
        f[1] += (double)(p * gy * so);
    }
    obs_block_sum(f, sh);
    if (threadIdx.x == 0) {
        const unsigned long long step = *A.step_rd;
        const Philox4 r = philox4x32_10(
            (uint32_t)(step & 0xFFFFFFFFull), (uint32_t)(step >> 32), (uint32_t)c.global_id,
            0xA5A5A5A5u, (uint32_t)(A.polarity_seed & 0xFFFFFFFFull),
            (uint32_t)(A.polarity_seed >> 32));
        float theta = c.theta;
        if (uniform53(r.v[0], r.v[1]) < A.p_tumble)
            theta = (float)(2.0 * kPi * uniform53(r.v[2], r.v[3]));
        double sd, cd;
        sincos_d((double)theta, &sd, &cd);
        vel[0] = (double)c.v_A * (double)(float)cd;
        vel[1] = (double)c.v_A * (double)(float)sd;
        vel[2] = (double)A.mot_coeff * sh[0];
        vel[3] = (double)A.mot_coeff * sh[kObsWaves];
    }
    __syncthreads();
    const double cx = c.Cx / c.V, cy = c.Cy / c.V;
    const double dw = (double)(A.bulk_scale * c.gamma);
    const double vol = (double)(float)(A.vol_scale * (A.A0 - c.V));
    const double mob = (double)(c.M_pf / 0.5f);
    // the advection uses the stored (binary32) velocity
    const double vx = (double)(float)(vel[0] + vel[2]);
    const double vy = (double)(float)(vel[1] + vel[3]);
    double m[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = (int)threadIdx.x; k < W; k += kObsNT) {
        const int y = k / sc.wx, x = k - y * sc.wx;
        const double p = (double)at(x, y);
        const double e = at(x + 1, y), w = at(x - 1, y);
        const double nn = at(x, y + 1), ss = at(x, y - 1);
        const double ne = at(x + 1, y + 1), nw = at(x - 1, y + 1);
        const double se = at(x + 1, y - 1), sw = at(x - 1, y - 1);
        const double lap =
            (((((4.0 * (((nn + ss) + e) + w) + ne) + nw) + se) + sw) - 20.0 * p) / 6.0;
        const double so = (double)other(x, y, (float)p);
        double src[4] = {(double)c.gamma * lap - ((dw * p) * (1.0 - p)) * (1.0 - 2.0 * p),
                         -((double)A.rep_coeff * p) * so, vol * p,
                         -(vx * (0.5 * (e - w)) + vy * (0.5 * (nn - ss)))};
        // the mobility scales the passive sources, never the advection
        if (mob != 1.0) {
            src[0] *= mob;
            src[1] *= mob;
            src[2] *= mob;
        }
        const double wxp = ((2.0 * ((double)x - cx)) * p) / c.V;
        const double wyp = ((2.0 * ((double)y - cy)) * p) / c.V;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            m[2 * j] += wxp * src[j];
            m[2 * j + 1] += wyp * src[j];
        }
    }
    obs_block_sum(m, sh);
    if (threadIdx.x == 0) {
        pfh_moment_accum& z = out[n];
        for (int j = 0; j < 4; ++j) z.q[j] += pdt * vel[j];
        for (int j = 0; j < 6; ++j) z.q[4 + j] += pdt * sh[j * kObsWaves];
        z.q[10] += pdt * ((sh[6 * kObsWaves] - vel[0]) - vel[2]);
        z.q[11] += pdt * ((sh[7 * kObsWaves] - vel[1]) - vel[3]);
        ++z.count;
    }
}

}  // namespace pfh

// -------------------------------------------------------------------------
// C ABI
// -------------------------------------------------------------------------
namespace {
thread_local char g_last_error[256] = {0};
}  // namespace

namespace pfh_host {
int hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return PFH_OK;
    snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what,
             hipGetErrorString(e));
    return PFH_ERR_HIP;
}
int launch_status(const char* what) {
    return hip_status(hipGetLastError(), what);
}
int invalid(const char* what) {
    snprintf(g_last_error, sizeof(g_last_error), "invalid argument: %s", what);
    return PFH_ERR_INVALID;
}
int unsupported(const char* what) {
    snprintf(g_last_error, sizeof(g_last_error), "%s", what);
    return PFH_ERR_UNSUPPORTED;
}
}  // namespace pfh_host
using namespace pfh_host;

extern "C" {

int pfh_abi_version(void) { return 10000; }

const char* pfh_last_error(void) { return g_last_error; }

int pfh2d_configure_step(void) {
    // k_step uses static LDS only; nothing to opt into on gfx950.
    return PFH_OK;
}

int pfh2d_configure_morton(int smem_bytes) {
    // The sort keeps its keys and index buffers in static LDS (sized for
    // 16384 cells); there is nothing to opt into on gfx950.
    if (smem_bytes <= 0 || smem_bytes > 160 * 1024) return invalid("morton smem");
    return PFH_OK;
}

int pfh2d_step_grid(int device) {
    int cus = 0;
    const int st = hip_status(
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device),
        "hipDeviceGetAttribute(multiProcessorCount)");
    if (st != PFH_OK) return st;
    return cus > 0 ? cus : 1;
}

namespace {
int step_launch(const pfh2d_step_args* a, const int* run_if, int grid, void* stream,
                bool check = false, int observation = PFH_OBSERVE_DISABLED) {
    if (!a) return invalid("args");
    if (observation == PFH_OBSERVE_SPATIAL && !a->moments)
        return invalid("a spatial observation needs the moment accumulators (moments)");
    if (grid <= 0 || a->N < 0 || a->L <= 0 || a->P < a->L || !a->phi_in ||
        !a->phi_out || !a->S_rd || !a->cell || !a->cell_cls ||
        !a->perm || !a->cursor_use || !a->cursor_clear || !a->step_rd ||
        !a->step_wr || !a->flags || (a->clear_ahead_words && !a->S_cl) ||
        a->parity_out < 0 || a->parity_out > 1)
        return invalid("step args");
    // Every window must fit strictly inside the periodic domain
    // (include/params.cuh:666-680), or S self-subtraction is not exact.
    if (a->L <= 286) return invalid("domain side must exceed 286");
    // the step addresses S with 32-bit byte offsets
    if (4ull * (unsigned long long)a->P * (unsigned long long)a->L > 0xFFFFFFFFull)
        return invalid("aggregate larger than 4 GiB");
    const pfh2d_replica* no_reps = nullptr;
    const uint16_t* no_map = nullptr;
    const dim3 g((unsigned)grid), b(pfh::kNT);
    if (check)
        hipLaunchKernelGGL((pfh::k_step<true, false>), g, b, 0, as_stream(stream), *a, run_if,
                           no_reps, no_map);
    else if (observation == PFH_OBSERVE_ADVECTION)
        hipLaunchKernelGGL((pfh::k_step<false, false, 1>), g, b, 0, as_stream(stream), *a,
                           run_if, no_reps, no_map);
    else if (observation == PFH_OBSERVE_SPATIAL)
        hipLaunchKernelGGL((pfh::k_step<false, false, 2>), g, b, 0, as_stream(stream), *a,
                           run_if, no_reps, no_map);
    else
        hipLaunchKernelGGL((pfh::k_step<false, false>), g, b, 0, as_stream(stream), *a, run_if,
                           no_reps, no_map);
    return launch_status("k_step");
}
}  // namespace

int pfh2d_launch_step(const pfh2d_step_args* a, int grid, void* stream,
                      const void* l2_base, size_t l2_bytes, float l2_hit_ratio,
                      int observation) {
    (void)l2_base; (void)l2_bytes; (void)l2_hit_ratio;
    if (!a) return invalid("args");
    if (observation < PFH_OBSERVE_DISABLED || observation > PFH_OBSERVE_SPATIAL)
        return invalid("observation is 0 (disabled), 1 (advection) or 2 (spatial)");
    return step_launch(a, nullptr, grid, stream, false, observation);
}

int pfh2d_velocity_moments(const pfh2d_step_args* a, pfh_moment_accum* out,
                           double physical_dt, void* stream) {
    if (!a || !out) return invalid("velocity_moments: args and out are required");
    if (a->N <= 0 || a->L <= 286 || a->P < a->L || !a->phi_in || !a->S_rd || !a->cell ||
        !a->step_rd)
        return invalid("velocity_moments: step args");
    hipLaunchKernelGGL(pfh::k_velocity_observe, dim3((unsigned)a->N), dim3(pfh::kObsNT), 0,
                       as_stream(stream), *a, out, physical_dt);
    return launch_status("k_velocity_observe");
}

int pfh2d_launch_step_if(const pfh2d_step_args* a, const int* run_if, int grid,
                         void* stream) {
    return step_launch(a, run_if, grid, stream);
}

int pfh2d_launch_check_step_if(const pfh2d_step_args* a, const int* run_if, int grid,
                               void* stream) {
    if (a && a->S_sc) return invalid("a check step scatters nothing: S_sc must be NULL");
    return step_launch(a, run_if, grid, stream, true);
}

int pfh2d_launch_batch_step(const pfh2d_step_args* a, const pfh2d_replica* replicas,
                            const uint16_t* cell_replica, int num_replicas, int max_side,
                            int max_pitch, int grid, void* stream) {
    if (!a || !replicas || !cell_replica) return invalid("batch step: null argument");
    if (num_replicas <= 0 || num_replicas > 65536 || grid <= 0 || a->N <= 0 ||
        !a->phi_in || !a->phi_out || !a->S_rd || !a->S_sc || !a->cell || !a->cell_cls ||
        !a->perm || !a->cursor_use || !a->cursor_clear || !a->step_rd || !a->step_wr ||
        !a->flags || (a->clear_ahead_words && !a->S_cl) || a->parity_out < 0 ||
        a->parity_out > 1)
        return invalid("batch step args");
    // every replica's side must exceed the classes it can use: 256 is the
    // smallest side (class 4, 224 wide, still fits; the 286-wide fallback
    // class is then unavailable and a cell that needs it is flagged)
    if (max_side < 256 || max_pitch < max_side)
        return invalid("batch step: replica sides must be at least 256");
    if (4ull * (unsigned long long)max_pitch * (unsigned long long)max_side > 0xFFFFFFFFull)
        return invalid("aggregate larger than 4 GiB");
    hipLaunchKernelGGL((pfh::k_step<false, true>), dim3((unsigned)grid), dim3(pfh::kNT), 0,
                       as_stream(stream), *a, (const int*)nullptr, replicas, cell_replica);
    return launch_status("k_step (batched)");
}

int pfh2d_query_step_stats(int device, pfh_kernel_stats* out) {
    if (!out) return invalid("out");
    hipFuncAttributes fa{};
    int st = hip_status(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(pfh::k_step<false, false>)),
                        "hipFuncGetAttributes(k_step)");
    if (st != PFH_OK) return st;
    int blocks = 0;
    st = hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &blocks, reinterpret_cast<const void*>(pfh::k_step<false, false>), pfh::kNT, 0),
                    "hipOccupancyMaxActiveBlocksPerMultiprocessor(k_step)");
    if (st != PFH_OK) return st;
    int max_threads = 0;
    if (hipDeviceGetAttribute(&max_threads, hipDeviceAttributeMaxThreadsPerMultiProcessor,
                              device) != hipSuccess || max_threads <= 0)
        max_threads = 2048;
    memset(out, 0, sizeof(*out));
    out->regs = fa.numRegs;
    out->local_bytes = fa.localSizeBytes;
    out->static_smem = (int)fa.sharedSizeBytes;
    out->dynamic_smem = 0;
    out->ctas_per_sm = blocks;
    out->warps_per_sm = blocks * pfh::kNT / pfh::kWave;
    out->occupancy = (double)(blocks * pfh::kNT) / (double)max_threads;
    return PFH_OK;
}

int pfh2d_init_tiles(float* phi_a, float* phi_b, pfh_cell_state* cell,
                     const uint8_t* cls, int N, int L, const float* seed_cx,
                     const float* seed_cy, float radius_eff, float kappa_iface,
                     void* stream) {
    if (N < 0 || L <= 0 || (N > 0 && (!phi_a || !phi_b || !cell || !cls ||
                                      !seed_cx || !seed_cy)))
        return invalid("init_tiles");
    if (N == 0) return PFH_OK;
    if ((uintptr_t)phi_a % 16 || (uintptr_t)phi_b % 16)
        return invalid("init_tiles: tiles must be 16 B aligned");
    hipLaunchKernelGGL(pfh::k_init_tiles, dim3((unsigned)N, pfh::kInitBands), dim3(256), 0,
                       as_stream(stream), phi_a, phi_b, cell, cls, L, seed_cx, seed_cy,
                       radius_eff, kappa_iface);
    return launch_status("k_init_tiles");
}

int pfh2d_init_moments(pfh_cell_state* cell, const uint8_t* cls,
                       const float* phi, int N, void* stream) {
    if (N < 0 || (N > 0 && (!cell || !cls || !phi))) return invalid("init_moments");
    if (N == 0) return PFH_OK;
    hipLaunchKernelGGL(pfh::k_init_moments, dim3(N), dim3(pfh::kNT), 0,
                       as_stream(stream), cell, cls, phi, N);
    return launch_status("k_init_moments");
}

int pfh2d_scatter_all(const float* phi, const pfh_cell_state* cell,
                      const uint8_t* cls, uint32_t* S, int N, int L, int P,
                      uint32_t* flags, void* stream) {
    if (N < 0 || L <= 0 || P < L ||
        (N > 0 && (!phi || !cell || !cls || !S || !flags)))
        return invalid("scatter_all");
    if (N == 0) return PFH_OK;
    hipLaunchKernelGGL(pfh::k_scatter_all, dim3((unsigned)N, pfh::kScatterBands), dim3(256), 0,
                       as_stream(stream), phi, cell, cls, S, L, P, flags);
    return launch_status("k_scatter_all");
}

int pfh2d_zero_u32(uint32_t* p, size_t n, void* stream) {
    if (n == 0) return PFH_OK;
    if (!p) return invalid("zero_u32");
    size_t blocks = (n / 4 + 255) / 256 + 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(pfh::k_fill_u32, dim3((unsigned)blocks), dim3(256), 0,
                       as_stream(stream), p, n, 0u);
    return launch_status("k_fill_u32");
}

int pfh2d_verify_cells(const float* phi, const pfh_cell_state* cell,
                       const uint8_t* cls, int N, double* out_V,
                       float* out_outside_max, void* stream) {
    if (N < 0 || (N > 0 && (!phi || !cell || !cls || !out_V || !out_outside_max)))
        return invalid("verify_cells");
    if (N == 0) return PFH_OK;
    if ((uintptr_t)phi % 16) return invalid("verify_cells: tiles must be 16 B aligned");
    hipLaunchKernelGGL(pfh::k_verify_cells, dim3((unsigned)N), dim3(pfh::kNT), 0,
                       as_stream(stream), phi, cls, out_V, out_outside_max);
    return launch_status("k_verify_cells");
}

int pfh2d_verify_S(const uint32_t* S, size_t n, uint32_t* out_max, void* stream) {
    if (!out_max || (n > 0 && !S)) return invalid("verify_S");
    if (n == 0) return PFH_OK;
    size_t blocks = (n / 4 + 255) / 256 + 1;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(pfh::k_verify_S, dim3((unsigned)blocks), dim3(256), 0,
                       as_stream(stream), S, n, out_max);
    return launch_status("k_verify_S");
}

int pfh_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return PFH_OK;
    if (!dst || !src || (bytes & 15u) || ((uintptr_t)dst & 15u) || ((uintptr_t)src & 15u))
        return invalid("stream_copy needs 16 B aligned pointers and sizes");
    const size_t nvec = bytes / 16;
    size_t blocks = (nvec + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(pfh::k_stream_copy, dim3((unsigned)blocks), dim3(256), 0,
                       as_stream(stream), reinterpret_cast<pfh::u32x4*>(dst),
                       reinterpret_cast<const pfh::u32x4*>(src), nvec);
    return launch_status("k_stream_copy");
}

int pfh2d_pack_traj(const pfh_cell_state* cell, const uint8_t* cls,
                    pfh_traj_packed_cell* out, int N, int L, void* stream) {
    if (N < 0 || L <= 0 || (N > 0 && (!cell || !cls || !out))) return invalid("pack_traj");
    if (N == 0) return PFH_OK;
    if ((uintptr_t)out % 16 || (uintptr_t)cell % 16)
        return invalid("pack_traj: cell and out must be 16 B aligned");
    hipLaunchKernelGGL(pfh::k_pack_traj, dim3((unsigned)((N + pfh::kWave - 1) / pfh::kWave)),
                       dim3(pfh::kWave), 0, as_stream(stream), cell, cls, out, N, L);
    return launch_status("k_pack_traj");
}

int pfh2d_morton_sort(const pfh_cell_state* cell, uint32_t* perm, int N, int M,
                      int L, void* stream) {
    if (N <= 0 || M < N || (M & (M - 1)) != 0 || M > pfh::kMortonMax || L <= 0 || !cell ||
        !perm)
        return invalid("morton_sort");
    // Digit passes needed for the key bits the domain can produce: both
    // coordinates are below min(L, 2^16).
    int coord_bits = 0;
    while (coord_bits < 16 && ((L - 1) >> coord_bits) != 0) ++coord_bits;
    const int passes = (2 * coord_bits + pfh::kDigitBits - 1) / pfh::kDigitBits;
    hipLaunchKernelGGL(pfh::k_morton_sort, dim3(1), dim3(pfh::kNT), 0, as_stream(stream),
                       cell, perm, N, L, passes);
    return launch_status("k_morton_sort");
}

}  // extern "C"
