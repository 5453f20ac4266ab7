// 3D per-cell phase-field step on gfx950 (periodic cube, cubic bricks).
//
// Reference behaviour (file:line):
//   measurement + commit        src/pf3d/kernels.cu:427-496,1359-1409,1591-1737
//   motion / tumble             src/pf3d/kernels.cu:1739-1788, include/pf3d/rng.cuh
//   moment-free update          src/pf3d/kernels.cu:498-793,
//                               src/pf3d/kernels_periodic_slab_walk.cu:54-318
//   repair/finalize/restore/
//   advance                     src/pf3d/kernels.cu:2128-2263
//   step sequence               src/pf3d/sim.cu:1460-1651
//
// Layout in HBM: phi is N bricks of B^3 binary32, x fastest (the reference's
// checkpointed layout); the aggregate S is L*L planes of P uint32 words.
//
// gfx950 mapping. A brick plane of B*B voxels is B*B/4 sixteen-byte x
// groups; one lane owns one group (B = 48: 576 lanes = 9 waves), so every
// global access of a plane is one coalesced 16 B load or store per lane.
//   k3_measure  grid N*shards, shard = a contiguous z range. Planes roll
//               through registers (z neighbours) and a two-buffer LDS ring
//               (x/y neighbours). Per-lane fp64 sums in z-then-x order, a
//               64-lane shuffle tree, waves in ascending order; the last
//               shard of a cell to arrive adds the shard partials in
//               ascending order and commits the cell (centroid shift,
//               support check) and its motion. The same launch zeroes the
//               next aggregate buffer and snapshots the accepted cell state.
//   k3_update   grid N*slabs*parts, slab = a contiguous range of
//               destination planes, part = a contiguous range of their rows
//               (two parts from B = 32). Source planes (the part's rows and
//               one halo row on each side) stream through a two-buffer LDS ring
//               with a zero margin of five, so the recentring shift (|s| <= 4)
//               reads zeros outside the brick without branches. Each plane
//               is reduced once to per-voxel plane sums (centre/edge/corner
//               weights of the 27-point stencil) held in registers, and a
//               destination plane is finished when its upper neighbour
//               plane arrives. Every destination voxel is written exactly
//               once (no separate clear).
//   k3_finish   repair (copy) on a rejected step, origin update, state
//               restore and the step counters in one launch.
// Binary32 arithmetic follows source order without contraction
// (-ffp-contract=off); the CPU oracle reproduces every sum bit for bit.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <random>
#include <vector>

#include "../../include/pfh.h"
#include "pfh_device.h"
#include "pfh_host.h"

namespace pfh3 {

using pfh::philox4x32_10;
using pfh::Philox4;
using pfh::q_of;
using pfh::sincos_d;
using pfh::tanh_d;
using pfh::uniform53;

constexpr uint32_t kFatal = PFH3D_FATAL_MASK;
constexpr float kEps = 1.0e-5f;
constexpr int kMargin = 5;        // LDS zero margin of the update ring (shift 4 + 1)
constexpr int kMarginG = 2;       // the same in x, in 16 B groups (x >= -8)

// Update ring rows: four sub-rows (one per x mod 4) of B/4 + 2 kMarginG words
// each, indexed by the group x >> 2. The pitch is the smallest one that holds
// them and is congruent to B/4 modulo 64, so that the lanes of a wave, B/4
// per row, address consecutive banks across rows too.
__host__ __device__ inline int update_row_pitch(int B) {
    const int span = 4 * (B / 4 + 2 * kMarginG);
    return span + ((B / 4 - span) % 64 + 64) % 64;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

#define PFH_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ PFH_GLOBAL T* gptr(T* p) {
    return (PFH_GLOBAL T*)p;
}

__device__ __forceinline__ int wrap_origin(int64_t o, int n) {
    if (o >= 0 && o < n) return (int)o;
    int64_t v = o % (int64_t)n;
    if (v < 0) v += n;
    return (int)v;
}
__device__ __forceinline__ int wrap_off(int o, int off, int n) {
    int v = o + off;
    if (v >= n) v -= n;
    return v;
}
__device__ __forceinline__ size_t s_index(int L, int P, int x, int y, int z) {
    return ((size_t)z * (size_t)L + (size_t)y) * (size_t)P + (size_t)x;
}

__device__ __forceinline__ void raise3(pfh3d_cell_state* c, uint32_t* flags, int f) {
    atomicOr(&c->flags, 1u << f);
    atomicAdd(&flags[f], 1u);
}

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool global_fatal(const uint32_t* flags) {
    bool f = false;
#pragma unroll
    for (int i = 0; i < PFH3D_FLAG_COUNT; ++i)
        if ((kFatal >> i) & 1u) f = f || ld_relaxed(&flags[i]) != 0u;
    return f;
}
__device__ __forceinline__ bool fatal_present(const pfh3d_cell_state* c,
                                              const uint32_t* flags) {
    return (ld_relaxed(&c->flags) & kFatal) != 0u || global_fatal(flags);
}

// other_field, src/pf3d/device_common.cuh:182-191
__device__ __forceinline__ float other_field(uint32_t agg, float phi,
                                             pfh3d_cell_state* c, uint32_t* flags) {
    const uint32_t own = q_of(phi);
    if (agg >= own) return (float)(agg - own) * pfh::kQInv;
    raise3(c, flags, PFH3D_FLAG_S_NEGATIVE);
    return 0.0f;
}

// scatter_value, src/pf3d/device_common.cuh:193-205
__device__ __forceinline__ void scatter_value(uint32_t* S, size_t idx, float phi,
                                              pfh3d_cell_state* c, uint32_t* flags) {
    const float sq = phi * phi;
    if (sq > pfh::kQClampPhiSq) raise3(c, flags, PFH3D_FLAG_Q_CLAMP);
    const uint32_t q = q_of(phi);
    if (q == 0u) return;
    const uint32_t old = __hip_atomic_fetch_add(gptr(S + idx), q, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    if (old > 0xffffffffu - q) raise3(c, flags, PFH3D_FLAG_S_OVERFLOW);
}

// ---------------------------------------------------------------------------
// commit + motion (one lane of the last-arriving shard)
// ---------------------------------------------------------------------------
struct Acc { double V, Cx, Cy, Cz, surf, Ix, Iy, Iz; };
struct Box { int lo[3], hi[3]; uint32_t pmax; };

__device__ __forceinline__ int clamp_shift(int p, int limit, int lo, int hi, int B) {
    p = max(-limit, min(limit, p));
    if (hi >= lo) {
        p = min(p, lo);
        p = max(p, hi - (B - 1));
    }
    return p;
}

// commit_measurement, src/pf3d/kernels.cu:427-496
struct CommitArgs {
    uint32_t* flags;
    int B, max_shift, support_margin;
};

__device__ __noinline__ void commit_cell(pfh3d_cell_state* s, const CommitArgs a, Acc t,
                                         Box b, int surface) {
    const int B = a.B;
    s->V = t.V; s->Cx = t.Cx; s->Cy = t.Cy; s->Cz = t.Cz;
    if (surface) s->surface = t.surf;
    s->Ix = t.Ix; s->Iy = t.Iy; s->Iz = t.Iz;
    s->bb_lo_x = b.lo[0]; s->bb_hi_x = b.hi[0];
    s->bb_lo_y = b.lo[1]; s->bb_hi_y = b.hi[1];
    s->bb_lo_z = b.lo[2]; s->bb_hi_z = b.hi[2];
    s->phi_max = __uint_as_float(b.pmax);
    int sh[3] = {0, 0, 0};
    if (t.V > 0.0 && isfinite(t.V)) {
        const double middle = 0.5 * (double)(B - 1);
        const double c[3] = {t.Cx, t.Cy, t.Cz};
        for (int i = 0; i < 3; ++i) {
            sh[i] = (int)rint(c[i] / t.V - middle);
            sh[i] = clamp_shift(sh[i], a.max_shift, b.lo[i], b.hi[i], B);
        }
    } else {
        raise3(s, a.flags, PFH3D_FLAG_V_NONPOSITIVE);
    }
    s->pending_shift_x = sh[0];
    s->pending_shift_y = sh[1];
    s->pending_shift_z = sh[2];
    if (b.hi[0] >= b.lo[0]) {
        const int lm = a.support_margin / 2;
        const int um = a.support_margin - lm;
        bool bad = false;
        for (int i = 0; i < 3; ++i) {
            const int ext = b.hi[i] - b.lo[i] + 1;
            bad = bad || ext + a.support_margin > B || b.lo[i] - sh[i] < lm ||
                  b.hi[i] - sh[i] >= B - um;
        }
        if (bad) raise3(s, a.flags, PFH3D_FLAG_SUPPORT_EXHAUSTED);
    }
}

// isotropic_unit_vector / tumble_sample, include/pf3d/rng.cuh:19-139
__device__ __forceinline__ void isotropic_unit(double au, double zu, double out[3]) {
    const double z = 2.0 * au - 1.0;
    const double rsq = 1.0 - z * z;
    const double radial = sqrt(rsq > 0.0 ? rsq : 0.0);
    const double az = 2.0 * pfh::kPi * zu;
    double s, c;
    sincos_d(az, &s, &c);
    out[0] = radial * c;
    out[1] = radial * s;
    out[2] = z;
}

// k_apply_cell_motion, src/pf3d/kernels.cu:1739-1788
struct MotionArgs {
    uint32_t* flags;
    const uint64_t* counters;
    uint64_t aging_steps, polarity_stream;
    double p_tumble;
    float motility_coeff;
};

__device__ __noinline__ void motion_cell(pfh3d_cell_state* s, const MotionArgs a,
                                         int apply_tumble) {
    if (fatal_present(s, a.flags)) return;
    const uint64_t step = a.counters ? a.counters[0] : 0u;
    const bool active = step >= a.aging_steps;
    double pol[3] = {s->polarity_x, s->polarity_y, s->polarity_z};
    bool tumbled = false;
    if (apply_tumble && active) {
        const uint64_t st = step - a.aging_steps;
        const uint64_t g = (uint64_t)s->global_id;
        const uint32_t k0 = (uint32_t)a.polarity_stream;
        const uint32_t k1 = (uint32_t)(a.polarity_stream >> 32);
        const Philox4 ev = philox4x32_10((uint32_t)st, (uint32_t)(st >> 32), (uint32_t)g,
                                         (uint32_t)(g >> 32) ^ 0x54554D45u, k0, k1);
        if (uniform53(ev.v[0], ev.v[1]) < a.p_tumble) {
            const Philox4 ve = philox4x32_10((uint32_t)st, (uint32_t)(st >> 32),
                                             (uint32_t)g,
                                             (uint32_t)(g >> 32) ^ 0x54554D56u, k0, k1);
            isotropic_unit(uniform53(ev.v[2], ev.v[3]), uniform53(ve.v[0], ve.v[1]), pol);
            tumbled = true;
        }
    }
    const float px = (float)pol[0], py = (float)pol[1], pz = (float)pol[2];
    const float speed = active ? s->v_A : 0.0f;
    const float mot = a.motility_coeff;
    const float vx = speed * px + mot * (float)s->Ix;
    const float vy = speed * py + mot * (float)s->Iy;
    const float vz = speed * pz + mot * (float)s->Iz;
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz) || !isfinite(vx) ||
        !isfinite(vy) || !isfinite(vz)) {
        raise3(s, a.flags, PFH3D_FLAG_NONFINITE);
        return;
    }
    s->polarity_x = px; s->polarity_y = py; s->polarity_z = pz;
    s->velocity_x = vx; s->velocity_y = vy; s->velocity_z = vz;
    if (tumbled) ++s->tumble_ctr;
}

// ---------------------------------------------------------------------------
// measurement
// ---------------------------------------------------------------------------
struct MeasureMode {
    int apply_tumble;
    int surface_mode;   // 0 off, 1 on, 2 every full_moment_every begun steps
    int apply_motion;
    int clear_out;      // zero S_out (the next aggregate) in this launch
    int snapshot;       // copy cells -> accepted before committing
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    return v;
}
__device__ __forceinline__ int wmin(int v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_down(v, d, 64));
    return v;
}
__device__ __forceinline__ int wmax(int v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_down(v, d, 64));
    return v;
}

__global__ __launch_bounds__(1024) void k3_measure(const pfh3d_step_args a,
                                                   const MeasureMode m) {
    extern __shared__ __align__(16) float lds[];
    __shared__ double red[16][8];
    __shared__ int ired[16][7];

    const int B = a.B, S = a.shards, L = a.L, P = a.P;
    const int NT = (int)blockDim.x, tid = (int)threadIdx.x;
    const int n = (int)blockIdx.x / S, sh = (int)blockIdx.x - n * S;

    if (m.clear_out) {
        const size_t total4 = (size_t)L * (size_t)L * (size_t)P / 4u;
        const u32x4 z4 = {0u, 0u, 0u, 0u};
        u32x4* dst = reinterpret_cast<u32x4*>(a.S_out);
        for (size_t i = (size_t)blockIdx.x * NT + tid; i < total4;
             i += (size_t)gridDim.x * NT)
            __builtin_nontemporal_store(z4, dst + i);
    }
    if (m.snapshot && sh == 0 && tid < 16) {
        const f32x4* s4 = reinterpret_cast<const f32x4*>(a.cells + n);
        f32x4* d4 = reinterpret_cast<f32x4*>(a.accepted + n);
        d4[tid] = s4[tid];
    }

    pfh3d_cell_state* cs = a.cells + n;
    const int ox = wrap_origin(cs->origin_x, L);
    const int oy = wrap_origin(cs->origin_y, L);
    const int oz = wrap_origin(cs->origin_z, L);
    int surface = m.surface_mode;
    if (surface == 2)
        surface = a.full_moment_every > 0 &&
                  a.counters[1] % (uint64_t)a.full_moment_every == 0u;

    const int zb = sh * B / S, ze = (sh + 1) * B / S;
    const int gpr = B / 4, G = B * B / 4;
    const bool act = tid < G;
    const int y = act ? tid / gpr : 0;
    const int x0 = act ? 4 * (tid - y * gpr) : 0;
    const int PX = B + 4, PL = (B + 2) * PX;
    const int own = (y + 1) * PX + 4 + x0;

    // A ring of four planes in LDS, each with a zero border (x/y neighbours
    // past the brick): the plane a lane measures, its two z neighbours, and
    // the plane being filled. Lanes only ever write their own 16 B of a
    // plane's interior, so the border stays zero.
    for (int i = tid; i < 4 * PL; i += NT) lds[i] = 0.0f;
    auto slot = [&](int p) -> float* { return lds + ((p - zb + 1) & 3) * PL; };

    // Every vector load of the plane loop is unconditional and the loop
    // stores nothing to global memory, so the compiler's wait before each
    // use counts only the loads issued after it: a plane arrives one
    // iteration after it is requested, the aggregate words one plane ahead
    // of their use, and no iteration waits for its own requests. Planes
    // outside the brick are read from a clamped address and replaced by
    // zeros; inactive lanes read group 0.
    const PFH_GLOBAL float* tile = gptr(a.phi_in) + (size_t)n * B * B * B;
    const PFH_GLOBAL uint32_t* Sin = gptr(a.S_in);
    const int lane_off = act ? 4 * tid : 0;
    // (the raw request, and the zeros applied where the plane is used, so
    // that nothing touches the loaded registers before then)
    auto request_plane = [&](int p) -> f32x4 {
        const int pc = p < 0 ? 0 : (p >= B ? B - 1 : p);
        return *reinterpret_cast<const PFH_GLOBAL f32x4*>(tile + (size_t)pc * B * B +
                                                          lane_off);
    };
    auto in_brick = [&](const f32x4& v, int p) -> f32x4 {
        const bool in = act && (unsigned)p < (unsigned)B;
        return f32x4{in ? v[0] : 0.f, in ? v[1] : 0.f, in ? v[2] : 0.f, in ? v[3] : 0.f};
    };
    auto load_plane = [&](int p) -> f32x4 { return in_brick(request_plane(p), p); };
    const int wy = wrap_off(oy, y, L);
    int wx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) wx[j] = wrap_off(ox, x0 + j, L);
    // S words of this lane's 4 voxels in plane z, read only where that
    // plane's phi is nonzero (the interaction reads S nowhere else, and most
    // of a brick is empty); the other requests go to word 0, one cache line
    // for the whole launch. The plane loop reads a word only where its
    // voxel's phi is nonzero, the same mask.
    auto load_s = [&](int z, const f32x4& ph, uint32_t (&sv)[4]) {
        const size_t srow = s_index(L, P, 0, wy, wrap_off(oz, z < B ? z : B - 1, L));
        const bool zin = act && z < ze;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t idx = (zin && ph[j] != 0.0f) ? srow + (size_t)wx[j] : 0u;
            sv[j] = Sin[idx];
        }
    };

    {
        const f32x4 Pm = load_plane(zb - 1);
        const f32x4 P0 = load_plane(zb);
        const f32x4 P1 = load_plane(zb + 1);
        __syncthreads();                        // the ring's zeros are written
        if (act) {
            *reinterpret_cast<f32x4*>(slot(zb - 1) + own) = Pm;
            *reinterpret_cast<f32x4*>(slot(zb) + own) = P0;
            *reinterpret_cast<f32x4*>(slot(zb + 1) + own) = P1;
        }
    }
    // Two planes in flight (z + 2 and z + 3 while plane z is measured) and
    // the aggregate words of two planes (z and z + 1): each request has two
    // iterations to arrive.
    // (issued in the order the loop issues them: S0, R0, S1, R1)
    uint32_t S0[4], S1[4];
    f32x4 R0, R1;
    int Rp0 = zb + 2, Rp1 = zb + 3;            // the planes' indices
    {
        const f32x4 P0 = *reinterpret_cast<const f32x4*>(slot(zb) + own);
        const f32x4 P1 = *reinterpret_cast<const f32x4*>(slot(zb + 1) + own);
        load_s(zb, P0, S0);
        R0 = request_plane(Rp0);
        load_s(zb + 1, P1, S1);
        R1 = request_plane(Rp1);
    }
    __syncthreads();

    double aV = 0.0, aCx = 0.0, aCy = 0.0, aCz = 0.0, aS = 0.0, aIx = 0.0, aIy = 0.0,
           aIz = 0.0;
    int lo0 = B, lo1 = B, lo2 = B, hi0 = -1, hi1 = -1, hi2 = -1;
    uint32_t pmax = 0u, n_nonfin = 0u, n_neg = 0u;

    // One plane: measure plane z with the aggregate words requested two
    // planes earlier (set Sz), store the plane that arrived (z + 2, in Rz)
    // into the ring, request its aggregate words into Sz and plane z + 4
    // into Rz.
    // This lane's own groups of planes z - 1 and z roll through registers
    // (read from the ring once each, as plane z + 1)
    f32x4 Am = *reinterpret_cast<const f32x4*>(slot(zb - 1) + own);
    f32x4 A0 = *reinterpret_cast<const f32x4*>(slot(zb) + own);
    auto plane_step = [&](int z, uint32_t (&Sz)[4], f32x4& Rz, int& Rpz) {
        const uint32_t (&Sc)[4] = Sz;
        const f32x4 Ap = *reinterpret_cast<const f32x4*>(slot(z + 1) + own);
        if (act) {
            const float* pl = slot(z);
            const f32x4 ym = *reinterpret_cast<const f32x4*>(pl + own - PX);
            const f32x4 yp = *reinterpret_cast<const f32x4*>(pl + own + PX);
            const float xl = pl[own - 1], xr = pl[own + 4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float c = A0[j];
                if (!isfinite(c)) {
                    ++n_nonfin;
                    continue;
                }
                const float xm = j == 0 ? xl : A0[j == 0 ? 0 : j - 1];
                const float xp = j == 3 ? xr : A0[j == 3 ? 3 : j + 1];
                const float gx = 0.5f * (xp - xm);
                const float gy = 0.5f * (yp[j] - ym[j]);
                const float gz = 0.5f * (Ap[j] - Am[j]);
                if (c == 0.0f && gx == 0.0f && gy == 0.0f && gz == 0.0f) continue;
                float other = 0.0f;
                const int x = x0 + j;
                if (c != 0.0f) {
                    // other_field (src/pf3d/device_common.cuh:182-191), its
                    // S_NEGATIVE event counted here and raised after the loop
                    const uint32_t agg = Sc[j], mine = q_of(c);
                    if (agg >= mine) other = (float)(agg - mine) * pfh::kQInv;
                    else ++n_neg;
                }
                const double sq = (double)c * (double)c;
                aV += sq;
                aCx += sq * (double)x;
                aCy += sq * (double)y;
                aCz += sq * (double)z;
                if (surface)
                    aS += sqrt((double)gx * (double)gx + (double)gy * (double)gy +
                               (double)gz * (double)gz);
                aIx += (double)(c * gx * other);
                aIy += (double)(c * gy * other);
                aIz += (double)(c * gz * other);
                const float ac = fabsf(c);
                pmax = max(pmax, __float_as_uint(ac));
                if (ac > kEps) {
                    lo0 = min(lo0, x); hi0 = max(hi0, x);
                    lo1 = min(lo1, y); hi1 = max(hi1, y);
                    lo2 = min(lo2, z); hi2 = max(hi2, z);
                }
            }
        }
        // plane z + 2 replaces plane z - 2, last read before the previous
        // barrier; a plane past the shard's last neighbour (ze) is never
        // read, and its request is the cached plane ze again
        Am = A0;
        A0 = Ap;
        const f32x4 An = in_brick(Rz, Rpz);
        if (act) *reinterpret_cast<f32x4*>(slot(z + 2) + own) = An;
        load_s(z + 2, An, Sz);
        Rpz = z + 4 <= ze ? z + 4 : ze;
        Rz = request_plane(Rpz);
        pfh::lds_barrier();   // the ring is the only data shared here
    };
    // pairs of planes, then an odd last plane: both steps of a pair run
    // unconditionally, so the request counts the waits see are fixed
    int z = zb;
    for (; z + 1 < ze; z += 2) {
        plane_step(z, S0, R0, Rp0);
        plane_step(z + 1, S1, R1, Rp1);
    }
    if (z < ze) plane_step(z, S0, R0, Rp0);
    if (n_nonfin) {
        atomicOr(&cs->flags, 1u << PFH3D_FLAG_NONFINITE);
        atomicAdd(&a.flags[PFH3D_FLAG_NONFINITE], n_nonfin);
    }
    if (n_neg) {
        atomicOr(&cs->flags, 1u << PFH3D_FLAG_S_NEGATIVE);
        atomicAdd(&a.flags[PFH3D_FLAG_S_NEGATIVE], n_neg);
    }

    // 64-lane trees, then waves in ascending order
    aV = wsum(aV); aCx = wsum(aCx); aCy = wsum(aCy); aCz = wsum(aCz);
    aS = wsum(aS); aIx = wsum(aIx); aIy = wsum(aIy); aIz = wsum(aIz);
    lo0 = wmin(lo0); lo1 = wmin(lo1); lo2 = wmin(lo2);
    hi0 = wmax(hi0); hi1 = wmax(hi1); hi2 = wmax(hi2);
    pmax = (uint32_t)wmax((int)pmax);   // |phi| bit patterns are non-negative ints
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0) {
        red[wave][0] = aV; red[wave][1] = aCx; red[wave][2] = aCy; red[wave][3] = aCz;
        red[wave][4] = aS; red[wave][5] = aIx; red[wave][6] = aIy; red[wave][7] = aIz;
        ired[wave][0] = lo0; ired[wave][1] = hi0; ired[wave][2] = lo1;
        ired[wave][3] = hi1; ired[wave][4] = lo2; ired[wave][5] = hi2;
        ired[wave][6] = (int)pmax;
    }
    __syncthreads();
    if (tid != 0) return;

    Acc t = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    Box b = {{B, B, B}, {-1, -1, -1}, 0u};
    for (int w = 0; w < NT / 64; ++w) {
        t.V += red[w][0]; t.Cx += red[w][1]; t.Cy += red[w][2]; t.Cz += red[w][3];
        t.surf += red[w][4]; t.Ix += red[w][5]; t.Iy += red[w][6]; t.Iz += red[w][7];
        b.lo[0] = min(b.lo[0], ired[w][0]); b.hi[0] = max(b.hi[0], ired[w][1]);
        b.lo[1] = min(b.lo[1], ired[w][2]); b.hi[1] = max(b.hi[1], ired[w][3]);
        b.lo[2] = min(b.lo[2], ired[w][4]); b.hi[2] = max(b.hi[2], ired[w][5]);
        b.pmax = max(b.pmax, (uint32_t)ired[w][6]);
    }

    Acc tot = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    Box tb = {{B, B, B}, {-1, -1, -1}, 0u};
    if (S > 1) {
        pfh3d_moment_partial* pp = a.partials + (size_t)n * S;
        pfh3d_moment_partial mine;
        mine.V = t.V; mine.Cx = t.Cx; mine.Cy = t.Cy; mine.Cz = t.Cz;
        mine.surface = t.surf; mine.Ix = t.Ix; mine.Iy = t.Iy; mine.Iz = t.Iz;
        mine.lo_x = b.lo[0]; mine.hi_x = b.hi[0]; mine.lo_y = b.lo[1];
        mine.hi_y = b.hi[1]; mine.lo_z = b.lo[2]; mine.hi_z = b.hi[2];
        mine.phi_max_bits = b.pmax; mine.padding = 0u;
        pp[sh] = mine;
        __threadfence();
        const uint32_t prev = atomicAdd(a.arrivals + n, 1u);
        if (prev != (uint32_t)(S - 1)) return;
        __threadfence();
        for (int k = 0; k < S; ++k) {
            const volatile pfh3d_moment_partial* q = pp + k;
            tot.V += q->V; tot.Cx += q->Cx; tot.Cy += q->Cy; tot.Cz += q->Cz;
            tot.surf += q->surface; tot.Ix += q->Ix; tot.Iy += q->Iy; tot.Iz += q->Iz;
            tb.lo[0] = min(tb.lo[0], (int)q->lo_x); tb.hi[0] = max(tb.hi[0], (int)q->hi_x);
            tb.lo[1] = min(tb.lo[1], (int)q->lo_y); tb.hi[1] = max(tb.hi[1], (int)q->hi_y);
            tb.lo[2] = min(tb.lo[2], (int)q->lo_z); tb.hi[2] = max(tb.hi[2], (int)q->hi_z);
            tb.pmax = max(tb.pmax, (uint32_t)q->phi_max_bits);
        }
        atomicExch(a.arrivals + n, 0u);
    } else {
        tot.V += t.V; tot.Cx += t.Cx; tot.Cy += t.Cy; tot.Cz += t.Cz;
        tot.surf += t.surf; tot.Ix += t.Ix; tot.Iy += t.Iy; tot.Iz += t.Iz;
        tb = b;
    }
    const CommitArgs ca{a.flags, B, a.max_shift, a.support_margin};
    commit_cell(cs, ca, tot, tb, surface);
    if (m.apply_motion) {
        const MotionArgs ma{a.flags, a.counters, a.aging_steps, a.polarity_stream,
                            a.p_tumble, a.motility_coeff};
        motion_cell(cs, ma, m.apply_tumble);
    }
}

// ---------------------------------------------------------------------------
// update
// ---------------------------------------------------------------------------
struct CellCoef {
    float gamma, bulk, volume, vx, vy, vz, dt, ic;
};

// rhs and new value (src/pf3d/kernels.cu:655-676); returns false (and
// flags) when the result is not finite.
__device__ __forceinline__ float rhs_next(const CellCoef& k, float c, float lap,
                                          float gx, float gy, float gz, float other) {
    const float rhs = k.gamma * lap
        - k.bulk * c * (1.0f - c) * (1.0f - 2.0f * c)
        + k.volume * c
        - k.ic * c * other
        - (k.vx * gx + k.vy * gy + k.vz * gz);
    return c + k.dt * rhs;
}

// Per-plane partial sums of the 27-point stencil at one voxel of the 3x3
// neighbourhood (r0: y-1 row, r1: y row, r2: y+1 row; index 1 is x).
struct PlaneSums { float A, q1, q0, gx, gy; };
__device__ __forceinline__ PlaneSums plane_sums(float r00, float r01, float r02,
                                                float r10, float r11, float r12,
                                                float r20, float r21, float r22) {
    PlaneSums o;
    const float e = (r10 + r12) + (r01 + r21);
    const float cc = (r00 + r02) + (r20 + r22);
    o.A = r11;
    o.q1 = (14.0f * r11 + 3.0f * e) + cc;
    o.q0 = (-128.0f * r11 + 14.0f * e) + 3.0f * cc;
    o.gx = 0.5f * (r12 - r10);
    o.gy = 0.5f * (r21 - r01);
    return o;
}

// Recentring clip check of source voxel (x,y,z) whose destination lies
// outside the brick (src/pf3d/kernels_periodic_slab_walk.cu:268-314), with
// the same stencil arithmetic as the main path, read from global memory.
struct ClipArgs {
    const uint32_t* S_in;
    uint32_t* flags;
    int B, L, P;
};

__device__ __noinline__ void clip_voxel(const ClipArgs a, pfh3d_cell_state* cs,
                                        const float* src, CellCoef k, int ox, int oy,
                                        int oz, int x, int y, int z) {
    const int B = a.B;
    auto f = [&](int xx, int yy, int zz) -> float {
        return ((unsigned)xx < (unsigned)B && (unsigned)yy < (unsigned)B &&
                (unsigned)zz < (unsigned)B)
                   ? src[((size_t)zz * B + yy) * B + xx] : 0.0f;
    };
    const float c = f(x, y, z);
    if (!isfinite(c)) {
        raise3(cs, a.flags, PFH3D_FLAG_NONFINITE);
        return;
    }
    PlaneSums ps[3];
    for (int p = 0; p < 3; ++p) {
        const int zz = z + p - 1;
        ps[p] = plane_sums(f(x - 1, y - 1, zz), f(x, y - 1, zz), f(x + 1, y - 1, zz),
                           f(x - 1, y, zz), f(x, y, zz), f(x + 1, y, zz),
                           f(x - 1, y + 1, zz), f(x, y + 1, zz), f(x + 1, y + 1, zz));
    }
    const float lap = ((ps[0].q1 + ps[1].q0) + ps[2].q1) * (float)(1.0 / 30.0);
    const float gx = ps[1].gx, gy = ps[1].gy;
    const float gz = 0.5f * (ps[2].A - ps[0].A);
    if (c == 0.0f && lap == 0.0f && gx == 0.0f && gy == 0.0f && gz == 0.0f) return;
    const int L = a.L;
    const uint32_t agg = a.S_in[s_index(L, a.P, wrap_off(ox, x, L), wrap_off(oy, y, L),
                                        wrap_off(oz, z, L))];
    const float other = other_field(agg, c, cs, a.flags);
    const float nx = rhs_next(k, c, lap, gx, gy, gz, other);
    if (!isfinite(nx)) {
        raise3(cs, a.flags, PFH3D_FLAG_NONFINITE);
        return;
    }
    if (fabsf(nx) > kEps) raise3(cs, a.flags, PFH3D_FLAG_DESTINATION_CLIP);
}

__global__ __launch_bounds__(1024) void k3_update(const pfh3d_step_args a, const int K,
                                                  const int H) {
    extern __shared__ __align__(16) float lds[];
    __shared__ int s_abort;

    const int B = a.B, L = a.L, P = a.P;
    const int NT = (int)blockDim.x, tid = (int)threadIdx.x;
    // workgroup = (cell, z slab, y part): destination planes [zb, ze) and
    // rows [yb, ye); parts share nothing but the input field and the atomics
    const int KH = K * H;
    const int n = (int)blockIdx.x / KH, rem = (int)blockIdx.x - n * KH;
    const int k = rem / H, h = rem - (rem / H) * H;
    const int zb = k * B / K, ze = (k + 1) * B / K;
    const int yb = h * B / H, ye = (h + 1) * B / H;
    pfh3d_cell_state* cs = a.cells + n;
    const size_t words = (size_t)B * B * B;
    const float* src = a.phi_in + (size_t)n * words;
    float* dst = a.phi_out + (size_t)n * words;
    const int gpr = B / 4, G = (ye - yb) * gpr;
    const bool act = tid < G;
    const int gy_ = act ? yb + tid / gpr : 0;
    const int gx0 = act ? 4 * (tid % gpr) : 0;
    const int ox = wrap_origin(cs->origin_x, L);
    const int oy = wrap_origin(cs->origin_y, L);
    const int oz = wrap_origin(cs->origin_z, L);
    const int sx = cs->pending_shift_x, sy = cs->pending_shift_y, sz = cs->pending_shift_z;
    if (tid == 0) {
        int ab = fatal_present(cs, a.flags) ? 1 : 0;
        if (!ab && (abs(sx) > PFH3D_MAX_SHIFT || abs(sy) > PFH3D_MAX_SHIFT ||
                    abs(sz) > PFH3D_MAX_SHIFT)) {
            // a pending shift the measurement cannot produce (|s| <= 4):
            // counted once per cell, every slab keeps the accepted field
            if (k == 0 && h == 0) raise3(cs, a.flags, PFH3D_FLAG_INVALID_GEOMETRY);
            ab = 1;
        }
        s_abort = ab;
    }
    __syncthreads();
    const PFH_GLOBAL float* gsrc = gptr(src);
    PFH_GLOBAL float* gdst = gptr(dst);

    if (s_abort) {
        // keep the accepted field; rebuild its aggregate contribution
        for (int z = zb; z < ze; ++z) {
            if (!act) break;
            const size_t o = ((size_t)z * B + gy_) * B + gx0;
            const f32x4 v = *reinterpret_cast<const PFH_GLOBAL f32x4*>(gsrc + o);
            *reinterpret_cast<PFH_GLOBAL f32x4*>(gdst + o) = v;
            const int wy = wrap_off(oy, gy_, L), wz = wrap_off(oz, z, L);
            for (int j = 0; j < 4; ++j)
                if (isfinite(v[j]))
                    scatter_value(a.S_out, s_index(L, P, wrap_off(ox, gx0 + j, L), wy, wz),
                                  v[j], cs, a.flags);
        }
        __syncthreads();
        if (k == 0 && h == 0 && tid == 0) {
            cs->pending_shift_x = 0;
            cs->pending_shift_y = 0;
            cs->pending_shift_z = 0;
        }
        return;
    }

    CellCoef kc;
    kc.gamma = cs->gamma;
    kc.bulk = a.bulk_scale * kc.gamma;
    kc.volume = (float)(a.volume_scale * (a.V0 - cs->V));
    kc.vx = cs->velocity_x;
    kc.vy = cs->velocity_y;
    kc.vz = cs->velocity_z;
    kc.dt = a.dt;
    kc.ic = a.interaction_coeff;

    const int GPU = B / 4 + 2 * kMarginG;            // words per sub-row
    const int PXU = update_row_pitch(B);
    const int PLU = (B + 2 * kMargin + 1) * PXU;      // B + 11 rows
    for (int i = tid; i < 2 * PLU; i += NT) lds[i] = 0.0f;

    // staging: lane tid owns source group tid % gpr of source row
    // yb + sy - 1 + tid / gpr (the part's rows as the shift reads them, with
    // one halo row on each side); element e of the group goes to sub-row e
    const int g = gx0 >> 2;
    const int srow = yb + sy - 1 + tid / gpr;
    const bool stg = tid < (ye - yb + 2) * gpr && (unsigned)srow < (unsigned)B;
    const int sg = tid % gpr;
    const int stage = (srow + kMargin) * PXU + kMarginG + sg;
    // reading: destination group (gx0.., gy_) reads source x = gx0 + sx + c - 1,
    // c = 0..5, of rows gy_ + sy - 1..gy_ + sy + 1: sub-row (sx + c - 1) & 3,
    // group g + ((sx + c - 1) >> 2), both offsets uniform
    const int rd = (gy_ + sy + kMargin) * PXU + kMarginG + g;
    int roff[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const int t = sx + c - 1;
        roff[c] = (t & 3) * GPU + (t >> 2);
    }
    auto stage_plane = [&](float* plane, const f32x4& v) {
        float* st = plane + stage;
        st[0] = v[0];
        st[GPU] = v[1];
        st[2 * GPU] = v[2];
        st[3 * GPU] = v[3];
    };
    auto load_plane = [&](int p) -> f32x4 {
        f32x4 v = {0.f, 0.f, 0.f, 0.f};
        if (stg && (unsigned)p < (unsigned)B)
            v = *reinterpret_cast<const PFH_GLOBAL f32x4*>(gsrc + (size_t)p * B * B +
                                                          (size_t)srow * B + 4 * sg);
        return v;
    };

    const int pbeg = zb + sz - 1;
    const int pcount = ze - zb + 2;
    f32x4 R = load_plane(pbeg);
    f32x4 Rn = load_plane(pbeg + 1);
    __syncthreads();
    if (stg) stage_plane(lds, R);
    __syncthreads();

    float T[4] = {0.f, 0.f, 0.f, 0.f}, Q1p[4] = {0.f, 0.f, 0.f, 0.f};
    float Am2[4] = {0.f, 0.f, 0.f, 0.f}, Am1[4] = {0.f, 0.f, 0.f, 0.f};
    float gxm[4] = {0.f, 0.f, 0.f, 0.f}, gym[4] = {0.f, 0.f, 0.f, 0.f};
    const int syv = gy_ + sy;
    const bool row_ok = act && (unsigned)syv < (unsigned)B;
    const int wyv = row_ok ? wrap_off(oy, syv, L) : 0;
    const PFH_GLOBAL uint32_t* Sin = gptr(a.S_in);
    // S words of iteration i's destination row (source centre plane pbeg + i - 1),
    // fetched one iteration ahead; indices exactly those the update reads, and
    // only where the voxel's own phi (cv, the source centre value) is nonzero:
    // at phi = 0 the interaction term ic phi S_other is a zero of the same
    // sign for any finite S_other, and S_other cannot be negative
    auto load_su = [&](int i, const float (&cv)[4], uint32_t (&sv)[4]) {
#pragma unroll
        for (int j = 0; j < 4; ++j) sv[j] = 0u;
        const int zc = pbeg + i - 1;
        if (i >= 2 && i < pcount && row_ok && (unsigned)zc < (unsigned)B) {
            const size_t srow = s_index(L, P, 0, wyv, wrap_off(oz, zc, L));
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int xs = gx0 + j + sx;
                if ((unsigned)xs < (unsigned)B && cv[j] != 0.0f)
                    sv[j] = Sin[srow + (size_t)wrap_off(ox, xs, L)];
            }
        }
    };
    uint32_t Su[4], Sun[4];
    {
        const float none[4] = {0.f, 0.f, 0.f, 0.f};
        load_su(0, none, Su);
    }
    // S_OVERFLOW test of the previous plane's scatters, taken one plane later
    // so that no atomic's return is waited for in the plane that issued it
    uint32_t pold[4] = {0u, 0u, 0u, 0u}, pq[4] = {0u, 0u, 0u, 0u};
    auto check_overflow = [&]() {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (pq[j] != 0u && pold[j] > 0xffffffffu - pq[j])
                raise3(cs, a.flags, PFH3D_FLAG_S_OVERFLOW);
            pq[j] = 0u;
        }
    };

    for (int i = 0; i < pcount; ++i) {
        const int p = pbeg + i;
        const int cur = i & 1;
        const f32x4 Rnn = (i + 2 < pcount) ? load_plane(p + 2) : f32x4{0.f, 0.f, 0.f, 0.f};
        float A[4] = {0.f, 0.f, 0.f, 0.f}, Q1[4] = {0.f, 0.f, 0.f, 0.f},
              Q0[4] = {0.f, 0.f, 0.f, 0.f}, GX[4] = {0.f, 0.f, 0.f, 0.f},
              GY[4] = {0.f, 0.f, 0.f, 0.f};
        if (act && (unsigned)p < (unsigned)B) {
            const float* pl = lds + cur * PLU + rd;
            float r0[6], r1[6], r2[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                r0[c] = pl[roff[c] - PXU];
                r1[c] = pl[roff[c]];
                r2[c] = pl[roff[c] + PXU];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const PlaneSums s = plane_sums(r0[j], r0[j + 1], r0[j + 2], r1[j], r1[j + 1],
                                               r1[j + 2], r2[j], r2[j + 1], r2[j + 2]);
                A[j] = s.A; Q1[j] = s.q1; Q0[j] = s.q0; GX[j] = s.gx; GY[j] = s.gy;
            }
        }
        // A holds this plane's centre values: the next destination row's phi
        load_su(i + 1, A, Sun);
        check_overflow();
        if (i >= 2 && act) {
            // destination plane d = p - 1 - sz from source centre plane p - 1
            const int d = p - 1 - sz;
            const int zc = p - 1;
            f32x4 out = {0.f, 0.f, 0.f, 0.f};
            if (row_ok && (unsigned)zc < (unsigned)B) {
                const int wz = wrap_off(oz, zc, L);
                const size_t srow = s_index(L, P, 0, wyv, wz);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int xs = gx0 + j + sx;
                    if ((unsigned)xs >= (unsigned)B) continue;
                    const float c = Am1[j];
                    if (!isfinite(c)) {
                        raise3(cs, a.flags, PFH3D_FLAG_NONFINITE);
                        continue;
                    }
                    const float lap = (T[j] + Q1[j]) * (float)(1.0 / 30.0);
                    const float gx = gxm[j], gy = gym[j];
                    const float gz = 0.5f * (A[j] - Am2[j]);
                    if (c == 0.0f && lap == 0.0f && gx == 0.0f && gy == 0.0f && gz == 0.0f)
                        continue;
                    const size_t si = srow + (size_t)wrap_off(ox, xs, L);
                    const float other = other_field(Su[j], c, cs, a.flags);
                    const float nx = rhs_next(kc, c, lap, gx, gy, gz, other);
                    if (!isfinite(nx)) {
                        raise3(cs, a.flags, PFH3D_FLAG_NONFINITE);
                        continue;
                    }
                    out[j] = nx;
                    // scatter_value (src/pf3d/device_common.cuh:193-205) with
                    // the overflow test deferred
                    if (nx * nx > pfh::kQClampPhiSq) raise3(cs, a.flags, PFH3D_FLAG_Q_CLAMP);
                    const uint32_t q = q_of(nx);
                    if (q != 0u) {
                        pold[j] = __hip_atomic_fetch_add(gptr(a.S_out + si), q, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                        pq[j] = q;
                    }
                }
            }
            *reinterpret_cast<PFH_GLOBAL f32x4*>(gdst + ((size_t)d * B + gy_) * B + gx0) = out;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T[j] = Q1p[j] + Q0[j];
            Q1p[j] = Q1[j];
            Am2[j] = Am1[j];
            Am1[j] = A[j];
            gxm[j] = GX[j];
            gym[j] = GY[j];
        }
        if (i + 1 < pcount && stg) stage_plane(lds + (cur ^ 1) * PLU, Rn);
        pfh::lds_barrier();   // the plane buffers are the only shared data here
        Rn = Rnn;
#pragma unroll
        for (int j = 0; j < 4; ++j) Su[j] = Sun[j];
    }
    check_overflow();

    if ((sx != 0 || sy != 0 || sz != 0) && act) {
        // source voxels of this slab's planes whose destination leaves the brick
        const ClipArgs ca{a.S_in, a.flags, B, L, P};
        for (int z = zb; z < ze; ++z)
            for (int j = 0; j < 4; ++j) {
                const int x = gx0 + j;
                if ((unsigned)(x - sx) < (unsigned)B && (unsigned)(gy_ - sy) < (unsigned)B &&
                    (unsigned)(z - sz) < (unsigned)B)
                    continue;
                clip_voxel(ca, cs, src, kc, ox, oy, oz, x, gy_, z);
            }
    }
}

// ---------------------------------------------------------------------------
// finish: repair, finalize origins, restore, advance
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k3_finish(const pfh3d_step_args a) {
    const int b = (int)blockIdx.x, tid = (int)threadIdx.x;
    const bool fatal = global_fatal(a.flags);
    if (fatal) {
        // Repair: every field keeps its step input; S_in is exactly the
        // aggregate of those fields, so the rebuilt aggregate is its copy.
        const size_t words = (size_t)a.B * a.B * a.B;
        if (b < a.N) {
            const f32x4* s4 = reinterpret_cast<const f32x4*>(a.phi_in + (size_t)b * words);
            f32x4* d4 = reinterpret_cast<f32x4*>(a.phi_out + (size_t)b * words);
            for (size_t i = tid; i < words / 4; i += blockDim.x) d4[i] = s4[i];
        }
        const size_t sw4 = (size_t)a.L * a.L * a.P / 4u;
        const u32x4* si = reinterpret_cast<const u32x4*>(a.S_in);
        u32x4* so = reinterpret_cast<u32x4*>(a.S_out);
        for (size_t i = (size_t)b * blockDim.x + tid; i < sw4; i += (size_t)gridDim.x * blockDim.x)
            so[i] = si[i];
        // Restore the accepted state (the snapshot has zero pending shifts
        // and the accepted origins).
        if (b < a.N && tid < 16) {
            const f32x4* s4 = reinterpret_cast<const f32x4*>(a.accepted + b);
            f32x4* d4 = reinterpret_cast<f32x4*>(a.cells + b);
            d4[tid] = s4[tid];
        }
    } else if (b < a.N && tid == 0) {
        pfh3d_cell_state* s = a.cells + b;
        if ((s->flags & kFatal) == 0u) {
            const int sx = s->pending_shift_x, sy = s->pending_shift_y, sz = s->pending_shift_z;
            int64_t nx, ny, nz;
            if (__builtin_add_overflow(s->origin_x, (int64_t)sx, &nx) ||
                __builtin_add_overflow(s->origin_y, (int64_t)sy, &ny) ||
                __builtin_add_overflow(s->origin_z, (int64_t)sz, &nz)) {
                raise3(s, a.flags, PFH3D_FLAG_INVALID_GEOMETRY);
            } else {
                s->origin_x = nx;
                s->origin_y = ny;
                s->origin_z = nz;
                if (sx != 0 || sy != 0 || sz != 0) ++s->shift_ctr;
            }
        }
        s->pending_shift_x = s->pending_shift_y = s->pending_shift_z = 0;
    }
    if (b == 0 && tid == 0) {
        if (!fatal) a.counters[0] += 1u;
        a.counters[1] += 1u;
    }
}

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------
// k_initialize_spheres, src/pf3d/kernels.cu:1071-1122
__global__ __launch_bounds__(256) void k3_init_spheres(float* phi_even, float* phi_odd,
                                                       pfh3d_cell_state* cells,
                                                       const double* centres, int B,
                                                       float lambda, float seed_radius) {
    const int n = (int)blockIdx.x;
    pfh3d_cell_state* s = cells + n;
    const double cx = centres[3 * n], cy = centres[3 * n + 1], cz = centres[3 * n + 2];
    const int64_t ox = llrint(cx - 0.5 * (double)(B - 1));
    const int64_t oy = llrint(cy - 0.5 * (double)(B - 1));
    const int64_t oz = llrint(cz - 0.5 * (double)(B - 1));
    const size_t words = (size_t)B * B * B;
    const float k = 2.7386127875258306f / lambda;
    float* even = phi_even + (size_t)n * words;
    float* odd = phi_odd ? phi_odd + (size_t)n * words : nullptr;
    for (size_t q = threadIdx.x; q < words; q += blockDim.x) {
        const int x = (int)(q % (size_t)B);
        const int y = (int)((q / (size_t)B) % (size_t)B);
        const int z = (int)(q / ((size_t)B * B));
        const double dx = (double)(ox + x) - cx;
        const double dy = (double)(oy + y) - cy;
        const double dz = (double)(oz + z) - cz;
        const float r = (float)sqrt(dx * dx + dy * dy + dz * dz);
        const float t = (float)tanh_d((double)(k * (r - seed_radius)));
        even[q] = 0.5f * (1.0f - t);
        if (odd) odd[q] = 0.0f;
    }
    if (threadIdx.x == 0) {
        s->origin_x = ox;
        s->origin_y = oy;
        s->origin_z = oz;
        s->shift_ctr = s->tumble_ctr = 0u;
        s->flags = 0u;
        s->pad0 = s->pad1 = s->pad2 = 0u;
        s->pending_shift_x = s->pending_shift_y = s->pending_shift_z = 0;
        s->V = s->Cx = s->Cy = s->Cz = 0.0;
        s->surface = s->Ix = s->Iy = s->Iz = 0.0;
        s->wall_overlap = s->outside_slit_volume = 0.0;
        s->phi_max = 0.0f;
        for (int i = 0; i < 10; ++i) s->reserved[i] = 0u;
    }
}

// k_scatter_current, src/pf3d/kernels.cu:1148-1186: one workgroup per plane
__global__ __launch_bounds__(256) void k3_scatter(const float* phi, pfh3d_cell_state* cells,
                                                  uint32_t* S, uint32_t* flags, int B,
                                                  int L, int P) {
    const int n = (int)blockIdx.x / B, z = (int)blockIdx.x - n * B;
    pfh3d_cell_state* s = cells + n;
    const float* pl = phi + (size_t)n * B * B * B + (size_t)z * B * B;
    const int ox = wrap_origin(s->origin_x, L), oy = wrap_origin(s->origin_y, L),
              oz = wrap_origin(s->origin_z, L);
    const int wz = wrap_off(oz, z, L);
    for (int q = (int)threadIdx.x; q < B * B; q += (int)blockDim.x) {
        const int y = q / B, x = q - y * B;
        const float v = pl[q];
        if (!isfinite(v)) {
            raise3(s, flags, PFH3D_FLAG_NONFINITE);
            continue;
        }
        scatter_value(S, s_index(L, P, wrap_off(ox, x, L), wrap_off(oy, y, L), wz), v, s,
                      flags);
    }
}

__global__ __launch_bounds__(256) void k3_clear(uint32_t* p, size_t n4) {
    const u32x4 z4 = {0u, 0u, 0u, 0u};
    u32x4* d = reinterpret_cast<u32x4*>(p);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (size_t)gridDim.x * blockDim.x)
        d[i] = z4;
}

// Trajectory records (k_pack_trajectory's fields, src/pf3d/kernels.cu:2323-2356,
// periodic geometry: no height offset) of 64 cells per one-wave workgroup:
// the 64 CellState3D records (16 KiB) arrive in LDS by 16 B loads, lane c
// packs cell c's record into LDS, and the 6.5 KiB of records leave by 8 B
// stores (a record is 104 B).
constexpr int kCell3Vecs = (int)(sizeof(pfh3d_cell_state) / 16);
constexpr int kTraj3Words = (int)(sizeof(pfh3d_traj_cell) / 8);
static_assert(sizeof(pfh3d_cell_state) % 16 == 0 && sizeof(pfh3d_traj_cell) % 8 == 0,
              "record vectors");

// ---------------------------------------------------------------------------
// brick growth: every cell's cube embedded, centred, in a larger cube
// (Sim3D::resize_cells, src/pf3d/adaptive_storage.cu:143-266: copy_centered
// :41-83, translate_state :85-106). new_local = old_local + off keeps every
// value at its world position; the voxels outside the old cube are zero.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k3_grow_fields(const float* src, float* dst, int Bo,
                                                      int Bn, int off) {
    const int n = (int)blockIdx.x / Bn, z = (int)blockIdx.x - n * Bn;   // grid N * Bn
    const int zs = z - off;
    float* d = dst + ((size_t)n * Bn + z) * Bn * Bn;
    const float* sp = src + (size_t)n * Bo * Bo * Bo;
    for (int i = (int)threadIdx.x; i < Bn * Bn; i += (int)blockDim.x) {
        const int y = i / Bn, x = i - y * Bn;
        const int ys = y - off, xs = x - off;
        float v = 0.0f;
        if ((unsigned)xs < (unsigned)Bo && (unsigned)ys < (unsigned)Bo &&
            (unsigned)zs < (unsigned)Bo)
            v = sp[((size_t)zs * Bo + ys) * Bo + xs];
        d[i] = v;
    }
}

// The state of a moved cube: origin - off, first moments + off V on each
// axis, a non-empty support box + off, no pending shift, the support flags
// cleared (in the cell and, by thread 0, the population counters).
__global__ __launch_bounds__(256) void k3_grow_state(pfh3d_cell_state* cells, uint32_t* flags,
                                                     int N, int off) {
    const int n = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (n == 0) {
        flags[PFH3D_FLAG_SUPPORT_EXHAUSTED] = 0u;
        flags[PFH3D_FLAG_SUPPORT_EDGE] = 0u;
    }
    if (n >= N) return;
    pfh3d_cell_state* c = cells + n;
    c->origin_x -= off;
    c->origin_y -= off;
    c->origin_z -= off;
    c->Cx += (double)off * c->V;
    c->Cy += (double)off * c->V;
    c->Cz += (double)off * c->V;
    if (c->bb_hi_x >= c->bb_lo_x) {
        c->bb_lo_x += off; c->bb_hi_x += off;
        c->bb_lo_y += off; c->bb_hi_y += off;
        c->bb_lo_z += off; c->bb_hi_z += off;
    }
    c->pending_shift_x = 0;
    c->pending_shift_y = 0;
    c->pending_shift_z = 0;
    c->flags &= ~((1u << PFH3D_FLAG_SUPPORT_EXHAUSTED) | (1u << PFH3D_FLAG_SUPPORT_EDGE));
}

__global__ __launch_bounds__(64) void k3_pack_traj(const pfh3d_cell_state* cells,
                                                   pfh3d_traj_cell* out, int N) {
    __shared__ u32x4 in[64 * kCell3Vecs];
    __shared__ uint64_t rec[64 * kTraj3Words];
    const int lane = (int)threadIdx.x;
    const int base = (int)blockIdx.x * 64;
    const int cnt = min(64, N - base);
    const u32x4* src = reinterpret_cast<const u32x4*>(cells + base);
    for (int i = lane; i < cnt * kCell3Vecs; i += 64) in[i] = src[i];
    __syncthreads();
    if (lane < cnt) {
        const pfh3d_cell_state& s =
            *reinterpret_cast<const pfh3d_cell_state*>(&in[lane * kCell3Vecs]);
        pfh3d_traj_cell& r = *reinterpret_cast<pfh3d_traj_cell*>(&rec[lane * kTraj3Words]);
        const double inv = s.V > 0.0 ? 1.0 / s.V : 0.0;
        r.global_id = s.global_id;
        r.x = (double)s.origin_x + s.Cx * inv;
        r.y = (double)s.origin_y + s.Cy * inv;
        r.z = (double)s.origin_z + s.Cz * inv;
        r.volume = s.V;
        r.surface = s.surface;
        r.wall_overlap = s.wall_overlap;
        r.outside_slit_volume = s.outside_slit_volume;
        r.velocity_x = s.velocity_x;
        r.velocity_y = s.velocity_y;
        r.velocity_z = s.velocity_z;
        r.polarity_x = s.polarity_x;
        r.polarity_y = s.polarity_y;
        r.polarity_z = s.polarity_z;
        r.gamma = s.gamma;
        r.v_A = s.v_A;
        r.phi_max = s.phi_max;
        reinterpret_cast<uint32_t*>(&r)[25] = 0u;      // the 4 B of tail padding
    }
    __syncthreads();
    uint64_t* dst = reinterpret_cast<uint64_t*>(out + base);
    for (int i = lane; i < cnt * kTraj3Words; i += 64) dst[i] = rec[i];
}

}  // namespace pfh3

// -------------------------------------------------------------------------
// C ABI
// -------------------------------------------------------------------------
using namespace pfh_host;

// Scratch of the allocation-free step: 8192 shard partials (96 B each)
// followed by 8192 arrival words, 800 KiB per device, held until the
// process ends. Every call on a device shares it.
constexpr int kNoAllocMaxPartials = 8192;
constexpr int kNoAllocMaxCells = 8192;
constexpr size_t kNoAllocPartialBytes =
    (size_t)kNoAllocMaxPartials * sizeof(pfh3d_moment_partial);
constexpr size_t kNoAllocBytes = kNoAllocPartialBytes + (size_t)kNoAllocMaxCells * 4u;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, resolved on the first
// call per device and kept.
int noalloc_scratch(void** out) {
    int dev = 0;
    int st = hip_status(hipGetDevice(&dev), "hipGetDevice");
    if (st != PFH_OK) return st;
    if (dev < 0 || dev >= kMaxDevices) return invalid("device index");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        st = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                        "hipGetSymbolAddress(scratch)");
        if (st != PFH_OK) return st;
        if (!p) return hip_status(hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return PFH_OK;
}


int block_threads(int B) {
    const int g = B * B / 4;
    return ((g + 63) / 64) * 64;
}
// The update's y parts per slab: bricks from 32 split their rows in two, so
// that a workgroup is at most 7 waves at B = 56 and two of them are resident
// on a compute unit (one 13-wave workgroup otherwise): 208 against 212 us per
// step at 256^3 / 128 cells.
int update_parts(int B) { return B >= 32 ? 2 : 1; }
int update_threads(int B) {
    const int g = (B / update_parts(B) + 2) * (B / 4);   // rows + two halo rows
    return ((g + 63) / 64) * 64;
}
size_t measure_lds(int B) { return 4u * (size_t)(B + 2) * (size_t)(B + 4) * sizeof(float); }
size_t update_lds(int B) {
    return 2u * (size_t)(B + 2 * pfh3::kMargin + 1) * (size_t)pfh3::update_row_pitch(B) *
           sizeof(float);
}

int check_geometry(const pfh3d_step_args* a) {
    if (!a) return invalid("args");
    if (a->N <= 0 || a->B < 8 || a->B > PFH3D_MAX_EDGE || a->B % 8 != 0)
        return invalid("brick edge must be a multiple of 8 in [8, 64]");
    if (a->L <= a->B || a->P < a->L || a->P % 4 != 0)
        return invalid("domain side must exceed the brick edge; pitch >= side, multiple of 4");
    return PFH_OK;
}

int launch_measure(const pfh3d_step_args* a, const pfh3::MeasureMode& m, void* stream) {
    int st = check_geometry(a);
    if (st != PFH_OK) return st;
    if (!a->phi_in || !a->S_in || !a->cells || !a->flags || !a->counters)
        return invalid("measure buffers");
    if (a->shards < 1 || a->shards > a->B) return invalid("shards must lie in [1, B]");
    if (a->shards > 1 && (!a->partials || !a->arrivals)) return invalid("partials/arrivals");
    if (a->max_shift < 0 || a->max_shift > PFH3D_MAX_SHIFT) return invalid("max_shift");
    if (m.clear_out && !a->S_out) return invalid("S_out");
    if (m.snapshot && !a->accepted) return invalid("accepted");
    // four ring planes: 55.7 KB at B = 56, 71.8 KB at B = 64 (above the
    // default 64 KB of dynamic LDS, so the limit is raised for the kernel)
    const size_t lds = measure_lds(a->B);
    if (lds > 65536u) {
        st = hip_status(hipFuncSetAttribute(reinterpret_cast<const void*>(&pfh3::k3_measure),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)lds),
                        "hipFuncSetAttribute(k3_measure)");
        if (st != PFH_OK) return st;
    }
    hipLaunchKernelGGL(pfh3::k3_measure, dim3((unsigned)(a->N * a->shards)),
                       dim3((unsigned)block_threads(a->B)), lds, as_stream(stream), *a, m);
    return launch_status("k3_measure");
}

}  // namespace

extern "C" {

int pfh3d_update_slabs(int N, int B, int device) {
    if (N <= 0 || B < 8 || B > PFH3D_MAX_EDGE || B % 8 != 0) return invalid("N/B");
    int cus = 0;
    const int st = hip_status(
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device),
        "hipDeviceGetAttribute(multiProcessorCount)");
    if (st != PFH_OK) return st;
    // one full round of two resident workgroups per compute unit, counting
    // the y parts of each slab: a partial round leaves part of the chip idle
    // (at 128 cells, two parts: 2 slabs 208 us per step, 4 slabs 218, 8
    // slabs 241); slabs of >= 4 planes
    int k = (2 * (cus > 0 ? cus : 1)) / (N * update_parts(B));
    if (k < 1) k = 1;
    if (k > B / 4) k = B / 4;
    return k;
}

int pfh3d_init_spheres(float* phi_even, float* phi_odd, pfh3d_cell_state* cells,
                       const double* centres, int N, int B, int L, float lambda,
                       float seed_radius, void* stream) {
    if (!phi_even || !cells || !centres || N <= 0 || B < 8 || B % 8 != 0 || L <= B ||
        !(lambda > 0.0f) || !(seed_radius > 0.0f))
        return invalid("init_spheres");
    hipLaunchKernelGGL(pfh3::k3_init_spheres, dim3((unsigned)N), dim3(256), 0,
                       as_stream(stream), phi_even, phi_odd, cells, centres, B, lambda,
                       seed_radius);
    return launch_status("k3_init_spheres");
}

int pfh3d_clear_S(uint32_t* S, int L, int P, void* stream) {
    if (!S || L <= 0 || P < L || P % 4 != 0) return invalid("clear_S");
    const size_t n4 = (size_t)L * L * P / 4u;
    hipLaunchKernelGGL(pfh3::k3_clear, dim3(1024), dim3(256), 0, as_stream(stream), S, n4);
    return launch_status("k3_clear");
}

int pfh3d_scatter_current(const float* phi, pfh3d_cell_state* cells, uint32_t* S,
                          uint32_t* flags, int N, int B, int L, int P, void* stream) {
    if (!phi || !cells || !S || !flags || N <= 0 || B < 8 || B % 8 != 0 || L <= B || P < L)
        return invalid("scatter_current");
    hipLaunchKernelGGL(pfh3::k3_scatter, dim3((unsigned)(N * B)), dim3(256), 0,
                       as_stream(stream), phi, cells, S, flags, B, L, P);
    return launch_status("k3_scatter");
}

int pfh3d_measure_cells(const pfh3d_step_args* a, int apply_tumble, int compute_surface,
                        int apply_motion, void* stream) {
    pfh3::MeasureMode m{apply_tumble ? 1 : 0, compute_surface ? 1 : 0, apply_motion ? 1 : 0,
                        0, 0};
    return launch_measure(a, m, stream);
}

int pfh3d_update(const pfh3d_step_args* a, int slabs, void* stream) {
    int st = check_geometry(a);
    if (st != PFH_OK) return st;
    if (!a->phi_in || !a->phi_out || a->phi_in == a->phi_out || !a->S_in || !a->S_out ||
        a->S_in == a->S_out || !a->cells || !a->flags)
        return invalid("update buffers");
    int k = slabs;
    if (k <= 0) {
        int dev = 0;
        st = hip_status(hipGetDevice(&dev), "hipGetDevice");
        if (st != PFH_OK) return st;
        k = pfh3d_update_slabs(a->N, a->B, dev);
        if (k <= 0) return k;
    }
    if (k > a->B) return invalid("slabs must lie in [1, B]");
    const size_t lds = update_lds(a->B);    // <= 46 KB: within the default 64 KB
    const int H = update_parts(a->B);
    hipLaunchKernelGGL(pfh3::k3_update, dim3((unsigned)(a->N * k * H)),
                       dim3((unsigned)update_threads(a->B)), lds, as_stream(stream), *a, k, H);
    return launch_status("k3_update");
}

int pfh3d_finish_step(const pfh3d_step_args* a, void* stream) {
    int st = check_geometry(a);
    if (st != PFH_OK) return st;
    if (!a->phi_in || !a->phi_out || !a->S_in || !a->S_out || !a->cells || !a->accepted ||
        !a->flags || !a->counters)
        return invalid("finish buffers");
    const int grid = a->N > 512 ? a->N : 512;
    hipLaunchKernelGGL(pfh3::k3_finish, dim3((unsigned)grid), dim3(256), 0, as_stream(stream),
                       *a);
    return launch_status("k3_finish");
}

int pfh3d_step(const pfh3d_step_args* a, int slabs, void* stream) {
    pfh3::MeasureMode m{1, 2, 1, 1, 1};
    int st = launch_measure(a, m, stream);
    if (st != PFH_OK) return st;
    st = pfh3d_update(a, slabs, stream);
    if (st != PFH_OK) return st;
    return pfh3d_finish_step(a, stream);
}

// ---- allocation-free step ----
// pfh3d_step with the shard partials and arrival words in the file-scope
// array `scratch` instead of caller buffers. Host code never passes the
// array's name to a kernel: it resolves the array's address on each device
// once (hipGetSymbolAddress, with that device current) and passes that.
int pfh3d_step_noalloc(const pfh3d_step_args* a, int slabs, void* stream) {
    if (!a) return invalid("args");
    if (a->N > kNoAllocMaxCells || (a->N > 0 && a->shards > 0 &&
                                     (long long)a->N * a->shards > kNoAllocMaxPartials))
        return invalid("step_noalloc: N must be <= 8192 and N * shards <= 8192 "
                       "(the allocation-free scratch holds 8192 shard partials)");
    void* base = nullptr;
    int st = noalloc_scratch(&base);
    if (st != PFH_OK) return st;
    pfh3d_step_args b = *a;
    b.partials = reinterpret_cast<pfh3d_moment_partial*>(base);
    b.arrivals = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(base) +
                                             kNoAllocPartialBytes);
    if (b.N > 0) {
        // the arrival words must start at zero: a fill kernel, so that the
        // call can be captured in a graph
        const size_t n4 = ((size_t)b.N + 3u) / 4u;
        hipLaunchKernelGGL(pfh3::k3_clear, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0,
                           as_stream(stream), b.arrivals, n4);
        st = launch_status("k3_clear(arrivals)");
        if (st != PFH_OK) return st;
    }
    return pfh3d_step(&b, slabs, stream);
}

int pfh3d_pack_trajectory(const pfh3d_cell_state* cells, pfh3d_traj_cell* out, int N,
                          void* stream) {
    if (!cells || !out || N <= 0) return invalid("pack_trajectory");
    if ((uintptr_t)cells % 16 || (uintptr_t)out % 8)
        return invalid("pack_trajectory: cells 16 B and out 8 B aligned");
    hipLaunchKernelGGL(pfh3::k3_pack_traj, dim3((unsigned)((N + 63) / 64)), dim3(64), 0,
                       as_stream(stream), cells, out, N);
    return launch_status("k3_pack_traj");
}

int pfh3d_grow_bricks(const float* phi_old, float* phi_new, pfh3d_cell_state* cells,
                      uint32_t* flags, int N, int B_old, int B_new, void* stream) {
    if (!phi_old || !phi_new || phi_old == phi_new || !cells || !flags || N <= 0)
        return invalid("grow_bricks buffers");
    if (B_old < 8 || B_old % 8 != 0 || B_new % 8 != 0 || B_new <= B_old ||
        B_new > PFH3D_MAX_EDGE)
        return invalid("grow_bricks: edges multiples of 8 with 8 <= B_old < B_new <= 64");
    const int off = (B_new - B_old) / 2;
    hipLaunchKernelGGL(pfh3::k3_grow_fields, dim3((unsigned)(N * B_new)), dim3(256), 0,
                       as_stream(stream), phi_old, phi_new, B_old, B_new, off);
    int st = launch_status("k3_grow_fields");
    if (st != PFH_OK) return st;
    hipLaunchKernelGGL(pfh3::k3_grow_state, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                       as_stream(stream), cells, flags, N, off);
    return launch_status("k3_grow_state");
}

// ---- host-side set-up ----
int pfh3d_next_brick_edge(int current, int L) {
    // next_brick_edge (include/pf3d/brick_classes.hpp:23-32): 32 more planes,
    // rounded up to 16, within the largest aligned edge below the side
    // (src/pf3d/sim.cu:291-293) and this build's largest brick
    if (current <= 0 || current % 8 != 0 || L <= 1) return 0;
    int maximum = ((L - 1) / 8) * 8;
    if (maximum > PFH3D_MAX_EDGE) maximum = PFH3D_MAX_EDGE;
    if (maximum <= current) return 0;
    const int rounded = ((current + 32 + 15) / 16) * 16;
    return rounded < maximum ? rounded : maximum;
}

int pfh3d_brick_edge(double radius, double lambda) {
    if (!(radius > 0.0) || !(lambda > 0.0)) return invalid("radius/lambda");
    const double k = 2.7386127875258306 / lambda;
    const double eff = radius + 0.5 / k;
    const double tail = log(1.0 / 1.0e-5 - 1.0) / (2.0 * k);
    const double req = 2.0 * (eff + tail) + 8.0;
    if (!isfinite(req) || req > 2147483000.0) return invalid("brick size");
    const int c = (int)ceil(req);
    return ((c + 7) / 8) * 8;
}

int pfh3d_domain_side(int num_cells, double radius, double rho) {
    if (num_cells <= 0 || !(radius > 0.0) || !(rho > 0.0)) return invalid("domain side");
    const double vol = (double)num_cells * (4.0 * pfh::kPi * radius * radius * radius / 3.0) / rho;
    const double side = ceil(cbrt(vol));
    if (!isfinite(side) || side < 1.0 || side > 2147483647.0) return invalid("domain side");
    return (int)side;
}

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

int pfh3d_calibrate_seed_radius(double radius, double lambda, int B, double* seed_radius) {
    if (!seed_radius || !(radius > 0.0) || !(lambda > 0.0) || B < 3)
        return invalid("calibrate_seed_radius");
    const double target = 4.0 * pfh::kPi * radius * radius * radius / 3.0;
    double lo = 0.0, hi = (double)B;
    if (seeded_volume(B, lo, lambda) > target || seeded_volume(B, hi, lambda) < target)
        return invalid("target volume is not bracketed by the brick");
    for (int it = 1; it <= 80; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double v = seeded_volume(B, mid, lambda);
        *seed_radius = mid;
        if (fabs(v - target) / target <= 1.0e-10) return it;
        if (v < target) lo = mid; else hi = mid;
    }
    return unsupported("seed-radius bisection did not converge");
}

int pfh3d_sequential_centres(int N, double side, double radius, uint64_t seed, float* x,
                             float* y, float* z) {
    if (N < 1 || !(side > 0.0) || !(radius > 0.0) || !x || !y || !z)
        return invalid("sequential_centres");
    auto min_image = [side](double d) { return d - side * floor(d / side + 0.5); };
    std::mt19937_64 rng(seed);
    auto u53 = [&rng]() { return (double)(rng() >> 11) * 0x1.0p-53; };
    x[0] = y[0] = z[0] = (float)(0.5 * side);
    int count = 1;
    long long drawn = 0;
    const double r2 = radius * radius;
    while (count < N) {
        if (++drawn > 1000000000LL) return unsupported("placement did not terminate");
        const double cx = side * u53();
        const double cy = side * u53();
        const double cz = side * u53();
        bool ok = true;
        for (int j = 0; j < count && ok; ++j) {
            const double dx = min_image(cx - x[j]), dy = min_image(cy - y[j]),
                         dz = min_image(cz - z[j]);
            ok = dx * dx + dy * dy + dz * dz >= r2;
        }
        if (!ok) continue;
        x[count] = (float)cx;
        y[count] = (float)cy;
        z[count] = (float)cz;
        ++count;
    }
    return drawn > 2147483647LL ? 2147483647 : (int)drawn;
}

int pfh3d_initial_polarity(int64_t global_id, uint64_t stream, double out[3]) {
    if (!out) return invalid("out");
    const uint64_t g = (uint64_t)global_id;
    const pfh::Philox4 r = pfh::philox4x32_10((uint32_t)g, (uint32_t)(g >> 32), 0x504F4C49u,
                                              0u, (uint32_t)stream, (uint32_t)(stream >> 32));
    const double au = pfh::uniform53(r.v[0], r.v[1]);
    const double zu = pfh::uniform53(r.v[2], r.v[3]);
    const double zz = 2.0 * au - 1.0;
    const double rsq = 1.0 - zz * zz;
    const double radial = sqrt(rsq > 0.0 ? rsq : 0.0);
    double s, c;
    pfh::sincos_d(2.0 * pfh::kPi * zu, &s, &c);
    out[0] = radial * c;
    out[1] = radial * s;
    out[2] = zz;
    return PFH_OK;
}

}  // extern "C"
