// Device primitives of the gfx950 phase-field kernels: counter-based RNG,
// Q5.27 quantization, shape classes, deterministic double-precision sin/cos
// and tanh, and wavefront (64-lane) reductions.
//
// Reference counterparts:
//   philox4x32_10 / philox_uniform53   common/philox.cuh:32-67
//   q_of, shape classes, flags         include/params.cuh:94-111,121-242,434-445
//   phase_field_rhs                    include/kernels.cuh:121-138
//
// sin/cos/tanh: the reference's device sincosf/tanhf are a closed library's
// and cannot be reproduced, so these are this project's own routines
// (Cody-Waite argument reduction with the fdlibm three-part pi/2 and ln2
// splits, then Taylor polynomials in Horner form with fused multiply-add,
// truncated below double rounding), evaluated in double and rounded to
// binary32 once. The CPU oracle executes the same double operations, so the
// results are bit-identical on both sides.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pfh.h"

namespace pfh {

constexpr double kPi = 3.14159265358979323846;

// ---- workgroup shape: 16 wavefronts of 64 lanes ----
constexpr int kNT = 1024;
constexpr int kWave = 64;
constexpr int kWaves = kNT / kWave;
// Threads that may own pixels in the LDS path (column bands).
constexpr int kControlLanes = 16;   // lanes 0-15 of wave 0: next-cell fetch
constexpr int kPixThreads = kNT - kControlLanes;

// ---- fixed tile and shape classes (include/params.cuh:121-147) ----
constexpr int kTilePitch = 288;
constexpr int kTileArea = kTilePitch * kTilePitch;
constexpr int kNumClasses = 6;
constexpr int kClassFallback = 5;
constexpr int kClassStagedMax = 3;   // classes staged in LDS on gfx950
constexpr int kPromoteSlack = 8;
constexpr int kDemoteSlack = 20;
constexpr int kDemoteDwell = 200;
constexpr int kMaxShiftPerStep = 4;
constexpr float kSupportEps = 1e-5f;

struct Shape { int wx, wy, tx0, ty0; };
constexpr Shape kShapes[kNumClasses] = {
    {144, 144, 64, 64}, {176, 144, 32, 64}, {144, 176, 64, 32},
    {160, 160, 32, 32}, {224, 224, 32, 32}, {286, 286, 1, 1},
};

__host__ __device__ inline Shape shape_of(int c) {
    // Literal selection keeps geometry in immediates for runtime class ids.
    return c == 1 ? kShapes[1] : c == 2 ? kShapes[2] : c == 3 ? kShapes[3]
         : c == 4 ? kShapes[4] : c == 5 ? kShapes[5] : kShapes[0];
}

__host__ __device__ inline int class_containing(int ex, int ey, int slack) {
    int best = -1, best_area = 0;
    for (int c = 0; c < kNumClasses; ++c) {
        const Shape s = shape_of(c);
        const int area = s.wx * s.wy;
        if (ex + slack <= s.wx && ey + slack <= s.wy &&
            (best < 0 || area < best_area)) {
            best = c;
            best_area = area;
        }
    }
    return best;
}

// Staged-class column bands: NB = 1024/WX bands of RB = ceil(WY/NB) rows.
template <int CLS> struct Staged {
    static constexpr int WX = kShapes[CLS].wx;
    static constexpr int WY = kShapes[CLS].wy;
    static constexpr int TX0 = kShapes[CLS].tx0;
    static constexpr int TY0 = kShapes[CLS].ty0;
    static constexpr int NB = kPixThreads / WX;
    static constexpr int RB = (WY + NB - 1) / NB;
    static constexpr int RFULL = WY - (NB - 1) * RB;    // rows every band holds
    static constexpr int NCH = WX / 4;                  // 16 B chunks per row
    static constexpr int NCHUNK = WY * NCH;
    static constexpr int KCH = (NCHUNK + kNT - 1) / kNT;
    static constexpr int PX = WX + 4;                   // LDS row pitch
    static constexpr int LDS_FLOATS = (WY + 2) * PX + 4;
    static_assert(WX % 4 == 0 && TX0 % 4 == 0, "16 B phi rows");
    static_assert(NB * RB >= WY, "bands must cover the window");
};
constexpr int max2(int a, int b) { return a > b ? a : b; }

// Class 0 keeps its LDS window in two pieces, so that the next cell's window
// can be copied in while the current one is still being read: piece 1 holds
// rows -1..R1 (read by bands 0..KB-1) in one of two slots, piece 2 rows
// P2_R0..WY (read by bands KB..NB-1). Rows P2_R0 and R1 are in both pieces.
// Each piece has the padded row layout of the whole window, with four floats
// after its last row (that row's east pad).
struct Split0 {
    using G = Staged<0>;
    static constexpr int KB = 5;
    static constexpr int R1 = KB * G::RB;              // first row of band KB
    static constexpr int P2_R0 = R1 - 1;
    static constexpr int P1_ROWS = R1 + 2;             // rows -1..R1
    static constexpr int P2_ROWS = G::WY + 1 - P2_R0;  // rows P2_R0..WY
    static constexpr int P1_FLOATS = P1_ROWS * G::PX + 4;
    static constexpr int P2_FLOATS = P2_ROWS * G::PX + 4;
    static constexpr int FLOATS = 2 * P1_FLOATS + P2_FLOATS;
    static_assert(KB < G::NB && (G::NB - 1) * G::RB < G::WY, "two non-empty pieces");
    static_assert(P1_FLOATS % 4 == 0 && P2_FLOATS % 4 == 0, "16 B piece alignment");
};

constexpr int kPhiLdsFloats =
    max2(max2(Split0::FLOATS, Staged<1>::LDS_FLOATS),
         max2(Staged<2>::LDS_FLOATS, Staged<3>::LDS_FLOATS));

// ---- Q5.27 ----
constexpr float kQScale = 134217728.0f;             // 2^27
constexpr float kQInv = 7.45058059692382812e-09f;   // 2^-27
constexpr float kQClampPhiSq = 4.0f;

__device__ __forceinline__ uint32_t q_of(float phi) {
    float s = phi * phi;
    s = fminf(s, kQClampPhiSq);
    return __float2uint_rn(s * kQScale);
}

__device__ __forceinline__ int wrapi(int v, int L) {
    v %= L;
    return v < 0 ? v + L : v;
}

// v in [0, 2L): a window coordinate (origin in [0, L), offset below the
// widest class window, which is narrower than L) folded onto the domain
// without a division.
__device__ __forceinline__ int wrap_once(int v, int L) { return v >= L ? v - L : v; }

// ---- Philox4x32-10 ----
struct Philox4 { uint32_t v[4]; };

__host__ __device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1,
                                                 uint32_t c2, uint32_t c3,
                                                 uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    Philox4 r;
    r.v[0] = c0; r.v[1] = c1; r.v[2] = c2; r.v[3] = c3;
    return r;
}

__host__ __device__ inline double uniform53(uint32_t a, uint32_t b) {
    const uint64_t m = (((uint64_t)a << 32) | (uint64_t)b) >> 11;
    return (double)m * (1.0 / 9007199254740992.0);
}

// ---- deterministic sin/cos/tanh (see header) ----
__host__ __device__ inline void sincos_d(double x, double* s, double* c) {
    const double P1 = 0x1.921fb544p+0, P2 = 0x1.0b4611a6p-34,
                 P3 = 0x1.3198a2e037073p-69;
    const double k = rint(x * 0x1.45f306dc9c883p-1);
    const double r = ((x - k * P1) - k * P2) - k * P3;
    const double r2 = r * r;
    double ps = 0x1.952c77030ad4ap-49;
    ps = fma(ps, r2, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, r2, 0x1.6124613a86d09p-33);
    ps = fma(ps, r2, -0x1.ae64567f544e4p-26);
    ps = fma(ps, r2, 0x1.71de3a556c734p-19);
    ps = fma(ps, r2, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, r2, 0x1.1111111111111p-7);
    ps = fma(ps, r2, -0x1.5555555555555p-3);
    const double sr = fma(r * r2, ps, r);
    double pc = -0x1.6827863b97d97p-53;
    pc = fma(pc, r2, 0x1.ae7f3e733b81fp-45);
    pc = fma(pc, r2, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, r2, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, r2, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, r2, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, r2, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, r2, 0x1.5555555555555p-5);
    pc = fma(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    switch ((long)k & 3) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}

__host__ __device__ inline double expm1_pos_d(double x) {
    if (x < 0.5) {
        double p = 0x1.ae7f3e733b81fp-45;
        p = fma(p, x, 0x1.ae7f3e733b81fp-41);
        p = fma(p, x, 0x1.93974a8c07c9dp-37);
        p = fma(p, x, 0x1.6124613a86d09p-33);
        p = fma(p, x, 0x1.1eed8eff8d898p-29);
        p = fma(p, x, 0x1.ae64567f544e4p-26);
        p = fma(p, x, 0x1.27e4fb7789f5cp-22);
        p = fma(p, x, 0x1.71de3a556c734p-19);
        p = fma(p, x, 0x1.a01a01a01a01ap-16);
        p = fma(p, x, 0x1.a01a01a01a01ap-13);
        p = fma(p, x, 0x1.6c16c16c16c17p-10);
        p = fma(p, x, 0x1.1111111111111p-7);
        p = fma(p, x, 0x1.5555555555555p-5);
        p = fma(p, x, 0x1.5555555555555p-3);
        p = fma(p, x, 0.5);
        return fma(x * x, p, x);
    }
    const double L1 = 0x1.62e42feep-1, L2 = 0x1.a39ef35793c76p-33;
    const double k = rint(x * 0x1.71547652b82fep+0);
    const double r = (x - k * L1) - k * L2;
    double p = 0x1.93974a8c07c9dp-37;
    p = fma(p, r, 0x1.6124613a86d09p-33);
    p = fma(p, r, 0x1.1eed8eff8d898p-29);
    p = fma(p, r, 0x1.ae64567f544e4p-26);
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = fma(p, r, 0x1.71de3a556c734p-19);
    p = fma(p, r, 0x1.a01a01a01a01ap-16);
    p = fma(p, r, 0x1.a01a01a01a01ap-13);
    p = fma(p, r, 0x1.6c16c16c16c17p-10);
    p = fma(p, r, 0x1.1111111111111p-7);
    p = fma(p, r, 0x1.5555555555555p-5);
    p = fma(p, r, 0x1.5555555555555p-3);
    p = fma(p, r, 0.5);
    const double em = fma(r * r, p, r);
    return ldexp(1.0 + em, (int)k) - 1.0;
}

__host__ __device__ inline double tanh_d(double y) {
    const double a = fabs(y);
    double t;
    if (!(a <= 20.0)) {
        t = (a > 20.0) ? 1.0 : a;
    } else {
        const double e = expm1_pos_d(2.0 * a);
        t = e / (e + 2.0);
    }
    return copysign(t, y);
}

// ---- workgroup barrier that orders LDS only ----
// __syncthreads() is a release/acquire fence over every address space, so
// each wave first waits for all its outstanding global loads, stores and
// atomics. Loops whose barrier only protects LDS buffers use this instead:
// global prefetches and scatters stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---- wavefront reductions: lane tree with offsets 32..1 ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_down(v, d, kWave);
    return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = min(v, __shfl_down(v, d, kWave));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = max(v, __shfl_down(v, d, kWave));
    return v;
}

// ---- moments of one cell's window with the update's thread mapping ----
// Column bands of pixel threads (pix = tid - kControlLanes, x = pix % wx,
// band = pix / wx, contiguous rows per band) for the LDS classes, flat pixel
// index for classes 4-5; per-thread
// fp64 sums in visiting order, the lane tree, then waves ascending from 0.
struct WindowMoments {
    double V, Cx, Cy;
    int blox, bhix, bloy, bhiy, pmaxb;
};

// Per-thread partial moments and their workgroup reduction (the lane tree,
// then waves ascending from 0.0); the result is valid in thread 0.
struct MomentPart {
    double V, Cx, Cy;
    int blox, bhix, bloy, bhiy, pmaxb;
};

__device__ __forceinline__ void moment_visit(MomentPart& m, float pn, int a, int b) {
    const double d = (double)pn * (double)pn;
    m.V += d;
    m.Cx += d * (double)a;
    m.Cy += d * (double)b;
    m.pmaxb = max(m.pmaxb, (int)__float_as_uint(fabsf(pn)));
    if (pn > kSupportEps) {
        m.blox = min(m.blox, a); m.bhix = max(m.bhix, a);
        m.bloy = min(m.bloy, b); m.bhiy = max(m.bhiy, b);
    }
}

__device__ inline WindowMoments reduce_moments(MomentPart p, int wx, int wy) {
    __shared__ double red[kWaves * 3];
    __shared__ int ired[kWaves * 5];
    const int tid = (int)threadIdx.x;
    const int lane = tid & (kWave - 1), wave = tid / kWave;
    p.V = wave_sum(p.V); p.Cx = wave_sum(p.Cx); p.Cy = wave_sum(p.Cy);
    p.blox = wave_min(p.blox); p.bhix = wave_max(p.bhix);
    p.bloy = wave_min(p.bloy); p.bhiy = wave_max(p.bhiy);
    p.pmaxb = wave_max(p.pmaxb);
    if (lane == 0) {
        red[wave * 3 + 0] = p.V; red[wave * 3 + 1] = p.Cx; red[wave * 3 + 2] = p.Cy;
        ired[wave * 5 + 0] = p.blox; ired[wave * 5 + 1] = p.bhix;
        ired[wave * 5 + 2] = p.bloy; ired[wave * 5 + 3] = p.bhiy;
        ired[wave * 5 + 4] = p.pmaxb;
    }
    __syncthreads();
    WindowMoments m = {0.0, 0.0, 0.0, wx, -1, wy, -1, 0};
    if (tid == 0) {
        for (int w = 0; w < kWaves; ++w) {
            m.V += red[w * 3 + 0]; m.Cx += red[w * 3 + 1]; m.Cy += red[w * 3 + 2];
            m.blox = min(m.blox, ired[w * 5 + 0]); m.bhix = max(m.bhix, ired[w * 5 + 1]);
            m.bloy = min(m.bloy, ired[w * 5 + 2]); m.bhiy = max(m.bhiy, ired[w * 5 + 3]);
            m.pmaxb = max(m.pmaxb, ired[w * 5 + 4]);
        }
    }
    return m;
}

// phi_at(a, b) returns phi at window pixel (a, b); every pixel is visited
// exactly once, so phi_at may also store or scatter. Workgroup of kNT
// threads; the result is valid in thread 0.
template <class F>
__device__ WindowMoments window_moments(int cls, F phi_at) {
    const Shape sc = shape_of(cls);
    const int wx = sc.wx, wy = sc.wy;
    const int tid = (int)threadIdx.x;
    MomentPart p = {0.0, 0.0, 0.0, wx, -1, wy, -1, 0};
    if (cls <= kClassStagedMax) {
        const int nb = kPixThreads / wx, rb = (wy + nb - 1) / nb;
        const int pix = tid - kControlLanes;
        const int x = pix % wx, band = pix / wx;
        if (pix >= 0 && band < nb)
            for (int y = band * rb; y < min(band * rb + rb, wy); ++y)
                moment_visit(p, phi_at(x, y), x, y);
    } else {
        for (int q = tid; q < wx * wy; q += kNT) moment_visit(p, phi_at(q % wx, q / wx), q % wx, q / wx);
    }
    return reduce_moments(p, wx, wy);
}

// Moments into the cell state; the interaction integral restarts at zero.
__device__ __forceinline__ void store_moments(pfh_cell_state* cs, const WindowMoments& m) {
    cs->V = m.V; cs->Cx = m.Cx; cs->Cy = m.Cy;
    cs->bb_lo_x = m.blox; cs->bb_hi_x = m.bhix;
    cs->bb_lo_y = m.bloy; cs->bb_hi_y = m.bhiy;
    cs->phi_max = __uint_as_float((uint32_t)m.pmaxb);
    cs->promote_ctr = 0u;
    cs->Ix = 0.0;
    cs->Iy = 0.0;
}

}  // namespace pfh
