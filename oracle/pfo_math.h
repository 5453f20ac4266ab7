/*
 * TEST INFRASTRUCTURE ONLY. This header belongs to the CPU oracle, which is
 * the checker for the HIP kernels: only tests, the smoke check and the CPU
 * baseline leg of the benchmark may load it. No product code includes it.
 *
 * Scalar primitives of the 2D/3D per-cell phase-field update, restated in C
 * from the reference:
 *   philox4x32_10, philox_uniform53   common/philox.cuh:32-67
 *   coefficient helpers               common/model_coefficients.cuh:16-53
 *   q_of (Q5.27)                      include/params.cuh:94-111
 *
 * The device sincosf/tanhf used by the reference cannot be reproduced (their
 * bit patterns are a closed library's), so this project evaluates sin, cos and
 * tanh with its own double-precision routines, identical on CPU and GPU, and
 * rounds to binary32 once:
 *   - argument reduction by Cody and Waite (1980) with a three-part pi/2
 *     (the fdlibm split, accurate to 1e-37) or ln2 split;
 *   - Taylor polynomials in Horner form with fused multiply-add, truncated
 *     far below double rounding on the reduced interval (|r| <= pi/4:
 *     sin to r^17, cos to r^18; |r| <= ln2/2: exp to r^14; expm1 below 0.5 to
 *     x^16).
 * Both sides execute the same correctly rounded double operations, so the
 * binary32 results are bit-identical between this oracle and the kernels.
 * tests/test_oracle_primitives.py checks them against libm.
 */
#ifndef PFO_MATH_H
#define PFO_MATH_H

#include <math.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint32_t v[4]; } pfo_philox4;

/* Philox4x32-10 (Salmon et al., SC'11). common/philox.cuh:32-58. */
static inline pfo_philox4 pfo_philox4x32_10(uint32_t c0, uint32_t c1,
                                            uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += W0; k1 += W1; }
        const uint64_t p0 = (uint64_t)M0 * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)M1 * (uint64_t)c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    pfo_philox4 r = {{c0, c1, c2, c3}};
    return r;
}

/* u = ((a<<32 | b) >> 11) * 2^-53 on [0,1). common/philox.cuh:61-67. */
static inline double pfo_uniform53(uint32_t a, uint32_t b) {
    const uint64_t m = (((uint64_t)a << 32) | (uint64_t)b) >> 11;
    return (double)m * (1.0 / 9007199254740992.0);
}

/* Q5.27: q = RN(min(phi*phi, 4) * 2^27). include/params.cuh:103-111.
 * The square is one binary32 multiply (no contraction), the scaling by 2^27
 * is exact, and rint rounds half to even like __float2uint_rn. */
#define PFO_QSCALE 134217728.0f
#define PFO_QINV   7.45058059692382812e-09f
static inline uint32_t pfo_q_of(float phi) {
    float s = phi * phi;
    s = fminf(s, 4.0f);
    return (uint32_t)rintf(s * PFO_QSCALE);
}
/* FP64 variant used only by the FP64 oracle: the square is taken in double. */
static inline uint32_t pfo_q_of_d(double phi) {
    double s = phi * phi;
    s = fmin(s, 4.0);
    return (uint32_t)rint(s * 134217728.0);
}

/* ---- project-written sin/cos/tanh in double (see header comment) ---- */
static inline void pfo_sincos_d(double x, double* s, double* c) {
    const double P1 = 0x1.921fb544p+0, P2 = 0x1.0b4611a6p-34,
                 P3 = 0x1.3198a2e037073p-69;
    const double k = rint(x * 0x1.45f306dc9c883p-1);   /* x * 2/pi */
    const double r = ((x - k * P1) - k * P2) - k * P3;
    const double r2 = r * r;
    /* sin r = r + r^3 * (-1/3! + r^2 (1/5! - ...)) up to r^17 */
    double ps = 0x1.952c77030ad4ap-49;                 /* 1/17! */
    ps = fma(ps, r2, -0x1.ae7f3e733b81fp-41);          /* -1/15! */
    ps = fma(ps, r2, 0x1.6124613a86d09p-33);           /* 1/13! */
    ps = fma(ps, r2, -0x1.ae64567f544e4p-26);          /* -1/11! */
    ps = fma(ps, r2, 0x1.71de3a556c734p-19);           /* 1/9! */
    ps = fma(ps, r2, -0x1.a01a01a01a01ap-13);          /* -1/7! */
    ps = fma(ps, r2, 0x1.1111111111111p-7);            /* 1/5! */
    ps = fma(ps, r2, -0x1.5555555555555p-3);           /* -1/3! */
    const double sr = fma(r * r2, ps, r);
    /* cos r = 1 + r^2 * (-1/2! + r^2 (1/4! - r^2 (1/6! - ...))) to r^18 */
    double pc = -0x1.6827863b97d97p-53;                /* -1/18! */
    pc = fma(pc, r2, 0x1.ae7f3e733b81fp-45);           /* 1/16! */
    pc = fma(pc, r2, -0x1.93974a8c07c9dp-37);          /* -1/14! */
    pc = fma(pc, r2, 0x1.1eed8eff8d898p-29);           /* 1/12! */
    pc = fma(pc, r2, -0x1.27e4fb7789f5cp-22);          /* -1/10! */
    pc = fma(pc, r2, 0x1.a01a01a01a01ap-16);           /* 1/8! */
    pc = fma(pc, r2, -0x1.6c16c16c16c17p-10);          /* -1/6! */
    pc = fma(pc, r2, 0x1.5555555555555p-5);            /* 1/4! */
    pc = fma(pc, r2, -0.5);                            /* -1/2! */
    const double cr = fma(r2, pc, 1.0);
    const long q = (long)k & 3;
    switch (q) {
        case 0: *s = sr;  *c = cr;  break;
        case 1: *s = cr;  *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}

/* Polarity vector (cos theta, sin theta) in binary32, replacing sincosf. */
static inline void pfo_sincosf_theta(float theta, float* s, float* c) {
    double sd, cd;
    pfo_sincos_d((double)theta, &sd, &cd);
    *s = (float)sd;
    *c = (float)cd;
}

/* expm1 for x >= 0 (x <= 64). */
static inline double pfo_expm1_pos_d(double x) {
    if (x < 0.5) {
        double p = 0x1.ae7f3e733b81fp-45;              /* 1/16! */
        p = fma(p, x, 0x1.ae7f3e733b81fp-41);          /* 1/15! */
        p = fma(p, x, 0x1.93974a8c07c9dp-37);          /* 1/14! */
        p = fma(p, x, 0x1.6124613a86d09p-33);          /* 1/13! */
        p = fma(p, x, 0x1.1eed8eff8d898p-29);          /* 1/12! */
        p = fma(p, x, 0x1.ae64567f544e4p-26);          /* 1/11! */
        p = fma(p, x, 0x1.27e4fb7789f5cp-22);          /* 1/10! */
        p = fma(p, x, 0x1.71de3a556c734p-19);          /* 1/9! */
        p = fma(p, x, 0x1.a01a01a01a01ap-16);          /* 1/8! */
        p = fma(p, x, 0x1.a01a01a01a01ap-13);          /* 1/7! */
        p = fma(p, x, 0x1.6c16c16c16c17p-10);          /* 1/6! */
        p = fma(p, x, 0x1.1111111111111p-7);           /* 1/5! */
        p = fma(p, x, 0x1.5555555555555p-5);           /* 1/4! */
        p = fma(p, x, 0x1.5555555555555p-3);           /* 1/3! */
        p = fma(p, x, 0.5);                            /* 1/2! */
        return fma(x * x, p, x);
    }
    const double L1 = 0x1.62e42feep-1, L2 = 0x1.a39ef35793c76p-33;
    const double k = rint(x * 0x1.71547652b82fep+0);   /* x / ln2 */
    const double r = (x - k * L1) - k * L2;
    double p = 0x1.93974a8c07c9dp-37;                  /* 1/14! */
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
    const double em = fma(r * r, p, r);                /* exp(r) - 1 */
    return ldexp(1.0 + em, (int)k) - 1.0;
}

static inline double pfo_tanh_d(double y) {
    const double a = fabs(y);
    double t;
    if (!(a <= 20.0)) {
        t = (a > 20.0) ? 1.0 : a;                      /* NaN propagates */
    } else {
        const double e = pfo_expm1_pos_d(2.0 * a);
        t = e / (e + 2.0);
    }
    return copysign(t, y);
}

/* ---- coefficient convention (common/model_coefficients.cuh:16-53) ---- */
static inline double pfo_bulk_coeff(double lambda) { return 30.0 / (lambda * lambda); }
static inline double pfo_interaction_coeff(double kappa, double lambda) {
    return 60.0 * kappa / (lambda * lambda);
}
static inline double pfo_motility_coeff(double kappa, double xi, double lambda) {
    return pfo_interaction_coeff(kappa, lambda) / xi;
}
static inline double pfo_volume_coeff(double mu, double target) { return 2.0 * mu / target; }
static inline double pfo_interface_k(double lambda) { return 2.7386127875258306 / lambda; }
static inline double pfo_init_radius(double r, double lambda) {
    return r + 0.5 / pfo_interface_k(lambda);
}

#ifdef __cplusplus
}
#endif
#endif /* PFO_MATH_H */
