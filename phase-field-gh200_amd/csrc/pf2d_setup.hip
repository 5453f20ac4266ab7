// Host-side set-up of the 2D path (no GPU work): the counter-based draws
// and the domain sizing that a host driver needs before the first launch,
// computed with the same Philox4x32-10 the kernels use.
//
// Reference counterparts:
//   domain_side_for, s_pitch_for      include/params.cuh:562-584
//   ic_theta, ic_v_A                  include/kernels.cuh:147-172
//   default grid + jitter placement   src/sim.cu:108-129
//   centre-table FNV-1a fingerprint   include/palmieri_initializer.hpp:105-122
#include <cmath>
#include <cstring>

#include "pfh.h"
#include "pfh_device.h"
#include "pfh_host.h"

using pfh_host::invalid;

namespace {
// Counter domains of the initial-condition draws (include/kernels.cuh:147-150).
constexpr uint32_t kDomainJitter = 1u;
constexpr uint32_t kDomainPolarity = 2u;
constexpr uint32_t kDomainVA = 3u;

pfh::Philox4 draw(uint32_t gid, uint32_t domain, uint64_t key) {
    return pfh::philox4x32_10(gid, domain, 0u, 0u, (uint32_t)(key & 0xFFFFFFFFull),
                              (uint32_t)(key >> 32));
}
}  // namespace

extern "C" {

int pfh_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    if (!ctr || !key || !out) return invalid("philox4x32_10");
    const pfh::Philox4 r = pfh::philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    std::memcpy(out, r.v, sizeof(r.v));
    return PFH_OK;
}

double pfh_uniform53(uint32_t a, uint32_t b) { return pfh::uniform53(a, b); }

int pfh2d_domain_side(int num_cells, double radius, double rho) {
    if (num_cells <= 0 || !(radius > 0.0) || !(rho > 0.0)) return invalid("domain_side");
    const double a = (double)num_cells * (pfh::kPi * radius * radius) / rho;
    const double side = std::ceil(std::sqrt(a));
    if (!std::isfinite(side) || side < 1.0 || side > 2147483647.0)
        return invalid("domain_side: outside the int range");
    return (int)side;
}

int pfh2d_s_pitch(int side) {
    if (side <= 0) return invalid("s_pitch: side must be positive");
    const long long p = 32LL * (((long long)side + 31LL) / 32LL);
    return p <= 2147483647LL ? (int)p : invalid("s_pitch: pitch overflows int");
}

float pfh2d_ic_theta(int32_t gid, uint64_t polarity_stream) {
    const pfh::Philox4 r = draw((uint32_t)gid, kDomainPolarity, polarity_stream);
    return (float)(2.0 * pfh::kPi * pfh::uniform53(r.v[0], r.v[1]));
}

double pfh2d_ic_v_A(int32_t gid, uint64_t seed, double v_A, double sigma) {
    if (!(sigma > 0.0)) return v_A;
    const pfh::Philox4 r = draw((uint32_t)gid, kDomainVA, seed);
    const double u1 = std::fmax(1e-300, pfh::uniform53(r.v[0], r.v[1]));
    const double u2 = pfh::uniform53(r.v[2], r.v[3]);
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * pfh::kPi * u2);
    return v_A * std::exp(sigma * z);
}

int pfh2d_grid_jitter_centres(int N, int L, uint64_t seed, float* cx, float* cy) {
    if (N <= 0 || L <= 0 || !cx || !cy) return invalid("grid_jitter_centres");
    // nx columns on rows of alternating half-spacing offset, each centre
    // jittered by +-5% of the spacing.
    const int nx = (int)std::ceil(std::sqrt((double)N));
    const double sp = (double)L / (double)nx;
    const double Ld = (double)L;
    for (int i = 0; i < N; ++i) {
        const int col = i % nx, row = i / nx;
        const pfh::Philox4 r = draw((uint32_t)i, kDomainJitter, seed);
        const double jx = (pfh::uniform53(r.v[0], r.v[1]) - 0.5) * 0.10 * sp;
        const double jy = (pfh::uniform53(r.v[2], r.v[3]) - 0.5) * 0.10 * sp;
        double x = ((double)col + 0.5 + ((row & 1) ? 0.5 : 0.0)) * sp + jx;
        double y = ((double)row + 0.5) * sp + jy;
        x -= std::floor(x / Ld) * Ld;
        y -= std::floor(y / Ld) * Ld;
        cx[i] = (float)x;
        cy[i] = (float)y;
    }
    return PFH_OK;
}

uint64_t pfh_centre_table_fnv1a64(const float* x, const float* y, int n) {
    uint64_t h = 0xcbf29ce484222325ull;
    if (!x || !y || n < 0) return 0;
    for (int i = 0; i < n; ++i) {
        const float xy[2] = {x[i], y[i]};
        for (float f : xy) {
            uint32_t bits;
            std::memcpy(&bits, &f, 4);
            for (int b = 0; b < 4; ++b) {
                h ^= (bits >> (8 * b)) & 0xFFu;
                h *= 0x100000001b3ull;
            }
        }
    }
    return h;
}

}  // extern "C"
