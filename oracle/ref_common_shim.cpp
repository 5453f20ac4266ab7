// TEST INFRASTRUCTURE ONLY: exposes the reference's own common/ primitives
// (compiled from the reference tree by oracle/ref.mk) through a C ABI so the
// oracle's restatement can be checked against them.
#include <cstdint>
#include <cstring>

#include "model_coefficients.cuh"
#include "philox.cuh"

extern "C" {

void ref_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const pf_common::Philox4 r = pf_common::philox4x32_10(
        ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    std::memcpy(out, r.v, sizeof(r.v));
}

double ref_u53(uint32_t a, uint32_t b) { return pf_common::philox_uniform53(a, b); }

// coefficients in double: bulk, interaction, motility, volume, k, R_eff
void ref_coeffs(double lambda, double kappa, double xi, double mu,
                double target, double radius, double out[6]) {
    out[0] = pf_common::bulk_coeff(lambda);
    out[1] = pf_common::interaction_coeff(kappa, lambda);
    out[2] = pf_common::motility_coeff(kappa, xi, lambda);
    out[3] = pf_common::volume_coeff(mu, target);
    out[4] = pf_common::interface_k(lambda);
    out[5] = pf_common::init_radius(radius, lambda);
}

}  // extern "C"
