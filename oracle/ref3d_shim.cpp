// TEST INFRASTRUCTURE ONLY: exposes the reference's own host-compilable 3D
// pieces (compiled from the reference tree by oracle/ref.mk) through a C ABI
// so the oracle's restatement and the product's host set-up can be checked
// against them: the dense FP64 one-step reference, the seed-radius
// bisection, the Philox polarity draws, the brick sizing and the sequential
// centre placement.
#include <cstdint>
#include <cstring>
#include <vector>

#include "palmieri_initializer.hpp"
#include "params.cuh"
#include "reference.hpp"
#include "rng.cuh"

extern "C" {

int ref3d_calibrate(double radius, double lambda, int edge, double tol,
                    int max_iter, double out[3]) {
    const auto r = pf3d::reference::calibrate_seed_radius(radius, lambda, edge,
                                                          tol, max_iter);
    out[0] = r.seed_radius;
    out[1] = r.measured_volume;
    out[2] = r.relative_error;
    return r.converged ? r.iterations : -r.iterations;
}

void ref3d_initial_polarity(int64_t gid, uint64_t stream, double out[3]) {
    const pf3d::Vec3 v = pf3d::initial_polarity(gid, stream);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

int ref3d_tumble(uint64_t step, int64_t gid, uint64_t stream, double p,
                 const double cur[3], double out[3]) {
    const pf3d::TumbleSample s = pf3d::tumble_sample(
        step, gid, stream, p, pf3d::Vec3{cur[0], cur[1], cur[2]});
    out[0] = s.polarity.x; out[1] = s.polarity.y; out[2] = s.polarity.z;
    return s.tumbled ? 1 : 0;
}

int ref3d_brick_edge(double radius, double lambda) {
    return pf3d::automatic_brick_edge(radius, lambda);
}

int ref3d_domain_side(int n, double radius, double rho) {
    return pf3d::domain_side_for(n, radius, rho);
}

uint32_t ref3d_q_of(float phi) { return pf3d::q_of(phi); }

int ref3d_sequential_centres(int n, double side, double radius, uint64_t seed,
                             float* x, float* y, float* z) {
    std::vector<float> vx, vy, vz;
    pf3d::PalmieriInit3DDiagnostics d{};
    try {
        pf3d::palmieri3d_sequential_centres(n, side, radius, seed, &vx, &vy, &vz, &d);
    } catch (...) {
        return -1;
    }
    std::memcpy(x, vx.data(), sizeof(float) * vx.size());
    std::memcpy(y, vy.data(), sizeof(float) * vy.size());
    std::memcpy(z, vz.data(), sizeof(float) * vz.size());
    return static_cast<int>(d.candidates_drawn);
}

// Dense periodic FP64 one Euler step (include/pf3d/reference.hpp:256-331).
// phi: cells * side^3 doubles; per-cell gamma, speed and polarity (3 each).
int ref3d_one_step(int side, int cells, const double* phi, const double* gamma,
                   const double* speed, const double* polarity, double dt,
                   double lambda, double radius, double kappa, double mu,
                   double xi, double* phi_next, double* volume,
                   double* integral, double* velocity, uint32_t* aggregate) {
    pf3d::reference::DenseState st;
    st.side = side;
    st.cells.resize(static_cast<std::size_t>(cells));
    const std::size_t words = static_cast<std::size_t>(side) * side * side;
    st.phi.assign(phi, phi + words * static_cast<std::size_t>(cells));
    for (int c = 0; c < cells; ++c) {
        st.cells[c].gamma = gamma[c];
        st.cells[c].active_speed = speed[c];
        st.cells[c].polarity = {polarity[3 * c], polarity[3 * c + 1], polarity[3 * c + 2]};
    }
    pf3d::reference::ModelParameters m;
    m.dt = dt; m.lambda = lambda; m.target_radius = radius; m.kappa = kappa;
    m.mu = mu; m.xi = xi;
    try {
        const auto r = pf3d::reference::one_euler_step(st, m);
        std::memcpy(phi_next, r.phi_next.data(), sizeof(double) * r.phi_next.size());
        std::memcpy(aggregate, r.aggregate_q.data(), sizeof(uint32_t) * r.aggregate_q.size());
        for (int c = 0; c < cells; ++c) {
            volume[c] = r.volume[c];
            integral[3 * c] = r.interaction_integral[c].x;
            integral[3 * c + 1] = r.interaction_integral[c].y;
            integral[3 * c + 2] = r.interaction_integral[c].z;
            velocity[3 * c] = r.velocity[c].x;
            velocity[3 * c + 1] = r.velocity[c].y;
            velocity[3 * c + 2] = r.velocity[c].z;
        }
    } catch (...) {
        return -1;
    }
    return 0;
}

}  // extern "C"
