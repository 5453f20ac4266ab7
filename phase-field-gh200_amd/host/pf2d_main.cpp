// pf2d: command-line interface of the 2D simulator (src/main.cu). The same
// flags, defaults, resume rules and exit codes as the reference simulator;
// the boundary, velocity-observer and mobility-map outputs are not part of
// this build and are rejected with a message.
#include <cerrno>
#include <cmath>
#include <csignal>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <limits>

#include "pf2d_driver.hpp"

using namespace pf2d;

static void usage(const char* argv0) {
    std::printf(
        "Active phase-field cell simulator, 2D, one gfx950 GPU\n\n"
        "usage: %s [options]\n\n"
        "physics\n"
        "  --N <int>              number of cells                    (288)\n"
        "  --radius <f>           target radius R                    (49)\n"
        "  --rho <f>              target packing fraction; sets domain (0.90)\n"
        "  --lambda <f>           interface width parameter          (7)\n"
        "  --kappa <f>            repulsion strength                 (10)\n"
        "  --mu <f>               volume-constraint strength         (1)\n"
        "  --xi <f>               friction                           (1500)\n"
        "  --tau <f>              run-and-tumble persistence time    (1e4)\n"
        "  --v-A <f>              active self-propulsion speed       (1e-2)\n"
        "  --v-A-sigma <f>        lognormal spread on per-cell v_A   (0)\n"
        "  --gamma <f>            gamma for normal cells             (1.0)\n"
        "  --gamma-cancer <f>     gamma for cancer cells             (0.35)\n"
        "  --cancer-fraction <f>  lowest-ID fraction given the soft gamma (0)\n"
        "  --phase-field-mobility <f>  uniform passive Allen-Cahn mobility (0.5)\n\n"
        "numerics / run control\n"
        "  --dt <f>               time step                          (0.01)\n"
        "  --t-end <f>            end time                           (100)\n"
        "  --seed <u64>           grid-placement/per-cell RNG seed   (1234)\n"
        "  --polarity-seed <u64>  initial-polarity and tumble stream; 0 follows --seed\n"
        "  --initial-centres <csv>  accepted centres with header global_id,x,y\n"
        "  --print-interval <int> steps between status lines         (100)\n"
        "  --full-moment <int>    steps between perimeter updates    (100)\n"
        "  --verify-every <int>   steps between strict-mode checks    (4096)\n"
        "  --device <int>         HIP device                         (0)\n\n"
        "performance options\n"
        "  --bench <int>          run <int> timed steps and exit\n"
        "  --no-graph             per-step launches instead of a 6-step HIP graph\n"
        "  --morton               Morton-order cell traversal\n"
        "  --no-l2                accepted; gfx950 has no L2 persistence window\n"
        "  --strict               run the invariant checks on the verify cadence\n\n"
        "output\n"
        "  --out <path>           plain-text trajectory output\n"
        "  --trajectory-samples <int>  evenly spaced trajectory frames (100)\n"
        "  --trajectory-interval <int> steps between frames (overrides samples)\n"
        "  --velocity-moments <path>  new PFVMOM4 velocity-component output\n"
        "  --velocity-spatial-stride <int>  spatial measurement stride, 1..1000 (100)\n"
        "  --velocity-dense-start <int>  initial densely measured steps (10000)\n"
        "  --velocity-moments-reference  independent dense observer for validation\n"
        "  --self-test            run coefficient, RNG, and force-sign checks\n"
        "  -h, --help             this message\n\n"
        "checkpointing (CELL schema 1)\n"
        "  -c, --checkpoint <path>     resume from a checkpoint; explicit options override\n"
        "                              stored parameters; --N and --rho cannot change\n"
        "  --checkpoint-interval <int> steps between rolling <dir>/checkpoint.bin\n"
        "  --save-interval <int>       steps between tagged <dir>/checkpoint_%%08d.bin\n"
        "  --checkpoint-dir <path>     where they go (default: directory of --out, else of -c)\n"
        "  --no-final-checkpoint       do not write <dir>/checkpoint.bin at exit\n",
        argv0);
}

static bool parse_d(const char* flag, const char* s, double* out, double lo, double hi) {
    char* end = nullptr;
    const double v = std::strtod(s, &end);
    if (end == s || *end != '\0' || !(v >= lo && v <= hi)) {
        std::fprintf(stderr, "[fatal] %s: expected a number in [%g, %g], got '%s'\n", flag, lo,
                     hi, s);
        return false;
    }
    *out = v;
    return true;
}

static bool parse_i(const char* flag, const char* s, long long* out, long long lo, long long hi) {
    char* end = nullptr;
    const long long v = std::strtoll(s, &end, 10);
    if (end == s || *end != '\0' || v < lo || v > hi) {
        std::fprintf(stderr, "[fatal] %s: expected an integer in [%lld, %lld], got '%s'\n", flag,
                     lo, hi, s);
        return false;
    }
    *out = v;
    return true;
}

static bool parse_u64(const char* flag, const char* s, unsigned long long* out) {
    errno = 0;
    char* end = nullptr;
    const unsigned long long v = s[0] == '-' ? 0 : std::strtoull(s, &end, 10);
    if (s[0] == '-' || errno == ERANGE || end == s || *end != '\0') {
        std::fprintf(stderr, "[fatal] %s: expected an unsigned integer, got '%s'\n", flag, s);
        return false;
    }
    *out = v;
    return true;
}

// ---- self-test gates (src/main.cu:144-330) ----
static int gate_laplacian() {
    // lap(x^2) = 2 at h = 1 and the weights sum to zero.
    const double f[3][3] = {{1, 0, 1}, {1, 0, 1}, {1, 0, 1}};
    const double lap = (4.0 * (f[2][1] + f[0][1] + f[1][2] + f[1][0]) +
                        (f[2][2] + f[2][0] + f[0][2] + f[0][0]) - 20.0 * f[1][1]) / 6.0;
    const bool ok = lap == 2.0 && (4 * 4 + 4 * 1 - 20) == 0;
    std::printf("  [%s] Laplacian stencil: lap(x^2) = %.17g (expected 2), weight sum = 0\n",
                ok ? "PASS" : "FAIL", lap);
    return ok ? 0 : 1;
}

static int gate_coefficients(const SimParams& p) {
    const double ratio = p.interaction() / p.motility();
    const bool ok = std::fabs(ratio - p.xi) <= 1e-12 * p.xi;
    std::printf("  [%s] coefficients: 30/l^2 = %.9g   60k/l^2 = %.9g   60k/(xi l^2) = %.9g   "
                "ratio = %.12g (want xi = %.12g)\n", ok ? "PASS" : "FAIL", p.bulk(),
                p.interaction(), p.motility(), ratio, p.xi);
    return ok ? 0 : 1;
}

static int gate_velocity_sign(const SimParams& p) {
    // Cell n left of cell m: grad_x(phi_n) < 0 in their overlap, so the
    // interaction integral of n must point away from m.
    const int L = 192;
    const double R = 20.0, k = p.interface_k();
    int bad = 0;
    for (int sep = 26; sep <= 40; sep += 4) {
        const double nx = 0.5 * L - 0.5 * sep, mx = 0.5 * L + 0.5 * sep;
        std::vector<double> pn((size_t)L * L), pm((size_t)L * L);
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const double dy = y - 0.5 * L;
                pn[(size_t)y * L + x] = 0.5 * (1.0 - std::tanh(k * (std::hypot(x - nx, dy) - R)));
                pm[(size_t)y * L + x] = 0.5 * (1.0 - std::tanh(k * (std::hypot(x - mx, dy) - R)));
            }
        double Ix = 0.0;
        for (int y = 0; y < L; ++y)
            for (int x = 1; x < L - 1; ++x) {
                const size_t i = (size_t)y * L + x;
                Ix += pn[i] * 0.5 * (pn[i + 1] - pn[i - 1]) * pm[i] * pm[i];
            }
        const bool ok = Ix < 0.0;
        bad += !ok;
        std::printf("  [%s] repulsive sign at separation %2d: Ix = %+.6e  vx = %+.6e (must be "
                    "negative: n moves away from m)\n", ok ? "PASS" : "FAIL", sep, Ix,
                    p.motility() * Ix);
    }
    return bad ? 1 : 0;
}

static int gate_tumble(const SimParams& p) {
    const double good = -std::expm1(-p.dt / p.tau);
    const double tau_good = -p.dt / std::log1p(-good);
    const bool ok = std::fabs(tau_good - p.tau) <= 1e-9 * p.tau;
    std::printf("  [%s] tumble probability: -expm1(-dt/tau) = %.9e -> tau_eff = %.4f\n",
                ok ? "PASS" : "FAIL", good, tau_good);
    // Uniformity of the tumble stream at a larger event probability.
    const double tau_s = 10.0, ps = -std::expm1(-p.dt / tau_s);
    const long long trials = 5000000;
    long long hits = 0;
    uint32_t out[4];
    const uint32_t key[2] = {(uint32_t)p.seed, 0u};
    for (long long i = 0; i < trials; ++i) {
        const uint32_t ctr[4] = {(uint32_t)(i & 0xFFFFFFFF), (uint32_t)(i >> 32), 7u, 0xA5A5A5A5u};
        pfh_philox4x32_10(ctr, key, out);
        hits += pfh_uniform53(out[0], out[1]) < ps;
    }
    const double phat = (double)hits / (double)trials;
    const double tau_meas = -p.dt / std::log1p(-phat);
    const double sigma = std::sqrt(ps * (1.0 - ps) / (double)trials) / ps;
    const bool ok2 = std::fabs(tau_meas - tau_s) <= 5.0 * sigma * tau_s;
    std::printf("  [%s] Philox sampling: %lld draws at p = %.6e, tau_eff = %.3f (want %.1f +- "
                "%.3f at 5 sigma)\n", ok2 ? "PASS" : "FAIL", trials, ps, tau_meas, tau_s,
                5.0 * sigma * tau_s);
    return ok && ok2 ? 0 : 1;
}

static int gate_geometry(const SimParams& p) {
    // Every class window must hold a relaxed cell's phi^2 tail down to half
    // a Q5.27 quantum on both axes.
    const double k = p.interface_k();
    const double circum = std::sqrt(p.area0() / (1.5 * std::sqrt(3.0)));
    const double d_q = std::log(1.0 / (0.5 * kQInv)) / (4.0 * k);
    const double need = circum + d_q + 1.0;
    bool ok = true;
    for (const Shape& s : kShapes) ok = ok && (double)(std::min(s.wx, s.wy) / 2) >= need;
    std::printf("  [%s] window geometry: k = %.6f  hex circumradius %.2f  phi^2 < half a quantum "
                "at %.2f; required half-width %.2f px\n", ok ? "PASS" : "FAIL", k, circum, d_q,
                need);
    return ok ? 0 : 1;
}

static int run_self_test(const SimParams& p) {
    std::printf("--- self-test checks ---\n");
    const int bad = gate_laplacian() + gate_coefficients(p) + gate_velocity_sign(p) +
                    gate_tumble(p) + gate_geometry(p);
    std::printf("--- %s ---\n", bad ? "SELF-TEST FAILED" : "self-test passed");
    return bad ? 1 : 0;
}

int main(int argc, char** argv) {
    SimParams p;
    RunOptions opt;
    SimOverrides ov;
    long long device = 0, iv = 0;
    bool self_test = false, no_final = false, samples_given = false, interval_given = false;
    bool velocity_given = false;
    std::string ckpt_in, ckpt_dir_flag;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        auto need = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "[fatal] %s requires a value\n", a);
                std::exit(2);
            }
            return argv[++i];
        };
        auto is = [&](const char* f) { return std::strcmp(a, f) == 0; };
        if (is("-h") || is("--help")) {
            usage(argv[0]);
            return 0;
        } else if (is("--N")) {
            if (!parse_i(a, need(), &iv, 1, 4000000)) return 2;
            p.num_cells = (int)iv;
            ov.num_cells = true;
        } else if (is("--radius")) {
            if (!parse_d(a, need(), &p.target_radius, 1e-6, 1e6)) return 2;
            ov.target_radius = true;
        } else if (is("--rho")) {
            if (!parse_d(a, need(), &p.rho, 1e-6, 0.999999)) return 2;
            ov.rho = true;
        } else if (is("--lambda")) {
            if (!parse_d(a, need(), &p.lambda, 1e-6, 1e6)) return 2;
            ov.lambda = true;
        } else if (is("--kappa")) {
            if (!parse_d(a, need(), &p.kappa, 0.0, 1e9)) return 2;
            ov.kappa = true;
        } else if (is("--mu")) {
            if (!parse_d(a, need(), &p.mu, 0.0, 1e9)) return 2;
            ov.mu = true;
        } else if (is("--xi")) {
            if (!parse_d(a, need(), &p.xi, 1e-9, 1e12)) return 2;
            ov.xi = true;
        } else if (is("--tau")) {
            if (!parse_d(a, need(), &p.tau, 1e-9, 1e12)) return 2;
            ov.tau = true;
        } else if (is("--v-A")) {
            if (!parse_d(a, need(), &p.v_A, 0.0, 1e6)) return 2;
            ov.v_A = true;
        } else if (is("--v-A-sigma")) {
            if (!parse_d(a, need(), &p.v_A_sigma, 0.0, 10.0)) return 2;
            ov.v_A_sigma = true;
        } else if (is("--gamma")) {
            if (!parse_d(a, need(), &p.gamma_normal, 1e-9, 1e6)) return 2;
            ov.gamma = true;
        } else if (is("--gamma-cancer")) {
            if (!parse_d(a, need(), &p.gamma_cancer, 1e-9, 1e6)) return 2;
            ov.gamma_cancer = true;
        } else if (is("--cancer-fraction")) {
            if (!parse_d(a, need(), &p.cancer_fraction, 0.0, 1.0)) return 2;
            ov.cancer_fraction = true;
        } else if (is("--phase-field-mobility")) {
            if (!parse_d(a, need(), &p.phase_field_mobility, 1e-9, 1e6)) return 2;
            ov.phase_field_mobility = true;
        } else if (is("--dt")) {
            if (!parse_d(a, need(), &p.dt, 1e-12, 1e6)) return 2;
            ov.dt = true;
        } else if (is("--t-end")) {
            if (!parse_d(a, need(), &p.t_end, 0.0, 1e12)) return 2;
            ov.t_end = true;
        } else if (is("--seed")) {
            if (!parse_u64(a, need(), &p.seed)) return 2;
            ov.seed = true;
        } else if (is("--polarity-seed")) {
            if (!parse_u64(a, need(), &p.polarity_seed)) return 2;
            ov.polarity_seed = true;
        } else if (is("--initial-centres")) {
            opt.initial_centres_path = need();
            if (opt.initial_centres_path.empty()) {
                std::fprintf(stderr, "[fatal] --initial-centres requires a non-empty path\n");
                return 2;
            }
        } else if (is("--print-interval")) {
            if (!parse_i(a, need(), &iv, 0, 1000000000)) return 2;
            p.print_interval = (int)iv;
            ov.print_interval = true;
        } else if (is("-c") || is("--checkpoint")) {
            ckpt_in = need();
        } else if (is("--checkpoint-interval")) {
            if (!parse_i(a, need(), &iv, 1, 1000000000000LL)) return 2;
            opt.ckpt_interval = iv;
        } else if (is("--save-interval")) {
            if (!parse_i(a, need(), &iv, 1, std::numeric_limits<int32_t>::max())) return 2;
            opt.save_interval = iv;
        } else if (is("--checkpoint-dir")) {
            ckpt_dir_flag = need();
        } else if (is("--no-final-checkpoint")) {
            no_final = true;
        } else if (is("--trajectory-samples")) {
            if (!parse_i(a, need(), &iv, 1, 1000000000)) return 2;
            opt.traj_samples = (int)iv;
            samples_given = true;
        } else if (is("--trajectory-interval")) {
            if (!parse_i(a, need(), &iv, 1, 1000000000)) return 2;
            opt.traj_interval = iv;
            interval_given = true;
        } else if (is("--full-moment")) {
            if (!parse_i(a, need(), &iv, 0, 1000000000)) return 2;
            p.full_moment_every = (int)iv;
            ov.full_moment = true;
        } else if (is("--verify-every")) {
            if (!parse_i(a, need(), &iv, 1, 1000000000)) return 2;
            p.verify_every = (int)iv;
            ov.verify_every = true;
        } else if (is("--device")) {
            if (!parse_i(a, need(), &device, 0, 63)) return 2;
        } else if (is("--bench")) {
            if (!parse_i(a, need(), &iv, 1, 100000000)) return 2;
            opt.bench_steps = (int)iv;
        } else if (is("--no-graph")) {
            opt.use_graph = false;
        } else if (is("--morton")) {
            opt.morton = true;
        } else if (is("--no-l2")) {
            // gfx950 has no L2 access-policy window: nothing to turn off.
        } else if (is("--strict")) {
            opt.strict = true;
        } else if (is("--out")) {
            opt.out_path = need();
        } else if (is("--self-test")) {
            self_test = true;
        } else if (is("--velocity-moments")) {
            opt.velocity.path = need();
            velocity_given = true;
        } else if (is("--velocity-spatial-stride")) {
            if (!parse_i(a, need(), &iv, 1, 1000)) return 2;
            opt.velocity.spatial_stride = (int)iv;
            velocity_given = true;
        } else if (is("--velocity-dense-start")) {
            if (!parse_i(a, need(), &iv, 0, 1000000000)) return 2;
            opt.velocity.dense_start = iv;
            velocity_given = true;
        } else if (is("--velocity-moments-reference")) {
            opt.velocity.reference = true;
            velocity_given = true;
        } else if (is("--phase-field-mobility-map") || is("--boundary-out") ||
                   is("--boundary-interval") || is("--boundary-compression")) {
            std::fprintf(stderr, "[fatal] %s is not supported by this build (boundary and "
                                 "mobility-map outputs are outside it)\n", a);
            return 2;
        } else {
            std::fprintf(stderr, "[fatal] unknown option '%s' (try --help)\n", a);
            return 2;
        }
    }
    if (velocity_given &&
        (opt.velocity.path.empty() || (opt.out_path.empty() && opt.bench_steps == 0))) {
        std::fprintf(stderr, "[fatal] velocity recording requires --velocity-moments and --out "
                             "(--bench measures the observer without writing files)\n");
        return 2;
    }
    if (!distinct_output_paths({opt.out_path, opt.velocity.path})) return 2;
    CheckpointData ckpt;
    if (!ckpt_in.empty()) {
        if (!opt.initial_centres_path.empty()) {
            std::fprintf(stderr, "[fatal] --initial-centres is valid only for a fresh start. A "
                                 "checkpoint already contains the complete initialised "
                                 "microstate.\n");
            return 2;
        }
        if (ov.num_cells || ov.rho) {
            std::fprintf(stderr, "[fatal] --N and --rho cannot be overridden on resume: they set "
                                 "the domain side, and every origin and phi tile in the "
                                 "checkpoint\n        is expressed in it. Drop the flag, or "
                                 "start a fresh run.\n");
            return 2;
        }
        if (!checkpoint_read(ckpt_in, &ckpt)) return 8;
        if (!samples_given && !interval_given) {
            opt.traj_samples = ckpt.trajectory_samples;
            opt.traj_interval = ckpt.trajectory_interval;
        }
        const SimParams cli = p;
        p = ckpt.params;
        ov.apply(p, cli);
        resolve_per_cell_scalars(p, ov, &ckpt);
    } else {
        p.Nx = p.Ny = pfh2d_domain_side(p.num_cells, p.target_radius, p.rho);
    }
    if (!validate(p, !ckpt_in.empty())) return 3;
    print_params(p, p.Nx, pfh2d_s_pitch(p.Nx));
    if (self_test && run_self_test(p)) return 4;
    if (!ckpt_dir_flag.empty() || opt.ckpt_interval > 0 || opt.save_interval > 0 ||
        !ckpt_in.empty()) {
        if (!ckpt_dir_flag.empty()) {
            opt.ckpt_dir = ckpt_dir_flag;
        } else {
            const std::filesystem::path src = !opt.out_path.empty()
                ? std::filesystem::path(opt.out_path) : std::filesystem::path(ckpt_in);
            opt.ckpt_dir = src.parent_path().empty() ? std::string(".")
                                                     : src.parent_path().string();
        }
        if (!distinct_output_paths({opt.out_path, opt.velocity.path,
                                    opt.ckpt_dir + "/checkpoint.bin",
                                    opt.ckpt_dir + "/checkpoint_failed.bin"}))
            return 2;
        std::error_code ec;
        std::filesystem::create_directories(opt.ckpt_dir, ec);
        if (ec) {
            std::fprintf(stderr, "[fatal] cannot create checkpoint directory %s: %s\n",
                         opt.ckpt_dir.c_str(), ec.message().c_str());
            return 2;
        }
        opt.final_checkpoint = !no_final;
        std::printf("  checkpoints      %s  (rolling every %lld steps, tagged every %lld, final "
                    "%s)\n", opt.ckpt_dir.c_str(), opt.ckpt_interval, opt.save_interval,
                    opt.final_checkpoint ? "yes" : "no");
    }
    Sim sim;
    const bool ok = ckpt_in.empty() ? sim.init(p, opt, (int)device)
                                    : sim.init_from_checkpoint(p, ckpt, opt, (int)device);
    if (!ok) {
        std::fprintf(stderr, "[fatal] initialisation failed\n");
        return 5;
    }
    ckpt = CheckpointData{};
    if (opt.bench_steps > 0) {
        double ms = 0.0;
        return sim.bench(opt.bench_steps, &ms) ? 0 : 6;
    }
    std::signal(SIGTERM, Sim::request_termination);
    std::signal(SIGINT, Sim::request_termination);
    return sim.run() ? 0 : 9;
}
