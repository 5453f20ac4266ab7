// pf3d host driver: the command-line simulator of the 3D periodic per-cell
// phase-field path on one gfx950 GPU, calling the kernels only through the C
// ABI of include/pfh.h. Keeps the reference simulator's flags for the
// periodic cube, its parameter convention, the PF3D checkpoint format and
// the periodic trajectory text contract (header line, column line, rows).
//
// Reference counterparts:
//   command line                 src/pf3d/main.cu
//   SimParams3D, validate        include/pf3d/params.cuh:236-441
//   Sim3D                        src/pf3d/sim.cu (allocate :202-762,
//                                initialize_fields :898-1177, begin_step /
//                                finish_step :1460-1651, trajectory
//                                :1782-2034, run and bench :2180-2310)
//   trajectory header            include/pf3d/trajectory_header.hpp
// Outside this build (rejected with a message): slab and channel geometries,
// adaptive (promoted) bricks and memory modes, heterogeneous populations,
// boundary output, the peer device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <csignal>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "pfh.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kLap27RowSumBound = 256.0 / 30.0;   // include/pf3d/params.cuh:77
constexpr int kSupportMargin = 8;                    // kBrickSafetyMargin
constexpr int kGraphBody = 2;                        // fields alternate every step
constexpr long long kFatalPollEvery = 1000;
constexpr const char* kColumns =
    "# columns: time cell_id x y z vx vy vz px py pz gamma v_A "
    "surface sphericity volume phi_max";

volatile std::sig_atomic_t g_terminate = 0;
void on_signal(int) { g_terminate = 1; }

#define HIP_OK(call)                                                                    \
    do {                                                                                \
        const hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "[hip] %s: %s\n", #call, hipGetErrorString(e_));       \
            return false;                                                               \
        }                                                                               \
    } while (0)
#define PFH_CALL(call)                                                                  \
    do {                                                                                \
        if ((call) != PFH_OK) {                                                         \
            std::fprintf(stderr, "[pfh] %s: %s\n", #call, pfh_last_error());            \
            return false;                                                               \
        }                                                                               \
    } while (0)

// The periodic subset of SimParams3D with the reference defaults.
struct Params3D {
    int num_cells = 288;
    int L = 0;                      // periodic cube side
    double rho = 0.90;
    double dt = 0.01, t_end = 100.0, aging_time = 0.0;
    double lambda = 7.0, target_radius = 49.0, kappa = 10.0, mu = 1.0, xi = 1500.0;
    double tau = 1.0e4, v_A = 1.0e-2, gamma_normal = 1.0, gamma_cancer = 0.35;
    double cancer_fraction = 0.0, v_A_sigma = 0.0;
    unsigned long long seed = 1234, polarity_seed = 0;
    int full_moment_every = 100, verify_every = 4096, print_interval = 100;
    int brick_edge = 0;             // 0: the automatic edge

    unsigned long long polarity_stream() const { return polarity_seed ? polarity_seed : seed; }
    double volume0() const { return 4.0 * kPi * target_radius * target_radius * target_radius / 3.0; }
    double bulk() const { return 30.0 / (lambda * lambda); }
    double interaction() const { return 60.0 * kappa / (lambda * lambda); }
    double motility() const { return interaction() / xi; }
    double volume() const { return 2.0 * mu / volume0(); }
    double p_tumble() const { return -std::expm1(-dt / tau); }
    long long total_steps() const {
        const double s = t_end / dt + 0.5;
        return std::isfinite(s) && s >= 0.0 && s < 9.0e18 ? (long long)s : -1;
    }
    long long aging_steps() const {
        const double s = aging_time / dt + 0.5;
        return std::isfinite(s) && s >= 0.0 && s <= 9.0e18 ? (long long)s : -1;
    }
    double realized_rho() const {
        return L > 0 ? (double)num_cells * volume0() / ((double)L * L * L) : 0.0;
    }
};

struct Options3D {
    std::string out_path, centres_path, ckpt_in, ckpt_dir;
    long long ckpt_interval = 0, save_interval = 0;
    bool final_checkpoint = true;
    int traj_samples = 100;
    long long traj_interval = 0;
    int bench_steps = 0;
    int shards = 0;
    int device = 0;
    bool use_graph = true, strict = false;
};

// Startup validation: the periodic subset of include/pf3d/params.cuh:332-441.
bool validate(const Params3D& p, int B, int automatic_B) {
    bool ok = true;
    auto fail = [&](const char* m) {
        std::printf("[fatal] %s\n", m);
        ok = false;
    };
    if (p.num_cells <= 0) fail("num_cells must be positive");
    if (!(p.dt > 0.0) || p.total_steps() < 0) fail("dt and t_end must give a finite step count");
    if (p.aging_steps() < 0) fail("aging_time must be finite and non-negative");
    if (B <= 0 || p.L <= B) fail("the brick must fit strictly inside the domain");
    if (p.brick_edge && (p.brick_edge < automatic_B || p.brick_edge % 8))
        fail("an explicit brick edge must be an aligned edge >= the automatic one");
    if (B > PFH3D_MAX_EDGE)
        fail("bricks above 64^3 (adaptive support) are outside this build; the automatic "
             "edge grows with R and lambda");
    if (p.dt * p.gamma_normal * kLap27RowSumBound >= 1.0)
        fail("explicit-Euler Laplacian stability bound is violated");
    if (p.dt * p.gamma_normal * p.bulk() >= 1.0)
        fail("explicit-Euler bulk-reaction timestep bound is violated");
    if (p.dt * 2.0 * p.mu >= 1.0) fail("explicit-Euler volume-restoring timestep bound is violated");
    if (p.cancer_fraction != 0.0 || p.v_A_sigma != 0.0)
        fail("heterogeneous cell populations (--cancer-fraction, --v-A-sigma) are outside "
             "this build");
    if (p.full_moment_every < 0 || p.print_interval < 0 || p.verify_every <= 0)
        fail("cadences must be non-negative (verify-every positive)");
    return ok;
}

// Trajectory metadata line (include/pf3d/trajectory_header.hpp, periodic
// branch): every token is part of the append contract.
std::string trajectory_header(const Params3D& p, int B, int shards, unsigned long long every) {
    char text[1536];
    const int n = std::snprintf(
        text, sizeof(text),
        "# phase_field_gh200 trajectory schema=1 dim=3 geometry=periodic-xyz "
        "N=%d L=%d B=%d "
        "dx=%a dt=%a aging_time=%a rho_V_target=%a rho_V_realized=%a "
        "R=%a lambda=%a kappa=%a mu=%a xi=%a "
        "tau=%a v_A=%a v_A_sigma=%a gamma_normal=%a gamma_soft=%a "
        "soft_fraction=%a seed=%llu polarity_stream=%llu init_hash=%016llx "
        "promoted_measure_reduction=%s promoted_measure_policy=%d "
        "promoted_measure_auto_wave_ctas=%d base_measure_shards=%d "
        "trajectory_interval=%llu",
        p.num_cells, p.L, B, 1.0, p.dt, p.aging_time, p.rho, p.realized_rho(),
        p.target_radius, p.lambda, p.kappa, p.mu, p.xi, p.tau, p.v_A, p.v_A_sigma,
        p.gamma_normal, p.gamma_cancer, p.cancer_fraction, p.seed, p.polarity_stream(),
        0ull, "one-cta", 0, 0, shards, every);
    return n > 0 && (size_t)n < sizeof(text) ? std::string(text, (size_t)n) : std::string();
}

double sphericity(double volume, double surface) {
    if (!(volume > 0.0) || !(surface > 0.0)) return 0.0;
    return std::cbrt(kPi) * std::pow(6.0 * volume, 2.0 / 3.0) / surface;
}

long long next_boundary(long long current, long long interval) {
    if (interval <= 0) return LLONG_MAX;
    const long long r = current % interval;
    return current + (r == 0 ? interval : interval - r);
}

class Sim3D {
public:
    ~Sim3D();
    bool init(const Params3D& p, const Options3D& opt);
    bool run();
    bool bench(int steps);

private:
    bool allocate();
    bool fresh_state();
    bool resume_state();
    bool rebuild_S_and_measure(int slot);
    pfh3d_step_args args_for(int slot) const;
    bool capture_graph();
    bool advance(long long next_stop);
    bool fatal_flags(bool print);
    bool open_trajectory();
    bool append_trajectory();
    bool save_checkpoint(const std::string& path);
    void print_line();
    double time() const { return (double)steps_done_ * p_.dt; }

    Params3D p_;
    Options3D opt_;
    int B_ = 0, P_ = 0, shards_ = 1, slabs_ = 1;
    size_t voxels_ = 0, words_ = 0;
    long long steps_done_ = 0, traj_every_ = 1, traj_frames_ = 0;
    hipStream_t stream_ = nullptr;
    float* d_phi_[2] = {nullptr, nullptr};
    uint32_t* d_S_[2] = {nullptr, nullptr};
    pfh3d_cell_state *d_cells_ = nullptr, *d_accepted_ = nullptr;
    uint32_t *d_flags_ = nullptr, *d_arrivals_ = nullptr;
    unsigned long long* d_counters_ = nullptr;
    pfh3d_moment_partial* d_partials_ = nullptr;
    pfh3d_traj_cell* d_traj_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    std::FILE* traj_fp_ = nullptr;
    double seed_radius_ = 0.0;
};

Sim3D::~Sim3D() {
    if (traj_fp_) std::fclose(traj_fp_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (void* q : {(void*)d_phi_[0], (void*)d_phi_[1], (void*)d_S_[0], (void*)d_S_[1],
                    (void*)d_cells_, (void*)d_accepted_, (void*)d_flags_, (void*)d_arrivals_,
                    (void*)d_counters_, (void*)d_partials_, (void*)d_traj_})
        if (q) (void)hipFree(q);
    if (stream_) (void)hipStreamDestroy(stream_);
}

pfh3d_step_args Sim3D::args_for(int slot) const {
    pfh3d_step_args a{};
    a.phi_in = d_phi_[slot];
    a.phi_out = d_phi_[1 - slot];
    a.S_in = d_S_[slot];
    a.S_out = d_S_[1 - slot];
    a.cells = d_cells_;
    a.accepted = d_accepted_;
    a.flags = d_flags_;
    a.counters = reinterpret_cast<uint64_t*>(d_counters_);
    a.partials = d_partials_;
    a.arrivals = d_arrivals_;
    a.N = p_.num_cells;
    a.B = B_;
    a.L = p_.L;
    a.P = P_;
    a.shards = shards_;
    a.max_shift = PFH3D_MAX_SHIFT;
    a.support_margin = kSupportMargin;
    a.full_moment_every = p_.full_moment_every;
    a.dt = (float)p_.dt;
    a.bulk_scale = (float)p_.bulk();
    a.interaction_coeff = (float)p_.interaction();
    a.motility_coeff = (float)p_.motility();
    a.V0 = p_.volume0();
    a.volume_scale = p_.volume();
    a.p_tumble = p_.p_tumble();
    a.polarity_stream = p_.polarity_stream();
    a.aging_steps = (uint64_t)p_.aging_steps();
    return a;
}

bool Sim3D::allocate() {
    HIP_OK(hipSetDevice(opt_.device));
    HIP_OK(hipStreamCreate(&stream_));
    const int N = p_.num_cells;
    voxels_ = (size_t)B_ * B_ * B_;
    P_ = 32 * ((p_.L + 31) / 32);
    words_ = (size_t)p_.L * p_.L * P_;
    hipDeviceProp_t prop{};
    HIP_OK(hipGetDeviceProperties(&prop, opt_.device));
    // measurement workgroups per cell: one pass over the compute units, at
    // most 4 and at least four planes per shard (include/pf3d/measure_shards.hpp:36-51)
    shards_ = opt_.shards > 0 ? opt_.shards
                              : std::max(1, std::min({4, prop.multiProcessorCount /
                                                             std::max(1, N), B_ / 4}));
    slabs_ = pfh3d_update_slabs(N, B_, opt_.device);
    if (slabs_ <= 0 || shards_ > B_) {
        std::fprintf(stderr, "[fatal] launch shape: %s\n", pfh_last_error());
        return false;
    }
    for (int k = 0; k < 2; ++k) {
        HIP_OK(hipMalloc(&d_phi_[k], voxels_ * N * sizeof(float)));
        HIP_OK(hipMemset(d_phi_[k], 0, voxels_ * N * sizeof(float)));
        HIP_OK(hipMalloc(&d_S_[k], words_ * sizeof(uint32_t)));
        HIP_OK(hipMemset(d_S_[k], 0, words_ * sizeof(uint32_t)));
    }
    HIP_OK(hipMalloc(&d_cells_, N * sizeof(pfh3d_cell_state)));
    HIP_OK(hipMalloc(&d_accepted_, N * sizeof(pfh3d_cell_state)));
    HIP_OK(hipMalloc(&d_flags_, PFH3D_FLAG_COUNT * sizeof(uint32_t)));
    HIP_OK(hipMemset(d_flags_, 0, PFH3D_FLAG_COUNT * sizeof(uint32_t)));
    HIP_OK(hipMalloc(&d_arrivals_, N * sizeof(uint32_t)));
    HIP_OK(hipMemset(d_arrivals_, 0, N * sizeof(uint32_t)));
    HIP_OK(hipMalloc(&d_counters_, 2 * sizeof(unsigned long long)));
    HIP_OK(hipMemset(d_counters_, 0, 2 * sizeof(unsigned long long)));
    HIP_OK(hipMalloc(&d_partials_, (size_t)N * shards_ * sizeof(pfh3d_moment_partial)));
    HIP_OK(hipMalloc(&d_traj_, N * sizeof(pfh3d_traj_cell)));
    std::printf("  device %d: %s, %d CUs; brick B=%d, %zu voxels/cell, S %d x %d x %d (pitch "
                "%d), %d measurement shards, %d update slabs\n", opt_.device, prop.name,
                prop.multiProcessorCount, B_, voxels_, p_.L, p_.L, p_.L, P_, shards_, slabs_);
    return true;
}

// Aggregate of the current cubes and a fresh measurement (no tumble, surface
// measured, motion planned), as the reference rebuilds them at start and
// resume (src/pf3d/sim.cu:1280-1368).
bool Sim3D::rebuild_S_and_measure(int slot) {
    PFH_CALL(pfh3d_clear_S(d_S_[slot], p_.L, P_, stream_));
    PFH_CALL(pfh3d_scatter_current(d_phi_[slot], d_cells_, d_S_[slot], d_flags_, p_.num_cells,
                                 B_, p_.L, P_, stream_));
    const pfh3d_step_args a = args_for(slot);
    PFH_CALL(pfh3d_measure_cells(&a, 0, 1, 1, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    return true;
}

bool Sim3D::fresh_state() {
    const int N = p_.num_cells;
    std::vector<float> x(N), y(N), z(N);
    if (!opt_.centres_path.empty()) {
        std::FILE* f = std::fopen(opt_.centres_path.c_str(), "r");
        if (!f) {
            std::fprintf(stderr, "[fatal] cannot read %s\n", opt_.centres_path.c_str());
            return false;
        }
        char line[512];
        int got = 0;
        bool header = false;
        while (std::fgets(line, sizeof(line), f)) {
            long long id = 0;
            double cx = 0, cy = 0, cz = 0;
            if (!header) {
                header = true;
                if (std::strncmp(line, "global_id,x,y,z", 15) == 0) continue;
            }
            if (std::sscanf(line, "%lld,%lf,%lf,%lf", &id, &cx, &cy, &cz) != 4 || id < 0 ||
                id >= N || id != got || !(cx >= 0 && cx < p_.L && cy >= 0 && cy < p_.L &&
                                          cz >= 0 && cz < p_.L)) {
                std::fprintf(stderr, "[fatal] %s: row %d must be %d,x,y,z inside the domain\n",
                             opt_.centres_path.c_str(), got + 1, got);
                std::fclose(f);
                return false;
            }
            x[got] = (float)cx;
            y[got] = (float)cy;
            z[got] = (float)cz;
            ++got;
        }
        std::fclose(f);
        if (got != N) {
            std::fprintf(stderr, "[fatal] %s has %d centres for %d cells\n",
                         opt_.centres_path.c_str(), got, N);
            return false;
        }
    } else if (pfh3d_sequential_centres(N, (double)p_.L, p_.target_radius, p_.seed, x.data(),
                                        y.data(), z.data()) < 0) {
        std::fprintf(stderr, "[fatal] sequential placement failed: %s\n", pfh_last_error());
        return false;
    }
    std::vector<pfh3d_cell_state> cells(N);
    std::memset(cells.data(), 0, N * sizeof(pfh3d_cell_state));
    for (int i = 0; i < N; ++i) {
        pfh3d_cell_state& c = cells[i];
        c.global_id = i;
        c.gamma = (float)p_.gamma_normal;
        c.v_A = (float)p_.v_A;
        c.R_tgt = (float)p_.target_radius;
        double pol[3];
        PFH_CALL(pfh3d_initial_polarity(i, p_.polarity_stream(), pol));
        c.polarity_x = (float)pol[0];
        c.polarity_y = (float)pol[1];
        c.polarity_z = (float)pol[2];
    }
    HIP_OK(hipMemcpy(d_cells_, cells.data(), N * sizeof(pfh3d_cell_state), hipMemcpyHostToDevice));
    std::vector<double> centres(3 * (size_t)N);
    for (int i = 0; i < N; ++i) {
        centres[3 * i] = x[i];
        centres[3 * i + 1] = y[i];
        centres[3 * i + 2] = z[i];
    }
    double* d_centres = nullptr;
    HIP_OK(hipMalloc(&d_centres, centres.size() * sizeof(double)));
    const bool ok =
        hipMemcpy(d_centres, centres.data(), centres.size() * sizeof(double),
                  hipMemcpyHostToDevice) == hipSuccess &&
        pfh3d_init_spheres(d_phi_[0], d_phi_[1], d_cells_, d_centres, N, B_, p_.L,
                           (float)p_.lambda, (float)seed_radius_, stream_) == PFH_OK &&
        hipStreamSynchronize(stream_) == hipSuccess;
    (void)hipFree(d_centres);
    if (!ok) {
        std::fprintf(stderr, "[fatal] sphere initialisation failed: %s\n", pfh_last_error());
        return false;
    }
    steps_done_ = 0;
    return rebuild_S_and_measure(0);
}

// Resume (src/pf3d/sim.cu:795-897): cubes and records at the file's step, the
// aggregate rebuilt from the cubes, the measurement refreshed.
bool Sim3D::resume_state() {
    pfh3d_ckpt_params pr{};
    int64_t step = 0;
    double t = 0.0;
    uint64_t shards = 0;
    int N = 0, B = 0;
    PFH_CALL(pfh3d_checkpoint_read(opt_.ckpt_in.c_str(), &pr, &step, &t, &shards, &N, &B, nullptr,
                                 nullptr, stream_));
    if (N != p_.num_cells || B != B_ || pr.Nx != p_.L) {
        std::fprintf(stderr, "[fatal] checkpoint geometry N=%d B=%d L=%lld does not match "
                             "this run (N=%d B=%d L=%d)\n", N, B, (long long)pr.Nx,
                     p_.num_cells, B_, p_.L);
        return false;
    }
    const int slot = (int)(step % 2);
    std::vector<pfh3d_cell_state> cells(N);
    PFH_CALL(pfh3d_checkpoint_read(opt_.ckpt_in.c_str(), &pr, &step, &t, &shards, &N, &B,
                                 cells.data(), d_phi_[slot], stream_));
    HIP_OK(hipMemcpy(d_cells_, cells.data(), N * sizeof(pfh3d_cell_state), hipMemcpyHostToDevice));
    const unsigned long long counters[2] = {(unsigned long long)step, (unsigned long long)step};
    HIP_OK(hipMemcpy(d_counters_, counters, sizeof(counters), hipMemcpyHostToDevice));
    steps_done_ = step;
    std::printf("[ckpt] loaded %s: step %lld (t = %.6f), %d cells, B=%d\n",
                opt_.ckpt_in.c_str(), (long long)step, t, N, B);
    return rebuild_S_and_measure(slot);
}

bool Sim3D::capture_graph() {
    HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeGlobal));
    bool ok = true;
    for (int slot = 0; slot < kGraphBody && ok; ++slot) {
        const pfh3d_step_args a = args_for(slot);
        ok = pfh3d_step(&a, slabs_, stream_) == PFH_OK;
    }
    const hipError_t e = hipStreamEndCapture(stream_, &graph_);
    if (!ok || e != hipSuccess) return false;
    HIP_OK(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
    std::printf("  HIP graph: %d steps cover field and aggregate parity\n", kGraphBody);
    return true;
}

bool Sim3D::init(const Params3D& p, const Options3D& opt) {
    p_ = p;
    opt_ = opt;
    const int automatic = pfh3d_brick_edge(p_.target_radius, p_.lambda);
    B_ = p_.brick_edge > 0 ? p_.brick_edge : automatic;
    if (!validate(p_, B_, automatic)) return false;
    if (pfh3d_calibrate_seed_radius(p_.target_radius, p_.lambda, B_, &seed_radius_) < 0) {
        std::fprintf(stderr, "[fatal] seed-radius calibration: %s\n", pfh_last_error());
        return false;
    }
    const long long total = p_.total_steps();
    traj_every_ = opt_.traj_interval > 0 ? opt_.traj_interval
                                         : std::max(1LL, total / std::max(1, opt_.traj_samples));
    std::printf("pf3d: periodic cube L=%d, N=%d cells, R=%g, lambda=%g, rho_V %.6f (realized "
                "%.6f), dt=%g, %lld steps, brick B=%d (automatic %d), seed radius %.6f\n",
                p_.L, p_.num_cells, p_.target_radius, p_.lambda, p_.rho, p_.realized_rho(),
                p_.dt, total, B_, automatic, seed_radius_);
    if (!allocate()) return false;
    if (!(opt_.ckpt_in.empty() ? fresh_state() : resume_state())) return false;
    if (fatal_flags(true)) return false;
    if (opt_.use_graph && !capture_graph())
        std::fprintf(stderr, "[warn] graph capture failed; using per-step launches\n");
    return true;
}

bool Sim3D::fatal_flags(bool print) {
    uint32_t f[PFH3D_FLAG_COUNT];
    if (hipMemcpy(f, d_flags_, sizeof(f), hipMemcpyDeviceToHost) != hipSuccess) return true;
    bool fatal = false;
    for (int i = 0; i < PFH3D_FLAG_COUNT; ++i) {
        if (!f[i] || !((PFH3D_FATAL_MASK >> i) & 1u)) continue;
        fatal = true;
        if (print) std::printf("[fatal] integrity flag %d raised %u times\n", i, f[i]);
    }
    if (fatal && print && f[PFH3D_FLAG_SUPPORT_EXHAUSTED])
        std::printf("[fatal] cell support reached the brick margin at B=%d; promotion to larger "
                    "bricks is outside this build: rerun with a larger --brick-edge (the step "
                    "was rejected, the accepted state is intact)\n", B_);
    return fatal;
}

// Advance to at most next_stop: one graph replay when aligned, else one step.
bool Sim3D::advance(long long next_stop) {
    if (graph_exec_ && steps_done_ % kGraphBody == 0 && steps_done_ + kGraphBody <= next_stop) {
        HIP_OK(hipGraphLaunch(graph_exec_, stream_));
        steps_done_ += kGraphBody;
        return true;
    }
    const pfh3d_step_args a = args_for((int)(steps_done_ % 2));
    PFH_CALL(pfh3d_step(&a, slabs_, stream_));
    steps_done_ += 1;
    return true;
}

void Sim3D::print_line() {
    std::vector<pfh3d_cell_state> h(p_.num_cells);
    if (hipMemcpy(h.data(), d_cells_, h.size() * sizeof(pfh3d_cell_state),
                  hipMemcpyDeviceToHost) != hipSuccess)
        return;
    double v = 0.0, vmin = 1e300, vmax = 0.0, speed = 0.0;
    for (const auto& c : h) {
        v += c.V;
        vmin = std::min(vmin, c.V);
        vmax = std::max(vmax, c.V);
        speed += std::sqrt((double)c.velocity_x * c.velocity_x +
                           (double)c.velocity_y * c.velocity_y +
                           (double)c.velocity_z * c.velocity_z);
    }
    std::printf("step %10lld  t %10.3f  <V> %.4f [%.4f, %.4f] / V0 %.4f  <|v|> %.4e\n",
                steps_done_, time(), v / h.size(), vmin, vmax, p_.volume0(), speed / h.size());
    std::fflush(stdout);
}

bool Sim3D::open_trajectory() {
    if (opt_.out_path.empty()) return true;
    const std::string header = trajectory_header(p_, B_, shards_, (unsigned long long)traj_every_);
    if (header.empty()) return false;
    traj_fp_ = std::fopen(opt_.out_path.c_str(), "a+");
    if (!traj_fp_) {
        std::fprintf(stderr, "[3d] cannot open trajectory %s\n", opt_.out_path.c_str());
        return false;
    }
    std::rewind(traj_fp_);
    char first[1600] = {0}, second[256] = {0};
    const bool has1 = std::fgets(first, sizeof(first), traj_fp_) != nullptr;
    const bool has2 = std::fgets(second, sizeof(second), traj_fp_) != nullptr;
    if (!has1) {
        std::clearerr(traj_fp_);
        if (std::fprintf(traj_fp_, "%s\n%s\n", header.c_str(), kColumns) < 0) return false;
        return append_trajectory();            // seeded at the starting state
    }
    first[std::strcspn(first, "\n")] = 0;
    second[std::strcspn(second, "\n")] = 0;
    if (!has2 || header != first || std::strcmp(second, kColumns) != 0) {
        std::fprintf(stderr, "[3d] existing trajectory %s has different metadata; refusing to "
                             "append\n", opt_.out_path.c_str());
        return false;
    }
    std::clearerr(traj_fp_);
    return std::fseek(traj_fp_, 0, SEEK_END) == 0;
}

bool Sim3D::append_trajectory() {
    if (!traj_fp_) return true;
    // the surface is measured on the full-moment cadence: refresh it on the
    // frame's state (no tumble, no motion: the next step re-plans both)
    const pfh3d_step_args a = args_for((int)(steps_done_ % 2));
    PFH_CALL(pfh3d_measure_cells(&a, 0, 1, 0, stream_));
    PFH_CALL(pfh3d_pack_trajectory(d_cells_, d_traj_, p_.num_cells, stream_));
    std::vector<pfh3d_traj_cell> h(p_.num_cells);
    HIP_OK(hipMemcpyAsync(h.data(), d_traj_, h.size() * sizeof(pfh3d_traj_cell),
                          hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    const double t = time();
    for (const auto& c : h) {
        if (std::fprintf(traj_fp_,
                         "%.17g %lld %.17g %.17g %.17g %.9g %.9g %.9g %.9g %.9g %.9g "
                         "%.9g %.9g %.17g %.17g %.17g %.9g\n",
                         t, (long long)c.global_id, c.x, c.y, c.z, (double)c.velocity_x,
                         (double)c.velocity_y, (double)c.velocity_z, (double)c.polarity_x,
                         (double)c.polarity_y, (double)c.polarity_z, (double)c.gamma,
                         (double)c.v_A, c.surface, sphericity(c.volume, c.surface), c.volume,
                         (double)c.phi_max) < 0) {
            std::fprintf(stderr, "[3d] failed to write trajectory frame\n");
            return false;
        }
    }
    if (std::fflush(traj_fp_) != 0) return false;
    ++traj_frames_;
    return true;
}

bool Sim3D::save_checkpoint(const std::string& path) {
    HIP_OK(hipStreamSynchronize(stream_));
    std::vector<pfh3d_cell_state> cells(p_.num_cells);
    HIP_OK(hipMemcpy(cells.data(), d_cells_, cells.size() * sizeof(pfh3d_cell_state),
                     hipMemcpyDeviceToHost));
    pfh3d_ckpt_params pr{};
    pr.Nx = pr.Ny = pr.Nz = p_.L;
    pr.dx = pr.dy = pr.dz = 1.0;
    pr.brick_edge = (uint32_t)B_;
    pr.boundary_flags = 7u;
    pr.dt = p_.dt;
    pr.t_end = p_.t_end;
    pr.lambda = p_.lambda;
    pr.gamma_normal = p_.gamma_normal;
    pr.gamma_soft = p_.gamma_cancer;
    pr.soft_fraction = p_.cancer_fraction;
    pr.kappa = p_.kappa;
    pr.target_radius = p_.target_radius;
    pr.mu = p_.mu;
    pr.v_A = p_.v_A;
    pr.v_A_sigma = p_.v_A_sigma;
    pr.xi = p_.xi;
    pr.tau = p_.tau;
    pr.volume_fraction = p_.rho;
    pr.seed = p_.seed;
    pr.polarity_seed = p_.polarity_stream();
    pr.print_interval = p_.print_interval;
    pr.full_moment_every = p_.full_moment_every;
    pr.verify_every = p_.verify_every;
    pr.integrator_id = 1;
    pr.stencil_id = 2;
    pr.rng_id = 1;
    pr.trajectory_interval = (uint32_t)traj_every_;
    pr.aging_time = p_.aging_time;
    PFH_CALL(pfh3d_checkpoint_write(path.c_str(), steps_done_, time(), &pr, cells.data(),
                                  p_.num_cells, B_, d_phi_[steps_done_ % 2], (uint64_t)shards_,
                                  stream_));
    std::printf("[ckpt] wrote %s at step %lld\n", path.c_str(), steps_done_);
    return true;
}

bool Sim3D::run() {
    const long long total = p_.total_steps();
    const long long pi = p_.print_interval > 0 ? p_.print_interval : std::max(1LL, total);
    const long long traj_every = !opt_.out_path.empty() ? traj_every_ : total + 1;
    long long next_traj = next_boundary(steps_done_, traj_every);
    const bool ckpt_on = !opt_.ckpt_dir.empty();
    long long next_ckpt = ckpt_on && opt_.ckpt_interval > 0 ? steps_done_ + opt_.ckpt_interval
                                                            : LLONG_MAX;
    long long next_save = ckpt_on && opt_.save_interval > 0 ? steps_done_ + opt_.save_interval
                                                            : LLONG_MAX;
    long long next_poll = steps_done_ + kFatalPollEvery;
    long long next_verify = opt_.strict ? next_boundary(steps_done_, p_.verify_every) : LLONG_MAX;
    if (!open_trajectory()) return false;
    bool failed = false;
    while (steps_done_ < total) {
        const long long next_stop = std::min({total, (steps_done_ / pi + 1) * pi, next_traj,
                                              next_ckpt, next_save, next_poll, next_verify});
        if (!advance(next_stop)) {
            failed = true;
            break;
        }
        if (steps_done_ < next_stop) continue;
        if (hipStreamSynchronize(stream_) != hipSuccess || fatal_flags(true)) {
            failed = true;
            break;
        }
        if (steps_done_ >= next_poll) next_poll = steps_done_ + kFatalPollEvery;
        if (steps_done_ % pi == 0 || steps_done_ >= total) print_line();
        if (steps_done_ >= next_traj) {
            if (!append_trajectory()) {
                failed = true;
                break;
            }
            next_traj = next_boundary(steps_done_, traj_every);
        }
        if (steps_done_ >= next_verify) {
            // launch_verify_cells / launch_verify_S (src/pf3d/sim.cu strict mode)
            pfh3d_verify_cell* d_v = nullptr;
            uint32_t* d_m = nullptr;
            pfh3d_layout lay{};
            lay.nx = lay.ny = lay.nz = p_.L;
            lay.pitch_x = P_;
            lay.boundary_flags = 7u;
            std::vector<pfh3d_verify_cell> v(p_.num_cells);
            uint32_t smax = 0;
            const bool ok =
                hipMalloc(&d_v, v.size() * sizeof(pfh3d_verify_cell)) == hipSuccess &&
                hipMalloc(&d_m, sizeof(uint32_t)) == hipSuccess &&
                hipMemset(d_m, 0, sizeof(uint32_t)) == hipSuccess &&
                pfh3d_verify_cells(d_phi_[steps_done_ % 2], d_cells_, d_v, p_.num_cells, B_,
                                   &lay, stream_) == PFH_OK &&
                pfh3d_verify_S(d_S_[steps_done_ % 2], &lay, d_m, stream_) == PFH_OK &&
                hipMemcpyAsync(v.data(), d_v, v.size() * sizeof(pfh3d_verify_cell),
                               hipMemcpyDeviceToHost, stream_) == hipSuccess &&
                hipMemcpyAsync(&smax, d_m, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_) ==
                    hipSuccess &&
                hipStreamSynchronize(stream_) == hipSuccess;
            if (d_v) (void)hipFree(d_v);
            if (d_m) (void)hipFree(d_m);
            if (ok) {
                double rel = 0.0;
                unsigned long long bad = 0;
                for (const auto& c : v) {
                    rel = std::max(rel, std::fabs(c.measured_V - c.state_V) /
                                            std::max(1e-300, std::fabs(c.measured_V)));
                    bad += c.nonfinite_count;
                }
                std::printf("  [verify] step %lld  max rel V error %.3e  non-finite %llu  max S "
                            "%.6f\n", steps_done_, rel, bad, (double)smax * 7.45058059692382812e-9);
            }
            next_verify = next_boundary(steps_done_, p_.verify_every);
        }
        const bool do_ckpt = steps_done_ >= next_ckpt, do_save = steps_done_ >= next_save;
        if (do_ckpt && !save_checkpoint(opt_.ckpt_dir + "/checkpoint.pf3d")) failed = true;
        if (do_save) {
            char name[64];
            std::snprintf(name, sizeof(name), "/checkpoint_%012lld.pf3d", steps_done_);
            if (!save_checkpoint(opt_.ckpt_dir + name)) failed = true;
        }
        if (failed) break;
        if (do_ckpt) next_ckpt = steps_done_ + opt_.ckpt_interval;
        if (do_save) next_save = steps_done_ + opt_.save_interval;
        if (g_terminate) {
            std::printf("[3d] termination requested; stopped at accepted step %lld\n",
                        steps_done_);
            break;
        }
    }
    (void)hipStreamSynchronize(stream_);
    if (!failed && fatal_flags(true)) failed = true;
    if (traj_fp_) {
        std::printf("trajectory -> %lld frames x %d cells\n", traj_frames_, p_.num_cells);
        if (std::fclose(traj_fp_) != 0) failed = true;
        traj_fp_ = nullptr;
    }
    if (ckpt_on && (opt_.final_checkpoint || failed)) {
        if (!save_checkpoint(opt_.ckpt_dir + (failed ? "/checkpoint_failed.pf3d"
                                                     : "/checkpoint.pf3d")))
            failed = true;
    }
    return !failed;
}

bool Sim3D::bench(int steps) {
    const int warm = std::min(steps, 20);
    for (int i = 0; i < warm; ++i)
        if (!advance(LLONG_MAX)) return false;
    while (steps_done_ % kGraphBody)
        if (!advance(LLONG_MAX)) return false;
    HIP_OK(hipStreamSynchronize(stream_));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    const long long from = steps_done_;
    HIP_OK(hipEventRecord(e0, stream_));
    while (steps_done_ - from < steps)
        if (!advance(LLONG_MAX)) return false;
    HIP_OK(hipEventRecord(e1, stream_));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const long long done = steps_done_ - from;
    std::printf("[3d] benchmark %lld steps: %.6f ms/step (%s, B=%d, N=%d, %d shards, %d slabs)\n",
                done, (double)ms / (double)done, graph_exec_ ? "graph" : "per-step launch", B_,
                p_.num_cells, shards_, slabs_);
    return !fatal_flags(true);
}

void usage(const char* argv0) {
    std::printf(
        "Three-dimensional active phase-field cell simulator (periodic cube, gfx950)\n\n"
        "usage: %s [options]\n\n"
        "model and initial condition (fresh runs)\n"
        "  --geometry periodic       boundary geometry (slab and channel are outside this build)\n"
        "  --N <int>                 number of cells                         (288)\n"
        "  --radius <f>              sphere radius R                         (49)\n"
        "  --rho <f>                 target volume fraction; sets the cube   (0.90)\n"
        "  --lambda <f>              interface-width parameter               (7)\n"
        "  --kappa <f>               cell-cell repulsion strength            (10)\n"
        "  --mu <f>                  volume-constraint strength              (1)\n"
        "  --xi <f>                  friction                                (1500)\n"
        "  --tau <f>                 mean run time                           (1e4)\n"
        "  --v-A <f>                 active speed                            (1e-2)\n"
        "  --gamma <f>               deformability coefficient               (1)\n"
        "  --seed <u64>              placement seed                          (1234)\n"
        "  --polarity-seed <u64>     polarity/tumble stream; 0 follows --seed (0)\n"
        "  --initial-centres <csv>   centres with header global_id,x,y,z\n\n"
        "integration\n"
        "  --dt <f>                  explicit-Euler time step                (0.01)\n"
        "  --t-end <f>               target simulation time                  (100)\n"
        "  --aging-time <f>          initial passive duration                (0)\n"
        "  --full-moment <int>       surface-measurement cadence             (100)\n"
        "  --verify-every <int>      strict-check cadence                    (4096)\n"
        "  --brick-edge <int>        aligned brick edge; 0 selects minimum   (0)\n"
        "  --measure-shards <int>    measurement workgroups per cell; 0 auto (0)\n"
        "  --device <int>            HIP device                              (0)\n"
        "  --bench <int>             time this many steps, then exit\n"
        "  --no-graph                per-step launches instead of a 2-step HIP graph\n"
        "  --strict                  run the invariant checks on the verify cadence\n"
        "  --print-interval <int>    steps between status lines              (100)\n\n"
        "output and checkpoints\n"
        "  --out <path>              plain-text trajectory (appends to a matching file)\n"
        "  --trajectory-samples <int>  evenly spaced frames                  (100)\n"
        "  --trajectory-interval <int> steps between frames\n"
        "  -c, --checkpoint <path>   resume from a PF3D checkpoint\n"
        "  --checkpoint-dir <path>   where checkpoints go\n"
        "  --checkpoint-interval <int> steps between rolling checkpoint.pf3d\n"
        "  --save-interval <int>     steps between tagged checkpoint_<step>.pf3d\n"
        "  --no-final-checkpoint     no checkpoint.pf3d at exit\n", argv0);
}

bool parse_d(const char* flag, const char* s, double* out) {
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(s, &end);
    if (errno || end == s || *end || !std::isfinite(v)) {
        std::printf("[fatal] %s expects a finite number, got '%s'\n", flag, s);
        return false;
    }
    *out = v;
    return true;
}

bool parse_ll(const char* flag, const char* s, long long lo, long long hi, long long* out) {
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(s, &end, 10);
    if (errno || end == s || *end || v < lo || v > hi) {
        std::printf("[fatal] %s expects an integer in [%lld, %lld], got '%s'\n", flag, lo, hi, s);
        return false;
    }
    *out = v;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    Params3D p;
    Options3D opt;
    long long iv = 0;
    bool no_final = false;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        auto need = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::printf("[fatal] %s requires a value\n", a);
                std::exit(2);
            }
            return argv[++i];
        };
        auto is = [&](const char* f) { return std::strcmp(a, f) == 0; };
        auto dbl = [&](double* dst) { return parse_d(a, need(), dst); };
        if (is("-h") || is("--help")) {
            usage(argv[0]);
            return 0;
        } else if (is("--geometry")) {
            const char* g = need();
            if (std::strcmp(g, "periodic") != 0) {
                std::printf("[fatal] --geometry %s is not supported by this build (periodic "
                            "only)\n", g);
                return 2;
            }
        } else if (is("--N")) {
            if (!parse_ll(a, need(), 1, 4000000, &iv)) return 2;
            p.num_cells = (int)iv;
        } else if (is("--radius")) {
            if (!dbl(&p.target_radius)) return 2;
        } else if (is("--rho")) {
            if (!dbl(&p.rho)) return 2;
        } else if (is("--lambda")) {
            if (!dbl(&p.lambda)) return 2;
        } else if (is("--kappa")) {
            if (!dbl(&p.kappa)) return 2;
        } else if (is("--mu")) {
            if (!dbl(&p.mu)) return 2;
        } else if (is("--xi")) {
            if (!dbl(&p.xi)) return 2;
        } else if (is("--tau")) {
            if (!dbl(&p.tau)) return 2;
        } else if (is("--v-A")) {
            if (!dbl(&p.v_A)) return 2;
        } else if (is("--v-A-sigma")) {
            if (!dbl(&p.v_A_sigma)) return 2;
        } else if (is("--gamma")) {
            if (!dbl(&p.gamma_normal)) return 2;
        } else if (is("--gamma-cancer")) {
            if (!dbl(&p.gamma_cancer)) return 2;
        } else if (is("--cancer-fraction")) {
            if (!dbl(&p.cancer_fraction)) return 2;
        } else if (is("--seed")) {
            if (!parse_ll(a, need(), 0, LLONG_MAX, &iv)) return 2;
            p.seed = (unsigned long long)iv;
        } else if (is("--polarity-seed")) {
            if (!parse_ll(a, need(), 0, LLONG_MAX, &iv)) return 2;
            p.polarity_seed = (unsigned long long)iv;
        } else if (is("--initial-centres")) {
            opt.centres_path = need();
        } else if (is("--dt")) {
            if (!dbl(&p.dt)) return 2;
        } else if (is("--t-end")) {
            if (!dbl(&p.t_end)) return 2;
        } else if (is("--aging-time")) {
            if (!dbl(&p.aging_time)) return 2;
        } else if (is("--full-moment")) {
            if (!parse_ll(a, need(), 0, 1000000000, &iv)) return 2;
            p.full_moment_every = (int)iv;
        } else if (is("--verify-every")) {
            if (!parse_ll(a, need(), 1, 1000000000, &iv)) return 2;
            p.verify_every = (int)iv;
        } else if (is("--print-interval")) {
            if (!parse_ll(a, need(), 0, 1000000000, &iv)) return 2;
            p.print_interval = (int)iv;
        } else if (is("--brick-edge")) {
            if (!parse_ll(a, need(), 0, 4096, &iv)) return 2;
            p.brick_edge = (int)iv;
        } else if (is("--measure-shards")) {
            if (!parse_ll(a, need(), 0, 1024, &iv)) return 2;
            opt.shards = (int)iv;
        } else if (is("--device")) {
            if (!parse_ll(a, need(), 0, 63, &iv)) return 2;
            opt.device = (int)iv;
        } else if (is("--bench")) {
            if (!parse_ll(a, need(), 1, 100000000, &iv)) return 2;
            opt.bench_steps = (int)iv;
        } else if (is("--no-graph")) {
            opt.use_graph = false;
        } else if (is("--strict")) {
            opt.strict = true;
        } else if (is("--out")) {
            opt.out_path = need();
        } else if (is("--trajectory-samples")) {
            if (!parse_ll(a, need(), 1, 1000000000, &iv)) return 2;
            opt.traj_samples = (int)iv;
        } else if (is("--trajectory-interval")) {
            if (!parse_ll(a, need(), 1, 1000000000, &iv)) return 2;
            opt.traj_interval = iv;
        } else if (is("-c") || is("--checkpoint")) {
            opt.ckpt_in = need();
        } else if (is("--checkpoint-dir")) {
            opt.ckpt_dir = need();
        } else if (is("--checkpoint-interval")) {
            if (!parse_ll(a, need(), 1, 1000000000, &iv)) return 2;
            opt.ckpt_interval = iv;
        } else if (is("--save-interval")) {
            if (!parse_ll(a, need(), 1, 1000000000, &iv)) return 2;
            opt.save_interval = iv;
        } else if (is("--no-final-checkpoint")) {
            no_final = true;
        } else if (is("--slab-height") || is("--channel-height") || is("--wall-kappa") ||
                   is("--wall-width") || is("--wall-padding") || is("--memory-mode") ||
                   is("--memory-fraction") || is("--scratch-slots") || is("--peer-device") ||
                   is("--promoted-shards") || is("--fast-base-shards") ||
                   is("--promoted-measure-shards") || is("--allow-promoted-measure-regroup") ||
                   is("--boundary-out") || is("--boundary-interval") ||
                   is("--boundary-compression") || is("--boundary-projection") ||
                   is("--bench-phases")) {
            std::printf("[fatal] %s is not supported by this build (bounded geometries, adaptive "
                        "storage, boundary output and the peer device are outside it)\n", a);
            return 2;
        } else {
            std::printf("[fatal] unknown option '%s' (try --help)\n", a);
            return 2;
        }
    }
    opt.final_checkpoint = !no_final;
    if ((opt.ckpt_interval > 0 || opt.save_interval > 0) && opt.ckpt_dir.empty()) {
        const std::filesystem::path src = !opt.out_path.empty()
            ? std::filesystem::path(opt.out_path) : std::filesystem::path(opt.ckpt_in);
        opt.ckpt_dir = src.parent_path().empty() ? "." : src.parent_path().string();
    }
    if (!opt.ckpt_dir.empty()) {
        std::error_code ec;
        std::filesystem::create_directories(opt.ckpt_dir, ec);
        if (ec) {
            std::printf("[fatal] cannot create %s: %s\n", opt.ckpt_dir.c_str(),
                        ec.message().c_str());
            return 2;
        }
    }
    if (!(p.target_radius > 0.0) || !(p.lambda > 0.0) || !(p.rho > 0.0)) {
        std::printf("[fatal] radius, lambda and rho must be positive\n");
        return 3;
    }
    if (!opt.ckpt_in.empty()) {
        // the cube side and the model come from the checkpoint
        pfh3d_ckpt_params pr{};
        int64_t step = 0;
        double t = 0.0;
        uint64_t sh = 0;
        int N = 0, B = 0;
        if (pfh3d_checkpoint_read(opt.ckpt_in.c_str(), &pr, &step, &t, &sh, &N, &B, nullptr,
                                  nullptr, nullptr) != PFH_OK) {
            std::printf("[fatal] %s\n", pfh_last_error());
            return 8;
        }
        p.num_cells = N;
        p.L = (int)pr.Nx;
        p.brick_edge = B;
        p.dt = pr.dt;
        p.lambda = pr.lambda;
        p.gamma_normal = pr.gamma_normal;
        p.kappa = pr.kappa;
        p.target_radius = pr.target_radius;
        p.mu = pr.mu;
        p.v_A = pr.v_A;
        p.xi = pr.xi;
        p.tau = pr.tau;
        p.rho = pr.volume_fraction;
        p.seed = pr.seed;
        p.polarity_seed = pr.polarity_seed;
        p.full_moment_every = (int)pr.full_moment_every;
        p.aging_time = pr.aging_time;
        if (opt.shards == 0 && sh > 0) opt.shards = (int)sh;   // the reduction contract
    } else {
        p.L = pfh3d_domain_side(p.num_cells, p.target_radius, p.rho);
        if (p.L <= 0) {
            std::printf("[fatal] domain side: %s\n", pfh_last_error());
            return 3;
        }
    }
    Sim3D sim;
    if (!sim.init(p, opt)) return 5;
    if (opt.bench_steps > 0) return sim.bench(opt.bench_steps) ? 0 : 6;
    std::signal(SIGTERM, on_signal);
    std::signal(SIGINT, on_signal);
    return sim.run() ? 0 : 9;
}
