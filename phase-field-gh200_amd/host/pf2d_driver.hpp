// pf2d host driver: the command-line simulator of the 2D per-cell
// phase-field path on one gfx950 GPU, calling the kernels only through the
// C ABI of include/pfh.h. It keeps the reference simulator's command-line
// flags, run loop, binary checkpoint schema (CELL v1) and plain-text
// trajectory contract, so existing runs, checkpoints and tools carry over.
//
// Reference counterparts:
//   SimParams, validate                 include/params.cuh:498-777
//   Sim (init, resume, run, bench)      include/sim.cuh, src/sim.cu
//   checkpoint schema and I/O           common/checkpoint_format.h,
//                                       src/checkpoint.cu
//   trajectory append contract          include/trajectory.hpp, src/trajectory.cpp
//   command line                        src/main.cu
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "pfh.h"

namespace pf2d {

constexpr double kPi = 3.14159265358979323846;
constexpr double kPhaseFieldM0 = 0.5;          // reference Allen-Cahn mobility
constexpr double kLapRowSumBound = 40.0 / 6.0; // 9-point Laplacian row-sum bound
constexpr double kQInv = 7.45058059692382812e-09;
constexpr int kGraphBody = 6;                  // phi parity x S rotation
constexpr int kDemoteDwell = 200;
constexpr int kPromoteSlack = 8;
constexpr float kSupportEps = 1e-5f;
constexpr long long kFatalPollEvery = 1000;

struct Shape { int wx, wy, tx0, ty0; };
constexpr Shape kShapes[PFH_NUM_CLASSES] = {
    {144, 144, 64, 64}, {176, 144, 32, 64}, {144, 176, 64, 32},
    {160, 160, 32, 32}, {224, 224, 32, 32}, {286, 286, 1, 1},
};
const char* flag_name(int i);
inline bool flag_is_fatal(int i) { return i != PFH_FLAG_SUPPORT_CLIP; }

// Model and run parameters with the reference defaults.
struct SimParams {
    int Nx = 0, Ny = 0;
    double dx = 1.0, dy = 1.0;
    int num_cells = 288;
    double rho = 0.90;
    double dt = 0.01;
    double t_end = 100.0;
    double lambda = 7.0;
    double target_radius = 49.0;
    double kappa = 10.0;
    double mu = 1.0;
    double xi = 1500.0;
    double tau = 1.0e4;
    double v_A = 1.0e-2;
    double gamma_normal = 1.0;
    double gamma_cancer = 0.35;
    double cancer_fraction = 0.0;
    double v_A_sigma = 0.0;
    double phase_field_mobility = kPhaseFieldM0;
    unsigned long long seed = 1234;
    unsigned long long polarity_seed = 0;
    unsigned long long initialization_hash = 0;
    int print_interval = 100;
    int full_moment_every = 100;
    int verify_every = 4096;

    unsigned long long polarity_stream() const { return polarity_seed ? polarity_seed : seed; }
    double area0() const { return kPi * target_radius * target_radius; }
    double bulk() const { return 30.0 / (lambda * lambda); }
    double interaction() const { return 60.0 * kappa / (lambda * lambda); }
    double motility() const { return interaction() / xi; }
    double volume() const { return 2.0 * mu / area0(); }
    double interface_k() const { return 2.7386127875258306 / lambda; }
    double init_radius() const { return target_radius + 0.5 / interface_k(); }
    double realized_rho() const {
        return Nx > 0 ? (double)num_cells * area0() / ((double)Nx * (double)Ny) : 0.0;
    }
    double p_tumble() const;
    long long total_steps() const;
};

// Startup validation (include/params.cuh:596-725); prints what is wrong.
bool validate(const SimParams& p, bool resolved_mobility);
bool validate_effective_stiffness(const SimParams& p, const float* gamma,
                                  const float* mobility, int n);
void print_params(const SimParams& p, int side, int pitch);

// Explicitly supplied options, which override a checkpoint's parameters.
struct SimOverrides {
    bool t_end = false, dt = false, v_A = false, tau = false;
    bool gamma = false, gamma_cancer = false, cancer_fraction = false;
    bool kappa = false, mu = false, xi = false, lambda = false;
    bool target_radius = false, v_A_sigma = false;
    bool phase_field_mobility = false;
    bool seed = false, polarity_seed = false;
    bool print_interval = false, full_moment = false, verify_every = false;
    bool num_cells = false, rho = false;
    bool gamma_policy_changed() const { return gamma || gamma_cancer || cancer_fraction; }
    bool v_A_policy_changed() const { return v_A || v_A_sigma; }
    void apply(SimParams& p, const SimParams& cli) const;
};

// Velocity-component recording (--velocity-moments and friends).
struct VelocityOptions {
    std::string path;
    int spatial_stride = 100;
    long long dense_start = 10000;
    bool reference = false;
};

// Which steps measure the spatial moments (include/velocity_sampling.hpp):
// every step a multiple of the stride, and every step of the dense start
// interval and the first step after (re)start, which graphs cannot capture.
// Graphs cover a period of lcm(6, stride) steps (phi parity, S rotation and
// the stride), split into pieces of `stride` steps for long strides.
class VelocitySampling {
public:
    void configure(bool enabled, bool reference, int stride);
    void start_run(long long step, long long dense_steps) { dense_until_ = step + dense_steps; }
    int graph_period() const { return period_; }
    int graph_steps() const { return graph_steps_; }
    int graph_count() const { return period_ / graph_steps_; }
    int graph_index(long long step) const { return (int)(step % period_) / graph_steps_; }
    bool can_replay(long long step) const {
        return !enabled_ || reference_ || (primed_ && step >= dense_until_);
    }
    bool spatial_sample(int slot, long long step, bool capture) const {
        return slot % stride_ == 0 || (!capture && (!primed_ || step < dense_until_));
    }
    void advanced() { primed_ = true; }

private:
    bool enabled_ = false, reference_ = false, primed_ = false;
    int stride_ = 1, period_ = kGraphBody, graph_steps_ = kGraphBody;
    long long dense_until_ = 0;
};

// PFVMOM4 velocity-moment output (include/velocity_output.cuh,
// common/velocity_format.h): a 56-byte header, then frames of an int64 step
// and cells x 26 float64 columns (id, centroid x/y, theta, gamma, M_pf, v_A,
// V, recorded steps, the 12 cumulative integrals, spatial samples, and the
// four instantaneous endpoint rates Gx, Gy, Kx, Ky).
class VelocityOutput {
public:
    VelocityOutput() = default;
    ~VelocityOutput();
    VelocityOutput(const VelocityOutput&) = delete;
    VelocityOutput& operator=(const VelocityOutput&) = delete;

    bool configure(const VelocityOptions& options, int cells, double dt);
    bool open(const pfh2d_step_args& state, long long step, void* stream);
    bool capture(const pfh2d_step_args& state, long long step, void* stream);
    bool close();
    bool enabled() const { return moments_ != nullptr; }
    bool reference() const { return options_.reference; }
    pfh_moment_accum* accumulators() const { return moments_; }
    long long last_frame() const { return last_frame_; }
    VelocitySampling& sampling() { return sampling_; }
    const VelocitySampling& sampling() const { return sampling_; }

private:
    bool readback(const pfh2d_step_args& state, void* stream);
    bool pack_frame(long long step, int side);
    bool write(const void* data, size_t bytes);

    VelocityOptions options_;
    VelocitySampling sampling_;
    double dt_ = 0.0;
    pfh_moment_accum* moments_ = nullptr;      // device, N
    pfh_moment_accum* endpoint_ = nullptr;     // device, N
    std::vector<pfh_cell_state> cells_;
    std::vector<pfh_moment_accum> integrals_, endpoints_;
    std::vector<double> rows_;
    std::FILE* file_ = nullptr;
    long long start_step_ = 0, last_frame_ = -1;
    bool failed_ = false;
};

struct RunOptions {
    std::string out_path;
    std::string initial_centres_path;
    std::string ckpt_dir;
    long long ckpt_interval = 0;
    long long save_interval = 0;
    bool final_checkpoint = true;
    int traj_samples = 100;
    long long traj_interval = 0;
    int bench_steps = 0;
    bool use_graph = true;
    bool morton = false;
    bool strict = false;
    VelocityOptions velocity;
};

// A checkpoint's state: cell records as device CellState images (with the
// active-window origin), the class array, tiles and the resolved schedule.
struct CheckpointData {
    SimParams params;
    int n = 0;
    long long step = 0;
    double t = 0.0;
    int trajectory_samples = 0;
    long long trajectory_interval = 0;
    bool had_mobi = false;
    std::vector<pfh_cell_state> cells;
    std::vector<uint8_t> cls;
    std::vector<float> phi;
};

bool checkpoint_read(const std::string& path, CheckpointData* out);
void resolve_per_cell_scalars(const SimParams& p, const SimOverrides& ov, CheckpointData* d);

struct CheckpointWriteView {
    const SimParams* p = nullptr;
    long long step = 0;
    double t = 0.0;
    int N = 0, L = 0;
    const pfh_cell_state* cell = nullptr;   // host copies
    const uint8_t* cls = nullptr;
    const float* d_phi = nullptr;           // device tiles of phi^step
    int trajectory_samples = 0;
    long long trajectory_interval = 0;
    int save_interval = 0;
};
bool checkpoint_write(const CheckpointWriteView& v, const std::vector<std::string>& paths);

// Trajectory text file: three header lines, then N rows per frame.
std::string trajectory_metadata(const SimParams& p, int side, long long interval);
bool open_trajectory(const std::string& path, const SimParams& p, int side,
                     long long current_step, long long interval, long long* frames,
                     std::FILE** file);

bool read_centres_csv(const std::string& path, int n, double side, double radius,
                      std::vector<float>* x, std::vector<float>* y, uint64_t* fnv,
                      std::string* error);
bool distinct_output_paths(const std::vector<std::string>& paths);

class Sim {
public:
    ~Sim();
    bool init(const SimParams& p, const RunOptions& opt, int device);
    bool init_from_checkpoint(const SimParams& p, const CheckpointData& d,
                              const RunOptions& opt, int device);
    bool run();
    bool bench(int steps, double* ms_per_step);
    static void request_termination(int);

private:
    bool alloc_device(const SimParams& p, const RunOptions& opt, int device);
    bool configure_and_capture();
    bool seed_positions(std::vector<float>& cx, std::vector<float>& cy);
    pfh2d_step_args args_for_slot(int slot) const;
    bool launch_one(int slot, bool capturing = false);
    bool build_graph();
    bool replay_graph(bool* replayed, long long next_stop);
    bool open_observations();
    bool close_observations(bool valid_state);
    void print_line();
    bool fatal_flag_set();
    void report_flags();
    bool verify(double* max_rel_V, float* max_outside, uint32_t* max_S);
    std::vector<std::string> checkpoint_paths(bool rolling, bool tagged) const;
    bool save_checkpoint(const std::vector<std::string>& paths);
    bool append_trajectory_frame(long long step_at);
    double time() const { return (double)steps_done_ * p_.dt; }

    SimParams p_;
    RunOptions opt_;
    int device_ = 0, grid_ = 0, side_ = 0, pitch_ = 0;
    size_t s_words_ = 0;
    long long steps_done_ = 0;
    long long trajectory_every_ = 1;
    void* stream_ = nullptr;
    float* d_phi_[2] = {nullptr, nullptr};
    uint32_t* d_S_ = nullptr;
    pfh_cell_state* d_cell_ = nullptr;
    uint8_t* d_cls_ = nullptr;
    uint32_t* d_perm_ = nullptr;
    unsigned long long* d_cursor_ = nullptr;
    unsigned long long* d_step_ = nullptr;
    uint32_t* d_flags_ = nullptr;
    double* d_vchk_ = nullptr;
    float* d_ochk_ = nullptr;
    uint32_t* d_smax_ = nullptr;
    pfh_traj_packed_cell* h_traj_ = nullptr;   // mapped pinned host memory
    pfh_traj_packed_cell* d_traj_ = nullptr;
    // hipGraph_t / hipGraphExec_t of each graph piece: one six-step body, or
    // the pieces of the observation period
    std::vector<void*> graphs_, graph_execs_;
    VelocityOutput velocity_;
    std::FILE* traj_fp_ = nullptr;
    long long traj_frames_ = 0;
};

}  // namespace pf2d
