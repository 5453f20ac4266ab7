// pf2d host driver: the simulation object. Owns every device buffer, builds
// the initial condition or restores a checkpoint, and advances the cells
// through the C ABI (pfh2d_launch_step and the set-up kernels), replaying a
// six-step HIP graph between output points.
//
// Reference counterpart: src/sim.cu (Sim::seed_positions :82-156,
// alloc_device :158-225, init :308-370, init_from_checkpoint :375-485,
// args_for_slot :578-618, launch_one :648-669, build_graph :671-687,
// verify :817-849, run :851-1003, bench :1006-1055, trajectory :1105-1141).
#include <algorithm>
#include <climits>
#include <cmath>
#include <csignal>
#include <cstring>

#include <hip/hip_runtime_api.h>

#include "pf2d_driver.hpp"

#define HIP_OK(expr)                                                                       \
    do {                                                                                   \
        const hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "[hip] %s:%d %s -> %s\n", __FILE__, __LINE__, #expr,      \
                         hipGetErrorString(e_));                                           \
            return false;                                                                  \
        }                                                                                  \
    } while (0)

#define PFH_OK_OR_FAIL(expr)                                                               \
    do {                                                                                   \
        const int s_ = (expr);                                                             \
        if (s_ != PFH_OK) {                                                                \
            std::fprintf(stderr, "[pfh] %s:%d %s -> status %d: %s\n", __FILE__, __LINE__,  \
                         #expr, s_, pfh_last_error());                                     \
            return false;                                                                  \
        }                                                                                  \
    } while (0)

namespace pf2d {

namespace {
volatile std::sig_atomic_t g_terminate = 0;

long long next_boundary(long long current, long long interval) {
    if (current < 0 || interval <= 0) return LLONG_MAX;
    const long long r = current % interval;
    const long long d = r == 0 ? interval : interval - r;
    return d <= LLONG_MAX - current ? current + d : LLONG_MAX;
}
}  // namespace

void Sim::request_termination(int) { g_terminate = 1; }

Sim::~Sim() {
    if (traj_fp_) std::fclose(traj_fp_);
    (void)velocity_.close();
    for (void* x : graph_execs_) (void)hipGraphExecDestroy((hipGraphExec_t)x);
    for (void* g : graphs_) (void)hipGraphDestroy((hipGraph_t)g);
    if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
    for (void* p : {(void*)d_phi_[0], (void*)d_phi_[1], (void*)d_S_, (void*)d_cell_,
                    (void*)d_cls_, (void*)d_perm_, (void*)d_cursor_, (void*)d_step_,
                    (void*)d_flags_, (void*)d_vchk_, (void*)d_ochk_, (void*)d_smax_})
        if (p) (void)hipFree(p);
    if (h_traj_) (void)hipHostFree(h_traj_);
}

bool Sim::seed_positions(std::vector<float>& cx, std::vector<float>& cy) {
    const int N = p_.num_cells;
    if (!opt_.initial_centres_path.empty()) {
        std::string err;
        uint64_t fnv = 0;
        if (!read_centres_csv(opt_.initial_centres_path, N, (double)side_, p_.target_radius,
                              &cx, &cy, &fnv, &err)) {
            std::fprintf(stderr, "[fatal] invalid --initial-centres '%s': %s\n",
                         opt_.initial_centres_path.c_str(), err.c_str());
            return false;
        }
        p_.initialization_hash = fnv;
        std::printf("  initial centres  %s  (%d rows, table FNV-1a %016llx)\n",
                    opt_.initial_centres_path.c_str(), N, (unsigned long long)fnv);
        return true;
    }
    cx.resize((size_t)N);
    cy.resize((size_t)N);
    PFH_OK_OR_FAIL(pfh2d_grid_jitter_centres(N, side_, p_.seed, cx.data(), cy.data()));
    p_.initialization_hash = pfh_centre_table_fnv1a64(cx.data(), cy.data(), N);
    std::printf("  initializer      grid+jitter (default), table FNV-1a %016llx\n",
                (unsigned long long)p_.initialization_hash);
    return true;
}

bool Sim::alloc_device(const SimParams& p, const RunOptions& opt, int device) {
    p_ = p;
    opt_ = opt;
    device_ = device;
    HIP_OK(hipSetDevice(device_));
    hipDeviceProp_t prop{};
    HIP_OK(hipGetDeviceProperties(&prop, device_));
    side_ = p_.Nx;
    pitch_ = pfh2d_s_pitch(side_);
    if (pitch_ <= 0) return false;
    const int N = p_.num_cells;
    s_words_ = (size_t)pitch_ * (size_t)side_;
    const size_t pool = (size_t)N * PFH_TILE_AREA * sizeof(float);
    std::printf("--- device ---\n");
    std::printf("  %s  %s  %d CUs  %.1f GiB HBM\n", prop.name, prop.gcnArchName,
                prop.multiProcessorCount, (double)prop.totalGlobalMem / 1073741824.0);
    std::printf("  phi pool  2 x %.2f MB    S  3 x %.2f MB    cells %.2f MB\n",
                (double)pool / 1048576.0, (double)s_words_ * 4.0 / 1048576.0,
                (double)N * sizeof(pfh_cell_state) / 1048576.0);
    HIP_OK(hipMalloc(&d_phi_[0], pool));
    HIP_OK(hipMalloc(&d_phi_[1], pool));
    HIP_OK(hipMalloc(&d_S_, 3 * s_words_ * sizeof(uint32_t)));
    HIP_OK(hipMalloc(&d_cell_, (size_t)N * sizeof(pfh_cell_state)));
    HIP_OK(hipMalloc(&d_cls_, (size_t)N));
    HIP_OK(hipMalloc(&d_perm_, (size_t)N * sizeof(uint32_t)));
    HIP_OK(hipMalloc(&d_cursor_, 2 * sizeof(unsigned long long)));
    HIP_OK(hipMalloc(&d_step_, 2 * sizeof(unsigned long long)));
    HIP_OK(hipMalloc(&d_flags_, PFH_FLAG_COUNT * sizeof(uint32_t)));
    HIP_OK(hipMalloc(&d_vchk_, (size_t)N * sizeof(double)));
    HIP_OK(hipMalloc(&d_ochk_, (size_t)N * sizeof(float)));
    HIP_OK(hipMalloc(&d_smax_, sizeof(uint32_t)));
    // Trajectory records are packed by the GPU straight into mapped host memory.
    HIP_OK(hipHostMalloc((void**)&h_traj_, (size_t)N * sizeof(pfh_traj_packed_cell),
                         hipHostMallocMapped));
    HIP_OK(hipHostGetDevicePointer((void**)&d_traj_, h_traj_, 0));
    hipStream_t s = nullptr;
    HIP_OK(hipStreamCreate(&s));
    stream_ = s;
    HIP_OK(hipMemset(d_cursor_, 0, 2 * sizeof(unsigned long long)));
    HIP_OK(hipMemset(d_step_, 0, 2 * sizeof(unsigned long long)));
    HIP_OK(hipMemset(d_flags_, 0, PFH_FLAG_COUNT * sizeof(uint32_t)));
    HIP_OK(hipMemset(d_S_, 0, 3 * s_words_ * sizeof(uint32_t)));
    grid_ = pfh2d_step_grid(device_);
    if (grid_ <= 0) {
        std::fprintf(stderr, "[pfh] pfh2d_step_grid: %s\n", pfh_last_error());
        return false;
    }
    return true;
}

bool Sim::configure_and_capture() {
    const long long total = p_.total_steps();
    if (!velocity_.configure(opt_.velocity, p_.num_cells, p_.dt)) return false;
    trajectory_every_ = opt_.traj_interval > 0
        ? opt_.traj_interval
        : std::max<long long>(1, total / std::max(1, opt_.traj_samples));
    PFH_OK_OR_FAIL(pfh2d_configure_step());
    pfh_kernel_stats ks{};
    if (pfh2d_query_step_stats(device_, &ks) == PFH_OK)
        std::printf("  k_step: 1024 thr  %d VGPRs  %d AGPRs  %d SGPRs  %d B LDS  %zu B "
                    "private/thread  %d workgroup/CU\n",
                    ks.regs, ks.agprs, ks.sgprs, ks.static_smem + ks.dynamic_smem,
                    ks.local_bytes, ks.ctas_per_sm);
    if (opt_.morton) {
        int M = 1;
        while (M < p_.num_cells) M <<= 1;
        if (pfh2d_configure_morton(M * 8) != PFH_OK) {
            std::fprintf(stderr, "[warn] Morton sort unavailable for N=%d (%s); using identity "
                                 "order.\n", p_.num_cells, pfh_last_error());
            opt_.morton = false;
        }
    }
    if (opt_.use_graph && !build_graph())
        std::fprintf(stderr, "[warn] graph capture failed; using per-step launches.\n");
    return true;
}

bool Sim::init(const SimParams& p, const RunOptions& opt, int device) {
    p_ = p;
    opt_ = opt;
    side_ = p.Nx;
    std::vector<float> cx, cy;
    if (!seed_positions(cx, cy)) return false;
    const int N = p_.num_cells;
    std::vector<pfh_cell_state> h((size_t)N);
    std::memset(h.data(), 0, h.size() * sizeof(pfh_cell_state));
    std::vector<float> g((size_t)N), m((size_t)N);
    const int n_soft = (int)std::llround(p_.cancer_fraction * (double)N);
    for (int i = 0; i < N; ++i) {
        pfh_cell_state& c = h[(size_t)i];
        c.global_id = i;
        c.gamma = (float)(i < n_soft ? p_.gamma_cancer : p_.gamma_normal);
        c.v_A = (float)pfh2d_ic_v_A(i, p_.seed, p_.v_A, p_.v_A_sigma);
        c.M_pf = (float)p_.phase_field_mobility;
        c.R_tgt = (float)p_.target_radius;
        c.theta = pfh2d_ic_theta(i, p_.polarity_stream());
        g[(size_t)i] = c.gamma;
        m[(size_t)i] = c.M_pf;
    }
    if (!validate_effective_stiffness(p_, g.data(), m.data(), N)) return false;
    const SimParams seeded = p_;
    if (!alloc_device(seeded, opt, device)) return false;
    std::vector<uint8_t> hc((size_t)N, 0);
    std::vector<uint32_t> perm((size_t)N);
    for (int i = 0; i < N; ++i) perm[(size_t)i] = (uint32_t)i;
    HIP_OK(hipMemcpy(d_cell_, h.data(), h.size() * sizeof(pfh_cell_state), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_cls_, hc.data(), hc.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_perm_, perm.data(), perm.size() * 4, hipMemcpyHostToDevice));
    float *d_cx = nullptr, *d_cy = nullptr;
    HIP_OK(hipMalloc(&d_cx, (size_t)N * 4));
    HIP_OK(hipMalloc(&d_cy, (size_t)N * 4));
    HIP_OK(hipMemcpy(d_cx, cx.data(), (size_t)N * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_cy, cy.data(), (size_t)N * 4, hipMemcpyHostToDevice));
    PFH_OK_OR_FAIL(pfh2d_init_tiles(d_phi_[0], d_phi_[1], d_cell_, d_cls_, N, side_, d_cx, d_cy,
                                    (float)p_.init_radius(), (float)p_.interface_k(), stream_));
    PFH_OK_OR_FAIL(pfh2d_init_moments(d_cell_, d_cls_, d_phi_[0], N, stream_));
    PFH_OK_OR_FAIL(pfh2d_scatter_all(d_phi_[0], d_cell_, d_cls_, d_S_, N, side_, pitch_,
                                     d_flags_, stream_));
    HIP_OK(hipStreamSynchronize((hipStream_t)stream_));
    (void)hipFree(d_cx);
    (void)hipFree(d_cy);
    return configure_and_capture();
}

bool Sim::init_from_checkpoint(const SimParams& p, const CheckpointData& d, const RunOptions& opt,
                               int device) {
    if ((int)d.cells.size() != d.n || d.phi.size() != (size_t)d.n * PFH_TILE_AREA ||
        p.num_cells != d.n || d.step < 0) {
        std::fprintf(stderr, "[ckpt] payload is internally inconsistent\n");
        return false;
    }
    std::vector<float> g((size_t)d.n), m((size_t)d.n);
    for (int i = 0; i < d.n; ++i) {
        g[(size_t)i] = d.cells[(size_t)i].gamma;
        m[(size_t)i] = d.cells[(size_t)i].M_pf;
    }
    if (!validate_effective_stiffness(p, g.data(), m.data(), d.n)) return false;
    if (p.total_steps() < d.step) {
        std::fprintf(stderr, "[ckpt] requested t_end ends at step %lld, before checkpoint step "
                             "%lld\n", p.total_steps(), d.step);
        return false;
    }
    if (!alloc_device(p, opt, device)) return false;
    const int N = p_.num_cells;
    steps_done_ = d.step;
    const int pin = (int)(steps_done_ % 2), rot = (int)(steps_done_ % 3);
    std::vector<uint32_t> perm((size_t)N);
    for (int i = 0; i < N; ++i) perm[(size_t)i] = (uint32_t)i;
    HIP_OK(hipMemcpy(d_cell_, d.cells.data(), (size_t)N * sizeof(pfh_cell_state),
                     hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_cls_, d.cls.data(), (size_t)N, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_perm_, perm.data(), perm.size() * 4, hipMemcpyHostToDevice));
    const size_t pool = (size_t)N * PFH_TILE_AREA * sizeof(float);
    HIP_OK(hipMemcpy(d_phi_[pin], d.phi.data(), pool, hipMemcpyHostToDevice));
    HIP_OK(hipMemset(d_phi_[1 - pin], 0, pool));
    // The step counter is the Philox counter: restore it in both parity slots.
    const unsigned long long hs[2] = {
        (unsigned long long)(pin == 0 ? steps_done_ : steps_done_ + 1),
        (unsigned long long)(pin == 1 ? steps_done_ : steps_done_ + 1)};
    HIP_OK(hipMemcpy(d_step_, hs, sizeof(hs), hipMemcpyHostToDevice));
    // Moments, support, class and demotion counter come from the file (a
    // recomputation would change a reduction and break restart equality);
    // only the aggregate is rebuilt, in the rotation slot of this step.
    PFH_OK_OR_FAIL(pfh2d_scatter_all(d_phi_[pin], d_cell_, d_cls_, d_S_ + (size_t)rot * s_words_,
                                     N, side_, pitch_, d_flags_, stream_));
    HIP_OK(hipStreamSynchronize((hipStream_t)stream_));
    if (std::fabs(d.t - time()) > 1e-6 * std::max(1.0, std::fabs(d.t)))
        std::fprintf(stderr, "[ckpt] warning: the file records t = %.6f at step %lld, but "
                             "step * dt = %.6f with the dt now in force; the run continues at "
                             "%.6f.\n", d.t, steps_done_, time(), time());
    std::printf("  resumed at step %lld (t = %.4f), phi parity %d, S rotation slot %d\n",
                steps_done_, time(), pin, rot);
    return configure_and_capture();
}

pfh2d_step_args Sim::args_for_slot(int slot) const {
    const int pin = slot % 2, pout = 1 - pin, rs = slot % 3;
    pfh2d_step_args a{};
    a.phi_in = d_phi_[pin];
    a.phi_out = d_phi_[pout];
    a.S_rd = d_S_ + (size_t)rs * s_words_;
    a.S_sc = d_S_ + (size_t)((rs + 1) % 3) * s_words_;
    a.S_cl = d_S_ + (size_t)((rs + 2) % 3) * s_words_;
    a.cell = d_cell_;
    a.cell_cls = d_cls_;
    a.perm = d_perm_;
    a.cursor_use = d_cursor_ + pin;
    a.cursor_clear = d_cursor_ + pout;
    a.step_rd = d_step_ + pin;
    a.step_wr = d_step_ + pout;
    a.flags = d_flags_;
    a.moments = velocity_.reference() ? nullptr : velocity_.accumulators();
    a.physical_dt = p_.dt;
    a.N = p_.num_cells;
    a.L = side_;
    a.P = pitch_;
    a.parity_out = pout;
    a.dt = (float)p_.dt;
    a.A0 = p_.area0();
    a.vol_scale = p_.volume();
    a.bulk_scale = (float)p_.bulk();
    a.rep_coeff = (float)p_.interaction();
    a.mot_coeff = (float)p_.motility();
    a.seed = p_.seed;
    a.polarity_seed = p_.polarity_stream();
    a.p_tumble = p_.p_tumble();
    a.full_moment_every = p_.full_moment_every;
    a.clear_ahead_words = (unsigned long long)s_words_;
    return a;
}

// slot: the step's index in the graph period (phi parity = slot % 2, S
// rotation = slot % 3, spatial sample when slot % stride == 0). With the
// dense reference observer, its measurement of the step's input precedes
// the plain step; otherwise the step itself records (src/sim.cu:648-669).
bool Sim::launch_one(int slot, bool capturing) {
    if (opt_.morton && slot % kGraphBody == 0) {
        int M = 1;
        while (M < p_.num_cells) M <<= 1;
        PFH_OK_OR_FAIL(pfh2d_morton_sort(d_cell_, d_perm_, p_.num_cells, M, side_, stream_));
    }
    const pfh2d_step_args a = args_for_slot(slot % kGraphBody);
    int observation = PFH_OBSERVE_DISABLED;
    if (velocity_.enabled()) {
        if (velocity_.reference())
            PFH_OK_OR_FAIL(pfh2d_velocity_moments(&a, velocity_.accumulators(), p_.dt, stream_));
        else
            observation = velocity_.sampling().spatial_sample(slot, steps_done_, capturing)
                              ? PFH_OBSERVE_SPATIAL : PFH_OBSERVE_ADVECTION;
    }
    PFH_OK_OR_FAIL(pfh2d_launch_step(&a, grid_, stream_, nullptr, 0, 0.0f, observation));
    if (!capturing) velocity_.sampling().advanced();
    return true;
}

bool Sim::build_graph() {
    hipStream_t s = (hipStream_t)stream_;
    const VelocitySampling& sampling = velocity_.sampling();
    const int steps = sampling.graph_steps();
    for (int piece = 0; piece < sampling.graph_count(); ++piece) {
        hipGraph_t g = nullptr;
        HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
        bool ok = true;
        for (int k = 0; k < steps && ok; ++k) ok = launch_one(piece * steps + k, true);
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (!ok || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return false;
        }
        hipGraphExec_t x = nullptr;
        if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            return false;
        }
        graphs_.push_back(g);
        graph_execs_.push_back(x);
    }
    std::printf("  HIP graphs: %d x %d steps cover phi parity, S rotation and the observation "
                "cadence%s\n", sampling.graph_count(), steps,
                opt_.morton ? "; Morton sort every six steps" : "");
    return true;
}

// One replay of the graph piece that starts at the current step, when the
// sampling allows it and the piece ends at or before next_stop.
bool Sim::replay_graph(bool* replayed, long long next_stop) {
    VelocitySampling& sampling = velocity_.sampling();
    const int steps = sampling.graph_steps();
    *replayed = false;
    if (graph_execs_.empty() || !sampling.can_replay(steps_done_) || steps_done_ % steps != 0 ||
        steps_done_ + steps > next_stop)
        return true;
    const hipGraphExec_t x = (hipGraphExec_t)graph_execs_[sampling.graph_index(steps_done_)];
    if (hipGraphLaunch(x, (hipStream_t)stream_) != hipSuccess) {
        std::fprintf(stderr, "[hip] graph launch failed\n");
        return false;
    }
    steps_done_ += steps;
    sampling.advanced();
    *replayed = true;
    return true;
}

bool Sim::open_observations() {
    return velocity_.open(args_for_slot((int)(steps_done_ % kGraphBody)), steps_done_, stream_);
}

bool Sim::close_observations(bool valid_state) {
    bool ok = true;
    if (valid_state && velocity_.enabled() && velocity_.last_frame() != steps_done_)
        ok = velocity_.capture(args_for_slot((int)(steps_done_ % kGraphBody)), steps_done_,
                               stream_);
    if (!velocity_.close()) ok = false;
    return ok;
}

void Sim::print_line() {
    std::vector<pfh_cell_state> h((size_t)p_.num_cells);
    if (hipMemcpy(h.data(), d_cell_, h.size() * sizeof(pfh_cell_state), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return;
    double vsum = 0.0, vmin = 1e300, vmax = -1e300, spd = 0.0, pmax = 0.0;
    long long shifts = 0, tumbles = 0;
    int counts[PFH_NUM_CLASSES] = {0}, bad = 0;
    for (const pfh_cell_state& c : h) {
        vsum += c.V;
        vmin = std::min(vmin, c.V);
        vmax = std::max(vmax, c.V);
        spd += std::sqrt((double)c.vx * c.vx + (double)c.vy * c.vy);
        pmax = std::max(pmax, (double)c.phi_max);
        shifts += c.shift_ctr;
        tumbles += c.tumble_ctr;
        if (c.cls < PFH_NUM_CLASSES) ++counts[c.cls]; else ++bad;
    }
    const double n = (double)p_.num_cells;
    std::printf("step %8lld  t %10.3f  <V>/A0 %.5f  V range [%.1f %.1f]  <|v|> %.4e  "
                "max|phi| %.5f  cls(r/w/t/sq/large/fallback) %d/%d/%d/%d/%d/%d bad=%d  "
                "shifts %lld  tumbles %lld\n",
                steps_done_, time(), vsum / n / p_.area0(), vmin, vmax, spd / n, pmax, counts[0],
                counts[1], counts[2], counts[3], counts[4], counts[5], bad, shifts, tumbles);
    std::fflush(stdout);
}

bool Sim::fatal_flag_set() {
    uint32_t f[PFH_FLAG_COUNT] = {0};
    const hipError_t e = hipMemcpy(f, d_flags_, sizeof(f), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        std::fprintf(stderr, "\n*** STOPPING AT STEP %lld: integrity-flag readback failed: %s ***\n",
                     steps_done_, hipGetErrorString(e));
        return true;
    }
    bool fatal = false;
    for (int i = 0; i < PFH_FLAG_COUNT; ++i) fatal = fatal || (f[i] && flag_is_fatal(i));
    if (!fatal) return false;
    std::fprintf(stderr, "\n*** STOPPING AT STEP %lld: FATAL INTEGRITY CHECK ***\n", steps_done_);
    for (int i = 0; i < PFH_FLAG_COUNT; ++i)
        if (f[i] && flag_is_fatal(i)) std::fprintf(stderr, "    %-18s %u\n", flag_name(i), f[i]);
    std::fprintf(stderr, "    The trajectory is invalid from the first flagged step. A final "
                         "diagnostic checkpoint will be attempted.\n");
    return true;
}

void Sim::report_flags() {
    uint32_t f[PFH_FLAG_COUNT] = {0};
    if (hipMemcpy(f, d_flags_, sizeof(f), hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "[checks] flag readback failed\n");
        return;
    }
    bool fatal = false, advisory = false;
    for (int i = 0; i < PFH_FLAG_COUNT; ++i) {
        if (!f[i]) continue;
        if (i == PFH_FLAG_SUPPORT_CLIP) advisory = true;
        else fatal = true;
    }
    if (!fatal && !advisory) {
        std::printf("integrity checks: all clear\n");
        return;
    }
    if (fatal) {
        std::printf("fatal integrity flags set; trajectory is invalid from the first flagged "
                    "step\n");
        for (int i = 0; i < PFH_FLAG_COUNT; ++i)
            if (f[i] && flag_is_fatal(i)) std::printf("  %-18s %u\n", flag_name(i), f[i]);
    } else {
        std::printf("integrity checks: no fatal flags\n");
    }
    if (advisory)
        std::printf("  advisory: %-18s %u\n", flag_name(PFH_FLAG_SUPPORT_CLIP),
                    f[PFH_FLAG_SUPPORT_CLIP]);
}

bool Sim::verify(double* max_rel_V, float* max_outside, uint32_t* max_S) {
    const int N = p_.num_cells;
    const int pin = (int)(steps_done_ % 2);
    HIP_OK(hipMemsetAsync(d_smax_, 0, sizeof(uint32_t), (hipStream_t)stream_));
    PFH_OK_OR_FAIL(pfh2d_verify_cells(d_phi_[pin], d_cell_, d_cls_, N, d_vchk_, d_ochk_, stream_));
    PFH_OK_OR_FAIL(pfh2d_verify_S(d_S_ + (size_t)(steps_done_ % 3) * s_words_, s_words_, d_smax_,
                                  stream_));
    HIP_OK(hipStreamSynchronize((hipStream_t)stream_));
    std::vector<double> v((size_t)N);
    std::vector<float> o((size_t)N);
    std::vector<pfh_cell_state> h((size_t)N);
    HIP_OK(hipMemcpy(v.data(), d_vchk_, v.size() * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(o.data(), d_ochk_, o.size() * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(h.data(), d_cell_, h.size() * sizeof(pfh_cell_state), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(max_S, d_smax_, 4, hipMemcpyDeviceToHost));
    double mr = 0.0;
    float mo = 0.0f;
    for (int i = 0; i < N; ++i) {
        const double den = std::max(1e-12, std::fabs(h[(size_t)i].V));
        mr = std::max(mr, std::fabs(v[(size_t)i] - h[(size_t)i].V) / den);
        mo = std::max(mo, o[(size_t)i]);
    }
    *max_rel_V = mr;
    *max_outside = mo;
    return true;
}

std::vector<std::string> Sim::checkpoint_paths(bool rolling, bool tagged) const {
    std::vector<std::string> paths;
    if (opt_.ckpt_dir.empty()) return paths;
    if (rolling) paths.push_back(opt_.ckpt_dir + "/checkpoint.bin");
    if (tagged) {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "/checkpoint_%08lld.bin", steps_done_);
        paths.push_back(opt_.ckpt_dir + buf);
    }
    return paths;
}

bool Sim::save_checkpoint(const std::vector<std::string>& paths) {
    if (paths.empty()) return true;
    const int N = p_.num_cells;
    std::vector<pfh_cell_state> h((size_t)N);
    std::vector<uint8_t> hc((size_t)N);
    HIP_OK(hipMemcpy(h.data(), d_cell_, h.size() * sizeof(pfh_cell_state), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(hc.data(), d_cls_, hc.size(), hipMemcpyDeviceToHost));
    CheckpointWriteView v;
    v.p = &p_;
    v.step = steps_done_;
    v.t = time();
    v.N = N;
    v.L = side_;
    v.cell = h.data();
    v.cls = hc.data();
    v.d_phi = d_phi_[steps_done_ % 2];
    v.trajectory_samples = opt_.traj_samples;
    v.trajectory_interval = trajectory_every_;
    v.save_interval = (int)opt_.save_interval;
    return checkpoint_write(v, paths);
}

bool Sim::append_trajectory_frame(long long step_at) {
    if (!traj_fp_) return true;
    const double Lw = (double)side_;
    auto wrap = [Lw](double v) {
        double m = std::fmod(v, Lw);
        return m < 0.0 ? m + Lw : m;
    };
    for (int i = 0; i < p_.num_cells; ++i) {
        const pfh_traj_packed_cell& c = h_traj_[i];
        const double th = (double)c.theta;
        const double l_n = (double)c.perim / (2.0 * kPi * p_.target_radius);
        if (std::fprintf(traj_fp_, "%.17g %d %.6f %.6f %.6f %.6f %.6f %.6f %.6f %.6f %.6f %.6f\n",
                         (double)step_at * p_.dt, c.global_id, wrap((double)c.cx),
                         wrap((double)c.cy), (double)c.vx, (double)c.vy, std::cos(th),
                         std::sin(th), th, (double)c.v_A, l_n, (double)c.volume) < 0) {
            std::fprintf(stderr, "[error] trajectory write failed at step %lld, cell %d\n",
                         step_at, i);
            return false;
        }
    }
    if (std::fflush(traj_fp_) != 0 || std::ferror(traj_fp_)) {
        std::fprintf(stderr, "[error] trajectory flush failed at step %lld\n", step_at);
        return false;
    }
    ++traj_frames_;
    return true;
}

bool Sim::run() {
    const long long total = p_.total_steps();
    const long long pi = p_.print_interval > 0 ? p_.print_interval : std::max(1LL, total);
    bool failed = false;
    long long next_verify = p_.verify_every > 0 ? next_boundary(steps_done_, p_.verify_every)
                                                : total + 1;
    // Frames sit on absolute step multiples, so a continuation keeps the cadence.
    const long long traj_every = !opt_.out_path.empty() ? trajectory_every_ : total + 1;
    long long next_traj = next_boundary(steps_done_, traj_every);
    const bool ckpt_on = !opt_.ckpt_dir.empty();
    if (ckpt_on && total > INT32_MAX) {
        std::fprintf(stderr, "[fatal] the current checkpoint schema supports at most %d steps\n",
                     INT32_MAX);
        return false;
    }
    const long long ckpt_every = ckpt_on ? opt_.ckpt_interval : 0;
    const long long save_every = ckpt_on ? opt_.save_interval : 0;
    long long next_ckpt = ckpt_every > 0 ? steps_done_ + ckpt_every : total + 1;
    long long next_save = save_every > 0 ? steps_done_ + save_every : total + 1;
    long long next_poll = steps_done_ + kFatalPollEvery;
    if (!opt_.out_path.empty() &&
        !open_trajectory(opt_.out_path, p_, side_, steps_done_, trajectory_every_, &traj_frames_,
                         &traj_fp_))
        return false;
    if (!open_observations()) return false;
    const int period = velocity_.sampling().graph_period();
    hipStream_t s = (hipStream_t)stream_;
    while (steps_done_ < total) {
        const long long next_stop = std::min({total, (steps_done_ / pi + 1) * pi, next_traj,
                                              next_ckpt, next_save, next_poll});
        bool replayed = false;
        if (!replay_graph(&replayed, next_stop)) {
            failed = true;
            break;
        }
        if (!replayed) {
            if (!launch_one((int)(steps_done_ % period))) {
                failed = true;
                break;
            }
            steps_done_ += 1;
        }
        const bool do_print = steps_done_ % pi == 0 || steps_done_ >= total;
        const bool do_traj = !opt_.out_path.empty() && steps_done_ >= next_traj;
        const bool do_ckpt = ckpt_every > 0 && steps_done_ >= next_ckpt;
        const bool do_save = save_every > 0 && steps_done_ >= next_save;
        const bool do_poll = steps_done_ >= next_poll;
        if (do_print || do_traj || do_ckpt || do_save || do_poll) {
            (void)hipStreamSynchronize(s);
            if (fatal_flag_set()) {
                failed = true;
                break;
            }
            if (do_poll) next_poll = steps_done_ + kFatalPollEvery;
            if (do_print) print_line();
            if (do_traj) {
                if (pfh2d_pack_traj(d_cell_, d_cls_, d_traj_, p_.num_cells, side_, stream_) !=
                        PFH_OK ||
                    hipStreamSynchronize(s) != hipSuccess || !append_trajectory_frame(steps_done_)) {
                    failed = true;
                    break;
                }
                next_traj = next_boundary(steps_done_, traj_every);
                // a velocity frame at every trajectory frame
                if (velocity_.enabled() &&
                    !velocity_.capture(args_for_slot((int)(steps_done_ % kGraphBody)),
                                       steps_done_, stream_)) {
                    failed = true;
                    break;
                }
            }
            if (do_ckpt || do_save) {
                if (!save_checkpoint(checkpoint_paths(do_ckpt, do_save))) {
                    std::fprintf(stderr, "[ckpt] checkpoint failed at step %lld; stopping.\n",
                                 steps_done_);
                    failed = true;
                    break;
                }
                if (do_ckpt) next_ckpt = steps_done_ + ckpt_every;
                if (do_save) next_save = steps_done_ + save_every;
            }
        }
        if (g_terminate) {
            (void)hipStreamSynchronize(s);
            std::printf("\n[signal] termination requested at step %lld (t = %.3f); %lld "
                        "trajectory frames already written\n", steps_done_, time(), traj_frames_);
            std::fflush(stdout);
            break;
        }
        if (opt_.strict && p_.verify_every > 0 && steps_done_ >= next_verify) {
            double mr = 0.0;
            float mo = 0.0f;
            uint32_t ms = 0;
            if (verify(&mr, &mo, &ms))
                std::printf("  [verify] step %lld  max rel V error %.3e   max|phi| outside window "
                            "%.3e   max S %.6f\n", steps_done_, mr, (double)mo, (double)ms * kQInv);
            next_verify = steps_done_ + p_.verify_every;
        }
    }
    (void)hipStreamSynchronize(s);
    if (!failed && fatal_flag_set()) failed = true;
    if (traj_fp_) {
        std::printf("trajectory -> %lld frames x %d cells (streamed)\n", traj_frames_,
                    p_.num_cells);
        const bool ok = std::fflush(traj_fp_) == 0 && !std::ferror(traj_fp_);
        if (std::fclose(traj_fp_) != 0 || !ok) {
            std::fprintf(stderr, "[error] failed to finalize trajectory output\n");
            failed = true;
        }
        traj_fp_ = nullptr;
    }
    if (!close_observations(!failed)) failed = true;
    if (ckpt_on && (opt_.final_checkpoint || failed)) {
        const std::vector<std::string> paths =
            failed ? std::vector<std::string>{opt_.ckpt_dir + "/checkpoint_failed.bin"}
                   : checkpoint_paths(true, false);
        if (!save_checkpoint(paths)) {
            std::fprintf(stderr, "[ckpt] FINAL CHECKPOINT FAILED. The run cannot be resumed from "
                                 "where it stopped.\n");
            failed = true;
        }
    } else if (failed && !ckpt_on) {
        std::fprintf(stderr, "[ckpt] fatal run has no checkpoint directory; no diagnostic "
                             "checkpoint can be written.\n");
    }
    report_flags();
    return !failed;
}

bool Sim::bench(int steps, double* ms_per_step) {
    hipStream_t s = (hipStream_t)stream_;
    const int period = velocity_.sampling().graph_period();
    const int gsteps = velocity_.sampling().graph_steps();
    const int warm = std::min(steps, 200);
    for (int i = 0; i < warm; ++i) {
        if (!launch_one((int)(steps_done_ % period))) return false;
        ++steps_done_;
    }
    while (steps_done_ % gsteps != 0) {
        if (!launch_one((int)(steps_done_ % period))) return false;
        ++steps_done_;
    }
    HIP_OK(hipStreamSynchronize(s));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    const long long from = steps_done_;
    HIP_OK(hipEventRecord(e0, s));
    while (steps_done_ - from < steps) {
        if (!graph_execs_.empty()) {
            const int g = velocity_.sampling().graph_index(steps_done_);
            HIP_OK(hipGraphLaunch((hipGraphExec_t)graph_execs_[g], s));
            steps_done_ += gsteps;
            velocity_.sampling().advanced();
        } else {
            if (!launch_one((int)(steps_done_ % period))) return false;
            ++steps_done_;
        }
    }
    HIP_OK(hipEventRecord(e1, s));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    const long long done = steps_done_ - from;
    *ms_per_step = (double)ms / (double)done;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::printf("bench: %lld steps in %.3f ms -> %.6f ms/step (%.2f us/step, %d cells, L=%d, "
                "%s)\n", done, (double)ms, *ms_per_step, *ms_per_step * 1000.0, p_.num_cells,
                side_, graph_execs_.empty() ? "per-step launch" : "graph");
    report_flags();
    return true;
}

}  // namespace pf2d
