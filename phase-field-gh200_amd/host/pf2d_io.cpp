// pf2d host driver: parameters, validation, the CELL v1 checkpoint schema
// and the trajectory append contract.
//
// Checkpoint schema (common/checkpoint_format.h:8-147): 44 B fixed prefix,
// 192 B parameter record, int32 tile pitch, 12 B rank trailer, then per
// cell a 96 B record header and the 288 x 288 binary32 tile, then the
// sidecars POLR, GAMA, VA_A, RADI, MOBI (8 B header + N binary32 each).
// Little-endian throughout; the layouts are static-asserted below because
// files written by the reference and by this driver must be interchangeable.
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>

#include <hip/hip_runtime_api.h>

#include "pf2d_driver.hpp"

namespace pf2d {

const char* flag_name(int i) {
    static const char* names[PFH_FLAG_COUNT] = {
        "S_overflow", "q_clamp", "support_clip", "class_exhausted",
        "S_other_negative", "phi_nonfinite", "V_nonpositive", "class_unsupported"};
    return (i >= 0 && i < PFH_FLAG_COUNT) ? names[i] : "unused";
}

double SimParams::p_tumble() const { return -std::expm1(-dt / tau); }

long long SimParams::total_steps() const {
    const double n = t_end / dt + 0.5;
    return (n >= 0.0 && n <= 9.0e18) ? (long long)n : -1LL;
}

bool validate(const SimParams& p, bool resolved_mobility) {
    bool ok = true;
    auto fail = [&](const char* msg) {
        std::fprintf(stderr, "[fatal] %s\n", msg);
        ok = false;
    };
    if (!(p.dx == 1.0 && p.dy == 1.0))
        fail("dx and dy must both be exactly 1.0 (the 9-point Laplacian, the central "
             "gradients and dA = 1 assume h = 1)");
    if (p.Nx != p.Ny) fail("Nx must equal Ny (the aggregate uses one side length)");
    if (p.Nx <= 0) fail("domain side is invalid or exceeds the supported integer range");
    else if (pfh2d_s_pitch(p.Nx) < 0)
        fail("domain side is too large for the aligned shared-field pitch");
    if (p.num_cells <= 0) fail("num_cells must be positive");
    if (!std::isfinite(p.lambda) || !(p.lambda > 0.0)) fail("lambda must be finite and positive");
    if (!std::isfinite(p.target_radius) || !(p.target_radius > 0.0))
        fail("radius must be finite and positive");
    if (!std::isfinite(p.dt) || !(p.dt > 0.0)) fail("dt must be finite and positive");
    if (!std::isfinite(p.t_end) || p.t_end < 0.0) fail("t_end must be finite and non-negative");
    if (p.total_steps() < 0) fail("t_end/dt does not fit in a 64-bit step count");
    if (!std::isfinite(p.xi) || !(p.xi > 0.0)) fail("xi must be finite and positive");
    if (!std::isfinite(p.tau) || !(p.tau > 0.0)) fail("tau must be finite and positive");
    if (!std::isfinite(p.kappa) || p.kappa < 0.0) fail("kappa must be finite and non-negative");
    if (!std::isfinite(p.mu) || p.mu < 0.0) fail("mu must be finite and non-negative");
    if (!std::isfinite(p.phase_field_mobility) ||
        !std::isfinite((float)p.phase_field_mobility) ||
        !((float)p.phase_field_mobility > 0.0f))
        fail("phase-field mobility must be finite and positive in binary32");
    if (!std::isfinite(p.v_A) || p.v_A < 0.0 || !std::isfinite(p.v_A_sigma) || p.v_A_sigma < 0.0)
        fail("v_A and v_A_sigma must be finite and non-negative");
    if (!std::isfinite(p.gamma_normal) || !std::isfinite(p.gamma_cancer) ||
        !(p.gamma_normal > 0.0) || !(p.gamma_cancer > 0.0))
        fail("both gamma values must be finite and positive");
    if (!std::isfinite(p.cancer_fraction) || !(p.cancer_fraction >= 0.0 && p.cancer_fraction <= 1.0))
        fail("cancer_fraction must be finite and lie in [0,1]");
    if (!std::isfinite(p.rho) || !(p.rho > 0.0 && p.rho < 1.0))
        fail("rho must be finite and lie in (0,1)");
    if (p.print_interval < 0 || p.full_moment_every < 0 || p.verify_every < 0)
        fail("output and verification intervals must be non-negative");
    // A window must fit strictly inside the domain: otherwise two window
    // pixels alias one lattice pixel and S - q(phi_n) is no longer exact.
    int wmax = 0;
    for (const Shape& s : kShapes) wmax = std::max(wmax, std::max(s.wx, s.wy));
    if (p.Nx > 0 && p.Nx <= wmax) {
        std::fprintf(stderr, "[fatal] domain side %d must exceed the largest rect dimension %d.\n",
                     p.Nx, wmax);
        ok = false;
    }
    const double gmax = std::max(p.gamma_normal, p.gamma_cancer);
    const double mrel = p.phase_field_mobility / kPhaseFieldM0;
    const double cfl = p.dt * gmax * mrel * kLapRowSumBound;
    if (cfl >= 1.0) {
        if (resolved_mobility) {
            std::fprintf(stderr, "[warn] dt*(M/M0)*gamma*laplacian_bound = %.4f exceeds the "
                                 "safety limit for the uniform mobility; the per-cell values "
                                 "are checked instead.\n", cfl);
        } else {
            std::fprintf(stderr, "[fatal] dt*(M/M0)*gamma*laplacian_bound = %.4f; the safety "
                                 "limit is 1. Reduce dt below %.5g.\n",
                         cfl, 1.0 / (gmax * mrel * kLapRowSumBound));
            ok = false;
        }
    } else if (cfl > 0.5) {
        std::fprintf(stderr, "[warn] dt*(M/M0)*gamma*laplacian_bound = %.4f; margin to the "
                             "safety limit is under 2x.\n", cfl);
    }
    const double ratio = p.interaction() / p.motility();
    if (!(std::fabs(ratio - p.xi) <= 1e-9 * p.xi)) {
        std::fprintf(stderr, "[fatal] interaction/motility = %.17g but xi = %.17g\n", ratio, p.xi);
        ok = false;
    }
    return ok;
}

bool validate_effective_stiffness(const SimParams& p, const float* gamma, const float* mobility,
                                  int n) {
    double geff = 0.0, mmax = 0.0;
    for (int i = 0; i < n; ++i) {
        const double g = gamma[i], m = mobility[i];
        if (!(g > 0.0) || !std::isfinite(g) || !(m > 0.0) || !std::isfinite(m)) {
            std::fprintf(stderr, "[fatal] cell %d has gamma = %.9g, phase-field mobility = %.9g; "
                                 "both must be positive and finite.\n", i, g, m);
            return false;
        }
        geff = std::fmax(geff, (m / kPhaseFieldM0) * g);
        mmax = std::fmax(mmax, m / kPhaseFieldM0);
    }
    const double cfl = p.dt * geff * kLapRowSumBound;
    if (cfl >= 1.0) {
        std::fprintf(stderr, "[fatal] max_i[(M_i/M0)*gamma_i] = %.6g gives dt*coeff*"
                             "laplacian_bound = %.4f; the safety limit is 1.\n", geff, cfl);
        return false;
    }
    if (cfl > 0.5)
        std::fprintf(stderr, "[warn] dt*coeff*laplacian_bound = %.4f leaves under 2x margin\n", cfl);
    if (p.dt * mmax * p.interaction() > 0.5)
        std::fprintf(stderr, "[warn] the overlap-repulsion term has under 2x explicit-Euler "
                             "margin at S_other ~ 1\n");
    if (p.dt * mmax * p.volume() * p.area0() > 0.5)
        std::fprintf(stderr, "[warn] the area-constraint term has under 2x explicit-Euler margin\n");
    return true;
}

void print_params(const SimParams& p, int side, int pitch) {
    std::printf("--- simulation configuration ---\n");
    std::printf("  cells            %d\n", p.num_cells);
    std::printf("  domain           %d x %d  (rho target/realized = %.6g / %.6g)\n", side, side,
                p.rho, p.realized_rho());
    std::printf("  S pitch          %d uint32  (%.2f MB/buffer, x3)\n", pitch,
                (double)pitch * side * 4.0 / 1048576.0);
    std::printf("  R, lambda        %.4g, %.4g   (init R_eff = %.6g)\n", p.target_radius,
                p.lambda, p.init_radius());
    std::printf("  dt, t_end        %.6g, %.6g  (%lld steps)\n", p.dt, p.t_end, p.total_steps());
    std::printf("  kappa, mu, xi    %.6g, %.6g, %.6g\n", p.kappa, p.mu, p.xi);
    std::printf("  gamma n/c        %.6g / %.6g  (cancer fraction %.4g)\n", p.gamma_normal,
                p.gamma_cancer, p.cancer_fraction);
    std::printf("  phase-field M    %.6g  (M/M0 = %.6g)\n", p.phase_field_mobility,
                p.phase_field_mobility / kPhaseFieldM0);
    std::printf("  v_A, tau         %.6g, %.6g\n", p.v_A, p.tau);
    std::printf("  A0               %.6f\n", p.area0());
    std::printf("  bulk  30/lambda^2             %.9g\n", p.bulk());
    std::printf("  inter 60*kappa/lambda^2       %.9g\n", p.interaction());
    std::printf("  motil 60*kappa/(xi*lambda^2)  %.9g\n", p.motility());
    std::printf("  vol   2mu/A0     %.9g\n", p.volume());
    std::printf("  p_tumble         %.9e   (-expm1(-dt/tau))\n", p.p_tumble());
    std::printf("  fatal checks     ENABLED\n");
}

void SimOverrides::apply(SimParams& p, const SimParams& cli) const {
    if (t_end) p.t_end = cli.t_end;
    if (dt) p.dt = cli.dt;
    if (v_A) p.v_A = cli.v_A;
    if (v_A_sigma) p.v_A_sigma = cli.v_A_sigma;
    if (phase_field_mobility) p.phase_field_mobility = cli.phase_field_mobility;
    if (tau) p.tau = cli.tau;
    if (gamma) p.gamma_normal = cli.gamma_normal;
    if (gamma_cancer) p.gamma_cancer = cli.gamma_cancer;
    if (cancer_fraction) p.cancer_fraction = cli.cancer_fraction;
    if (kappa) p.kappa = cli.kappa;
    if (mu) p.mu = cli.mu;
    if (xi) p.xi = cli.xi;
    if (lambda) p.lambda = cli.lambda;
    if (target_radius) p.target_radius = cli.target_radius;
    if (seed) p.seed = cli.seed;
    if (polarity_seed) p.polarity_seed = cli.polarity_seed;
    if (print_interval) p.print_interval = cli.print_interval;
    if (full_moment) p.full_moment_every = cli.full_moment_every;
    if (verify_every) p.verify_every = cli.verify_every;
}

// ---------------------------------------------------------------------------
// CELL v1 checkpoint records
// ---------------------------------------------------------------------------
namespace {

constexpr uint32_t kMagic = 0x43454C4Cu;      // 'CELL'
constexpr uint32_t kFormat = 1u;
constexpr int32_t kPitch = PFH_TILE_PITCH;
constexpr uint32_t kPOLR = 0x504F4C52u, kGAMA = 0x47414D41u, kVA_A = 0x56415F41u,
                   kRADI = 0x52414449u, kMOBI = 0x4D4F4249u;

#pragma pack(push, 1)
struct FixedPrefix {
    uint32_t magic, version;
    int32_t step;
    double cur_time;
    int32_t num_cells_local, save_interval, reserved, trajectory_samples;
    uint8_t bools[4];
    uint32_t sp_sz;
};
struct RankTrailer { int32_t num_ranks, rank_id, num_cells_global; };
struct CellRecord {
    int32_t cell_id, origin_x, origin_y, shape_class;
    uint32_t promote_ctr, reserved0;
    float cx, cy, vx, vy;
    double volume_moment, moment_x, moment_y, perimeter;
    int32_t support_lo_x, support_hi_x, support_lo_y, support_hi_y;
    float phi_max;
    uint32_t reserved1;
};
struct SidecarHeader { uint32_t magic; int32_t count; };
struct ParamsRecord {
    int32_t Nx, Ny;
    double dx, dy, dt, t_end, rho, lambda, gamma_normal, gamma_soft, soft_fraction, kappa,
        target_radius, mu, v_A, v_A_sigma, xi, tau;
    uint64_t seed, polarity_seed, initialization_hash;
    int32_t print_interval, full_moment_every, verify_every, save_interval, trajectory_samples,
        reserved;
    int64_t trajectory_interval;
};
#pragma pack(pop)
static_assert(sizeof(FixedPrefix) == 44, "FixedPrefix layout");
static_assert(sizeof(RankTrailer) == 12, "RankTrailer layout");
static_assert(sizeof(CellRecord) == 96, "CellRecordHeader layout");
static_assert(sizeof(SidecarHeader) == 8, "SidecarBlockHeader layout");
static_assert(sizeof(ParamsRecord) == 192, "CheckpointParamsRecord layout");
static_assert(offsetof(ParamsRecord, rho) == 40 && offsetof(ParamsRecord, lambda) == 48 &&
              offsetof(ParamsRecord, gamma_normal) == 56 &&
              offsetof(ParamsRecord, target_radius) == 88 && offsetof(ParamsRecord, v_A) == 104 &&
              offsetof(ParamsRecord, tau) == 128 && offsetof(ParamsRecord, seed) == 136 &&
              offsetof(ParamsRecord, polarity_seed) == 144 &&
              offsetof(ParamsRecord, initialization_hash) == 152 &&
              offsetof(ParamsRecord, print_interval) == 160 &&
              offsetof(ParamsRecord, trajectory_interval) == 184,
              "parameter record offsets");
static_assert(offsetof(CellRecord, volume_moment) == 40 && offsetof(CellRecord, perimeter) == 64 &&
              offsetof(CellRecord, phi_max) == 88, "cell record offsets");

int wrapi(long long v, int L) {
    v %= L;
    return (int)(v < 0 ? v + L : v);
}
double wrapd(double v, double L) {
    double m = std::fmod(v, L);
    return m < 0.0 ? m + L : m;
}

bool read_exact(std::FILE* f, void* dst, size_t bytes, const char* what) {
    if (bytes == 0) return true;
    if (std::fread(dst, 1, bytes, f) != bytes) {
        std::fprintf(stderr, "[ckpt] short read of %s (%zu B)\n", what, bytes);
        return false;
    }
    return true;
}

void params_from_record(const ParamsRecord& f, int n, SimParams* p) {
    p->Nx = f.Nx; p->Ny = f.Ny; p->dx = f.dx; p->dy = f.dy; p->dt = f.dt; p->t_end = f.t_end;
    p->rho = f.rho; p->lambda = f.lambda; p->gamma_normal = f.gamma_normal;
    p->gamma_cancer = f.gamma_soft; p->cancer_fraction = f.soft_fraction; p->kappa = f.kappa;
    p->target_radius = f.target_radius; p->mu = f.mu; p->v_A = f.v_A;
    p->v_A_sigma = f.v_A_sigma; p->xi = f.xi; p->tau = f.tau; p->num_cells = n;
    p->seed = f.seed; p->polarity_seed = f.polarity_seed;
    p->initialization_hash = f.initialization_hash; p->print_interval = f.print_interval;
    p->full_moment_every = f.full_moment_every; p->verify_every = f.verify_every;
}

ParamsRecord params_to_record(const SimParams& p, int traj_samples, long long traj_interval,
                              int save_interval) {
    ParamsRecord f;
    std::memset(&f, 0, sizeof(f));
    f.Nx = p.Nx; f.Ny = p.Ny; f.dx = p.dx; f.dy = p.dy; f.dt = p.dt; f.t_end = p.t_end;
    f.rho = p.rho; f.lambda = p.lambda; f.gamma_normal = p.gamma_normal;
    f.gamma_soft = p.gamma_cancer; f.soft_fraction = p.cancer_fraction; f.kappa = p.kappa;
    f.target_radius = p.target_radius; f.mu = p.mu; f.v_A = p.v_A; f.v_A_sigma = p.v_A_sigma;
    f.xi = p.xi; f.tau = p.tau; f.seed = p.seed;
    f.polarity_seed = p.polarity_stream();   // the resolved stream
    f.initialization_hash = p.initialization_hash; f.print_interval = p.print_interval;
    f.full_moment_every = p.full_moment_every; f.verify_every = p.verify_every;
    f.save_interval = save_interval; f.trajectory_samples = traj_samples;
    f.trajectory_interval = traj_interval;
    return f;
}

// Support box (phi > kSupportEps) and nonzero box of a whole tile.
bool tile_boxes(const float* t, int sup_lo[2], int sup_hi[2], int nz_lo[2], int nz_hi[2]) {
    for (int d = 0; d < 2; ++d) {
        sup_lo[d] = nz_lo[d] = kPitch;
        sup_hi[d] = nz_hi[d] = -1;
    }
    for (int y = 0; y < kPitch; ++y)
        for (int x = 0; x < kPitch; ++x) {
            const float v = t[(size_t)y * kPitch + x];
            if (v == 0.0f) continue;
            nz_lo[0] = std::min(nz_lo[0], x); nz_hi[0] = std::max(nz_hi[0], x);
            nz_lo[1] = std::min(nz_lo[1], y); nz_hi[1] = std::max(nz_hi[1], y);
            if (v <= kSupportEps) continue;
            sup_lo[0] = std::min(sup_lo[0], x); sup_hi[0] = std::max(sup_hi[0], x);
            sup_lo[1] = std::min(sup_lo[1], y); sup_hi[1] = std::max(sup_hi[1], y);
        }
    return sup_hi[0] >= 0;
}

bool read_sidecars(std::FILE* f, CheckpointData* out) {
    const int n = out->n;
    bool seen_va = false, seen_gamma = false, seen_radius = false, seen_polr = false;
    std::vector<float> buf;
    for (;;) {
        SidecarHeader h{};
        const size_t got = std::fread(&h, 1, sizeof(h), f);
        if (got == 0) break;
        if (got != sizeof(h)) {
            std::fprintf(stderr, "[ckpt] truncated sidecar header\n");
            return false;
        }
        bool* seen = h.magic == kVA_A ? &seen_va : h.magic == kGAMA ? &seen_gamma
                   : h.magic == kRADI ? &seen_radius : h.magic == kPOLR ? &seen_polr
                   : h.magic == kMOBI ? &out->had_mobi : nullptr;
        if (!seen) {
            std::fprintf(stderr, "[ckpt] unsupported sidecar 0x%08X in current checkpoint\n",
                         h.magic);
            return false;
        }
        if (h.count != n) {
            std::fprintf(stderr, "[ckpt] sidecar 0x%08X has %d entries; expected %d\n", h.magic,
                         h.count, n);
            return false;
        }
        buf.assign((size_t)n, 0.0f);
        if (!read_exact(f, buf.data(), (size_t)n * 4, "sidecar payload")) return false;
        for (int i = 0; i < n; ++i) {
            const float v = buf[(size_t)i];
            const bool valid = std::isfinite(v) &&
                (h.magic == kPOLR || (h.magic == kVA_A ? v >= 0.0f : v > 0.0f));
            if (!valid) {
                std::fprintf(stderr, "[ckpt] sidecar 0x%08X has an invalid value at cell %d\n",
                             h.magic, i);
                return false;
            }
        }
        if (*seen) {
            const char tag[5] = {(char)(h.magic >> 24), (char)(h.magic >> 16),
                                 (char)(h.magic >> 8), (char)h.magic, 0};
            std::fprintf(stderr, "[ckpt] duplicate %s sidecar\n", tag);
            return false;
        }
        *seen = true;
        for (int i = 0; i < n; ++i) {
            pfh_cell_state& c = out->cells[(size_t)i];
            const float v = buf[(size_t)i];
            if (h.magic == kVA_A) c.v_A = v;
            else if (h.magic == kGAMA) c.gamma = v;
            else if (h.magic == kRADI) c.R_tgt = v;
            else if (h.magic == kPOLR) c.theta = v;
            else c.M_pf = v;
        }
    }
    if (!seen_va || !seen_gamma || !seen_radius || !seen_polr) {
        std::fprintf(stderr, "[ckpt] current checkpoint is missing required sidecars:%s%s%s%s\n",
                     seen_polr ? "" : " POLR", seen_gamma ? "" : " GAMA", seen_va ? "" : " VA_A",
                     seen_radius ? "" : " RADI");
        return false;
    }
    return true;
}

// One payload into every path through <path>.tmp files, renamed only after
// every write and close succeeded.
class FanOut {
public:
    bool open(const std::vector<std::string>& paths) {
        for (const std::string& p : paths) {
            std::FILE* f = std::fopen((p + ".tmp").c_str(), "wb");
            if (!f) {
                std::fprintf(stderr, "[ckpt] cannot open %s.tmp: %s\n", p.c_str(),
                             std::strerror(errno));
                discard();
                return false;
            }
            final_.push_back(p);
            fps_.push_back(f);
        }
        return !fps_.empty();
    }
    bool write(const void* src, size_t bytes, const char* what) {
        for (size_t i = 0; i < fps_.size(); ++i)
            if (std::fwrite(src, 1, bytes, fps_[i]) != bytes) {
                std::fprintf(stderr, "[ckpt] short write of %s to %s.tmp: %s\n", what,
                             final_[i].c_str(), std::strerror(errno));
                return false;
            }
        return true;
    }
    bool commit() {
        bool ok = true;
        for (size_t i = 0; i < fps_.size(); ++i) {
            if (std::fclose(fps_[i]) != 0) {
                std::fprintf(stderr, "[ckpt] failed to close %s.tmp\n", final_[i].c_str());
                ok = false;
            }
            fps_[i] = nullptr;
        }
        fps_.clear();
        if (!ok) {
            unlink_tmp();
            return false;
        }
        for (const std::string& p : final_) {
            std::error_code ec;
            std::filesystem::rename(p + ".tmp", p, ec);
            if (ec) {
                std::fprintf(stderr, "[ckpt] rename %s.tmp -> %s failed: %s\n", p.c_str(),
                             p.c_str(), ec.message().c_str());
                ok = false;
            }
        }
        return ok;
    }
    void discard() {
        for (std::FILE* f : fps_)
            if (f) std::fclose(f);
        fps_.clear();
        unlink_tmp();
    }
    ~FanOut() {
        for (std::FILE* f : fps_)
            if (f) std::fclose(f);
    }

private:
    void unlink_tmp() {
        std::error_code ec;
        for (const std::string& p : final_) std::filesystem::remove(p + ".tmp", ec);
    }
    std::vector<std::FILE*> fps_;
    std::vector<std::string> final_;
};

constexpr int kStageCells = 64;

}  // namespace

bool checkpoint_read(const std::string& path, CheckpointData* out) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "[ckpt] cannot open %s: %s\n", path.c_str(), std::strerror(errno));
        return false;
    }
    struct Closer { std::FILE* f; ~Closer() { std::fclose(f); } } closer{f};

    FixedPrefix pre{};
    if (!read_exact(f, &pre, sizeof(pre), "fixed prefix")) return false;
    if (pre.magic != kMagic) {
        std::fprintf(stderr, "[ckpt] %s is not a cell checkpoint (magic 0x%08X, expected 0x%08X)\n",
                     path.c_str(), pre.magic, kMagic);
        return false;
    }
    if (pre.version != kFormat) {
        std::fprintf(stderr, "[ckpt] %s uses unsupported 2D checkpoint schema %u; expected %u\n",
                     path.c_str(), pre.version, kFormat);
        return false;
    }
    if (pre.step < 0 || pre.reserved != 0 || pre.save_interval < 0 ||
        pre.trajectory_samples < 0 || pre.bools[0] || pre.bools[1] || pre.bools[2] ||
        pre.bools[3]) {
        std::fprintf(stderr, "[ckpt] fixed checkpoint metadata is invalid or unsupported\n");
        return false;
    }
    if (pre.sp_sz != sizeof(ParamsRecord)) {
        std::fprintf(stderr, "[ckpt] schema %u parameter record is %u B; expected %zu B. "
                             "The record layout is incompatible.\n",
                     pre.version, pre.sp_sz, sizeof(ParamsRecord));
        return false;
    }
    if (pre.num_cells_local <= 0 || pre.num_cells_local > 4000000) {
        std::fprintf(stderr, "[ckpt] implausible cell count %d\n", pre.num_cells_local);
        return false;
    }
    ParamsRecord sp{};
    if (!read_exact(f, &sp, sizeof(sp), "parameter record")) return false;
    if (sp.reserved != 0 || sp.save_interval != pre.save_interval ||
        sp.trajectory_samples != pre.trajectory_samples || sp.trajectory_samples <= 0 ||
        sp.trajectory_interval < 0) {
        std::fprintf(stderr, "[ckpt] checkpoint parameter metadata is internally inconsistent\n");
        return false;
    }
    int32_t pitch = 0;
    if (!read_exact(f, &pitch, 4, "tile pitch")) return false;
    if (pitch != kPitch) {
        std::fprintf(stderr, "[ckpt] tile pitch %d does not match schema %u pitch %d\n", pitch,
                     pre.version, kPitch);
        return false;
    }
    RankTrailer tr{};
    if (!read_exact(f, &tr, sizeof(tr), "rank trailer")) return false;
    if (tr.num_ranks != 1 || tr.rank_id != 0 || tr.num_cells_global != pre.num_cells_local) {
        std::fprintf(stderr, "[ckpt] %s is not a complete single-GPU checkpoint (ranks=%d, "
                             "rank=%d, global cells=%d, local cells=%d)\n",
                     path.c_str(), tr.num_ranks, tr.rank_id, tr.num_cells_global,
                     pre.num_cells_local);
        return false;
    }
    const int n = pre.num_cells_local;
    out->n = n;
    out->step = pre.step;
    out->t = pre.cur_time;
    out->trajectory_samples = sp.trajectory_samples;
    out->trajectory_interval = sp.trajectory_interval;
    out->params = SimParams{};
    params_from_record(sp, n, &out->params);
    const double expect_t = (double)pre.step * out->params.dt;
    if (!std::isfinite(pre.cur_time) || !std::isfinite(expect_t) ||
        std::fabs(pre.cur_time - expect_t) > 1.0e-12 * std::max(1.0, std::fabs(expect_t))) {
        std::fprintf(stderr, "[ckpt] checkpoint time does not match step multiplied by dt\n");
        return false;
    }
    if (!validate(out->params, true)) {
        std::fprintf(stderr, "[ckpt] checkpoint parameters are invalid\n");
        return false;
    }
    const int L = out->params.Nx;
    out->cells.assign((size_t)n, pfh_cell_state{});
    out->cls.assign((size_t)n, 0);
    out->phi.assign((size_t)n * PFH_TILE_AREA, 0.0f);
    for (int i = 0; i < n; ++i) {
        pfh_cell_state& c = out->cells[(size_t)i];
        c.gamma = (float)out->params.gamma_normal;
        c.v_A = (float)out->params.v_A;
        c.R_tgt = (float)out->params.target_radius;
        c.M_pf = (float)kPhaseFieldM0;
    }
    int hist[PFH_NUM_CLASSES] = {0};
    int max_ext[2] = {0, 0};
    for (int i = 0; i < n; ++i) {
        CellRecord rec{};
        float* tile = out->phi.data() + (size_t)i * PFH_TILE_AREA;
        if (!read_exact(f, &rec, sizeof(rec), "cell record header") ||
            !read_exact(f, tile, (size_t)PFH_TILE_AREA * 4, "phi tile"))
            return false;
        if (rec.cell_id != i) {
            std::fprintf(stderr, "[ckpt] cell slot %d carries global id %d; current checkpoints "
                                 "require canonical IDs 0..N-1\n", i, rec.cell_id);
            return false;
        }
        if (rec.shape_class < 0 || rec.shape_class >= PFH_NUM_CLASSES ||
            rec.promote_ctr >= (uint32_t)kDemoteDwell || rec.reserved0 || rec.reserved1 ||
            rec.origin_x < 0 || rec.origin_x >= L || rec.origin_y < 0 || rec.origin_y >= L ||
            !std::isfinite(rec.cx) || !std::isfinite(rec.cy) || !std::isfinite(rec.vx) ||
            !std::isfinite(rec.vy) || !std::isfinite(rec.volume_moment) ||
            !(rec.volume_moment > 0.0) || !std::isfinite(rec.moment_x) ||
            !std::isfinite(rec.moment_y) || !std::isfinite(rec.perimeter) ||
            rec.perimeter < 0.0 || !std::isfinite(rec.phi_max) || rec.phi_max < 0.0f) {
            std::fprintf(stderr, "[ckpt] cell %d has invalid state metadata\n", rec.cell_id);
            return false;
        }
        const Shape sc = kShapes[rec.shape_class];
        if (rec.support_lo_x < 0 || rec.support_hi_x < rec.support_lo_x ||
            rec.support_hi_x >= sc.wx || rec.support_lo_y < 0 ||
            rec.support_hi_y < rec.support_lo_y || rec.support_hi_y >= sc.wy) {
            std::fprintf(stderr, "[ckpt] cell %d support bounds do not fit shape class %d\n",
                         rec.cell_id, rec.shape_class);
            return false;
        }
        float pmax = 0.0f;
        for (size_t k = 0; k < (size_t)PFH_TILE_AREA; ++k) {
            if (!std::isfinite(tile[k])) {
                std::fprintf(stderr, "[ckpt] cell %d phase field contains a non-finite value\n",
                             rec.cell_id);
                return false;
            }
            pmax = std::max(pmax, std::fabs(tile[k]));
        }
        int slo[2], shi[2], nlo[2], nhi[2];
        if (!tile_boxes(tile, slo, shi, nlo, nhi)) {
            std::fprintf(stderr, "[ckpt] cell %d has no phase-field support above %.1e\n",
                         rec.cell_id, (double)kSupportEps);
            return false;
        }
        if (slo[0] != sc.tx0 + rec.support_lo_x || shi[0] != sc.tx0 + rec.support_hi_x ||
            slo[1] != sc.ty0 + rec.support_lo_y || shi[1] != sc.ty0 + rec.support_hi_y ||
            nlo[0] < sc.tx0 || nhi[0] >= sc.tx0 + sc.wx || nlo[1] < sc.ty0 ||
            nhi[1] >= sc.ty0 + sc.wy || pmax != rec.phi_max) {
            std::fprintf(stderr, "[ckpt] cell %d phase field disagrees with its stored class or "
                                 "moments\n", rec.cell_id);
            return false;
        }
        const float ecx = (float)wrapd((double)rec.origin_x + sc.tx0 +
                                           rec.moment_x / rec.volume_moment, (double)L);
        const float ecy = (float)wrapd((double)rec.origin_y + sc.ty0 +
                                           rec.moment_y / rec.volume_moment, (double)L);
        if (rec.cx != ecx || rec.cy != ecy) {
            std::fprintf(stderr, "[ckpt] cell %d derived centroid is internally inconsistent\n",
                         rec.cell_id);
            return false;
        }
        max_ext[0] = std::max(max_ext[0], rec.support_hi_x - rec.support_lo_x + 1);
        max_ext[1] = std::max(max_ext[1], rec.support_hi_y - rec.support_lo_y + 1);
        pfh_cell_state& c = out->cells[(size_t)i];
        c.global_id = rec.cell_id;
        c.gx0 = wrapi((long long)rec.origin_x + sc.tx0, L);
        c.gy0 = wrapi((long long)rec.origin_y + sc.ty0, L);
        c.cls = (uint8_t)rec.shape_class;
        c.cls_written[0] = c.cls_written[1] = c.cls;
        c.vx = rec.vx;
        c.vy = rec.vy;
        c.V = rec.volume_moment;
        c.Cx = rec.moment_x;
        c.Cy = rec.moment_y;
        c.perim = rec.perimeter;
        c.bb_lo_x = rec.support_lo_x;
        c.bb_hi_x = rec.support_hi_x;
        c.bb_lo_y = rec.support_lo_y;
        c.bb_hi_y = rec.support_hi_y;
        c.promote_ctr = rec.promote_ctr;
        c.phi_max = rec.phi_max;
        out->cls[(size_t)i] = c.cls;
        ++hist[rec.shape_class];
    }
    if (!read_sidecars(f, out)) return false;
    std::vector<float> g((size_t)n), m((size_t)n);
    for (int i = 0; i < n; ++i) {
        g[(size_t)i] = out->cells[(size_t)i].gamma;
        m[(size_t)i] = out->cells[(size_t)i].M_pf;
    }
    if (!validate_effective_stiffness(out->params, g.data(), m.data(), n)) return false;
    std::printf("[ckpt] loaded %s: schema %u, step %d, t %.4f, %d cells, L=%d, tile %d\n"
                "       classes round/wide/tall/square/large/fallback %d/%d/%d/%d/%d/%d; "
                "stored fields: polarity gamma v_A radius%s\n",
                path.c_str(), pre.version, pre.step, pre.cur_time, n, L, pitch, hist[0], hist[1],
                hist[2], hist[3], hist[4], hist[5], out->had_mobi ? " mobility" : "");
    std::printf("       maximum support (phi > %.0e): %d x %d px\n", (double)kSupportEps,
                max_ext[0], max_ext[1]);
    return true;
}

void resolve_per_cell_scalars(const SimParams& p, const SimOverrides& ov, CheckpointData* d) {
    const int n = d->n;
    if (ov.gamma_policy_changed()) {
        const int n_soft = (int)std::llround(p.cancer_fraction * (double)n);
        for (int i = 0; i < n; ++i) {
            pfh_cell_state& c = d->cells[(size_t)i];
            c.gamma = (float)(c.global_id < n_soft ? p.gamma_cancer : p.gamma_normal);
        }
        std::printf("[ckpt] per-cell gamma reassigned: %d of %d cells at gamma_cancer = %.6g, "
                    "the rest at %.6g (stored values overridden)\n",
                    n_soft, n, p.gamma_cancer, p.gamma_normal);
    }
    bool redo_vA = ov.v_A_policy_changed();
    if (!redo_vA && p.v_A > 0.0) {
        double sum = 0.0;
        for (int i = 0; i < n; ++i) sum += std::fabs((double)d->cells[(size_t)i].v_A);
        if (sum <= 1e-12) {
            std::printf("[ckpt] stored per-cell v_A values are zero; reassigning for "
                        "v_A = %.6g.\n", p.v_A);
            redo_vA = true;
        }
    }
    if (redo_vA) {
        for (int i = 0; i < n; ++i) {
            pfh_cell_state& c = d->cells[(size_t)i];
            c.v_A = (float)pfh2d_ic_v_A(c.global_id, p.seed, p.v_A, p.v_A_sigma);
        }
        std::printf("[ckpt] per-cell v_A reassigned: median %.6g, lognormal sigma %.6g\n", p.v_A,
                    p.v_A_sigma);
    }
    if (ov.target_radius)
        for (int i = 0; i < n; ++i) d->cells[(size_t)i].R_tgt = (float)p.target_radius;
    if (ov.phase_field_mobility)
        for (int i = 0; i < n; ++i) d->cells[(size_t)i].M_pf = (float)p.phase_field_mobility;
}

bool checkpoint_write(const CheckpointWriteView& v, const std::vector<std::string>& paths) {
    if (paths.empty()) return true;
    if (!v.p || !v.cell || !v.cls || !v.d_phi || v.N <= 0 || v.L <= 0) {
        std::fprintf(stderr, "[ckpt] incomplete write view\n");
        return false;
    }
    if (v.trajectory_samples <= 0 || v.trajectory_interval < 0 || v.save_interval < 0) {
        std::fprintf(stderr, "[ckpt] invalid output schedule metadata\n");
        return false;
    }
    if (v.p->num_cells != v.N || v.p->Nx != v.L || v.p->Ny != v.L || !validate(*v.p, true)) {
        std::fprintf(stderr, "[ckpt] write view and simulation parameters disagree\n");
        return false;
    }
    std::vector<float> g((size_t)v.N), m((size_t)v.N);
    for (int i = 0; i < v.N; ++i) {
        g[(size_t)i] = v.cell[i].gamma;
        m[(size_t)i] = v.cell[i].M_pf;
    }
    if (!validate_effective_stiffness(*v.p, g.data(), m.data(), v.N)) return false;
    if (v.step < 0 || v.step > 2147483647LL) {
        std::fprintf(stderr, "[ckpt] step %lld exceeds the checkpoint int32 field; checkpoint "
                             "not written.\n", v.step);
        return false;
    }
    FanOut w;
    if (!w.open(paths)) return false;
    FixedPrefix pre{};
    pre.magic = kMagic;
    pre.version = kFormat;
    pre.step = (int32_t)v.step;
    pre.cur_time = v.t;
    pre.num_cells_local = v.N;
    pre.save_interval = v.save_interval;
    pre.trajectory_samples = v.trajectory_samples;
    pre.sp_sz = (uint32_t)sizeof(ParamsRecord);
    const ParamsRecord sp =
        params_to_record(*v.p, v.trajectory_samples, v.trajectory_interval, v.save_interval);
    const RankTrailer tr{1, 0, v.N};
    if (!w.write(&pre, sizeof(pre), "fixed prefix") || !w.write(&sp, sizeof(sp), "parameters") ||
        !w.write(&kPitch, 4, "tile pitch") || !w.write(&tr, sizeof(tr), "rank trailer")) {
        w.discard();
        return false;
    }
    std::vector<float> stage((size_t)kStageCells * PFH_TILE_AREA);
    for (int base = 0; base < v.N; base += kStageCells) {
        const int cnt = std::min(kStageCells, v.N - base);
        const hipError_t e = hipMemcpy(stage.data(), v.d_phi + (size_t)base * PFH_TILE_AREA,
                                       (size_t)cnt * PFH_TILE_AREA * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            std::fprintf(stderr, "[ckpt] D2H of cells %d..%d failed: %s\n", base,
                         base + cnt - 1, hipGetErrorString(e));
            w.discard();
            return false;
        }
        for (int k = 0; k < cnt; ++k) {
            const int i = base + k;
            const pfh_cell_state& c = v.cell[i];
            const int cls = (int)c.cls;
            if (c.global_id != i) {
                std::fprintf(stderr, "[ckpt] cell slot %d carries global id %d; refusing a "
                                     "noncanonical checkpoint\n", i, c.global_id);
                w.discard();
                return false;
            }
            if (!std::isfinite(c.gamma) || !(c.gamma > 0.0f) || !std::isfinite(c.v_A) ||
                c.v_A < 0.0f || !std::isfinite(c.R_tgt) || !(c.R_tgt > 0.0f) ||
                !std::isfinite(c.theta) || !std::isfinite(c.vx) || !std::isfinite(c.vy) ||
                !std::isfinite(c.V) || !std::isfinite(c.Cx) || !std::isfinite(c.Cy) ||
                !std::isfinite(c.perim) || c.perim < 0.0 || !std::isfinite(c.phi_max) ||
                c.phi_max < 0.0f || !(c.V > 0.0) || c.gx0 < 0 || c.gx0 >= v.L || c.gy0 < 0 ||
                c.gy0 >= v.L || c.promote_ctr >= (uint32_t)kDemoteDwell) {
                std::fprintf(stderr, "[ckpt] cell %d has invalid state; checkpoint not written\n",
                             i);
                w.discard();
                return false;
            }
            if (cls < 0 || cls >= PFH_NUM_CLASSES || (int)v.cls[i] != cls) {
                std::fprintf(stderr, "[ckpt] cell %d carries shape class %d (cell_cls says %d); "
                                     "refusing to write.\n", i, cls, (int)v.cls[i]);
                w.discard();
                return false;
            }
            const Shape sc = kShapes[cls];
            if (c.cls_written[v.step & 1LL] != c.cls || c.bb_lo_x < 0 ||
                c.bb_hi_x < c.bb_lo_x || c.bb_hi_x >= sc.wx || c.bb_lo_y < 0 ||
                c.bb_hi_y < c.bb_lo_y || c.bb_hi_y >= sc.wy) {
                std::fprintf(stderr, "[ckpt] cell %d class layout or support bounds are invalid\n",
                             i);
                w.discard();
                return false;
            }
            CellRecord rec{};
            rec.cell_id = c.global_id;
            rec.origin_x = wrapi((long long)c.gx0 - sc.tx0, v.L);   // tile origin
            rec.origin_y = wrapi((long long)c.gy0 - sc.ty0, v.L);
            rec.shape_class = cls;
            rec.promote_ctr = c.promote_ctr;
            const double invV = 1.0 / c.V;
            rec.cx = (float)wrapd((double)c.gx0 + c.Cx * invV, (double)v.L);
            rec.cy = (float)wrapd((double)c.gy0 + c.Cy * invV, (double)v.L);
            rec.vx = c.vx;
            rec.vy = c.vy;
            rec.volume_moment = c.V;
            rec.moment_x = c.Cx;
            rec.moment_y = c.Cy;
            rec.perimeter = c.perim;
            rec.support_lo_x = c.bb_lo_x;
            rec.support_hi_x = c.bb_hi_x;
            rec.support_lo_y = c.bb_lo_y;
            rec.support_hi_y = c.bb_hi_y;
            rec.phi_max = c.phi_max;
            if (!w.write(&rec, sizeof(rec), "cell record") ||
                !w.write(stage.data() + (size_t)k * PFH_TILE_AREA, (size_t)PFH_TILE_AREA * 4,
                         "phi tile")) {
                w.discard();
                return false;
            }
        }
    }
    std::vector<float> col((size_t)v.N);
    auto emit = [&](uint32_t magic, float pfh_cell_state::*field) {
        for (int i = 0; i < v.N; ++i) col[(size_t)i] = v.cell[i].*field;
        const SidecarHeader h{magic, v.N};
        return w.write(&h, sizeof(h), "sidecar header") &&
               w.write(col.data(), col.size() * 4, "sidecar payload");
    };
    if (!emit(kPOLR, &pfh_cell_state::theta) || !emit(kGAMA, &pfh_cell_state::gamma) ||
        !emit(kVA_A, &pfh_cell_state::v_A) || !emit(kRADI, &pfh_cell_state::R_tgt) ||
        !emit(kMOBI, &pfh_cell_state::M_pf)) {
        w.discard();
        return false;
    }
    if (!w.commit()) return false;
    std::printf("[ckpt] saved step %lld  t %.4f  %d cells  tile pitch %d ->", v.step, v.t, v.N,
                (int)kPitch);
    for (const std::string& p : paths) std::printf(" %s", p.c_str());
    std::printf("\n");
    std::fflush(stdout);
    return true;
}

// ---------------------------------------------------------------------------
// Trajectory text output
// ---------------------------------------------------------------------------
namespace {
const char* kTrajTitle = "# Trajectory data";
const char* kTrajColumns = "# Format: time cell_id x y vx vy px py theta v_A_i L_n volume";

bool get_line(std::FILE* f, std::string* line, bool* too_long) {
    char buf[4096];
    if (!std::fgets(buf, sizeof(buf), f)) return false;
    size_t len = std::strlen(buf);
    if (len > 0 && buf[len - 1] == '\n') {
        buf[--len] = '\0';
        if (len > 0 && buf[len - 1] == '\r') buf[--len] = '\0';
    } else if (!std::feof(f)) {
        *too_long = true;
        return false;
    }
    line->assign(buf, len);
    return true;
}
}  // namespace

std::string trajectory_metadata(const SimParams& p, int side, long long interval) {
    char line[2048];
    const double rho_real = side > 0 ? (double)p.num_cells * p.area0() / ((double)side * side)
                                     : std::numeric_limits<double>::quiet_NaN();
    const int w = std::snprintf(
        line, sizeof(line),
        "# trajectory_schema=1 dim=2 model=run_tumble N=%d Lx=%d Ly=%d dx=%.17g dy=%.17g "
        "dt=%.17g rho_target=%.17g rho_realized=%.17g lambda=%.17g R=%.17g kappa=%.17g "
        "mu=%.17g xi=%.17g tau=%.17g v_A=%.17g v_A_sigma=%.17g gamma_normal=%.17g "
        "gamma_soft=%.17g soft_fraction=%.17g soft_assignment=lowest_global_ids seed=%llu "
        "polarity_seed=%llu initialization_hash=%016llx full_moment=%d perim_offset=%.17g "
        "trajectory_interval=%lld",
        p.num_cells, side, side, p.dx, p.dy, p.dt, p.rho, rho_real, p.lambda, p.target_radius,
        p.kappa, p.mu, p.xi, p.tau, p.v_A, p.v_A_sigma, p.gamma_normal, p.gamma_cancer,
        p.cancer_fraction, p.seed, p.polarity_stream(), p.initialization_hash,
        p.full_moment_every, kPi / p.interface_k(), interval);
    if (w < 0 || (size_t)w >= sizeof(line)) return {};
    return std::string(line, (size_t)w);
}

bool open_trajectory(const std::string& path, const SimParams& p, int side,
                     long long current_step, long long interval, long long* frames,
                     std::FILE** file) {
    *frames = 0;
    *file = nullptr;
    if (path.empty() || side <= 0 || current_step < 0 || interval <= 0 || p.num_cells <= 0)
        return false;
    const std::string meta = trajectory_metadata(p, side, interval);
    if (meta.empty()) {
        std::fprintf(stderr, "[error] 2D trajectory metadata is too long\n");
        return false;
    }
    const std::filesystem::path out(path);
    std::error_code ec;
    if (!out.parent_path().empty()) std::filesystem::create_directories(out.parent_path(), ec);
    if (ec) {
        std::fprintf(stderr, "[error] cannot create trajectory directory: %s\n",
                     ec.message().c_str());
        return false;
    }
    std::FILE* f = std::fopen(path.c_str(), "a+");
    if (!f) {
        std::fprintf(stderr, "[error] cannot open %s for append\n", path.c_str());
        return false;
    }
    auto fail = [&](const char* msg) {
        std::fprintf(stderr, "[error] %s\n", msg);
        std::fclose(f);
        return false;
    };
    std::rewind(f);
    const int first = std::fgetc(f);
    if (first == EOF && std::ferror(f)) return fail("cannot inspect the trajectory file");
    std::clearerr(f);
    if (first == EOF) {   // new file: the three header lines
        if (std::fseek(f, 0, SEEK_END) != 0) return fail("cannot seek the trajectory file");
        if (std::fprintf(f, "%s\n%s\n%s\n", kTrajTitle, kTrajColumns, meta.c_str()) < 0 ||
            std::fflush(f) != 0 || std::ferror(f))
            return fail("failed to write the trajectory header");
        *file = f;
        return true;
    }
    // Existing file: the header must match this run, the payload must hold
    // whole frames in canonical cell order with increasing times, and the
    // last frame must not lie beyond the initialised state.
    std::rewind(f);
    std::string line;
    bool too_long = false;
    for (const std::string& want : {std::string(kTrajTitle), std::string(kTrajColumns), meta})
        if (!get_line(f, &line, &too_long) || line != want)
            return fail("refusing to append: trajectory metadata does not match this run");
    unsigned long long rows = 0;
    double frame_t = 0.0, prev_t = -1.0;
    const unsigned long long N = (unsigned long long)p.num_cells;
    while (get_line(f, &line, &too_long)) {
        if (line.empty()) return fail("blank line in trajectory payload");
        double v[11];
        long long id = 0;
        char extra = 0;
        const int got = std::sscanf(line.c_str(), "%lf %lld %lf %lf %lf %lf %lf %lf %lf %lf %lf "
                                                  "%lf %c", &v[0], &id, &v[1], &v[2], &v[3],
                                    &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &extra);
        bool valid = got == 12 && id >= INT32_MIN && id <= INT32_MAX;
        for (double x : v) valid = valid && std::isfinite(x);
        if (!valid) {
            std::fprintf(stderr, "[error] malformed trajectory row %llu\n", rows + 1);
            std::fclose(f);
            return false;
        }
        const unsigned long long within = rows % N;
        if (within == 0) {
            frame_t = v[0];
            if (rows != 0 && !(frame_t > prev_t))
                return fail("trajectory frame times are not strictly increasing");
        } else if (v[0] != frame_t) {
            return fail("trajectory contains a frame with mixed times");
        }
        if (id != (long long)within) return fail("trajectory cells are not in canonical ID order");
        ++rows;
        if (rows % N == 0) prev_t = frame_t;
    }
    if (too_long || std::ferror(f) || rows % N != 0)
        return fail("refusing to append to a partial trajectory frame");
    *frames = (long long)(rows / N);
    if (*frames > 0) {
        const double sv = prev_t / p.dt;
        if (!std::isfinite(sv) || sv < 0.0 || sv > 9.0e18)
            return fail("trajectory time is outside step range");
        const long long last = std::llround(sv);
        if (std::fabs((double)last * p.dt - prev_t) > 1.0e-10 * std::max(1.0, std::fabs(prev_t)) ||
            last > current_step)
            return fail("trajectory extends beyond the initialized state");
    }
    std::clearerr(f);
    if (std::fseek(f, 0, SEEK_END) != 0) return fail("cannot seek the trajectory file");
    *file = f;
    return true;
}

// ---------------------------------------------------------------------------
// Accepted centre table (global_id,x,y CSV), include/palmieri_initializer.hpp:124-305
// ---------------------------------------------------------------------------
bool read_centres_csv(const std::string& path, int n, double side, double radius,
                      std::vector<float>* x, std::vector<float>* y, uint64_t* fnv,
                      std::string* error) {
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        *error = "cannot open centre CSV: " + path;
        return false;
    }
    std::string line;
    if (!std::getline(in, line) || line != "global_id,x,y") {
        *error = "CSV header must be exactly global_id,x,y with LF line endings";
        return false;
    }
    x->clear();
    y->clear();
    for (int want = 0; want < n; ++want) {
        if (!std::getline(in, line)) {
            *error = "centre CSV ended before expected N rows";
            return false;
        }
        const size_t c1 = line.find(','), c2 = c1 == std::string::npos ? c1 : line.find(',', c1 + 1);
        if (c1 == std::string::npos || c2 == std::string::npos ||
            line.find(',', c2 + 1) != std::string::npos) {
            *error = "centre CSV row must contain exactly three fields";
            return false;
        }
        int id = -1;
        float xv = 0.0f, yv = 0.0f;
        const char* s = line.data();
        const auto a = std::from_chars(s, s + c1, id);
        const auto b = std::from_chars(s + c1 + 1, s + c2, xv);
        const auto c = std::from_chars(s + c2 + 1, s + line.size(), yv);
        if (a.ec != std::errc{} || a.ptr != s + c1 || b.ec != std::errc{} || b.ptr != s + c2 ||
            c.ec != std::errc{} || c.ptr != s + line.size()) {
            *error = "centre CSV contains a malformed field";
            return false;
        }
        if (id != want) {
            *error = "global_id must be the ordered sequence 0..N-1";
            return false;
        }
        x->push_back(xv);
        y->push_back(yv);
    }
    if (std::getline(in, line)) {
        *error = "centre CSV contains more than expected N rows";
        return false;
    }
    const float mid = (float)(0.5 * side);
    if ((*x)[0] != mid || (*y)[0] != mid) {
        *error = "global_id 0 is not exactly at the box centre";
        return false;
    }
    auto min_image = [side](double d) { return d - side * std::floor(d / side + 0.5); };
    for (int i = 0; i < n; ++i) {
        const float xi = (*x)[(size_t)i], yi = (*y)[(size_t)i];
        if (!std::isfinite(xi) || !std::isfinite(yi) || !(xi >= 0.0f && xi < side) ||
            !(yi >= 0.0f && yi < side)) {
            *error = "centre is non-finite or outside [0,L)";
            return false;
        }
        for (int j = 0; j < i; ++j) {
            const double dx = min_image((double)xi - (*x)[(size_t)j]);
            const double dy = min_image((double)yi - (*y)[(size_t)j]);
            if (dx * dx + dy * dy < radius * radius) {
                *error = "centre table violates the periodic minimum-separation rule";
                return false;
            }
        }
    }
    *fnv = pfh_centre_table_fnv1a64(x->data(), y->data(), n);
    return true;
}

bool distinct_output_paths(const std::vector<std::string>& paths) {
    std::vector<std::filesystem::path> seen;
    for (const std::string& p : paths) {
        if (p.empty()) continue;
        std::error_code ec;
        auto abs = std::filesystem::absolute(p, ec);
        auto can = ec ? abs : std::filesystem::weakly_canonical(abs, ec);
        if (ec) {
            std::fprintf(stderr, "[fatal] cannot resolve output path %s: %s\n", p.c_str(),
                         ec.message().c_str());
            return false;
        }
        for (const auto& q : seen)
            if (can == q) {
                std::fprintf(stderr, "[fatal] output files must use distinct paths\n");
                return false;
            }
        seen.push_back(can);
    }
    return true;
}

}  // namespace pf2d
