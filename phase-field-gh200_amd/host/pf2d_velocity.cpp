// pf2d host driver: velocity-component recording. Schedules the observed
// step variants (advection on every step, spatial moments on the sampling
// cadence), reads the accumulators back at output frames and writes the
// PFVMOM4 file.
//
// Reference counterparts: include/velocity_sampling.hpp (cadence),
// src/velocity_output.cu (VelocityOutput), common/velocity_moments.h
// (finish_snapshot and the discrete corrections), common/velocity_format.h
// (PFVMOM4 header and columns).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

#include "pf2d_driver.hpp"

namespace pf2d {

namespace {

constexpr uint32_t kColumns = 26;
constexpr uint32_t kDiscreteLinearQuadrature = 2;
constexpr int kIntegralColumn = 9;
constexpr int kSpatialCountColumn = 21;
constexpr int kEndpointColumn = 22;
constexpr int kInterfacial = 4;       // first interfacial column of q
constexpr int kIntegralValues = 12;
constexpr int kSpatialValues = 4;

#pragma pack(push, 1)
struct FileHeader {
    char magic[8];             // "PFVMOM4\0"
    uint32_t cells;
    uint32_t columns;
    double dt;
    int64_t start_step;
    double side;
    uint32_t spatial_stride;
    uint32_t quadrature;
    uint64_t dense_start;
};
#pragma pack(pop)
static_assert(sizeof(FileHeader) == 56, "PFVMOM4 header is 56 bytes");
static_assert(offsetof(FileHeader, dense_start) == 48, "PFVMOM4 header layout");

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "[velocity] %s: %s\n", what, hipGetErrorString(e));
    return false;
}

// A+I (x, y), then the stored-velocity rounding (x, y), integrated so far.
void advection_totals(const pfh_advection_accum& a, double (&t)[4]) {
    t[0] = a.q[0] + a.q[2];
    t[1] = a.q[1] + a.q[3];
    t[2] = a.q[4];
    t[3] = a.q[5];
}

// The advection correction of the interval since the last spatial sample:
// (beta - 1) x delta(A + I) + beta x delta(rounding), beta held.
void add_advection_correction(pfh_moment_accum& m, const double (&t)[4]) {
    for (int xy = 0; xy < 2; ++xy) {
        const double beta = m.held[kSpatialValues + xy];
        m.q[10 + xy] += (beta - 1.0) * (t[xy] - m.last_totals[xy]) +
                        beta * (t[2 + xy] - m.last_totals[2 + xy]);
    }
}

double linear_correction(double dt, unsigned long long steps, double first, double last) {
    return steps ? dt * 0.5 * (double)(steps - 1ull) * (last - first) : 0.0;
}

// Host readback only: completes a copy of the accumulators up to the
// current step, leaving the device state intact. endpoint holds dt times the
// rates measured now by the dense observer.
void finish_snapshot(pfh_moment_accum& m, const pfh_advection_accum& a,
                     const pfh_moment_accum& endpoint, double dt) {
    m.count = a.count;
    for (int j = 0; j < kInterfacial; ++j) m.q[j] = a.q[j];
    if (!m.spatial_count) return;
    const unsigned long long tail = a.count - m.last_sample;
    for (int j = 0; j < kSpatialValues; ++j) {
        m.q[kInterfacial + j] += m.correction[j] + dt * (double)tail * m.held[j];
        m.q[kInterfacial + j] +=
            linear_correction(dt, tail, m.held[j], endpoint.q[kInterfacial + j] / dt);
    }
    double t[4];
    advection_totals(a, t);
    add_advection_correction(m, t);
}

}  // namespace

void VelocitySampling::configure(bool enabled, bool reference, int stride) {
    enabled_ = enabled;
    reference_ = reference;
    stride_ = stride;
    period_ = enabled && !reference ? std::lcm(kGraphBody, stride) : kGraphBody;
    // a long period is cut at the spatial samples, so frequent output does
    // not break a long replay; the pieces together hold the same launches
    graph_steps_ = enabled && !reference && stride >= kGraphBody ? stride : period_;
}

VelocityOutput::~VelocityOutput() {
    (void)close();
    if (moments_) (void)hipFree(moments_);
    if (endpoint_) (void)hipFree(endpoint_);
}

bool VelocityOutput::configure(const VelocityOptions& options, int cells, double dt) {
    if (enabled()) return false;
    options_ = options;
    dt_ = dt;
    if (options.path.empty()) return true;
    if (cells <= 0 || !std::isfinite(dt) || dt <= 0.0 || options.spatial_stride < 1 ||
        options.spatial_stride > 1000 || options.dense_start < 0) {
        std::fprintf(stderr, "[velocity] invalid geometry or sampling cadence\n");
        return false;
    }
    sampling_.configure(true, options.reference, options.spatial_stride);
    cells_.resize((size_t)cells);
    integrals_.resize((size_t)cells);
    endpoints_.resize((size_t)cells);
    rows_.resize((size_t)cells * kColumns);
    const size_t bytes = (size_t)cells * sizeof(pfh_moment_accum);
    if (!hip_ok(hipMalloc(&moments_, bytes), "hipMalloc(moments)") ||
        !hip_ok(hipMalloc(&endpoint_, bytes), "hipMalloc(endpoint)") ||
        !hip_ok(hipMemset(moments_, 0, bytes), "hipMemset(moments)"))
        return false;
    return true;
}

bool VelocityOutput::write(const void* data, size_t bytes) {
    if (std::fwrite(data, 1, bytes, file_) == bytes) return true;
    std::perror("[velocity] write");
    failed_ = true;
    return false;
}

bool VelocityOutput::open(const pfh2d_step_args& state, long long step, void* stream) {
    if (!enabled()) return true;
    if (file_ || failed_ || step < 0 ||
        options_.dense_start > std::numeric_limits<long long>::max() - step)
        return false;
    // a new file only: an existing path is never overwritten
    file_ = std::fopen(options_.path.c_str(), "wbx");
    if (!file_) {
        std::perror("[velocity] cannot create new output file");
        failed_ = true;
        return false;
    }
    start_step_ = step;
    const long long dense = reference() ? 0 : options_.dense_start;
    sampling_.start_run(step, dense);
    FileHeader h{};
    std::memcpy(h.magic, "PFVMOM4", 8);
    h.cells = (uint32_t)cells_.size();
    h.columns = kColumns;
    h.dt = dt_;
    h.start_step = step;
    h.side = state.L;
    h.spatial_stride = reference() ? 1u : (uint32_t)options_.spatial_stride;
    h.quadrature = kDiscreteLinearQuadrature;
    h.dense_start = (uint64_t)dense;
    return write(&h, sizeof(h)) && capture(state, step, stream);
}

bool VelocityOutput::readback(const pfh2d_step_args& state, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = integrals_.size() * sizeof(pfh_moment_accum);
    if (!hip_ok(hipMemsetAsync(endpoint_, 0, bytes, s), "hipMemsetAsync(endpoint)"))
        return false;
    if (pfh2d_velocity_moments(&state, endpoint_, dt_, stream) != PFH_OK) {
        std::fprintf(stderr, "[velocity] dense observer: %s\n", pfh_last_error());
        return false;
    }
    return hip_ok(hipMemcpyAsync(cells_.data(), state.cell,
                                 cells_.size() * sizeof(pfh_cell_state), hipMemcpyDeviceToHost,
                                 s), "copy cells") &&
           hip_ok(hipMemcpyAsync(integrals_.data(), moments_, bytes, hipMemcpyDeviceToHost, s),
                  "copy moments") &&
           hip_ok(hipMemcpyAsync(endpoints_.data(), endpoint_, bytes, hipMemcpyDeviceToHost, s),
                  "copy endpoint") &&
           hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
}

bool VelocityOutput::pack_frame(long long step, int side) {
    for (size_t i = 0; i < cells_.size(); ++i) {
        const pfh_cell_state& c = cells_[i];
        pfh_moment_accum& m = integrals_[i];
        if (!reference()) finish_snapshot(m, c.advection, endpoints_[i], dt_);
        if (m.count != (unsigned long long)(step - start_step_)) {
            std::fprintf(stderr, "[velocity] missed observation at step %lld, cell %d\n", step,
                         c.global_id);
            return false;
        }
        double* row = rows_.data() + i * kColumns;
        row[0] = c.global_id;
        row[1] = std::fmod((double)c.gx0 + c.Cx / c.V, (double)side);
        row[2] = std::fmod((double)c.gy0 + c.Cy / c.V, (double)side);
        row[3] = c.theta;
        row[4] = c.gamma;
        row[5] = c.M_pf;
        row[6] = c.v_A;
        row[7] = c.V;
        row[8] = (double)m.count;
        for (int j = 0; j < kIntegralValues; ++j) row[kIntegralColumn + j] = m.q[j];
        row[kSpatialCountColumn] = (double)(reference() ? m.count : m.spatial_count);
        for (int j = 0; j < kSpatialValues; ++j)
            row[kEndpointColumn + j] = endpoints_[i].q[kInterfacial + j] / dt_;
        for (uint32_t j = 0; j < kColumns; ++j) {
            if (!std::isfinite(row[j])) {
                std::fprintf(stderr, "[velocity] non-finite column %u at step %lld, cell %d\n", j,
                             step, c.global_id);
                return false;
            }
        }
    }
    return true;
}

bool VelocityOutput::capture(const pfh2d_step_args& state, long long step, void* stream) {
    if (!enabled()) return true;
    if (!file_ || failed_ || step <= last_frame_ || state.N != (int)cells_.size()) return false;
    if (!readback(state, stream) || !pack_frame(step, state.L)) {
        failed_ = true;
        return false;
    }
    const int64_t disk_step = step;
    if (!write(&disk_step, sizeof(disk_step)) ||
        !write(rows_.data(), rows_.size() * sizeof(double)))
        return false;
    if (std::fflush(file_) != 0 || std::ferror(file_)) {
        std::perror("[velocity] flush");
        failed_ = true;
        return false;
    }
    last_frame_ = step;
    return true;
}

bool VelocityOutput::close() {
    if (!file_) return !failed_;
    bool ok = std::fflush(file_) == 0 && !std::ferror(file_);
    if (std::fclose(file_) != 0) ok = false;
    file_ = nullptr;
    if (!ok) {
        std::perror("[velocity] close");
        failed_ = true;
    }
    return !failed_;
}

}  // namespace pf2d
