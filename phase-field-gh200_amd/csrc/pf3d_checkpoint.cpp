// PF3D schema-1 checkpoint files of the 3D periodic path: write and read
// with the CRC-64/ECMA-182 over the whole file.
//
// Layout (common/checkpoint_format_3d.h:140-391): a 96 B file header, the
// 288 B parameter record, then per cell a 256 B record and its B^3 binary32
// cube (x fastest). Little-endian, packed; every offset the reference pins is
// static-asserted below, because files must be interchangeable with the
// reference's own (src/pf3d/checkpoint.cu:116-420, 819-1235). The aggregate
// field is not stored: a reader rebuilds it from the cubes.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "pfh.h"
#include "pfh_host.h"

using pfh_host::hip_status;
using pfh_host::invalid;

namespace {

constexpr uint32_t kMagic = 0x44334650u;      // "PF3D"
constexpr uint16_t kFormat = 1;
constexpr uint32_t kEndian = 0x01020304u;
constexpr uint32_t kFlags = 3u;               // payload CRC-64 | variable cell storage
constexpr uint32_t kPeriodicXYZ = 7u;
constexpr uint32_t kIntegratorEuler = 1, kStencil27 = 2, kRngPhilox = 1;
constexpr uint64_t kPolicyAuto = UINT64_MAX;

#pragma pack(push, 1)
struct Header {
    uint32_t magic;
    uint16_t version, header_bytes;
    uint32_t endian_marker, dimensions, scalar_format, flags;
    int64_t step;
    double time;
    uint64_t num_cells;
    uint32_t params_bytes, cell_record_bytes, brick_edge, phase_order;
    uint64_t phase_values_per_cell, phase_bytes_per_cell, file_crc64, base_measure_shards;
};
struct Record {
    int64_t global_id, origin_x, origin_y, origin_z;
    float polarity_x, polarity_y, polarity_z, velocity_x, velocity_y, velocity_z;
    float gamma, v_A, target_radius, phi_max;
    double volume, moment_x, moment_y, moment_z, surface, ix, iy, iz;
    uint32_t shift_counter, tumble_counter, state_flags;
    int32_t lo_x, hi_x, lo_y, hi_y, lo_z, hi_z;
    int32_t pend_x, pend_y, pend_z;
    uint32_t storage_edge, reserved32;
    uint64_t reserved64[8];
};
#pragma pack(pop)
static_assert(sizeof(Header) == 96 && offsetof(Header, step) == 24 &&
              offsetof(Header, num_cells) == 40 && offsetof(Header, brick_edge) == 56 &&
              offsetof(Header, phase_values_per_cell) == 64 &&
              offsetof(Header, file_crc64) == 80 && offsetof(Header, base_measure_shards) == 88,
              "FileHeader3D layout");
static_assert(sizeof(pfh3d_ckpt_params) == 288 && offsetof(pfh3d_ckpt_params, dx) == 24 &&
              offsetof(pfh3d_ckpt_params, brick_edge) == 48 &&
              offsetof(pfh3d_ckpt_params, dt) == 56 && offsetof(pfh3d_ckpt_params, seed) == 168 &&
              offsetof(pfh3d_ckpt_params, polarity_seed) == 176 &&
              offsetof(pfh3d_ckpt_params, integrator_id) == 208 &&
              offsetof(pfh3d_ckpt_params, trajectory_interval) == 220 &&
              offsetof(pfh3d_ckpt_params, initialization_hash) == 224 &&
              offsetof(pfh3d_ckpt_params, aging_time) == 232 &&
              offsetof(pfh3d_ckpt_params, promoted_measure_policy) == 240 &&
              offsetof(pfh3d_ckpt_params, promoted_measure_auto_wave_ctas) == 248 &&
              offsetof(pfh3d_ckpt_params, channel_height) == 256 &&
              offsetof(pfh3d_ckpt_params, wall_kappa) == 272,
              "ParamsRecord3D layout");
static_assert(sizeof(Record) == 256 && offsetof(Record, origin_z) == 24 &&
              offsetof(Record, polarity_x) == 32 && offsetof(Record, velocity_x) == 44 &&
              offsetof(Record, volume) == 72 && offsetof(Record, shift_counter) == 136 &&
              offsetof(Record, state_flags) == 144 && offsetof(Record, lo_x) == 148 &&
              offsetof(Record, pend_x) == 172 && offsetof(Record, storage_edge) == 184 &&
              offsetof(Record, reserved64) == 192,
              "CellRecord3D layout");

// CRC-64/ECMA-182: polynomial 0x42F0E1EBA9EA3693, MSB first, init 0, no
// reflection, no final xor; the table is the polynomial's byte remainders.
struct Crc64 {
    uint64_t t[256];
    Crc64() {
        for (int i = 0; i < 256; ++i) {
            uint64_t v = (uint64_t)i << 56;
            for (int b = 0; b < 8; ++b)
                v = (v >> 63) ? (v << 1) ^ 0x42F0E1EBA9EA3693ull : (v << 1);
            t[i] = v;
        }
    }
    uint64_t update(uint64_t crc, const void* p, size_t n) const {
        const uint8_t* s = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) crc = t[(uint8_t)((crc >> 56) ^ s[i])] ^ (crc << 8);
        return crc;
    }
};
const Crc64& crc64() {
    static const Crc64 c;
    return c;
}

char g_msg[320];
int fail(const char* path, const char* why) {
    std::snprintf(g_msg, sizeof(g_msg), "[ckpt3d] %s: %s", path, why);
    return invalid(g_msg);
}

Record record_of(const pfh3d_cell_state& c, int B) {
    Record r;
    std::memset(&r, 0, sizeof(r));
    r.global_id = c.global_id;
    r.origin_x = c.origin_x; r.origin_y = c.origin_y; r.origin_z = c.origin_z;
    r.polarity_x = c.polarity_x; r.polarity_y = c.polarity_y; r.polarity_z = c.polarity_z;
    r.velocity_x = c.velocity_x; r.velocity_y = c.velocity_y; r.velocity_z = c.velocity_z;
    r.gamma = c.gamma; r.v_A = c.v_A; r.target_radius = c.R_tgt; r.phi_max = c.phi_max;
    r.volume = c.V; r.moment_x = c.Cx; r.moment_y = c.Cy; r.moment_z = c.Cz;
    r.surface = c.surface; r.ix = c.Ix; r.iy = c.Iy; r.iz = c.Iz;
    r.shift_counter = c.shift_ctr; r.tumble_counter = c.tumble_ctr; r.state_flags = c.flags;
    r.lo_x = c.bb_lo_x; r.hi_x = c.bb_hi_x; r.lo_y = c.bb_lo_y;
    r.hi_y = c.bb_hi_y; r.lo_z = c.bb_lo_z; r.hi_z = c.bb_hi_z;
    r.pend_x = c.pending_shift_x; r.pend_y = c.pending_shift_y; r.pend_z = c.pending_shift_z;
    r.storage_edge = (uint32_t)(c.storage_edge ? c.storage_edge : (uint32_t)B);
    return r;
}

pfh3d_cell_state cell_of(const Record& r, int B) {
    pfh3d_cell_state c;
    std::memset(&c, 0, sizeof(c));
    c.global_id = r.global_id;
    c.origin_x = r.origin_x; c.origin_y = r.origin_y; c.origin_z = r.origin_z;
    c.polarity_x = r.polarity_x; c.polarity_y = r.polarity_y; c.polarity_z = r.polarity_z;
    c.velocity_x = r.velocity_x; c.velocity_y = r.velocity_y; c.velocity_z = r.velocity_z;
    c.gamma = r.gamma; c.v_A = r.v_A; c.R_tgt = r.target_radius; c.phi_max = r.phi_max;
    c.V = r.volume; c.Cx = r.moment_x; c.Cy = r.moment_y; c.Cz = r.moment_z;
    c.surface = r.surface; c.Ix = r.ix; c.Iy = r.iy; c.Iz = r.iz;
    c.shift_ctr = r.shift_counter; c.tumble_ctr = r.tumble_counter; c.flags = r.state_flags;
    c.bb_lo_x = r.lo_x; c.bb_hi_x = r.hi_x; c.bb_lo_y = r.lo_y;
    c.bb_hi_y = r.hi_y; c.bb_lo_z = r.lo_z; c.bb_hi_z = r.hi_z;
    c.pending_shift_x = r.pend_x; c.pending_shift_y = r.pend_y; c.pending_shift_z = r.pend_z;
    c.storage_edge = r.storage_edge == (uint32_t)B ? 0u : r.storage_edge;
    return c;
}

// A record this build can continue: a base-brick cell (adaptive promotion
// is outside this build), finite measurements, support inside the cube,
// no fatal state flag, unit polarity, zero reserved words.
bool record_ok(const Record& r, int B) {
    bool zero = r.reserved32 == 0;
    for (uint64_t v : r.reserved64) zero = zero && v == 0;
    const double n2 = (double)r.polarity_x * r.polarity_x + (double)r.polarity_y * r.polarity_y +
                      (double)r.polarity_z * r.polarity_z;
    return zero && r.storage_edge == (uint32_t)B && std::isfinite(r.volume) && r.volume > 0.0 &&
           std::isfinite(r.moment_x) && std::isfinite(r.moment_y) &&
           std::isfinite(r.moment_z) && std::isfinite(r.surface) && r.surface >= 0.0 &&
           std::isfinite(r.ix) && std::isfinite(r.iy) && std::isfinite(r.iz) &&
           std::isfinite(r.velocity_x) && std::isfinite(r.velocity_y) &&
           std::isfinite(r.velocity_z) && std::fabs(n2 - 1.0) <= 2.0e-5 &&
           std::isfinite(r.gamma) && r.gamma > 0.0f && std::isfinite(r.v_A) && r.v_A >= 0.0f &&
           std::isfinite(r.target_radius) && r.target_radius > 0.0f &&
           std::isfinite(r.phi_max) && r.phi_max >= 0.0f &&
           (r.state_flags & ~((1u << PFH3D_FLAG_COUNT) - 1u)) == 0 &&
           (r.state_flags & PFH3D_FATAL_MASK) == 0 && r.lo_x >= 0 && r.lo_x <= r.hi_x &&
           r.hi_x < B && r.lo_y >= 0 && r.lo_y <= r.hi_y && r.hi_y < B && r.lo_z >= 0 &&
           r.lo_z <= r.hi_z && r.hi_z < B && r.pend_x >= -B && r.pend_x <= B &&
           r.pend_y >= -B && r.pend_y <= B && r.pend_z >= -B && r.pend_z <= B;
}

const char* header_problem(const Header& h) {
    if (h.magic != kMagic) return "not a PF3D checkpoint (dimension/magic mismatch)";
    if (h.version != kFormat) return "unsupported PF3D checkpoint format";
    if (h.header_bytes != sizeof(Header) || h.params_bytes != sizeof(pfh3d_ckpt_params) ||
        h.cell_record_bytes != sizeof(Record))
        return "record-size fields do not match the PF3D contract";
    if (h.endian_marker != kEndian || h.dimensions != 3 || h.scalar_format != 1 ||
        h.phase_order != 1)
        return "incompatible scalar, dimension, endian, or phase order";
    if (h.flags != kFlags) return "missing required checksum flag";
    if (h.base_measure_shards > 64) return "base measurement shard count is outside 0..64";
    if (h.step < 0 || !std::isfinite(h.time) || h.time < 0.0)
        return "negative step/time or non-finite time";
    if (h.num_cells == 0 || h.num_cells > 2147483647ull)
        return "cell count is outside the supported range";
    if (h.brick_edge < 8 || h.brick_edge > 4096 || h.brick_edge % 8)
        return "brick edge is outside the corruption guard";
    const uint64_t v = (uint64_t)h.brick_edge * h.brick_edge * h.brick_edge;
    if (h.phase_values_per_cell != v || h.phase_bytes_per_cell != 4 * v)
        return "phase payload size is inconsistent with B^3 float32";
    return nullptr;
}

const char* params_problem(const pfh3d_ckpt_params& p, const Header& h) {
    if (p.Nx <= 0 || p.Nx > 2147483647 || p.Ny != p.Nx || p.Nz != p.Nx ||
        p.boundary_flags != kPeriodicXYZ)
        return "only the periodic-XYZ cube is supported by this build";
    const double s[] = {p.dx, p.dy, p.dz, p.dt, p.t_end, p.lambda, p.gamma_normal, p.gamma_soft,
                        p.soft_fraction, p.kappa, p.target_radius, p.mu, p.v_A, p.v_A_sigma,
                        p.xi, p.tau, p.volume_fraction, p.aging_time};
    for (double v : s)
        if (!std::isfinite(v)) return "checkpoint contains a non-finite parameter";
    if (p.brick_edge != h.brick_edge || (int64_t)p.brick_edge >= p.Nx)
        return "parameter and header brick edges disagree";
    if (p.integrator_id != kIntegratorEuler || p.stencil_id != kStencil27 ||
        p.rng_id != kRngPhilox)
        return "integrator, stencil, or RNG contract is incompatible";
    if (p.print_interval < 0 || p.full_moment_every < 0 || p.verify_every < 0 ||
        p.print_interval > 2147483647 || p.full_moment_every > 2147483647 ||
        p.verify_every > 2147483647)
        return "stored output or verification cadence is out of range";
    const bool contract = p.promoted_measure_policy == kPolicyAuto
        ? (p.promoted_measure_auto_wave_ctas > 0 &&
           p.promoted_measure_auto_wave_ctas <= 2147483647ull)
        : (p.promoted_measure_policy <= 64 && p.promoted_measure_auto_wave_ctas == 0);
    if (!contract) return "invalid promoted-measurement reduction contract";
    if (p.channel_height < 0 || p.channel_padding < 0 || !std::isfinite(p.wall_kappa) ||
        !std::isfinite(p.wall_width))
        return "invalid resolved-wall checkpoint parameters";
    return nullptr;
}

}  // namespace

extern "C" {

uint64_t pfh_crc64_ecma(uint64_t crc, const void* data, size_t bytes) {
    if (!data && bytes) return crc;
    return crc64().update(crc, data, bytes);
}

int pfh3d_checkpoint_write(const char* path, int64_t step, double time,
                           const pfh3d_ckpt_params* params, const pfh3d_cell_state* cells,
                           int N, int B, const float* phi, uint64_t base_measure_shards,
                           void* stream) {
    if (!path || !params || !cells || !phi || N <= 0 || B < 8 || B % 8 || step < 0 ||
        base_measure_shards > 64)
        return invalid("checkpoint_write");
    Header h;
    std::memset(&h, 0, sizeof(h));
    h.magic = kMagic;
    h.version = kFormat;
    h.header_bytes = sizeof(Header);
    h.endian_marker = kEndian;
    h.dimensions = 3;
    h.scalar_format = 1;
    h.flags = kFlags;
    h.step = step;
    h.time = time;
    h.num_cells = (uint64_t)N;
    h.params_bytes = sizeof(pfh3d_ckpt_params);
    h.cell_record_bytes = sizeof(Record);
    h.brick_edge = (uint32_t)B;
    h.phase_order = 1;
    h.phase_values_per_cell = (uint64_t)B * B * B;
    h.phase_bytes_per_cell = 4 * h.phase_values_per_cell;
    h.base_measure_shards = base_measure_shards;
    const char* bad = header_problem(h);
    if (!bad) bad = params_problem(*params, h);
    if (bad) return fail(path, bad);
    const std::string tmp = std::string(path) + ".tmp";
    std::FILE* f = std::fopen(tmp.c_str(), "wb+");
    if (!f) return fail(path, std::strerror(errno));
    const size_t vox = (size_t)B * B * B;
    std::vector<float> cube(vox);
    uint64_t crc = crc64().update(0, &h, sizeof(h));
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
              std::fwrite(params, sizeof(*params), 1, f) == 1;
    crc = crc64().update(crc, params, sizeof(*params));
    for (int n = 0; ok && n < N; ++n) {
        const Record r = record_of(cells[n], B);
        if (cells[n].storage_edge != 0) {
            std::fclose(f);
            std::remove(tmp.c_str());
            return fail(path, "promoted cells (adaptive storage) are outside this build");
        }
        const int st = hip_status(hipMemcpyAsync(cube.data(), phi + (size_t)n * vox, vox * 4,
                                                 hipMemcpyDeviceToHost, (hipStream_t)stream),
                                  "checkpoint_write: cube D2H");
        if (st != PFH_OK || hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
            std::fclose(f);
            std::remove(tmp.c_str());
            return st != PFH_OK ? st : hip_status(hipErrorUnknown, "checkpoint_write: sync");
        }
        crc = crc64().update(crc, &r, sizeof(r));
        crc = crc64().update(crc, cube.data(), vox * 4);
        ok = std::fwrite(&r, sizeof(r), 1, f) == 1 && std::fwrite(cube.data(), 4, vox, f) == vox;
    }
    // the checksum covers the file with its own field zeroed
    h.file_crc64 = crc;
    ok = ok && std::fseek(f, 0, SEEK_SET) == 0 && std::fwrite(&h, sizeof(h), 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        std::remove(tmp.c_str());
        return fail(path, "short write");
    }
    std::error_code ec;
    std::filesystem::rename(tmp, path, ec);
    if (ec) return fail(path, "rename of the temporary file failed");
    return PFH_OK;
}

int pfh3d_checkpoint_read(const char* path, pfh3d_ckpt_params* params_out, int64_t* step,
                          double* time, uint64_t* base_measure_shards, int* N, int* B,
                          pfh3d_cell_state* cells, float* phi, void* stream) {
    if (!path || !params_out || !step || !time || !base_measure_shards || !N || !B)
        return invalid("checkpoint_read");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(path, std::strerror(errno));
    struct Closer { std::FILE* f; ~Closer() { std::fclose(f); } } closer{f};
    Header h;
    pfh3d_ckpt_params p;
    if (std::fread(&h, sizeof(h), 1, f) != 1) return fail(path, "short read of the header");
    const char* bad = header_problem(h);
    if (bad) return fail(path, bad);
    if (std::fread(&p, sizeof(p), 1, f) != 1) return fail(path, "short read of the parameters");
    bad = params_problem(p, h);
    if (bad) return fail(path, bad);
    const int n = (int)h.num_cells, b = (int)h.brick_edge;
    if (b > PFH3D_MAX_EDGE) return fail(path, "bricks above 64^3 are outside this build");
    const size_t vox = (size_t)b * b * b;
    std::error_code ec;
    const uintmax_t size = std::filesystem::file_size(path, ec);
    if (ec || size != sizeof(Header) + sizeof(p) + (uintmax_t)n * (sizeof(Record) + 4 * vox))
        return fail(path, "file size does not match the base-brick payload");
    *params_out = p;
    *step = h.step;
    *time = h.time;
    *base_measure_shards = h.base_measure_shards;
    *N = n;
    *B = b;
    if (!cells && !phi) return PFH_OK;   // probe
    if (!cells || !phi) return invalid("checkpoint_read: cells and phi go together");
    Header canon = h;
    canon.file_crc64 = 0;
    uint64_t crc = crc64().update(0, &canon, sizeof(canon));
    crc = crc64().update(crc, &p, sizeof(p));
    std::vector<float> cube(vox);
    std::vector<pfh3d_cell_state> staged((size_t)n);
    for (int i = 0; i < n; ++i) {
        Record r;
        if (std::fread(&r, sizeof(r), 1, f) != 1 || std::fread(cube.data(), 4, vox, f) != vox)
            return fail(path, "short read of a cell");
        if (!record_ok(r, b) || r.global_id != i) {
            std::snprintf(g_msg, sizeof(g_msg), "[ckpt3d] %s: invalid cell record %d", path, i);
            return invalid(g_msg);
        }
        crc = crc64().update(crc, &r, sizeof(r));
        crc = crc64().update(crc, cube.data(), vox * 4);
        staged[(size_t)i] = cell_of(r, b);
        int st = hip_status(hipMemcpyAsync(phi + (size_t)i * vox, cube.data(), vox * 4,
                                           hipMemcpyHostToDevice, (hipStream_t)stream),
                            "checkpoint_read: cube H2D");
        if (st == PFH_OK)
            st = hip_status(hipStreamSynchronize((hipStream_t)stream), "checkpoint_read: sync");
        if (st != PFH_OK) return st;
    }
    if (crc != h.file_crc64) return fail(path, "file CRC-64/ECMA mismatch");
    std::memcpy(cells, staged.data(), staged.size() * sizeof(pfh3d_cell_state));
    return PFH_OK;
}

}  // extern "C"
