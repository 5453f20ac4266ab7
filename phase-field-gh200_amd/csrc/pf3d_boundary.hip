// The 3D launch API in the reference's own argument records (InitArgs3D,
// ScatterArgs3D, MeasureArgs3D, UpdateArgs3D; include/pf3d/kernels.cuh:
// 153-316), so a host that builds those records binds without repacking,
// and the 3D verification kernels (launch_verify_cells / launch_verify_S,
// kernels.cuh:478-489; src/pf3d/kernels.cu:2410-2527).
//
// The records are translated into this library's step arguments and run by
// the same kernels as pfh3d_step. This build covers the periodic cube with
// cubic base bricks: a record that asks for bounded-z walls, compact
// storage, a cell selection or a partial tile pass is rejected with
// PFH_ERR_UNSUPPORTED.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include "../../include/pfh.h"
#include "pfh_device.h"
#include "pfh_host.h"

using pfh_host::as_stream;
using pfh_host::invalid;
using pfh_host::launch_status;
using pfh_host::unsupported;

namespace pfh3v {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Per cell (one 256-thread workgroup): V = sum phi^2 in fp64, max |phi|,
// non-finite voxels, and voxels above the support threshold on a face of the
// cube (support touching the brick edge), read as 16 B vectors.
__global__ __launch_bounds__(256) void k3_verify_cells(const float* phi,
                                                       const pfh3d_cell_state* cells,
                                                       pfh3d_verify_cell* out, int B) {
    __shared__ double vred[4];
    __shared__ uint32_t mred[4], bred[4], ered[4];
    const int n = (int)blockIdx.x;
    const size_t vox = (size_t)B * B * B;
    const f32x4* cube = reinterpret_cast<const f32x4*>(phi + (size_t)n * vox);
    double v = 0.0;
    uint32_t mbits = 0u, bad = 0u, edge = 0u;
    for (size_t g = threadIdx.x; g < vox / 4; g += 256) {
        const f32x4 q = cube[g];
        const size_t i0 = 4 * g;
        const int z = (int)(i0 / ((size_t)B * B));
        const int y = (int)((i0 / (size_t)B) % (size_t)B);
        const int x0 = (int)(i0 % (size_t)B);
        const bool face_yz = y == 0 || y == B - 1 || z == 0 || z == B - 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float f = q[k];
            if (!isfinite(f)) {
                ++bad;
                continue;
            }
            v += (double)f * (double)f;
            const float a = fabsf(f);
            mbits = max(mbits, __float_as_uint(a));
            const int x = x0 + k;
            if (a > pfh::kSupportEps && (face_yz || x == 0 || x == B - 1)) ++edge;
        }
    }
    v = pfh::wave_sum(v);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        mbits = max(mbits, (uint32_t)__shfl_xor((int)mbits, d, 64));
        bad += (uint32_t)__shfl_xor((int)bad, d, 64);
        edge += (uint32_t)__shfl_xor((int)edge, d, 64);
    }
    const int w = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        vred[w] = v;
        mred[w] = mbits;
        bred[w] = bad;
        ered[w] = edge;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        pfh3d_verify_cell r;
        r.measured_V = ((vred[0] + vred[1]) + vred[2]) + vred[3];
        r.state_V = cells[n].V;
        r.max_abs_phi = __uint_as_float(max(max(mred[0], mred[1]), max(mred[2], mred[3])));
        r.nonfinite_count = bred[0] + bred[1] + bred[2] + bred[3];
        r.support_edge_count = ered[0] + ered[1] + ered[2] + ered[3];
        r.verified = 1u;
        out[n] = r;
    }
}

// Largest aggregate word: 16 B loads, a wave maximum, one atomic per wave.
__global__ __launch_bounds__(256) void k3_verify_S(const uint32_t* S, size_t n4,
                                                   uint32_t* out_max) {
    uint32_t m = 0u;
    const u32x4* s4 = reinterpret_cast<const u32x4*>(S);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4;
         i += (size_t)gridDim.x * 256) {
        const u32x4 q = s4[i];
        m = max(m, max(max(q.x, q.y), max(q.z, q.w)));
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d, 64));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out_max, m);
}

}  // namespace pfh3v

namespace {

int check_layout(const pfh3d_layout& l, int B) {
    if (l.boundary_flags != 7u || l.nx != l.ny || l.nx != l.nz)
        return unsupported("only the periodic cube (boundary_flags 7, nx = ny = nz) is "
                           "supported by this build");
    if (l.nx <= B || l.pitch_x < l.nx || l.pitch_x % 4 != 0)
        return invalid("layout: side must exceed the brick edge; pitch >= side, multiple of 4");
    return PFH_OK;
}

int check_common(const pfh3d_layout& l, int B, const pfh3d_storage& st,
                 const pfh3d_selection& sel) {
    if (st.z_cap != 0) return unsupported("compact phase storage is outside this build");
    if (sel.ids != nullptr) return unsupported("cell selections are outside this build");
    return check_layout(l, B);
}

}  // namespace

extern "C" {

int pfh3d_launch_initialize_spheres(const pfh3d_init_args* a, void* stream) {
    if (!a) return invalid("init args");
    if (a->storage.z_cap != 0) return unsupported("compact phase storage is outside this build");
    const int st = check_layout(a->layout, a->B);
    if (st != PFH_OK) return st;
    return pfh3d_init_spheres(a->phi_even, a->phi_odd, a->cells,
                              reinterpret_cast<const double*>(a->centres), a->N, a->B,
                              a->layout.nx, a->lambda, a->seed_radius, stream);
}

int pfh3d_launch_scatter_current(const pfh3d_scatter_args* a, void* stream) {
    if (!a) return invalid("scatter args");
    const int st = check_common(a->layout, a->B, a->storage, a->selection);
    if (st != PFH_OK) return st;
    return pfh3d_scatter_current(a->phi, a->cells, a->S, a->global_flags, a->N, a->B,
                                 a->layout.nx, a->layout.pitch_x, stream);
}

int pfh3d_launch_measure(const pfh3d_measure_args* a, uint32_t* arrivals, int apply_motion,
                         float dt, void* stream) {
    if (!a) return invalid("measure args");
    int st = check_common(a->layout, a->B, a->storage, a->selection);
    if (st != PFH_OK) return st;
    if (a->wall.values != nullptr || a->wall.planes != 0)
        return unsupported("wall fields belong to bounded-z geometries");
    if (a->shards > 1 && !arrivals) return invalid("arrivals: N zeroed words for shards > 1");
    pfh3d_step_args s;
    memset(&s, 0, sizeof(s));
    s.phi_in = a->phi;
    s.S_in = a->S;
    s.cells = a->cells;
    s.flags = a->global_flags;
    // the measurement reads the accepted-step counter only (tumble, aging)
    s.counters = const_cast<uint64_t*>(a->step);
    s.partials = a->partials;
    s.arrivals = arrivals;
    s.N = a->N;
    s.B = a->B;
    s.L = a->layout.nx;
    s.P = a->layout.pitch_x;
    s.shards = a->shards;
    s.max_shift = a->max_shift;
    s.support_margin = a->support_margin;
    s.dt = dt;
    s.motility_coeff = a->motility_coeff;
    s.p_tumble = a->p_tumble;
    s.polarity_stream = a->polarity_stream;
    s.aging_steps = a->aging_steps;
    return pfh3d_measure_cells(&s, a->apply_tumble ? 1 : 0, a->compute_surface ? 1 : 0,
                               apply_motion ? 1 : 0, stream);
}

int pfh3d_launch_update(const pfh3d_update_args* a, int slabs, void* stream) {
    if (!a) return invalid("update args");
    int st = check_common(a->layout, a->B, a->storage, a->selection);
    if (st != PFH_OK) return st;
    if (a->wall.values != nullptr || a->wall.planes != 0)
        return unsupported("wall fields belong to bounded-z geometries");
    if (a->compute_surface)
        return unsupported("the update measures nothing; measure the new field instead");
    if (a->tile_pass != 0 || a->exchange_region.y_count != 0)
        return unsupported("partial tile passes belong to the two-GPU runner");
    if (!a->S_out) return unsupported("an update without a scatter target");
    pfh3d_step_args s;
    memset(&s, 0, sizeof(s));
    s.phi_in = a->phi_in;
    s.phi_out = a->phi_out;
    s.S_in = a->S_in;
    s.S_out = a->S_out;
    s.cells = a->cells;
    s.flags = a->global_flags;
    s.N = a->N;
    s.B = a->B;
    s.L = a->layout.nx;
    s.P = a->layout.pitch_x;
    s.shards = 1;
    s.dt = a->dt;
    s.V0 = a->V0;
    s.volume_scale = a->volume_scale;
    s.bulk_scale = a->bulk_scale;
    s.interaction_coeff = a->interaction_coeff;
    return pfh3d_update(&s, slabs, stream);
}

int pfh3d_verify_cells(const float* phi, const pfh3d_cell_state* cells, pfh3d_verify_cell* out,
                       int N, int B, const pfh3d_layout* layout, void* stream) {
    if (!phi || !cells || !out || !layout || N <= 0 || B < 8 || B % 8 != 0)
        return invalid("verify_cells");
    const int st = check_layout(*layout, B);
    if (st != PFH_OK) return st;
    if ((uintptr_t)phi % 16) return invalid("verify_cells: cubes must be 16 B aligned");
    hipLaunchKernelGGL(pfh3v::k3_verify_cells, dim3((unsigned)N), dim3(256), 0,
                       as_stream(stream), phi, cells, out, B);
    return launch_status("k3_verify_cells");
}

int pfh3d_verify_S(const uint32_t* S, const pfh3d_layout* layout, uint32_t* out_max,
                   void* stream) {
    if (!S || !layout || !out_max || (uintptr_t)S % 16) return invalid("verify_S");
    const int st = check_layout(*layout, 8);
    if (st != PFH_OK) return st;
    const size_t n4 = (size_t)layout->pitch_x * layout->ny * layout->nz / 4;
    hipLaunchKernelGGL(pfh3v::k3_verify_S, dim3(1024), dim3(256), 0, as_stream(stream), S, n4,
                       out_max);
    return launch_status("k3_verify_S");
}

}  // extern "C"
