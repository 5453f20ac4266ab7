// Learned surrogate of the 2D step on gfx950: per-cell crops, a 4-layer 3x3
// CNN (4 -> 32 -> 32 -> 32 -> 1 channels, residual output) in BF16 with FP32
// accumulation, its training kernels, and the accept/fallback kernels.
//
// The reference has no learned component; the crop builds on its tile layout
// (include/params.cuh:121-147) and CellState (include/kernels.cuh:27-57), and
// the accepted fields are re-measured with the plain path's own kernels
// (k_scatter_all / k_init_moments semantics, src/kernels.cu:1098-1190).
//
// Layout: activations NHWC bf16, one 96x96 crop per cell, channels fastest
// (8 for the input: phi, S_other, cos theta, sin theta, 4 zero channels; 32
// for hidden layers). Weights [out][tap][in] bf16 for the MFMA layers, tap =
// 3 ky + kx.
//
// k_sur_conv: implicit GEMM on MFMA (v_mfma_f32_32x32x16_bf16 for 32 input
// channels, v_mfma_f32_32x32x8_bf16 for 8): M = 32 pixels of a row, N = 32
// output channels, K = 9 taps x input channels. Persistent workgroups keep
// the weights in LDS and walk (crop, 4-row block) items; each item stages a
// 6 x 98 pixel halo tile in LDS, each wave owns one output row (three
// 32-pixel tiles sharing every B fragment). The same kernel computes the
// backward-data pass with tap-flipped, transposed weights and a ReLU mask.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>

#include "../../include/pfh.h"
#include "pfh_device.h"
#include "pfh_host.h"

namespace sur {

typedef float fx4 __attribute__((ext_vector_type(4)));
constexpr int kCrop = 96;
constexpr int kCropPix = kCrop * kCrop;
constexpr int kC0 = 8;     // input channels (4 used)
constexpr int kCH = 32;    // hidden channels
constexpr int kRowsPerItem = 4;
constexpr int kItemsPerCrop = kCrop / kRowsPerItem;   // 24
constexpr int kHaloCols = kCrop + 2;                  // 98

typedef uint16_t bf16_t;
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// binary32 -> bfloat16, round to nearest even (NaN stays NaN)
__host__ __device__ inline bf16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (bf16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Two binary32 values -> packed bfloat16 (low half = a) with gfx950's
// conversion instruction, round to nearest even (the same values as f2bf).
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// ---------------------------------------------------------------------------
// crops
// ---------------------------------------------------------------------------
// Crop origin: the cell's rounded centroid minus 48, wrapped into [0, L).
__device__ inline void crop_origin(const pfh_cell_state& c, int L, int* ox, int* oy) {
    double cx = (double)c.gx0, cy = (double)c.gy0;
    if (c.V > 0.0) {
        cx += c.Cx / c.V;
        cy += c.Cy / c.V;
    }
    *ox = pfh::wrapi((int)rint(cx) - kCrop / 2, L);
    *oy = pfh::wrapi((int)rint(cy) - kCrop / 2, L);
}

// phi of cell n at world pixel (wx, wy): the tile's active window, else 0
__device__ inline float cell_phi(const float* tile, const pfh_cell_state& c, int cls, int L,
                                 int wx, int wy) {
    const pfh::Shape s = pfh::shape_of(cls);
    int a = wx - c.gx0;
    if (a < 0) a += L;
    int b = wy - c.gy0;
    if (b < 0) b += L;
    if (a >= s.wx || b >= s.wy) return 0.0f;
    return tile[(size_t)(s.ty0 + b) * pfh::kTilePitch + s.tx0 + a];
}

// One workgroup per cell: network input (bf16 NHWC, 8 channels), the exact
// phi crop (fp32) and the crop origin.
__global__ __launch_bounds__(256) void k_sur_crop(const float* phi, const pfh_cell_state* cells,
                                                  const uint8_t* cls, const uint32_t* S,
                                                  int L, int P, bf16_t* X, float* phic,
                                                  int* origin) {
    const int n = (int)blockIdx.x;
    const pfh_cell_state c = cells[n];
    int ox, oy;
    crop_origin(c, L, &ox, &oy);
    if (threadIdx.x == 0) {
        origin[2 * n] = ox;
        origin[2 * n + 1] = oy;
    }
    double sn, cs;
    pfh::sincos_d((double)c.theta, &sn, &cs);
    const bf16_t bc = f2bf((float)cs), bs = f2bf((float)sn);
    const float* tile = phi + (size_t)n * pfh::kTileArea;
    const int k = (int)cls[n];
    for (int p = (int)threadIdx.x; p < kCropPix; p += (int)blockDim.x) {
        const int i = p % kCrop, j = p / kCrop;
        int wx = ox + i;
        if (wx >= L) wx -= L;
        int wy = oy + j;
        if (wy >= L) wy -= L;
        const float v = cell_phi(tile, c, k, L, wx, wy);
        const uint32_t q = pfh::q_of(v);
        const uint32_t s = S[(size_t)wy * P + wx];
        const float other = s >= q ? (float)(s - q) * pfh::kQInv : 0.0f;
        bf16_t* o = X + ((size_t)n * kCropPix + p) * kC0;
        u32x4 w;
        w[0] = (uint32_t)f2bf(v) | ((uint32_t)f2bf(other) << 16);
        w[1] = (uint32_t)bc | ((uint32_t)bs << 16);
        w[2] = 0u;
        w[3] = 0u;
        *reinterpret_cast<u32x4*>(o) = w;
        phic[(size_t)n * kCropPix + p] = v;
    }
}

// The same window of the field after the surrogate interval: phi crop at the
// recorded origins (the cell's window may have moved meanwhile).
__global__ __launch_bounds__(256) void k_sur_recrop(const float* phi, const pfh_cell_state* cells,
                                                    const uint8_t* cls, const int* origin, int L,
                                                    float* out) {
    const int n = (int)blockIdx.x;
    const pfh_cell_state c = cells[n];
    const int ox = origin[2 * n], oy = origin[2 * n + 1];
    const float* tile = phi + (size_t)n * pfh::kTileArea;
    const int k = (int)cls[n];
    for (int p = (int)threadIdx.x; p < kCropPix; p += (int)blockDim.x) {
        const int i = p % kCrop, j = p / kCrop;
        int wx = ox + i;
        if (wx >= L) wx -= L;
        int wy = oy + j;
        if (wy >= L) wy -= L;
        out[(size_t)n * kCropPix + p] = cell_phi(tile, c, k, L, wx, wy);
    }
}

// ---------------------------------------------------------------------------
// MFMA convolution
// ---------------------------------------------------------------------------
template <int CIN>
struct ConvShape {
    static constexpr int K = 9 * CIN;
    static constexpr int KS = CIN == 8 ? 8 : 16;      // k per MFMA
    static constexpr int NKS = K / KS;
    static constexpr int PIX = CIN + 8;               // LDS pixel pitch (bf16)
    static constexpr int WROW = K + 8;                // LDS weight row pitch (bf16)
    static constexpr int HALO = (kRowsPerItem + 2) * kHaloCols;
};

template <int CIN>
__device__ __forceinline__ f32x16 mfma_step(const bf16_t* a, const bf16_t* b, f32x16 acc) {
    if constexpr (CIN == 8) {
        const s16x4 av = *reinterpret_cast<const s16x4*>(a);
        const s16x4 bv = *reinterpret_cast<const s16x4*>(b);
        return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    } else {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(a);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(b);
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
}

// Y = act(conv3x3(X, W) + bias) [* (mask > 0)], zero padding at the crop edge.
// X: [ncrops][96][96][CIN] bf16; W: [32][9*CIN] bf16; Y: [ncrops][96][96][32].
template <int CIN>
__global__ __launch_bounds__(256) void k_sur_conv(const bf16_t* X, const bf16_t* W,
                                                  const float* bias, const bf16_t* mask,
                                                  bf16_t* Y, int ncrops, int relu) {
    using SH = ConvShape<CIN>;
    __shared__ __align__(16) bf16_t wl[kCH * SH::WROW];
    __shared__ __align__(16) bf16_t xl[SH::HALO * SH::PIX];
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, col = lane & 31;

    // weights -> LDS once (16 B chunks)
    for (int i = tid; i < kCH * (SH::K / 8); i += 256) {
        const int co = i / (SH::K / 8), k8 = i - co * (SH::K / 8);
        *reinterpret_cast<u32x4*>(wl + co * SH::WROW + 8 * k8) =
            *reinterpret_cast<const u32x4*>(W + (size_t)co * SH::K + 8 * k8);
    }
    float bco = bias ? bias[col] : 0.0f;

    const int items = ncrops * kItemsPerCrop;
    for (int item = (int)blockIdx.x; item < items; item += (int)gridDim.x) {
        const int crop = item / kItemsPerCrop;
        const int y0 = (item - crop * kItemsPerCrop) * kRowsPerItem;
        __syncthreads();
        // stage rows y0-1 .. y0+4, columns -1 .. 96 (8-channel chunks)
        constexpr int CH8 = CIN / 8;
        for (int i = tid; i < SH::HALO * CH8; i += 256) {
            const int px = i / CH8, c8 = i - px * CH8;
            const int r = px / kHaloCols, cc = px - r * kHaloCols;
            const int y = y0 - 1 + r, x = cc - 1;
            u32x4 v = {0u, 0u, 0u, 0u};
            if ((unsigned)y < (unsigned)kCrop && (unsigned)x < (unsigned)kCrop)
                v = *reinterpret_cast<const u32x4*>(
                    X + (((size_t)crop * kCrop + y) * kCrop + x) * CIN + 8 * c8);
            *reinterpret_cast<u32x4*>(xl + px * SH::PIX + 8 * c8) = v;
        }
        __syncthreads();

        f32x16 acc0 = {}, acc1 = {}, acc2 = {};
#pragma unroll 3
        for (int s = 0; s < SH::NKS; ++s) {
            const int k0 = s * SH::KS;
            const int t = k0 / CIN, c0 = k0 - t * CIN;
            const int ky = t / 3, kx = t - 3 * ky;
            const bf16_t* bp = wl + col * SH::WROW + k0 + (SH::KS / 2) * half;
            const bf16_t* ap = xl + ((wave + ky) * kHaloCols + col + kx) * SH::PIX + c0 +
                               (SH::KS / 2) * half;
            acc0 = mfma_step<CIN>(ap, bp, acc0);
            acc1 = mfma_step<CIN>(ap + 32 * SH::PIX, bp, acc1);
            acc2 = mfma_step<CIN>(ap + 64 * SH::PIX, bp, acc2);
        }

        // epilogue: lane holds output channel `col`, pixel rows 8j + 4 half + i
        const int y = y0 + wave;
        const size_t rowbase = ((size_t)crop * kCrop + y) * kCrop;
        for (int tile = 0; tile < 3; ++tile) {
            const f32x16 a = tile == 0 ? acc0 : tile == 1 ? acc1 : acc2;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int px = 32 * tile + 8 * (r >> 2) + 4 * half + (r & 3);
                const size_t o = (rowbase + px) * kCH + col;
                float v = a[r] + bco;
                if (mask && !(bf2f(mask[o]) > 0.0f)) v = 0.0f;
                if (relu) v = fmaxf(v, 0.0f);
                Y[o] = f2bf(v);
            }
        }
    }
}

// The conversion k_sur_infer applies to hidden activations, exported so that
// tests can pin its rounding against round-to-nearest-even.
__global__ void k_sur_round_bf16(const float* x, uint32_t* y, int npairs) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < npairs) y[i] = pack_bf16(x[2 * i], x[2 * i + 1]);
}

// ---------------------------------------------------------------------------
// output layer (32 -> 1), loss and their gradients (VALU, fp32)
// ---------------------------------------------------------------------------
// dphi = b4 + sum_{tap, ci} W4[tap][ci] h3[p + off(tap)][ci]; out = phic + dphi.
// Persistent over (crop, 4-row block) items like k_sur_conv: the 6 x 98
// halo tile of h3 is staged in LDS, each lane then forms 288-term dot
// products in fp32 (taps in order, channels in order).
__global__ __launch_bounds__(256) void k_sur_out(const bf16_t* H, const float* W4, float b4,
                                                 const float* phic, float* dphi, float* out,
                                                 int ncrops) {
    constexpr int PIXP = kCH + 8;
    __shared__ float w[9 * kCH];
    __shared__ __align__(16) bf16_t hl[(kRowsPerItem + 2) * kHaloCols * PIXP];
    const int tid = (int)threadIdx.x;
    for (int i = tid; i < 9 * kCH; i += 256) w[i] = W4[i];
    const int items = ncrops * kItemsPerCrop;
    for (int item = (int)blockIdx.x; item < items; item += (int)gridDim.x) {
        const int crop = item / kItemsPerCrop;
        const int y0 = (item - crop * kItemsPerCrop) * kRowsPerItem;
        __syncthreads();
        for (int i = tid; i < (kRowsPerItem + 2) * kHaloCols * 4; i += 256) {
            const int px = i >> 2, c8 = i & 3;
            const int r = px / kHaloCols, cc = px - r * kHaloCols;
            const int y = y0 - 1 + r, x = cc - 1;
            u32x4 v = {0u, 0u, 0u, 0u};
            if ((unsigned)y < (unsigned)kCrop && (unsigned)x < (unsigned)kCrop)
                v = *reinterpret_cast<const u32x4*>(
                    H + (((size_t)crop * kCrop + y) * kCrop + x) * kCH + 8 * c8);
            *reinterpret_cast<u32x4*>(hl + px * PIXP + 8 * c8) = v;
        }
        __syncthreads();
        for (int q = tid; q < kRowsPerItem * kCrop; q += 256) {
            const int r = q / kCrop, x = q - r * kCrop;
            float acc = 0.0f;
            for (int t = 0; t < 9; ++t) {
                const int ky = t / 3, kx = t - 3 * ky;
                const bf16_t* h = hl + ((r + ky) * kHaloCols + x + kx) * PIXP;
                for (int c8 = 0; c8 < 4; ++c8) {
                    const u32x4 v = *reinterpret_cast<const u32x4*>(h + 8 * c8);
                    for (int k = 0; k < 4; ++k) {
                        acc += w[t * kCH + 8 * c8 + 2 * k] * bf2f((bf16_t)(v[k] & 0xFFFFu));
                        acc += w[t * kCH + 8 * c8 + 2 * k + 1] * bf2f((bf16_t)(v[k] >> 16));
                    }
                }
            }
            const size_t g = (size_t)crop * kCropPix + (size_t)(y0 + r) * kCrop + x;
            const float d = acc + b4;
            if (dphi) dphi[g] = d;
            if (out) out[g] = phic[g] + d;
        }
    }
}

// ---------------------------------------------------------------------------
// fused inference: all four layers of one crop in one workgroup
// ---------------------------------------------------------------------------
// The layer-by-layer kernels stream every 32-channel activation map through
// HBM (96 x 96 x 32 bf16 = 590 KB per crop and layer). k_sur_infer keeps them
// on chip: one workgroup walks a crop top to bottom two rows per step, each
// layer trailing the one before by three rows, and holds the last six rows of
// X, h1, h2 and of the output layer's per-tap partials in LDS rings. Every
// stage of a step reads only rows finished in earlier steps and writes a ring
// slot nobody reads in that step, so one barrier per step suffices.
//
//   step s: X rows   [2s+3, 2s+5)   written to the ring (fetched in step s-1)
//           h1 rows  [2s,   2s+2)   conv1, waves 0-5
//           h2 rows  [2s-3, 2s-1)   conv2, waves 0-5
//           h3 rows  [2s-6, 2s-4)   conv3 + output partials, waves 6-11
//           out rows [2s-9, 2s-7)   sum of partials, waves 0-2
//
// MFMA operands: A = weights (row m holds output channel perm(m), so that the
// accumulator of lane (pixel, half h) holds channels 16h .. 16h+15 in order),
// B = activations (lane = pixel). Each wave keeps its layer's A fragments in
// registers for the whole kernel. The output layer (32 -> 1, 3x3) is an MFMA
// on conv3's accumulators: P[t][p] = sum_c W4[t][c] h3[p][c] for the nine
// taps t, with W4 split into bf16 hi + lo parts (16 significant bits), and
// dphi[y][x] = b4 + sum_t P[t][(y, x) + off(t)] in tap order.
// Rows outside the crop are written as zeros (the zero padding of every
// layer's input); ring columns 0 and 97 stay zero.
constexpr int kIT = 768;                       // 12 waves
constexpr int kRing = 6;                       // rows per ring
constexpr int kHP = kCH + 8;                   // h ring pixel pitch (bf16), 80 B
constexpr int kRowH = kHaloCols * kHP;
constexpr int kXP = 4;                         // X ring: 4 channels per pixel
constexpr int kRowX = kHaloCols * kXP;
constexpr int kRowP = 9 * kHaloCols;           // P ring row: [tap][column] fp32
constexpr int kFirstStep = -3, kLastStep = 52;

// Where a crop's input comes from when the kernel builds it itself
struct CropCtx {
    int ox, oy;           // crop origin (world pixel of crop (0, 0))
    int gx0, gy0;         // window origin
    int wx, wy, tx0, ty0; // window shape and its place in the tile
    uint32_t cs;          // bf16 cos theta | bf16 sin theta << 16
};

struct InferLDS {
    bf16_t h1[kRing * kRowH];
    bf16_t h2[kRing * kRowH];
    bf16_t x[kRing * kRowX];
    float p[kRing * kRowP];
    alignas(16) float bias[3 * kCH];
    CropCtx ctx;
};

struct InferArgs {
    // inputs as buffers (pfh_sur_conv layout) ...
    const bf16_t* X;
    const float* phic;
    // ... or built from the field (k_sur_crop's definition)
    const float* phi;
    const pfh_cell_state* cells;
    const uint8_t* cls;
    const uint32_t* S;
    int L, P;
    int* origin;
    // network
    const bf16_t *W1, *W2, *W3;
    const float *b1, *b2, *b3, *W4;
    float b4;
    float* dphi;
    float* pred;
    int ncrops;
    const int* go;        // NULL, or nothing is done unless *go != 0
};

// phi of the crop at crop pixel (x, y): the cell's active window, else 0
__device__ __forceinline__ float ctx_phi(const float* tile, const CropCtx& c, int L, int x, int y,
                                         int* wxo, int* wyo) {
    int wx = c.ox + x;
    if (wx >= L) wx -= L;
    int wy = c.oy + y;
    if (wy >= L) wy -= L;
    *wxo = wx;
    *wyo = wy;
    int a = wx - c.gx0;
    if (a < 0) a += L;
    int b = wy - c.gy0;
    if (b < 0) b += L;
    return (a < c.wx && b < c.wy) ? tile[(size_t)(c.ty0 + b) * pfh::kTilePitch + c.tx0 + a]
                                  : 0.0f;
}

__device__ __forceinline__ int ring(int row) { return (row + 16 * kRing) % kRing; }
__device__ __forceinline__ bool in_crop(int row) { return (unsigned)row < (unsigned)kCrop; }
// MFMA row m -> output channel, so that accumulator r of lane half h is
// channel 16 h + r
__device__ __forceinline__ int perm_row(int m) {
    return 16 * ((m >> 2) & 1) + 4 * (m >> 3) + (m & 3);
}

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// A lane's 16 biases (channels 16 h .. 16 h + 15) as the starting value of
// its accumulator: the convolution sums its products onto the bias
__device__ __forceinline__ f32x16 bias_acc(const float* bias) {
    f32x16 a;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const fx4 v = *reinterpret_cast<const fx4*>(bias + 4 * k);
        a[4 * k] = v[0];
        a[4 * k + 1] = v[1];
        a[4 * k + 2] = v[2];
        a[4 * k + 3] = v[3];
    }
    return a;
}

// ReLU + bf16 of a lane's 16 channels, packed as two 8-channel chunks. The
// ReLU is a signed 16-bit max with 0 on the packed values: a negative bf16 is
// a negative int16, and rounding never turns a negative value positive, so
// this is the bf16 of max(x, 0) (NaN inputs aside) at one instruction per
// two channels.
__device__ __forceinline__ void relu_pack(const f32x16& acc, u32x4* lo, u32x4* hi) {
    uint32_t w[8];
    const s16x2 zero = {0, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const s16x2 v = __builtin_bit_cast(s16x2, pack_bf16(acc[2 * k], acc[2 * k + 1]));
        w[k] = __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, zero));
    }
    *lo = u32x4{w[0], w[1], w[2], w[3]};
    *hi = u32x4{w[4], w[5], w[6], w[7]};
}

// 32 -> 32 convolution of one 32-pixel tile of row y from ring `src`, summed
// onto `acc`. The activation fragments are read kPre k-steps ahead of their
// MFMA, so each MFMA finds its operand in registers instead of waiting out
// an LDS round trip. The scheduling groups pin that order (mask 0x100: LDS
// reads, 0x008: MFMA); the first group also holds the caller's four reads
// of the bias into `acc`.
constexpr int kPre = 4;
__device__ __forceinline__ f32x16 conv32_tile(const bf16_t* src, const bf16x8* wa, int y,
                                              int tile, int col, int half, f32x16 acc) {
    const bf16_t* base = src + (32 * tile + col) * kHP + 8 * half;
    const int r0 = ring(y - 1) * kRowH, r1 = ring(y) * kRowH, r2 = ring(y + 1) * kRowH;
    auto frag = [&](int s) {
        const int t = s >> 1, ky = t / 3, kx = t - 3 * ky;
        const int ro = ky == 0 ? r0 : ky == 1 ? r1 : r2;
        return *reinterpret_cast<const bf16x8*>(base + ro + kx * kHP + 16 * (s & 1));
    };
    bf16x8 b[kPre];
#pragma unroll
    for (int s = 0; s < kPre; ++s) b[s] = frag(s);
    __builtin_amdgcn_sched_group_barrier(0x100, kPre + 4, 0);
#pragma unroll
    for (int s = 0; s < 18; ++s) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[s], b[s % kPre], acc, 0, 0, 0);
        if (s + kPre < 18) b[s % kPre] = frag(s + kPre);
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (s + kPre < 18) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    return acc;
}

// 4 -> 32 convolution (k = 4 tap + channel, 36 of 48 used) of one tile. In
// k-step s, lane half 0 holds taps 4s, 4s+1 and half 1 taps 4s+2, 4s+3 (4
// channels each); the taps are compile-time constants per half, the three
// ring rows are wave-uniform.
__device__ __forceinline__ f32x16 conv4_tile(const bf16_t* xr, const bf16x8* wa, int y, int tile,
                                             int col, int half, f32x16 acc) {
    const int rb0 = ring(y - 1) * kRowX, rb1 = ring(y) * kRowX, rb2 = ring(y + 1) * kRowX;
    const bf16_t* base = xr + (32 * tile + col) * kXP;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        uint32_t v[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int t0 = 4 * s + j, t1 = 4 * s + 2 + j;
            const int ky0 = t0 / 3, kx0 = t0 % 3, ky1 = t1 / 3, kx1 = t1 % 3;
            const int o0 = (ky0 == 0 ? rb0 : ky0 == 1 ? rb1 : rb2) + kx0 * kXP;
            const int o1 = (ky1 == 0 ? rb0 : ky1 == 1 ? rb1 : rb2) + kx1 * kXP;
            const bool ok = half ? (t1 < 9) : (t0 < 9);
            if (ok) {
                const u32x2 q = *reinterpret_cast<const u32x2*>(base + (half ? o1 : o0));
                v[2 * j] = q[0];
                v[2 * j + 1] = q[1];
            }
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[s], as_bf16x8(u32x4{v[0], v[1], v[2], v[3]}),
                                                      acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ void store_h(bf16_t* dst, int y, int tile, int col, int half, u32x4 lo,
                                        u32x4 hi) {
    bf16_t* o = dst + ring(y) * kRowH + (32 * tile + col + 1) * kHP + 16 * half;
    *reinterpret_cast<u32x4*>(o) = lo;
    *reinterpret_cast<u32x4*>(o + 8) = hi;
}

template <bool kCells>
__global__ __launch_bounds__(kIT, 1) void k_sur_infer(const InferArgs A) {
    __shared__ InferLDS sm;
    if (A.go && *A.go == 0) return;
    const bf16_t *W1 = A.W1, *W2 = A.W2, *W3 = A.W3;
    const float *b1 = A.b1, *b2 = A.b2, *b3 = A.b3, *W4 = A.W4;
    const float b4 = A.b4;
    float* dphi = A.dphi;
    float* pred = A.pred;
    const int ncrops = A.ncrops, L = A.L, P = A.P;
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63;
    // wave-uniform: keeps the step's row and ring arithmetic on the scalar unit
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int half = lane >> 5, col = lane & 31;
    const int m = perm_row(col);

    // zero every ring once: the halo columns are never written again
    {
        uint32_t* z = reinterpret_cast<uint32_t*>(&sm);
        for (int i = tid; i < (int)(sizeof(InferLDS) / 4); i += kIT) z[i] = 0u;
    }
    __syncthreads();
    for (int i = tid; i < kCH; i += kIT) {
        sm.bias[i] = b1[i];
        sm.bias[kCH + i] = b2[i];
        sm.bias[2 * kCH + i] = b3[i];
    }

    // register-resident A fragments
    const bool early = wave < 6;           // waves 0-5: conv1 + conv2; 6-11: conv3 + out
    bf16x8 wa[18];
    {
        const bf16_t* wr = (early ? W2 : W3) + (size_t)m * (9 * kCH);
#pragma unroll
        for (int s = 0; s < 18; ++s)
            wa[s] = *reinterpret_cast<const bf16x8*>(wr + 16 * s + 8 * half);
    }
    // wx: conv1 fragments (waves 0-5: k-steps 0-2) or the output layer's two
    // (waves 6-11: channels 0-15 and 16-31, each with the weights' bf16 high
    // parts in rows 0-8 and low parts in rows 16-24); one array keeps both in
    // one set of registers
    bf16x8 wx[3];
    if (early) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            uint32_t v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t pr = 0u;
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int kk = 16 * s + 8 * half + 2 * k + e;
                    const int t = kk >> 2, ci = kk & 3;
                    const uint32_t b = t < 9 ? (uint32_t)W1[(size_t)m * (9 * kC0) + t * kC0 + ci]
                                             : 0u;
                    pr |= b << (16 * e);
                }
                v[k] = pr;
            }
            wx[s] = as_bf16x8(u32x4{v[0], v[1], v[2], v[3]});
        }
    } else {
        // Output layer, channels 16 j .. 16 j + 15 per fragment: rows 0-8
        // hold the bf16 high parts of the nine taps' weights, rows 16-24 the
        // low parts (w - hi, rounded to bf16), so one MFMA yields both
        // partial sums of a tap (accumulators r and r + 8 of a lane).
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            uint32_t v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t pr = 0u;
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int c = 16 * half + 8 * j + 2 * k + e;
                    const int t = col < 16 ? col : col - 16;
                    const float w = t < 9 ? W4[t * kCH + c] : 0.0f;
                    const bf16_t hb = f2bf(w);
                    const bf16_t part = col < 16 ? hb : f2bf(w - bf2f(hb));
                    pr |= (uint32_t)part << (16 * e);
                }
                v[k] = pr;
            }
            wx[j] = as_bf16x8(u32x4{v[0], v[1], v[2], v[3]});
        }
    }
    // Retire the fragment loads here: otherwise the counter-based waits in
    // the step loop also wait for that step's X / phi prefetches.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // job coordinates within a step: waves 0-5 and 6-11 each cover 2 rows x 3 tiles
    const int jw = early ? wave : wave - 6;
    const int jrow = jw / 3, jtile = jw - 3 * jrow;
    // X fetch (waves 3-5) and output (waves 0-2): one pixel per lane, 2 rows
    const int io = early ? (wave < 3 ? tid : tid - 192) : 0;
    const int iorow = io / kCrop, iox = io - iorow * kCrop;

    for (int crop = (int)blockIdx.x; crop < ncrops; crop += (int)gridDim.x) {
        const bf16_t* Xc = kCells ? nullptr : A.X + (size_t)crop * kCropPix * kC0;
        const float* phicc = kCells ? nullptr : A.phic + (size_t)crop * kCropPix;
        const float* tile = kCells ? A.phi + (size_t)crop * pfh::kTileArea : nullptr;
        if (kCells) {
            // the crop of k_sur_crop: centred on the rounded centroid
            if (tid == 0) {
                const pfh_cell_state c = A.cells[crop];
                const pfh::Shape sh = pfh::shape_of((int)A.cls[crop]);
                CropCtx x;
                crop_origin(c, L, &x.ox, &x.oy);
                x.gx0 = c.gx0;
                x.gy0 = c.gy0;
                x.wx = sh.wx;
                x.wy = sh.wy;
                x.tx0 = sh.tx0;
                x.ty0 = sh.ty0;
                double sn, cs;
                pfh::sincos_d((double)c.theta, &sn, &cs);
                x.cs = (uint32_t)f2bf((float)cs) | ((uint32_t)f2bf((float)sn) << 16);
                sm.ctx = x;
                A.origin[2 * crop] = x.ox;
                A.origin[2 * crop + 1] = x.oy;
            }
            __syncthreads();
        }
        // the crop context in scalar registers
        CropCtx cx = {};
        if (kCells) {
            cx.ox = __builtin_amdgcn_readfirstlane(sm.ctx.ox);
            cx.oy = __builtin_amdgcn_readfirstlane(sm.ctx.oy);
            cx.gx0 = __builtin_amdgcn_readfirstlane(sm.ctx.gx0);
            cx.gy0 = __builtin_amdgcn_readfirstlane(sm.ctx.gy0);
            cx.wx = __builtin_amdgcn_readfirstlane(sm.ctx.wx);
            cx.wy = __builtin_amdgcn_readfirstlane(sm.ctx.wy);
            cx.tx0 = __builtin_amdgcn_readfirstlane(sm.ctx.tx0);
            cx.ty0 = __builtin_amdgcn_readfirstlane(sm.ctx.ty0);
            cx.cs = (uint32_t)__builtin_amdgcn_readfirstlane((int)sm.ctx.cs);
        }
        // X rows of the first step are outside the crop
        u32x2 xreg = {0u, 0u};
        float xv = 0.0f;
        uint32_t xs = 0u;
        bool xin = false;           // the fetched row lies inside the crop
        float preg = 0.0f;
        for (int s = kFirstStep; s <= kLastStep; ++s) {
            if (early && wave >= 3) {
                // X rows [2s+3, 2s+5): store what step s-1 fetched, fetch [2s+5, 2s+7)
                const int yw = 2 * s + 3 + iorow;
                if (kCells) {
                    // phi, S_other = (S - q(phi)) 2^-27 exact (0 when S < q),
                    // cos theta, sin theta
                    // rows outside the crop are zero in every channel
                    const uint32_t q = pfh::q_of(xv);
                    const float other = xs >= q ? (float)(xs - q) * pfh::kQInv : 0.0f;
                    xreg = xin ? u32x2{pack_bf16(xv, other), cx.cs} : u32x2{0u, 0u};
                }
                *reinterpret_cast<u32x2*>(sm.x + ring(yw) * kRowX + (iox + 1) * kXP) = xreg;
                const int yf = yw + 2;
                if (kCells) {
                    xv = 0.0f;
                    xs = 0u;
                    xin = in_crop(yf);
                    if (xin) {
                        int wxp, wyp;
                        xv = ctx_phi(tile, cx, L, iox, yf, &wxp, &wyp);
                        xs = A.S[(size_t)wyp * P + wxp];
                    }
                } else {
                    xreg = in_crop(yf) ? *reinterpret_cast<const u32x2*>(
                                             Xc + ((size_t)yf * kCrop + iox) * kC0)
                                       : u32x2{0u, 0u};
                }
            }
            float pnext = 0.0f;
            if (early && wave < 3) {
                const int yn = 2 * s - 7 + iorow;     // output row of step s+1
                if (in_crop(yn)) {
                    if (kCells) {
                        int wxp, wyp;
                        pnext = ctx_phi(tile, cx, L, iox, yn, &wxp, &wyp);
                    } else {
                        pnext = phicc[yn * kCrop + iox];
                    }
                }
            }
            if (early) {
                // conv1: h1 row 2s + jrow
                const int y1 = 2 * s + jrow;
                u32x4 lo = {0u, 0u, 0u, 0u}, hi = {0u, 0u, 0u, 0u};
                if (in_crop(y1)) {
                    const f32x16 a = conv4_tile(sm.x, wx, y1, jtile, col, half,
                                                bias_acc(sm.bias + 16 * half));
                    relu_pack(a, &lo, &hi);
                }
                store_h(sm.h1, y1, jtile, col, half, lo, hi);
                // conv2: h2 row 2s - 3 + jrow
                const int y2 = 2 * s - 3 + jrow;
                lo = u32x4{0u, 0u, 0u, 0u};
                hi = lo;
                if (in_crop(y2)) {
                    const f32x16 a = conv32_tile(sm.h1, wa, y2, jtile, col, half,
                                                 bias_acc(sm.bias + kCH + 16 * half));
                    relu_pack(a, &lo, &hi);
                }
                store_h(sm.h2, y2, jtile, col, half, lo, hi);
                // output rows [2s-9, 2s-7): b4 + partials in tap order
                if (wave < 3) {
                    const int yo = 2 * s - 9 + iorow;
                    if (in_crop(yo)) {
                        float d = b4;
#pragma unroll
                        for (int t = 0; t < 9; ++t) {
                            const int ky = t / 3, kx = t - 3 * ky;
                            d += sm.p[ring(yo + ky - 1) * kRowP + t * kHaloCols + iox + kx];
                        }
                        const size_t g = (size_t)crop * kCropPix + (size_t)yo * kCrop + iox;
                        if (dphi) dphi[g] = d;
                        pred[g] = preg + d;
                    }
                }
            } else {
                // conv3 (h3 row 2s - 6 + jrow) and the output layer's partials
                const int y3 = 2 * s - 6 + jrow;
                f32x16 pacc = {};
                if (in_crop(y3)) {
                    const f32x16 a = conv32_tile(sm.h2, wa, y3, jtile, col, half,
                                                 bias_acc(sm.bias + 2 * kCH + 16 * half));
                    u32x4 lo, hi;
                    relu_pack(a, &lo, &hi);
                    pacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wx[0], as_bf16x8(lo), pacc,
                                                                   0, 0, 0);
                    pacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wx[1], as_bf16x8(hi), pacc,
                                                                   0, 0, 0);
                }
                // lane (pixel, half 0) holds taps 0-3 and 8, half 1 taps 4-7:
                // high parts in accumulators r, low parts in r + 8
                float* pr = sm.p + ring(y3) * kRowP + 32 * jtile + col + 1;
#pragma unroll
                for (int r = 0; r < 4; ++r) pr[(4 * half + r) * kHaloCols] = pacc[r] + pacc[r + 8];
                if (half == 0) pr[8 * kHaloCols] = pacc[4] + pacc[12];
            }
            preg = pnext;
            // the rings are the only data shared between lanes here
            pfh::lds_barrier();
        }
    }
}

// Per-crop loss partial and g = 2 (dphi - target) / count
__global__ __launch_bounds__(256) void k_sur_loss(const float* dphi, const float* target,
                                                  float* g, double* crop_loss, float scale) {
    __shared__ double red[4];
    const int n = (int)blockIdx.x;
    double acc = 0.0;
    for (int p = (int)threadIdx.x; p < kCropPix; p += 256) {
        const size_t i = (size_t)n * kCropPix + p;
        const float e = dphi[i] - target[i];
        g[i] = scale * e;
        acc += (double)e * (double)e;
    }
    acc = pfh::wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) crop_loss[n] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Backward of the output layer: dH3 (masked by h3 > 0, bf16), and per-crop
// partials of dW4 [9][32] followed by db4.
__global__ __launch_bounds__(256) void k_sur_out_bwd(const float* g, const bf16_t* H,
                                                     const float* W4, bf16_t* dH,
                                                     float* dW4_part) {
    __shared__ float w[9 * kCH];
    const int n = (int)blockIdx.x, tid = (int)threadIdx.x;
    for (int i = tid; i < 9 * kCH; i += 256) w[i] = W4[i];
    __syncthreads();
    const float* gn = g + (size_t)n * kCropPix;
    // dH[q][ci] = sum_t g[q - off(t)] W4[t][ci]: one thread per (q, 8 channels)
    for (int i = tid; i < kCropPix * 4; i += 256) {
        const int q = i >> 2, c8 = i & 3;
        const int x = q % kCrop, y = q / kCrop;
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int t = 0; t < 9; ++t) {
            const int yy = y - (t / 3 - 1), xx = x - (t % 3 - 1);
            if ((unsigned)yy >= (unsigned)kCrop || (unsigned)xx >= (unsigned)kCrop) continue;
            const float gv = gn[yy * kCrop + xx];
            for (int k = 0; k < 8; ++k) acc[k] += gv * w[t * kCH + 8 * c8 + k];
        }
        // the 8 channels' h3 in one 16 B read, their dH in one 16 B write
        const size_t o = ((size_t)n * kCropPix + q) * kCH + 8 * c8;
        const u32x4 hv = *reinterpret_cast<const u32x4*>(H + o);
        u32x4 dv;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float h0 = bf2f((bf16_t)(hv[k] & 0xFFFFu)), h1 = bf2f((bf16_t)(hv[k] >> 16));
            const uint32_t d0 = f2bf(h0 > 0.0f ? acc[2 * k] : 0.0f);
            const uint32_t d1 = f2bf(h1 > 0.0f ? acc[2 * k + 1] : 0.0f);
            dv[k] = d0 | (d1 << 16);
        }
        *reinterpret_cast<u32x4*>(dH + o) = dv;
    }
    // dW4[t][ci] = sum_p g[p] h3[p + off(t)][ci] = sum_q g[q - off(t)] h3[q][ci]
    // (entries 0..287) and db4 = sum_p g[p] (entry 288), a row q_y at a time:
    // the three rows of g it pairs with (zero outside the crop) and its row
    // of h3 staged in LDS, each thread summing its entries over the row in
    // pixel order
    __shared__ float gl[3 * kHaloCols];
    __shared__ float hl[kCrop * kCH];
    float a0 = 0.0f, a1 = 0.0f;                    // entries tid and tid + 256
    for (int y = 0; y < kCrop; ++y) {
        __syncthreads();
        for (int i = tid; i < 3 * kHaloCols; i += 256) {
            const int r = i / kHaloCols, xx = i - r * kHaloCols - 1;
            const int yy = y + 1 - r;                   // row r pairs with ky = r
            gl[i] = ((unsigned)yy < (unsigned)kCrop && (unsigned)xx < (unsigned)kCrop)
                        ? gn[yy * kCrop + xx] : 0.0f;
        }
        for (int i = tid; i < kCrop * kCH; i += 256)
            hl[i] = bf2f(H[((size_t)n * kCrop + y) * kCrop * kCH + i]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int e = tid + 256 * k;
            if (e > 9 * kCH) break;
            float a = k == 0 ? a0 : a1;
            if (e < 9 * kCH) {
                const int t = e / kCH, cc = e - t * kCH;
                const int ky = t / 3, kx = t - 3 * ky;
                // g[y - ky + 1][x - kx + 1] sits at gl[ky][x - kx + 1 + 1]
                const float* gr = gl + ky * kHaloCols + 2 - kx;
                for (int x = 0; x < kCrop; ++x) a += gr[x] * hl[x * kCH + cc];
            } else {
                const float* gr = gl + kHaloCols + 1;   // g row y
                for (int x = 0; x < kCrop; ++x) a += gr[x];
            }
            if (k == 0) a0 = a; else a1 = a;
        }
    }
    dW4_part[(size_t)n * (9 * kCH + 1) + tid] = a0;
    if (tid + 256 <= 9 * kCH) dW4_part[(size_t)n * (9 * kCH + 1) + tid + 256] = a1;
}

// Per-crop partials of dW[co][tap][ci] and db[co] for an MFMA layer:
// dW = sum_p dY[p][co] X[p + off(tap)][ci], db = sum_p dY[p][co].
// dW is a GEMM on MFMA (v_mfma_f32_32x32x16_bf16, FP32 accumulation) with
// M = 32 output channels, N = 9 CIN (tap, input channel) in 32-column tiles
// and K = the crop's pixels, one crop row (six 16-pixel steps) at a time.
// Every operand is eight consecutive pixels of one channel, one 16 B LDS
// read: dY's row is staged channel-major, and each row of X is staged once
// as three x-shifted, channel-major copies (one per kx) in a ring of three
// rows, so that the taps of output row y read X rows y - 1, y, y + 1 from
// the ring without rebuilding them. db is summed on the VALU in pixel
// order, one lane per channel.
template <int CIN>
__global__ __launch_bounds__(256) void k_sur_wgrad(const bf16_t* dY, const bf16_t* X,
                                                   float* part) {
    constexpr int N = 9 * CIN;                  // columns: tap x input channel
    constexpr int NTL = (N + 31) / 32;          // 32-column tiles: 9 or 3
    constexpr int RP = kCrop + 8;               // LDS row pitch (bf16): 208 B
    constexpr int BLK = 3 * CIN * RP;           // one X row: [kx][ci][pixel]
    __shared__ __align__(16) bf16_t xb[3 * BLK];
    __shared__ __align__(16) bf16_t dyT[kCH * RP];
    __shared__ __align__(16) bf16_t zrow[RP];
    const int crop = (int)blockIdx.x, tid = (int)threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, col = lane & 31;
    // The ring starts zero: entries (kx = 0, pixel 0) and (kx = 2, pixel 95)
    // are the crop's zero padding and are never written again.
    for (int i = tid; i < 3 * BLK; i += 256) xb[i] = 0;
    for (int i = tid; i < RP; i += 256) zrow[i] = 0;
    const bf16_t* Xc = X + (size_t)crop * kCropPix * CIN;
    const bf16_t* dYc = dY + (size_t)crop * kCropPix * kCH;
    __syncthreads();

    // X row r (zeros outside the crop) into ring slot (r + 3) % 3: pixel xx
    // of channel ci lands at column xx - kx + 1 of copy kx
    auto build = [&](int r) {
        bf16_t* b = xb + ((r + 3) % 3) * BLK;
        constexpr int C8 = CIN / 8;
        const bool in = (unsigned)r < (unsigned)kCrop;
        for (int i = tid; i < kCrop * C8; i += 256) {
            const int xx = i / C8, c8 = i - (i / C8) * C8;
            u32x4 v = {0u, 0u, 0u, 0u};
            if (in)
                v = *reinterpret_cast<const u32x4*>(Xc + ((size_t)r * kCrop + xx) * CIN + 8 * c8);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const bf16_t h = (bf16_t)(v[k >> 1] >> (16 * (k & 1)));
                const int ci = 8 * c8 + k;
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) {
                    const int px = xx - kx + 1;
                    if ((unsigned)px < (unsigned)kCrop) b[(kx * CIN + ci) * RP + px] = h;
                }
            }
        }
    };
    // dY row y, channel-major
    auto stage_dy = [&](int y) {
        for (int i = tid; i < kCrop * (kCH / 8); i += 256) {
            const int px = i >> 2, c8 = i & 3;
            const u32x4 v = *reinterpret_cast<const u32x4*>(
                dYc + ((size_t)y * kCrop + px) * kCH + 8 * c8);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                dyT[(8 * c8 + k) * RP + px] = (bf16_t)(v[k >> 1] >> (16 * (k & 1)));
        }
    };

    f32x16 acc0 = {}, acc1 = {}, acc2 = {};    // tiles wave, wave + 4, wave + 8
    float db = 0.0f;
    build(-1);
    build(0);
    for (int y = 0; y < kCrop; ++y) {
        __syncthreads();            // row y - 1's operand reads are done
        build(y + 1);               // into the slot of row y - 2
        stage_dy(y);
        __syncthreads();
        if (tid < kCH) {
            for (int px = 0; px < kCrop; ++px) db += bf2f(dyT[tid * RP + px]);
        }
        const bf16_t* brow[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int nn = 32 * (wave + 4 * j) + col;
            brow[j] = zrow;
            if (nn < N) {
                const int t = nn / CIN, ci = nn - (nn / CIN) * CIN;
                const int ky = t / 3, kx = t - 3 * (t / 3);
                brow[j] = xb + ((y + ky + 2) % 3) * BLK + (kx * CIN + ci) * RP;
            }
        }
        const bf16_t* arow = dyT + col * RP;
#pragma unroll
        for (int ks = 0; ks < kCrop / 16; ++ks) {
            const int k0 = 16 * ks + 8 * half;
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(arow + k0);
            if (wave < NTL)
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    av, *reinterpret_cast<const bf16x8*>(brow[0] + k0), acc0, 0, 0, 0);
            if (wave + 4 < NTL)
                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    av, *reinterpret_cast<const bf16x8*>(brow[1] + k0), acc1, 0, 0, 0);
            if (wave + 8 < NTL)
                acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    av, *reinterpret_cast<const bf16x8*>(brow[2] + k0), acc2, 0, 0, 0);
        }
    }

    // lane holds column 32 T + col of tile T, output channels 8 (r >> 2) +
    // 4 half + (r & 3)
    float* out = part + (size_t)crop * (kCH * N + kCH);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int T = wave + 4 * j;
        const int nn = 32 * T + col;
        if (T >= NTL || nn >= N) continue;
        const f32x16 a = j == 0 ? acc0 : (j == 1 ? acc1 : acc2);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int co = 8 * (r >> 2) + 4 * half + (r & 3);
            out[(size_t)co * N + nn] = a[r];
        }
    }
    if (tid < kCH) out[kCH * N + tid] = db;
}

// Fixed-order sum of per-crop partials: out[e] = sum_n part[n][e] (ascending n)
__global__ void k_sur_reduce(const float* part, int ncrops, int width, float* out) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= width) return;
    float a = 0.0f;
    for (int n = 0; n < ncrops; ++n) a += part[(size_t)n * width + e];
    out[e] = a;
}

// W'[ci][tap][co] = W[co][8 - tap][ci] (32 x 32 layers): weights of the
// backward-data convolution
__global__ void k_sur_flip(const bf16_t* W, bf16_t* Wt) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= kCH * 9 * kCH) return;
    const int ci = e / (9 * kCH), rem = e - ci * 9 * kCH;
    const int t = rem / kCH, co = rem - t * kCH;
    Wt[e] = W[((size_t)co * 9 + (8 - t)) * kCH + ci];
}

// Adam on fp32 master weights, bf16 copy for compute (wb may be null)
__global__ void k_sur_adam(float* w, const float* g, float* m, float* v, bf16_t* wb, int n,
                           float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

// ---------------------------------------------------------------------------
// acceptance and write-back
// ---------------------------------------------------------------------------
// Per-crop relative L2 of a against b: sqrt(sum (a-b)^2 / sum b^2)
__global__ __launch_bounds__(256) void k_sur_rel_l2(const float* a, const float* b,
                                                    float* out) {
    __shared__ double r1[4], r2[4];
    const int n = (int)blockIdx.x;
    double e = 0.0, s = 0.0;
    for (int p = (int)threadIdx.x; p < kCropPix; p += 256) {
        const size_t i = (size_t)n * kCropPix + p;
        const double d = (double)a[i] - (double)b[i];
        e += d * d;
        s += (double)b[i] * (double)b[i];
    }
    e = pfh::wave_sum(e);
    s = pfh::wave_sum(s);
    if ((threadIdx.x & 63) == 0) {
        r1[threadIdx.x >> 6] = e;
        r2[threadIdx.x >> 6] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double E = ((r1[0] + r1[1]) + r1[2]) + r1[3];
        const double T = ((r2[0] + r2[1]) + r2[2]) + r2[3];
        out[n] = T > 0.0 ? (float)sqrt(E / T) : (E > 0.0 ? 1e30f : 0.0f);
    }
}

// Accepted cells: write the predicted crop into the tile's active window
// (window pixels outside the crop become 0). Cells with accept[n] == 0 are
// left alone.
__global__ __launch_bounds__(256) void k_sur_writeback(float* phi, const pfh_cell_state* cells,
                                                       const uint8_t* cls, const int* origin,
                                                       const float* pred, const int* accept,
                                                       int L) {
    const int n = (int)blockIdx.x;
    if (!accept[n]) return;
    const pfh_cell_state c = cells[n];
    const pfh::Shape s = pfh::shape_of((int)cls[n]);
    float* tile = phi + (size_t)n * pfh::kTileArea;
    const int ox = origin[2 * n], oy = origin[2 * n + 1];
    for (int p = (int)threadIdx.x; p < s.wx * s.wy; p += (int)blockDim.x) {
        const int a = p % s.wx, b = p / s.wx;
        int wx = c.gx0 + a;
        if (wx >= L) wx -= L;
        int wy = c.gy0 + b;
        if (wy >= L) wy -= L;
        int i = wx - ox;
        if (i < 0) i += L;
        int j = wy - oy;
        if (j < 0) j += L;
        float v = 0.0f;
        if (i < kCrop && j < kCrop) v = pred[(size_t)n * kCropPix + j * kCrop + i];
        tile[(size_t)(s.ty0 + b) * pfh::kTilePitch + s.tx0 + a] = v;
    }
}

// Acceptance test of every cell in one pass: the check field is read from
// the check step's tile on the recorded crop window (as k_sur_recrop), err[n]
// = relative L2 of the prediction against it (fp64 sums, lane tree, waves
// ascending), and rejected[0] counts the cells with err > tol.
__global__ __launch_bounds__(256) void k_sur_check(const float* phi_chk,
                                                   const pfh_cell_state* cells,
                                                   const uint8_t* cls, const int* origin, int L,
                                                   const float* pred, float tol, float* err,
                                                   uint32_t* rejected, uint32_t* err_max,
                                                   const int* go) {
    if (go && *go == 0) return;
    __shared__ double r1[4], r2[4];
    const int n = (int)blockIdx.x;
    const int gx0 = cells[n].gx0, gy0 = cells[n].gy0;
    const int ox = origin[2 * n], oy = origin[2 * n + 1];
    const float* tile = phi_chk + (size_t)n * pfh::kTileArea;
    const pfh::Shape sh = pfh::shape_of((int)cls[n]);
    // crop column 0 and row 0 relative to the cell's active window
    int a0 = ox - gx0;
    if (a0 < 0) a0 += L;
    int b0 = oy - gy0;
    if (b0 < 0) b0 += L;
    double e = 0.0, s = 0.0;
    // item t = 4 pixels of one crop row (one 16 B load of the prediction);
    // 96 x 24 items, 9 per thread, all in flight together
    constexpr int kQuads = kCrop / 4, kItems = kCrop * kQuads;
    static_assert(kItems % 256 == 0, "items per thread");
#pragma unroll
    for (int r = 0; r < kItems / 256; ++r) {
        const int t = (int)threadIdx.x + 256 * r;
        const int j = t / kQuads, i0 = 4 * (t - j * kQuads);
        const fx4 pv = *reinterpret_cast<const fx4*>(pred + (size_t)n * kCropPix + j * kCrop + i0);
        int b = b0 + j;
        if (b >= L) b -= L;
        const bool brow = b < sh.wy;
        const float* trow = tile + (size_t)(sh.ty0 + (brow ? b : 0)) * pfh::kTilePitch + sh.tx0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int a = a0 + i0 + u;
            if (a >= L) a -= L;
            const float bv = (brow && a < sh.wx) ? trow[a] : 0.0f;
            const double bd = (double)bv;
            const double d = (double)pv[u] - bd;
            e += d * d;
            s += bd * bd;
        }
    }
    e = pfh::wave_sum(e);
    s = pfh::wave_sum(s);
    if ((threadIdx.x & 63) == 0) {
        r1[threadIdx.x >> 6] = e;
        r2[threadIdx.x >> 6] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double E = ((r1[0] + r1[1]) + r1[2]) + r1[3];
        const double T = ((r2[0] + r2[1]) + r2[2]) + r2[3];
        const float r = T > 0.0 ? (float)sqrt(E / T) : (E > 0.0 ? 1e30f : 0.0f);
        err[n] = r;
        if (!(r <= tol)) atomicAdd(rejected, 1u);
        if (err_max) atomicMax(err_max, __float_as_uint(r));   // r >= 0: bits order as values
    }
}

// Per-cell acceptance of a block (one workgroup of 256 threads per cell).
// The committed field is the prediction on the crop and zero elsewhere, so
// its error against the check field is measured over the crop and over the
// check window outside the crop (a cell larger than the crop cannot pass):
//   E = sum_crop (pred - chk)^2 + sum_{window \ crop} chk^2,  T = sum chk^2,
//   I = sum_crop (chk - phi_n)^2  (the block's own increment).
// A cell is accepted when sqrt(E/T) <= tol, sqrt(E/I) <= tol_inc (the
// prediction carries the increment: predicting no change scores 1) and its
// accumulated error acc[n] + sqrt(E/T) stays within budget; acc[n] grows by
// the committed error of every accepted block. fp64 sums, fixed order.

// Value of world pixel (wx, wy) in a cell's tile (0 outside its window).
__device__ __forceinline__ float window_value(const float* tile, const pfh::Shape& sh, int gx0,
                                              int gy0, int wx, int wy, int L) {
    int a = wx - gx0;
    if (a < 0) a += L;
    int b = wy - gy0;
    if (b < 0) b += L;
    return (a < sh.wx && b < sh.wy)
        ? tile[(size_t)(sh.ty0 + b) * pfh::kTilePitch + sh.tx0 + a] : 0.0f;
}

// One cell's decision, computed by a whole workgroup of NT threads.
template <int NT>
__device__ void decide_cell(
    int n, const float* phi_chk, const pfh_cell_state* cells_chk, const uint8_t* cls_chk,
    const float* phi_now, const pfh_cell_state* cells_now, const uint8_t* cls_now,
    const int* origin, int L, const float* pred, float tol, float tol_inc, float budget,
    float* acc, float* err, int* accept, uint32_t* rejected, uint32_t* err_max) {
    constexpr int NW = NT / 64;
    __shared__ double red[3][NW];
    const int ox = origin[2 * n], oy = origin[2 * n + 1];
    const pfh::Shape sc = pfh::shape_of((int)cls_chk[n]);
    const pfh::Shape sn = pfh::shape_of((int)cls_now[n]);
    const int gxc = cells_chk[n].gx0, gyc = cells_chk[n].gy0;
    const int gxn = cells_now[n].gx0, gyn = cells_now[n].gy0;
    const float* tc = phi_chk + (size_t)n * pfh::kTileArea;
    const float* tn = phi_now + (size_t)n * pfh::kTileArea;
    const float* pr = pred + (size_t)n * kCropPix;
    double e = 0.0, t = 0.0, inc = 0.0;
    // Each thread visits its pixels in ascending order, U at a time: the U
    // pixels' loads are issued together, then added in order.
    constexpr int U = 4;
    for (int p0 = (int)threadIdx.x; p0 < kCropPix; p0 += U * NT) {
        float cf[U], vf[U], pf[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = p0 + u * NT;
            cf[u] = vf[u] = pf[u] = 0.0f;
            if (p < kCropPix) {
                const int j = p / kCrop, i = p - j * kCrop;
                int wx = ox + i, wy = oy + j;
                if (wx >= L) wx -= L;
                if (wy >= L) wy -= L;
                cf[u] = window_value(tc, sc, gxc, gyc, wx, wy, L);
                vf[u] = window_value(tn, sn, gxn, gyn, wx, wy, L);
                pf[u] = pr[p];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (p0 + u * NT < kCropPix) {
                const double c = (double)cf[u], v = (double)vf[u];
                const double d = (double)pf[u] - c;
                e += d * d;
                t += c * c;
                inc += (c - v) * (c - v);
            }
        }
    }
    // gxc, ox in [0, L) and a < wx < L, so i lies in (-L, 2L): one wrap
    const int wpix = sc.wx * sc.wy;
    for (int p0 = (int)threadIdx.x; p0 < wpix; p0 += U * NT) {
        float cf[U];
        bool on[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = p0 + u * NT;
            on[u] = false;
            cf[u] = 0.0f;
            if (p < wpix) {
                const int b = p / sc.wx, a = p - b * sc.wx;
                int i = gxc + a - ox, j = gyc + b - oy;
                i += i < 0 ? L : (i >= L ? -L : 0);
                j += j < 0 ? L : (j >= L ? -L : 0);
                on[u] = !(i < kCrop && j < kCrop);    // crop pixels are counted above
                if (on[u]) cf[u] = tc[(size_t)(sc.ty0 + b) * pfh::kTilePitch + sc.tx0 + a];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (on[u]) {
                const double c = (double)cf[u];
                e += c * c;
                t += c * c;
            }
        }
    }
    e = pfh::wave_sum(e);
    t = pfh::wave_sum(t);
    inc = pfh::wave_sum(inc);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = e;
        red[1][threadIdx.x >> 6] = t;
        red[2][threadIdx.x >> 6] = inc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double E = 0.0, T = 0.0, I = 0.0;
        for (int w = 0; w < NW; ++w) {
            E += red[0][w];
            T += red[1][w];
            I += red[2][w];
        }
        const float r = T > 0.0 ? (float)sqrt(E / T) : (E > 0.0 ? 1e30f : 0.0f);
        const float ri = I > 0.0 ? (float)sqrt(E / I) : (E > 0.0 ? 1e30f : 0.0f);
        const bool ok = r <= tol && ri <= tol_inc && acc[n] + r <= budget;
        err[n] = r;
        accept[n] = ok ? 1 : 0;
        if (ok) acc[n] += r;
        else atomicAdd(rejected, 1u);
        if (err_max) atomicMax(err_max, __float_as_uint(r));
    }
}

__global__ __launch_bounds__(256) void k_sur_check_cells(
    const float* phi_chk, const pfh_cell_state* cells_chk, const uint8_t* cls_chk,
    const float* phi_now, const pfh_cell_state* cells_now, const uint8_t* cls_now,
    const int* origin, int L, const float* pred, float tol, float tol_inc, float budget,
    float* acc, float* err, int* accept, uint32_t* rejected, uint32_t* err_max) {
    decide_cell<256>((int)blockIdx.x, phi_chk, cells_chk, cls_chk, phi_now, cells_now, cls_now,
                     origin, L, pred, tol, tol_inc, budget, acc, err, accept, rejected,
                     err_max);
}

// Accepted block, one workgroup of 1024 threads per cell: writes the
// prediction into the tile's active window (window pixels outside the crop
// become 0), scatters q(phi) into the zeroed aggregate S and stores the
// moments, visiting the window with the update's thread mapping so that V,
// Cx, Cy equal k_init_moments' sums bit for bit (k_scatter_all /
// k_init_moments semantics, src/kernels.cu:1098-1190).
// Staged classes: each pixel thread walks its rows of one column (the update's
// mapping), keeps the atomics' previous values in registers and tests them for
// S_OVERFLOW after the loop, so no row waits for its scatter.
template <int CLS>
__device__ void commit_staged(float* t, const float* pr, int L, int P, uint32_t* S, int gx0,
                              int gy0, int ox, int oy, pfh::MomentPart& mp, bool& ovf,
                              bool& clampf, const float* src = nullptr) {
    constexpr int WX = pfh::kShapes[CLS].wx, WY = pfh::kShapes[CLS].wy;
    constexpr int NB = pfh::kPixThreads / WX, RB = (WY + NB - 1) / NB;
    const int pix = (int)threadIdx.x - pfh::kControlLanes;
    if (pix < 0 || pix / WX >= NB) return;
    const int x = pix % WX, y0 = (pix / WX) * RB;
    const int wx = pfh::wrap_once(gx0 + x, L);
    int i = wx - ox;
    if (i < 0) i += L;
    const bool icrop = i < kCrop;
    // rows in chunks of 8: a chunk's scatters are in flight together and their
    // previous values are tested at the chunk's end (every addend is at most
    // 2^29, so a wrap needs old >= 2^32 - 2^29; only such candidates
    // recompute q from the value this thread stored)
    constexpr int CH = 8;
    for (int k0 = 0; k0 < RB; k0 += CH) {
        uint32_t olds[CH];
        // the chunk's values are read before any of its stores (the tile
        // written may be the buffer read, as far as the compiler knows), so
        // that the reads are in flight together: the prediction on the crop,
        // or (src) a whole window's field
        float vals[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            vals[k] = 0.0f;
            const int y = y0 + k0 + k;
            if (k0 + k < RB && y < WY) {
                const int wy = pfh::wrap_once(gy0 + y, L);
                int j = wy - oy;
                if (j < 0) j += L;
                vals[k] = src ? src[(size_t)y * pfh::kTilePitch + x]
                              : ((icrop && j < kCrop) ? pr[j * kCrop + i] : 0.0f);
            }
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            olds[k] = 0u;
            const int y = y0 + k0 + k;
            if (k0 + k < RB && y < WY) {
                const int wy = pfh::wrap_once(gy0 + y, L);
                const float v = vals[k];
                t[(size_t)y * pfh::kTilePitch + x] = v;
                const uint32_t q = pfh::q_of(v);
                if (q) {
                    olds[k] = atomicAdd(&S[(size_t)wy * P + wx], q);
                    clampf |= v * v > pfh::kQClampPhiSq;
                }
                pfh::moment_visit(mp, v, x, y);
            }
        }
#pragma unroll
        for (int k = 0; k < CH; ++k)
            if (olds[k] >= 0xE0000000u)
                ovf |= olds[k] >
                       0xFFFFFFFFu - pfh::q_of(t[(size_t)(y0 + k0 + k) * pfh::kTilePitch + x]);
    }
}

__global__ __launch_bounds__(pfh::kNT) void k_sur_commit(float* phi, pfh_cell_state* cells,
                                                         const uint8_t* cls, const int* origin,
                                                         const float* pred, int L, int P,
                                                         uint32_t* S, uint32_t* flags,
                                                         const int* skip_if) {
    if (skip_if && *skip_if) return;
    const int n = (int)blockIdx.x;
    const int k = (int)cls[n];
    const pfh::Shape sc = pfh::shape_of(k);
    const int gx0 = cells[n].gx0, gy0 = cells[n].gy0;
    const int ox = origin[2 * n], oy = origin[2 * n + 1];
    float* t = phi + (size_t)n * pfh::kTileArea + (size_t)sc.ty0 * pfh::kTilePitch + sc.tx0;
    const float* pr = pred + (size_t)n * kCropPix;
    bool ovf = false, clampf = false;
    pfh::WindowMoments m;
    if (k <= pfh::kClassStagedMax) {
        pfh::MomentPart mp = {0.0, 0.0, 0.0, sc.wx, -1, sc.wy, -1, 0};
        switch (k) {
            case 0: commit_staged<0>(t, pr, L, P, S, gx0, gy0, ox, oy, mp, ovf, clampf); break;
            case 1: commit_staged<1>(t, pr, L, P, S, gx0, gy0, ox, oy, mp, ovf, clampf); break;
            case 2: commit_staged<2>(t, pr, L, P, S, gx0, gy0, ox, oy, mp, ovf, clampf); break;
            default: commit_staged<3>(t, pr, L, P, S, gx0, gy0, ox, oy, mp, ovf, clampf); break;
        }
        m = pfh::reduce_moments(mp, sc.wx, sc.wy);
    } else {
        m = pfh::window_moments(k, [&](int a, int b) {
            const int wx = pfh::wrap_once(gx0 + a, L), wy = pfh::wrap_once(gy0 + b, L);
            int i = wx - ox;
            if (i < 0) i += L;
            int j = wy - oy;
            if (j < 0) j += L;
            const float v = (i < kCrop && j < kCrop) ? pr[j * kCrop + i] : 0.0f;
            t[(size_t)b * pfh::kTilePitch + a] = v;
            const uint32_t q = pfh::q_of(v);
            if (q) {
                const uint32_t old = atomicAdd(&S[(size_t)wy * P + wx], q);
                ovf |= old > 0xFFFFFFFFu - q;
                clampf |= v * v > pfh::kQClampPhiSq;
            }
            return v;
        });
    }
    if (ovf) atomicOr(&flags[PFH_FLAG_S_OVERFLOW], 1u);
    if (clampf) atomicOr(&flags[PFH_FLAG_Q_CLAMP], 1u);
    if (threadIdx.x == 0) pfh::store_moments(&cells[n], m);
}

// A block with per-cell fallback: an accepted cell (accept[n] != 0) takes
// the prediction in its current window; a rejected cell takes the check
// step's field (the plain stencil's own step of K dt) with the check step's
// window position and class. Either way the field is written, q(phi) is
// scattered into the zeroed aggregate and the moments are recomputed as in
// k_sur_commit.
__device__ void commit_cell(
    int n, bool take, float* phi, pfh_cell_state* cells, uint8_t* cls, const int* origin,
    const float* pred, const float* phi_chk, const pfh_cell_state* cells_chk,
    const uint8_t* cls_chk, int parity, int L, int P, uint32_t* S, uint32_t* flags) {
    const int kold = (int)cls[n];
    const int k = take ? kold : (int)cls_chk[n];
    const pfh::Shape sc = pfh::shape_of(k);
    const int gx0 = take ? cells[n].gx0 : cells_chk[n].gx0;
    const int gy0 = take ? cells[n].gy0 : cells_chk[n].gy0;
    const int ox = origin[2 * n], oy = origin[2 * n + 1];
    float* tile = phi + (size_t)n * pfh::kTileArea;
    if (k != kold) {
        // the window moves to another class layout: clear the tile first
        for (int i = (int)threadIdx.x; i < pfh::kTileArea; i += pfh::kNT) tile[i] = 0.0f;
        __syncthreads();
    }
    float* t = tile + (size_t)sc.ty0 * pfh::kTilePitch + sc.tx0;
    const float* src = take ? nullptr
        : phi_chk + (size_t)n * pfh::kTileArea + (size_t)sc.ty0 * pfh::kTilePitch + sc.tx0;
    const float* pr = pred + (size_t)n * kCropPix;
    bool ovf = false, clampf = false;
    pfh::WindowMoments m;
    if (k <= pfh::kClassStagedMax) {
        pfh::MomentPart mp = {0.0, 0.0, 0.0, sc.wx, -1, sc.wy, -1, 0};
        switch (k) {
            case 0: commit_staged<0>(t, pr, L, P, S, gx0, gy0, ox, oy, mp, ovf, clampf, src); break;
            case 1: commit_staged<1>(t, pr, L, P, S, gx0, gy0, ox, oy, mp, ovf, clampf, src); break;
            case 2: commit_staged<2>(t, pr, L, P, S, gx0, gy0, ox, oy, mp, ovf, clampf, src); break;
            default: commit_staged<3>(t, pr, L, P, S, gx0, gy0, ox, oy, mp, ovf, clampf, src); break;
        }
        m = pfh::reduce_moments(mp, sc.wx, sc.wy);
    } else {
        m = pfh::window_moments(k, [&](int a, int b) {
            const int wx = pfh::wrap_once(gx0 + a, L), wy = pfh::wrap_once(gy0 + b, L);
            float v;
            if (src) {
                v = src[(size_t)b * pfh::kTilePitch + a];
            } else {
                int i = wx - ox;
                if (i < 0) i += L;
                int j = wy - oy;
                if (j < 0) j += L;
                v = (i < kCrop && j < kCrop) ? pr[j * kCrop + i] : 0.0f;
            }
            t[(size_t)b * pfh::kTilePitch + a] = v;
            const uint32_t q = pfh::q_of(v);
            if (q) {
                const uint32_t old = atomicAdd(&S[(size_t)wy * P + wx], q);
                ovf |= old > 0xFFFFFFFFu - q;
                clampf |= v * v > pfh::kQClampPhiSq;
            }
            return v;
        });
    }
    if (ovf) atomicOr(&flags[PFH_FLAG_S_OVERFLOW], 1u);
    if (clampf) atomicOr(&flags[PFH_FLAG_Q_CLAMP], 1u);
    if (threadIdx.x == 0) {
        pfh_cell_state* cs = &cells[n];
        pfh::store_moments(cs, m);
        if (!take) {
            cs->gx0 = gx0;
            cs->gy0 = gy0;
            cs->cls = (uint8_t)k;
            cs->cls_written[parity] = (uint8_t)k;
            cls[n] = (uint8_t)k;
        }
    }
}

__global__ __launch_bounds__(pfh::kNT) void k_sur_commit_cells(
    float* phi, pfh_cell_state* cells, uint8_t* cls, const int* origin, const float* pred,
    const int* accept, const float* phi_chk, const pfh_cell_state* cells_chk,
    const uint8_t* cls_chk, int parity, int L, int P, uint32_t* S, uint32_t* flags) {
    const int n = (int)blockIdx.x;
    commit_cell(n, accept[n] != 0, phi, cells, cls, origin, pred, phi_chk, cells_chk, cls_chk,
                parity, L, P, S, flags);
}

// The polarity of K skipped steps: the plain step's tumble draw
// (src/kernels.cu:98-114) for steps step0 .. step0+K-1, in order.
__device__ void tumble_cell(pfh_cell_state* c, unsigned long long step0, int K,
                            unsigned long long seed, double p_tumble) {
    float theta = c->theta;
    uint32_t ctr = c->tumble_ctr;
    for (int k = 0; k < K; ++k) {
        const unsigned long long step = step0 + (unsigned long long)k;
        const pfh::Philox4 r = pfh::philox4x32_10(
            (uint32_t)(step & 0xFFFFFFFFull), (uint32_t)(step >> 32), (uint32_t)c->global_id,
            0xA5A5A5A5u, (uint32_t)(seed & 0xFFFFFFFFull), (uint32_t)(seed >> 32));
        if (pfh::uniform53(r.v[0], r.v[1]) < p_tumble) {
            theta = (float)(2.0 * pfh::kPi * pfh::uniform53(r.v[2], r.v[3]));
            ++ctr;
        }
    }
    c->theta = theta;
    c->tumble_ctr = ctr;
}

__global__ void k_sur_tumble(pfh_cell_state* cells, int N, const long long* step0_dev, int K,
                             unsigned long long seed, double p_tumble, const int* skip_if) {
    if (skip_if && *skip_if) return;
    const int n = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (n >= N) return;
    tumble_cell(cells + n, (unsigned long long)*step0_dev, K, seed, p_tumble);
}

// Two aggregate slots cleared for an accepted block (nothing when *skip_if).
__global__ void k_sur_zero2(uint32_t* a, uint32_t* b, size_t words, const int* skip_if) {
    if (skip_if && *skip_if) return;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t nvec = words >> 2;
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride) {
        reinterpret_cast<u32x4*>(a)[i] = z;
        reinterpret_cast<u32x4*>(b)[i] = z;
    }
    for (size_t i = (nvec << 2) + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += stride) {
        a[i] = 0u;
        b[i] = 0u;
    }
}

// End of a block. Accepted: the step counters advance by K and the queue
// cursors restart. Rejected while running: *go drops to 0 (every later
// gated launch does nothing), abort_step records the block's first step and
// rejected_blocks counts it; the caller then runs the plain steps itself.
__device__ void advance_counters(long long* stepc, unsigned long long* cursor, int K, int pin,
                                 int* go, const int* rejected, long long* abort_step,
                                 uint32_t* rejected_blocks) {
    if (*rejected) {
        if (*go) {
            *go = 0;
            *abort_step = stepc[pin];
            if (rejected_blocks) *rejected_blocks += 1u;
        }
        return;
    }
    stepc[0] += K;
    stepc[1] += K;
    cursor[0] = 0ull;
    cursor[1] = 0ull;
}

__global__ void k_sur_advance(long long* stepc, unsigned long long* cursor, int K, int pin,
                              int* go, const int* rejected, long long* abort_step,
                              uint32_t* rejected_blocks) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    advance_counters(stepc, cursor, K, pin, go, rejected, abort_step, rejected_blocks);
}

// The end of a block in one launch: k_sur_tumble's draws (nothing when the
// block has rejected cells) and k_sur_advance (thread 0 of workgroup 0).
// The draws take the block's first step from step_first (the check step's
// private counter, which prep set and the advance does not write), so they
// do not race with the advance of stepc.
__global__ void k_sur_finish(pfh_cell_state* cells, int N, const long long* step_first, int K,
                             unsigned long long seed, double p_tumble, long long* stepc,
                             unsigned long long* cursor, int pin, int* go, const int* rejected,
                             long long* abort_step, uint32_t* rejected_blocks) {
    const int n = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (*rejected == 0 && n < N)
        tumble_cell(cells + n, (unsigned long long)*step_first, K, seed, p_tumble);
    if (n == 0) advance_counters(stepc, cursor, K, pin, go, rejected, abort_step, rejected_blocks);
}

// Start of a block (nothing when *go == 0): the check step's private copies
// of the cell states and classes, its zeroed aggregate, queue cursors and
// step counter, and the rejected-cell count reset.
__global__ __launch_bounds__(256) void k_sur_prep(const int* go, int* rejected,
                                                  const pfh_cell_state* cells,
                                                  pfh_cell_state* cells_chk, const uint8_t* cls,
                                                  uint8_t* cls_chk, int N, uint32_t* S_chk,
                                                  size_t words, unsigned long long* cur_chk,
                                                  long long* step_chk, const long long* step_now) {
    if (*go == 0) return;
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const u32x4 z = {0u, 0u, 0u, 0u};
    u32x4* sv = reinterpret_cast<u32x4*>(S_chk);
    for (size_t i = i0; i < (words >> 2); i += stride) sv[i] = z;
    for (size_t i = ((words >> 2) << 2) + i0; i < words; i += stride) S_chk[i] = 0u;
    constexpr size_t kCellVec = sizeof(pfh_cell_state) / sizeof(u32x4);
    const u32x4* cs = reinterpret_cast<const u32x4*>(cells);
    u32x4* cd = reinterpret_cast<u32x4*>(cells_chk);
    for (size_t i = i0; i < (size_t)N * kCellVec; i += stride) cd[i] = cs[i];
    for (size_t i = i0; i < (size_t)N; i += stride) cls_chk[i] = cls[i];
    if (i0 == 0) {
        const long long st = *step_now;
        cur_chk[0] = 0ull;
        cur_chk[1] = 0ull;
        step_chk[0] = st;
        step_chk[1] = st;
        *rejected = 0;
    }
}

}  // namespace sur

// -------------------------------------------------------------------------
// C ABI
// -------------------------------------------------------------------------
using namespace pfh_host;

namespace {
int sur_grid() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 512;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 512;
    return 2 * (cus > 0 ? cus : 256);
}
}  // namespace

extern "C" {

int pfh_sur_crop(const float* phi, const pfh_cell_state* cells, const uint8_t* cls,
                 const uint32_t* S, int N, int L, int P, uint16_t* X, float* phic,
                 int* origin, void* stream) {
    if (!phi || !cells || !cls || !S || !X || !phic || !origin || N <= 0 || L < sur::kCrop ||
        P < L)
        return invalid("sur_crop");
    hipLaunchKernelGGL(sur::k_sur_crop, dim3((unsigned)N), dim3(256), 0, as_stream(stream),
                       phi, cells, cls, S, L, P, X, phic, origin);
    return launch_status("k_sur_crop");
}

int pfh_sur_recrop(const float* phi, const pfh_cell_state* cells, const uint8_t* cls,
                   const int* origin, int N, int L, float* out, void* stream) {
    if (!phi || !cells || !cls || !origin || !out || N <= 0 || L < sur::kCrop)
        return invalid("sur_recrop");
    hipLaunchKernelGGL(sur::k_sur_recrop, dim3((unsigned)N), dim3(256), 0, as_stream(stream),
                       phi, cells, cls, origin, L, out);
    return launch_status("k_sur_recrop");
}

int pfh_sur_conv(int cin, const uint16_t* X, const uint16_t* W, const float* bias,
                 const uint16_t* mask, uint16_t* Y, int ncrops, int relu, void* stream) {
    if (!X || !W || !Y || ncrops <= 0 || (cin != 8 && cin != 32)) return invalid("sur_conv");
    const int grid = sur_grid();
    if (cin == 8)
        hipLaunchKernelGGL(sur::k_sur_conv<8>, dim3((unsigned)grid), dim3(256), 0,
                           as_stream(stream), X, W, bias, mask, Y, ncrops, relu);
    else
        hipLaunchKernelGGL(sur::k_sur_conv<32>, dim3((unsigned)grid), dim3(256), 0,
                           as_stream(stream), X, W, bias, mask, Y, ncrops, relu);
    return launch_status("k_sur_conv");
}

int pfh_sur_out(const uint16_t* H, const float* W4, float b4, const float* phic, float* dphi,
                float* out, int ncrops, void* stream) {
    if (!H || !W4 || ncrops <= 0 || (out && !phic)) return invalid("sur_out");
    hipLaunchKernelGGL(sur::k_sur_out, dim3((unsigned)sur_grid()), dim3(256), 0,
                       as_stream(stream), H, W4, b4, phic, dphi, out, ncrops);
    return launch_status("k_sur_out");
}

namespace {
int infer_launch(const sur::InferArgs& a, bool cells, void* stream) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    const int grid = a.ncrops < cus ? a.ncrops : cus;
    if (cells)
        hipLaunchKernelGGL(sur::k_sur_infer<true>, dim3((unsigned)grid), dim3(sur::kIT), 0,
                           as_stream(stream), a);
    else
        hipLaunchKernelGGL(sur::k_sur_infer<false>, dim3((unsigned)grid), dim3(sur::kIT), 0,
                           as_stream(stream), a);
    return launch_status("k_sur_infer");
}

bool weights_ok(const uint16_t* W1, const float* b1, const uint16_t* W2, const float* b2,
                const uint16_t* W3, const float* b3, const float* W4) {
    return W1 && b1 && W2 && b2 && W3 && b3 && W4 &&
           !((((uintptr_t)W2) | ((uintptr_t)W3)) & 15u);
}
}  // namespace

int pfh_sur_infer(const uint16_t* X, const uint16_t* W1, const float* b1, const uint16_t* W2,
                  const float* b2, const uint16_t* W3, const float* b3, const float* W4, float b4,
                  const float* phic, float* dphi, float* pred, int ncrops, const int* go,
                  void* stream) {
    if (!X || !phic || !pred || ncrops <= 0 || !weights_ok(W1, b1, W2, b2, W3, b3, W4) ||
        (((uintptr_t)X) & 15u))
        return invalid("sur_infer");
    sur::InferArgs a = {};
    a.X = X;
    a.phic = phic;
    a.W1 = W1; a.W2 = W2; a.W3 = W3;
    a.b1 = b1; a.b2 = b2; a.b3 = b3; a.W4 = W4; a.b4 = b4;
    a.dphi = dphi;
    a.pred = pred;
    a.ncrops = ncrops;
    a.go = go;
    return infer_launch(a, false, stream);
}

int pfh_sur_infer_cells(const float* phi, const pfh_cell_state* cells, const uint8_t* cls,
                        const uint32_t* S, int N, int L, int P, const uint16_t* W1,
                        const float* b1, const uint16_t* W2, const float* b2, const uint16_t* W3,
                        const float* b3, const float* W4, float b4, int* origin, float* dphi,
                        float* pred, const int* go, void* stream) {
    if (!phi || !cells || !cls || !S || !origin || !pred || N <= 0 || L < sur::kCrop ||
        P < L || !weights_ok(W1, b1, W2, b2, W3, b3, W4))
        return invalid("sur_infer_cells");
    sur::InferArgs a = {};
    a.phi = phi;
    a.cells = cells;
    a.cls = cls;
    a.S = S;
    a.L = L;
    a.P = P;
    a.origin = origin;
    a.W1 = W1; a.W2 = W2; a.W3 = W3;
    a.b1 = b1; a.b2 = b2; a.b3 = b3; a.W4 = W4; a.b4 = b4;
    a.dphi = dphi;
    a.pred = pred;
    a.ncrops = N;
    a.go = go;
    return infer_launch(a, true, stream);
}

int pfh_sur_round_bf16(const float* x, uint16_t* y, int n, void* stream) {
    if (!x || !y || n <= 0 || (n & 1) || (((uintptr_t)y) & 3u)) return invalid("sur_round_bf16");
    const int np = n / 2;
    hipLaunchKernelGGL(sur::k_sur_round_bf16, dim3((unsigned)((np + 255) / 256)), dim3(256), 0,
                       as_stream(stream), x, reinterpret_cast<uint32_t*>(y), np);
    return launch_status("k_sur_round_bf16");
}

int pfh_sur_loss(const float* dphi, const float* target, float* g, double* crop_loss,
                 int ncrops, float scale, void* stream) {
    if (!dphi || !target || !g || !crop_loss || ncrops <= 0) return invalid("sur_loss");
    hipLaunchKernelGGL(sur::k_sur_loss, dim3((unsigned)ncrops), dim3(256), 0,
                       as_stream(stream), dphi, target, g, crop_loss, scale);
    return launch_status("k_sur_loss");
}

int pfh_sur_out_bwd(const float* g, const uint16_t* H, const float* W4, uint16_t* dH,
                    float* dW4_part, int ncrops, void* stream) {
    if (!g || !H || !W4 || !dH || !dW4_part || ncrops <= 0) return invalid("sur_out_bwd");
    hipLaunchKernelGGL(sur::k_sur_out_bwd, dim3((unsigned)ncrops), dim3(256), 0,
                       as_stream(stream), g, H, W4, dH, dW4_part);
    return launch_status("k_sur_out_bwd");
}

int pfh_sur_wgrad(int cin, const uint16_t* dY, const uint16_t* X, float* part, int ncrops,
                  void* stream) {
    if (!dY || !X || !part || ncrops <= 0 || (cin != 8 && cin != 32)) return invalid("sur_wgrad");
    if (cin == 8)
        hipLaunchKernelGGL(sur::k_sur_wgrad<8>, dim3((unsigned)ncrops), dim3(256), 0,
                           as_stream(stream), dY, X, part);
    else
        hipLaunchKernelGGL(sur::k_sur_wgrad<32>, dim3((unsigned)ncrops), dim3(256), 0,
                           as_stream(stream), dY, X, part);
    return launch_status("k_sur_wgrad");
}

int pfh_sur_reduce(const float* part, int ncrops, int width, float* out, void* stream) {
    if (!part || !out || ncrops <= 0 || width <= 0) return invalid("sur_reduce");
    hipLaunchKernelGGL(sur::k_sur_reduce, dim3((unsigned)((width + 255) / 256)), dim3(256), 0,
                       as_stream(stream), part, ncrops, width, out);
    return launch_status("k_sur_reduce");
}

int pfh_sur_flip(const uint16_t* W, uint16_t* Wt, void* stream) {
    if (!W || !Wt) return invalid("sur_flip");
    hipLaunchKernelGGL(sur::k_sur_flip, dim3((32 * 9 * 32 + 255) / 256), dim3(256), 0,
                       as_stream(stream), W, Wt);
    return launch_status("k_sur_flip");
}

int pfh_sur_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, int n, float lr,
                 float b1, float b2, float eps, float c1, float c2, void* stream) {
    if (!w || !g || !m || !v || n <= 0) return invalid("sur_adam");
    hipLaunchKernelGGL(sur::k_sur_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), w, g, m, v, wb, n, lr, b1, b2, eps, c1, c2);
    return launch_status("k_sur_adam");
}

int pfh_sur_rel_l2(const float* a, const float* b, float* out, int ncrops, void* stream) {
    if (!a || !b || !out || ncrops <= 0) return invalid("sur_rel_l2");
    hipLaunchKernelGGL(sur::k_sur_rel_l2, dim3((unsigned)ncrops), dim3(256), 0,
                       as_stream(stream), a, b, out);
    return launch_status("k_sur_rel_l2");
}

int pfh_sur_writeback(float* phi, const pfh_cell_state* cells, const uint8_t* cls,
                      const int* origin, const float* pred, const int* accept, int N, int L,
                      void* stream) {
    if (!phi || !cells || !cls || !origin || !pred || !accept || N <= 0 || L < sur::kCrop)
        return invalid("sur_writeback");
    hipLaunchKernelGGL(sur::k_sur_writeback, dim3((unsigned)N), dim3(256), 0, as_stream(stream),
                       phi, cells, cls, origin, pred, accept, L);
    return launch_status("k_sur_writeback");
}

int pfh_sur_check(const float* phi_chk, const pfh_cell_state* cells, const uint8_t* cls,
                  const int* origin, int N, int L, const float* pred, float tol, float* err,
                  uint32_t* rejected, uint32_t* err_max, const int* go, void* stream) {
    if (!phi_chk || !cells || !cls || !origin || !pred || !err || !rejected || N <= 0 ||
        L < sur::kCrop)
        return invalid("sur_check");
    hipLaunchKernelGGL(sur::k_sur_check, dim3((unsigned)N), dim3(256), 0, as_stream(stream),
                       phi_chk, cells, cls, origin, L, pred, tol, err, rejected, err_max, go);
    return launch_status("k_sur_check");
}

int pfh_sur_check_cells(const float* phi_chk, const pfh_cell_state* cells_chk,
                        const uint8_t* cls_chk, const float* phi_now,
                        const pfh_cell_state* cells_now, const uint8_t* cls_now,
                        const int* origin, int N, int L, const float* pred, float tol,
                        float tol_inc, float budget, float* acc, float* err, int* accept,
                        uint32_t* rejected, uint32_t* err_max, void* stream) {
    if (!phi_chk || !cells_chk || !cls_chk || !phi_now || !cells_now || !cls_now || !origin ||
        !pred || !acc || !err || !accept || !rejected || N <= 0 || L < sur::kCrop)
        return invalid("sur_check_cells");
    hipLaunchKernelGGL(sur::k_sur_check_cells, dim3((unsigned)N), dim3(256), 0,
                       as_stream(stream), phi_chk, cells_chk, cls_chk, phi_now, cells_now,
                       cls_now, origin, L, pred, tol, tol_inc, budget, acc, err, accept,
                       rejected, err_max);
    return launch_status("k_sur_check_cells");
}

int pfh_sur_commit_cells(float* phi, pfh_cell_state* cells, uint8_t* cls, const int* origin,
                         const float* pred, const int* accept, const float* phi_chk,
                         const pfh_cell_state* cells_chk, const uint8_t* cls_chk, int parity,
                         int N, int L, int P, uint32_t* S, uint32_t* flags, void* stream) {
    if (!phi || !cells || !cls || !origin || !pred || !accept || !phi_chk || !cells_chk ||
        !cls_chk || !S || !flags || N <= 0 || L < sur::kCrop || P < L || parity < 0 ||
        parity > 1)
        return invalid("sur_commit_cells");
    hipLaunchKernelGGL(sur::k_sur_commit_cells, dim3((unsigned)N), dim3(pfh::kNT), 0,
                       as_stream(stream), phi, cells, cls, origin, pred, accept, phi_chk,
                       cells_chk, cls_chk, parity, L, P, S, flags);
    return launch_status("k_sur_commit_cells");
}

int pfh_sur_commit(float* phi, pfh_cell_state* cells, const uint8_t* cls, const int* origin,
                   const float* pred, int N, int L, int P, uint32_t* S, uint32_t* flags,
                   const int* skip_if, void* stream) {
    if (!phi || !cells || !cls || !origin || !pred || !S || !flags || N <= 0 ||
        L < sur::kCrop || P < L)
        return invalid("sur_commit");
    hipLaunchKernelGGL(sur::k_sur_commit, dim3((unsigned)N), dim3(pfh::kNT), 0,
                       as_stream(stream), phi, cells, cls, origin, pred, L, P, S, flags,
                       skip_if);
    return launch_status("k_sur_commit");
}

int pfh_sur_tumble(pfh_cell_state* cells, int N, const long long* step0, int K,
                   unsigned long long polarity_seed, double p_tumble, const int* skip_if,
                   void* stream) {
    if (!cells || !step0 || N <= 0 || K < 0) return invalid("sur_tumble");
    if (K == 0) return PFH_OK;
    hipLaunchKernelGGL(sur::k_sur_tumble, dim3((unsigned)((N + 127) / 128)), dim3(128), 0,
                       as_stream(stream), cells, N, step0, K, polarity_seed, p_tumble, skip_if);
    return launch_status("k_sur_tumble");
}

int pfh_sur_zero2(uint32_t* a, uint32_t* b, unsigned long long words, const int* skip_if,
                  void* stream) {
    if (!a || !b || (((uintptr_t)a | (uintptr_t)b) & 15u)) return invalid("sur_zero2");
    if (words == 0ull) return PFH_OK;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    hipLaunchKernelGGL(sur::k_sur_zero2, dim3((unsigned)(4 * cus)), dim3(256), 0,
                       as_stream(stream), a, b, (size_t)words, skip_if);
    return launch_status("k_sur_zero2");
}

int pfh_sur_advance(long long* stepc, unsigned long long* cursor, int K, int pin, int* go,
                    const int* rejected, long long* abort_step, uint32_t* rejected_blocks,
                    void* stream) {
    if (!stepc || !cursor || !go || !rejected || !abort_step || K <= 0 || (pin & ~1))
        return invalid("sur_advance");
    hipLaunchKernelGGL(sur::k_sur_advance, dim3(1), dim3(64), 0, as_stream(stream), stepc, cursor,
                       K, pin, go, rejected, abort_step, rejected_blocks);
    return launch_status("k_sur_advance");
}

int pfh_sur_finish(pfh_cell_state* cells, int N, const long long* step_first, int K,
                   unsigned long long polarity_seed, double p_tumble, long long* stepc,
                   unsigned long long* cursor, int pin, int* go, const int* rejected,
                   long long* abort_step, uint32_t* rejected_blocks, void* stream) {
    if (!cells || !step_first || !stepc || !cursor || !go || !rejected || !abort_step ||
        N <= 0 || K <= 0 || (pin & ~1))
        return invalid("sur_finish");
    hipLaunchKernelGGL(sur::k_sur_finish, dim3((unsigned)((N + 127) / 128)), dim3(128), 0,
                       as_stream(stream), cells, N, step_first, K, polarity_seed, p_tumble,
                       stepc, cursor, pin, go, rejected, abort_step, rejected_blocks);
    return launch_status("k_sur_finish");
}

int pfh_sur_prep(const int* go, int* rejected, const pfh_cell_state* cells,
                 pfh_cell_state* cells_chk, const uint8_t* cls, uint8_t* cls_chk, int N,
                 uint32_t* S_chk, unsigned long long words, unsigned long long* cur_chk,
                 long long* step_chk, const long long* step_now, void* stream) {
    if (!go || !rejected || !cells || !cells_chk || !cls || !cls_chk || N <= 0 ||
        (words && !S_chk) || !cur_chk || !step_chk || !step_now ||
        ((((uintptr_t)cells) | ((uintptr_t)cells_chk) | ((uintptr_t)S_chk)) & 15u))
        return invalid("sur_prep");
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    hipLaunchKernelGGL(sur::k_sur_prep, dim3((unsigned)(4 * cus)), dim3(256), 0,
                       as_stream(stream), go, rejected, cells, cells_chk, cls, cls_chk, N, S_chk,
                       (size_t)words, cur_chk, step_chk, step_now);
    return launch_status("k_sur_prep");
}

}  // extern "C"
