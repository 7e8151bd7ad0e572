// The learned flux-correction network as one fused gfx950 kernel (inference), run by the library on the
// solver's own state: features -> four 3x3 periodic convolutions of 32 channels with tanh-GELU -> a 3x3
// linear head of 8 channels -> out_scale * tanh, written as the dimensionless face factors the stage
// kernels scale and add to their face fluxes (corr_x: face left of each cell, corr_y: face below it).
// No counterpart in the reference (its networks are JAX); the closest analog is the periodic 3x3
// convolution of surrogate/models.py:16-29.  The training path (npde/learned/conv.py) runs the same
// network layer by layer on the kernels of conv_mfma.hip.
//
// One workgroup (4 wavefronts) per 16 x 16 tile of output cells.  The five layers need a halo of 5 cells,
// so the workgroup computes the features on a 26 x 26 region and each layer's output on a region 2 cells
// smaller (24, 22, 20, 18, 16), keeping every activation in LDS as bfloat16 (ping-pong between two
// buffers, 66 KB: two workgroups per CU).  Each layer is an implicit GEMM on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation: M = 16 cells of the layer's region, N = 16 output channels, K = 32 -- one tap's 32
// input channels (hidden layers and head), or four taps' 8 features (first layer).  A wavefront keeps the
// layer's weight fragments in registers and walks the region's 16-cell blocks; the A fragments come
// from LDS with 16-byte reads.  Epilogues (bias, GELU or tanh) run in FP32; activations are rounded to
// bfloat16 (round to nearest even), as the training kernels round them at their MFMA inputs.
#include <hip/hip_runtime.h>

#include "npde_device.hpp"
#include "run_state.hpp"

namespace npde {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int kNetTile = 16;
constexpr int kNetThreads = 256;
constexpr int kNetWaves = kNetThreads / kWave;
constexpr int kFeatW = kNetTile + 10;  // the features' region: the tile and a 5-cell halo

// LDS: 32-channel activations as 64-byte cell records of four 16-byte chunks; the chunk holding channels
// 8c .. 8c+7 of cell p sits at chunk c ^ ((p >> 1) & 3), so the 16-byte reads of consecutive cells by one
// lane group fall on different banks.  The features use 16-byte records (8 channels).
__device__ __forceinline__ int chunk_at(int p, int c) { return (p << 5) + ((c ^ ((p >> 1) & 3)) << 3); }

__device__ __forceinline__ unsigned short bf16_bits(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// tanh-approximate GELU, 0.5 x (1 + tanh(y)) with y = sqrt(2/pi) (x + 0.044715 x^3), evaluated as the
// equal x * sigmoid(2y) = x / (1 + exp(-2y)) on the hardware's base-2 exponential and reciprocal: finite
// for every finite x (exp overflows to +inf: 0; underflows to 0: x), within ~1e-7 relative of the tanh
// form, far below the bfloat16 rounding of the activations it produces.
__device__ __forceinline__ float gelu_tanh(float x) {
    const float z = x * __builtin_fmaf(0.0713548162726009f, x * x, 1.5957691216057308f);  // 2y
    const float e = __builtin_amdgcn_exp2f(z * -1.4426950408889634f);                   // exp(-2y)
    return x * __builtin_amdgcn_rcpf(1.0f + e);
}

// two floats -> two bfloat16 (round to nearest even) in one dword, low half first (v_cvt_pk_bf16_f32)
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pack_bf16(float lo, float hi) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// Input features of one cell, as npde.learned.model.features computes them (binary64 primitives, then
// (value - offset) * (1 / scale) rounded to float32): rho, rho*u, rho*v, E, u, v, p, a.
__device__ __forceinline__ void cell_features(double d, double m, double n, double e, float* f) {
    const double u = m / d, v = n / d;
    const double p = (1.4 - 1.0) * (e - (0.5 * d) * (u * u + v * v));
    const double q = (1.4 * p) / d;
    const double a = __builtin_sqrt(q < 0.0 ? 0.0 : q);
    f[0] = float((d - 1.0) * 1.0);
    f[1] = float(m * 1.0);
    f[2] = float(n * 1.0);
    f[3] = float((e - 2.5) * (1.0 / 2.5));
    f[4] = float(u * 1.0);
    f[5] = float(v * 1.0);
    f[6] = float((p - 1.0) * 1.0);
    f[7] = float((a - 1.2) * (1.0 / 0.6));
}

// One convolution layer of the stack: input region Win x Win (cells of `in`), output region Wout x Wout
// (Wout = Win - 2) into `out` (hidden) or the face factors (head).  KB k-blocks of 32, NT n-tiles of 16.
// The weight fragment is the MFMA's A operand (16 output channels x 32 k) and the cells' activations its
// B operand (32 k x 16 cells), so lane (m, g) accumulates output channels 16t + 4g .. 16t + 4g + 3 of cell
// 16 blk + m and writes them as one 8-byte record piece.
//   FIRST: the input is the 8-feature region (k-block b, lane group g: tap 4b + g, 8 features; taps
//          beyond the ninth are zero);
//   HEAD:  8 output channels, out_scale * tanh written as binary64 factors into the instance's
//          correction planes (channels 0-3: face left of the cell, 4-7: face below it).
template <int KB, int NT, bool FIRST, bool HEAD, int Win>
__device__ __forceinline__ void net_layer(const unsigned short* __restrict__ in, unsigned short* __restrict__ out,
                                          const u32x4* __restrict__ frags, const float* __restrict__ bias,
                                          int wave, int lane, const InstDesc& D, int gx0, int gy0, float out_scale) {
    constexpr int Wout = Win - 2;  // compile-time region widths: the cell -> (row, column) divisions are constant
    const int ncell = Wout * Wout;
    const int nblk = (ncell + 15) >> 4;
    const int m = lane & 15, g = lane >> 4;
    bf16x8 w[KB][NT];
#pragma unroll
    for (int b = 0; b < KB; ++b)
#pragma unroll
        for (int t = 0; t < NT; ++t) w[b][t] = as_bf16x8(frags[(b * NT + t) * kWave + lane]);
    float bb[NT][4];  // the head has 8 biases: lane groups 2 and 3 (channels 8-15, no output) read none
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) bb[t][i] = (HEAD && g >= 2) ? 0.f : bias[t * 16 + 4 * g + i];

    for (int blk = wave; blk < nblk; blk += kNetWaves) {
        // this lane's cell (the last block's spare columns read the last cell and store nothing)
        const int c = (blk << 4) + m;
        const int o = c < ncell ? c : ncell - 1;
        const int oy = o / Wout, ox = o - oy * Wout;
        f32x4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < KB; ++b) {
            u32x4 a;
            if constexpr (FIRST) {
                const int tap = 4 * b + g;
                const int pin = (oy + tap / 3) * Win + ox + tap % 3;
                a = tap < 9 ? *reinterpret_cast<const u32x4*>(in + (pin << 3)) : u32x4{0u, 0u, 0u, 0u};
            } else {
                const int pin = (oy + b / 3) * Win + ox + b % 3;
                a = *reinterpret_cast<const u32x4*>(in + chunk_at(pin, g));
            }
            const bf16x8 av = as_bf16x8(a);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[b][t], av, acc[t], 0, 0, 0);
        }
        if (c >= ncell) continue;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if constexpr (HEAD) {
                if (g < 2) {  // channels 4g .. 4g+3: the x-face factors (g = 0) or the y-face factors (g = 1)
                    const int cy = c / Wout, cx = c - cy * Wout;
                    double* dst = const_cast<double*>(g == 0 ? D.corr_x : D.corr_y) +
                                  (long long)(gy0 + cy) * D.nx + gx0 + cx;
#pragma unroll
                    for (int i = 0; i < 4; ++i) dst[i * D.plane] = double(tanhf(acc[t][i] + bb[t][i]) * out_scale);
                }
            } else {
                const float v0 = gelu_tanh(acc[t][0] + bb[t][0]), v1 = gelu_tanh(acc[t][1] + bb[t][1]);
                const float v2 = gelu_tanh(acc[t][2] + bb[t][2]), v3 = gelu_tanh(acc[t][3] + bb[t][3]);
                typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
                *reinterpret_cast<u32x2*>(out + chunk_at(c, 2 * t + (g >> 1)) + 4 * (g & 1)) =
                    u32x2{pack_bf16(v0, v1), pack_bf16(v2, v3)};
            }
        }
    }
}

// Weight fragments (1 KB each: 64 lanes x 8 bfloat16) of the five layers, in launch order: first layer
// 3 k-blocks x 2 n-tiles, three hidden layers 9 x 2 each, head 9 x 1; biases 32, 32, 32, 32, 8.
// 69 fragments and 136 biases in all.
constexpr int kFragFirst = 0, kFragHidden = 6, kFragHead = 60;

// One tile: features of the 26 x 26 region around it, the five layers, the factors of its 16 x 16 cells.
__device__ __forceinline__ void net_tile(const InstDesc& D, const NetTask& tk, unsigned short* bufA,
                                         unsigned short* bufB, const u32x4* __restrict__ frags,
                                         const float* __restrict__ bias, float out_scale, unsigned* __restrict__ peak) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nx = D.nx, ny = D.ny;
    const long long P = D.plane;

    // features of the 26 x 26 region around the tile (periodic wrap), 16 bytes per cell; and the largest
    // density of the tile's own cells (the error indicator the caller reads)
    float dmax = 0.f;
    for (int c = threadIdx.x; c < kFeatW * kFeatW; c += kNetThreads) {
        const int ry = c / kFeatW, rx = c - ry * kFeatW;
        int gy = tk.y0 + ry - 5, gx = tk.x0 + rx - 5;
        gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
        const long long off = (long long)gy * nx + gx;
        float f[8];
        const double d = D.U[off];
        cell_features(d, D.U[off + P], D.U[off + 2 * P], D.U[off + 3 * P], f);
        // NaN compares false and leaves dmax alone: a non-finite state fails the solver's own check
        if (ry >= 5 && ry < 5 + kNetTile && rx >= 5 && rx < 5 + kNetTile && float(d) > dmax) dmax = float(d);
        const u32x4 rec{pack_bf16(f[0], f[1]), pack_bf16(f[2], f[3]), pack_bf16(f[4], f[5]), pack_bf16(f[6], f[7])};
        *reinterpret_cast<u32x4*>(bufB + (c << 3)) = rec;
    }
    // non-negative floats order as their bit patterns: one integer maximum per wavefront
    for (int o = 32; o > 0; o >>= 1) dmax = fmaxf(dmax, __shfl_xor(dmax, o, kWave));
    if (peak && lane == 0) atomicMax(peak + tk.inst, __float_as_uint(dmax));
    __syncthreads();
    net_layer<3, 2, true, false, 26>(bufB, bufA, frags + kFragFirst * kWave, bias, wave, lane, D, 0, 0, 0.f);
    __syncthreads();
    net_layer<9, 2, false, false, 24>(bufA, bufB, frags + kFragHidden * kWave, bias + 32, wave, lane, D, 0, 0, 0.f);
    __syncthreads();
    net_layer<9, 2, false, false, 22>(bufB, bufA, frags + (kFragHidden + 18) * kWave, bias + 64, wave, lane, D, 0, 0,
                                      0.f);
    __syncthreads();
    net_layer<9, 2, false, false, 20>(bufA, bufB, frags + (kFragHidden + 36) * kWave, bias + 96, wave, lane, D, 0, 0,
                                      0.f);
    __syncthreads();
    net_layer<9, 1, false, true, 18>(bufB, nullptr, frags + kFragHead * kWave, bias + 128, wave, lane, D, tk.x0, tk.y0,
                                     out_scale);
}

// Workgroups stride over the tile list.  counters == null: every instance's tiles.  Otherwise (the device
// loop of a flux-network run) nothing happens until every instance that is not done holds at a frame
// landing; then the tiles of the holding instances are evaluated, and the last workgroup to finish
// releases the holding instances (each decides its next step) and clears the counts.  Every workgroup
// read the counts and the hold flags before it counted itself in counters[kCountNetGroups], so the
// release cannot overtake a reader.
__global__ __launch_bounds__(kNetThreads) void k_flux_net(const InstDesc* __restrict__ desc, Ctl* __restrict__ ctl,
                                                          const NetTask* __restrict__ tasks, int ntasks, int n,
                                                          int* __restrict__ counters,
                                                          const u32x4* __restrict__ frags,
                                                          const float* __restrict__ bias, float out_scale,
                                                          unsigned* __restrict__ peak) {
    if (counters) {
        const int held = counters[kCountHeld];
        if (!(held > 0 && held == n - counters[kCountDone])) return;
    }
    __shared__ __align__(16) unsigned short bufA[24 * 24 * 32];  // first and third layers' outputs
    __shared__ __align__(16) unsigned short bufB[22 * 22 * 32];  // features, second and fourth layers' outputs
    __shared__ int last;
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const NetTask tk = tasks[t];
        if (counters && !ctl[tk.inst].hold) continue;
        const InstDesc& D = desc[tk.inst];
        if (D.corr_x == nullptr || D.corr_y == nullptr) continue;
        if (counters && ctl[tk.inst].net_off) {  // switched off at this landing or before: zero factors
            double* cx = const_cast<double*>(D.corr_x);
            double* cy = const_cast<double*>(D.corr_y);
            for (int c = threadIdx.x; c < kNetTile * kNetTile; c += blockDim.x) {
                const long long off = (long long)(tk.y0 + c / kNetTile) * D.nx + tk.x0 + c % kNetTile;
                for (int q = 0; q < 4; ++q) {
                    cx[off + q * D.plane] = 0.0;
                    cy[off + q * D.plane] = 0.0;
                }
            }
            continue;
        }
        net_tile(D, tk, bufA, bufB, frags, bias, out_scale, peak);
        __syncthreads();  // the next tile's features overwrite the buffer the head read
    }
    if (!counters) return;
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(counters + kCountNetGroups, 1) == int(gridDim.x) - 1;
    __syncthreads();
    if (!last) return;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        Ctl* c = ctl + k;
        if (!c->hold) continue;
        c->hold = 0;
        if (decide_next(desc[k], c)) atomicAdd(counters + kCountDone, 1);
    }
    if (threadIdx.x == 0) {
        counters[kCountHeld] = 0;
        counters[kCountNetGroups] = 0;
    }
}

// counters: the run's counters in the device loop (see k_flux_net), or null for every instance.  At most
// kNetBlocks workgroups (two per CU at this kernel's LDS footprint), so a launch that finds nothing to do
// costs little.
constexpr int kNetBlocks = 512;
hipError_t launch_flux_net(const InstDesc* desc, Ctl* ctl, const NetTask* tasks, int ntasks, int n, int* counters,
                           const void* frags, const float* bias, float out_scale, unsigned* peak, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    const int blocks = ntasks < kNetBlocks ? ntasks : kNetBlocks;
    hipLaunchKernelGGL(k_flux_net, dim3(unsigned(blocks)), dim3(kNetThreads), 0, s, desc, ctl, tasks, ntasks, n,
                       counters, (const u32x4*)frags, bias, out_scale, peak);
    return hipGetLastError();
}

// One 16x16x32 BF16 MFMA with the operand layout the kernel above assumes (lane l: A row l % 16 and B
// column l % 16 hold the 8 k-values 8 (l / 16) .. +7; D row 4 (l / 16) + i, column l % 16), on row-major
// float inputs A (16 x 32) and B (32 x 16) rounded to bfloat16; D (16 x 16) written row-major.
__global__ void k_mfma32_probe(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ d) {
    const int l = threadIdx.x;
    if (l >= kWave) return;
    const int m = l & 15, g = l >> 4;
    unsigned short av[8], bv[8];
    for (int j = 0; j < 8; ++j) {
        av[j] = bf16_bits(a[m * 32 + 8 * g + j]);
        bv[j] = bf16_bits(b[(8 * g + j) * 16 + m]);
    }
    u32x4 ap, bp;
    ap.x = av[0] | (unsigned(av[1]) << 16); ap.y = av[2] | (unsigned(av[3]) << 16);
    ap.z = av[4] | (unsigned(av[5]) << 16); ap.w = av[6] | (unsigned(av[7]) << 16);
    bp.x = bv[0] | (unsigned(bv[1]) << 16); bp.y = bv[2] | (unsigned(bv[3]) << 16);
    bp.z = bv[4] | (unsigned(bv[5]) << 16); bp.w = bv[6] | (unsigned(bv[7]) << 16);
    const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ap), as_bf16x8(bp), f32x4{0.f, 0.f, 0.f, 0.f},
                                                             0, 0, 0);
    for (int i = 0; i < 4; ++i) d[(4 * g + i) * 16 + m] = acc[i];
}

hipError_t launch_mfma32_probe(const float* a, const float* b, float* d, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma32_probe, dim3(1), dim3(kWave), 0, s, a, b, d);
    return hipGetLastError();
}

}  // namespace npde
