// 3x3 periodic convolution on BF16 MFMA with FP32 accumulation (gfx950): the layer kernels of the learned
// flux correction's training path and of the full-state surrogates (ResNet, U-Net).  No counterpart in the
// reference (its networks are JAX); the closest analog is the periodic 3x3 convolution of
// surrogate/models.py:16-29.
//
// Activations are NHWC float32 in memory (channel count a multiple of 8); every value enters an MFMA
// rounded to bfloat16 (round to nearest even), so the arithmetic is BF16 x BF16 products accumulated in
// FP32.
//
// k_conv_fwd (forward and data gradient): an implicit GEMM on v_mfma_f32_16x16x32_bf16.  A workgroup of
// 4 wavefronts owns a tile of up to 128 output pixels of one image (8 x 16, 16 x 8 or 8 x 8 rows x
// columns).  For every 32 input channels it stages the tile's input region -- the tile and a 1-pixel
// periodic halo -- in LDS as bfloat16, once (64-byte pixel records, swizzled against bank conflicts), and
// the wavefronts then take all nine taps from there: M = 16 output channels (the weights, pre-packed in
// MFMA operand order and staged in LDS with the region, are the A operand), N = 16 pixels of the tile,
// K = the 32 staged channels of one tap (their B operand, one 16-byte LDS read per lane).  Epilogue: + bias, optional tanh-GELU, optional
// pre-activation output; or, for the data gradient, multiplication by GELU'(pre-activation of the
// previous layer).  Weights: [tap][cin/32][cout/16][lane][8] (tap = (dy+1)*3 + (dx+1), cross-correlation,
// circular padding; channels beyond cin or cout zero).
//
// k_conv_wgrad: M = input channels of one tap, N = output channels, K = pixels of a band of rows, on
// v_mfma_f32_16x16x16_bf16; one partial per (band, tap) -> k_wgrad_reduce sums the bands in a fixed order.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace npde_nn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ short bf16_rne(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return short((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
}

__device__ __forceinline__ v4s to_bf16x4(float a, float b, float c, float d) {
    v4s r;
    r.x = bf16_rne(a);
    r.y = bf16_rne(b);
    r.z = bf16_rne(c);
    r.w = bf16_rne(d);
    return r;
}

// two floats -> two bfloat16 (round to nearest even) in one dword, low half first (v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned pack_bf16(float lo, float hi) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}

constexpr float kGeluC0 = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kGeluC1 = 0.044715f;

__device__ __forceinline__ float gelu(float x) {
    const float t = tanhf(kGeluC0 * (x + kGeluC1 * x * x * x));
    return 0.5f * x * (1.0f + t);
}

__device__ __forceinline__ float gelu_grad(float x) {
    const float t = tanhf(kGeluC0 * (x + kGeluC1 * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGeluC0 * (1.0f + 3.0f * kGeluC1 * x * x);
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

enum Epilogue : int { kLinear = 0, kGelu = 1, kGeluGrad = 2 };

constexpr int kConvThreads = 256;
constexpr int kConvWaves = kConvThreads / 64;
constexpr int kTilePixels = 128;          // output pixels per workgroup at most
constexpr int kRegionMax = 18 * 10;       // staged pixels: (rows + 2) x (columns + 2) of an 8 x 16 or 16 x 8 tile
constexpr int kMaxBlocksPerWave = kTilePixels / 16 / kConvWaves;

// LDS pixel records: 32 bfloat16 channels = four 16-byte pieces; piece c of pixel p sits at c ^ ((p >> 1) & 3)
// (in units of 8 channels), so that the 16 lanes of a group, reading piece g of 16 consecutive pixels, spread
// over the banks.  Offsets in bfloat16 elements.
__device__ __forceinline__ int rec_at(int p, int c) { return (p << 5) + ((c ^ ((p >> 1) & 3)) << 3); }

// The weight fragments of a chunk go to LDS when they fit beside the region in 64 KiB (up to 64 output
// channels: 36 KiB); wider layers read them from global memory (L2) at each tap.
template <int NT>
constexpr bool kWeightsInLds = NT <= 4;

// One workgroup per tile of TH x TW output pixels (TH * TW a multiple of 16, at most 128; H % TH == 0,
// W % TW == 0); NT = cout / 16 output-channel tiles.
template <int NT>
__global__ __launch_bounds__(kConvThreads) void k_conv_fwd(const float* __restrict__ x, int B, int H, int W,
                                                           int cin, int TH, int TW, const u32x4* __restrict__ wp,
                                                           const float* __restrict__ bias, int cout_store,
                                                           float* __restrict__ y, float* __restrict__ pre, int epi,
                                                           const float* __restrict__ gpre) {
    __shared__ __align__(16) unsigned short s_in[kRegionMax * 32];
    extern __shared__ __align__(16) u32x4 s_w[];  // the chunk's 9 x NT weight fragments, 64 lanes each
    const int tiles_x = W / TW, tiles_y = H / TH;
    const int tile = blockIdx.x;
    const int b = tile / (tiles_x * tiles_y);
    if (b >= B) return;
    const int y0 = (tile / tiles_x) % tiles_y * TH, x0 = tile % tiles_x * TW;
    const int RW = TW + 2, nreg = (TH + 2) * RW;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = lane & 15, g = lane >> 4;
    const int nmb = TH * TW / 16;
    const int nchunk = (cin + 31) >> 5;

    v4f acc[kMaxBlocksPerWave][NT];
#pragma unroll
    for (int i = 0; i < kMaxBlocksPerWave; ++i)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[i][t] = v4f{0.f, 0.f, 0.f, 0.f};
    // region pixel of this lane's tile pixel in each of its M blocks (tap 0); the tap adds dy * RW + dx
    int rp0[kMaxBlocksPerWave];
#pragma unroll
    for (int i = 0; i < kMaxBlocksPerWave; ++i) {
        const int pix = (wave + kConvWaves * i) * 16 + m;
        const int py = pix / TW;
        rp0[i] = py * RW + (pix - py * TW);
    }

    for (int c = 0; c < nchunk; ++c) {
        // stage the chunk's weight fragments (tap t, output tile nt at slot t * NT + nt) and channels
        // 32c .. 32c + 31 of the region (channels past cin read as zero); both sets of loads are in flight
        // together, and the taps then read everything from LDS
        if constexpr (kWeightsInLds<NT>) {
            for (int it = threadIdx.x; it < 9 * NT * 64; it += kConvThreads) {
                const int frag = it >> 6, l = it & 63;
                const int t = frag / NT, nt = frag - t * NT;
                s_w[it] = wp[(size_t(t * nchunk + c) * NT + nt) * 64 + l];
            }
        }
        for (int it = threadIdx.x; it < nreg * 4; it += kConvThreads) {
            const int p = it >> 2, q = it & 3;
            const int ry = p / RW, rx = p - ry * RW;
            const int gy = wrap(y0 + ry - 1, H), gx = wrap(x0 + rx - 1, W);
            const int ch = (c << 5) + (q << 3);
            u32x4 v{0u, 0u, 0u, 0u};
            if (ch < cin) {
                const float4* src = reinterpret_cast<const float4*>(x + ((size_t(b) * H + gy) * W + gx) * cin + ch);
                const float4 lo = src[0], hi = src[1];
                v = u32x4{pack_bf16(lo.x, lo.y), pack_bf16(lo.z, lo.w), pack_bf16(hi.x, hi.y), pack_bf16(hi.z, hi.w)};
            }
            *reinterpret_cast<u32x4*>(s_in + rec_at(p, q)) = v;
        }
        __syncthreads();
#pragma unroll 3
        for (int t = 0; t < 9; ++t) {
            const int toff = (t / 3) * RW + (t % 3);
            bf16x8 bv[kMaxBlocksPerWave];
#pragma unroll
            for (int i = 0; i < kMaxBlocksPerWave; ++i)
                if (wave + kConvWaves * i < nmb)
                    bv[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(s_in + rec_at(rp0[i] + toff, g)));
            const u32x4* wt = kWeightsInLds<NT> ? s_w + t * NT * 64 + lane : wp + (size_t(t * nchunk + c) * NT) * 64 + lane;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const bf16x8 av = __builtin_bit_cast(bf16x8, wt[nt * 64]);
#pragma unroll
                for (int i = 0; i < kMaxBlocksPerWave; ++i)
                    if (wave + kConvWaves * i < nmb)
                        acc[i][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[i], acc[i][nt], 0, 0, 0);
            }
        }
        __syncthreads();  // the next chunk overwrites the region
    }

    // lane (m, g) holds output channels 16 nt + 4g .. + 3 of tile pixel 16 mb + m
    const int cout = NT * 16;
    const bool vec_store = (cout_store & 3) == 0 && (reinterpret_cast<uintptr_t>(y) & 15) == 0;
#pragma unroll
    for (int i = 0; i < kMaxBlocksPerWave; ++i) {
        const int mb = wave + kConvWaves * i;
        if (mb >= nmb) continue;
        const int pix = mb * 16 + m;
        const int py = pix / TW, px = pix - py * TW;
        const size_t gp = (size_t(b) * H + y0 + py) * W + x0 + px;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = nt * 16 + 4 * g;
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float z = acc[i][nt][e] + (bias ? bias[co + e] : 0.f);
                if (pre) pre[gp * cout + co + e] = z;
                if (epi == kGelu) z = gelu(z);
                else if (epi == kGeluGrad) z *= gelu_grad(gpre[gp * cout + co + e]);
                v[e] = z;
            }
            if (vec_store && co + 4 <= cout_store) {
                *reinterpret_cast<float4*>(y + gp * cout_store + co) = float4{v[0], v[1], v[2], v[3]};
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (co + e < cout_store) y[gp * cout_store + co + e] = v[e];
            }
        }
    }
}

// Weight gradient partials: one wavefront per (band of `rows` image rows, tap, group of NCI x 16 input
// channels; NCI = 4 when cin is a multiple of 64, else 2 or 1).  dW[t][ci][co] = sum over the band's pixels p of x[p + off(t)][ci] * g[p][co], 16 pixels
// at a time in row-major order within the band (rows * W a multiple of 16, W a multiple of 4, so that each
// lane's 4 consecutive pixels of a K block lie in one image row); the (centre tap, first group)
// wavefront also sums g over the band (bias gradient).  M = 16 input channels, N = 16 output channels,
// K = 16 pixels per v_mfma_f32_16x16x16_bf16; the gradient values of a K block are loaded once for all the
// group's channel tiles.
template <int NCI, int NCOT>
__global__ __launch_bounds__(256) void k_conv_wgrad(const float* __restrict__ x, const float* __restrict__ g,
                                                    int B, int H, int W, int cin, int rows, float* __restrict__ part,
                                                    float* __restrict__ bpart) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int nc16 = cin >> 4;
    const int ngroups = nc16 / NCI;
    const int bands_per_img = H / rows;
    const int nbands = B * bands_per_img;
    if (wave >= nbands * 9 * ngroups) return;
    const int grp = wave % ngroups;
    const int t = (wave / ngroups) % 9;
    const int band = wave / (ngroups * 9);
    const int b = band / bands_per_img, r0 = (band % bands_per_img) * rows;
    const int Cout = NCOT * 16;
    const int m = lane & 15, kq = lane >> 4;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int ct0 = grp * NCI;
    const bool want_bias = t == 4 && grp == 0;
    v4f acc[NCI][NCOT];
    float bsum[NCOT];
#pragma unroll
    for (int j = 0; j < NCOT; ++j) {
#pragma unroll
        for (int i = 0; i < NCI; ++i) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
        bsum[j] = 0.f;
    }
    // the K blocks walk the band 16 pixels at a time in row-major order; (r, xl) is the first of this lane's
    // 4 pixels of the current block (k index 4 kq + i: row r, columns xl + i), advanced by 16 pixels per block
    int r = r0 + (4 * kq) / W, xl = (4 * kq) % W;
    for (int blk = rows * W / 16; blk > 0; --blk) {
        size_t ps[4], pd[4];
        const size_t rd = (size_t(b) * H + r) * W, rs = (size_t(b) * H + wrap(r + dy, H)) * W;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int xd = xl + i;
            pd[i] = rd + xd;
            ps[i] = rs + wrap(xd + dx, W);
        }
        xl += 16;
        while (xl >= W) {
            xl -= W;
            ++r;
        }
        v4s av[NCI];
#pragma unroll
        for (int i = 0; i < NCI; ++i) {
            const int ci = (ct0 + i) * 16 + m;
            av[i] = to_bf16x4(x[ps[0] * cin + ci], x[ps[1] * cin + ci], x[ps[2] * cin + ci], x[ps[3] * cin + ci]);
        }
#pragma unroll
        for (int j = 0; j < NCOT; ++j) {
            const int co = j * 16 + m;
            const float g0 = g[pd[0] * Cout + co], g1 = g[pd[1] * Cout + co];
            const float g2 = g[pd[2] * Cout + co], g3 = g[pd[3] * Cout + co];
            if (want_bias) bsum[j] += (g0 + g1) + (g2 + g3);
            const v4s bv = to_bf16x4(g0, g1, g2, g3);
#pragma unroll
            for (int i = 0; i < NCI; ++i)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[i], bv, acc[i][j], 0, 0, 0);
        }
    }
    float* out = part + (size_t(band) * 9 + t) * cin * Cout;
#pragma unroll
    for (int i = 0; i < NCI; ++i) {
#pragma unroll
        for (int j = 0; j < NCOT; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) out[((ct0 + i) * 16 + kq * 4 + e) * Cout + j * 16 + m] = acc[i][j][e];
    }
    if (want_bias) {  // bias gradient of this band: sum the 4 lane groups holding the same channel
#pragma unroll
        for (int j = 0; j < NCOT; ++j) {
            float s = bsum[j];
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            if (kq == 0) bpart[size_t(band) * Cout + j * 16 + m] = s;
        }
    }
}

// LDS-staged weight gradient on v_mfma_f32_16x16x32_bf16, for bands of kWgradRows rows and up to 64 output
// channels: one workgroup per (band, 32 input channels).  For each chunk of TW columns (16, or 8 when W = 8)
// it stages, as bfloat16 and channel-major, the band's input region (kWgradRows + 2 rows x TW + 2 columns,
// periodic) in three copies shifted by one column each -- so that a tap's 8 consecutive pixels are one
// aligned 16-byte read whatever its column offset -- and the band's output gradient; the wavefronts then
// accumulate dW[tap][ci][co] over the chunk's pixels, 32 at a time: M = 16 input channels (the A operand,
// from the copy of the tap's column shift), N = 16 output channels (the B operand), K = 32 pixels.  The
// (tap, input tile, output tile) combinations are dealt round-robin to the 4 wavefronts.  Channel strides
// are padded to 4 dwords past a multiple of 32 so that the 16 lanes of a group spread over the banks.
constexpr int kWgradRows = 8;
constexpr int kXStride = 200;  // bfloat16 per input channel and copy: (8 + 2) rows x 16 columns, padded
constexpr int kGStride = 136;  // bfloat16 per output channel: 8 rows x 16 columns, padded
template <int NCOT>
__global__ __launch_bounds__(256) void k_conv_wgrad_lds(const float* __restrict__ x, const float* __restrict__ g,
                                                        int B, int H, int W, int cin, float* __restrict__ part) {
    constexpr int Cout = NCOT * 16;
    constexpr int kCombos = 9 * 2 * NCOT;
    constexpr int kPerWave = (kCombos + 3) / 4;
    __shared__ __align__(16) unsigned short xs[3 * 32 * kXStride];
    __shared__ __align__(16) unsigned short gs[Cout * kGStride];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = lane & 15, gq = lane >> 4;
    const int nchunks = (cin + 31) >> 5;
    const int bands_per_img = H / kWgradRows;
    const int band = blockIdx.x / nchunks, chunk = blockIdx.x - band * nchunks;
    if (band >= B * bands_per_img) return;
    const int b = band / bands_per_img, r0 = (band - b * bands_per_img) * kWgradRows;
    const int TW = W % 16 == 0 ? 16 : 8;
    const int RW = TW + 2;
    const int ci0 = chunk << 5;
    // this lane's 8 pixels of a K block: rows kb * (32 / TW) + prow, columns pcol .. pcol + 7
    const int prow = (8 * gq) / TW, pcol = (8 * gq) % TW;
    v4f acc[kPerWave];
#pragma unroll
    for (int i = 0; i < kPerWave; ++i) acc[i] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int c0 = 0; c0 < W; c0 += TW) {
        __syncthreads();  // every wavefront has read the previous chunk
        for (int it = threadIdx.x; it < (kWgradRows + 2) * RW * 4; it += 256) {
            const int q = it & 3, pix = it >> 2;
            const int ry = pix / RW, rx = pix - ry * RW;
            const int gy = wrap(r0 + ry - 1, H), gx = wrap(c0 + rx - 1, W);
            const int ch = ci0 + 8 * q;
            float4 lo{0.f, 0.f, 0.f, 0.f}, hi{0.f, 0.f, 0.f, 0.f};
            if (ch < cin) {
                const float4* src = reinterpret_cast<const float4*>(x + ((size_t(b) * H + gy) * W + gx) * cin + ch);
                lo = src[0];
                hi = src[1];
            }
            const short v[8] = {bf16_rne(lo.x), bf16_rne(lo.y), bf16_rne(lo.z), bf16_rne(lo.w),
                                bf16_rne(hi.x), bf16_rne(hi.y), bf16_rne(hi.z), bf16_rne(hi.w)};
#pragma unroll
            for (int dxc = 0; dxc < 3; ++dxc) {
                const int j = rx - dxc;  // the column of copy dxc that holds this region column
                if (j < 0 || j >= TW) continue;
                unsigned short* dst = xs + (dxc * 32 + 8 * q) * kXStride + ry * TW + j;
#pragma unroll
                for (int e = 0; e < 8; ++e) dst[e * kXStride] = (unsigned short)v[e];
            }
        }
        for (int it = threadIdx.x; it < kWgradRows * TW * (Cout / 8); it += 256) {
            const int q = it % (Cout / 8), pix = it / (Cout / 8);
            const int ry = pix / TW, cx = pix - ry * TW;
            const float4* src =
                reinterpret_cast<const float4*>(g + ((size_t(b) * H + r0 + ry) * W + c0 + cx) * Cout + 8 * q);
            const float4 lo = src[0], hi = src[1];
            const short v[8] = {bf16_rne(lo.x), bf16_rne(lo.y), bf16_rne(lo.z), bf16_rne(lo.w),
                                bf16_rne(hi.x), bf16_rne(hi.y), bf16_rne(hi.z), bf16_rne(hi.w)};
#pragma unroll
            for (int e = 0; e < 8; ++e) gs[(8 * q + e) * kGStride + pix] = (unsigned short)v[e];
        }
        __syncthreads();
        const int nkb = kWgradRows * TW / 32;
        for (int kb = 0; kb < nkb; ++kb) {
            const int row = kb * (32 / TW) + prow;
#pragma unroll
            for (int i = 0; i < kPerWave; ++i) {
                const int cmb = wave + 4 * i;
                if (cmb < kCombos) {  // wave-uniform
                    const int t = cmb / (2 * NCOT), rest = cmb - t * (2 * NCOT);
                    const int cit = rest / NCOT, cot = rest - cit * NCOT;
                    const int dy = t / 3, dxc = t - dy * 3;
                    const u32x4 a = *reinterpret_cast<const u32x4*>(
                        xs + (dxc * 32 + 16 * cit + m) * kXStride + (row + dy) * TW + pcol);
                    const u32x4 bv = *reinterpret_cast<const u32x4*>(gs + (16 * cot + m) * kGStride + kb * 32 + 8 * gq);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                                    __builtin_bit_cast(bf16x8, bv), acc[i], 0, 0, 0);
                }
            }
        }
    }
    // lane (m, gq) holds dW[t][ci0 + 16 cit + 4 gq + e][16 cot + m]
#pragma unroll
    for (int i = 0; i < kPerWave; ++i) {
        const int cmb = wave + 4 * i;
        if (cmb >= kCombos) break;
        const int t = cmb / (2 * NCOT), rest = cmb - t * (2 * NCOT);
        const int cit = rest / NCOT, cot = rest - cit * NCOT;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ci = ci0 + 16 * cit + 4 * gq + e;
            if (ci < cin) part[((size_t(band) * 9 + t) * cin + ci) * Cout + 16 * cot + m] = acc[i][e];
        }
    }
}

// Bias-gradient partials of the LDS path, one workgroup of 256 threads per band: thread (slice, co) sums g over
// the band's pixels p = slice, slice + nslice, ... (nslice = 256 / cout), then the slices' sums are added in
// slice order -- a fixed order, so the partials are deterministic.
__global__ __launch_bounds__(256) void k_bias_partials(const float* __restrict__ g, int H, int W, int cout,
                                                       float* __restrict__ bpart) {
    __shared__ float sums[256];
    const int nslice = 256 / cout;
    const int slice = threadIdx.x / cout, co = threadIdx.x - slice * cout;
    const int bands_per_img = H / kWgradRows;
    const int b = blockIdx.x / bands_per_img, r0 = (blockIdx.x - b * bands_per_img) * kWgradRows;
    const float* src = g + (size_t(b) * H + r0) * W * cout + co;
    float acc = 0.f;
    if (slice < nslice)
        for (int p = slice; p < kWgradRows * W; p += nslice) acc += src[size_t(p) * cout];
    sums[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < cout) {
        float t = 0.f;
        for (int k = 0; k < nslice; ++k) t += sums[k * cout + threadIdx.x];
        bpart[size_t(blockIdx.x) * cout + threadIdx.x] = t;
    }
}

// dW[k] = sum over bands (in band order) of part[band][k]; db likewise.  The band loop is unrolled so
// that many loads are in flight while the additions keep their fixed order (deterministic gradients).
__global__ void k_wgrad_reduce(const float* __restrict__ part, const float* __restrict__ bpart, int nbands,
                               int nw, int nb, float* __restrict__ dw, float* __restrict__ db) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nw) {
        float s = 0.f;
#pragma unroll 16
        for (int band = 0; band < nbands; ++band) s += part[size_t(band) * nw + k];
        dw[k] = s;
    } else if (k < nw + nb) {
        const int c = k - nw;
        float s = 0.f;
#pragma unroll 16
        for (int band = 0; band < nbands; ++band) s += bpart[size_t(band) * nb + c];
        db[c] = s;
    }
}

// MFMA operand-layout probe: D = A(16x16) B(16x16) with A[r][k] = a[r*16+k], B[k][c] = b[k*16+c] given
// row-major; lanes load the fragments the way the kernels above assume, and D is written back
// row-major from the accumulator layout they assume.
__global__ void k_mfma_probe(const float* __restrict__ a, const float* __restrict__ bm, float* __restrict__ d) {
    const int l = threadIdx.x;
    if (l >= 64) return;
    const int m = l & 15, kq = l >> 4;
    const v4s av = to_bf16x4(a[m * 16 + kq * 4], a[m * 16 + kq * 4 + 1], a[m * 16 + kq * 4 + 2], a[m * 16 + kq * 4 + 3]);
    const v4s bv = to_bf16x4(bm[(kq * 4) * 16 + m], bm[(kq * 4 + 1) * 16 + m], bm[(kq * 4 + 2) * 16 + m],
                             bm[(kq * 4 + 3) * 16 + m]);
    const v4f acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    for (int i = 0; i < 4; ++i) d[(kq * 4 + i) * 16 + m] = acc[i];
}

}  // namespace npde_nn

using namespace npde_nn;

extern "C" {

// Forward / data-gradient convolution (see the file header).  cin a multiple of 8 (x and wpack 16-byte
// aligned); cout 16 .. 128 or 192 in steps of 16; W a multiple of 8.  A tile is 16 columns wide when 16
// divides W, else 8, and takes the most rows (up to 128 pixels) that divide H with a multiple of 16 pixels.
int npde_conv3x3(const float* x, int B, int H, int W, int cin, const void* wpack, const float* bias, int cout,
                 int cout_store, float* y, float* pre, int epilogue, const float* gpre, void* stream) {
    if (!x || !wpack || !y || B <= 0 || H <= 0 || W <= 0 || (W % 8) || (cin % 8) || cin <= 0 ||
        ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(wpack)) & 15) || cout_store <= 0 ||
        cout_store > cout ||
        (epilogue == kGeluGrad && !gpre))
        return 1;
    // tile: 16 columns when the width allows, else 8; as many rows as divide H up to 128 pixels
    const int TW = W % 16 == 0 ? 16 : 8;
    int TH = kTilePixels / TW;
    while (TH > 1 && (H % TH != 0 || (TH * TW) % 16 != 0)) --TH;
    if ((TH * TW) % 16 != 0 || H % TH != 0) return 1;
    const dim3 grid(unsigned(B * (H / TH) * (W / TW))), block(kConvThreads);
    hipStream_t s = (hipStream_t)stream;
    const u32x4* wp = (const u32x4*)wpack;
#define NPDE_CONV_CASE(NT)                                                                          \
    case NT * 16:                                                                                   \
        hipLaunchKernelGGL(k_conv_fwd<NT>, grid, block, kWeightsInLds<NT> ? 9 * NT * 64 * sizeof(u32x4) : 0, \
                           s, x, B, H, W, cin, TH, TW, wp, bias, cout_store, y, pre, epilogue, gpre);   \
        break;
    switch (cout) {
        NPDE_CONV_CASE(1)
        NPDE_CONV_CASE(2)
        NPDE_CONV_CASE(3)
        NPDE_CONV_CASE(4)
        NPDE_CONV_CASE(5)
        NPDE_CONV_CASE(6)
        NPDE_CONV_CASE(7)
        NPDE_CONV_CASE(8)
        NPDE_CONV_CASE(12)
        default: return 1;
    }
#undef NPDE_CONV_CASE
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

// Weight and bias gradients.  cin a multiple of 16 up to 256, cout a multiple of 16 up to 128; W a multiple
// of 4; rows divides H and rows * W is a multiple of 16.  part must hold (B*H/rows) * 9 * cin * cout floats
// and bpart (B*H/rows) * cout; dw is written as [tap][ci][co], db as [co].  The LDS-staged kernel takes
// bands of 8 rows, up to 64 output channels, widths that are multiples of 8 and 16-byte aligned x and g;
// other shapes take the global-load kernel.
int npde_conv3x3_wgrad(const float* x, const float* g, int B, int H, int W, int cin, int cout, int rows,
                       float* part, float* bpart, float* dw, float* db, void* stream) {
    if (!x || !g || !part || !bpart || !dw || !db || rows <= 0 || B <= 0 || H <= 0 || W <= 0 || (W % 4) ||
        (H % rows) || ((rows * W) % 16) || cin <= 0 || cout <= 0 || (cin % 16) || (cout % 16) || cin > 256 ||
        cout > 128)
        return 1;
    const int nbands = B * (H / rows);
    hipStream_t s = (hipStream_t)stream;
    const int nw = 9 * cin * cout;
    const bool aligned = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(g)) & 15) == 0;
    if (rows == kWgradRows && cout <= 64 && W % 8 == 0 && aligned) {  // the LDS-staged path
        const dim3 grid(unsigned(nbands * ((cin + 31) / 32))), block(256);
        switch (cout) {
            case 16: hipLaunchKernelGGL(k_conv_wgrad_lds<1>, grid, block, 0, s, x, g, B, H, W, cin, part); break;
            case 32: hipLaunchKernelGGL(k_conv_wgrad_lds<2>, grid, block, 0, s, x, g, B, H, W, cin, part); break;
            case 48: hipLaunchKernelGGL(k_conv_wgrad_lds<3>, grid, block, 0, s, x, g, B, H, W, cin, part); break;
            default: hipLaunchKernelGGL(k_conv_wgrad_lds<4>, grid, block, 0, s, x, g, B, H, W, cin, part); break;
        }
        hipLaunchKernelGGL(k_bias_partials, dim3(unsigned(nbands)), dim3(256), 0, s, g, H, W, cout, bpart);
        hipLaunchKernelGGL(k_wgrad_reduce, dim3((nw + cout + 255) / 256), dim3(256), 0, s, part, bpart, nbands, nw,
                           cout, dw, db);
        return hipGetLastError() == hipSuccess ? 0 : 3;
    }
    const int nc16 = cin / 16;
    const int nci = nc16 % 4 == 0 ? 4 : (nc16 % 2 == 0 ? 2 : 1);  // input-channel tiles per wavefront
    const int waves = nbands * 9 * (nc16 / nci);
    const dim3 grid(unsigned((waves + 3) / 4)), block(256);
    using WgradFn = void (*)(const float*, const float*, int, int, int, int, int, float*, float*);
#define NPDE_WGRAD_ROW(NCI) \
    {k_conv_wgrad<NCI, 1>, k_conv_wgrad<NCI, 2>, k_conv_wgrad<NCI, 3>, k_conv_wgrad<NCI, 4>, \
     k_conv_wgrad<NCI, 5>, k_conv_wgrad<NCI, 6>, k_conv_wgrad<NCI, 7>, k_conv_wgrad<NCI, 8>}
    static constexpr WgradFn table[3][8] = {NPDE_WGRAD_ROW(1), NPDE_WGRAD_ROW(2), NPDE_WGRAD_ROW(4)};
#undef NPDE_WGRAD_ROW
    hipLaunchKernelGGL(table[nci == 4 ? 2 : nci - 1][cout / 16 - 1], grid, block, 0, s, x, g, B, H, W, cin, rows, part,
                       bpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((nw + cout + 255) / 256), dim3(256), 0, s, part, bpart, nbands, nw, cout,
                       dw, db);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

int npde_mfma_probe(const float* a, const float* b, float* d, void* stream) {
    if (!a || !b || !d) return 1;
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, a, b, d);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

}  // extern "C"
