// 3x3 periodic convolution on BF16 MFMA with FP32 accumulation (gfx950), for the learned flux
// correction network.  No counterpart in the reference (its networks are JAX); the closest analog is
// the periodic 3x3 convolution of surrogate/models.py:16-29.
//
// Layout: activations NHWC float32 in memory, channel count a multiple of 16; every value is rounded
// to bfloat16 (round to nearest even) when it enters an MFMA, so the arithmetic is BF16 x BF16 products
// accumulated in FP32.  Weights are pre-packed as BF16 in the MFMA operand order
// [tap][ci/16][co/16][lane][4] (tap = (dy+1)*3 + (dx+1), cross-correlation, circular padding).
//
// Implicit GEMM with v_mfma_f32_16x16x16_bf16 (A 16x16 M x K, B 16x16 K x N, D 16x16; lane l holds
// A[l%16][4(l/16)..+3], B[4(l/16)..+3][l%16], D[4(l/16)..+3][l%16]):
//   k_conv_fwd   M = 16 pixels of one image row, N = output channels, K = 9 taps x input channels;
//                epilogue: + bias, optional tanh-GELU, optional pre-activation output; or, for the
//                data gradient, multiplication by GELU'(pre-activation of the previous layer)
//   k_conv_wgrad M = input channels of one tap, N = output channels, K = pixels of a band of rows;
//                one partial per (band, tap) -> k_wgrad_reduce sums the bands in a fixed order
#include <hip/hip_runtime.h>

namespace npde_nn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

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

// One wavefront: 16 consecutive pixels of one image row x all output channels (NCOT tiles of 16).
template <int NCOT>
__global__ __launch_bounds__(256) void k_conv_fwd(const float* __restrict__ x, int B, int H, int W, int Cin,
                                                  const v4s* __restrict__ wp, const float* __restrict__ bias,
                                                  int cout_store, float* __restrict__ y, float* __restrict__ pre,
                                                  int epi, const float* __restrict__ gpre) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int xblocks = W >> 4;
    if (wave >= B * H * xblocks) return;
    const int xb = wave % xblocks;
    const int row = (wave / xblocks) % H;
    const int b = wave / (xblocks * H);
    const int m = lane & 15, kq = lane >> 4;
    const int x0 = xb << 4;
    const int nc16 = Cin >> 4;
    v4f acc[NCOT];
    for (int c = 0; c < NCOT; ++c) acc[c] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 9; ++t) {
        const int yy = wrap(row + t / 3 - 1, H);
        const int xx = wrap(x0 + m + t % 3 - 1, W);
        const float* src = x + ((size_t(b) * H + yy) * W + xx) * Cin + kq * 4;
        for (int c16 = 0; c16 < nc16; ++c16) {
            const float4 a4 = *reinterpret_cast<const float4*>(src + c16 * 16);
            const v4s a = to_bf16x4(a4.x, a4.y, a4.z, a4.w);
            const v4s* wt = wp + (size_t(t * nc16 + c16) * NCOT) * 64 + lane;
#pragma unroll
            for (int c = 0; c < NCOT; ++c) acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, wt[c * 64], acc[c], 0, 0, 0);
        }
    }
    const int cout = NCOT * 16;
#pragma unroll
    for (int c = 0; c < NCOT; ++c) {
        const int co = c * 16 + m;
        const float bb = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const size_t pix = (size_t(b) * H + row) * W + x0 + kq * 4 + i;
            float v = acc[c][i] + bb;
            if (pre) pre[pix * cout + co] = v;
            if (epi == kGelu) v = gelu(v);
            else if (epi == kGeluGrad) v *= gelu_grad(gpre[pix * cout + co]);
            if (co < cout_store) y[pix * cout_store + co] = v;
        }
    }
}

// Weight gradient partials: one wavefront per (band of `rows` image rows, tap).  dW[t][ci][co] =
// sum over pixels p of x[p + off(t)][ci] * g[p][co]; the centre-tap wavefront also sums g (bias grad).
template <int NC16, int NCOT>
__global__ __launch_bounds__(256) void k_conv_wgrad(const float* __restrict__ x, const float* __restrict__ g,
                                                    int B, int H, int W, int rows, float* __restrict__ part,
                                                    float* __restrict__ bpart) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int bands_per_img = H / rows;
    const int nbands = B * bands_per_img;
    if (wave >= nbands * 9) return;
    const int t = wave % 9;
    const int band = wave / 9;
    const int b = band / bands_per_img, r0 = (band % bands_per_img) * rows;
    const int Cin = NC16 * 16, Cout = NCOT * 16;
    const int m = lane & 15, kq = lane >> 4;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    v4f acc[NC16][NCOT];
    for (int i = 0; i < NC16; ++i)
        for (int j = 0; j < NCOT; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    float bsum[NCOT];
    for (int j = 0; j < NCOT; ++j) bsum[j] = 0.f;
    for (int r = r0; r < r0 + rows; ++r) {
        const int ys = wrap(r + dy, H);
        for (int x0 = 0; x0 < W; x0 += 16) {
            // this lane's 4 pixels (k index 4*kq + i) of the 16-pixel K block
            size_t ps[4], pd[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int xd = x0 + kq * 4 + i;
                pd[i] = (size_t(b) * H + r) * W + xd;
                ps[i] = (size_t(b) * H + ys) * W + wrap(xd + dx, W);
            }
#pragma unroll
            for (int j = 0; j < NCOT; ++j) {
                const int co = j * 16 + m;
                const float g0 = g[pd[0] * Cout + co], g1 = g[pd[1] * Cout + co];
                const float g2 = g[pd[2] * Cout + co], g3 = g[pd[3] * Cout + co];
                const v4s bv = to_bf16x4(g0, g1, g2, g3);
                bsum[j] += (g0 + g1) + (g2 + g3);
#pragma unroll
                for (int i = 0; i < NC16; ++i) {
                    const int ci = i * 16 + m;
                    const v4s av = to_bf16x4(x[ps[0] * Cin + ci], x[ps[1] * Cin + ci], x[ps[2] * Cin + ci],
                                             x[ps[3] * Cin + ci]);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[i][j], 0, 0, 0);
                }
            }
        }
    }
    float* out = part + (size_t(band) * 9 + t) * Cin * Cout;
#pragma unroll
    for (int i = 0; i < NC16; ++i)
#pragma unroll
        for (int j = 0; j < NCOT; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) out[(i * 16 + kq * 4 + e) * Cout + j * 16 + m] = acc[i][j][e];
    if (t == 4) {  // bias gradient of this band: sum the 4 lane groups holding the same channel
#pragma unroll
        for (int j = 0; j < NCOT; ++j) {
            float s = bsum[j];
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            if (kq == 0) bpart[size_t(band) * Cout + j * 16 + m] = s;
        }
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

// Forward / data-gradient convolution (see the file header).  cout = 16 or 32; cin a multiple of 16.
int npde_conv3x3(const float* x, int B, int H, int W, int cin, const void* wpack, const float* bias, int cout,
                 int cout_store, float* y, float* pre, int epilogue, const float* gpre, void* stream) {
    if (!x || !wpack || !y || B <= 0 || H <= 0 || W <= 0 || (W % 16) || (cin % 16) || cin <= 0 ||
        cout_store <= 0 || cout_store > cout || (epilogue == kGeluGrad && !gpre))
        return 1;
    const int waves = B * H * (W / 16);
    const dim3 grid((waves + 3) / 4), block(256);
    hipStream_t s = (hipStream_t)stream;
    const v4s* wp = (const v4s*)wpack;
    switch (cout) {
        case 16: hipLaunchKernelGGL(k_conv_fwd<1>, grid, block, 0, s, x, B, H, W, cin, wp, bias, cout_store, y, pre, epilogue, gpre); break;
        case 32: hipLaunchKernelGGL(k_conv_fwd<2>, grid, block, 0, s, x, B, H, W, cin, wp, bias, cout_store, y, pre, epilogue, gpre); break;
        case 48: hipLaunchKernelGGL(k_conv_fwd<3>, grid, block, 0, s, x, B, H, W, cin, wp, bias, cout_store, y, pre, epilogue, gpre); break;
        case 64: hipLaunchKernelGGL(k_conv_fwd<4>, grid, block, 0, s, x, B, H, W, cin, wp, bias, cout_store, y, pre, epilogue, gpre); break;
        default: return 1;
    }
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

// Weight and bias gradients.  part must hold (B*H/rows) * 9 * cin * cout floats and bpart
// (B*H/rows) * cout; dw is written as [tap][ci][co], db as [co].
int npde_conv3x3_wgrad(const float* x, const float* g, int B, int H, int W, int cin, int cout, int rows,
                       float* part, float* bpart, float* dw, float* db, void* stream) {
    if (!x || !g || !part || !bpart || !dw || !db || rows <= 0 || (H % rows) || (W % 16) || B <= 0) return 1;
    const int nbands = B * (H / rows);
    const dim3 grid((nbands * 9 + 3) / 4), block(256);
    hipStream_t s = (hipStream_t)stream;
    if (cin <= 0 || cout <= 0 || (cin % 16) || (cout % 16) || cin > 64 || cout > 64) return 1;
    using WgradFn = void (*)(const float*, const float*, int, int, int, int, float*, float*);
    static constexpr WgradFn table[4][4] = {
        {k_conv_wgrad<1, 1>, k_conv_wgrad<1, 2>, k_conv_wgrad<1, 3>, k_conv_wgrad<1, 4>},
        {k_conv_wgrad<2, 1>, k_conv_wgrad<2, 2>, k_conv_wgrad<2, 3>, k_conv_wgrad<2, 4>},
        {k_conv_wgrad<3, 1>, k_conv_wgrad<3, 2>, k_conv_wgrad<3, 3>, k_conv_wgrad<3, 4>},
        {k_conv_wgrad<4, 1>, k_conv_wgrad<4, 2>, k_conv_wgrad<4, 3>, k_conv_wgrad<4, 4>}};
    hipLaunchKernelGGL(table[cin / 16 - 1][cout / 16 - 1], grid, block, 0, s, x, g, B, H, W, rows, part, bpart);
    const int nw = 9 * cin * cout;
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
