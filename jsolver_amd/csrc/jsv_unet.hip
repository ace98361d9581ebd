// Kernels of the learned warm start: a small fully convolutional U-Net in
// BF16 (weights and activations) with FP32 accumulation on CDNA4 matrix cores.
//
// Activations are NHWC bf16.  A 3x3 (or 1x1) convolution is an implicit GEMM
//   out[m][co] = sum_k A(m, k) * Wt[co][k],   m = (n, h, w),  k = (tap, ci)
// with A gathered on the fly from the zero-padded input (no im2col buffer).
// One wave computes a 16-pixel x 16-channel tile with
// v_mfma_f32_16x16x16_bf16 (__builtin_amdgcn_mfma_f32_16x16x16bf16_1k):
//   A fragment: lane l holds A[l % 16][4 (l / 16) + i], i = 0..3
//   B fragment: lane l holds B[4 (l / 16) + i][l % 16]
//   C fragment: lane l holds C[4 (l / 16) + i][l % 16]
// Cin is a multiple of 4, so the 4 consecutive k of a lane share one tap and
// are one 8-byte load.  The input gradient (dgrad) is the same kernel on the
// output gradient with the weights flipped and transposed; the weight
// gradient (wgrad) reduces over pixels with the same instruction and writes
// per-chunk partial sums that a second kernel adds in a fixed order.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

namespace {

thread_local std::string u_err;

int ufail(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    u_err = buf;
    return 1;
}

#define ULAUNCH()                                                                      \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess) return ufail("kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b)
{
    return __uint_as_float(((uint32_t)b) << 16);
}
// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------------------
// implicit-GEMM convolution, KS = 3 (padding 1) or 1
// x: [N][H][W][Cin] bf16; wt: [Cout][Kp] bf16 with Kp = roundup(KS*KS*Cin, 16),
// k = tap * Cin + ci, tap = (dy + 1) * KS + (dx + 1); zero padded beyond K.
// out: bf16 [M][Cout] (y16) and/or fp32 (y32); bias fp32 or null; relu 0/1.
// grid: (ceil(M / 64), ceil(Cout / 16)); block 256 = 4 waves x 16 pixels.
// ---------------------------------------------------------------------------
template <int KS>
__global__ void __launch_bounds__(256)
k_conv_igemm(const uint16_t* __restrict__ x, const uint16_t* __restrict__ wt,
             const float* __restrict__ bias, uint16_t* __restrict__ y16, float* __restrict__ y32,
             int N, int H, int W, int Cin, int Cout, int Kp, int relu)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long M = (long)N * H * W;
    const long m0 = ((long)blockIdx.x * 4 + wave) * 16;
    const int n0 = blockIdx.y * 16;
    const int r = lane & 15;
    const int q = lane >> 4;
    const long ma = m0 + r;                    // pixel of this lane's A row
    int an = 0, ah = 0, aw = 0;
    const bool avalid = ma < M;
    if (avalid) {
        aw = (int)(ma % W);
        ah = (int)((ma / W) % H);
        an = (int)(ma / ((long)W * H));
    }
    const int co_b = n0 + r;                   // output channel of this lane's B column
    const int K = KS * KS * Cin;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < Kp; kb += 16) {
        const int k = kb + 4 * q;
        bf16x4 a = {0, 0, 0, 0};
        if (avalid && k < K) {
            const int tap = k / Cin;
            const int ci = k - tap * Cin;
            const int dy = (KS == 3) ? tap / 3 - 1 : 0;
            const int dx = (KS == 3) ? tap % 3 - 1 : 0;
            const int hh = ah + dy, ww = aw + dx;
            if (hh >= 0 && hh < H && ww >= 0 && ww < W)
                a = *reinterpret_cast<const bf16x4*>(x + (((long)an * H + hh) * W + ww) * Cin + ci);
        }
        bf16x4 b = {0, 0, 0, 0};
        if (co_b < Cout) b = *reinterpret_cast<const bf16x4*>(wt + (long)co_b * Kp + k);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int co = n0 + r;
    if (co >= Cout) return;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long m = m0 + 4 * q + i;
        if (m >= M) continue;
        float v = acc[i] + bv;
        if (relu) v = v > 0.f ? v : 0.f;
        if (y16) y16[m * Cout + co] = f2bf(v);
        if (y32) y32[m * Cout + co] = v;
    }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[co][k] = sum_m dY[m][co] A(m, k).
// grid: (ceil(Kp / 64), ceil(Cout / 16), chunks); wave w handles k tile
// blockIdx.x * 4 + w; pixels [chunk * per, (chunk + 1) * per).
// part: [chunks][Cout][Kp] fp32 partial sums.
// ---------------------------------------------------------------------------
template <int KS>
__global__ void __launch_bounds__(256)
k_conv_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
             float* __restrict__ part, int N, int H, int W, int Cin, int Cout, int Kp, long per)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int k0 = (blockIdx.x * 4 + wave) * 16;
    if (k0 >= Kp) return;
    const int co0 = blockIdx.y * 16;
    const long M = (long)N * H * W;
    const long p_begin = (long)blockIdx.z * per;
    long p_end = p_begin + per;
    if (p_end > M) p_end = M;
    const int r = lane & 15, q = lane >> 4;
    const int K = KS * KS * Cin;
    // B column of this lane: k = k0 + r -> (tap, ci, dy, dx)
    const int kb = k0 + r;
    const bool kvalid = kb < K;
    int ci = 0, dyy = 0, dxx = 0;
    if (kvalid) {
        const int tap = kb / Cin;
        ci = kb - tap * Cin;
        dyy = (KS == 3) ? tap / 3 - 1 : 0;
        dxx = (KS == 3) ? tap % 3 - 1 : 0;
    }
    const int coa = co0 + r;                  // A row: output channel
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long p0 = p_begin; p0 < p_end; p0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long m = p0 + 4 * q + i;
            short av = 0, bv = 0;
            if (m < p_end) {
                if (coa < Cout) av = (short)dy[m * Cout + coa];
                if (kvalid) {
                    const int w = (int)(m % W);
                    const int h = (int)((m / W) % H);
                    const long n = m / ((long)W * H);
                    const int hh = h + dyy, ww = w + dxx;
                    if (hh >= 0 && hh < H && ww >= 0 && ww < W)
                        bv = (short)x[((n * H + hh) * W + ww) * Cin + ci];
                }
            }
            a[i] = av;
            b[i] = bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    float* out = part + (long)blockIdx.z * Cout * Kp;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int co = co0 + 4 * q + i;
        const int k = k0 + r;
        if (co < Cout && k < Kp) out[(long)co * Kp + k] = acc[i];
    }
}

// sum partials over chunks (fixed order) into dw; bias gradient from dY
__global__ void __launch_bounds__(256)
k_reduce_chunks(const float* __restrict__ part, float* __restrict__ dw, long n, int chunks)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int c = 0; c < chunks; ++c) s += part[(long)c * n + i];
    dw[i] = s;
}

__global__ void __launch_bounds__(256)
k_bias_grad(const uint16_t* __restrict__ dy, float* __restrict__ db, long M, int C)
{
    // one workgroup per channel; fixed-order tree sum
    __shared__ float sh[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (long m = threadIdx.x; m < M; m += 256) s += bf2f(dy[m * C + c]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = sh[0];
}

// ---------------------------------------------------------------------------
// elementwise / layout kernels (NHWC bf16)
// ---------------------------------------------------------------------------
__global__ void k_avgpool2(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int N, int H,
                           int W, int C)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    const long tot = (long)N * Ho * Wo * C;
    if (i >= tot) return;
    const int c = (int)(i % C);
    const int w = (int)((i / C) % Wo);
    const int h = (int)((i / ((long)C * Wo)) % Ho);
    const long n = i / ((long)C * Wo * Ho);
    const long b = ((n * H + 2 * h) * W + 2 * w) * C + c;
    const float s = bf2f(x[b]) + bf2f(x[b + C]) + bf2f(x[b + (long)W * C]) + bf2f(x[b + (long)W * C + C]);
    y[i] = f2bf(0.25f * s);
}

__global__ void k_avgpool2_bwd(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx, int N,
                               int H, int W, int C)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long tot = (long)N * H * W * C;
    if (i >= tot) return;
    const int c = (int)(i % C);
    const int w = (int)((i / C) % W);
    const int h = (int)((i / ((long)C * W)) % H);
    const long n = i / ((long)C * W * H);
    const int Ho = H / 2, Wo = W / 2;
    dx[i] = f2bf(0.25f * bf2f(dy[((n * Ho + h / 2) * Wo + w / 2) * C + c]));
}

// nearest 2x upsample of x (C1 channels) into channels [0, C1) of y (Ct
// channels), and copy of skip (C2 channels) into [C1, C1 + C2) -- the
// decoder's concatenated input.
__global__ void k_up_concat(const uint16_t* __restrict__ x, const uint16_t* __restrict__ skip,
                            uint16_t* __restrict__ y, int N, int H, int W, int C1, int C2)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int Ct = C1 + C2;
    const long tot = (long)N * H * W * Ct;
    if (i >= tot) return;
    const int c = (int)(i % Ct);
    const long pix = i / Ct;
    const int w = (int)(pix % W);
    const int h = (int)((pix / W) % H);
    const long n = pix / ((long)W * H);
    if (c < C1) y[i] = x[((n * (H / 2) + h / 2) * (W / 2) + w / 2) * C1 + c];
    else y[i] = skip[pix * C2 + (c - C1)];
}

__global__ void k_up_concat_bwd(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx,
                                uint16_t* __restrict__ dskip, int N, int H, int W, int C1, int C2)
{
    // dx: [N][H/2][W/2][C1] = sum of the 4 upsampled copies; dskip: slice of dy
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int Ct = C1 + C2;
    const int Ho = H / 2, Wo = W / 2;
    const long tot_lo = (long)N * Ho * Wo * C1;
    if (i < tot_lo) {
        const int c = (int)(i % C1);
        const int w = (int)((i / C1) % Wo);
        const int h = (int)((i / ((long)C1 * Wo)) % Ho);
        const long n = i / ((long)C1 * Wo * Ho);
        float s = 0.f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                s += bf2f(dy[(((n * H + 2 * h + a) * W) + 2 * w + b) * Ct + c]);
        dx[i] = f2bf(s);
    }
    const long tot_s = (long)N * H * W * C2;
    if (i < tot_s && dskip) {
        const int c = (int)(i % C2);
        const long pix = i / C2;
        dskip[i] = dy[pix * Ct + C1 + c];
    }
}

// dy <- dy * (y > 0) (ReLU backward, y = the layer's output)
__global__ void k_relu_bwd(uint16_t* __restrict__ dy, const uint16_t* __restrict__ y, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (!(bf2f(y[i]) > 0.f)) dy[i] = 0;
}

// a += b (bf16, fp32 arithmetic): skip-gradient accumulation
__global__ void k_add_bf16(uint16_t* __restrict__ a, const uint16_t* __restrict__ b, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    a[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    b[i] = f2bf(a[i]);
}

// Adam on fp32 master weights, refreshing the bf16 copy.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ w16, long n, float lr, float b1,
                       float b2, float eps, float bc1, float bc2, float wd)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}

inline dim3 g1(long n) { return dim3((unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1)); }

}  // namespace

extern "C" {

const char* jsv_unet_last_error(void) { return u_err.c_str(); }

int jsv_conv_fwd(const void* x, const void* wt, const float* bias, void* y16, float* y32, int N,
                 int H, int W, int Cin, int Cout, int ksize, int Kp, int relu, void* stream)
{
    if (Cin % 4 != 0) return ufail("Cin must be a multiple of 4 (got %d)", Cin);
    if (Kp % 16 != 0 || Kp < ksize * ksize * Cin) return ufail("bad Kp %d", Kp);
    if (ksize != 1 && ksize != 3) return ufail("ksize must be 1 or 3");
    const long M = (long)N * H * W;
    dim3 gr((unsigned)((M + 63) / 64), (unsigned)((Cout + 15) / 16));
    hipStream_t s = (hipStream_t)stream;
    if (ksize == 3)
        hipLaunchKernelGGL(k_conv_igemm<3>, gr, dim3(256), 0, s, (const uint16_t*)x, (const uint16_t*)wt,
                           bias, (uint16_t*)y16, y32, N, H, W, Cin, Cout, Kp, relu);
    else
        hipLaunchKernelGGL(k_conv_igemm<1>, gr, dim3(256), 0, s, (const uint16_t*)x, (const uint16_t*)wt,
                           bias, (uint16_t*)y16, y32, N, H, W, Cin, Cout, Kp, relu);
    ULAUNCH();
    return 0;
}

// part: chunks * Cout * Kp floats of workspace
int jsv_conv_wgrad(const void* x, const void* dy, float* dw, float* db, float* part, int chunks,
                   int N, int H, int W, int Cin, int Cout, int ksize, int Kp, void* stream)
{
    if (Cin % 4 != 0 || Kp % 16 != 0 || chunks < 1) return ufail("bad wgrad arguments");
    const long M = (long)N * H * W;
    long per = (M + chunks - 1) / chunks;
    per = (per + 15) / 16 * 16;
    dim3 gr((unsigned)((Kp + 63) / 64), (unsigned)((Cout + 15) / 16), (unsigned)chunks);
    hipStream_t s = (hipStream_t)stream;
    if (ksize == 3)
        hipLaunchKernelGGL(k_conv_wgrad<3>, gr, dim3(256), 0, s, (const uint16_t*)x, (const uint16_t*)dy,
                           part, N, H, W, Cin, Cout, Kp, per);
    else
        hipLaunchKernelGGL(k_conv_wgrad<1>, gr, dim3(256), 0, s, (const uint16_t*)x, (const uint16_t*)dy,
                           part, N, H, W, Cin, Cout, Kp, per);
    const long n = (long)Cout * Kp;
    hipLaunchKernelGGL(k_reduce_chunks, g1(n), dim3(256), 0, s, part, dw, n, chunks);
    if (db) hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)Cout), dim3(256), 0, s, (const uint16_t*)dy, db, M, Cout);
    ULAUNCH();
    return 0;
}

int jsv_avgpool2(const void* x, void* y, int N, int H, int W, int C, int backward, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (H % 2 || W % 2) return ufail("pooling needs even H, W");
    if (!backward)
        hipLaunchKernelGGL(k_avgpool2, g1((long)N * (H / 2) * (W / 2) * C), dim3(256), 0, s,
                           (const uint16_t*)x, (uint16_t*)y, N, H, W, C);
    else
        hipLaunchKernelGGL(k_avgpool2_bwd, g1((long)N * H * W * C), dim3(256), 0, s,
                           (const uint16_t*)x, (uint16_t*)y, N, H, W, C);
    ULAUNCH();
    return 0;
}

int jsv_up_concat(const void* x, const void* skip, void* y, int N, int H, int W, int C1, int C2,
                  void* stream)
{
    hipLaunchKernelGGL(k_up_concat, g1((long)N * H * W * (C1 + C2)), dim3(256), 0, (hipStream_t)stream,
                       (const uint16_t*)x, (const uint16_t*)skip, (uint16_t*)y, N, H, W, C1, C2);
    ULAUNCH();
    return 0;
}

int jsv_up_concat_bwd(const void* dy, void* dx, void* dskip, int N, int H, int W, int C1, int C2,
                      void* stream)
{
    const long n = (long)N * H * W * (C2 > C1 ? C2 : C1);
    hipLaunchKernelGGL(k_up_concat_bwd, g1(n), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)dy,
                       (uint16_t*)dx, (uint16_t*)dskip, N, H, W, C1, C2);
    ULAUNCH();
    return 0;
}

int jsv_relu_bwd(void* dy, const void* y, int64_t n, void* stream)
{
    hipLaunchKernelGGL(k_relu_bwd, g1(n), dim3(256), 0, (hipStream_t)stream, (uint16_t*)dy,
                       (const uint16_t*)y, (long)n);
    ULAUNCH();
    return 0;
}

int jsv_add_bf16(void* a, const void* b, int64_t n, void* stream)
{
    hipLaunchKernelGGL(k_add_bf16, g1(n), dim3(256), 0, (hipStream_t)stream, (uint16_t*)a,
                       (const uint16_t*)b, (long)n);
    ULAUNCH();
    return 0;
}

int jsv_f32_to_bf16(const float* a, void* b, int64_t n, void* stream)
{
    hipLaunchKernelGGL(k_f32_to_bf16, g1(n), dim3(256), 0, (hipStream_t)stream, a, (uint16_t*)b, (long)n);
    ULAUNCH();
    return 0;
}

int jsv_adam(float* w, const float* g, float* m, float* v, void* w16, int64_t n, float lr, float b1,
             float b2, float eps, int step, float wd, void* stream)
{
    const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, g1(n), dim3(256), 0, (hipStream_t)stream, w, g, m, v, (uint16_t*)w16,
                       (long)n, lr, b1, b2, eps, bc1, bc2, wd);
    ULAUNCH();
    return 0;
}

}  // extern "C"
