// Kernels of the learned warm start: a small fully convolutional U-Net in
// BF16 (weights and activations) with FP32 accumulation on CDNA4 matrix cores.
//
// Activations are NHWC bf16.  A 3x3 (or 1x1) convolution is an implicit GEMM
//   out[m][co] = sum_k A(m, k) * Wt[co][k],   m = (n, h, w),  k = (tap, ci)
// with A gathered on the fly from the zero-padded input (no im2col buffer).
// The forward convolution uses gfx950's v_mfma_f32_16x16x32_bf16
// (__builtin_amdgcn_mfma_f32_16x16x32_bf16), lane layout checked on the GPU:
//   A fragment: lane l holds A[l % 16][8 (l / 16) + i], i = 0..7
//   B fragment: lane l holds B[8 (l / 16) + i][l % 16]
//   C fragment: lane l holds C[4 (l / 16) + i][l % 16], i = 0..3
// one wave computing 16 pixels x every output channel.  The input gradient
// (dgrad) is the same kernel on the output gradient with the weights flipped
// and transposed.  The weight gradient (wgrad) reduces over pixels with the
// same instruction on operands staged transposed in LDS, and writes
// per-chunk partial sums that a second kernel adds in a fixed order.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>

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
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
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
// x: [N][H][W][Cin] bf16; wt: [Cout][Kp] bf16 with Kp = roundup(KS*KS*Cin, 32),
// k = tap * Cin + ci, tap = (dy + 1) * KS + (dx + 1); zero padded beyond K.
// out: bf16 [M][Cout] (y16) and/or fp32 (y32); bias fp32 or null; relu 0/1.
// One wave computes 16 pixels x all Cout channels (NCT tiles of 16): the A
// fragment (the pixels' input patch, gathered once per K step) feeds NCT
// v_mfma_f32_16x16x32_bf16, each with its own weight fragment.  A lane's 8
// consecutive k are two groups of 4 (Cin is a multiple of 4, so a group never
// straddles a tap): two 8-byte loads.  The (tap, ci) of a lane's first k
// advance by 32 per step without divisions.
// grid: (ceil(M / 64)); block 256 = 4 waves x 16 pixels.
// ---------------------------------------------------------------------------
// four consecutive input channels ci..ci+3 of one tap as raw bf16 bits
template <int KS>
__device__ __forceinline__ u32x2 patch4(const uint16_t* __restrict__ x, bool valid, int tap, int ci, int an,
                                        int ah, int aw, int H, int W, int Cin)
{
    u32x2 v = {0u, 0u};
    if (!valid || tap >= KS * KS) return v;
    const int ty = (KS == 3) ? (tap >= 6 ? 1 : (tap >= 3 ? 0 : -1)) : 0;
    const int tx = (KS == 3) ? tap - 3 * (ty + 1) - 1 : 0;
    const int hh = ah + ty, ww = aw + tx;
    if (hh >= 0 && hh < H && ww >= 0 && ww < W)
        v = *reinterpret_cast<const u32x2*>(x + (((long)an * H + hh) * W + ww) * Cin + ci);
    return v;
}

template <int KS, int NCT>
__global__ void __launch_bounds__(256)
k_conv_igemm(const uint16_t* __restrict__ x, const uint16_t* __restrict__ wt,
             const float* __restrict__ bias, uint16_t* __restrict__ y16, float* __restrict__ y32,
             int N, int H, int W, int Cin, int Cout, int Kp, int relu)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int M = N * H * W;
    const int m0 = (blockIdx.x * 4 + wave) * 16;
    if (m0 >= M) return;
    const int r = lane & 15;
    const int q = lane >> 4;
    const int ma = m0 + r;                     // pixel of this lane's A row
    const bool avalid = ma < M;
    const int aw = ma % W, ah = (ma / W) % H, an = ma / (W * H);
    // (tap, ci) of k = kb + 8q, stepped with kb
    int tap = (8 * q) / Cin;
    int ci = 8 * q - tap * Cin;
    f32x4 acc[NCT];
#pragma unroll
    for (int t = 0; t < NCT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < Kp; kb += 32) {
        const bool second_next = ci + 4 >= Cin;
        const u32x2 lo = patch4<KS>(x, avalid, tap, ci, an, ah, aw, H, W, Cin);
        const u32x2 hi = patch4<KS>(x, avalid, second_next ? tap + 1 : tap, second_next ? ci + 4 - Cin : ci + 4,
                                    an, ah, aw, H, W, Cin);
        const u32x4 raw = {lo.x, lo.y, hi.x, hi.y};
        const bf16x8 a = __builtin_bit_cast(bf16x8, raw);
        const int k = kb + 8 * q;
#pragma unroll
        for (int t = 0; t < NCT; ++t) {
            const int co = 16 * t + r;
            bf16x8 b = {};
            if (co < Cout) b = *reinterpret_cast<const bf16x8*>(wt + (long)co * Kp + k);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
        ci += 32;
        while (ci >= Cin) {
            ci -= Cin;
            ++tap;
        }
    }
#pragma unroll
    for (int t = 0; t < NCT; ++t) {
        const int co = 16 * t + r;
        if (co >= Cout) continue;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = m0 + 4 * q + i;
            if (m >= M) continue;
            float v = acc[t][i] + bv;
            if (relu) v = v > 0.f ? v : 0.f;
            if (y16) y16[(long)m * Cout + co] = f2bf(v);
            if (y32) y32[(long)m * Cout + co] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[co][tap * Cin + ci] = sum_m dY[m][co] X(m shifted by tap)[ci]
// A GEMM whose reduction runs over pixels, while both operands are stored
// channel-contiguous per pixel, so each workgroup stages WG_PIX pixels of dY
// and of the tap-shifted input in LDS transposed ([channel][pixel]); a
// fragment of 8 consecutive pixels is then one 16-byte LDS read, and
// v_mfma_f32_16x16x32_bf16 reduces 32 pixels per instruction.
// grid: (KS * KS taps, chunks); block 256 = 4 waves; the Cout x Cin tile of
// the tap is split into 16 x 16 MFMA tiles spread over the waves (at most
// WG_TILES per wave).  part: [chunks][Cout][Kp] fp32 partial sums; the
// workgroups of tap 0 also write the zero padding k >= KS*KS*Cin and the
// chunk's bias partials: [chunks][Cout] after the weight partials.
// ---------------------------------------------------------------------------
constexpr int WG_TILES = 8;         // MFMA tiles per wave (Cout/16 * Cin/16 <= 32)

// v[j] holds channels c..c+3 of pixel j (bf16 pairs in two words); writes
// row c + i, pixels 0..3, for i = 0..3 (rows LD apart), 8 bytes each
template <int LD>
__device__ __forceinline__ void store_t4(uint16_t* dst, const u32x2 v[4])
{
    // channel (low or high half of word w) of pixels j and j + 1 in one word
    auto lo2 = [](unsigned a, unsigned b) { return (a & 0xffffu) | (b << 16); };
    auto hi2 = [](unsigned a, unsigned b) { return (a >> 16) | (b & 0xffff0000u); };
    const u32x2 r0 = {lo2(v[0].x, v[1].x), lo2(v[2].x, v[3].x)};
    const u32x2 r1 = {hi2(v[0].x, v[1].x), hi2(v[2].x, v[3].x)};
    const u32x2 r2 = {lo2(v[0].y, v[1].y), lo2(v[2].y, v[3].y)};
    const u32x2 r3 = {hi2(v[0].y, v[1].y), hi2(v[2].y, v[3].y)};
    *reinterpret_cast<u32x2*>(dst) = r0;
    *reinterpret_cast<u32x2*>(dst + LD) = r1;
    *reinterpret_cast<u32x2*>(dst + 2 * LD) = r2;
    *reinterpret_cast<u32x2*>(dst + 3 * LD) = r3;
}

template <int KS, int WG_PIX>
__global__ void __launch_bounds__(256)
k_conv_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
             float* __restrict__ part, int N, int H, int W, int Cin, int Cout, int Kp, int per)
{
    constexpr int WG_LD = WG_PIX + 8;   // LDS row (bf16): 16-byte aligned, rows offset in banks
    // dynamic LDS: [16 cot][WG_LD] dY, [16 cit][WG_LD] input, [WG_PIX] source pixels
    extern __shared__ __attribute__((aligned(16))) uint16_t u_smem[];
    uint16_t* sD = u_smem;
    uint16_t* sX = sD + 16 * ((Cout + 15) / 16) * WG_LD;
    int* sSrc = reinterpret_cast<int*>(sX + 16 * ((Cin + 15) / 16) * WG_LD);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int tap = blockIdx.x;
    const int ty = (KS == 3) ? tap / 3 - 1 : 0, tx = (KS == 3) ? tap % 3 - 1 : 0;
    const int M = N * H * W;
    const int p_begin = blockIdx.y * per;
    int p_end = p_begin + per;
    if (p_end > M) p_end = M;
    const int cot = (Cout + 15) / 16, cit = (Cin + 15) / 16;
    const int ntile = cot * cit;
    f32x4 acc[WG_TILES];
#pragma unroll
    for (int j = 0; j < WG_TILES; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // tap 0 also sums dY over its pixels (the bias gradient): dY^T times a
    // column of ones, channel tile `wave`
    const bool bias_wave = tap == 0 && wave < cot;
    f32x4 accb = {0.f, 0.f, 0.f, 0.f};
    const u32x4 ones_raw = {0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u};   // bf16 1.0
    const bf16x8 ones = __builtin_bit_cast(bf16x8, ones_raw);
    const int c4o = Cout / 4, c4i = Cin / 4;
    for (int p0 = p_begin; p0 < p_end; p0 += WG_PIX) {
        __syncthreads();                      // the previous step's fragments are read
        if (tid < WG_PIX) {
            const int m = p0 + tid;
            int src = -1;
            if (m < p_end) {
                const int w = m % W, h = (m / W) % H, n = m / (W * H);
                const int hh = h + ty, ww = w + tx;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W) src = (n * H + hh) * W + ww;
            }
            sSrc[tid] = src;
        }
        // dY rows of the step stored transposed: an item is 4 pixels x 4
        // channels (four 8-byte loads, four 8-byte LDS stores)
        for (int it = tid; it < (WG_PIX / 4) * c4o; it += 256) {
            const int pq = it / c4o, c = 4 * (it - pq * c4o);
            u32x2 v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = p0 + 4 * pq + j;
                v[j] = u32x2{0u, 0u};
                if (m < p_end) v[j] = *reinterpret_cast<const u32x2*>(dy + (long)m * Cout + c);
            }
            store_t4<WG_LD>(sD + c * WG_LD + 4 * pq, v);
        }
        // padding rows of a partial channel tile hold zeros
        for (int it = tid; it < (16 * cot - Cout) * WG_PIX; it += 256)
            sD[(Cout + it / WG_PIX) * WG_LD + it % WG_PIX] = 0;
        __syncthreads();                      // sSrc ready
        for (int it = tid; it < (WG_PIX / 4) * c4i; it += 256) {
            const int pq = it / c4i, c = 4 * (it - pq * c4i);
            u32x2 v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int src = sSrc[4 * pq + j];
                v[j] = u32x2{0u, 0u};
                if (src >= 0) v[j] = *reinterpret_cast<const u32x2*>(x + (long)src * Cin + c);
            }
            store_t4<WG_LD>(sX + c * WG_LD + 4 * pq, v);
        }
        for (int it = tid; it < (16 * cit - Cin) * WG_PIX; it += 256)
            sX[(Cin + it / WG_PIX) * WG_LD + it % WG_PIX] = 0;
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < WG_PIX; kk += 32) {
            if (bias_wave) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(sD + (16 * wave + r) * WG_LD + kk + 8 * q);
                accb = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, ones, accb, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < WG_TILES; ++j) {
                const int t = wave + 4 * j;
                if (t < ntile) {
                    const int ct = t / cit, ctile = t - ct * cit;
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(sD + (16 * ct + r) * WG_LD + kk + 8 * q);
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(sX + (16 * ctile + r) * WG_LD + kk + 8 * q);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
                }
            }
        }
    }
    float* out = part + (long)blockIdx.y * Cout * Kp;
#pragma unroll
    for (int j = 0; j < WG_TILES; ++j) {
        const int t = wave + 4 * j;
        if (t >= ntile) continue;
        const int ct = t / cit, ctile = t - ct * cit;
        const int ci = 16 * ctile + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = 16 * ct + 4 * q + i;
            if (co < Cout && ci < Cin) out[(long)co * Kp + tap * Cin + ci] = acc[j][i];
        }
    }
    if (bias_wave && r == 0) {
        // every column of the product holds the channel sums: lane r = 0 writes
        float* pb = part + (long)gridDim.y * Cout * Kp + (long)blockIdx.y * Cout;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = 16 * wave + 4 * q + i;
            if (co < Cout) pb[co] = accb[i];
        }
    }
    if (tap == 0) {
        const int K = KS * KS * Cin;
        for (int it = tid; it < Cout * (Kp - K); it += 256) {
            const int co = it / (Kp - K);
            out[(long)co * Kp + K + (it - co * (Kp - K))] = 0.f;
        }
    }
}

// sum partials over chunks (fixed order) into dw (and the bias partials into db)
// few outputs, many chunks: one workgroup per output, a fixed-order tree
__global__ void __launch_bounds__(256)
k_reduce_chunks_tree(const float* __restrict__ part, float* __restrict__ dw, long n, int chunks)
{
    __shared__ float sh[256];
    const long i = blockIdx.x;
    float v = 0.f;
    for (int c = threadIdx.x; c < chunks; c += 256) v += part[(long)c * n + i];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) dw[i] = sh[0];
}

__global__ void __launch_bounds__(256)
k_reduce_chunks(const float* __restrict__ part, float* __restrict__ dw, long n, int chunks)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int c = 0; c < chunks; ++c) s += part[(long)c * n + i];
    dw[i] = s;
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
    if (Kp % 32 != 0 || Kp < ksize * ksize * Cin) return ufail("bad Kp %d (a multiple of 32 >= K)", Kp);
    if (ksize != 1 && ksize != 3) return ufail("ksize must be 1 or 3");
    if (Cout < 1 || Cout > 128) return ufail("Cout must be in [1, 128] (got %d)", Cout);
    const long M = (long)N * H * W;
    if (M > (1L << 30)) return ufail("too many pixels (%ld)", M);
    if ((((uintptr_t)wt) & 15) != 0 || (((uintptr_t)x) & 7) != 0) return ufail("misaligned conv operands");
    const dim3 gr((unsigned)((M + 63) / 64));
    hipStream_t s = (hipStream_t)stream;
    const int nct = (Cout + 15) / 16;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, gr, dim3(256), 0, s, (const uint16_t*)x, (const uint16_t*)wt, bias,
                           (uint16_t*)y16, y32, N, H, W, Cin, Cout, Kp, relu);
    };
    auto pick = [&](auto ks) {
        constexpr int KS = decltype(ks)::value;
        switch (nct) {
        case 1: go(k_conv_igemm<KS, 1>); break;
        case 2: go(k_conv_igemm<KS, 2>); break;
        case 3: go(k_conv_igemm<KS, 3>); break;
        case 4: go(k_conv_igemm<KS, 4>); break;
        case 5: case 6: go(k_conv_igemm<KS, 6>); break;
        default: go(k_conv_igemm<KS, 8>); break;
        }
    };
    if (ksize == 3) pick(std::integral_constant<int, 3>{});
    else pick(std::integral_constant<int, 1>{});
    ULAUNCH();
    return 0;
}

// part: chunks * Cout * Kp floats of workspace
int jsv_conv_wgrad(const void* x, const void* dy, float* dw, float* db, float* part, int chunks,
                   int N, int H, int W, int Cin, int Cout, int ksize, int Kp, void* stream)
{
    if (Cin % 4 != 0 || Kp % 32 != 0 || chunks < 1) return ufail("bad wgrad arguments");
    if (ksize != 1 && ksize != 3) return ufail("ksize must be 1 or 3");
    if (Cout < 1 || Cout > 64 || Cout % 4 != 0 || Cin > 128 || Kp < ksize * ksize * Cin)
        return ufail("wgrad needs Cout in [4, 64] and Cin <= 128, multiples of 4 (got %d, %d)", Cout, Cin);
    if (((Cout + 15) / 16) * ((Cin + 15) / 16) > 4 * WG_TILES) return ufail("wgrad tile count too large");
    if ((((uintptr_t)x) & 7) != 0 || (((uintptr_t)dy) & 7) != 0) return ufail("misaligned wgrad operands");
    const long M = (long)N * H * W;
    if (M > (1L << 30)) return ufail("too many pixels (%ld)", M);
    const long per = (M + chunks - 1) / chunks;
    const dim3 gr((unsigned)(ksize * ksize), (unsigned)chunks);
    hipStream_t s = (hipStream_t)stream;
    // LDS sized by the layer's channel tiles, so narrow layers fit more
    // workgroups per CU (their steps are latency-bound)
    constexpr int pix = 64;
    const int rows = 16 * ((Cout + 15) / 16) + 16 * ((Cin + 15) / 16);
    const size_t lds = (size_t)rows * (pix + 8) * 2 + (size_t)pix * 4;
    if (ksize == 3)
        hipLaunchKernelGGL((k_conv_wgrad<3, pix>), gr, dim3(256), lds, s, (const uint16_t*)x, (const uint16_t*)dy,
                           part, N, H, W, Cin, Cout, Kp, (int)per);
    else
        hipLaunchKernelGGL((k_conv_wgrad<1, pix>), gr, dim3(256), lds, s, (const uint16_t*)x, (const uint16_t*)dy,
                           part, N, H, W, Cin, Cout, Kp, (int)per);
    const long n = (long)Cout * Kp;
    // few outputs: a workgroup each (the chunk loop would be one long serial chain)
    auto reduce = [&](const float* src, float* dst, long cnt) {
        if (cnt < 16384)
            hipLaunchKernelGGL(k_reduce_chunks_tree, dim3((unsigned)cnt), dim3(256), 0, s, src, dst, cnt, chunks);
        else
            hipLaunchKernelGGL(k_reduce_chunks, g1(cnt), dim3(256), 0, s, src, dst, cnt, chunks);
    };
    reduce(part, dw, n);
    if (db) reduce(part + (long)chunks * n, db, (long)Cout);
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
