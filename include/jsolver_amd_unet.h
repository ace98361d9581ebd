/*
 * jsolver_amd learned warm start: kernels of a small BF16 U-Net (NHWC
 * activations, FP32 accumulation on MFMA).  No counterpart in the reference
 * (jsolver has no learned component); they serve the warm-start guess that
 * enters solve_2d through its existing `phi` argument (solver_2d.py:212).
 * Pointers are device pointers; bf16 buffers are passed as void*.
 * Every function returns 0 or nonzero; jsv_unet_last_error() gives the message.
 */
#ifndef JSOLVER_AMD_UNET_H
#define JSOLVER_AMD_UNET_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char *jsv_unet_last_error(void);
/* y = act(conv(x, wt) + bias), 3x3 (padding 1) or 1x1; x [N][H][W][Cin] bf16,
 * wt [Cout][Kp] bf16 (k = tap * Cin + ci, zero padded to Kp, a multiple of 32),
 * bias fp32 (nullable); outputs bf16 y16 and/or fp32 y32 [N][H][W][Cout];
 * Cin a multiple of 4, Cout <= 128. */
int jsv_conv_fwd(const void *x, const void *wt, const float *bias, void *y16, float *y32, int N,
                 int H, int W, int Cin, int Cout, int ksize, int Kp, int relu, void *stream);
/* dw [Cout][Kp] = sum over pixels of dy x im2col(x); db [Cout] = sum dy (nullable);
 * part: chunks * Cout * (Kp + 1) floats of workspace (deterministic two-pass
 * sum); Cout in [4, 64] and Cin <= 128, multiples of 4. */
int jsv_conv_wgrad(const void *x, const void *dy, float *dw, float *db, float *part, int chunks,
                   int N, int H, int W, int Cin, int Cout, int ksize, int Kp, void *stream);
/* 2x2 average pooling (backward = 1: x is dy of the pooled map, y is dx) */
int jsv_avgpool2(const void *x, void *y, int N, int H, int W, int C, int backward, void *stream);
/* y[N][H][W][C1+C2] = nearest-2x-upsample(x[N][H/2][W/2][C1]) ++ skip[N][H][W][C2] */
int jsv_up_concat(const void *x, const void *skip, void *y, int N, int H, int W, int C1, int C2,
                  void *stream);
int jsv_up_concat_bwd(const void *dy, void *dx, void *dskip, int N, int H, int W, int C1, int C2,
                      void *stream);
int jsv_relu_bwd(void *dy, const void *y, int64_t n, void *stream);
int jsv_add_bf16(void *a, const void *b, int64_t n, void *stream);
int jsv_f32_to_bf16(const float *a, void *b, int64_t n, void *stream);
/* Adam on fp32 master weights, refreshing the bf16 copy w16 (nullable) */
int jsv_adam(float *w, const float *g, float *m, float *v, void *w16, int64_t n, float lr, float b1,
             float b2, float eps, int step, float wd, void *stream);

#ifdef __cplusplus
}
#endif
#endif
