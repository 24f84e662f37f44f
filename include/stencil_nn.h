/*
 * stencil_nn.h - C ABI of the learned warm-start kernels (3D U-Net on
 * channels-last BF16 tensors with FP32 accumulation, gfx950 MFMA).
 *
 * No counterpart in mrfarstad/thesis: the learned warm start is an addition
 * of this library.  All functions launch on `stream` and return at once;
 * BF16 tensors are arrays of 16-bit words.  Shapes are (N, D, H, W, C) with
 * C fastest; convolution weights are [Co][Kpad] BF16 with
 * Kpad = st_nn_kpad(Ci) = 27*Ci rounded up to 32 and k = tap*Ci + ci,
 * tap = 9*(dz+1) + 3*(dy+1) + (dx+1); padding columns are zero.
 */
#ifndef THESIS_AMD_STENCIL_NN_H
#define THESIS_AMD_STENCIL_NN_H

#include "stencil_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

int st_nn_kpad(int ci);

/* 3x3x3 convolution, zero padding 1: out[p][co] = bias[co] + sum_k W[co][k] in[p+off][ci],
 * optional ReLU; BF16 and/or FP32 output (NULL to skip), FP32 optionally
 * accumulated.  Ci a multiple of 8, Co a multiple of 4.  With the weights of
 * st_nn_wt_transpose and Ci/Co swapped it is the backward-data pass. */
st_status st_nn_conv3(const void *act, const void *w, const float *bias, void *out_bf16,
                      float *out_f32, int N, int D, int H, int W, int Ci, int Co, int relu,
                      int accumulate, void *stream);
/* dW[co][k] += sum_p dY[p][co] X[p+off(tap)][ci]  (FP32, atomic accumulation).
 * With W a multiple of 8 (and Ci, Co <= 128) both operands are first copied
 * channel-major into the library's workspace (P*(Ci+Co) 16-bit words, kept
 * per device and stream). */
st_status st_nn_conv3_wgrad(const void *x, const void *dy, float *dw, int N, int D, int H, int W,
                            int Ci, int Co, void *stream);
/* backward-data weights: Wt[ci][tap'*Co + co] = W[co][(26 - tap')*Ci + ci] */
st_status st_nn_wt_transpose(const void *w, void *wt, int Ci, int Co, void *stream);
st_status st_nn_f32_to_bf16(const float *a, void *b, long long n, void *stream);
st_status st_nn_bf16_to_f32(const void *a, float *b, long long n, void *stream);
/* out(BF16) = g(FP32) where y(BF16) > 0, else 0 */
st_status st_nn_relu_bwd(const float *g, const void *y, void *out, long long n, void *stream);
/* db[c] += sum_p g[p][c]  (C divides 256) */
st_status st_nn_bias_grad(const void *g, float *db, long long P, int C, void *stream);
/* 2x2x2 average pool (output shape given) and its backward (FP32, accumulates) */
st_status st_nn_pool_fwd(const void *in, void *out, int N, int D, int H, int W, int C, void *stream);
st_status st_nn_pool_bwd(const float *gout, float *gin, int N, int D, int H, int W, int C, void *stream);
/* nearest x2 upsample of a C-channel tensor (input shape given) into channels
 * [c0, c0+C) of a Ct-channel output, and its backward */
st_status st_nn_up_fwd(const void *in, void *out, int N, int D, int H, int W, int C, int Ct, int c0,
                       void *stream);
st_status st_nn_up_bwd(const float *gout, float *gin, int N, int D, int H, int W, int C, int Ct,
                       int c0, void *stream);
/* dst[p][dc0+c] (=|+=) src[p][sc0+c] for c < C (FP32 when is_f32, else BF16) */
st_status st_nn_chan_copy(const void *src, void *dst, long long P, int C, int Cs, int sc0, int Cd,
                          int dc0, int is_f32, int accumulate, void *stream);
/* 1x1 head y[p] = b + sum_c w[c] x[p][c] and its backward (C = 16) */
st_status st_nn_head_fwd(const void *x, const float *w, const float *b, float *y, long long P, int C,
                         void *stream);
st_status st_nn_head_bwd(const void *x, const float *w, const float *g, float *gx, float *dw,
                         float *db, long long P, int C, void *stream);
/* u = base + y on the interior of each n^3 sample (base on the shell) into
 * pred; with a target: loss += sum (u - t)^2 (FP64) and g = 2*scale*(u - t) */
st_status st_nn_mse(const float *base, const float *y, const float *target, float *g, double *loss,
                    float *pred, int N, int n, double scale, void *stream);
/* out = a*x + b*y over n floats */
st_status st_nn_axpby(float *out, float a, const float *x, float b, const float *y, long long n,
                      void *stream);
/* loss += sum x^2 (FP64 accumulation); g = 2*scale*x when g is not NULL */
st_status st_nn_sumsq(const float *x, float *g, double *loss, long long n, double scale, void *stream);
/* Adam step (bias-corrected) with an optional BF16 copy of the weights */
st_status st_nn_adam(float *w, const float *g, float *m, float *v, void *wbf, long long n, float lr,
                     float b1, float b2, float eps, int step, void *stream);
/* input features of N n^3 fields (shell filled): ch0 transfinite blend,
 * ch1 shell value, others 0; base = blend inside, boundary on the shell */
st_status st_nn_features(const float *u, void *feat, float *base, int N, int n, int C, void *stream);
/* trilinear interpolation of a node-centred m^3 field into the interior of an M^3 field */
st_status st_nn_trilinear(const float *in, float *out, int m, int M, void *stream);
/* Dirichlet faces sum_m a sin(pi kx s) sin(pi ky t) for N n^3 fields;
 * modes[(b*6 + face)*M*3 + 3*m + {0,1,2}] = (a, kx, ky), faces x0 x1 y0 y1 z0 z1 */
st_status st_nn_fourier_faces(float *u, const float *modes, int N, int n, int M, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* THESIS_AMD_STENCIL_NN_H */
