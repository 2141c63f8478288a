/* libyapt_denoise: BF16 matrix-core kernels of the learned denoiser (no counterpart in the reference,
 * whose denoiser is a hand-written bilateral filter, shaders/postprocess/denoise.comp).
 *
 * Tensors are device pointers. Activations are NHWC bfloat16 (uint16 storage); conv weights are
 * [Cout][3][3][Cin] bfloat16; gradients of weights and biases are FP32. Every function returns 0 or a
 * negative YDN_E_* code (ydn_last_error() describes it) and enqueues its work on `stream`
 * (a hipStream_t; NULL = the default stream). */
#ifndef YAPT_DENOISE_H
#define YAPT_DENOISE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define YDN_OK 0
#define YDN_E_INVALID (-1)
#define YDN_E_DEVICE (-2)

/* out = act(conv3x3(in, w) + bias), zero padding, stride 1; act = ReLU when relu != 0.
 * Cin % 16 == 0, Cout % 32 == 0, Cout <= 256; bias may be NULL. */
int ydn_conv3x3(const uint16_t* in, int N, int H, int W, int Cin, const uint16_t* w, const float* bias, int Cout, int relu,
                uint16_t* out, void* stream);
/* ydn_conv3x3 that also writes pooled[N][H/2][W/2][Cout] = ydn_pool2(out), from the tile in registers
 * (the same values as the two calls; H and W even) */
int ydn_conv3x3_pool2(const uint16_t* in, int N, int H, int W, int Cin, const uint16_t* w, const float* bias, int Cout,
                      int relu, uint16_t* out, uint16_t* pooled, void* stream);
/* the conv of the concatenation [nearest 2x upsample of coarse[N][H/2][W/2][C1], skip[N][H][W][C2]] with
 * w[Cout][3][3][C1 + C2]: ydn_up2_concat followed by ydn_conv3x3, without the concatenated tensor.
 * H and W even, C1 and C2 multiples of 16. */
int ydn_conv3x3_up2cat(const uint16_t* coarse, int C1, const uint16_t* skip, int C2, int N, int H, int W, const uint16_t* w,
                       const float* bias, int Cout, int relu, uint16_t* out, void* stream);
/* dw[Cout][9][Cin] += sum over pixels of dout x input patches; db[Cout] += sum of dout (db may be NULL) */
int ydn_conv3x3_wgrad(const uint16_t* in, const uint16_t* dout, int N, int H, int W, int Cin, int Cout, float* dw, float* db,
                      void* stream);
/* wt[Cin][3][3][Cout] = w[Cout][2-ky][2-kx][Cin]: ydn_conv3x3(dout, wt) is the input gradient */
int ydn_flip_weights(const uint16_t* w, int Cout, int Cin, uint16_t* wt, void* stream);
/* g[i] = 0 where out[i] <= 0 (ReLU backward on the layer's output) */
int ydn_relu_grad(const uint16_t* out, uint16_t* g, long n, void* stream);
/* 2x2 average pooling and its backward */
int ydn_pool2(const uint16_t* in, int N, int H, int W, int C, uint16_t* out, void* stream);
int ydn_pool2_grad(const uint16_t* gout, int N, int H, int W, int C, uint16_t* gin, void* stream);
/* out[N][H][W][C1 + C2] = (nearest 2x upsample of in[N][H/2][W/2][C1], skip[N][H][W][C2]) and its
 * backward (gskip = or += the skip slice of g) */
int ydn_up2_concat(const uint16_t* in, int C1, const uint16_t* skip, int C2, int N, int H, int W, uint16_t* out, void* stream);
int ydn_up2_concat_grad(const uint16_t* g, int C1, int C2, int N, int H, int W, uint16_t* gin, uint16_t* gskip,
                        int accumulate_skip, void* stream);
int ydn_add(uint16_t* a, const uint16_t* b, long n, void* stream);
/* network input from the film (float4 per pixel) and first-hit features (12 floats per pixel):
 * log(1 + colour), albedo, normal, zeros up to C channels */
int ydn_pack_input(const float* xyzw, const float* features, long P, int C, uint16_t* out, void* stream);
/* prediction = log(1 + noisy) + net[:, 0:3]; loss += mean |prediction - log(1 + ref)|; gnet = its
 * gradient (NULL: none); pred = exp(prediction) - 1 as float4 per pixel (NULL: none) */
int ydn_l1_log_loss(const uint16_t* net, int C, const float* noisy, const float* ref, long P, uint16_t* gnet, float* loss,
                    float* pred, void* stream);
/* Adam step `step` (from 1) on FP32 weights w with gradient g and moments m, v; refreshes w16 (may be NULL) */
int ydn_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps, int step,
             uint16_t* w16, void* stream);
int ydn_to_bf16(const float* a, uint16_t* b, long n, void* stream);
const char* ydn_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
