// libyapt_denoise: BF16 MFMA kernels of the learned denoiser (a convolutional U-Net over the noisy
// film and first-hit features), forward and backward, plus the element-wise pieces of training.
//
// Layout: activations NHWC in bfloat16 (stored as uint16), weights [Cout][3][3][Cin] in bfloat16
// with an FP32 master copy, accumulation in FP32 on the matrix cores
// (v_mfma_f32_32x32x16_bf16: a 32x16 by 16x32 product per wavefront and instruction).
//
// conv3x3 forward is an implicit GEMM: output [pixels x Cout] = input patches [pixels x 9*Cin] times
// weights [9*Cin x Cout]. A wavefront owns a 32-output-channel x 32-pixel tile and walks K in steps
// of 16 (one tap, 16 input channels). Its operands come straight from global memory in 16-byte
// lane loads (each lane holds 8 consecutive K values of one row or column), which the L1/L2 caches
// share between the wavefronts of a workgroup.
// The data gradient is the same kernel on the gradient with flipped, transposed weights; the weight
// gradient is a GEMM over pixels ([Cout x P] times [P x 9*Cin]) split into pixel chunks whose FP32
// partial tiles are added atomically.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/yapt_denoise.h"

namespace {

thread_local std::string g_err;
int fail(int code, const char* m) {
  g_err = m;
  return code;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even; NaN stays NaN
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

// ---------------------------------------------------------------- conv3x3 forward (implicit GEMM)
// grid: (ceil(P / 32), Cout / 32 / WPB), block: 64 * WPB threads; wave w of a block takes output
// channels [co0 + 32 w, +32). Requirements: Cin % 16 == 0, Cout % 32 == 0.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_conv3x3(const uint16_t* __restrict__ in, int N, int H, int W, int Cin,
                                                      const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                                      int Cout, int relu, uint16_t* __restrict__ out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int co0 = (blockIdx.y * WPB + wave) * 32;
  if (co0 >= Cout) return;
  const long P = (long)N * H * W;
  const long p0 = (long)blockIdx.x * 32;
  const int col = lane & 31, half = lane >> 5;
  // B operand: this lane's pixel
  long p = p0 + col;
  bool pv = p < P;
  int n = 0, y = 0, x = 0;
  if (pv) {
    n = (int)(p / ((long)H * W));
    int r = (int)(p - (long)n * H * W);
    y = r / W;
    x = r - y * W;
  }
  // A operand: this lane's output channel row
  const uint16_t* wrow = w + (size_t)(co0 + col) * 9 * Cin + half * 8;
  f32x16 acc;
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  for (int tap = 0; tap < 9; ++tap) {
    int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
    bool v = pv && yy >= 0 && yy < H && xx >= 0 && xx < W;
    const uint16_t* src = in + (((size_t)n * H + (v ? yy : 0)) * W + (v ? xx : 0)) * Cin + half * 8;
    const uint16_t* wt = wrow + tap * Cin;
    for (int c = 0; c < Cin; c += 16) {
      bf16x8 a = load8(wt + c);
      bf16x8 b = v ? load8(src + c) : zero8();
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
  }
  if (!pv) return;
  // acc[r]: output channel co0 + 8 (r / 4) + 4 half + r % 4 of pixel p
  uint16_t* dst = out + (size_t)p * Cout + co0 + 4 * half;
  for (int g = 0; g < 4; ++g) {
    uint16_t o[4];
    for (int j = 0; j < 4; ++j) {
      int co = co0 + 8 * g + 4 * half + j;
      float v = acc[4 * g + j] + (bias ? bias[co] : 0.f);
      if (relu && v < 0.f) v = 0.f;
      o[j] = f2bf(v);
    }
    uint2 pk = make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
    *reinterpret_cast<uint2*>(dst + 8 * g) = pk;
  }
}

// ---------------------------------------------------------------- weight gradient
// dW[co][tap][ci] += sum over pixels of dOut[p][co] * in[p + tap][ci]: M = Cout, N = 9 Cin, K = pixels.
// grid: (ceil(9 Cin / 32), Cout / 32, pixel chunks), one wave per block; chunk = 16 * steps pixels.
__global__ __launch_bounds__(64) void k_conv3x3_wgrad(const uint16_t* __restrict__ in, const uint16_t* __restrict__ dout,
                                                      int N, int H, int W, int Cin, int Cout, int steps,
                                                      float* __restrict__ dw) {
  const int lane = threadIdx.x, col = lane & 31, half = lane >> 5;
  const int K = 9 * Cin;
  const int k = blockIdx.x * 32 + col;  // B column: (tap, ci)
  const int co0 = blockIdx.y * 32;
  const long P = (long)N * H * W;
  const long pbase = (long)blockIdx.z * 16 * steps;
  const bool kv = k < K;
  const int tap = kv ? k / Cin : 0, ci = kv ? k - tap * Cin : 0;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  f32x16 acc;
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  for (int s = 0; s < steps; ++s) {
    if (pbase + 16 * s >= P) break;  // uniform across the wave: the matrix op needs every lane
    long pb = pbase + 16 * s + 8 * half;
    uint16_t av[8], bv[8];
    for (int j = 0; j < 8; ++j) {
      long p = pb + j;
      uint16_t a = 0, b = 0;
      if (p < P) {
        a = dout[(size_t)p * Cout + co0 + col];
        int n = (int)(p / ((long)H * W));
        int r = (int)(p - (long)n * H * W);
        int yy = r / W + dy, xx = r % W + dx;
        if (kv && yy >= 0 && yy < H && xx >= 0 && xx < W) b = in[(((size_t)n * H + yy) * W + xx) * Cin + ci];
      }
      av[j] = a;
      bv[j] = b;
    }
    uint32_t aw[4], bw[4];
    for (int j = 0; j < 4; ++j) {
      aw[j] = (uint32_t)av[2 * j] | ((uint32_t)av[2 * j + 1] << 16);
      bw[j] = (uint32_t)bv[2 * j] | ((uint32_t)bv[2 * j + 1] << 16);
    }
    uint4 ap = make_uint4(aw[0], aw[1], aw[2], aw[3]);
    uint4 bp = make_uint4(bw[0], bw[1], bw[2], bw[3]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ap), __builtin_bit_cast(bf16x8, bp), acc, 0, 0, 0);
  }
  if (!kv) return;
  for (int r = 0; r < 16; ++r) {
    int co = co0 + 8 * (r / 4) + 4 * half + (r % 4);
    atomicAdd(&dw[(size_t)co * K + k], acc[r]);
  }
}

// dBias[co] += sum over pixels of dOut[p][co]; one thread per channel and pixel chunk
__global__ void k_bias_grad(const uint16_t* dout, long P, int C, int chunk, float* db) {
  int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= C) return;
  long p0 = (long)blockIdx.x * chunk;
  float s = 0.f;
  for (long p = p0; p < p0 + chunk && p < P; ++p) s += bf2f(dout[(size_t)p * C + c]);
  atomicAdd(&db[c], s);
}

// w'[ci][8 - tap][co] = w[co][tap][ci]: the weights of the data-gradient convolution
__global__ void k_flip_weights(const uint16_t* w, int Cout, int Cin, uint16_t* wt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long)Cout * 9 * Cin;
  if (i >= n) return;
  int ci = (int)(i % Cin);
  int tap = (int)((i / Cin) % 9);
  int co = (int)(i / ((long)9 * Cin));
  wt[((size_t)ci * 9 + (8 - tap)) * Cout + co] = w[i];
}

// ---------------------------------------------------------------- element-wise pieces
__global__ void k_relu_grad(const uint16_t* out, uint16_t* g, long n) {  // g *= (out > 0)
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(bf2f(out[i]) > 0.f)) g[i] = 0;
}

__global__ void k_pool2(const uint16_t* in, int N, int H, int W, int C, uint16_t* out) {  // 2x2 average
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int Ho = H / 2, Wo = W / 2;
  long n_ = (long)N * Ho * Wo * C;
  if (i >= n_) return;
  int c = (int)(i % C);
  long q = i / C;
  int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), n = (int)(q / ((long)Wo * Ho));
  const uint16_t* b = in + (((size_t)n * H + 2 * yo) * W + 2 * xo) * C + c;
  float s = bf2f(b[0]) + bf2f(b[C]) + bf2f(b[(size_t)W * C]) + bf2f(b[(size_t)W * C + C]);
  out[i] = f2bf(0.25f * s);
}

__global__ void k_pool2_grad(const uint16_t* gout, int N, int H, int W, int C, uint16_t* gin) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n_ = (long)N * H * W * C;
  if (i >= n_) return;
  int c = (int)(i % C);
  long q = i / C;
  int x = (int)(q % W), y = (int)((q / W) % H), n = (int)(q / ((long)W * H));
  gin[i] = f2bf(0.25f * bf2f(gout[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]));
}

// nearest 2x upsample of `in` (N, H/2, W/2, C1) into channels [0, C1) of `out` (N, H, W, C1 + C2) and
// `skip` (N, H, W, C2) into channels [C1, C1 + C2): the decoder's upsample + concatenation
__global__ void k_up2_concat(const uint16_t* in, int C1, const uint16_t* skip, int C2, int N, int H, int W, uint16_t* out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int C = C1 + C2;
  long n_ = (long)N * H * W * C;
  if (i >= n_) return;
  int c = (int)(i % C);
  long q = i / C;
  int x = (int)(q % W), y = (int)((q / W) % H), n = (int)(q / ((long)W * H));
  if (c < C1) out[i] = in[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C1 + c];
  else out[i] = skip[(size_t)q * C2 + (c - C1)];
}

// backward of k_up2_concat: gin (N, H/2, W/2, C1) = sum of the 4 upsampled gradients; gskip += its slice
__global__ void k_up2_concat_grad(const uint16_t* g, int C1, int C2, int N, int H, int W, uint16_t* gin, uint16_t* gskip,
                                  int accumulateSkip) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int C = C1 + C2;
  int Hi = H / 2, Wi = W / 2;
  long nIn = (long)N * Hi * Wi * C1;
  if (i < nIn) {
    int c = (int)(i % C1);
    long q = i / C1;
    int x = (int)(q % Wi), y = (int)((q / Wi) % Hi), n = (int)(q / ((long)Wi * Hi));
    float s = 0.f;
    for (int d = 0; d < 4; ++d)
      s += bf2f(g[(((size_t)n * H + 2 * y + d / 2) * W + 2 * x + d % 2) * C + c]);
    gin[i] = f2bf(s);
  }
  long nSkip = (long)N * H * W * C2;
  if (i < nSkip) {
    int c = (int)(i % C2);
    long q = i / C2;
    float v = bf2f(g[(size_t)q * C + C1 + c]);
    if (accumulateSkip) v += bf2f(gskip[i]);
    gskip[i] = f2bf(v);
  }
}

__global__ void k_add_bf16(uint16_t* a, const uint16_t* b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// network input: log(1 + max(x, 0)) of the noisy colour, albedo, normal, zero padding up to C
__global__ void k_pack_input(const float* xyzw, const float* features, long P, int C, uint16_t* out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * C) return;
  long p = i / C;
  int c = (int)(i % C);
  float v = 0.f;
  if (c < 3) v = logf(1.f + fmaxf(xyzw[p * 4 + c], 0.f));
  else if (c < 6) v = features[p * 12 + 3 + (c - 3)];        // albedo
  else if (c < 9) v = features[p * 12 + (c - 6)];            // normal
  out[i] = f2bf(v);
}

// prediction = log(1 + noisy) + net[0..2]; loss = mean |prediction - log(1 + reference)| over
// P x 3 values; gradient into channels 0..2 of gnet (others 0); loss partial sums into loss[0]
__global__ void k_l1_log_loss(const uint16_t* net, int C, const float* noisy, const float* ref, long P, uint16_t* gnet,
                              float* loss, float* pred) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (p < P) {
    float inv = 1.f / (3.f * (float)P);
    for (int c = 0; c < C; ++c) {
      if (c < 3) {
        float pr = logf(1.f + fmaxf(noisy[p * 4 + c], 0.f)) + bf2f(net[p * C + c]);
        float t = logf(1.f + fmaxf(ref[p * 4 + c], 0.f));
        float d = pr - t;
        l += fabsf(d);
        if (gnet) gnet[p * C + c] = f2bf(d > 0.f ? inv : (d < 0.f ? -inv : 0.f));
        if (pred) pred[p * 4 + c] = expf(pr) - 1.f;
      } else if (gnet) {
        gnet[p * C + c] = 0;
      }
    }
    if (pred) pred[p * 4 + 3] = 1.f;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l / (3.f * (float)P));
}

// Adam on FP32 master weights; refreshes the BF16 copy
__global__ void k_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, uint16_t* w16) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = f2bf(wi);
}

__global__ void k_f32_to_bf16(const float* a, uint16_t* b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

// the full-resolution shape (N, H, W) of a pooling or upsampling step: positive, H and W even (the
// half-resolution tensor has H/2 x W/2 texels, so an odd size would reach past it)
int check_half_shape(int N, int H, int W, int C) {
  if (N <= 0 || H <= 0 || W <= 0 || C <= 0) return fail(YDN_E_INVALID, "empty shape");
  if ((H | W) & 1) return fail(YDN_E_INVALID, "pooling and upsampling need even H and W");
  return YDN_OK;
}

unsigned blocks(long n, int b) { return (unsigned)((n + b - 1) / b); }

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(YDN_E_DEVICE, hipGetErrorString(e));
  return YDN_OK;
}

}  // namespace

extern "C" {

const char* ydn_last_error(void) { return g_err.c_str(); }

int ydn_conv3x3(const uint16_t* in, int N, int H, int W, int Cin, const uint16_t* w, const float* bias, int Cout, int relu,
                uint16_t* out, void* stream) {
  if (!in || !w || !out) return fail(YDN_E_INVALID, "null tensor");
  if (N <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0) return fail(YDN_E_INVALID, "empty shape");
  if (Cin % 16 != 0) return fail(YDN_E_INVALID, "Cin must be a multiple of 16");
  if (Cout % 32 != 0) return fail(YDN_E_INVALID, "Cout must be a multiple of 32");
  long P = (long)N * H * W;
  hipStream_t s = (hipStream_t)stream;
  if (Cout % 128 == 0)
    hipLaunchKernelGGL(k_conv3x3<4>, dim3(blocks(P, 32), Cout / 128), dim3(256), 0, s, in, N, H, W, Cin, w, bias, Cout, relu, out);
  else if (Cout % 64 == 0)
    hipLaunchKernelGGL(k_conv3x3<2>, dim3(blocks(P, 32), Cout / 64), dim3(128), 0, s, in, N, H, W, Cin, w, bias, Cout, relu, out);
  else
    hipLaunchKernelGGL(k_conv3x3<1>, dim3(blocks(P, 32), Cout / 32), dim3(64), 0, s, in, N, H, W, Cin, w, bias, Cout, relu, out);
  return check_launch();
}

int ydn_conv3x3_wgrad(const uint16_t* in, const uint16_t* dout, int N, int H, int W, int Cin, int Cout, float* dw, float* db,
                      void* stream) {
  if (!in || !dout || !dw) return fail(YDN_E_INVALID, "null tensor");
  if (Cin % 16 != 0 || Cout % 32 != 0) return fail(YDN_E_INVALID, "Cin % 16 and Cout % 32 must be 0");
  long P = (long)N * H * W;
  hipStream_t s = (hipStream_t)stream;
  const int steps = 32;  // 512 pixels per chunk
  long chunks = (P + 16 * steps - 1) / (16 * steps);
  if (chunks > 65535) return fail(YDN_E_INVALID, "too many pixels for one weight-gradient launch");
  hipLaunchKernelGGL(k_conv3x3_wgrad, dim3(blocks(9L * Cin, 32), Cout / 32, (unsigned)chunks), dim3(64), 0, s, in, dout, N, H, W,
                     Cin, Cout, steps, dw);
  if (db) {
    const int chunk = 1024;
    hipLaunchKernelGGL(k_bias_grad, dim3(blocks(P, chunk), blocks(Cout, 64)), dim3(64), 0, s, dout, P, Cout, chunk, db);
  }
  return check_launch();
}

int ydn_flip_weights(const uint16_t* w, int Cout, int Cin, uint16_t* wt, void* stream) {
  long n = (long)Cout * 9 * Cin;
  hipLaunchKernelGGL(k_flip_weights, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, w, Cout, Cin, wt);
  return check_launch();
}

int ydn_relu_grad(const uint16_t* out, uint16_t* g, long n, void* stream) {
  hipLaunchKernelGGL(k_relu_grad, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, out, g, n);
  return check_launch();
}

int ydn_pool2(const uint16_t* in, int N, int H, int W, int C, uint16_t* out, void* stream) {
  if (!in || !out) return fail(YDN_E_INVALID, "null tensor");
  if (int rc = check_half_shape(N, H, W, C)) return rc;
  long n = (long)N * (H / 2) * (W / 2) * C;
  hipLaunchKernelGGL(k_pool2, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, in, N, H, W, C, out);
  return check_launch();
}

int ydn_pool2_grad(const uint16_t* gout, int N, int H, int W, int C, uint16_t* gin, void* stream) {
  if (!gout || !gin) return fail(YDN_E_INVALID, "null tensor");
  if (int rc = check_half_shape(N, H, W, C)) return rc;
  long n = (long)N * H * W * C;
  hipLaunchKernelGGL(k_pool2_grad, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, gout, N, H, W, C, gin);
  return check_launch();
}

int ydn_up2_concat(const uint16_t* in, int C1, const uint16_t* skip, int C2, int N, int H, int W, uint16_t* out, void* stream) {
  if (!in || !skip || !out) return fail(YDN_E_INVALID, "null tensor");
  if (C2 <= 0) return fail(YDN_E_INVALID, "empty shape");
  if (int rc = check_half_shape(N, H, W, C1)) return rc;
  long n = (long)N * H * W * (C1 + C2);
  hipLaunchKernelGGL(k_up2_concat, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, in, C1, skip, C2, N, H, W, out);
  return check_launch();
}

int ydn_up2_concat_grad(const uint16_t* g, int C1, int C2, int N, int H, int W, uint16_t* gin, uint16_t* gskip,
                        int accumulate_skip, void* stream) {
  if (!g || !gin || !gskip) return fail(YDN_E_INVALID, "null tensor");
  if (C2 <= 0) return fail(YDN_E_INVALID, "empty shape");
  if (int rc = check_half_shape(N, H, W, C1)) return rc;
  long n = (long)N * H * W * (C1 > C2 ? C1 : C2);
  hipLaunchKernelGGL(k_up2_concat_grad, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, g, C1, C2, N, H, W, gin,
                     gskip, accumulate_skip);
  return check_launch();
}

int ydn_add(uint16_t* a, const uint16_t* b, long n, void* stream) {
  hipLaunchKernelGGL(k_add_bf16, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, a, b, n);
  return check_launch();
}

int ydn_pack_input(const float* xyzw, const float* features, long P, int C, uint16_t* out, void* stream) {
  if (C < 9) return fail(YDN_E_INVALID, "input needs at least 9 channels");
  hipLaunchKernelGGL(k_pack_input, dim3(blocks(P * C, 256)), dim3(256), 0, (hipStream_t)stream, xyzw, features, P, C, out);
  return check_launch();
}

int ydn_l1_log_loss(const uint16_t* net, int C, const float* noisy, const float* ref, long P, uint16_t* gnet, float* loss,
                    float* pred, void* stream) {
  hipLaunchKernelGGL(k_l1_log_loss, dim3(blocks(P, 256)), dim3(256), 0, (hipStream_t)stream, net, C, noisy, ref, P, gnet, loss,
                     pred);
  return check_launch();
}

int ydn_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps, int step,
             uint16_t* w16, void* stream) {
  float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, b1, b2, eps, c1, c2, w16);
  return check_launch();
}

int ydn_to_bf16(const float* a, uint16_t* b, long n, void* stream) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, a, b, n);
  return check_launch();
}

}  // extern "C"
