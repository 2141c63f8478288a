// libyapt_denoise: BF16 MFMA kernels of the learned denoiser (a convolutional U-Net over the noisy
// film and first-hit features), forward and backward, plus the element-wise pieces of training.
//
// Layout: activations NHWC in bfloat16 (stored as uint16), weights [Cout][3][3][Cin] in bfloat16
// with an FP32 master copy, accumulation in FP32 on the matrix cores
// (v_mfma_f32_32x32x16_bf16: a 32x16 by 16x32 product per wavefront and instruction).
//
// conv3x3 forward is an implicit GEMM: output [pixels x Cout] = input patches [pixels x 9*Cin] times
// weights [9*Cin x Cout], tiled through LDS (k_conv3x3: a workgroup stages an input patch with its halo
// and the tile's weights, chunk by chunk of input channels, and its wavefronts multiply from LDS). The
// decoder's nearest-neighbour upsample and channel concatenation happen in that staging
// (ydn_conv3x3_up2cat), so the concatenated tensor is never written.
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
// The input of a conv: one NHWC tensor, or the channel concatenation of a half-resolution tensor
// upsampled 2x by nearest neighbour (channels [0, Ca)) and a full-resolution one (channels [Ca, Ca + Cb)),
// which fuses the decoder's upsample-and-concatenate into the conv's staging.
struct ConvIn {
  const uint16_t* a;
  int Ca;
  const uint16_t* b;  // null: the input is `a` alone
  int Cb;
  int upA;            // a is at half resolution
};

// the value of the neighbouring lane (lane ^ 1): a DPP quad permutation [1, 0, 3, 2] in the VALU,
// where a shuffle would go through the LDS crossbar. Both lanes of the pair must be active (the
// pooled epilogue runs with the whole wavefront).
__device__ __forceinline__ float xor1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

__device__ __forceinline__ bf16x8 lds8(const uint8_t* p) {
  uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// A workgroup of four wavefronts computes tiles of TH = 4 RPW output rows x 32 columns x 32 COT output
// channels of one image, as an implicit GEMM on the matrix cores: C[co][pixel] += W[co][(tap, ci)] x
// patch[(tap, ci)][pixel]. Input channels go in chunks of CIC: the chunk's (TH + 2) x 34 patch with its
// halo (zeros outside the image, the conv's zero padding) and the tile's weights for the chunk are staged
// in LDS; then each wavefront runs 9 taps x CIC / 16 k-steps of v_mfma_f32_32x32x16_bf16 on its RPW x COT
// accumulators (RPW output rows, COT channel tiles), with operands read from LDS: a lane holds 8
// consecutive k of one weight row (A) or one pixel (B). Staged pixels and weight rows are padded by 16
// bytes, so the 16-byte reads of 16 consecutive lanes fall in distinct LDS banks.
// The grid is persistent: a workgroup walks the (tile, chunk) steps of its jobs (job = tile + tiles
// (n + N co-block), strided by the grid), and the 16-byte global loads of the next step are issued into
// registers before the current step's matrix work, so their latency hides behind it. When the input has
// one chunk, the weights stay in LDS from one job to the next of the same channel block.
// Requirements: Cin % 16 == 0 (CIC divides Cin), Ca % 8 == 0, Cout % 32 == 0. block 256.
constexpr int CONV_MAX_COUT = 256;  // output channels of one conv (the bias staged in LDS)

template <int CIC, int COT, int RPW, int OCC>
__global__ __launch_bounds__(256, OCC) void k_conv3x3(ConvIn src, int N, int H, int W, const uint16_t* __restrict__ w,
                                                    const float* __restrict__ bias, int Cout, int relu,
                                                    uint16_t* __restrict__ out, uint16_t* __restrict__ pooled) {
  static_assert(RPW % 2 == 0, "the pooled epilogue pairs the output rows of a wavefront");
  constexpr int TH = 4 * RPW, TW = 32, PW = TW + 2, PH = TH + 2;
  constexpr int PIX = CIC * 2 + 16;       // LDS bytes per staged pixel
  constexpr int WROW = 9 * CIC * 2 + 16;  // LDS bytes per staged output channel
  constexpr int VPP = CIC / 8;            // 16-byte vectors per pixel and chunk
  constexpr int NVI = PH * PW * VPP;      // input vectors per step
  constexpr int NVW = 32 * COT * 9 * VPP; // weight vectors per step
  constexpr int NIN = (NVI + 255) / 256, NW = (NVW + 255) / 256;
  __shared__ __attribute__((aligned(16))) uint8_t sIn[PH * PW * PIX];
  __shared__ __attribute__((aligned(16))) uint8_t sW[32 * COT * WROW];
  // the bias, read by every epilogue: staged once (the first step's barrier orders it). The
  // epilogue then issues no global load, whose wait (the vector-memory counter drains in order)
  // would also wait for the next step's staging loads issued before the matrix work.
  __shared__ float sBias[CONV_MAX_COUT];
  for (int c = threadIdx.x; c < CONV_MAX_COUT; c += 256) sBias[c] = (bias && c < Cout) ? bias[c] : 0.f;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, col = lane & 31, half = lane >> 5;
  const int tilesX = (W + TW - 1) / TW, tiles = tilesX * ((H + TH - 1) / TH);
  const int Cin = src.Ca + src.Cb, nChunks = Cin / CIC;
  const long jobs = (long)tiles * N * ((Cout + 32 * COT - 1) / (32 * COT));
  const int Ha = src.upA ? H >> 1 : H, Wa = src.upA ? W >> 1 : W;

  long job = blockIdx.x;
  if (job >= jobs) return;
  int chunk = 0;
  int tx0, ty0, n, co0;
  auto decode = [&](long j, int& x0, int& y0, int& img, int& c0) {
    const int tile = (int)(j % tiles);
    const long rest = j / tiles;
    img = (int)(rest % N);
    c0 = (int)(rest / N) * 32 * COT;
    x0 = (tile % tilesX) * TW;
    y0 = (tile / tilesX) * TH;
  };
  uint4 pin[NIN], pw[NW];
  auto load_in = [&](int x0, int y0, int img, int ci0) {
#pragma unroll
    for (int k = 0; k < NIN; ++k) {
      const int i = tid + 256 * k;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (i < NVI) {
        const int pix = i / VPP, part = i - pix * VPP;
        const int py = pix / PW, px = pix - py * PW;
        const int gy = y0 + py - 1, gx = x0 + px - 1;
        if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
          const int c = ci0 + part * 8;
          const uint16_t* q;
          if (c < src.Ca) {
            const int ay = src.upA ? gy >> 1 : gy, ax = src.upA ? gx >> 1 : gx;
            q = src.a + (((size_t)img * Ha + ay) * Wa + ax) * src.Ca + c;
          } else {
            q = src.b + (((size_t)img * H + gy) * W + gx) * src.Cb + (c - src.Ca);
          }
          v = *reinterpret_cast<const uint4*>(q);
        }
      }
      pin[k] = v;
    }
  };
  auto load_w = [&](int c0, int ci0) {
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int i = tid + 256 * k;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (i < NVW) {
        const int co = i / (9 * VPP), rem = i - co * (9 * VPP);
        const int tap = rem / VPP, part = rem - tap * VPP;
        if (c0 + co < Cout) v = *reinterpret_cast<const uint4*>(w + ((size_t)(c0 + co) * 9 + tap) * Cin + ci0 + part * 8);
      }
      pw[k] = v;
    }
  };

  decode(job, tx0, ty0, n, co0);
  load_in(tx0, ty0, n, 0);
  load_w(co0, 0);
  bool stageW = true;
  int wCo = co0;  // channel block of the weights last loaded
  f32x16 acc[RPW][COT];
  for (int r = 0; r < RPW; ++r)
    for (int t = 0; t < COT; ++t)
      for (int i = 0; i < 16; ++i) acc[r][t][i] = 0.f;
  while (true) {
    __syncthreads();  // every wavefront is done reading the previous step
#pragma unroll
    for (int k = 0; k < NIN; ++k) {
      const int i = tid + 256 * k;
      if (i < NVI) {
        const int pix = i / VPP, part = i - pix * VPP;
        *reinterpret_cast<uint4*>(sIn + pix * PIX + part * 16) = pin[k];
      }
    }
    if (stageW) {
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        const int i = tid + 256 * k;
        if (i < NVW) {
          const int co = i / (9 * VPP), rem = i - co * (9 * VPP);
          const int tap = rem / VPP, part = rem - tap * VPP;
          *reinterpret_cast<uint4*>(sW + co * WROW + tap * (CIC * 2) + part * 16) = pw[k];
        }
      }
    }
    __syncthreads();
    // the next step's loads, in flight during this step's matrix work
    long nJob = job;
    int nChunk = chunk + 1;
    if (nChunk == nChunks) {
      nChunk = 0;
      nJob = job + gridDim.x;
    }
    const bool more = nJob < jobs;
    int nx0 = 0, ny0 = 0, nn = 0, nco = 0;
    bool nStageW = false;
    if (more) {
      decode(nJob, nx0, ny0, nn, nco);
      load_in(nx0, ny0, nn, nChunk * CIC);
      nStageW = !(nChunks == 1 && nco == wCo);
      if (nStageW) {
        load_w(nco, nChunk * CIC);
        wCo = nco;
      }
    }
#pragma unroll 1
    for (int tap = 0; tap < 9; ++tap) {
      const int ky = tap / 3, kx = tap % 3;
#pragma unroll
      for (int kk = 0; kk < CIC / 16; ++kk) {
        bf16x8 a[COT];
#pragma unroll
        for (int t = 0; t < COT; ++t) a[t] = lds8(sW + (t * 32 + col) * WROW + tap * (CIC * 2) + kk * 32 + half * 16);
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
          const int row = wave * RPW + r;
          bf16x8 b = lds8(sIn + ((row + ky) * PW + col + kx) * PIX + kk * 32 + half * 16);
#pragma unroll
          for (int t = 0; t < COT; ++t) acc[r][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[t], b, acc[r][t], 0, 0, 0);
        }
      }
    }
    if (chunk == nChunks - 1) {
      // acc[r][t][i]: output channel co0 + 32 t + 8 (i / 4) + 4 half + i % 4 of pixel (row r of this wave, column col).
      // Rows go in pairs so that the pooled output (2x2 average of the stored values in k_pool2's sum order:
      // rows r, r + 1, columns col (even) and col + 1, the next lane) is made from registers.
      const int x = tx0 + col;
      const int Ho = H >> 1, Wo = W >> 1;
      for (int r = 0; r < RPW; r += 2) {
        const int y = ty0 + wave * RPW + r;
        for (int t = 0; t < COT; ++t) {
          const int cb = co0 + 32 * t;
          if (cb >= Cout) break;
          for (int g = 0; g < 4; ++g) {
            const int co = cb + 8 * g + 4 * half;
            uint16_t o0[4], o1[4];
            for (int j = 0; j < 4; ++j) {
              const float bj = sBias[co + j];
              float v0 = acc[r][t][4 * g + j] + bj, v1 = acc[r + 1][t][4 * g + j] + bj;
              if (relu && v0 < 0.f) v0 = 0.f;
              if (relu && v1 < 0.f) v1 = 0.f;
              o0[j] = f2bf(v0);
              o1[j] = f2bf(v1);
            }
            if (x < W) {
              if (y < H)
                *reinterpret_cast<uint2*>(out + (((size_t)n * H + y) * W + x) * Cout + co) =
                    make_uint2((uint32_t)o0[0] | ((uint32_t)o0[1] << 16), (uint32_t)o0[2] | ((uint32_t)o0[3] << 16));
              if (y + 1 < H)
                *reinterpret_cast<uint2*>(out + (((size_t)n * H + y + 1) * W + x) * Cout + co) =
                    make_uint2((uint32_t)o1[0] | ((uint32_t)o1[1] << 16), (uint32_t)o1[2] | ((uint32_t)o1[3] << 16));
            }
            if (pooled) {
              uint16_t q[4];
              for (int j = 0; j < 4; ++j) {
                const float x00 = bf2f(o0[j]), x10 = bf2f(o1[j]);
                const float x01 = xor1(x00), x11 = xor1(x10);
                q[j] = f2bf(0.25f * (((x00 + x01) + x10) + x11));
              }
              if ((col & 1) == 0 && y < H && x < W)
                *reinterpret_cast<uint2*>(pooled + (((size_t)n * Ho + (y >> 1)) * Wo + (x >> 1)) * Cout + co) =
                    make_uint2((uint32_t)q[0] | ((uint32_t)q[1] << 16), (uint32_t)q[2] | ((uint32_t)q[3] << 16));
            }
          }
        }
      }
      for (int r = 0; r < RPW; ++r)
        for (int t = 0; t < COT; ++t)
          for (int i = 0; i < 16; ++i) acc[r][t][i] = 0.f;
    }
    if (!more) break;
    job = nJob;
    chunk = nChunk;
    tx0 = nx0;
    ty0 = ny0;
    n = nn;
    co0 = nco;
    stageW = nStageW;
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

// persistent grid: `occ` workgroups per CU (as many as their LDS and registers allow), at most one per job
unsigned conv_grid(long jobs, int occ) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  long g = (long)occ * cus;
  return (unsigned)(jobs < g ? jobs : g);
}

template <int COT, int RPW, int OCC>
void launch_conv_t(const ConvIn& src, int N, int H, int W, const uint16_t* w, const float* bias, int Cout, int relu,
                   uint16_t* out, uint16_t* pooled, hipStream_t s) {
  const long jobs = (long)((W + 31) / 32) * ((H + 4 * RPW - 1) / (4 * RPW)) * N * ((Cout + 32 * COT - 1) / (32 * COT));
  dim3 grid(conv_grid(jobs, OCC));
  if ((src.Ca + src.Cb) % 32 == 0)
    hipLaunchKernelGGL((k_conv3x3<32, COT, RPW, OCC>), grid, dim3(256), 0, s, src, N, H, W, w, bias, Cout, relu, out, pooled);
  else
    hipLaunchKernelGGL((k_conv3x3<16, COT, RPW, OCC>), grid, dim3(256), 0, s, src, N, H, W, w, bias, Cout, relu, out, pooled);
}

// One 32-channel tile and two output rows per wavefront (8 x 32 pixels per workgroup), three workgroups
// per CU: of the shapes measured on the 1080p U-Net forward pass (rows 2 to 8 per wavefront, one or two
// channel tiles, one to three workgroups per CU), the one with the most workgroups in flight was
// fastest, as each step's global loads wait on the memory system rather than on the matrix cores.
int launch_conv(const ConvIn& src, int N, int H, int W, const uint16_t* w, const float* bias, int Cout, int relu,
                uint16_t* out, hipStream_t s, uint16_t* pooled = nullptr) {
  launch_conv_t<1, 2, 3>(src, N, H, W, w, bias, Cout, relu, out, pooled, s);
  return check_launch();
}

int check_conv(int N, int H, int W, int Cin, int Cout) {
  if (N <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0) return fail(YDN_E_INVALID, "empty shape");
  if (N > 65535) return fail(YDN_E_INVALID, "more than 65535 images");
  if (Cin % 16 != 0) return fail(YDN_E_INVALID, "Cin must be a multiple of 16");
  if (Cout % 32 != 0) return fail(YDN_E_INVALID, "Cout must be a multiple of 32");
  if (Cout > CONV_MAX_COUT) return fail(YDN_E_INVALID, "Cout must be at most 256");
  if ((long)((W + 31) / 32) * ((H + 7) / 8) > 0x7FFFFFFFL) return fail(YDN_E_INVALID, "image too large");
  return YDN_OK;
}

}  // namespace

extern "C" {

const char* ydn_last_error(void) { return g_err.c_str(); }

int ydn_conv3x3(const uint16_t* in, int N, int H, int W, int Cin, const uint16_t* w, const float* bias, int Cout, int relu,
                uint16_t* out, void* stream) {
  if (!in || !w || !out) return fail(YDN_E_INVALID, "null tensor");
  if (int rc = check_conv(N, H, W, Cin, Cout)) return rc;
  ConvIn src{in, Cin, nullptr, 0, 0};
  return launch_conv(src, N, H, W, w, bias, Cout, relu, out, (hipStream_t)stream);
}

int ydn_conv3x3_pool2(const uint16_t* in, int N, int H, int W, int Cin, const uint16_t* w, const float* bias, int Cout, int relu,
                      uint16_t* out, uint16_t* pooled, void* stream) {
  if (!in || !w || !out || !pooled) return fail(YDN_E_INVALID, "null tensor");
  if (int rc = check_half_shape(N, H, W, Cout)) return rc;
  if (int rc = check_conv(N, H, W, Cin, Cout)) return rc;
  ConvIn src{in, Cin, nullptr, 0, 0};
  return launch_conv(src, N, H, W, w, bias, Cout, relu, out, (hipStream_t)stream, pooled);
}

int ydn_conv3x3_up2cat(const uint16_t* coarse, int C1, const uint16_t* skip, int C2, int N, int H, int W, const uint16_t* w,
                       const float* bias, int Cout, int relu, uint16_t* out, void* stream) {
  if (!coarse || !skip || !w || !out) return fail(YDN_E_INVALID, "null tensor");
  if (int rc = check_half_shape(N, H, W, C1)) return rc;
  if (C2 <= 0 || C1 % 16 != 0 || C2 % 16 != 0) return fail(YDN_E_INVALID, "C1 and C2 must be positive multiples of 16");
  if (int rc = check_conv(N, H, W, C1 + C2, Cout)) return rc;
  ConvIn src{coarse, C1, skip, C2, 1};
  return launch_conv(src, N, H, W, w, bias, Cout, relu, out, (hipStream_t)stream);
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
