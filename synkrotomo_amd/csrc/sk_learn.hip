// sk_learn.hip -- kernels of the learned SIRT correction (no reference counterpart) and the
// GPU phantom generator used to train it.
//
// Network: y = x + conv4(relu(conv3(relu(conv2(relu(conv1(x))))))), 3x3 convolutions with
// stride 1 and zero padding 1; conv1: 1 -> 32 channels, conv2/conv3: 32 -> 32, conv4: 32 -> 1.
// Activations are BF16 in NHWC layout ([image][row][col][32]); images, biases and every
// accumulation are FP32; weights are BF16 in [Cout][ky][kx][Cin] layout.
//
//   k_conv32      32 -> 32 (the FLOP-heavy layers) as an implicit GEMM on gfx950 MFMA
//                 (v_mfma_f32_16x16x32_bf16): M = 16 output channels, N = 16 output pixels
//                 of one row, K = 32 input channels of one tap; 9 taps accumulate.  Persistent
//                 blocks compute 16x16-pixel x 32-channel tiles from an LDS-staged 18x18x32
//                 input patch (the next tile's patch loaded during the MFMAs) and the 32x9x32
//                 weights, staged once per block.  The same kernel computes the data
//                 gradient (weights flipped and transposed on the host side), with the ReLU
//                 derivative of the layer's output fused into the staging of its input.
//   k_conv_in     1 or 2 -> 32 (288 MACs per pixel and input channel) on VALU with fused bias and
//                 ReLU, in FP32 on the FP32 input (the image, or the image and the SIRT update
//                 direction interleaved); the data gradients of both thin layers likewise.
//   k_conv_out_mfma  32 -> 1: per 16x16 tile, the 9 tap products of every halo pixel in one
//                 MFMA per 16 pixels, then a 9-term shifted sum per output, fused bias and
//                 residual.
//   k_wgrad32_mfma  weight and bias gradients of the 32 -> 32 layers on MFMA (K = pixels),
//                 channel-major operands staged in LDS; per-block partial sums (deterministic),
//                 then k_reduce_partials.
//   k_wgrad       weight and bias gradients of the thin layers (1 -> 32, 32 -> 1) on VALU, the
//                 same per-block partials.
//   k_phantom     random-ellipse phantoms, evaluated in FP64 exactly as the numpy
//                 generator (synkrotomo_amd/phantom.py) does.
#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "sk_common.h"

namespace sk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CT = 16;    // conv tile side (pixels)
constexpr int CP = 40;    // LDS channel stride (bf16 elements): 80 B, keeps 16 B alignment
constexpr int NCH = 32;   // hidden channels

// ---------------------------------------------------------------------- 32 -> 32, MFMA
// Persistent blocks: each stages the 32x9x32 weights in LDS once, then walks output tiles
// blockIdx.x, blockIdx.x + gridDim.x, ...  A tile's 18x18x32 input patch is loaded into
// registers while the previous tile's MFMAs run, then stored to LDS.  The MFMA's M is the 16
// output channels of a channel block (A = weights), N the 16 pixels of a row (B = input), so
// each lane's accumulator holds 4 consecutive channels of one pixel: 8-byte stores.
constexpr int PATCH_Q = (CT + 2) * (CT + 2) * 4;      // 16-byte quads of a patch
constexpr int PATCH_PER_T = (PATCH_Q + 255) / 256;    // per thread

template <bool MASK_IN, bool BIAS_RELU>
__global__ __launch_bounds__(256) void k_conv32(int H, int W, long ntiles, const __bf16* __restrict__ x,
                                                const __bf16* __restrict__ mask,
                                                const __bf16* __restrict__ w,
                                                const float* __restrict__ bias,
                                                __bf16* __restrict__ y) {
  __shared__ __attribute__((aligned(16))) __bf16 sx[(CT + 2) * (CT + 2) * CP];
  __shared__ __attribute__((aligned(16))) __bf16 sw[NCH * 9 * CP];
  const int tiles_x = (W + CT - 1) / CT;
  const long tiles_img = (long)tiles_x * ((H + CT - 1) / CT);
  for (int c = threadIdx.x; c < NCH * 9 * 4; c += 256) {
    const int row = c >> 2, q = c & 3;  // row = co*9 + tap
    *(bf16x8*)(sw + row * CP + q * 8) = *(const bf16x8*)(w + row * NCH + q * 8);
  }
  bf16x8 pf[PATCH_PER_T];
  auto load_patch = [&](long t) {
    const long b = t / tiles_img;
    const int r = (int)(t - b * tiles_img);
    const int tx0 = (r % tiles_x) * CT, ty0 = (r / tiles_x) * CT;
    const size_t img = (size_t)b * H * W;
#pragma unroll
    for (int k = 0; k < PATCH_PER_T; ++k) {
      const int c = threadIdx.x + 256 * k;
      bf16x8 v;
      for (int i = 0; i < 8; ++i) v[i] = (__bf16)0.0f;
      const int pix = c >> 2, q = c & 3;
      const int py = pix / (CT + 2), px = pix - py * (CT + 2);
      const int gy = ty0 + py - 1, gx = tx0 + px - 1;
      if (c < PATCH_Q && gy >= 0 && gy < H && gx >= 0 && gx < W) {
        const size_t off = (img + (size_t)gy * W + gx) * NCH + q * 8;
        v = *(const bf16x8*)(x + off);
        if constexpr (MASK_IN) {
          const bf16x8 m = *(const bf16x8*)(mask + off);
          for (int i = 0; i < 8; ++i)
            if (!((float)m[i] > 0.0f)) v[i] = (__bf16)0.0f;
        }
      }
      pf[k] = v;
    }
  };
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  long t = blockIdx.x;
  if (t < ntiles) load_patch(t);
  for (; t < ntiles; t += gridDim.x) {  // uniform over the block: every thread reaches each barrier
    __syncthreads();  // the previous tile's LDS reads (and, first, the weight stores) are done
#pragma unroll
    for (int k = 0; k < PATCH_PER_T; ++k) {
      const int c = threadIdx.x + 256 * k;
      if (c < PATCH_Q) *(bf16x8*)(sx + (c >> 2) * CP + (c & 3) * 8) = pf[k];
    }
    __syncthreads();
    const long tn = t + gridDim.x;
    if (tn < ntiles) load_patch(tn);  // in flight during this tile's MFMAs
    const long b = t / tiles_img;
    const int r = (int)(t - b * tiles_img);
    const int tx0 = (r % tiles_x) * CT, ty0 = (r / tiles_x) * CT;
    const size_t img = (size_t)b * H * W;
    f32x4 acc[4][2];
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 2; ++c) acc[a][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) {
      const int ky = tp / 3, kx = tp - 3 * (tp / 3);
      // A: weights of output channels nb*16 + m, input channels 8kq..8kq+7 of tap tp
      const bf16x8 w0 = *(const bf16x8*)(sw + ((0 * 16 + m) * 9 + tp) * CP + 8 * kq);
      const bf16x8 w1 = *(const bf16x8*)(sw + ((1 * 16 + m) * 9 + tp) * CP + 8 * kq);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const int oy = wv * 4 + mb;
        // B: input channels 8kq..8kq+7 of pixel m of output row oy, shifted by the tap
        const bf16x8 a = *(const bf16x8*)(sx + ((oy + ky) * (CT + 2) + m + kx) * CP + 8 * kq);
        acc[mb][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, a, acc[mb][0], 0, 0, 0);
        acc[mb][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, a, acc[mb][1], 0, 0, 0);
      }
    }
    // C: lane holds output channels nb*16 + 4kq + i (i < 4) of pixel m of row oy.
    const int ox = tx0 + m;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const int oy = ty0 + wv * 4 + mb;
      if (oy >= H || ox >= W) continue;
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        const int co = nb * 16 + 4 * kq;
        bf16x4 o;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float v = acc[mb][nb][i];
          if constexpr (BIAS_RELU) v = fmaxf(v + bias[co + i], 0.0f);
          o[i] = (__bf16)v;
        }
        *(bf16x4*)(y + (img + (size_t)oy * W + ox) * NCH + co) = o;
      }
    }
  }
}

// ---------------------------------------------------------------------- CIN -> 32, VALU
// h[p][co] = relu(b[co] + sum_t sum_c w[co][t][c] x[p + off_t][c]); x: FP32 [B][H][W][CIN]
// (CIN = 1: an image; CIN = 2: the image and the SIRT update direction, interleaved), w: BF16
// [32][9][CIN].  Grid (pixel blocks of one image, images).  The weights are read as BF16 pairs
// through readfirstlane (wave-uniform values: scalar loads into SGPRs, unpacked on the scalar
// unit, operands of the FMAs): no LDS reads for them and no per-lane conversion.  Same
// operations in the same order as a plain per-pixel loop: bias, then taps 0..8 (channels in
// order within a tap).  The block's 256 x 32 BF16 outputs (16 KB, contiguous in NHWC) go out
// through LDS so that every store instruction writes 1 KB of consecutive bytes.  w2: the
// weights as 144 * CIN words.
template <int CIN>
__global__ __launch_bounds__(256) void k_conv_in(int H, int W, const float* __restrict__ x,
                                                 const uint32_t* __restrict__ w2,
                                                 const float* __restrict__ bias,
                                                 __bf16* __restrict__ h) {
  __shared__ bf16x8 st[256 * 4];
  const int r0 = blockIdx.x * 256;
  const int r = r0 + threadIdx.x;
  const bool live = r < H * W;
  const size_t img = (size_t)blockIdx.y * H * W;
  const int py = r / W, px = r - py * W;
  float v[9][CIN];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int yy = py + t / 3 - 1, xx = px + t % 3 - 1;
    const bool in = live && yy >= 0 && yy < H && xx >= 0 && xx < W;
#pragma unroll
    for (int c = 0; c < CIN; ++c) v[t][c] = in ? x[(img + (size_t)yy * W + xx) * CIN + c] : 0.0f;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int co = q * 8 + i;
      float s = bias[co];
#pragma unroll
      for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int c = 0; c < CIN; ++c) {
          const int k = (co * 9 + t) * CIN + c;
          const uint32_t pr = __builtin_amdgcn_readfirstlane(w2[k >> 1]);
          s = fmaf(__builtin_bit_cast(float, (k & 1) ? (pr & 0xffff0000u) : (pr << 16)), v[t][c], s);
        }
      o[i] = (__bf16)fmaxf(s, 0.0f);
    }
    st[threadIdx.x * 4 + q] = o;
  }
  __syncthreads();
  bf16x8* out = (bf16x8*)(h + (img + r0) * NCH);
  const int n = min(256, H * W - r0) * 4;  // 16-byte chunks of this block's pixels
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int c = threadIdx.x + 256 * k;
    if (c < n) out[c] = st[c];
  }
}

// 32 -> 1 on MFMA: for a 16x16 output tile, the 18x18 input pixels of its halo each get
// U[q][t] = sum_ci h[q][ci] w[t][ci] for the 9 taps t in one GEMM ([pixels x 32] x [32 x 16],
// taps 9..15 zero: one v_mfma_f32_16x16x32_bf16 per 16 pixels), kept in LDS; each output
// pixel then sums its 9 shifted taps: y[p] = res[p] + (sum_t U[p + off_t][t] + bias).
// Each activation is read once per tile (the halo adds 27%).
constexpr int CO_SLOTS = 21 * 16;  // 18*18 = 324 halo pixels in groups of 16

__global__ __launch_bounds__(256) void k_conv_out_mfma(int H, int W, const __bf16* __restrict__ h,
                                                       const __bf16* __restrict__ w,
                                                       const float* __restrict__ bias,
                                                       const float* __restrict__ res,
                                                       float* __restrict__ y) {
  __shared__ float U[CO_SLOTS * 9];
  const int tiles_x = (W + CT - 1) / CT;
  const int tx0 = (blockIdx.x % tiles_x) * CT, ty0 = (blockIdx.x / tiles_x) * CT;
  const size_t img = (size_t)blockIdx.y * H * W;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  // B: w[tap m][8kq .. 8kq+7] for taps 0..8, zero columns for m >= 9
  bf16x8 wb;
  for (int i = 0; i < 8; ++i) wb[i] = (__bf16)0.0f;
  if (m < 9) wb = *(const bf16x8*)(w + m * NCH + 8 * kq);
  // wave wv takes pixel groups wv, wv + 4, ...: all of its loads are issued before the MFMAs
  constexpr int NG = CO_SLOTS / 16, GPW = (NG + 3) / 4;
  bf16x8 a[GPW];
#pragma unroll
  for (int k = 0; k < GPW; ++k) {
    const int slot = (wv + 4 * k) * 16 + m;  // A row m: halo pixel `slot`
    const int hy = slot / (CT + 2), hx = slot - hy * (CT + 2);
    const int gy = ty0 + hy - 1, gx = tx0 + hx - 1;
    for (int i = 0; i < 8; ++i) a[k][i] = (__bf16)0.0f;
    if (slot < (CT + 2) * (CT + 2) && gy >= 0 && gy < H && gx >= 0 && gx < W)
      a[k] = *(const bf16x8*)(h + (img + (size_t)gy * W + gx) * NCH + 8 * kq);
  }
#pragma unroll
  for (int k = 0; k < GPW; ++k) {
    const int g = wv + 4 * k;
    if (g >= NG) break;  // wave-uniform
    const f32x4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[k], wb, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
    // C: lane holds U[pixel g*16 + 4kq + i][tap m]
    if (m < 9)
      for (int i = 0; i < 4; ++i) U[(g * 16 + 4 * kq + i) * 9 + m] = c[i];
  }
  __syncthreads();
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  const int oy = ty0 + ty, ox = tx0 + tx;
  if (oy >= H || ox >= W) return;
  float sum = 0.0f;
#pragma unroll
  for (int t = 0; t < 9; ++t) sum += U[((ty + t / 3) * (CT + 2) + tx + t % 3) * 9 + t];
  const size_t p = img + (size_t)oy * W + ox;
  y[p] = res[p] + (sum + bias[0]);
}

// d h[q][ci] = sum_t w[t][ci] g[q - off_t]   (data gradient of the 32 -> 1 layer)
__global__ __launch_bounds__(256) void k_conv_out_dgrad(int H, int W, long npix, const float* __restrict__ g,
                                                        const __bf16* __restrict__ w,
                                                        __bf16* __restrict__ dh) {
  __shared__ float sw[9 * NCH];
  for (int i = threadIdx.x; i < 9 * NCH; i += 256) sw[i] = (float)w[i];
  __syncthreads();
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  const long b = p / ((long)H * W);
  const int r = (int)(p - b * H * W);
  const int py = r / W, px = r - py * W;
  float gv[9];
  for (int t = 0; t < 9; ++t) {
    const int yy = py - (t / 3 - 1), xx = px - (t % 3 - 1);
    gv[t] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? g[(b * H + yy) * W + xx] : 0.0f;
  }
  __bf16* out = dh + p * NCH;
  for (int q = 0; q < 4; ++q) {
    bf16x8 o;
    for (int i = 0; i < 8; ++i) {
      const int ci = q * 8 + i;
      float s = 0.0f;
      for (int t = 0; t < 9; ++t) s = fmaf(sw[t * NCH + ci], gv[t], s);
      o[i] = (__bf16)s;
    }
    *(bf16x8*)(out + q * 8) = o;
  }
}

// dx[q][c] = sum_{t,co} w[co][t][c] g1[q - off_t][co],  g1 = dh * (h > 0)   (data gradient of
// k_conv_in; dx FP32 [B][H][W][CIN])
template <int CIN>
__global__ __launch_bounds__(256) void k_conv_in_dgrad(int H, int W, long npix, const __bf16* __restrict__ dh,
                                                       const __bf16* __restrict__ hmask,
                                                       const __bf16* __restrict__ w,
                                                       float* __restrict__ dx) {
  __shared__ float sw[NCH * 9 * CIN];
  for (int i = threadIdx.x; i < NCH * 9 * CIN; i += 256) sw[i] = (float)w[i];
  __syncthreads();
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  const long b = p / ((long)H * W);
  const int r = (int)(p - b * H * W);
  const int py = r / W, px = r - py * W;
  float s[CIN];
#pragma unroll
  for (int c = 0; c < CIN; ++c) s[c] = 0.0f;
  for (int t = 0; t < 9; ++t) {
    const int yy = py - (t / 3 - 1), xx = px - (t % 3 - 1);
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    const size_t off = ((b * H + yy) * W + xx) * NCH;
    for (int q = 0; q < 4; ++q) {
      const bf16x8 gv = *(const bf16x8*)(dh + off + q * 8);
      const bf16x8 mv = *(const bf16x8*)(hmask + off + q * 8);
      for (int i = 0; i < 8; ++i)
        if ((float)mv[i] > 0.0f)
#pragma unroll
          for (int c = 0; c < CIN; ++c) s[c] = fmaf(sw[((q * 8 + i) * 9 + t) * CIN + c], (float)gv[i], s[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < CIN; ++c) dx[p * CIN + c] = s[c];
}

// ---------------------------------------------------------------------- weight gradients
// dW[co][t][ci] = sum_p G[p][co] X[p + off_t][ci],  db[co] = sum_p G[p][co]
// G: CO channels (bf16 NHWC, optionally masked by M > 0) or an FP32 image (CO == 1).
// X: CI = 32 channels (bf16 NHWC) or CI <= 2 FP32 channels ([B][H][W][CI]).  Each block
// accumulates over a grid-stride set of 16x16 tiles and writes CO*9*CI + CO partial sums.
template <int CO, int CI>
__global__ __launch_bounds__(256) void k_wgrad(int H, int W, int B, const void* __restrict__ Gp,
                                               const __bf16* __restrict__ M,
                                               const void* __restrict__ Xp,
                                               float* __restrict__ partials) {
  constexpr int NOUT = CO * 9 * CI + CO;
  constexpr int PER = (NOUT + 255) / 256;
  // 32-channel operands are BF16 already and stay BF16 in LDS; FP32 images stay FP32.
  using GT = typename std::conditional<CO == 1, float, __bf16>::type;
  using XT = typename std::conditional<CI <= 2, float, __bf16>::type;
  __shared__ GT sg[CT * CT * CO];
  __shared__ XT sx[(CT + 2) * (CT + 2) * CI];
  float acc[PER];
  for (int k = 0; k < PER; ++k) acc[k] = 0.0f;
  const int tiles_x = (W + CT - 1) / CT, tiles_y = (H + CT - 1) / CT;
  const long ntiles = (long)tiles_x * tiles_y * B;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int b = (int)(tile / ((long)tiles_x * tiles_y));
    const int tr = (int)(tile - (long)b * tiles_x * tiles_y);
    const int ty0 = (tr / tiles_x) * CT, tx0 = (tr % tiles_x) * CT;
    for (int e = threadIdx.x; e < CT * CT * CO; e += 256) {
      const int pix = e / CO, co = e - pix * CO;
      const int gy = ty0 + pix / CT, gx = tx0 + pix % CT;
      float v = 0.0f;
      if (gy < H && gx < W) {
        const size_t p = ((size_t)b * H + gy) * W + gx;
        if constexpr (CO == 1) {
          v = ((const float*)Gp)[p];
        } else {
          v = (float)((const __bf16*)Gp)[p * CO + co];
          if (M && !((float)M[p * CO + co] > 0.0f)) v = 0.0f;
        }
      }
      sg[e] = (GT)v;
    }
    for (int e = threadIdx.x; e < (CT + 2) * (CT + 2) * CI; e += 256) {
      const int pix = e / CI, ci = e - pix * CI;
      const int py = pix / (CT + 2), px = pix % (CT + 2);
      const int gy = ty0 + py - 1, gx = tx0 + px - 1;
      float v = 0.0f;
      if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
        const size_t p = ((size_t)b * H + gy) * W + gx;
        if constexpr (CI <= 2)
          v = ((const float*)Xp)[p * CI + ci];
        else
          v = (float)((const __bf16*)Xp)[p * CI + ci];
      }
      sx[e] = (XT)v;
    }
    __syncthreads();
    for (int k = 0; k < PER; ++k) {
      const int o = threadIdx.x + 256 * k;
      if (o >= NOUT) break;
      float s = 0.0f;
      if (o < CO * 9 * CI) {
        const int co = o / (9 * CI), rem = o - co * 9 * CI, t = rem / CI, ci = rem - t * CI;
        const int ky = t / 3, kx = t % 3;
        for (int py = 0; py < CT; ++py)
          for (int px = 0; px < CT; ++px)
            s = fmaf((float)sg[(py * CT + px) * CO + co],
                     (float)sx[((py + ky) * (CT + 2) + px + kx) * CI + ci], s);
      } else {
        const int co = o - CO * 9 * CI;
        for (int pix = 0; pix < CT * CT; ++pix) s += (float)sg[pix * CO + co];
      }
      acc[k] += s;
    }
    __syncthreads();
  }
  for (int k = 0; k < PER; ++k) {
    const int o = threadIdx.x + 256 * k;
    if (o < NOUT) partials[(size_t)blockIdx.x * NOUT + o] = acc[k];
  }
}

// 32 -> 32 weight and bias gradients on MFMA: for each tap t, dW[co][t][ci] is a GEMM with M = 16
// output channels, N = 16 input channels and K = pixels (v_mfma_f32_16x16x32_bf16, K = 32 pixels
// per step); db[co] is the same GEMM against a column of ones.  The MFMA takes 8 consecutive K
// values per lane from one row of A (and of B), so the NHWC operands are staged in LDS channel-
// major: gT[co][pixel] for the 16x16 tile's gradient (masked by M > 0), and three copies of the
// 18x18 input halo, xs[kx][ci][halo row][16 columns] with column j holding halo column j + kx,
// so that a tap's shifted rows start 16-byte aligned.  A K step covers two pixel rows of the
// tile (lane group kq: row 2s + kq/2, columns 8(kq%2) .. +7); wave w takes K steps 2w and 2w+1 of
// every tile of its block and keeps all 36 tap x channel-block accumulators.  At the end the four
// waves' sums are added in LDS in wave order and the block writes one partial (deterministic).
constexpr int WG_GS = 256 + 8;       // gT row (elements): 256 pixels + pad (528 B, 16 B multiple)
constexpr int WG_XS = 18 * 16 + 8;   // xs channel row: 18 halo rows x 16 columns + pad (592 B)
constexpr int WG_NOUT = NCH * 9 * NCH + NCH;

__global__ __launch_bounds__(256) void k_wgrad32_mfma(int H, int W, int B, const __bf16* __restrict__ G,
                                                      const __bf16* __restrict__ M,
                                                      const __bf16* __restrict__ X,
                                                      float* __restrict__ partials) {
  __shared__ __attribute__((aligned(16))) __bf16 gT[NCH * WG_GS];
  __shared__ __attribute__((aligned(16))) __bf16 xs[3 * NCH * WG_XS];
  static_assert(sizeof(__bf16) * 3 * NCH * WG_XS >= sizeof(float) * WG_NOUT, "reduction buffer");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  f32x4 acc[9][2][2];
  f32x4 accb[2];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[t][a][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  accb[0] = accb[1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  bf16x8 ones;
#pragma unroll
  for (int i = 0; i < 8; ++i) ones[i] = (__bf16)1.0f;
  const int tiles_x = (W + CT - 1) / CT, tiles_y = (H + CT - 1) / CT;
  const long ntiles = (long)tiles_x * tiles_y * B;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {  // uniform over the block
    const long b = tile / ((long)tiles_x * tiles_y);
    const int tr = (int)(tile - b * tiles_x * tiles_y);
    const int ty0 = (tr / tiles_x) * CT, tx0 = (tr % tiles_x) * CT;
    const size_t img = (size_t)b * H * W;
    __syncthreads();  // the previous tile's LDS reads are done
    for (int c = threadIdx.x; c < CT * CT * 4; c += 256) {  // gradient: 256 pixels x 4 quads
      const int pix = c >> 2, q = c & 3;
      const int gy = ty0 + (pix >> 4), gx = tx0 + (pix & 15);
      bf16x8 v;
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (__bf16)0.0f;
      if (gy < H && gx < W) {
        const size_t off = (img + (size_t)gy * W + gx) * NCH + 8 * q;
        v = *(const bf16x8*)(G + off);
        if (M) {
          const bf16x8 mk = *(const bf16x8*)(M + off);
#pragma unroll
          for (int i = 0; i < 8; ++i)
            if (!((float)mk[i] > 0.0f)) v[i] = (__bf16)0.0f;
        }
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) gT[(8 * q + i) * WG_GS + pix] = v[i];
    }
    for (int c = threadIdx.x; c < (CT + 2) * (CT + 2) * 4; c += 256) {  // input halo: 324 pixels x 4 quads
      const int pix = c >> 2, q = c & 3;
      const int hy = pix / (CT + 2), hx = pix - hy * (CT + 2);
      const int gy = ty0 + hy - 1, gx = tx0 + hx - 1;
      bf16x8 v;
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (__bf16)0.0f;
      if (gy >= 0 && gy < H && gx >= 0 && gx < W) v = *(const bf16x8*)(X + (img + (size_t)gy * W + gx) * NCH + 8 * q);
#pragma unroll
      for (int kx = 0; kx < 3; ++kx) {
        const int j = hx - kx;
        if (j < 0 || j >= CT) continue;
#pragma unroll
        for (int i = 0; i < 8; ++i) xs[(kx * NCH + 8 * q + i) * WG_XS + hy * CT + j] = v[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ss = 0; ss < 2; ++ss) {
      const int py = 2 * (2 * wv + ss) + (kq >> 1), px0 = (kq & 1) * 8;
      // A: gradient of output channels cb*16 + m at the 8 pixels (py, px0 .. px0 + 7)
      const bf16x8 a0 = *(const bf16x8*)(gT + m * WG_GS + py * CT + px0);
      const bf16x8 a1 = *(const bf16x8*)(gT + (16 + m) * WG_GS + py * CT + px0);
      accb[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, ones, accb[0], 0, 0, 0);
      accb[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, ones, accb[1], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const int ky = t / 3, kx = t - 3 * (t / 3);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
          // B: input channel nb*16 + m at the same pixels shifted by the tap
          const bf16x8 bv = *(const bf16x8*)(xs + (kx * NCH + nb * 16 + m) * WG_XS + (py + ky) * CT + px0);
          acc[t][0][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bv, acc[t][0][nb], 0, 0, 0);
          acc[t][1][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bv, acc[t][1][nb], 0, 0, 0);
        }
      }
    }
  }
  // C: lane (m, kq) holds rows 4kq + i (output channels cb*16 + 4kq + i) of column m (input
  // channel nb*16 + m).  Waves add their sums in order 0, 1, 2, 3.
  float* red = (float*)xs;
  for (int w = 0; w < 4; ++w) {
    __syncthreads();
    if (wv == w) {
#pragma unroll
      for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
          for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int o = ((cb * 16 + 4 * kq + i) * 9 + t) * NCH + nb * 16 + m;
              red[o] = (w ? red[o] : 0.0f) + acc[t][cb][nb][i];
            }
      if (m == 0)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int o = NCH * 9 * NCH + cb * 16 + 4 * kq + i;
            red[o] = (w ? red[o] : 0.0f) + accb[cb][i];
          }
    }
  }
  __syncthreads();
  for (int o = threadIdx.x; o < WG_NOUT; o += 256) partials[(size_t)blockIdx.x * WG_NOUT + o] = red[o];
}

// out[o] = sum over blocks b of partials[b][o] in a fixed order (deterministic): a workgroup
// takes 64 consecutive outputs (coalesced rows of the partials), wave w sums the blocks
// b = w, w + 4, ... in increasing order, then the four sums are added in wave order.
__global__ __launch_bounds__(256) void k_reduce_partials(const float* __restrict__ partials, int nblocks,
                                                         int nout, float* __restrict__ out) {
  __shared__ float sh[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int o = blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (o < nout)
    for (int b = w; b < nblocks; b += 4) s += partials[(size_t)b * nout + o];
  sh[w][lane] = s;
  __syncthreads();
  if (w == 0 && o < nout) out[o] = ((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane];
}

// ---------------------------------------------------------------------- phantoms
// params[b][e] = (cx, cy, a, b, cos t, sin t, value); pixel centres at (2j+1)/N - 1.
__global__ __launch_bounds__(256) void k_phantom(int N, int E, long npix, const double* __restrict__ params,
                                                 float* __restrict__ out) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  const long b = p / ((long)N * N);
  const int r = (int)(p - b * N * N);
  const int i = r / N, j = r - i * N;
  const double x = (2.0 * j + 1.0) / N - 1.0, y = (2.0 * i + 1.0) / N - 1.0;
  double acc = 0.0;
  const double* P = params + b * E * 7;
  for (int e = 0; e < E; ++e) {
    const double* q = P + e * 7;
    const double dx = x - q[0], dy = y - q[1];
    const double u = (dx * q[4] + dy * q[5]) / q[2];
    const double w = (-dx * q[5] + dy * q[4]) / q[3];
    acc += (u * u + w * w <= 1.0) ? q[6] : 0.0;
  }
  acc = acc < 0.0 ? 0.0 : (acc > 1.0 ? 1.0 : acc);
  out[p] = (float)acc;
}

}  // namespace sk

using namespace sk;

static int check_ctx(sk_ctx* ctx) {
  if (!ctx) return SK_EINVAL;
  SK_HIP(&ctx->c, hipSetDevice(ctx->c.device));
  return SK_OK;
}

extern "C" {

int sk_conv32(sk_ctx* ctx, int B, int H, int W, const void* x, const void* mask, const void* w,
              const float* bias, void* y, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (B < 1 || B > 65535 || H < 1 || W < 1 || !x || !w || !y)
    return fail(&ctx->c, SK_EINVAL, "conv32: bad arguments");
  const bool relu = bias != nullptr;
  const long ntiles = (long)((W + CT - 1) / CT) * ((H + CT - 1) / CT) * B;
  // persistent grid: the resident blocks of the chip (queried once per process and device)
  static int per_cu[64] = {0}, cus[64] = {0};
  const int dev = ctx->c.device < 64 ? ctx->c.device : 63;
  if (!per_cu[dev]) {
    SK_HIP(&ctx->c, hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, ctx->c.device));
    SK_HIP(&ctx->c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[dev], k_conv32<true, true>, 256, 0));
    per_cu[dev] = std::max(1, per_cu[dev]);
  }
  const long resident = (long)cus[dev] * per_cu[dev];
  dim3 grid((unsigned)std::min(ntiles, resident));
  hipStream_t s = (hipStream_t)stream;
  const __bf16 *X = (const __bf16*)x, *Mk = (const __bf16*)mask, *Wt = (const __bf16*)w;
  __bf16* Y = (__bf16*)y;
  if (mask && relu)
    hipLaunchKernelGGL((k_conv32<true, true>), grid, dim3(256), 0, s, H, W, ntiles, X, Mk, Wt, bias, Y);
  else if (mask)
    hipLaunchKernelGGL((k_conv32<true, false>), grid, dim3(256), 0, s, H, W, ntiles, X, Mk, Wt, bias, Y);
  else if (relu)
    hipLaunchKernelGGL((k_conv32<false, true>), grid, dim3(256), 0, s, H, W, ntiles, X, Mk, Wt, bias, Y);
  else
    hipLaunchKernelGGL((k_conv32<false, false>), grid, dim3(256), 0, s, H, W, ntiles, X, Mk, Wt, bias, Y);
  return launched(&ctx->c, "conv32 kernel");
}

static int conv_in_go(sk_ctx* ctx, int cin, int B, int H, int W, const float* x, const void* w, const float* bias,
                      void* h, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (B < 1 || H < 1 || W < 1 || !x || !w || !bias || !h) return fail(&ctx->c, SK_EINVAL, "conv_in: bad arguments");
  if (B > 65535 || (long)H * W >= (1L << 31)) return fail(&ctx->c, SK_EINVAL, "conv_in: batch or image too large");
  if ((uintptr_t)w % 4) return fail(&ctx->c, SK_EINVAL, "conv_in: weights must be 4-byte aligned");
  const dim3 grid((unsigned)((H * W + 255) / 256), B);
  if (cin == 1)
    hipLaunchKernelGGL(k_conv_in<1>, grid, dim3(256), 0, (hipStream_t)stream, H, W, x, (const uint32_t*)w, bias,
                       (__bf16*)h);
  else
    hipLaunchKernelGGL(k_conv_in<2>, grid, dim3(256), 0, (hipStream_t)stream, H, W, x, (const uint32_t*)w, bias,
                       (__bf16*)h);
  return launched(&ctx->c, "conv_in kernel");
}

int sk_conv_in(sk_ctx* ctx, int B, int H, int W, const float* x, const void* w, const float* bias,
               void* h, void* stream) {
  return conv_in_go(ctx, 1, B, H, W, x, w, bias, h, stream);
}

int sk_conv_in2(sk_ctx* ctx, int B, int H, int W, const float* x2, const void* w, const float* bias,
                void* h, void* stream) {
  return conv_in_go(ctx, 2, B, H, W, x2, w, bias, h, stream);
}

int sk_conv_out(sk_ctx* ctx, int B, int H, int W, const void* h, const void* w, const float* bias,
                const float* res, float* y, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (B < 1 || H < 1 || W < 1 || !h || !w || !bias || !res || !y)
    return fail(&ctx->c, SK_EINVAL, "conv_out: bad arguments");
  if (B > 65535) return fail(&ctx->c, SK_EINVAL, "conv_out: batch must be <= 65535");
  dim3 grid(((W + CT - 1) / CT) * ((H + CT - 1) / CT), B);
  hipLaunchKernelGGL(k_conv_out_mfma, grid, dim3(256), 0, (hipStream_t)stream, H, W, (const __bf16*)h,
                     (const __bf16*)w, bias, res, y);
  return launched(&ctx->c, "conv_out kernel");
}

int sk_conv_out_dgrad(sk_ctx* ctx, int B, int H, int W, const float* g, const void* w, void* dh,
                      void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (B < 1 || H < 1 || W < 1 || !g || !w || !dh) return fail(&ctx->c, SK_EINVAL, "conv_out_dgrad: bad arguments");
  const long npix = (long)B * H * W;
  hipLaunchKernelGGL(k_conv_out_dgrad, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     H, W, npix, g, (const __bf16*)w, (__bf16*)dh);
  return launched(&ctx->c, "conv_out_dgrad kernel");
}

static int conv_in_dgrad_go(sk_ctx* ctx, int cin, int B, int H, int W, const void* dh, const void* hmask,
                            const void* w, float* dx, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (B < 1 || H < 1 || W < 1 || !dh || !hmask || !w || !dx)
    return fail(&ctx->c, SK_EINVAL, "conv_in_dgrad: bad arguments");
  const long npix = (long)B * H * W;
  const dim3 grid((unsigned)((npix + 255) / 256));
  if (cin == 1)
    hipLaunchKernelGGL(k_conv_in_dgrad<1>, grid, dim3(256), 0, (hipStream_t)stream, H, W, npix, (const __bf16*)dh,
                       (const __bf16*)hmask, (const __bf16*)w, dx);
  else
    hipLaunchKernelGGL(k_conv_in_dgrad<2>, grid, dim3(256), 0, (hipStream_t)stream, H, W, npix, (const __bf16*)dh,
                       (const __bf16*)hmask, (const __bf16*)w, dx);
  return launched(&ctx->c, "conv_in_dgrad kernel");
}

int sk_conv_in_dgrad(sk_ctx* ctx, int B, int H, int W, const void* dh, const void* hmask,
                     const void* w, float* dx, void* stream) {
  return conv_in_dgrad_go(ctx, 1, B, H, W, dh, hmask, w, dx, stream);
}

int sk_conv_in2_dgrad(sk_ctx* ctx, int B, int H, int W, const void* dh, const void* hmask,
                      const void* w, float* dx2, void* stream) {
  return conv_in_dgrad_go(ctx, 2, B, H, W, dh, hmask, w, dx2, stream);
}

// kind 0: conv1 (G = 32-ch gradient masked by h1, X = fp32 image) -> dW [32][9], db [32]
// kind 1: conv2/3 (G masked, X = 32-ch activation)                  -> dW [32][9][32], db [32]
// kind 2: conv4 (G = fp32 image gradient, X = 32-ch activation)     -> dW [9][32], db [1]
// kind 3: conv1 with two input channels (G masked, X = fp32 [B][H][W][2]) -> dW [32][9][2], db [32]
// out: CO*9*CI weight gradients followed by CO bias gradients (fp32).
int sk_conv_wgrad(sk_ctx* ctx, int kind, int B, int H, int W, const void* g, const void* mask,
                  const void* x, float* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  Ctx* c = &ctx->c;
  if (B < 1 || H < 1 || W < 1 || !g || !x || !out || kind < 0 || kind > 3)
    return fail(c, SK_EINVAL, "conv_wgrad: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const long ntiles = (long)((W + CT - 1) / CT) * ((H + CT - 1) / CT) * B;
  long cap = 1024;
  if (kind == 1) {  // the MFMA kernel: one resident set of blocks (queried once per device)
    static int wg_resident[64] = {0};
    const int dev = c->device < 64 ? c->device : 63;
    if (!wg_resident[dev]) {
      int cus = 0, per_cu = 0;
      SK_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
      SK_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_wgrad32_mfma, 256, 0));
      wg_resident[dev] = std::max(1, cus * std::max(1, per_cu));
    }
    cap = wg_resident[dev];
  }
  const int nblocks = (int)std::min<long>(ntiles, cap);
  const int nout = kind == 0 ? NCH * 9 + NCH
                             : (kind == 1 ? NCH * 9 * NCH + NCH : (kind == 2 ? 9 * NCH + 1 : NCH * 9 * 2 + NCH));
  void* ws = nullptr;
  rc = scratch(c, sizeof(float) * (size_t)nblocks * nout, s, &ws);
  if (rc) return rc;
  float* part = (float*)ws;
  const __bf16* Mk = (const __bf16*)mask;
  if (kind == 0)
    hipLaunchKernelGGL((k_wgrad<NCH, 1>), dim3(nblocks), dim3(256), 0, s, H, W, B, g, Mk, x, part);
  else if (kind == 1)
    hipLaunchKernelGGL(k_wgrad32_mfma, dim3(nblocks), dim3(256), 0, s, H, W, B, (const __bf16*)g, Mk,
                       (const __bf16*)x, part);
  else if (kind == 2)
    hipLaunchKernelGGL((k_wgrad<1, NCH>), dim3(nblocks), dim3(256), 0, s, H, W, B, g, Mk, x, part);
  else
    hipLaunchKernelGGL((k_wgrad<NCH, 2>), dim3(nblocks), dim3(256), 0, s, H, W, B, g, Mk, x, part);
  rc = launched(c, "wgrad kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(k_reduce_partials, dim3((nout + 63) / 64), dim3(256), 0, s, part, nblocks, nout, out);
  return launched(c, "partials reduction kernel");
}

int sk_phantoms(sk_ctx* ctx, int B, int N, int E, const double* d_params, float* d_out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (B < 1 || N < 1 || E < 0 || !d_params || !d_out) return fail(&ctx->c, SK_EINVAL, "phantoms: bad arguments");
  const long npix = (long)B * N * N;
  hipLaunchKernelGGL(k_phantom, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N, E,
                     npix, d_params, d_out);
  return launched(&ctx->c, "phantom kernel");
}

}  // extern "C"
