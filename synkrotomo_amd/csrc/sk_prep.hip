// sk_prep.hip -- measured detector data to the sinogram the projector expects.
//
// The reference reconstructs from sinograms that are already line integrals (its SIRT
// benchmark reads data/tooth.in, SIRT.fut:1-2); the steps before that are the caller's.
// These kernels do them on the GPU, next to the projector:
//
//   sk_normalize      flat/dark-field normalisation and -log: per bin r, the flat and dark
//                     frames are averaged (k_field_means), then for every entry
//                     T = (raw - dark[r]) / (flat[r] - dark[r]) and
//                     b = -log(max(T, min_transmission)) / scale.  One read and one write of
//                     the sinogram stream: HBM-bound elementwise work.
//   sk_row_centroid   the centre of mass sum_r r*v / sum_r v of every sinogram row, summed in
//                     FP64 (one wave per row, a fixed reduction order: deterministic).  For a
//                     parallel-beam scan the centroids trace c0 + p cos(theta) + q sin(theta);
//                     the host fits c0, the detector position of the rotation axis, and
//                     centres the geometry by rhozero = -c0 * deltarho (no resampling).
#include "sk_common.h"

namespace sk {

int launched(Ctx* c, const char* what);

__global__ __launch_bounds__(256) void k_field_means(int R, int nf, int nd, const float* __restrict__ flat,
                                                     const float* __restrict__ dark, float* __restrict__ fd) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  double f = 0.0, d = 0.0;
  for (int k = 0; k < nf; ++k) f += (double)flat[(size_t)k * R + r];
  for (int k = 0; k < nd; ++k) d += (double)dark[(size_t)k * R + r];
  const float fm = (float)(f / nf);
  const float dm = nd > 0 ? (float)(d / nd) : 0.0f;
  const float range = fm - dm;
  fd[r] = dm;  // dark mean
  // reciprocal of the open-beam range; 0 for a dead bin (flat mean not above the dark mean),
  // which then reads T = 0 -> min_transmission: fully absorbing
  fd[R + r] = range > 0.0f ? 1.0f / range : 0.0f;
}

// Grid (chunks of a row, rows): thread t of block (bx, by) handles entry (or 4-entry group)
// bx*256 + t of rows by, by + gridDim.y, ...  Four consecutive entries per thread when R is a
// multiple of 4 and both arrays are 16-byte aligned (16-byte loads and stores).
// raw and out may be the same array (each entry is read, then written, by one thread).
template <bool VEC4>
__global__ __launch_bounds__(256) void k_normalize(long rows, int R, const float* raw,
                                                   const float* __restrict__ fd, float tmin, float inv_scale,
                                                   float* out) {
  const float* dark = fd;
  const float* irange = fd + R;
  auto one = [&](float v, int r) {
    const float t = (v - dark[r]) * irange[r];
    return -logf(fmaxf(t, tmin)) * inv_scale;  // fmaxf(NaN, tmin) = tmin
  };
  const int q = blockIdx.x * 256 + threadIdx.x;
  if constexpr (VEC4) {
    const int R4 = R >> 2;
    if (q >= R4) return;
    const int r = q << 2;
    const float d0 = dark[r], d1 = dark[r + 1], d2 = dark[r + 2], d3 = dark[r + 3];
    const float i0 = irange[r], i1 = irange[r + 1], i2 = irange[r + 2], i3 = irange[r + 3];
    for (long row = blockIdx.y; row < rows; row += gridDim.y) {
      const float4 v = reinterpret_cast<const float4*>(raw + row * R)[q];
      float4 o;
      o.x = -logf(fmaxf((v.x - d0) * i0, tmin)) * inv_scale;
      o.y = -logf(fmaxf((v.y - d1) * i1, tmin)) * inv_scale;
      o.z = -logf(fmaxf((v.z - d2) * i2, tmin)) * inv_scale;
      o.w = -logf(fmaxf((v.w - d3) * i3, tmin)) * inv_scale;
      reinterpret_cast<float4*>(out + row * R)[q] = o;
    }
  } else {
    if (q >= R) return;
    for (long row = blockIdx.y; row < rows; row += gridDim.y) out[row * R + q] = one(raw[row * R + q], q);
  }
}

// One wave per row: lanes stride the row, then a butterfly over the wave (fixed order).
__global__ __launch_bounds__(256) void k_row_centroid(long rows, int R, const float* __restrict__ sino,
                                                      double* __restrict__ com) {
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  const float* p = sino + row * R;
  double m0 = 0.0, m1 = 0.0;
  for (int r = lane; r < R; r += 64) {
    const double v = (double)p[r];
    m0 += v;
    m1 += v * (double)r;
  }
  for (int off = 32; off > 0; off >>= 1) {
    m0 += __shfl_xor(m0, off, 64);
    m1 += __shfl_xor(m1, off, 64);
  }
  if (lane == 0) com[row] = m0 != 0.0 ? m1 / m0 : __builtin_nan("");
}

}  // namespace sk

using namespace sk;

extern "C" {

int sk_normalize(sk_ctx* ctx, int batch, int A, int R, const float* d_raw, int n_flat, const float* d_flat,
                 int n_dark, const float* d_dark, float min_transmission, float scale, float* d_out,
                 void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (batch < 1 || A < 1 || R < 1) return fail(c, SK_EINVAL, "batch, angles and bins must be >= 1");
  if (n_flat < 1 || n_dark < 0) return fail(c, SK_EINVAL, "need >= 1 flat frame and >= 0 dark frames");
  if (!d_raw || !d_flat || !d_out || (n_dark > 0 && !d_dark))
    return fail(c, SK_EINVAL, "device pointer is NULL");
  if (!(min_transmission > 0.0f) || !(scale > 0.0f))
    return fail(c, SK_EINVAL, "min_transmission and scale must be > 0");
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  void* ws = nullptr;
  int rc = scratch(c, sizeof(float) * 2 * (size_t)R, s, &ws);
  if (rc) return rc;
  float* fd = (float*)ws;
  hipLaunchKernelGGL(k_field_means, dim3((R + 255) / 256), dim3(256), 0, s, R, n_flat, n_dark, d_flat, d_dark, fd);
  const long rows = (long)batch * A;
  const bool vec4 = (R % 4 == 0) && ((uintptr_t)d_raw % 16 == 0) && ((uintptr_t)d_out % 16 == 0);
  const int per_row = vec4 ? R / 4 : R;
  const dim3 grid((unsigned)((per_row + 255) / 256), (unsigned)std::min<long>(rows, 65535));
  if (vec4)
    hipLaunchKernelGGL(k_normalize<true>, grid, dim3(256), 0, s, rows, R, d_raw, fd, min_transmission,
                       1.0f / scale, d_out);
  else
    hipLaunchKernelGGL(k_normalize<false>, grid, dim3(256), 0, s, rows, R, d_raw, fd, min_transmission,
                       1.0f / scale, d_out);
  return launched(c, "normalisation kernels");
}

int sk_row_centroid(sk_ctx* ctx, int64_t rows, int R, const float* d_sino, double* d_com, void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (rows < 1 || R < 1) return fail(c, SK_EINVAL, "rows and bins must be >= 1");
  if (!d_sino || !d_com) return fail(c, SK_EINVAL, "device pointer is NULL");
  SK_HIP(c, hipSetDevice(c->device));
  const long blocks = (rows + 3) / 4;
  if (blocks > 0x7fffffffL) return fail(c, SK_EINVAL, "too many rows");
  hipLaunchKernelGGL(k_row_centroid, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (long)rows, R,
                     d_sino, d_com);
  return launched(c, "centroid kernel");
}

}  // extern "C"
