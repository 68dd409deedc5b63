// sk_proj.hip -- the per-call projector kernels and their C-ABI entry points.
//
//   k_fp  A.x   (futhark/forwardprojection.fut:21-84).  Ray-parallel: one lane per ray (or
//               two, variant 1), a wavefront = 64 consecutive bins of one line.  All lanes
//               of a wave step the same frame row together, so the two gathers per step
//               land in one or two cache lines of that row; each row is addressed through
//               a buffer descriptor whose range check returns 0 outside the row, the
//               reference's contribution of an out-of-bounds pixel.  The row range is
//               clipped per wave to the rows where any of its rays can touch the image
//               (skipped rows contribute exactly +0 in the reference).  Flat lines read a
//               transposed copy of the image, the reference's own frame trick
//               (forwardprojection.fut:80-82).
//   k_fp_staged A.x, staged (variants 5, 6): a workgroup = 256 consecutive bins of one
//               line.  It walks the frame rows in blocks of FPS_ROWS (12): the four waves load the
//               block's image-row segment (the columns any of its rays can touch) with
//               coalesced dword loads into LDS, then every lane forms its ray's row steps
//               from LDS.  Same per-step expressions as k_fp (bitwise equal results).
//   k_fp_grouped A.x, grouped staged (variant 7): a workgroup = 64 consecutive bins of four
//               lines adjacent in slope, one wave per line, so the staged segment of a
//               block is two to four chunks of 64 columns instead of six; blocks of 16
//               rows.  Same walk.  k_fp_grouped_res: its SIRT-residual instance.
//   k_bp  B.r   (futhark/backprojection.fut:20-93).  Pixel-parallel gather, no atomics: a
//               block owns a 16 x 16*PPT pixel tile and walks the lines in order, staging
//               for each block of 32 lines the tile's rho-window of the sinogram in LDS
//               together with the per-bin intercepts r/cost.
//
// The work items themselves live in sk_kern.h (shared with the persistent kernels).
// With SK_OPT_EXACT the results are bitwise those of the sequential FP32 restatement of
// the reference.  By default (fast mode) the per-step quotients use the hardware
// reciprocal, the forward projector accumulates the fractions of each row's length and
// scales by the line length once per ray, and the backprojector folds the line length into
// the staged sinogram value; pixel indices and candidate sets are unchanged.
#include "sk_kern.h"

namespace sk {

// ------------------------------------------------------------------------ kernels
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ in,
                                                   float* __restrict__ out, int N) {
  __shared__ float t[32][33];
  const size_t off = (size_t)blockIdx.z * N * N;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int y = by + k, x = bx + tx;
    if (x < N && y < N) t[k][tx] = in[off + (size_t)y * N + x];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int y = bx + k, x = by + tx;
    if (x < N && y < N) out[off + (size_t)y * N + x] = t[tx][k];
  }
}

// Workgroup -> (line, bin block) orders.  ORDER 0: grid (bin blocks, lines, images).
// ORDER 1: grid (lines, bin blocks, images): the workgroups resident together are one bin
// block of (nearly) every angle -- one ring of the image at a time, with each row touched by
// the rays of many neighbouring angles while it is cached.  (Permuting the lines so that
// each XCD gets a contiguous run of angles measured 1.2x slower, and giving each XCD whole
// bin blocks 1.6x slower, from the imbalance between short edge rays and long central ones.)
template <bool EXACT, int MODE, int WEIGHTS, bool DR1, int RPT, int ORDER = 0>
__global__ __launch_bounds__(FP_BLOCK) void k_fp(FpParams P) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if constexpr (ORDER == 0)
    fp_item<EXACT, MODE, WEIGHTS, DR1, RPT>(P, blockIdx.y, blockIdx.x * FP_WAVES + wave, blockIdx.z);
  else
    fp_item<EXACT, MODE, WEIGHTS, DR1, RPT>(P, blockIdx.x, blockIdx.y * FP_WAVES + wave, blockIdx.z);
}

// ------------------------------------------------------------- staged A.x (fast mode)
// A workgroup holds FPS_RAYS consecutive bins of one line, one ray per lane.  Frame rows go
// in blocks of FPS_ROWS: for each row the columns any of the workgroup's rays can touch (a
// segment of at most FPS_MAXW floats starting one column left of the lowest ray) are staged
// in LDS with coalesced loads, zeros outside the image and past the last row, and every
// lane reads its two pixels of the row from LDS.  The gathers thus leave the texture path,
// which bounds the per-wave kernel.  The row arithmetic (fp_row_add, xbot from the
// broadcast i*slope) is that of the per-wave kernel, and rows where a ray is outside the
// image add exactly 0 in both, so the two give bitwise equal results.  NBUF = 2 double-
// buffers the staged block (one barrier per block), NBUF = 1 halves the LDS per workgroup
// (two barriers per block).
// Grid: (lines, bin blocks of FPS_RAYS, images).  Dynamic LDS: NBUF * FPS_CS * FPS_MAXW floats.
template <int MODE, int NBUF>
__global__ __launch_bounds__(FPS_RAYS) void k_fp_staged(FpParams P) {
  extern __shared__ float4 seg4[];
  fps_item<MODE, NBUF>(P, blockIdx.x, blockIdx.y, blockIdx.z, (float*)seg4);
}

// Grouped staged A.x (variant 7): grid (line groups, bin blocks of FPG_BINS, images).
// Dynamic LDS: NBUF * (FPG_ROWS + 1) * 64 * (the geometry's largest CPL) floats.
template <int MODE, int NBUF>
__device__ __forceinline__ void fpg_block(const FpParams& P) {
  extern __shared__ float4 seg4[];
  const int4 grp = P.groups[blockIdx.x];
  const int cpl = grp.x >> 24;  // uniform
  if (cpl <= 2)
    fpg_item<MODE, NBUF, 2>(P, grp, blockIdx.y, blockIdx.z, (float*)seg4);
  else if (cpl == 3)
    fpg_item<MODE, NBUF, 3>(P, grp, blockIdx.y, blockIdx.z, (float*)seg4);
  else
    fpg_item<MODE, NBUF, 4>(P, grp, blockIdx.y, blockIdx.z, (float*)seg4);
}

template <int MODE, int NBUF>
__global__ __launch_bounds__(FPS_RAYS, 8) void k_fp_grouped(FpParams P) {
  fpg_block<MODE, NBUF>(P);
}

// The same for the SIRT residual (FP_SIRT), whose epilogue keeps two more pointers across the
// walk: under the 8-workgroups-per-CU bound's scalar-register budget (about 80) they spilled
// 14 scalar registers into a vector register's lanes and one vector register to scratch,
// reloaded once per row block.  64 vector registers (still 8 waves per SIMD) without that
// bound's scalar limit: 84 scalar registers, no spill.
template <int NBUF>
__global__ __launch_bounds__(FPS_RAYS) __attribute__((amdgpu_num_vgpr(64))) void k_fp_grouped_res(FpParams P) {
  fpg_block<FP_SIRT, NBUF>(P);
}

// Grid: (tiles across, tiles down, images).  Dynamic LDS: BP_LB * W float2.
template <bool EXACT, int MODE, int WEIGHTS, int NC, bool DR1, int PPT>
__global__ __launch_bounds__(256, NC > 0 ? 8 : 1) void k_bp(BpParams P) {
  extern __shared__ float2 win[];
  __shared__ BpLineLds lsh[BP_LB];
  int tx = blockIdx.x, ty = blockIdx.y;
  if (P.order) {
    const unsigned t = P.order[blockIdx.y * gridDim.x + blockIdx.x];
    tx = (int)(t & 0xffffu);
    ty = (int)(t >> 16);
  }
  bp_item<EXACT, MODE, WEIGHTS, NC, DR1, PPT>(P, tx, ty, blockIdx.z, win, lsh);
}

// ------------------------------------------------------------------------ launchers
int launched(Ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(c, e, what);
  return SK_OK;
}

int launch_transpose(Ctx* c, const float* in, float* out, int N, int batch, hipStream_t s) {
  dim3 grid((N + 31) / 32, (N + 31) / 32, batch);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, in, out, N);
  return launched(c, "transpose kernel");
}

FpParams fp_params(const Geom* g, const FpArgs& a) {
  return FpParams{g->d_lines, g->A, g->R, g->N, g->ncand, g->rho0, g->drho,
                  a.img, a.imgT, a.out, a.b, a.rowinv, g->d_fpg, g->d_fp_steps};
}

// fp variants: 0/1/2 = 1/2/4 rays per lane, bin blocks fastest; 3/4 = 1/2 rays per lane,
// lines fastest.
int fp_rays_per_lane(int variant) {
  return (variant == 1 || variant == 4) ? 2 : (variant == 2 ? 4 : 1);
}
int bp_pixels_per_lane(int variant) { return variant == 1 ? 2 : (variant == 2 ? 4 : 1); }

template <bool EXACT, int MODE, int WEIGHTS, bool DR1, int RPT, int ORDER = 0>
static void fp_go(const Geom* g, const FpArgs& a, hipStream_t s) {
  const int per_block = FP_BLOCK * RPT;
  const unsigned nb = (unsigned)((g->R + per_block - 1) / per_block);
  const dim3 grid = ORDER ? dim3(g->A, nb, a.batch) : dim3(nb, g->A, a.batch);
  hipLaunchKernelGGL((k_fp<EXACT, MODE, WEIGHTS, DR1, RPT, ORDER>), grid, dim3(FP_BLOCK), 0, s,
                     fp_params(g, a));
}

// Exact mode runs one ray per lane (its IEEE divisions make more rays per lane no faster) in
// either grid order.
template <bool EXACT, int MODE>
static void fp_go_v(int variant, const Geom* g, const FpArgs& a, hipStream_t s) {
  if (variant == 3 || (EXACT && variant == 4))
    fp_go<EXACT, MODE, W_REF, true, 1, 1>(g, a, s);
  else if (!EXACT && variant == 1)
    fp_go<EXACT, MODE, W_REF, true, 2>(g, a, s);
  else if (!EXACT && variant == 2)
    fp_go<EXACT, MODE, W_REF, true, 4>(g, a, s);
  else if (!EXACT && variant == 4)
    fp_go<EXACT, MODE, W_REF, true, 2, 1>(g, a, s);
  else
    fp_go<EXACT, MODE, W_REF, true, 1>(g, a, s);
}

template <int MODE, int NBUF>
static void fp_go_staged(const Geom* g, const FpArgs& a, hipStream_t s) {
  const unsigned nb = (unsigned)((g->R + FPS_RAYS - 1) / FPS_RAYS);
  const size_t lds = sizeof(float) * NBUF * FPS_CS * (size_t)FPS_MAXW;
  hipLaunchKernelGGL((k_fp_staged<MODE, NBUF>), dim3(g->A, nb, a.batch), dim3(FPS_RAYS), lds, s,
                     fp_params(g, a));
}

template <int MODE, int NBUF>
static void fp_go_grouped(const Geom* g, const FpArgs& a, hipStream_t s) {
  const unsigned nb = (unsigned)((g->R + FPG_BINS - 1) / FPG_BINS);
  const size_t lds = sizeof(float) * NBUF * (FPG_ROWS + 1) * 64 * (size_t)g->fpg_cpl;
  if constexpr (MODE == FP_SIRT)
    hipLaunchKernelGGL((k_fp_grouped_res<NBUF>), dim3(g->n_fpg, nb, a.batch), dim3(FPS_RAYS), lds, s,
                       fp_params(g, a));
  else
    hipLaunchKernelGGL((k_fp_grouped<MODE, NBUF>), dim3(g->n_fpg, nb, a.batch), dim3(FPS_RAYS), lds, s,
                       fp_params(g, a));
}

int launch_fp(Ctx* c, const Geom* g, FpMode mode, const FpArgs& a, hipStream_t s) {
  const bool ex = c->exact != 0;
  int v = g->fp_variant >= 0 ? g->fp_variant : c->fp_variant;
  if (v == 7) {
    if (!ex && g->d_fpg && mode != FP_ONES_INV) {
      // two buffers, one barrier per block (17 KB of LDS at CPL 2: 8 workgroups per CU)
      if (mode == FP_PLAIN)
        fp_go_grouped<FP_PLAIN, 2>(g, a, s);
      else
        fp_go_grouped<FP_SIRT, 2>(g, a, s);
      return launched(c, "forward projection kernel");
    }
    v = 6;  // no group table for this geometry: the one-line staged kernel
  }
  if (v == 5 || v == 6) {
    if (!ex && g->fps_w > 0 && g->fps_w <= FPS_MAXW && mode != FP_ONES_INV) {
      if (mode == FP_PLAIN)
        v == 5 ? fp_go_staged<FP_PLAIN, 2>(g, a, s) : fp_go_staged<FP_PLAIN, 1>(g, a, s);
      else
        v == 5 ? fp_go_staged<FP_SIRT, 2>(g, a, s) : fp_go_staged<FP_SIRT, 1>(g, a, s);
      return launched(c, "forward projection kernel");
    }
    v = 3;  // exact mode, or a geometry whose segments do not fit: the per-wave kernel
  }
  switch (mode) {
    case FP_PLAIN: ex ? fp_go_v<true, FP_PLAIN>(v, g, a, s) : fp_go_v<false, FP_PLAIN>(v, g, a, s); break;
    case FP_ONES_INV:
      ex ? fp_go<true, FP_ONES_INV, W_REF, true, 1>(g, a, s) : fp_go<false, FP_ONES_INV, W_REF, true, 1>(g, a, s);
      break;
    case FP_SIRT: ex ? fp_go_v<true, FP_SIRT>(v, g, a, s) : fp_go_v<false, FP_SIRT>(v, g, a, s); break;
  }
  return launched(c, "forward projection kernel");
}

int launch_fp_bpt(Ctx* c, const Geom* g, const FpArgs& a, hipStream_t s) {
  const bool ex = c->exact != 0;
  const bool dr1 = g->drho == 1.0f;
  if (ex)
    dr1 ? fp_go<true, FP_PLAIN, W_ADJ, true, 1>(g, a, s) : fp_go<true, FP_PLAIN, W_ADJ, false, 1>(g, a, s);
  else
    dr1 ? fp_go<false, FP_PLAIN, W_ADJ, true, 1>(g, a, s) : fp_go<false, FP_PLAIN, W_ADJ, false, 1>(g, a, s);
  return launched(c, "backprojection-transpose kernel");
}

BpParams bp_params(const Geom* g, const BpArgs& a, int ppt) {
  const unsigned* order =
      g->d_bp_order ? g->d_bp_order + g->bp_order_off[ppt == 1 ? 0 : (ppt == 2 ? 1 : 2)] : nullptr;
  return BpParams{g->d_lines, g->A, g->R, g->N, g->ncand, bp_window_bins(ppt, g->drho, g->ncand),
                  g->n_steep, g->rho0, g->drho, a.sino, a.out, a.outT, a.colinv, g->d_inter, order};
}

template <bool EXACT, int MODE, int WEIGHTS, int NC, bool DR1, int PPT>
static void bp_go(const Geom* g, const BpArgs& a, hipStream_t s) {
  const BpParams P = bp_params(g, a, PPT);
  const int TH = 16 * PPT;
  dim3 grid((g->N + BP_TW - 1) / BP_TW, (g->N + TH - 1) / TH, a.batch);
  const size_t lds = sizeof(float2) * BP_LB * (size_t)P.W;
  hipLaunchKernelGGL((k_bp<EXACT, MODE, WEIGHTS, NC, DR1, PPT>), grid, dim3(256), lds, s, P);
}

template <bool EXACT, int MODE, int WEIGHTS, int PPT>
static void bp_go_nc(const Geom* g, const BpArgs& a, hipStream_t s) {
  if (g->bp_two_cand)
    bp_go<EXACT, MODE, WEIGHTS, 2, true, PPT>(g, a, s);
  else
    bp_go<EXACT, MODE, WEIGHTS, 0, false, PPT>(g, a, s);
}

template <bool EXACT, int MODE, int WEIGHTS>
static void bp_go_v(int variant, const Geom* g, const BpArgs& a, hipStream_t s) {
  if (WEIGHTS != W_REF || variant == 0)
    bp_go_nc<EXACT, MODE, WEIGHTS, 1>(g, a, s);
  else if (variant == 1)
    bp_go_nc<EXACT, MODE, WEIGHTS, 2>(g, a, s);
  else
    bp_go_nc<EXACT, MODE, WEIGHTS, 4>(g, a, s);
}

template <int MODE, int WEIGHTS>
static void bp_go_ex(bool ex, int variant, const Geom* g, const BpArgs& a, hipStream_t s) {
  if (ex)
    bp_go_v<true, MODE, WEIGHTS>(variant, g, a, s);
  else
    bp_go_v<false, MODE, WEIGHTS>(variant, g, a, s);
}

int launch_bp(Ctx* c, const Geom* g, BpMode mode, BpWeights w, const BpArgs& a, hipStream_t s) {
  const bool ex = c->exact != 0;
  int v = g->bp_variant >= 0 ? g->bp_variant : c->bp_variant;
  if (sizeof(float2) * BP_LB * (size_t)bp_window_bins(bp_pixels_per_lane(v), g->drho, g->ncand) > 60 * 1024)
    v = 0;
  if (sizeof(float2) * BP_LB * (size_t)bp_window_bins(1, g->drho, g->ncand) > 60 * 1024)
    return fail(c, SK_EINVAL, "backprojection window exceeds LDS (deltarho too small)");
  if (w == W_REF) {
    switch (mode) {
      case BP_PLAIN: bp_go_ex<BP_PLAIN, W_REF>(ex, v, g, a, s); break;
      case BP_ONES_INV: bp_go_ex<BP_ONES_INV, W_REF>(ex, 0, g, a, s); break;
      case BP_SIRT: bp_go_ex<BP_SIRT, W_REF>(ex, v, g, a, s); break;
    }
  } else {
    if (mode != BP_PLAIN) return fail(c, SK_EINVAL, "adjoint weights support plain mode only");
    bp_go_ex<BP_PLAIN, W_ADJ>(ex, 0, g, a, s);
  }
  return launched(c, "backprojection kernel");
}

}  // namespace sk

using namespace sk;

static int check_common(sk_ctx* ctx, sk_geom* g, int batch, const void* p1, const void* p2) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (!g) return fail(c, SK_EINVAL, "geometry is NULL");
  if (batch < 1 || batch > 65535) return fail(c, SK_EINVAL, "batch must be in [1, 65535]");
  if (!p1 || !p2) return fail(c, SK_EINVAL, "device pointer is NULL");
  if (g->g.device != c->device) return fail(c, SK_EINVAL, "geometry belongs to another device");
  SK_HIP(c, hipSetDevice(c->device));
  return SK_OK;
}

extern "C" {

int sk_fp(sk_ctx* ctx, sk_geom* g, int batch, const float* d_img, float* d_sino, void* stream) {
  int rc = check_common(ctx, g, batch, d_img, d_sino);
  if (rc) return rc;
  Ctx* c = &ctx->c;
  const Geom& G = g->g;
  hipStream_t s = (hipStream_t)stream;
  const float* imgT = d_img;
  if (G.n_steep < G.A) {
    void* ws = nullptr;
    rc = scratch(c, sizeof(float) * (size_t)batch * G.N * G.N, s, &ws);
    if (rc) return rc;
    rc = launch_transpose(c, d_img, (float*)ws, G.N, batch, s);
    if (rc) return rc;
    imgT = (const float*)ws;
  }
  FpArgs a{d_img, imgT, d_sino, nullptr, nullptr, batch};
  return launch_fp(c, &G, FP_PLAIN, a, s);
}

int sk_bp(sk_ctx* ctx, sk_geom* g, int batch, const float* d_sino, float* d_img, void* stream) {
  int rc = check_common(ctx, g, batch, d_sino, d_img);
  if (rc) return rc;
  BpArgs a{d_sino, d_img, nullptr, nullptr, batch};
  return launch_bp(&ctx->c, &g->g, BP_PLAIN, W_REF, a, (hipStream_t)stream);
}

int sk_bp_transpose(sk_ctx* ctx, sk_geom* g, int batch, const float* d_img, float* d_sino,
                    void* stream) {
  int rc = check_common(ctx, g, batch, d_img, d_sino);
  if (rc) return rc;
  Ctx* c = &ctx->c;
  const Geom& G = g->g;
  hipStream_t s = (hipStream_t)stream;
  const float* imgT = d_img;
  if (G.n_steep < G.A) {
    void* ws = nullptr;
    rc = scratch(c, sizeof(float) * (size_t)batch * G.N * G.N, s, &ws);
    if (rc) return rc;
    rc = launch_transpose(c, d_img, (float*)ws, G.N, batch, s);
    if (rc) return rc;
    imgT = (const float*)ws;
  }
  FpArgs a{d_img, imgT, d_sino, nullptr, nullptr, batch};
  return launch_fp_bpt(c, &G, a, s);
}

int sk_bp_adjoint(sk_ctx* ctx, sk_geom* g, int batch, const float* d_sino, float* d_img,
                  void* stream) {
  int rc = check_common(ctx, g, batch, d_sino, d_img);
  if (rc) return rc;
  BpArgs a{d_sino, d_img, nullptr, nullptr, batch};
  return launch_bp(&ctx->c, &g->g, BP_PLAIN, W_ADJ, a, (hipStream_t)stream);
}

}  // extern "C"
