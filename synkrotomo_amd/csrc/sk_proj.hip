// sk_proj.hip -- the projector kernels for gfx950 and their C-ABI entry points.
//
//   k_fp  A.x   (futhark/forwardprojection.fut:21-84).  Ray-parallel: one lane per ray,
//               a wavefront = 64 consecutive bins of one line.  All lanes of a wave step
//               the same frame row together, so the two gathers per step land in one or
//               two cache lines of that row.  The row range is clipped per wave to the
//               rows where any of its rays can touch the image (skipped rows contribute
//               exactly +0 in the reference).  Flat lines read a transposed copy of the
//               image, the reference's own frame trick (forwardprojection.fut:80-82).
//   k_bp  B.r   (futhark/backprojection.fut:20-93).  Pixel-parallel gather, no atomics: a
//               block owns a 16x16 pixel tile and walks the lines in order, staging for
//               each block of 32 lines the tile's rho-window of the sinogram in LDS
//               together with the per-bin intercepts r/cost.  Steep and flat lines are
//               accumulated separately and summed at the end, as the reference adds the
//               transposed flat half (backprojection.fut:75-78).
//
// Arithmetic follows the reference expression by expression (compiled with
// -ffp-contract=off so nothing is fused).  With SK_OPT_EXACT the results are bitwise
// those of the sequential FP32 restatement; by default the per-step quotients
// (Xmax-xbot)/xdiff and m/diff use the hardware reciprocal.
#include "sk_common.h"

namespace sk {

constexpr int FP_BLOCK = 256;
constexpr int BP_T = 16;    // pixel tile side
constexpr int BP_LB = 32;   // lines per staged block

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Conservative frame-row range [i0, i1) outside which both candidate pixels of a ray with
// this slope and intercept are out of bounds (|x| > h + 2 with a 2-row margin).  Both
// ends move monotonically (in the same direction) with the intercept, so the hull of the
// ranges of a wave's first and last rays covers every ray between them.  Near-zero slopes
// (theta ~ 0, pi/2) take the full range: there a ray is either in for every row or in for
// none, and an "empty" range would break that monotonicity.
__device__ __forceinline__ void row_clip(float slope, float icpt, int h, int& i0, int& i1) {
  const float lo = -(float)h - 2.0f, hi = (float)h + 2.0f;
  if (fabsf(slope) < 1e-6f) {
    i0 = -h;
    i1 = h;
    return;
  }
  const float ta = (lo - icpt) / slope, tb = (hi - icpt) / slope;
  float tmin = fminf(ta, tb), tmax = fmaxf(ta, tb);
  tmin = fmaxf(tmin, -(float)h - 4.0f);
  tmax = fminf(tmax, (float)h + 4.0f);
  i0 = max(-h, (int)floorf(tmin) - 2);
  i1 = min(h, (int)ceilf(tmax) + 3);
}

// ------------------------------------------------------------------------ transpose
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

// ------------------------------------------------------------------------ forward projector
// First candidate ray of a pixel (backprojection.fut:51-54).
template <bool DR1>
__device__ __forceinline__ float cand_s(const Line& L, float xmin, float ymin, float rho0,
                                        float drho) {
  const float px = (xmin + 0.5f) - L.fc;
  const float py = (ymin + 0.5f) - L.fs;
  const float rho = L.cost * px + L.sint * py;
  float t = rho - rho0;
  if constexpr (!DR1) t = t / drho;
  return ceilf(t);
}

// Length B[pixel, ray] of the reference backprojector for the pixel in frame column X of
// frame row y (= the ray's row): intersect_fact of the ray's [xbot, xtop] with [X, X+1),
// if the ray is one of the pixel's candidates s..s+ncand-1 (backprojection.fut:50-62).
template <bool EXACT, bool DR1>
__device__ __forceinline__ float bp_entry(const Line& L, int X, float y, float xbot, float xtop,
                                          int r, int ncand, float rho0, float drho) {
  const float xmin = (float)X;
  const float s = cand_s<DR1>(L, xmin, y, rho0, drho);
  const int si = (int)s;
  const float maxx = fmaxf(xbot, xtop), minx = fminf(xbot, xtop);
  const float lo = fmaxf(maxx - xmin, 0.0f);
  const float hi = fmaxf((xmin + 1.0f) - minx, 0.0f);
  const float diff = maxx - minx;
  const float m = fminf(lo, hi);
  float p;
  if constexpr (EXACT)
    p = (m == 0.0f) ? 0.0f : m / diff;
  else
    p = (m == 0.0f) ? 0.0f : m * fast_rcp(diff);
  const float l = fminf(p, 1.0f) * L.lbase;
  return (r >= si && r < si + ncand) ? l : 0.0f;
}

// WEIGHTS == W_REF: A (the forward projector).  WEIGHTS == W_ADJ here means B^T: the same
// ray walk with the backprojector's lengths (the exact adjoint of k_bp).
template <bool EXACT, int MODE, int WEIGHTS = W_REF, bool DR1 = true>
__global__ __launch_bounds__(FP_BLOCK) void k_fp(const Line* __restrict__ lines, int A, int R,
                                                 int N, float rho0, float drho, int ncand,
                                                 const float* __restrict__ img,
                                                 const float* __restrict__ imgT,
                                                 float* __restrict__ out,
                                                 const float* __restrict__ bsino,
                                                 const float* __restrict__ rowinv) {
  const int j = blockIdx.y;
  const int bz = blockIdx.z;
  const int r = blockIdx.x * FP_BLOCK + threadIdx.x;
  const int wave_first = blockIdx.x * FP_BLOCK + (threadIdx.x & ~63);
  if (wave_first >= R) return;  // whole wave past the last bin
  const int wave_last = min(wave_first + 63, R - 1);
  const Line L = lines[j];
  const int h = N >> 1;
  const float cost = L.cost, slope = L.slope, lbase = L.lbase;
  const bool active = r < R;
  // forwardprojection.fut:33-34
  const float icpt = (rho0 + (float)r * drho) / cost;

  int i0, i1;
  {
    int a0, a1, b0, b1;
    row_clip(slope, (rho0 + (float)wave_first * drho) / cost, h, a0, a1);
    row_clip(slope, (rho0 + (float)wave_last * drho) / cost, h, b0, b1);
    i0 = __builtin_amdgcn_readfirstlane(min(a0, b0));
    i1 = __builtin_amdgcn_readfirstlane(max(a1, b1));
  }

  const float* src = ((L.flags & LINE_FLAT) ? imgT : img) + (size_t)bz * N * N + h;
  float sum = 0.0f;
  for (int i = i0; i < i1; ++i) {
    // forwardprojection.fut:37-56
    const float xbot = (float)i * slope + icpt;
    const float xtop = xbot + slope;
    const float fXb = floorf(xbot);
    const float fXt = floorf(xtop);
    const int Xb = (int)fXb;
    const int Xt = (int)fXt;
    const float Xmax = fmaxf(fXb, fXt);
    const float xdiff = xtop - xbot;
    const bool single = Xb == Xt;
    const bool inb = active && (unsigned)(Xb + h) < (unsigned)N;
    const bool intop = active && !single && (unsigned)(Xt + h) < (unsigned)N;
    float lb, lt;
    if constexpr (WEIGHTS == W_ADJ) {
      lb = bp_entry<EXACT, DR1>(L, Xb, (float)i, xbot, xtop, r, ncand, rho0, drho);
      lt = bp_entry<EXACT, DR1>(L, Xt, (float)i, xbot, xtop, r, ncand, rho0, drho);
    } else if constexpr (EXACT) {
      lb = single ? lbase : ((Xmax - xbot) / xdiff) * lbase;
      lt = ((xtop - Xmax) / xdiff) * lbase;
    } else {
      const float rd = fast_rcp(xdiff);
      lb = single ? lbase : ((Xmax - xbot) * rd) * lbase;
      lt = ((xtop - Xmax) * rd) * lbase;
    }
    const float* row = src + (size_t)(i + h) * N;
    float bot = 0.0f, top = 0.0f;
    if constexpr (MODE == FP_ONES_INV) {
      if (inb) bot = lb * 1.0f;
      if (intop) top = lt * 1.0f;
    } else {
      if (inb) bot = lb * row[Xb];
      if (intop) top = lt * row[Xt];
    }
    sum = sum + (bot + top);
  }
  if (!active) return;
  const size_t o = (size_t)bz * A * R + (size_t)L.orig * R + r;
  if constexpr (MODE == FP_PLAIN) {
    out[o] = sum;
  } else if constexpr (MODE == FP_ONES_INV) {
    out[o] = (sum == 0.0f) ? 0.0f : 1.0f / sum;  // sirtlib.fut:9-10
  } else {
    // SIRT.fut:33-36: R * (b - A x)
    out[o] = rowinv[(size_t)L.orig * R + r] * (bsino[o] - sum);
  }
}

// ------------------------------------------------------------------------ backprojector
// Length of one candidate ray in one pixel.
//   W_REF: intersect_fact(max, min, xmin, xmin+1) * lbase (backprojection.fut:22-33,59-62)
//   W_ADJ: the forward projector's length for this (ray, row) if its Xb or Xt is this
//          column (forwardprojection.fut:37-54), i.e. the entry of A.
template <bool EXACT, int WEIGHTS>
__device__ __forceinline__ float cand_len(float xbot, float slope, float lbase, float xmin,
                                          float xmin1, bool degen) {
  const float xtop = xbot + slope;
  if constexpr (WEIGHTS == W_REF) {
    const float maxx = fmaxf(xbot, xtop);
    const float minx = fminf(xbot, xtop);
    const float diff = maxx - minx;
    float p;
    if constexpr (EXACT) {
      const float lo = fmaxf(maxx - xmin, 0.0f);
      const float hi = fmaxf(xmin1 - minx, 0.0f);
      const float m = fminf(lo, hi);
      p = (m == 0.0f) ? 0.0f : m / diff;
    } else {
      // min(max(a,0), max(b,0)) == max(min(a,b), 0)
      const float m = fmaxf(fminf(maxx - xmin, xmin1 - minx), 0.0f);
      // diff > 0 on non-degenerate lines, where m == 0 already gives 0.
      p = m * fast_rcp(diff);
      if (degen) p = (m == 0.0f) ? 0.0f : p;
    }
    return fminf(p, 1.0f) * lbase;
  } else {
    const float fXb = floorf(xbot);
    const float fXt = floorf(xtop);
    const bool single = fXb == fXt;
    const float Xmax = fmaxf(fXb, fXt);
    const float xdiff = xtop - xbot;
    float lb, lt;
    if constexpr (EXACT) {
      lb = single ? lbase : ((Xmax - xbot) / xdiff) * lbase;
      lt = ((xtop - Xmax) / xdiff) * lbase;
    } else {
      const float rd = fast_rcp(xdiff);
      lb = single ? lbase : ((Xmax - xbot) * rd) * lbase;
      lt = ((xtop - Xmax) * rd) * lbase;
    }
    return (fXb == xmin) ? lb : ((!single && fXt == xmin) ? lt : 0.0f);
  }
}

template <bool EXACT, int MODE, int WEIGHTS, int NC, bool DR1>
__global__ __launch_bounds__(256) void k_bp(const Line* __restrict__ lines, int A, int R, int N,
                                            float rho0, float drho, int ncand_rt, int W,
                                            const float* __restrict__ sino,
                                            float* __restrict__ out, float* __restrict__ outT,
                                            const float* __restrict__ colinv) {
  extern __shared__ float2 win[];  // [BP_LB][W]: (r/cost, sinogram value)
  __shared__ int w0s[BP_LB];
  const int ncand = NC > 0 ? NC : ncand_rt;
  const int tx = threadIdx.x & (BP_T - 1), ty = threadIdx.x / BP_T;
  const int X0 = blockIdx.x * BP_T, Y0 = blockIdx.y * BP_T;
  const int bz = blockIdx.z;
  const int h = N >> 1;
  const int X = X0 + tx, Y = Y0 + ty;
  const float* sb = sino + (size_t)bz * A * R;
  const float xs = (float)(X - h), ys = (float)(Y - h);

  float acc_s = 0.0f, acc_f = 0.0f;
  for (int l0 = 0; l0 < A; l0 += BP_LB) {
    const int nl = min(BP_LB, A - l0);
    if ((int)threadIdx.x < nl) {
      // Window origin: s is monotone in each frame coordinate, so its minimum over the
      // tile is at a corner.
      const Line L = lines[l0 + threadIdx.x];
      const bool fl = L.flags & LINE_FLAT;
      const float cx0 = (float)((fl ? Y0 : X0) - h), cy0 = (float)((fl ? X0 : Y0) - h);
      const float cx1 = cx0 + (float)(BP_T - 1), cy1 = cy0 + (float)(BP_T - 1);
      float smin = cand_s<DR1>(L, cx0, cy0, rho0, drho);
      smin = fminf(smin, cand_s<DR1>(L, cx1, cy0, rho0, drho));
      smin = fminf(smin, cand_s<DR1>(L, cx0, cy1, rho0, drho));
      smin = fminf(smin, cand_s<DR1>(L, cx1, cy1, rho0, drho));
      w0s[threadIdx.x] = (int)smin;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nl * W; e += 256) {
      const int jj = e / W, k = e - jj * W;
      const Line L = lines[l0 + jj];
      const int sp = w0s[jj] + k;
      // backprojection.fut:57-58: r = s'*drho + rho0, then r/cost
      const float inter = ((float)sp * drho + rho0) / L.cost;
      float p = 0.0f;
      if (sp >= 0 && sp < R) p = (MODE == BP_ONES_INV) ? 1.0f : sb[(size_t)L.orig * R + sp];
      win[jj * W + k] = make_float2(inter, p);
    }
    __syncthreads();
    for (int jj = 0; jj < nl; ++jj) {
      const Line L = lines[l0 + jj];
      const bool fl = L.flags & LINE_FLAT;
      const bool degen = L.flags & LINE_DEGEN;
      const float xmin = fl ? ys : xs;
      const float ymin = fl ? xs : ys;
      const float s = cand_s<DR1>(L, xmin, ymin, rho0, drho);
      const float xbase = ymin * L.slope;  // backprojection.fut:55
      // s - w0 lies in [0, W - ncand] (monotone s, window sized on the host); the clamp
      // only guards geometries so far off-centre that rho rounds by whole bins.
      const int off = min(max((int)s - w0s[jj], 0), W - ncand);
      const float2* wrow = win + jj * W + off;
      float inner = 0.0f;
#pragma unroll
      for (int k = 0; k < (NC > 0 ? NC : 16); ++k) {
        if (NC == 0 && k >= ncand) break;
        const float2 e = wrow[k];
        const float l =
            cand_len<EXACT, WEIGHTS>(xbase + e.x, L.slope, L.lbase, xmin, xmin + 1.0f, degen);
        inner = inner + l * e.y;
      }
      if (fl)
        acc_f = acc_f + inner;
      else
        acc_s = acc_s + inner;
    }
    __syncthreads();
  }
  if (X >= N || Y >= N) return;
  const size_t p = (size_t)bz * N * N + (size_t)Y * N + X;
  const float v = acc_s + acc_f;  // backprojection.fut:78
  if constexpr (MODE == BP_PLAIN) {
    out[p] = v;
  } else if constexpr (MODE == BP_ONES_INV) {
    out[p] = (v == 0.0f) ? 0.0f : 1.0f / v;
  } else {
    // SIRT.fut:37-39: x + C * B(q)
    const float xn = out[p] + colinv[(size_t)Y * N + X] * v;
    out[p] = xn;
    outT[(size_t)bz * N * N + (size_t)X * N + Y] = xn;
  }
}

// ------------------------------------------------------------------------ launchers
static int launched(Ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(c, e, what);
  return SK_OK;
}

int launch_transpose(Ctx* c, const float* in, float* out, int N, int batch, hipStream_t s) {
  dim3 grid((N + 31) / 32, (N + 31) / 32, batch);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, in, out, N);
  return launched(c, "transpose kernel");
}

template <bool EXACT, int MODE, int WEIGHTS = W_REF, bool DR1 = true>
static void fp_go(const Geom* g, const FpArgs& a, hipStream_t s) {
  dim3 grid((g->R + FP_BLOCK - 1) / FP_BLOCK, g->A, a.batch);
  hipLaunchKernelGGL((k_fp<EXACT, MODE, WEIGHTS, DR1>), grid, dim3(FP_BLOCK), 0, s, g->d_lines,
                     g->A, g->R, g->N, g->rho0, g->drho, g->ncand, a.img, a.imgT, a.out, a.b,
                     a.rowinv);
}

int launch_fp_bpt(Ctx* c, const Geom* g, const FpArgs& a, hipStream_t s) {
  const bool ex = c->exact != 0;
  const bool dr1 = g->drho == 1.0f;
  if (ex)
    dr1 ? fp_go<true, FP_PLAIN, W_ADJ, true>(g, a, s) : fp_go<true, FP_PLAIN, W_ADJ, false>(g, a, s);
  else
    dr1 ? fp_go<false, FP_PLAIN, W_ADJ, true>(g, a, s) : fp_go<false, FP_PLAIN, W_ADJ, false>(g, a, s);
  return launched(c, "backprojection-transpose kernel");
}

int launch_fp(Ctx* c, const Geom* g, FpMode mode, const FpArgs& a, hipStream_t s) {
  const bool ex = c->exact != 0;
  switch (mode) {
    case FP_PLAIN: ex ? fp_go<true, FP_PLAIN>(g, a, s) : fp_go<false, FP_PLAIN>(g, a, s); break;
    case FP_ONES_INV:
      ex ? fp_go<true, FP_ONES_INV>(g, a, s) : fp_go<false, FP_ONES_INV>(g, a, s);
      break;
    case FP_SIRT: ex ? fp_go<true, FP_SIRT>(g, a, s) : fp_go<false, FP_SIRT>(g, a, s); break;
  }
  return launched(c, "forward projection kernel");
}

// LDS window width (in bins) that covers every candidate of a 16x16 tile.
static int bp_window(const Geom* g) {
  const double span = (BP_T - 1) * 1.41422 / (double)g->drho;
  int W = (int)std::ceil(span) + g->ncand + 3;
  return (W + 1) & ~1;
}

template <bool EXACT, int MODE, int WEIGHTS, int NC, bool DR1>
static void bp_go(const Geom* g, const BpArgs& a, int W, hipStream_t s) {
  dim3 grid((g->N + BP_T - 1) / BP_T, (g->N + BP_T - 1) / BP_T, a.batch);
  const size_t lds = sizeof(float2) * BP_LB * (size_t)W;
  hipLaunchKernelGGL((k_bp<EXACT, MODE, WEIGHTS, NC, DR1>), grid, dim3(256), lds, s, g->d_lines,
                     g->A, g->R, g->N, g->rho0, g->drho, g->ncand, W, a.sino, a.out, a.outT,
                     a.colinv);
}

template <bool EXACT, int MODE, int WEIGHTS>
static void bp_go_nc(const Geom* g, const BpArgs& a, int W, hipStream_t s) {
  if (g->drho == 1.0f && g->ncand == 2)
    bp_go<EXACT, MODE, WEIGHTS, 2, true>(g, a, W, s);
  else
    bp_go<EXACT, MODE, WEIGHTS, 0, false>(g, a, W, s);
}

template <int MODE, int WEIGHTS>
static void bp_go_ex(bool ex, const Geom* g, const BpArgs& a, int W, hipStream_t s) {
  if (ex)
    bp_go_nc<true, MODE, WEIGHTS>(g, a, W, s);
  else
    bp_go_nc<false, MODE, WEIGHTS>(g, a, W, s);
}

int launch_bp(Ctx* c, const Geom* g, BpMode mode, BpWeights w, const BpArgs& a, hipStream_t s) {
  const bool ex = c->exact != 0;
  const int W = bp_window(g);
  if (sizeof(float2) * BP_LB * (size_t)W > 60 * 1024)
    return fail(c, SK_EINVAL, "backprojection window of %d bins exceeds LDS", W);
  if (w == W_REF) {
    switch (mode) {
      case BP_PLAIN: bp_go_ex<BP_PLAIN, W_REF>(ex, g, a, W, s); break;
      case BP_ONES_INV: bp_go_ex<BP_ONES_INV, W_REF>(ex, g, a, W, s); break;
      case BP_SIRT: bp_go_ex<BP_SIRT, W_REF>(ex, g, a, W, s); break;
    }
  } else {
    if (mode != BP_PLAIN) return fail(c, SK_EINVAL, "adjoint weights support plain mode only");
    bp_go_ex<BP_PLAIN, W_ADJ>(ex, g, a, W, s);
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
