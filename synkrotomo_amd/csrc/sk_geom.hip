// sk_geom.hip -- context, errors, scratch memory and the per-angle line table.
//
// The line table replaces sirtlib.preprocess (futhark/sirtlib.fut:28-44), which the
// reference runs as a map + partition + scatter on the device at every call.  Here it is
// built once per geometry on the host (A entries of 32 bytes) and uploaded; the kernels
// index the sinogram through Line::orig, so neither the sinogram partition
// (fix_projections, sirtlib.fut:16-21) nor the output scatter (postprocess_fp,
// sirtlib.fut:24-25) exists.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "sk_common.h"
#include "sk_kern.h"

namespace sk {

int fail(Ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what) {
  int code = (e == hipErrorOutOfMemory) ? SK_ENOMEM : SK_EHIP;
  return fail(c, code, "%s failed: %s", what, hipGetErrorString(e));
}

int scratch(Ctx* c, size_t bytes, hipStream_t stream, void** out) {
  if (bytes > c->scratch_bytes) {
    // Previously enqueued work may still use the old buffer.
    SK_HIP(c, hipStreamSynchronize(stream));
    if (c->scratch) SK_HIP(c, hipFree(c->scratch));
    c->scratch = nullptr;
    c->scratch_bytes = 0;
    size_t want = bytes + (bytes >> 3);
    SK_HIP(c, hipMalloc(&c->scratch, want));
    c->scratch_bytes = want;
  }
  *out = c->scratch;
  return SK_OK;
}

// Host line table.  Same operations and rounding points as sirtlib.fut:29-41:
// cos/sin of the f32 angle, l = sqrt(1 + (a/b)**2) in f32, flat <=> |sin| >= |cos|,
// stable partition steep-then-flat.
// |slope| from which the fast backprojector takes 1/(x_top - x_bot) by one Newton step from
// 1/|slope|.  The candidates that get a fraction strictly between 0 and 1 cross their pixel, so
// |x_bot|, |x_top| <= N/2 + 4.  x_top = fl(x_bot + slope) is off slope + x_bot by at most half an
// ulp of N/2 + 4, and the difference fl(x_top - x_bot) adds a relative rounding of at most 2^-24
// (none when x_top and x_bot are within a factor 2 of each other, by Sterbenz; otherwise it
// rounds once).  So d = |x_top - x_bot| differs from |slope| by a relative
// eps <= ulp(N/2 + 4)/2/|slope| + 2^-24.  With eps <= 4e-4 the step leaves eps^2 <= 1.6e-7, on par
// with v_rcp_f32's 1 ulp.  (0.15 at N = 2048.)
static float newton_min_slope(int N) {
  const float x = (float)(N / 2 + 4);
  const float half_ulp = 0.5f * (nextafterf(x, INFINITY) - x);
  return half_ulp / 4e-4f;
}

// Launch orders of the backprojector's tiles (BP_TW x 16*ppt pixels) for ppt = 1, 2, 4, dearest
// first.  A tile's lines take the exact step (5 VALU per candidate instead of 10) where its
// frame columns lie in one binade (tile_binade: steep lines over the tile's 16 columns, flat
// lines over its 64-row band), so the tiles of a launch differ in cost; starting the dear ones
// first leaves cheap blocks for the end of the launch (longest-first list scheduling).
// Line groups of the grouped staged forward projector: the lines of each frame and sign of
// cost in slope order, in groups of FPG_LINES.  A group's staged width for FPG_BINS consecutive bins covers the spread
// of its rays' x = i*slope + (rho0 + r*drho)/cost over the FPG_ROWS rows of any block: the
// spread (a maximum minus a minimum of functions linear in the block's first row) is convex
// in that row, so the blocks at the image's first and last rows bound it; plus the rounding
// of the intercepts and of x_bot, the pair and the floors (as for fps_w).  Entry: x = first
// line | CPL << 24 (chunks of 64 columns), y, z, w = the other lines or -1.  Empty when a
// group would need more than FPG_MAXCPL chunks.
static std::vector<int4> fpg_groups(const std::vector<Line>& lines, int n_steep, int N, int R, float rho0,
                                    float drho, int* cpl_max) {
  std::vector<int4> out;
  *cpl_max = 0;
  const int A = (int)lines.size(), h = N / 2, nb = (R + FPG_BINS - 1) / FPG_BINS;
  for (int cls = 0; cls < 4; ++cls) {  // frame x sign of cost: lines of opposite directions
    const int frame = cls >> 1;          // put bin r on opposite sides of the frame
    const bool pos = cls & 1;
    std::vector<int> idx;
    for (int j = frame ? n_steep : 0; j < (frame ? A : n_steep); ++j)
      if ((lines[j].cost > 0.0f) == pos) idx.push_back(j);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return lines[a].slope < lines[b].slope; });
    for (size_t g0 = 0; g0 < idx.size(); g0 += FPG_LINES) {
      const int n = (int)std::min<size_t>(FPG_LINES, idx.size() - g0);
      double spread = 0.0, inv = 0.0;
      for (int b = 0; b < nb; ++b) {
        const double r0 = (double)b * FPG_BINS, r1 = r0 + FPG_BINS - 1;
        for (const int ib : {-h, h - 1}) {
          double lo = 1e300, hi = -1e300;
          for (int k = 0; k < n; ++k) {
            const Line& L = lines[idx[g0 + k]];
            inv = std::max(inv, 1.0 / std::fabs((double)L.cost));
            for (const int i : {ib, ib + FPG_ROWS - 1})
              for (const double r : {r0, r1}) {
                const double x = i * (double)L.slope + ((double)rho0 + r * drho) / (double)L.cost;
                lo = std::min(lo, x);
                hi = std::max(hi, x);
              }
          }
          spread = std::max(spread, hi - lo);
        }
      }
      const double round = 4.0 * (std::fabs((double)rho0) + (double)R * drho) * inv * 0x1p-23 + (h + 16) * 0x1p-22;
      const int w = (int)std::ceil(spread + round + 0.01) + 5;
      const int cpl = std::max(2, (w + 63) / 64);
      if (cpl > FPG_MAXCPL) {
        out.clear();
        *cpl_max = 0;
        return out;
      }
      *cpl_max = std::max(*cpl_max, cpl);
      int l[FPG_LINES];
      for (int k = 0; k < FPG_LINES; ++k) l[k] = k < n ? idx[g0 + k] : -1;
      out.push_back(make_int4(l[0] | (cpl << 24), l[1], l[2], l[3]));
    }
  }
  return out;
}

static std::vector<unsigned> bp_tile_orders(int N, int n_steep, int A, size_t off[3]) {
  std::vector<unsigned> all;
  const int h = N / 2, ntx = (N + BP_TW - 1) / BP_TW;
  const int ppts[3] = {1, 2, 4};
  for (int i = 0; i < 3; ++i) {
    const int TH = 16 * ppts[i], nty = (N + TH - 1) / TH;
    std::vector<std::pair<long, unsigned>> t;
    t.reserve((size_t)ntx * nty);
    for (int ty = 0; ty < nty; ++ty)
      for (int tx = 0; tx < ntx; ++tx) {
        // per slot: about 61 VALU per line and lane with the exact step, 100 without
        const int Y64 = (ty * TH) & ~63;
        const long cs = (tile_binade(tx * BP_TW - h, tx * BP_TW - h + BP_TW - 1) >= 0 ? 61 : 100) * ppts[i];
        int ef, band;
        bp_band_rule(Y64 - h, ef, band);
        const long nex = __builtin_popcount((unsigned)((band >> ((ty * TH - Y64) >> 4)) & ((1 << ppts[i]) - 1)));
        const long cf = 61 * nex + 100 * (ppts[i] - nex);
        t.push_back({-(cs * n_steep + cf * (A - n_steep)), ((unsigned)ty << 16) | (unsigned)tx});
      }
    std::stable_sort(t.begin(), t.end(),
                     [](const std::pair<long, unsigned>& a, const std::pair<long, unsigned>& b) {
                       return a.first < b.first;
                     });
    off[i] = all.size();
    for (const auto& p : t) all.push_back(p.second);
  }
  return all;
}

// N > 0: also mark the lines of LINE_NEWTON for an N x N image.
static int build_lines(const float* ang, int A, std::vector<Line>& out, int N = 0) {
  const float fact = sqrtf(2.0f) / 2.0f;
  std::vector<Line> steep, flat;
  steep.reserve(A);
  flat.reserve(A);
  for (int i = 0; i < A; i++) {
    const float c = cosf(ang[i]);
    const float s = sinf(ang[i]);
    const float lsteep = sqrtf(1.0f + powf(s / c, 2.0f));
    const float lflat = sqrtf(1.0f + powf(c / s, 2.0f));
    const bool is_flat = fabsf(s) >= fabsf(c);
    Line L{};
    if (is_flat) {
      L.cost = -s;
      L.sint = c;
      L.lbase = lflat;
    } else {
      L.cost = c;
      L.sint = -s;
      L.lbase = lsteep;
    }
    L.slope = -L.sint / L.cost;
    L.fc = fact * L.cost;
    L.fs = fact * L.sint;
    L.orig = i;
    L.flags = (is_flat ? LINE_FLAT : 0) | (L.slope >= 0.0f ? LINE_SLOPE_POS : 0) |
              (fabsf(L.slope) < 1e-3f ? LINE_DEGEN : 0);
    if (N > 0 && !(L.flags & LINE_DEGEN) && fabsf(L.slope) >= newton_min_slope(N)) L.flags |= LINE_NEWTON;
    (is_flat ? flat : steep).push_back(L);
  }
  out = steep;
  out.insert(out.end(), flat.begin(), flat.end());
  // run lengths (consecutive lines with equal key bits), counted from the back
  for (int j = A - 1; j >= 0; j--) {
    const bool same = j + 1 < A && (out[j + 1].flags & LINE_KEY_BITS) == (out[j].flags & LINE_KEY_BITS);
    const int run = same ? (out[j + 1].flags >> LINE_RUN_SHIFT) + 1 : 1;
    out[j].flags = (out[j].flags & LINE_KEY_BITS) | (run << LINE_RUN_SHIFT);
  }
  return (int)steep.size();
}

// The forward projector's exact per-block step per line j and binade e (fps_walk): (D, 1/D)
// with D = binade_step(slope, e) -- 0 where |slope| is a rounding tie or below half a unit at
// e -- and 1/D by the expression fp_row_add evaluates for a row whose |x_top - x_bot| is D
// (the Newton step from 1/|slope| on LINE_NEWTON lines, else the hardware reciprocal), so a
// block that takes the exact step forms every fraction bit for bit as the per-row step does.
__global__ __launch_bounds__(256) void k_fp_step_table(const Line* __restrict__ lines, int A,
                                                       float2* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A * FP_STEP_BINADES) return;
  const int j = i / FP_STEP_BINADES, e = i - j * FP_STEP_BINADES;
  const Line L = lines[j];
  const float D = binade_step(L.slope, e);
  float r = 0.0f;
  if (D > 0.0f) {
    if (L.flags & LINE_NEWTON) {
      const float rs = 1.0f / fabsf(L.slope);
      r = fmaf(rs, fmaf(-D, rs, 1.0f), rs);
    } else {
      r = fast_rcp(D);
    }
  }
  out[i] = make_float2(D, r);
}

// inter[j][s] = (s*drho + rho0) / cost_j (backprojection.fut:57-58), IEEE division: the values
// the backprojector stages for its candidate windows.
__global__ __launch_bounds__(256) void k_inter_table(const Line* __restrict__ lines, int R, float rho0,
                                                     float drho, float* __restrict__ out) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= R) return;
  out[(size_t)blockIdx.y * R + s] = ((float)s * drho + rho0) / lines[blockIdx.y].cost;
}

}  // namespace sk

using namespace sk;

extern "C" {

int sk_create(int device, sk_ctx** out) {
  if (!out) return SK_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SK_EHIP;
  if (device < 0 || device >= n) return SK_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return SK_EHIP;
  sk_ctx* c = new sk_ctx();
  c->c.device = device;
  if (hipStreamCreateWithFlags(&c->c.capture_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SK_EHIP;
  }
  if (noalloc_init(&c->c, device) != SK_OK) {
    (void)hipStreamDestroy(c->c.capture_stream);
    delete c;
    return SK_EHIP;
  }
  *out = c;
  return SK_OK;
}

void sk_destroy(sk_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  (void)hipDeviceSynchronize();  // work of this context may still use its graphs and scratch
  for (auto& bg : ctx->c.batch_graphs)
    if (bg.exec) (void)hipGraphExecDestroy(bg.exec);
  if (ctx->c.scratch) (void)hipFree(ctx->c.scratch);
  if (ctx->c.capture_stream) (void)hipStreamDestroy(ctx->c.capture_stream);
  delete ctx;
}

const char* sk_last_error(const sk_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int sk_set_option(sk_ctx* ctx, int option, int value) {
  if (!ctx) return SK_EINVAL;
  switch (option) {
    case SK_OPT_EXACT: ctx->c.exact = value ? 1 : 0; return SK_OK;
    case SK_OPT_FP_VARIANT: ctx->c.fp_variant = value; return SK_OK;
    case SK_OPT_BP_VARIANT: ctx->c.bp_variant = value; return SK_OK;
    default: return fail(&ctx->c, SK_EINVAL, "unknown option %d", option);
  }
}

int sk_get_option(const sk_ctx* ctx, int option, int* value) {
  if (!ctx || !value) return SK_EINVAL;
  switch (option) {
    case SK_OPT_EXACT: *value = ctx->c.exact; return SK_OK;
    case SK_OPT_FP_VARIANT: *value = ctx->c.fp_variant; return SK_OK;
    case SK_OPT_BP_VARIANT: *value = ctx->c.bp_variant; return SK_OK;
    default: return SK_EINVAL;
  }
}

// Host-only: the line table as 8 values per line (cost, sint, lbase, slope, fc, fs,
// orig, flags), steep lines first.  Returns the number of steep lines, or -1.
int sk_lines_compute(const float* angles_host, int A, float* table_out) {
  if (!angles_host || A < 1 || !table_out) return -1;
  std::vector<Line> lines;
  int ns = build_lines(angles_host, A, lines);
  for (int j = 0; j < A; j++) {
    const Line& L = lines[j];
    float* t = table_out + 8 * j;
    t[0] = L.cost; t[1] = L.sint; t[2] = L.lbase; t[3] = L.slope;
    t[4] = L.fc; t[5] = L.fs; t[6] = (float)L.orig; t[7] = (float)(L.flags & LINE_FLAG_BITS);
  }
  return ns;
}

int sk_geom_create(sk_ctx* ctx, const float* angles_host, int A, float rho0, float drho, int R,
                   int N, sk_geom** out) {
  if (!ctx || !out) return SK_EINVAL;
  Ctx* c = &ctx->c;
  *out = nullptr;
  if (!angles_host) return fail(c, SK_EINVAL, "angles is NULL");
  if (A < 1 || A > 65535) return fail(c, SK_EINVAL, "angle count must be in [1, 65535] (got %d)", A);
  if (R < 1) return fail(c, SK_EINVAL, "numrhos must be positive (got %d)", R);
  if (N < 2 || (N & 1)) return fail(c, SK_EINVAL, "image size must be even and >= 2 (got %d)", N);
  if (N > 16384) return fail(c, SK_EINVAL, "image size %d exceeds 16384", N);
  if (!(drho > 0.0f) || !std::isfinite(drho)) return fail(c, SK_EINVAL, "deltarho must be > 0");
  if (!std::isfinite(rho0) || fabsf(rho0) > 1e6f)
    return fail(c, SK_EINVAL, "rhozero must be finite and |rhozero| <= 1e6");
  if ((double)R * drho > 4e6) return fail(c, SK_EINVAL, "detector extent numrhos*deltarho too large");
  for (int i = 0; i < A; i++)
    if (!std::isfinite(angles_host[i])) return fail(c, SK_EINVAL, "angle %d is not finite", i);
  const int ncand = (int)ceilf(sqrtf(2.0f) / drho);  // backprojection.fut:87
  if (ncand > 16) return fail(c, SK_EINVAL, "deltarho %g too small (more than 16 rays per pixel)", drho);
  SK_HIP(c, hipSetDevice(c->device));
  static std::atomic<unsigned long long> next_uid{1};
  sk_geom* g = new sk_geom();
  g->g.uid = next_uid++;
  Geom& G = g->g;
  G.A = A; G.R = R; G.N = N; G.rho0 = rho0; G.drho = drho; G.ncand = ncand;
  G.bp_two_cand = drho == 1.0f && ncand == 2 && std::fabs((double)rho0) + N < (double)(1 << 20);
  G.device = c->device;
  G.n_steep = build_lines(angles_host, A, G.lines, N);
  {
    // Staged forward projector: the intercepts of 256 consecutive bins span 255*drho/|cost|
    // columns, plus their rounding; a block of FPS_ROWS rows shares one segment start, so the
    // segment adds the drift of FPS_ROWS - 1 rows and 5 columns (the pair and the floors).
    double inv = 0.0, slope = 0.0;
    for (const Line& L : G.lines) {
      inv = std::max(inv, 1.0 / std::fabs((double)L.cost));
      slope = std::max(slope, std::fabs((double)L.slope));
    }
    const double span = 255.0 * drho * inv + 4.0 * (std::fabs((double)rho0) + (double)R * drho) * inv * 0x1p-23;
    G.fps_w = (int)std::ceil(span + (FPS_ROWS - 1) * slope + 0.01) + 5;
  }
  hipError_t e = hipMalloc(&G.d_lines, sizeof(Line) * (size_t)A);
  if (e == hipSuccess)
    e = hipMemcpy(G.d_lines, G.lines.data(), sizeof(Line) * (size_t)A, hipMemcpyHostToDevice);
  if (e == hipSuccess && (long)A * R < (1L << 31) && A <= 65535) {
    e = hipMalloc(&G.d_inter, sizeof(float) * (size_t)A * R);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_inter_table, dim3((R + 255) / 256, A), dim3(256), 0, nullptr, G.d_lines, R, rho0, drho,
                         G.d_inter);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  }
  if (e == hipSuccess && G.bp_two_cand && N / BP_TW < 65536) {
    const std::vector<unsigned> ord = bp_tile_orders(N, G.n_steep, A, G.bp_order_off);
    e = hipMalloc(&G.d_bp_order, sizeof(unsigned) * ord.size());
    if (e == hipSuccess)
      e = hipMemcpy(G.d_bp_order, ord.data(), sizeof(unsigned) * ord.size(), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && A < (1 << 24)) {
    const std::vector<int4> grp = fpg_groups(G.lines, G.n_steep, N, R, rho0, drho, &G.fpg_cpl);
    if (!grp.empty() && grp.size() < 65536) {
      e = hipMalloc(&G.d_fpg, sizeof(int4) * grp.size());
      if (e == hipSuccess) e = hipMemcpy(G.d_fpg, grp.data(), sizeof(int4) * grp.size(), hipMemcpyHostToDevice);
      G.n_fpg = (int)grp.size();
      if (e == hipSuccess) e = hipMalloc(&G.d_fp_steps, sizeof(float2) * (size_t)A * FP_STEP_BINADES);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fp_step_table, dim3((A * FP_STEP_BINADES + 255) / 256), dim3(256), 0, nullptr,
                           G.d_lines, A, G.d_fp_steps);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    }
  }
  if (e != hipSuccess) {
    if (G.d_lines) (void)hipFree(G.d_lines);
    if (G.d_inter) (void)hipFree(G.d_inter);
    if (G.d_bp_order) (void)hipFree(G.d_bp_order);
    if (G.d_fpg) (void)hipFree(G.d_fpg);
    if (G.d_fp_steps) (void)hipFree(G.d_fp_steps);
    delete g;
    return hip_fail(c, e, "geometry upload");
  }
  *out = g;
  return SK_OK;
}

void sk_geom_destroy(sk_geom* g) {
  if (!g) return;
  (void)hipSetDevice(g->g.device);
  (void)hipDeviceSynchronize();
  for (auto& sg : g->g.sirt_graphs)
    if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
  if (g->g.d_lines) (void)hipFree(g->g.d_lines);
  if (g->g.d_inter) (void)hipFree(g->g.d_inter);
  if (g->g.d_bp_order) (void)hipFree(g->g.d_bp_order);
  if (g->g.d_fpg) (void)hipFree(g->g.d_fpg);
  if (g->g.d_fp_steps) (void)hipFree(g->g.d_fp_steps);
  if (g->g.d_rowinv) (void)hipFree(g->g.d_rowinv);
  if (g->g.d_colinv) (void)hipFree(g->g.d_colinv);
  delete g;
}

int sk_geom_query(const sk_geom* g, int* A, int* R, int* N, int* n_steep, int* ncand) {
  if (!g) return SK_EINVAL;
  if (A) *A = g->g.A;
  if (R) *R = g->g.R;
  if (N) *N = g->g.N;
  if (n_steep) *n_steep = g->g.n_steep;
  if (ncand) *ncand = g->g.ncand;
  return SK_OK;
}

}  // extern "C"
