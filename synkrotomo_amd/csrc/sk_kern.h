// sk_kern.h -- device code of the projector: one work item of A.x (a block of rays of one
// line) and one work item of B.r (a pixel tile), shared by the per-call kernels
// (sk_proj.hip) and the persistent multi-problem kernels (sk_batch.hip).
//
// Arithmetic follows the reference expression by expression; this code is compiled with
// -ffp-contract=off so every multiply and add stays a separate IEEE operation wherever a
// pixel index or candidate index depends on it.
#pragma once

#include <type_traits>
#include <utility>

#include "sk_common.h"

namespace sk {

constexpr int FP_BLOCK = 256;              // threads per fp work item
constexpr int FP_WAVES = FP_BLOCK / 64;
constexpr int BP_TW = 16;                  // pixel tile width; height = 16 * PPT
constexpr int BP_LB = 32;                  // lines per staged LDS block
constexpr unsigned BUF_DWORD3 = 0x00020000u;  // gfx9 raw buffer descriptor word 3

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// floor(x) and floor(-x) = -ceil(x) as an int in one instruction (v_cvt_flr_i32_f32).  The
// compiler only selects this instruction under fast-math flags, which would also license
// contractions elsewhere; the arguments here are finite and far inside the int range.
__device__ __forceinline__ int floor_to_int(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}
__device__ __forceinline__ int neg_ceil_to_int(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, -%1" : "=v"(r) : "v"(x));
  return r;
}

// A buffer descriptor covering one image row of n floats: loads at columns outside [0, n)
// return 0 (hardware range check), which is exactly the reference's contribution of an
// out-of-bounds pixel.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const float* row, int n) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)row, 0, n * 4, BUF_DWORD3);
}
__device__ __forceinline__ float row_load(__amdgpu_buffer_rsrc_t r, int col) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, col << 2, 0, 0));
}

// Conservative frame-row range [i0, i1) outside which both candidate pixels of a ray with
// this slope and intercept are out of bounds (|x| > h + 2 with a 2-row margin).  Both ends
// move monotonically, in the same direction, with the intercept, so the hull of the ranges
// of a wave's first and last rays covers every ray between them.  Near-zero slopes
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

// Length of one candidate ray in one pixel, in the reference's own formulation.
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
      p = m * fast_rcp(diff);  // diff > 0 on non-degenerate lines, where m == 0 gives 0
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

// ---------------------------------------------------------------------------- A.x item
struct FpParams {
  const Line* lines;
  int A, R, N, ncand;
  float rho0, drho;
  const float* img;     // [batch][N][N]
  const float* imgT;    // transposed copy (flat lines)
  float* out;           // [batch][A][R]
  const float* b;       // FP_SIRT: measured sinogram
  const float* rowinv;  // FP_SIRT: [A][R]
  const int4* groups;   // grouped staged variant: the geometry's line groups (else null)
  const float2* steps;  // grouped staged variant: per line, (D_e, 1/D_e) for binades e (else null)
};

// Binades e = 0 .. FP_STEP_BINADES-1 of the forward projector's exact per-block step table.
constexpr int FP_STEP_BINADES = 16;

// Lane k of every 16-lane row of v, broadcast to the whole wave when v holds the same 16
// values in each row (DPP row_newbcast; the compiler folds it into the consuming add).
template <int K>
__device__ __forceinline__ float bcast16(float v) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + K, 0xf, 0xf, true));
}

// Fast-mode contribution of one row (forwardprojection.fut:37-56) given the row's xbot,
// xtop and the two pixels the row can touch: vb = v(Xb), Xb = floor(xbot), and vo = its
// neighbour on the slope's side, v(Xb + 1) for slope >= 0 and v(Xb - 1) otherwise.  The
// row adds vb + fo*(vo - vb), fo = |xtop - Xmax| / |xtop - xbot| the fraction of its length
// in the neighbour (the reference's Xmax = floor(xbot) + 1, resp. floor(xbot), when the
// row crosses into it).  When the row stays in pixel Xb (the reference's full-length case)
// the numerator has the wrong sign and the clamp to [0, 1] makes fo exactly 0, without a
// select; a zero-length range (x/0 = inf, or 0*inf) also clamps to 0.  The sums are
// fractions of the row length; the caller scales by lbase.  NT (LINE_NEWTON lines): the
// reciprocal of |xtop - xbot| by one Newton step from rs = 1/|slope| (two FMAs) instead of
// the quarter-rate v_rcp_f32 (error <= 1.6e-7 relative, see newton_min_slope).
template <bool POS, bool NT = false>
__device__ __forceinline__ float fp_row_add(float sum, float xb, float xt, float vb, float vo, float rs = 0.0f) {
  const float fx = floorf(xb);
  const float d = POS ? xt - xb : xb - xt;
  const float r = NT ? fmaf(rs, fmaf(-d, rs, 1.0f), rs) : fast_rcp(d);
  float fo;
  if constexpr (POS) {  // (xtop - (fx + 1)) / (xtop - xbot) with one rounding: xtop - fx is exact
    fo = fmaf(xt - fx, r, -r);
  } else {              // (fx - xtop) / (xbot - xtop)
    fo = (fx - xt) * r;
  }
  fo = fminf(fmaxf(fo, 0.0f), 1.0f);
  return fmaf(fo, vo - vb, sum + vb);
}

// Fast-mode row walk for RPT rays of one line, the slope sign a template argument.  Both
// pixels of a row come from one column offset: the buffer loads at c and c + 1 (c = Xb
// for slope >= 0, Xb - 1 otherwise), each range-checked by the row's descriptor.
//
// Rows go in blocks of 16.  The wave-uniform products i*slope of a block's rows are formed
// once, one per lane of a 16-lane row, and each row's xbot = i*slope + icpt broadcasts its
// lane into the add (one instruction per row instead of convert, multiply and add).  The
// walk is software-pipelined three rows deep (four register sets): rows i+1..i+3 are in
// flight while row i's fraction is formed.  Rows past i1 (the last block's tail and the
// prefetches) read through an empty descriptor: no memory access, both values 0, and such
// a row adds exactly 0.
template <bool POS, int RPT, bool NT = false>
__device__ __forceinline__ void fp_rows_fast(const float* src, int N, int h, int i0, int i1,
                                             float slope, const float* icpt, float* sum, float rs = 0.0f) {
  if (i0 >= i1) return;
  constexpr int D = 3;  // rows in flight
  const int l16 = (int)(threadIdx.x & 15);
  const int cofs = h - (POS ? 0 : 1);
  float xb[D + 1][RPT], xt[D + 1][RPT], va[D + 1][RPT], vc[D + 1][RPT];
  auto gather = [&](int i, float yslope, int q) {
    const __amdgpu_buffer_rsrc_t row = row_rsrc(src + (size_t)(i + h) * N, i < i1 ? N : 0);
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      xb[q][k] = yslope + icpt[k];
      xt[q][k] = xb[q][k] + slope;
      const int c = (floor_to_int(xb[q][k]) + cofs) << 2;
      // Two separate dword loads, each range-checked on its own offset (c may be -4): the
      // empty asm hides the adjacency so they are not merged into one 8-byte access.
      int c1 = c + 4;
      asm volatile("" : "+v"(c1));
      va[q][k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(row, c, 0, 0));
      vc[q][k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(row, c1, 0, 0));
    }
  };
  auto consume = [&](int q) {
#pragma unroll
    for (int k = 0; k < RPT; ++k)
      sum[k] = POS ? fp_row_add<true, NT>(sum[k], xb[q][k], xt[q][k], va[q][k], vc[q][k], rs)
                   : fp_row_add<false, NT>(sum[k], xb[q][k], xt[q][k], vc[q][k], va[q][k], rs);
  };
  auto stage = [&](auto kc, int i, float tab) {
    constexpr int K = decltype(kc)::value;
    gather(i + D + K, bcast16<K>(tab), (K + D) % (D + 1));
    __builtin_amdgcn_sched_barrier(0);
    consume(K % (D + 1));
    __builtin_amdgcn_sched_barrier(0);
  };
#pragma unroll
  for (int d = 0; d < D; ++d) {
    gather(i0 + d, (float)(i0 + d) * slope, d);
    __builtin_amdgcn_sched_barrier(0);  // issue in row order: the loop waits on the oldest
  }
  for (int i = i0; i < i1; i += 16) {
    // sched_barrier keeps the compiler from regrouping the stages (it would otherwise
    // cluster the loads and wait on all of them at the top of the loop).
    const float tab = (float)(i + D + l16) * slope;  // rows gathered in this block
    stage(std::integral_constant<int, 0>{}, i, tab);
    stage(std::integral_constant<int, 1>{}, i, tab);
    stage(std::integral_constant<int, 2>{}, i, tab);
    stage(std::integral_constant<int, 3>{}, i, tab);
    stage(std::integral_constant<int, 4>{}, i, tab);
    stage(std::integral_constant<int, 5>{}, i, tab);
    stage(std::integral_constant<int, 6>{}, i, tab);
    stage(std::integral_constant<int, 7>{}, i, tab);
    stage(std::integral_constant<int, 8>{}, i, tab);
    stage(std::integral_constant<int, 9>{}, i, tab);
    stage(std::integral_constant<int, 10>{}, i, tab);
    stage(std::integral_constant<int, 11>{}, i, tab);
    stage(std::integral_constant<int, 12>{}, i, tab);
    stage(std::integral_constant<int, 13>{}, i, tab);
    stage(std::integral_constant<int, 14>{}, i, tab);
    stage(std::integral_constant<int, 15>{}, i, tab);
  }
}

// One wave's work item: line j, bins [wb*64*RPT, (wb+1)*64*RPT) of image bz (lane l: bins
// first + l + 64k), all lanes stepping the same frame row.  WEIGHTS == W_ADJ walks the same
// rays with the backprojector's lengths (B^T).  No barriers: waves are independent.
template <bool EXACT, int MODE, int WEIGHTS, bool DR1, int RPT>
__device__ __forceinline__ void fp_item(const FpParams& P, int j, int wb, int bz) {
  const int lane = threadIdx.x & 63;
  const int first = wb * 64 * RPT;
  if (first >= P.R) return;  // whole wave past the last bin
  const int last = min(first + 64 * RPT - 1, P.R - 1);
  const Line L = P.lines[j];
  const int N = P.N, h = N >> 1;
  const float cost = L.cost, slope = L.slope, lbase = L.lbase;

  int r[RPT];
  float icpt[RPT], sum[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    r[k] = first + 64 * k + lane;
    icpt[k] = (P.rho0 + (float)r[k] * P.drho) / cost;  // forwardprojection.fut:33-34
    sum[k] = 0.0f;
  }
  int i0, i1;
  {
    int a0, a1, b0, b1;
    row_clip(slope, (P.rho0 + (float)first * P.drho) / cost, h, a0, a1);
    row_clip(slope, (P.rho0 + (float)last * P.drho) / cost, h, b0, b1);
    i0 = __builtin_amdgcn_readfirstlane(min(a0, b0));
    i1 = __builtin_amdgcn_readfirstlane(max(a1, b1));
  }
  const float* src = ((L.flags & LINE_FLAT) ? P.imgT : P.img) + (size_t)bz * N * N;

  constexpr bool FASTW = !(WEIGHTS == W_ADJ || EXACT || MODE == FP_ONES_INV);
  if constexpr (FASTW) {
    const bool nt = L.flags & LINE_NEWTON;
    const float rs = nt ? 1.0f / fabsf(slope) : 0.0f;
    if ((L.flags & LINE_SLOPE_POS) && nt)
      fp_rows_fast<true, RPT, true>(src, N, h, i0, i1, slope, icpt, sum, rs);
    else if (L.flags & LINE_SLOPE_POS)
      fp_rows_fast<true, RPT>(src, N, h, i0, i1, slope, icpt, sum);
    else if (nt)
      fp_rows_fast<false, RPT, true>(src, N, h, i0, i1, slope, icpt, sum, rs);
    else
      fp_rows_fast<false, RPT>(src, N, h, i0, i1, slope, icpt, sum);
  } else {
    for (int i = i0; i < i1; ++i) {
      const __amdgpu_buffer_rsrc_t row = row_rsrc(src + (size_t)(i + h) * N, N);
      const float yslope = (float)i * slope;
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        // forwardprojection.fut:37-56
        const float xbot = yslope + icpt[k];
        const float xtop = xbot + slope;
        const float fXb = floorf(xbot);
        const float fXt = floorf(xtop);
        const int Xb = (int)fXb;
        const int Xt = (int)fXt;
        const bool single = Xb == Xt;
        float vb = 1.0f, vt = 1.0f;
        if constexpr (MODE != FP_ONES_INV) {
          vb = row_load(row, Xb + h);
          vt = row_load(row, Xt + h);
        }
        {
          float lb, lt;
          if constexpr (WEIGHTS == W_ADJ) {
            lb = bp_entry<EXACT, DR1>(L, Xb, (float)i, xbot, xtop, r[k], P.ncand, P.rho0, P.drho);
            lt = bp_entry<EXACT, DR1>(L, Xt, (float)i, xbot, xtop, r[k], P.ncand, P.rho0, P.drho);
          } else {
            const float Xmax = fmaxf(fXb, fXt);
            const float xdiff = xtop - xbot;
            if constexpr (EXACT) {
              lb = single ? lbase : ((Xmax - xbot) / xdiff) * lbase;
              lt = ((xtop - Xmax) / xdiff) * lbase;
            } else {
              const float rd = fast_rcp(xdiff);
              lb = single ? lbase : ((Xmax - xbot) * rd) * lbase;
              lt = ((xtop - Xmax) * rd) * lbase;
            }
          }
          const bool inb = (unsigned)(Xb + h) < (unsigned)N;
          const bool intop = !single && (unsigned)(Xt + h) < (unsigned)N;
          const float bot = inb ? lb * vb : 0.0f;
          const float top = intop ? lt * vt : 0.0f;
          sum[k] = sum[k] + (bot + top);
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    if (r[k] >= P.R) continue;
    float s = sum[k];
    if constexpr (!(WEIGHTS == W_ADJ || EXACT || MODE == FP_ONES_INV)) s = s * lbase;
    const size_t o = (size_t)bz * P.A * P.R + (size_t)L.orig * P.R + r[k];
    if constexpr (MODE == FP_PLAIN) {
      P.out[o] = s;
    } else if constexpr (MODE == FP_ONES_INV) {
      P.out[o] = (s == 0.0f) ? 0.0f : 1.0f / s;  // sirtlib.fut:9-10
    } else {
      // SIRT.fut:33-36: R * (b - A x)
      P.out[o] = P.rowinv[(size_t)L.orig * P.R + r[k]] * (P.b[o] - s);
    }
  }
}

// ---------------------------------------------------------------------------- B.r item
struct BpParams {
  const Line* lines;
  int A, R, N, ncand, W;
  int nsteep;           // lines [0, nsteep) are steep, the rest flat
  float rho0, drho;
  const float* sino;    // [batch][A][R]
  float* out;           // [batch][N][N]  (BP_SIRT: the image, updated in place)
  float* outT;          // BP_SIRT: transposed copy of the updated image
  const float* colinv;  // BP_SIRT: [N][N]
  const float* inter;   // [A][R] (s*drho + rho0)/cost per table line, or null
  // launch order of the tiles, (tile_y << 16) | tile_x per block in grid order, or null (grid
  // order): the tiles whose lines take the exact step (tile_binade) are the cheapest, so the
  // dearer ones go first and the last blocks to start are cheap ones
  const unsigned* order;
};

// LDS window width (bins) covering every candidate of a BP_TW x (16*PPT) tile.
__host__ __device__ inline int bp_window_bins(int ppt, float drho, int ncand) {
  const double w = BP_TW - 1, hgt = 16.0 * ppt - 1;
  const double span = sqrt(w * w + hgt * hgt) * 1.00001 / (double)drho;
  const int W = (int)ceil(span) + ncand + 3;
  return (W + 1) & ~1;
}

// Per-line values of one line block, in LDS.  w0: the tile's rho-window origin.  Fast mode
// only: ci, si, cst give the candidate index t = ci*(xmin+0.5) + si*(ymin+0.5) + cst (the
// reference's ((cost*x + sint*y) - rho0)/drho, x = xmin+0.5-fact*cost, y = ymin+0.5-fact*sint,
// with the constant terms folded).  t does not depend on the tile, so every tile shape
// selects the same candidates.
struct alignas(16) BpLineLds {
  int w0;
  float ci, si, cst;
  float dx;      // hoisted fast path: the exact |x_top - x_bot| of every candidate of the tile
                 // (tile_binade), else 0
  float slope;   // the line's (read here with the rest of the record: no scalar load per line)
  float rs;      // 1/|slope| on LINE_NEWTON lines (fast path), else 0
  float cost;    // the line's (window staging)
  float wscale;  // factor of the staged values: lbase, or lbase/dx when dx > 0
  int orig;
  int tie;       // hoisted fast path, tile with one binade: the step is a rounding tie (dx = 0)
};

// Binade [2^e, 2^(e+1)) holding the magnitudes of every x_bot and x_top of a tile's candidates
// that can have a nonzero length, and of the pixel edges moved by the step (xmin - d,
// xmin + 1 + d), or -1.  x0..x1: the tile's frame columns (steep lines: image columns; flat
// lines: image rows).  A candidate with a nonzero length has x_bot, x_top in (xmin - 1,
// xmin + 2) (backprojection.fut:59-62 with |slope| <= 1); the range [x0 - 4, x1 + 5] also
// covers every candidate within 3 of its pixel, so one farther away (whose length is 0 in the
// reference) stays at 0 here.
__host__ __device__ inline int tile_binade(int x0, int x1) {
  const int lo = x0 - 4, hi = x1 + 5;
  if (lo <= 0 && hi >= 0) return -1;
  const int a = lo > 0 ? lo : -hi, b = lo > 0 ? hi : -lo;  // 1 <= a <= b
  const int ea = 31 - __builtin_clz((unsigned)a), eb = 31 - __builtin_clz((unsigned)b);
  return ea == eb ? ea : -1;
}

// Flat lines' exact-step rule for the 64-row band whose first frame row is y0: e = the binade
// of its first 16-row sub-band that lies in one (tile_binade), mask bit s = sub-band s lies in
// binade e.  Only the masks a band starting at a multiple of 64 can have -- 1111, and 0111,
// 1110, 0110 where a power of two lies at an edge of the band -- are kept; any other gives
// e = -1, mask 0.  Every tile shape derives its pixels' choice from the same band, so the
// variants stay bitwise equal.
__host__ __device__ inline void bp_band_rule(int y0, int& e, int& mask) {
  e = -1;
  mask = 0;
  for (int s = 0; s < 4; ++s) {
    const int es = tile_binade(y0 + 16 * s, y0 + 16 * s + 15);
    if (es < 0) continue;
    if (e < 0) e = es;
    if (es == e) mask |= 1 << s;
  }
  if (mask != 15 && mask != 7 && mask != 14 && mask != 6) {
    e = -1;
    mask = 0;
  }
}

// The exact step d = |fl(x_bot + slope) - x_bot| of every x_bot in binade e whose x_top stays
// in it: x_bot is a multiple of u = 2^(e-23), so fl(x_bot + slope) = x_bot + RN_u(slope), and d
// = RN_u(|slope|) whatever x_bot is -- unless |slope| lies exactly halfway between two multiples
// of u, where round-half-even makes d depend on x_bot's last bit (returns 0 then, and for
// d = 0).  |slope|/u and its rounding are exact (|slope| <= 1, u a power of two, e >= 0).
__device__ __forceinline__ float binade_step(float slope, int e) {
  const float t = ldexpf(fabsf(slope), 23 - e);
  const float ft = floorf(t);
  if (t - ft == 0.5f) return 0.0f;
  return ldexpf(rintf(t), e - 23);
}

// One work item: pixel tile (tile_x, tile_y) of image bz, 16 columns x 16*PPT rows; thread
// (ty, tx) owns pixels (Y0 + ty + 16k, X0 + tx).  Lines are walked in order in blocks of
// BP_LB: the block finds each line's rho-window origin from the tile's corners (the
// candidate index is monotone in each frame coordinate), stages the window -- (s*drho +
// rho0)/cost and the sinogram value -- in LDS, then every lane evaluates its candidates
// from LDS.  Steep and flat lines accumulate separately and are added at the end
// (backprojection.fut:75-78).  win: BP_LB*W float2, lsh: BP_LB entries, both in LDS.
//
// Exact mode evaluates every expression of backprojection.fut:49-62 in the reference's
// order.  Fast mode keeps the reference's xbot/xtop (so each length fraction is the
// reference's up to the reciprocal) but computes the candidate index with fewer roundings:
// a pixel whose rounded index moves by one bin gains or loses only a bin at distance
// >= sqrt(2)/2 from its centre, where the intersection length is zero.
template <bool EXACT, int MODE, int WEIGHTS, int NC, bool DR1, int PPT>
__device__ __forceinline__ void bp_item(const BpParams& P, int tile_x, int tile_y, int bz,
                                        float2* win, BpLineLds* lsh) {
  // Fast reference weights: slope sign hoisted per line, lbase folded into the staged value.
  constexpr bool FASTREF = !EXACT && WEIGHTS == W_REF;
  // The hoisted fast path keeps each line's window in descending bin order, so that the
  // candidate address comes from floor(-t) (one conversion) instead of ceil(t).
  constexpr bool REV = FASTREF && NC > 0;
  const int ncand = NC > 0 ? NC : P.ncand;
  const int W = P.W;
  const int tx = threadIdx.x & (BP_TW - 1), ty = threadIdx.x / BP_TW;
  constexpr int TH = 16 * PPT;
  const int X0 = tile_x * BP_TW, Y0 = tile_y * TH;
  const int N = P.N, h = N >> 1;
  const float* sb = P.sino + (size_t)bz * P.A * P.R;
  const int X = X0 + tx;
  float xs[PPT], ys[PPT], acc_s[PPT], acc_f[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    xs[k] = (float)(X - h);
    ys[k] = (float)(Y0 + ty + 16 * k - h);
    acc_s[k] = 0.0f;
    acc_f[k] = 0.0f;
  }

  // hoisted fast path: the tile's binade for steep lines (frame x = image column; -1 where
  // the tile's columns straddle a power of two or zero) and for flat lines (frame x = image
  // row): the binade of the 64-row band holding the tile (bp_band_rule), whose 16-row
  // sub-bands in it take the exact step -- fslots bit q: the sub-band of slot q (rows
  // Y0 + 16q ..).  The other slots of such a tile take the per-candidate step scaled by the
  // step (their staged values carry lbase/dx as well).
  const int Y64 = Y0 & ~63;
  int eflat = -1, fband = 0;
  if (REV) bp_band_rule(Y64 - h, eflat, fband);
  const int fslots = (fband >> ((Y0 - Y64) >> 4)) & ((1 << PPT) - 1);
  const int ebin[2] = {REV ? tile_binade(X0 - h, X0 - h + BP_TW - 1) : -1, eflat};

  for (int l0 = 0; l0 < P.A; l0 += BP_LB) {
    const int nl = min(BP_LB, P.A - l0);
    if ((int)threadIdx.x < nl) {
      const Line L = P.lines[l0 + threadIdx.x];
      const bool fl = L.flags & LINE_FLAT;
      const float cx0 = (float)((fl ? Y0 : X0) - h), cy0 = (float)((fl ? X0 : Y0) - h);
      const float cx1 = cx0 + (float)((fl ? TH : BP_TW) - 1);
      const float cy1 = cy0 + (float)((fl ? BP_TW : TH) - 1);
      float smin = cand_s<DR1>(L, cx0, cy0, P.rho0, P.drho);
      smin = fminf(smin, cand_s<DR1>(L, cx1, cy0, P.rho0, P.drho));
      smin = fminf(smin, cand_s<DR1>(L, cx0, cy1, P.rho0, P.drho));
      smin = fminf(smin, cand_s<DR1>(L, cx1, cy1, P.rho0, P.drho));
      BpLineLds e;
      // The hoisted fast path starts the window one bin lower: its index may round one bin
      // below the reference's at the tile's lowest corner (the top keeps a one-bin margin).
      e.w0 = (int)smin - (REV ? 1 : 0);
      e.ci = e.si = e.cst = 0.0f;
      if constexpr (FASTREF) {
        const float idr = DR1 ? 1.0f : 1.0f / P.drho;
        e.ci = L.cost * idr;
        e.si = L.sint * idr;
        e.cst = -(fmaf(L.cost, L.fc, L.sint * L.fs) + P.rho0) * idr;
      }
      e.cost = L.cost;
      e.slope = L.slope;
      e.orig = L.orig;
      e.rs = (L.flags & LINE_NEWTON) ? 1.0f / fabsf(L.slope) : 0.0f;
      const int eb = ebin[fl ? 1 : 0];
      const bool exd = REV && eb >= 0 && !(L.flags & LINE_DEGEN);
      e.dx = exd ? binade_step(L.slope, eb) : 0.0f;
      e.tie = exd && e.dx == 0.0f;
      e.wscale = e.dx > 0.0f ? L.lbase / e.dx : L.lbase;
      lsh[threadIdx.x] = e;
    }
    __syncthreads();
    // Window staging, flattened over (line, bin): the 256 threads take entries e = jj*W + k of
    // the block's nl windows in turn (a wave per line would leave most lanes of its second
    // pass over a 70-bin window idle).  e / W by a multiply-high: exact for e < 2^32 / W.
    {
      const int total = nl * W;
      const unsigned Mw = 0xffffffffu / (unsigned)W + 1u;
      for (int e = threadIdx.x; e < total; e += 256) {
        const int jj = (int)__umulhi((unsigned)e, Mw);
        const int k = e - jj * W;
        const int sp = lsh[jj].w0 + k;
        const bool in = sp >= 0 && sp < P.R;
        // backprojection.fut:57-58: r = s'*drho + rho0, then r/cost (from the geometry's table
        // for bins of the sinogram; bins outside it are rare: divide)
        float inter;
        if (in && P.inter)
          inter = P.inter[(l0 + jj) * P.R + sp];
        else
          inter = ((float)sp * P.drho + P.rho0) / lsh[jj].cost;
        float p;
        if constexpr (MODE == BP_ONES_INV) {
          p = in ? 1.0f : 0.0f;
        } else {
          const float v = sb[(size_t)lsh[jj].orig * P.R + (in ? sp : 0)];  // in bounds either way
          p = in ? v : 0.0f;
        }
        if constexpr (FASTREF) p = p * lsh[jj].wscale;
        win[jj * W + (REV ? W - 1 - k : k)] = make_float2(inter, p);
      }
    }
    __syncthreads();
    // Steep lines precede flat lines in the table, so the block's lines split into (at
    // most) one steep and one flat run; each run uses its own frame coordinates.
    const int nsteep = min(max(P.nsteep - l0, 0), nl);
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const int j_begin = part == 0 ? 0 : nsteep, j_end = part == 0 ? nsteep : nl;
      if (j_begin >= j_end) continue;
      float xm[PPT], ym[PPT], xh[PPT], yh[PPT], x1[PPT], acc[PPT];
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        xm[q] = part ? ys[q] : xs[q];
        ym[q] = part ? xs[q] : ys[q];
        xh[q] = xm[q] + 0.5f;
        yh[q] = ym[q] + 0.5f;
        x1[q] = xm[q] + 1.0f;
        acc[q] = part ? acc_f[q] : acc_s[q];
      }
      if constexpr (REV) {
        // The table keeps the lines of each part in angle order, so the lines of one slope
        // sign (degenerate or not) form runs; Line::flags >> LINE_RUN_SHIFT is the number of
        // lines from a line to the end of its run.  Each run is walked by a loop whose body
        // has the sign and the degenerate case resolved at compile time (no per-line branch).
        // Kinds: 0: slope >= 0, 1: slope < 0, 2: degenerate; 3, 4: as 0, 1 on LINE_NEWTON
        // lines; 5, 6: as 0, 1 in a tile with one binade (tile_binade), with the exact step.
        // SLOTS (kinds 0, 1, 3, 4 of a flat run whose band has a binade, bp_band_rule): bit q
        // set: slot q takes the exact step as kinds 5, 6 do; the other slots take the kind's
        // own step, times dx (the staged value is lbase/dx).  -1: the kind's step, unscaled.
        auto line_body = [&](auto kind, auto slots, int jj) {
          constexpr int KIND = decltype(kind)::value;
          constexpr int SLOTS = decltype(slots)::value;
          constexpr bool POS = KIND == 0 || KIND == 3 || KIND == 5;
          const BpLineLds lc = lsh[jj];
          const float slope = lc.slope;
          // steep run: frame x (the image column) is shared by a lane's pixels; flat run: y.
          const float tb = part == 0 ? fmaf(lc.ci, xh[0], lc.cst) : fmaf(lc.si, yh[0], lc.cst);
          // All of the lane's pixels first compute their candidate index and read their
          // candidates: the LDS reads of all pixels are in flight together.  Bin s sits at
          // window position W-1-(s-w0); wrev[-s] is the entry of the last candidate of the
          // pixel whose first candidate is s (w0 is uniform: scalar base arithmetic).
          const float2* wrev = win + __builtin_amdgcn_readfirstlane(jj * W + W - NC + lc.w0);
          float2 ev[PPT][NC];
          float xbase[PPT], edge[PPT];
          bool ex[PPT];  // the slot takes the exact step (constant once the loops unroll)
#pragma unroll
          for (int q = 0; q < PPT; ++q) {
            ex[q] = KIND == 5 || KIND == 6 || (KIND != 2 && SLOTS >= 0 && ((SLOTS >> q) & 1));
            // t is within 0.4 bin of the reference's index (|rho0| + N < 2^20 for this
            // path), so its ceiling lies in [w0, w0 + W - NC]: no clamp.
            const float t = part == 0 ? fmaf(lc.si, yh[q], tb) : fmaf(lc.ci, xh[q], tb);
            const float2* wrow = wrev + neg_ceil_to_int(t);
            xbase[q] = ym[q] * slope;  // backprojection.fut:55
            // x_top - xmin = x_bot - (xmin - d), resp. xmin + 1 - x_top = (xmin + 1 + d) -
            // x_bot, exactly: x_bot, x_top and the moved edge are multiples of u in the
            // tile's binade
            if (ex[q]) edge[q] = POS ? xm[q] - lc.dx : x1[q] + lc.dx;
#pragma unroll
            for (int k = 0; k < NC; ++k) ev[q][k] = wrow[NC - 1 - k];
          }
          // Non-degenerate lines: max(min(a, b), 0) == med3(a, b, 0): a and b cannot both
          // be negative when the ray's x-range has positive length; m/diff is then >= 0,
          // so min(., 1) is a clamp to [0, 1].
#pragma unroll
          for (int q = 0; q < PPT; ++q)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
              const float xb = xbase[q] + ev[q][k].x;
              if constexpr (KIND == 2) {
                // degenerate line: general form with the m == 0 case, lbase already in e.y
                const float l = cand_len<false, W_REF>(xb, slope, 1.0f, xm[q], x1[q], true);
                acc[q] = fmaf(l, ev[q][k].y, acc[q]);
              } else if (ex[q]) {
                // min(m/d, 1) = min(m, d)/d with d = dx exact and 1/d in the staged value:
                // max(min(a, b, d), 0) is one v_min3 with the output clamp (d <= 1).  A line
                // whose step is a rounding tie has dx = 0 and adds 0 here (tie pass below).
                const float a = POS ? xb - edge[q] : xb - xm[q];
                const float b = POS ? x1[q] - xb : edge[q] - xb;
                const float g = __builtin_amdgcn_fmed3f(fminf(fminf(a, b), lc.dx), 0.0f, 1.0f);
                acc[q] = fmaf(g, ev[q][k].y, acc[q]);
              } else {
                const float xt = xb + slope;
                const float d = POS ? xt - xb : xb - xt;
                const float m = POS ? __builtin_amdgcn_fmed3f(xt - xm[q], x1[q] - xb, 0.0f)
                                    : __builtin_amdgcn_fmed3f(xb - xm[q], x1[q] - xt, 0.0f);
                float f;
                if constexpr (KIND == 3 || KIND == 4) {
                  // 1/d by one Newton step from 1/|slope| (d = |x_top - x_bot| is |slope| to
                  // within the rounding of x_bot): two full-rate FMAs for the quarter-rate
                  // v_rcp_f32
                  const float r = fmaf(lc.rs, fmaf(-d, lc.rs, 1.0f), lc.rs);
                  f = fminf(fmaxf(m * r, 0.0f), 1.0f);
                } else {
                  f = fminf(fmaxf(m * fast_rcp(d), 0.0f), 1.0f);
                }
                // a tie line (dx = 0) adds 0 here as well
                if constexpr (SLOTS >= 0) f = f * lc.dx;
                acc[q] = fmaf(f, ev[q][k].y, acc[q]);
              }
            }
        };
        auto run = [&](auto kind, auto slots, int jb, int je) {
          for (int jj = jb; jj < je; ++jj) line_body(kind, slots, jj);
        };
        using none = std::integral_constant<int, -1>;
        // a flat run in a tile whose band has a binade but not every slot in it: the slot
        // masks each tile height can meet (bp_band_rule's bands)
        auto run_mixed = [&](auto kind, int jb, int je) {
          if constexpr (PPT == 4) {
            if (fslots == 7)
              run(kind, std::integral_constant<int, 7>{}, jb, je);
            else if (fslots == 14)
              run(kind, std::integral_constant<int, 14>{}, jb, je);
            else
              run(kind, std::integral_constant<int, 6>{}, jb, je);
          } else if constexpr (PPT == 2) {
            if (fslots == 1)
              run(kind, std::integral_constant<int, 1>{}, jb, je);
            else
              run(kind, std::integral_constant<int, 2>{}, jb, je);
          } else {
            run(kind, std::integral_constant<int, 0>{}, jb, je);
          }
        };
        const bool exact_step = ebin[part] >= 0;  // uniform: the tile's binade for this part
        const bool all_slots = part == 0 || fslots == (1 << PPT) - 1;  // uniform
        for (int jr = j_begin; jr < j_end;) {  // uniform: scalar loop over the runs
          const int f = P.lines[l0 + jr].flags;
          const int je = min(jr + (f >> LINE_RUN_SHIFT), j_end);
          if (f & LINE_DEGEN)
            run(std::integral_constant<int, 2>{}, none{}, jr, je);
          else if (exact_step && all_slots && (f & LINE_SLOPE_POS))
            run(std::integral_constant<int, 5>{}, none{}, jr, je);
          else if (exact_step && all_slots)
            run(std::integral_constant<int, 6>{}, none{}, jr, je);
          else if (exact_step && (f & LINE_NEWTON) && (f & LINE_SLOPE_POS))
            run_mixed(std::integral_constant<int, 3>{}, jr, je);
          else if (exact_step && (f & LINE_NEWTON))
            run_mixed(std::integral_constant<int, 4>{}, jr, je);
          else if (exact_step && (f & LINE_SLOPE_POS))
            run_mixed(std::integral_constant<int, 0>{}, jr, je);
          else if (exact_step)
            run_mixed(std::integral_constant<int, 1>{}, jr, je);
          else if ((f & LINE_NEWTON) && (f & LINE_SLOPE_POS))
            run(std::integral_constant<int, 3>{}, none{}, jr, je);
          else if (f & LINE_NEWTON)
            run(std::integral_constant<int, 4>{}, none{}, jr, je);
          else if (f & LINE_SLOPE_POS)
            run(std::integral_constant<int, 0>{}, none{}, jr, je);
          else
            run(std::integral_constant<int, 1>{}, none{}, jr, je);
          jr = je;
        }
        if (exact_step) {
          // tie pass: the part's lines whose step depends on x_bot's last bit, with the
          // hardware reciprocal (their staged values carry lbase only)
          const int lane = threadIdx.x & 63;
          const bool tie = lane >= j_begin && lane < j_end && lsh[lane < BP_LB ? lane : 0].tie;
          unsigned long long tm = __ballot(tie);
          while (tm) {  // uniform
            const int jj = __builtin_ctzll(tm);
            tm &= tm - 1;
            if (P.lines[l0 + jj].flags & LINE_SLOPE_POS)
              line_body(std::integral_constant<int, 0>{}, none{}, jj);
            else
              line_body(std::integral_constant<int, 1>{}, none{}, jj);
          }
        }
      }
      if constexpr (!REV) for (int jj = j_begin; jj < j_end; ++jj) {
        const Line L = P.lines[l0 + jj];
        const BpLineLds lc = lsh[jj];
        const bool degen = L.flags & LINE_DEGEN;
        const bool pos = L.flags & LINE_SLOPE_POS;
        // steep run: frame x (the image column) is shared by a lane's pixels; flat run: y.
        const float tb = part == 0 ? fmaf(lc.ci, xh[0], lc.cst) : fmaf(lc.si, yh[0], lc.cst);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          const float xmin = xm[q], ymin = ym[q], xmin1 = x1[q];
          const float xbase = ymin * L.slope;  // backprojection.fut:55
          const float2* wrow;
          if constexpr (FASTREF) {
            // general path (any geometry): the fast index, clamped into the window
            const float t = part == 0 ? fmaf(lc.si, yh[q], tb) : fmaf(lc.ci, xh[q], tb);
            wrow = win + jj * W + min(max((int)ceilf(t) - lc.w0, 0), W - ncand);
          } else {
            // backprojection.fut:51-54 with (xmin + 0.5), (ymin + 0.5) hoisted
            const float px = xh[q] - L.fc;
            const float py = yh[q] - L.fs;
            const float rho = L.cost * px + L.sint * py;
            float t = rho - P.rho0;
            if constexpr (!DR1) t = t / P.drho;
            // s - w0 lies in [0, W - ncand] (monotone s, window sized on the host); the
            // clamp only guards geometries so far off-centre that rho rounds by whole bins.
            wrow = win + jj * W + min(max((int)ceilf(t) - lc.w0, 0), W - ncand);
          }
          constexpr int KC = NC > 0 ? NC : 16;
          float2 ev[KC];  // fast mode: all candidates read before the per-line branches
          if constexpr (FASTREF) {
#pragma unroll
            for (int k = 0; k < KC; ++k)
              if (NC > 0 || k < ncand) ev[k] = wrow[k];
          }
          if (FASTREF && !degen) {
            // Non-degenerate lines: max(min(a, b), 0) == med3(a, b, 0) because a and b
            // cannot both be negative when the ray's x-range has positive length; the
            // fraction m/diff is then >= 0, so min(., 1) is a clamp to [0, 1].
            if (pos) {
#pragma unroll
              for (int k = 0; k < KC; ++k) {
                if (NC == 0 && k >= ncand) break;
                const float xb = xbase + ev[k].x;
                const float xt = xb + L.slope;
                const float m = __builtin_amdgcn_fmed3f(xt - xmin, xmin1 - xb, 0.0f);
                const float f = fminf(fmaxf(m * fast_rcp(xt - xb), 0.0f), 1.0f);
                acc[q] = fmaf(f, ev[k].y, acc[q]);
              }
            } else {
#pragma unroll
              for (int k = 0; k < KC; ++k) {
                if (NC == 0 && k >= ncand) break;
                const float xb = xbase + ev[k].x;
                const float xt = xb + L.slope;
                const float m = __builtin_amdgcn_fmed3f(xb - xmin, xmin1 - xt, 0.0f);
                const float f = fminf(fmaxf(m * fast_rcp(xb - xt), 0.0f), 1.0f);
                acc[q] = fmaf(f, ev[k].y, acc[q]);
              }
            }
          } else if (FASTREF) {
            // degenerate line: general form with the m == 0 case, lbase already in e.y
#pragma unroll
            for (int k = 0; k < KC; ++k) {
              if (NC == 0 && k >= ncand) break;
              const float l =
                  cand_len<false, W_REF>(xbase + ev[k].x, L.slope, 1.0f, xmin, xmin1, true);
              acc[q] = fmaf(l, ev[k].y, acc[q]);
            }
          } else {
            // reference order: the line's candidates summed, then added to the pixel
            float inner = 0.0f;
#pragma unroll
            for (int k = 0; k < (NC > 0 ? NC : 16); ++k) {
              if (NC == 0 && k >= ncand) break;
              const float2 e = wrow[k];
              const float l =
                  cand_len<EXACT, WEIGHTS>(xbase + e.x, L.slope, L.lbase, xmin, xmin1, degen);
              inner = inner + l * e.y;
            }
            acc[q] = acc[q] + inner;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        if (part)
          acc_f[q] = acc[q];
        else
          acc_s[q] = acc[q];
      }
    }
    __syncthreads();
  }
  if constexpr (MODE == BP_SIRT) {
    // SIRT.fut:37-39: x + C * B(q), written to x and to its transpose.  The transposed copy
    // goes out through LDS (the window buffer, free after the line loop's last barrier) when
    // the tile fits: each column of the tile becomes a contiguous run of x^T instead of a
    // scatter with stride N.
    constexpr int TS = 17;  // LDS row stride of the staged tile (floats)
    float* st = reinterpret_cast<float*>(win);
    const bool staged = 2 * BP_LB * W >= TH * TS;  // uniform
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const int Y = Y0 + ty + 16 * q;
      if (X >= N || Y >= N) continue;
      const size_t p = (size_t)bz * N * N + (size_t)Y * N + X;
      const float v = acc_s[q] + acc_f[q];  // backprojection.fut:78
      const float xn = P.out[p] + P.colinv[(size_t)Y * N + X] * v;
      P.out[p] = xn;
      if (staged)
        st[(ty + 16 * q) * TS + tx] = xn;
      else
        P.outT[(size_t)bz * N * N + (size_t)X * N + Y] = xn;
    }
    if (staged) {
      __syncthreads();
      // thread -> column x = tid / 16 of the tile, rows PPT*(tid % 16) .. + PPT-1
      const int cx = threadIdx.x >> 4, cy = (threadIdx.x & 15) * PPT;
      const int XT = X0 + cx;
      if (XT < N) {
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          const int YT = Y0 + cy + k;
          if (YT < N) P.outT[(size_t)bz * N * N + (size_t)XT * N + YT] = st[(cy + k) * TS + cx];
        }
      }
    }
    return;
  }
  if (X >= N) return;
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int Y = Y0 + ty + 16 * q;
    if (Y >= N) continue;
    const size_t p = (size_t)bz * N * N + (size_t)Y * N + X;
    const float v = acc_s[q] + acc_f[q];  // backprojection.fut:78
    if constexpr (MODE == BP_PLAIN) {
      P.out[p] = v;
    } else {
      P.out[p] = (v == 0.0f) ? 0.0f : 1.0f / v;
    }
  }
}

// ------------------------------------------------------------------ staged forward projection
// A workgroup = FPS_RAYS consecutive bins of one line.  The frame rows are walked in blocks of
// FPS_ROWS: the waves load the block's image-row segment (the columns any of the workgroup's
// rays can touch, at most FPS_MAXW) into LDS with coalesced dword loads, then every lane forms
// its ray's row steps from LDS with the fast-mode step of fp_row_add.
constexpr int FPS_RAYS = 256;
constexpr int FPS_ROWS = 12;
constexpr int FPS_CPL = 6;                       // staged floats per lane and row
constexpr int FPS_MAXW = FPS_CPL * 64;           // widest segment (floats)
constexpr int FPS_CS = FPS_ROWS + 1;             // LDS stride of a staged column (floats)
constexpr int FPS_PF = FPS_ROWS / 2;             // rows of LDS reads in flight ahead of the sums

// f(integral_constant<int, I>) for I in the sequence, in order (compile-time unrolling).
template <int... I, class F>
__device__ __forceinline__ void unroll_seq(std::integer_sequence<int, I...>, F&& f) {
  (f(std::integral_constant<int, I>{}), ...);
}

typedef __attribute__((address_space(3))) float lds_float;

template <int K>
__device__ __forceinline__ int bcast16i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xf, 0xf, true);
}

// The staged walk of one ray per lane over the workgroup's rows i0 .. i1 (uniform) in blocks
// of ROWS (<= 16, a multiple of 4); the wave computes the blocks that meet its own rows i0w ..
// i1w.  CPL: staged chunks of 64 columns per row.  seg_start(ib, tab) (uniform): the block's
// segment start, one column left of floor(xbot) of the lowest ray of the workgroup over the
// block's rows.  steps: the line's exact step per binade (fp_step_table), or null.
template <bool POS, int NBUF, bool NT, int CPL, int ROWS, class SegStart>
__device__ __forceinline__ float fps_walk(const float* src, int N, int h, int i0, int i1,
                                          int i0w, int i1w, float slope, float icpt,
                                          SegStart seg_start, float* seg, float rs = 0.0f,
                                          const float2* __restrict__ steps = nullptr) {
  const int tid = threadIdx.x;
  const int l16 = tid & 15;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int RPW = ROWS / (FPS_RAYS / 64);  // staged rows per wave
  constexpr int CS = ROWS + 1;                 // LDS stride of a staged column (floats)
  constexpr int PF = ROWS / 2;                 // rows of LDS reads in flight ahead of the sums
  constexpr int W = 64 * CPL;                  // staged columns
  float st[RPW][CPL];
  // Per block, lane l16 holds row ib + l16's i*slope (tab); the block's segment starts at
  // image column sb = seg_start(ib, tab).
  auto block_rows = [&](int ib, float& tab, int& sb) {
    tab = (float)(ib + l16) * slope;
    sb = seg_start(ib, tab);
  };
  // Wave w stages rows RPW*w .. RPW*w + RPW-1 of the block, lane l columns s0 + l + 64m:
  // coalesced dword loads through the row's descriptor (zero outside the image row and
  // for rows past i1).
  auto load = [&](int ib, int s0) {
#pragma unroll
    for (int t = 0; t < RPW; ++t) {
      const int i = ib + RPW * wave + t;
      const __amdgpu_buffer_rsrc_t rr = row_rsrc(src + (size_t)(i < i1 ? i + h : 0) * N, i < i1 ? N : 0);
      const int vo = (s0 + lane) << 2;
      if (s0 >= 0) {  // wave-uniform: non-negative offsets, the constant parts as immediates
#pragma unroll
        for (int m = 0; m < CPL; ++m)
          st[t][m] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rr, vo + 256 * m, 0, 0));
      } else {
        // segment starting left of the image: each load's full offset in its own register
        // (a negative one is out of range as an unsigned offset and reads 0)
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
          int o = vo + 256 * m;
          asm volatile("" : "+v"(o));
          st[t][m] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rr, o, 0, 0));
        }
      }
    }
  };
  // LDS layout of a staged block: column-interleaved, column j of row slot r at j*CS + r
  // (CS = ROWS + 1 spreads consecutive columns over the banks), so a lane's pair in row K
  // is one address plus the constant offsets K and K + CS.
  auto store = [&](int buf) {
#pragma unroll
    for (int t = 0; t < RPW; ++t) {
      float* base = seg + (size_t)buf * W * CS + lane * CS + RPW * wave + t;
#pragma unroll
      for (int m = 0; m < CPL; ++m) base[64 * CS * m] = st[t][m];
    }
  };
  float acc = 0.0f;
  const int seg_lds = (int)(uintptr_t)(lds_float*)seg;  // LDS address of the block
  // Exact per-block step.  When every x_bot and x_top of the wave's rays over the block's rows
  // lies in one binade [2^e, 2^(e+1)), e >= 0 (tile_binade of the block's column range), all
  // of them are multiples of u = 2^(e-23), so x_top = x_bot + RN_u(slope) and the row's
  // |x_top - x_bot| is the line's constant D_e (binade_step; no rounding tie, D_e > 0).  The
  // row then needs no x_top, difference or reciprocal: with fr = fract(x_bot) (exact for
  // |x_bot| >= 1), x_top - floor(x_bot) = fr + D_e and floor(x_bot) - x_top = D_e - fr
  // exactly, and 1/D_e is fp_row_add's reciprocal of D_e (from the table) -- so each row's
  // fraction, and the sum, are bit for bit fp_row_add's.  The wave's x_bot extremes over the
  // block are at its extreme rays and rows (x_bot is monotone in the intercept and the row).
  // The block's column range from 16.16 fixed-point bounds on the scalar unit: x_bot =
  // RN(RN(i*slope) + icpt) lies within 3e-3 of i*slope + icpt (|i| <= 4112, |icpt| < 2^14),
  // the fixed-point products and intercepts within 2^-17*|i| + 2^-16 (< 0.04) of theirs, so
  // the floors of the fixed-point extremes, widened by one column each way, bound every lane's
  // and row's floor(x_bot).
  // Walks with |intercept| >= 2^14 or images over 8192 rows (where the products could leave
  // 32 bits) take the per-row step.
  const float ic0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, icpt), 0));
  const float ic1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, icpt), 63));
  const float ic_lo = fminf(ic0, ic1), ic_hi = fmaxf(ic0, ic1);
  const bool xok = steps && N <= 8192 && ic_lo > -16384.0f && ic_hi < 16384.0f;
  const int s_fx = __builtin_amdgcn_readfirstlane(xok ? (int)rintf(slope * 65536.0f) : 0);
  const int lo_fx = __builtin_amdgcn_readfirstlane(xok ? (int)floorf(ic_lo * 65536.0f) : 0);
  const int hi_fx = __builtin_amdgcn_readfirstlane(xok ? (int)ceilf(ic_hi * 65536.0f) : 0);
  auto block_step = [&](int ib, float2& dr) -> bool {
    if (!xok) return false;
    const int a = ib * s_fx, z = (ib + ROWS - 1) * s_fx;
    const int x0 = ((min(a, z) + lo_fx) >> 16) - 1, x1 = ((max(a, z) + hi_fx) >> 16) + 1;
    const int e = tile_binade(x0, x1);
    if (e < 0 || e >= FP_STEP_BINADES) return false;
    dr = steps[e];
    return dr.x > 0.0f;
  };
  // The LDS reads run PF rows ahead of the sums (rolling window of PF row registers).
  // xbs: the row's x_bot (per-row step) or fract(x_bot) (exact step).
  float xbs[PF], vas[PF], vcs[PF];
  auto fetch = [&](auto ex, auto kc, float tab, int bbase) {
    constexpr int K = decltype(kc)::value;
    constexpr int S = K % PF;
    const float xb = bcast16<K>(tab) + icpt;
    // LDS byte address of the pair's first column in row slot 0: 4*CS*floor(xbot) + bbase
    unsigned off;
    if constexpr (decltype(ex)::value) {
      xbs[S] = __builtin_amdgcn_fractf(xb);
      // floor(xbot) = xbot - fract(xbot) exactly; the address in FP32 is exact (integers
      // below 2^24)
      off = (unsigned)(int)fmaf(xb - xbs[S], (float)(4 * CS), (float)bbase);
    } else {
      xbs[S] = xb;
      off = (unsigned)(__mul24(floor_to_int(xb), 4 * CS) + bbase);
    }
    const lds_float* p = (const lds_float*)(uintptr_t)off;
    vas[S] = p[K];
    vcs[S] = p[K + CS];
  };
  auto sum_row = [&](auto ex, int S, float2 dr) {
    if constexpr (decltype(ex)::value) {
      const float fr = xbs[S], D = dr.x, rD = dr.y;
      float fo = POS ? fmaf(fr + D, rD, -rD) : (D - fr) * rD;
      fo = fminf(fmaxf(fo, 0.0f), 1.0f);
      const float vb = POS ? vas[S] : vcs[S], vo = POS ? vcs[S] : vas[S];
      acc = fmaf(fo, vo - vb, acc + vb);
    } else {
      const float xb = xbs[S], xt = xb + slope;
      acc = POS ? fp_row_add<true, NT>(acc, xb, xt, vas[S], vcs[S], rs)
                : fp_row_add<false, NT>(acc, xb, xt, vcs[S], vas[S], rs);
    }
  };
  auto step = [&](auto ex, auto kc, float tab, int bbase, float2 dr) {  // sum row K, fetch K + PF
    constexpr int K = decltype(kc)::value;
    sum_row(ex, K, dr);
    fetch(ex, std::integral_constant<int, K + PF>{}, tab, bbase);
    __builtin_amdgcn_sched_barrier(0);
  };
  auto compute = [&](auto ex, int buf, float tab, int sb, float2 dr) {
    // LDS byte address of the pair's first column in the block's row 0: 4*CS*floor(xbot) + bbase
    const int bbase = (buf * W * CS + (h - sb - (POS ? 0 : 1)) * CS) * 4 + seg_lds;
    unroll_seq(std::make_integer_sequence<int, PF>{}, [&](auto kc) { fetch(ex, kc, tab, bbase); });
    __builtin_amdgcn_sched_barrier(0);
    unroll_seq(std::make_integer_sequence<int, ROWS - PF>{}, [&](auto kc) { step(ex, kc, tab, bbase, dr); });
#pragma unroll
    for (int S = 0; S < PF; ++S) sum_row(ex, S, dr);
  };
  float tab, tabn;
  int sb, sbn;
  block_rows(i0, tab, sb);
  load(i0, sb);
  store(0);
  __syncthreads();
  int b = 0;
  for (int ib = i0; ib < i1; ib += ROWS, ++b) {  // i0, i1 uniform: every lane reaches each barrier
    const bool more = ib + ROWS < i1;
    if (more) {
      block_rows(ib + ROWS, tabn, sbn);
      load(ib + ROWS, sbn);
    }
    __builtin_amdgcn_sched_barrier(0);  // the next block's loads stay ahead of the compute
    // A wave skips blocks outside its own rays' row range: there they add exactly 0.
    if (ib + ROWS > i0w && ib < i1w) {
      float2 dr = make_float2(0.0f, 0.0f);
      if (block_step(ib, dr))
        compute(std::true_type{}, NBUF == 2 ? (b & 1) : 0, tab, sb, dr);
      else
        compute(std::false_type{}, NBUF == 2 ? (b & 1) : 0, tab, sb, dr);
    }
    if constexpr (NBUF == 1) __syncthreads();
    if (more) store(NBUF == 2 ? ((b + 1) & 1) : 0);
    __syncthreads();
    tab = tabn;
    sb = sbn;
  }
  return acc;
}

// One staged work item: bins blk*FPS_RAYS .. +FPS_RAYS-1 of line j of image bz, one ray per
// thread of a FPS_RAYS-thread workgroup.  seg: NBUF * FPS_CS * FPS_MAXW floats of LDS.  All
// barriers are workgroup-uniform; on return every wave has passed the walk's last barrier,
// so the caller may reuse seg at once.
template <int MODE, int NBUF>
__device__ __forceinline__ void fps_item(const FpParams& P, int j, int blk, int bz, float* seg) {
  const int first = blk * FPS_RAYS;
  const int r = first + (int)threadIdx.x;
  const Line L = P.lines[j];
  const int N = P.N, h = N >> 1;
  const float cost = L.cost, slope = L.slope;
  const float icpt = (P.rho0 + (float)r * P.drho) / cost;  // forwardprojection.fut:33-34
  // The workgroup's intercepts are monotone in r: the extremes are lanes 0 and FPS_RAYS-1.
  const float ia = (P.rho0 + (float)first * P.drho) / cost;
  const float iz = (P.rho0 + (float)(first + FPS_RAYS - 1) * P.drho) / cost;
  const float icpt_lo = fminf(ia, iz);
  int i0, i1, i0w, i1w;
  {
    // the workgroup's row range (hull over its rays) and the wave's own
    int a0, a1, b0, b1;
    const int last = min(first + FPS_RAYS - 1, P.R - 1);
    row_clip(slope, ia, h, a0, a1);
    row_clip(slope, (P.rho0 + (float)last * P.drho) / cost, h, b0, b1);
    i0 = __builtin_amdgcn_readfirstlane(min(a0, b0));
    i1 = __builtin_amdgcn_readfirstlane(max(a1, b1));
    const int wfirst = first + (int)(threadIdx.x & ~63u);
    const int wlast = min(wfirst + 63, P.R - 1);
    if (wfirst > wlast) {
      i0w = i1w = i0;  // no valid ray in this wave
    } else {
      row_clip(slope, (P.rho0 + (float)wfirst * P.drho) / cost, h, a0, a1);
      row_clip(slope, (P.rho0 + (float)wlast * P.drho) / cost, h, b0, b1);
      i0w = min(a0, b0);
      i1w = max(a1, b1);
    }
    i0w = __builtin_amdgcn_readfirstlane(i0w);
    i1w = __builtin_amdgcn_readfirstlane(i1w);
  }
  const float* src = ((L.flags & LINE_FLAT) ? P.imgT : P.img) + (size_t)bz * N * N;
  // The block's segment start: xbot is rounded from i*slope + icpt exactly as each lane
  // rounds it and rounding is monotone, so the lowest ray's floor(xbot) - 1 over the block's
  // first and last rows lies at or left of every lane's.
  auto seg_start = [&](int ib, float tab) {
    const int sv = floor_to_int(tab + icpt_lo) + h - 1;
    return min(__builtin_amdgcn_readlane(sv, 0), __builtin_amdgcn_readlane(sv, FPS_ROWS - 1));
  };
  float acc = 0.0f;
  if (i0 < i1) {
    const bool nt = L.flags & LINE_NEWTON;
    const float rs = nt ? 1.0f / fabsf(slope) : 0.0f;
    if ((L.flags & LINE_SLOPE_POS) && nt)
      acc = fps_walk<true, NBUF, true, FPS_CPL, FPS_ROWS>(src, N, h, i0, i1, i0w, i1w, slope, icpt, seg_start, seg, rs);
    else if (L.flags & LINE_SLOPE_POS)
      acc = fps_walk<true, NBUF, false, FPS_CPL, FPS_ROWS>(src, N, h, i0, i1, i0w, i1w, slope, icpt, seg_start, seg);
    else if (nt)
      acc = fps_walk<false, NBUF, true, FPS_CPL, FPS_ROWS>(src, N, h, i0, i1, i0w, i1w, slope, icpt, seg_start, seg, rs);
    else
      acc = fps_walk<false, NBUF, false, FPS_CPL, FPS_ROWS>(src, N, h, i0, i1, i0w, i1w, slope, icpt, seg_start, seg);
  }
  if (r >= P.R) return;  // after the walk: every lane took part in its barriers
  const float s = acc * L.lbase;
  const size_t o = (size_t)bz * P.A * P.R + (size_t)L.orig * P.R + r;
  if constexpr (MODE == FP_PLAIN)
    P.out[o] = s;
  else
    P.out[o] = P.rowinv[(size_t)L.orig * P.R + r] * (P.b[o] - s);  // SIRT.fut:33-36
}

// ------------------------------------------------------- grouped staged forward projection
// A workgroup = FPG_BINS (64) consecutive bins of up to FPG_LINES (4) lines of one frame that
// are adjacent in slope (the geometry's group table); wave w walks line w's 64 rays.  The
// rays of such lines at the same bins cross a frame row within a few columns of each other,
// so a block's staged segment (every column any of the 256 rays touches) is CPL = 2..4 chunks
// of 64 columns, where 256 bins of one line need 6: the staging per ray-row drops to a third
// or a half, and blocks of FPG_ROWS = 16 rows fit (fewer per-block steps per row).  The walk
// and each ray's sums are fps_walk's (bitwise equal results).
constexpr int FPG_LINES = FPS_RAYS / 64;
constexpr int FPG_BINS = 64;
constexpr int FPG_MAXCPL = 4;
constexpr int FPG_ROWS = 16;  // rows per staged block (8.5 KB of LDS per buffer at CPL 2)

// A group table entry: the lines' table indices (x: first line | CPL << 24; y, z, w: the
// others, -1 past the group's size).
__device__ __forceinline__ int fpg_line(const int4& g, int k) {
  return k == 0 ? (g.x & 0xffffff) : (k == 1 ? g.y : (k == 2 ? g.z : g.w));
}

// XSTEP: take the exact per-block step where the table allows (off: the per-row step only).
template <int MODE, int NBUF, int CPL, bool XSTEP = true>
__device__ __forceinline__ void fpg_item(const FpParams& P, const int4 grp, int blk, int bz, float* seg) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int nl = 1 + (grp.y >= 0) + (grp.z >= 0) + (grp.w >= 0);  // lines fill slots 0 .. nl-1
  const bool active = wave < nl;
  const int N = P.N, h = N >> 1;
  const int first = blk * FPG_BINS, r = first + lane;
  const int last = min(first + FPG_BINS - 1, P.R - 1);
  const Line L = P.lines[fpg_line(grp, active ? wave : 0)];
  const float cost = L.cost, slope = L.slope;
  const float icpt = (P.rho0 + (float)r * P.drho) / cost;  // forwardprojection.fut:33-34
  // Lane k < 2*nl describes line k/2 at its first (k even) or last ray: the workgroup's row
  // hull (valid rays) and the block segment starts (every lane's ray, valid or not: the
  // intercepts are monotone in the bin, so the extremes are the block's first and last bin).
  // Lanes past 2*nl repeat lane 0.
  const int kk = lane < 2 * nl ? lane : 0;
  const Line Lk = P.lines[fpg_line(grp, kk >> 1)];
  const float ia = (P.rho0 + (float)first * P.drho) / Lk.cost;
  const float iz = (P.rho0 + (float)(first + FPG_BINS - 1) * P.drho) / Lk.cost;
  const float sl_k = Lk.slope, ilo_k = fminf(ia, iz);
  const int ro_k = (kk & 1) ? FPG_ROWS - 1 : 0;
  int i0 = 0, i1 = 0, i0w = 0, i1w = 0;
  {
    int c0, c1;
    row_clip(sl_k, (kk & 1) ? (P.rho0 + (float)last * P.drho) / Lk.cost : ia, h, c0, c1);
    i0 = __builtin_amdgcn_readlane(c0, 0);
    i1 = __builtin_amdgcn_readlane(c1, 0);
#pragma unroll
    for (int k = 1; k < 2 * FPG_LINES; ++k) {
      i0 = min(i0, __builtin_amdgcn_readlane(c0, k));
      i1 = max(i1, __builtin_amdgcn_readlane(c1, k));
    }
    // the wave's own rows: its line's lanes (2*wave, 2*wave + 1)
    if (active) {
      i0w = min(__builtin_amdgcn_readlane(c0, 2 * wave), __builtin_amdgcn_readlane(c0, 2 * wave + 1));
      i1w = max(__builtin_amdgcn_readlane(c1, 2 * wave), __builtin_amdgcn_readlane(c1, 2 * wave + 1));
    } else {
      i0w = i1w = i0;  // no line: stages its rows, computes nothing
    }
  }
  // Segment start of the block at row ib: the lowest floor(xbot) - 1 over the group's lines
  // at the block's first and last rows, xbot rounded exactly as the lanes round it.
  auto seg_start = [&](int ib, float) {
    const int sv = floor_to_int((float)(ib + ro_k) * sl_k + ilo_k) + h - 1;
    int sb = __builtin_amdgcn_readlane(sv, 0);
#pragma unroll
    for (int k = 1; k < 2 * FPG_LINES; ++k) sb = min(sb, __builtin_amdgcn_readlane(sv, k));
    return sb;
  };
  const float* src = ((L.flags & LINE_FLAT) ? P.imgT : P.img) + (size_t)bz * N * N;
  const float2* lsteps =
      XSTEP && P.steps ? P.steps + (size_t)fpg_line(grp, active ? wave : 0) * FP_STEP_BINADES : nullptr;
  float acc = 0.0f;
  if (i0 < i1) {
    const bool nt = L.flags & LINE_NEWTON;
    const float rs = nt ? 1.0f / fabsf(slope) : 0.0f;
    if ((L.flags & LINE_SLOPE_POS) && nt)
      acc = fps_walk<true, NBUF, true, CPL, FPG_ROWS>(src, N, h, i0, i1, i0w, i1w, slope, icpt, seg_start, seg, rs, lsteps);
    else if (L.flags & LINE_SLOPE_POS)
      acc = fps_walk<true, NBUF, false, CPL, FPG_ROWS>(src, N, h, i0, i1, i0w, i1w, slope, icpt, seg_start, seg, 0.0f, lsteps);
    else if (nt)
      acc = fps_walk<false, NBUF, true, CPL, FPG_ROWS>(src, N, h, i0, i1, i0w, i1w, slope, icpt, seg_start, seg, rs, lsteps);
    else
      acc = fps_walk<false, NBUF, false, CPL, FPG_ROWS>(src, N, h, i0, i1, i0w, i1w, slope, icpt, seg_start, seg, 0.0f, lsteps);
  }
  if (!active || r >= P.R) return;  // after the walk: every lane took part in its barriers
  const float s = acc * L.lbase;
  const size_t o = (size_t)bz * P.A * P.R + (size_t)L.orig * P.R + r;
  if constexpr (MODE == FP_PLAIN)
    P.out[o] = s;
  else
    P.out[o] = P.rowinv[(size_t)L.orig * P.R + r] * (P.b[o] - s);  // SIRT.fut:33-36
}


}  // namespace sk
