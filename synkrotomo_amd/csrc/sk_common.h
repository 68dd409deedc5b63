// sk_common.h -- internal declarations shared by the libsynkro translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/synkro.h"

namespace sk {

// Flags of a frame line.
enum : int {
  LINE_FLAT = 1,      // flat angle: the line acts on the transposed image
  LINE_SLOPE_POS = 2, // slope >= 0
  LINE_DEGEN = 4,     // |slope| < 1e-3: x_top may round onto x_bot (theta ~ 0, pi/2)
  LINE_FLAG_BITS = 7, // the flags above (the exported line table carries these)
  // |slope| large enough for the fast backprojector to take 1/(x_top - x_bot) by one Newton
  // step from 1/|slope| (newton_min_slope in sk_geom.hip); set only on geometry tables
  LINE_NEWTON = 8,
  LINE_KEY_BITS = 15,
  // bits >= LINE_RUN_SHIFT: number of lines from this one to the end of its run of
  // consecutive table entries with equal LINE_KEY_BITS (>= 1)
  LINE_RUN_SHIFT = 8
};

// One line in its frame (futhark/sirtlib.fut:39-41): steep angles keep the image frame
// with (cost, sint) = (cos, -sin); flat angles use the transposed frame with
// (cost, sint) = (-sin, cos).  32 bytes, read with scalar loads by the kernels.
struct alignas(32) Line {
  float cost, sint, lbase, slope;  // slope = -sint/cost (forwardprojection.fut:29)
  float fc, fs;                    // fact*cost, fact*sint, fact = sqrt(2)/2 (backprojection.fut:20)
  int orig;                        // sinogram row (caller's angle index)
  int flags;
};

struct Geom {
  int A = 0, R = 0, N = 0, n_steep = 0, ncand = 0;
  float rho0 = 0.f, drho = 1.f;
  // Unit bins with two candidates per pixel and |rho0| + N < 2^20, where the fast candidate
  // index (rounding error < 0.4 bin) needs no clamp: the specialised backprojector applies.
  bool bp_two_cand = false;
  // Row-segment width (floats) the staged forward projector needs for 256 consecutive bins.
  int fps_w = 0;
  // Grouped staged forward projector: line groups (int4 entries, fpg_groups in sk_geom.hip),
  // their count and the largest staged chunk count among them (null: not applicable).
  int4* d_fpg = nullptr;
  int n_fpg = 0, fpg_cpl = 0;
  // and its exact per-block step per line and binade (k_fp_step_table in sk_geom.hip)
  float2* d_fp_steps = nullptr;
  std::vector<Line> lines;   // host copy, steep lines then flat lines
  Line* d_lines = nullptr;
  // backprojector intercepts (s*drho + rho0)/cost per table line j and bin s, [A][R] (null when
  // A*R >= 2^31: the kernels then divide)
  float* d_inter = nullptr;
  // backprojector tile launch orders for 1, 2 and 4 pixels per lane (bp_tile_orders in
  // sk_geom.hip; null where the hoisted fast path does not apply), one buffer
  unsigned* d_bp_order = nullptr;
  size_t bp_order_off[3] = {0, 0, 0};
  // SIRT weights, computed on first use (SIRT.fut:22-25)
  float* d_rowinv = nullptr;  // [A][R]
  float* d_colinv = nullptr;  // [N][N]
  int weights_exact = -1;     // SK_OPT_EXACT value the weights were computed with
  int fp_variant = -1;        // tuned kernel variants (-1: the context's default)
  int bp_variant = -1;
  int device = 0;
  unsigned long long uid = 0;  // unique per geometry created in this process (graph cache keys)
  // Cached SIRT iteration graphs and the arguments each was captured with (a few, so that
  // callers alternating between buffer sets -- double-buffered streaming -- reuse them).
  struct SirtGraph {
    hipGraphExec_t exec = nullptr;
    const float* img = nullptr;
    const float* b = nullptr;
    const void* ws = nullptr;
    int batch = 0, exact = -1, iters = 0, fpv = -2, bpv = -2;
    unsigned long long used = 0;  // last use (LRU eviction)
  };
  static constexpr int SIRT_GRAPHS = 4;
  SirtGraph sirt_graphs[SIRT_GRAPHS];
  unsigned long long sirt_clock = 0;
};

struct Ctx {
  int device = 0;
  std::string err;
  int exact = 0;
  int fp_variant = 0;
  int bp_variant = 0;
  // growable scratch (stream-ordered use; grown after a stream sync)
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  hipStream_t capture_stream = nullptr;  // private stream used for hipGraph capture
  // Executable graphs of sk_sirt_batched, keyed by every value their kernel arguments bake in
  // (problem pointers, geometry ids, weights, scratch address, iterations, mode; LRU).
  struct BatchGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned long long> key;
    unsigned long long used = 0;
  };
  static constexpr int BATCH_GRAPHS = 4;
  BatchGraph batch_graphs[BATCH_GRAPHS];
  unsigned long long batch_clock = 0;
};

// Error helpers: record a message in the context and return the code.
int fail(Ctx* c, int code, const char* fmt, ...);
int hip_fail(Ctx* c, hipError_t e, const char* what);

#define SK_HIP(ctx, expr)                                   \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return ::sk::hip_fail(ctx, _e, #expr); \
  } while (0)

// Scratch memory of at least `bytes`, ordered after `stream`.
int scratch(Ctx* c, size_t bytes, hipStream_t stream, void** out);

// Device address of sk_fp_noalloc's fixed scratch on `device` (current), looked up once per
// device (sk_noalloc.hip).
int noalloc_init(Ctx* c, int device);

// Kernel launchers (sk_proj.hip).  batch >= 1; strides are per-batch element counts.
enum FpMode { FP_PLAIN = 0, FP_ONES_INV = 1, FP_SIRT = 2 };
enum BpMode { BP_PLAIN = 0, BP_ONES_INV = 1, BP_SIRT = 2 };
enum BpWeights { W_REF = 0, W_ADJ = 1 };

struct FpArgs {
  const float* img;    // [batch][N][N]  (unused in FP_ONES_INV)
  const float* imgT;   // [batch][N][N] transposed copy (flat lines)
  float* out;          // [batch][A][R]
  const float* b;      // FP_SIRT: measured sinogram [batch][A][R]
  const float* rowinv; // FP_SIRT: [A][R]
  int batch;
};
struct BpArgs {
  const float* sino;   // [batch][A][R]   (unused in BP_ONES_INV)
  float* out;          // [batch][N][N]   (BP_SIRT: the image, updated in place)
  float* outT;         // BP_SIRT: transposed copy of the updated image
  const float* colinv; // BP_SIRT: [N][N]
  int batch;
};

int launched(Ctx* c, const char* what);
int launch_transpose(Ctx* c, const float* in, float* out, int N, int batch, hipStream_t s);
int fp_rays_per_lane(int variant);
int bp_pixels_per_lane(int variant);
int launch_fp(Ctx* c, const Geom* g, FpMode mode, const FpArgs& a, hipStream_t s);
int launch_fp_bpt(Ctx* c, const Geom* g, const FpArgs& a, hipStream_t s);
int launch_bp(Ctx* c, const Geom* g, BpMode mode, BpWeights w, const BpArgs& a, hipStream_t s);

}  // namespace sk

struct sk_ctx {
  sk::Ctx c;
};
struct sk_geom {
  sk::Geom g;
};
