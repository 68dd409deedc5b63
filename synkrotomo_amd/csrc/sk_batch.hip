// sk_batch.hip -- many SIRT problems of uneven size in persistent launches, and the
// kernel-variant autotuner.
//
// Batched SIRT (no reference counterpart; generalizes SIRT3D.fut:18-28, which maps SIRT over
// the slices of one geometry): a problem set holds groups of independent problems, each
// group with its own geometry (image size, angles, bins).  One iteration is two persistent
// launches sized to fill the chip once: in k_fpg_persist (fast mode) every workgroup pulls
// (problem, image, line group, 64-bin block) items of the grouped staged forward projector
// from a global work counter (k_fps_persist: (problem, image, line, 256-bin block) items of
// the one-line staged kernel) (k_fp_persist: waves pull 64-bin items of the per-wave walk, exact mode or
// segments wider than the staging buffer); in k_bp_persist every block pulls (problem,
// image, tile) items, queued largest-problem first so the tail of the queue is made of the
// shortest items.  The fp kernel resets the bp counter and the
// bp kernel the fp slot, so the whole iteration sequence (a reset kernel, then fp/bp pairs)
// is captured once into a hipGraph and launched once.
//
// Autotuner: times the fp variants (per-wave kernels with 1, 2 or 4 rays per lane, staged
// kernels) and the bp variants (16x16,
// 16x32 or 16x64 pixel tiles) on a geometry with HIP events and keeps the fastest.  All
// variants evaluate the same expressions in the same order per ray and per pixel, so the
// choice never changes results.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "sk_kern.h"

namespace sk {

int launched(Ctx* c, const char* what);
int ensure_weights(Ctx* c, Geom* g, hipStream_t s);
FpParams fp_params(const Geom* g, const FpArgs& a);
BpParams bp_params(const Geom* g, const BpArgs& a, int ppt);

struct BatchProb {
  FpParams fp;
  BpParams bp;
  int batch;
  int fp_units;       // forward items per bin group: lines (line groups for the grouped kernel)
  int fp_wb;          // wave-sized bin groups per line
  int bp_tx, bp_ty;   // tiles across / down
  long fp_begin, bp_begin;
};

__device__ __forceinline__ int u32(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The problem table travels by value in the kernel arguments (kernarg memory is read with
// scalar loads), so every per-problem value -- and every row pointer built from it -- stays
// in scalar registers.
constexpr int MAX_GROUPS = 16;
struct BatchArgs {
  BatchProb p[MAX_GROUPS];
  int n;
  long fp_total, bp_total;
};

__global__ void k_reset_counters(unsigned* ctr) {
  if (threadIdx.x < 3) ctr[threadIdx.x] = 0u;
}

// Uniform lookup of the problem owning global item `item`.
template <bool FP>
__device__ __forceinline__ int find_prob(const BatchArgs& a, long item) {
  int p = 0;
  while (p + 1 < a.n && item >= (FP ? a.p[p + 1].fp_begin : a.p[p + 1].bp_begin)) ++p;
  return p;
}

// Forward items are one wave each (64*RPT bins of one line of one image).  DYN: waves claim
// CLAIM consecutive items at a time from a global counter (items vary in cost -- rays near
// the detector edge are clipped short -- and the counter balances them); otherwise items
// are strided statically over the grid's waves.  No barriers either way.  Consecutive
// items are the same bins of neighbouring lines (angles), whose rays cross nearly the same
// pixels at the same time, so concurrent gathers share cache lines.
constexpr int CLAIM = 2;

template <bool EXACT, int RPT, bool DYN>
__global__ __launch_bounds__(FP_BLOCK) void k_fp_persist(const BatchArgs a, unsigned* __restrict__ ctr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr[1] = 0u;  // the bp queue of this iteration
  const long nwaves = (long)gridDim.x * FP_WAVES;
  const int wave = u32(threadIdx.x >> 6);  // uniform per wave
  const int lane = threadIdx.x & 63;
  long item = DYN ? 0 : (long)blockIdx.x * FP_WAVES + wave;
  int left = 0;
  for (;;) {
    if (DYN) {
      if (left == 0) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(&ctr[2], (unsigned)CLAIM);
        item = (long)(unsigned)u32((int)base);
        left = CLAIM;
      }
      --left;
    }
    if (item >= a.fp_total) break;
    const BatchProb& B = a.p[find_prob<true>(a, item)];
    const long local = item - B.fp_begin;
    const long per_img = (long)B.fp.A * B.fp_wb;
    const int bz = (int)(local / per_img);
    const long rem = local - (long)bz * per_img;
    const int wb = (int)(rem / B.fp.A);
    const int j = (int)(rem - (long)wb * B.fp.A);
    fp_item<EXACT, FP_SIRT, W_REF, true, RPT>(B.fp, u32(j), u32(wb), u32(bz));
    item += DYN ? 1 : nwaves;
  }
}

// Staged forward items (fast mode, every group's segments fit FPS_MAXW): one workgroup per
// item of FPS_RAYS bins of one line of one image, claimed one at a time from the global
// counter (the claim is broadcast through LDS; its two barriers are workgroup-uniform), in
// the same lines-fastest order.  NBUF = 1: FPS_CS * FPS_MAXW floats of dynamic LDS.
__global__ __launch_bounds__(FPS_RAYS, 8) void k_fps_persist(const BatchArgs a, unsigned* __restrict__ ctr) {
  extern __shared__ float4 seg4[];
  __shared__ unsigned item_sh;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr[1] = 0u;  // the bp queue of this iteration
  for (;;) {
    if (threadIdx.x == 0) item_sh = atomicAdd(&ctr[2], 1u);
    __syncthreads();
    const long item = item_sh;
    __syncthreads();
    if (item >= a.fp_total) break;
    const BatchProb& B = a.p[find_prob<true>(a, item)];
    const long local = item - B.fp_begin;
    const long per_img = (long)B.fp.A * B.fp_wb;
    const int bz = (int)(local / per_img);
    const long rem = local - (long)bz * per_img;
    const int wb = (int)(rem / B.fp.A);
    const int j = (int)(rem - (long)wb * B.fp.A);
    fps_item<FP_SIRT, 1>(B.fp, u32(j), u32(wb), u32(bz), (float*)seg4);
  }
}

// Grouped staged forward items (fast mode, every group has a line-group table): one
// workgroup per item of FPG_BINS bins of one line group of one image, groups fastest, claimed
// as in k_fps_persist; the per-row step only (the exact per-block step's registers would
// spill here).  Two buffers of (FPG_ROWS + 1) * 64 * CPL floats of dynamic LDS for the
// largest CPL of the set.
__global__ __launch_bounds__(FPS_RAYS, 8) void k_fpg_persist(const BatchArgs a, unsigned* __restrict__ ctr) {
  extern __shared__ float4 seg4[];
  __shared__ unsigned item_sh;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr[1] = 0u;  // the bp queue of this iteration
  for (;;) {
    if (threadIdx.x == 0) item_sh = atomicAdd(&ctr[2], 1u);
    __syncthreads();
    const long item = item_sh;
    __syncthreads();
    if (item >= a.fp_total) break;
    const BatchProb& B = a.p[find_prob<true>(a, item)];
    const long local = item - B.fp_begin;
    const long per_img = (long)B.fp_units * B.fp_wb;
    const int bz = (int)(local / per_img);
    const long rem = local - (long)bz * per_img;
    const int wb = (int)(rem / B.fp_units);
    const int g = (int)(rem - (long)wb * B.fp_units);
    const int4 grp = B.fp.groups[u32(g)];
    const int cpl = u32(grp.x >> 24);
    if (cpl <= 2)
      fpg_item<FP_SIRT, 2, 2, false>(B.fp, grp, u32(wb), u32(bz), (float*)seg4);
    else if (cpl == 3)
      fpg_item<FP_SIRT, 2, 3, false>(B.fp, grp, u32(wb), u32(bz), (float*)seg4);
    else
      fpg_item<FP_SIRT, 2, 4, false>(B.fp, grp, u32(wb), u32(bz), (float*)seg4);
  }
}

template <bool EXACT, int NC, bool DR1, int PPT>
__global__ __launch_bounds__(256) void k_bp_persist(const BatchArgs a, unsigned* __restrict__ ctr) {
  extern __shared__ float2 win[];
  __shared__ BpLineLds lsh[BP_LB];
  __shared__ unsigned item_sh;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr[2] = 0u;  // the fp queue of the next iteration
  for (;;) {
    if (threadIdx.x == 0) item_sh = atomicAdd(&ctr[1], 1u);
    __syncthreads();
    const long item = item_sh;
    __syncthreads();
    if (item >= a.bp_total) break;
    const BatchProb& B = a.p[find_prob<false>(a, item)];
    const long local = item - B.bp_begin;
    const long per_img = (long)B.bp_tx * B.bp_ty;
    const int bz = (int)(local / per_img);
    const long rem = local - (long)bz * per_img;
    int ty = (int)(rem / B.bp_tx);
    int tx = (int)(rem - (long)ty * B.bp_tx);
    if (B.bp.order) {  // the image's tiles dearest first, as in k_bp
      const unsigned t = B.bp.order[rem];
      tx = (int)(t & 0xffffu);
      ty = (int)(t >> 16);
    }
    bp_item<EXACT, BP_SIRT, W_REF, NC, DR1, PPT>(B.bp, u32(tx), u32(ty), u32(bz), win, lsh);
  }
}

struct BatchPlan {
  BatchArgs args{};
  // Per-wave forward kernel: bit0 2 rays per lane, bit1 dynamic claiming.  2 measured fastest
  // of those on the batched configuration (3.82 s vs 4.02 s for mode 3 and 5.10 s for mode 0:
  // static striding leaves waves with the costliest items idle-waiting at the end).
  // SYNKRO_BATCH_FP_MODE overrides it; 4 = the staged kernel, the default in fast mode when
  // every group's segments fit; 5 = the grouped staged kernel, the default in fast mode when
  // every group has a line-group table.
  int fp_mode = 2;
  int fpg_cpl = 0;  // mode 5: the largest chunk count of the set's line groups
  int bp_ppt = 4;
  unsigned* d_ctr = nullptr;
  int Wmax = 0;
  bool dr1 = true;
};

static size_t fpg_lds(int cpl) { return sizeof(float) * 2 * (FPG_ROWS + 1) * 64 * (size_t)cpl; }

// Pixels per lane of the persistent backprojector (16 x 16*PPT tiles): 4 by default (16x64
// tiles: 3.11 s vs 3.23 s with 16x32 on the four large groups of the batched configuration,
// equal on small groups); SYNKRO_BATCH_BP_PPT=2 selects 16x32 tiles.

template <bool EXACT, int NC, bool DR1>
static int launch_iteration(Ctx* c, const BatchPlan& plan, int fp_grid, int bp_grid, hipStream_t s) {
  if (plan.fp_mode == 5)
    hipLaunchKernelGGL(k_fpg_persist, dim3(fp_grid), dim3(FPS_RAYS), fpg_lds(plan.fpg_cpl), s, plan.args,
                       plan.d_ctr);
  else if (plan.fp_mode == 4)
    hipLaunchKernelGGL(k_fps_persist, dim3(fp_grid), dim3(FPS_RAYS), sizeof(float) * FPS_CS * FPS_MAXW, s,
                       plan.args, plan.d_ctr);
  else if (plan.fp_mode == 0)
    hipLaunchKernelGGL((k_fp_persist<EXACT, 1, false>), dim3(fp_grid), dim3(FP_BLOCK), 0, s, plan.args, plan.d_ctr);
  else if (plan.fp_mode == 1)
    hipLaunchKernelGGL((k_fp_persist<EXACT, 2, false>), dim3(fp_grid), dim3(FP_BLOCK), 0, s, plan.args, plan.d_ctr);
  else if (plan.fp_mode == 2)
    hipLaunchKernelGGL((k_fp_persist<EXACT, 1, true>), dim3(fp_grid), dim3(FP_BLOCK), 0, s, plan.args, plan.d_ctr);
  else
    hipLaunchKernelGGL((k_fp_persist<EXACT, 2, true>), dim3(fp_grid), dim3(FP_BLOCK), 0, s, plan.args, plan.d_ctr);
  const size_t lds = sizeof(float2) * BP_LB * (size_t)plan.Wmax;
  if (plan.bp_ppt == 4)
    hipLaunchKernelGGL((k_bp_persist<EXACT, NC, DR1, 4>), dim3(bp_grid), dim3(256), lds, s, plan.args, plan.d_ctr);
  else
    hipLaunchKernelGGL((k_bp_persist<EXACT, NC, DR1, 2>), dim3(bp_grid), dim3(256), lds, s, plan.args, plan.d_ctr);
  return launched(c, "persistent SIRT kernels");
}

static int persistent_grid(Ctx* c, const void* kernel, int block, size_t lds, int* out) {
  int dev = c->device, cus = 0, per_cu = 0;
  SK_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  SK_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds));
  *out = std::max(1, cus * std::max(1, per_cu));
  return SK_OK;
}

}  // namespace sk

using namespace sk;

extern "C" {

int sk_sirt_batched(sk_ctx* ctx, const sk_problem* problems, int count, int iters, void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (!problems || count < 1) return fail(c, SK_EINVAL, "need at least one problem group");
  if (count > MAX_GROUPS) return fail(c, SK_EINVAL, "at most %d problem groups per call", MAX_GROUPS);
  if (iters < 0) return fail(c, SK_EINVAL, "iterations must be >= 0 (got %d)", iters);
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < count; i++) {
    const sk_problem& P = problems[i];
    if (!P.geom || !P.d_img || !P.d_sino) return fail(c, SK_EINVAL, "problem %d: NULL argument", i);
    if (P.batch < 1) return fail(c, SK_EINVAL, "problem %d: batch must be >= 1", i);
    if (P.geom->g.device != c->device) return fail(c, SK_EINVAL, "problem %d: other device", i);
  }
  if (iters == 0) return SK_OK;
  for (int i = 0; i < count; i++) {
    int rc = ensure_weights(c, &problems[i].geom->g, s);
    if (rc) return rc;
  }

  // Largest problems first (their items are the longest).
  std::vector<int> order(count);
  for (int i = 0; i < count; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const Geom &ga = problems[a].geom->g, &gb = problems[b].geom->g;
    return (long)ga.N * ga.A > (long)gb.N * gb.A;
  });

  // Scratch: counters | per group: x^T, q
  size_t ws_floats = 0;
  for (int i = 0; i < count; i++) {
    const Geom& G = problems[i].geom->g;
    ws_floats += (size_t)problems[i].batch * ((size_t)G.N * G.N + (size_t)G.A * G.R);
  }
  void* ws = nullptr;
  int rc = scratch(c, 256 + sizeof(float) * ws_floats, s, &ws);
  if (rc) return rc;
  BatchPlan plan;
  plan.d_ctr = (unsigned*)ws;
  float* fws = (float*)((char*)ws + 256);
  BatchArgs& args = plan.args;
  args.n = count;
  const bool ex = c->exact != 0;
  bool staged_fits = !ex, grouped_fits = !ex;
  for (int i = 0; i < count; i++) {
    const Geom& G = problems[i].geom->g;
    staged_fits = staged_fits && G.fps_w > 0 && G.fps_w <= FPS_MAXW;
    grouped_fits = grouped_fits && G.d_fpg != nullptr;
    plan.fpg_cpl = std::max(plan.fpg_cpl, G.fpg_cpl);
  }
  if (staged_fits) plan.fp_mode = 4;
  if (grouped_fits) plan.fp_mode = 5;
  if (const char* m = getenv("SYNKRO_BATCH_FP_MODE")) {
    const int want = atoi(m);
    if (want == 5 || want == 4)
      plan.fp_mode = (want == 5 ? grouped_fits : staged_fits) ? want : plan.fp_mode;
    else
      plan.fp_mode = want & 3;
  }
  if (const char* m = getenv("SYNKRO_BATCH_BP_PPT")) plan.bp_ppt = atoi(m) == 2 ? 2 : 4;
  const int ppt = plan.bp_ppt;
  const int fp_item_bins =
      plan.fp_mode == 5 ? FPG_BINS : (plan.fp_mode == 4 ? FPS_RAYS : 64 * ((plan.fp_mode & 1) ? 2 : 1));
  for (int oi = 0; oi < count; oi++) {
    const sk_problem& P = problems[order[oi]];
    const Geom& G = P.geom->g;
    float* imgT = fws;
    fws += (size_t)P.batch * G.N * G.N;
    float* q = fws;
    fws += (size_t)P.batch * G.A * G.R;
    BatchProb B{};
    B.fp = fp_params(&G, FpArgs{P.d_img, imgT, q, P.d_sino, G.d_rowinv, P.batch});
    B.bp = bp_params(&G, BpArgs{q, P.d_img, imgT, G.d_colinv, P.batch}, ppt);
    B.batch = P.batch;
    B.fp_units = plan.fp_mode == 5 ? G.n_fpg : G.A;
    B.fp_wb = (G.R + fp_item_bins - 1) / fp_item_bins;
    B.bp_tx = (G.N + BP_TW - 1) / BP_TW;
    B.bp_ty = (G.N + 16 * ppt - 1) / (16 * ppt);
    B.fp_begin = args.fp_total;
    B.bp_begin = args.bp_total;
    args.fp_total += (long)P.batch * B.fp_units * B.fp_wb;
    args.bp_total += (long)P.batch * B.bp_tx * B.bp_ty;
    plan.Wmax = std::max(plan.Wmax, B.bp.W);
    plan.dr1 = plan.dr1 && G.bp_two_cand;
    args.p[oi] = B;
  }
  if (args.bp_total >= (1L << 32))
    return fail(c, SK_EINVAL, "problem set too large for one work queue");
  if (sizeof(float2) * BP_LB * (size_t)plan.Wmax > 60 * 1024)
    return fail(c, SK_EINVAL, "backprojection window exceeds LDS");

  const size_t lds = sizeof(float2) * BP_LB * (size_t)plan.Wmax;
  int fp_grid = 0, bp_grid = 0;
  // occupancy (and so the persistent grid) of the forward kernel that will run
  const void* kfp_ex[4] = {(const void*)k_fp_persist<true, 1, false>, (const void*)k_fp_persist<true, 2, false>,
                           (const void*)k_fp_persist<true, 1, true>, (const void*)k_fp_persist<true, 2, true>};
  const void* kfp_fast[4] = {(const void*)k_fp_persist<false, 1, false>, (const void*)k_fp_persist<false, 2, false>,
                             (const void*)k_fp_persist<false, 1, true>, (const void*)k_fp_persist<false, 2, true>};
  const void* kfp = plan.fp_mode == 5   ? (const void*)k_fpg_persist
                    : plan.fp_mode == 4 ? (const void*)k_fps_persist
                                        : (ex ? kfp_ex[plan.fp_mode & 3] : kfp_fast[plan.fp_mode & 3]);
  const void* kbp2 =
      plan.dr1 ? (ex ? (const void*)k_bp_persist<true, 2, true, 2> : (const void*)k_bp_persist<false, 2, true, 2>)
               : (ex ? (const void*)k_bp_persist<true, 0, false, 2> : (const void*)k_bp_persist<false, 0, false, 2>);
  const void* kbp4 =
      plan.dr1 ? (ex ? (const void*)k_bp_persist<true, 2, true, 4> : (const void*)k_bp_persist<false, 2, true, 4>)
               : (ex ? (const void*)k_bp_persist<true, 0, false, 4> : (const void*)k_bp_persist<false, 0, false, 4>);
  const void* kbp = plan.bp_ppt == 4 ? kbp4 : kbp2;
  rc = plan.fp_mode == 5   ? persistent_grid(c, kfp, FPS_RAYS, fpg_lds(plan.fpg_cpl), &fp_grid)
       : plan.fp_mode == 4 ? persistent_grid(c, kfp, FPS_RAYS, sizeof(float) * FPS_CS * FPS_MAXW, &fp_grid)
                           : persistent_grid(c, kfp, FP_BLOCK, 0, &fp_grid);
  if (!rc) rc = persistent_grid(c, kbp, 256, lds, &bp_grid);
  if (rc) return rc;
  if (getenv("SYNKRO_VERBOSE"))
    fprintf(stderr, "sk_sirt_batched: %d groups, fp mode %d items %ld (grid %d), bp 16x%d tiles items %ld (grid %d), W %d\n",
            count, plan.fp_mode, args.fp_total, fp_grid, 16 * plan.bp_ppt, args.bp_total, bp_grid, plan.Wmax);

  // Everything the graph's kernel arguments bake in: a later call with the same key replays it.
  std::vector<unsigned long long> key;
  key.reserve(8 + 6 * (size_t)count);
  for (unsigned long long v : {(unsigned long long)count, (unsigned long long)iters, (unsigned long long)ex,
                               (unsigned long long)plan.fp_mode, (unsigned long long)plan.bp_ppt,
                               (unsigned long long)(uintptr_t)ws, (unsigned long long)fp_grid,
                               (unsigned long long)bp_grid})
    key.push_back(v);
  for (int i = 0; i < count; i++) {
    const sk_problem& P = problems[i];
    const Geom& G = P.geom->g;
    for (unsigned long long v : {G.uid, (unsigned long long)P.batch, (unsigned long long)(uintptr_t)P.d_img,
                                 (unsigned long long)(uintptr_t)P.d_sino, (unsigned long long)(uintptr_t)G.d_rowinv,
                                 (unsigned long long)(uintptr_t)G.d_colinv})
      key.push_back(v);
  }
  for (auto& bg : c->batch_graphs) {
    if (bg.exec && bg.key == key) {
      bg.used = ++c->batch_clock;
      SK_HIP(c, hipGraphLaunch(bg.exec, s));
      return SK_OK;
    }
  }

  // Capture the transposed copies, the counter reset and iters x (fp, bp) once; keep the graph.
  hipStream_t cs = c->capture_stream;
  SK_HIP(c, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  for (int oi = 0; oi < count && rc == SK_OK; oi++) {
    const BatchProb& B = args.p[oi];
    rc = launch_transpose(c, B.fp.img, const_cast<float*>(B.fp.imgT), B.fp.N, B.batch, cs);
  }
  if (rc == SK_OK) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, cs, plan.d_ctr);
    rc = launched(c, "counter reset");
  }
  for (int it = 0; it < iters && rc == SK_OK; it++) {
    if (plan.dr1)
      rc = ex ? launch_iteration<true, 2, true>(c, plan, fp_grid, bp_grid, cs)
              : launch_iteration<false, 2, true>(c, plan, fp_grid, bp_grid, cs);
    else
      rc = ex ? launch_iteration<true, 0, false>(c, plan, fp_grid, bp_grid, cs)
              : launch_iteration<false, 0, false>(c, plan, fp_grid, bp_grid, cs);
  }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(cs, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
  Ctx::BatchGraph* slot = nullptr;  // an empty slot, else the least recently used
  for (auto& bg : c->batch_graphs) {
    if (!bg.exec) {
      slot = &bg;
      break;
    }
    if (!slot || bg.used < slot->used) slot = &bg;
  }
  if (slot->exec) {
    // An evicted graph may still run on some stream: wait for the device before freeing it.
    hipError_t es = hipDeviceSynchronize();
    if (es != hipSuccess) {
      (void)hipGraphExecDestroy(exec);
      return hip_fail(c, es, "batched SIRT (graph eviction)");
    }
    (void)hipGraphExecDestroy(slot->exec);
  }
  slot->exec = exec;
  slot->key = std::move(key);
  slot->used = ++c->batch_clock;
  SK_HIP(c, hipGraphLaunch(exec, s));
  return SK_OK;
}

int sk_geom_set_variants(sk_geom* g, int fp_variant, int bp_variant) {
  if (!g) return SK_EINVAL;
  if (fp_variant < -1 || fp_variant > 7 || bp_variant < -1 || bp_variant > 2) return SK_EINVAL;
  g->g.fp_variant = fp_variant;
  g->g.bp_variant = bp_variant;
  return SK_OK;
}

int sk_autotune(sk_ctx* ctx, sk_geom* g, int* fp_variant, int* bp_variant, void* stream) {
  return sk_autotune_batch(ctx, g, 1, fp_variant, bp_variant, stream);
}

int sk_autotune_batch(sk_ctx* ctx, sk_geom* g, int batch, int* fp_variant, int* bp_variant, void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (!g) return fail(c, SK_EINVAL, "geometry is NULL");
  if (batch < 1 || batch > 1024) return fail(c, SK_EINVAL, "autotune batch must be 1..1024 (got %d)", batch);
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  Geom& G = g->g;
  // batch images and sinograms: the variants are timed on the calls' own batch size (small
  // images in a batch fill the chip where one does not, and favour other shapes)
  const size_t nimg = (size_t)batch * G.N * G.N, nsino = (size_t)batch * G.A * G.R;
  // own buffers (the projector calls below use the context scratch for x^T)
  float *img = nullptr, *sino = nullptr;
  SK_HIP(c, hipMalloc(&img, sizeof(float) * (nimg + nsino)));
  sino = img + nimg;
  hipError_t e = hipMemsetAsync(img, 0, sizeof(float) * (nimg + nsino), s);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int rc = (e == hipSuccess) ? SK_OK : hip_fail(c, e, "autotune setup");
  const int saved_fp = G.fp_variant, saved_bp = G.bp_variant;
  int best_fp = 0, best_bp = 0;
  auto time_it = [&](bool fp, int v, float* ms) -> int {
    if (fp) G.fp_variant = v; else G.bp_variant = v;
    const int reps = 3;
    for (int k = 0; k <= reps; k++) {
      if (k == 1) SK_HIP(c, hipEventRecord(e0, s));
      int r = fp ? sk_fp(ctx, g, batch, img, sino, stream) : sk_bp(ctx, g, batch, sino, img, stream);
      if (r) return r;
    }
    SK_HIP(c, hipEventRecord(e1, s));
    SK_HIP(c, hipEventSynchronize(e1));
    SK_HIP(c, hipEventElapsedTime(ms, e0, e1));
    return SK_OK;
  };
  const bool verbose = getenv("SYNKRO_VERBOSE") != nullptr;
  float best = 1e30f, ms = 0.0f;
  for (int v = 0; v <= 7 && !rc; v++) {
    rc = time_it(true, v, &ms);
    if (!rc && ms < best) { best = ms; best_fp = v; }
    if (!rc && verbose) fprintf(stderr, "sk_autotune: fp variant %d: %.3f ms / 3\n", v, ms);
  }
  best = 1e30f;
  for (int v = 0; v <= 2 && !rc; v++) {
    rc = time_it(false, v, &ms);
    if (!rc && ms < best) { best = ms; best_bp = v; }
    if (!rc && verbose) fprintf(stderr, "sk_autotune: bp variant %d: %.3f ms / 3\n", v, ms);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipFree(img);
  if (rc) {
    G.fp_variant = saved_fp;
    G.bp_variant = saved_bp;
    return rc;
  }
  G.fp_variant = best_fp;
  G.bp_variant = best_bp;
  if (fp_variant) *fp_variant = best_fp;
  if (bp_variant) *bp_variant = best_bp;
  return SK_OK;
}

}  // extern "C"
