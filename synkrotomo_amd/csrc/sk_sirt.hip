// sk_sirt.hip -- SIRT (futhark/SIRT.fut:9-49) on the fused projector kernels.
//
// Per iteration the reference runs a transpose, two fp kernels, four elementwise maps, two
// bp kernels, a transpose and two more maps (SIRT.fut:27-39).  Here an iteration is two
// kernels: k_fp in FP_SIRT mode writes q = R*(b - A x) directly (the maps fused into its
// epilogue), and k_bp in BP_SIRT mode writes x += C*B(q) together with the transposed
// copy of x that the next k_fp reads for flat lines.  A block of iterations is captured
// once into a hipGraph and replayed.
#include "sk_common.h"

namespace sk {

// Row and column sums (SIRT.fut:22-25): fp of an all-ones image and bp of an all-ones
// sinogram, with the safe inverse fused into the kernels' epilogues.
int ensure_weights(Ctx* c, Geom* g, hipStream_t s) {
  if (g->d_rowinv && g->d_colinv && g->weights_exact == c->exact) return SK_OK;
  if (!g->d_rowinv) SK_HIP(c, hipMalloc(&g->d_rowinv, sizeof(float) * (size_t)g->A * g->R));
  if (!g->d_colinv) SK_HIP(c, hipMalloc(&g->d_colinv, sizeof(float) * (size_t)g->N * g->N));
  FpArgs fa{nullptr, nullptr, g->d_rowinv, nullptr, nullptr, 1};
  int rc = launch_fp(c, g, FP_ONES_INV, fa, s);
  if (rc) return rc;
  BpArgs ba{nullptr, g->d_colinv, nullptr, nullptr, 1};
  rc = launch_bp(c, g, BP_ONES_INV, W_REF, ba, s);
  if (rc) return rc;
  g->weights_exact = c->exact;
  return SK_OK;
}

static int sirt_iteration(Ctx* c, const Geom* g, int batch, float* img, float* imgT,
                          const float* b, const float* rowinv, const float* colinv, float* q,
                          hipStream_t s) {
  FpArgs fa{img, imgT, q, b, rowinv, batch};
  int rc = launch_fp(c, g, FP_SIRT, fa, s);
  if (rc) return rc;
  BpArgs ba{q, img, imgT, colinv, batch};
  return launch_bp(c, g, BP_SIRT, W_REF, ba, s);
}


// Capture `k` iterations into an executable graph.
static int capture(Ctx* c, Geom* g, int k, int batch, float* img, float* imgT, const float* b,
                   float* q, hipGraphExec_t* out) {
  hipStream_t cs = c->capture_stream;
  SK_HIP(c, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  int rc = SK_OK;
  for (int it = 0; it < k && rc == SK_OK; it++)
    rc = sirt_iteration(c, g, batch, img, imgT, b, g->d_rowinv, g->d_colinv, q, cs);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(cs, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
  e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
  return SK_OK;
}

}  // namespace sk

using namespace sk;

extern "C" {

int sk_sirt_weights(sk_ctx* ctx, sk_geom* g, float* d_rowinv, float* d_colinv, void* stream) {
  if (!ctx || !g) return SK_EINVAL;
  Ctx* c = &ctx->c;
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_weights(c, &g->g, s);
  if (rc) return rc;
  const Geom& G = g->g;
  if (d_rowinv)
    SK_HIP(c, hipMemcpyAsync(d_rowinv, G.d_rowinv, sizeof(float) * (size_t)G.A * G.R,
                             hipMemcpyDeviceToDevice, s));
  if (d_colinv)
    SK_HIP(c, hipMemcpyAsync(d_colinv, G.d_colinv, sizeof(float) * (size_t)G.N * G.N,
                             hipMemcpyDeviceToDevice, s));
  return SK_OK;
}

int sk_sirt(sk_ctx* ctx, sk_geom* g, int batch, float* d_img, const float* d_sino, int iters,
            void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (!g) return fail(c, SK_EINVAL, "geometry is NULL");
  if (!d_img || !d_sino) return fail(c, SK_EINVAL, "device pointer is NULL");
  if (batch < 1 || batch > 65535) return fail(c, SK_EINVAL, "batch must be in [1, 65535]");
  if (iters < 0) return fail(c, SK_EINVAL, "iterations must be >= 0 (got %d)", iters);
  SK_HIP(c, hipSetDevice(c->device));
  if (iters == 0) return SK_OK;
  Geom& G = g->g;
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_weights(c, &G, s);
  if (rc) return rc;
  const size_t nimg = (size_t)batch * G.N * G.N, nsino = (size_t)batch * G.A * G.R;
  void* ws = nullptr;
  rc = scratch(c, sizeof(float) * (nimg + nsino), s, &ws);
  if (rc) return rc;
  float* imgT = (float*)ws;
  float* q = imgT + nimg;
  rc = launch_transpose(c, d_img, imgT, G.N, batch, s);
  if (rc) return rc;

  const int K = iters < 8 ? iters : 8;
  const int fpv = G.fp_variant >= 0 ? G.fp_variant : c->fp_variant;
  const int bpv = G.bp_variant >= 0 ? G.bp_variant : c->bp_variant;
  Geom::SirtGraph* slot = nullptr;
  for (auto& sg : G.sirt_graphs)
    if (sg.exec && sg.img == d_img && sg.b == d_sino && sg.ws == ws && sg.batch == batch &&
        sg.exact == c->exact && sg.iters == K && sg.fpv == fpv && sg.bpv == bpv)
      slot = &sg;
  if (!slot) {
    // an empty slot, else the least recently used one (its graph may still run, on any
    // stream: wait for the device)
    slot = &G.sirt_graphs[0];
    for (auto& sg : G.sirt_graphs)
      if (!sg.exec || (slot->exec && sg.used < slot->used)) slot = &sg;
    if (slot->exec) {
      SK_HIP(c, hipDeviceSynchronize());
      (void)hipGraphExecDestroy(slot->exec);
      *slot = Geom::SirtGraph{};
    }
    hipGraphExec_t ex = nullptr;
    rc = capture(c, &G, K, batch, d_img, imgT, d_sino, q, &ex);
    if (rc) return rc;
    slot->exec = ex;
    slot->img = d_img;
    slot->b = d_sino;
    slot->ws = ws;
    slot->batch = batch;
    slot->exact = c->exact;
    slot->iters = K;
    slot->fpv = fpv;
    slot->bpv = bpv;
  }
  slot->used = ++G.sirt_clock;
  for (int done = 0; done + K <= iters; done += K) SK_HIP(c, hipGraphLaunch(slot->exec, s));
  for (int it = 0; it < iters % K; it++) {
    rc = sirt_iteration(c, &G, batch, d_img, imgT, d_sino, G.d_rowinv, G.d_colinv, q, s);
    if (rc) return rc;
  }
  return SK_OK;
}

int sk_sirt_step(sk_ctx* ctx, sk_geom* g, int batch, float* d_img, const float* d_sino,
                 const float* d_rowinv, const float* d_colinv, float* d_q, void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (!g || !d_img || !d_sino || !d_rowinv || !d_colinv || !d_q)
    return fail(c, SK_EINVAL, "NULL argument");
  if (batch < 1 || batch > 65535) return fail(c, SK_EINVAL, "batch must be in [1, 65535]");
  SK_HIP(c, hipSetDevice(c->device));
  const Geom& G = g->g;
  hipStream_t s = (hipStream_t)stream;
  void* ws = nullptr;
  int rc = scratch(c, sizeof(float) * (size_t)batch * G.N * G.N, s, &ws);
  if (rc) return rc;
  float* imgT = (float*)ws;
  rc = launch_transpose(c, d_img, imgT, G.N, batch, s);
  if (rc) return rc;
  return sirt_iteration(c, &G, batch, d_img, imgT, d_sino, d_rowinv, d_colinv, d_q, s);
}

}  // extern "C"
