// sk_noalloc.hip -- sk_fp_noalloc: the forward projector A.x with no device allocation.
//
// sk_fp stages the transposed image its flat lines read (forwardprojection.fut:80-82) in the
// context's growable scratch, which is (re)allocated with hipMalloc on first use and whenever
// a larger batch arrives.  Callers that may not allocate (code under hipGraph capture, tight
// serving loops, harnesses that forbid hipMalloc) use sk_fp_noalloc instead: the same kernels,
// with the transposed copy in `scratch`, a fixed device array that lives as long as the
// process.  Its device address is looked up once per device (sk_create, with the device
// current) and passed to the kernels; the array's name is never used as a device address.
//
// One array per GPU, shared by every call on it: calls that overlap on one GPU (two streams,
// two host threads) race on it.  Each call overwrites the part it uses before reading it, so
// no clearing is needed.
#include <mutex>

#include "sk_common.h"

// Maximum floats of transposed image per call: batch * N * N <= 16 Mi (64 MiB), e.g. four
// 2048^2 images or one 4096^2 image.
#define SK_NOALLOC_FLOATS (16u << 20)
#define SK_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) float scratch[SK_NOALLOC_FLOATS];

namespace {
std::mutex g_mu;
float* g_scratch_ptr[SK_NOALLOC_MAX_DEVICES] = {};  // device address of `scratch`, per device
}  // namespace

namespace sk {

// With `device` current: look up this device's address of `scratch` (once per device).
int noalloc_init(Ctx* c, int device) {
  if (device < 0 || device >= SK_NOALLOC_MAX_DEVICES)
    return fail(c, SK_EINVAL, "device %d beyond the %d devices the no-alloc scratch supports", device,
                SK_NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch_ptr[device]) return SK_OK;
  void* p = nullptr;
  // ::scratch, the array (inside namespace sk the plain name is the context-scratch function)
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
  if (e != hipSuccess || !p) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = (float*)p;
  return SK_OK;
}

}  // namespace sk

extern "C" {

int sk_fp_noalloc(sk_ctx* ctx, sk_geom* g, int batch, const float* d_img, float* d_sino, void* stream) {
  if (!ctx) return SK_EINVAL;
  sk::Ctx* c = &ctx->c;
  if (!g) return sk::fail(c, SK_EINVAL, "geometry is NULL");
  if (batch < 1 || batch > 65535) return sk::fail(c, SK_EINVAL, "batch must be in [1, 65535]");
  if (!d_img || !d_sino) return sk::fail(c, SK_EINVAL, "device pointer is NULL");
  const sk::Geom& G = g->g;
  if (G.device != c->device) return sk::fail(c, SK_EINVAL, "geometry belongs to another device");
  const unsigned long long need = (unsigned long long)batch * G.N * G.N;
  if (G.n_steep < G.A && need > SK_NOALLOC_FLOATS)
    return sk::fail(c, SK_EINVAL,
                    "sk_fp_noalloc: batch * N * N = %llu floats exceeds its fixed scratch of %u floats "
                    "(use sk_fp)", need, SK_NOALLOC_FLOATS);
  float* ws = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    ws = (c->device >= 0 && c->device < SK_NOALLOC_MAX_DEVICES) ? g_scratch_ptr[c->device] : nullptr;
  }
  if (!ws) return sk::fail(c, SK_EHIP, "sk_fp_noalloc: scratch address of device %d not set up", c->device);
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const float* imgT = d_img;
  if (G.n_steep < G.A) {
    int rc = sk::launch_transpose(c, d_img, ws, G.N, batch, s);
    if (rc) return rc;
    imgT = ws;
  }
  sk::FpArgs a{d_img, imgT, d_sino, nullptr, nullptr, batch};
  return sk::launch_fp(c, &G, sk::FP_PLAIN, a, s);
}

}  // extern "C"
