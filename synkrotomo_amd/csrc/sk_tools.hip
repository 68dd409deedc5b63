// sk_tools.hip -- the reference's comparison tools (futhark/tools/*.fut) in one pass.
//
//   sanityCheck.main (tools/sanityCheck.fut:1-4): every |sane - test| < 0.05
//   diff.main        (tools/diff.fut:1-7):        the number of indices i with
//                                                 |sane - test| > 0.05, after the
//                                                 reference's filter (> 0), which also
//                                                 drops index 0
//   mse.main         (tools/mse.fut:1-3):         sum((sane - test)**2) / n
//
// One read of both arrays: each block folds a grid-stride slice into (entries failing the
// sanity bound, entries counted by diff, sum of f32 squares in FP64); a second one-block
// kernel folds the block partials in block order, so the result is deterministic.  The
// tolerance is a parameter (the reference hard-codes 0.05).
#include "sk_common.h"

namespace sk {

struct ComparePartial {
  unsigned long long n_fail;  // entries with !(|d| < tol) -- NaN included
  unsigned long long n_over;  // entries with |d| > tol and index > 0
  double sumsq;               // sum of the f32 squares d*d
};

constexpr int CMP_THREADS = 256;
constexpr int CMP_MAX_BLOCKS = 1024;

__global__ __launch_bounds__(CMP_THREADS) void k_compare_partial(long n, const float* __restrict__ sane,
                                                                 const float* __restrict__ test, float tol,
                                                                 ComparePartial* __restrict__ part) {
  unsigned long long nf = 0, no = 0;
  double ss = 0.0;
  for (long i = (long)blockIdx.x * CMP_THREADS + threadIdx.x; i < n; i += (long)gridDim.x * CMP_THREADS) {
    const float d = sane[i] - test[i];
    const float ad = fabsf(d);
    nf += (ad < tol) ? 0 : 1;
    no += (ad > tol && i > 0) ? 1 : 0;
    ss += (double)(d * d);  // mse.fut:2, (s-t)**2 in f32
  }
  __shared__ unsigned long long snf[CMP_THREADS], sno[CMP_THREADS];
  __shared__ double sss[CMP_THREADS];
  snf[threadIdx.x] = nf;
  sno[threadIdx.x] = no;
  sss[threadIdx.x] = ss;
  __syncthreads();
  for (int off = CMP_THREADS / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      snf[threadIdx.x] += snf[threadIdx.x + off];
      sno[threadIdx.x] += sno[threadIdx.x + off];
      sss[threadIdx.x] += sss[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = ComparePartial{snf[0], sno[0], sss[0]};
}

__global__ __launch_bounds__(CMP_THREADS) void k_compare_final(int nparts, const ComparePartial* __restrict__ part,
                                                               ComparePartial* __restrict__ out) {
  __shared__ unsigned long long snf[CMP_THREADS], sno[CMP_THREADS];
  __shared__ double sss[CMP_THREADS];
  unsigned long long nf = 0, no = 0;
  double ss = 0.0;
  for (int k = threadIdx.x; k < nparts; k += CMP_THREADS) {
    nf += part[k].n_fail;
    no += part[k].n_over;
    ss += part[k].sumsq;
  }
  snf[threadIdx.x] = nf;
  sno[threadIdx.x] = no;
  sss[threadIdx.x] = ss;
  __syncthreads();
  for (int off = CMP_THREADS / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      snf[threadIdx.x] += snf[threadIdx.x + off];
      sno[threadIdx.x] += sno[threadIdx.x + off];
      sss[threadIdx.x] += sss[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = ComparePartial{snf[0], sno[0], sss[0]};
}

}  // namespace sk

using namespace sk;

extern "C" int sk_compare(sk_ctx* ctx, int64_t n, const float* d_sane, const float* d_test, float tol,
                          int* all_within, int64_t* n_over, float* mse, void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (n < 1) return fail(c, SK_EINVAL, "n must be >= 1 (got %lld)", (long long)n);
  if (!d_sane || !d_test) return fail(c, SK_EINVAL, "device pointer is NULL");
  if (!(tol >= 0.0f)) return fail(c, SK_EINVAL, "tolerance must be >= 0");
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const long want = (n + CMP_THREADS * 8 - 1) / (CMP_THREADS * 8);
  const int nb = (int)(want < CMP_MAX_BLOCKS ? want : CMP_MAX_BLOCKS);
  void* ws = nullptr;
  int rc = scratch(c, sizeof(ComparePartial) * (size_t)(nb + 1), s, &ws);
  if (rc) return rc;
  ComparePartial* part = (ComparePartial*)ws;
  hipLaunchKernelGGL(k_compare_partial, dim3(nb), dim3(CMP_THREADS), 0, s, (long)n, d_sane, d_test, tol, part);
  hipLaunchKernelGGL(k_compare_final, dim3(1), dim3(CMP_THREADS), 0, s, nb, part, part + nb);
  rc = launched(c, "compare kernels");
  if (rc) return rc;
  ComparePartial h{};
  SK_HIP(c, hipMemcpyAsync(&h, part + nb, sizeof(h), hipMemcpyDeviceToHost, s));
  SK_HIP(c, hipStreamSynchronize(s));
  if (all_within) *all_within = h.n_fail == 0 ? 1 : 0;
  if (n_over) *n_over = (int64_t)h.n_over;
  if (mse) *mse = (float)(h.sumsq / (double)n);
  return SK_OK;
}
