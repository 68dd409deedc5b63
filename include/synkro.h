/*
 * synkro.h -- C ABI of libsynkro.so, the MI355X (gfx950) projector / SIRT library.
 *
 * This is the drop-in boundary for tomograph/synkrotomo's projector and SIRT path.  The
 * reference exposes each Futhark entry point as a generated PyOpenCL class with one
 * `main` method (Makefile:1-4, README.md:21-25); this library exposes the same
 * operations as plain C functions over device pointers, and synkrotomo_amd/futhark/
 * rebuilds the reference's module/class/`main` surface on top of it with ctypes.
 *
 * Conventions
 *   - Device pointers (d_*) are caller-owned (normally torch tensors); the library never
 *     frees them.  Host pointers (*_host) are read during the call only.
 *   - Image:    float32 [N][N], row-major (row = y + N/2, column = x + N/2).
 *   - Sinogram: float32 [A][R], angle-major, bin r <-> rho = rho0 + r*drho.
 *   - Every launch goes on the caller's stream (hipStream_t passed as void*; NULL = the
 *     null stream).  Calls return when the work is enqueued, except where noted.
 *   - Return value: SK_OK, or an error code; sk_last_error() gives the message.
 *   - One sk_ctx per host thread (or external locking): a context owns scratch memory
 *     that its calls share.
 */
#ifndef SYNKRO_H
#define SYNKRO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sk_ctx sk_ctx;
typedef struct sk_geom sk_geom;
typedef struct sk_mat sk_mat;

enum {
  SK_OK = 0,
  SK_EINVAL = 1, /* bad argument value or shape (Python: ValueError) */
  SK_ETYPE = 2,  /* bad argument kind (Python: TypeError) */
  SK_EHIP = 3,   /* HIP runtime error (Python: RuntimeError) */
  SK_ENOMEM = 4  /* device allocation failed (Python: MemoryError) */
};

/* Options for sk_set_option. */
enum {
  /* 1: every division in the inner loops is an IEEE division, so results are bitwise
   *    those of the sequential FP32 restatement of the reference (slower).
   * 0 (default): the per-step quotients use the hardware reciprocal (<= 2 ulp per
   *    length); pixel indices and candidate sets are unchanged. */
  SK_OPT_EXACT = 1,
  /* Tile / rays-per-thread variant for the projector kernels (0 = autotuned default). */
  SK_OPT_FP_VARIANT = 2,
  SK_OPT_BP_VARIANT = 3
};

/* Operators that can be made explicit by sk_sparse_build. */
enum {
  SK_MAT_FP = 0, /* A: rows = rays (angle-major), columns = pixels */
  SK_MAT_BP = 1, /* B: rows = pixels, columns = sinogram entries (the backprojector) */
  SK_MAT_ADJ = 2 /* A^T (the CSC form of A): rows = pixels, columns = rays */
};

/* ---------------------------------------------------------------- context */
/* No reference counterpart (replaces the generated class constructors, which create an
 * OpenCL context and queue; Futhark convention, absent from the reference tree). */
int sk_create(int device, sk_ctx** out);
void sk_destroy(sk_ctx* ctx);
const char* sk_last_error(const sk_ctx* ctx);
int sk_set_option(sk_ctx* ctx, int option, int value);
int sk_get_option(const sk_ctx* ctx, int option, int* value);

/* ---------------------------------------------------------------- geometry */
/* Replaces sirtlib.preprocess (futhark/sirtlib.fut:28-44), which every `main` runs on the
 * device: per-angle cos/sin, line lengths, steep/flat split.  Built once on the host and
 * uploaded.  Rejects A < 1, R < 1, odd or non-positive N, drho <= 0. */
int sk_geom_create(sk_ctx* ctx, const float* angles_host, int A, float rho0, float drho, int R,
                   int N, sk_geom** out);
void sk_geom_destroy(sk_geom* g);
/* n_steep = number of angles with |sin| < |cos|; ncand = ceil(sqrt(2)/drho). */
int sk_geom_query(const sk_geom* g, int* A, int* R, int* N, int* n_steep, int* ncand);

/* Host only (no device needed): the line table sk_geom_create builds, 8 floats per line in
 * steep-then-flat order: cost, sint, lbase, slope, fact*cost, fact*sint, orig, flags.
 * Returns the number of steep lines (sirtlib.fut:28-44), or -1 on bad arguments. */
int sk_lines_compute(const float* angles_host, int A, float* table_out);

/* ---------------------------------------------------------------- dense projector */
/* A.x: forwardprojection.main (futhark/forwardprojection.fut:71-84).
 * d_img [batch][N][N] -> d_sino [batch][A][R]. */
int sk_fp(sk_ctx* ctx, sk_geom* g, int batch, const float* d_img, float* d_sino, void* stream);

/* A.x with no device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same kernels and results as sk_fp; the transposed image that the flat
 * lines read goes to a fixed 64 MiB device array of the library instead of the context's
 * growable scratch, so batch * N * N must be at most 16,777,216 floats when any angle is flat
 * (SK_EINVAL otherwise).  One array per GPU, shared by every call on it: calls that overlap on
 * one GPU race. */
int sk_fp_noalloc(sk_ctx* ctx, sk_geom* g, int batch, const float* d_img, float* d_sino, void* stream);

/* B.r: backprojection.main (futhark/backprojection.fut:82-93).
 * d_sino [batch][A][R] -> d_img [batch][N][N]. */
int sk_bp(sk_ctx* ctx, sk_geom* g, int batch, const float* d_sino, float* d_img, void* stream);

/* A^T.r, the exact adjoint of sk_fp (same per-entry lengths as A).  No reference
 * counterpart: the reference's backprojection equals A^T except on lines whose slope
 * rounds to zero (theta = 0, pi/2), where it drops rays lying on pixel edges.  Used as
 * the backward of the differentiable projector. */
int sk_bp_adjoint(sk_ctx* ctx, sk_geom* g, int batch, const float* d_sino, float* d_img,
                  void* stream);

/* B^T.v, the exact adjoint of sk_bp: the ray walk of sk_fp with the backprojector's
 * lengths.  d_img [batch][N][N] -> d_sino [batch][A][R].  No reference counterpart; used by
 * the backward of the differentiable SIRT step. */
int sk_bp_transpose(sk_ctx* ctx, sk_geom* g, int batch, const float* d_img, float* d_sino,
                    void* stream);

/* ---------------------------------------------------------------- SIRT */
/* SIRT.main (futhark/SIRT.fut:9-49): x <- x + C*B(R*(b - A x)) `iters` times, in place
 * on d_img [batch][N][N], with R = 1/rowsum(A), C = 1/colsum(B) (0 where the sum is 0).
 * The iteration (two fused kernels) is captured once into a hipGraph and replayed.
 * batch > 1 runs independent slices of one geometry (SIRT3D.main,
 * futhark/SIRT3D.fut:18-28). */
int sk_sirt(sk_ctx* ctx, sk_geom* g, int batch, float* d_img, const float* d_sino, int iters,
            void* stream);

/* The SIRT weights: d_rowinv [A][R] = 1/rowsum, d_colinv [N][N] = 1/colsum (SIRT.fut:22-25). */
int sk_sirt_weights(sk_ctx* ctx, sk_geom* g, float* d_rowinv, float* d_colinv, void* stream);

/* One SIRT iteration with caller-provided weights (no graph, no state):
 * d_q [batch][A][R] is scratch.  Used by the learned/unrolled paths. */
int sk_sirt_step(sk_ctx* ctx, sk_geom* g, int batch, float* d_img, const float* d_sino,
                 const float* d_rowinv, const float* d_colinv, float* d_q, void* stream);

/* ---------------------------------------------------------------- batched SIRT */
/* A group of `batch` independent problems sharing one geometry: d_img [batch][N][N]
 * (initial images, updated in place), d_sino [batch][A][R]. */
typedef struct {
  sk_geom* geom;
  int batch;
  float* d_img;
  const float* d_sino;
} sk_problem;

/* No reference counterpart (generalizes SIRT3D.main, futhark/SIRT3D.fut:18-28, which maps
 * SIRT over slices of one geometry): advance every problem of every group `iters` SIRT
 * iterations.  Each iteration is two persistent launches -- one work queue over all
 * problems' forward-projection items, one over all their backprojection tiles -- and the
 * whole sequence runs as one hipGraph.  The executable graph is kept (the context holds the
 * four most recently used), keyed by the problems' geometries, batches and device pointers and
 * the iteration count, so a repeated call replays it without capturing again.  Returns once the
 * work is enqueued on `stream`. */
int sk_sirt_batched(sk_ctx* ctx, const sk_problem* problems, int count, int iters, void* stream);

/* No reference counterpart: time the kernel variants on this geometry and keep the fastest
 * for it.  fp: 0/1/2 = a wave per 64/128/256 bins (1, 2 or 4 rays per lane), bin blocks
 * fastest in the grid; 3/4 = 1 or 2 rays per lane, lines fastest; 5/6 = 256 bins per
 * workgroup with each block of 12 image rows staged in LDS (double / single buffered; fast
 * mode, deltarho up to about 1, otherwise they run variant 3); 7 = 64 bins of four lines
 * adjacent in slope per workgroup, staged the same way (fast mode, geometries whose line
 * groups fit four chunks of 64 columns, otherwise variant 6).  bp: 16x16, 16x32, 16x64 pixel
 * tiles.  Variants compute identical results.  Synchronizes `stream`. */
int sk_autotune(sk_ctx* ctx, sk_geom* g, int* fp_variant, int* bp_variant, void* stream);
/* No reference counterpart: sk_autotune timed on calls of `batch` images (1..1024) -- the
 * batch the caller will run, since batches of small images favour other shapes than one
 * image (sk_autotune is batch 1).  Keeps the fastest pair for the geometry. */
int sk_autotune_batch(sk_ctx* ctx, sk_geom* g, int batch, int* fp_variant, int* bp_variant, void* stream);
/* Apply previously tuned variants (-1 = the context default). */
int sk_geom_set_variants(sk_geom* g, int fp_variant, int bp_variant);

/* ---------------------------------------------------------------- learned correction */
/* No reference counterpart.  Kernels of the correction network interleaved with SIRT:
 * y = x + conv4(relu(conv3(relu(conv2(relu(conv1(x))))))), 3x3 convolutions, stride 1, zero
 * padding 1.  Images: FP32 [B][H][W].  Activations: BF16 NHWC [B][H][W][32].  Weights: BF16
 * [Cout][3][3][Cin].  Biases and all accumulation: FP32. */
/* 32 -> 32 on MFMA.  bias != NULL: y = relu(conv(x) + bias).  bias == NULL: y = conv(x) (the
 * data gradient, with flipped/transposed weights); mask != NULL zeroes x where mask <= 0. */
int sk_conv32(sk_ctx* ctx, int B, int H, int W, const void* x, const void* mask, const void* w,
              const float* bias, void* y, void* stream);
/* 1 -> 32: h = relu(conv(x) + bias), w [32][3][3]. */
int sk_conv_in(sk_ctx* ctx, int B, int H, int W, const float* x, const void* w, const float* bias,
               void* h, void* stream);
/* 32 -> 1 with residual: y = res + conv(h) + bias[0], w [3][3][32]. */
int sk_conv_out(sk_ctx* ctx, int B, int H, int W, const void* h, const void* w, const float* bias,
                const float* res, float* y, void* stream);
/* 2 -> 32: h = relu(conv(x2) + bias) on two FP32 input channels x2 [B][H][W][2] (the image and
 * its SIRT update direction), w [32][3][3][2]. */
int sk_conv_in2(sk_ctx* ctx, int B, int H, int W, const float* x2, const void* w, const float* bias,
                void* h, void* stream);
/* Data gradients of sk_conv_out (dh from dy) and of sk_conv_in (dx from dh masked by h > 0). */
int sk_conv_out_dgrad(sk_ctx* ctx, int B, int H, int W, const float* g, const void* w, void* dh,
                      void* stream);
int sk_conv_in_dgrad(sk_ctx* ctx, int B, int H, int W, const void* dh, const void* hmask,
                     const void* w, float* dx, void* stream);
/* Data gradient of sk_conv_in2: dx2 [B][H][W][2]. */
int sk_conv_in2_dgrad(sk_ctx* ctx, int B, int H, int W, const void* dh, const void* hmask,
                      const void* w, float* dx2, void* stream);
/* Weight and bias gradients (FP32, deterministic): kind 0 = conv1, 1 = conv2/conv3 (on MFMA),
 * 2 = conv4, 3 = conv1 with two input channels (x = [B][H][W][2] FP32).  out = Cout*9*Cin weight
 * gradients then Cout bias gradients. */
int sk_conv_wgrad(sk_ctx* ctx, int kind, int B, int H, int W, const void* g, const void* mask,
                  const void* x, float* out, void* stream);
/* Random-ellipse phantoms: params [B][E][7] = (cx, cy, a, b, cos t, sin t, value) in FP64 on
 * [-1,1]^2 pixel centres, sum clipped to [0,1] -> d_out [B][N][N] FP32. */
int sk_phantoms(sk_ctx* ctx, int B, int N, int E, const double* d_params, float* d_out, void* stream);

/* ---------------------------------------------------------------- comparison tools */
/* sanityCheck.main, diff.main and mse.main (futhark/tools/sanityCheck.fut:1-4, diff.fut:1-7,
 * mse.fut:1-3) in one pass over two device arrays of n floats, with the reference's tolerance
 * (0.05) as the parameter `tol`:
 *   *all_within = every |sane - test| < tol (a NaN difference fails);
 *   *n_over     = number of indices i > 0 with |sane - test| > tol (diff.fut's filter (> 0)
 *                 also drops index 0);
 *   *mse        = sum of the f32 squares (sane - test)**2 / n, summed in FP64.
 * Any output pointer may be NULL.  Synchronizes `stream`. */
int sk_compare(sk_ctx* ctx, int64_t n, const float* d_sane, const float* d_test, float tol,
               int* all_within, int64_t* n_over, float* mse, void* stream);

/* ---------------------------------------------------------------- sparse path */
/* forwardprojection_sparse.main / backprojection_sparse.main
 * (futhark/forwardprojection_sparse.fut:26-40, futhark/backprojection_sparse.fut:26-37):
 * same signatures as the dense mains, run on an explicit CSR matrix built on the GPU.
 *   SK_MAT_FP  A: rows = rays (angle*R + bin), cols = pixels; entries in row-step order with
 *              the forward projector's lengths (bitwise equal to the FP32 restatement).
 *   SK_MAT_BP  B: rows = pixels, cols = sinogram entries; the backprojector's non-zero
 *              intersect_fact lengths, lines in steep-then-flat order.
 *   SK_MAT_ADJ A^T, i.e. A in CSC form: rows = pixels, cols = rays; exactly A's entries of
 *              each pixel (explicit zeros included), values bit for bit A's.  No reference
 *              counterpart (the reference's backprojection_sparse is B); A^T.r is the exact
 *              adjoint of the sparse forward projection.
 * 64-bit row pointers, 32-bit columns, f32 values, all owned by the matrix.  The call
 * synchronizes `stream` once (to size the column/value arrays). */
int sk_sparse_build(sk_ctx* ctx, sk_geom* g, int which, sk_mat** out, void* stream);
/* nnz of an SK_MAT_* matrix without building it (synchronizes `stream`). */
int sk_sparse_count(sk_ctx* ctx, sk_geom* g, int which, int64_t* nnz_out, void* stream);
int sk_mat_info(const sk_mat* m, int64_t* rows, int64_t* cols, int64_t* nnz);
/* Device pointers of the CSR arrays (valid until sk_mat_destroy). */
int sk_mat_arrays(const sk_mat* m, const int64_t** d_rowptr, const int32_t** d_cols,
                  const float** d_vals);
/* Copy the CSR arrays into caller-owned device buffers (rows+1, nnz, nnz entries). */
int sk_mat_copy(sk_ctx* ctx, const sk_mat* m, int64_t* d_rowptr, int32_t* d_cols, float* d_vals,
                void* stream);
/* SpMV kernel: 0 = one wave per row, 1 = one lane per row (64 adjacent rows in lockstep),
 * 2 = one wave per row with 16-byte column/value streams; SK_MAT_FP only: 3, 4 = as 0, 2 with
 * the gathers of steep rays (whose pixels step through image rows) from a transposed copy of
 * the input, where they are contiguous; 7 = one lane per row on a SELL-64 copy of the matrix
 * (64-row slices stored column-major, both streams coalesced; built on first use, about the
 * size of the CSR arrays); SK_MAT_FP only: 8 = as 7 with the flat rays' gathers from the
 * transposed input; 9 = as 7 with the columns stored as 16-bit steps from the previous entry
 * of the row (6 bytes per entry instead of 8; for SK_MAT_BP / SK_MAT_ADJ the steps run over the
 * rays in line-table order, steep lines then flat lines, and each call first permutes the
 * input's rows into that order; unavailable, and sk_spmv fails with EINVAL, when two
 * consecutive columns of a row are more than 32767 apart); 10 = row slots, one lane per row on a
 * copy that keeps the 64 rows of a slice in step: for SK_MAT_FP a slice is 64 bins of one angle
 * and record k holds each ray's two entries in frame row k (gathered from the image, or its
 * transpose for flat angles: one frame row per wave), for SK_MAT_BP a slice is 64 pixels of an
 * image row and record j holds each pixel's two candidates of the line at table position j
 * (deltarho in [0.71, 1.41) only); holes where an entry does not exist; built on first use (not
 * for SK_MAT_ADJ: EINVAL).  Kernels that sum a row in the same order give the same bits: 0, 3;
 * 2, 4; 1, 7, 8, 9, 10.  -1 (default) = time the candidates
 * (0, 2; 3, 4 for SK_MAT_FP; 7, 8 for SK_MAT_FP and 9, if their copies fit in half the free
 * device memory) on the first sk_spmv call and keep the fastest.  EINVAL for 5, 6, and for 3,
 * 4, 8 on other matrices. */
int sk_mat_set_spmv_variant(sk_mat* m, int variant);
/* The SpMV kernel in use (-1 before the first sk_spmv call of an untimed matrix). */
int sk_mat_spmv_variant(const sk_mat* m);
/* d_out[batch][rows] = M . d_in[batch][cols] */
int sk_spmv(sk_ctx* ctx, const sk_mat* m, int batch, const float* d_in, float* d_out, void* stream);
void sk_mat_destroy(sk_mat* m);

/* ---------------------------------------------------------------- measured data */
/* No reference counterpart: the reference reconstructs from sinograms that are already line
 * integrals (its SIRT benchmark reads data/tooth.in, futhark/SIRT.fut:1-2).
 * Flat/dark-field normalisation and -log of raw detector counts [batch][A][R]:
 *   dark[r] = mean of the n_dark dark frames d_dark[n_dark][R] (0 if n_dark == 0),
 *   flat[r] = mean of the n_flat open-beam frames d_flat[n_flat][R] (both summed in FP64),
 *   T = (raw - dark[r]) / (flat[r] - dark[r]),
 *   out = -log(max(T, min_transmission)) / scale.
 * A bin whose flat mean does not exceed its dark mean reads T = 0 (fully absorbing); a NaN
 * count reads T = min_transmission. 
 * d_out may equal d_raw. */
int sk_normalize(sk_ctx* ctx, int batch, int A, int R, const float* d_raw, int n_flat,
                 const float* d_flat, int n_dark, const float* d_dark, float min_transmission,
                 float scale, float* d_out, void* stream);
/* Centre of mass of each of `rows` sinogram rows of R bins: d_com[i] = sum_r r*v / sum_r v in
 * FP64 (NaN where sum_r v == 0).  The rotation axis sits at the bin c0 of the fit
 * com(theta) = c0 + p*cos(theta) + q*sin(theta); rhozero = -c0*deltarho centres a geometry. */
int sk_row_centroid(sk_ctx* ctx, int64_t rows, int R, const float* d_sino, double* d_com, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SYNKRO_H */
