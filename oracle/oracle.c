/*
 * oracle.c -- CPU restatement of tomograph/synkrotomo's projector and SIRT path.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load liboracle.so as the checker.  The product (synkrotomo_amd) never links it.
 *
 * The reference is Futhark (the futhark/ .fut files); no Futhark compiler exists in this image, so
 * the reference itself cannot be built.  This file restates its arithmetic in C with the
 * semantics of Futhark's sequential C backend (left folds from 0, no FMA contraction,
 * libm cosf/sinf/powf/sqrtf).  Built with -ffp-contract=off.
 *
 * Exported entry points mirror the reference's `main`s:
 *   orc_fp      forwardprojection.fut:71-84
 *   orc_bp      backprojection.fut:82-93
 *   orc_sirt    SIRT.fut:9-49
 *   orc_sirt3d  SIRT3D.fut:18-28 with its argument defect fixed (rho0 = rhos[0],
 *               drho = rhos[1]-rhos[0], numrhos = len(rhos), as
 *               archive/projection_lib.fut:114-115 derives them)
 * plus FP64 twins (_f64), the explicit sparse structures of both operators, the exact
 * adjoint of the forward projector, and per-ray / per-pixel helpers for known-answer tests.
 *
 * Parity pins (see tests/): closed-form chords, the archive's known-answer tests
 * (archive/line_lib.fut:1-57, archive/projection_lib.fut:1-41) after the +sin -> -sin
 * remap, brute-force exact-geometry FP64 matrices on small grids, the adjoint identity,
 * SIRT invariants.  Futhark-OpenCL's reduction tree and FMA use are parity unpinned.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SUFFIX _f32
#define RCOS cosf
#define RSIN sinf
#define RSQRT sqrtf
#define RPOW powf
#define RFLOOR floorf
#define RCEIL ceilf
#define RFABS fabsf
#define RMAX fmaxf
#define RMIN fminf
#include "oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef RCOS
#undef RSIN
#undef RSQRT
#undef RPOW
#undef RFLOOR
#undef RCEIL
#undef RFABS
#undef RMAX
#undef RMIN

#define REAL double
#define SUFFIX _f64
#define RCOS cos
#define RSIN sin
#define RSQRT sqrt
#define RPOW pow
#define RFLOOR floor
#define RCEIL ceil
#define RFABS fabs
#define RMAX fmax
#define RMIN fmin
#include "oracle_impl.h"

/* ------------------------------------------------------------------ geometry */

int orc_preprocess(const float* angles, int A, float* cost, float* sint, float* lbase,
                   int* orig) {
  line_t_f32* L = (line_t_f32*)malloc(sizeof(line_t_f32) * (size_t)A);
  int ns = preprocess_f32(angles, A, L);
  for (int j = 0; j < A; j++) {
    cost[j] = L[j].cost;
    sint[j] = L[j].sint;
    lbase[j] = L[j].lbase;
    orig[j] = L[j].orig;
  }
  free(L);
  return ns;
}

/* ------------------------------------------------------------------ FP32 mains */

/* forwardprojection.fut:71-84 */
void orc_fp(const float* angles, int A, float rho0, float drho, int R, const float* img, int n,
            float* out, int nt) {
  int N = (int)sqrtf((float)n);
  line_t_f32* L = (line_t_f32*)malloc(sizeof(line_t_f32) * (size_t)A);
  int ns = preprocess_f32(angles, A, L);
  forward_f32(L, A, ns, rho0, drho, R, N, img, out, nt);
  free(L);
}

/* backprojection.fut:82-93 */
void orc_bp(const float* angles, int A, float rho0, float drho, int size, const float* proj,
            int p, float* out, int nt) {
  int ncand = (int)ceilf(sqrtf(2.0f) / drho);
  int R = p / A;
  line_t_f32* L = (line_t_f32*)malloc(sizeof(line_t_f32) * (size_t)A);
  int ns = preprocess_f32(angles, A, L);
  float* split = (float*)malloc(sizeof(float) * (size_t)A * R);
  split_sino_f32(L, A, R, proj, split);
  backward_f32(L, A, ns, rho0, drho, ncand, R, size, split, out, nt);
  free(split);
  free(L);
}

/* SIRT.fut:43-49 */
void orc_sirt(const float* angles, int A, float rho0, float drho, float* img, int n,
              const float* proj, int p, int iters, int nt) {
  sirt_f32(angles, A, rho0, drho, img, n, proj, p, iters, nt);
}

/* SIRT3D.fut:18-28 with the rhos -> (rho0, drho) fix; out is [size][size*size] */
void orc_sirt3d(const float* angles, int A, const float* rhos, int r, const float* proj,
                int iters, int size, float* out, int nt) {
  float rho0 = rhos[0];
  float drho = rhos[1] - rhos[0];
  long n = (long)size * size;
  for (int i = 0; i < size; i++) {
    float* img = out + (long)i * n;
    memset(img, 0, sizeof(float) * (size_t)n);
    sirt_f32(angles, A, rho0, drho, img, (int)n, proj + (long)i * r * A, r * A, iters, nt);
  }
}

/* ------------------------------------------------------------------ FP64 twins */

void orc_fp_f64(const float* angles, int A, double rho0, double drho, int R, const double* img,
                int n, double* out, int nt) {
  int N = (int)sqrt((double)n);
  line_t_f64* L = (line_t_f64*)malloc(sizeof(line_t_f64) * (size_t)A);
  int ns = preprocess_f64(angles, A, L);
  forward_f64(L, A, ns, rho0, drho, R, N, img, out, nt);
  free(L);
}

void orc_bp_f64(const float* angles, int A, double rho0, double drho, int size,
                const double* proj, int p, double* out, int nt) {
  int ncand = (int)ceil(sqrt(2.0) / drho);
  int R = p / A;
  line_t_f64* L = (line_t_f64*)malloc(sizeof(line_t_f64) * (size_t)A);
  int ns = preprocess_f64(angles, A, L);
  double* split = (double*)malloc(sizeof(double) * (size_t)A * R);
  split_sino_f64(L, A, R, proj, split);
  backward_f64(L, A, ns, rho0, drho, ncand, R, size, split, out, nt);
  free(split);
  free(L);
}

void orc_sirt_f64(const float* angles, int A, double rho0, double drho, double* img, int n,
                  const double* proj, int p, int iters, int nt) {
  sirt_f64(angles, A, rho0, drho, img, n, proj, p, iters, nt);
}

/* ------------------------------------------------------------------ helpers for KATs */

float orc_fp_ray_line(float cost, float sint, float lbase, int r, float rho0, float drho, int h,
                      const float* frame_img) {
  line_t_f32 L = {cost, sint, lbase, 0};
  return fp_ray_f32(&L, r, rho0, drho, h, frame_img);
}

float orc_bp_pixel_line(float cost, float sint, float lbase, float xmin, float ymin, float rho0,
                        float drho, int ncand, int R, const float* proj_row, int* s_out,
                        float* l_out) {
  line_t_f32 L = {cost, sint, lbase, 0};
  return bp_pixel_line_f32(&L, xmin, ymin, rho0, drho, ncand, R, proj_row, s_out, l_out);
}

/* ------------------------------------------------------------------ sparse structure */

/*
 * Entries of one ray of A (forwardprojection.fut:31-57): in row-step order, pixel Xb when
 * in bounds, then pixel Xt when it differs from Xb and is in bounds.  Columns are indices
 * into the caller's (untransposed) image.  Values are the f32 lengths lb, lt.
 */
static int fp_ray_entries(const line_t_f32* L, int flat, int r, float rho0, float drho, int h,
                          int32_t* cols, float* vals) {
  const int N = 2 * h;
  const float cost = L->cost, sint = L->sint, lbase = L->lbase;
  const float slope = -sint / cost;
  const float rho = rho0 + (float)r * drho;
  const float intercept = rho / cost;
  int cnt = 0;
  for (int i = -h; i < h; i++) {
    float xbot = (float)i * slope + intercept;
    float xtop = xbot + slope;
    int Xb = (int)floorf(xbot);
    int Xt = (int)floorf(xtop);
    float Xmax = (float)(Xb > Xt ? Xb : Xt);
    float xdiff = xtop - xbot;
    int single = Xb == Xt;
    int inb = Xb >= -h && Xb < h;
    int intop = (!single) && Xt >= -h && Xt < h;
    float lb = single ? lbase : ((Xmax - xbot) / xdiff) * lbase;
    float lt = ((xtop - Xmax) / xdiff) * lbase;
    if (inb) {
      if (cols) {
        cols[cnt] = flat ? (Xb + h) * N + (i + h) : (i + h) * N + (Xb + h);
        vals[cnt] = lb;
      }
      cnt++;
    }
    if (intop) {
      if (cols) {
        cols[cnt] = flat ? (Xt + h) * N + (i + h) : (i + h) * N + (Xt + h);
        vals[cnt] = lt;
      }
      cnt++;
    }
  }
  return cnt;
}

/*
 * CSR of A: rows = rays in angle-major order (angle*R + bin), columns = pixels.
 * Pass rowptr == NULL to only count.  Returns nnz.
 */
int64_t orc_fp_csr(const float* angles, int A, float rho0, float drho, int R, int N,
                   int64_t* rowptr, int32_t* cols, float* vals) {
  int h = N / 2;
  line_t_f32* L = (line_t_f32*)malloc(sizeof(line_t_f32) * (size_t)A);
  int ns = preprocess_f32(angles, A, L);
  long rows = (long)A * R;
  int32_t* cnt = (int32_t*)calloc((size_t)rows, sizeof(int32_t));
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < A; j++)
    for (int r = 0; r < R; r++)
      cnt[(long)L[j].orig * R + r] = fp_ray_entries(&L[j], j >= ns, r, rho0, drho, h, NULL, NULL);
  int64_t nnz = 0;
  for (long k = 0; k < rows; k++) {
    if (rowptr) rowptr[k] = nnz;
    nnz += cnt[k];
  }
  if (rowptr) {
    rowptr[rows] = nnz;
    for (int j = 0; j < A; j++)
      for (int r = 0; r < R; r++) {
        long row = (long)L[j].orig * R + r;
        fp_ray_entries(&L[j], j >= ns, r, rho0, drho, h, cols + rowptr[row], vals + rowptr[row]);
      }
  }
  free(cnt);
  free(L);
  return nnz;
}

/*
 * Entries of one pixel of the backprojection operator B (backprojection.fut:44-78):
 * steep lines in order, then flat lines (transposed frame), candidates in order; only
 * candidates inside [0, R) with non-zero length.  Columns are sinogram indices
 * (angle*R + bin).
 */
static int bp_pixel_entries(const line_t_f32* L, int A, int ns, int Y, int X, int h, float rho0,
                            float drho, int ncand, int R, int32_t* cols, float* vals) {
  int cnt = 0;
  int s_buf[64];
  float l_buf[64];
  for (int j = 0; j < A; j++) {
    float xmin = (float)(j < ns ? X - h : Y - h);
    float ymin = (float)(j < ns ? Y - h : X - h);
    bp_pixel_line_f32(&L[j], xmin, ymin, rho0, drho, ncand, R, NULL, s_buf, l_buf);
    for (int k = 0; k < ncand; k++) {
      if (s_buf[k] < 0 || s_buf[k] >= R || l_buf[k] == 0.0f) continue;
      if (cols) {
        cols[cnt] = L[j].orig * R + s_buf[k];
        vals[cnt] = l_buf[k];
      }
      cnt++;
    }
  }
  return cnt;
}

/* CSR of B: rows = pixels (row-major), columns = sinogram entries.  Returns nnz. */
int64_t orc_bp_csr(const float* angles, int A, float rho0, float drho, int R, int N,
                   int64_t* rowptr, int32_t* cols, float* vals) {
  int h = N / 2;
  int ncand = (int)ceilf(sqrtf(2.0f) / drho);
  if (ncand > 64) return -1;
  line_t_f32* L = (line_t_f32*)malloc(sizeof(line_t_f32) * (size_t)A);
  int ns = preprocess_f32(angles, A, L);
  int64_t nnz = 0;
  for (int Y = 0; Y < N; Y++)
    for (int X = 0; X < N; X++) {
      long row = (long)Y * N + X;
      if (rowptr) rowptr[row] = nnz;
      nnz += bp_pixel_entries(L, A, ns, Y, X, h, rho0, drho, ncand, R,
                              cols ? cols + nnz : NULL, vals ? vals + nnz : NULL);
    }
  if (rowptr) rowptr[(long)N * N] = nnz;
  free(L);
  return nnz;
}

/*
 * Exact adjoint of orc_fp: out = A^T sino, accumulated in double from A's entries in
 * CSR order.  (The reference's backprojection is not exactly A^T: on lines whose slope
 * rounds to zero it drops rays that lie on pixel edges; see intersect_fact.)
 */
void orc_fp_adjoint(const float* angles, int A, float rho0, float drho, int R, int N,
                    const float* sino, double* out) {
  int h = N / 2;
  line_t_f32* L = (line_t_f32*)malloc(sizeof(line_t_f32) * (size_t)A);
  int ns = preprocess_f32(angles, A, L);
  int32_t* cols = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * N + 2));
  float* vals = (float*)malloc(sizeof(float) * (size_t)(2 * N + 2));
  memset(out, 0, sizeof(double) * (size_t)N * N);
  for (int j = 0; j < A; j++)
    for (int r = 0; r < R; r++) {
      int c = fp_ray_entries(&L[j], j >= ns, r, rho0, drho, h, cols, vals);
      double y = sino[(long)L[j].orig * R + r];
      for (int k = 0; k < c; k++) out[cols[k]] += (double)vals[k] * y;
    }
  free(cols);
  free(vals);
  free(L);
}

/*
 * Number of bp candidates that fall outside [0, R) yet have non-zero length (the
 * reference would read another line's data there).  Zero for centred geometries.
 */
int64_t orc_bp_oob_nonzero(const float* angles, int A, float rho0, float drho, int R, int N) {
  int h = N / 2;
  int ncand = (int)ceilf(sqrtf(2.0f) / drho);
  if (ncand > 64) return -1;
  line_t_f32* L = (line_t_f32*)malloc(sizeof(line_t_f32) * (size_t)A);
  int ns = preprocess_f32(angles, A, L);
  int s_buf[64];
  float l_buf[64];
  int64_t bad = 0;
  for (int Y = 0; Y < N; Y++)
    for (int X = 0; X < N; X++)
      for (int j = 0; j < A; j++) {
        float xmin = (float)(j < ns ? X - h : Y - h);
        float ymin = (float)(j < ns ? Y - h : X - h);
        bp_pixel_line_f32(&L[j], xmin, ymin, rho0, drho, ncand, R, NULL, s_buf, l_buf);
        for (int k = 0; k < ncand; k++)
          if ((s_buf[k] < 0 || s_buf[k] >= R) && l_buf[k] != 0.0f) bad++;
      }
  free(L);
  return bad;
}

/*
 * Comparison tools (futhark/tools/{sanityCheck,diff,mse}.fut), sequential as Futhark C runs them.
 * sanityCheck.fut:1-4: reduce (&&) true over |s - t| < 0.05.
 */
int orc_sanity_check(const float* sane, const float* test, int64_t n, float tol) {
  int ok = 1;
  for (int64_t i = 0; i < n; i++) ok = ok && (fabsf(sane[i] - test[i]) < tol);
  return ok;
}

/* diff.fut:1-7: length (filter (> 0) (map3 (\s t i -> if |s-t| > 0.05 then i else -1))). */
int64_t orc_diff(const float* sane, const float* test, int64_t n, float tol) {
  int64_t count = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t v = fabsf(sane[i] - test[i]) > tol ? i : -1;
    if (v > 0) count++;
  }
  return count;
}

/* mse.fut:1-3: (reduce (+) 0 (map2 (\s t -> (s-t)**2))) / r32(n), all in f32. */
float orc_mse(const float* sane, const float* test, int64_t n) {
  float acc = 0.0f;
  for (int64_t i = 0; i < n; i++) acc = acc + powf(sane[i] - test[i], 2.0f);
  return acc / (float)n;
}
