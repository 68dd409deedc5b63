/*
 * oracle_impl.h -- body of the CPU restatement, instantiated twice by oracle.c:
 * once with REAL=float (the FP32 restatement that the HIP path must match) and once
 * with REAL=double (the FP64 twin used to state errors against exact-ish arithmetic).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in synkrotomo_amd/ may link or call this code.
 *
 * Every function follows the reference line by line; the file:line citations point
 * into tomograph/synkrotomo.  Evaluation order, rounding points and reductions are the
 * ones Futhark's sequential C backend gives: left-to-right folds starting from 0,
 * no FMA contraction (this file is compiled with -ffp-contract=off), libm for
 * cos/sin/pow/sqrt.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

typedef struct {
  REAL cost, sint, lbase;
  int orig; /* index of the angle in the caller's order */
} FN(line_t);

/* sirtlib.fut:6-7 */
static int FN(is_flat)(REAL c, REAL s) { return RFABS(s) >= RFABS(c); }

/* sirtlib.fut:9-10 */
static REAL FN(safe_inverse)(REAL v) { return v == (REAL)0 ? (REAL)0 : (REAL)1 / v; }

/*
 * sirtlib.fut:28-44 (preprocess).  Per angle: cos, sin, l_steep = sqrt(1+(sin/cos)**2),
 * l_flat = sqrt(1+(cos/sin)**2), flat = |sin| >= |cos|.  Stable partition: steep lines
 * (cos, -sin, l_steep) first, then flat lines (-sin, cos, l_flat), each in the caller's
 * relative order.  Returns the number of steep lines.
 */
static int FN(preprocess)(const float* angles, int A, FN(line_t) * lines) {
  int ns = 0;
  for (int i = 0; i < A; i++) {
    REAL c = RCOS((REAL)angles[i]), s = RSIN((REAL)angles[i]);
    if (!FN(is_flat)(c, s)) ns++;
  }
  int js = 0, jf = ns;
  for (int i = 0; i < A; i++) {
    REAL c = RCOS((REAL)angles[i]);
    REAL s = RSIN((REAL)angles[i]);
    REAL lsteep = RSQRT((REAL)1 + RPOW(s / c, (REAL)2));
    REAL lflat = RSQRT((REAL)1 + RPOW(c / s, (REAL)2));
    if (FN(is_flat)(c, s)) {
      lines[jf].cost = -s;
      lines[jf].sint = c;
      lines[jf].lbase = lflat;
      lines[jf].orig = i;
      jf++;
    } else {
      lines[js].cost = c;
      lines[js].sint = -s;
      lines[js].lbase = lsteep;
      lines[js].orig = i;
      js++;
    }
  }
  return ns;
}

/*
 * forwardprojection.fut:21-59 (fpTlib.fp), one ray: line (cost, sint, lbase), bin r.
 * Walks rows i in [-h, h) of the line's frame; in each row the ray covers pixel Xb and,
 * when it crosses a column boundary, pixel Xt, with the length split at the crossing.
 * img is the frame image (the caller's image for steep lines, its transpose for flat).
 */
static REAL FN(fp_ray)(const FN(line_t) * L, int r, REAL rho0, REAL drho, int h,
                       const REAL* img) {
  const int size = 2 * h;
  const int pixcorr = h * (1 + size);
  const REAL cost = L->cost, sint = L->sint, lbase = L->lbase;
  const REAL slope = -sint / cost;
  const REAL rho = rho0 + (REAL)r * drho;
  const REAL intercept = rho / cost;
  REAL sum = 0;
  for (int i = -h; i < h; i++) {
    REAL xbot = (REAL)i * slope + intercept;
    REAL xtop = xbot + slope;
    int Xb = (int)RFLOOR(xbot);
    int Xt = (int)RFLOOR(xtop);
    REAL Xmax = (REAL)(Xb > Xt ? Xb : Xt);
    REAL xdiff = xtop - xbot;
    int single = Xb == Xt;
    int inb = Xb >= -h && Xb < h;
    int intop = (!single) && Xt >= -h && Xt < h;
    REAL lbfac = (Xmax - xbot) / xdiff;
    REAL lb = single ? lbase : lbfac * lbase;
    REAL lt = ((xtop - Xmax) / xdiff) * lbase;
    REAL bot = inb ? lb * img[Xb + i * size + pixcorr] : (REAL)0;
    REAL top = intop ? lt * img[Xt + i * size + pixcorr] : (REAL)0;
    sum = sum + (bot + top);
  }
  return sum;
}

/* forwardprojection.fut:21-59 over a list of lines: out[j*R + r] (line-major) */
static void FN(fp_lines)(const FN(line_t) * lines, int nl, REAL rho0, REAL drho, int R, int h,
                         const REAL* img, REAL* out, int nt) {
  long total = (long)nl * R;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64) if (nt > 1)
  for (long k = 0; k < total; k++) {
    int j = (int)(k / R), r = (int)(k % R);
    out[k] = FN(fp_ray)(&lines[j], r, rho0, drho, h, img);
  }
}

/* backprojection.fut:22-33 */
static REAL FN(intersect_fact)(REAL plus, REAL minus, REAL mini, REAL maxi) {
  REAL lo = RMAX(plus - mini, (REAL)0);
  REAL hi = RMAX(maxi - minus, (REAL)0);
  REAL diff = plus - minus;
  REAL m = RMIN(lo, hi);
  REAL p = (m == (REAL)0) ? (REAL)0 : m / diff;
  return RMIN(p, (REAL)1);
}

/*
 * backprojection.fut:48-65, one (pixel, line) pair: candidate rays s, s+1, ... (ncand of
 * them) starting at the first ray whose rho is not below the pixel's circumscribed
 * circle.  Writes the candidate bin indices and their lengths; returns the inner sum
 * over candidates of length * proj[bin] (proj = this line's sinogram row, or NULL).
 * Candidates outside [0, R) contribute nothing (the reference reads them unchecked,
 * with zero length for centred geometries).
 */
static REAL FN(bp_pixel_line)(const FN(line_t) * L, REAL xmin, REAL ymin, REAL rho0, REAL drho,
                              int ncand, int R, const REAL* proj_row, int* s_out,
                              REAL* l_out) {
  const REAL fact = RSQRT((REAL)2) / (REAL)2; /* backprojection.fut:20 */
  const REAL cost = L->cost, sint = L->sint, lbase = L->lbase;
  const REAL slope = -sint / cost;
  REAL x = xmin + (REAL)0.5 - fact * cost;
  REAL y = ymin + (REAL)0.5 - fact * sint;
  REAL rho = cost * x + sint * y;
  REAL s = RCEIL((rho - rho0) / drho);
  REAL xbase = ymin * slope;
  REAL inner = 0;
  for (int k = 0; k < ncand; k++) {
    REAL sp = s + (REAL)k;
    REAL r = sp * drho + rho0;
    REAL xbot = xbase + (r / cost);
    REAL xtop = xbot + slope;
    REAL maxx = RMAX(xbot, xtop);
    REAL minx = RMIN(xbot, xtop);
    REAL l = FN(intersect_fact)(maxx, minx, xmin, xmin + (REAL)1) * lbase;
    int si = (int)sp;
    if (s_out) s_out[k] = si;
    if (l_out) l_out[k] = l;
    REAL v = 0;
    if (proj_row && si >= 0 && si < R) v = l * proj_row[si];
    inner = inner + v;
  }
  return inner;
}

/*
 * backprojection.fut:36-68 (bpTlib.bp) for one frame: out[(irow+h)*N + (icol+h)],
 * xmin = icol, ymin = irow, lines in the given order, proj[idx*R + s].
 */
static void FN(bp_lines)(const FN(line_t) * lines, int nl, REAL rho0, REAL drho, int ncand, int R,
                         int h, const REAL* proj, REAL* out, int nt) {
  const int N = 2 * h;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1) if (nt > 1)
  for (int irow = -h; irow < h; irow++) {
    for (int icol = -h; icol < h; icol++) {
      REAL xmin = (REAL)icol, ymin = (REAL)irow;
      REAL acc = 0;
      for (int idx = 0; idx < nl; idx++)
        acc = acc + FN(bp_pixel_line)(&lines[idx], xmin, ymin, rho0, drho, ncand, R,
                                      proj + (long)idx * R, NULL, NULL);
      out[(long)(irow + h) * N + (icol + h)] = acc;
    }
  }
}

/* flatten <| transpose <| unflatten size size x */
static void FN(transpose)(const REAL* in, REAL* out, int N) {
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) out[(long)i * N + j] = in[(long)j * N + i];
}

/* sirtlib.fut:16-21 (fix_projections): sinogram rows in steep-then-flat line order */
static void FN(split_sino)(const FN(line_t) * lines, int A, int R, const REAL* proj, REAL* out) {
  for (int j = 0; j < A; j++)
    for (int r = 0; r < R; r++) out[(long)j * R + r] = proj[(long)lines[j].orig * R + r];
}

/*
 * forwardprojection.fut:61-84: fp steep on x, fp flat on x^T, then postprocess_fp
 * (sirtlib.fut:24-25) scatters each line back to its original angle row.
 */
static void FN(forward)(const FN(line_t) * lines, int A, int ns, REAL rho0, REAL drho, int R,
                        int N, const REAL* img, REAL* out, int nt) {
  int h = N / 2;
  REAL* imgT = (REAL*)malloc(sizeof(REAL) * (size_t)N * N);
  REAL* tmp = (REAL*)malloc(sizeof(REAL) * (size_t)A * R);
  FN(transpose)(img, imgT, N);
  FN(fp_lines)(lines, ns, rho0, drho, R, h, img, tmp, nt);
  FN(fp_lines)(lines + ns, A - ns, rho0, drho, R, h, imgT, tmp + (long)ns * R, nt);
  for (int j = 0; j < A; j++)
    for (int r = 0; r < R; r++) out[(long)lines[j].orig * R + r] = tmp[(long)j * R + r];
  free(imgT);
  free(tmp);
}

/*
 * backprojection.fut:70-78: bp steep + transpose(bp flat).  proj_split is the sinogram in
 * steep-then-flat line order (fix_projections).
 */
static void FN(backward)(const FN(line_t) * lines, int A, int ns, REAL rho0, REAL drho, int ncand,
                         int R, int N, const REAL* proj_split, REAL* out, int nt) {
  int h = N / 2;
  REAL* s = (REAL*)malloc(sizeof(REAL) * (size_t)N * N);
  REAL* f = (REAL*)malloc(sizeof(REAL) * (size_t)N * N);
  FN(bp_lines)(lines, ns, rho0, drho, ncand, R, h, proj_split, s, nt);
  FN(bp_lines)(lines + ns, A - ns, rho0, drho, ncand, R, h, proj_split + (long)ns * R, f, nt);
  for (int Y = 0; Y < N; Y++)
    for (int X = 0; X < N; X++)
      out[(long)Y * N + X] = s[(long)Y * N + X] + f[(long)X * N + Y];
  free(s);
  free(f);
}

/* SIRT.fut:9-41, in place on img */
static void FN(sirt)(const float* angles, int A, REAL rho0, REAL drho, REAL* img, int n,
                     const REAL* proj, int p, int iters, int nt) {
  const int ncand = (int)RCEIL(RSQRT((REAL)2) / drho);
  const int N = (int)RSQRT((REAL)n);
  const int R = p / A;
  FN(line_t)* lines = (FN(line_t)*)malloc(sizeof(FN(line_t)) * (size_t)A);
  int ns = FN(preprocess)(angles, A, lines);
  size_t P = (size_t)A * R, NN = (size_t)N * N;
  REAL* bsplit = (REAL*)malloc(sizeof(REAL) * P);
  REAL* rw = (REAL*)malloc(sizeof(REAL) * P);
  REAL* cw = (REAL*)malloc(sizeof(REAL) * NN);
  REAL* ones_i = (REAL*)malloc(sizeof(REAL) * NN);
  REAL* ones_p = (REAL*)malloc(sizeof(REAL) * P);
  REAL* imgT = (REAL*)malloc(sizeof(REAL) * NN);
  REAL* q = (REAL*)malloc(sizeof(REAL) * P);
  REAL* bpv = (REAL*)malloc(sizeof(REAL) * NN);
  FN(split_sino)(lines, A, R, proj, bsplit);
  for (size_t k = 0; k < NN; k++) ones_i[k] = 1;
  for (size_t k = 0; k < P; k++) ones_p[k] = 1;
  int h = N / 2;
  /* SIRT.fut:22-25: row sums (fp of ones, per half) and column sums (bp of ones) */
  FN(fp_lines)(lines, ns, rho0, drho, R, h, ones_i, rw, nt);
  FN(fp_lines)(lines + ns, A - ns, rho0, drho, R, h, ones_i, rw + (long)ns * R, nt);
  for (size_t k = 0; k < P; k++) rw[k] = FN(safe_inverse)(rw[k]);
  FN(backward)(lines, A, ns, rho0, drho, ncand, R, N, ones_p, cw, nt);
  for (size_t k = 0; k < NN; k++) cw[k] = FN(safe_inverse)(cw[k]);
  /* SIRT.fut:27-39 */
  for (int it = 0; it < iters; it++) {
    FN(transpose)(img, imgT, N);
    FN(fp_lines)(lines, ns, rho0, drho, R, h, img, q, nt);
    FN(fp_lines)(lines + ns, A - ns, rho0, drho, R, h, imgT, q + (long)ns * R, nt);
    for (size_t k = 0; k < P; k++) q[k] = rw[k] * (bsplit[k] - q[k]);
    FN(backward)(lines, A, ns, rho0, drho, ncand, R, N, q, bpv, nt);
    for (size_t k = 0; k < NN; k++) img[k] = img[k] + cw[k] * bpv[k];
  }
  free(lines); free(bsplit); free(rw); free(cw); free(ones_i); free(ones_p);
  free(imgT); free(q); free(bpv);
}
