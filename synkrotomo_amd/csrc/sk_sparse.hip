// sk_sparse.hip -- explicit system matrices built on the GPU, and their SpMV.
//
// The reference's *_sparse entry points (futhark/forwardprojection_sparse.fut:26-40,
// futhark/backprojection_sparse.fut:26-37) have the dense bodies; here they run on
// explicit matrices:
//   A (SK_MAT_FP): rows = rays in angle-major order (angle*R + bin), columns = pixels of the
//                  caller's image.  Entries per ray in row-step order: pixel Xb when in
//                  bounds, then Xt when it differs and is in bounds, with the lengths lb, lt
//                  of forwardprojection.fut:37-54 (IEEE division -> bitwise those of fp).
//   B (SK_MAT_BP): rows = pixels, columns = sinogram entries (angle*R + bin).  Entries per
//                  pixel: steep lines then flat lines, candidates in order, those inside
//                  [0, R) with non-zero intersect_fact length (backprojection.fut:44-78).
//                  B is the reference's backprojector, which is A^T except for edge rays on
//                  lines whose slope rounds to zero.
//   A^T (SK_MAT_ADJ): the CSC form of A, stored as the CSR of A^T: rows = pixels, columns =
//                  rays.  Entries per pixel: steep lines then flat lines, rays in bin order,
//                  exactly A's entries of that pixel (explicit zeros included) with A's
//                  values bit for bit: the same xbot = y*slope + (s*drho + rho0)/cost and
//                  forwardprojection.fut:37-54 lengths, evaluated for the backprojector's
//                  candidate window widened by one bin on each side (a ray with an entry in
//                  the pixel touches the pixel square, so it lies within sqrt(2)/2 of the
//                  pixel centre; the margin absorbs the rounding of the window origin).
// CSR, 64-bit row pointers (nnz exceeds 2^31 at 2048^2), 32-bit columns, f32 values.
//
// Build = count pass -> 64-bit exclusive scan -> fill pass.  Both passes are
// wavefront-cooperative: one wave per row, lanes over the row's steps (A: 64 image rows at
// a time; B: 64 lines at a time); ballots and mbcnt place each lane's entries, so the
// entries come out in the sequential order above and the writes are contiguous.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "sk_common.h"

struct sk_mat {
  int kind = 0;
  int spmv_variant = -1;  // 0: wave per row, 1: lane per row (lockstep), 2: wave per row with
                          // 16-byte streams; 3, 4: as 0, 2 with the gathers of steep rays
                          // from x^T (SK_MAT_FP only); 7: SELL-64 copy; 9: SELL-64 with
                          // 16-bit column steps; 10: row slots (SK_MAT_FP only); -1: not
                          // yet timed
  int64_t rows = 0, cols = 0, nnz = 0;
  // SK_MAT_FP: per original angle, 1 if its line is steep (its rays step through image rows:
  // in x^T their pixels are contiguous), and the geometry's R and N
  unsigned char* steep = nullptr;
  int R = 0, N = 0;
  // SELL-64 copy (variant 7, built on first use): rows in slices of 64; slice k holds
  // L_k = its longest row's length columns of 64 entries, entry j of row 64k + l at
  // sell_off[k] + 64j + l (zero padding past each row's end), so a wave reads both streams
  // coalesced and its 64 gathers belong to 64 neighbouring rows.
  int64_t* sell_off = nullptr;  // [slices + 1]
  int32_t* sell_col = nullptr;
  float* sell_val = nullptr;
  int32_t* sell_len = nullptr;  // [rows] row lengths
  bool sell_failed = false;     // no memory for the copy: variant 7 unavailable
  // Variant 9: the SELL-64 values with the columns as 16-bit steps d_j = col_j - col_{j-1}
  // (col_{-1} = the row's first column, kept in sell_first; padding repeats the last column,
  // d = 0), two entries per 32-bit word: word sell_off[k]/2 + 64*(j/2) + l, entry j in bits
  // 16*(j&1) ... 16*(j&1)+15.  6 bytes per entry instead of 8.  Unavailable (d16_failed) when a
  // row has a step outside [-32768, 32767] or there is no memory for it.
  // For SK_MAT_BP / SK_MAT_ADJ the steps run over columns in line-table order, j*R + r for the
  // line at table position j (a pixel's entries come steep lines first, then flat lines, so in
  // the caller's angle order its columns jump by up to A*R once or twice; in table order each
  // step is about R); sk_spmv permutes the input's rows into that order first.
  uint32_t* sell_d16 = nullptr;
  int32_t* sell_first = nullptr;  // [rows]
  bool d16_failed = false;
  int32_t* line_pos = nullptr;    // SK_MAT_BP / SK_MAT_ADJ: [A] table position of each angle,
  int32_t* line_orig = nullptr;   //   then [A] the angle at each table position
  int A = 0;
  // Variant 10 (SK_MAT_FP): row slots.  Slice s = caller's angle a x block kb of 64 bins (rays
  // a*R + 64kb + l, one per lane l); its records k = 0 .. rs_len-1 are the frame rows
  // rs_lo + k.  Record (s, k, l) holds the ray's two entries of that row, (lb, lt) at
  // rs_val[(rs_off[s] + k)*64 + l], NaN where an entry does not exist, and a control byte in byte
  // k%4 of rs_ctrl[(rs_off[s]/4 + k/4)*64 + l]: low nibble Xb(i) - Xb(i-1), high nibble
  // Xt(i) - Xb(i), both signed 4-bit.  The frame index (i + N/2)*N + Xb + N/2 of record k is
  // rs_first[s*64 + l] + (k+1)*N + the nibble sum: the lanes of a wave gather from one frame row
  // (the image for steep angles, its transpose for flat ones) instead of drifting apart.
  int64_t* rs_off = nullptr;    // [slices + 1], a multiple of 8 records per slice
  uint32_t* rs_ctrl = nullptr;
  float2* rs_val = nullptr;
  int32_t* rs_first = nullptr;  // [slices * 64]
  int32_t* rs_lpos = nullptr;   // [A] table position of each angle's line
  sk::Line* rs_lines = nullptr; // the geometry's line table (the slots are built on first use)
  float rho0 = 0.0f, drho = 1.0f;
  int ncand = 0;
  // For SK_MAT_BP the row slots are per line instead: slice s = image row Y x block xb of 64
  // pixels (pixels Y*N + 64xb + l), record j = the line at table position j (A rounded up to 8
  // records), holding the pixel's two candidates (s', s'+1) of that line, rs_val as above; the
  // control is a 16-bit step of the base column j*R + s' (line-table order) per record, two per
  // word: word (s*A8/2 + j/2)*64 + l; rs_first holds the base column before record 0.
  bool rs_failed = false;
  int64_t* rowptr = nullptr;  // [rows + 1]
  int32_t* col = nullptr;     // [nnz]
  float* val = nullptr;       // [nnz]
  int device = 0;
};

namespace sk {

constexpr int SP_BLOCK = 256;  // 4 waves, one row each

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ int prefix_of(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// ------------------------------------------------------------------ A: one wave per ray
template <bool FILL>
__global__ __launch_bounds__(SP_BLOCK) void k_fp_rows(const Line* __restrict__ lines, int A, int R,
                                                      int N, float rho0, float drho,
                                                      int32_t* __restrict__ counts,
                                                      const int64_t* __restrict__ rowptr,
                                                      int32_t* __restrict__ cols,
                                                      float* __restrict__ vals) {
  const int lane = lane_id();
  const long ray = (long)blockIdx.x * (SP_BLOCK / 64) + (threadIdx.x >> 6);
  if (ray >= (long)A * R) return;  // whole wave
  const int j = (int)(ray / R), r = (int)(ray - (long)j * R);
  const Line L = lines[j];
  const int h = N >> 1;
  const bool flat = L.flags & LINE_FLAT;
  const float slope = L.slope, lbase = L.lbase;
  const float icpt = (rho0 + (float)r * drho) / L.cost;  // forwardprojection.fut:33-34
  const long row = (long)L.orig * R + r;
  long pos = FILL ? rowptr[row] : 0;
  int total = 0;
  for (int i0 = -h; i0 < h; i0 += 64) {
    const int i = i0 + lane;
    const bool live = i < h;
    const float xbot = (float)i * slope + icpt;
    const float xtop = xbot + slope;
    const int Xb = (int)floorf(xbot), Xt = (int)floorf(xtop);
    const bool single = Xb == Xt;
    const bool inb = live && Xb >= -h && Xb < h;
    const bool intop = live && !single && Xt >= -h && Xt < h;
    const unsigned long long b1 = __ballot(inb || intop);
    const unsigned long long b2 = __ballot(inb && intop);
    if (FILL) {
      const float Xmax = (float)(Xb > Xt ? Xb : Xt);
      const float xdiff = xtop - xbot;
      const float lb = single ? lbase : ((Xmax - xbot) / xdiff) * lbase;
      const float lt = ((xtop - Xmax) / xdiff) * lbase;
      long p = pos + prefix_of(b1) + prefix_of(b2);
      if (inb) {
        cols[p] = flat ? (Xb + h) * N + (i + h) : (i + h) * N + (Xb + h);
        vals[p] = lb;
        p++;
      }
      if (intop) {
        cols[p] = flat ? (Xt + h) * N + (i + h) : (i + h) * N + (Xt + h);
        vals[p] = lt;
      }
    }
    const int n = __popcll(b1) + __popcll(b2);
    pos += n;
    total += n;
  }
  if (!FILL && lane == 0) counts[row] = total;
}

// ------------------------------------------------------------------ B, A^T: one wave per pixel
template <bool FILL, bool ADJ>
__global__ __launch_bounds__(SP_BLOCK) void k_bp_rows(const Line* __restrict__ lines, int A, int R,
                                                      int N, float rho0, float drho, int ncand,
                                                      int32_t* __restrict__ counts,
                                                      const int64_t* __restrict__ rowptr,
                                                      int32_t* __restrict__ cols,
                                                      float* __restrict__ vals) {
  const int lane = lane_id();
  const long pix = (long)blockIdx.x * (SP_BLOCK / 64) + (threadIdx.x >> 6);
  if (pix >= (long)N * N) return;
  const int Y = (int)(pix / N), X = (int)(pix - (long)Y * N);
  const int h = N >> 1;
  long pos = FILL ? rowptr[pix] : 0;
  int total = 0;
  for (int j0 = 0; j0 < A; j0 += 64) {
    const int j = j0 + lane;
    const bool live = j < A;
    Line L = lines[live ? j : 0];
    const bool fl = L.flags & LINE_FLAT;
    const float xmin = (float)((fl ? Y : X) - h);
    const float ymin = (float)((fl ? X : Y) - h);
    // backprojection.fut:50-54
    const float px = (xmin + 0.5f) - L.fc;
    const float py = (ymin + 0.5f) - L.fs;
    const float rho = L.cost * px + L.sint * py;
    const float s = ceilf((rho - rho0) / drho);
    const float xbase = ymin * L.slope;
    int before = 0;  // entries of lower lanes
    int mine = 0;    // entries of this lane so far
    constexpr int KMAX = 16 + (ADJ ? 2 : 0);
    const int nk = ncand + (ADJ ? 2 : 0);
    unsigned long long masks[KMAX];
    float lens[KMAX];
    int sps[KMAX];
    for (int k = 0; k < nk; ++k) {
      // backprojection.fut:56-64
      const float sp = s + (float)(ADJ ? k - 1 : k);
      const float rr = sp * drho + rho0;
      const float xbot = xbase + (rr / L.cost);
      const float xtop = xbot + L.slope;
      float l;
      bool member;
      if constexpr (ADJ) {
        // the entry of A for (ray sp, this pixel): forwardprojection.fut:37-54
        const float fXb = floorf(xbot), fXt = floorf(xtop);
        const bool single = fXb == fXt;
        const float Xmax = fmaxf(fXb, fXt);
        const float xdiff = xtop - xbot;
        const float lb = single ? L.lbase : ((Xmax - xbot) / xdiff) * L.lbase;
        const float lt = ((xtop - Xmax) / xdiff) * L.lbase;
        member = fXb == xmin || (!single && fXt == xmin);
        l = fXb == xmin ? lb : lt;
      } else {
        const float maxx = fmaxf(xbot, xtop), minx = fminf(xbot, xtop);
        const float lo = fmaxf(maxx - xmin, 0.0f);
        const float hi = fmaxf((xmin + 1.0f) - minx, 0.0f);
        const float diff = maxx - minx;
        const float m = fminf(lo, hi);
        const float p = (m == 0.0f) ? 0.0f : m / diff;
        l = fminf(p, 1.0f) * L.lbase;
        member = l != 0.0f;
      }
      const int si = (int)sp;
      const bool ok = live && si >= 0 && si < R && member;
      masks[k] = __ballot(ok);
      lens[k] = l;
      sps[k] = ok ? si : -1;
      before += prefix_of(masks[k]);
    }
    if (FILL) {
      long p = pos + before;
      for (int k = 0; k < nk; ++k) {
        if (sps[k] >= 0) {
          cols[p + mine] = L.orig * R + sps[k];
          vals[p + mine] = lens[k];
          mine++;
        }
      }
    }
    int n = 0;
    for (int k = 0; k < nk; ++k) n += __popcll(masks[k]);
    pos += n;
    total += n;
  }
  if (!FILL && lane == 0) counts[pix] = total;
}

// ------------------------------------------------------------------ 64-bit exclusive scan
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = 256 * SCAN_ITEMS;

__device__ __forceinline__ long block_excl_scan(long v, long* sh, long* total) {
  // Hillis-Steele over 256 threads in LDS (scan of per-thread sums).
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    long add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  long incl = sh[t];
  *total = sh[255];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(256) void k_scan_tiles(const int32_t* __restrict__ in, long n,
                                                    int64_t* __restrict__ out,
                                                    int64_t* __restrict__ tile_sums) {
  __shared__ long sh[256];
  const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
  long v[SCAN_ITEMS];
  long s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  long total;
  long run = block_excl_scan(s, sh, &total);
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One block scans the tile sums (any count) in place to exclusive offsets; writes the
// grand total to *grand.
__global__ __launch_bounds__(256) void k_scan_sums(int64_t* __restrict__ sums, long ntiles,
                                                   int64_t* __restrict__ grand) {
  __shared__ long sh[256];
  long carry = 0;
  for (long b0 = 0; b0 < ntiles; b0 += 256) {
    const long i = b0 + threadIdx.x;
    const long v = i < ntiles ? sums[i] : 0;
    long total;
    const long ex = block_excl_scan(v, sh, &total);
    if (i < ntiles) sums[i] = carry + ex;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) *grand = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(int64_t* __restrict__ out, long n,
                                                  const int64_t* __restrict__ sums,
                                                  const int64_t* __restrict__ grand) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] += sums[i / SCAN_TILE];
  if (i == 0) out[n] = *grand;
}

// ------------------------------------------------------------------ SpMV: one wave per row
// Gathers of one row: from x itself, or (TR: a steep ray of A, whose pixels p = y*N + x step
// through image rows) from x^T at x*N + y, where they are contiguous.  x^T holds exactly x's
// values, so both read the same numbers: results are bitwise equal.
struct SpGather {
  const float* x;
  int N;
  double invN;  // 1/N
  template <bool TR>
  __device__ __forceinline__ float at(int c) const {
    if constexpr (!TR) {
      return x[c];
    } else {
      // y = floor(c / N): (c + 0.5)/N is at least 0.5/N away from an integer, far beyond the
      // FP64 product's rounding for c < 2^31
      const int yq = (int)(((double)c + 0.5) * invN);
      return x[(c - yq * N) * N + yq];
    }
  }
};

template <int UNROLL, bool TR>
__device__ __forceinline__ float spmv_row(const int32_t* __restrict__ col, const float* __restrict__ val,
                                          long beg, long end, int lane, const SpGather& g) {
  float acc = 0.0f;
  long e = beg + lane;
  for (; e + 64 * (UNROLL - 1) < end; e += 64 * UNROLL) {
    int c[UNROLL];
    float v[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      c[u] = __builtin_nontemporal_load(col + e + 64 * u);
      v[u] = __builtin_nontemporal_load(val + e + 64 * u);
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) acc = fmaf(v[u], g.at<TR>(c[u]), acc);
  }
  for (; e < end; e += 64)
    acc = fmaf(__builtin_nontemporal_load(val + e), g.at<TR>(__builtin_nontemporal_load(col + e)), acc);
  return acc;
}

// TR: steep rows (steep[row / R]) gather from xT; the others from x.
template <int UNROLL, bool TR>
__global__ __launch_bounds__(256) void k_spmv(const int64_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ col,
                                              const float* __restrict__ val, long rows, long ncols,
                                              const float* __restrict__ x, float* __restrict__ y,
                                              const float* __restrict__ xT, const unsigned char* __restrict__ steep,
                                              int R, int N, double invN) {
  const int lane = lane_id();
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const size_t boff = (size_t)blockIdx.y * ncols;
  const long beg = rowptr[row], end = rowptr[row + 1];
  float acc;
  if (TR && steep[row / R])  // wave-uniform
    acc = spmv_row<UNROLL, TR>(col, val, beg, end, lane, SpGather{xT + boff, N, invN});
  else
    acc = spmv_row<UNROLL, false>(col, val, beg, end, lane, SpGather{x + boff, N, invN});
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) y[(size_t)blockIdx.y * rows + row] = acc;
}

// ------------------------------------------------------------------ SpMV: one wave per row, wide
// As k_spmv, but each lane streams aligned 4-entry chunks (16-byte column and value loads)
// and keeps U chunks -- 4U gathers -- in flight per iteration; the at most 3 entries before
// the first aligned chunk and after the last are handled by single lanes.
template <int U, bool TR>
__device__ __forceinline__ float spmv_row_wide(const int32_t* __restrict__ col, const float* __restrict__ val,
                                               long beg, long end, int lane, const SpGather& xb);

template <int U, bool TR>
__global__ __launch_bounds__(256) void k_spmv_wide(const int64_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col,
                                                   const float* __restrict__ val, long rows,
                                                   long ncols, const float* __restrict__ x,
                                                   float* __restrict__ y, const float* __restrict__ xT,
                                                   const unsigned char* __restrict__ steep, int R, int N,
                                                   double invN) {
  const int lane = lane_id();
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const size_t boff = (size_t)blockIdx.y * ncols;
  const long beg = rowptr[row], end = rowptr[row + 1];
  float acc;
  if (TR && steep[row / R])  // wave-uniform
    acc = spmv_row_wide<U, TR>(col, val, beg, end, lane, SpGather{xT + boff, N, invN});
  else
    acc = spmv_row_wide<U, false>(col, val, beg, end, lane, SpGather{x + boff, N, invN});
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) y[(size_t)blockIdx.y * rows + row] = acc;
}

template <int U, bool TR>
__device__ __forceinline__ float spmv_row_wide(const int32_t* __restrict__ col, const float* __restrict__ val,
                                               long beg, long end, int lane, const SpGather& xb) {
  const long a0 = min((beg + 3) & ~3L, end);  // first aligned entry (or end)
  const long a1 = max(a0, end & ~3L);         // end of the aligned body
  float acc = 0.0f;
  if (beg + lane < a0) acc = fmaf(val[beg + lane], xb.at<TR>(col[beg + lane]), acc);
  typedef int i32x4 __attribute__((ext_vector_type(4)));
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const i32x4* c4 = reinterpret_cast<const i32x4*>(col);
  const f32x4* v4 = reinterpret_cast<const f32x4*>(val);
  const long q1 = a1 >> 2;
  for (long q = (a0 >> 2) + lane; q < q1; q += 64 * U) {
    i32x4 c[U];
    f32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = q + 64 * u < q1;
      c[u] = ok ? __builtin_nontemporal_load(c4 + q + 64 * u) : i32x4{0, 0, 0, 0};
      v[u] = ok ? __builtin_nontemporal_load(v4 + q + 64 * u) : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      acc = fmaf(v[u].x, xb.at<TR>(c[u].x), acc);
      acc = fmaf(v[u].y, xb.at<TR>(c[u].y), acc);
      acc = fmaf(v[u].z, xb.at<TR>(c[u].z), acc);
      acc = fmaf(v[u].w, xb.at<TR>(c[u].w), acc);
    }
  }
  if (a1 + lane < end) acc = fmaf(val[a1 + lane], xb.at<TR>(col[a1 + lane]), acc);
  return acc;
}

// ------------------------------------------------------------------ SELL-64 (variant 7)
// Per slice: 64 * (longest row, rounded up to even) entries; per row: its length.
__global__ __launch_bounds__(256) void k_sell_sizes(const int64_t* __restrict__ rowptr, long rows,
                                                    int32_t* __restrict__ slice_entries,
                                                    int32_t* __restrict__ len) {
  const long row = (long)blockIdx.x * 256 + threadIdx.x;  // 4 slices per block
  const int n = row < rows ? (int)(rowptr[row + 1] - rowptr[row]) : 0;
  if (row < rows) len[row] = n;
  int m = n;
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  m = (m + 1) & ~1;  // even: the 16-bit steps of variant 9 pack two entries per word
  if ((threadIdx.x & 63) == 0 && row < rows) slice_entries[row >> 6] = 64 * m;
}

// One one-wave block per slice, through a 64 x 64 LDS tile: for entries j0 .. j0+63, the wave
// reads each of the slice's rows coalesced (lane = entry), then writes the tile transposed
// (lane = row), so both the CSR reads and the SELL writes are contiguous.
__global__ __launch_bounds__(64) void k_sell_fill(const int64_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ col,
                                                  const float* __restrict__ val, long rows,
                                                  const int64_t* __restrict__ off,
                                                  int32_t* __restrict__ scol, float* __restrict__ sval) {
  __shared__ int32_t tc[64][65];
  __shared__ float tv[64][65];
  __shared__ long sbeg[64];
  __shared__ int slen[64];
  const long slice = blockIdx.x;
  const int lane = threadIdx.x;
  const long myrow = slice * 64 + lane;
  sbeg[lane] = myrow < rows ? rowptr[myrow] : 0;
  slen[lane] = myrow < rows ? (int)(rowptr[myrow + 1] - rowptr[myrow]) : 0;
  const long base = off[slice];
  const long L = (off[slice + 1] - base) >> 6;
  __syncthreads();
  for (long j0 = 0; j0 < L; j0 += 64) {
    const long j = j0 + lane;
#pragma unroll 8
    for (int r = 0; r < 64; ++r) {
      const bool in = j < slen[r];
      tc[r][lane] = in ? col[sbeg[r] + j] : 0;
      tv[r][lane] = in ? val[sbeg[r] + j] : 0.0f;
    }
    __syncthreads();
    const int nj = (int)min(64L, L - j0);
    for (int jj = 0; jj < nj; ++jj) {
      scol[base + (j0 + jj) * 64 + lane] = tc[lane][jj];
      sval[base + (j0 + jj) * 64 + lane] = tv[lane][jj];
    }
    __syncthreads();
  }
}

// One wave per slice, one lane per row, entries in CSR order (the lockstep kernel's order and
// operations: bitwise equal to variant 1); entries past a row's end are skipped.  TRF (A only):
// the rows of flat rays gather from x^T -- the 64 parallel rays of a slice then read
// neighbouring pixels at each step for steep and flat angles alike.
template <int U, bool TRF>
__global__ __launch_bounds__(256) void k_spmv_sell(const int64_t* __restrict__ off,
                                                   const int32_t* __restrict__ scol,
                                                   const float* __restrict__ sval,
                                                   const int32_t* __restrict__ len, long rows,
                                                   long slices, long ncols, const float* __restrict__ x,
                                                   float* __restrict__ y, const float* __restrict__ xT,
                                                   const unsigned char* __restrict__ steep, int R, int N,
                                                   double invN) {
  const long slice = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (slice >= slices) return;
  const int lane = threadIdx.x & 63;
  const long row = slice * 64 + lane;
  const int n = row < rows ? len[row] : 0;
  const bool tr = TRF && row < rows && !steep[row / R];
  const float* xb = (tr ? xT : x) + (size_t)blockIdx.y * ncols;
  auto at = [&](int c) -> float {
    if (!tr) return xb[c];
    const int yq = (int)(((double)c + 0.5) * invN);  // c = yq*N + xq, exact (see SpGather)
    return xb[(c - yq * N) * N + yq];
  };
  const long base = off[slice];
  const int L = (int)((off[slice + 1] - base) >> 6);
  const int32_t* c = scol + base + lane;
  const float* v = sval + base + lane;
  float acc = 0.0f;
  int j = 0;
  if (U <= L) {
    // Software-pipelined: the next block's column and value loads are issued after this
    // block's gathers, so they are in flight while the gathered values are summed.  (U = 8
    // measured fastest: A.x 23.5 ms with 8, 26.6 ms with 4, 33.2 ms with 16 at 2048^2.)
    int cc[U];
    float vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cc[u] = __builtin_nontemporal_load(c + 64 * u);
      vv[u] = __builtin_nontemporal_load(v + 64 * u);
    }
    for (; j + U <= L; j += U) {
      float g[U];
#pragma unroll
      for (int u = 0; u < U; ++u) g[u] = at(cc[u]);  // padding columns are 0: in bounds
      const bool more = j + 2 * U <= L;  // uniform over the wave
      int cn[U];
      float vn[U];
      __builtin_amdgcn_sched_barrier(0);
      if (more) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          cn[u] = __builtin_nontemporal_load(c + 64 * (j + U + u));
          vn[u] = __builtin_nontemporal_load(v + 64 * (j + U + u));
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j + u < n) acc = fmaf(vv[u], g[u], acc);
      if (more) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          cc[u] = cn[u];
          vv[u] = vn[u];
        }
      }
    }
  }
  for (; j < L; ++j)
    if (j < n) acc = fmaf(v[64 * j], at(c[64 * j]), acc);
  if (row < rows) y[(size_t)blockIdx.y * rows + row] = acc;
}

// ------------------------------------------------------------------ 16-bit steps (variant 9)
__device__ inline int step_lo(uint32_t w) { return (int)(int16_t)(w & 0xffffu); }
__device__ inline int step_hi(uint32_t w) { return (int)w >> 16; }

// One wave per slice, one lane per row: the SELL-64 columns of a row as 16-bit steps (layout
// at sk_mat::sell_d16).  Sets *bad when a step does not fit.
// pos (B, A^T): columns a*R + r are taken as pos[a]*R + r, their line-table order.
__global__ __launch_bounds__(256) void k_sell_steps(const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ scol,
                                                    const int32_t* __restrict__ len, long rows, long slices,
                                                    const int32_t* __restrict__ pos, int R,
                                                    uint32_t* __restrict__ d16, int32_t* __restrict__ first,
                                                    unsigned* __restrict__ bad) {
  const long slice = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (slice >= slices) return;
  const int lane = threadIdx.x & 63;
  const long row = slice * 64 + lane;
  const int n = row < rows ? len[row] : 0;
  const long base = off[slice];
  const int L = (int)((off[slice + 1] - base) >> 6);  // even
  const int32_t* c = scol + base + lane;
  uint32_t* d = d16 + (base >> 1) + lane;
  auto colm = [&](int j) -> long {
    const long v = c[64 * (long)j];
    if (!pos) return v;
    const long a = v / R;
    return (long)pos[a] * R + (v - a * R);
  };
  long prev = n > 0 ? colm(0) : 0;
  if (row < rows) first[row] = (int32_t)prev;
  bool over = false;
  for (int j = 0; j < L; j += 2) {
    const long c0 = j < n ? colm(j) : prev;
    const long c1 = j + 1 < n ? colm(j + 1) : c0;
    const long d0 = c0 - prev, d1 = c1 - c0;
    over = over || d0 < -32768 || d0 > 32767 || d1 < -32768 || d1 > 32767;
    d[64 * (long)(j >> 1)] = (uint32_t)(uint16_t)(int16_t)d0 | ((uint32_t)(uint16_t)(int16_t)d1 << 16);
    prev = c1;
  }
  if (over) atomicOr(bad, 1u);
}

// As k_spmv_sell<U, false> (the same entries, gathers and FMAs in the same order: bitwise
// equal to variants 7 and 1), with the columns rebuilt from the 16-bit steps by a running sum
// per lane.  Padding steps are 0, so its gathers stay on the row's last column: in bounds.
template <int U>
__global__ __launch_bounds__(256) void k_spmv_sell16(const int64_t* __restrict__ off,
                                                     const uint32_t* __restrict__ d16,
                                                     const float* __restrict__ sval,
                                                     const int32_t* __restrict__ len,
                                                     const int32_t* __restrict__ first, long rows, long slices,
                                                     long ncols, const float* __restrict__ x,
                                                     float* __restrict__ y) {
  static_assert(U % 2 == 0, "two entries per step word");
  constexpr int H = U / 2;
  const long slice = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (slice >= slices) return;
  const int lane = threadIdx.x & 63;
  const long row = slice * 64 + lane;
  const int n = row < rows ? len[row] : 0;
  const float* xb = x + (size_t)blockIdx.y * ncols;
  const long base = off[slice];
  const int L = (int)((off[slice + 1] - base) >> 6);  // even
  const uint32_t* dp = d16 + (base >> 1) + lane;
  const float* v = sval + base + lane;
  int cur = row < rows ? first[row] : 0;
  float acc = 0.0f;
  int j = 0;
  if (U <= L) {
    uint32_t ww[H];
    float vv[U];
#pragma unroll
    for (int h = 0; h < H; ++h) ww[h] = __builtin_nontemporal_load(dp + 64 * h);
#pragma unroll
    for (int u = 0; u < U; ++u) vv[u] = __builtin_nontemporal_load(v + 64 * u);
    for (; j + U <= L; j += U) {
      float g[U];
#pragma unroll
      for (int h = 0; h < H; ++h) {
        cur += step_lo(ww[h]);
        g[2 * h] = xb[cur];
        cur += step_hi(ww[h]);
        g[2 * h + 1] = xb[cur];
      }
      const bool more = j + 2 * U <= L;  // uniform over the wave
      uint32_t wn[H];
      float vn[U];
      __builtin_amdgcn_sched_barrier(0);
      if (more) {
#pragma unroll
        for (int h = 0; h < H; ++h) wn[h] = __builtin_nontemporal_load(dp + 64 * ((j + U) / 2 + h));
#pragma unroll
        for (int u = 0; u < U; ++u) vn[u] = __builtin_nontemporal_load(v + 64 * (j + U + u));
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j + u < n) acc = fmaf(vv[u], g[u], acc);
      if (more) {
#pragma unroll
        for (int h = 0; h < H; ++h) ww[h] = wn[h];
#pragma unroll
        for (int u = 0; u < U; ++u) vv[u] = vn[u];
      }
    }
  }
  for (; j < L; j += 2) {
    const uint32_t w = dp[64 * (j >> 1)];
    cur += step_lo(w);
    const float g0 = xb[cur];
    cur += step_hi(w);
    const float g1 = xb[cur];
    if (j < n) acc = fmaf(v[64 * j], g0, acc);
    if (j + 1 < n) acc = fmaf(v[64 * (j + 1)], g1, acc);
  }
  if (row < rows) y[(size_t)blockIdx.y * rows + row] = acc;
}

// out row b*A + j = in row b*A + orig[j] (R floats each): sinograms into line-table order.
// One block per row.
__global__ __launch_bounds__(256) void k_rows_permute(const float* __restrict__ in, const int32_t* __restrict__ orig,
                                                      int A, int R, float* __restrict__ out) {
  const long row = blockIdx.x;
  const long b = row / A;
  const int j = (int)(row - b * A);
  const float* src = in + (b * A + orig[j]) * (long)R;
  float* dst = out + row * (long)R;
  for (int r = threadIdx.x; r < R; r += 256) dst[r] = src[r];
}

// counts -> rowptr (rows + 1 entries).  Returns nnz via host sync.
static int scan_counts(Ctx* c, const int32_t* counts, long rows, int64_t* rowptr, int64_t* nnz,
                       hipStream_t s) {
  const long ntiles = (rows + SCAN_TILE - 1) / SCAN_TILE;
  void* ws = nullptr;
  int rc = scratch(c, sizeof(int64_t) * (size_t)(ntiles + 1), s, &ws);
  if (rc) return rc;
  int64_t* sums = (int64_t*)ws;
  int64_t* grand = sums + ntiles;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(256), 0, s, counts, rows, rowptr, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, sums, ntiles, grand);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, rowptr, rows,
                     sums, grand);
  rc = launched(c, "scan kernels");
  if (rc) return rc;
  SK_HIP(c, hipMemcpyAsync(nnz, rowptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  SK_HIP(c, hipStreamSynchronize(s));
  return SK_OK;
}

// Lockstep SpMV: one lane per row, the 64 rows of a wave advance through their entries
// together.  Adjacent rows of A are rays of adjacent bins of one angle and adjacent rows
// of B are neighbouring pixels, so at each step the wave's x gathers fall on neighbouring
// addresses; each lane streams its own row's columns and values, summed in CSR order.
__global__ __launch_bounds__(256) void k_spmv_lockstep(const int64_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const float* __restrict__ val, long rows,
                                                       long ncols, const float* __restrict__ x,
                                                       float* __restrict__ y) {
  const long row = (long)blockIdx.x * 256 + threadIdx.x;
  const float* xb = x + (size_t)blockIdx.y * ncols;
  long beg = 0, end = 0;
  if (row < rows) {
    beg = rowptr[row];
    end = rowptr[row + 1];
  }
  const int len = (int)(end - beg);
  int maxlen = len;
  for (int off = 32; off > 0; off >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, off));
  float acc = 0.0f;
  int k = 0;
  for (; k + 4 <= maxlen; k += 4) {
    float v[4];
    int c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool ok = k + u < len;
      c[u] = ok ? col[beg + k + u] : 0;
      v[u] = ok ? val[beg + k + u] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc = fmaf(v[u], xb[c[u]], acc);
  }
  for (; k < maxlen; ++k)
    if (k < len) acc = fmaf(val[beg + k], xb[col[beg + k]], acc);
  if (row < rows) y[(size_t)blockIdx.y * rows + row] = acc;
}

// ------------------------------------------------------------------ row slots (variant 10, A)
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// The row step of one ray (k_fp_rows's expressions: forwardprojection.fut:31-54).
struct RsStep {
  int Xb, Xt;
  bool inb, intop;
  float lb, lt;
};
__device__ __forceinline__ RsStep rs_step(int i, int h, float slope, float icpt, float lbase, bool live) {
  RsStep t;
  const float xbot = (float)i * slope + icpt;
  const float xtop = xbot + slope;
  t.Xb = (int)floorf(xbot);
  t.Xt = (int)floorf(xtop);
  const bool single = t.Xb == t.Xt;
  t.inb = live && t.Xb >= -h && t.Xb < h;
  t.intop = live && !single && t.Xt >= -h && t.Xt < h;
  const float Xmax = (float)(t.Xb > t.Xt ? t.Xb : t.Xt);
  const float xdiff = xtop - xbot;
  t.lb = single ? lbase : ((Xmax - xbot) / xdiff) * lbase;
  t.lt = ((xtop - Xmax) / xdiff) * lbase;
  return t;
}

// One wave per slice: the frame rows [lo, lo + len) holding an entry of any of its rays (len a
// multiple of 8; 0 for a slice without entries).
__global__ __launch_bounds__(SP_BLOCK) void k_rs_count(const Line* __restrict__ lines, const int32_t* __restrict__ lpos,
                                                       int A, int R, int N, float rho0, float drho, int nbb,
                                                       int32_t* __restrict__ lo_out, int32_t* __restrict__ len_out) {
  const long s = (long)blockIdx.x * (SP_BLOCK / 64) + (threadIdx.x >> 6);
  if (s >= (long)A * nbb) return;  // whole wave
  const int lane = lane_id();
  const int a = (int)(s / nbb), r = (int)(s - (long)a * nbb) * 64 + lane;
  const Line L = lines[lpos[a]];
  const int h = N >> 1;
  int first = 1 << 30, last = -(1 << 30);
  if (r < R) {
    const float icpt = (rho0 + (float)r * drho) / L.cost;  // forwardprojection.fut:33-34
    for (int i = -h; i < h; ++i) {
      const RsStep t = rs_step(i, h, L.slope, icpt, L.lbase, true);
      if (t.inb || t.intop) {
        first = min(first, i);
        last = i;
      }
    }
  }
  first = wave_min(first);
  last = wave_max(last);
  if (lane == 0) {
    lo_out[s] = first <= last ? first : 0;
    len_out[s] = first <= last ? ((last - first + 1 + 7) & ~7) : 0;
  }
}

// One wave per slice: fill its records (lane = ray).  Sets *bad when a nibble does not fit.
__global__ __launch_bounds__(SP_BLOCK) void k_rs_fill(const Line* __restrict__ lines, const int32_t* __restrict__ lpos,
                                                      int A, int R, int N, float rho0, float drho, int nbb,
                                                      const int32_t* __restrict__ lo, const int64_t* __restrict__ off,
                                                      uint32_t* __restrict__ ctrl, float2* __restrict__ val,
                                                      int32_t* __restrict__ first, unsigned* __restrict__ bad) {
  const long s = (long)blockIdx.x * (SP_BLOCK / 64) + (threadIdx.x >> 6);
  if (s >= (long)A * nbb) return;
  const int lane = lane_id();
  const int a = (int)(s / nbb), r = (int)(s - (long)a * nbb) * 64 + lane;
  const Line L = lines[lpos[a]];
  const int h = N >> 1;
  const long o = off[s];
  const int len = (int)(off[s + 1] - o);
  const int i0 = lo[s];
  const bool valid = r < R;
  const float icpt = valid ? (rho0 + (float)r * drho) / L.cost : 0.0f;
  const float nan = __builtin_nanf("");
  bool over = false;
  int prev = 0;
  uint32_t word = 0;
  for (int k = 0; k < len; ++k) {
    const int i = i0 + k;
    int dX = 0, ot = 0;
    float vb = nan, vt = nan;
    if (valid) {
      const RsStep t = rs_step(i, h, L.slope, icpt, L.lbase, i < h);
      if (k == 0) first[s * 64 + lane] = (i + h) * N + (t.Xb + h) - N;
      dX = k == 0 ? 0 : t.Xb - prev;
      ot = t.Xt - t.Xb;
      prev = t.Xb;
      if (t.inb) vb = t.lb;
      if (t.intop) vt = t.lt;
      over = over || dX < -8 || dX > 7 || ot < -8 || ot > 7;
    } else if (k == 0) {
      first[s * 64 + lane] = -N;
    }
    word |= ((uint32_t)(dX & 15) | ((uint32_t)(ot & 15) << 4)) << (8 * (k & 3));
    if ((k & 3) == 3) {
      ctrl[(o / 4 + k / 4) * 64 + lane] = word;
      word = 0;
    }
    val[(o + k) * 64 + lane] = make_float2(vb, vt);
  }
  if (len == 0) first[s * 64 + lane] = -N;
  if (over) atomicOr(bad, 1u);
}

__device__ __forceinline__ int nib(uint32_t w, int sh) { return ((int)(w << (28 - sh))) >> 28; }
typedef float vf2 __attribute__((ext_vector_type(2)));

// y = A x on the row slots: one wave per slice, one ray per lane, records in row order, the
// entries of a record in the order Xb, Xt -- the CSR order, so the sums are bitwise those of
// variants 1, 7 and 9.  Gathers through a buffer descriptor over the frame image: frame
// indices outside it (a hole's neighbour) read 0.  Eight records per step: the next step's
// control words are loaded first and its values after them, while this step's sixteen gathers
// are summed.
__global__ __launch_bounds__(256) void k_spmv_rs(const int64_t* __restrict__ off, const uint32_t* __restrict__ ctrl,
                                                 const float2* __restrict__ val, const int32_t* __restrict__ first,
                                                 const unsigned char* __restrict__ steep, long slices, int nbb, int R,
                                                 int N, long ncols, long nrows, const float* __restrict__ x,
                                                 const float* __restrict__ xT, float* __restrict__ y) {
  // the slice, its angle and frame are wave-uniform (scalar registers: the buffer descriptor
  // below must be uniform)
  const long s = (long)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (s >= slices) return;
  const int lane = threadIdx.x & 63;
  const int a = (int)(s / nbb), r = (int)(s - (long)a * nbb) * 64 + lane;
  const int st = __builtin_amdgcn_readfirstlane((int)steep[a]);
  const float* src = (st ? x : xT) + (size_t)blockIdx.y * ncols;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)src, 0, (int)(ncols * 4), 0x00020000u);
  const long o = off[s];
  const int len = (int)(off[s + 1] - o);  // a multiple of 8
  const uint32_t* cp = ctrl + (o / 4) * 64 + lane;
  const vf2* vp = (const vf2*)(val + o * 64 + lane);
  int cur = first[s * 64 + lane];
  float acc = 0.0f;
  if (len > 0) {
    uint32_t w[2];
    vf2 v[8];
#pragma unroll
    for (int h = 0; h < 2; ++h) w[h] = __builtin_nontemporal_load(cp + 64 * h);
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(vp + 64 * u);
    for (int k = 0; k < len; k += 8) {
      float gb[8], gt[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t ww = w[u >> 2];
        const int sh = 8 * (u & 3);
        cur += N + nib(ww, sh);
        const int ct = cur + nib(ww, sh + 4);
        gb[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, cur << 2, 0, 0));
        gt[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, ct << 2, 0, 0));
      }
      const bool more = k + 8 < len;  // uniform over the wave
      uint32_t wn[2];
      vf2 vn[8];
      __builtin_amdgcn_sched_barrier(0);
      if (more) {
#pragma unroll
        for (int h = 0; h < 2; ++h) wn[h] = __builtin_nontemporal_load(cp + 64 * ((k + 8) / 4 + h));
#pragma unroll
        for (int u = 0; u < 8; ++u) vn[u] = __builtin_nontemporal_load(vp + 64 * (k + 8 + u));
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (v[u].x == v[u].x) acc = fmaf(v[u].x, gb[u], acc);
        if (v[u].y == v[u].y) acc = fmaf(v[u].y, gt[u], acc);
      }
      if (more) {
#pragma unroll
        for (int h = 0; h < 2; ++h) w[h] = wn[h];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = vn[u];
      }
    }
  }
  if (r < R) y[(size_t)blockIdx.y * nrows + (size_t)a * R + r] = acc;
}

// ------------------------------------------------------------------ row slots for B (variant 10)
// One wave per slice (64 pixels of an image row), lanes = pixels, records = lines in table order:
// the two candidates of backprojection.fut:44-65 with k_bp_rows's expressions (so the values and
// the member test are the CSR's).  Sets *bad when a column step does not fit in 16 bits.
__global__ __launch_bounds__(SP_BLOCK) void k_rsb_fill(const Line* __restrict__ lines, int A, int A8, int R, int N,
                                                       float rho0, float drho, uint32_t* __restrict__ ctrl,
                                                       float2* __restrict__ val, int32_t* __restrict__ first,
                                                       unsigned* __restrict__ bad) {
  const int nxb = (N + 63) / 64;
  const long sl = (long)blockIdx.x * (SP_BLOCK / 64) + (threadIdx.x >> 6);
  if (sl >= (long)N * nxb) return;
  const int lane = lane_id();
  const int Y = (int)(sl / nxb), X = (int)(sl - (long)Y * nxb) * 64 + lane;
  const bool valid = X < N;
  const int h = N >> 1;
  const long o = sl * A8;
  const float nan = __builtin_nanf("");
  bool over = false;
  long prev = 0;
  uint32_t word = 0;
  for (int j = 0; j < A8; ++j) {
    long step = 0;
    float v0 = nan, v1 = nan;
    if (valid && j < A) {
      const Line L = lines[j];
      const bool fl = L.flags & LINE_FLAT;
      const float xmin = (float)((fl ? Y : X) - h);
      const float ymin = (float)((fl ? X : Y) - h);
      // backprojection.fut:50-54
      const float px = (xmin + 0.5f) - L.fc;
      const float py = (ymin + 0.5f) - L.fs;
      const float rho = L.cost * px + L.sint * py;
      const float sf = ceilf((rho - rho0) / drho);
      const float xbase = ymin * L.slope;
      float lk[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        // backprojection.fut:56-64
        const float sp = sf + (float)k;
        const float rr = sp * drho + rho0;
        const float xbot = xbase + (rr / L.cost);
        const float xtop = xbot + L.slope;
        const float maxx = fmaxf(xbot, xtop), minx = fminf(xbot, xtop);
        const float lo = fmaxf(maxx - xmin, 0.0f);
        const float hi = fmaxf((xmin + 1.0f) - minx, 0.0f);
        const float diff = maxx - minx;
        const float m = fminf(lo, hi);
        const float p = (m == 0.0f) ? 0.0f : m / diff;
        const float l = fminf(p, 1.0f) * L.lbase;
        const int si = (int)sp;
        lk[k] = (si >= 0 && si < R && l != 0.0f) ? l : nan;
      }
      v0 = lk[0];
      v1 = lk[1];
      const long col = (long)j * R + (long)(int)sf;
      if (j == 0) first[sl * 64 + lane] = (int32_t)(col - R);
      step = j == 0 ? R : col - prev;
      prev = col;
      over = over || step < -32768 || step > 32767;
    } else if (j == 0) {
      first[sl * 64 + lane] = 0;
    }
    word |= (uint32_t)(uint16_t)(int16_t)step << (16 * (j & 1));
    if (j & 1) {
      ctrl[(o / 2 + j / 2) * 64 + lane] = word;
      word = 0;
    }
    val[(o + j) * 64 + lane] = make_float2(v0, v1);
  }
  if (over) atomicOr(bad, 1u);
}

// y = B r on the row slots of B: one wave per slice of 64 pixels, records in line-table order
// (the CSR order: bitwise equal to variants 1, 7 and 9).  r is the input with its rows in
// line-table order; gathers through a buffer descriptor over it (out-of-range columns read 0).
__global__ __launch_bounds__(256) void k_spmv_rsb(const uint32_t* __restrict__ ctrl, const float2* __restrict__ val,
                                                  const int32_t* __restrict__ first, long slices, int A8, int N,
                                                  long ncols, long nrows, const float* __restrict__ in,
                                                  float* __restrict__ y) {
  const long sl = (long)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (sl >= slices) return;
  const int lane = threadIdx.x & 63;
  const int nxb = (N + 63) / 64;
  const int Y = (int)(sl / nxb), X = (int)(sl - (long)Y * nxb) * 64 + lane;
  const float* src = in + (size_t)blockIdx.y * ncols;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)src, 0, (int)(ncols * 4), 0x00020000u);
  const long o = sl * A8;
  const uint32_t* cp = ctrl + (o / 2) * 64 + lane;
  const vf2* vp = (const vf2*)(val + o * 64 + lane);
  int cur = first[sl * 64 + lane];
  float acc = 0.0f;
  uint32_t w[4];
  vf2 v[8];
#pragma unroll
  for (int hh = 0; hh < 4; ++hh) w[hh] = __builtin_nontemporal_load(cp + 64 * hh);
#pragma unroll
  for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(vp + 64 * u);
  for (int j = 0; j < A8; j += 8) {
    float g0[8], g1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t ww = w[u >> 1];
      cur += (u & 1) ? ((int)ww >> 16) : (int)(int16_t)(ww & 0xffffu);
      g0[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, cur << 2, 0, 0));
      g1[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (cur + 1) << 2, 0, 0));
    }
    const bool more = j + 8 < A8;  // uniform
    uint32_t wn[4];
    vf2 vn[8];
    __builtin_amdgcn_sched_barrier(0);
    if (more) {
#pragma unroll
      for (int hh = 0; hh < 4; ++hh) wn[hh] = __builtin_nontemporal_load(cp + 64 * ((j + 8) / 2 + hh));
#pragma unroll
      for (int u = 0; u < 8; ++u) vn[u] = __builtin_nontemporal_load(vp + 64 * (j + 8 + u));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (v[u].x == v[u].x) acc = fmaf(v[u].x, g0[u], acc);
      if (v[u].y == v[u].y) acc = fmaf(v[u].y, g1[u], acc);
    }
    if (more) {
#pragma unroll
      for (int hh = 0; hh < 4; ++hh) w[hh] = wn[hh];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = vn[u];
    }
  }
  if (X < N) y[(size_t)blockIdx.y * nrows + (size_t)Y * N + X] = acc;
}

static void launch_rows(const Geom& G, int kind, bool fill, int32_t* counts, const int64_t* rowptr,
                        int32_t* cols, float* vals, hipStream_t s) {
  const long rows = kind == SK_MAT_FP ? (long)G.A * G.R : (long)G.N * G.N;
  const unsigned blocks = (unsigned)((rows + 3) / 4);
  if (kind == SK_MAT_FP) {
    if (fill)
      hipLaunchKernelGGL(k_fp_rows<true>, dim3(blocks), dim3(SP_BLOCK), 0, s, G.d_lines, G.A, G.R, G.N,
                         G.rho0, G.drho, counts, rowptr, cols, vals);
    else
      hipLaunchKernelGGL(k_fp_rows<false>, dim3(blocks), dim3(SP_BLOCK), 0, s, G.d_lines, G.A, G.R, G.N,
                         G.rho0, G.drho, counts, rowptr, cols, vals);
  } else if (kind == SK_MAT_BP) {
    if (fill)
      hipLaunchKernelGGL((k_bp_rows<true, false>), dim3(blocks), dim3(SP_BLOCK), 0, s, G.d_lines, G.A, G.R,
                         G.N, G.rho0, G.drho, G.ncand, counts, rowptr, cols, vals);
    else
      hipLaunchKernelGGL((k_bp_rows<false, false>), dim3(blocks), dim3(SP_BLOCK), 0, s, G.d_lines, G.A, G.R,
                         G.N, G.rho0, G.drho, G.ncand, counts, rowptr, cols, vals);
  } else {
    if (fill)
      hipLaunchKernelGGL((k_bp_rows<true, true>), dim3(blocks), dim3(SP_BLOCK), 0, s, G.d_lines, G.A, G.R,
                         G.N, G.rho0, G.drho, G.ncand, counts, rowptr, cols, vals);
    else
      hipLaunchKernelGGL((k_bp_rows<false, true>), dim3(blocks), dim3(SP_BLOCK), 0, s, G.d_lines, G.A, G.R,
                         G.N, G.rho0, G.drho, G.ncand, counts, rowptr, cols, vals);
  }
}

}  // namespace sk

using namespace sk;

static void mat_free(sk_mat* m) {
  if (!m) return;
  if (m->rowptr) (void)hipFree(m->rowptr);
  if (m->col) (void)hipFree(m->col);
  if (m->val) (void)hipFree(m->val);
  if (m->steep) (void)hipFree(m->steep);
  if (m->sell_off) (void)hipFree(m->sell_off);
  if (m->sell_col) (void)hipFree(m->sell_col);
  if (m->sell_val) (void)hipFree(m->sell_val);
  if (m->sell_len) (void)hipFree(m->sell_len);
  if (m->sell_d16) (void)hipFree(m->sell_d16);
  if (m->sell_first) (void)hipFree(m->sell_first);
  if (m->line_pos) (void)hipFree(m->line_pos);
  if (m->rs_off) (void)hipFree(m->rs_off);
  if (m->rs_ctrl) (void)hipFree(m->rs_ctrl);
  if (m->rs_val) (void)hipFree(m->rs_val);
  if (m->rs_first) (void)hipFree(m->rs_first);
  if (m->rs_lpos) (void)hipFree(m->rs_lpos);
  if (m->rs_lines) (void)hipFree(m->rs_lines);
  delete m;
}

// The SELL-64 copy of a matrix (variant 7).  With `required` false it is built only if it
// takes at most half the free device memory (the autotuner's candidate); otherwise it is
// marked unavailable.  Synchronizes `s` (sizes).
static int build_sell(Ctx* c, sk_mat* m, bool required, hipStream_t s) {
  if (m->sell_col || (m->sell_failed && !required)) return SK_OK;
  const long slices = (m->rows + 63) / 64;
  int32_t* sizes = nullptr;
  SK_HIP(c, hipMalloc(&sizes, sizeof(int32_t) * (size_t)slices));
  hipError_t e = hipSuccess;
  if (!m->sell_len) e = hipMalloc(&m->sell_len, sizeof(int32_t) * (size_t)m->rows);
  if (e == hipSuccess && !m->sell_off) e = hipMalloc(&m->sell_off, sizeof(int64_t) * (size_t)(slices + 1));
  if (e != hipSuccess) {
    (void)hipFree(sizes);
    return hip_fail(c, e, "SELL index allocation");
  }
  hipLaunchKernelGGL(k_sell_sizes, dim3((unsigned)((m->rows + 255) / 256)), dim3(256), 0, s, m->rowptr,
                     (long)m->rows, sizes, m->sell_len);
  int rc = launched(c, "SELL size kernel");
  int64_t total = 0;
  if (!rc) rc = scan_counts(c, sizes, slices, m->sell_off, &total, s);
  (void)hipFree(sizes);
  if (rc) return rc;
  size_t free_b = 0, total_b = 0;
  SK_HIP(c, hipMemGetInfo(&free_b, &total_b));
  const size_t need = (size_t)total * (sizeof(int32_t) + sizeof(float));
  if (!required && need > free_b / 2) {
    m->sell_failed = true;
    return SK_OK;
  }
  e = hipMalloc(&m->sell_col, sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1));
  if (e == hipSuccess) e = hipMalloc(&m->sell_val, sizeof(float) * (size_t)std::max<int64_t>(total, 1));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (m->sell_col) (void)hipFree(m->sell_col);
    m->sell_col = nullptr;
    m->sell_val = nullptr;
    m->sell_failed = true;
    return required ? hip_fail(c, e, "SELL copy allocation") : SK_OK;
  }
  if (getenv("SYNKRO_VERBOSE"))
    fprintf(stderr, "sk_spmv: SELL-64 copy, %lld entries for %lld nonzeros (%.1f%% padding)\n", (long long)total,
            (long long)m->nnz, 100.0 * (double)(total - m->nnz) / (double)std::max<int64_t>(m->nnz, 1));
  hipLaunchKernelGGL(k_sell_fill, dim3((unsigned)slices), dim3(64), 0, s, m->rowptr, m->col, m->val,
                     (long)m->rows, m->sell_off, m->sell_col, m->sell_val);
  return launched(c, "SELL fill kernel");
}

// The 16-bit steps of variant 9, from the SELL-64 copy (built first if needed).  With
// `required` false it is built only if it takes at most half the free device memory.  A matrix
// with a step beyond 16 bits (consecutive columns of a row more than 32767 apart) marks it
// unavailable.  Synchronizes `s`.
static int build_d16(Ctx* c, sk_mat* m, bool required, hipStream_t s) {
  if (m->sell_d16) return SK_OK;
  auto unavailable = [&]() {
    m->d16_failed = true;
    return required ? fail(c, SK_EINVAL,
                           "SpMV variant 9 unavailable: a row steps more than 32767 columns (or no memory)")
                    : SK_OK;
  };
  if (m->d16_failed) return unavailable();
  int rc = build_sell(c, m, required, s);
  if (rc) return rc;
  if (!m->sell_col) return unavailable();
  const long slices = (m->rows + 63) / 64;
  int64_t total = 0;
  SK_HIP(c, hipMemcpyAsync(&total, m->sell_off + slices, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  SK_HIP(c, hipStreamSynchronize(s));
  size_t free_b = 0, total_b = 0;
  SK_HIP(c, hipMemGetInfo(&free_b, &total_b));
  const size_t need = sizeof(uint16_t) * (size_t)total + sizeof(int32_t) * (size_t)m->rows;
  if (!required && need > free_b / 2) return unavailable();
  unsigned* bad = nullptr;
  hipError_t e = hipMalloc(&m->sell_d16, sizeof(uint16_t) * (size_t)std::max<int64_t>(total, 2));
  if (e == hipSuccess) e = hipMalloc(&m->sell_first, sizeof(int32_t) * (size_t)m->rows);
  if (e == hipSuccess) e = hipMalloc(&bad, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(unsigned), s);
  unsigned hbad = 1;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sell_steps, dim3((unsigned)((slices + 3) / 4)), dim3(256), 0, s, m->sell_off,
                       m->sell_col, m->sell_len, (long)m->rows, slices, m->line_pos, m->R, m->sell_d16,
                       m->sell_first, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (bad) (void)hipFree(bad);
  if (e != hipSuccess || hbad) {
    if (e != hipSuccess) (void)hipGetLastError();
    if (m->sell_d16) (void)hipFree(m->sell_d16);
    if (m->sell_first) (void)hipFree(m->sell_first);
    m->sell_d16 = nullptr;
    m->sell_first = nullptr;
    if (e != hipSuccess && required) {
      m->d16_failed = true;
      return hip_fail(c, e, "SpMV step copy");
    }
    return unavailable();
  }
  return SK_OK;
}

// The row slots of variant 10 (SK_MAT_FP), built from the geometry's line table with the
// expressions of the CSR fill (so every value is bitwise the CSR's).  With `required` false it
// is built only if it takes at most half the free device memory.  Synchronizes `s`.
static int build_rs(Ctx* c, sk_mat* m, bool required, hipStream_t s) {
  if (m->rs_val) return SK_OK;
  auto unavailable = [&](const char* why) {
    m->rs_failed = true;
    return required ? fail(c, SK_EINVAL, "SpMV variant 10 unavailable: %s", why) : SK_OK;
  };
  if (m->kind == SK_MAT_ADJ || !m->rs_lines) return unavailable("row slots exist for A and B only");
  if (m->rs_failed) return unavailable("a step does not fit its field (or no memory)");
  if (m->kind == SK_MAT_BP) {
    // B: a fixed A8 records (lines) per slice of 64 pixels, no count pass
    if (m->ncand != 2) return unavailable("B's row slots hold two candidates per line (deltarho in [0.71, 1.41))");
    const int nxb = (m->N + 63) / 64;
    const long slices = (long)m->N * nxb;
    const long A8 = (m->A + 7) & ~7;
    size_t free_b = 0, total_b = 0;
    SK_HIP(c, hipMemGetInfo(&free_b, &total_b));
    const size_t recs = (size_t)slices * A8 * 64;
    const size_t need = recs * (sizeof(float2) + 2) + sizeof(int32_t) * 64 * (size_t)slices;
    if (!required && need > free_b / 2) return unavailable("no room for the copy");
    unsigned* bad = nullptr;
    hipError_t e = hipMalloc(&m->rs_val, sizeof(float2) * recs);
    if (e == hipSuccess) e = hipMalloc(&m->rs_ctrl, sizeof(uint32_t) * (recs / 2));
    if (e == hipSuccess) e = hipMalloc(&m->rs_first, sizeof(int32_t) * 64 * (size_t)slices);
    if (e == hipSuccess) e = hipMalloc(&bad, sizeof(unsigned));
    if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(unsigned), s);
    unsigned hbad = 1;
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_rsb_fill, dim3((unsigned)((slices + 3) / 4)), dim3(SP_BLOCK), 0, s, m->rs_lines, m->A,
                         (int)A8, m->R, m->N, m->rho0, m->drho, m->rs_ctrl, m->rs_val, m->rs_first, bad);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (bad) (void)hipFree(bad);
    if (e != hipSuccess || hbad) {
      if (e != hipSuccess) (void)hipGetLastError();
      for (void* q : {(void*)m->rs_val, (void*)m->rs_ctrl, (void*)m->rs_first})
        if (q) (void)hipFree(q);
      m->rs_val = nullptr;
      m->rs_ctrl = nullptr;
      m->rs_first = nullptr;
      if (e != hipSuccess && required) return hip_fail(c, e, "row-slot copy");
      return unavailable(e != hipSuccess ? "no memory for the copy" : "a column step does not fit 16 bits");
    }
    return SK_OK;
  }
  const int nbb = (m->R + 63) / 64;
  const long slices = (long)m->A * nbb;
  const unsigned blocks = (unsigned)((slices + 3) / 4);
  int32_t *lo = nullptr, *len = nullptr;
  SK_HIP(c, hipMalloc(&lo, sizeof(int32_t) * (size_t)slices));
  hipError_t e = hipMalloc(&len, sizeof(int32_t) * (size_t)slices);
  if (e == hipSuccess && !m->rs_off) e = hipMalloc(&m->rs_off, sizeof(int64_t) * (size_t)(slices + 1));
  if (e != hipSuccess) {
    (void)hipFree(lo);
    if (len) (void)hipFree(len);
    return hip_fail(c, e, "row-slot index allocation");
  }
  hipLaunchKernelGGL(k_rs_count, dim3(blocks), dim3(SP_BLOCK), 0, s, m->rs_lines, m->rs_lpos, m->A, m->R, m->N,
                     m->rho0, m->drho, nbb, lo, len);
  int rc = launched(c, "row-slot count kernel");
  int64_t total = 0;
  if (!rc) rc = scan_counts(c, len, slices, m->rs_off, &total, s);
  (void)hipFree(len);
  if (rc) {
    (void)hipFree(lo);
    return rc;
  }
  size_t free_b = 0, total_b = 0;
  e = hipMemGetInfo(&free_b, &total_b);
  const size_t need = (size_t)total * 64 * (sizeof(float2) + 1) + sizeof(int32_t) * 64 * (size_t)slices;
  if (e != hipSuccess || (!required && need > free_b / 2)) {
    (void)hipFree(lo);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemGetInfo");
    return unavailable("no room for the copy");
  }
  unsigned* bad = nullptr;
  e = hipMalloc(&m->rs_val, sizeof(float2) * 64 * (size_t)std::max<int64_t>(total, 4));
  if (e == hipSuccess) e = hipMalloc(&m->rs_ctrl, sizeof(uint32_t) * 16 * (size_t)std::max<int64_t>(total, 4));
  if (e == hipSuccess) e = hipMalloc(&m->rs_first, sizeof(int32_t) * 64 * (size_t)slices);
  if (e == hipSuccess) e = hipMalloc(&bad, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(unsigned), s);
  unsigned hbad = 1;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rs_fill, dim3(blocks), dim3(SP_BLOCK), 0, s, m->rs_lines, m->rs_lpos, m->A, m->R, m->N,
                       m->rho0, m->drho, nbb, lo, m->rs_off, m->rs_ctrl, m->rs_val, m->rs_first, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(lo);
  if (bad) (void)hipFree(bad);
  if (e != hipSuccess || hbad) {
    if (e != hipSuccess) (void)hipGetLastError();
    for (void* p : {(void*)m->rs_val, (void*)m->rs_ctrl, (void*)m->rs_first})
      if (p) (void)hipFree(p);
    m->rs_val = nullptr;
    m->rs_ctrl = nullptr;
    m->rs_first = nullptr;
    if (e != hipSuccess && required) return hip_fail(c, e, "row-slot copy");
    return unavailable(e != hipSuccess ? "no memory for the copy" : "a row step does not fit its 4-bit field");
  }
  if (getenv("SYNKRO_VERBOSE"))
    fprintf(stderr, "sk_spmv: row slots, %lld records of 64 lanes for %lld nonzeros\n", (long long)total,
            (long long)m->nnz);
  return SK_OK;
}

extern "C" {

int sk_sparse_build(sk_ctx* ctx, sk_geom* g, int which, sk_mat** out, void* stream) {
  if (!ctx || !out) return SK_EINVAL;
  Ctx* c = &ctx->c;
  *out = nullptr;
  if (!g) return fail(c, SK_EINVAL, "geometry is NULL");
  if (which != SK_MAT_FP && which != SK_MAT_BP && which != SK_MAT_ADJ)
    return fail(c, SK_EINVAL, "unknown matrix kind %d", which);
  const Geom& G = g->g;
  if ((long)G.A * G.R >= (1L << 31) || (long)G.N * G.N >= (1L << 31))
    return fail(c, SK_EINVAL, "matrix dimensions exceed 32-bit column indices");
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  sk_mat* m = new sk_mat();
  m->kind = which;
  m->device = c->device;
  m->rows = which == SK_MAT_FP ? (long)G.A * G.R : (long)G.N * G.N;
  m->cols = which == SK_MAT_FP ? (long)G.N * G.N : (long)G.A * G.R;
  m->R = G.R;
  m->N = G.N;
  hipError_t e = hipSuccess;
  if (which == SK_MAT_FP) {
    std::vector<unsigned char> st((size_t)G.A, 0);
    for (const Line& L : G.lines) st[(size_t)L.orig] = (L.flags & LINE_FLAT) ? 0 : 1;
    e = hipMalloc(&m->steep, st.size());
    if (e == hipSuccess) e = hipMemcpy(m->steep, st.data(), st.size(), hipMemcpyHostToDevice);
    std::vector<int32_t> lp((size_t)G.A, 0);
    for (size_t j = 0; j < G.lines.size(); ++j) lp[(size_t)G.lines[j].orig] = (int32_t)j;
    if (e == hipSuccess) e = hipMalloc(&m->rs_lpos, sizeof(int32_t) * lp.size());
    if (e == hipSuccess) e = hipMemcpy(m->rs_lpos, lp.data(), sizeof(int32_t) * lp.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&m->rs_lines, sizeof(Line) * (size_t)G.A);
    if (e == hipSuccess) e = hipMemcpy(m->rs_lines, G.lines.data(), sizeof(Line) * (size_t)G.A, hipMemcpyHostToDevice);
    m->rho0 = G.rho0;
    m->drho = G.drho;
    if (e != hipSuccess) {
      mat_free(m);
      return hip_fail(c, e, "steep-angle table");
    }
  } else {
    std::vector<int32_t> t(2 * (size_t)G.A, 0);
    for (size_t j = 0; j < G.lines.size(); ++j) {
      t[(size_t)G.lines[j].orig] = (int32_t)j;
      t[(size_t)G.A + j] = G.lines[j].orig;
    }
    e = hipMalloc(&m->line_pos, sizeof(int32_t) * t.size());
    if (e == hipSuccess) e = hipMemcpy(m->line_pos, t.data(), sizeof(int32_t) * t.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && which == SK_MAT_BP) e = hipMalloc(&m->rs_lines, sizeof(Line) * (size_t)G.A);
    if (e == hipSuccess && which == SK_MAT_BP)
      e = hipMemcpy(m->rs_lines, G.lines.data(), sizeof(Line) * (size_t)G.A, hipMemcpyHostToDevice);
    m->rho0 = G.rho0;
    m->drho = G.drho;
    m->ncand = G.ncand;
    if (e != hipSuccess) {
      mat_free(m);
      return hip_fail(c, e, "line-order table");
    }
    m->line_orig = m->line_pos + G.A;
  }
  m->A = G.A;
  e = hipMalloc(&m->rowptr, sizeof(int64_t) * (size_t)(m->rows + 1));
  if (e != hipSuccess) {
    mat_free(m);
    return hip_fail(c, e, "row pointer allocation");
  }
  int32_t* counts = nullptr;
  e = hipMalloc(&counts, sizeof(int32_t) * (size_t)m->rows);
  if (e != hipSuccess) {
    mat_free(m);
    return hip_fail(c, e, "count allocation");
  }
  launch_rows(G, which, false, counts, nullptr, nullptr, nullptr, s);
  int rc = launched(c, "sparse count kernel");
  int64_t nnz = 0;
  if (!rc) rc = scan_counts(c, counts, m->rows, m->rowptr, &nnz, s);
  (void)hipFree(counts);
  if (rc) {
    mat_free(m);
    return rc;
  }
  m->nnz = nnz;
  if (nnz > 0) {
    e = hipMalloc(&m->col, sizeof(int32_t) * (size_t)nnz);
    if (e == hipSuccess) e = hipMalloc(&m->val, sizeof(float) * (size_t)nnz);
    if (e != hipSuccess) {
      mat_free(m);
      return hip_fail(c, e, "matrix allocation");
    }
    launch_rows(G, which, true, nullptr, m->rowptr, m->col, m->val, s);
    rc = launched(c, "sparse fill kernel");
    if (rc) {
      mat_free(m);
      return rc;
    }
  }
  *out = m;
  return SK_OK;
}

int sk_sparse_count(sk_ctx* ctx, sk_geom* g, int which, int64_t* nnz_out, void* stream) {
  if (!ctx || !nnz_out) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (!g) return fail(c, SK_EINVAL, "geometry is NULL");
  if (which != SK_MAT_FP && which != SK_MAT_BP && which != SK_MAT_ADJ)
    return fail(c, SK_EINVAL, "unknown matrix kind %d", which);
  const Geom& G = g->g;
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const long rows = which == SK_MAT_FP ? (long)G.A * G.R : (long)G.N * G.N;
  int32_t* counts = nullptr;
  int64_t* rowptr = nullptr;
  SK_HIP(c, hipMalloc(&counts, sizeof(int32_t) * (size_t)rows));
  hipError_t e = hipMalloc(&rowptr, sizeof(int64_t) * (size_t)(rows + 1));
  if (e != hipSuccess) {
    (void)hipFree(counts);
    return hip_fail(c, e, "row pointer allocation");
  }
  launch_rows(G, which, false, counts, nullptr, nullptr, nullptr, s);
  int rc = launched(c, "sparse count kernel");
  if (!rc) rc = scan_counts(c, counts, rows, rowptr, nnz_out, s);
  (void)hipFree(counts);
  (void)hipFree(rowptr);
  return rc;
}

int sk_mat_spmv_variant(const sk_mat* m) { return m ? m->spmv_variant : -1; }

int sk_mat_set_spmv_variant(sk_mat* m, int variant) {
  if (!m || variant < -1 || variant > 10 || variant == 5 || variant == 6) return SK_EINVAL;
  if ((variant == 3 || variant == 4 || variant == 8) && !(m->kind == SK_MAT_FP && m->steep)) return SK_EINVAL;
  if (variant == 10 && m->kind == SK_MAT_ADJ) return SK_EINVAL;
  m->spmv_variant = variant;
  return SK_OK;
}

int sk_mat_info(const sk_mat* m, int64_t* rows, int64_t* cols, int64_t* nnz) {
  if (!m) return SK_EINVAL;
  if (rows) *rows = m->rows;
  if (cols) *cols = m->cols;
  if (nnz) *nnz = m->nnz;
  return SK_OK;
}

int sk_mat_arrays(const sk_mat* m, const int64_t** d_rowptr, const int32_t** d_cols,
                  const float** d_vals) {
  if (!m) return SK_EINVAL;
  if (d_rowptr) *d_rowptr = m->rowptr;
  if (d_cols) *d_cols = m->col;
  if (d_vals) *d_vals = m->val;
  return SK_OK;
}

int sk_mat_copy(sk_ctx* ctx, const sk_mat* m, int64_t* d_rowptr, int32_t* d_cols, float* d_vals,
                void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (!m) return fail(c, SK_EINVAL, "matrix is NULL");
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  if (d_rowptr)
    SK_HIP(c, hipMemcpyAsync(d_rowptr, m->rowptr, sizeof(int64_t) * (size_t)(m->rows + 1),
                             hipMemcpyDeviceToDevice, s));
  if (d_cols && m->nnz)
    SK_HIP(c, hipMemcpyAsync(d_cols, m->col, sizeof(int32_t) * (size_t)m->nnz, hipMemcpyDeviceToDevice, s));
  if (d_vals && m->nnz)
    SK_HIP(c, hipMemcpyAsync(d_vals, m->val, sizeof(float) * (size_t)m->nnz, hipMemcpyDeviceToDevice, s));
  return SK_OK;
}

int sk_spmv(sk_ctx* ctx, const sk_mat* m, int batch, const float* d_in, float* d_out, void* stream) {
  if (!ctx) return SK_EINVAL;
  Ctx* c = &ctx->c;
  if (!m) return fail(c, SK_EINVAL, "matrix is NULL");
  if (!d_in || !d_out) return fail(c, SK_EINVAL, "device pointer is NULL");
  if (batch < 1 || batch > 65535) return fail(c, SK_EINVAL, "batch must be in [1, 65535]");
  if (m->device != c->device) return fail(c, SK_EINVAL, "matrix belongs to another device");
  SK_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  // Variants 3, 4 (SK_MAT_FP): the steep rays gather from x^T, a transposed copy of the input
  // images in the context scratch (one transpose per call, ~10 us at 2048^2).
  const bool fpT = m->kind == SK_MAT_FP && m->steep != nullptr;
  float* xT = nullptr;
  auto need_xT = [&]() -> int {
    if (xT) return SK_OK;
    void* ws = nullptr;
    int rc = scratch(c, sizeof(float) * (size_t)batch * m->cols, s, &ws);
    if (rc) return rc;
    xT = (float*)ws;
    return launch_transpose(c, d_in, xT, m->N, batch, s);
  };
  const double invN = m->N > 0 ? 1.0 / (double)m->N : 0.0;
  auto run = [&](int v) -> int {
    dim3 grid((unsigned)((m->rows + 3) / 4), batch);
    if (v == 10 && m->kind == SK_MAT_BP) {
      int rc = build_rs(c, const_cast<sk_mat*>(m), true, s);
      if (rc) return rc;
      void* ws = nullptr;  // the input's rows in line-table order
      rc = scratch(c, sizeof(float) * (size_t)batch * m->cols, s, &ws);
      if (rc) return rc;
      hipLaunchKernelGGL(k_rows_permute, dim3((unsigned)((long)batch * m->A)), dim3(256), 0, s, d_in,
                         m->line_orig, m->A, m->R, (float*)ws);
      const long slices = (long)m->N * ((m->N + 63) / 64);
      hipLaunchKernelGGL(k_spmv_rsb, dim3((unsigned)((slices + 3) / 4), batch), dim3(256), 0, s, m->rs_ctrl,
                         m->rs_val, m->rs_first, slices, (m->A + 7) & ~7, m->N, (long)m->cols, (long)m->rows,
                         (const float*)ws, d_out);
    } else if (v == 10) {
      int rc = build_rs(c, const_cast<sk_mat*>(m), true, s);
      if (rc) return rc;
      rc = need_xT();
      if (rc) return rc;
      const int nbb = (m->R + 63) / 64;
      const long slices = (long)m->A * nbb;
      hipLaunchKernelGGL(k_spmv_rs, dim3((unsigned)((slices + 3) / 4), batch), dim3(256), 0, s, m->rs_off,
                         m->rs_ctrl, m->rs_val, m->rs_first, m->steep, slices, nbb, m->R, m->N, (long)m->cols,
                         (long)m->rows, d_in, xT, d_out);
    } else if (v == 9) {
      int rc = build_d16(c, const_cast<sk_mat*>(m), true, s);
      if (rc) return rc;
      const long slices = (m->rows + 63) / 64;
      const float* in = d_in;
      if (m->line_orig) {  // B, A^T: the input's rows in line-table order
        void* ws = nullptr;
        rc = scratch(c, sizeof(float) * (size_t)batch * m->cols, s, &ws);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rows_permute, dim3((unsigned)((long)batch * m->A)), dim3(256), 0, s, d_in,
                           m->line_orig, m->A, m->R, (float*)ws);
        in = (const float*)ws;
      }
      hipLaunchKernelGGL(k_spmv_sell16<8>, dim3((unsigned)((slices + 3) / 4), batch), dim3(256), 0, s,
                         m->sell_off, m->sell_d16, m->sell_val, m->sell_len, m->sell_first, (long)m->rows,
                         slices, (long)m->cols, in, d_out);
    } else if (v == 7 || v == 8) {
      int rc = build_sell(c, const_cast<sk_mat*>(m), true, s);
      if (rc) return rc;
      const long slices = (m->rows + 63) / 64;
      const dim3 g7((unsigned)((slices + 3) / 4), batch);
      if (v == 7) {
        hipLaunchKernelGGL((k_spmv_sell<8, false>), g7, dim3(256), 0, s, m->sell_off, m->sell_col, m->sell_val,
                           m->sell_len, (long)m->rows, slices, (long)m->cols, d_in, d_out, nullptr, nullptr, 1, 0, 0.0);
      } else {
        rc = need_xT();
        if (rc) return rc;
        hipLaunchKernelGGL((k_spmv_sell<8, true>), g7, dim3(256), 0, s, m->sell_off, m->sell_col, m->sell_val,
                           m->sell_len, (long)m->rows, slices, (long)m->cols, d_in, d_out, xT, m->steep, m->R, m->N,
                           invN);
      }
    } else if (v == 3 || v == 4) {
      int rc = need_xT();
      if (rc) return rc;
      if (v == 3)
        hipLaunchKernelGGL((k_spmv<4, true>), grid, dim3(256), 0, s, m->rowptr, m->col, m->val,
                           (long)m->rows, (long)m->cols, d_in, d_out, xT, m->steep, m->R, m->N, invN);
      else
        hipLaunchKernelGGL((k_spmv_wide<4, true>), grid, dim3(256), 0, s, m->rowptr, m->col, m->val,
                           (long)m->rows, (long)m->cols, d_in, d_out, xT, m->steep, m->R, m->N, invN);
    } else if (v == 2) {
      hipLaunchKernelGGL((k_spmv_wide<4, false>), grid, dim3(256), 0, s, m->rowptr, m->col, m->val,
                         (long)m->rows, (long)m->cols, d_in, d_out, nullptr, nullptr, 1, 0, 0.0);
    } else if (v == 1) {
      dim3 g1((unsigned)((m->rows + 255) / 256), batch);
      hipLaunchKernelGGL(k_spmv_lockstep, g1, dim3(256), 0, s, m->rowptr, m->col, m->val,
                         (long)m->rows, (long)m->cols, d_in, d_out);
    } else {
      hipLaunchKernelGGL((k_spmv<4, false>), grid, dim3(256), 0, s, m->rowptr, m->col, m->val,
                         (long)m->rows, (long)m->cols, d_in, d_out, nullptr, nullptr, 1, 0, 0.0);
    }
    return SK_OK;
  };
  if (m->spmv_variant < 0) {
    // First use: time the wave-per-row kernels once on this matrix and keep the fastest.
    // (The lockstep kernel, variant 1, measured 6-9x slower on both matrices at 2048^2:
    // each of its load instructions touches 64 rows' cache lines.  It is selectable only.)
    // SELL-64 (7) is a candidate when its copy fits in half the free memory (built here,
    // outside the timed region).
    // Its 16-bit-step form (9) likewise, when every step fits.
    // The row slots (10) come first: where memory allows only one copy beside the CSR arrays
    // (2048^2: about 75-90 GB each), it is theirs.
    {
      int rc = m->kind == SK_MAT_ADJ ? SK_OK : build_rs(c, const_cast<sk_mat*>(m), false, s);
      if (!rc) rc = build_sell(c, const_cast<sk_mat*>(m), false, s);
      if (!rc && m->sell_col) rc = build_d16(c, const_cast<sk_mat*>(m), false, s);
      if (rc) return rc;
    }
    const bool sell = m->sell_col != nullptr;
    const bool d16 = m->sell_d16 != nullptr;
    const bool rs = m->rs_val != nullptr;
    // (8 loads in flight per lane instead of 4 measured no faster: 45.8 vs 43.9 ms for A,
    // 63.5 vs 62.6 ms for B at 2048^2.)
    int cand[8] = {0, 2, 3, 4};
    int NV = fpT ? 4 : 2;
    if (sell) cand[NV++] = 7;
    if (sell && fpT) cand[NV++] = 8;
    if (d16) cand[NV++] = 9;
    if (rs) cand[NV++] = 10;
    hipEvent_t e[9];
    for (int v = 0; v <= NV; ++v) SK_HIP(c, hipEventCreate(&e[v]));
    if (fpT) {
      int rc = need_xT();  // outside the timed region of variants 3, 4 (as for every later call)
      if (rc) return rc;
    }
    if ((d16 || rs) && m->line_orig) {  // variants 9, 10: the permuted input's scratch grown before timing
      void* ws = nullptr;
      int rc = scratch(c, sizeof(float) * (size_t)batch * m->cols, s, &ws);
      if (rc) return rc;
    }
    SK_HIP(c, hipEventRecord(e[0], s));
    for (int v = 0; v < NV; ++v) {
      int rc = run(cand[v]);
      if (rc) return rc;
      SK_HIP(c, hipEventRecord(e[v + 1], s));
    }
    SK_HIP(c, hipEventSynchronize(e[NV]));
    int best = cand[0];
    float tbest = 0.0f;
    for (int v = 0; v < NV; ++v) {
      float t = 0.0f;
      SK_HIP(c, hipEventElapsedTime(&t, e[v], e[v + 1]));
      if (getenv("SYNKRO_VERBOSE")) fprintf(stderr, "sk_spmv: variant %d: %.3f ms\n", cand[v], t);
      if (v == 0 || t < tbest) {
        tbest = t;
        best = cand[v];
      }
    }
    for (int v = 0; v <= NV; ++v) (void)hipEventDestroy(e[v]);
    const_cast<sk_mat*>(m)->spmv_variant = best;
    // The kernels sum a row in different orders; leave the chosen one's result, so that
    // every call on this matrix returns the same bits.
    if (best != cand[NV - 1]) {
      int rc = run(best);
      if (rc) return rc;
    }
    return launched(c, "SpMV kernel");
  }
  int rc = run(m->spmv_variant);
  if (rc) return rc;
  return launched(c, "SpMV kernel");
}

void sk_mat_destroy(sk_mat* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  (void)hipDeviceSynchronize();
  mat_free(m);
}

}  // extern "C"
