// Batched CG-Lanczos Fiedler solver: every segment of a bisection level runs
// the reference's restarted CG-Lanczos (fiedler.c:434-621) at once.
//
// Rows of the level are grouped in tiles of RSB_TILE rows that never cross a
// segment, one workgroup per tile.  Each workgroup writes the partial sums of
// its tile (fixed reduction tree), so results do not depend on block
// scheduling.  The segment totals and scalar updates are done by a small
// k_lz_fin launch after each reduction (one workgroup per segment), so no
// workgroup ever waits on another inside a launch.  One CG-Lanczos
// iteration is two launches over all rows of the level, each followed by
// its k_lz_fin:
//   k_lz_apply    p = beta (p_old - mean) + r formed on the fly for each row
//                 and its neighbours (staged in LDS per tile), w = L p from the
//                 tile's 12- or 16-bit entries (the SELL-64 global form where a
//                 tile's halo does not fit), pap = w . (p - mean(p))
//   k_lz_rupdate  r = r - alpha w, rtr = r . r, tridiagonal entries, stop test
// (FP32 passes on few tiles fold the update into the next apply, below).
// The Lanczos basis vector of iteration it is r_it / |r_it|: the residuals
// themselves are kept, r_it in basis slot it, and the Ritz step scales them.
// The mean of p is removed lazily (fiedler.c:20-30 `ortho`): w = L p is
// evaluated in the difference form sum_j w_ij (p_i - p_j), which the mean
// does not change, and the stored p keeps its mean, subtracted when used.
// Segments that converged or are between passes skip every launch at the
// first instruction.
//
// Precision: vectors are T (double, or float); every scalar and the
// tridiagonal eigen-solve are FP64.  With T = float, dot products are
// compensated (Dot2: exact products and error-free sums in FP32) and their
// block partials are combined in FP64.
#include <algorithm>
#include <vector>

#include <cfloat>

#include "state.hpp"

struct lz_args {
  const int *seg;        // 5*nseg: cstart, n, p, r0, pstart
  const int *tile;       // 3*ntiles: seg, row0, row1
  const int *seg_tiles;  // 2*nseg: first tile, tile count
  const int *row_elem;
  int nseg;
  int *status, *iter, *k, *npass, *niter, *miter;
  double *sc, *T, *y, *partial;
  double *rn;  // (miter_max + 1) per segment: 1 / |r_j|
  int *running;
  int mpass;
  int mode;  // start vector: 0 reference, 1 caller per element, 2 learned per row
  int max_n; // mode 2: segments with more rows take the reference start
  double tol;
  double stall;  // > 0: a pass whose Ritz value moved by at most this (relative) ends the solve
  int stall_min; // ... from this pass on
  double plateau; // > 0: a pass whose Ritz residual is at most plateau_max and not below
                  // this times the one two passes back ends the solve (from the fourth pass)
  double plateau_max;
  int miter_max;
  const int *act_tiles;  // if set: blockIdx.x -> tile (compacted active list)
  const int *act_segs;   // if set: blockIdx.x -> segment, for k_lz_fin
  int pstride;           // partial is slot-major: slot k of tile t at k * pstride + t
};

#define SC(s, j) a.sc[(size_t)(s) * LZ_NSC + (j)]

template <typename T>
struct acc_t;
template <>
struct acc_t<double> {
  double v = 0.0;
  __device__ inline void dot(double x, double y) { v += x * y; }
  __device__ inline void add(double x) { v += x; }
  __device__ inline double value() const { return v; }
};
template <>
struct acc_t<float> {
  float s = 0.f, c = 0.f;
  __device__ inline void dot(float x, float y) {
    float p = x * y;
    float ep = __builtin_fmaf(x, y, -p);  // exact product error
    float t, et;
    two_sum(s, p, t, et);
    s = t;
    c += ep + et;
  }
  __device__ inline void add(float x) {
    float t, et;
    two_sum(s, x, t, et);
    s = t;
    c += et;
  }
  __device__ inline double value() const { return (double)s + (double)c; }
};

// Tile prologue: returns false if the block has nothing to do.
struct tile_t {
  int ti, s, row0, row1, pstart, n, first, count;
};
// Workgroups are handed to the XCDs round robin (workgroup b to XCD b % 8),
// so consecutive tiles, which are neighbours in space, would land in eight
// different L2 caches.  With nlist > 0 the list of nlist tiles is split in
// eight contiguous parts, one per XCD, so a tile's halo rows are mostly rows
// its XCD has just read (grid: nlist rounded up to a multiple of 8).
#define LZ_XCDS 8
// running tiles up to which an FP32 pass folds its r-update into the apply
#define LZ_FOLD_TILES 4096
// segments of more tiles than this take 512-thread finishing workgroups
#define LZ_FIN_WIDE 1024
// profiling samples one SpMV launch in this many (coprime with the 8
// iterations between list compactions and the 50 of a pass)
#define LZ_PROF_EVERY 11
__device__ inline int list_index(int nlist) {
  if (nlist <= 0) return blockIdx.x;
  const int per = (int)(gridDim.x / LZ_XCDS);
  const int i = (int)(blockIdx.x % LZ_XCDS) * per + (int)(blockIdx.x / LZ_XCDS);
  return i < nlist ? i : -1;
}
__device__ inline bool tile_of(const lz_args &a, int want_status, tile_t &t, int nlist = 0) {
  const int li = list_index(nlist);
  if (li < 0) return false;
  const int ti = a.act_tiles ? a.act_tiles[li] : li;
  t.ti = ti;
  t.s = a.tile[3 * ti];
  if (a.status[t.s] != want_status) return false;
  t.row0 = a.tile[3 * ti + 1];
  t.row1 = a.tile[3 * ti + 2];
  t.n = a.seg[a.nseg + t.s];
  t.pstart = a.seg[4 * a.nseg + t.s];
  t.first = a.seg_tiles[2 * t.s];
  t.count = a.seg_tiles[2 * t.s + 1];
  return true;
}

// Write the block's partial sums (a fixed reduction tree) for the tile.
// The segment totals are formed by k_lz_fin in the next launch: no
// cross-workgroup handshake inside a kernel, so no device-scope fences.
template <int NV>
__device__ inline void publish(const lz_args &a, const tile_t &t, double (&v)[NV]) {
  __shared__ double red[(RSB_BLOCK / RSB_WAVE) * (NV > 4 ? NV : 4)];
  block_sum<NV>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; k++) a.partial[(size_t)k * a.pstride + t.ti] = v[k];
}

// ---------------------------------------------------------------------------
// Start vector (fiedler.c:570-585): x_i = i + 1, plus 1000 n on the first half
// of the pre-partition order; then mean removed and normalised (FP64).
// mode 0: reference start; mode 1: caller-supplied per-element start;
// mode 2: per-row start (the learned warm start's prediction).
__device__ inline double start_value(int mode, int i, int n, const float *start, int e,
                                     int row) {
  if (mode == 1) return (double)start[e];
  if (mode == 2) return (double)start[row];
  double v = i + 1.0;
  if (i < n / 2) v += (double)(1000 * (long long)n);
  return v;
}

template <typename T>
__global__ void __launch_bounds__(RSB_BLOCK) k_lz_start(lz_args a, int phase, int mode,
                                                        const float *start, T *f) {
  tile_t t;
  if (!tile_of(a, LZ_SETUP, t)) return;
  double m = phase > 0 ? SC(t.s, SC_M0) : 0.0;
  double sc = phase > 1 ? SC(t.s, SC_N0) : 0.0;
  double v[1] = {0.0};
  for (int row = t.row0 + threadIdx.x; row < t.row1; row += blockDim.x) {
    int i = row - t.pstart;
    double x = 0.0;
    if (i < t.n) {
      x = start_value(mode, i, t.n, start, a.row_elem[row], row);
      if (phase == 0) v[0] += x;
      else if (phase == 1) v[0] += (x - m) * (x - m);
    }
    if (phase == 2) f[row] = i < t.n ? (T)((x - m) * sc) : (T)0;
  }
  if (phase == 2) return;
  publish<1>(a, t, v);
}

// Pass setup (fiedler.c:441-456): r = f - mean_f; then r -= mean(r);
// rtr, |r|, target tol |r|.
template <typename T>
__global__ void __launch_bounds__(RSB_BLOCK) k_lz_setup(lz_args a, int phase, const T *f, T *r,
                                                        T *p) {
  tile_t t;
  if (!tile_of(a, LZ_SETUP, t)) return;
  acc_t<T> acc;
  if (phase == 0) {
    T mf = (T)SC(t.s, SC_MEANF);
    for (int row = t.row0 + threadIdx.x; row < t.row1; row += blockDim.x) {
      int i = row - t.pstart;
      T x = i < t.n ? (T)(f[row] - mf) : (T)0;
      r[row] = x;
      p[row] = (T)0;
      acc.add(x);
    }
  } else {
    T mr = (T)SC(t.s, SC_TMP);
    for (int row = t.row0 + threadIdx.x; row < t.row1; row += blockDim.x) {
      int i = row - t.pstart;
      if (i < t.n) {
        T x = r[row] - mr;
        r[row] = x;
        acc.dot(x, x);
      }
    }
  }
  double v[1] = {acc.value()};
  publish<1>(a, t, v);
}

// ---------------------------------------------------------------------------
// One row of L q in the difference form sum_j w_ij (q_i - q_j), where q is
// formed on the fly from the previous search direction and the residual:
// q = beta (p - m) + r (q = r on the first iteration of a pass).  The packed
// entries of a SELL-64 slice are read 8 at a time so that each thread keeps 8
// gathers in flight.
// The residual of iteration it: r itself, or with FOLD (the FP32 path)
// r_prev - alpha w_prev, the update of the previous iteration, formed where
// it is read (fiedler.c:486-488).
// r - alpha w with one rounding (the r-update of both forms)
__device__ inline float rsub(float r, float alpha, float w) { return __builtin_fmaf(-alpha, w, r); }
__device__ inline double rsub(double r, double alpha, double w) { return __builtin_fma(-alpha, w, r); }
template <typename T, bool FOLD>
__device__ inline T rval(const T *__restrict__ r, const T *__restrict__ wp, uint32_t c, T alpha) {
  return FOLD ? rsub(r[c], alpha, wp[c]) : r[c];
}
template <typename T, bool FIRST, bool FOLD>
__device__ inline T qval(const T *__restrict__ p, const T *__restrict__ r,
                         const T *__restrict__ wp, uint32_t c, T beta, T m, T alpha) {
  return FIRST ? r[c] : beta * (p[c] - m) + rval<T, FOLD>(r, wp, c, alpha);
}

template <typename T, bool FIRST, bool FOLD>
__device__ inline T lap_row_q(const uint32_t *__restrict__ e, int wd, T qi,
                              const T *__restrict__ p, const T *__restrict__ r,
                              const T *__restrict__ wp, T beta, T m, T alpha) {
  T sum = (T)0;
  int j = 0;
  for (; j + 8 <= wd; j += 8) {
    uint32_t q[8];
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) q[u] = __builtin_nontemporal_load(e + (size_t)(j + u) * RSB_WAVE);
#pragma unroll
    for (int u = 0; u < 8; u++)
      v[u] = qval<T, FIRST, FOLD>(p, r, wp, q[u] & 0x0fffffffu, beta, m, alpha);
#pragma unroll
    for (int u = 0; u < 8; u++) sum += (T)(q[u] >> 28) * (qi - v[u]);
  }
  for (; j < wd; j++) {
    uint32_t q = __builtin_nontemporal_load(e + (size_t)j * RSB_WAVE);
    sum += (T)(q >> 28) * (qi - qval<T, FIRST, FOLD>(p, r, wp, q & 0x0fffffffu, beta, m, alpha));
  }
  return sum;
}

// What the folded apply adds: the residual of iteration it written for the
// tile's rows, and the sums r.r, r.w and w.w from which the finishing step
// forms |r - alpha w|^2, the next residual's norm.
template <typename T>
struct fold_t {
  const T *wp;  // w of the previous iteration
  T *rn;        // r of this iteration (written)
  T alpha;      // alpha of the previous iteration
  acc_t<T> rr, rw, ww;
};

// Fused search-direction update and Laplacian apply (fiedler.c:460-475):
//   q = beta (p_old - mean_p) + r,  w = L q  (r: basis slot it)
// and the segment sums of q, w.q and w, from which the last block forms
// mean_p and pap = w.(q - mean_p), then alpha = rtz1 / pap.
template <typename T, bool FIRST, bool FOLD>
__device__ inline void lz_apply_rows(const lz_args &a, const tile_t &t, const T *__restrict__ r,
                                     const T *__restrict__ pold, T *__restrict__ pnew,
                                     T *__restrict__ w, T beta, T m,
                                     const int64_t *__restrict__ slice_off,
                                     const int *__restrict__ slice_w,
                                     const uint32_t *__restrict__ ell, acc_t<T> &sq,
                                     acc_t<T> &swq, acc_t<T> &sw, fold_t<T> &fd) {
  constexpr bool UPD = FOLD && !FIRST;
  for (int row = t.row0 + threadIdx.x; row < t.row1; row += blockDim.x) {
    const bool real = row - t.pstart < t.n;
    const T ri = rval<T, UPD>(r, fd.wp, (uint32_t)row, fd.alpha);
    T qi = (T)0, wi = (T)0;
    if (real) {
      qi = FIRST ? ri : beta * (pold[row] - m) + ri;
      const int sl = row >> 6;
      wi = lap_row_q<T, FIRST, UPD>(ell + slice_off[sl] + (row & 63), slice_w[sl], qi, pold, r,
                                    fd.wp, beta, m, fd.alpha);
      sq.add(qi);
      swq.dot(wi, qi);
      sw.add(wi);
      if (FOLD) fd.rr.dot(ri, ri), fd.rw.dot(ri, wi), fd.ww.dot(wi, wi);
    }
    pnew[row] = qi;
    w[row] = wi;
    if (UPD) fd.rn[row] = ri;
  }
}

// LDS-tiled rows (tiles.hip): q for the tile's rows and its halo is staged
// in LDS once; every neighbour is then read from LDS through a 16-bit local
// index, 8 entries per 16-byte load.
struct tiled_args {
  const int64_t *slice_off;
  const int *slice_w;
  const uint32_t *ell;
  const int *nhalo;
  const int64_t *halo_off;
  const int *halo;
  const int64_t *off16;
  const int *w16;
  const uint16_t *e16;
  // the 12-bit form (tiles.hip), or e12 == nullptr
  const int64_t *off12;
  const int *w12;
  const uint32_t *e12;
  const int *c12;
};

#define TB_LOCAL 4096
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ inline void acc_entry(uint32_t e, T qi, const T *ql, T &sum) {
  sum += (T)(e >> 12) * (qi - ql[e & 0xfffu]);
}

// One row of the tiled apply: its 16-bit entries (groups of 8, then groups
// of 4 and 2) against the LDS window; writes p', w and the basis entry.
template <typename T, bool FOLD>
__device__ inline void apply_row_tiled(int row, const tile_t &t, const tiled_args &ta,
                                       const T *ql, const T *rl, T *__restrict__ pnew,
                                       T *__restrict__ w, int nreal, acc_t<T> &sq,
                                       acc_t<T> &swq, acc_t<T> &sw, fold_t<T> &fd) {
  const int li = row - t.row0, sl = row >> 6;
  const uint16_t *sbase = ta.e16 + ta.off16[sl];
  const u32x4 *ep = (const u32x4 *)(sbase + (row & 63) * 8);
  const int w16 = ta.w16[sl], nb = w16 >> 3;
  const T qi = ql[li];
  T sum = (T)0;
  for (int jb = 0; jb < nb; jb++) {
    const u32x4 v = __builtin_nontemporal_load(ep + (size_t)jb * RSB_WAVE);
    acc_entry(v.x & 0xffffu, qi, ql, sum);
    acc_entry(v.x >> 16, qi, ql, sum);
    acc_entry(v.y & 0xffffu, qi, ql, sum);
    acc_entry(v.y >> 16, qi, ql, sum);
    acc_entry(v.z & 0xffffu, qi, ql, sum);
    acc_entry(v.z >> 16, qi, ql, sum);
    acc_entry(v.w & 0xffffu, qi, ql, sum);
    acc_entry(v.w >> 16, qi, ql, sum);
  }
  if (w16 & 4) {  // a group of 4 entries
    const u32x2 v =
        __builtin_nontemporal_load((const u32x2 *)(sbase + nb * 512 + (row & 63) * 4));
    acc_entry(v.x & 0xffffu, qi, ql, sum);
    acc_entry(v.x >> 16, qi, ql, sum);
    acc_entry(v.y & 0xffffu, qi, ql, sum);
    acc_entry(v.y >> 16, qi, ql, sum);
  }
  if (w16 & 2) {  // a group of 2 entries
    const unsigned v = __builtin_nontemporal_load(
        (const unsigned *)(sbase + nb * 512 + (w16 & 4) * 64 + (row & 63) * 2));
    acc_entry(v & 0xffffu, qi, ql, sum);
    acc_entry(v >> 16, qi, ql, sum);
  }
  pnew[row] = qi;
  w[row] = sum;
  const T ri = FOLD ? rl[li] : (T)0;  // the tile's r is kept in LDS only to fold
  if (li < nreal) {
    sq.add(qi);
    swq.dot(sum, qi);
    sw.add(sum);
    if (FOLD) fd.rr.dot(ri, ri), fd.rw.dot(ri, sum), fd.ww.dot(sum, sum);
  }
}

// R rows per thread and step: the entries of all R rows are loaded before
// any is summed, so each lane keeps up to 4R 16-byte loads in flight.  Rows
// whose slice is wider than 32 entries take apply_row_tiled.  Missing groups
// are zero words (weight 0), which add exact zeros.
template <typename T, int R, bool FOLD>
__device__ inline void apply_rows_batched(const tile_t &t, const tiled_args &ta, const T *ql,
                                          const T *rl, T *__restrict__ pnew, T *__restrict__ w,
                                          int nreal, acc_t<T> &sq, acc_t<T> &swq, acc_t<T> &sw,
                                          fold_t<T> &fd) {
  for (int row0 = t.row0 + threadIdx.x; row0 < t.row1; row0 += R * blockDim.x) {
    int rows[R], wd[R];
    const uint16_t *base[R];
    bool wide = false;
#pragma unroll
    for (int k = 0; k < R; k++) {
      rows[k] = row0 + k * blockDim.x;
      const bool has = rows[k] < t.row1;
      const int sl = (has ? rows[k] : row0) >> 6;
      base[k] = ta.e16 + ta.off16[sl];
      wd[k] = has ? ta.w16[sl] : 0;
      wide |= wd[k] > 32;
    }
    if (wide) {
#pragma unroll
      for (int k = 0; k < R; k++)
        if (rows[k] < t.row1)
          apply_row_tiled<T, FOLD>(rows[k], t, ta, ql, rl, pnew, w, nreal, sq, swq, sw, fd);
      continue;
    }
    u32x4 v[R][4];
    u32x2 tl[R];
    unsigned t2[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
#pragma unroll
      for (int g = 0; g < 4; g++)
        v[k][g] = g < (wd[k] >> 3)
                      ? __builtin_nontemporal_load((const u32x4 *)(base[k] + (rows[k] & 63) * 8) +
                                                   g * RSB_WAVE)
                      : (u32x4){0u, 0u, 0u, 0u};
      const uint16_t *tail = base[k] + (wd[k] >> 3) * 512;
      tl[k] = (wd[k] & 4) ? __builtin_nontemporal_load((const u32x2 *)(tail + (rows[k] & 63) * 4))
                          : (u32x2){0u, 0u};
      t2[k] = (wd[k] & 2) ? __builtin_nontemporal_load(
                                (const unsigned *)(tail + (wd[k] & 4) * 64 + (rows[k] & 63) * 2))
                          : 0u;
    }
#pragma unroll
    for (int k = 0; k < R; k++) {
      if (rows[k] >= t.row1) continue;
      const int li = rows[k] - t.row0;
      const T qi = ql[li];
      T sum = (T)0;
#pragma unroll
      for (int g = 0; g < 4; g++) {
        acc_entry(v[k][g].x & 0xffffu, qi, ql, sum);
        acc_entry(v[k][g].x >> 16, qi, ql, sum);
        acc_entry(v[k][g].y & 0xffffu, qi, ql, sum);
        acc_entry(v[k][g].y >> 16, qi, ql, sum);
        acc_entry(v[k][g].z & 0xffffu, qi, ql, sum);
        acc_entry(v[k][g].z >> 16, qi, ql, sum);
        acc_entry(v[k][g].w & 0xffffu, qi, ql, sum);
        acc_entry(v[k][g].w >> 16, qi, ql, sum);
      }
      acc_entry(tl[k].x & 0xffffu, qi, ql, sum);
      acc_entry(tl[k].x >> 16, qi, ql, sum);
      acc_entry(tl[k].y & 0xffffu, qi, ql, sum);
      acc_entry(tl[k].y >> 16, qi, ql, sum);
      acc_entry(t2[k] & 0xffffu, qi, ql, sum);
      acc_entry(t2[k] >> 16, qi, ql, sum);
      pnew[rows[k]] = qi;
      w[rows[k]] = sum;
      const T ri = FOLD ? rl[li] : (T)0;
      if (li < nreal) {
        sq.add(qi);
        swq.dot(sum, qi);
        sw.add(sum);
        if (FOLD) fd.rr.dot(ri, ri), fd.rw.dot(ri, sum), fd.ww.dot(sum, sum);
      }
    }
  }
}

// 12-bit index k (0..20) of a 32-byte block held in eight words
template <int K>
__device__ inline unsigned e12_at(const unsigned (&wd)[8]) {
  constexpr int bit = 12 * K, i = bit / 32, sh = bit % 32;
  if constexpr (sh <= 20) return (wd[i] >> sh) & 0xfffu;
  else return __builtin_amdgcn_alignbit(wd[i + 1], wd[i], sh) & 0xfffu;
}

// One slot of the 12-bit form: its weight 2^e from the slice's class
// bounds (slots ordered 4s, 2s, 1s; n42 = n4 + n2; both uniform in the
// wave, so e is scalar work).  Scaling by 2^e is exact: the sum rounds as
// sum + w (qi - q) would.
__device__ inline float ldexp_t(float x, int e) { return __builtin_ldexpf(x, e); }
__device__ inline double ldexp_t(double x, int e) { return __builtin_ldexp(x, e); }
// m4 and m42 hold bit s for the slots s < n4 and s < n42 (shifts and masks
// of wave-uniform words: scalar work).
template <typename T>
__device__ inline void acc12(unsigned idx, int s, uint64_t m4, uint64_t m42, T qi, const T *ql,
                             T &sum) {
  const int e = (int)((m4 >> s) & 1u) + (int)((m42 >> s) & 1u);
  sum += ldexp_t(qi - ql[idx], e);
}

template <typename T, int K0, int N>
__device__ inline void acc12_words(const unsigned (&wd)[8], int s0, uint64_t m4, uint64_t m42,
                                   T qi, const T *ql, T &sum) {
  if constexpr (N > 0) {
    acc12(e12_at<K0>(wd), s0 + K0, m4, m42, qi, ql, sum);
    acc12_words<T, K0 + 1, N - 1>(wd, s0, m4, m42, qi, ql, sum);
  }
}

// The tiled rows on the 12-bit form: R rows per thread and step, every
// word of their blocks and tails loaded before any is summed.  The 64 rows
// of a wave are one slice, so block and tail counts are uniform in a wave.
template <typename T, int R, bool FOLD>
__device__ inline void apply_rows_e12(const tile_t &t, const tiled_args &ta, const T *ql,
                                      const T *rl, T *__restrict__ pnew, T *__restrict__ w,
                                      int nreal, acc_t<T> &sq, acc_t<T> &swq, acc_t<T> &sw,
                                      fold_t<T> &fd) {
  for (int row0 = t.row0 + threadIdx.x; row0 < t.row1; row0 += R * blockDim.x) {
    int rows[R], nb[R], tw[R], cc[R];
    u32x4 bw[R][2][2], tl[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
      rows[k] = row0 + k * blockDim.x;
      const bool has = rows[k] < t.row1;
      const int sl = (has ? rows[k] : row0) >> 6, lane = rows[k] & 63;
      const int info = has ? ta.w12[sl] : 0;
      nb[k] = info & 0xff, tw[k] = info >> 8;
      const uint32_t *base = ta.e12 + ta.off12[sl];
      cc[k] = has ? ta.c12[sl] : 0;
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int h = 0; h < 2; h++)
          bw[k][b][h] = b < nb[k] ? __builtin_nontemporal_load(
                                        (const u32x4 *)(base + (size_t)(2 * b + h) * 256 + lane * 4))
                                  : (u32x4){0u, 0u, 0u, 0u};
      const uint32_t *tail = base + (size_t)nb[k] * 512;
      if (tw[k] == 4)
        tl[k] = __builtin_nontemporal_load((const u32x4 *)(tail + lane * 4));
      else if (tw[k] == 2) {
        const u32x2 v = __builtin_nontemporal_load((const u32x2 *)(tail + lane * 2));
        tl[k] = (u32x4){v.x, v.y, 0u, 0u};
      } else {
        tl[k] = (u32x4){0u, 0u, 0u, 0u};
      }
    }
#pragma unroll
    for (int k = 0; k < R; k++) {
      if (rows[k] >= t.row1) continue;
      const int li = rows[k] - t.row0;
      const T qi = ql[li];
      // the slice's class bounds: every active lane holds the same value
      const int c = __builtin_amdgcn_readfirstlane(cc[k]);
      const int n4 = c & 0xff, n42 = n4 + ((c >> 8) & 0xff);  // at most 42 each
      const uint64_t m4 = (1ull << n4) - 1u, m42 = (1ull << n42) - 1u;
      T sum = (T)0;
#pragma unroll
      for (int b = 0; b < 2; b++)
        if (b < nb[k]) {
          const unsigned wd[8] = {bw[k][b][0].x, bw[k][b][0].y, bw[k][b][0].z, bw[k][b][0].w,
                                  bw[k][b][1].x, bw[k][b][1].y, bw[k][b][1].z, bw[k][b][1].w};
          acc12_words<T, 0, 21>(wd, 21 * b, m4, m42, qi, ql, sum);
        }
      if (tw[k]) {
        const unsigned wd[8] = {tl[k].x, tl[k].y, tl[k].z, tl[k].w, 0u, 0u, 0u, 0u};
        if (tw[k] == 4) acc12_words<T, 0, 10>(wd, 21 * nb[k], m4, m42, qi, ql, sum);
        else acc12_words<T, 0, 5>(wd, 21 * nb[k], m4, m42, qi, ql, sum);
      }
      pnew[rows[k]] = qi;
      w[rows[k]] = sum;
      const T ri = FOLD ? rl[li] : (T)0;
      if (li < nreal) {
        sq.add(qi);
        swq.dot(sum, qi);
        sw.add(sum);
        if (FOLD) fd.rr.dot(ri, ri), fd.rw.dot(ri, sum), fd.ww.dot(sum, sum);
      }
    }
  }
}

// Stages q = beta (p_old - m) + r (q = r on the first iteration of a pass)
// for the tile's rows and its nh halo rows into LDS (tile rows first), and r
// of the tile's rows; rows past the segment's end get q = 0.  FP32 without
// the folded update: four consecutive tile rows per thread in 16-byte
// accesses (tiles are whole 64-row slices), then the halo rows.  SU indices
// per thread and step, all their loads issued before any value is used (the
// halo rows are gathers).  Ends with a barrier.
template <typename T, bool FIRST, bool KEEP_R = true, bool FOLD = false>
__device__ inline void stage_window(const tile_t &t, const int *__restrict__ halo, int nh, T *ql,
                                    T *rl, const T *__restrict__ r, const T *__restrict__ pold,
                                    T beta, T m, const T *__restrict__ wp = nullptr,
                                    T alpha = (T)0, T *__restrict__ rn = nullptr) {
  constexpr bool UPD = FOLD && !FIRST;
  const int nrows = t.row1 - t.row0;
  const int nreal = min(nrows, t.pstart + t.n - t.row0);
  const int tot = nrows + nh;
  constexpr int SU = 4;
  int ibeg = threadIdx.x;
  if constexpr (sizeof(T) == 4) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    for (int i = 4 * (int)threadIdx.x; i < nrows; i += 4 * (int)blockDim.x) {
      f32x4 rv = *(const f32x4 *)(r + t.row0 + i);
      if (UPD) {
        const f32x4 wv = *(const f32x4 *)(wp + t.row0 + i);
        rv.x = rsub(rv.x, alpha, wv.x), rv.y = rsub(rv.y, alpha, wv.y);
        rv.z = rsub(rv.z, alpha, wv.z), rv.w = rsub(rv.w, alpha, wv.w);
        if (rn) *(f32x4 *)(rn + t.row0 + i) = rv;  // r_it to its basis slot
      }
      f32x4 q = rv;
      if (!FIRST) {
        const f32x4 pv = *(const f32x4 *)(pold + t.row0 + i);
        q.x = beta * (pv.x - m) + rv.x, q.y = beta * (pv.y - m) + rv.y;
        q.z = beta * (pv.z - m) + rv.z, q.w = beta * (pv.w - m) + rv.w;
      }
      if (i + 0 >= nreal) q.x = 0.f;
      if (i + 1 >= nreal) q.y = 0.f;
      if (i + 2 >= nreal) q.z = 0.f;
      if (i + 3 >= nreal) q.w = 0.f;
      *(f32x4 *)(ql + i) = q;
      if (KEEP_R) *(f32x4 *)(rl + i) = rv;
    }
    ibeg = nrows + threadIdx.x;
  }
  for (int i0 = ibeg; i0 < tot; i0 += SU * blockDim.x) {
    int gr[SU];
    T rg[SU], po[SU], wg[SU];
#pragma unroll
    for (int k = 0; k < SU; k++) {
      const int idx = i0 + k * blockDim.x;
      gr[k] = idx < nrows ? t.row0 + idx : (idx < tot ? halo[idx - nrows] : t.row0);
    }
#pragma unroll
    for (int k = 0; k < SU; k++) {
      rg[k] = r[gr[k]];
      po[k] = FIRST ? (T)0 : pold[gr[k]];
      wg[k] = UPD ? wp[gr[k]] : (T)0;
    }
    if (UPD)
#pragma unroll
      for (int k = 0; k < SU; k++) rg[k] = rsub(rg[k], alpha, wg[k]);
#pragma unroll
    for (int k = 0; k < SU; k++) {
      const int idx = i0 + k * blockDim.x;
      if (idx >= tot) break;
      T q = FIRST ? rg[k] : beta * (po[k] - m) + rg[k];
      if (idx < nrows) {
        if (KEEP_R) rl[idx] = rg[k];
        if (UPD && rn) rn[gr[k]] = rg[k];  // r_it to its basis slot
        if (idx >= nreal) q = (T)0;
      }
      ql[idx] = q;
    }
  }
  __syncthreads();
}

template <typename T, bool FIRST, bool FOLD = false>
__device__ inline void lz_apply_tiled(const lz_args &a, const tile_t &t, const tiled_args &ta,
                                      int nh, T *ql, T *rl, const T *__restrict__ r,
                                      const T *__restrict__ pold, T *__restrict__ pnew,
                                      T *__restrict__ w, T beta, T m, acc_t<T> &sq,
                                      acc_t<T> &swq, acc_t<T> &sw, fold_t<T> &fd) {
  const int nrows = t.row1 - t.row0;
  const int nreal = min(nrows, t.pstart + t.n - t.row0);
  // the tile's r in LDS: the folded update's row phase reads it; the FP32
  // kernel writes it anyway, its 4 KB holding the kernel at seven
  // workgroups per CU (at eight, with rl dropped, the level-0 apply of the
  // 256^3 box took 279 us against 269); the FP64 kernel drops it and runs
  // four workgroups per CU instead of three (0.217 -> 0.209 ms per launch)
  constexpr bool KEEP_R = FOLD || sizeof(T) == 4;
  stage_window<T, FIRST, KEEP_R, FOLD>(t, ta.halo + ta.halo_off[t.ti], nh, ql, rl, r, pold, beta,
                                       m, fd.wp, fd.alpha, fd.rn);
  // two rows per thread and step (measured 15% faster than one at a time
  // on the 256^3 workload: more loads in flight per lane)
  if (ta.e12)
    apply_rows_e12<T, 2, FOLD>(t, ta, ql, rl, pnew, w, nreal, sq, swq, sw, fd);
  else
    apply_rows_batched<T, 2, FOLD>(t, ta, ql, rl, pnew, w, nreal, sq, swq, sw, fd);
}

// The residuals of a pass live in the basis, unscaled: r_it in slot it
// (the Ritz step scales slot j by 1 / |r_j|), so the apply reads its r
// there and writes no basis vector.
// FOLD (FP32): the residual update of the previous iteration is formed in
// the staging, r_it = r_{it-1} - alpha w_{it-1} (w_prev: the w buffer of
// the other parity), written to slot it, and the launch adds the sums r.r,
// r.w, w.w (partial slots 3..5) for FIN_APPLYR.  No separate r-update pass
// and one finishing step per iteration.
template <typename T, bool FOLD>
__global__ void __launch_bounds__(RSB_BLOCK) k_lz_apply(lz_args a, const T *__restrict__ pold,
                                                        T *__restrict__ pnew, T *__restrict__ w,
                                                        T *__restrict__ basis, int64_t mpad,
                                                        tiled_args ta, int nlist,
                                                        const T *__restrict__ wprev) {
  __shared__ __attribute__((aligned(16))) T ql[TB_LOCAL];
  // the tile's r (read only to fold the update; see lz_apply_tiled)
  __shared__ __attribute__((aligned(16))) T rl[RSB_TILE];
  tile_t t;
  if (!tile_of(a, LZ_RUNNING, t, nlist)) return;
  const int s = t.s, it = a.iter[s];
  T *cur = basis + (size_t)it * mpad;
  // what the staging reads: r_it, or r_{it-1} to update (FOLD)
  const T *__restrict__ r = (FOLD && it > 0) ? cur - mpad : cur;
  const T beta = (T)SC(s, SC_BETA), m = (T)SC(s, SC_MEANP);
  acc_t<T> sq, swq, sw;
  fold_t<T> fd;
  fd.wp = wprev, fd.rn = it == 0 ? nullptr : cur, fd.alpha = (T)SC(s, SC_ALPHA);
  const int nh = ta.nhalo[t.ti];
  if (nh >= 0) {
    if (it == 0)
      lz_apply_tiled<T, true, FOLD>(a, t, ta, nh, ql, rl, r, pold, pnew, w, beta, m, sq, swq, sw,
                                    fd);
    else
      lz_apply_tiled<T, false, FOLD>(a, t, ta, nh, ql, rl, r, pold, pnew, w, beta, m, sq, swq,
                                     sw, fd);
  } else if (it == 0) {
    lz_apply_rows<T, true, FOLD>(a, t, r, pold, pnew, w, beta, m, ta.slice_off, ta.slice_w,
                                 ta.ell, sq, swq, sw, fd);
  } else {
    lz_apply_rows<T, false, FOLD>(a, t, r, pold, pnew, w, beta, m, ta.slice_off, ta.slice_w,
                                  ta.ell, sq, swq, sw, fd);
  }
  if (FOLD) {
    double v[6] = {sq.value(), swq.value(), sw.value(),
                   fd.rr.value(), fd.rw.value(), fd.ww.value()};
    publish<6>(a, t, v);
  } else {
    double v[3] = {sq.value(), swq.value(), sw.value()};
    publish<3>(a, t, v);
  }
}

// Vertex form of the tiles (tiles.hip, rsb_build_vtiles): the reference's
// gather-scatter form (laplacian.c:159-205), (L q)_e = sum over the slots of
// e of sum over the slots at their vertex v of (q_e - q_e'), evaluated as
// mult(v) (q_e - q_f(v)) - S(v), f(v) the first slot at v and
// S(v) = sum over the slots at v of (q_e' - q_f(v)): every term a difference
// of neighbouring values, as in the neighbour-entry form (summing q itself
// loses the low-frequency part of L q in FP32).  One aligned load per vertex
// fetches its nv window indices (ascending, padded with VT_MAXW); LDS keeps
// S(v) and one 16-bit word f(v) | mult(v) << 12 per vertex; fixed order of
// summation.
struct vtile_args {
  const int *vt_n;
  const int64_t *vt_loff;
  const uint16_t *list, *rowv;
  const int *nhalo;
  const int64_t *halo_off;
  const int *halo;
  int nv;
};

template <typename T, int W>
__device__ inline void vt_sums(const uint16_t *__restrict__ Ls, int nvert, const T *ql, T *S,
                               uint16_t *F) {
  // two vertices per thread and step, both loads issued first
  for (int v0 = threadIdx.x; v0 < nvert; v0 += 2 * blockDim.x) {
    const int v1 = v0 + blockDim.x;
    uint16_t e[2][W];
    if (W == 8) {
      const u32x4 a = __builtin_nontemporal_load((const u32x4 *)(Ls + (int64_t)v0 * 8));
      const u32x4 b = v1 < nvert ? __builtin_nontemporal_load((const u32x4 *)(Ls + (int64_t)v1 * 8))
                                 : (u32x4){0u, 0u, 0u, 0u};
      const unsigned wa[4] = {a.x, a.y, a.z, a.w}, wb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        e[0][2 * k] = (uint16_t)(wa[k] & 0xffffu), e[0][2 * k + 1] = (uint16_t)(wa[k] >> 16);
        e[1][2 * k] = (uint16_t)(wb[k] & 0xffffu), e[1][2 * k + 1] = (uint16_t)(wb[k] >> 16);
      }
    } else {
      const u32x2 a = __builtin_nontemporal_load((const u32x2 *)(Ls + (int64_t)v0 * 4));
      const u32x2 b = v1 < nvert ? __builtin_nontemporal_load((const u32x2 *)(Ls + (int64_t)v1 * 4))
                                 : (u32x2){0u, 0u};
      const unsigned wa[2] = {a.x, a.y}, wb[2] = {b.x, b.y};
#pragma unroll
      for (int k = 0; k < 2; k++) {
        e[0][2 * k] = (uint16_t)(wa[k] & 0xffffu), e[0][2 * k + 1] = (uint16_t)(wa[k] >> 16);
        e[1][2 * k] = (uint16_t)(wb[k] & 0xffffu), e[1][2 * k + 1] = (uint16_t)(wb[k] >> 16);
      }
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int v = u ? v1 : v0;
      if (v >= nvert) break;
      const T ref = ql[e[u][0]];  // every list has at least one slot
      T sum = (T)0;
      int cnt = 0;
#pragma unroll
      for (int k = 0; k < W; k++) {
        const bool real = e[u][k] != VT_MAXW;
        sum += real ? ql[e[u][k]] - ref : (T)0;
        cnt += real;
      }
      S[v] = sum;
      F[v] = (uint16_t)(e[u][0] | (cnt << 12));
    }
  }
}

template <typename T, bool FIRST>
__device__ inline void lz_apply_vertex(const lz_args &a, const tile_t &t, const vtile_args &va,
                                       T *ql, T *S, uint16_t *F, const T *__restrict__ r,
                                       const T *__restrict__ pold, T *__restrict__ pnew,
                                       T *__restrict__ w, T beta, T m, acc_t<T> &sq,
                                       acc_t<T> &swq, acc_t<T> &sw) {
  const int nrows = t.row1 - t.row0;
  const int nreal = min(nrows, t.pstart + t.n - t.row0);
  if (threadIdx.x == 0) ql[VT_MAXW] = (T)0;  // what the lists' padding reads
  stage_window<T, FIRST, false>(t, va.halo + va.halo_off[t.ti], va.nhalo[t.ti], ql, nullptr, r,
                                pold, beta, m);
  const int nvert = va.vt_n[t.ti];
  const uint16_t *Ls = va.list + va.vt_loff[t.ti];
  if (va.nv == 8) vt_sums<T, 8>(Ls, nvert, ql, S, F);
  else vt_sums<T, 4>(Ls, nvert, ql, S, F);
  __syncthreads();
  // the (up to) four rows of a thread: their index words loaded first
  constexpr int R = RSB_TILE / RSB_BLOCK;
  u32x4 pks[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int row = t.row0 + threadIdx.x + k * blockDim.x;
    pks[k] = row < t.row1 ? __builtin_nontemporal_load((const u32x4 *)(va.rowv + (int64_t)row * 8))
                          : (u32x4){0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
  }
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int row = t.row0 + threadIdx.x + k * blockDim.x;
    if (row >= t.row1) continue;
    const int li = row - t.row0;
    const u32x4 pk = pks[k];
    const unsigned id[8] = {pk.x & 0xffffu, pk.x >> 16, pk.y & 0xffffu, pk.y >> 16,
                            pk.z & 0xffffu, pk.z >> 16, pk.w & 0xffffu, pk.w >> 16};
    const T qi = ql[li];
    T sum = (T)0;
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (id[j] != 0xffffu) {
        const unsigned f = F[id[j]];
        sum += (T)(f >> 12) * (qi - ql[f & 0xfffu]) - S[id[j]];
      }
    pnew[row] = qi;
    w[row] = sum;
    if (li < nreal) {
      sq.add(qi);
      swq.dot(sum, qi);
      sw.add(sum);
    }
  }
}

// FP32 only (its LDS: window + per-vertex sums and first-slot / count words
// = 28 KB, five workgroups per CU); a tile without the vertex form (vt_n < 0) takes the
// neighbour entries of the same window.  The level has no wide tiles.
template <typename T>
__global__ void __launch_bounds__(RSB_BLOCK) k_lz_apply_v(lz_args a, const T *__restrict__ pold,
                                                          T *__restrict__ pnew, T *__restrict__ w,
                                                          const T *__restrict__ basis,
                                                          int64_t mpad, vtile_args va,
                                                          tiled_args ta, int nlist) {
  __shared__ __attribute__((aligned(16))) T ql[VT_MAXW + 1];
  __shared__ T S[VT_MAXV];  // the neighbour-entry fallback keeps r of the tile rows here
  __shared__ uint16_t F[VT_MAXV];
  static_assert(VT_MAXV >= RSB_TILE, "the fallback's r of the tile rows lives in S");
  tile_t t;
  if (!tile_of(a, LZ_RUNNING, t, nlist)) return;
  const int s = t.s, it = a.iter[s];
  const T *__restrict__ r = basis + (size_t)it * mpad;  // r_it (its basis slot)
  const T beta = (T)SC(s, SC_BETA), m = (T)SC(s, SC_MEANP);
  acc_t<T> sq, swq, sw;
  if (va.vt_n[t.ti] >= 0) {
    if (it == 0)
      lz_apply_vertex<T, true>(a, t, va, ql, S, F, r, pold, pnew, w, beta, m, sq, swq, sw);
    else
      lz_apply_vertex<T, false>(a, t, va, ql, S, F, r, pold, pnew, w, beta, m, sq, swq, sw);
  } else {
    const int nh = ta.nhalo[t.ti];
    fold_t<T> fd;
    fd.wp = nullptr, fd.rn = nullptr, fd.alpha = (T)0;
    if (it == 0)
      lz_apply_tiled<T, true>(a, t, ta, nh, ql, S, r, pold, pnew, w, beta, m, sq, swq, sw, fd);
    else
      lz_apply_tiled<T, false>(a, t, ta, nh, ql, S, r, pold, pnew, w, beta, m, sq, swq, sw, fd);
  }
  double v[3] = {sq.value(), swq.value(), sw.value()};
  publish<3>(a, t, v);
}

template <typename T>
// r_{it+1} = r_it - alpha w into basis slot it + 1; partial r.r
__global__ void __launch_bounds__(RSB_BLOCK) k_lz_rupdate(lz_args a, T *__restrict__ basis,
                                                          int64_t mpad, const T *__restrict__ w) {
  tile_t t;
  if (!tile_of(a, LZ_RUNNING, t)) return;
  const int s = t.s;
  const T *__restrict__ r = basis + (size_t)a.iter[s] * mpad;
  T *__restrict__ rn = basis + (size_t)(a.iter[s] + 1) * mpad;
  const T alpha = (T)SC(s, SC_ALPHA);
  acc_t<T> acc;
  if constexpr (sizeof(T) == 4) {
    // FP32: four consecutive rows per thread in one 16-byte access each
    // (tile bounds and vector bases are multiples of 64 rows, so a group
    // lies wholly inside or outside the tile)
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    for (int row = t.row0 + 4 * (int)threadIdx.x; row < t.row1; row += 4 * (int)blockDim.x) {
      const f32x4 rv = *(const f32x4 *)(r + row), wv = *(const f32x4 *)(w + row);
      f32x4 x;
      x.x = rsub(rv.x, alpha, wv.x), x.y = rsub(rv.y, alpha, wv.y);
      x.z = rsub(rv.z, alpha, wv.z), x.w = rsub(rv.w, alpha, wv.w);
      *(f32x4 *)(rn + row) = x;
      acc.dot(x.x, x.x), acc.dot(x.y, x.y), acc.dot(x.z, x.z), acc.dot(x.w, x.w);
    }
    double v[1] = {acc.value()};
    publish<1>(a, t, v);
    return;
  }
  // the (up to) four rows of a thread: all loads first
  constexpr int R = RSB_TILE / RSB_BLOCK;
  T rv[R], wv[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int row = t.row0 + threadIdx.x + k * blockDim.x;
    rv[k] = row < t.row1 ? r[row] : (T)0;
    wv[k] = row < t.row1 ? w[row] : (T)0;
  }
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int row = t.row0 + threadIdx.x + k * blockDim.x;
    if (row >= t.row1) continue;
    const T x = rsub(rv[k], alpha, wv[k]);
    rn[row] = x;
    acc.dot(x, x);
  }
  double v[1] = {acc.value()};
  publish<1>(a, t, v);
}

// ---------------------------------------------------------------------------
// Tridiagonal eigen-solve of each finished pass: implicit QL with shifts, the
// steps of fiedler.c:328-432 (FP64, contraction off so every operation rounds
// as written).  One wavefront per segment: every lane runs the same scalar
// recurrence; the diagonal and off-diagonal are held in registers (below) so
// the serial chain never waits on LDS; the plane rotations of the
// eigenvector matrix (in LDS) are spread over the lanes (lane = matrix row,
// each lane touches only its own rows).  The coefficients y of the
// eigenvector with the smallest |eigenvalue| are kept.
#pragma clang fp contract(off)
__device__ inline double sgn_abs(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// d and e live in registers: lane k holds entries k and k + 64 (n <= 128).
// The serial QL chain reads them with readlane (a uniform index), the owning
// lane updates them, and the deflation scan (first m >= l with
// |e_m| / (|d_m| + |d_m+1|) < 1e-12) tests every m at once and takes the
// first hit from a ballot: the same tests, in parallel, give the same m.
__device__ inline double lane_get(double v, int i) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), i);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), i);
  return __longlong_as_double(
      (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}
struct lanes2 {
  double a, b;  // entries lane and lane + 64
  __device__ inline double get(int i) const { return i < RSB_WAVE ? lane_get(a, i) : lane_get(b, i - RSB_WAVE); }
  __device__ inline void set(int i, double x, int lane) {
    if (i < RSB_WAVE) {
      if (lane == i) a = x;
    } else if (lane == i - RSB_WAVE) {
      b = x;
    }
  }
};

__device__ void tqli_wave(int n, double *d, double *e, double *z, double *y, double *lam_out) {
  const int lane = threadIdx.x;
  for (int i = lane; i < n * n; i += RSB_WAVE) z[i] = 0.0;
  __syncthreads();
  for (int i = lane; i < n; i += RSB_WAVE) z[i * n + i] = 1.0;
  lanes2 D, E;
  D.a = lane < n ? d[lane] : 0.0;
  D.b = lane + RSB_WAVE < n ? d[lane + RSB_WAVE] : 0.0;
  E.a = lane < n ? e[lane] : 0.0;
  E.b = lane + RSB_WAVE < n ? e[lane + RSB_WAVE] : 0.0;
  __syncthreads();
  int failed = 0;
  for (int l = 0; l < n && !failed; l++) {
    int sweeps = 0, m;
    while (true) {
      // d[m + 1] next to d[m]: lane 63's neighbour is entry 64
      double dn_a = __shfl_down(D.a, 1, RSB_WAVE);
      const double d64 = lane_get(D.b, 0);
      if (lane == RSB_WAVE - 1) dn_a = d64;
      const double dn_b = __shfl_down(D.b, 1, RSB_WAVE);
      const int ma = lane, mb = lane + RSB_WAVE;
      const bool ta = ma >= l && ma < n - 1 && fabs(E.a) / (fabs(D.a) + fabs(dn_a)) < 1e-12;
      const bool tb = mb >= l && mb < n - 1 && fabs(E.b) / (fabs(D.b) + fabs(dn_b)) < 1e-12;
      const unsigned long long ba = __ballot(ta), bb = __ballot(tb);
      m = ba ? __ffsll((long long)ba) - 1 : bb ? RSB_WAVE + __ffsll((long long)bb) - 1 : n - 1;
      if (m == l) break;
      if (sweeps++ == 30) {
        failed = 1;
        break;
      }
      const double dl = D.get(l), el = E.get(l);
      double g = (D.get(l + 1) - dl) / (2.0 * el);
      double r = sqrt(g * g + 1.0);
      g = D.get(m) - dl + el / (g + sgn_abs(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      for (int i = m - 1; i >= l; i--) {
        const double ei = E.get(i);
        double f = s * ei, b = c * ei;
        double ei1;
        if (fabs(f) >= fabs(g)) {
          c = g / f;
          r = sqrt(c * c + 1.0);
          ei1 = f * r;
          s = 1.0 / r;
          c = c * s;
        } else {
          s = f / g;
          r = sqrt(s * s + 1.0);
          ei1 = g * r;
          c = 1.0 / r;
          s = s * c;
        }
        g = D.get(i + 1) - p;
        r = (D.get(i) - g) * s + 2.0 * c * b;
        p = s * r;
        const double di1 = g + p;
        g = c * r - b;
        E.set(i + 1, ei1, lane);
        D.set(i + 1, di1, lane);
        for (int k = lane; k < n; k += RSB_WAVE) {
          double zk1 = z[k * n + i + 1], zk = z[k * n + i];
          z[k * n + i + 1] = s * zk + c * zk1;
          z[k * n + i] = c * zk - s * zk1;
        }
      }
      D.set(l, D.get(l) - p, lane);
      E.set(l, g, lane);
      E.set(m, 0.0, lane);
    }
  }
  __syncthreads();  // rows of z written by other lanes are read below
  // eigenvalue of smallest magnitude (first on ties, fiedler.c:538-545)
  int imin = 0;
  double emin = fabs(D.get(0));
  for (int i = 1; i < n; i++) {
    const double v = fabs(D.get(i));
    if (v < emin) emin = v, imin = i;
  }
  if (failed) {
    // the reference leaves the rotations untransposed and unnormalised
    for (int j = lane; j < n; j += RSB_WAVE) y[j] = z[imin * n + j];
  } else {
    // eigenvector imin = column imin of z, normalised (sum in row order)
    if (lane == 0) {
      double nrm = 0.0;
      for (int i = 0; i < n; i++) nrm += z[i * n + imin] * z[i * n + imin];
      if (nrm > 0.0) nrm = sqrt(fabs(nrm));
      e[0] = 1.0 / nrm;
    }
    __syncthreads();
    double scale = e[0];
    for (int j = lane; j < n; j += RSB_WAVE) y[j] = z[j * n + imin] * scale;
  }
  const double lam = D.get(imin);
  if (lane == 0) *lam_out = lam;
  __syncthreads();
}
// The same QL steps for passes longer than 128 iterations: d and e in LDS
// (every lane computes each update and writes the same value), the
// eigenvector matrix z (n x n, row-major) in global memory, each lane
// rotating its own rows.  Operations and their order are those above, so
// results are bit-identical to it (and to the oracle).
__device__ void tqli_wave_mem(int n, double *d, double *e, double *z, double *y,
                              double *lam_out) {
  const int lane = threadIdx.x;
  for (int64_t i = lane; i < (int64_t)n * n; i += RSB_WAVE) z[i] = 0.0;
  __syncthreads();
  for (int i = lane; i < n; i += RSB_WAVE) z[(int64_t)i * n + i] = 1.0;
  __syncthreads();
  int failed = 0;
  for (int l = 0; l < n && !failed; l++) {
    int sweeps = 0, m;
    while (true) {
      for (m = l; m < n - 1; m++)
        if (fabs(e[m]) / (fabs(d[m]) + fabs(d[m + 1])) < 1e-12) break;
      if (m == l) break;
      if (sweeps++ == 30) {
        failed = 1;
        break;
      }
      const double dl = d[l], el = e[l];
      double g = (d[l + 1] - dl) / (2.0 * el);
      double r = sqrt(g * g + 1.0);
      g = d[m] - dl + el / (g + sgn_abs(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      for (int i = m - 1; i >= l; i--) {
        const double ei = e[i];
        double f = s * ei, b = c * ei;
        double ei1;
        if (fabs(f) >= fabs(g)) {
          c = g / f;
          r = sqrt(c * c + 1.0);
          ei1 = f * r;
          s = 1.0 / r;
          c = c * s;
        } else {
          s = f / g;
          r = sqrt(s * s + 1.0);
          ei1 = g * r;
          c = 1.0 / r;
          s = s * c;
        }
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        const double di1 = g + p;
        g = c * r - b;
        e[i + 1] = ei1;
        d[i + 1] = di1;
        for (int k = lane; k < n; k += RSB_WAVE) {
          double *zk = z + (int64_t)k * n;
          const double zk1 = zk[i + 1], zk0 = zk[i];
          zk[i + 1] = s * zk0 + c * zk1;
          zk[i] = c * zk0 - s * zk1;
        }
      }
      d[l] = d[l] - p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  __syncthreads();  // rows of z written by other lanes are read below
  int imin = 0;
  double emin = fabs(d[0]);
  for (int i = 1; i < n; i++) {
    const double v = fabs(d[i]);
    if (v < emin) emin = v, imin = i;
  }
  if (failed) {
    for (int j = lane; j < n; j += RSB_WAVE) y[j] = z[(int64_t)imin * n + j];
  } else {
    __shared__ double scale;
    if (lane == 0) {
      double nrm = 0.0;
      for (int i = 0; i < n; i++) nrm += z[(int64_t)i * n + imin] * z[(int64_t)i * n + imin];
      if (nrm > 0.0) nrm = sqrt(fabs(nrm));
      scale = 1.0 / nrm;
    }
    __syncthreads();
    for (int j = lane; j < n; j += RSB_WAVE) y[j] = z[(int64_t)j * n + imin] * scale;
  }
  if (lane == 0) *lam_out = d[imin];
  __syncthreads();
}

#pragma clang fp contract(fast)

// Fast variant for the FP32 path: the eigenpair of smallest |eigenvalue| of
// the tridiagonal T directly, instead of all of them.  Eigenvalue by
// multisection on Sturm counts (each of the 64 lanes counts the eigenvalues
// below its own shift; 65 sub-intervals per step), eigenvector by inverse
// iteration with the LDL^T factorisation of T - a I, a the lower end of the
// final interval (Parlett, The Symmetric Eigenvalue Problem, 3.3 and 4.3;
// the count recurrence with the pivot guard of LAPACK dstebz).  Sign: the
// component along the first Lanczos vector (the pass's start) is made
// non-negative.  It agrees with the QL solve to rounding, not bit for bit.
__device__ inline int sturm_count(int n, const lanes2 &D, const lanes2 &E2, double x,
                                  double pivmin) {
  int c = 0;
  double q = D.get(0) - x;
  if (fabs(q) < pivmin) q = -pivmin;
  c += q < 0.0;
  for (int i = 1; i < n; i++) {
    q = (D.get(i) - x) - E2.get(i - 1) / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

// eigenvalue number j (ascending) in [lo, hi]; returns the final interval
__device__ inline void multisect(int n, const lanes2 &D, const lanes2 &E2, double pivmin, int j,
                                 double &lo, double &hi) {
  const int lane = threadIdx.x;
  for (int step = 0; step < 16; step++) {
    const double w = hi - lo;
    if (!(w > 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin)) break;
    const double x = lo + w * (double)(lane + 1) / (double)(RSB_WAVE + 1);
    const int c = sturm_count(n, D, E2, x, pivmin);
    // lanes whose shift is at or below eigenvalue j: count <= j
    const unsigned long long below = __ballot(c <= j);
    // shifts increase with the lane: the highest such lane is the new lower
    // end, the next lane the new upper end
    const int k = below ? 63 - __clzll((long long)below) : -1;
    const double xk = k >= 0 ? __shfl(x, k, RSB_WAVE) : lo;
    const double xk1 = k + 1 < RSB_WAVE ? __shfl(x, k + 1, RSB_WAVE) : hi;
    lo = xk, hi = xk1;
  }
}

__device__ void eig_min_wave(int n, const double *d, const double *e, double *y,
                             double *lam_out, double *work) {
  const int lane = threadIdx.x;
  lanes2 D, E2;
  D.a = lane < n ? d[lane] : 0.0;
  D.b = lane + RSB_WAVE < n ? d[lane + RSB_WAVE] : 0.0;
  const double ea = lane < n - 1 ? e[lane] : 0.0, eb = lane + RSB_WAVE < n - 1 ? e[lane + RSB_WAVE] : 0.0;
  E2.a = ea * ea, E2.b = eb * eb;
  // Gershgorin bounds and the pivot guard
  double glo = DBL_MAX, ghi = -DBL_MAX, emax = 0.0, tnorm = 0.0;
  for (int i = lane; i < n; i += RSB_WAVE) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    glo = fmin(glo, d[i] - r), ghi = fmax(ghi, d[i] + r);
    tnorm = fmax(tnorm, fabs(d[i]) + r);
    if (i < n - 1) emax = fmax(emax, e[i] * e[i]);
  }
  for (int o = RSB_WAVE / 2; o > 0; o >>= 1) {
    glo = fmin(glo, __shfl_xor(glo, o, RSB_WAVE));
    ghi = fmax(ghi, __shfl_xor(ghi, o, RSB_WAVE));
    emax = fmax(emax, __shfl_xor(emax, o, RSB_WAVE));
    tnorm = fmax(tnorm, __shfl_xor(tnorm, o, RSB_WAVE));
  }
  const double pivmin = DBL_MIN * fmax(1.0, emax);
  const double span = fmax(ghi - glo, DBL_MIN);
  glo -= 2.0 * DBL_EPSILON * span, ghi += 2.0 * DBL_EPSILON * span;
  // smallest eigenvalue; if it is negative, compare the eigenvalues either
  // side of 0 by magnitude
  double lo = glo, hi = ghi;
  multisect(n, D, E2, pivmin, 0, lo, hi);
  double lam = 0.5 * (lo + hi), shift = lo;
  if (lam < 0.0) {
    const int k = sturm_count(n, D, E2, 0.0, pivmin);
    double l1 = glo, h1 = ghi;
    multisect(n, D, E2, pivmin, k - 1, l1, h1);
    lam = 0.5 * (l1 + h1), shift = l1;
    if (k < n) {
      double l2 = glo, h2 = ghi;
      multisect(n, D, E2, pivmin, k, l2, h2);
      const double lam2 = 0.5 * (l2 + h2);
      if (fabs(lam2) < fabs(lam)) lam = lam2, shift = l2;
    }
  }
  // inverse iteration (two solves) on lane 0: (T - shift I) = L D L^T, a
  // pivot below eps * |T|_inf replaced by +-eps * |T|_inf; each solve is
  // rescaled by its largest entry before normalising
  if (lane == 0) {
    double *piv = work, *x = work + n;
    const double tau = DBL_EPSILON * fmax(tnorm, DBL_MIN);
    double q = d[0] - shift;
    if (fabs(q) < tau) q = q < 0.0 ? -tau : tau;
    piv[0] = q;
    for (int i = 1; i < n; i++) {
      q = (d[i] - shift) - e[i - 1] * e[i - 1] / q;
      if (fabs(q) < tau) q = q < 0.0 ? -tau : tau;
      piv[i] = q;
    }
    for (int i = 0; i < n; i++) x[i] = 1.0;
    for (int it = 0; it < 2; it++) {
      // L z = x (unit lower, l_i = e_i / piv_i), D w = z, L^T v = w
      for (int i = 1; i < n; i++) x[i] -= e[i - 1] / piv[i - 1] * x[i - 1];
      for (int i = 0; i < n; i++) x[i] /= piv[i];
      for (int i = n - 2; i >= 0; i--) x[i] -= e[i] / piv[i] * x[i + 1];
      double amax = 0.0;
      for (int i = 0; i < n; i++) amax = fmax(amax, fabs(x[i]));
      if (!(amax > 0.0) || amax > DBL_MAX) {
        for (int i = 0; i < n; i++) x[i] = i == 0 ? 1.0 : 0.0;
        break;
      }
      double nrm = 0.0;
      for (int i = 0; i < n; i++) {
        x[i] /= amax;
        nrm += x[i] * x[i];
      }
      nrm = sqrt(nrm);
      int f = 0;
      while (f < n - 1 && x[f] == 0.0) f++;
      const double sc = (x[f] < 0.0 ? -1.0 : 1.0) / nrm;
      for (int i = 0; i < n; i++) x[i] *= sc;
    }
    *lam_out = lam;
  }
  __syncthreads();
  for (int j = lane; j < n; j += RSB_WAVE) y[j] = work[n + j];
  __syncthreads();
}

__global__ void __launch_bounds__(RSB_WAVE) k_lz_eigmin(lz_args a) {
  const int s = blockIdx.x;
  if (a.status[s] != LZ_PASSDONE) return;
  extern __shared__ double lds[];
  const int n = a.k[s], M = a.miter_max;
  double *d = lds, *e = d + M, *work = e + M;
  const double *Td = a.T + (size_t)s * 2 * M, *Tu = Td + M;
  for (int i = threadIdx.x; i < n; i += RSB_WAVE) {
    d[i] = Td[i];
    e[i] = (i < n - 1) ? Tu[i] : 0.0;
  }
  __syncthreads();
  eig_min_wave(n, d, e, a.y + (size_t)s * M, &SC(s, SC_LAMBDA), work);
}

__global__ void __launch_bounds__(RSB_WAVE) k_eigmin_test(int n, const double *dd,
                                                          const double *ee, double *y,
                                                          double *lam) {
  extern __shared__ double lds[];
  double *d = lds, *e = d + n, *work = e + n;
  for (int i = threadIdx.x; i < n; i += RSB_WAVE) d[i] = dd[i], e[i] = i < n - 1 ? ee[i] : 0.0;
  __syncthreads();
  eig_min_wave(n, d, e, y, lam, work);
}

__global__ void __launch_bounds__(RSB_WAVE) k_lz_tqli(lz_args a) {
  const int s = blockIdx.x;
  if (a.status[s] != LZ_PASSDONE) return;
  extern __shared__ double lds[];
  const int n = a.k[s], M = a.miter_max;
  double *d = lds, *e = d + M, *z = e + M;
  const double *Td = a.T + (size_t)s * 2 * M, *Tu = Td + M;
  for (int i = threadIdx.x; i < n; i += RSB_WAVE) {
    d[i] = Td[i];
    e[i] = (i < n - 1) ? Tu[i] : 0.0;
  }
  __syncthreads();
  tqli_wave(n, d, e, z, a.y + (size_t)s * M, &SC(s, SC_LAMBDA));
}

// Passes longer than 128 iterations (both precisions): segments s0 +
// blockIdx.x, each with its n x n eigenvector matrix at zbuf + blockIdx.x *
// M * M.
__global__ void __launch_bounds__(RSB_WAVE) k_lz_tqli_g(lz_args a, int s0, double *zbuf) {
  const int s = s0 + blockIdx.x;
  if (s >= a.nseg || a.status[s] != LZ_PASSDONE) return;
  extern __shared__ double lds[];
  const int n = a.k[s], M = a.miter_max;
  double *d = lds, *e = d + M;
  const double *Td = a.T + (size_t)s * 2 * M, *Tu = Td + M;
  for (int i = threadIdx.x; i < n; i += RSB_WAVE) {
    d[i] = Td[i];
    e[i] = (i < n - 1) ? Tu[i] : 0.0;
  }
  __syncthreads();
  tqli_wave_mem(n, d, e, zbuf + (size_t)blockIdx.x * M * M, a.y + (size_t)s * M,
                &SC(s, SC_LAMBDA));
}

__global__ void __launch_bounds__(RSB_WAVE) k_tqli_test_g(int n, const double *dd,
                                                          const double *ee, double *y,
                                                          double *lam, double *z) {
  extern __shared__ double lds[];
  double *d = lds, *e = d + n;
  for (int i = threadIdx.x; i < n; i += RSB_WAVE) d[i] = dd[i], e[i] = i < n - 1 ? ee[i] : 0.0;
  __syncthreads();
  tqli_wave_mem(n, d, e, z, y, lam);
}

// Stand-alone tridiagonal solve for the parity tests.
__global__ void __launch_bounds__(RSB_WAVE) k_tqli_test(int n, const double *dd,
                                                        const double *ee, double *y,
                                                        double *lam) {
  extern __shared__ double lds[];
  double *d = lds, *e = d + n, *z = e + n;
  for (int i = threadIdx.x; i < n; i += RSB_WAVE) d[i] = dd[i], e[i] = i < n - 1 ? ee[i] : 0.0;
  __syncthreads();
  tqli_wave(n, d, e, z, y, lam);
}

// Ritz vector (fiedler.c:547-553): f = sum_j (r_j / |r_j|) y_j, the basis
// slots holding the residuals r_j unscaled (FP64 accumulation of r_j times
// y_j / |r_j|, j ascending), then its mean (ortho).  Decides restart or
// finish.
template <typename T>
__global__ void __launch_bounds__(RSB_BLOCK) k_lz_ritz(lz_args a, const T *__restrict__ basis,
                                                       int64_t mpad, T *__restrict__ f) {
  tile_t t;
  if (!tile_of(a, LZ_PASSDONE, t)) return;
  const int s = t.s, k = a.k[s];
  const double *y = a.y + (size_t)s * a.miter_max;
  const double *rn = a.rn + (size_t)s * (a.miter_max + 1);
  double v[1] = {0.0};
  if constexpr (sizeof(T) == 4) {
    // FP32: four consecutive rows per thread, one 16-byte load per basis
    // vector, eight vectors in flight (vectors past k read vector k - 1 and
    // are skipped); each row sums j ascending, as below
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    constexpr int U = 8;
    for (int row = t.row0 + 4 * (int)threadIdx.x; row < t.row1; row += 4 * (int)blockDim.x) {
      double x[4] = {0.0, 0.0, 0.0, 0.0};
      for (int j0 = 0; j0 < k; j0 += U) {
        f32x4 b[U];
#pragma unroll
        for (int u = 0; u < U; u++)
          b[u] = *(const f32x4 *)(basis + (size_t)min(j0 + u, k - 1) * mpad + row);
#pragma unroll
        for (int u = 0; u < U; u++)
          if (j0 + u < k) {
            const double yj = y[j0 + u] * rn[j0 + u];
            x[0] += (double)b[u].x * yj, x[1] += (double)b[u].y * yj;
            x[2] += (double)b[u].z * yj, x[3] += (double)b[u].w * yj;
          }
      }
      f32x4 o;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const float xq = row + q - t.pstart < t.n ? (float)x[q] : 0.f;
        o[q] = xq;
        v[0] += (double)xq;
      }
      *(f32x4 *)(f + row) = o;
    }
    publish<1>(a, t, v);
    return;
  }
  // The (up to) four rows of a thread together and eight basis vectors at a
  // time, every load unconditional (rows past the tile read its last row,
  // vectors past k read vector k - 1 and are skipped): 32 loads in flight
  // per lane, no branches between them.  Each row sums j ascending.
  constexpr int R = RSB_TILE / RSB_BLOCK, U = 8;
  int rows[R];
  double x[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    rows[q] = min(t.row0 + (int)threadIdx.x + q * (int)blockDim.x, t.row1 - 1);
    x[q] = 0.0;
  }
  for (int j0 = 0; j0 < k; j0 += U) {
    T b[U][R];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const T *bj = basis + (size_t)min(j0 + u, k - 1) * mpad;
#pragma unroll
      for (int q = 0; q < R; q++) b[u][q] = bj[rows[q]];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (j0 + u < k) {
        const double yj = y[j0 + u] * rn[j0 + u];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] += (double)b[u][q] * yj;
      }
  }
#pragma unroll
  for (int q = 0; q < R; q++) {
    const int row = t.row0 + (int)threadIdx.x + q * (int)blockDim.x;
    if (row >= t.row1) continue;
    const double xq = row - t.pstart < t.n ? x[q] : 0.0;
    f[row] = (T)xq;
    v[0] += (double)(T)xq;
  }
  publish<1>(a, t, v);
}

// Final normalisation (fiedler.c:606-616) and scatter to compact positions.
template <typename T>
__global__ void __launch_bounds__(RSB_BLOCK) k_lz_finish(lz_args a, int phase,
                                                         const T *__restrict__ f,
                                                         double *__restrict__ fied,
                                                         double *__restrict__ xrow) {
  tile_t t;
  if (!tile_of(a, LZ_FINISHED, t)) return;
  const int s = t.s;
  const double mf = SC(s, SC_MEANF);
  if (phase == 1) {
    const double ni = SC(s, SC_NORMF);
    const int cstart = a.seg[s];
    for (int row = t.row0 + threadIdx.x; row < t.row1; row += blockDim.x) {
      int i = row - t.pstart;
      double x = i < t.n ? ((double)f[row] - mf) * ni : 0.0;
      xrow[row] = x;
      if (i < t.n) fied[cstart + i] = x;
    }
    return;
  }
  double v[1] = {0.0};
  for (int row = t.row0 + threadIdx.x; row < t.row1; row += blockDim.x) {
    int i = row - t.pstart;
    if (i < t.n) {
      double x = (double)f[row] - mf;
      v[0] += x * x;
    }
  }
  publish<1>(a, t, v);
}

// Eigen-residual of the final vector x (FP64): lambda_R = x.Lx, then
// |Lx - lambda_R x| / lambda_R.
__global__ void __launch_bounds__(RSB_BLOCK) k_lz_resid(lz_args a, int phase,
                                                        const double *__restrict__ x,
                                                        const int64_t *__restrict__ slice_off,
                                                        const int *__restrict__ slice_w,
                                                        const uint32_t *__restrict__ ell) {
  tile_t t;
  if (!tile_of(a, LZ_FINISHED, t)) return;
  const int s = t.s;
  const double lam = phase ? SC(s, SC_RES) : 0.0;
  double v[1] = {0.0};
  for (int row = t.row0 + threadIdx.x; row < t.row1; row += blockDim.x) {
    const int sl = row >> 6, lane = row & 63;
    const uint32_t *e = ell + slice_off[sl] + lane;
    const double xi = x[row];
    double sum = 0.0;
    for (int j = 0; j < slice_w[sl]; j++) {
      uint32_t q = e[(size_t)j * RSB_WAVE];
      sum += (double)(q >> 28) * (xi - x[q & 0x0fffffffu]);
    }
    if (phase == 0) v[0] += sum * xi;
    else {
      double d = sum - lam * xi;
      v[0] += d * d;
    }
  }
  publish<1>(a, t, v);
}

// ---------------------------------------------------------------------------
// Segment finalisation: the sums of a segment's tile partials (fixed order)
// and the scalar updates of the phase that produced them.  One workgroup per
// segment; segments not in the phase's state leave at once.
enum {
  FIN_START0 = 0, FIN_START1, FIN_SETUP0, FIN_SETUP1, FIN_APPLY, FIN_RUPDATE, FIN_RITZ,
  FIN_FINISH, FIN_RES0, FIN_RES1, FIN_APPLYR
};

__device__ inline int fin_status(int op) {
  switch (op) {
    case FIN_START0: case FIN_START1: case FIN_SETUP0: case FIN_SETUP1: return LZ_SETUP;
    case FIN_APPLY: case FIN_RUPDATE: case FIN_APPLYR: return LZ_RUNNING;
    case FIN_RITZ: return LZ_PASSDONE;
    default: return LZ_FINISHED;
  }
}

// Totals of partial slots k0 .. k0 + nv - 1 (nv <= 3) of a segment: 8 tiles
// per thread in flight (large segments have up to 16k tiles), a fixed order
// of summation (each slot's partials are contiguous).  Thread 0 holds them.
template <int NV>
__device__ inline void fin_sums(const lz_args &a, int first, int count, int k0,
                                double (&tot)[NV], double *red) {
#pragma unroll
  for (int k = 0; k < NV; k++) tot[k] = 0.0;
  for (int i0 = threadIdx.x; i0 < count; i0 += 8 * blockDim.x) {
    double v[NV][8];
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int i = i0 + u * blockDim.x;
        v[k][u] = i < count ? a.partial[(size_t)(k0 + k) * a.pstride + first + i] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
      for (int k = 0; k < NV; k++) tot[k] += v[k][u];
  }
  block_sum<NV>(tot, red);
}

// The stop test and the next coefficients after the residual update of
// iteration it, from rtr = r.r of the new residual (fiedler.c:486-509).
__device__ inline void rupdate_step(const lz_args &a, int s, double rtr) {
  const int it = a.iter[s], mi = a.miter[s];
  double rnorm = sqrt(rtr);
  double rtz1 = SC(s, SC_RTZ1), rtz2 = SC(s, SC_RTZ2), beta = SC(s, SC_BETA);
  double pap = SC(s, SC_PAP), pap_old = SC(s, SC_PAPOLD);
  double *Td = a.T + (size_t)s * 2 * a.miter_max, *Tu = Td + a.miter_max;
  if (it == 0) {
    Td[0] = pap / rtz1;
  } else {
    Td[it] = (beta * beta * pap_old + pap) / rtz1;
    Tu[it - 1] = -beta * pap_old / sqrt(rtz2 * rtz1);
  }
  SC(s, SC_RTR) = rtr;
  SC(s, SC_RNORM) = rnorm;
  SC(s, SC_RNI) = 1.0 / rnorm;
  a.rn[(size_t)s * (a.miter_max + 1) + it + 1] = 1.0 / rnorm;
  if (rnorm < SC(s, SC_RTOL) || it + 1 == mi) {
    a.k[s] = it + 1;
    a.status[s] = LZ_PASSDONE;
    atomicSub(&a.running[0], 1);
  } else {
    a.iter[s] = it + 1;
    SC(s, SC_RTZ2) = rtz1;
    SC(s, SC_RTZ1) = rtr;
    SC(s, SC_BETA) = rtr / rtz1;
  }
}

// NT threads per segment: 256, or 512 for the per-iteration sums of levels
// whose segments hold more than LZ_FIN_WIDE tiles (more partials in flight;
// on thousands of small segments the wider workgroup costs more than it
// saves)
template <int NT>
__global__ void __launch_bounds__(NT) k_lz_fin(lz_args a, int op) {
  __shared__ double red[(NT / RSB_WAVE) * 6];
  const int s = a.act_segs ? a.act_segs[blockIdx.x] : blockIdx.x;
  if (a.status[s] != fin_status(op)) return;
  const int first = a.seg_tiles[2 * s], count = a.seg_tiles[2 * s + 1];
  // only the slots the phase wrote are read
  // one pass over the slots the phase wrote (all of them in flight together)
  double tot[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (op == FIN_APPLYR) {
    fin_sums<6>(a, first, count, 0, tot, red);
  } else if (op == FIN_APPLY) {
    double t3[3];
    fin_sums<3>(a, first, count, 0, t3, red);
    tot[0] = t3[0], tot[1] = t3[1], tot[2] = t3[2];
  } else {
    double t1[1];
    fin_sums<1>(a, first, count, 0, t1, red);
    tot[0] = t1[0];
  }
  if (threadIdx.x != 0) return;
  const double n = a.seg[a.nseg + s];
  switch (op) {
    case FIN_START0: SC(s, SC_M0) = tot[0] / n; break;
    case FIN_START1:
      if (a.mode == 2 && a.max_n > 0 && a.seg[a.nseg + s] > a.max_n) {
        // above the size limit: no learned attempt, reference start next
        SC(s, SC_N0) = 0.0;
        SC(s, SC_FALLBACK) = -1.0;
        a.k[s] = a.miter[s];
        a.status[s] = LZ_FINISHED;
      } else if (a.mode == 2 && !(tot[0] > 0.0 && tot[0] < INFINITY)) {
        // constant or non-finite prediction: no pass, straight to the fallback
        SC(s, SC_N0) = 0.0;
        a.k[s] = a.miter[s];
        a.status[s] = LZ_FINISHED;
      } else {
        SC(s, SC_N0) = 1.0 / sqrt(tot[0]);
      }
      break;
    case FIN_SETUP0: SC(s, SC_TMP) = tot[0] / n; break;
    case FIN_SETUP1: {  // fiedler.c:441-456
      double rtr = tot[0], rn = sqrt(rtr);
      SC(s, SC_RTR) = rtr;
      SC(s, SC_RNORM) = rn;
      SC(s, SC_RNORM0) = rn;
      SC(s, SC_RTOL) = rn * a.tol;
      SC(s, SC_RNI) = 1.0 / rn;
      a.rn[(size_t)s * (a.miter_max + 1)] = 1.0 / rn;
      SC(s, SC_RTZ1) = rtr;  // iteration 0: rtz2 = 1 (initial rtz1), rtz1 = rtr
      SC(s, SC_RTZ2) = 1.0;
      SC(s, SC_BETA) = 0.0;
      SC(s, SC_PAP) = 0.0;
      SC(s, SC_MEANP) = 0.0;
      a.iter[s] = 0;
      a.status[s] = LZ_RUNNING;
      atomicAdd(&a.running[0], 1);
      break;
    }
    case FIN_APPLY: {  // mean of p, pap = w.(p - mean), alpha (fiedler.c:470-478)
      double mq = tot[0] / n;
      double pap = tot[1] - mq * tot[2];
      SC(s, SC_MEANP) = mq;
      SC(s, SC_PAPOLD) = SC(s, SC_PAP);
      SC(s, SC_PAP) = pap;
      SC(s, SC_ALPHA) = SC(s, SC_RTZ1) / pap;
      break;
    }
    case FIN_RUPDATE: rupdate_step(a, s, tot[0]); break;  // tot: r.r
    case FIN_APPLYR: {  // FIN_APPLY, then the update's r.r from the sums r.r, r.w, w.w
      double mq = tot[0] / n;
      double pap = tot[1] - mq * tot[2];
      SC(s, SC_MEANP) = mq;
      SC(s, SC_PAPOLD) = SC(s, SC_PAP);
      SC(s, SC_PAP) = pap;
      const double alpha = SC(s, SC_RTZ1) / pap;
      SC(s, SC_ALPHA) = alpha;
      // |r - af w|^2 with af = alpha as the next launch rounds it: the FP64
      // sums of exact FP32 products, from this iteration's stored r and w,
      // so the value does not drift over the iterations; it differs from
      // the sum of squares of the rounded new r by at most 2^-23 relative
      const double af = (double)(float)alpha;
      double rtr = tot[3] - 2.0 * af * tot[4] + af * af * tot[5];
      if (!(rtr > 0.0)) rtr = 0.0;
      SC(s, SC_NFOLD) += 1.0;
      rupdate_step(a, s, rtr);
      break;
    }
    case FIN_RITZ: {  // mean of the Ritz vector; restart or finish (fiedler.c:529-553)
      const int k = a.k[s];
      SC(s, SC_MEANF) = tot[0] / n;
      int np = a.npass[s] + 1;
      a.npass[s] = np;
      a.niter[s] = (np - 1) * a.miter[s] + k;
      const double lam = SC(s, SC_LAMBDA), lprev = SC(s, SC_LAMPREV);
      SC(s, SC_LAMPREV) = lam;
      bool stalled = a.stall > 0.0 && np >= max(2, a.stall_min) &&
                     fabs(lam - lprev) <= a.stall * fabs(lam);
      if (a.plateau > 0.0 && k >= 2) {
        // the Ritz pair's residual |L f - lam f| / (lam |f|) from the Lanczos
        // relation: |t_{k+1,k}| |y_k| / (|y| lam), t_{k+1,k} the next
        // off-diagonal of T, -beta_k pap_{k-1} / sqrt(rtz_{k-1} rtz_k) with
        // beta_k = rtz_k / rtz_{k-1} (the state the last update left)
        const double rk = SC(s, SC_RTR), rk1 = SC(s, SC_RTZ1), pap = SC(s, SC_PAP);
        const double t = (rk / rk1) * pap / sqrt(rk1 * rk);
        const double *y = a.y + (size_t)s * a.miter_max;
        double yy = 0.0;
        for (int j = 0; j < k; j++) yy += y[j] * y[j];
        const double rho = fabs(t * y[k - 1]) / (sqrt(yy) * fabs(lam));
        stalled |= np >= 4 && rho <= a.plateau_max && rho >= a.plateau * SC(s, SC_RHO2);
        SC(s, SC_RHO2) = SC(s, SC_RHO1);
        SC(s, SC_RHO1) = rho;
      }
      if (k == a.miter[s] && np < a.mpass && !stalled) {
        a.status[s] = LZ_SETUP;
        atomicAdd(&a.running[1], 1);
      } else {
        a.status[s] = LZ_FINISHED;
      }
      break;
    }
    case FIN_FINISH: SC(s, SC_NORMF) = 1.0 / sqrt(tot[0]); break;
    case FIN_RES0: SC(s, SC_RES) = tot[0]; break;
    case FIN_RES1: SC(s, SC_TMP) = sqrt(tot[0]) / SC(s, SC_RES); break;
  }
}

// Lists of the tiles and segments in a given state (order irrelevant: every
// tile writes its partial at its own index and k_lz_fin sums them in tile
// order).
__global__ void k_lz_compact(lz_args a, int ntiles, int want, int *__restrict__ tiles,
                             int *__restrict__ segs, int *__restrict__ counts) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntiles && a.status[a.tile[3 * i]] == want) tiles[atomicAdd(&counts[0], 1)] = i;
  if (i < a.nseg && a.status[i] == want) segs[atomicAdd(&counts[1], 1)] = i;
}

__global__ void k_lz_init_state(lz_args a, int miter_opt) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nseg) return;
  int n = a.seg[a.nseg + s], p = a.seg[2 * a.nseg + s];
  a.status[s] = (p > 1 && n >= 2) ? LZ_SETUP : LZ_IDLE;
  a.miter[s] = n < miter_opt ? n : miter_opt;  // fiedler.c:522
  a.npass[s] = 0, a.niter[s] = 0, a.iter[s] = 0, a.k[s] = 0;
  for (int j = 0; j < LZ_NSC; j++) SC(s, j) = 0.0;
}

// Learned start: segments that missed tol within their passes (k == miter
// on the last pass, as the restart rule of fiedler.c:529-553 reads it) start
// over from the reference start; the iterations spent so far are kept.
__global__ void k_lz_fallback(lz_args a, int *__restrict__ count) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nseg) return;
  if (a.status[s] != LZ_FINISHED || a.k[s] != a.miter[s]) return;
  const bool skipped = SC(s, SC_FALLBACK) < 0.0;  // above the size limit
  SC(s, SC_WARMIT) = skipped ? 0.0 : a.niter[s];
  SC(s, SC_FALLBACK) = skipped ? -1.0 : 1.0;
  SC(s, SC_MEANF) = 0.0;
  a.npass[s] = 0, a.niter[s] = 0, a.iter[s] = 0, a.k[s] = 0;
  a.status[s] = LZ_SETUP;
  atomicAdd(count, 1);
}

// Segments that cannot be bisected spectrally (fewer than 2 elements) keep
// Fiedler value 0.
__global__ void k_lz_zero_small(lz_args a, double *fied) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nseg) return;
  int n = a.seg[a.nseg + s], p = a.seg[2 * a.nseg + s];
  if (p > 1 && n == 1) fied[a.seg[s]] = 0.0;
}

// ---------------------------------------------------------------------------
template <typename T>
static int run_lanczos(rsbgpu *g, int miter_opt, int mpass, double tol, int init_mode,
                       rsbgpu_fiedler_out *out) {
  const int nseg = g->nseg, ntiles = g->ntiles;
  const int64_t mpad = g->mpad;
  int miter_max = 1;
  for (int s = 0; s < nseg; s++)
    if (g->h_p[s] > 1 && g->h_n[s] >= 2) miter_max = std::max(miter_max, std::min(miter_opt, g->h_n[s]));
  // passes of more than 128 iterations take the tridiagonal solve with its
  // eigenvector matrix in global memory (k_lz_tqli_g)
  const bool long_pass = miter_max > 128;
  // buffers
  size_t vb = (size_t)mpad * sizeof(T);
  if (vb > g->vec_bytes) {
    for (void **q : {&g->d_p, &g->d_p2, &g->d_w, &g->d_f, &g->d_w2}) {
      if (*q) RSB_TRY(rsb_dev_free(*q));
      *q = nullptr;
      RSB_TRY(rsb_dev_malloc(q, vb + 256));
    }
    g->vec_bytes = vb;
  }
  // the pass's residuals r_0 .. r_miter (the last one written, not used)
  size_t bb = vb * ((size_t)miter_max + 1);
  if (bb > g->basis_bytes) {
    if (g->d_basis) RSB_TRY(rsb_dev_free(g->d_basis));
    g->d_basis = nullptr;
    RSB_TRY(rsb_dev_malloc(&g->d_basis, bb + 256));
    g->basis_bytes = bb;
  }
  double *xrow = nullptr;
  RSB_TRY(RSB_SCRATCH(xrow, sizeof(double) * (mpad + 1)));
  lz_state &L = g->lz;
  if (nseg > g->lz_nseg_cap || miter_max > g->lz_miter_cap) {
    // iter, k, npass, niter and miter are carved from the status block
    for (void *q : {(void *)L.status, (void *)L.sc, (void *)L.T, (void *)L.y, (void *)L.rn})
      if (q) RSB_TRY(rsb_dev_free(q));
    int cap = std::max(nseg, 1);
    RSB_TRY(rsb_dev_malloc((void **)&L.status, sizeof(int) * cap * 6));
    L.iter = L.status + cap, L.k = L.status + 2 * cap, L.npass = L.status + 3 * cap;
    L.niter = L.status + 4 * cap;
    RSB_TRY(rsb_dev_malloc((void **)&L.sc, sizeof(double) * cap * LZ_NSC));
    RSB_TRY(rsb_dev_malloc((void **)&L.T, sizeof(double) * cap * 2 * miter_max));
    RSB_TRY(rsb_dev_malloc((void **)&L.y, sizeof(double) * cap * miter_max));
    RSB_TRY(rsb_dev_malloc((void **)&L.rn, sizeof(double) * cap * ((size_t)miter_max + 1)));
    g->lz_nseg_cap = cap, g->lz_miter_cap = miter_max;
  }
  if (ntiles > g->lz_tile_cap) {
    if (L.partial) RSB_TRY(rsb_dev_free(L.partial));
    RSB_TRY(rsb_dev_malloc((void **)&L.partial, sizeof(double) * 6 * std::max(ntiles, 1)));
    g->lz_tile_cap = ntiles;
  }
  if (!L.running) RSB_TRY(rsb_dev_malloc((void **)&L.running, sizeof(int) * 2));
  lz_args a;
  a.act_tiles = nullptr, a.act_segs = nullptr;
  a.seg = g->d_seg, a.tile = g->d_tile, a.seg_tiles = g->d_seg_tiles, a.row_elem = g->d_row_elem;
  a.nseg = nseg;
  a.status = L.status, a.iter = L.iter, a.k = L.k, a.npass = L.npass, a.niter = L.niter;
  a.miter = L.status + 5 * std::max(g->lz_nseg_cap, 1);
  a.sc = L.sc, a.T = L.T, a.y = L.y, a.partial = L.partial, a.rn = L.rn,
  a.running = L.running;
  a.mpass = mpass, a.mode = init_mode, a.tol = tol, a.miter_max = miter_max;
  {
    const char *sv = getenv("PARRSB_AMD_LZ_STALL"), *sm = getenv("PARRSB_AMD_LZ_STALL_MIN");
    a.stall = sv ? atof(sv) : 0.0;
    a.stall_min = sm ? atoi(sm) : 2;
    const char *pv = getenv("PARRSB_AMD_LZ_PLATEAU"), *pm = getenv("PARRSB_AMD_LZ_PLATEAU_MAX");
    a.plateau = pv ? atof(pv) : 0.0;
    a.plateau_max = pm ? atof(pm) : 1e-2;
  }
  a.pstride = std::max(g->lz_tile_cap, 1);
  a.max_n = g->lws_max_n;
  T *p = (T *)g->d_p, *p2 = (T *)g->d_p2, *w = (T *)g->d_w, *f = (T *)g->d_f,
    *basis = (T *)g->d_basis;
  hipStream_t st = g->stream;
#define FIN(op)                                           \
  do {                                                    \
    k_lz_fin<RSB_BLOCK><<<nseg, RSB_BLOCK, 0, st>>>(a, (op)); \
    RSB_LAUNCH_CHECK();                                   \
  } while (0)
  // compacted lists of running tiles / segments for the iteration launches
  int *act = nullptr;
  RSB_TRY(RSB_SCRATCH(act, sizeof(int) * ((size_t)ntiles + nseg + 2)));
  int *act_t = act, *act_s = act + ntiles, *act_n = act + ntiles + nseg;
  int nact[2] = {ntiles, nseg};
  auto compact = [&](int want) -> int {
    RSB_TRY(hipMemsetAsync(act_n, 0, sizeof(int) * 2, st));
    k_lz_compact<<<nblocks(std::max(ntiles, nseg), 256), 256, 0, st>>>(a, ntiles, want, act_t,
                                                                        act_s, act_n);
    RSB_LAUNCH_CHECK();
    RSB_TRY(hipMemcpyAsync(nact, act_n, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
    RSB_TRY(hipStreamSynchronize(st));
    return 0;
  };
  lz_args ai = a;  // iteration launches over the running lists
  auto xgrid = [](int n) { return (unsigned)((n + LZ_XCDS - 1) / LZ_XCDS * LZ_XCDS); };
  ai.act_tiles = act_t, ai.act_segs = act_s;
  // The FP64 apply reads the 16-bit entries: with the 12-bit form it took
  // 0.253 against 0.233 ms per launch on the 256^3 box (FP32: 0.163 against
  // 0.177), the class sums in doubles costing more than the bytes they save.
  const bool use12 = g->e12_ok && (sizeof(T) == 4 || g->e12_f64);
  tiled_args ta{g->d_slice_off, g->d_slice_w, g->d_ell,  g->d_tnhalo,
                g->d_thalo_off, g->d_halo,    g->d_off16, g->d_w16,
                g->d_e16,       g->d_off12,   g->d_w12,   use12 ? g->d_e12 : nullptr,
                g->d_c12};
  vtile_args va{g->d_vt_n,  g->d_vt_loff,    g->d_vt_list, g->d_vt_rowv,
                g->d_tnhalo, g->d_thalo_off, g->d_halo,    g->nv};
  // the vertex form serves the FP32 path (vectors of T = float)
  const bool vform = g->vt_ok && sizeof(T) == 4;
  // FP32 with the neighbour entries: a pass that starts with at most
  // fold_tiles running tiles folds the r-update into the next apply
  // (k_lz_apply<float, true>, FIN_APPLYR: two launches per iteration instead
  // of four; w alternates between two buffers by iteration parity).
  // On more tiles the extra halo gathers of w outweigh the saved pass.
  // PARRSB_AMD_LZ_FOLD_TILES sets the limit (0: never).
  const char *fev = getenv("PARRSB_AMD_LZ_FOLD_TILES");
  const int fold_tiles = (sizeof(T) == 4 && !vform) ? (fev ? atoi(fev) : LZ_FOLD_TILES) : 0;
  bool fold = false;
  T *const W[2] = {w, (T *)g->d_w2};
  // one iteration (it: the pass iteration, where every running segment is):
  // the apply launch, then the finishing step(s) and the r-update
  auto apply_launch = [&](int it) {
    T *pold = (it & 1) ? p2 : p, *pnew = (it & 1) ? p : p2;
    const unsigned grid = xgrid(nact[0]);
    if (fold)
      k_lz_apply<T, true><<<grid, RSB_BLOCK, 0, st>>>(ai, pold, pnew, W[it & 1], basis, mpad, ta,
                                                      nact[0], W[(it + 1) & 1]);
    else if (vform)
      k_lz_apply_v<float><<<grid, RSB_BLOCK, 0, st>>>(ai, (const float *)pold, (float *)pnew,
                                                      (float *)w, (const float *)basis, mpad, va,
                                                      ta, nact[0]);
    else
      k_lz_apply<T, false><<<grid, RSB_BLOCK, 0, st>>>(ai, pold, pnew, w, basis, mpad, ta,
                                                       nact[0], nullptr);
  };
  // the iteration sums: wide workgroups where the segments are big
  int max_seg_tiles = 0;
  for (int q = 0; q < nseg; q++) max_seg_tiles = std::max(max_seg_tiles, g->h_seg_tiles[2 * q + 1]);
  const bool wide_fin = max_seg_tiles > LZ_FIN_WIDE;
  auto fin_launch = [&](int op) {
    if (wide_fin) k_lz_fin<2 * RSB_BLOCK><<<nact[1], 2 * RSB_BLOCK, 0, st>>>(ai, op);
    else k_lz_fin<RSB_BLOCK><<<nact[1], RSB_BLOCK, 0, st>>>(ai, op);
  };
  auto after_apply = [&]() {
    if (fold) {
      fin_launch(FIN_APPLYR);
    } else {
      fin_launch(FIN_APPLY);
      k_lz_rupdate<T><<<nact[0], RSB_BLOCK, 0, st>>>(ai, basis, mpad, w);
      fin_launch(FIN_RUPDATE);
    }
  };

  RSB_TRY(hipMemsetAsync(L.running, 0, sizeof(int) * 2, st));
  k_lz_init_state<<<nblocks(nseg, 256), 256, 0, st>>>(a, miter_opt);
  RSB_LAUNCH_CHECK();
  k_lz_zero_small<<<nblocks(nseg, 256), 256, 0, st>>>(a, g->d_fied);
  RSB_LAUNCH_CHECK();
  if (ntiles == 0) {
    return 0;
  }
  // start vector of the segments in LZ_SETUP, normalised into f
  auto start = [&](int mode, const float *sv) -> int {
    a.mode = ai.mode = mode;
    for (int ph = 0; ph < 3; ph++) {
      k_lz_start<T><<<ntiles, RSB_BLOCK, 0, st>>>(a, ph, mode, sv, f);
      RSB_LAUNCH_CHECK();
      if (ph < 2) FIN(ph == 0 ? FIN_START0 : FIN_START1);
    }
    return 0;
  };
  // restarted CG-Lanczos passes of the segments in LZ_SETUP (f holds the
  // normalised start vector; SC_MEANF is 0 for the first pass)
  const size_t tq_lds = sizeof(double) * (2 * miter_max + miter_max * miter_max);
  auto passes = [&](int np_max) -> int {
    a.mpass = ai.mpass = np_max;
    // the running lists of a pass: compacted at the end of the pass before
    // (the segments that restart are LZ_SETUP there and RUNNING once set up,
    // which saves a device wait per pass), at its start for the first
    bool lists_ok = false;
    for (int pass = 0; pass < np_max; pass++) {
      // r_0 into basis slot 0
      k_lz_setup<T><<<ntiles, RSB_BLOCK, 0, st>>>(a, 0, f, basis, p);
      RSB_LAUNCH_CHECK();
      FIN(FIN_SETUP0);
      k_lz_setup<T><<<ntiles, RSB_BLOCK, 0, st>>>(a, 1, f, basis, p);
      RSB_LAUNCH_CHECK();
      FIN(FIN_SETUP1);
      if (!lists_ok && compact(LZ_RUNNING)) return 2;
      fold = nact[0] <= fold_tiles;
      RSB_TRY(hipEventRecord(g->lz_ev0, st));
      int it, nrec = 0;
      const char *gv = getenv("PARRSB_AMD_LZ_GRAPH");
      const bool use_graph = !g->prof_spmv && gv && atoi(gv) == 1;
      hipGraphExec_t gexec = nullptr;
      int gn[2] = {-1, -1};
      if (g->prof_spmv && (int)g->spmv_ev.size() < 2 * miter_max) {
        size_t old = g->spmv_ev.size();
        g->spmv_ev.resize(2 * miter_max);
        for (size_t q = old; q < g->spmv_ev.size(); q++) RSB_TRY(hipEventCreate(&g->spmv_ev[q]));
      }
      // the running lists are compacted 8 iterations into the pass and then
      // at intervals that double (up to 32) while no segment has finished:
      // each compaction waits on the device, and a finished segment's tiles
      // only cost launches whose workgroups return at once
      int next_c = 7, step = 8;
      for (it = 0; it < miter_max; it++) {
        if (nact[0] == 0) break;
        if (use_graph && (it & 7) == 0 && it + 8 < miter_max) {
          // eight iterations as one graph replay (same kernels and grids as
          // the launches below; the grids follow the compacted lists, so
          // the chunk is captured again when they changed)
          if (!gexec || gn[0] != nact[0] || gn[1] != nact[1]) {
            hipGraph_t gr = nullptr;
            RSB_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            // (it is a multiple of 8: the chunk's buffer parities are q's)
            for (int q = 0; q < 8; q++) {
              apply_launch(q);
              after_apply();
            }
            RSB_TRY(hipStreamEndCapture(st, &gr));
            if (gexec) RSB_TRY(hipGraphExecDestroy(gexec));
            gexec = nullptr;
            RSB_TRY(hipGraphInstantiateWithFlags(&gexec, gr, 0));
            RSB_TRY(hipGraphDestroy(gr));
            gn[0] = nact[0], gn[1] = nact[1];
          }
          RSB_TRY(hipGraphLaunch(gexec, st));
          g->lz_launches += 8;
          it += 7;
          if (it + 1 < miter_max) {
            if (compact(LZ_RUNNING)) return 2;
            if (nact[1] == 0) break;
          }
          continue;
        }
        // profiling: every LZ_PROF_EVERY-th launch is bracketed by events (each
        // pair of records costs several microseconds of stream time, so
        // timing every launch slowed the whole call by ~5%)
        const bool samp = g->prof_spmv && (g->spmv_total++ % LZ_PROF_EVERY) == 0;
        if (samp) RSB_TRY(hipEventRecord(g->spmv_ev[2 * nrec], st));
        apply_launch(it);
        if (samp) {
          RSB_TRY(hipEventRecord(g->spmv_ev[2 * nrec + 1], st));
          nrec++;
        }
        after_apply();
        RSB_LAUNCH_CHECK();
        g->lz_launches++;
        if (it == next_c && it + 1 < miter_max) {
          // drop segments that finished their pass from the launch lists
          const int before = nact[1];
          if (compact(LZ_RUNNING)) return 2;
          if (nact[1] == 0) break;
          step = nact[1] == before ? std::min(2 * step, 32) : 8;
          next_c = it + step;
        }
      }
      RSB_TRY(hipEventRecord(g->lz_ev1, st));
      RSB_TRY(hipMemsetAsync(L.running + 1, 0, sizeof(int), st));
      if (long_pass) {
        // segments in chunks whose matrices take at most 1 GiB together
        const size_t zb = sizeof(double) * (size_t)miter_max * miter_max;
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nseg, (1ull << 30) / zb));
        double *zbuf = nullptr;
        RSB_TRY(RSB_SCRATCH(zbuf, zb * (size_t)chunk));
        for (int s0 = 0; s0 < nseg; s0 += chunk) {
          k_lz_tqli_g<<<std::min(chunk, nseg - s0), RSB_WAVE, sizeof(double) * 2 * miter_max,
                        st>>>(a, s0, zbuf);
          RSB_LAUNCH_CHECK();
        }
      } else if (sizeof(T) == 4) {
        k_lz_eigmin<<<nseg, RSB_WAVE, sizeof(double) * 4 * miter_max, st>>>(a);
      } else {
        k_lz_tqli<<<nseg, RSB_WAVE, tq_lds, st>>>(a);
      }
      RSB_LAUNCH_CHECK();
      k_lz_ritz<T><<<ntiles, RSB_BLOCK, 0, st>>>(a, basis, mpad, f);
      RSB_LAUNCH_CHECK();
      FIN(FIN_RITZ);
      if (compact(LZ_SETUP)) return 2;  // waits on the device
      lists_ok = true;
      int again = 0;
      RSB_TRY(hipMemcpyAsync(&again, L.running + 1, sizeof(int), hipMemcpyDeviceToHost, st));
      RSB_TRY(hipStreamSynchronize(st));
      // the pass's events are complete now (one wait per pass end)
      if (gexec) RSB_TRY(hipGraphExecDestroy(gexec));
      float ms = 0.f;
      RSB_TRY(hipEventElapsedTime(&ms, g->lz_ev0, g->lz_ev1));
      g->lz_ms += ms;
      for (int q = 0; q < nrec; q++) {
        RSB_TRY(hipEventElapsedTime(&ms, g->spmv_ev[2 * q], g->spmv_ev[2 * q + 1]));
        g->spmv_ms += ms;
      }
      g->spmv_launches += nrec;
      if (again == 0) break;
    }
    return 0;
  };
  int rc = 0;
  // learned start: nothing to do if no segment is within the size limit
  bool learned_any = false;
  if (init_mode == 2)
    for (int s = 0; s < nseg; s++)
      if (g->h_p[s] > 1 && g->h_n[s] >= 2 && (g->lws_max_n == 0 || g->h_n[s] <= g->lws_max_n))
        learned_any = true;
  if (init_mode == 2 && learned_any) {
    // learned start (per row, written into w, which the passes overwrite
    // only after the start has been normalised into f), then the fallback
    float *srow = (float *)w;
    if ((rc = rsb_lws_predict(g, g->lws_params, g->lws_level, srow))) return rc;
    if ((rc = start(2, srow)) || (rc = passes(std::min(g->lws_passes, mpass)))) return rc;
    RSB_TRY(hipMemsetAsync(L.running + 1, 0, sizeof(int), st));
    k_lz_fallback<<<nblocks(nseg, 256), 256, 0, st>>>(a, L.running + 1);
    RSB_LAUNCH_CHECK();
    int nfb = 0;
    RSB_TRY(hipMemcpyAsync(&nfb, L.running + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipStreamSynchronize(st));
    if (nfb > 0 && ((rc = start(0, nullptr)) || (rc = passes(mpass)))) return rc;
  } else {
    if ((rc = start(init_mode == 2 ? 0 : init_mode, g->d_start)) || (rc = passes(mpass)))
      return rc;
  }
  k_lz_finish<T><<<ntiles, RSB_BLOCK, 0, st>>>(a, 0, f, g->d_fied, xrow);
  RSB_LAUNCH_CHECK();
  FIN(FIN_FINISH);
  k_lz_finish<T><<<ntiles, RSB_BLOCK, 0, st>>>(a, 1, f, g->d_fied, xrow);
  RSB_LAUNCH_CHECK();
  if (out && out->resid) {
    k_lz_resid<<<ntiles, RSB_BLOCK, 0, st>>>(a, 0, xrow, g->d_slice_off, g->d_slice_w, g->d_ell);
    RSB_LAUNCH_CHECK();
    FIN(FIN_RES0);
    k_lz_resid<<<ntiles, RSB_BLOCK, 0, st>>>(a, 1, xrow, g->d_slice_off, g->d_slice_w, g->d_ell);
    RSB_LAUNCH_CHECK();
    FIN(FIN_RES1);
  }
  // read back per-segment results
  std::vector<int> ni(nseg), np(nseg);
  std::vector<double> sc((size_t)nseg * LZ_NSC);
  RSB_TRY(hipMemcpyAsync(ni.data(), a.niter, sizeof(int) * nseg, hipMemcpyDeviceToHost, st));
  RSB_TRY(hipMemcpyAsync(np.data(), a.npass, sizeof(int) * nseg, hipMemcpyDeviceToHost, st));
  RSB_TRY(hipMemcpyAsync(sc.data(), a.sc, sizeof(double) * nseg * LZ_NSC,
                         hipMemcpyDeviceToHost, st));
  RSB_TRY(hipStreamSynchronize(st));
  // Algorithmic bytes, per iteration and segment of n rows and nnz
  // off-diagonals: 4(n+1) row pointers + 4 nnz packed entries + 7 s n vector
  // streams for the whole iteration; for k_lz_apply alone 4(n+1) + 4 nnz +
  // 4 s n (p_old and r read once, p_new and w written; r stays in its basis
  // slot).  With the folded update the apply is the iteration: 6 s n (r, w
  // and p_old of the previous iteration read; r, p_new and w written).
  // Counted iterations, real (unpadded) sizes.
  for (int s = 0; s < nseg; s++) {
    const double *scs = &sc[(size_t)s * LZ_NSC];
    ni[s] += (int)scs[SC_WARMIT];
    if (g->h_p[s] > 1 && g->h_n[s] >= 2) {
      double n = g->h_n[s], nz = (double)g->h_seg_nnz[s], sz = sizeof(T);
      const double nf = scs[SC_NFOLD], nu = ni[s] - nf;
      g->lz_bytes += nu * (4 * (n + 1) + 4 * nz + 7 * sz * n) + nf * (4 * (n + 1) + 4 * nz + 6 * sz * n);
      g->spmv_bytes += nu * (4 * (n + 1) + 4 * nz + 4 * sz * n) + nf * (4 * (n + 1) + 4 * nz + 6 * sz * n);
    }
    if (out) {
      if (out->niter) out->niter[s] = ni[s];
      if (out->npass) out->npass[s] = np[s];
      if (out->lambda) out->lambda[s] = sc[(size_t)s * LZ_NSC + SC_LAMBDA];
      if (out->rnorm) {
        out->rnorm[3 * s] = sc[(size_t)s * LZ_NSC + SC_RNORM0];
        out->rnorm[3 * s + 1] = sc[(size_t)s * LZ_NSC + SC_RTOL];
        out->rnorm[3 * s + 2] = sc[(size_t)s * LZ_NSC + SC_RNORM];
      }
      if (out->resid) out->resid[s] = sc[(size_t)s * LZ_NSC + SC_TMP];
      if (out->fallback)
        out->fallback[s] = (init_mode == 2 && !learned_any && g->h_p[s] > 1 && g->h_n[s] >= 2) ||
                                   scs[SC_FALLBACK] < 0.0
                               ? -1
                           : scs[SC_FALLBACK] > 0.0 ? 1 + (int)scs[SC_WARMIT]
                                                    : 0;
    }
  }
  return 0;
}

extern "C" int rsbgpu_set_start(rsbgpu *g, const float *d_start) {
  RSB_TRY(hipSetDevice(g->device));
  RSB_TRY(dev_alloc(g, g->d_start, (size_t)g->nel));
  RSB_TRY(hipMemcpyAsync(g->d_start, d_start, sizeof(float) * (size_t)g->nel,
                         hipMemcpyDeviceToDevice, g->stream));
  return 0;
}

extern "C" int rsbgpu_fiedler(rsbgpu *g, int precision, int max_iter, int max_passes,
                              double tol, int init_mode, rsbgpu_fiedler_out *out) {
  RSB_TRY(hipSetDevice(g->device));
  if (max_iter < 1 || max_passes < 1) RSB_FAIL(3, "rsb_max_iter and rsb_max_passes must be >= 1");
  if (init_mode == 1 && !g->d_start) RSB_FAIL(3, "no start vector set");
  if (init_mode == 2 && !g->lws_params) RSB_FAIL(3, "no warm-start parameters set");
  if (init_mode < 0 || init_mode > 2) RSB_FAIL(3, "unknown init_mode %d", init_mode);
  rsb_phase_begin(g);
  int rc;
  if (precision == 32) rc = run_lanczos<float>(g, max_iter, max_passes, tol, init_mode, out);
  else rc = run_lanczos<double>(g, max_iter, max_passes, tol, init_mode, out);
  rsb_phase_end(g, PH_LANCZOS);
  return rc;
}

static int rsb_test_tridiag(const double *d, const double *e, int n, double *y, double *lam,
                            int fast);

extern "C" int rsbgpu_test_tqli(const double *d, const double *e, int n, double *y,
                                double *lam) {
  return rsb_test_tridiag(d, e, n, y, lam, 0);
}

extern "C" int rsbgpu_test_eigmin(const double *d, const double *e, int n, double *y,
                                  double *lam) {
  return rsb_test_tridiag(d, e, n, y, lam, 1);
}

static int rsb_test_tridiag(const double *d, const double *e, int n, double *y, double *lam,
                            int fast) {
  double *dd = nullptr, *de = nullptr, *dy = nullptr, *dl = nullptr, *dz = nullptr;
  if (n < 1 || n > 4096 || (fast && n > 128)) return 3;
  hipError_t err = rsb_dev_malloc((void **)&dd, sizeof(double) * n);
  if (err == hipSuccess) err = rsb_dev_malloc((void **)&de, sizeof(double) * n);
  if (err == hipSuccess) err = rsb_dev_malloc((void **)&dy, sizeof(double) * n);
  if (err == hipSuccess) err = rsb_dev_malloc((void **)&dl, sizeof(double));
  if (err == hipSuccess) err = hipMemcpy(dd, d, sizeof(double) * n, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemset(de, 0, sizeof(double) * n);
  if (err == hipSuccess && n > 1)
    err = hipMemcpy(de, e, sizeof(double) * (n - 1), hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    if (fast) {
      k_eigmin_test<<<1, RSB_WAVE, sizeof(double) * 4 * n, 0>>>(n, dd, de, dy, dl);
    } else if (n > 128) {
      err = rsb_dev_malloc((void **)&dz, sizeof(double) * (size_t)n * n);
      if (err == hipSuccess)
        k_tqli_test_g<<<1, RSB_WAVE, sizeof(double) * 2 * n, 0>>>(n, dd, de, dy, dl, dz);
    } else {
      size_t lds = sizeof(double) * (2 * n + n * n);
      k_tqli_test<<<1, RSB_WAVE, lds, 0>>>(n, dd, de, dy, dl);
    }
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipDeviceSynchronize();
  if (err == hipSuccess) err = hipMemcpy(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(lam, dl, sizeof(double), hipMemcpyDeviceToHost);
  if (dd) (void)rsb_dev_free(dd);
  if (de) (void)rsb_dev_free(de);
  if (dy) (void)rsb_dev_free(dy);
  if (dl) (void)rsb_dev_free(dl);
  if (dz) (void)rsb_dev_free(dz);
  return err == hipSuccess ? 0 : 2;
}
