// Fiedler vectors by MG-preconditioned inverse iteration (rsb_algo = 1;
// fiedler.c:94-321, multigrid.c:184-317, the matrices of mat.c:17-61 and
// 704-764), for every segment of a level at once, with the optional
// Gram-matrix refinement of the start vectors (rsb_mg_grammian = 1,
// fiedler.c:250-305; see k_mg_gjac).
//
// Level 0 of the hierarchy is the level Laplacian already assembled for the
// Lanczos path (padded rows, SELL-64 entries with 4-bit weights, diagonal =
// row sum of the weights).  Level l + 1 aggregates `factor` consecutive rows
// of level l inside each segment (row i -> i / factor); its matrix is built
// on the device: every entry of the children rows is keyed (coarse row,
// coarse column), one radix sort groups equal keys, entries inside an
// aggregate are dropped and equal keys add up (integer-valued, so the sums
// are exact in any order).  Coarse rows are compact CSR in FP64.
//
// The solve is the reference's, segment by segment in lockstep: every
// reduction is a per-tile partial sum (fixed tree) finished by a
// per-segment kernel (tiles in order), so results do not depend on
// scheduling.  Segments that finished skip every launch.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "state.hpp"

#define MG_NP 64  // partial-sum slots per tile (rsb_max_iter <= 60 dots + spare)

// per-segment FP64 scalars
enum { MS_RZ1 = 0, MS_ALPHA, MS_PWINV, MS_BETA, MS_RTOL, MS_RR, MS_MEAN, MS_NI, MS_LAMBDA,
       MS_RES, MS_GRNI, MS_NSC };
// per-segment ints (MI_NB: vectors in the Gram basis, MI_GADD: the pass's
// vector joins the basis)
enum { MI_INV = 0, MI_PCG, MI_IT, MI_PPFI, MI_ITERS, MI_NPASS, MI_NC, MI_NB, MI_GADD, MI_NI };
// what a finishing kernel does with a segment's totals
enum { FIN_MEAN = 0, FIN_NORM, FIN_PASS0, FIN_RZ1, FIN_ALPHA, FIN_CONV, FIN_BETA,
       FIN_PASSEND, FIN_RAYLEIGH };
// which segments a launch works on (GATE_GRAM: continuing segments whose
// basis grows this pass; GATE_GMIX: continuing segments with >= 2 basis
// vectors)
enum { GATE_ELIG = 0, GATE_INV, GATE_PCG, GATE_GRAM, GATE_GMIX };

struct mg_args {
  int nseg;
  const int *seg;        // 5*nseg: cstart, n, p, r0, pstart
  const int *tile;       // 3*ntiles: seg, row0, row1
  const int *seg_tiles;  // 2*nseg: first tile, count
  double *sc;            // MS_NSC per segment
  int *st;               // MI_NI per segment
  double *partial;       // MG_NP per tile
  int *active;           // count written by finishing kernels
  double tol;
  int miter, mpass;
  int ntiles;            // partial is slot-major: slot k of tile t at k * ntiles + t
};

#define MSC(s, k) a.sc[(size_t)(s) * MS_NSC + (k)]
#define MST(s, k) a.st[(size_t)(s) * MI_NI + (k)]

__device__ inline bool mg_elig(const mg_args &a, int s) {
  return a.seg[2 * a.nseg + s] > 1 && a.seg[a.nseg + s] >= 2;
}
__device__ inline bool mg_gate(const mg_args &a, int s, int gate) {
  if (gate == GATE_ELIG) return mg_elig(a, s);
  if (gate == GATE_INV) return MST(s, MI_INV) != 0;
  if (gate == GATE_PCG) return MST(s, MI_PCG) != 0;
  if (gate == GATE_GRAM) return MST(s, MI_INV) != 0 && MST(s, MI_GADD) != 0;
  return MST(s, MI_INV) != 0 && MST(s, MI_NB) >= 2;
}

// Tile prologue: rows [row0, row1) of segment s, valid while local < n.
struct mg_tile {
  int ti, s, row0, row1, pstart, n;
};
__device__ inline bool mg_tile_of(const mg_args &a, int gate, mg_tile &t) {
  t.ti = blockIdx.x;
  t.s = a.tile[3 * t.ti];
  if (!mg_gate(a, t.s, gate)) return false;
  t.row0 = a.tile[3 * t.ti + 1];
  t.row1 = a.tile[3 * t.ti + 2];
  t.n = a.seg[a.nseg + t.s];
  t.pstart = a.seg[4 * a.nseg + t.s];
  return true;
}
#define MG_ROWS(t, row) \
  for (int row = (t).row0 + (int)threadIdx.x; row < (t).row1 && row - (t).pstart < (t).n; row += RSB_BLOCK)

template <int NV>
__device__ inline void mg_publish(const mg_args &a, const mg_tile &t, double (&v)[NV]) {
  __shared__ double red[(RSB_BLOCK / RSB_WAVE) * 4];
  block_sum<NV>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; k++) a.partial[(size_t)k * a.ntiles + t.ti] = v[k];
}

// Level-0 Laplacian row: sum_j w_ij x_j over the SELL-64 entries; the row
// sum of the weights is the diagonal.
struct mg_sell {
  const int64_t *slice_off;
  const uint32_t *ell;
  const int *rowlen;
  const double *diag;
  // the level's LDS tiles (16-bit entries: local index | weight << 12 over
  // the tile's rows and then its halo rows), or nhalo == nullptr
  const int *nhalo;       // per tile, -1 if its halo does not fit
  const int64_t *halo_off;
  const int *halo;
  const int64_t *off16;
  const int *w16;
  const uint16_t *e16;
};

typedef unsigned mg_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned mg_u32x2 __attribute__((ext_vector_type(2)));
#define MG_TB_LOCAL 4096

// Stage x over the tile's rows and halo rows into LDS; returns false (and
// stages nothing) if the tile has no LDS form.
__device__ inline bool mg_stage(const mg_sell &L, const mg_tile &t, const double *x, double *xl) {
  if (!L.nhalo) return false;
  const int nh = L.nhalo[t.ti];
  if (nh < 0) return false;
  const int nrows = t.row1 - t.row0, tot = nrows + nh;
  const int64_t ho = L.halo_off[t.ti];
  for (int i = threadIdx.x; i < tot; i += blockDim.x)
    xl[i] = x[i < nrows ? t.row0 + i : L.halo[ho + i - nrows]];
  __syncthreads();
  return true;
}

// sum_j w_ij x_j over the row's 16-bit entries (the same entries in the same
// order as the 32-bit form, then weight-0 padding, which adds exact zeros)
__device__ inline double tile_offsum(const mg_sell &L, int row, const double *xl) {
  const int sl = row >> 6;
  const uint16_t *sbase = L.e16 + L.off16[sl];
  const mg_u32x4 *ep = (const mg_u32x4 *)(sbase + (row & 63) * 8);
  const int w16 = L.w16[sl], nb = w16 >> 3;
  double s = 0.0;
  for (int jb = 0; jb < nb; jb++) {
    const mg_u32x4 v = ep[(size_t)jb * RSB_WAVE];
    const unsigned e[8] = {v.x & 0xffffu, v.x >> 16, v.y & 0xffffu, v.y >> 16,
                           v.z & 0xffffu, v.z >> 16, v.w & 0xffffu, v.w >> 16};
#pragma unroll
    for (int k = 0; k < 8; k++) s += (double)(e[k] >> 12) * xl[e[k] & 0xfffu];
  }
  if (w16 & 4) {
    const mg_u32x2 v = *(const mg_u32x2 *)(sbase + nb * 512 + (row & 63) * 4);
    const unsigned e[4] = {v.x & 0xffffu, v.x >> 16, v.y & 0xffffu, v.y >> 16};
#pragma unroll
    for (int k = 0; k < 4; k++) s += (double)(e[k] >> 12) * xl[e[k] & 0xfffu];
  }
  if (w16 & 2) {
    const unsigned v = *(const unsigned *)(sbase + nb * 512 + (w16 & 4) * 64 + (row & 63) * 2);
    const unsigned e[2] = {v & 0xffffu, v >> 16};
#pragma unroll
    for (int k = 0; k < 2; k++) s += (double)(e[k] >> 12) * xl[e[k] & 0xfffu];
  }
  return s;
}
// Eight entries and the eight values they point at are loaded before any is
// used (one entry, then its gather, at a time left the smoother
// latency-bound); the sum still runs over j in order.
__device__ inline double sell_offsum(const mg_sell &L, int row, const double *x) {
  const int64_t base = L.slice_off[row / RSB_WAVE] + (row % RSB_WAVE);
  const int len = L.rowlen[row];
  double s = 0.0;
  for (int j0 = 0; j0 < len; j0 += 8) {
    uint32_t e[8];
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
      e[u] = j0 + u < len ? L.ell[base + (int64_t)(j0 + u) * RSB_WAVE] : (uint32_t)row;
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = x[e[u] & 0x0fffffffu];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (j0 + u < len) s += (double)(e[u] >> 28) * v[u];
  }
  return s;
}

// ---------------------------------------------------------------------------
// Totals of partial slots k0 .. k0 + ns - 1 (ns <= 4) of a segment, formed
// by the whole workgroup: thread t sums tiles t, t + blockDim, ... in that
// order (four tiles of every slot in flight; a slot's partials are
// contiguous), then wave shuffles and the waves in order.  Every thread
// returns the same totals.  (One wave walking the tile-major records of
// 16k tiles took ~0.1 ms per slot on a 16.7M-row segment.)
__device__ inline void seg_totals(const mg_args &a, int first, int count, int k0, int ns,
                                  double (&tot)[4], double *red) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int t0 = threadIdx.x; t0 < count; t0 += 4 * (int)blockDim.x) {
    double x[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int t = t0 + u * (int)blockDim.x;
#pragma unroll
      for (int k = 0; k < 4; k++)
        x[u][k] = (k < ns && t < count) ? a.partial[(size_t)(k0 + k) * a.ntiles + first + t] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int k = 0; k < 4; k++) v[k] += x[u][k];
  }
  const int lane = threadIdx.x & (RSB_WAVE - 1), wid = threadIdx.x / RSB_WAVE;
  const int nw = blockDim.x / RSB_WAVE;
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = wave_sum(v[k]);
  __syncthreads();  // red may still be read from the previous call
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) red[wid * 4 + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double t = red[k];
    for (int w = 1; w < nw; w++) t += red[w * 4 + k];
    tot[k] = t;
  }
}

// Finishing kernel: one workgroup per segment forms the totals of the tile
// partials and applies the scalar step.
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_fin(mg_args a, int op, int gate, int nslots) {
  __shared__ double red[(RSB_BLOCK / RSB_WAVE) * 4];
  const int s = blockIdx.x;
  if (!mg_gate(a, s, gate)) return;
  double tot[4];
  const int first = a.seg_tiles[2 * s], count = a.seg_tiles[2 * s + 1];
  seg_totals(a, first, count, 0, nslots, tot, red);
  if (threadIdx.x != 0) return;
  const double n = (double)a.seg[a.nseg + s];
  switch (op) {
  case FIN_MEAN: MSC(s, MS_MEAN) = tot[0] / n; break;
  case FIN_NORM: MSC(s, MS_NI) = 1.0 / sqrt(tot[0]); break;
  case FIN_PASS0: {  // tot: sum r, r.r
    MSC(s, MS_RR) = tot[1];
    MSC(s, MS_RTOL) = tot[1] * a.tol * a.tol;
    MSC(s, MS_MEAN) = tot[0] / n;
    MST(s, MI_PCG) = 1, MST(s, MI_IT) = 0;
    break;
  }
  case FIN_RZ1: MSC(s, MS_RZ1) = tot[0]; break;
  case FIN_ALPHA: {  // tot: p.w
    const double pw = tot[0];
    MSC(s, MS_ALPHA) = MSC(s, MS_RZ1) / pw;
    MSC(s, MS_PWINV) = 1.0 / sqrt(pw);
    break;
  }
  case FIN_CONV: {  // tot: r.r after the update of iteration i
    const double rr = tot[0];
    const int i = MST(s, MI_IT);
    MSC(s, MS_RR) = rr;
    if (rr < MSC(s, MS_RTOL) || sqrt(rr) < a.tol) {
      MST(s, MI_PCG) = 0, MST(s, MI_PPFI) = i + 1;
    } else if (i + 1 == a.miter) {
      MST(s, MI_PCG) = 0, MST(s, MI_PPFI) = a.miter;
    } else {
      atomicAdd(a.active, 1);
    }
    break;
  }
  case FIN_BETA: {  // tot: r.z, r.(z - z0)
    const double rzt = MSC(s, MS_RZ1);
    MSC(s, MS_RZ1) = tot[0];
    MSC(s, MS_BETA) = tot[1] / rzt;
    break;
  }
  case FIN_PASSEND: {  // tot: sum of the new start vector (after ortho of y)
    MSC(s, MS_MEAN) = tot[0] / n;
    const int ppfi = MST(s, MI_PPFI);
    MST(s, MI_ITERS) += ppfi;
    MST(s, MI_NPASS) += 1;
    if (ppfi == 1 || MST(s, MI_NPASS) == a.mpass) MST(s, MI_INV) = 0;
    else atomicAdd(a.active, 1);
    break;
  }
  case FIN_RAYLEIGH: {  // tot: f.Lf, Lf.Lf, f.f
    const double lam = tot[0] / tot[2];
    MSC(s, MS_LAMBDA) = lam;
    const double r2 = tot[1] - 2.0 * lam * tot[0] + lam * lam * tot[2];
    MSC(s, MS_RES) = sqrt(r2 > 0.0 ? r2 : 0.0) / sqrt(tot[2]) / (lam != 0.0 ? fabs(lam) : 1.0);
    break;
  }
  }
}

// Totals of the re-orthogonalisation dots W[j].p (j <= i) into the
// per-segment coefficient array read by k_mg_pcorr; the iteration ends here.
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_coef(mg_args a, double *coef) {
  __shared__ double red[(RSB_BLOCK / RSB_WAVE) * 4];
  const int s = blockIdx.x;
  if (!mg_gate(a, s, GATE_PCG)) return;
  const int first = a.seg_tiles[2 * s], count = a.seg_tiles[2 * s + 1];
  const int nc = MST(s, MI_IT) + 1;
  for (int k0 = 0; k0 < nc; k0 += 4) {
    double tot[4];
    const int ns = min(4, nc - k0);
    seg_totals(a, first, count, k0, ns, tot, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < ns; k++) coef[(size_t)s * MG_NP + k0 + k] = tot[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) MST(s, MI_NC) = nc, MST(s, MI_IT) = nc;
}

// ---------------------------------------------------------------------------
// Vector kernels over the tiles of level 0.

// start vector (fiedler.c:572-578) into b; partial sum
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_start(mg_args a, double *b) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_ELIG, t)) return;
  double v[1] = {0.0};
  MG_ROWS(t, row) {
    const int i = row - t.pstart;
    double x = i + 1.0;
    if (i < t.n / 2) x += (double)(1000 * (long long)t.n);
    b[row] = x;
    v[0] += x;
  }
  mg_publish<1>(a, t, v);
}

// partial sum of v
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_sum(mg_args a, int gate, const double *v) {
  mg_tile t;
  if (!mg_tile_of(a, gate, t)) return;
  double acc[1] = {0.0};
  MG_ROWS(t, row) acc[0] += v[row];
  mg_publish<1>(a, t, acc);
}

// v -= mean; partial v.v
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_center_dot(mg_args a, int gate, double *v) {
  mg_tile t;
  if (!mg_tile_of(a, gate, t)) return;
  const double m = MSC(t.s, MS_MEAN);
  double acc[1] = {0.0};
  MG_ROWS(t, row) {
    const double x = v[row] - m;
    v[row] = x;
    acc[0] += x * x;
  }
  mg_publish<1>(a, t, acc);
}

// dst = v * ni; partial sum of dst
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_scale_sum(mg_args a, int gate, const double *v,
                                                           double *dst) {
  mg_tile t;
  if (!mg_tile_of(a, gate, t)) return;
  const double ni = MSC(t.s, MS_NI);
  double acc[1] = {0.0};
  MG_ROWS(t, row) {
    const double x = v[row] * ni;
    dst[row] = x;
    acc[0] += x;
  }
  mg_publish<1>(a, t, acc);
}

// v -= mean (no partials)
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_center(mg_args a, int gate, double *v) {
  mg_tile t;
  if (!mg_tile_of(a, gate, t)) return;
  const double m = MSC(t.s, MS_MEAN);
  MG_ROWS(t, row) v[row] -= m;
}

// projection start (fiedler.c:110-126): x = 0, r = b; partials sum r, r.r
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_pass0(mg_args a, const double *b, double *x,
                                                       double *r) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_INV, t)) return;
  double acc[2] = {0.0, 0.0};
  MG_ROWS(t, row) {
    const double v = b[row];
    x[row] = 0.0, r[row] = v;
    acc[0] += v, acc[1] += v * v;
  }
  mg_publish<2>(a, t, acc);
}

// z = r - mean(r), p = z; partial z.z
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_pass1(mg_args a, const double *r, double *z,
                                                       double *p) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  const double m = MSC(t.s, MS_MEAN);
  double acc[1] = {0.0};
  MG_ROWS(t, row) {
    const double v = r[row] - m;
    z[row] = v, p[row] = v;
    acc[0] += v * v;
  }
  mg_publish<1>(a, t, acc);
}

// w = L p; partial p.w
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_apply(mg_args a, int gate, mg_sell L,
                                                       const double *p, double *w) {
  __shared__ double xl[MG_TB_LOCAL];
  mg_tile t;
  if (!mg_tile_of(a, gate, t)) return;
  double acc[1] = {0.0};
  if (mg_stage(L, t, p, xl)) {
    MG_ROWS(t, row) {
      const double pi = xl[row - t.row0];
      const double v = L.diag[row] * pi - tile_offsum(L, row, xl);
      w[row] = v;
      acc[0] += pi * v;
    }
  } else {
    MG_ROWS(t, row) {
      const double v = L.diag[row] * p[row] - sell_offsum(L, row, p);
      w[row] = v;
      acc[0] += p[row] * v;
    }
  }
  mg_publish<1>(a, t, acc);
}

// W[i] = pwinv w, P[i] = pwinv p, x += alpha p, r -= alpha w; partial r.r
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_update(mg_args a, int64_t mpad,
                                                        const double *p, const double *w,
                                                        double *x, double *r, double *Wb,
                                                        double *Pb) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  const int i = MST(t.s, MI_IT);
  const double alpha = MSC(t.s, MS_ALPHA), pwinv = MSC(t.s, MS_PWINV);
  double acc[1] = {0.0};
  MG_ROWS(t, row) {
    const double pv = p[row], wv = w[row];
    Wb[(size_t)i * mpad + row] = pwinv * wv;
    Pb[(size_t)i * mpad + row] = pwinv * pv;
    x[row] += alpha * pv;
    const double rv = r[row] - alpha * wv;
    r[row] = rv;
    acc[0] += rv * rv;
  }
  mg_publish<1>(a, t, acc);
}

// z -= mean(z); partials r.z, r.(z - z0)
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_zdots(mg_args a, const double *r, double *z,
                                                       const double *z0) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  const double m = MSC(t.s, MS_MEAN);
  double acc[2] = {0.0, 0.0};
  MG_ROWS(t, row) {
    const double zv = z[row] - m;
    z[row] = zv;
    acc[0] += r[row] * zv;
    acc[1] += r[row] * (zv - z0[row]);
  }
  mg_publish<2>(a, t, acc);
}

// p = z + beta p; partials W[j].p for j <= i
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_pnew(mg_args a, int64_t mpad, const double *z,
                                                      double *p, const double *Wb) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  const double beta = MSC(t.s, MS_BETA);
  const int i = MST(t.s, MI_IT);
  double pv[RSB_TILE / RSB_BLOCK];
  int k = 0;
  MG_ROWS(t, row) {
    const double v = z[row] + beta * p[row];
    p[row] = v;
    pv[k++] = v;
  }
  __shared__ double red[(RSB_BLOCK / RSB_WAVE) * 4];
  for (int j = 0; j <= i; j++) {
    double acc[1] = {0.0};
    int q = 0;
    MG_ROWS(t, row) acc[0] += Wb[(size_t)j * mpad + row] * pv[q++];
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) a.partial[(size_t)j * a.ntiles + t.ti] = acc[0];
    __syncthreads();
  }
}

// p -= sum_j c_j P[j]
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_pcorr(mg_args a, int64_t mpad, double *p,
                                                       const double *Pb, const double *coef) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  const int nc = MST(t.s, MI_NC);
  const double *c = coef + (size_t)t.s * MG_NP;
  MG_ROWS(t, row) {
    double acc = 0.0;
    for (int j = 0; j < nc; j++) acc += c[j] * Pb[(size_t)j * mpad + row];
    p[row] -= acc;
  }
}

__global__ void k_mg_zero(int *c) { *c = 0; }

// z0 = z (the previous preconditioned residual)
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_copy(mg_args a, const double *z, double *z0) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  MG_ROWS(t, row) z0[row] = z[row];
}

// fiedler[cstart + i] = x * ni (fiedler.c:606-616)
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_out(mg_args a, const double *x, double *fied) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_ELIG, t)) return;
  const double ni = MSC(t.s, MS_NI);
  const int cstart = a.seg[t.s];
  MG_ROWS(t, row) fied[cstart + (row - t.pstart)] = x[row] * ni;
}

// f.Lf, Lf.Lf, f.f for the Rayleigh quotient and residual of the result
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_rayleigh(mg_args a, mg_sell L, const double *x) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_ELIG, t)) return;
  double acc[3] = {0.0, 0.0, 0.0};
  MG_ROWS(t, row) {
    const double f = x[row];
    const double lf = L.diag[row] * f - sell_offsum(L, row, x);
    acc[0] += f * lf, acc[1] += lf * lf, acc[2] += f * f;
  }
  mg_publish<3>(a, t, acc);
}

// ---------------------------------------------------------------------------
// Gram-matrix refinement of the next pass's start vector (rsb_mg_grammian =
// 1).  The reference (fiedler.c:250-305) keeps a basis Z of the normalised
// iterates, forms M = Z^T L Z and takes the smallest Ritz vector as the next
// start; its buffers are laid out so that rows 1.. of Z overlap M (Z = err +
// lelt, M = Z + lelt, fiedler.c:221-222), so the steps are those its comments
// state, made stable: the new iterate is orthogonalised against Z twice
// (classical Gram-Schmidt, twice), it joins Z only if what is left has norm
// > 1e-10, M grows by the new vector's row and column, and the Ritz vector is
// the smallest eigenvector of sym(M) by cyclic Jacobi, signed to point along
// the iterate (sum_j v_j (Z_j . z) >= 0).  Segments whose iteration stops
// this pass skip the step: it only changes the start of a pass that follows.

struct mg_gram {
  double *Z, *GZ;  // cap vectors of mpad rows each: the basis and L times it
  double *M;       // cap x cap per segment: M[k][j] = Z_k . L Z_j
  double *cg;      // MG_NP per segment: the current Gram-Schmidt sweep's coefficients
  double *c;       // MG_NP per segment: Z_j . z over both sweeps
  double *v;       // MG_NP per segment: the Ritz vector's coefficients
  int cap;         // basis capacity (rsb_max_passes - 1 <= MG_NP - 1)
};

// slots k0 .. k0 + ns - 1 of the tile's partials <- v, summed over the workgroup
__device__ inline void mg_publish_at(const mg_args &a, const mg_tile &t, double (&v)[4], int k0,
                                     int ns) {
  __shared__ double red[(RSB_BLOCK / RSB_WAVE) * 4];
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < ns; k++) a.partial[(size_t)(k0 + k) * a.ntiles + t.ti] = v[k];
}

// partial slot j = X_j . y for j < MI_NB + extra, X_j = X + j mpad; with
// xlast given, X_nb is xlast (the newest vector), which is also stored to
// keep + nb mpad when keep is given
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_gdots(mg_args a, int gate, int64_t mpad,
                                                       const double *X, const double *xlast,
                                                       double *keep, int extra, const double *y) {
#pragma clang fp contract(off)
  mg_tile t;
  if (!mg_tile_of(a, gate, t)) return;
  const int nb = MST(t.s, MI_NB), cnt = nb + extra;
  if (keep) MG_ROWS(t, row) keep[(size_t)nb * mpad + row] = xlast[row];
  for (int j0 = 0; j0 < cnt; j0 += 4) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    MG_ROWS(t, row) {
      const double yv = y[row];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int j = j0 + k;
        if (j < cnt) {
          const double xv = (xlast && j == nb) ? xlast[row] : X[(size_t)j * mpad + row];
          acc[k] += xv * yv;
        }
      }
    }
    mg_publish_at(a, t, acc, j0, min(4, cnt - j0));
  }
}

// dst = src - sum_j cg_j Z_j (j in order); partial dst.dst
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_gsub(mg_args a, mg_gram G, int64_t mpad,
                                                      const double *src, double *dst) {
#pragma clang fp contract(off)
  mg_tile t;
  if (!mg_tile_of(a, GATE_INV, t)) return;
  const int nb = MST(t.s, MI_NB);
  const double *cg = G.cg + (size_t)t.s * MG_NP;
  double acc[1] = {0.0};
  MG_ROWS(t, row) {
    double v = src[row];
    for (int j = 0; j < nb; j++) v = v - cg[j] * G.Z[(size_t)j * mpad + row];
    dst[row] = v;
    acc[0] += v * v;
  }
  mg_publish<1>(a, t, acc);
}

// the new basis vector Z_nb = r / |r| (also into p for L p); partial Z_nb . z
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_gnew(mg_args a, mg_gram G, int64_t mpad,
                                                      const double *r, const double *z,
                                                      double *p) {
#pragma clang fp contract(off)
  mg_tile t;
  if (!mg_tile_of(a, GATE_GRAM, t)) return;
  const int nb = MST(t.s, MI_NB);
  const double rni = MSC(t.s, MS_GRNI);
  double acc[1] = {0.0};
  MG_ROWS(t, row) {
    const double zi = r[row] * rni;
    p[row] = zi;
    G.Z[(size_t)nb * mpad + row] = zi;
    acc[0] += zi * z[row];
  }
  mg_publish<1>(a, t, acc);
}

// z = sum_j v_j Z_j (j in order); partial sum z
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_gcomb(mg_args a, mg_gram G, int64_t mpad,
                                                       double *z) {
#pragma clang fp contract(off)
  mg_tile t;
  if (!mg_tile_of(a, GATE_GMIX, t)) return;
  const int nb = MST(t.s, MI_NB);
  const double *v = G.v + (size_t)t.s * MG_NP;
  double acc[1] = {0.0};
  MG_ROWS(t, row) {
    double x = 0.0;
    for (int j = 0; j < nb; j++) x += G.Z[(size_t)j * mpad + row] * v[j];
    z[row] = x;
    acc[0] += x;
  }
  mg_publish<1>(a, t, acc);
}

// what k_mg_gfin does with a segment's totals
enum { GF_CGS0 = 0, GF_CGS1, GF_NORM, GF_CNEW, GF_MROW, GF_MCOL };

__global__ void __launch_bounds__(RSB_BLOCK) k_mg_gfin(mg_args a, mg_gram G, int op) {
  __shared__ double red[(RSB_BLOCK / RSB_WAVE) * 4];
  const int s = blockIdx.x;
  if (!mg_gate(a, s, op <= GF_NORM ? GATE_INV : GATE_GRAM)) return;
  const int first = a.seg_tiles[2 * s], count = a.seg_tiles[2 * s + 1];
  const int nb = MST(s, MI_NB);
  const int cnt = op == GF_MROW ? nb + 1 : (op == GF_NORM || op == GF_CNEW) ? 1 : nb;
  double *cg = G.cg + (size_t)s * MG_NP, *c = G.c + (size_t)s * MG_NP;
  double *M = G.M + (size_t)s * G.cap * G.cap;
  for (int k0 = 0; k0 < cnt; k0 += 4) {
    double tot[4];
    const int ns = min(4, cnt - k0);
    seg_totals(a, first, count, k0, ns, tot, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < ns; k++) {
        const int j = k0 + k;
        const double x = tot[k];
        switch (op) {
        case GF_CGS0: cg[j] = x, c[j] = x; break;
        case GF_CGS1: cg[j] = x, c[j] += x; break;
        case GF_NORM: {
          const double rn = sqrt(x);
          MST(s, MI_GADD) = rn > 1e-10;
          MSC(s, MS_GRNI) = 1.0 / rn;
          break;
        }
        case GF_CNEW: c[nb] = x; break;
        case GF_MROW: M[(size_t)nb * G.cap + j] = x; break;
        case GF_MCOL: M[(size_t)j * G.cap + nb] = x; break;
        }
      }
  }
  __syncthreads();
  if (threadIdx.x == 0 && op == GF_MCOL) MST(s, MI_NB) = nb + 1;
}

// Smallest eigenpair of the symmetric N x N matrix A (row-major in LDS,
// overwritten; V: N x N LDS workspace) by cyclic Jacobi rotations (Golub &
// Van Loan, Matrix Computations, 4th ed., Alg. 8.5.1 and 8.5.3), one wave:
// sweeps over p < q in order, a rotation skipped when |a_pq| <= eps^2
// (|a_pp| + |a_qq|), until a sweep rotates nothing (at most 64 sweeps).  The
// threads split the rows of each rotation; every entry takes the same
// operations in the same order as a serial sweep, so the result is the
// serial one bit for bit.  v <- the column of the smallest diagonal entry
// (lowest index on ties), unit length; returns that eigenvalue.
__device__ double gram_jacobi(double *A, double *V, int N, double *v) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < N * N; i += nt) V[i] = (i / N == i % N) ? 1.0 : 0.0;
  __syncthreads();
  for (int sweep = 0; sweep < 64; sweep++) {
    bool rotated = false;
    for (int p = 0; p < N - 1; p++)
      for (int q = p + 1; q < N; q++) {
        const double apq = A[p * N + q];
        if (apq == 0.0) continue;
        const double app = A[p * N + p], aqq = A[q * N + q];
        const double scale = fabs(app) + fabs(aqq);
        if (fabs(apq) <= DBL_EPSILON * DBL_EPSILON * scale) continue;
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                    : -1.0 / (-tau + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        __syncthreads();  // a_pq, a_pp, a_qq read by every thread
        for (int k = tid; k < N; k += nt) {
          const double akp = A[k * N + p], akq = A[k * N + q];
          A[k * N + p] = c * akp - s * akq;
          A[k * N + q] = s * akp + c * akq;
        }
        __syncthreads();
        for (int k = tid; k < N; k += nt) {
          const double apk = A[p * N + k], aqk = A[q * N + k];
          A[p * N + k] = c * apk - s * aqk;
          A[q * N + k] = s * apk + c * aqk;
          const double vkp = V[k * N + p], vkq = V[k * N + q];
          V[k * N + p] = c * vkp - s * vkq;
          V[k * N + q] = s * vkp + c * vkq;
        }
        __syncthreads();
        if (tid == 0) A[p * N + q] = A[q * N + p] = 0.0;
        __syncthreads();
        rotated = true;
      }
    if (!rotated) break;
  }
  int imin = 0;
  for (int i = 1; i < N; i++)
    if (A[i * N + i] < A[imin * N + imin]) imin = i;
  double nrm = 0.0;
  for (int k = 0; k < N; k++) nrm += V[k * N + imin] * V[k * N + imin];
  nrm = 1.0 / sqrt(nrm);
  for (int k = tid; k < N; k += nt) v[k] = V[k * N + imin] * nrm;
  __syncthreads();
  return A[imin * N + imin];
}

// Ritz vector of every segment with >= 2 basis vectors: smallest eigenvector
// of sym(M), signed so that sum_j v_j c_j >= 0.  One wave per segment; LDS
// (2 N^2 + N) doubles.
__global__ void __launch_bounds__(RSB_WAVE) k_mg_gjac(mg_args a, mg_gram G) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  const int s = blockIdx.x;
  if (!mg_gate(a, s, GATE_GMIX)) return;
  const int N = MST(s, MI_NB), cap = G.cap;
  double *A = lds, *V = lds + N * N, *v = V + N * N;
  const double *M = G.M + (size_t)s * cap * cap;
  for (int i = threadIdx.x; i < N * N; i += blockDim.x) {
    const int k = i / N, j = i % N;
    A[i] = 0.5 * (M[(size_t)k * cap + j] + M[(size_t)j * cap + k]);
  }
  __syncthreads();
  (void)gram_jacobi(A, V, N, v);
  if (threadIdx.x == 0) {
    const double *c = G.c + (size_t)s * MG_NP;
    double dir = 0.0;
    for (int j = 0; j < N; j++) dir += v[j] * c[j];
    double *out = G.v + (size_t)s * MG_NP;
    for (int j = 0; j < N; j++) out[j] = dir < 0.0 ? -v[j] : v[j];
  }
}

// parity hook of gram_jacobi on one matrix
__global__ void __launch_bounds__(RSB_WAVE) k_gjac_test(int n, const double *Ain, double *v,
                                                        double *lam) {
  extern __shared__ double lds[];
  double *A = lds, *V = lds + n * n, *vv = V + n * n;
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) A[i] = Ain[i];
  __syncthreads();
  const double l = gram_jacobi(A, V, n, vv);
  for (int k = threadIdx.x; k < n; k += blockDim.x) v[k] = vv[k];
  if (threadIdx.x == 0) *lam = l;
}

// ---------------------------------------------------------------------------
// V-cycle.  Weighted Jacobi step k (multigrid.c:249-268): s = sigma_k r / D
// (0 where D is 0), u = s on the first step and u += s after, r -= D s - A s.
// The step's s is formed once per row by the kernel before (k_mg_jac* for
// the first, the residual kernel of step k - 1 for the others), so the
// residual kernel reads its neighbours' s and writes the next step's s
// into the other buffer.

__global__ void __launch_bounds__(RSB_BLOCK) k_mg_jac0(mg_args a, mg_sell L, double sigma,
                                                      const double *r, double *sv, double *u) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  MG_ROWS(t, row) {
    const double d = L.diag[row];
    const double v = d != 0.0 ? sigma * r[row] / d : 0.0;
    sv[row] = v, u[row] = v;
  }
}

// rout = rin - (D s - A s); then, if sigma_next > 0, the next step's s
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_res0(mg_args a, mg_sell L, const double *rin,
                                                      double *rout, const double *sv,
                                                      double sigma_next, double *snext,
                                                      double *u) {
  __shared__ double xl[MG_TB_LOCAL];
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  const bool tiled = mg_stage(L, t, sv, xl);
  MG_ROWS(t, row) {
    const double d = L.diag[row];
    const double r = tiled ? rin[row] - (d * xl[row - t.row0] - tile_offsum(L, row, xl))
                           : rin[row] - (d * sv[row] - sell_offsum(L, row, sv));
    rout[row] = r;
    if (sigma_next > 0.0) {
      const double v = d != 0.0 ? sigma_next * r / d : 0.0;
      snext[row] = v, u[row] += v;
    }
  }
}

// coarse level: rows [0, m) with row -> segment rseg; CSR (rp, ci, w), diag
struct mg_csr {
  int64_t m;
  const int *rseg;
  const int64_t *soff;  // SELL-64 slices
  const int *rlen, *sci, *sw;
  const double *diag;
  const uint32_t *spk;  // if set: packed entries col | weight << 24 (sci, sw unused)
};

__global__ void __launch_bounds__(RSB_BLOCK) k_mg_jac_c(mg_args a, mg_csr M, double sigma,
                                                       const double *r, double *sv, double *u) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.m || !MST(M.rseg[i], MI_PCG)) return;
  const double d = M.diag[i];
  const double v = d != 0.0 ? sigma * r[i] / d : 0.0;
  sv[i] = v, u[i] = v;
}

__global__ void __launch_bounds__(RSB_BLOCK) k_mg_res_c(mg_args a, mg_csr M, double *r,
                                                       const double *sv, double sigma_next,
                                                       double *snext, double *u) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.m || !MST(M.rseg[i], MI_PCG)) return;
  double as = 0.0;
  const int64_t base = M.soff[i / RSB_WAVE] + (i % RSB_WAVE);
  const int len = M.rlen[i];
  // eight entries and their gathers in flight, summed in order
  for (int j0 = 0; j0 < len; j0 += 8) {
    int c[8], wt[8];
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int64_t k = base + (int64_t)(j0 + u) * RSB_WAVE;
      const bool in = j0 + u < len;
      if (M.spk) {
        const uint32_t e = in ? M.spk[k] : (uint32_t)i;
        c[u] = (int)(e & 0xffffffu), wt[u] = (int)(e >> 24);
      } else {
        c[u] = in ? M.sci[k] : (int)i;
        wt[u] = in ? M.sw[k] : 0;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = sv[c[u]];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (j0 + u < len) as += (double)wt[u] * v[u];
  }
  const double d = M.diag[i];
  const double rv = r[i] - (d * sv[i] - as);
  r[i] = rv;
  if (sigma_next > 0.0) {
    const double v = d != 0.0 ? sigma_next * rv / d : 0.0;
    snext[i] = v, u[i] += v;
  }
}

// per-level segment layout: off[s], n[s] (n = 0: the segment has no rows on
// this level); level 0 uses pstart and the segment size
struct mg_lvl {
  const int *off, *n;
};

// restriction: coarse row c = sum of its children (multigrid.c:271-272)
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_restrict(mg_args a, int64_t mc,
                                                          const int *rseg_c, mg_lvl F,
                                                          mg_lvl Cl, int factor,
                                                          const double *rf, double *rc) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= mc) return;
  const int s = rseg_c[c];
  if (!MST(s, MI_PCG)) return;
  const int k = (int)(c - Cl.off[s]);
  const int nf = F.n[s], f0 = F.off[s];
  const int i0 = k * factor, i1 = min(i0 + factor, nf);
  double v = 0.0;
  for (int i = i0; i < i1; i++) v += rf[f0 + i];
  rc[c] = v;
}

// prolongation, coarse levels: r_f = over * r_c (0 on a one-row level) + u_f
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_prolong_c(mg_args a, int64_t mf,
                                                           const int *rseg_f, mg_lvl F,
                                                           mg_lvl Cl, int factor, double over,
                                                           const double *rc, const double *u,
                                                           double *rf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mf) return;
  const int s = rseg_f[i];
  if (!MST(s, MI_PCG) || F.n[s] <= 1) return;
  const int k = (int)(i - F.off[s]);
  const double cv = Cl.n[s] == 1 ? 0.0 : rc[Cl.off[s] + k / factor];
  rf[i] = over * cv + u[i];
}

// prolongation onto level 0 (tiles), result into z
__global__ void __launch_bounds__(RSB_BLOCK) k_mg_prolong0(mg_args a, mg_lvl Cl, int factor,
                                                          double over, const double *rc,
                                                          const double *u, double *z) {
  mg_tile t;
  if (!mg_tile_of(a, GATE_PCG, t)) return;
  const int coff = Cl.off[t.s], cn = Cl.n[t.s];
  MG_ROWS(t, row) {
    const int k = row - t.pstart;
    const double cv = cn == 1 ? 0.0 : rc[coff + k / factor];
    z[row] = over * cv + u[row];
  }
}

// ---------------------------------------------------------------------------
// Hierarchy build.

__global__ void k_mg_diag0(int64_t mpad, mg_sell L, double *diag) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= mpad) return;
  const int64_t base = L.slice_off[row / RSB_WAVE] + (row % RSB_WAVE);
  int s = 0;
  for (int j = 0; j < L.rowlen[row]; j++) s += (int)(L.ell[base + (int64_t)j * RSB_WAVE] >> 28);
  diag[row] = (double)s;
}

__global__ void k_mg_rseg(int nseg, const int *off, const int *n, int *rseg) {
  const int s = blockIdx.x;
  for (int i = threadIdx.x; i < n[s]; i += blockDim.x) rseg[off[s] + i] = s;
}

// raw entry count of every coarse row (sum of its children's row lengths)
__global__ void k_mg_rawcnt(int64_t mc, const int *rseg_c, mg_lvl F, mg_lvl Cl, int factor,
                            int level, const int *rowlen0, const int64_t *rp_f,
                            int64_t *cnt) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > mc) return;
  if (c == mc) {
    cnt[c] = 0;
    return;
  }
  const int s = rseg_c[c];
  const int k = (int)(c - Cl.off[s]);
  const int i0 = k * factor, i1 = min(i0 + factor, F.n[s]);
  int64_t v = 0;
  for (int i = i0; i < i1; i++) {
    const int64_t row = (int64_t)F.off[s] + i;
    v += level == 0 ? rowlen0[row] : rp_f[row + 1] - rp_f[row];
  }
  cnt[c] = v;
}

// raw entries: key (coarse row << 32 | coarse column), weight, index
__global__ void k_mg_rawfill(int64_t mc, const int *rseg_c, mg_lvl F, mg_lvl Cl, int factor,
                             int level, mg_sell L0, const int64_t *rp_f, const int *ci_f,
                             const double *w_f, const int64_t *off, uint64_t *key, int *idx,
                             double *wr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= mc) return;
  const int s = rseg_c[c];
  const int k = (int)(c - Cl.off[s]);
  const int i0 = k * factor, i1 = min(i0 + factor, F.n[s]);
  const int foff = F.off[s], coff = Cl.off[s];
  int64_t o = off[c];
  for (int i = i0; i < i1; i++) {
    const int64_t row = (int64_t)foff + i;
    if (level == 0) {
      const int64_t base = L0.slice_off[row / RSB_WAVE] + (row % RSB_WAVE);
      for (int j = 0; j < L0.rowlen[row]; j++) {
        const uint32_t e = L0.ell[base + (int64_t)j * RSB_WAVE];
        const int cc = coff + ((int)(e & 0x0fffffffu) - foff) / factor;
        key[o] = ((uint64_t)c << 32) | (uint32_t)cc, idx[o] = (int)o, wr[o] = (double)(e >> 28);
        o++;
      }
    } else {
      for (int64_t j = rp_f[row]; j < rp_f[row + 1]; j++) {
        const int cc = coff + (ci_f[j] - foff) / factor;
        key[o] = ((uint64_t)c << 32) | (uint32_t)cc, idx[o] = (int)o, wr[o] = w_f[j];
        o++;
      }
    }
  }
}

// keep the first of each run of equal keys that is off the diagonal; its
// weight is the run's sum
__global__ void k_mg_heads(int64_t nr, const uint64_t *key, const int *idx, const double *wr,
                           int *flag, double *wsum, int *rowcnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  const uint64_t k = key[i];
  const bool head = i == 0 || key[i - 1] != k;
  const bool diag = (uint32_t)(k >> 32) == (uint32_t)k;
  const int keep = head && !diag;
  flag[i] = keep;
  if (keep) {
    double v = 0.0;
    for (int64_t j = i; j < nr && key[j] == k; j++) v += wr[idx[j]];
    wsum[i] = v;
    atomicAdd(&rowcnt[k >> 32], 1);
  }
}

__global__ void k_mg_compact(int64_t nr, const uint64_t *key, const int *flag, const int64_t *pos,
                             const double *wsum, int *ci, double *w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr || !flag[i]) return;
  ci[pos[i]] = (int)(uint32_t)key[i];
  w[pos[i]] = wsum[i];
}

__global__ void k_mg_widen(int64_t n, const int *in, int64_t *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// SELL-64 layout of a coarse level: rlen per row (0 past the last row),
// bytes[slice] = 64 * longest row of the slice (entries per slice)
__global__ void k_mg_slice_len(int64_t m, int64_t nsl, const int64_t *rp, int *rlen,
                               int64_t *ents) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsl * RSB_WAVE) {
    if (i == nsl * RSB_WAVE) ents[nsl] = 0;
    return;
  }
  const int len = i < m ? (int)(rp[i + 1] - rp[i]) : 0;
  rlen[i] = len;
  int mx = len;
  for (int o = RSB_WAVE / 2; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, RSB_WAVE));
  if ((threadIdx.x & (RSB_WAVE - 1)) == 0) ents[i / RSB_WAVE] = (int64_t)mx * RSB_WAVE;
}

// spk != nullptr: one packed word per entry (col | weight << 24; the caller
// has checked m <= 2^24 and every weight < 256), else two int arrays
__global__ void k_mg_sell_fill(int64_t m, int64_t nsl, const int64_t *rp, const int *ci,
                               const double *w, const int64_t *soff, int *sci, int *sw,
                               uint32_t *spk) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsl * RSB_WAVE) return;
  const int64_t sl = i / RSB_WAVE, lane = i % RSB_WAVE;
  const int width = (int)((soff[sl + 1] - soff[sl]) / RSB_WAVE);
  const int64_t base = soff[sl] + lane;
  int j = 0;
  if (i < m)
    for (int64_t k = rp[i]; k < rp[i + 1]; k++, j++) {
      const int64_t o = base + (int64_t)j * RSB_WAVE;
      if (spk) spk[o] = (uint32_t)ci[k] | ((uint32_t)(int)w[k] << 24);
      else sci[o] = ci[k], sw[o] = (int)w[k];
    }
  const int self = i < m ? (int)i : 0;
  for (; j < width; j++) {
    const int64_t o = base + (int64_t)j * RSB_WAVE;
    if (spk) spk[o] = (uint32_t)self;
    else sci[o] = self, sw[o] = 0;
  }
}

// largest weight of a coarse matrix (non-negative integers)
__global__ void k_mg_wmax(int64_t nnz, const double *w, int *out) {
  int mx = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x)
    mx = max(mx, w[k] < 2147483647.0 ? (int)w[k] : 2147483647);
  for (int o = RSB_WAVE / 2; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, RSB_WAVE));
  if ((threadIdx.x & (RSB_WAVE - 1)) == 0) atomicMax(out, mx);
}

__global__ void k_mg_rowdiag(int64_t m, const int64_t *rp, const double *w, double *diag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int64_t k = rp[i]; k < rp[i + 1]; k++) s += w[k];
  diag[i] = s;
}

// ---------------------------------------------------------------------------
// Host side.

namespace {

struct mg_level_dev {
  int64_t m = 0;
  int *off = nullptr, *n = nullptr, *rseg = nullptr;
  int64_t *rp = nullptr;
  int *ci = nullptr;
  double *w = nullptr, *diag = nullptr;
  // the same matrix as SELL-64 for the smoother: slices of 64 rows, entries
  // column-major within a slice (a wave's loads are contiguous), padded with
  // weight-0 entries to the slice's longest row; weights are integers
  int64_t *soff = nullptr;
  int *rlen = nullptr, *sci = nullptr, *sw = nullptr;
  uint32_t *spk = nullptr;  // packed col | weight << 24 where it fits (half the bytes)
  double *r0 = nullptr, *sa = nullptr, *sb = nullptr, *u = nullptr;  // V-cycle vectors
};

struct mg_hier {
  std::vector<mg_level_dev> lv;
  std::vector<std::vector<int>> h_off, h_n;
  ~mg_hier() {
    for (auto &l : lv) {
      rsb_dev_free(l.off), rsb_dev_free(l.n), rsb_dev_free(l.rseg), rsb_dev_free(l.rp), rsb_dev_free(l.ci);
      rsb_dev_free(l.w), rsb_dev_free(l.diag), rsb_dev_free(l.r0), rsb_dev_free(l.sa), rsb_dev_free(l.sb), rsb_dev_free(l.u);
      rsb_dev_free(l.soff), rsb_dev_free(l.rlen), rsb_dev_free(l.sci), rsb_dev_free(l.sw), rsb_dev_free(l.spk);
    }
  }
};

double sigma_cheb(int k, int n, double lmin, double lmax) {
  k = (k - 1) % n + 1;
  const double theta = M_PI * (k - 0.5) / n;
  return 1.0 / (lmin + 0.5 * (lmax - lmin) * (cos(theta) + 1.0));
}

}  // namespace

// The hierarchy of the last rsbgpu_fiedler_mg call stays with the handle
// (for the parity hook rsbgpu_mg_level_csr) until the next call or close.
void rsb_mg_free(rsbgpu *g) {
  delete (mg_hier *)g->mg;
  g->mg = nullptr;
}

#define MG_TRY(x)             \
  do {                        \
    int rc_ = (x);            \
    if (rc_) return rc_;      \
  } while (0)

// Device buffers of one scope (a level of mg_build, or one rsbgpu_fiedler_mg
// call with its CG-iteration graph): freed on every return path.  A capture still open when the scope ends (an
// error between begin and end) is ended and its graph dropped, so the
// handle's stream is never left capturing.  The buffers are freed after the
// stream has drained (an early return may leave launches queued).
struct mg_scope {
  hipStream_t st;
  std::vector<void *> bufs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool capturing = false;
  explicit mg_scope(hipStream_t s) : st(s) {}
  template <typename T>
  hipError_t alloc(T *&ptr, size_t bytes) {
    hipError_t e = rsb_dev_malloc((void **)&ptr, bytes);
    if (e == hipSuccess) bufs.push_back((void *)ptr);
    else ptr = nullptr;
    return e;
  }
  ~mg_scope() {
    if (capturing) {
      hipGraph_t left = nullptr;
      (void)hipStreamEndCapture(st, &left);
      if (left) (void)hipGraphDestroy(left);
    }
    (void)hipStreamSynchronize(st);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void *q : bufs) (void)rsb_dev_free(q);
  }
};

static int mg_build(rsbgpu *g, int factor, mg_hier &H, double *diag0) {
  hipStream_t st = g->stream;
  const int nseg = g->nseg;
  // level 0 layout: pstart / n of the eligible segments
  std::vector<int> off0(nseg), n0(nseg);
  for (int s = 0; s < nseg; s++) {
    const bool el = g->h_p[s] > 1 && g->h_n[s] >= 2;
    off0[s] = el ? g->h_pstart[s] : 0;
    n0[s] = el ? g->h_n[s] : 0;
  }
  H.h_off.push_back(off0), H.h_n.push_back(n0);
  if (H.lv.empty()) H.lv.emplace_back();
  {
    mg_level_dev &L0 = H.lv[0];
    L0.m = g->mpad;
    RSB_TRY(rsb_dev_malloc((void **)&L0.off, sizeof(int) * nseg));
    RSB_TRY(rsb_dev_malloc((void **)&L0.n, sizeof(int) * nseg));
    RSB_TRY(hipMemcpyAsync(L0.off, off0.data(), sizeof(int) * nseg, hipMemcpyHostToDevice, st));
    RSB_TRY(hipMemcpyAsync(L0.n, n0.data(), sizeof(int) * nseg, hipMemcpyHostToDevice, st));
    L0.diag = diag0;  // owned by the hierarchy (the caller put it there)
    RSB_TRY(rsb_dev_malloc((void **)&L0.r0, sizeof(double) * (size_t)std::max<int64_t>(g->mpad, 1)));
    RSB_TRY(rsb_dev_malloc((void **)&L0.sa, sizeof(double) * (size_t)std::max<int64_t>(g->mpad, 1)));
    RSB_TRY(rsb_dev_malloc((void **)&L0.sb, sizeof(double) * (size_t)std::max<int64_t>(g->mpad, 1)));
    RSB_TRY(rsb_dev_malloc((void **)&L0.u, sizeof(double) * (size_t)std::max<int64_t>(g->mpad, 1)));
  }
  mg_sell S{g->d_slice_off, g->d_ell, g->d_rowlen, diag0, nullptr, nullptr, nullptr, nullptr,
            nullptr,          nullptr};
  for (int l = 0;; l++) {
    const std::vector<int> nf = H.h_n[l];
    std::vector<int> nc(nseg, 0), oc(nseg, 0);
    int64_t mc = 0;
    for (int s = 0; s < nseg; s++) {
      nc[s] = nf[s] > 1 ? (nf[s] + factor - 1) / factor : 0;
      oc[s] = (int)mc;
      mc += nc[s];
    }
    if (mc == 0) break;
    if (mc >= (int64_t)1 << 31) RSB_FAIL(2, "MG level too large");
    H.h_off.push_back(oc), H.h_n.push_back(nc);
    H.lv.emplace_back();
    mg_level_dev &F = H.lv[l], &C = H.lv[l + 1];
    C.m = mc;
    RSB_TRY(rsb_dev_malloc((void **)&C.off, sizeof(int) * nseg));
    RSB_TRY(rsb_dev_malloc((void **)&C.n, sizeof(int) * nseg));
    RSB_TRY(hipMemcpyAsync(C.off, oc.data(), sizeof(int) * nseg, hipMemcpyHostToDevice, st));
    RSB_TRY(hipMemcpyAsync(C.n, nc.data(), sizeof(int) * nseg, hipMemcpyHostToDevice, st));
    RSB_TRY(rsb_dev_malloc((void **)&C.rseg, sizeof(int) * (size_t)mc));
    k_mg_rseg<<<nseg, 256, 0, st>>>(nseg, C.off, C.n, C.rseg);
    RSB_LAUNCH_CHECK();
    RSB_TRY(rsb_dev_malloc((void **)&C.r0, sizeof(double) * (size_t)mc));
    RSB_TRY(rsb_dev_malloc((void **)&C.sa, sizeof(double) * (size_t)mc));
    RSB_TRY(rsb_dev_malloc((void **)&C.sb, sizeof(double) * (size_t)mc));
    RSB_TRY(rsb_dev_malloc((void **)&C.u, sizeof(double) * (size_t)mc));
    RSB_TRY(rsb_dev_malloc((void **)&C.diag, sizeof(double) * (size_t)mc));
    RSB_TRY(rsb_dev_malloc((void **)&C.rp, sizeof(int64_t) * (size_t)(mc + 1)));
    // raw entries
    mg_scope tmp(st);  // this level's temporaries
    int64_t *cnt = nullptr, *roff = nullptr;
    RSB_TRY(tmp.alloc(cnt, sizeof(int64_t) * (size_t)(mc + 1)));
    RSB_TRY(tmp.alloc(roff, sizeof(int64_t) * (size_t)(mc + 1)));
    mg_lvl FL{F.off, F.n}, CL{C.off, C.n};
    k_mg_rawcnt<<<nblocks(mc + 1, 256), 256, 0, st>>>(mc, C.rseg, FL, CL, factor, l,
                                                       g->d_rowlen, F.rp, cnt);
    RSB_LAUNCH_CHECK();
    MG_TRY(rsb_exclusive_scan_i64(g, cnt, roff, mc + 1));
    int64_t nr = 0;
    RSB_TRY(hipMemcpyAsync(&nr, roff + mc, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipStreamSynchronize(st));
    if (nr >= (int64_t)1 << 31) RSB_FAIL(2, "MG level has too many entries");
    uint64_t *key = nullptr;
    int *idx = nullptr, *flag = nullptr, *rowcnt = nullptr;
    double *wr = nullptr, *wsum = nullptr;
    int64_t *pos = nullptr, *flag64 = nullptr, *rc64 = nullptr;
    const size_t nra = (size_t)std::max<int64_t>(nr, 1);
    RSB_TRY(tmp.alloc(key, sizeof(uint64_t) * nra));
    RSB_TRY(tmp.alloc(idx, sizeof(int) * nra));
    RSB_TRY(tmp.alloc(wr, sizeof(double) * nra));
    k_mg_rawfill<<<nblocks(mc, 256), 256, 0, st>>>(mc, C.rseg, FL, CL, factor, l, S, F.rp, F.ci,
                                                    F.w, roff, key, idx, wr);
    RSB_LAUNCH_CHECK();
    int cbits = 1;
    while (((int64_t)1 << cbits) < mc) cbits++;
    RSB_TRY(tmp.alloc(flag, sizeof(int) * nra));
    RSB_TRY(tmp.alloc(wsum, sizeof(double) * nra));
    RSB_TRY(tmp.alloc(rowcnt, sizeof(int) * (size_t)(mc + 1)));
    RSB_TRY(hipMemsetAsync(rowcnt, 0, sizeof(int) * (size_t)(mc + 1), st));
    RSB_TRY(tmp.alloc(flag64, sizeof(int64_t) * nra));
    RSB_TRY(tmp.alloc(pos, sizeof(int64_t) * nra));
    if (nr > 0) {
      MG_TRY(rsb_sort_pairs_u64(g, key, idx, nr, 0, 32 + cbits));
      k_mg_heads<<<nblocks(nr, 256), 256, 0, st>>>(nr, key, idx, wr, flag, wsum, rowcnt);
      RSB_LAUNCH_CHECK();
      k_mg_widen<<<nblocks(nr, 256), 256, 0, st>>>(nr, flag, flag64);
      MG_TRY(rsb_exclusive_scan_i64(g, flag64, pos, nr));
    }
    RSB_TRY(tmp.alloc(rc64, sizeof(int64_t) * (size_t)(mc + 1)));
    k_mg_widen<<<nblocks(mc + 1, 256), 256, 0, st>>>(mc + 1, rowcnt, rc64);
    MG_TRY(rsb_exclusive_scan_i64(g, rc64, C.rp, mc + 1));
    int64_t nnz = 0;
    RSB_TRY(hipMemcpyAsync(&nnz, C.rp + mc, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipStreamSynchronize(st));
    RSB_TRY(rsb_dev_malloc((void **)&C.ci, sizeof(int) * (size_t)std::max<int64_t>(nnz, 1)));
    RSB_TRY(rsb_dev_malloc((void **)&C.w, sizeof(double) * (size_t)std::max<int64_t>(nnz, 1)));
    if (nr > 0) k_mg_compact<<<nblocks(nr, 256), 256, 0, st>>>(nr, key, flag, pos, wsum, C.ci, C.w);
    k_mg_rowdiag<<<nblocks(mc, 256), 256, 0, st>>>(mc, C.rp, C.w, C.diag);
    RSB_LAUNCH_CHECK();
    // SELL-64 copy for the smoother
    {
      const int64_t nsl = (mc + RSB_WAVE - 1) / RSB_WAVE;
      int64_t *sbytes = nullptr;
      RSB_TRY(rsb_dev_malloc((void **)&C.rlen, sizeof(int) * (size_t)(nsl * RSB_WAVE)));
      RSB_TRY(rsb_dev_malloc((void **)&C.soff, sizeof(int64_t) * (size_t)(nsl + 1)));
      RSB_TRY(tmp.alloc(sbytes, sizeof(int64_t) * (size_t)(nsl + 1)));
      k_mg_slice_len<<<nblocks(nsl * RSB_WAVE, 256), 256, 0, st>>>(mc, nsl, C.rp, C.rlen, sbytes);
      RSB_LAUNCH_CHECK();
      MG_TRY(rsb_exclusive_scan_i64(g, sbytes, C.soff, nsl + 1));
      int64_t ns = 0;
      RSB_TRY(hipMemcpyAsync(&ns, C.soff + nsl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
      RSB_TRY(hipStreamSynchronize(st));
      // one packed word per entry where the row index fits in 24 bits and
      // the weights in 8 (the large coarse levels), two ints otherwise
      int *dmax = nullptr, hmax = 0;
      RSB_TRY(tmp.alloc(dmax, sizeof(int)));
      RSB_TRY(hipMemsetAsync(dmax, 0, sizeof(int), st));
      if (nnz > 0) {
        k_mg_wmax<<<std::min<int64_t>(nblocks(nnz, 256), 2048), 256, 0, st>>>(nnz, C.w, dmax);
        RSB_LAUNCH_CHECK();
      }
      RSB_TRY(hipMemcpyAsync(&hmax, dmax, sizeof(int), hipMemcpyDeviceToHost, st));
      RSB_TRY(hipStreamSynchronize(st));
      if (mc <= (1 << 24) && hmax < 256) {
        RSB_TRY(rsb_dev_malloc((void **)&C.spk, sizeof(uint32_t) * (size_t)std::max<int64_t>(ns, 1)));
      } else {
        RSB_TRY(rsb_dev_malloc((void **)&C.sci, sizeof(int) * (size_t)std::max<int64_t>(ns, 1)));
        RSB_TRY(rsb_dev_malloc((void **)&C.sw, sizeof(int) * (size_t)std::max<int64_t>(ns, 1)));
      }
      k_mg_sell_fill<<<nblocks(nsl * RSB_WAVE, 256), 256, 0, st>>>(mc, nsl, C.rp, C.ci, C.w,
                                                                    C.soff, C.sci, C.sw, C.spk);
      RSB_LAUNCH_CHECK();
      RSB_TRY(hipStreamSynchronize(st));
    }
  }
  return 0;
}

extern "C" int rsbgpu_fiedler_mg(rsbgpu *g, int max_iter, int max_passes, double tol,
                                 int factor, int grammian, rsbgpu_fiedler_out *out) {
  RSB_TRY(hipSetDevice(g->device));
  if (factor < 2) RSB_FAIL(3, "rsb_mg_factor must be > 1");
  if (max_iter < 1 || max_iter > MG_NP - 4) RSB_FAIL(3, "rsb_max_iter must be in 1..%d", MG_NP - 4);
  if (grammian && max_passes > MG_NP)
    RSB_FAIL(3, "rsb_max_passes must be <= %d with rsb_mg_grammian = 1", MG_NP);
  const int nseg = g->nseg;
  rsb_phase_begin(g);
  hipStream_t st = g->stream;
  const int64_t mpad = g->mpad;
  const int ntiles = g->ntiles;
  for (int s = 0; s < nseg; s++) {
    if (out->niter) out->niter[s] = 0;
    if (out->npass) out->npass[s] = 0;
    if (out->lambda) out->lambda[s] = 0.0;
    if (out->rnorm) out->rnorm[3 * s] = out->rnorm[3 * s + 1] = out->rnorm[3 * s + 2] = 0.0;
    if (out->resid) out->resid[s] = 0.0;
    if (out->fallback) out->fallback[s] = 0;
  }
  // segments that cannot be bisected get Fiedler value 0 (as the Lanczos path)
  RSB_TRY(dev_alloc(g, g->d_fied, (size_t)std::max(g->nel, 1)));
  RSB_TRY(hipMemsetAsync(g->d_fied, 0, sizeof(double) * (size_t)std::max(g->nel, 1), st));
  bool any = false;
  for (int s = 0; s < nseg; s++) any |= g->h_p[s] > 1 && g->h_n[s] >= 2;
  if (!any || mpad == 0) {
    rsb_phase_end(g, PH_LANCZOS);
    return 0;
  }
  // vectors and state: owned by `own`, whose destructor frees them (and the
  // graph) on every return path; diag0 belongs to the new hierarchy, which
  // stays with the handle
  const size_t vb = sizeof(double) * (size_t)mpad;
  mg_scope own(st);
  rsb_mg_free(g);
  mg_hier *hp = new mg_hier();
  g->mg = hp;
  mg_hier &H = *hp;
  H.lv.emplace_back();
  RSB_TRY(rsb_dev_malloc((void **)&H.lv[0].diag, vb));
  double *const diag0 = H.lv[0].diag;
  double *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *z0 = nullptr, *p = nullptr,
         *w = nullptr, *Wb = nullptr, *Pb = nullptr, *sc = nullptr, *partial = nullptr,
         *coef = nullptr;
  int *stt = nullptr, *active = nullptr;
  RSB_TRY(own.alloc(b, vb));
  RSB_TRY(own.alloc(x, vb));
  RSB_TRY(own.alloc(r, vb));
  RSB_TRY(own.alloc(z, vb));
  RSB_TRY(own.alloc(z0, vb));
  RSB_TRY(own.alloc(p, vb));
  RSB_TRY(own.alloc(w, vb));
  RSB_TRY(own.alloc(Wb, vb * (size_t)max_iter));
  RSB_TRY(own.alloc(Pb, vb * (size_t)max_iter));
  RSB_TRY(own.alloc(sc, sizeof(double) * MS_NSC * (size_t)nseg));
  RSB_TRY(own.alloc(stt, sizeof(int) * MI_NI * (size_t)nseg));
  RSB_TRY(own.alloc(partial, sizeof(double) * MG_NP * (size_t)std::max(ntiles, 1)));
  RSB_TRY(own.alloc(coef, sizeof(double) * MG_NP * (size_t)nseg));
  RSB_TRY(own.alloc(active, sizeof(int)));
  // Gram basis: at most one vector per pass that another pass follows
  mg_gram G{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, max_passes - 1};
  const bool gram = grammian && G.cap >= 1;
  if (gram) {
    RSB_TRY(own.alloc(G.Z, vb * (size_t)G.cap));
    RSB_TRY(own.alloc(G.GZ, vb * (size_t)G.cap));
    RSB_TRY(own.alloc(G.M, sizeof(double) * (size_t)G.cap * G.cap * (size_t)nseg));
    RSB_TRY(own.alloc(G.cg, sizeof(double) * MG_NP * (size_t)nseg));
    RSB_TRY(own.alloc(G.c, sizeof(double) * MG_NP * (size_t)nseg));
    RSB_TRY(own.alloc(G.v, sizeof(double) * MG_NP * (size_t)nseg));
  }
  RSB_TRY(hipMemsetAsync(sc, 0, sizeof(double) * MS_NSC * (size_t)nseg, st));
  // every eligible segment starts in the inverse iteration
  {
    std::vector<int> h((size_t)MI_NI * nseg, 0);
    for (int s = 0; s < nseg; s++) h[(size_t)s * MI_NI + MI_INV] = g->h_p[s] > 1 && g->h_n[s] >= 2;
    RSB_TRY(hipMemcpyAsync(stt, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, st));
  }
  mg_args a{nseg, g->d_seg, g->d_tile, g->d_seg_tiles, sc, stt, partial, active, tol,
            max_iter, max_passes, ntiles};
  mg_sell S{g->d_slice_off, g->d_ell, g->d_rowlen, diag0,   g->d_tnhalo, g->d_thalo_off,
            g->d_halo,      g->d_off16, g->d_w16,     g->d_e16};
  k_mg_diag0<<<nblocks(mpad, 256), 256, 0, st>>>(mpad, S, diag0);
  RSB_LAUNCH_CHECK();
  MG_TRY(mg_build(g, factor, H, diag0));
  const int nlev = (int)H.lv.size();
  auto fin = [&](int op, int gate, int nslots) {
    k_mg_fin<<<nseg, RSB_BLOCK, 0, st>>>(a, op, gate, nslots);
  };
  auto count_active = [&](int &n) -> int {
    RSB_TRY(hipMemcpyAsync(&n, active, sizeof(int), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipStreamSynchronize(st));
    return 0;
  };
  // start vector, orthogonalised and normalised (fiedler.c:572-585)
  k_mg_start<<<ntiles, RSB_BLOCK, 0, st>>>(a, b);
  fin(FIN_MEAN, GATE_ELIG, 1);
  k_mg_center_dot<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_ELIG, b);
  fin(FIN_NORM, GATE_ELIG, 1);
  k_mg_scale_sum<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_ELIG, b, b);
  RSB_LAUNCH_CHECK();
  const double s1 = sigma_cheb(1, 4, 1.0, 2.0), s2 = sigma_cheb(2, 4, 1.0, 2.0),
               s3 = sigma_cheb(3, 4, 1.0, 2.0);
  const double sig[3] = {s1, s2, s3};
  // One CG iteration is part_a (w = L p, alpha, the update of x and r, the
  // stop test) and part_b (z = V-cycle(r), beta, the new direction made
  // L-orthogonal to the earlier ones).  The host reads the count of running
  // segments after each part_a; part_b + part_a of the next iteration run as
  // one captured graph (no memset nodes: counters are cleared by a kernel).
  auto part_a = [&]() -> int {
    k_mg_apply<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_PCG, S, p, w);
    fin(FIN_ALPHA, GATE_PCG, 1);
    k_mg_update<<<ntiles, RSB_BLOCK, 0, st>>>(a, mpad, p, w, x, r, Wb, Pb);
    k_mg_zero<<<1, 1, 0, st>>>(active);
    fin(FIN_CONV, GATE_PCG, 1);
    RSB_LAUNCH_CHECK();
    return 0;
  };
  auto part_b = [&]() -> int {
    k_mg_copy<<<ntiles, RSB_BLOCK, 0, st>>>(a, z, z0);
    // down
    for (int l = 0; l + 1 < nlev; l++) {
      mg_level_dev &F = H.lv[l];
      // the residual ends in F.r0 (level 0 starts from the CG residual r)
      if (l == 0) {
        k_mg_jac0<<<ntiles, RSB_BLOCK, 0, st>>>(a, S, sig[0], r, F.sa, F.u);
        k_mg_res0<<<ntiles, RSB_BLOCK, 0, st>>>(a, S, r, F.r0, F.sa, sig[1], F.sb, F.u);
        k_mg_res0<<<ntiles, RSB_BLOCK, 0, st>>>(a, S, F.r0, F.r0, F.sb, sig[2], F.sa, F.u);
        k_mg_res0<<<ntiles, RSB_BLOCK, 0, st>>>(a, S, F.r0, F.r0, F.sa, 0.0, nullptr, F.u);
      } else {
        const mg_csr M{F.m, F.rseg, F.soff, F.rlen, F.sci, F.sw, F.diag, F.spk};
        const unsigned nb = nblocks(F.m, RSB_BLOCK);
        k_mg_jac_c<<<nb, RSB_BLOCK, 0, st>>>(a, M, sig[0], F.r0, F.sa, F.u);
        k_mg_res_c<<<nb, RSB_BLOCK, 0, st>>>(a, M, F.r0, F.sa, sig[1], F.sb, F.u);
        k_mg_res_c<<<nb, RSB_BLOCK, 0, st>>>(a, M, F.r0, F.sb, sig[2], F.sa, F.u);
        k_mg_res_c<<<nb, RSB_BLOCK, 0, st>>>(a, M, F.r0, F.sa, 0.0, nullptr, F.u);
      }
      mg_level_dev &C = H.lv[l + 1];
      k_mg_restrict<<<nblocks(C.m, RSB_BLOCK), RSB_BLOCK, 0, st>>>(
          a, C.m, C.rseg, mg_lvl{F.off, F.n}, mg_lvl{C.off, C.n}, factor, F.r0, C.r0);
    }
    // up: each coarse level's result goes to its r0
    for (int l = nlev - 2; l >= 1; l--) {
      mg_level_dev &F = H.lv[l], &C = H.lv[l + 1];
      k_mg_prolong_c<<<nblocks(F.m, RSB_BLOCK), RSB_BLOCK, 0, st>>>(
          a, F.m, F.rseg, mg_lvl{F.off, F.n}, mg_lvl{C.off, C.n}, factor, 1.5, C.r0, F.u, F.r0);
    }
    mg_level_dev &F0 = H.lv[0], &C1 = H.lv[1];
    k_mg_prolong0<<<ntiles, RSB_BLOCK, 0, st>>>(a, mg_lvl{C1.off, C1.n}, factor, 1.33333, C1.r0,
                                                 F0.u, z);
    k_mg_sum<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_PCG, z);
    fin(FIN_MEAN, GATE_PCG, 1);
    k_mg_zdots<<<ntiles, RSB_BLOCK, 0, st>>>(a, r, z, z0);
    fin(FIN_BETA, GATE_PCG, 2);
    k_mg_pnew<<<ntiles, RSB_BLOCK, 0, st>>>(a, mpad, z, p, Wb);
    k_mg_coef<<<nseg, RSB_BLOCK, 0, st>>>(a, coef);
    k_mg_pcorr<<<ntiles, RSB_BLOCK, 0, st>>>(a, mpad, p, Pb, coef);
    RSB_LAUNCH_CHECK();
    return 0;
  };
  {
    const char *ev = getenv("PARRSB_AMD_MG_GRAPH");
    if (!(ev && atoi(ev) == 0)) {
      RSB_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      own.capturing = true;
      int rb = part_b(), ra = part_a();
      own.capturing = false;
      hipError_t ec = hipStreamEndCapture(st, &own.graph);
      if (rb || ra || ec != hipSuccess) RSB_FAIL(2, "MG graph capture failed");
      RSB_TRY(hipGraphInstantiateWithFlags(&own.exec, own.graph, 0));
    }
  }
  hipGraphExec_t exec = own.exec;
  for (int pass = 0; pass < max_passes; pass++) {
    // projection (fiedler.c:94-187)
    k_mg_pass0<<<ntiles, RSB_BLOCK, 0, st>>>(a, b, x, r);
    fin(FIN_PASS0, GATE_INV, 2);
    k_mg_pass1<<<ntiles, RSB_BLOCK, 0, st>>>(a, r, z, p);
    fin(FIN_RZ1, GATE_PCG, 1);
    RSB_LAUNCH_CHECK();
    MG_TRY(part_a());
    int nact = 0;
    MG_TRY(count_active(nact));
    for (int it = 1; it < max_iter && nact > 0; it++) {
      if (exec) {
        RSB_TRY(hipGraphLaunch(exec, st));
      } else {
        MG_TRY(part_b());
        MG_TRY(part_a());
      }
      MG_TRY(count_active(nact));
    }
    // end of pass (fiedler.c:231-247): y orthogonalised, z = y / |y|, z orthogonalised
    k_mg_sum<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_INV, x);
    fin(FIN_MEAN, GATE_INV, 1);
    k_mg_center_dot<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_INV, x);
    fin(FIN_NORM, GATE_INV, 1);
    k_mg_scale_sum<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_INV, x, b);
    k_mg_zero<<<1, 1, 0, st>>>(active);
    fin(FIN_PASSEND, GATE_INV, 1);
    k_mg_center<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_INV, b);
    RSB_LAUNCH_CHECK();
    if (gram) {
      // z (in b) orthogonalised against the basis twice into r; r / |r|
      // joins the basis, M grows by its row and column; the Ritz vector
      // becomes the next start (r, p, w are free between passes)
      auto gfin = [&](int op) { k_mg_gfin<<<nseg, RSB_BLOCK, 0, st>>>(a, G, op); };
      k_mg_gdots<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_INV, mpad, G.Z, nullptr, nullptr, 0, b);
      gfin(GF_CGS0);
      k_mg_gsub<<<ntiles, RSB_BLOCK, 0, st>>>(a, G, mpad, b, r);
      k_mg_gdots<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_INV, mpad, G.Z, nullptr, nullptr, 0, r);
      gfin(GF_CGS1);
      k_mg_gsub<<<ntiles, RSB_BLOCK, 0, st>>>(a, G, mpad, r, r);
      gfin(GF_NORM);
      k_mg_gnew<<<ntiles, RSB_BLOCK, 0, st>>>(a, G, mpad, r, b, p);
      gfin(GF_CNEW);
      k_mg_apply<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_GRAM, S, p, w);
      k_mg_gdots<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_GRAM, mpad, G.GZ, w, G.GZ, 1, p);
      gfin(GF_MROW);
      k_mg_gdots<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_GRAM, mpad, G.Z, nullptr, nullptr, 0, w);
      gfin(GF_MCOL);
      const size_t lds = sizeof(double) * (2 * (size_t)G.cap * G.cap + G.cap);
      k_mg_gjac<<<nseg, RSB_WAVE, lds, st>>>(a, G);
      k_mg_gcomb<<<ntiles, RSB_BLOCK, 0, st>>>(a, G, mpad, b);
      fin(FIN_MEAN, GATE_GMIX, 1);
      k_mg_center<<<ntiles, RSB_BLOCK, 0, st>>>(a, GATE_GMIX, b);
      RSB_LAUNCH_CHECK();
    }
    int nrun = 0;
    MG_TRY(count_active(nrun));
    if (nrun == 0) break;
  }
  // result: the last y, normalised; Rayleigh quotient for the report
  k_mg_out<<<ntiles, RSB_BLOCK, 0, st>>>(a, x, g->d_fied);
  k_mg_rayleigh<<<ntiles, RSB_BLOCK, 0, st>>>(a, S, x);
  fin(FIN_RAYLEIGH, GATE_ELIG, 3);
  RSB_LAUNCH_CHECK();
  std::vector<int> hs((size_t)MI_NI * nseg);
  std::vector<double> hsc((size_t)MS_NSC * nseg);
  RSB_TRY(hipMemcpyAsync(hs.data(), stt, sizeof(int) * hs.size(), hipMemcpyDeviceToHost, st));
  RSB_TRY(hipMemcpyAsync(hsc.data(), sc, sizeof(double) * hsc.size(), hipMemcpyDeviceToHost, st));
  RSB_TRY(hipStreamSynchronize(st));
  for (int s = 0; s < nseg; s++) {
    if (!(g->h_p[s] > 1 && g->h_n[s] >= 2)) continue;
    if (out->niter) out->niter[s] = hs[(size_t)s * MI_NI + MI_ITERS];
    if (out->npass) out->npass[s] = hs[(size_t)s * MI_NI + MI_NPASS];
    if (out->lambda) out->lambda[s] = hsc[(size_t)s * MS_NSC + MS_LAMBDA];
    if (out->resid) out->resid[s] = hsc[(size_t)s * MS_NSC + MS_RES];
  }
  // buffers and graph: freed by `own`; diag0 stays with the kept hierarchy
  rsb_phase_end(g, PH_LANCZOS);
  return 0;
}

// Parity hook: level `level` of the hierarchy built by the last
// rsbgpu_fiedler_mg call, rows of segment `seg` only, in that segment's local
// numbering: rowptr[n+1], col[], w[], diag[n].  Level 0 is the level
// Laplacian itself (not exported here).  Returns the row count through *n
// (and nnz through *nnz) when the arrays are NULL.
extern "C" int rsbgpu_mg_level_csr(rsbgpu *g, int level, int seg, int *n, int64_t *nnz,
                                   int64_t *rowptr, int *col, double *w, double *diag) {
  RSB_TRY(hipSetDevice(g->device));
  mg_hier *hp = (mg_hier *)g->mg;
  if (!hp || level < 1 || level >= (int)hp->lv.size() || seg < 0 ||
      seg >= (int)hp->h_n[level].size())
    RSB_FAIL(3, "no such MG level");
  mg_hier &H = *hp;
  const mg_level_dev &L = H.lv[level];
  const int rn = H.h_n[level][seg], ro = H.h_off[level][seg];
  std::vector<int64_t> rp((size_t)rn + 1);
  if (rn > 0)
    RSB_TRY(hipMemcpy(rp.data(), L.rp + ro, sizeof(int64_t) * (rn + 1), hipMemcpyDeviceToHost));
  *n = rn;
  *nnz = rn > 0 ? rp[rn] - rp[0] : 0;
  if (!rowptr) return 0;
  for (int i = 0; i <= rn; i++) rowptr[i] = rp[i] - rp[0];
  if (*nnz > 0) {
    RSB_TRY(hipMemcpy(col, L.ci + rp[0], sizeof(int) * (size_t)*nnz, hipMemcpyDeviceToHost));
    RSB_TRY(hipMemcpy(w, L.w + rp[0], sizeof(double) * (size_t)*nnz, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < *nnz; k++) col[k] -= ro;
  }
  if (rn > 0)
    RSB_TRY(hipMemcpy(diag, L.diag + ro, sizeof(double) * rn, hipMemcpyDeviceToHost));
  return 0;
}

// Parity hook: the Ritz step's Jacobi eigen-solve on one symmetric n x n
// matrix (row-major, 1 <= n <= MG_NP - 1): v = unit eigenvector of the
// smallest eigenvalue *lam.
extern "C" int rsbgpu_test_sym_smallest(const double *A, int n, double *v, double *lam) {
  if (n < 1 || n > MG_NP - 1) return 3;
  double *dA = nullptr, *dv = nullptr, *dl = nullptr;
  hipError_t err = rsb_dev_malloc((void **)&dA, sizeof(double) * (size_t)n * n);
  if (err == hipSuccess) err = rsb_dev_malloc((void **)&dv, sizeof(double) * n);
  if (err == hipSuccess) err = rsb_dev_malloc((void **)&dl, sizeof(double));
  if (err == hipSuccess)
    err = hipMemcpy(dA, A, sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    k_gjac_test<<<1, RSB_WAVE, sizeof(double) * (2 * (size_t)n * n + n), 0>>>(n, dA, dv, dl);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipDeviceSynchronize();
  if (err == hipSuccess) err = hipMemcpy(v, dv, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(lam, dl, sizeof(double), hipMemcpyDeviceToHost);
  if (dA) (void)rsb_dev_free(dA);
  if (dv) (void)rsb_dev_free(dv);
  if (dl) (void)rsb_dev_free(dl);
  return err == hipSuccess ? 0 : 2;
}
