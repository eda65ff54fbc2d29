// Device state behind the opaque rsbgpu handle (internal to the HIP layer).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <map>
#include <unordered_map>
#include <vector>

#include "rsbgpu.h"
#include "common.hpp"

// ---------------------------------------------------------------------------
// Device memory of the library.  Every device allocation goes through these
// two functions: hipMalloc / hipFree normally; inside an allocation-free call
// (noalloc.hip) a bump allocator over a fixed device array, whose frees are
// no-ops until the call ends.
hipError_t rsb_dev_malloc(void **p, size_t bytes);
hipError_t rsb_dev_free(const void *p);
template <typename T>
inline hipError_t rsb_dev_malloc(T **p, size_t bytes) {
  return rsb_dev_malloc((void **)p, bytes);
}
// Inside an allocation-free call on `device`: the stream and the four timing
// events set up once for that device (returns false otherwise).
bool rsb_arena_stream(int device, hipStream_t *stream, hipEvent_t *ev);

#define RSB_TILE 1024   // Lanczos rows per block (4 rows per thread)
#define RSB_BLOCK 256
#define RSB_MAXC 64     // max candidate neighbours gathered per element

// Phases timed with events (rsbgpu_phase_times).
enum { PH_SETUP = 0, PH_PRE, PH_LAP, PH_LANCZOS, PH_TQLI, PH_SORT, PH_COMP, PH_BAL, PH_N };

// Per-segment Lanczos state (device arrays of length nseg).
struct lz_state {
  int *status;     // LZ_* below
  int *iter;       // iteration index within the pass
  int *k;          // iterations of the finished pass
  int *npass, *niter;
  double *sc;      // LZ_NSC scalars per segment
  double *T;       // 2*miter per segment: diag[miter], upper[miter]
  double *y;       // miter per segment: Ritz coefficients
  double *partial; // 6 per tile
  double *rn;      // miter + 1 per segment: 1 / |r_j| of the pass's residuals
  int *running;    // count of RUNNING segments
};
enum { LZ_IDLE = 0, LZ_SETUP = 1, LZ_RUNNING = 2, LZ_PASSDONE = 3, LZ_FINISHED = 4 };
// scalar slots
enum {
  SC_RTR = 0, SC_RTZ1, SC_RTZ2, SC_PAP, SC_PAPOLD, SC_BETA, SC_ALPHA, SC_MEANP,
  SC_RNI, SC_RNORM, SC_RTOL, SC_RNORM0, SC_MEANF, SC_LAMBDA, SC_NORMF, SC_TMP, SC_RES,
  SC_M0, SC_N0, SC_WARMIT, SC_FALLBACK, SC_NFOLD, SC_LAMPREV, SC_RHO1, SC_RHO2,
  LZ_NSC
};

struct rsbgpu {
  int device = 0;
  hipStream_t stream = nullptr;
  bool borrowed = false;  // stream and events belong to the allocation-free set-up
  char err[512] = {0};

  // mesh
  int nel = 0, nv = 0, ndim = 0;
  const long long *d_vtx = nullptr;
  const double *d_xyz = nullptr;
  long long *own_vtx = nullptr;
  double *own_xyz = nullptr;
  double *d_coord = nullptr;  // SoA x[nel], y[nel], z[nel]
  uint32_t *d_rank = nullptr; // SoA dense coordinate ranks per axis
  int rank_bits = 1;
  int has_coord = 0;

  // weighted dual graph of the whole mesh (CSR, columns ascending)
  int64_t *d_growptr = nullptr;
  uint32_t *d_gadj = nullptr; // col | weight << 28
  int64_t gnnz = 0, gwsum = 0;

  // level state
  int *d_order = nullptr, *d_order2 = nullptr; // element at compact position
  int *d_segof = nullptr;                      // segment of each element
  int nseg = 0;
  std::vector<int> h_cstart, h_n, h_p, h_r0, h_pstart;
  int *d_seg = nullptr;  // 5*nseg: cstart, n, p, r0, pstart
  int64_t mpad = 0;
  int *d_row_elem = nullptr, *d_elem_row = nullptr;
  // SELL-64 level Laplacian
  int64_t nslices = 0, lnnz_pad = 0, lnnz = 0;
  int64_t *d_slice_off = nullptr;
  int *d_slice_w = nullptr;
  uint32_t *d_ell = nullptr;
  int *d_rowlen = nullptr;
  // LDS-tiled form (tiles.hip): halo rows per tile and 16-bit local entries
  int *d_tnhalo = nullptr, *d_halo = nullptr, *d_w16 = nullptr;
  int64_t *d_thalo_off = nullptr, *d_off16 = nullptr;
  uint16_t *d_e16 = nullptr;
  int64_t nhalo_total = 0, e16_count = 0;
  // the same entries as 12-bit local indices ordered by weight class
  // (tiles.hip): per slice its block count and tail words, its slots per
  // class (n4 | n2 << 8 | n1 << 16, the same for its 64 rows), word offsets
  // and the packed words; e12_ok = 0 when a weight is not 1, 2 or 4 or a
  // slice needs more than 42 slots.  The FP32 solve
  // reads this form; the FP64 solve reads the 16-bit entries unless e12_f64
  int *d_w12 = nullptr, *d_e12_bad = nullptr;
  int64_t *d_off12 = nullptr;
  uint32_t *d_e12 = nullptr;
  int *d_c12 = nullptr;
  int e12_ok = 0, e12_f64 = 0;
  int nwide = 0;
  // vertex form of the tiles (tiles.hip, rsb_build_vtiles): per tile, the
  // distinct vertex ids of its rows, for each the nv window-local indices of
  // the slots at it (padded); per row its nv local vertex indices
  int *d_vt_n = nullptr;            // local vertices per tile (-1: none)
  int64_t *d_vt_loff = nullptr;     // per tile: start of its lists
  uint16_t *d_vt_list = nullptr, *d_vt_rowv = nullptr;
  int vt_ok = 0;  // the vertex form was built (tiles with vt_n < 0 use the entries)
  int64_t *d_seg_nnz = nullptr;
  std::vector<int64_t> h_seg_nnz;
  // tiles of the level (a tile never straddles a segment)
  int ntiles = 0;
  int *d_tile = nullptr;       // 3*ntiles: seg, row0, row1
  int *d_seg_tiles = nullptr;  // 2*nseg: first tile, tile count
  std::vector<int> h_tile, h_seg_tiles;

  // Lanczos vectors (padded rows)
  void *d_p = nullptr, *d_p2 = nullptr, *d_w = nullptr, *d_f = nullptr;
  void *d_w2 = nullptr;  // the other parity's w (FP32 folded update)
  void *d_basis = nullptr;
  size_t vec_bytes = 0, basis_bytes = 0;
  double *d_fied = nullptr;     // Fiedler value per compact position
  float *d_start = nullptr;     // optional start vector per padded row
  const float *lws_params = nullptr;  // learned warm start parameters (device)
  float lws_level = 0.f;              // bisection level fed to the network
  int lws_passes = 1;                 // passes granted to the learned start
  int lws_max_n = 0;                  // larger segments skip it (0: no limit)
  void *d_lws_ws = nullptr;
  size_t lws_ws_bytes = 0;
  lz_state lz{};
  void *mg = nullptr;  // MG hierarchy of the last rsbgpu_fiedler_mg (mg.hip)
  int lz_nseg_cap = 0, lz_tile_cap = 0, lz_miter_cap = 0;

  // sort scratch
  void *d_sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  uint64_t *d_k0 = nullptr, *d_k1 = nullptr;
  int *d_v1 = nullptr;
  size_t sort_cap = 0;

  // generic scratch
  int *d_iscratch = nullptr;
  size_t iscratch_n = 0;

  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double ph_ms[PH_N] = {0};
  double lz_bytes = 0, lz_ms = 0;      // iteration kernels: model bytes, loop time
  long long lz_launches = 0;           // iterations launched (3 kernels each)
  double spmv_bytes = 0, spmv_ms = 0;  // SpMV kernel alone: model bytes, kernel time
  long long spmv_launches = 0;         // launches timed
  long long spmv_total = 0;            // launches while profiling (timed or not)
  int prof_spmv = 0;                   // bracket every LZ_PROF_EVERY-th SpMV launch with events
  std::vector<hipEvent_t> spmv_ev;
  hipEvent_t lz_ev0 = nullptr, lz_ev1 = nullptr;
  std::unordered_map<const void *, size_t> cap;  // bytes held by each dev_alloc slot
  // per-call scratch buffers, one per call site (file, line), kept across
  // calls while large enough and freed with the handle (rsb_scratch)
  std::map<std::pair<const void *, int>, std::pair<void *, size_t>> scratch;
};

// ---------------------------------------------------------------------------
// error helpers
#define RSB_TRY(expr)                                                           \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      snprintf(g->err, sizeof(g->err), "%s:%d: %s: %s", __FILE__, __LINE__,     \
               #expr, hipGetErrorString(e_));                                   \
      return 2;                                                                 \
    }                                                                           \
  } while (0)

#define RSB_LAUNCH_CHECK() RSB_TRY(hipGetLastError())

#define RSB_FAIL(code, ...)                                                     \
  do {                                                                          \
    snprintf(g->err, sizeof(g->err), __VA_ARGS__);                             \
    return (code);                                                              \
  } while (0)

// grow-only device allocation
template <typename T>
inline hipError_t dev_ensure(T *&p, size_t &cap_elems, size_t need) {
  if (need <= cap_elems && p) return hipSuccess;
  if (p) {
    hipError_t e = rsb_dev_free(p);
    if (e != hipSuccess) return e;
    p = nullptr;
  }
  size_t n = need ? need : 1;
  hipError_t e = rsb_dev_malloc((void **)&p, n * sizeof(T));
  if (e == hipSuccess) cap_elems = n;
  return e;
}

template <typename T>
inline hipError_t dev_alloc(rsbgpu *g, T *&p, size_t n) {
  // a buffer slot of the handle keeps its allocation while it is large
  // enough: the level loop re-sizes the same slots every level, and freeing
  // and mapping the multi-GB level matrices again each time cost ~10 ms a
  // level at 16.7M elements (contents are never assumed to be fresh)
  const size_t bytes = (n ? n : 1) * sizeof(T);
  auto it = g->cap.find((const void *)&p);
  if (p && it != g->cap.end() && it->second >= bytes) return hipSuccess;
  if (p) {
    hipError_t e = rsb_dev_free(p);
    if (e != hipSuccess) return e;
    p = nullptr;
  }
  hipError_t e = rsb_dev_malloc((void **)&p, bytes);
  if (e == hipSuccess) g->cap[(const void *)&p] = bytes;
  return e;
}

// Scratch buffer of one call site: the level loop used to hipMalloc and
// hipFree ~25 buffers a level, ~75 us each (about a third of a 1M-element
// call).  The buffer of a site lives until the handle is closed and is
// reused while large enough; its contents are never assumed.  Work queued
// on the handle's stream is ordered with the next use of the same site.
inline hipError_t rsb_scratch(rsbgpu *g, const void *file, int line, size_t bytes, void **out) {
  auto &s = g->scratch[std::make_pair(file, line)];
  if (s.first && s.second >= bytes) {
    *out = s.first;
    return hipSuccess;
  }
  if (s.first) {
    hipError_t e = rsb_dev_free(s.first);
    s.first = nullptr, s.second = 0;
    if (e != hipSuccess) return e;
  }
  hipError_t e = rsb_dev_malloc(&s.first, bytes ? bytes : 1);
  if (e != hipSuccess) {
    s.first = nullptr;
    return e;
  }
  s.second = bytes;
  *out = s.first;
  return hipSuccess;
}
#define RSB_SCRATCH(ptr, bytes) \
  rsb_scratch(g, (const void *)__FILE__, __LINE__, (size_t)(bytes), (void **)&(ptr))

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int rsb_build_tiles(rsbgpu *g);   // tiles.hip
int rsb_build_vtiles(rsbgpu *g);  // tiles.hip
// limits of the vertex form: window rows (tile + halo, the LDS window of the
// neighbour-entry form) and local vertices
#define VT_MAXW 4096
#define VT_MAXV 2048
void rsb_mg_free(rsbgpu *g);     // mg.hip
int rsb_lws_predict(rsbgpu *g, const float *P, float level, float *d_rows);  // lws.hip

// sort helpers (sort.hip)
int rsb_sort_pairs_u64(rsbgpu *g, uint64_t *keys, int *vals, int64_t n, int begin_bit,
                       int end_bit);
int rsb_sort_pairs_u32(rsbgpu *g, uint32_t *keys, int *vals, int64_t n, int begin_bit,
                       int end_bit);
int rsb_exclusive_scan_i64(rsbgpu *g, const int64_t *in, int64_t *out, int64_t n);
int rsb_exclusive_scan_i32(rsbgpu *g, const int *in, int *out, int64_t n);

// phase timing
void rsb_phase_begin(rsbgpu *g);
void rsb_phase_end(rsbgpu *g, int phase);
