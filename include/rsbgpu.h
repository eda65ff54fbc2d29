/*
 * rsbgpu.h -- the thin C ABI between the C host driver and the HIP device
 * layer.  Plain pointers and sizes only; the device state lives behind the
 * opaque `rsbgpu` handle.  Every function returns 0 on success and a
 * non-zero code on failure (the message is in rsbgpu_error()).
 *
 * Data model (one handle = one partitioning call on one GPU):
 *   elements      e in [0, nel), centroids and vertex ids on the device
 *   order         the current element order (a permutation of [0, nel));
 *                 segment s is the contiguous range
 *                 [cstart[s], cstart[s] + n[s]) of it, spread over the
 *                 virtual ranks [r0[s], r0[s] + p[s])
 *   rows          the Lanczos vectors of a level: segment s owns the padded
 *                 row range [pstart[s], pstart[s] + ceil64(n[s]))
 */
#ifndef PARRSB_AMD_RSBGPU_H
#define PARRSB_AMD_RSBGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsbgpu rsbgpu;

/* Per-level results the host reads back (arrays of length nseg). */
typedef struct {
  int *niter;       /* Lanczos iterations (RSB_FIEDLER_CALC_NITER) */
  int *npass;       /* Lanczos passes */
  double *lambda;   /* Ritz value of the last pass */
  double *rnorm;    /* 3 per segment: initial, target, final CG residual */
  double *resid;    /* relative eigen-residual ||Lf - lambda f|| / lambda */
  int *fallback;    /* init_mode 2: 0 if the learned start met tol; 1 + the
                       iterations of the learned attempt if the segment was
                       solved again from the reference start; -1 if it was
                       above the size limit and took the reference start */
} rsbgpu_fiedler_out;

int rsbgpu_open(rsbgpu **g, int device);
void rsbgpu_close(rsbgpu *g);
const char *rsbgpu_error(const rsbgpu *g);
/* Milliseconds spent in each phase so far (events on the device stream):
 * t[0] setup (copies, centroids, ranks, graph), t[1] pre-partition,
 * t[2] Laplacian assembly, t[3] Lanczos, t[4] TQLI + Ritz, t[5] sort,
 * t[6] components, t[7] balance + split. */
int rsbgpu_phase_times(rsbgpu *g, double *t8);
/* Lanczos traffic accumulated over the handle's life:
 *   v6[0] algorithmic bytes of the CG-Lanczos iterations, v6[1] ms of the
 *   iteration loops, v6[2] iterations launched;
 *   v6[3] algorithmic bytes of the SpMV kernel alone, v6[4] ms inside SpMV
 *   launches (only with profiling on: the mean of the timed launches times
 *   v6[5]), v6[5] SpMV launches while profiling. */
int rsbgpu_lanczos_traffic(rsbgpu *g, double *v6);
/* 1 = bracket every 11th SpMV launch with events (for the roofline figure;
 * timing every launch costs ~5% of the call). */
int rsbgpu_set_profiling(rsbgpu *g, int spmv_events);
/* Parity hook: tridiagonal eigen-solve of one (d, e) on the device; y gets
 * the coefficients of the eigenvector of smallest |eigenvalue|. */
int rsbgpu_test_tqli(const double *d, const double *e, int n, double *y, double *lam);
/* HBM calibration: best GB/s of plain streaming kernels (16-byte vector
 * accesses, grids of 1024..16384 workgroups, plain and non-temporal) over
 * bytes-sized arrays on device `device`, reps timed launches each:
 * gbs[0] read, gbs[1] write, gbs[2] copy (read + written bytes). */
int rsbgpu_stream_test(int device, size_t bytes, int reps, double *gbs);
/* The FP32 path's eigen-solve (multisection + inverse iteration) on the same
 * input: eigenpair of smallest |eigenvalue|, vector signed y[0] >= 0. */
int rsbgpu_test_eigmin(const double *d, const double *e, int n, double *y, double *lam);

/* Load the mesh.  If on_device != 0 the pointers are device pointers that
 * stay valid for the lifetime of the handle; else they are copied.  Computes
 * centroids (if xyz != NULL), per-axis coordinate ranks and the weighted
 * dual graph.  Sets the order to the identity. */
int rsbgpu_set_mesh(rsbgpu *g, const long long *vtx, const double *xyz, int nel, int nv,
                    int on_device);
/* Graph statistics: number of directed off-diagonal entries and sum of
 * weights. */
int rsbgpu_graph_stats(rsbgpu *g, int64_t *nnz, int64_t *wsum);

/* Install the segment table of the current level (host arrays). */
int rsbgpu_set_segments(rsbgpu *g, int nseg, const int *cstart, const int *n, const int *p,
                        const int *r0);

/* Pre-partition every segment with p > 1: 1 = RCB (rcb.c), 0 = sort by
 * global id (rsb.c:284-287), 2 = RIB (rib.c, deterministic axis). */
int rsbgpu_prepartition(rsbgpu *g, int rsb_pre);

/* Assemble the shared-vertex Laplacian of every segment with p > 1 (cut
 * edges dropped) in padded row space. */
int rsbgpu_level_laplacian(rsbgpu *g);

/* Fiedler vectors of every segment with p > 1.  precision: 64 or 32.
 * init_mode: 0 = the reference's start vector from the pre-partition
 * position; 1 = a start vector set by rsbgpu_set_start (per element);
 * 2 = the learned warm start (rsbgpu_lws_set_params): the network's
 * prediction gets the passes set there (max_iter iterations each); a
 * segment that misses tol within them, or whose prediction is constant or
 * not finite, starts over from the reference start with the full
 * max_passes budget, and niter counts both attempts.
 * The result is held on the device per compact position. */
int rsbgpu_fiedler(rsbgpu *g, int precision, int max_iter, int max_passes, double tol,
                   int init_mode, rsbgpu_fiedler_out *out);
int rsbgpu_set_start(rsbgpu *g, const float *d_start_per_element);
/* Fiedler vectors by MG-preconditioned inverse iteration (rsb_algo = 1,
 * fiedler.c:94-321 and 186-312; multigrid.c:184-317), FP64, for every
 * segment with p > 1 and at least 2 elements: each projection is at most
 * max_iter (<= 60) preconditioned CG iterations to tol, at most max_passes
 * passes, stopping after a projection of one iteration.  factor (> 1) is the
 * MG coarsening factor.  grammian != 0 refines each next start vector by a
 * Rayleigh-Ritz step over the normalised iterates of the earlier passes
 * (fiedler.c:250-305; needs max_passes <= 64).  out->niter = summed CG
 * iterations, out->npass = passes, out->lambda / out->resid = Rayleigh
 * quotient and relative eigen-residual of the result.  Same result layout
 * as rsbgpu_fiedler. */
int rsbgpu_fiedler_mg(rsbgpu *g, int max_iter, int max_passes, double tol, int factor,
                      int grammian, rsbgpu_fiedler_out *out);
/* Parity hook: the Ritz step's eigen-solve (cyclic Jacobi) on one symmetric
 * n x n matrix A (row-major, 1 <= n <= 63): v = unit eigenvector of the
 * smallest eigenvalue *lam. */
int rsbgpu_test_sym_smallest(const double *A, int n, double *v, double *lam);
/* Parity hook: MG level `level` (>= 1) of segment `seg` from the last
 * rsbgpu_fiedler_mg call, in the segment's local numbering: rowptr[n+1],
 * col[nnz], w[nnz] (positive off-diagonal weights), diag[n].  With NULL
 * arrays only *n and *nnz are returned. */
int rsbgpu_mg_level_csr(rsbgpu *g, int level, int seg, int *n, int64_t *nnz, int64_t *rowptr,
                        int *col, double *w, double *diag);

/* Stable sort of every segment with p > 1 by its Fiedler values. */
int rsbgpu_fiedler_sort(rsbgpu *g);

/* Bisect every segment with p > 1 after the Fiedler sort.
 *   nlower[s]   in: size of the lower half after the sort;
 *               out: size after balancing
 *   sender[s]   -1 none, 0 lower half sends, 1 upper half sends
 *   scount[s]   number of interface elements to move
 *   find_comps  count components of both halves before balancing
 *   ncomp[2s], ncomp[2s+1]  out: components of lower / upper half
 *   moved[s]    out: 1 if elements were moved */
int rsbgpu_bisect(rsbgpu *g, int *nlower, const int *sender, const int *scount,
                  int find_comps, int *ncomp, int *moved);

/* part[e] = r0 of the segment holding e.  part_is_device selects whether
 * `part` is a device pointer. */
int rsbgpu_write_part(rsbgpu *g, int *part, int part_is_device);
/* seq[e] = position of e within its segment in the current order (after the
 * last cut: the order of the element within its part). */
int rsbgpu_write_seq(rsbgpu *g, int *seq, int seq_is_device);

/* Tagged partitioning (parrsb.c:379-452): reorder the element order so that
 * the elements of each distinct tag are contiguous, tags ascending, elements
 * of one tag in their previous order (stable).  *ngroups = number of distinct
 * tags; counts[0..ngroups) = elements per tag, filled only when
 * ngroups <= max_groups (the caller checks). */
int rsbgpu_group_by_tag(rsbgpu *g, const int *tag, int tag_is_device, int max_groups,
                        int *ngroups, int *counts);

/* Allocation-free calls (noalloc.hip): between begin and end, every device
 * allocation of this thread's handles is carved from the device's fixed
 * array and rsbgpu_open borrows the stream and events set up for the device
 * on its first such call.  begin returns 0, 2 (device set-up failed) or 3
 * (already inside such a call), with a message in msg; end drains the
 * borrowed stream and returns the most bytes the call held at once (*high:
 * the highest offset of the array it touched). */
int rsbgpu_noalloc_begin(int device, char *msg, int msglen);
long long rsbgpu_noalloc_end(long long *high);

/* Copy the current order / Fiedler values (per compact position) to the
 * host; used by the parity tests. */
int rsbgpu_get_order(rsbgpu *g, int *order_host);
int rsbgpu_get_fiedler(rsbgpu *g, double *fiedler_host);
/* Form of the current level's Laplacian apply: *wide_tiles = tiles whose
 * halo does not fit LDS (global 32-bit path), *vertex_tiles = tiles with the
 * vertex (gather-scatter) form, which FP32 Lanczos uses (0 unless
 * PARRSB_AMD_VTILES=1); for the parity tests. */
int rsbgpu_level_form(rsbgpu *g, int *wide_tiles, int *vertex_tiles);
/* 1 if the level's tiles also carry the 12-bit form (indices ordered by
 * weight class; every weight 1, 2 or 4), read by the tiled Lanczos apply
 * unless PARRSB_AMD_E12=0; else 0. */
int rsbgpu_level_e12(rsbgpu *g);
/* Overwrite the Fiedler value of every compact position (host array of nel
 * values): lets the parity tests drive the balance step with chosen values
 * (ties, either sender) through rsbgpu_bisect. */
int rsbgpu_set_fiedler(rsbgpu *g, const double *fiedler_host);
/* Copy the assembled level Laplacian in padded-row CSR form (host arrays,
 * row-major, columns ascending): rowptr[mpad+1], col[], wgt[].  Returns the
 * padded row count and nnz through the pointers when the arrays are NULL. */
int rsbgpu_get_level_csr(rsbgpu *g, int64_t *mpad, int64_t *nnz, int64_t *rowptr, int *col,
                         int *wgt, int *row_elem);
int rsbgpu_get_graph(rsbgpu *g, int64_t *rowptr, int *col, int *wgt);
int rsbgpu_get_centroids(rsbgpu *g, double *coord);

/* Quality of a partition held on the device (part[] device pointer):
 * out[0] weighted edge cut, out[1] cut edges, out[2] max part, out[3] min
 * part, out[4] components over all parts, out[5] max components in a part,
 * out[6] comm volume, out[7] empty parts. */
int rsbgpu_quality(rsbgpu *g, const int *d_part, int nparts, int64_t *out);

/* Learned warm start (a message-passing network over the level graph, see
 * lws.hip for the architecture).  No counterpart in the reference: it only
 * replaces the start vector of fiedler.c:570-585, and the reference start is
 * the fallback.  Parameters are LWS_NPARAM floats in device memory. */
int rsbgpu_lws_nparams(void);
/* Parameters and inputs of init_mode 2: the bisection level (0 = first
 * cut), the passes the learned start may use before the fallback, and the
 * largest segment (elements) it is tried on (0: all; larger segments take
 * the reference start with the full budget). */
int rsbgpu_lws_set_params(rsbgpu *g, const float *d_params, float level, int passes,
                          int max_n);
/* Network output for every row of the current level (after
 * rsbgpu_level_laplacian), d_rows[mpad]. */
int rsbgpu_lws_predict(rsbgpu *g, const float *d_params, float level, float *d_rows);
/* One training step on the current level, after a plain rsbgpu_fiedler and
 * before rsbgpu_fiedler_sort: the loss R(y) / lambda - 1 per segment (R the
 * Rayleigh quotient of the prediction y, lambda the segment's Fiedler value)
 * averaged over the level's segments, its gradient written to d_grad. */
int rsbgpu_lws_train_step(rsbgpu *g, const float *d_params, float *d_grad, float level,
                          double *loss);
/* Adam update of the parameters (device arrays of LWS_NPARAM floats). */
int rsbgpu_lws_adam(float *d_params, const float *d_grad, float *d_m, float *d_v, int step,
                    float lr);
/* Parity hook: forward + backward on one segment given on the host. */
int rsbgpu_lws_test(int M, const double *coord, float level, const int *rowptr, const int *col,
                    const int *wgt, double lam, const float *params, float *y, float *grad,
                    double *loss);

/* Number of other current segments that share a vertex with each segment
 * (statistics.c:17-113), nbrs[nseg] on the host. */
int rsbgpu_neighbors(rsbgpu *g, int *nbrs);

/* Connectivity (parrsb_conn_mesh): vtx_out[nel*nv] 1-based ids in the
 * symmetric corner order from coord[nel*nv*ndim] (preprocessor order), on
 * device `device`; host or device arrays (vtx_out and coord).  pinfo (host,
 * may be NULL when npinfo == 0) holds npinfo periodic face records
 * (element, face, partner element, partner face), 1-based, preprocessor face
 * numbers (con.c:41-47).  0, 1 (element/face check failed, or periodic faces
 * that do not match), 2 (device failure, message in msg), 3 (bad input). */
int rsbgpu_conn_mesh(int device, long long *vtx_out, const double *coord, int nel, int ndim,
                     double tol, const long long *pinfo, int npinfo, int on_device, char *msg,
                     int msg_len);

/* Small helpers for callers holding host data: copy n ints (or `bytes`
 * bytes) to a new device buffer (freed with rsbgpu_free). */
int rsbgpu_upload_ints(rsbgpu *g, const int *host, int n, int **dev_out);
int rsbgpu_upload(rsbgpu *g, const void *host, size_t bytes, void **dev_out);
int rsbgpu_download(rsbgpu *g, const void *dev, size_t bytes, void *host);
void rsbgpu_free(void *dev);

#ifdef __cplusplus
}
#endif
#endif
