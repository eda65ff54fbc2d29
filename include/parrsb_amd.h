/*
 * parrsb_amd.h -- entry points of the MI355X build beyond the drop-in
 * parrsb_part_mesh (parRSB.h).  None of them has a counterpart in the
 * reference; they serve callers whose mesh already lives in GPU memory,
 * benchmarks and diagnostics.
 */
#ifndef PARRSB_AMD_H
#define PARRSB_AMD_H

#include <stdint.h>

#include "parRSB.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PARRSB_AMD_MAX_LEVELS 32

/* Report of the last partitioning call (per bisection level = "cut", as the
 * reference's metric stack records them, metrics.c:40-48 and 68-106). */
typedef struct {
  int nlevels, nparts, nel, precision;
  double seconds;             /* wall time from entry to return */
  double phase_ms[8];         /* setup, pre-partition, Laplacian, Lanczos,
                                 TQLI+Ritz, sort, components, balance */
  double lanczos_bytes;       /* algorithmic bytes of all CG-Lanczos iterations */
  double lanczos_ms;          /* time of the iteration loops */
  double lanczos_iters;       /* iterations launched (3 kernels each) */
  double spmv_bytes;          /* algorithmic bytes of the SpMV kernel */
  double spmv_ms;             /* time inside SpMV launches (profiling on: every
                                 11th launch is timed, the mean times all) */
  double spmv_launches;       /* SpMV launches while profiling */
  int64_t graph_nnz, graph_wsum; /* dual graph: off-diagonals, sum of weights */
  int nsegs[PARRSB_AMD_MAX_LEVELS];      /* segments bisected at the level */
  int niter_min[PARRSB_AMD_MAX_LEVELS];  /* RSB_FIEDLER_CALC_NITER */
  int niter_max[PARRSB_AMD_MAX_LEVELS];
  int npass_max[PARRSB_AMD_MAX_LEVELS];
  int ncomp_min[PARRSB_AMD_MAX_LEVELS];  /* RSB_COMPONENTS_NCOMP */
  int ncomp_max[PARRSB_AMD_MAX_LEVELS];
  int nbalanced[PARRSB_AMD_MAX_LEVELS];  /* segments where balancing moved elements */
  int nconverged_fail[PARRSB_AMD_MAX_LEVELS]; /* segments at max_iter*max_passes */
  double lambda_min[PARRSB_AMD_MAX_LEVELS], lambda_max[PARRSB_AMD_MAX_LEVELS];
  double resid_max[PARRSB_AMD_MAX_LEVELS]; /* max ||Lf - lambda f|| / lambda */
  double resid_med[PARRSB_AMD_MAX_LEVELS]; /* median over the level's segments */
  double niter_mean[PARRSB_AMD_MAX_LEVELS]; /* element-weighted mean iterations */
  double level_ms[PARRSB_AMD_MAX_LEVELS];  /* wall time of the level */
  int nfallback[PARRSB_AMD_MAX_LEVELS];    /* learned start: segments that fell back */
  int warm_start;                          /* 1 if the learned start was used */
  /* RSB_NEIGHBORS (rsb.c:342-343): parts adjacent to a rank's part after the
     cut, min / max over segments, mean over ranks */
  int nbrs_min[PARRSB_AMD_MAX_LEVELS], nbrs_max[PARRSB_AMD_MAX_LEVELS];
  double nbrs_mean[PARRSB_AMD_MAX_LEVELS];
  /* TOL_INIT, TOL_TGT, TOL_FNL (fiedler.c:450-451, 506): the CG residual at
     the start of the last pass, its target and the final residual, each the
     minimum over the level's segments that stopped at rsb_max_iter *
     rsb_max_passes (rsb.c:79-90); 0 where every segment converged */
  double tol_init[PARRSB_AMD_MAX_LEVELS], tol_target[PARRSB_AMD_MAX_LEVELS];
  double tol_final[PARRSB_AMD_MAX_LEVELS];
  /* allocation-free calls: the most bytes of the fixed device array held at
     once, and the highest offset touched (freed blocks are reused only when
     the free part of the array runs short) */
  double noalloc_bytes, noalloc_high;
} parrsb_amd_report;

/* Allocation-free partitioning (no counterpart in the reference; for callers
 * that cannot allocate device memory: code around HIP graph capture, tight
 * serving loops, harnesses that forbid it).  Same contract as
 * parrsb_amd_part_mesh_dev, but the call makes no hipMalloc / hipFree: every
 * device buffer lives in a fixed array of PARRSB_AMD_NOALLOC_BYTES per GPU
 * (allocated with the library's code object, held until the process ends)
 * and the call's stream and timing events are created once per GPU, on its
 * first allocation-free call.  Meshes of more than PARRSB_AMD_NOALLOC_MAX_NEL
 * elements (PARRSB_AMD_NOALLOC_MAX_NEL_MG with rsb_algo = 1, whose stored
 * search directions take the most memory per element), or more parts than
 * PARRSB_AMD_NOALLOC_MAX_NEL, are rejected with 3, and so are tagged mode
 * and rsb_mg_grammian = 1 with rsb_algo = 1 (its Gram matrices take
 * rsb_max_passes^2 doubles per segment); at
 * these limits a call holds at most about a third (Lanczos) or half (MG) of
 * the array.  The array keeps the last call's contents and is shared
 * by every allocation-free call on that GPU: calls that overlap on one GPU
 * (two host threads) race.  Host bookkeeping still uses the C heap. */
#define PARRSB_AMD_NOALLOC_BYTES (256ull << 20)
#define PARRSB_AMD_NOALLOC_MAX_NEL 65536
#define PARRSB_AMD_NOALLOC_MAX_NEL_MG 32768
int parrsb_amd_part_mesh_dev_noalloc(int *d_part, const long long *d_vtx, const double *d_xyz,
                                     int nel, int nv, parrsb_options *options);

/* Same contract as parrsb_part_mesh, with every array in GPU memory of the
 * selected device (PARRSB_AMD_DEVICE, default 0): d_part[nel] out,
 * d_vtx[nel*nv], d_xyz[nel*nv*ndim] or NULL.  No MPI: one caller, one GPU;
 * the number of parts is options->nparts (>= 1). */
int parrsb_amd_part_mesh_dev(int *d_part, const long long *d_vtx, const double *d_xyz,
                             int nel, int nv, parrsb_options *options);

/* parrsb_part_mesh that also returns seq[i], the order of element i within
 * part[i] after partitioning (the part/seq form of the reference README,
 * README.md:46-66): the element's position in its part's final order (for
 * RSB the Fiedler order of the last cut). */
int parrsb_part_mesh_seq(int *part, int *seq, const long long *vtx, const double *xyz,
                         const int *tag, int nel, int nv, parrsb_options *options,
                         MPI_Comm comm);

/* parrsb_conn_mesh with device arrays: d_vtx[nel*nv] out, d_coord in. */
int parrsb_amd_conn_mesh_dev(long long *d_vtx, const double *d_coord, int nel, int ndim,
                             double tol);

/* Copy the report of the last call made in this process. */
int parrsb_amd_last_report(parrsb_amd_report *rep);

/* Message of the last failing call ("" if none). */
const char *parrsb_amd_last_error(void);

/* Lanczos vector precision for later calls: 64 (reference, default) or 32
 * (FP32 vectors, compensated dot products, FP64 scalars and eigen-solve).
 * Overridden by PARRSB_AMD_PRECISION. */
int parrsb_amd_set_precision(int bits);

/* 1 = time every SpMV launch with events (roofline measurements). */
int parrsb_amd_set_profiling(int on);

/* Learned warm start for later RSB calls: `params` are the network's
 * parrsb_amd_warm_start_nparams() FP32 parameters (host memory, copied), or
 * NULL to switch it off.  The Fiedler solve of every segment of at most
 * `max_n` elements (0: every segment) then starts from the network's
 * prediction, which gets `passes` restarted passes of rsb_max_iter
 * iterations; a segment that misses rsb_tol within them is solved again from
 * the reference start.  Larger segments take the reference start.
 * PARRSB_AMD_WARM_START=<file> loads parameters from a file written by
 * parrsb_amd.lws_train (PARRSB_AMD_WARM_PASSES and PARRSB_AMD_WARM_MAX_N set
 * `passes` and `max_n`, default 1 and 1024). */
int parrsb_amd_warm_start_nparams(void);
int parrsb_amd_set_warm_start(const float *params, int passes, int max_n);

/* Quality of a partition (host arrays): out[0] weighted edge cut, out[1] cut
 * edges, out[2] max part size, out[3] min part size, out[4] components over
 * all parts, out[5] max components of a part, out[6] comm volume
 * sum_v k_v (k_v - 1), out[7] empty parts.  Computed on the GPU. */
int parrsb_amd_partition_quality(const int *part, const long long *vtx, int nel, int nv,
                                 int nparts, int64_t *out);
/* Same, device arrays. */
int parrsb_amd_partition_quality_dev(const int *d_part, const long long *d_vtx, int nel,
                                     int nv, int nparts, int64_t *out);

/* free() for the arrays parrsb_read_mesh allocates, for callers that cannot
 * reach the C library's free (bindings from other languages). */
void parrsb_amd_free(void *p);

/* .ma2 for one rank holding all parts: the rank field of element e is
 * part[e] (what a P-rank genmap run writes after parrsb_dist_mesh; call
 * parrsb_dist_mesh first to get the same record order).  No counterpart in
 * the reference. */
int parrsb_amd_dump_map_parts(const char *name, unsigned nelt, unsigned nv, const long long *vl,
                              const int *part);

#ifdef __cplusplus
}
#endif
#endif
