/*
 * CPU oracle: a plain-C, FP64 restatement of the reference's recursive
 * spectral bisection path, used ONLY as the checker by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py.  The product never links,
 * loads or calls anything here.
 *
 * The reference assigns one part per MPI rank (rsb.c:244-260).  The oracle
 * emulates P "virtual ranks" inside one process: a segment is a contiguous
 * range of an element order plus the range of virtual ranks it is spread
 * over.  Every distributed step of the reference is restated on that
 * representation:
 *   - element distribution after a global sort       sort.c:48-70, 106-117
 *   - RCB order with virtual-rank split points        rcb.c:7-137
 *   - RIB order (deterministic axis)                  rib.c:6-111
 *   - Lanczos start vector from the global position   fiedler.c:563-585
 *   - CG-Lanczos, TQLI, Ritz restarts                 fiedler.c:13-30, 323-561
 *   - shared-vertex Laplacian (gather-scatter form)   laplacian.c:159-205
 *   - component counts of each half                   components.c:160-328
 *   - exact-count balance of the two halves           rsb.c:118-229
 *   - neighbour-part count per part                   statistics.c:17-113
 * Ties in every sort are broken by the current order (stable sorts).  The
 * reference breaks ties between ranks by crystal-router arrival order, which
 * no fixture pins; the stable rule is the convention both the oracle and the
 * device path follow.
 */
#ifndef PARRSB_ORACLE_H
#define PARRSB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int find_disconnected_comps;
  int rsb_pre; /* 0 none (sort by global id), 1 RCB, 2 RIB */
  int max_iter, max_passes;
  double tol;
  int nparts;
  int rsb_algo;  /* 0 Lanczos, 1 MG-preconditioned inverse iteration */
  int mg_factor; /* MG coarsening factor (rsb_mg_factor, > 1) */
  int mg_grammian; /* rsb_mg_grammian: Gram-matrix refinement of the iterates */
} oracle_opts;

/* Per-cut (level) statistics, as the reference's metric stack records them
 * (metrics.c:40-48): min/max over the segments that were cut at that level. */
#define ORACLE_MAX_LEVELS 32
typedef struct {
  int nlevels;
  int niter_min[ORACLE_MAX_LEVELS], niter_max[ORACLE_MAX_LEVELS];
  int ncomp_min[ORACLE_MAX_LEVELS], ncomp_max[ORACLE_MAX_LEVELS];
  int nbrs_max[ORACLE_MAX_LEVELS];
  int nsegs[ORACLE_MAX_LEVELS];
  int nbalanced[ORACLE_MAX_LEVELS]; /* segments where interface elements moved */
  double lambda_min[ORACLE_MAX_LEVELS], lambda_max[ORACLE_MAX_LEVELS];
  int nbrs_min[ORACLE_MAX_LEVELS];    /* neighbouring segments after the cut */
  double nbrs_mean[ORACLE_MAX_LEVELS]; /* mean over virtual ranks */
} oracle_report;

/* Centroids: FP64 mean of the nv corners, summed in corner order
 * (parrsb.c:139-145).  coord is nel x 3; coord[e*3+2] = 0 in 2D. */
void oracle_centroids(const double *xyz, int nel, int nv, double *coord);

/* Number of elements on virtual rank r when n elements are spread over p
 * ranks after a global sort: the first p - (n mod p) ranks get floor(n/p),
 * the others one more (sort.c:48-70). */
int64_t oracle_rank_count(int64_t n, int p, int r);
/* Elements on the lower ceil(p/2) ranks. */
int64_t oracle_lower_count(int64_t n, int p);
/* Expected size of the lower (bin 0) / upper (bin 1) half after balancing
 * (rsb.c:138-140). */
int64_t oracle_balance_expected(int64_t n, int p, int bin);

/* RCB order of a segment (rcb.c:116-137): `order` holds n element indices in
 * the current order and is permuted in place into the RCB order for p
 * virtual ranks.  coord is indexed by element (x3). */
void oracle_rcb_order(const double *coord, int ndim, int *order, int n, int p);

/* RIB order of a segment (rib.c:6-111) with the deterministic axis of
 * oracle_rib_axis; same split points as RCB. */
void oracle_rib_order(const double *coord, int ndim, int *order, int n, int p);
/* Mean and dominant covariance eigenvector of the centroids order[0..n). */
void oracle_rib_axis(const double *coord, int ndim, const int *order, int n, double *mean,
                     double *ev);
/* Dominant eigenvector of a symmetric n x n (n <= 3) matrix stored 3 x 3. */
void oracle_sym_dominant(const double *A, int n, double *ev);
/* Smallest eigenpair of a symmetric n x n matrix (n <= 64, A overwritten) by
 * cyclic Jacobi; v of unit length; returns the eigenvalue. */
double oracle_sym_smallest(double *A, int n, double *v);

/* Weighted dual graph of the whole mesh: w_ij = number of shared vertex ids
 * (mat.c:17-61, 203-249).  CSR with columns sorted ascending, no diagonal.
 * Arrays are malloc'ed; free with oracle_free. */
int64_t oracle_dual_graph(const long long *vtx, int nel, int nv, int64_t **rowptr,
                          int **col, int **wgt);
void oracle_free(void *p);

/* Threads of the OpenMP build (liboracle_mt.so) for later calls; n <= 0
 * leaves the count.  Returns the count in use (1 in the serial build). */
int oracle_threads(int n);
/* 1: oracle_part_mesh reports each finished cut on stderr. */
void oracle_set_verbose(int v);

/* Apply the shared-vertex Laplacian of the sub-mesh `elems[0..n)` (the
 * gather-scatter form of laplacian.c:159-205) to u, writing v. */
void oracle_laplacian_apply(const long long *vtx, int nv, const int *elems, int n,
                            const double *u, double *v);

/* Implicit QL on a symmetric tridiagonal (fiedler.c:328-432).  evec is n x n,
 * row k = eigenvector k (normalised), eval[n].  Returns 0, or 1 after 30
 * sweeps without deflation (then evec is left as the reference leaves it). */
int oracle_tqli(const double *diag, const double *upper, int n, double *evec,
                double *eval);

/* Fiedler vector of the sub-mesh elems[0..n) (given in pre-partition order)
 * by the reference's CG-Lanczos with restarts (fiedler.c:513-621).
 * fiedler[n] is normalised.  Returns the iteration count the reference
 * records (RSB_FIEDLER_CALC_NITER).  lambda_out (may be NULL) receives the
 * Ritz value of the last pass; rnorm_out (may be NULL) receives
 * {initial, target, final} residual norms of the last pass. */
int oracle_fiedler(const long long *vtx, int nv, const int *elems, int n,
                   int max_iter, int max_passes, double tol, double *fiedler,
                   double *lambda_out, double *rnorm_out);

/* Same Lanczos, from a caller-supplied start vector (length n, need not be
 * normalised or orthogonal to 1). */
int oracle_fiedler_from(const long long *vtx, int nv, const int *elems, int n,
                        int max_iter, int max_passes, double tol,
                        const double *initv, double *fiedler, double *lambda_out,
                        double *rnorm_out);

/* One CG-Lanczos pass (fiedler.c:434-511) from start vector f (length n):
 * writes diag[k], upper[k-1] and the normalised residual basis
 * basis[(k+1) x n]; returns k, the number of iterations taken. */
int oracle_lanczos_pass(const long long *vtx, int nv, const int *elems, int n, int niter,
                        double tol, const double *f, double *diag, double *upper,
                        double *basis);

/* Fiedler vector by MG-preconditioned inverse iteration (rsb_algo = 1,
 * fiedler.c:94-321 with rsb_mg_grammian = 0, multigrid.c:184-317,
 * mat.c:17-61, 704-764) of the sub-mesh elems[0..n), from the same start
 * vector as the Lanczos path.  fiedler[n] normalised; returns the summed PCG
 * iterations (RSB_FIEDLER_CALC_NITER); passes_out (may be NULL) receives the
 * number of inverse-iteration passes. */
int oracle_fiedler_mg(const long long *vtx, int nv, const int *elems, int n, int max_iter,
                      int max_passes, double tol, int factor, int grammian, double *fiedler,
                      int *passes_out);

/* Pieces of it, for the tests: the MG hierarchy's level sizes and matrices
 * (CSR of positive off-diagonal weights plus the diagonal, level by level,
 * concatenated), one V-cycle, and one projection (PCG solve of L x = b). */
int oracle_mg_levels(const long long *vtx, int nv, const int *elems, int n, int factor,
                     int *sizes, int64_t *nnz);
int oracle_mg_level_matrix(const long long *vtx, int nv, const int *elems, int n, int factor,
                           int level, int64_t *rowptr, int *col, double *w, double *diag);
void oracle_mg_vcycle(const long long *vtx, int nv, const int *elems, int n, int factor,
                      const double *rhs, double *out);
int oracle_mg_project(const long long *vtx, int nv, const int *elems, int n, int factor,
                      int miter, double tol, const double *b, double *x);

/* Connected components of the sub-mesh elems[0..n) under shared-vertex
 * adjacency (components.c:160-328 returns the same count). */
int oracle_components(const long long *vtx, int nv, const int *elems, int n);

/* One balance step (rsb.c:118-229) on a segment in ascending Fiedler order
 * (elements el, values fv, lower half el[0..nl)): the sending side hands
 * scnt of its interface elements nearest the cut to the other side if it has
 * that many; both arrays are permuted in place, each half keeping its
 * members' order.  Returns the new lower-half size. */
int oracle_balance_step(const long long *vtx, int nv, int *el, double *fv, int n, int nl,
                        int sender_lower, int64_t scnt);

/* Full partitioner: part[nel] in [0, nparts).  Returns 0 on success. */
int oracle_part_mesh(int *part, const long long *vtx, const double *xyz, int nel,
                     int nv, const oracle_opts *opts, oracle_report *rep);

/* Tagged partitioning (parrsb.c:379-561): each distinct tag partitioned on
 * its own into nparts / ntags parts, offset by its rank among the tags.
 * Returns 3 when nparts is not a multiple of the number of tags. */
int oracle_part_mesh_tagged(int *part, const long long *vtx, const double *xyz,
                            const int *tag, int nel, int nv, const oracle_opts *opts);

/* Connectivity (parrsb_conn_mesh, con.c:189-295): vtx[nel*nv] (symmetric
 * corner order, 1-based ids) from coord[nel][nv][ndim] (preprocessor corner
 * order) with relative tolerance tol; returns 1 if the element or face check
 * fails (con-check.c), 0 otherwise. */
int oracle_conn_mesh(long long *vtx, const double *coord, int nel, int ndim, double tol);

/* Same with periodic faces (con-periodic.c): pinfo[4*npinfo] = (element,
 * face, partner element, partner face), 1-based, preprocessor face order.
 * Returns 1 if a face pair does not match within 1e-3 of the faces' size,
 * 3 if a record names an element or face out of range. */
int oracle_conn_mesh_periodic(long long *vtx, const double *coord, int nel, int ndim,
                              double tol, const long long *pinfo, int npinfo);

/* Partition quality.  out[0] weighted edge cut (sum over undirected edges of
 * shared-vertex weight), out[1] unweighted cut edges, out[2] max part size,
 * out[3] min part size, out[4] total components over parts, out[5] max
 * components in one part, out[6] comm volume sum_v k_v (k_v - 1), out[7]
 * number of empty parts. */
void oracle_quality(const int *part, const long long *vtx, int nel, int nv,
                    int nparts, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif
