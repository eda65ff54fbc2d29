/*
 * CPU oracle (test infrastructure only; see oracle.h).  FP64 throughout,
 * compiled without floating-point contraction like the reference's
 * `-std=c99` build (Makefile:2), so that sums and products round exactly as
 * written.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef M_PI
#define M_PI 3.141592653589793
#endif

/* The same source built with -fopenmp -DORACLE_THREADS (liboracle_mt.so)
 * times the CPU baseline on all host cores: while a level has fewer
 * segments than threads, the vector loops and the Laplacian of each segment
 * are split over the threads (the reference's ranks sharing a segment); once
 * it has at least as many, whole segments run on their own threads.  Sums
 * then depend on the thread count, so the parity checks use the serial
 * build, where OMP() expands to nothing. */
#ifdef ORACLE_THREADS
#include <omp.h>
#define OMP(x) _Pragma(#x)
#else
#define OMP(x)
#endif
#define ORACLE_PAR_MIN 65536

void oracle_free(void *p) { free(p); }

/* 1: oracle_part_mesh prints one line per cut to stderr (the long CPU
 * baseline shows its progress) */
static int oracle_verbose = 0;
void oracle_set_verbose(int v) { oracle_verbose = v; }

int oracle_threads(int n) {
#ifdef ORACLE_THREADS
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
    abort();
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* Stable merge sort of (key, id) records by key; equal keys keep their
 * relative order.  -0.0 and +0.0 compare equal, as C's `<` has them. */
typedef struct {
  double key;
  int id;
} kv_t;

static void merge_sort_kv(kv_t *a, kv_t *tmp, int n) {
  if (n < 2) return;
  for (int w = 1; w < n; w *= 2) {
    for (int lo = 0; lo < n; lo += 2 * w) {
      int mid = lo + w < n ? lo + w : n;
      int hi = lo + 2 * w < n ? lo + 2 * w : n;
      int i = lo, j = mid, k = lo;
      while (i < mid && j < hi) {
        if (a[j].key < a[i].key) tmp[k++] = a[j++];
        else tmp[k++] = a[i++];
      }
      while (i < mid) tmp[k++] = a[i++];
      while (j < hi) tmp[k++] = a[j++];
    }
    memcpy(a, tmp, sizeof(kv_t) * (size_t)n);
  }
}

/* Stable sort of order[0..n) by key(order[i]) where keys are given per
 * position (pkey[i] belongs to order[i]).  pkey is permuted alongside. */
static void stable_sort_by_pos_key(int *order, double *pkey, int n) {
  kv_t *a = xmalloc(sizeof(kv_t) * (size_t)n), *t = xmalloc(sizeof(kv_t) * (size_t)n);
  for (int i = 0; i < n; i++) a[i].key = pkey[i], a[i].id = i;
  merge_sort_kv(a, t, n);
  int *o2 = xmalloc(sizeof(int) * (size_t)n);
  double *k2 = xmalloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) o2[i] = order[a[i].id], k2[i] = pkey[a[i].id];
  memcpy(order, o2, sizeof(int) * (size_t)n);
  memcpy(pkey, k2, sizeof(double) * (size_t)n);
  free(a), free(t), free(o2), free(k2);
}

/* ------------------------------------------------------------------------ */
/* Centroids, parrsb.c:139-145: start from 0.0, add the corners in order,
 * divide by nv. */
void oracle_centroids(const double *xyz, int nel, int nv, double *coord) {
  int ndim = (nv == 8) ? 3 : 2;
  for (int e = 0; e < nel; e++) {
    double c[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < nv; v++)
      for (int d = 0; d < ndim; d++) c[d] += xyz[(size_t)e * ndim * nv + v * ndim + d];
    for (int d = 0; d < ndim; d++) c[d] /= nv;
    coord[3 * (size_t)e + 0] = c[0];
    coord[3 * (size_t)e + 1] = c[1];
    coord[3 * (size_t)e + 2] = c[2];
  }
}

/* ------------------------------------------------------------------------ */
/* Virtual-rank arithmetic. */
int64_t oracle_rank_count(int64_t n, int p, int r) {
  int64_t q = n / p, rem = n - q * p;
  return (r < p - rem) ? q : q + 1; /* sort.c:58-66 */
}

int64_t oracle_lower_count(int64_t n, int p) {
  int64_t q = n / p, rem = n - q * p;
  int64_t hl = (p + 1) / 2; /* ranks [0, ceil(p/2)) form bin 0, rsb.c:241 */
  int64_t extra = hl - (p - rem);
  return q * hl + (extra > 0 ? extra : 0);
}

int64_t oracle_balance_expected(int64_t n, int p, int bin) {
  int64_t q = n / p, rem = n - q * p; /* rsb.c:138-140 */
  int64_t np_half = bin == 0 ? (p + 1) / 2 : p / 2;
  return q * np_half + rem / 2 + (rem % 2) * (1 - bin);
}

/* ------------------------------------------------------------------------ */
/* RCB, rcb.c:7-137.  For p > 1 the segment is sorted along the axis of
 * largest extent and split at the boundary of the lower ceil(p/2) virtual
 * ranks; at p == 1 the local recursion splits at floor(n/2).  Ties in the
 * extent go to the lower axis (strict `>`). */
static void rcb_rec(const double *coord, int ndim, int *order, double *key, int n,
                    int p) {
  if (n <= 1) return;
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = 0; i < n; i++) {
    const double *c = coord + 3 * (size_t)order[i];
    for (int d = 0; d < ndim; d++) {
      if (c[d] < mn[d]) mn[d] = c[d];
      if (c[d] > mx[d]) mx[d] = c[d];
    }
  }
  double len[3];
  for (int d = 0; d < ndim; d++) len[d] = mx[d] - mn[d];
  int axis = 0;
  if (len[1] > len[0]) axis = 1;
  if (ndim == 3 && len[2] > len[axis]) axis = 2;

  for (int i = 0; i < n; i++) key[i] = coord[3 * (size_t)order[i] + axis];
  stable_sort_by_pos_key(order, key, n);

  if (p > 1) {
    int nl = (int)oracle_lower_count(n, p);
    rcb_rec(coord, ndim, order, key, nl, (p + 1) / 2);
    rcb_rec(coord, ndim, order + nl, key + nl, n - nl, p / 2);
  } else {
    int mid = n / 2;
    rcb_rec(coord, ndim, order, key, mid, 1);
    rcb_rec(coord, ndim, order + mid, key + mid, n - mid, 1);
  }
}

void oracle_rcb_order(const double *coord, int ndim, int *order, int n, int p) {
  double *key = xmalloc(sizeof(double) * (size_t)n);
  rcb_rec(coord, ndim, order, key, n, p);
  free(key);
}

/* ------------------------------------------------------------------------ */
/* Smallest eigenpair of a symmetric n x n matrix A (row-major, n <= 64,
 * overwritten) by cyclic Jacobi rotations (Golub & Van Loan, Matrix
 * Computations, 4th ed., Alg. 8.5.1 and 8.5.3): sweeps over p < q in order,
 * a rotation skipped when |a_pq| <= eps^2 (|a_pp| + |a_qq|), until a sweep
 * rotates nothing (at most 64 sweeps).  v = the column of the smallest
 * diagonal entry (lowest index on ties), of unit length, sign left as the
 * rotations give it.  Returns that eigenvalue. */
double oracle_sym_smallest(double *A, int n, double *v) {
  double *V = xmalloc(sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) V[i * n + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 64; sweep++) {
    int rotated = 0;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p * n + q];
        if (apq == 0.0) continue;
        const double scale = fabs(A[p * n + p]) + fabs(A[q * n + q]);
        if (fabs(apq) <= DBL_EPSILON * DBL_EPSILON * scale) continue;
        const double tau = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                    : -1.0 / (-tau + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        A[p * n + q] = A[q * n + p] = 0.0;
        for (int k = 0; k < n; k++) {
          const double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
        rotated = 1;
      }
    if (!rotated) break;
  }
  int imin = 0;
  for (int i = 1; i < n; i++)
    if (A[i * n + i] < A[imin * n + imin]) imin = i;
  double nrm = 0.0;
  for (int k = 0; k < n; k++) nrm += V[k * n + imin] * V[k * n + imin];
  nrm = 1.0 / sqrt(nrm);
  for (int k = 0; k < n; k++) v[k] = V[k * n + imin] * nrm;
  const double lam = A[imin * n + imin];
  free(V);
  return lam;
}

/* ------------------------------------------------------------------------ */
/* Dominant eigenvector of a symmetric n x n matrix (n = 2, 3) by cyclic
 * Jacobi rotations (Golub & Van Loan, Matrix Computations, 4th ed.,
 * Alg. 8.5.1 sym.schur2 and 8.5.3 cyclic Jacobi).  The eigenvector of the
 * largest diagonal entry after convergence (lowest index on ties), signed so
 * that its largest-magnitude component (lowest index on ties) is positive. */
void oracle_sym_dominant(const double *Ain, int n, double *ev) {
  double A[3][3], V[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      A[i][j] = (i < n && j < n) ? Ain[i * 3 + j] : 0.0;
      V[i][j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 32; sweep++) {
    int rotated = 0;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p][q];
        if (apq == 0.0) continue;
        double scale = fabs(A[p][p]) + fabs(A[q][q]);
        if (fabs(apq) <= DBL_EPSILON * DBL_EPSILON * scale) continue;
        double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
        double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                              : -1.0 / (-tau + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        /* A <- J^T A J, V <- V J with J_pp = J_qq = c, J_pq = s, J_qp = -s */
        for (int k = 0; k < n; k++) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        A[p][q] = A[q][p] = 0.0;
        for (int k = 0; k < n; k++) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
        rotated = 1;
      }
    if (!rotated) break;
  }
  int im = 0;
  for (int i = 1; i < n; i++)
    if (A[i][i] > A[im][im]) im = i;
  int dm = 0;
  for (int d = 1; d < n; d++)
    if (fabs(V[d][im]) > fabs(V[dm][im])) dm = d;
  double sg = V[dm][im] < 0.0 ? -1.0 : 1.0;
  for (int d = 0; d < 3; d++) ev[d] = d < n ? sg * V[d][im] : 0.0;
}

/* RIB axis of the elements order[0..n) (rib.c:6-57): mean of the centroids
 * (summed in order, then divided), centred covariance (summed in order),
 * dominant eigenvector.  The reference takes <= 100 power steps from a
 * random start (fiedler.c:38-73), which fixes neither the sign nor, to better
 * than ~1e-6, the direction; this restatement takes the exact dominant
 * eigenvector with a fixed sign.  In 2D the covariance is 2 x 2 (the
 * reference passes the 3 x 3 array with N = 2, rib.c:47). */
void oracle_rib_axis(const double *coord, int ndim, const int *order, int n, double *mean,
                     double *ev) {
  double m[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < ndim; d++) m[d] += coord[3 * (size_t)order[i] + d];
  for (int d = 0; d < ndim; d++) m[d] /= n;
  double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; i++) {
    const double *c = coord + 3 * (size_t)order[i];
    double x[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ndim; d++) x[d] = c[d] - m[d];
    for (int a = 0; a < ndim; a++)
      for (int b = 0; b < ndim; b++) I[a * 3 + b] += x[a] * x[b];
  }
  oracle_sym_dominant(I, ndim, ev);
  for (int d = 0; d < 3; d++) mean[d] = m[d];
}

static void rib_rec(const double *coord, int ndim, int *order, double *key, int n, int p) {
  if (n <= 1) return;
  double m[3], ev[3];
  oracle_rib_axis(coord, ndim, order, n, m, ev);
  for (int i = 0; i < n; i++) {
    const double *c = coord + 3 * (size_t)order[i];
    double k = 0.0;
    for (int d = 0; d < ndim; d++) k += (c[d] - m[d]) * ev[d];
    key[i] = k;
  }
  stable_sort_by_pos_key(order, key, n);
  /* split as RCB: virtual-rank boundary for p > 1 (rib.c:73-105), half for
     the local recursion (rib.c:59-71) */
  if (p > 1) {
    int nl = (int)oracle_lower_count(n, p);
    rib_rec(coord, ndim, order, key, nl, (p + 1) / 2);
    rib_rec(coord, ndim, order + nl, key + nl, n - nl, p / 2);
  } else {
    int mid = n / 2;
    rib_rec(coord, ndim, order, key, mid, 1);
    rib_rec(coord, ndim, order + mid, key + mid, n - mid, 1);
  }
}

void oracle_rib_order(const double *coord, int ndim, int *order, int n, int p) {
  double *key = xmalloc(sizeof(double) * (size_t)n);
  rib_rec(coord, ndim, order, key, n, p);
  free(key);
}

/* ------------------------------------------------------------------------ */
/* Slot grouping: slots (e, j) of a sub-mesh grouped by vertex id.  Vertex
 * id 0 is "no connection" in gather-scatter: such a slot forms its own
 * group. */
typedef struct {
  long long vid;
  int slot;
} vs_t;

static int cmp_vs(const void *a, const void *b) {
  const vs_t *x = a, *y = b;
  if (x->vid != y->vid) return x->vid < y->vid ? -1 : 1;
  return x->slot - y->slot;
}

/* group[slot] = group id; returns number of groups; order of slots within
 * a group is ascending slot index. gstart/gslots describe groups (CSR). */
static int group_slots(const long long *vtx, int nv, const int *elems, int n, int *group,
                       int **gstart_out, int **gslots_out) {
  int ns = n * nv;
  vs_t *a = xmalloc(sizeof(vs_t) * (size_t)ns);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < nv; j++) {
      a[i * nv + j].vid = vtx[(size_t)elems[i] * nv + j];
      a[i * nv + j].slot = i * nv + j;
    }
  qsort(a, (size_t)ns, sizeof(vs_t), cmp_vs);
  int *gstart = xmalloc(sizeof(int) * (size_t)(ns + 1)), *gslots = xmalloc(sizeof(int) * (size_t)ns);
  int ng = 0;
  for (int s = 0; s < ns; s++) {
    if (s == 0 || a[s].vid != a[s - 1].vid || a[s].vid == 0) gstart[ng++] = s;
    group[a[s].slot] = ng - 1;
    gslots[s] = a[s].slot;
  }
  gstart[ng] = ns;
  free(a);
  if (gstart_out) *gstart_out = gstart; else free(gstart);
  if (gslots_out) *gslots_out = gslots; else free(gslots);
  return ng;
}

/* ------------------------------------------------------------------------ */
/* Gather-scatter Laplacian of a sub-mesh, laplacian.c:159-205:
 *   mult(v)  = number of slots of the sub-mesh with vertex id v
 *   diag_e   = sum over the slots of e of mult(v)
 *   (Lu)_e   = diag_e u_e - sum over the slots of e of (sum of u over the
 *              slots sharing that vertex id, own slot included). */
typedef struct {
  int n, nv, ng;
  int *group;   /* per slot */
  int *gstart, *gslots;
  int *gel;     /* element of each slot, in group order (gslots / nv) */
  double *diag; /* per element */
  double *gsum; /* per group scratch */
} gs_lap;

static void gs_lap_init(gs_lap *L, const long long *vtx, int nv, const int *elems, int n) {
  L->n = n, L->nv = nv;
  L->group = xmalloc(sizeof(int) * (size_t)n * nv);
  L->ng = group_slots(vtx, nv, elems, n, L->group, &L->gstart, &L->gslots);
  L->gel = xmalloc(sizeof(int) * (size_t)n * nv);
  for (int k = 0; k < n * nv; k++) L->gel[k] = L->gslots[k] / nv;
  L->diag = xmalloc(sizeof(double) * (size_t)n);
  L->gsum = xmalloc(sizeof(double) * (size_t)L->ng);
  for (int i = 0; i < n; i++) {
    double d = 0.0;
    for (int j = 0; j < nv; j++) {
      int g = L->group[i * nv + j];
      d += (double)(L->gstart[g + 1] - L->gstart[g]);
    }
    L->diag[i] = d;
  }
}

static void gs_lap_apply(const gs_lap *L, const double *u, double *v) {
  const int nv = L->nv;
  OMP(omp parallel for if (L->n > ORACLE_PAR_MIN))
  for (int g = 0; g < L->ng; g++) {
    double s = 0.0;
    for (int k = L->gstart[g]; k < L->gstart[g + 1]; k++) s += u[L->gel[k]];
    L->gsum[g] = s;
  }
  OMP(omp parallel for if (L->n > ORACLE_PAR_MIN))
  for (int i = 0; i < L->n; i++) {
    double r = L->diag[i] * u[i];
    for (int j = 0; j < nv; j++) r -= L->gsum[L->group[i * nv + j]];
    v[i] = r;
  }
}

static void gs_lap_free(gs_lap *L) {
  free(L->group), free(L->gstart), free(L->gslots), free(L->gel), free(L->diag), free(L->gsum);
}

void oracle_laplacian_apply(const long long *vtx, int nv, const int *elems, int n,
                            const double *u, double *v) {
  gs_lap L;
  gs_lap_init(&L, vtx, nv, elems, n);
  gs_lap_apply(&L, u, v);
  gs_lap_free(&L);
}

/* ------------------------------------------------------------------------ */
/* Dual graph of the whole mesh (mat.c:17-61, 203-249 semantics): for each
 * element, the elements sharing at least one vertex id, with weight = the
 * number of shared ids. */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int64_t oracle_dual_graph(const long long *vtx, int nel, int nv, int64_t **rowptr_out,
                          int **col_out, int **wgt_out) {
  int *elems = xmalloc(sizeof(int) * (size_t)nel);
  for (int i = 0; i < nel; i++) elems[i] = i;
  int *group = xmalloc(sizeof(int) * (size_t)nel * nv), *gstart, *gslots;
  group_slots(vtx, nv, elems, nel, group, &gstart, &gslots);
  int64_t *rowptr = xmalloc(sizeof(int64_t) * ((size_t)nel + 1));
  /* Two passes: count distinct neighbours, then fill. */
  int cap = 64;
  int *cand = xmalloc(sizeof(int) * (size_t)cap);
  int64_t nnz = 0;
  for (int pass = 0; pass < 2; pass++) {
    int *col = NULL, *wgt = NULL;
    if (pass == 1) {
      col = xmalloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
      wgt = xmalloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
    }
    int64_t pos = 0;
    for (int e = 0; e < nel; e++) {
      int nc = 0;
      for (int j = 0; j < nv; j++) {
        int g = group[(size_t)e * nv + j];
        for (int k = gstart[g]; k < gstart[g + 1]; k++) {
          int e2 = gslots[k] / nv;
          if (e2 == e) continue;
          if (nc == cap) cap *= 2, cand = realloc(cand, sizeof(int) * (size_t)cap);
          cand[nc++] = e2;
        }
      }
      qsort(cand, (size_t)nc, sizeof(int), cmp_int);
      if (pass == 0) rowptr[e] = pos;
      for (int k = 0; k < nc;) {
        int k2 = k;
        while (k2 < nc && cand[k2] == cand[k]) k2++;
        if (pass == 1) col[pos] = cand[k], wgt[pos] = k2 - k;
        pos++;
        k = k2;
      }
    }
    if (pass == 0) {
      nnz = pos;
      rowptr[nel] = nnz;
    } else {
      *col_out = col;
      *wgt_out = wgt;
    }
  }
  free(cand), free(group), free(gstart), free(gslots), free(elems);
  *rowptr_out = rowptr;
  return nnz;
}

/* ------------------------------------------------------------------------ */
/* Implicit QL with shifts (tridiagonal eigen-solve), the steps of
 * fiedler.c:328-432: deflation test |e_m| / (|d_m| + |d_m+1|) < 1e-12, at
 * most 30 sweeps per eigenvalue, eigenvectors accumulated by plane
 * rotations on the columns, then transposed to rows and normalised. */
static double sign_of(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

int oracle_tqli(const double *diag, const double *upper, int n, double *z, double *eval) {
  if (n == 0) return 0;
  double *d = xmalloc(sizeof(double) * 2 * (size_t)n), *e = d + n;
  for (int i = 0; i < n; i++) d[i] = diag[i];
  for (int i = 0; i < n - 1; i++) e[i] = upper[i];
  e[n - 1] = 0.0;
  for (int i = 0; i < n * n; i++) z[i] = 0.0;
  for (int i = 0; i < n; i++) z[i * n + i] = 1.0;

  for (int l = 0; l < n; l++) {
    int sweeps = 0, m;
    do {
      for (m = l; m < n - 1; m++) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) / dd < 1e-12) break;
      }
      if (m == l) break;
      if (sweeps++ == 30) {
        for (int i = 0; i < n; i++) eval[i] = d[i];
        free(d);
        return 1;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = sqrt(g * g + 1.0);
      g = d[m] - d[l] + e[l] / (g + sign_of(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      for (i = m - 1; i >= l; i--) {
        double f = s * e[i], b = c * e[i];
        if (fabs(f) >= fabs(g)) {
          c = g / f;
          r = sqrt(c * c + 1.0);
          e[i + 1] = f * r;
          s = 1.0 / r;
          c = c * s;
        } else {
          s = f / g;
          r = sqrt(s * s + 1.0);
          e[i + 1] = g * r;
          c = 1.0 / r;
          s = s * c;
        }
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < n; k++) {
          double zk1 = z[k * n + i + 1], zk = z[k * n + i];
          z[k * n + i + 1] = s * zk + c * zk1;
          z[k * n + i] = c * zk - s * zk1;
        }
      }
      /* The loop always ends with i == l - 1, so the reference's early
       * `continue` (r < tol && i >= l) can never fire. */
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    } while (1);
  }
  /* columns -> rows */
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) {
      double t = z[i * n + j];
      z[i * n + j] = z[j * n + i];
      z[j * n + i] = t;
    }
  for (int k = 0; k < n; k++) {
    double nrm = 0.0;
    for (int i = 0; i < n; i++) nrm += z[k * n + i] * z[k * n + i];
    if (nrm > 0.0) nrm = sqrt(fabs(nrm));
    double scale = 1.0 / nrm;
    for (int i = 0; i < n; i++) z[k * n + i] *= scale;
  }
  for (int i = 0; i < n; i++) eval[i] = d[i];
  free(d);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CG-Lanczos, fiedler.c:13-30 (dot, ortho) and 434-561. */
static double dotp(const double *y, const double *x, int n) {
  double s = 0.0;
  OMP(omp parallel for reduction(+ : s) if (n > ORACLE_PAR_MIN))
  for (int i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

static void ortho(double *q, int n) {
  double s = 0.0;
  OMP(omp parallel for reduction(+ : s) if (n > ORACLE_PAR_MIN))
  for (int i = 0; i < n; i++) s += q[i];
  s /= n;
  OMP(omp parallel for if (n > ORACLE_PAR_MIN))
  for (int i = 0; i < n; i++) q[i] -= s;
}

static int lanczos_pass(double *diag, double *upper, double *rr, int n, int niter,
                        double tol, const double *f, const gs_lap *L, double *rn3) {
  double *r = xmalloc(sizeof(double) * 3 * (size_t)n), *p = r + n, *w = p + n;
  OMP(omp parallel for if (n > ORACLE_PAR_MIN))
  for (int i = 0; i < n; i++) r[i] = f[i], p[i] = 0.0;
  ortho(r, n);
  double rtz1 = 1.0, pap = 0.0, alpha, beta, rtz2, pap_old;
  double rtr = dotp(r, r, n);
  double rnorm = sqrt(rtr), rtol = rnorm * tol;
  rn3[0] = rnorm, rn3[1] = rtol;
  double rni = 1.0 / rnorm;
  OMP(omp parallel for if (n > ORACLE_PAR_MIN))
  for (int i = 0; i < n; i++) rr[i] = r[i] * rni;
  int iter;
  for (iter = 0; iter < niter; iter++) {
    rtz2 = rtz1, rtz1 = rtr;
    beta = rtz1 / rtz2;
    if (iter == 0) beta = 0.0;
    OMP(omp parallel for if (n > ORACLE_PAR_MIN))
    for (int i = 0; i < n; i++) p[i] = beta * p[i] + r[i];
    ortho(p, n);
    gs_lap_apply(L, p, w);
    pap_old = pap, pap = dotp(w, p, n);
    alpha = rtz1 / pap;
    OMP(omp parallel for if (n > ORACLE_PAR_MIN))
    for (int i = 0; i < n; i++) r[i] = r[i] - alpha * w[i];
    rtr = dotp(r, r, n);
    rnorm = sqrt(rtr), rni = 1.0 / rnorm;
    OMP(omp parallel for if (n > ORACLE_PAR_MIN))
    for (int i = 0; i < n; i++) rr[(size_t)(iter + 1) * n + i] = r[i] * rni;
    if (iter == 0) {
      diag[iter] = pap / rtz1;
    } else {
      diag[iter] = (beta * beta * pap_old + pap) / rtz1;
      upper[iter - 1] = -beta * pap_old / sqrt(rtz2 * rtz1);
    }
    if (rnorm < rtol) {
      iter++;
      break;
    }
  }
  rn3[2] = rnorm;
  free(r);
  return iter;
}

static int lanczos_run(double *fied, const gs_lap *L, int n, double *initv, int miter,
                       int mpass, double tol, double *lambda_out, double *rn3) {
  if (n < miter) miter = n;
  double *alpha = xmalloc(sizeof(double) * (size_t)(2 * miter)), *beta = alpha + miter;
  double *rr = xmalloc(sizeof(double) * (size_t)(miter + 1) * n);
  double *ev = xmalloc(sizeof(double) * (size_t)miter * miter);
  double *el = xmalloc(sizeof(double) * (size_t)miter);
  int iter = miter, ipass;
  double lam = 0.0;
  for (ipass = 0; iter == miter && ipass < mpass; ipass++) {
    iter = lanczos_pass(alpha, beta, rr, n, miter, tol, initv, L, rn3);
    oracle_tqli(alpha, beta, iter, ev, el);
    double emin = fabs(el[0]);
    int imin = 0;
    for (int i = 1; i < iter; i++)
      if (fabs(el[i]) < emin) emin = fabs(el[i]), imin = i;
    lam = el[imin];
    OMP(omp parallel for if (n > ORACLE_PAR_MIN))
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = 0; j < iter; j++) s += rr[(size_t)j * n + i] * ev[imin * iter + j];
      fied[i] = s;
    }
    ortho(fied, n);
    OMP(omp parallel for if (n > ORACLE_PAR_MIN))
    for (int i = 0; i < n; i++) initv[i] = fied[i];
  }
  free(alpha), free(rr), free(ev), free(el);
  if (lambda_out) *lambda_out = lam;
  return (ipass - 1) * miter + iter;
}

static int fiedler_common(const long long *vtx, int nv, const int *elems, int n,
                          int max_iter, int max_passes, double tol, double *initv,
                          double *fiedler, double *lambda_out, double *rnorm_out) {
  double rn3[3] = {0, 0, 0};
  ortho(initv, n);
  double rtr = dotp(initv, initv, n), rni = 1.0 / sqrt(rtr);
  for (int i = 0; i < n; i++) initv[i] *= rni;
  gs_lap L;
  gs_lap_init(&L, vtx, nv, elems, n);
  int it = lanczos_run(fiedler, &L, n, initv, max_iter, max_passes, tol, lambda_out, rn3);
  gs_lap_free(&L);
  double nrm = 0.0;
  for (int i = 0; i < n; i++) nrm += fiedler[i] * fiedler[i];
  double ni = 1.0 / sqrt(nrm);
  for (int i = 0; i < n; i++) fiedler[i] *= ni;
  if (rnorm_out) rnorm_out[0] = rn3[0], rnorm_out[1] = rn3[1], rnorm_out[2] = rn3[2];
  return it;
}

int oracle_fiedler(const long long *vtx, int nv, const int *elems, int n, int max_iter,
                   int max_passes, double tol, double *fiedler, double *lambda_out,
                   double *rnorm_out) {
  /* Start vector, fiedler.c:572-578: global position + 1, plus 1000 n on
   * the first half. */
  double *initv = xmalloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) {
    initv[i] = i + 1.0;
    if (i < n / 2) initv[i] += (double)(1000 * (long long)n);
  }
  int it = fiedler_common(vtx, nv, elems, n, max_iter, max_passes, tol, initv, fiedler,
                          lambda_out, rnorm_out);
  free(initv);
  return it;
}

int oracle_fiedler_from(const long long *vtx, int nv, const int *elems, int n,
                        int max_iter, int max_passes, double tol, const double *init,
                        double *fiedler, double *lambda_out, double *rnorm_out) {
  double *initv = xmalloc(sizeof(double) * (size_t)n);
  memcpy(initv, init, sizeof(double) * (size_t)n);
  int it = fiedler_common(vtx, nv, elems, n, max_iter, max_passes, tol, initv, fiedler,
                          lambda_out, rnorm_out);
  free(initv);
  return it;
}

int oracle_lanczos_pass(const long long *vtx, int nv, const int *elems, int n, int niter,
                        double tol, const double *f, double *diag, double *upper,
                        double *basis) {
  gs_lap L;
  gs_lap_init(&L, vtx, nv, elems, n);
  double rn3[3];
  int k = lanczos_pass(diag, upper, basis, n, niter, tol, f, &L, rn3);
  gs_lap_free(&L);
  return k;
}

/* ------------------------------------------------------------------------ */
/* MG-preconditioned inverse iteration, rsb_algo = 1.
 *
 * Matrices (mat.c:17-61, 704-764): CSR of the off-diagonal weights, stored
 * positive (the reference stores them negated), and the diagonal, which is
 * always the row sum of those weights.  Level 0 is the segment's
 * shared-vertex Laplacian.  Level l + 1 aggregates `factor` consecutive rows
 * of level l (row i -> i / factor, multigrid.c:147-160): entries between two
 * aggregates add up, entries inside one aggregate vanish (the diagonal is
 * reset to the row sum of what remains).  Levels continue until one row is
 * left (multigrid.c:221-235). */
typedef struct {
  int n;
  int64_t *rp;
  int *ci;
  double *w, *diag;
} mgm_t;

typedef struct {
  int nl, factor;
  mgm_t *M;
} mgh_t;

typedef struct {
  int r, c;
  double w;
} rcw_t;

static int cmp_rcw(const void *a, const void *b) {
  const rcw_t *x = a, *y = b;
  if (x->r != y->r) return x->r - y->r;
  return x->c - y->c;
}

/* CSR from (row, col, weight) triples: duplicates add, row == col dropped. */
static void mgm_from_triples(mgm_t *M, int n, rcw_t *t, int64_t nt) {
  qsort(t, (size_t)nt, sizeof(rcw_t), cmp_rcw);
  M->n = n;
  M->rp = xmalloc(sizeof(int64_t) * (size_t)(n + 1));
  M->ci = xmalloc(sizeof(int) * (size_t)(nt ? nt : 1));
  M->w = xmalloc(sizeof(double) * (size_t)(nt ? nt : 1));
  M->diag = xmalloc(sizeof(double) * (size_t)(n ? n : 1));
  int64_t k = 0, i = 0;
  for (int r = 0; r < n; r++) {
    M->rp[r] = k;
    double d = 0.0;
    while (i < nt && t[i].r == r) {
      if (t[i].c == r) {
        i++;
        continue;
      }
      if (k > M->rp[r] && M->ci[k - 1] == t[i].c) {
        M->w[k - 1] += t[i].w;
      } else {
        M->ci[k] = t[i].c, M->w[k] = t[i].w, k++;
      }
      i++;
    }
    for (int64_t j = M->rp[r]; j < k; j++) d += M->w[j];
    M->diag[r] = d;
  }
  M->rp[n] = k;
}

static void mgm_free(mgm_t *M) { free(M->rp), free(M->ci), free(M->w), free(M->diag); }

/* y = M x = diag x - A x (mat.c:829-846) */
static void mgm_apply(const mgm_t *M, const double *x, double *y) {
  for (int i = 0; i < M->n; i++) {
    double s = M->diag[i] * x[i];
    for (int64_t j = M->rp[i]; j < M->rp[i + 1]; j++) s -= M->w[j] * x[M->ci[j]];
    y[i] = s;
  }
}

static void mg_setup(mgh_t *h, const gs_lap *L, int factor) {
  /* level 0 from the vertex groups: one unit of weight per shared id */
  int64_t nt = 0, cap = 1024;
  rcw_t *t = xmalloc(sizeof(rcw_t) * (size_t)cap);
  for (int g = 0; g < L->ng; g++)
    for (int a = L->gstart[g]; a < L->gstart[g + 1]; a++)
      for (int b = L->gstart[g]; b < L->gstart[g + 1]; b++) {
        const int ea = L->gslots[a] / L->nv, eb = L->gslots[b] / L->nv;
        if (ea == eb) continue;
        if (nt == cap) cap *= 2, t = realloc(t, sizeof(rcw_t) * (size_t)cap);
        t[nt].r = ea, t[nt].c = eb, t[nt].w = 1.0, nt++;
      }
  h->factor = factor;
  h->nl = 1;
  h->M = xmalloc(sizeof(mgm_t));
  mgm_from_triples(&h->M[0], L->n, t, nt);
  while (h->M[h->nl - 1].n > 1) {
    const mgm_t *F = &h->M[h->nl - 1];
    const int nc = (F->n + factor - 1) / factor;
    nt = 0;
    for (int i = 0; i < F->n; i++)
      for (int64_t j = F->rp[i]; j < F->rp[i + 1]; j++) {
        if (nt == cap) cap *= 2, t = realloc(t, sizeof(rcw_t) * (size_t)cap);
        t[nt].r = i / factor, t[nt].c = F->ci[j] / factor, t[nt].w = F->w[j], nt++;
      }
    h->M = realloc(h->M, sizeof(mgm_t) * (size_t)(h->nl + 1));
    mgm_from_triples(&h->M[h->nl], nc, t, nt);
    h->nl++;
  }
  free(t);
}

static void mg_free(mgh_t *h) {
  for (int l = 0; l < h->nl; l++) mgm_free(&h->M[l]);
  free(h->M);
}

/* Chebyshev-spaced Jacobi weights 1/lambda_k on [1, 2] (multigrid.c:24-30) */
static double sigma_cheb(int k, int n, double lmin, double lmax) {
  k = (k - 1) % n + 1;
  const double theta = M_PI * (k - 0.5) / n;
  return 1.0 / (lmin + 0.5 * (lmax - lmin) * (cos(theta) + 1.0));
}

/* One V-cycle (multigrid.c:235-303): on every level but the last, three
 * weighted Jacobi steps from zero (weights sigma_cheb(1..3, 4, 1, 2)) and
 * the residual summed into the aggregates; on the one-row level the solve
 * (zero, since that row's diagonal is zero); going up, each level's result
 * is its smoothed correction plus over times its aggregate's result (over =
 * 4/3 on level 0 as 1.33333, 1.5 above).  A zero diagonal gives a zero
 * Jacobi step here (the reference divides by it). */
static void mg_vcycle_run(const mgh_t *h, const double *rhs, double *out) {
  double **r = xmalloc(sizeof(double *) * (size_t)h->nl), **u = xmalloc(sizeof(double *) * (size_t)h->nl);
  for (int l = 0; l < h->nl; l++) {
    r[l] = calloc((size_t)h->M[l].n + 1, sizeof(double));
    u[l] = calloc((size_t)h->M[l].n + 1, sizeof(double));
  }
  memcpy(r[0], rhs, sizeof(double) * (size_t)h->M[0].n);
  double *s = xmalloc(sizeof(double) * (size_t)(h->M[0].n + 1));
  double *gs = xmalloc(sizeof(double) * (size_t)(h->M[0].n + 1));
  for (int l = 0; l + 1 < h->nl; l++) {
    const mgm_t *M = &h->M[l];
    const int n = M->n;
    for (int step = 0; step < 3; step++) {
      const double sg = sigma_cheb(step + 1, 4, 1.0, 2.0);
      for (int i = 0; i < n; i++) s[i] = M->diag[i] != 0.0 ? sg * r[l][i] / M->diag[i] : 0.0;
      for (int i = 0; i < n; i++) u[l][i] = step == 0 ? s[i] : u[l][i] + s[i];
      mgm_apply(M, s, gs);
      for (int i = 0; i < n; i++) r[l][i] -= gs[i];
    }
    for (int i = 0; i < n; i++) r[l + 1][i / h->factor] += r[l][i];
  }
  {
    const mgm_t *M = &h->M[h->nl - 1];
    if (M->n == 1) {
      const double v = fabs(M->diag[0]) > 1e-6 ? r[h->nl - 1][0] / M->diag[0] : 0.0;
      r[h->nl - 1][0] = v;
    }
  }
  for (int l = h->nl - 2; l >= 0; l--) {
    const double over = l == 0 ? 1.33333 : 1.5;
    for (int i = 0; i < h->M[l].n; i++) r[l][i] = over * r[l + 1][i / h->factor] + u[l][i];
  }
  memcpy(out, r[0], sizeof(double) * (size_t)h->M[0].n);
  for (int l = 0; l < h->nl; l++) free(r[l]), free(u[l]);
  free(r), free(u), free(s), free(gs);
}

/* Projection (fiedler.c:94-187): x ~ L^+ b by conjugate gradients with the
 * V-cycle as preconditioner, its result kept orthogonal to the constant
 * vector; the new direction uses r.(z - z_old) / (r.z_old) and is made
 * L-orthogonal to every earlier one.  Stops when |r|^2 < tol^2 |b|^2 or
 * |r| < tol; returns the iterations taken (miter if it never stopped). */
static int mg_project(double *x, int n, const double *b, const gs_lap *L, const mgh_t *h,
                      int miter, double tol) {
  if (n == 0) return 0;
  double *z = xmalloc(sizeof(double) * 6 * (size_t)n);
  double *w = z + n, *r = w + n, *p = r + n, *z0 = p + n, *dz = z0 + n;
  double *P = xmalloc(sizeof(double) * 2 * (size_t)(miter + 1) * n), *W = P + (size_t)n * (miter + 1);
  for (int i = 0; i < n; i++) x[i] = 0.0, r[i] = b[i];
  double rr = dotp(r, r, n);
  const double rtol = rr * tol * tol;
  memcpy(z, r, sizeof(double) * (size_t)n);
  ortho(z, n);
  double rz1 = dotp(z, z, n), rzt, rz2, alpha, beta;
  memcpy(p, z, sizeof(double) * (size_t)n);
  int i;
  for (i = 0; i < miter; i++) {
    gs_lap_apply(L, p, w);
    double pw = dotp(p, w, n);
    alpha = rz1 / pw;
    pw = 1.0 / sqrt(pw);
    for (int j = 0; j < n; j++) W[(size_t)i * n + j] = pw * w[j], P[(size_t)i * n + j] = pw * p[j];
    for (int j = 0; j < n; j++) x[j] += alpha * p[j], r[j] -= alpha * w[j];
    rr = dotp(r, r, n);
    if (rr < rtol || sqrt(rr) < tol) break;
    memcpy(z0, z, sizeof(double) * (size_t)n);
    mg_vcycle_run(h, r, z);
    rzt = rz1;
    ortho(z, n);
    rz1 = dotp(r, z, n);
    for (int j = 0; j < n; j++) dz[j] = z[j] - z0[j];
    rz2 = dotp(r, dz, n);
    beta = rz2 / rzt;
    for (int j = 0; j < n; j++) p[j] = z[j] + beta * p[j];
    double *acc = P + (size_t)miter * n;
    for (int k = 0; k < n; k++) acc[k] = 0.0;
    for (int j = 0; j <= i; j++) {
      const double c = dotp(W + (size_t)j * n, p, n);
      for (int k = 0; k < n; k++) acc[k] += c * P[(size_t)j * n + k];
    }
    for (int k = 0; k < n; k++) p[k] -= acc[k];
  }
  free(z), free(P);
  return i == miter ? i : i + 1;
}

/* Inverse iteration (fiedler.c:191-321): y = project(z); orthogonalise y;
 * z = y / |y| orthogonalised; stop after a projection that took a single
 * iteration (z was already an eigenvector to tol).  With grammian = 1
 * (fiedler.c:263-307) the normalised iterates span a basis Z and each pass
 * replaces z by the Rayleigh-Ritz vector of the Laplacian over Z:
 * M = Z L Z^T (M[k][j] = Z_k . L Z_j), z = sum_j v_j Z_j orthogonalised, v
 * the eigenvector of M of smallest eigenvalue.  Where the reference differs:
 *   - it orthogonalises z against Z once (classical Gram-Schmidt) and always
 *     adds the result; here twice (Giraud, Langou, Rozloznik, 2005: "twice is
 *     enough"), and the residual joins Z only if its norm exceeds 1e-10 (z
 *     has unit length): a z already in span(Z) otherwise adds a normalised
 *     rounding residual, Z stops being orthonormal and M's smallest
 *     eigenvalue then belongs to a near-cancelling combination;
 *   - it finds v by inverting M (LAPACK, BLAS builds) and 100 power steps
 *     from a srand(time) start; here v is the exact smallest eigenpair of
 *     (M + M^T) / 2 (cyclic Jacobi), signed so that sum_j v_j (Z_j . z) >= 0
 *     (the refined vector points the way the iterate does). */
static int mg_inverse(double *y, double *z, int n, const gs_lap *L, const mgh_t *h, int miter,
                      int mpass, double tol, int grammian, int *passes) {
  int iters = 0, i, nb = 0;
  double *Z = NULL, *GZ = NULL, *M = NULL, *Ms = NULL, *c = NULL, *v = NULL, *r = NULL;
  if (grammian) {
    Z = xmalloc(sizeof(double) * (size_t)mpass * n);
    GZ = xmalloc(sizeof(double) * (size_t)mpass * n);
    M = xmalloc(sizeof(double) * (size_t)mpass * mpass);
    Ms = xmalloc(sizeof(double) * (size_t)mpass * mpass);
    c = xmalloc(sizeof(double) * (size_t)mpass);
    v = xmalloc(sizeof(double) * (size_t)mpass);
    r = xmalloc(sizeof(double) * (size_t)n);
  }
  for (i = 0; i < mpass; i++) {
    const int ppfi = mg_project(y, n, z, L, h, miter, tol);
    iters += ppfi;
    ortho(y, n);
    const double nrm = dotp(y, y, n), ni = 1.0 / sqrt(nrm);
    for (int j = 0; j < n; j++) z[j] = y[j] * ni;
    ortho(z, n);
    if (grammian) {
      /* r = z minus its projection on Z, twice; c_j = Z_j . z */
      for (int l = 0; l < n; l++) r[l] = z[l];
      for (int j = 0; j < nb; j++) c[j] = 0.0;
      for (int g = 0; g < 2; g++) {
        for (int j = 0; j < nb; j++) Ms[j] = dotp(Z + (size_t)j * n, r, n);
        for (int j = 0; j < nb; j++) {
          c[j] += Ms[j];
          for (int l = 0; l < n; l++) r[l] = r[l] - Ms[j] * Z[(size_t)j * n + l];
        }
      }
      const double rn = sqrt(dotp(r, r, n));
      if (rn > 1e-10) {
        double *Zi = Z + (size_t)nb * n;
        const double rni = 1.0 / rn;
        for (int l = 0; l < n; l++) Zi[l] = r[l] * rni;
        c[nb] = dotp(Zi, z, n);
        gs_lap_apply(L, Zi, GZ + (size_t)nb * n);
        /* M grows by its new row and column (older entries: the same
           vectors as before) */
        const int N = nb + 1;
        for (int k = nb - 1; k >= 0; k--)
          for (int j = nb - 1; j >= 0; j--) M[k * N + j] = M[k * nb + j];
        for (int j = 0; j < N; j++) M[nb * N + j] = dotp(Zi, GZ + (size_t)j * n, n);
        for (int k = 0; k < nb; k++) M[k * N + nb] = dotp(Z + (size_t)k * n, GZ + (size_t)nb * n, n);
        nb = N;
      }
      if (nb >= 2) {
        for (int k = 0; k < nb; k++)
          for (int j = 0; j < nb; j++) Ms[k * nb + j] = 0.5 * (M[k * nb + j] + M[j * nb + k]);
        oracle_sym_smallest(Ms, nb, v);
        double dir = 0.0;
        for (int j = 0; j < nb; j++) dir += v[j] * c[j];
        if (dir < 0.0)
          for (int j = 0; j < nb; j++) v[j] = -v[j];
        for (int l = 0; l < n; l++) z[l] = 0.0;
        for (int j = 0; j < nb; j++)
          for (int l = 0; l < n; l++) z[l] += Z[(size_t)j * n + l] * v[j];
        ortho(z, n);
      }
    }
    if (ppfi == 1) {
      i++;
      break;
    }
  }
  if (passes) *passes = i;
  free(Z), free(GZ), free(M), free(Ms), free(c), free(v), free(r);
  return iters;
}

int oracle_fiedler_mg(const long long *vtx, int nv, const int *elems, int n, int max_iter,
                      int max_passes, double tol, int factor, int grammian, double *fiedler,
                      int *passes_out) {
  double *z = xmalloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) {
    z[i] = i + 1.0;
    if (i < n / 2) z[i] += (double)(1000 * (long long)n);
  }
  ortho(z, n);
  const double rtr = dotp(z, z, n), rni = 1.0 / sqrt(rtr);
  for (int i = 0; i < n; i++) z[i] *= rni;
  gs_lap L;
  gs_lap_init(&L, vtx, nv, elems, n);
  mgh_t h;
  mg_setup(&h, &L, factor);
  const int it =
      mg_inverse(fiedler, z, n, &L, &h, max_iter, max_passes, tol, grammian, passes_out);
  mg_free(&h);
  gs_lap_free(&L);
  double nrm = 0.0;
  for (int i = 0; i < n; i++) nrm += fiedler[i] * fiedler[i];
  const double ni = 1.0 / sqrt(nrm);
  for (int i = 0; i < n; i++) fiedler[i] *= ni;
  free(z);
  return it;
}

int oracle_mg_levels(const long long *vtx, int nv, const int *elems, int n, int factor,
                     int *sizes, int64_t *nnz) {
  gs_lap L;
  gs_lap_init(&L, vtx, nv, elems, n);
  mgh_t h;
  mg_setup(&h, &L, factor);
  for (int l = 0; l < h.nl; l++) {
    if (sizes) sizes[l] = h.M[l].n;
    if (nnz) nnz[l] = h.M[l].rp[h.M[l].n];
  }
  const int nl = h.nl;
  mg_free(&h);
  gs_lap_free(&L);
  return nl;
}

int oracle_mg_level_matrix(const long long *vtx, int nv, const int *elems, int n, int factor,
                           int level, int64_t *rowptr, int *col, double *w, double *diag) {
  gs_lap L;
  gs_lap_init(&L, vtx, nv, elems, n);
  mgh_t h;
  mg_setup(&h, &L, factor);
  int rc = level < 0 || level >= h.nl;
  if (!rc) {
    const mgm_t *M = &h.M[level];
    memcpy(rowptr, M->rp, sizeof(int64_t) * (size_t)(M->n + 1));
    memcpy(col, M->ci, sizeof(int) * (size_t)M->rp[M->n]);
    memcpy(w, M->w, sizeof(double) * (size_t)M->rp[M->n]);
    memcpy(diag, M->diag, sizeof(double) * (size_t)M->n);
  }
  mg_free(&h);
  gs_lap_free(&L);
  return rc;
}

void oracle_mg_vcycle(const long long *vtx, int nv, const int *elems, int n, int factor,
                      const double *rhs, double *out) {
  gs_lap L;
  gs_lap_init(&L, vtx, nv, elems, n);
  mgh_t h;
  mg_setup(&h, &L, factor);
  mg_vcycle_run(&h, rhs, out);
  mg_free(&h);
  gs_lap_free(&L);
}

int oracle_mg_project(const long long *vtx, int nv, const int *elems, int n, int factor,
                      int miter, double tol, const double *b, double *x) {
  gs_lap L;
  gs_lap_init(&L, vtx, nv, elems, n);
  mgh_t h;
  mg_setup(&h, &L, factor);
  const int it = mg_project(x, n, b, &L, &h, miter, tol);
  mg_free(&h);
  gs_lap_free(&L);
  return it;
}

/* ------------------------------------------------------------------------ */
/* Components by union-find over shared vertex ids. */
static int uf_find(int *par, int x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

static void uf_union(int *par, int a, int b) {
  a = uf_find(par, a), b = uf_find(par, b);
  if (a == b) return;
  if (a < b) par[b] = a;
  else par[a] = b;
}

int oracle_components(const long long *vtx, int nv, const int *elems, int n) {
  if (n == 0) return 0;
  int *group = xmalloc(sizeof(int) * (size_t)n * nv), *gstart, *gslots;
  int ng = group_slots(vtx, nv, elems, n, group, &gstart, &gslots);
  int *par = xmalloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) par[i] = i;
  for (int g = 0; g < ng; g++)
    for (int k = gstart[g] + 1; k < gstart[g + 1]; k++)
      uf_union(par, gslots[gstart[g]] / nv, gslots[k] / nv);
  int nc = 0;
  for (int i = 0; i < n; i++) nc += (uf_find(par, i) == i);
  free(group), free(gstart), free(gslots), free(par);
  return nc;
}

/* ------------------------------------------------------------------------ */
/* Interface flags for a bisected segment: flag[i] = 1 if element elems[i]
 * of side `side` shares a vertex id with an element of the other side
 * (rsb.c:145-165: gather-scatter add over the parent's ranks). */
static void interface_flags(const long long *vtx, int nv, const int *elems, int n, int nl,
                            int sender_is_lower, int *flag) {
  int *group = xmalloc(sizeof(int) * (size_t)n * nv), *gstart, *gslots;
  int ng = group_slots(vtx, nv, elems, n, group, &gstart, &gslots);
  /* receivers touching each group */
  int *recv = xmalloc(sizeof(int) * (size_t)ng);
  for (int g = 0; g < ng; g++) {
    int c = 0;
    for (int k = gstart[g]; k < gstart[g + 1]; k++) {
      int i = gslots[k] / nv;
      int lower = i < nl;
      if (lower != sender_is_lower) c++;
    }
    recv[g] = c;
  }
  for (int i = 0; i < n; i++) {
    flag[i] = 0;
    for (int j = 0; j < nv; j++)
      if (recv[group[i * nv + j]] > 0) {
        flag[i] = 1;
        break;
      }
  }
  free(group), free(gstart), free(gslots), free(recv);
}

/* Distinct other segments sharing a vertex id with each segment
 * (statistics.c:17-113); returns the maximum over segments. */
typedef struct {
  long long vid;
  int seg;
} vseg_t;

static int cmp_vseg(const void *a, const void *b) {
  const vseg_t *x = a, *y = b;
  if (x->vid != y->vid) return x->vid < y->vid ? -1 : 1;
  return (x->seg > y->seg) - (x->seg < y->seg);
}

static int cmp_ll(const void *a, const void *b) {
  long long x = *(const long long *)a, y = *(const long long *)b;
  return (x > y) - (x < y);
}

/* Neighbouring segments after a cut (statistics.c:17-113): for every
 * segment, the number of other segments sharing a vertex with it; min / max
 * over segments and the mean over virtual ranks (each rank reports its
 * segment's count, metrics.c:53-62). */
static void neighbor_stats(const long long *vtx, int nv, const int *segof, int nel, int nseg,
                           const int *segp, int *mn_out, int *mx_out, double *mean_out) {
  size_t ns = (size_t)nel * nv;
  vseg_t *a = xmalloc(sizeof(vseg_t) * ns);
  for (size_t s = 0; s < ns; s++) a[s].vid = vtx[s], a[s].seg = segof[s / nv];
  qsort(a, ns, sizeof(vseg_t), cmp_vseg);
  size_t cap = 1024, np = 0;
  long long *pairs = xmalloc(sizeof(long long) * cap);
  for (size_t s = 0; s < ns;) {
    size_t t = s;
    while (t < ns && a[t].vid == a[s].vid && a[s].vid != 0) t++;
    if (t == s) t = s + 1;
    for (size_t i = s; i < t; i++)
      for (size_t j = s; j < t; j++)
        if (a[i].seg != a[j].seg) {
          if (np == cap) cap *= 2, pairs = realloc(pairs, sizeof(long long) * cap);
          pairs[np++] = (long long)a[i].seg * nseg + a[j].seg;
        }
    s = t;
  }
  qsort(pairs, np, sizeof(long long), cmp_ll);
  int *cnt = calloc((size_t)nseg, sizeof(int));
  for (size_t i = 0; i < np; i++)
    if (i == 0 || pairs[i] != pairs[i - 1]) cnt[pairs[i] / nseg]++;
  int mx = 0, mn = 1 << 30;
  double sum = 0.0, ranks = 0.0;
  for (int s = 0; s < nseg; s++) {
    mx = cnt[s] > mx ? cnt[s] : mx;
    mn = cnt[s] < mn ? cnt[s] : mn;
    sum += (double)cnt[s] * segp[s], ranks += segp[s];
  }
  *mn_out = nseg ? mn : 0, *mx_out = mx, *mean_out = ranks > 0 ? sum / ranks : 0.0;
  free(a), free(pairs), free(cnt);
}

/* ------------------------------------------------------------------------ */
/* One balance step (rsb.c:118-229).  el[0..n) are a segment's elements in
 * ascending Fiedler order (values fv), el[0..nl) its lower half.  If the
 * sending side (lower when sender_lower = 1) has at least scnt interface
 * elements (rsb.c:145-165, 194), it hands over scnt of them: ordered by
 * mul * fiedler with ties by position (rsb.c:168-187), the upper side offers
 * its smallest values first, the lower side its largest.  Each half then
 * keeps its members in the current order (the reference re-sorts each half
 * by Fiedler value, rsb.c:219-220, which gives the same order up to ties).
 * el and fv are permuted in place; returns the new size of the lower half
 * (nl itself when nothing moves). */
int oracle_balance_step(const long long *vtx, int nv, int *el, double *fv, int n, int nl,
                        int sender_lower, int64_t scnt) {
  if (scnt <= 0 || n <= 0) return nl;
  int *flag = xmalloc(sizeof(int) * (size_t)n);
  interface_flags(vtx, nv, el, n, nl, sender_lower, flag);
  int lo = sender_lower ? 0 : nl, hi = sender_lower ? nl : n;
  int nif = 0;
  for (int i = lo; i < hi; i++) nif += flag[i];
  if (nif < scnt) {
    free(flag);
    return nl;
  }
  int *cand = xmalloc(sizeof(int) * (size_t)nif);
  double *ck = xmalloc(sizeof(double) * (size_t)nif);
  int k = 0;
  for (int i = lo; i < hi; i++)
    if (flag[i]) cand[k] = i, ck[k] = sender_lower ? -fv[i] : fv[i], k++;
  stable_sort_by_pos_key(cand, ck, nif);
  /* membership after the move (1 = lower), reusing flag */
  for (int i = 0; i < n; i++) flag[i] = i < nl;
  for (int j = 0; j < scnt; j++) flag[cand[j]] = !sender_lower;
  int *o2 = xmalloc(sizeof(int) * (size_t)n);
  double *f2 = xmalloc(sizeof(double) * (size_t)n);
  int a = 0;
  for (int i = 0; i < n; i++)
    if (flag[i]) o2[a] = el[i], f2[a] = fv[i], a++;
  const int nl2 = a;
  for (int i = 0; i < n; i++)
    if (!flag[i]) o2[a] = el[i], f2[a] = fv[i], a++;
  memcpy(el, o2, sizeof(int) * (size_t)n);
  memcpy(fv, f2, sizeof(double) * (size_t)n);
  free(o2), free(f2), free(cand), free(ck), free(flag);
  return nl2;
}

/* ------------------------------------------------------------------------ */
typedef struct {
  int start, n, p, r0;
} seg_t;

int oracle_part_mesh(int *part, const long long *vtx, const double *xyz, int nel, int nv,
                     const oracle_opts *o, oracle_report *rep) {
  int P = o->nparts;
  if (P < 1 || nel < 0) return 3;
  int ndim = (nv == 8) ? 3 : 2;
  double *coord = xmalloc(sizeof(double) * 3 * (size_t)(nel ? nel : 1));
  int rsb_pre = o->rsb_pre;
  if (xyz) oracle_centroids(xyz, nel, nv, coord);
  else rsb_pre = 0;
  int *order = xmalloc(sizeof(int) * (size_t)(nel ? nel : 1));
  for (int i = 0; i < nel; i++) order[i] = i;
  seg_t *segs = xmalloc(sizeof(seg_t) * (size_t)P), *nsegs = xmalloc(sizeof(seg_t) * (size_t)P);
  int nseg = 1;
  segs[0].start = 0, segs[0].n = nel, segs[0].p = P, segs[0].r0 = 0;
  int ncuts = 0;
  while ((1 << ncuts) < P) ncuts++;
  if (rep) memset(rep, 0, sizeof(*rep)), rep->nlevels = ncuts;
  double *fv = xmalloc(sizeof(double) * (size_t)(nel ? nel : 1));
  int *segof = xmalloc(sizeof(int) * (size_t)(nel ? nel : 1));

  int *coff = xmalloc(sizeof(int) * (size_t)P);
  double t_start = 0.0;
  if (oracle_verbose) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    t_start = ts.tv_sec + 1e-9 * ts.tv_nsec;
  }
  for (int cut = 0; cut < ncuts; cut++) {
    int nn = 0;
    int itmin = 1 << 30, itmax = 0, cmin = 1 << 30, cmax = 0, nbal = 0, ncut = 0;
    double lmin = DBL_MAX, lmax = -DBL_MAX;
    /* children slots: a final segment keeps one, a bisected one takes two */
    for (int s = 0; s < nseg; s++) coff[s] = nn, nn += segs[s].p == 1 ? 1 : 2;
    int nthr = 1;
#ifdef ORACLE_THREADS
    nthr = omp_get_max_threads();
#endif
    (void)nthr;
    OMP(omp parallel for schedule(dynamic, 1) if (nseg >= nthr && nthr > 1)
        reduction(min : itmin, cmin, lmin) reduction(max : itmax, cmax, lmax)
        reduction(+ : nbal, ncut))
    for (int s = 0; s < nseg; s++) {
      seg_t sg = segs[s];
      if (sg.p == 1) {
        nsegs[coff[s]] = sg;
        continue;
      }
      ncut++;
      int *el = order + sg.start, n = sg.n, p = sg.p;
      double *fs = fv + sg.start; /* this segment's share of the scratch */
      /* pre-partition (rsb.c:282-291) */
      if (rsb_pre == 1) {
        oracle_rcb_order(coord, ndim, el, n, p);
      } else if (rsb_pre == 2) {
        oracle_rib_order(coord, ndim, el, n, p);
      } else {
        for (int i = 0; i < n; i++) fs[i] = (double)el[i];
        stable_sort_by_pos_key(el, fs, n);
      }
      /* Fiedler (fiedler.c:563-621) */
      int it = 0;
      double lam = 0.0;
      if (n >= 2 && o->rsb_algo == 1) {
        it = oracle_fiedler_mg(vtx, nv, el, n, o->max_iter, o->max_passes, o->tol,
                               o->mg_factor, o->mg_grammian, fs, NULL);
      } else if (n >= 2) {
        it = oracle_fiedler(vtx, nv, el, n, o->max_iter, o->max_passes, o->tol, fs, &lam,
                            NULL);
      } else {
        for (int i = 0; i < n; i++) fs[i] = 0.0;
      }
      itmin = it < itmin ? it : itmin, itmax = it > itmax ? it : itmax;
      lmin = lam < lmin ? lam : lmin, lmax = lam > lmax ? lam : lmax;
      /* sort by Fiedler value (rsb.c:308) */
      stable_sort_by_pos_key(el, fs, n);
      int nl = (int)oracle_lower_count(n, p);
      /* components of each half before balancing (rsb.c:320-327) */
      if (o->find_disconnected_comps) {
        int c0 = oracle_components(vtx, nv, el, nl);
        int c1 = oracle_components(vtx, nv, el + nl, n - nl);
        int c = c0 < c1 ? c0 : c1, C = c0 > c1 ? c0 : c1;
        cmin = c < cmin ? c : cmin, cmax = C > cmax ? C : cmax;
      }
      /* balance (rsb.c:118-229) */
      int64_t expl = oracle_balance_expected(n, p, 0), expu = oracle_balance_expected(n, p, 1);
      int sender_lower = -1;
      int64_t scnt = 0;
      if (nl > expl) sender_lower = 1, scnt = nl - expl;
      else if (n - nl > expu) sender_lower = 0, scnt = (n - nl) - expu;
      if (scnt > 0) {
        int nl2 = oracle_balance_step(vtx, nv, el, fs, n, nl, sender_lower, scnt);
        if (nl2 != nl) nbal++;
        nl = nl2;
      }
      seg_t *c = nsegs + coff[s];
      c[0].start = sg.start, c[0].n = nl, c[0].p = (p + 1) / 2, c[0].r0 = sg.r0;
      c[1].start = sg.start + nl, c[1].n = n - nl, c[1].p = p / 2, c[1].r0 = sg.r0 + (p + 1) / 2;
    }
    seg_t *t = segs;
    segs = nsegs, nsegs = t, nseg = nn;
    if (oracle_verbose) {
      struct timespec ts;
      clock_gettime(CLOCK_MONOTONIC, &ts);
      fprintf(stderr, "oracle: cut %d/%d done (%d segments) at %.1f s\n", cut + 1, ncuts, ncut,
              ts.tv_sec + 1e-9 * ts.tv_nsec - t_start);
      fflush(stderr);
    }
    if (rep && cut < ORACLE_MAX_LEVELS) {
      rep->niter_min[cut] = ncut ? itmin : 0, rep->niter_max[cut] = itmax;
      rep->ncomp_min[cut] = o->find_disconnected_comps ? cmin : 0;
      rep->ncomp_max[cut] = cmax;
      rep->nsegs[cut] = ncut;
      rep->nbalanced[cut] = nbal;
      rep->lambda_min[cut] = ncut ? lmin : 0, rep->lambda_max[cut] = ncut ? lmax : 0;
      int *segp = xmalloc(sizeof(int) * (size_t)nseg);
      for (int s = 0; s < nseg; s++) {
        segp[s] = segs[s].p;
        for (int i = 0; i < segs[s].n; i++) segof[order[segs[s].start + i]] = s;
      }
      neighbor_stats(vtx, nv, segof, nel, nseg, segp, &rep->nbrs_min[cut], &rep->nbrs_max[cut],
                     &rep->nbrs_mean[cut]);
      free(segp);
    }
  }
  for (int s = 0; s < nseg; s++)
    for (int i = 0; i < segs[s].n; i++) part[order[segs[s].start + i]] = segs[s].r0;
  free(coord), free(order), free(segs), free(nsegs), free(fv), free(segof), free(coff);
  return 0;
}

/* Tagged partitioning, parrsb.c:379-561.  The distinct tags, ascending, get
 * numbers 0..ntags-1 (379-421); nparts must be a multiple of ntags (423-431);
 * each tag's elements, in input order, are partitioned on their own into
 * chunk = nparts / ntags parts (parrsb_part_mesh_v0 on a communicator of
 * chunk ranks, 532-541) and tag number t adds t * chunk to its part ids
 * (543-547).  Returns 3 when nparts is not a multiple of ntags. */
typedef struct {
  unsigned tag;
  int idx;
} ti_t;

static int cmp_ti(const void *a, const void *b) {
  const ti_t *x = a, *y = b;
  if (x->tag != y->tag) return (x->tag > y->tag) - (x->tag < y->tag);
  return (x->idx > y->idx) - (x->idx < y->idx);
}

int oracle_part_mesh_tagged(int *part, const long long *vtx, const double *xyz,
                            const int *tag, int nel, int nv, const oracle_opts *o) {
  if (nel <= 0) return 0;
  ti_t *t = xmalloc(sizeof(ti_t) * (size_t)nel);
  for (int i = 0; i < nel; i++) t[i].tag = (unsigned)tag[i], t[i].idx = i;
  qsort(t, (size_t)nel, sizeof(ti_t), cmp_ti);
  int ntags = 1;
  for (int i = 1; i < nel; i++) ntags += t[i].tag != t[i - 1].tag;
  if (o->nparts % ntags != 0) {
    free(t);
    return 3;
  }
  const int chunk = o->nparts / ntags, ndim = nv == 8 ? 3 : 2;
  oracle_opts so = *o;
  so.nparts = chunk;
  long long *lv = xmalloc(sizeof(long long) * (size_t)nel * nv);
  double *lx = xyz ? xmalloc(sizeof(double) * (size_t)nel * nv * ndim) : NULL;
  int *lp = xmalloc(sizeof(int) * (size_t)nel);
  int rc = 0;
  for (int s = 0, tn = 0; s < nel && !rc; tn++) {
    int e = s + 1;
    while (e < nel && t[e].tag == t[s].tag) e++;
    const int m = e - s;
    for (int j = 0; j < m; j++) {
      const int i = t[s + j].idx;
      memcpy(lv + (size_t)j * nv, vtx + (size_t)i * nv, sizeof(long long) * nv);
      if (lx) memcpy(lx + (size_t)j * nv * ndim, xyz + (size_t)i * nv * ndim,
                     sizeof(double) * nv * ndim);
    }
    rc = oracle_part_mesh(lp, lv, lx, m, nv, &so, NULL);
    for (int j = 0; j < m && !rc; j++) part[t[s + j].idx] = lp[j] + tn * chunk;
    s = e;
  }
  free(t), free(lv), free(lx), free(lp);
  return rc;
}

/* ------------------------------------------------------------------------ */
typedef struct {
  long long vid;
  int part;
} vp_t;

static int cmp_vp(const void *a, const void *b) {
  const vp_t *x = a, *y = b;
  if (x->vid != y->vid) return x->vid < y->vid ? -1 : 1;
  return (x->part > y->part) - (x->part < y->part);
}

void oracle_quality(const int *part, const long long *vtx, int nel, int nv, int nparts,
                    int64_t *out) {
  int64_t *rowptr;
  int *col, *wgt;
  oracle_dual_graph(vtx, nel, nv, &rowptr, &col, &wgt);
  int64_t wcut = 0, ucut = 0;
  int *par = xmalloc(sizeof(int) * (size_t)(nel ? nel : 1));
  for (int i = 0; i < nel; i++) par[i] = i;
  for (int i = 0; i < nel; i++)
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
      int j = col[k];
      if (part[i] != part[j]) {
        if (j > i) wcut += wgt[k], ucut++;
      } else {
        uf_union(par, i, j);
      }
    }
  int64_t *sz = calloc((size_t)nparts, sizeof(int64_t));
  int64_t *nc = calloc((size_t)nparts, sizeof(int64_t));
  for (int i = 0; i < nel; i++) {
    sz[part[i]]++;
    if (uf_find(par, i) == i) nc[part[i]]++;
  }
  int64_t mx = 0, mn = INT64_MAX, tot = 0, cmx = 0, empty = 0;
  for (int q = 0; q < nparts; q++) {
    mx = sz[q] > mx ? sz[q] : mx;
    mn = sz[q] < mn ? sz[q] : mn;
    tot += nc[q];
    cmx = nc[q] > cmx ? nc[q] : cmx;
    empty += sz[q] == 0;
  }
  size_t ns = (size_t)nel * nv;
  vp_t *a = xmalloc(sizeof(vp_t) * (ns ? ns : 1));
  for (size_t s = 0; s < ns; s++) a[s].vid = vtx[s], a[s].part = part[s / nv];
  qsort(a, ns, sizeof(vp_t), cmp_vp);
  int64_t vol = 0;
  for (size_t s = 0; s < ns;) {
    size_t t = s;
    int64_t k = 0;
    while (t < ns && a[t].vid == a[s].vid) {
      if (t == s || a[t].part != a[t - 1].part) k++;
      t++;
    }
    if (a[s].vid != 0) vol += k * (k - 1);
    s = t;
  }
  out[0] = wcut, out[1] = ucut, out[2] = mx, out[3] = mn, out[4] = tot, out[5] = cmx;
  out[6] = vol, out[7] = empty;
  free(rowptr), free(col), free(wgt), free(par), free(sz), free(nc), free(a);
}

/* ------------------------------------------------------------------------ */
/* Connectivity: global vertex ids from corner coordinates (con.c:189-295,
 * con-unique-vertices.c:445-505, con-check.c).  One process: every corner
 * point starts in one segment; ndim times over the ndim axes, each segment
 * is sorted by that coordinate and split wherever consecutive points differ
 * by more than tol^2 * min(dx) in squared distance, dx being the point's
 * smallest squared distance to the corners joined to it by an element edge
 * (con.c:70-101).  A point's id is the index of its final segment in the
 * final layout (con.c:120-145).  The reference sorts segments with a heap
 * sort (ties in unspecified order); this restatement sorts stably. */
static const int CON_PRE_TO_SYM[8] = {0, 1, 3, 2, 4, 5, 7, 6};
static const int CON_NEIGHBOR[8][3] = {{1, 2, 4}, {0, 3, 5}, {0, 3, 6}, {1, 2, 7},
                                       {0, 5, 6}, {1, 4, 7}, {2, 4, 7}, {3, 5, 6}};
static const int CON_FACES3D[6][4] = {{1, 5, 7, 3}, {2, 4, 8, 6}, {1, 2, 6, 5},
                                      {3, 7, 8, 4}, {1, 3, 4, 2}, {5, 6, 8, 7}};

static int conn_periodic(long long *vtx, const double *coord, int nel, int ndim,
                         const long long *pinfo, int npinfo);

int oracle_conn_mesh(long long *vtx, const double *coord, int nel, int ndim, double tol) {
  return oracle_conn_mesh_periodic(vtx, coord, nel, ndim, tol, NULL, 0);
}

int oracle_conn_mesh_periodic(long long *vtx, const double *coord, int nel, int ndim,
                              double tol, const long long *pinfo, int npinfo) {
  const int nv = ndim == 3 ? 8 : 4;
  const int64_t np = (int64_t)nel * nv;
  double *x = xmalloc(sizeof(double) * 3 * (size_t)(np ? np : 1));
  double *dx = xmalloc(sizeof(double) * (size_t)(np ? np : 1));
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < nv; k++) {
      const int j = CON_PRE_TO_SYM[k];
      for (int d = 0; d < 3; d++)
        x[3 * ((int64_t)e * nv + k) + d] = d < ndim ? coord[((int64_t)e * nv + j) * ndim + d] : 0.0;
    }
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < nv; k++) {
      const double *a = x + 3 * ((int64_t)e * nv + k);
      double m = DBL_MAX;
      for (int q = 0; q < ndim; q++) {
        const double *b = x + 3 * ((int64_t)e * nv + CON_NEIGHBOR[k][q]);
        double s = 0.0;
        for (int d = 0; d < ndim; d++) s += (a[d] - b[d]) * (a[d] - b[d]);
        m = s < m ? s : m;
      }
      dx[(int64_t)e * nv + k] = m;
    }
  int *lay = xmalloc(sizeof(int) * (size_t)(np ? np : 1));   /* layout -> point */
  int *seg = xmalloc(sizeof(int) * (size_t)(np ? np : 1));   /* layout -> segment */
  for (int64_t i = 0; i < np; i++) lay[i] = (int)i, seg[i] = 0;
  const double tol2 = tol * tol;
  double *key = xmalloc(sizeof(double) * (size_t)(np ? np : 1));
  for (int t = 0; t < ndim; t++)
    for (int d = 0; d < ndim; d++) {
      /* sort every segment by x[d] (segments are contiguous) */
      int64_t s0 = 0;
      while (s0 < np) {
        int64_t s1 = s0 + 1;
        while (s1 < np && seg[s1] == seg[s0]) s1++;
        for (int64_t i = s0; i < s1; i++) key[i] = x[3 * (int64_t)lay[i] + d];
        stable_sort_by_pos_key(lay + s0, key + s0, (int)(s1 - s0));
        s0 = s1;
      }
      /* split and renumber in layout order */
      int cur = 0;
      int prev_old = np ? seg[0] : 0;
      for (int64_t i = 1; i < np; i++) {
        const int pa = lay[i - 1], pb = lay[i];
        const double df = x[3 * (int64_t)pb + d] - x[3 * (int64_t)pa + d];
        const double lim = (dx[pb] < dx[pa] ? dx[pb] : dx[pa]) * tol2;
        const int old = seg[i];
        if (old != prev_old || df * df > lim) cur++;
        prev_old = old;
        seg[i] = cur;
      }
    }
  int64_t *gid = xmalloc(sizeof(int64_t) * (size_t)(np ? np : 1));
  for (int64_t i = 0; i < np; i++) gid[lay[i]] = seg[i];
  for (int64_t p = 0; p < np; p++) vtx[p] = gid[p] + 1;
  free(x), free(dx), free(lay), free(seg), free(key), free(gid);
  /* element check (con-check.c:262-288): no repeated id in an element */
  for (int e = 0; e < nel; e++)
    for (int a = 0; a < nv; a++)
      for (int b = a + 1; b < nv; b++)
        if (vtx[(int64_t)e * nv + a] == vtx[(int64_t)e * nv + b]) return 1;
  /* face check (con-check.c:197-260): no element holds exactly 3 of the 4
     corners of a face (quads: faces have 2 corners, never 3) */
  if (ndim == 3) {
    /* vertex -> elements holding it */
    int64_t nvid = 0;
    for (int64_t p = 0; p < np; p++) nvid = vtx[p] > nvid ? vtx[p] : nvid;
    int64_t *cnt = calloc((size_t)nvid + 2, sizeof(int64_t));
    for (int64_t p = 0; p < np; p++) cnt[vtx[p] + 1]++;
    for (int64_t v = 0; v <= nvid; v++) cnt[v + 1] += cnt[v];
    int *els = xmalloc(sizeof(int) * (size_t)(np ? np : 1));
    int64_t *fill = xmalloc(sizeof(int64_t) * (size_t)(nvid + 1));
    for (int64_t v = 0; v <= nvid; v++) fill[v] = cnt[v];
    for (int64_t p = 0; p < np; p++) els[fill[vtx[p]]++] = (int)(p / nv);
    int err = 0;
    for (int e = 0; e < nel && !err; e++)
      for (int f = 0; f < 6 && !err; f++) {
        /* elements of the face's corners, with multiplicity per corner */
        for (int c = 0; c < 4 && !err; c++) {
          const long long v = vtx[(int64_t)e * nv + CON_FACES3D[f][c] - 1];
          for (int64_t l = cnt[v]; l < cnt[v + 1] && !err; l++) {
            const int el = els[l];
            int m = 0;
            for (int c2 = 0; c2 < 4; c2++) {
              const long long v2 = vtx[(int64_t)e * nv + CON_FACES3D[f][c2] - 1];
              for (int64_t l2 = cnt[v2]; l2 < cnt[v2 + 1]; l2++) m += els[l2] == el;
            }
            if (m == 3) err = 1;
          }
        }
      }
    free(cnt), free(els), free(fill);
    if (err) return 1;
  }
  return conn_periodic(vtx, coord, nel, ndim, pinfo, npinfo);
}

/* Periodic faces, con-periodic.c:8-288.  pinfo[4i..4i+3] = (element, face,
 * element, face), 1-based, faces in the preprocessor numbering (con.c:41-47).
 * A face whose (element, face) is named by a later record's partner is matched
 * with that record's face: both are centred, the partner's corners are taken
 * in reverse order under the cyclic shift with the least sum of squared
 * distances (first such shift), and the match fails (returns 1; the reference
 * exits) if the root of that sum exceeds 1e-3 times the largest absolute
 * coordinate of the two faces.  Every matched corner pair joins two ids; each
 * id then takes the smallest id of its class, and the ids are renumbered
 * densely in ascending order (1-based). */
static const int CON_FACE_PRE_TO_SYM[6] = {2, 1, 3, 0, 4, 5};
static const int CON_FACES2D[4][2] = {{3, 1}, {2, 4}, {1, 2}, {4, 3}};

static int conn_periodic(long long *vtx, const double *coord, int nel, int ndim,
                         const long long *pinfo, int npinfo) {
  if (npinfo <= 0) return 0;
  const int nv = ndim == 3 ? 8 : 4, nvf = nv / 2;
  for (int i = 0; i < npinfo; i++) {
    const long long *r = pinfo + 4 * (int64_t)i;
    if (r[0] < 1 || r[0] > nel || r[2] < 1 || r[2] > nel || r[1] < 1 || r[1] > 2 * ndim ||
        r[3] < 1 || r[3] > 2 * ndim)
      return 3;
  }
  /* corners of each record's face: coordinates and current ids */
  double *fx = xmalloc(sizeof(double) * (size_t)npinfo * 4 * 3);
  long long *fid = xmalloc(sizeof(long long) * (size_t)npinfo * 4);
  for (int i = 0; i < npinfo; i++) {
    const int e = (int)pinfo[4 * (int64_t)i] - 1;
    const int f = CON_FACE_PRE_TO_SYM[pinfo[4 * (int64_t)i + 1] - 1];
    for (int j = 0; j < nvf; j++) {
      const int s = (ndim == 3 ? CON_FACES3D[f][j] : CON_FACES2D[f][j]) - 1;
      const int pre = CON_PRE_TO_SYM[s];
      for (int d = 0; d < 3; d++)
        fx[((int64_t)i * 4 + j) * 3 + d] =
            d < ndim ? coord[((int64_t)e * nv + pre) * ndim + d] : 0.0;
      fid[(int64_t)i * 4 + j] = vtx[(int64_t)e * nv + s];
    }
  }
  /* matched id pairs */
  int64_t npair = 0, cap = 16;
  long long *pa = xmalloc(sizeof(long long) * (size_t)cap),
            *pb = xmalloc(sizeof(long long) * (size_t)cap);
  int rc = 0;
  for (int i = 0; i < npinfo - 1 && !rc; i++) {
    const long long ei = pinfo[4 * (int64_t)i], fi = CON_FACE_PRE_TO_SYM[pinfo[4 * (int64_t)i + 1] - 1];
    for (int j = i + 1; j < npinfo && !rc; j++) {
      const long long *rj = pinfo + 4 * (int64_t)j;
      if (rj[2] != ei || CON_FACE_PRE_TO_SYM[rj[3] - 1] != fi) continue;
      double F[4][3], G[4][3], fmax = 0.0, gmax = 0.0;
      for (int d = 0; d < ndim; d++) {
        double mf = 0.0, mg = 0.0;
        for (int c = 0; c < nvf; c++) {
          F[c][d] = fx[((int64_t)i * 4 + c) * 3 + d], G[c][d] = fx[((int64_t)j * 4 + c) * 3 + d];
          fmax = fabs(F[c][d]) > fmax ? fabs(F[c][d]) : fmax;
          gmax = fabs(G[c][d]) > gmax ? fabs(G[c][d]) : gmax;
          mf += F[c][d], mg += G[c][d];
        }
        for (int c = 0; c < nvf; c++) F[c][d] -= mf / nvf, G[c][d] -= mg / nvf;
      }
      int shift = 0;
      double best = 1e20;
      for (int sh = 0; sh < nvf; sh++) {
        double d2 = 0.0;
        for (int c = 0; c < nvf; c++) {
          const int k = nvf - 1 - (c + sh) % nvf;
          for (int d = 0; d < ndim; d++) d2 += (F[c][d] - G[k][d]) * (F[c][d] - G[k][d]);
        }
        if (d2 < best) best = d2, shift = sh;
      }
      if (sqrt(best) > 1e-3 * (fmax > gmax ? fmax : gmax)) {
        rc = 1;
        break;
      }
      for (int c = 0; c < nvf; c++) {
        const int k = nvf - 1 - (c + shift) % nvf;
        const long long a = fid[(int64_t)i * 4 + c], b = fid[(int64_t)j * 4 + k];
        if (npair == cap) {
          cap *= 2;
          pa = realloc(pa, sizeof(long long) * (size_t)cap);
          pb = realloc(pb, sizeof(long long) * (size_t)cap);
        }
        pa[npair] = a < b ? a : b, pb[npair] = a < b ? b : a, npair++;
      }
    }
  }
  if (!rc && npair > 0) {
    /* smallest id of each class: relax the pairs until nothing changes */
    const int64_t np = (int64_t)nel * nv;
    long long nid = 0;
    for (int64_t p = 0; p < np; p++) nid = vtx[p] > nid ? vtx[p] : nid;
    long long *lab = xmalloc(sizeof(long long) * (size_t)(nid + 1));
    for (long long v = 0; v <= nid; v++) lab[v] = v;
    for (int changed = 1; changed;) {
      changed = 0;
      for (int64_t q = 0; q < npair; q++) {
        const long long m = lab[pa[q]] < lab[pb[q]] ? lab[pa[q]] : lab[pb[q]];
        if (lab[pa[q]] != m || lab[pb[q]] != m) lab[pa[q]] = lab[pb[q]] = m, changed = 1;
      }
    }
    /* dense renumbering in ascending order of the surviving ids */
    long long *rank = xmalloc(sizeof(long long) * (size_t)(nid + 1));
    long long r = 0;
    for (long long v = 1; v <= nid; v++) rank[v] = lab[v] == v ? ++r : 0;
    for (int64_t p = 0; p < np; p++) vtx[p] = rank[lab[vtx[p]]];
    free(lab), free(rank);
  }
  free(fx), free(fid), free(pa), free(pb);
  return rc;
}
