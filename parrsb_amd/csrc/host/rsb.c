/*
 * Level loop of recursive spectral bisection on one GPU (the reference's
 * rsb(), rsb.c:262-350, with P virtual ranks in place of P MPI ranks).
 *
 * A segment is a contiguous range of the device-held element order together
 * with the range of virtual ranks [r0, r0 + p) it is spread over.  Every cut
 * bisects all segments with p > 1 at once:
 *   pre-partition (RCB order)  rsb.c:282-291      rsbgpu_prepartition
 *   Laplacian of the segment   fiedler.c:519       rsbgpu_level_laplacian
 *   Fiedler vector             rsb.c:301           rsbgpu_fiedler
 *   sort by Fiedler value      rsb.c:308           rsbgpu_fiedler_sort
 *   split after ceil(p/2) ranks' share, components, balance
 *                              rsb.c:313-334       rsbgpu_bisect
 * The split arithmetic (how many elements the lower ranks hold after a sort,
 * how many each half should hold) is done here, on the host.
 */
#define _POSIX_C_SOURCE 200809L
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <time.h>

#include "parrsb_host.h"

static double wall(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int cmp_dbl(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* Elements on the lower ceil(p/2) virtual ranks after n elements are spread
 * over p ranks: the first p - (n mod p) ranks get floor(n/p), the others one
 * more (sort.c:48-70). */
long long rsb_lower_count(long long n, int p) {
  long long q = n / p, rem = n - q * p, hl = (p + 1) / 2;
  long long extra = hl - (p - rem);
  return q * hl + (extra > 0 ? extra : 0);
}

/* Expected size of bin 0 / bin 1 after balancing (rsb.c:138-140). */
long long rsb_balance_expected(long long n, int p, int bin) {
  long long q = n / p, rem = n - q * p;
  long long np_half = bin == 0 ? (p + 1) / 2 : p / 2;
  return q * np_half + rem / 2 + (rem % 2) * (1 - bin);
}

int rsb_ncuts(int nparts) {
  int c = 0;
  while ((1ll << c) < nparts) c++;
  return c;
}

typedef struct {
  int *cstart, *n, *p, *r0;
} segtab;

/* RSB_NEIGHBORS of the segments just set (rsb.c:342-343, statistics.c:17-113):
 * min / max over segments, mean over the virtual ranks. */
static int level_neighbors(rsbgpu *g, int nseg, const int *p, int *nb, parrsb_amd_report *rep,
                           int lv) {
  int rc = rsbgpu_neighbors(g, nb);
  if (rc) return rc;
  int mn = 1 << 30, mx = 0;
  double sum = 0.0, ranks = 0.0;
  for (int s = 0; s < nseg; s++) {
    mn = nb[s] < mn ? nb[s] : mn;
    mx = nb[s] > mx ? nb[s] : mx;
    sum += (double)nb[s] * p[s], ranks += p[s];
  }
  rep->nbrs_min[lv] = nseg ? mn : 0, rep->nbrs_max[lv] = mx;
  rep->nbrs_mean[lv] = ranks > 0 ? sum / ranks : 0.0;
  return 0;
}

static int segtab_alloc(segtab *t, int cap) {
  t->cstart = (int *)calloc((size_t)cap, sizeof(int));
  t->n = (int *)calloc((size_t)cap, sizeof(int));
  t->p = (int *)calloc((size_t)cap, sizeof(int));
  t->r0 = (int *)calloc((size_t)cap, sizeof(int));
  return t->cstart && t->n && t->p && t->r0;
}

static void segtab_free(segtab *t) { free(t->cstart), free(t->n), free(t->p), free(t->r0); }

#define TRY(x)                  \
  do {                          \
    int rc_ = (x);              \
    if (rc_) {                  \
      rc = rc_;                 \
      goto done;                \
    }                           \
  } while (0)

/* Initial segments: one per tag group with nparts / ngroups virtual ranks
   each (parrsb.c:423-431, 532-541), or one segment of all elements. */
static int initial_segments(segtab *t, int nel, int P, int ngroups, const int *count) {
  if (ngroups <= 1) {
    t->cstart[0] = 0, t->n[0] = nel, t->p[0] = P, t->r0[0] = 0;
    return 1;
  }
  const int chunk = P / ngroups;
  for (int i = 0, c = 0; i < ngroups; c += count[i], i++)
    t->cstart[i] = c, t->n[i] = count[i], t->p[i] = chunk, t->r0[i] = i * chunk;
  return ngroups;
}

int rsb_run(rsbgpu *g, int nel, const rsb_params *prm, parrsb_amd_report *rep) {
  int rc = 0, P = prm->nparts;
  int ncuts = rsb_ncuts(prm->ngroups > 1 ? P / prm->ngroups : P);
  segtab A, B;
  int ok = segtab_alloc(&A, P) & segtab_alloc(&B, P);
  int *nl = (int *)calloc((size_t)P, sizeof(int)), *snd = (int *)calloc((size_t)P, sizeof(int));
  int *scnt = (int *)calloc((size_t)P, sizeof(int)), *moved = (int *)calloc((size_t)P, sizeof(int));
  int *ncomp = (int *)calloc(2 * (size_t)P, sizeof(int));
  int *niter = (int *)calloc((size_t)P, sizeof(int)), *npass = (int *)calloc((size_t)P, sizeof(int));
  double *lam = (double *)calloc((size_t)P, sizeof(double));
  double *rn = (double *)calloc(3 * (size_t)P, sizeof(double));
  double *res = (double *)calloc((size_t)P, sizeof(double));
  double *rsort = (double *)calloc((size_t)P, sizeof(double));
  int *fb = (int *)calloc((size_t)P, sizeof(int)), *nbr = (int *)calloc((size_t)P, sizeof(int));
  if (!ok || !nl || !snd || !scnt || !moved || !ncomp || !niter || !npass || !lam || !rn || !res ||
      !rsort || !fb || !nbr) {
    rc = 2;
    goto done;
  }
  int nseg = initial_segments(&A, nel, P, prm->ngroups, prm->group_count);
  rep->nlevels = ncuts;
  for (int cut = 0; cut < ncuts; cut++) {
    double tl = wall();
    TRY(rsbgpu_set_segments(g, nseg, A.cstart, A.n, A.p, A.r0));
    if (cut > 0 && cut - 1 < PARRSB_AMD_MAX_LEVELS)
      TRY(level_neighbors(g, nseg, A.p, nbr, rep, cut - 1));
    TRY(rsbgpu_prepartition(g, prm->rsb_pre));
    TRY(rsbgpu_level_laplacian(g));
    rsbgpu_fiedler_out fo = {niter, npass, lam, rn, prm->residuals ? res : NULL, fb};
    int mode = 0;
    if (prm->warm && prm->rsb_algo == 0) {
      TRY(rsbgpu_lws_set_params(g, prm->warm, (float)cut, prm->warm_passes, prm->warm_max_n));
      mode = 2;
    }
    memset(fb, 0, sizeof(int) * (size_t)nseg);
    if (prm->rsb_algo == 1)
      TRY(rsbgpu_fiedler_mg(g, prm->max_iter, prm->max_passes, prm->tol, prm->mg_factor,
                             prm->mg_grammian, &fo));
    else
      TRY(rsbgpu_fiedler(g, prm->precision, prm->max_iter, prm->max_passes, prm->tol, mode,
                         &fo));
    TRY(rsbgpu_fiedler_sort(g));
    for (int s = 0; s < nseg; s++) {
      snd[s] = -1, scnt[s] = 0, nl[s] = A.n[s];
      if (A.p[s] <= 1) continue;
      long long n = A.n[s];
      long long l = rsb_lower_count(n, A.p[s]);
      long long e0 = rsb_balance_expected(n, A.p[s], 0), e1 = rsb_balance_expected(n, A.p[s], 1);
      nl[s] = (int)l;
      if (l > e0) snd[s] = 0, scnt[s] = (int)(l - e0);
      else if (n - l > e1) snd[s] = 1, scnt[s] = (int)(n - l - e1);
    }
    TRY(rsbgpu_bisect(g, nl, snd, scnt, prm->find_comps, ncomp, moved));
    /* level statistics */
    int lv = cut < PARRSB_AMD_MAX_LEVELS ? cut : PARRSB_AMD_MAX_LEVELS - 1;
    int itmin = 1 << 30, itmax = 0, pmax = 0, cmin = 1 << 30, cmax = 0, nb = 0, ncut = 0, nfail = 0;
    double lmin = DBL_MAX, lmax = -DBL_MAX, rmax = 0.0, work = 0.0, nwork = 0.0;
    double t_init = DBL_MAX, t_tgt = DBL_MAX, t_fnl = DBL_MAX;
    int nres = 0;
    for (int s = 0; s < nseg; s++) {
      if (A.p[s] <= 1) continue;
      ncut++;
      if (A.n[s] >= 2) {
        itmin = niter[s] < itmin ? niter[s] : itmin;
        itmax = niter[s] > itmax ? niter[s] : itmax;
        pmax = npass[s] > pmax ? npass[s] : pmax;
        lmin = lam[s] < lmin ? lam[s] : lmin;
        lmax = lam[s] > lmax ? lam[s] : lmax;
        if (prm->residuals) {
          if (res[s] > rmax) rmax = res[s];
          rsort[nres++] = res[s];
        }
        work += (double)niter[s] * A.n[s];
        nwork += A.n[s];
        /* the test of rsb.c:66-71 on the iterations of the reference-start
           solve (a fallback counts the learned attempt's fb[s] - 1
           iterations in front of them): rsb_max_iter * rsb_max_passes for
           Lanczos, rsb_max_passes for the inverse iteration */
        int fail;
        if (prm->rsb_algo == 1) fail = niter[s] == prm->max_passes;
        else fail = niter[s] - (fb[s] > 0 ? fb[s] - 1 : 0) == prm->max_iter * prm->max_passes;
        if (fail) {
          nfail++;
          if (rn[3 * s] < t_init) t_init = rn[3 * s];
          if (rn[3 * s + 1] < t_tgt) t_tgt = rn[3 * s + 1];
          if (rn[3 * s + 2] < t_fnl) t_fnl = rn[3 * s + 2];
        }
      }
      if (prm->find_comps) {
        for (int h = 0; h < 2; h++) {
          int c = ncomp[2 * s + h];
          cmin = c < cmin ? c : cmin;
          cmax = c > cmax ? c : cmax;
        }
      }
      nb += moved[s];
      rep->nfallback[lv] += fb[s] > 0;
    }
    rep->nsegs[lv] = ncut;
    rep->niter_min[lv] = ncut ? itmin : 0, rep->niter_max[lv] = itmax;
    rep->npass_max[lv] = pmax;
    rep->ncomp_min[lv] = prm->find_comps && ncut ? cmin : 0;
    rep->ncomp_max[lv] = prm->find_comps ? cmax : 0;
    rep->nbalanced[lv] = nb;
    rep->nconverged_fail[lv] = nfail;
    rep->lambda_min[lv] = ncut ? lmin : 0.0, rep->lambda_max[lv] = ncut ? lmax : 0.0;
    rep->resid_max[lv] = rmax;
    if (nres > 0) {
      qsort(rsort, (size_t)nres, sizeof(double), cmp_dbl);
      rep->resid_med[lv] = rsort[nres / 2];
    }
    if (nfail > 0 && prm->rsb_algo == 0)
      rep->tol_init[lv] = t_init, rep->tol_target[lv] = t_tgt, rep->tol_final[lv] = t_fnl;
    rep->niter_mean[lv] = nwork > 0 ? work / nwork : 0.0;
    if (prm->verbose > 1)
      printf("rsb: cut %d/%d segments=%d niter=%d..%d ncomp=%d..%d balanced=%d\n", cut + 1,
             ncuts, ncut, rep->niter_min[lv], itmax, rep->ncomp_min[lv], cmax, nb);
    /* children: lower half keeps the first ceil(p/2) ranks (rsb.c:231-242) */
    int nn = 0;
    for (int s = 0; s < nseg; s++) {
      if (A.p[s] <= 1) {
        B.cstart[nn] = A.cstart[s], B.n[nn] = A.n[s], B.p[nn] = A.p[s], B.r0[nn] = A.r0[s];
        nn++;
        continue;
      }
      int pl = (A.p[s] + 1) / 2, pu = A.p[s] / 2;
      B.cstart[nn] = A.cstart[s], B.n[nn] = nl[s], B.p[nn] = pl, B.r0[nn] = A.r0[s], nn++;
      B.cstart[nn] = A.cstart[s] + nl[s], B.n[nn] = A.n[s] - nl[s], B.p[nn] = pu;
      B.r0[nn] = A.r0[s] + pl, nn++;
    }
    segtab t = A;
    A = B, B = t, nseg = nn;
    rep->level_ms[lv] = 1e3 * (wall() - tl);
  }
  TRY(rsbgpu_set_segments(g, nseg, A.cstart, A.n, A.p, A.r0));
  if (ncuts > 0 && ncuts - 1 < PARRSB_AMD_MAX_LEVELS)
    TRY(level_neighbors(g, nseg, A.p, nbr, rep, ncuts - 1));
done:
  segtab_free(&A), segtab_free(&B);
  free(nl), free(snd), free(scnt), free(moved), free(ncomp), free(niter), free(npass), free(lam),
      free(rn), free(res), free(rsort), free(fb), free(nbr);
  return rc;
}

/* partitioner = 1: RCB alone, 2: RIB alone (parrsb.c:275-280).  The element
 * order after the virtual-rank levels gives each rank its share; the ranks'
 * ranges follow from the same split arithmetic. */
static void rcb_ranges(int *cstart, int *n, int start, int len, int p, int r0) {
  if (p == 1) {
    cstart[r0] = start, n[r0] = len;
    return;
  }
  int l = (int)rsb_lower_count(len, p);
  rcb_ranges(cstart, n, start, l, (p + 1) / 2, r0);
  rcb_ranges(cstart, n, start + l, len - l, p / 2, r0 + (p + 1) / 2);
}

int rcb_run(rsbgpu *g, int nel, int P, int pre, int ngroups, const int *group_count) {
  int rc = 0;
  segtab G;
  int *cs = (int *)calloc((size_t)P, sizeof(int)), *n = (int *)calloc((size_t)P, sizeof(int));
  int *p = (int *)calloc((size_t)P, sizeof(int)), *r0 = (int *)calloc((size_t)P, sizeof(int));
  int ok = segtab_alloc(&G, P);
  if (!cs || !n || !p || !r0 || !ok) {
    rc = 2;
    goto done;
  }
  int ng = initial_segments(&G, nel, P, ngroups, group_count);
  TRY(rsbgpu_set_segments(g, ng, G.cstart, G.n, G.p, G.r0));
  TRY(rsbgpu_prepartition(g, pre));
  for (int i = 0; i < ng; i++) rcb_ranges(cs, n, G.cstart[i], G.n[i], G.p[i], G.r0[i]);
  for (int r = 0; r < P; r++) p[r] = 1, r0[r] = r;
  TRY(rsbgpu_set_segments(g, P, cs, n, p, r0));
done:
  segtab_free(&G);
  free(cs), free(n), free(p), free(r0);
  return rc;
}
