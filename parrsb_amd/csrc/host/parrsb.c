/*
 * Public entry points: parrsb_part_mesh (drop-in for parrsb.c:853-906) and
 * the device-pointer variant.  Host code in C; the device work goes through
 * the rsbgpu C ABI.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "parrsb_host.h"

/* MPI is used only to learn the size of `comm`, and only if the process has
 * MPI loaded and initialised; the library itself does not link MPI. */
#if defined(PARRSB_NO_MPI)
int MPI_Initialized(int *flag) __attribute__((weak));
int MPI_Comm_size(MPI_Comm comm, int *size) __attribute__((weak));
#else
#pragma weak MPI_Initialized
#pragma weak MPI_Comm_size
#endif

/* Defaults of parrsb.c:26-44, plus nparts = 0 (one part per rank). */
parrsb_options parrsb_default_options = {
    .partitioner = 0,
    .tagged = 0,
    .levels = 2,
    .find_disconnected_comps = 1,
    .repair = 0,
    .verbose_level = 1,
    .profile_level = 0,
    .rsb_algo = 0,
    .rsb_pre = 1,
    .rsb_max_iter = 50,
    .rsb_max_passes = 50,
    .rsb_tol = 1e-5,
    .rsb_dump_stats = 0,
    .rsb_mg_grammian = 0,
    .rsb_mg_factor = 2,
    .nparts = 0};

static parrsb_amd_report last_report;
static char last_error[512];
static int cfg_precision = 64, cfg_profiling = 0;
static float *cfg_warm = NULL; /* host copy of the learned-start parameters */
static int cfg_warm_passes = 1, cfg_warm_max_n = 1024;
static char cfg_warm_file[512];

const char *parrsb_amd_last_error(void) { return last_error; }

int parrsb_amd_last_report(parrsb_amd_report *rep) {
  if (!rep) return 1;
  *rep = last_report;
  return 0;
}

int parrsb_amd_set_precision(int bits) {
  if (bits != 32 && bits != 64) return 3;
  cfg_precision = bits;
  return 0;
}

int parrsb_amd_set_profiling(int on) {
  cfg_profiling = on != 0;
  return 0;
}

void parrsb_host_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(last_error, sizeof(last_error), fmt, ap);
  va_end(ap);
}
#define set_error parrsb_host_error

int parrsb_host_comm_size(MPI_Comm comm) {
  int np = 1;
  if (MPI_Initialized && MPI_Comm_size) {
    int init = 0;
    MPI_Initialized(&init);
    if (init) MPI_Comm_size(comm, &np);
  }
  return np;
}

int parrsb_amd_warm_start_nparams(void) { return rsbgpu_lws_nparams(); }

int parrsb_amd_set_warm_start(const float *params, int passes, int max_n) {
  if (params && (passes < 1 || max_n < 0)) return 3;
  free(cfg_warm);
  cfg_warm = NULL;
  cfg_warm_file[0] = 0;
  if (!params) return 0;
  size_t n = (size_t)rsbgpu_lws_nparams();
  cfg_warm = (float *)malloc(n * sizeof(float));
  if (!cfg_warm) return 2;
  memcpy(cfg_warm, params, n * sizeof(float));
  cfg_warm_passes = passes;
  cfg_warm_max_n = max_n;
  return 0;
}

/* Parameter file: "LWS1", int32 count, count float32 values (little endian,
 * as written by parrsb_amd.lws_train). */
static int load_warm_file(const char *path) {
  if (cfg_warm && strcmp(path, cfg_warm_file) == 0) return 0;
  FILE *fp = fopen(path, "rb");
  if (!fp) {
    set_error("cannot open warm-start file %s", path);
    return 3;
  }
  char magic[4];
  int32_t n = 0;
  int ok = fread(magic, 1, 4, fp) == 4 && memcmp(magic, "LWS1", 4) == 0 &&
           fread(&n, sizeof(n), 1, fp) == 1 && n == rsbgpu_lws_nparams();
  float *buf = ok ? (float *)malloc(sizeof(float) * (size_t)n) : NULL;
  ok = ok && buf && fread(buf, sizeof(float), (size_t)n, fp) == (size_t)n;
  fclose(fp);
  if (!ok) {
    free(buf);
    set_error("warm-start file %s is not a parameter file of this network", path);
    return 3;
  }
  const char *v = getenv("PARRSB_AMD_WARM_PASSES"), *m = getenv("PARRSB_AMD_WARM_MAX_N");
  int rc = parrsb_amd_set_warm_start(buf, v ? atoi(v) : cfg_warm_passes,
                                     m ? atoi(m) : cfg_warm_max_n);
  free(buf);
  if (rc) {
    set_error("bad PARRSB_AMD_WARM_PASSES or PARRSB_AMD_WARM_MAX_N");
    return 3;
  }
  snprintf(cfg_warm_file, sizeof(cfg_warm_file), "%s", path);
  return 0;
}

/* PARRSB_* environment overrides written into the caller's struct
 * (parrsb.c:48-77), plus PARRSB_NPARTS. */
static void update_options(parrsb_options *o) {
#define UPD(field, name, is_int)                 \
  do {                                           \
    const char *v_ = getenv(name);               \
    if (v_) {                                    \
      if (is_int) o->field = atoi(v_);           \
      else o->field = atof(v_);                  \
    }                                            \
  } while (0)
  UPD(partitioner, "PARRSB_PARTITIONER", 1);
  UPD(tagged, "PARRSB_TAGGED", 1);
  UPD(levels, "PARRSB_LEVELS", 1);
  UPD(find_disconnected_comps, "PARRSB_FIND_DISCONNECTED_COMPONENTS", 1);
  UPD(repair, "PARRSB_REPAIR", 1);
  UPD(verbose_level, "PARRSB_VERBOSE_LEVEL", 1);
  UPD(profile_level, "PARRSB_PROFILE_LEVEL", 1);
  UPD(rsb_algo, "PARRSB_RSB_ALGO", 1);
  UPD(rsb_pre, "PARRSB_RSB_PRE", 1);
  UPD(rsb_max_iter, "PARRSB_RSB_MAX_ITER", 1);
  UPD(rsb_max_passes, "PARRSB_RSB_MAX_PASSES", 1);
  UPD(rsb_tol, "PARRSB_RSB_TOL", 0);
  UPD(rsb_mg_grammian, "PARRSB_RSB_MG_GRAMMIAN", 1);
  UPD(rsb_mg_factor, "PARRSB_RSB_MG_FACTOR", 1);
  UPD(nparts, "PARRSB_NPARTS", 1);
#undef UPD
}

static void print_options(const parrsb_options *o) {
  if (o->verbose_level <= 0) return;
#define PO(field, name, fmt) printf("%s = " fmt "\n", name, o->field)
  PO(partitioner, "PARRSB_PARTITIONER", "%d");
  PO(tagged, "PARRSB_TAGGED", "%d");
  PO(levels, "PARRSB_LEVELS", "%d");
  PO(find_disconnected_comps, "PARRSB_FIND_DISCONNECTED_COMPONENTS", "%d");
  PO(repair, "PARRSB_REPAIR", "%d");
  PO(verbose_level, "PARRSB_VERBOSE_LEVEL", "%d");
  PO(profile_level, "PARRSB_PROFILE_LEVEL", "%d");
  PO(rsb_algo, "PARRSB_RSB_ALGO", "%d");
  PO(rsb_pre, "PARRSB_RSB_PRE", "%d");
  PO(rsb_max_iter, "PARRSB_RSB_MAX_ITER", "%d");
  PO(rsb_max_passes, "PARRSB_RSB_MAX_PASSES", "%d");
  PO(rsb_tol, "PARRSB_RSB_TOL", "%lf");
  PO(rsb_mg_grammian, "PARRSB_RSB_MG_GRAMMIAN", "%d");
  PO(rsb_mg_factor, "PARRSB_RSB_MG_FACTOR", "%d");
  PO(nparts, "PARRSB_NPARTS", "%d");
#undef PO
  fflush(stdout);
}

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int device_index(void) {
  const char *v = getenv("PARRSB_AMD_DEVICE");
  return v ? atoi(v) : 0;
}

static int precision_bits(void) {
  const char *v = getenv("PARRSB_AMD_PRECISION");
  if (v) return atoi(v) == 32 ? 32 : 64;
  return cfg_precision;
}

/* Warnings of check_rsb_partition (rsb.c:61-111). */
static void check_partition(const parrsb_amd_report *r, const parrsb_options *o) {
  for (int l = 0; l < r->nlevels && l < PARRSB_AMD_MAX_LEVELS; l++) {
    if (r->nconverged_fail[l] > 0 && o->rsb_algo == 1)
      fprintf(stderr, "Warning: Inverse iteration didn't converge after %d iterations in "
                      "Level = %d\n",
              o->rsb_max_passes, l);
    else if (r->nconverged_fail[l] > 0)
      fprintf(stderr,
              "Warning: Lanczos reached a residual of %lf (target: %lf) after %d x %d "
              "iterations in Level=%d!\n",
              r->tol_final[l], r->tol_target[l], o->rsb_max_passes, o->rsb_max_iter, l);
    if (o->find_disconnected_comps == 1 && r->ncomp_max[l] > 1)
      fprintf(stderr,
              "Warning: Partition created %d/%d (min/max) disconnected components in "
              "Level=%d!\n",
              r->ncomp_min[l], r->ncomp_max[l], l);
  }
  fflush(stderr);
}

static void print_profile(const parrsb_amd_report *r) {
  /* the reference's timer names (metrics.h:10-27): the level Laplacian is the
     Fiedler setup, the solve its calculation */
  printf("phase ms: setup %.3f RSB_PRE %.3f RSB_FIEDLER_SETUP %.3f RSB_FIEDLER_CALC %.3f "
         "RSB_SORT %.3f RSB_COMPONENTS %.3f RSB_BALANCE %.3f\n",
         r->phase_ms[0], r->phase_ms[1], r->phase_ms[2], r->phase_ms[3], r->phase_ms[5],
         r->phase_ms[6], r->phase_ms[7]);
  for (int l = 0; l < r->nlevels && l < PARRSB_AMD_MAX_LEVELS; l++)
    printf("level=%02d segments=%d RSB_FIEDLER_CALC_NITER=%d/%d RSB_COMPONENTS_NCOMP=%d/%d "
           "RSB_NEIGHBORS=%g/%d/%d lambda=%.6e/%.6e\n",
           l, r->nsegs[l], r->niter_min[l], r->niter_max[l], r->ncomp_min[l], r->ncomp_max[l],
           r->nbrs_mean[l], r->nbrs_min[l], r->nbrs_max[l], r->lambda_min[l], r->lambda_max[l]);
  fflush(stdout);
}

/* Shared body of the host- and device-pointer entry points. */
static int run(int *part, int *seq, const long long *vtx, const double *xyz, const int *tag,
               int nel, int nv, parrsb_options *options, int nparts, int on_device) {
  double t0 = now();
  const int verbose = options->verbose_level, profile = options->profile_level;
  memset(&last_report, 0, sizeof(last_report));
  last_error[0] = 0;
  if (vtx == NULL && xyz == NULL) {
    set_error("both vertices and coordinates are NULL");
    return 4;
  }
  if (nv != 4 && nv != 8) {
    set_error("nv must be 4 or 8");
    return 3;
  }
  if (options->partitioner < 0 || options->partitioner > 2) {
    set_error("partitioner must be 0 (RSB), 1 (RCB) or 2 (RIB)");
    return 3;
  }
  if (options->partitioner != 0 && xyz == NULL) {
    set_error("RCB and RIB need coordinates");
    return 3;
  }
  if (options->partitioner == 0 && options->rsb_algo != 0 && options->rsb_algo != 1) {
    set_error("rsb_algo must be 0 (Lanczos) or 1 (MG inverse iteration)");
    return 3;
  }
  if (options->partitioner == 0 && options->rsb_algo == 1 && options->rsb_mg_factor < 2) {
    set_error("MG inverse iteration needs rsb_mg_factor > 1");
    return 3;
  }
  /* the refinement is on for rsb_mg_grammian == 1 only (fiedler.c:250) */
  if (options->partitioner == 0 && options->rsb_algo == 1 && options->rsb_mg_grammian == 1 &&
      options->rsb_max_passes > 64) {
    set_error("rsb_mg_grammian = 1 takes at most 64 passes (rsb_max_passes = %d)",
              options->rsb_max_passes);
    return 3;
  }
  if (vtx == NULL && options->partitioner == 0) {
    set_error("RSB needs vertex ids");
    return 3;
  }
  if (xyz == NULL) options->rsb_pre = 0; /* parrsb.c:250 */
  if (options->levels > 2) options->levels = 2; /* parrsb.c:228-229 */
  if (nparts < 1) {
    set_error("number of parts must be >= 1");
    return 3;
  }
  if (nel == 0) {
    /* nothing to partition: no device work and no output (in tagged mode the
       reference would count 0 tags and divide by that, parrsb.c:436) */
    if (tag) options->verbose_level = 0, options->profile_level = 0;
    last_report.nparts = nparts, last_report.precision = precision_bits();
    last_report.seconds = now() - t0;
    return 0;
  }
  const char *wf = getenv("PARRSB_AMD_WARM_START");
  if (wf && *wf) {
    int wrc = load_warm_file(wf);
    if (wrc) return wrc;
  }
  rsbgpu *g = NULL;
  int rc = rsbgpu_open(&g, device_index());
  if (rc) {
    set_error("no usable GPU (device %d)", device_index());
    return 2;
  }
  rsbgpu_set_profiling(g, cfg_profiling);
  rc = rsbgpu_set_mesh(g, vtx, xyz, nel, nv, on_device);
  parrsb_amd_report *rep = &last_report;
  rep->nparts = nparts, rep->nel = nel, rep->precision = precision_bits();
  /* tagged partitioning (parrsb.c:379-561): each distinct tag is partitioned
     on its own into nparts / ntags parts; tag i (ascending) owns parts
     [i * nparts / ntags, (i + 1) * nparts / ntags) */
  int ngroups = 0, *gcount = NULL;
  if (!rc && tag) {
    gcount = (int *)calloc((size_t)nparts, sizeof(int));
    if (!gcount) rc = 2;
    if (!rc) rc = rsbgpu_group_by_tag(g, tag, on_device, nparts, &ngroups, gcount);
    if (!rc && verbose > 0) printf("Num tags: %d\n", ngroups), fflush(stdout);
    if (!rc && (ngroups > nparts || nparts % ngroups != 0)) {
      free(gcount);
      rsbgpu_close(g);
      set_error("number of parts must be a multiple of number of tags: parts = %d, tags = %d",
                nparts, ngroups);
      return 3;
    }
    if (!rc && verbose > 0) printf("Processes per tag: %d\n", nparts / ngroups), fflush(stdout);
    /* the reference zeroes both on the caller's struct (parrsb.c:535-536) */
    options->verbose_level = 0, options->profile_level = 0;
  }
  if (!rc) {
    int64_t nnz = 0, ws = 0;
    rsbgpu_graph_stats(g, &nnz, &ws);
    rep->graph_nnz = nnz, rep->graph_wsum = ws;
    if (options->partitioner == 1 || options->partitioner == 2) {
      rc = rcb_run(g, nel, nparts, options->partitioner == 1 ? 1 : 2, ngroups, gcount);
    } else {
      rsb_params prm;
      prm.nparts = nparts, prm.rsb_pre = options->rsb_pre;
      prm.max_iter = options->rsb_max_iter, prm.max_passes = options->rsb_max_passes;
      prm.find_comps = options->find_disconnected_comps, prm.precision = precision_bits();
      prm.verbose = options->verbose_level, prm.tol = options->rsb_tol;
      prm.residuals = 1;
      prm.warm = NULL, prm.warm_passes = cfg_warm_passes, prm.warm_max_n = cfg_warm_max_n;
      prm.ngroups = ngroups, prm.group_count = gcount;
      prm.rsb_algo = options->rsb_algo, prm.mg_factor = options->rsb_mg_factor;
      prm.mg_grammian = options->rsb_mg_grammian == 1;
      void *dw = NULL;
      if (cfg_warm) {
        rc = rsbgpu_upload(g, cfg_warm, sizeof(float) * (size_t)rsbgpu_lws_nparams(), &dw);
        prm.warm = (const float *)dw;
        rep->warm_start = 1;
      }
      if (!rc) rc = rsb_run(g, nel, &prm, rep);
      if (!rc) rc = rsbgpu_write_part(g, part, on_device);
      if (!rc && seq) rc = rsbgpu_write_seq(g, seq, on_device);
      rsbgpu_free(dw);
    }
  }
  if (!rc && options->partitioner != 0) rc = rsbgpu_write_part(g, part, on_device);
  if (!rc && options->partitioner != 0 && seq) rc = rsbgpu_write_seq(g, seq, on_device);
  if (rc) set_error("%s", rsbgpu_error(g));
  rsbgpu_phase_times(g, rep->phase_ms);
  double tr[6];
  rsbgpu_lanczos_traffic(g, tr);
  rep->lanczos_bytes = tr[0], rep->lanczos_ms = tr[1], rep->lanczos_iters = tr[2];
  rep->spmv_bytes = tr[3], rep->spmv_ms = tr[4], rep->spmv_launches = tr[5];
  rsbgpu_close(g);
  free(gcount);
  rep->seconds = now() - t0;
  if (rc) return rc;
  if (options->partitioner == 0) check_partition(rep, options);
  if (verbose > 0) {
    static const char *algo[3] = {"RSB", "RCB", "RIB"};
    printf("par%s finished in %g seconds.\n", algo[options->partitioner % 3], rep->seconds);
    fflush(stdout);
  }
  if (profile > 0 && !tag) print_profile(rep);
  return 0;
}

static int part_mesh_host(int *part, int *seq, const long long *const vtx,
                          const double *const xyz, const int *const tag, const int nel,
                          const int nv, parrsb_options *const options, MPI_Comm comm) {
  update_options(options);
  int np = 1;
  if (MPI_Initialized && MPI_Comm_size) {
    int init = 0;
    MPI_Initialized(&init);
    if (init) MPI_Comm_size(comm, &np);
  }
  if (options->verbose_level > 0) {
    printf("Running parRSB ..., nv = %d, nelg = %d\n", nv, nel);
    fflush(stdout);
  }
  print_options(options);
  if (options->tagged == 1 && !tag) { /* parrsb.c:873-877 */
    set_error("tagged partitioning requested but tag array is NULL");
    return 1;
  }
  if (np > 1) {
    set_error("all elements must be passed from one rank (communicator size %d)", np);
    return 3;
  }
  int nparts = options->nparts > 0 ? options->nparts : np;
  return run(part, seq, vtx, xyz, options->tagged == 1 ? tag : NULL, nel, nv, options, nparts,
             0);
}

int parrsb_part_mesh(int *part, const long long *const vtx, const double *const xyz,
                     const int *const tag, const int nel, const int nv,
                     parrsb_options *const options, MPI_Comm comm) {
  return part_mesh_host(part, NULL, vtx, xyz, tag, nel, nv, options, comm);
}

/* part + seq form of README.md:46-66: seq[i] = order of element i within
 * part[i] after partitioning (its position in the final in-part order). */
int parrsb_part_mesh_seq(int *part, int *seq, const long long *vtx, const double *xyz,
                         const int *tag, int nel, int nv, parrsb_options *options,
                         MPI_Comm comm) {
  if (!seq) {
    set_error("seq is NULL");
    return 3;
  }
  return part_mesh_host(part, seq, vtx, xyz, tag, nel, nv, options, comm);
}

int parrsb_amd_part_mesh_dev(int *d_part, const long long *d_vtx, const double *d_xyz,
                             int nel, int nv, parrsb_options *options) {
  update_options(options);
  if (options->tagged == 1) {
    set_error("tagged partitioning takes host arrays (parrsb_part_mesh)");
    return 3;
  }
  return run(d_part, NULL, d_vtx, d_xyz, NULL, nel, nv, options, options->nparts, 1);
}

int parrsb_amd_part_mesh_dev_noalloc(int *d_part, const long long *d_vtx, const double *d_xyz,
                                     int nel, int nv, parrsb_options *options) {
  update_options(options);
  memset(&last_report, 0, sizeof(last_report));
  if (options->tagged == 1) {
    set_error("tagged partitioning takes host arrays (parrsb_part_mesh)");
    return 3;
  }
  if (nel < 0 || nel > PARRSB_AMD_NOALLOC_MAX_NEL) {
    set_error("allocation-free partitioning takes at most %d elements (got %d)",
              PARRSB_AMD_NOALLOC_MAX_NEL, nel);
    return 3;
  }
  if (options->partitioner == 0 && options->rsb_algo == 1 &&
      nel > PARRSB_AMD_NOALLOC_MAX_NEL_MG) {
    set_error("allocation-free partitioning with rsb_algo = 1 takes at most %d elements "
              "(got %d)",
              PARRSB_AMD_NOALLOC_MAX_NEL_MG, nel);
    return 3;
  }
  if (options->partitioner == 0 && options->rsb_algo == 1 && options->rsb_mg_grammian == 1) {
    /* its Gram matrices take rsb_max_passes^2 doubles per segment */
    set_error("allocation-free partitioning does not take rsb_mg_grammian = 1");
    return 3;
  }
  if (options->nparts > PARRSB_AMD_NOALLOC_MAX_NEL) {
    set_error("allocation-free partitioning makes at most %d parts (got %d)",
              PARRSB_AMD_NOALLOC_MAX_NEL, options->nparts);
    return 3;
  }
  char msg[256] = {0};
  int rc = rsbgpu_noalloc_begin(device_index(), msg, (int)sizeof(msg));
  if (rc) {
    set_error("%s", msg);
    return rc;
  }
  rc = run(d_part, NULL, d_vtx, d_xyz, NULL, nel, nv, options, options->nparts, 1);
  long long high = 0;
  last_report.noalloc_bytes = (double)rsbgpu_noalloc_end(&high);
  last_report.noalloc_high = (double)high;
  return rc;
}

int parrsb_conn_mesh(long long *vtx, double *coord, unsigned nel, unsigned nDim,
                     long long *periodicInfo, int nPeriodicFaces, double tol, MPI_Comm comm) {
  last_error[0] = 0;
  int verbose = 1;
  const char *val = getenv("PARRSB_VERBOSE_LEVEL");
  if (val) verbose = atoi(val);
  int np = 1;
  if (MPI_Initialized && MPI_Comm_size) {
    int init = 0;
    MPI_Initialized(&init);
    if (init) MPI_Comm_size(comm, &np);
  }
  if (np > 1) {
    set_error("all elements must be passed from one rank (communicator size %d)", np);
    return 3;
  }
  if (nPeriodicFaces < 0 || (nPeriodicFaces > 0 && !periodicInfo)) {
    set_error("periodicInfo must hold nPeriodicFaces records");
    return 3;
  }
  if (nDim != 2 && nDim != 3) {
    set_error("nDim must be 2 or 3");
    return 3;
  }
  if (nel > 0 && (!vtx || !coord)) {
    set_error("vtx and coord are required");
    return 3;
  }
  if (verbose > 0) printf("Running parCon ...\n"), fflush(stdout);
  double t0 = now();
  char msg[256] = {0};
  int rc = rsbgpu_conn_mesh(device_index(), vtx, coord, (int)nel, (int)nDim, tol, periodicInfo,
                            nPeriodicFaces, 0, msg,
                            (int)sizeof(msg));
  if (rc == 1 && msg[0]) set_error("%s", msg);
  else if (rc == 1) set_error("connectivity check failed (repeated vertex in an element, or a "
                              "face sharing exactly three corners with another element)");
  else if (rc) set_error("%s", msg[0] ? msg : "connectivity failed");
  if (verbose > 0 && rc == 0)
    printf("parCon (tol = %e) finished in %g s\n", tol, now() - t0), fflush(stdout);
  return rc;
}

void fparrsb_conn_mesh_(long long *vtx, double *coord, int *nel, int *nDim,
                        long long *periodicInfo, int *nPeriodicFaces, double *tol,
                        MPI_Fint *fcomm, int *err) {
  *err = 1;
  if (*nel < 0) {
    set_error("nel must be >= 0");
    *err = 3;
    return;
  }
  *err = parrsb_conn_mesh(vtx, coord, (unsigned)*nel, (unsigned)*nDim, periodicInfo,
                          *nPeriodicFaces, *tol, MPI_Comm_f2c(*fcomm));
}

void fparrsb_conn_mesh(long long *vtx, double *coord, int *nel, int *nDim,
                       long long *periodicInfo, int *nPeriodicFaces, double *tol,
                       MPI_Fint *fcomm, int *err) {
  fparrsb_conn_mesh_(vtx, coord, nel, nDim, periodicInfo, nPeriodicFaces, tol, fcomm, err);
}

int parrsb_amd_conn_mesh_dev(long long *d_vtx, const double *d_coord, int nel, int ndim,
                             double tol) {
  last_error[0] = 0;
  char msg[256] = {0};
  int rc = rsbgpu_conn_mesh(device_index(), d_vtx, d_coord, nel, ndim, tol, NULL, 0, 1, msg,
                            (int)sizeof(msg));
  if (rc == 1) set_error("connectivity check failed");
  else if (rc) set_error("%s", msg[0] ? msg : "connectivity failed");
  return rc;
}

static int quality(const int *part, const long long *vtx, int nel, int nv, int nparts,
                   int64_t *out, int on_device) {
  rsbgpu *g = NULL;
  if (rsbgpu_open(&g, device_index())) return 2;
  int rc = rsbgpu_set_mesh(g, vtx, NULL, nel, nv, on_device);
  const int *dp = part;
  int *tmp = NULL;
  if (!rc && !on_device) {
    rc = rsbgpu_upload_ints(g, part, nel, &tmp);
    dp = tmp;
  }
  if (!rc) rc = rsbgpu_quality(g, dp, nparts, out);
  if (rc) set_error("%s", rsbgpu_error(g));
  if (tmp) rsbgpu_free(tmp);
  rsbgpu_close(g);
  return rc;
}

int parrsb_amd_partition_quality(const int *part, const long long *vtx, int nel, int nv,
                                 int nparts, int64_t *out) {
  return quality(part, vtx, nel, nv, nparts, out, 0);
}

int parrsb_amd_partition_quality_dev(const int *d_part, const long long *d_vtx, int nel,
                                     int nv, int nparts, int64_t *out) {
  return quality(d_part, d_vtx, nel, nv, nparts, out, 1);
}
