/*
 * Nek5000 mesh files and the helpers around them (the reference's io.c and
 * helpers.c), for a single rank that holds the whole mesh:
 *   parrsb_read_mesh   .re2 geometry + periodic faces, .co2 connectivity
 *                      (io.c:40-337)
 *   parrsb_dump_con    .co2 writer (io.c:339-400)
 *   parrsb_dump_map    .ma2 writer (io.c:402-470)
 *   parrsb_dist_mesh   redistribution by part (helpers.c:35-92)
 *   parrsb_setup_mesh  read + connectivity + partition + redistribution
 *                      (helpers.c:94-130)
 *   command-line options of helpers.c:245-315.
 *
 * File layouts (all little-endian, as the byte-test float checks):
 *   .re2  80-byte text header "#v00x nelgt ndim nelgv ...", float 6.54321,
 *         per element one double (group) then the corner coordinates as
 *         doubles, x[nv] y[nv] (z[nv]), preprocessor corner order; then
 *         ncurves (double) and 8 doubles per curve; then nbcs (double) and
 *         per boundary face 8 doubles: element, face, 5 parameters and the
 *         boundary type in the first 3 bytes of the last one.  A face of type
 *         "P  " is periodic; its first two parameters name the partner
 *         element and face.
 *   .co2  132-byte header "%5s%12d%12d%12d" (version, nelgt, nelgv, nv),
 *         float 6.54321, per element an int (1-based element id) and nv ints
 *         (vertex ids).
 *   .ma2  132-byte header "%5s%12d%12d%12d%12d%12d%12d%12d" (version, nelgt,
 *         nactive, depth, 2^depth, nelgt*nv, nrank, noutflow), float
 *         6.54321, per element an int (rank) and nv ints (vertex ids).
 */
#define _POSIX_C_SOURCE 200809L
#include <getopt.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "parrsb_host.h"

#define RE2_HEADER 80
#define CO2_HEADER 132
#define BYTE_TEST 6.54321f

static int byte_test_ok(float t) { return fabs((double)t - 6.543210) <= 1e-7; }

static char *with_suffix(const char *name, const char *suffix) {
  size_t n = strlen(name), m = strlen(suffix);
  char *s = (char *)malloc(n + m + 1);
  if (s) memcpy(s, name, n), memcpy(s + n, suffix, m + 1);
  return s;
}

/* .re2: header, corner coordinates, periodic boundary faces. */
static int read_re2(const char *fname, unsigned *nel, unsigned *nv, double **coord,
                    unsigned *nbcs, long long **bcs) {
  FILE *f = fopen(fname, "rb");
  if (!f) {
    parrsb_host_error("cannot open %s", fname);
    return 1;
  }
  int rc = 1;
  char hdr[RE2_HEADER + 1] = {0};
  char version[6] = {0};
  long long nelgt = 0, nelgv = 0;
  int ndim = 0;
  float bt = 0.0f;
  double *buf = NULL, *xyz = NULL;
  long long *rec = NULL;
  if (fread(hdr, 1, RE2_HEADER, f) != RE2_HEADER ||
      sscanf(hdr, "%5s %lld %d %lld", version, &nelgt, &ndim, &nelgv) != 4 ||
      (ndim != 2 && ndim != 3) || nelgt < 0 || nelgt > INT_MAX) {
    parrsb_host_error("%s: not a .re2 header", fname);
    goto done;
  }
  if (fread(&bt, sizeof(float), 1, f) != 1 || !byte_test_ok(bt)) {
    parrsb_host_error("%s: byte test failed", fname);
    goto done;
  }
  const int v = ndim == 3 ? 8 : 4;
  const size_t per = 1 + (size_t)v * ndim; /* doubles per element record */
  buf = (double *)malloc(sizeof(double) * per * (size_t)(nelgt ? nelgt : 1));
  xyz = (double *)malloc(sizeof(double) * (size_t)v * ndim * (size_t)(nelgt ? nelgt : 1));
  if (!buf || !xyz) {
    parrsb_host_error("out of host memory");
    goto done;
  }
  if (fread(buf, sizeof(double) * per, (size_t)nelgt, f) != (size_t)nelgt) {
    parrsb_host_error("%s: truncated element data", fname);
    goto done;
  }
  for (long long e = 0; e < nelgt; e++) {
    const double *r = buf + (size_t)e * per + 1; /* skip the group id */
    for (int k = 0; k < v; k++)
      for (int d = 0; d < ndim; d++) xyz[((size_t)e * v + k) * ndim + d] = r[d * v + k];
  }
  double t = 0.0;
  if (fread(&t, sizeof(double), 1, f) != 1 || t < 0) {
    parrsb_host_error("%s: missing curve count", fname);
    goto done;
  }
  if (fseek(f, (long)(8 * sizeof(double)) * (long)t, SEEK_CUR) != 0 ||
      fread(&t, sizeof(double), 1, f) != 1 || t < 0) {
    parrsb_host_error("%s: missing boundary count", fname);
    goto done;
  }
  const long long nb = (long long)t;
  long long np = 0;
  rec = (long long *)malloc(sizeof(long long) * 4 * (size_t)(nb ? nb : 1));
  if (!rec) {
    parrsb_host_error("out of host memory");
    goto done;
  }
  for (long long i = 0; i < nb; i++) {
    double r[8];
    if (fread(r, sizeof(double), 8, f) != 8) {
      parrsb_host_error("%s: truncated boundary data", fname);
      goto done;
    }
    char cbc[4] = {0};
    memcpy(cbc, &r[7], 3);
    if (strcmp(cbc, "P  ") == 0) {
      rec[4 * np + 0] = (long long)r[0], rec[4 * np + 1] = (long long)r[1];
      rec[4 * np + 2] = (long long)r[2], rec[4 * np + 3] = (long long)r[3];
      np++;
    }
  }
  *nel = (unsigned)nelgt, *nv = (unsigned)v;
  if (coord) *coord = xyz, xyz = NULL;
  if (nbcs) *nbcs = (unsigned)np;
  if (bcs) *bcs = rec, rec = NULL;
  rc = 0;
done:
  fclose(f);
  free(buf), free(xyz), free(rec);
  return rc;
}

/* .co2: vertex ids per element. */
static int read_co2(const char *fname, unsigned *nel, unsigned *nv, long long **vl) {
  FILE *f = fopen(fname, "rb");
  if (!f) {
    parrsb_host_error("cannot open %s", fname);
    return 1;
  }
  int rc = 1;
  char hdr[CO2_HEADER + 1] = {0};
  char version[6] = {0};
  long long nelgt = 0, nelgv = 0;
  unsigned v = 0;
  float bt = 0.0f;
  int *buf = NULL;
  long long *out = NULL;
  if (fread(hdr, 1, CO2_HEADER, f) != CO2_HEADER ||
      sscanf(hdr, "%5s %12lld %12lld %u", version, &nelgt, &nelgv, &v) != 4 ||
      (v != 4 && v != 8) || nelgt < 0 || nelgt > INT_MAX) {
    parrsb_host_error("%s: not a .co2 header", fname);
    goto done;
  }
  if (*nv != 0 && *nv != v) {
    parrsb_host_error("nv values don't match: %u %u", *nv, v);
    goto done;
  }
  if (*nel != 0 && *nel != (unsigned)nelgt) {
    parrsb_host_error("nelt values don't match: %u %lld", *nel, nelgt);
    goto done;
  }
  if (fread(&bt, sizeof(float), 1, f) != 1 || !byte_test_ok(bt)) {
    parrsb_host_error("%s: byte test failed", fname);
    goto done;
  }
  buf = (int *)malloc(sizeof(int) * (v + 1) * (size_t)(nelgt ? nelgt : 1));
  out = (long long *)malloc(sizeof(long long) * v * (size_t)(nelgt ? nelgt : 1));
  if (!buf || !out) {
    parrsb_host_error("out of host memory");
    goto done;
  }
  if (fread(buf, sizeof(int) * (v + 1), (size_t)nelgt, f) != (size_t)nelgt) {
    parrsb_host_error("%s: truncated connectivity", fname);
    goto done;
  }
  for (long long e = 0; e < nelgt; e++)
    for (unsigned k = 0; k < v; k++) out[(size_t)e * v + k] = buf[(size_t)e * (v + 1) + 1 + k];
  *nel = (unsigned)nelgt, *nv = v;
  if (vl) *vl = out, out = NULL;
  rc = 0;
done:
  fclose(f);
  free(buf), free(out);
  return rc;
}

int parrsb_read_mesh(unsigned *nel, unsigned *nv, long long **vl, double **coord,
                     unsigned *nbcs, long long **bcs, char *name, MPI_Comm comm, int read) {
  if (!nel || !nv || !name) {
    parrsb_host_error("nel, nv and name are required");
    return 1;
  }
  if (parrsb_host_comm_size(comm) > 1) {
    parrsb_host_error("the mesh is read by one rank (communicator size %d)",
                      parrsb_host_comm_size(comm));
    return 1;
  }
  *nv = 0, *nel = 0;
  int rc = 0;
  if (read & 1) {
    char *fn = with_suffix(name, ".re2");
    rc = fn ? read_re2(fn, nel, nv, coord, nbcs, bcs) : 1;
    free(fn);
    if (rc) return rc;
  }
  if (read & 2) {
    char *fn = with_suffix(name, ".co2");
    rc = fn ? read_co2(fn, nel, nv, vl) : 1;
    free(fn);
  }
  return rc;
}

static int write_file(const char *fname, const char *hdr, unsigned nelt, unsigned nv,
                      const long long *vl, const int *first) {
  FILE *f = fopen(fname, "wb");
  if (!f) {
    parrsb_host_error("cannot open %s for writing", fname);
    return 1;
  }
  const float bt = BYTE_TEST;
  int ok = fwrite(hdr, 1, CO2_HEADER, f) == CO2_HEADER && fwrite(&bt, sizeof(float), 1, f) == 1;
  int rec[9];
  for (unsigned e = 0; e < nelt && ok; e++) {
    rec[0] = first[e];
    for (unsigned k = 0; k < nv; k++) rec[1 + k] = (int)vl[(size_t)e * nv + k];
    ok = fwrite(rec, sizeof(int), nv + 1, f) == nv + 1;
  }
  ok &= fclose(f) == 0;
  if (!ok) parrsb_host_error("write to %s failed", fname);
  return ok ? 0 : 1;
}

static void pad_header(char *hdr) {
  size_t n = strlen(hdr);
  memset(hdr + n, ' ', CO2_HEADER - n);
}

int parrsb_dump_con(char *name, unsigned nelt, unsigned nv, long long *vl, MPI_Comm comm) {
  if (parrsb_host_comm_size(comm) > 1 || (nv != 4 && nv != 8) || !name) {
    parrsb_host_error("dump_con: one rank, nv 4 or 8 and a name are required");
    return 1;
  }
  char hdr[CO2_HEADER + 1];
  snprintf(hdr, sizeof(hdr), "%5s%12d%12d%12d", "#v001", (int)nelt, (int)nelt, (int)nv);
  pad_header(hdr);
  int *ids = (int *)malloc(sizeof(int) * (nelt ? nelt : 1));
  char *fn = with_suffix(name, ".co2");
  int rc = 1;
  if (ids && fn) {
    for (unsigned e = 0; e < nelt; e++) ids[e] = (int)e + 1;
    rc = write_file(fn, hdr, nelt, nv, vl, ids);
  }
  free(ids), free(fn);
  return rc;
}

static int floor_log2(long long n) {
  int k = 0;
  while (n > 1) n /= 2, k++;
  return k;
}

/* .ma2 whose rank field is rank[e] (NULL: 0, the calling rank). */
static int dump_map(const char *name, unsigned nelt, unsigned nv, const long long *vl,
                    const int *rank, int nrank) {
  if ((nv != 4 && nv != 8) || !name) {
    parrsb_host_error("dump_map: nv 4 or 8 and a name are required");
    return 1;
  }
  const int depth = floor_log2(nelt);
  char hdr[CO2_HEADER + 1];
  snprintf(hdr, sizeof(hdr), "%5s%12d%12d%12d%12d%12d%12d%12d", "#v001", (int)nelt, (int)nelt,
           depth, 1 << depth, (int)(nelt * nv), nrank, 0);
  pad_header(hdr);
  int *r = (int *)calloc(nelt ? nelt : 1, sizeof(int));
  char *fn = with_suffix(name, ".ma2");
  int rc = 1;
  if (r && fn) {
    if (rank) memcpy(r, rank, sizeof(int) * nelt);
    rc = write_file(fn, hdr, nelt, nv, vl, r);
  }
  free(r), free(fn);
  return rc;
}

int parrsb_dump_map(char *name, unsigned nelt, unsigned nv, long long *vl, MPI_Comm comm) {
  if (parrsb_host_comm_size(comm) > 1) {
    parrsb_host_error("dump_map: one rank");
    return 1;
  }
  /* the reference writes nrnk = nelgt (io.c:416) */
  return dump_map(name, nelt, nv, vl, NULL, (int)nelt);
}

void parrsb_amd_free(void *p) { free(p); }

int parrsb_amd_dump_map_parts(const char *name, unsigned nelt, unsigned nv, const long long *vl,
                              const int *part) {
  if (!part) {
    parrsb_host_error("part is required");
    return 1;
  }
  return dump_map(name, nelt, nv, vl, part, (int)nelt);
}

/* One rank holds all parts: elements (with their vertices, coordinates and
 * part ids) are reordered by part, ascending, keeping their order within a
 * part -- the concatenation, in rank order, of what each rank receives in the
 * reference's transfer. */
int parrsb_dist_mesh(unsigned *nelt_, long long **vl_, double **coord_, int *part, unsigned nv,
                     MPI_Comm comm) {
  if (parrsb_host_comm_size(comm) > 1) {
    parrsb_host_error("dist_mesh: one rank");
    return 1;
  }
  if (!nelt_ || !vl_ || !part || (nv != 4 && nv != 8)) {
    parrsb_host_error("dist_mesh: nelt, vl, part and nv (4 or 8) are required");
    return 1;
  }
  const unsigned n = *nelt_, ndim = nv == 8 ? 3 : 2;
  int pmax = -1;
  for (unsigned e = 0; e < n; e++) {
    if (part[e] < 0) {
      parrsb_host_error("dist_mesh: negative part id");
      return 1;
    }
    pmax = part[e] > pmax ? part[e] : pmax;
  }
  size_t *cnt = (size_t *)calloc((size_t)pmax + 2, sizeof(size_t));
  unsigned *perm = (unsigned *)malloc(sizeof(unsigned) * (n ? n : 1));
  long long *v2 = (long long *)malloc(sizeof(long long) * nv * (n ? n : 1));
  double *c2 = coord_ && *coord_ ? (double *)malloc(sizeof(double) * nv * ndim * (n ? n : 1))
                                 : NULL;
  int *p2 = (int *)malloc(sizeof(int) * (n ? n : 1));
  int rc = 1;
  if (!cnt || !perm || !v2 || !p2 || (coord_ && *coord_ && !c2)) {
    parrsb_host_error("out of host memory");
    goto done;
  }
  for (unsigned e = 0; e < n; e++) cnt[part[e] + 1]++;
  for (int q = 0; q <= pmax; q++) cnt[q + 1] += cnt[q];
  for (unsigned e = 0; e < n; e++) perm[cnt[part[e]]++] = e;
  for (unsigned i = 0; i < n; i++) {
    const unsigned e = perm[i];
    memcpy(v2 + (size_t)i * nv, *vl_ + (size_t)e * nv, sizeof(long long) * nv);
    if (c2)
      memcpy(c2 + (size_t)i * nv * ndim, *coord_ + (size_t)e * nv * ndim,
             sizeof(double) * nv * ndim);
    p2[i] = part[e];
  }
  memcpy(*vl_, v2, sizeof(long long) * nv * n);
  if (c2) memcpy(*coord_, c2, sizeof(double) * nv * ndim * n);
  memcpy(part, p2, sizeof(int) * n);
  rc = 0;
done:
  free(cnt), free(perm), free(v2), free(c2), free(p2);
  return rc;
}

int parrsb_setup_mesh(unsigned *nelt, unsigned *nv, long long **vl, double **coord,
                      parrsb_cmd_line_opts *opts, MPI_Comm comm) {
  if (!opts || !opts->mesh) {
    parrsb_host_error("setup_mesh: options with a mesh name are required");
    return 1;
  }
  unsigned nbcs = 0;
  long long *bcs = NULL;
  int rc = parrsb_read_mesh(nelt, nv, NULL, coord, &nbcs, &bcs, opts->mesh, comm, 1);
  if (rc) return rc;
  *vl = (long long *)calloc((size_t)*nelt * *nv + 1, sizeof(long long));
  int *part = (int *)calloc((size_t)*nelt + 1, sizeof(int));
  if (!*vl || !part) {
    parrsb_host_error("out of host memory");
    free(bcs), free(part);
    return 1;
  }
  const unsigned ndim = *nv == 8 ? 3 : 2;
  rc = parrsb_conn_mesh(*vl, *coord, *nelt, ndim, bcs, (int)nbcs, opts->tol, comm);
  if (!rc) {
    parrsb_options o = parrsb_default_options;
    rc = parrsb_part_mesh(part, *vl, *coord, NULL, (int)*nelt, (int)*nv, &o, comm);
  }
  if (!rc) rc = parrsb_dist_mesh(nelt, vl, coord, part, *nv, comm);
  free(part), free(bcs);
  return rc;
}

parrsb_cmd_line_opts *parrsb_parse_cmd_opts(int argc, char *argv[]) {
  parrsb_cmd_line_opts *in = (parrsb_cmd_line_opts *)calloc(1, sizeof(parrsb_cmd_line_opts));
  if (!in) return NULL;
  in->mesh = NULL, in->tol = 0.2, in->test = 0, in->dump = 0, in->verbose = 0;
  in->nactive = INT_MAX;
  static struct option longopts[] = {{"mesh", required_argument, 0, 'm'},
                                     {"tol", optional_argument, 0, 't'},
                                     {"test", optional_argument, 0, 'T'},
                                     {"dump", optional_argument, 0, 'd'},
                                     {"nactive", optional_argument, 0, 'n'},
                                     {"verbose", optional_argument, 0, 'v'},
                                     {"help", optional_argument, 0, 'h'},
                                     {0, 0, 0, 0}};
  optind = 1;
  for (;;) {
    const int c = getopt_long(argc, argv, "", longopts, NULL);
    if (c == -1) break;
    switch (c) {
    case 'm':
      free(in->mesh);
      in->mesh = strdup(optarg);
      break;
    case 't': in->tol = optarg ? atof(optarg) : in->tol; break;
    case 'T': in->test = 1; break;
    case 'd': in->dump = 1; break;
    case 'n': in->nactive = optarg ? atoi(optarg) : in->nactive; break;
    case 'v': in->verbose = optarg ? atoi(optarg) : in->verbose; break;
    case 'h': break;
    default: parrsb_cmd_opts_free(in); return NULL;
    }
  }
  if (!in->mesh) {
    fprintf(stderr, "Required argument `--mesh` was not found !\n");
    parrsb_cmd_opts_free(in);
    return NULL;
  }
  return in;
}

void parrsb_cmd_opts_free(parrsb_cmd_line_opts *opts) {
  if (opts) free(opts->mesh), free(opts);
}
