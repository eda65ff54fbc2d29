/*
 * parRSB.h -- drop-in public interface of the MI355X recursive spectral
 * bisection partitioner.
 *
 * Mirrors the reference interface src/parRSB.h:21-48:
 *   - parrsb_options: the same fields in the same order (parRSB.h:21-41),
 *     plus one appended field `nparts` (see below);
 *   - parrsb_default_options (parRSB.h:43, values from parrsb.c:26-44);
 *   - parrsb_part_mesh(part, vtx, xyz, tag, nel, nv, options, comm)
 *     (parRSB.h:45-48, behaviour of parrsb.c:853-906).
 *
 * Differences a caller can observe:
 *   - `nparts` (appended, so callers that fill fields by name keep working;
 *     callers compiled against the reference header must be recompiled since
 *     the struct grew by one int).  0 means "one part per rank of comm", the
 *     reference behaviour; P > 0 makes one rank produce P parts on one GPU.
 *     The environment variable PARRSB_NPARTS overrides it, like every other
 *     PARRSB_* variable (parrsb.c:48-77).
 *   - All element data of the call must live on one rank (the one that owns
 *     the GPU).  A communicator with more than one rank is rejected with
 *     return code 3.
 *   - A device (HIP) failure returns 2 instead of aborting.
 */
#ifndef PARRSB_AMD_PARRSB_H
#define PARRSB_AMD_PARRSB_H

#if defined(PARRSB_NO_MPI)
/* Callers without MPI: MPICH's ABI (MPI_Comm is an int handle). */
typedef int MPI_Comm;
typedef int MPI_Fint;
#define MPI_Comm_f2c(comm) ((MPI_Comm)(comm))
#ifndef MPI_COMM_WORLD
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#define MPI_COMM_SELF ((MPI_Comm)0x44000001)
#endif
#else
#include <mpi.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* General options (parRSB.h:23-30) */
  int partitioner; /* 0 - RSB, 1 - RCB, 2 - RIB (default 0) */
  int tagged;      /* tagged partitioning: 0 - no, 1 - yes (default 0) */
  int levels;      /* 1 or 2 (default 2); clamped to <= 2 on return */
  int find_disconnected_comps; /* count components of each half (default 1) */
  int repair;        /* parsed, unused, as in the reference (default 0) */
  int verbose_level; /* default 1 */
  int profile_level; /* default 0 */
  /* RSB options (parRSB.h:32-37) */
  int rsb_algo;       /* 0 - Lanczos, 1 - MG-preconditioned inverse iteration */
  int rsb_pre;        /* 0 - none, 1 - RCB, 2 - RIB (default 1) */
  int rsb_max_iter;   /* Lanczos iterations per pass (default 50) */
  int rsb_max_passes; /* Lanczos restarts (default 50) */
  double rsb_tol;     /* CG residual reduction factor (default 1e-5) */
  int rsb_dump_stats; /* parsed, unused (default 0) */
  /* RSB MG options (parRSB.h:39-40), used with rsb_algo = 1 */
  int rsb_mg_grammian; /* Gram-matrix refinement of the iterates (default 0) */
  int rsb_mg_factor;   /* aggregation factor of the MG hierarchy, >= 2 (default 2) */
  /* Appended: number of parts; 0 => size of comm (reference behaviour). */
  int nparts;
} parrsb_options;

extern parrsb_options parrsb_default_options;

/*
 * Partition nel elements with nv vertices each (nv = 8: hexahedra in 3D,
 * nv = 4: quadrilaterals in 2D) into P parts.
 *   part[nel]            out: part id in [0, P) for each input element
 *   vtx[nel*nv]          global vertex ids (non-zero), element-major
 *   xyz[nel*nv*ndim]     per-vertex coordinates (ndim = 3 if nv == 8 else 2),
 *                        or NULL (then options->rsb_pre is set to 0)
 *   tag[nel]             only used when options->tagged == 1: the elements
 *                        of each distinct tag are partitioned on their own
 *                        into P / ntags parts, the i-th smallest tag taking
 *                        parts [i * P / ntags, (i + 1) * P / ntags)
 *                        (parrsb.c:379-561)
 *   options              read and MUTATED: PARRSB_* environment overrides,
 *                        levels clamp, rsb_pre = 0 when xyz == NULL,
 *                        verbose_level = profile_level = 0 in tagged mode
 * Returns 0 on success, 1 if tagged == 1 and tag == NULL (parrsb.c:873-877),
 * 2 on a device failure, 3 on unsupported input (see header comment; also
 * P not a multiple of the number of tags, where the reference exits,
 * parrsb.c:436-444), 4 if both vtx and xyz are NULL (the reference aborts,
 * parrsb.c:244-249).
 */
int parrsb_part_mesh(int *part, const long long *const vtx,
                     const double *const xyz, const int *const tag,
                     const int nel, const int nv, parrsb_options *const options,
                     MPI_Comm comm);

/*
 * Global vertex ids from corner coordinates (parRSB.h:62-64 of the
 * reference, con.c:189-295): coord[nel][nv][nDim] in the preprocessor corner
 * order (nv = 8 for nDim = 3, 4 for nDim = 2) -> vtx[nel][nv], 1-based ids,
 * symmetric corner order.  Corners closer than tol times their element's
 * shortest incident edge share an id.  periodicInfo holds nPeriodicFaces
 * records (element, face, partner element, partner face), 1-based,
 * preprocessor face numbers, as parrsb_read_mesh returns them: matched faces
 * share their corner ids (con-periodic.c) and the ids are renumbered densely.
 * Returns 0, 1 if the element or face check of the result fails
 * (con-check.c) or a periodic face does not match its partner (the
 * reference exits there), 2 on a device failure, 3 for a communicator of
 * more than one rank, bad sizes or a periodic record out of range.
 */
int parrsb_conn_mesh(long long *vtx, double *coord, unsigned nel, unsigned nDim,
                     long long *periodicInfo, int nPeriodicFaces, double tol,
                     MPI_Comm comm);

/*
 * Fortran binding of parrsb_conn_mesh (parRSB.h:66-70, con.c:297-307 of the
 * reference), the entry Nek5000 calls: every argument by reference, the
 * communicator as a Fortran handle, the return code in *err.  Exported under
 * the usual Fortran manglings: fparrsb_conn_mesh_ (gfortran and most
 * compilers, gslib's default) and fparrsb_conn_mesh.
 */
void fparrsb_conn_mesh_(long long *vtx, double *coord, int *nel, int *nDim,
                        long long *periodicInfo, int *nPeriodicFaces, double *tol,
                        MPI_Fint *fcomm, int *err);
void fparrsb_conn_mesh(long long *vtx, double *coord, int *nel, int *nDim,
                       long long *periodicInfo, int *nPeriodicFaces, double *tol,
                       MPI_Fint *fcomm, int *err);

/*
 * Nek5000 mesh files (parRSB.h:71-83 of the reference, io.c), read and
 * written by the one rank that holds the mesh (a larger communicator
 * returns 1).
 *   parrsb_read_mesh  read & 1: <name>.re2 -> *nel, *nv, *coord (malloc'd,
 *                     [nel][nv][ndim] preprocessor order), *nbcs periodic
 *                     face records in *bcs (malloc'd, 4 per record as
 *                     parrsb_conn_mesh takes them); read & 2: <name>.co2 ->
 *                     *vl (malloc'd vertex ids; nel and nv must agree with
 *                     the .re2 when both are read).  Returns 0, or 1 with a
 *                     message (parrsb_amd_last_error) on a missing, truncated
 *                     or foreign-endian file, where the reference exits.
 *   parrsb_dump_con   <name>.co2 from vl[nelt][nv].
 *   parrsb_dump_map   <name>.ma2 from vl[nelt][nv], rank field 0 (the
 *                     calling rank; see parrsb_amd_dump_map_parts for one
 *                     rank holding all parts).
 */
int parrsb_read_mesh(unsigned *nel, unsigned *nv, long long **vl, double **coord,
                     unsigned *nbcs, long long **bcs, char *name, MPI_Comm comm, int read);
int parrsb_dump_con(char *name, unsigned nelt, unsigned nv, long long *vl, MPI_Comm comm);
int parrsb_dump_map(char *name, unsigned nelt, unsigned nv, long long *vl, MPI_Comm comm);

/* Command-line options of the example drivers (parRSB.h:88-99 of the
 * reference): --mesh=<name> (required), --tol=<gencon tolerance, 0.2>,
 * --test, --dump, --nactive=<ranks>, --verbose=<level>.  NULL when --mesh is
 * missing or an option is unknown. */
typedef struct {
  char *mesh;
  double tol;
  int test;
  int dump;
  int nactive;
  int verbose;
} parrsb_cmd_line_opts;

parrsb_cmd_line_opts *parrsb_parse_cmd_opts(int argc, char *argv[]);
void parrsb_cmd_opts_free(parrsb_cmd_line_opts *opts);

/*
 * Redistribution after partitioning (helpers.c:35-92).  One rank holds all
 * parts here, so the elements -- vl, coord (may be NULL) and part itself --
 * are reordered by part id, ascending, keeping their order within a part:
 * the concatenation, in rank order, of what each rank receives in the
 * reference.  *nelt is unchanged.  Returns 1 on a negative part id.
 */
int parrsb_dist_mesh(unsigned *nelt, long long **vl, double **coord, int *part, unsigned nv,
                     MPI_Comm comm);

/* Read <opts->mesh>.re2, find the connectivity, partition with the default
 * options (PARRSB_NPARTS sets the part count) and redistribute
 * (helpers.c:94-130). */
int parrsb_setup_mesh(unsigned *nelt, unsigned *nv, long long **vl, double **coord,
                      parrsb_cmd_line_opts *opts, MPI_Comm comm);

#ifdef __cplusplus
}
#endif

#endif
