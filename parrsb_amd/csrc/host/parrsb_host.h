/* Internal declarations of the C host layer. */
#ifndef PARRSB_AMD_HOST_H
#define PARRSB_AMD_HOST_H

#include "parrsb_amd.h"
#include "rsbgpu.h"

typedef struct {
  int nparts, rsb_pre, max_iter, max_passes, find_comps, precision, verbose, residuals;
  int rsb_algo;  /* 0 Lanczos, 1 MG-preconditioned inverse iteration */
  int mg_factor; /* coarsening factor of the MG hierarchy (rsb_algo = 1) */
  int mg_grammian; /* 1 = Gram-matrix refinement of the MG start vectors */
  double tol;
  const float *warm;  /* device parameters of the learned start, or NULL */
  int warm_passes, warm_max_n;
  /* tagged partitioning: ngroups > 1 groups of group_count[i] consecutive
     elements (ascending tag), each partitioned into nparts / ngroups parts */
  int ngroups;
  const int *group_count;
} rsb_params;

/* last-error message of the library (parrsb_amd_last_error) */
void parrsb_host_error(const char *fmt, ...);
/* size of comm if the process has MPI loaded and initialised, else 1 */
int parrsb_host_comm_size(MPI_Comm comm);

long long rsb_lower_count(long long n, int p);
long long rsb_balance_expected(long long n, int p, int bin);
int rsb_ncuts(int nparts);
int rsb_run(rsbgpu *g, int nel, const rsb_params *prm, parrsb_amd_report *rep);
/* partitioner 1 (pre = 1, RCB) or 2 (pre = 2, RIB); ngroups / group_count
   as in rsb_params (ngroups <= 1: one group of nel elements) */
int rcb_run(rsbgpu *g, int nel, int nparts, int pre, int ngroups, const int *group_count);

#endif
