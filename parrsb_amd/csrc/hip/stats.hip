// Neighbouring segments (statistics.c:17-113, recorded as RSB_NEIGHBORS at
// rsb.c:342-343): for every current segment, the number of other segments
// that share a vertex with it.  Vertex sharing is the dual graph's relation,
// so the pairs come from the graph rows: each element emits the distinct
// foreign segments among its neighbours (most elements emit none), the
// (segment, foreign segment) pairs are sorted, and distinct pairs are counted
// per segment.
#include <algorithm>

#include "state.hpp"

// Distinct segments t != s among the row's neighbours, appended as
// (s << 32 | t) keys at an atomically taken offset (nothing is written past
// `cap`; the caller compares the total with it).  The distinct values found
// so far (at most NB_FOUND, the usual case by far) live in registers; a row
// with more re-checks against its own prefix.
#define NB_FOUND 8
__global__ void k_nb_pairs(int nel, const int64_t *__restrict__ growptr,
                           const uint32_t *__restrict__ gadj, const int *__restrict__ segof,
                           unsigned long long *__restrict__ total, int64_t cap,
                           uint64_t *__restrict__ keys) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  const int s = segof[e];
  const int64_t r0 = growptr[e], r1 = growptr[e + 1];
  int found[NB_FOUND];
  int nf = 0;
  bool many = false;
  // eight entries and their segments in flight, then examined in order
  for (int64_t k0 = r0; k0 < r1 && !many; k0 += 8) {
    int tt[8];
#pragma unroll
    for (int u = 0; u < 8; u++) tt[u] = k0 + u < r1 ? (int)(gadj[k0 + u] & 0x0fffffffu) : e;
#pragma unroll
    for (int u = 0; u < 8; u++) tt[u] = segof[tt[u]];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int t = tt[u];
      if (t == s || many) continue;  // padding entries are e itself: t == s
      bool seen = false;
#pragma unroll
      for (int j = 0; j < NB_FOUND; j++) seen |= j < nf && found[j] == t;
      if (seen) continue;
      if (nf < NB_FOUND) {
#pragma unroll
        for (int j = 0; j < NB_FOUND; j++)
          if (j == nf) found[j] = t;
        nf++;
      } else {
        many = true;
      }
    }
  }
  if (nf == 0) return;
  if (!many) {
    const unsigned long long o = atomicAdd(total, (unsigned long long)nf);
#pragma unroll
    for (int j = 0; j < NB_FOUND; j++)
      if (j < nf && (int64_t)(o + j) < cap) keys[o + j] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)found[j];
    return;
  }
  // more than NB_FOUND foreign segments: count, then write, against the prefix
  unsigned long long o = 0;
  for (int pass = 0; pass < 2; pass++) {
    int c = 0;
    for (int64_t k = r0; k < r1; k++) {
      const int t = segof[gadj[k] & 0x0fffffffu];
      if (t == s) continue;
      bool seen = false;
      for (int64_t j = r0; j < k && !seen; j++) seen = segof[gadj[j] & 0x0fffffffu] == t;
      if (seen) continue;
      if (pass == 1 && (int64_t)(o + c) < cap)
        keys[o + c] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)t;
      c++;
    }
    if (pass == 0) o = atomicAdd(total, (unsigned long long)c);
  }
}

__global__ void k_nb_count(int64_t n, const uint64_t *__restrict__ keys, int *__restrict__ nb) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0 || keys[i] != keys[i - 1]) atomicAdd(&nb[keys[i] >> 32], 1);
}

extern "C" int rsbgpu_neighbors(rsbgpu *g, int *nbrs) {
  RSB_TRY(hipSetDevice(g->device));
  const int nel = g->nel, nseg = g->nseg;
  for (int s = 0; s < nseg; s++) nbrs[s] = 0;
  if (nel == 0 || nseg <= 1) return 0;
  hipStream_t st = g->stream;
  // one pass appends into a buffer of nel / 2 keys; if it overflows (or an
  // element has more than NB_CACHE foreign segments) the exact count sizes
  // a second pass
  int64_t cap = std::max<int64_t>(nel / 2, 1024);
  unsigned long long *total = nullptr;
  uint64_t *keys = nullptr;
  int *vals = nullptr, *nb = nullptr;
  RSB_TRY(RSB_SCRATCH(total, sizeof(unsigned long long)));
  RSB_TRY(RSB_SCRATCH(nb, sizeof(int) * (size_t)nseg));
  unsigned long long npairs = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    RSB_TRY(RSB_SCRATCH(keys, sizeof(uint64_t) * (size_t)cap));
    RSB_TRY(hipMemsetAsync(total, 0, sizeof(unsigned long long), st));
    k_nb_pairs<<<nblocks(nel, 256), 256, 0, st>>>(nel, g->d_growptr, g->d_gadj,
                                                        g->d_segof, total, cap, keys);
    RSB_LAUNCH_CHECK();
    RSB_TRY(hipMemcpyAsync(&npairs, total, sizeof(npairs), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipStreamSynchronize(st));
    if ((int64_t)npairs <= cap) break;
    if (attempt == 1 || npairs >= (1ull << 31))
      RSB_FAIL(2, "too many neighbour pairs (%llu)", (unsigned long long)npairs);
    cap = (int64_t)npairs;
  }
  RSB_TRY(hipMemsetAsync(nb, 0, sizeof(int) * (size_t)nseg, st));
  if (npairs > 0) {
    RSB_TRY(RSB_SCRATCH(vals, sizeof(int) * (size_t)npairs));
    int sb = 1;
    while ((1ll << sb) < nseg) sb++;
    int rc = rsb_sort_pairs_u64(g, keys, vals, (int64_t)npairs, 0, 32 + sb);
    if (rc) return rc;
    k_nb_count<<<nblocks((int64_t)npairs, 256), 256, 0, st>>>((int64_t)npairs, keys, nb);
    RSB_LAUNCH_CHECK();
  }
  RSB_TRY(hipMemcpyAsync(nbrs, nb, sizeof(int) * (size_t)nseg, hipMemcpyDeviceToHost, st));
  RSB_TRY(hipStreamSynchronize(st));
  return 0;
}
