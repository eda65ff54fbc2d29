// Neighbouring segments (statistics.c:17-113, recorded as RSB_NEIGHBORS at
// rsb.c:342-343): for every current segment, the number of other segments
// that share a vertex with it.  Vertex sharing is the dual graph's relation,
// so the pairs come from the graph rows: each element emits the distinct
// foreign segments among its neighbours (most elements emit none), the
// (segment, foreign segment) pairs are sorted, and distinct pairs are counted
// per segment.
#include "state.hpp"

// Distinct segments t != s among the row's neighbours; emit = false counts.
template <bool EMIT>
__global__ void k_nb_pairs(int nel, const int64_t *__restrict__ growptr,
                           const uint32_t *__restrict__ gadj, const int *__restrict__ segof,
                           int *__restrict__ cnt, const int64_t *__restrict__ off,
                           uint64_t *__restrict__ keys) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  const int s = segof[e];
  const int64_t r0 = growptr[e], r1 = growptr[e + 1];
  int c = 0;
  int64_t o = EMIT ? off[e] : 0;
  for (int64_t k = r0; k < r1; k++) {
    const int t = segof[gadj[k] & 0x0fffffffu];
    if (t == s) continue;
    bool seen = false;
    for (int64_t j = r0; j < k && !seen; j++) seen = segof[gadj[j] & 0x0fffffffu] == t;
    if (seen) continue;
    if (EMIT) keys[o + c] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)t;
    c++;
  }
  if (!EMIT) cnt[e] = c;
}

__global__ void k_nb_count(int64_t n, const uint64_t *__restrict__ keys, int *__restrict__ nb) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0 || keys[i] != keys[i - 1]) atomicAdd(&nb[keys[i] >> 32], 1);
}

__global__ void k_widen(int n, const int *__restrict__ in, int64_t *__restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

extern "C" int rsbgpu_neighbors(rsbgpu *g, int *nbrs) {
  RSB_TRY(hipSetDevice(g->device));
  const int nel = g->nel, nseg = g->nseg;
  for (int s = 0; s < nseg; s++) nbrs[s] = 0;
  if (nel == 0 || nseg <= 1) return 0;
  hipStream_t st = g->stream;
  int *cnt = nullptr, *nb = nullptr;
  int64_t *off = nullptr, *wide = nullptr;
  RSB_TRY(hipMalloc((void **)&cnt, sizeof(int) * (size_t)nel));
  RSB_TRY(hipMalloc((void **)&off, sizeof(int64_t) * ((size_t)nel + 1)));
  RSB_TRY(hipMalloc((void **)&wide, sizeof(int64_t) * (size_t)nel));
  RSB_TRY(hipMalloc((void **)&nb, sizeof(int) * (size_t)nseg));
  k_nb_pairs<false><<<nblocks(nel, 256), 256, 0, st>>>(nel, g->d_growptr, g->d_gadj, g->d_segof,
                                                       cnt, nullptr, nullptr);
  RSB_LAUNCH_CHECK();
  // 64-bit offsets: widen the counts and scan them
  k_widen<<<nblocks(nel, 256), 256, 0, st>>>(nel, cnt, wide);
  RSB_LAUNCH_CHECK();
  int rc = rsb_exclusive_scan_i64(g, wide, off, nel);
  if (rc) return rc;
  int64_t last_off = 0;
  int last_cnt = 0;
  RSB_TRY(hipMemcpyAsync(&last_off, off + nel - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RSB_TRY(hipMemcpyAsync(&last_cnt, cnt + nel - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  RSB_TRY(hipStreamSynchronize(st));
  const int64_t npairs = last_off + last_cnt;
  RSB_TRY(hipMemsetAsync(nb, 0, sizeof(int) * (size_t)nseg, st));
  if (npairs > 0) {
    if (npairs >= (int64_t)1 << 31) RSB_FAIL(2, "too many neighbour pairs (%lld)", (long long)npairs);
    uint64_t *keys = nullptr;
    int *vals = nullptr;
    RSB_TRY(hipMalloc((void **)&keys, sizeof(uint64_t) * (size_t)npairs));
    RSB_TRY(hipMalloc((void **)&vals, sizeof(int) * (size_t)npairs));
    k_nb_pairs<true><<<nblocks(nel, 256), 256, 0, st>>>(nel, g->d_growptr, g->d_gadj,
                                                        g->d_segof, nullptr, off, keys);
    RSB_LAUNCH_CHECK();
    int sb = 1;
    while ((1ll << sb) < nseg) sb++;
    if ((rc = rsb_sort_pairs_u64(g, keys, vals, npairs, 0, 32 + sb))) return rc;
    k_nb_count<<<nblocks(npairs, 256), 256, 0, st>>>(npairs, keys, nb);
    RSB_LAUNCH_CHECK();
    RSB_TRY(hipStreamSynchronize(st));
    hipFree(keys), hipFree(vals);
  }
  RSB_TRY(hipMemcpyAsync(nbrs, nb, sizeof(int) * (size_t)nseg, hipMemcpyDeviceToHost, st));
  RSB_TRY(hipStreamSynchronize(st));
  hipFree(cnt), hipFree(off), hipFree(wide), hipFree(nb);
  return 0;
}
