// Fiedler sort, bisection, component counts, balance, part output and
// partition quality.
#include <algorithm>
#include <vector>

#include "state.hpp"

// ---------------------------------------------------------------------------
// Fiedler sort (rsb.c:308): stable sort of every segment with p > 1 by its
// Fiedler values; other segments keep their order.  Two stable passes: by
// value (inactive positions get one common key and so keep their order),
// then by segment.
__global__ void k_fkeys(const double *__restrict__ fied, const int *__restrict__ order,
                        const int *__restrict__ segof, const int *__restrict__ seg, int nseg,
                        int nel, uint64_t *__restrict__ key, int *__restrict__ val) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int s = segof[order[i]];
  key[i] = seg[2 * nseg + s] > 1 ? dkey(fied[i]) : 0;
  val[i] = i;
}

__global__ void k_segkeys(const int *__restrict__ perm, const int *__restrict__ order,
                          const int *__restrict__ segof, int nel, uint64_t *__restrict__ key) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nel) return;
  key[j] = (uint64_t)segof[order[perm[j]]];
}

__global__ void k_permute(const int *__restrict__ perm, const int *__restrict__ order,
                          const double *__restrict__ fied, int nel, int *__restrict__ order2,
                          double *__restrict__ fied2) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nel) return;
  int i = perm[j];
  order2[j] = order[i];
  fied2[j] = fied[i];
}

extern "C" int rsbgpu_fiedler_sort(rsbgpu *g) {
  RSB_TRY(hipSetDevice(g->device));
  const int nel = g->nel;
  if (nel == 0) return 0;
  rsb_phase_begin(g);
  uint64_t *key = nullptr;
  int *perm = nullptr;
  double *fied2 = nullptr;
  RSB_TRY(RSB_SCRATCH(key, sizeof(uint64_t) * (nel + 1)));
  RSB_TRY(RSB_SCRATCH(perm, sizeof(int) * (nel + 1)));
  RSB_TRY(RSB_SCRATCH(fied2, sizeof(double) * (nel + 1)));
  k_fkeys<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_fied, g->d_order, g->d_segof, g->d_seg,
                                                    g->nseg, nel, key, perm);
  RSB_LAUNCH_CHECK();
  int rc = rsb_sort_pairs_u64(g, key, perm, nel, 0, 64);
  if (rc) return rc;
  if (g->nseg > 1) {
    k_segkeys<<<nblocks(nel, 256), 256, 0, g->stream>>>(perm, g->d_order, g->d_segof, nel, key);
    RSB_LAUNCH_CHECK();
    int sb = 1;
    while ((1ll << sb) < g->nseg) sb++;
    rc = rsb_sort_pairs_u64(g, key, perm, nel, 0, sb);
    if (rc) return rc;
  }
  k_permute<<<nblocks(nel, 256), 256, 0, g->stream>>>(perm, g->d_order, g->d_fied, nel,
                                                      g->d_order2, fied2);
  RSB_LAUNCH_CHECK();
  std::swap(g->d_order, g->d_order2);
  RSB_TRY(hipMemcpyAsync(g->d_fied, fied2, sizeof(double) * nel, hipMemcpyDeviceToDevice,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  rsb_phase_end(g, PH_SORT);
  return 0;
}

// ---------------------------------------------------------------------------
// Half labels: hseg[e] = 2 s + (position >= nlower[s]) for elements of
// segments with p > 1, -1 otherwise.
__global__ void k_halves(const int *__restrict__ order, const int *__restrict__ segof,
                         const int *__restrict__ seg, int nseg, const int *__restrict__ nlower,
                         int nel, int *__restrict__ hseg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int e = order[i], s = segof[e];
  hseg[e] = seg[2 * nseg + s] > 1 ? 2 * s + ((i - seg[s]) >= nlower[s] ? 1 : 0) : -1;
}

// Lock-free union-find over level rows (parents always have smaller
// indices, so following parents ends at a root; hooking is a CAS on a root).
__device__ inline int uf_rep(int *par, int x) {
  int cur = __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (cur != x) {
    int prev = x, next;
    while (cur > (next = __hip_atomic_load(&par[cur], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT))) {
      __hip_atomic_store(&par[prev], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      prev = cur;
      cur = next;
    }
  }
  return cur;
}

__device__ inline void uf_union(int *par, int u, int v) {
  int ru = uf_rep(par, u), rv = uf_rep(par, v);
  while (ru != rv) {
    if (ru < rv) {
      int t = ru;
      ru = rv;
      rv = t;
    }
    int old = atomicCAS(&par[ru], ru, rv);
    if (old == ru) break;
    ru = uf_rep(par, old);
    rv = uf_rep(par, rv);
  }
}

__global__ void k_uf_init(int64_t mpad, int *__restrict__ par) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < mpad) par[i] = (int)i;
}

// Half of every row (-1 for padding rows), so the edge pass gathers one
// array instead of two in a chain.
__global__ void k_uf_hrow(int64_t mpad, const int *__restrict__ row_elem,
                          const int *__restrict__ hseg, int *__restrict__ hrow) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= mpad) return;
  const int e = row_elem[row];
  hrow[row] = e >= 0 ? hseg[e] : -1;
}

// Union-find in two steps.  Edges inside a tile (RSB_TILE consecutive rows
// of one segment, most edges of a level) are united in LDS by the tile's
// workgroup, which then writes every row's parent as the tile-local root (a
// set's root is its smallest row, so parents stay below their rows); only
// edges that leave the tile go through the global atomics.  One edge pass
// over the global forest took ~9 ms a level at 16.7M rows, most of it in
// agent-scope atomic loads of the parent chains.
__device__ inline int luf_rep(int *lp, int x) {
  int cur = __hip_atomic_load(&lp[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (cur != x) {
    const int nx = __hip_atomic_load(&lp[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (nx != cur)  // path splitting: each visited row's parent becomes its grandparent
      __hip_atomic_store(&lp[x], nx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = cur;
    cur = nx;
  }
  return cur;
}

__device__ inline void luf_union(int *lp, int u, int v) {
  int ru = luf_rep(lp, u), rv = luf_rep(lp, v);
  while (ru != rv) {
    if (ru < rv) {
      const int t = ru;
      ru = rv;
      rv = t;
    }
    int old = ru;
    __hip_atomic_compare_exchange_strong(&lp[ru], &old, rv, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == ru) break;
    ru = luf_rep(lp, old);
    rv = luf_rep(lp, rv);
  }
}

__global__ void __launch_bounds__(RSB_BLOCK) k_uf_tile(const int *__restrict__ tile,
                                                       const int *__restrict__ hrow,
                                                       const int64_t *__restrict__ slice_off,
                                                       const int *__restrict__ slice_w,
                                                       const uint32_t *__restrict__ ell,
                                                       int *__restrict__ par) {
  __shared__ int lp[RSB_TILE], lh[RSB_TILE];
  const int row0 = tile[3 * blockIdx.x + 1], row1 = tile[3 * blockIdx.x + 2];
  const int n = row1 - row0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    lp[i] = i;
    lh[i] = hrow[row0 + i];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int h = lh[i];
    if (h < 0) continue;
    const int row = row0 + i;
    const int64_t sl = row >> 6;
    const uint32_t *q = ell + slice_off[sl] + (row & 63);
    const int w = slice_w[sl];
    for (int j0 = 0; j0 < w; j0 += 8) {
      uint32_t a[8];
#pragma unroll
      for (int u = 0; u < 8; u++) a[u] = j0 + u < w ? q[(int64_t)(j0 + u) * RSB_WAVE] : 0u;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int c = (int)(a[u] & 0x0fffffffu);
        if ((a[u] >> 28) != 0 && c >= row0 && c < row && lh[c - row0] == h)
          luf_union(lp, i, c - row0);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) par[row0 + i] = row0 + luf_rep(lp, i);
}

// The edges that leave their tile (to a lower row of the same half).
__global__ void __launch_bounds__(RSB_BLOCK) k_uf_cross(const int *__restrict__ tile,
                                                        const int *__restrict__ hrow,
                                                        const int64_t *__restrict__ slice_off,
                                                        const int *__restrict__ slice_w,
                                                        const uint32_t *__restrict__ ell,
                                                        int *__restrict__ par) {
  const int row0 = tile[3 * blockIdx.x + 1], row1 = tile[3 * blockIdx.x + 2];
  for (int row = row0 + threadIdx.x; row < row1; row += blockDim.x) {
    const int h = hrow[row];
    if (h < 0) continue;
    const int64_t sl = row >> 6;
    const uint32_t *q = ell + slice_off[sl] + (row & 63);
    const int w = slice_w[sl];
    for (int j0 = 0; j0 < w; j0 += 8) {
      uint32_t a[8];
      int hc[8];
#pragma unroll
      for (int u = 0; u < 8; u++) a[u] = j0 + u < w ? q[(int64_t)(j0 + u) * RSB_WAVE] : 0u;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int c = (int)(a[u] & 0x0fffffffu);
        hc[u] = ((a[u] >> 28) != 0 && c < row0) ? hrow[c] : -2;
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (hc[u] == h) uf_union(par, row, (int)(a[u] & 0x0fffffffu));
    }
  }
}

__global__ void k_uf_count(int64_t mpad, const int *__restrict__ row_elem,
                           const int *__restrict__ hseg, const int *__restrict__ par,
                           int *__restrict__ ncomp) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= mpad) return;
  int e = row_elem[row];
  if (e < 0) return;
  if (par[row] == row) atomicAdd(&ncomp[hseg[e]], 1);
}

// Interface flag of the sending half (rsb.c:145-182): an element of the
// sender that shares a vertex with an element of the other half of the same
// segment.
__global__ void k_iface(const int *__restrict__ order, const int *__restrict__ segof,
                        const int *__restrict__ seg, int nseg, const int *__restrict__ nlower,
                        const int *__restrict__ sender, const int *__restrict__ hseg,
                        const int *__restrict__ elem_row, const int *__restrict__ row_elem,
                        const int64_t *__restrict__ slice_off, const int *__restrict__ slice_w,
                        const uint32_t *__restrict__ ell, int nel, int *__restrict__ flag,
                        int *__restrict__ nif) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int e = order[i], s = segof[e];
  int f = 0;
  if (seg[2 * nseg + s] > 1 && sender[s] >= 0) {
    int side = (i - seg[s]) >= nlower[s] ? 1 : 0;
    if (side == sender[s]) {
      int row = elem_row[e], h = hseg[e];
      int64_t sl = row >> 6;
      const uint32_t *q = ell + slice_off[sl] + (row & 63);
      for (int j = 0; j < slice_w[sl] && !f; j++) {
        uint32_t a = q[(int64_t)j * RSB_WAVE];
        if ((a >> 28) == 0) continue;
        if (hseg[row_elem[a & 0x0fffffffu]] != h) f = 1;
      }
    }
  }
  flag[i] = f;
  if (f) atomicAdd(&nif[s], 1);
}

// Pick the interface elements that move (rsb.c:167-200): the upper half
// offers its smallest Fiedler values first, the lower half its largest; ties
// by position.  C = exclusive prefix count of flags over positions.
__global__ void k_pick(const int *__restrict__ order, const int *__restrict__ segof,
                       const int *__restrict__ seg, int nseg, const int *__restrict__ nlower,
                       const int *__restrict__ sender, const int *__restrict__ scount,
                       const int *__restrict__ nif, const double *__restrict__ fied,
                       const int *__restrict__ flag, const int *__restrict__ C, int nel,
                       int *__restrict__ lower) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int e = order[i], s = segof[e];
  if (seg[2 * nseg + s] <= 1) {
    lower[i] = 0;
    return;
  }
  int cs = seg[s], n = seg[nseg + s], nl = nlower[s];
  int side = (i - cs) >= nl ? 1 : 0;
  int low = side == 0;
  int snd = sender[s];
  if (snd >= 0 && flag[i] && nif[s] >= scount[s]) {
    int rank;
    if (snd == 1) {
      rank = C[i] - C[cs + nl];
    } else {
      int lo = i, hi = i;
      while (lo > cs && fied[lo - 1] == fied[i]) lo--;
      while (hi + 1 < cs + nl && fied[hi + 1] == fied[i]) hi++;
      rank = (C[cs + nl] - C[hi + 1]) + (C[i] - C[lo]);
    }
    if (rank < scount[s]) low = !low;
  }
  (void)n;
  lower[i] = low;
}

__global__ void k_partition(const int *__restrict__ order, const double *__restrict__ fied,
                            const int *__restrict__ segof, const int *__restrict__ seg, int nseg,
                            const int *__restrict__ lower, const int *__restrict__ Lx, int nel,
                            int *__restrict__ order2, double *__restrict__ fied2,
                            int *__restrict__ nl_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int e = order[i], s = segof[e];
  int j = i;
  if (seg[2 * nseg + s] > 1) {
    int cs = seg[s], n = seg[nseg + s];
    int before = Lx[i] - Lx[cs], nl = Lx[cs + n] - Lx[cs];
    j = lower[i] ? cs + before : cs + nl + (i - cs - before);
    if (i == cs) nl_out[s] = nl;
  }
  order2[j] = e;
  fied2[j] = fied[i];
}

extern "C" int rsbgpu_bisect(rsbgpu *g, int *nlower, const int *sender, const int *scount,
                             int find_comps, int *ncomp, int *moved) {
  RSB_TRY(hipSetDevice(g->device));
  const int nel = g->nel, nseg = g->nseg;
  if (nel == 0) return 0;
  int *d_nl = nullptr, *d_snd = nullptr, *d_sc = nullptr, *d_nif = nullptr, *hseg = nullptr;
  int *d_ncomp = nullptr, *par = nullptr, *flag = nullptr, *C = nullptr, *low = nullptr;
  int *d_nlo = nullptr;
  double *fied2 = nullptr;
  RSB_TRY(RSB_SCRATCH(d_nl, sizeof(int) * nseg * 6));
  d_snd = d_nl + nseg, d_sc = d_nl + 2 * nseg, d_nif = d_nl + 3 * nseg, d_nlo = d_nl + 4 * nseg;
  RSB_TRY(RSB_SCRATCH(d_ncomp, sizeof(int) * 2 * nseg));
  RSB_TRY(RSB_SCRATCH(hseg, sizeof(int) * nel));
  RSB_TRY(hipMemcpyAsync(d_nl, nlower, sizeof(int) * nseg, hipMemcpyHostToDevice, g->stream));
  RSB_TRY(hipMemcpyAsync(d_snd, sender, sizeof(int) * nseg, hipMemcpyHostToDevice, g->stream));
  RSB_TRY(hipMemcpyAsync(d_sc, scount, sizeof(int) * nseg, hipMemcpyHostToDevice, g->stream));
  RSB_TRY(hipMemsetAsync(d_nif, 0, sizeof(int) * nseg, g->stream));
  RSB_TRY(hipMemsetAsync(d_ncomp, 0, sizeof(int) * 2 * nseg, g->stream));
  k_halves<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, g->d_segof, g->d_seg, nseg, d_nl,
                                                     nel, hseg);
  RSB_LAUNCH_CHECK();
  // components of each half before balancing (rsb.c:320-327)
  if (find_comps && g->mpad > 0) {
    rsb_phase_begin(g);
    RSB_TRY(RSB_SCRATCH(par, sizeof(int) * 2 * g->mpad));
    int *hrow = par + g->mpad;
    k_uf_hrow<<<nblocks(g->mpad, 256), 256, 0, g->stream>>>(g->mpad, g->d_row_elem, hseg, hrow);
    k_uf_tile<<<g->ntiles, RSB_BLOCK, 0, g->stream>>>(g->d_tile, hrow, g->d_slice_off,
                                                      g->d_slice_w, g->d_ell, par);
    k_uf_cross<<<g->ntiles, RSB_BLOCK, 0, g->stream>>>(g->d_tile, hrow, g->d_slice_off,
                                                       g->d_slice_w, g->d_ell, par);
    k_uf_count<<<nblocks(g->mpad, 256), 256, 0, g->stream>>>(g->mpad, g->d_row_elem, hseg, par,
                                                             d_ncomp);
    RSB_LAUNCH_CHECK();
    rsb_phase_end(g, PH_COMP);
  }
  rsb_phase_begin(g);
  RSB_TRY(RSB_SCRATCH(flag, sizeof(int) * (nel + 1)));
  RSB_TRY(RSB_SCRATCH(C, sizeof(int) * (nel + 1)));
  RSB_TRY(RSB_SCRATCH(low, sizeof(int) * (nel + 1)));
  RSB_TRY(RSB_SCRATCH(fied2, sizeof(double) * (nel + 1)));
  RSB_TRY(hipMemsetAsync(flag + nel, 0, sizeof(int), g->stream));
  RSB_TRY(hipMemsetAsync(low + nel, 0, sizeof(int), g->stream));
  k_iface<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, g->d_segof, g->d_seg, nseg, d_nl,
                                                    d_snd, hseg, g->d_elem_row, g->d_row_elem,
                                                    g->d_slice_off, g->d_slice_w, g->d_ell, nel,
                                                    flag, d_nif);
  RSB_LAUNCH_CHECK();
  int rc = rsb_exclusive_scan_i32(g, flag, C, nel + 1);
  if (rc) return rc;
  k_pick<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, g->d_segof, g->d_seg, nseg, d_nl,
                                                   d_snd, d_sc, d_nif, g->d_fied, flag, C, nel,
                                                   low);
  RSB_LAUNCH_CHECK();
  rc = rsb_exclusive_scan_i32(g, low, C, nel + 1);
  if (rc) return rc;
  RSB_TRY(hipMemcpyAsync(d_nlo, d_nl, sizeof(int) * nseg, hipMemcpyDeviceToDevice, g->stream));
  k_partition<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, g->d_fied, g->d_segof,
                                                        g->d_seg, nseg, low, C, nel,
                                                        g->d_order2, fied2, d_nlo);
  RSB_LAUNCH_CHECK();
  std::swap(g->d_order, g->d_order2);
  RSB_TRY(hipMemcpyAsync(g->d_fied, fied2, sizeof(double) * nel, hipMemcpyDeviceToDevice,
                         g->stream));
  std::vector<int> nlo(nseg), nif(nseg);
  RSB_TRY(hipMemcpyAsync(nlo.data(), d_nlo, sizeof(int) * nseg, hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipMemcpyAsync(nif.data(), d_nif, sizeof(int) * nseg, hipMemcpyDeviceToHost,
                         g->stream));
  if (ncomp)
    RSB_TRY(hipMemcpyAsync(ncomp, d_ncomp, sizeof(int) * 2 * nseg, hipMemcpyDeviceToHost,
                           g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  rsb_phase_end(g, PH_BAL);
  for (int s = 0; s < nseg; s++) {
    int mv = 0;
    if (g->h_p[s] > 1) {
      mv = (sender[s] >= 0 && scount[s] > 0 && nif[s] >= scount[s]);
      nlower[s] = nlo[s];
    }
    if (moved) moved[s] = mv;
  }
  return 0;
}

// ---------------------------------------------------------------------------
__global__ void k_write_part(const int *__restrict__ segof, const int *__restrict__ seg,
                             int nseg, int nel, int *__restrict__ part) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  part[e] = seg[3 * nseg + segof[e]];
}

extern "C" int rsbgpu_write_part(rsbgpu *g, int *part, int part_is_device) {
  RSB_TRY(hipSetDevice(g->device));
  const int nel = g->nel;
  if (nel == 0) return 0;
  int *dp = part;
  if (!part_is_device) RSB_TRY(rsb_dev_malloc((void **)&dp, sizeof(int) * nel));
  k_write_part<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_segof, g->d_seg, g->nseg, nel, dp);
  RSB_LAUNCH_CHECK();
  if (!part_is_device) {
    RSB_TRY(hipMemcpyAsync(part, dp, sizeof(int) * nel, hipMemcpyDeviceToHost, g->stream));
    RSB_TRY(hipStreamSynchronize(g->stream));
    rsb_dev_free(dp);
  } else {
    RSB_TRY(hipStreamSynchronize(g->stream));
  }
  return 0;
}

// seq[e] = position of element e within its final segment (its part), in the
// final element order.
__global__ void k_write_seq(const int *__restrict__ order, const int *__restrict__ segof,
                            const int *__restrict__ seg, int nel, int *__restrict__ seq) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  const int e = order[i];
  seq[e] = i - seg[segof[e]];
}

extern "C" int rsbgpu_write_seq(rsbgpu *g, int *seq, int seq_is_device) {
  RSB_TRY(hipSetDevice(g->device));
  const int nel = g->nel;
  if (nel == 0) return 0;
  int *dp = seq;
  if (!seq_is_device) RSB_TRY(rsb_dev_malloc((void **)&dp, sizeof(int) * nel));
  k_write_seq<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, g->d_segof, g->d_seg, nel, dp);
  RSB_LAUNCH_CHECK();
  if (!seq_is_device) {
    RSB_TRY(hipMemcpyAsync(seq, dp, sizeof(int) * nel, hipMemcpyDeviceToHost, g->stream));
    RSB_TRY(hipStreamSynchronize(g->stream));
    rsb_dev_free(dp);
  } else {
    RSB_TRY(hipStreamSynchronize(g->stream));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Tagged partitioning (parrsb.c:379-561): the elements of each distinct tag
// become one group, groups in ascending order of the tag read as an unsigned
// 32-bit value (the reference keeps it in a uint field and sorts that,
// parrsb.c:388-402, so negative tags come after every non-negative one),
// elements of a group in their current order.  The tag's bits are the key of
// a stable radix sort of the element order; group heads are the positions
// where the sorted key changes.
__global__ void k_tag_keys(int nel, const int *__restrict__ order, const int *__restrict__ tag,
                           uint32_t *__restrict__ key) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nel) key[i] = (uint32_t)tag[order[i]];
}

__global__ void k_tag_heads(int nel, const uint32_t *__restrict__ key, int cap,
                            int *__restrict__ nheads, int *__restrict__ heads) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel || (i > 0 && key[i] == key[i - 1])) return;
  const int o = atomicAdd(nheads, 1);
  if (o < cap) heads[o] = i;
}

extern "C" int rsbgpu_group_by_tag(rsbgpu *g, const int *tag, int tag_is_device, int max_groups,
                                   int *ngroups, int *counts) {
  RSB_TRY(hipSetDevice(g->device));
  const int nel = g->nel;
  *ngroups = 0;
  if (nel == 0) return 0;
  if (max_groups < 1) RSB_FAIL(3, "max_groups must be >= 1");
  hipStream_t st = g->stream;
  const int *dt = tag;
  int *tmp = nullptr, *heads = nullptr, *nh = nullptr;
  uint32_t *key = nullptr;
  if (!tag_is_device) {
    RSB_TRY(rsb_dev_malloc((void **)&tmp, sizeof(int) * (size_t)nel));
    RSB_TRY(hipMemcpyAsync(tmp, tag, sizeof(int) * (size_t)nel, hipMemcpyHostToDevice, st));
    dt = tmp;
  }
  RSB_TRY(rsb_dev_malloc((void **)&key, sizeof(uint32_t) * (size_t)nel));
  RSB_TRY(rsb_dev_malloc((void **)&heads, sizeof(int) * (size_t)max_groups));
  RSB_TRY(rsb_dev_malloc((void **)&nh, sizeof(int)));
  k_tag_keys<<<nblocks(nel, 256), 256, 0, st>>>(nel, g->d_order, dt, key);
  RSB_LAUNCH_CHECK();
  int rc = rsb_sort_pairs_u32(g, key, g->d_order, nel, 0, 32);
  if (rc) return rc;
  RSB_TRY(hipMemsetAsync(nh, 0, sizeof(int), st));
  k_tag_heads<<<nblocks(nel, 256), 256, 0, st>>>(nel, key, max_groups, nh, heads);
  RSB_LAUNCH_CHECK();
  int n = 0;
  RSB_TRY(hipMemcpyAsync(&n, nh, sizeof(int), hipMemcpyDeviceToHost, st));
  RSB_TRY(hipStreamSynchronize(st));
  *ngroups = n;
  if (n <= max_groups) {
    std::vector<int> h((size_t)n);
    RSB_TRY(hipMemcpy(h.data(), heads, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
    std::sort(h.begin(), h.end());
    for (int i = 0; i < n; i++) counts[i] = (i + 1 < n ? h[i + 1] : nel) - h[i];
  }
  rsb_dev_free(tmp), rsb_dev_free(key), rsb_dev_free(heads), rsb_dev_free(nh);
  return 0;
}

// ---------------------------------------------------------------------------
// Partition quality on the global dual graph.
// Part ids outside [0, nparts) are reported (bad) and left out.
__global__ void k_q_check(int nel, const int *__restrict__ part, int nparts,
                          int *__restrict__ bad) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nel && (part[e] < 0 || part[e] >= nparts)) atomicOr(bad, 1);
}

__global__ void k_q_edges(int nel, const int64_t *__restrict__ rowptr,
                          const uint32_t *__restrict__ adj, const int *__restrict__ part,
                          unsigned long long *__restrict__ cut, int *__restrict__ par,
                          int *__restrict__ size) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long w = 0, c = 0;
  if (e < nel) {
    int pe = part[e];
    atomicAdd(&size[pe], 1);
    for (int64_t k = rowptr[e]; k < rowptr[e + 1]; k++) {
      int j = (int)(adj[k] & 0x0fffffffu);
      if (j <= e) continue;
      if (part[j] != pe) w += adj[k] >> 28, c++;
      else uf_union(par, j, e);
    }
  }
  w = wave_sum(w);
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) atomicAdd(&cut[0], w), atomicAdd(&cut[1], c);
}

__global__ void k_q_roots(int nel, const int *__restrict__ part, const int *__restrict__ par,
                          int *__restrict__ ncomp) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nel && par[e] == e) atomicAdd(&ncomp[part[e]], 1);
}

__global__ void k_q_vkeys(const long long *__restrict__ vtx, const int *__restrict__ part,
                          int64_t ns, int nv, uint64_t *__restrict__ key, int *__restrict__ val) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  key[s] = ((uint64_t)vtx[s] << 24) | (uint32_t)part[s / nv];
  val[s] = 0;
}

__global__ void k_q_volume(const uint64_t *__restrict__ key, int64_t ns,
                           unsigned long long *__restrict__ vol) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (i < ns && (i == 0 || (key[i] >> 24) != (key[i - 1] >> 24)) && (key[i] >> 24) != 0) {
    unsigned long long k = 0;
    for (int64_t j = i; j < ns && (key[j] >> 24) == (key[i] >> 24); j++)
      if (j == i || key[j] != key[j - 1]) k++;
    v = k * (k - 1);
  }
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) atomicAdd(vol, v);
}

extern "C" int rsbgpu_quality(rsbgpu *g, const int *d_part, int nparts, int64_t *out) {
  RSB_TRY(hipSetDevice(g->device));
  const int nel = g->nel;
  for (int i = 0; i < 8; i++) out[i] = 0;
  if (nel == 0) return 0;
  if (nparts < 1 || nparts >= (1 << 24)) RSB_FAIL(3, "nparts out of range for the quality report");
  if (!g->d_vtx) RSB_FAIL(3, "quality needs vertex ids");
  {
    int *bad = nullptr, hb = 0;
    RSB_TRY(rsb_dev_malloc((void **)&bad, sizeof(int)));
    RSB_TRY(hipMemsetAsync(bad, 0, sizeof(int), g->stream));
    k_q_check<<<nblocks(nel, 256), 256, 0, g->stream>>>(nel, d_part, nparts, bad);
    RSB_LAUNCH_CHECK();
    RSB_TRY(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    RSB_TRY(hipStreamSynchronize(g->stream));
    rsb_dev_free(bad);
    if (hb) RSB_FAIL(3, "part ids outside [0, %d)", nparts);
  }
  unsigned long long *acc = nullptr;
  int *par = nullptr, *size = nullptr, *ncomp = nullptr;
  RSB_TRY(rsb_dev_malloc((void **)&acc, sizeof(unsigned long long) * 4));
  RSB_TRY(rsb_dev_malloc((void **)&par, sizeof(int) * nel));
  RSB_TRY(rsb_dev_malloc((void **)&size, sizeof(int) * nparts * 2));
  ncomp = size + nparts;
  RSB_TRY(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * 4, g->stream));
  RSB_TRY(hipMemsetAsync(size, 0, sizeof(int) * nparts * 2, g->stream));
  k_uf_init<<<nblocks(nel, 256), 256, 0, g->stream>>>(nel, par);
  k_q_edges<<<nblocks(nel, 256), 256, 0, g->stream>>>(nel, g->d_growptr, g->d_gadj, d_part, acc,
                                                      par, size);
  k_q_roots<<<nblocks(nel, 256), 256, 0, g->stream>>>(nel, d_part, par, ncomp);
  RSB_LAUNCH_CHECK();
  int64_t ns = (int64_t)nel * g->nv;
  uint64_t *key = nullptr;
  int *val = nullptr;
  RSB_TRY(rsb_dev_malloc((void **)&key, sizeof(uint64_t) * ns));
  RSB_TRY(rsb_dev_malloc((void **)&val, sizeof(int) * ns));
  k_q_vkeys<<<nblocks(ns, 256), 256, 0, g->stream>>>(g->d_vtx, d_part, ns, g->nv, key, val);
  RSB_LAUNCH_CHECK();
  int rc = rsb_sort_pairs_u64(g, key, val, ns, 0, 64);
  if (rc) return rc;
  k_q_volume<<<nblocks(ns, 256), 256, 0, g->stream>>>(key, ns, acc + 2);
  RSB_LAUNCH_CHECK();
  std::vector<int> hs(2 * (size_t)nparts);
  unsigned long long ha[4];
  RSB_TRY(hipMemcpyAsync(hs.data(), size, sizeof(int) * 2 * nparts, hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipMemcpyAsync(ha, acc, sizeof(ha), hipMemcpyDeviceToHost, g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  int64_t mx = 0, mn = INT64_MAX, tot = 0, cmx = 0, empty = 0;
  for (int q = 0; q < nparts; q++) {
    mx = std::max<int64_t>(mx, hs[q]);
    mn = std::min<int64_t>(mn, hs[q]);
    tot += hs[nparts + q];
    cmx = std::max<int64_t>(cmx, hs[nparts + q]);
    empty += hs[q] == 0;
  }
  out[0] = (int64_t)ha[0], out[1] = (int64_t)ha[1], out[2] = mx, out[3] = mn, out[4] = tot;
  out[5] = cmx, out[6] = (int64_t)ha[2], out[7] = empty;
  rsb_dev_free(acc), rsb_dev_free(par), rsb_dev_free(size), rsb_dev_free(key), rsb_dev_free(val);
  return 0;
}
