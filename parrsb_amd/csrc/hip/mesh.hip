// Handle lifetime, mesh upload, centroids, coordinate ranks and the weighted
// element dual graph.
#include <string.h>

#include <algorithm>

#include "state.hpp"

// ---------------------------------------------------------------------------
// timing
void rsb_phase_begin(rsbgpu *g) { hipEventRecord(g->ev0, g->stream); }
void rsb_phase_end(rsbgpu *g, int ph) {
  hipEventRecord(g->ev1, g->stream);
  if (hipEventSynchronize(g->ev1) == hipSuccess) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g->ev0, g->ev1) == hipSuccess) g->ph_ms[ph] += ms;
  }
}

extern "C" int rsbgpu_phase_times(rsbgpu *g, double *t) {
  for (int i = 0; i < PH_N; i++) t[i] = g->ph_ms[i];
  return 0;
}

extern "C" int rsbgpu_lanczos_traffic(rsbgpu *g, double *v6) {
  v6[0] = g->lz_bytes, v6[1] = g->lz_ms, v6[2] = (double)g->lz_launches;
  // the timed launches are a sample (every LZ_PROF_EVERY-th): their mean
  // duration, times all launches
  v6[3] = g->spmv_bytes;
  v6[4] = g->spmv_launches ? g->spmv_ms * (double)g->spmv_total / (double)g->spmv_launches : 0.0;
  v6[5] = (double)g->spmv_total;
  return 0;
}

extern "C" int rsbgpu_set_profiling(rsbgpu *g, int spmv_events) {
  g->prof_spmv = spmv_events;
  return 0;
}

extern "C" const char *rsbgpu_error(const rsbgpu *g) { return g ? g->err : "no handle"; }

extern "C" int rsbgpu_open(rsbgpu **out, int device) {
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return 2;
  if (device < 0 || device >= ndev) return 2;
  if (hipSetDevice(device) != hipSuccess) return 2;
  (void)hipGetLastError();  // start from a clean error state
  rsbgpu *g = new rsbgpu();
  g->device = device;
  hipEvent_t ev[4];
  if (rsb_arena_stream(device, &g->stream, ev)) {
    // allocation-free call: the stream and events set up for this device
    g->ev0 = ev[0], g->ev1 = ev[1], g->lz_ev0 = ev[2], g->lz_ev1 = ev[3];
    g->borrowed = true;
    *out = g;
    return 0;
  }
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess ||
      hipEventCreate(&g->lz_ev0) != hipSuccess || hipEventCreate(&g->lz_ev1) != hipSuccess) {
    delete g;
    return 2;
  }
  *out = g;
  return 0;
}

#define FREE(p)                 \
  do {                          \
    if (p) rsb_dev_free((void *)p);  \
    p = nullptr;                \
  } while (0)

extern "C" void rsbgpu_close(rsbgpu *g) {
  if (!g) return;
  hipSetDevice(g->device);
  if (g->stream) hipStreamSynchronize(g->stream);
  rsb_mg_free(g);
  FREE(g->own_vtx);
  FREE(g->own_xyz);
  FREE(g->d_coord);
  FREE(g->d_rank);
  FREE(g->d_growptr);
  FREE(g->d_gadj);
  FREE(g->d_order);
  FREE(g->d_order2);
  FREE(g->d_segof);
  FREE(g->d_seg);
  FREE(g->d_row_elem);
  FREE(g->d_elem_row);
  FREE(g->d_slice_off);
  FREE(g->d_slice_w);
  FREE(g->d_ell);
  FREE(g->d_rowlen);
  FREE(g->d_tnhalo);
  FREE(g->d_halo);
  FREE(g->d_w16);
  FREE(g->d_thalo_off);
  FREE(g->d_off16);
  FREE(g->d_e16);
  FREE(g->d_w12);
  FREE(g->d_e12_bad);
  FREE(g->d_off12);
  FREE(g->d_e12);
  FREE(g->d_c12);
  FREE(g->d_seg_nnz);
  FREE(g->d_vt_n);
  FREE(g->d_vt_loff);
  FREE(g->d_vt_list);
  FREE(g->d_vt_rowv);
  for (hipEvent_t ev : g->spmv_ev) hipEventDestroy(ev);
  FREE(g->d_tile);
  FREE(g->d_seg_tiles);
  FREE(g->d_p);
  FREE(g->d_p2);
  FREE(g->d_w);
  FREE(g->d_w2);
  FREE(g->d_f);
  FREE(g->d_basis);
  FREE(g->d_fied);
  FREE(g->d_start);
  FREE(g->d_lws_ws);
  FREE(g->lz.status);  // iter, k, npass, niter, miter live in the same block
  FREE(g->lz.sc);
  FREE(g->lz.T);
  FREE(g->lz.y);
  FREE(g->lz.partial);
  FREE(g->lz.rn);
  FREE(g->lz.running);
  FREE(g->d_sort_tmp);
  FREE(g->d_k0);
  FREE(g->d_k1);
  FREE(g->d_v1);
  FREE(g->d_iscratch);
  for (auto &kv : g->scratch)
    if (kv.second.first) rsb_dev_free(kv.second.first);
  g->scratch.clear();
  if (!g->borrowed) {
    if (g->ev0) hipEventDestroy(g->ev0);
    if (g->ev1) hipEventDestroy(g->ev1);
    if (g->lz_ev0) hipEventDestroy(g->lz_ev0);
    if (g->lz_ev1) hipEventDestroy(g->lz_ev1);
    if (g->stream) hipStreamDestroy(g->stream);
  }
  delete g;
}

// ---------------------------------------------------------------------------
// Centroids (parrsb.c:139-145): FP64, corners added in order starting from
// 0.0, then divided by nv.  Plain adds and one correctly rounded division:
// bit-identical to the host computation.
__global__ void k_centroids(const double *__restrict__ xyz, int nel, int nv, int ndim,
                            double *__restrict__ coord) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  double c[3] = {0.0, 0.0, 0.0};
  const double *x = xyz + (size_t)e * nv * ndim;
  for (int v = 0; v < nv; v++)
    for (int d = 0; d < ndim; d++) c[d] = c[d] + x[v * ndim + d];
  for (int d = 0; d < 3; d++) coord[(size_t)d * nel + e] = (d < ndim) ? c[d] / (double)nv : 0.0;
}

__global__ void k_coord_keys(const double *__restrict__ c, int nel, uint64_t *__restrict__ key,
                             int *__restrict__ val) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  key[e] = dkey(c[e]);
  val[e] = e;
}

__global__ void k_diff_flags(const uint64_t *__restrict__ key, int64_t n, int *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i > 0 && key[i] != key[i - 1]) ? 1 : 0;
}

__global__ void k_scatter_rank(const int *__restrict__ excl, const int *__restrict__ flag,
                               const int *__restrict__ val, int n, uint32_t *__restrict__ rank) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rank[val[i]] = (uint32_t)(excl[i] + flag[i]);
}

// ---------------------------------------------------------------------------
// Dual graph.  Slots (e, j) are sorted by vertex id; a group is a run of equal
// ids (id 0 never groups: "no connection" in gather-scatter).  An element's
// neighbours are the union of its nv groups, which are each sorted by
// element: an nv-way merge yields them sorted and counts the shared ids.
__global__ void k_slot_keys(const long long *__restrict__ vtx, int64_t ns,
                            uint64_t *__restrict__ key, int *__restrict__ val,
                            unsigned long long *__restrict__ vmax) {
  // grid-stride over the slots; one atomic per workgroup (one per wave on
  // a single address serialised ~2M atomics at 16.7M hexes: 24 ms)
  __shared__ unsigned long long wm[RSB_BLOCK / RSB_WAVE];
  unsigned long long m = 0;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns;
       s += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long v = (unsigned long long)vtx[s];
    key[s] = (uint64_t)v;
    val[s] = (int)s;
    m = v > m ? v : m;
  }
  for (int o = RSB_WAVE / 2; o > 0; o >>= 1) {
    unsigned long long t = __shfl_xor(m, o, RSB_WAVE);
    m = t > m ? t : m;
  }
  if ((threadIdx.x & (RSB_WAVE - 1)) == 0) wm[threadIdx.x / RSB_WAVE] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x / RSB_WAVE); w++) m = wm[w] > m ? wm[w] : m;
    atomicMax(vmax, m);
  }
}

__global__ void k_group_flags(const uint64_t *__restrict__ key, int64_t n, int *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || key[i] != key[i - 1] || key[i] == 0) ? 1 : 0;
}

__global__ void k_group_scatter(const int *__restrict__ flag, const int *__restrict__ gexcl,
                                const int *__restrict__ val, int64_t n,
                                int *__restrict__ slot_group, int *__restrict__ gstart) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int gid = gexcl[i] + flag[i] - 1;
  slot_group[val[i]] = gid;
  if (flag[i]) gstart[gid] = (int)i;
}

template <bool FILL>
__global__ void k_dual_rows(int nel, int nv, const int *__restrict__ slot_group,
                            const int *__restrict__ gstart, const int *__restrict__ sorted_slot,
                            int64_t *__restrict__ rowlen, const int64_t *__restrict__ rowptr,
                            uint32_t *__restrict__ adj, int *__restrict__ overflow) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  int pos[8], end[8], hv[8];
  for (int j = 0; j < 8; j++) {
    if (j < nv) {
      int gidx = slot_group[(int64_t)e * nv + j];
      pos[j] = gstart[gidx];
      end[j] = gstart[gidx + 1];
    } else {
      pos[j] = end[j] = 0;
    }
  }
  for (int j = 0; j < 8; j++) hv[j] = pos[j] < end[j] ? sorted_slot[pos[j]] / nv : INT_MAX;
  int64_t out = FILL ? rowptr[e] : 0;
  int cnt = 0;
  while (true) {
    int m = INT_MAX;
    for (int j = 0; j < 8; j++) m = hv[j] < m ? hv[j] : m;
    if (m == INT_MAX) break;
    int w = 0;
    for (int j = 0; j < 8; j++) {
      // a group lists each element's slots consecutively; consume them all
      while (hv[j] == m) {
        w++;
        pos[j]++;
        hv[j] = pos[j] < end[j] ? sorted_slot[pos[j]] / nv : INT_MAX;
      }
    }
    if (m == e) continue;
    if (FILL) {
      if (w > 15) atomicOr(overflow, 1);
      adj[out + cnt] = (uint32_t)m | ((uint32_t)(w & 15) << 28);
    }
    cnt++;
  }
  if (!FILL) rowlen[e] = cnt;
}

__global__ void k_wsum(const uint32_t *__restrict__ adj, int64_t nnz,
                       unsigned long long *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long s = 0;
  for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) s += adj[i] >> 28;
  s = wave_sum(s);
  if ((threadIdx.x & (RSB_WAVE - 1)) == 0) atomicAdd(out, s);
}

__global__ void k_iota(int *__restrict__ a, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

static int build_ranks(rsbgpu *g) {
  int n = g->nel;
  uint64_t *key = nullptr;
  int *val = nullptr, *flag = nullptr, *excl = nullptr;
  RSB_TRY(rsb_dev_malloc((void **)&key, sizeof(uint64_t) * (n + 1)));
  RSB_TRY(rsb_dev_malloc((void **)&val, sizeof(int) * (n + 1)));
  RSB_TRY(rsb_dev_malloc((void **)&flag, sizeof(int) * (n + 1)));
  RSB_TRY(rsb_dev_malloc((void **)&excl, sizeof(int) * (n + 1)));
  RSB_TRY(hipMemsetAsync(g->d_rank, 0, sizeof(uint32_t) * 3 * (size_t)n, g->stream));
  uint32_t maxrank = 0;
  for (int d = 0; d < g->ndim; d++) {
    k_coord_keys<<<nblocks(n, 256), 256, 0, g->stream>>>(g->d_coord + (size_t)d * n, n, key,
                                                         val);
    RSB_LAUNCH_CHECK();
    int rc = rsb_sort_pairs_u64(g, key, val, n, 0, 64);
    if (rc) return rc;
    k_diff_flags<<<nblocks(n, 256), 256, 0, g->stream>>>(key, n, flag);
    RSB_LAUNCH_CHECK();
    rc = rsb_exclusive_scan_i32(g, flag, excl, n);
    if (rc) return rc;
    k_scatter_rank<<<nblocks(n, 256), 256, 0, g->stream>>>(excl, flag, val, n,
                                                          g->d_rank + (size_t)d * n);
    RSB_LAUNCH_CHECK();
    int last[2] = {0, 0};
    RSB_TRY(hipMemcpyAsync(&last[0], excl + n - 1, sizeof(int), hipMemcpyDeviceToHost,
                           g->stream));
    RSB_TRY(hipMemcpyAsync(&last[1], flag + n - 1, sizeof(int), hipMemcpyDeviceToHost,
                           g->stream));
    RSB_TRY(hipStreamSynchronize(g->stream));
    uint32_t mr = (uint32_t)(last[0] + last[1]);
    maxrank = std::max(maxrank, mr);
  }
  int bits = 1;
  while (bits < 32 && (1u << bits) <= maxrank) bits++;
  g->rank_bits = bits;
  rsb_dev_free(key), rsb_dev_free(val), rsb_dev_free(flag), rsb_dev_free(excl);
  return 0;
}

static int build_graph(rsbgpu *g) {
  int64_t ns = (int64_t)g->nel * g->nv;
  int n = g->nel;
  uint64_t *key = nullptr;
  int *val = nullptr, *flag = nullptr, *excl = nullptr, *slot_group = nullptr, *gstart = nullptr;
  unsigned long long *vmax = nullptr;
  int64_t *rowlen = nullptr;
  int *ovf = nullptr;
  RSB_TRY(rsb_dev_malloc((void **)&key, sizeof(uint64_t) * (ns + 1)));
  RSB_TRY(rsb_dev_malloc((void **)&val, sizeof(int) * (ns + 1)));
  RSB_TRY(rsb_dev_malloc((void **)&flag, sizeof(int) * (ns + 1)));
  RSB_TRY(rsb_dev_malloc((void **)&excl, sizeof(int) * (ns + 1)));
  RSB_TRY(rsb_dev_malloc((void **)&vmax, sizeof(unsigned long long) * 2));
  RSB_TRY(hipMemsetAsync(vmax, 0, sizeof(unsigned long long) * 2, g->stream));
  k_slot_keys<<<std::min<int64_t>(nblocks(ns, RSB_BLOCK), 4096), RSB_BLOCK, 0, g->stream>>>(
      g->d_vtx, ns, key, val, vmax);
  RSB_LAUNCH_CHECK();
  unsigned long long hmax = 0;
  RSB_TRY(hipMemcpyAsync(&hmax, vmax, sizeof(hmax), hipMemcpyDeviceToHost, g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  int bits = 1;
  while (bits < 64 && (1ull << bits) <= hmax) bits++;
  int rc = rsb_sort_pairs_u64(g, key, val, ns, 0, bits);
  if (rc) return rc;
  k_group_flags<<<nblocks(ns, 256), 256, 0, g->stream>>>(key, ns, flag);
  RSB_LAUNCH_CHECK();
  rc = rsb_exclusive_scan_i32(g, flag, excl, ns);
  if (rc) return rc;
  int ng_parts[2];
  RSB_TRY(hipMemcpyAsync(&ng_parts[0], excl + ns - 1, sizeof(int), hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipMemcpyAsync(&ng_parts[1], flag + ns - 1, sizeof(int), hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  int ng = ng_parts[0] + ng_parts[1];
  RSB_TRY(rsb_dev_malloc((void **)&slot_group, sizeof(int) * (ns + 1)));
  RSB_TRY(rsb_dev_malloc((void **)&gstart, sizeof(int) * (ng + 1)));
  k_group_scatter<<<nblocks(ns, 256), 256, 0, g->stream>>>(flag, excl, val, ns, slot_group,
                                                           gstart);
  RSB_LAUNCH_CHECK();
  int nsi = (int)ns;
  RSB_TRY(hipMemcpyAsync(gstart + ng, &nsi, sizeof(int), hipMemcpyHostToDevice, g->stream));
  rsb_dev_free(key), rsb_dev_free(flag), rsb_dev_free(excl);
  key = nullptr;
  // count, scan, fill
  RSB_TRY(rsb_dev_malloc((void **)&rowlen, sizeof(int64_t) * (n + 1)));
  RSB_TRY(dev_alloc(g, g->d_growptr, (size_t)n + 1));
  RSB_TRY(rsb_dev_malloc((void **)&ovf, sizeof(int)));
  RSB_TRY(hipMemsetAsync(ovf, 0, sizeof(int), g->stream));
  RSB_TRY(hipMemsetAsync(rowlen + n, 0, sizeof(int64_t), g->stream));
  k_dual_rows<false><<<nblocks(n, 128), 128, 0, g->stream>>>(
      n, g->nv, slot_group, gstart, val, rowlen, nullptr, nullptr, ovf);
  RSB_LAUNCH_CHECK();
  rc = rsb_exclusive_scan_i64(g, rowlen, g->d_growptr, (int64_t)n + 1);
  if (rc) return rc;
  RSB_TRY(hipMemcpyAsync(&g->gnnz, g->d_growptr + n, sizeof(int64_t), hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  RSB_TRY(dev_alloc(g, g->d_gadj, (size_t)g->gnnz));
  k_dual_rows<true><<<nblocks(n, 128), 128, 0, g->stream>>>(
      n, g->nv, slot_group, gstart, val, nullptr, g->d_growptr, g->d_gadj, ovf);
  RSB_LAUNCH_CHECK();
  RSB_TRY(hipMemsetAsync(vmax, 0, sizeof(unsigned long long), g->stream));
  if (g->gnnz > 0) {
    k_wsum<<<1024, 256, 0, g->stream>>>(g->d_gadj, g->gnnz, vmax);
    RSB_LAUNCH_CHECK();
  }
  int hovf = 0;
  unsigned long long ws = 0;
  RSB_TRY(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, g->stream));
  RSB_TRY(hipMemcpyAsync(&ws, vmax, sizeof(ws), hipMemcpyDeviceToHost, g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  g->gwsum = (int64_t)ws;
  rsb_dev_free(val), rsb_dev_free(slot_group), rsb_dev_free(gstart), rsb_dev_free(rowlen), rsb_dev_free(ovf),
      rsb_dev_free(vmax);
  if (hovf) RSB_FAIL(3, "two elements share more than 15 vertex ids");
  return 0;
}

extern "C" int rsbgpu_set_mesh(rsbgpu *g, const long long *vtx, const double *xyz, int nel,
                               int nv, int on_device) {
  RSB_TRY(hipSetDevice(g->device));
  if (nv != 4 && nv != 8) RSB_FAIL(3, "nv must be 4 or 8 (got %d)", nv);
  if (nel < 0 || nel >= (1 << 28)) RSB_FAIL(3, "nel out of range (%d)", nel);
  rsb_phase_begin(g);
  g->nel = nel, g->nv = nv, g->ndim = (nv == 8) ? 3 : 2;
  size_t ns = (size_t)nel * nv;
  if (on_device) {
    g->d_vtx = vtx, g->d_xyz = xyz;
  } else {
    if (vtx) {
      RSB_TRY(dev_alloc(g, g->own_vtx, ns));
      RSB_TRY(hipMemcpyAsync(g->own_vtx, vtx, ns * sizeof(long long), hipMemcpyHostToDevice,
                             g->stream));
      g->d_vtx = g->own_vtx;
    } else {
      g->d_vtx = nullptr;
    }
    if (xyz) {
      RSB_TRY(dev_alloc(g, g->own_xyz, ns * g->ndim));
      RSB_TRY(hipMemcpyAsync(g->own_xyz, xyz, ns * g->ndim * sizeof(double),
                             hipMemcpyHostToDevice, g->stream));
      g->d_xyz = g->own_xyz;
    } else {
      g->d_xyz = nullptr;
    }
  }
  g->has_coord = xyz != nullptr;
  RSB_TRY(dev_alloc(g, g->d_coord, 3 * (size_t)nel));
  RSB_TRY(dev_alloc(g, g->d_rank, 3 * (size_t)nel));
  if (g->has_coord && nel > 0) {
    k_centroids<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_xyz, nel, nv, g->ndim,
                                                          g->d_coord);
    RSB_LAUNCH_CHECK();
    int rc = build_ranks(g);
    if (rc) return rc;
  } else {
    RSB_TRY(hipMemsetAsync(g->d_coord, 0, 3 * sizeof(double) * (size_t)nel, g->stream));
    RSB_TRY(hipMemsetAsync(g->d_rank, 0, 3 * sizeof(uint32_t) * (size_t)nel, g->stream));
  }
  if (nel > 0 && g->d_vtx) {
    int rc = build_graph(g);
    if (rc) return rc;
  }
  RSB_TRY(dev_alloc(g, g->d_order, (size_t)nel));
  RSB_TRY(dev_alloc(g, g->d_order2, (size_t)nel));
  RSB_TRY(dev_alloc(g, g->d_segof, (size_t)nel));
  RSB_TRY(dev_alloc(g, g->d_fied, (size_t)nel));
  if (nel > 0) {
    k_iota<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, nel);
    RSB_LAUNCH_CHECK();
  }
  rsb_phase_end(g, PH_SETUP);
  return 0;
}

extern "C" int rsbgpu_graph_stats(rsbgpu *g, int64_t *nnz, int64_t *wsum) {
  *nnz = g->gnnz, *wsum = g->gwsum;
  return 0;
}

__global__ void k_unpack_adj(const uint32_t *__restrict__ adj, int64_t n, int *__restrict__ col,
                             int *__restrict__ w) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  col[i] = (int)(adj[i] & 0x0fffffffu);
  w[i] = (int)(adj[i] >> 28);
}

extern "C" int rsbgpu_get_graph(rsbgpu *g, int64_t *rowptr, int *col, int *wgt) {
  RSB_TRY(hipSetDevice(g->device));
  RSB_TRY(hipMemcpyAsync(rowptr, g->d_growptr, sizeof(int64_t) * ((size_t)g->nel + 1),
                         hipMemcpyDeviceToHost, g->stream));
  if (g->gnnz > 0) {
    int *dc = nullptr, *dw = nullptr;
    RSB_TRY(rsb_dev_malloc((void **)&dc, sizeof(int) * g->gnnz));
    RSB_TRY(rsb_dev_malloc((void **)&dw, sizeof(int) * g->gnnz));
    k_unpack_adj<<<nblocks(g->gnnz, 256), 256, 0, g->stream>>>(g->d_gadj, g->gnnz, dc, dw);
    RSB_LAUNCH_CHECK();
    RSB_TRY(hipMemcpyAsync(col, dc, sizeof(int) * g->gnnz, hipMemcpyDeviceToHost, g->stream));
    RSB_TRY(hipMemcpyAsync(wgt, dw, sizeof(int) * g->gnnz, hipMemcpyDeviceToHost, g->stream));
    RSB_TRY(hipStreamSynchronize(g->stream));
    rsb_dev_free(dc), rsb_dev_free(dw);
  }
  RSB_TRY(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" int rsbgpu_get_centroids(rsbgpu *g, double *coord) {
  RSB_TRY(hipSetDevice(g->device));
  RSB_TRY(hipMemcpyAsync(coord, g->d_coord, sizeof(double) * 3 * (size_t)g->nel,
                         hipMemcpyDeviceToHost, g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" int rsbgpu_get_order(rsbgpu *g, int *order) {
  RSB_TRY(hipSetDevice(g->device));
  RSB_TRY(hipMemcpyAsync(order, g->d_order, sizeof(int) * (size_t)g->nel, hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" int rsbgpu_get_fiedler(rsbgpu *g, double *f) {
  RSB_TRY(hipSetDevice(g->device));
  RSB_TRY(hipMemcpyAsync(f, g->d_fied, sizeof(double) * (size_t)g->nel, hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" int rsbgpu_set_fiedler(rsbgpu *g, const double *f) {
  RSB_TRY(hipSetDevice(g->device));
  if (g->nel > 0 && !g->d_fied) RSB_FAIL(3, "no mesh set");
  RSB_TRY(hipMemcpyAsync(g->d_fied, f, sizeof(double) * (size_t)g->nel, hipMemcpyHostToDevice,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" int rsbgpu_upload_ints(rsbgpu *g, const int *host, int n, int **dev_out) {
  RSB_TRY(hipSetDevice(g->device));
  int *d = nullptr;
  RSB_TRY(rsb_dev_malloc((void **)&d, sizeof(int) * (size_t)(n > 0 ? n : 1)));
  if (n > 0) RSB_TRY(hipMemcpy(d, host, sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
  *dev_out = d;
  return 0;
}

extern "C" int rsbgpu_upload(rsbgpu *g, const void *host, size_t bytes, void **dev_out) {
  RSB_TRY(hipSetDevice(g->device));
  void *d = nullptr;
  RSB_TRY(rsb_dev_malloc(&d, bytes > 0 ? bytes : 1));
  if (bytes > 0) RSB_TRY(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice));
  *dev_out = d;
  return 0;
}

extern "C" int rsbgpu_download(rsbgpu *g, const void *dev, size_t bytes, void *host) {
  RSB_TRY(hipSetDevice(g->device));
  if (bytes > 0) RSB_TRY(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" void rsbgpu_free(void *dev) {
  if (dev) rsb_dev_free(dev);
}
