// Segment tables and the pre-partition orders (RCB, global id).
//
// RCB (rcb.c:7-137) on one GPU: every segment with p > 1 virtual ranks is a
// root node of a bisection tree.  A node of n elements spread over p ranks is
// sorted stably along its axis of largest extent and split after the lower
// ceil(p/2) ranks' share (sort.c:48-70 distribution); at p == 1 the split is at
// floor(n/2) (rcb_local).  All nodes of a tree depth are processed in one
// round: one segmented min/max for the extents, one device-wide stable radix
// sort of the composite key (node index, dense coordinate rank), one pass to
// form the children.  Coordinates are replaced by their dense ranks per axis
// (equal coordinates share a rank), so ties stay ties and a stable sort keeps
// the current order among them.
#include <algorithm>

#include "state.hpp"

__device__ __host__ inline int64_t lower_count_dev(int64_t n, int p) {
  int64_t q = n / p, rem = n - q * p;
  int64_t hl = (p + 1) / 2;
  int64_t extra = hl - (p - rem);
  return q * hl + (extra > 0 ? extra : 0);
}

// ---------------------------------------------------------------------------
__global__ void k_segof(const int *__restrict__ seg, int nseg, const int *__restrict__ order,
                        int nel, int *__restrict__ segof) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  const int *cstart = seg;
  int lo = 0, hi = nseg - 1;  // last segment with cstart <= i (segments may be empty)
  while (lo < hi) {
    int mid = (lo + hi + 1) / 2;
    if (cstart[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  // skip empty segments that share the same start
  while (lo + 1 < nseg && cstart[lo + 1] == cstart[lo] && seg[nseg + lo] == 0) lo++;
  segof[order[i]] = lo;
}

extern "C" int rsbgpu_set_segments(rsbgpu *g, int nseg, const int *cstart, const int *n,
                                   const int *p, const int *r0) {
  RSB_TRY(hipSetDevice(g->device));
  g->nseg = nseg;
  g->h_cstart.assign(cstart, cstart + nseg);
  g->h_n.assign(n, n + nseg);
  g->h_p.assign(p, p + nseg);
  g->h_r0.assign(r0, r0 + nseg);
  g->h_pstart.assign(nseg, -1);
  g->h_seg_tiles.assign(2 * (size_t)nseg, 0);
  g->h_tile.clear();
  int64_t rows = 0;
  for (int s = 0; s < nseg; s++) {
    if (p[s] <= 1 || n[s] <= 0) continue;
    g->h_pstart[s] = (int)rows;
    int64_t padded = ((int64_t)n[s] + RSB_WAVE - 1) / RSB_WAVE * RSB_WAVE;
    g->h_seg_tiles[2 * s] = (int)(g->h_tile.size() / 3);
    int nt = 0;
    for (int64_t r = 0; r < padded; r += RSB_TILE, nt++) {
      g->h_tile.push_back(s);
      g->h_tile.push_back((int)(rows + r));
      g->h_tile.push_back((int)(rows + std::min<int64_t>(padded, r + RSB_TILE)));
    }
    g->h_seg_tiles[2 * s + 1] = nt;
    rows += padded;
  }
  if (rows >= (1ll << 28)) RSB_FAIL(3, "too many rows (%lld)", (long long)rows);
  g->mpad = rows;
  g->ntiles = (int)(g->h_tile.size() / 3);
  std::vector<int> h(5 * (size_t)nseg);
  for (int s = 0; s < nseg; s++) {
    h[s] = cstart[s];
    h[nseg + s] = n[s];
    h[2 * nseg + s] = p[s];
    h[3 * nseg + s] = r0[s];
    h[4 * nseg + s] = g->h_pstart[s];
  }
  RSB_TRY(dev_alloc(g->d_seg, h.size()));
  RSB_TRY(hipMemcpyAsync(g->d_seg, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice,
                         g->stream));
  RSB_TRY(dev_alloc(g->d_tile, g->h_tile.size()));
  if (!g->h_tile.empty())
    RSB_TRY(hipMemcpyAsync(g->d_tile, g->h_tile.data(), sizeof(int) * g->h_tile.size(),
                           hipMemcpyHostToDevice, g->stream));
  RSB_TRY(dev_alloc(g->d_seg_tiles, g->h_seg_tiles.size()));
  RSB_TRY(hipMemcpyAsync(g->d_seg_tiles, g->h_seg_tiles.data(),
                         sizeof(int) * g->h_seg_tiles.size(), hipMemcpyHostToDevice, g->stream));
  if (g->nel > 0) {
    k_segof<<<nblocks(g->nel, 256), 256, 0, g->stream>>>(g->d_seg, nseg, g->d_order, g->nel,
                                                         g->d_segof);
    RSB_LAUNCH_CHECK();
  }
  // host tables must stay alive until the async copies finished
  RSB_TRY(hipStreamSynchronize(g->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// RCB rounds.  Node table: start, len, p, frozen (1 = needs no more work).
struct nodes_t {
  int *start, *len, *p, *frozen;
};

__global__ void k_nodes_from_segments(const int *__restrict__ seg, int nseg,
                                      nodes_t nd) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  nd.start[s] = seg[s];
  nd.len[s] = seg[nseg + s];
  nd.p[s] = seg[2 * nseg + s];
  nd.frozen[s] = (seg[2 * nseg + s] <= 1 || seg[nseg + s] <= 1) ? 1 : 0;
}

__global__ void k_nodeof_init(const int *__restrict__ seg, int nseg, int nel,
                              int *__restrict__ nodeof) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) / 2;
    if (seg[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  while (lo + 1 < nseg && seg[lo + 1] == seg[lo] && seg[nseg + lo] == 0) lo++;
  nodeof[i] = lo;
}

// min/max per node of the three coordinates (as order-preserving keys)
__global__ void k_node_extent_init(int K, uint64_t *__restrict__ ext) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  for (int d = 0; d < 3; d++) {
    ext[(size_t)(2 * d) * K + k] = ~0ull;    // min
    ext[(size_t)(2 * d + 1) * K + k] = 0ull; // max
  }
}

__global__ void __launch_bounds__(256) k_node_extent(const int *__restrict__ order,
                                                     const int *__restrict__ nodeof,
                                                     const int *__restrict__ frozen,
                                                     const double *__restrict__ coord, int nel,
                                                     int ndim, int K,
                                                     unsigned long long *__restrict__ ext) {
  __shared__ uint64_t smin[4][3], smax[4][3];
  __shared__ int block_node;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool act = i < nel;
  int node = act ? nodeof[i] : -1;
  if (act && frozen[node]) act = false;
  int e = act ? order[i] : 0;
  // The whole block inside one node (large nodes): reduce in the block and
  // issue one atomic per extent.
  if (threadIdx.x == 0) {
    int last = min((int)(blockIdx.x * blockDim.x + blockDim.x), nel) - 1;
    block_node = (act && nodeof[last] == node) ? node : -1;
  }
  __syncthreads();
  const int lane = threadIdx.x & (RSB_WAVE - 1), wid = threadIdx.x / RSB_WAVE;
  const bool whole = block_node >= 0;
  for (int d = 0; d < ndim; d++) {
    uint64_t kv = act ? dkey(coord[(size_t)d * nel + e]) : 0;
    uint64_t mn = act ? kv : ~0ull, mx = act ? kv : 0ull;
    if (whole) {
      for (int o = RSB_WAVE / 2; o > 0; o >>= 1) {
        uint64_t a = __shfl_xor(mn, o, RSB_WAVE), b = __shfl_xor(mx, o, RSB_WAVE);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
      }
      if (lane == 0) smin[wid][d] = mn, smax[wid][d] = mx;
      continue;
    }
    // runs of equal node within the wave, then one atomic per run
    for (int o = 1; o < RSB_WAVE; o <<= 1) {
      int nb = __shfl_down(node, o, RSB_WAVE);
      uint64_t a = __shfl_down(mn, o, RSB_WAVE), b = __shfl_down(mx, o, RSB_WAVE);
      if (lane + o < RSB_WAVE && nb == node) {
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
      }
    }
    int prev = __shfl_up(node, 1, RSB_WAVE);
    bool head = act && (lane == 0 || prev != node);
    if (head) {
      atomicMin(&ext[(size_t)(2 * d) * K + node], (unsigned long long)mn);
      atomicMax(&ext[(size_t)(2 * d + 1) * K + node], (unsigned long long)mx);
    }
  }
  if (!whole) return;
  __syncthreads();
  if (threadIdx.x < ndim) {
    int d = threadIdx.x;
    uint64_t mn = smin[0][d], mx = smax[0][d];
    for (int w = 1; w < (int)blockDim.x / RSB_WAVE; w++) {
      mn = smin[w][d] < mn ? smin[w][d] : mn;
      mx = smax[w][d] > mx ? smax[w][d] : mx;
    }
    atomicMin(&ext[(size_t)(2 * d) * K + block_node], (unsigned long long)mn);
    atomicMax(&ext[(size_t)(2 * d + 1) * K + block_node], (unsigned long long)mx);
  }
}

__global__ void k_node_axis(int K, int ndim, const int *__restrict__ frozen,
                            const uint64_t *__restrict__ ext, int *__restrict__ axis) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  if (frozen[k]) {
    axis[k] = -1;
    return;
  }
  double len[3];
  for (int d = 0; d < ndim; d++)
    len[d] = dkey_inv(ext[(size_t)(2 * d + 1) * K + k]) - dkey_inv(ext[(size_t)(2 * d) * K + k]);
  int a = 0;
  if (len[1] > len[0]) a = 1;
  if (ndim == 3 && len[2] > len[a]) a = 2;
  axis[k] = a;
}

__global__ void k_rcb_keys(const int *__restrict__ order, const int *__restrict__ nodeof,
                           const int *__restrict__ axis, const uint32_t *__restrict__ rank,
                           int nel, int rank_bits, uint64_t *__restrict__ key,
                           int *__restrict__ val) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int node = nodeof[i], e = order[i], a = axis[node];
  uint64_t r = a < 0 ? 0 : rank[(size_t)a * nel + e];
  key[i] = ((uint64_t)node << rank_bits) | r;
  val[i] = e;
}

__global__ void k_child_count(int K, const nodes_t nd, int *__restrict__ cnt) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  cnt[k] = nd.frozen[k] ? 1 : 2;
}

__global__ void k_children(int K, const nodes_t nd, const int *__restrict__ first, nodes_t ch,
                           int *__restrict__ split) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  int c = first[k];
  if (nd.frozen[k]) {
    ch.start[c] = nd.start[k], ch.len[c] = nd.len[k], ch.p[c] = nd.p[k], ch.frozen[c] = 1;
    split[k] = INT_MAX;
    return;
  }
  int n = nd.len[k], p = nd.p[k];
  int nl = p > 1 ? (int)lower_count_dev(n, p) : n / 2;
  int pl = p > 1 ? (p + 1) / 2 : 1, pu = p > 1 ? p / 2 : 1;
  ch.start[c] = nd.start[k], ch.len[c] = nl, ch.p[c] = pl;
  ch.frozen[c] = nl <= 1;
  ch.start[c + 1] = nd.start[k] + nl, ch.len[c + 1] = n - nl, ch.p[c + 1] = pu;
  ch.frozen[c + 1] = (n - nl) <= 1;
  split[k] = nd.start[k] + nl;
}

__global__ void k_nodeof_next(int nel, const int *__restrict__ first,
                              const int *__restrict__ split, int *__restrict__ nodeof) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int k = nodeof[i];
  nodeof[i] = first[k] + (i >= split[k] ? 1 : 0);
}

// Depth of the RCB tree of (n, p): the larger child is never shallower.
static int rcb_depth(int64_t n, int p) {
  int d = 0;
  while (n > 1) {
    int64_t nl = p > 1 ? lower_count_dev(n, p) : n / 2;
    int pl = p > 1 ? (p + 1) / 2 : 1, pu = p > 1 ? p / 2 : 1;
    if (nl >= n - nl) n = nl, p = pl;
    else n = n - nl, p = pu;
    d++;
  }
  return d;
}

static int run_rcb(rsbgpu *g) {
  const int nel = g->nel, nseg = g->nseg;
  int rounds = 0;
  for (int s = 0; s < nseg; s++)
    if (g->h_p[s] > 1) rounds = std::max(rounds, rcb_depth(g->h_n[s], g->h_p[s]));
  if (rounds == 0) return 0;
  // node tables (two generations), sized for the worst case of nel nodes
  // Every round adds at most one node per non-frozen node (len >= 2), so a
  // round adds at most nel / 2 nodes; empty children (n < p) keep the count
  // growing by at most that bound.
  size_t cap = (size_t)nseg + (size_t)nel * (size_t)std::min(rounds, 4) / 2 + (size_t)nel + 16;
  int *buf = nullptr;
  RSB_TRY(hipMalloc((void **)&buf, sizeof(int) * cap * 13 + sizeof(int) * (size_t)nel));
  nodes_t A{buf, buf + cap, buf + 2 * cap, buf + 3 * cap};
  nodes_t B{buf + 4 * cap, buf + 5 * cap, buf + 6 * cap, buf + 7 * cap};
  int *cnt = buf + 8 * cap, *first = buf + 9 * cap, *split = buf + 10 * cap;
  int *axis = buf + 11 * cap, *nodeof = buf + 13 * cap;
  uint64_t *ext = nullptr, *key = nullptr;
  RSB_TRY(hipMalloc((void **)&ext, sizeof(uint64_t) * 6 * cap));
  RSB_TRY(hipMalloc((void **)&key, sizeof(uint64_t) * ((size_t)nel + 1)));
  int K = nseg;
  k_nodes_from_segments<<<nblocks(nseg, 256), 256, 0, g->stream>>>(g->d_seg, nseg, A);
  RSB_LAUNCH_CHECK();
  k_nodeof_init<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_seg, nseg, nel, nodeof);
  RSB_LAUNCH_CHECK();
  for (int r = 0; r < rounds; r++) {
    k_node_extent_init<<<nblocks(K, 256), 256, 0, g->stream>>>(K, ext);
    RSB_LAUNCH_CHECK();
    k_node_extent<<<nblocks(nel, 256), 256, 0, g->stream>>>(
        g->d_order, nodeof, A.frozen, g->d_coord, nel, g->ndim, K,
        (unsigned long long *)ext);
    RSB_LAUNCH_CHECK();
    k_node_axis<<<nblocks(K, 256), 256, 0, g->stream>>>(K, g->ndim, A.frozen, ext, axis);
    RSB_LAUNCH_CHECK();
    k_rcb_keys<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, nodeof, axis, g->d_rank,
                                                         nel, g->rank_bits, key, g->d_order);
    RSB_LAUNCH_CHECK();
    int node_bits = 1;
    while ((1ll << node_bits) < K) node_bits++;
    int rc = rsb_sort_pairs_u64(g, key, g->d_order, nel, 0, g->rank_bits + node_bits);
    if (rc) return rc;
    // children
    k_child_count<<<nblocks(K, 256), 256, 0, g->stream>>>(K, A, cnt);
    RSB_LAUNCH_CHECK();
    rc = rsb_exclusive_scan_i32(g, cnt, first, K);
    if (rc) return rc;
    int last[2];
    RSB_TRY(hipMemcpyAsync(&last[0], first + K - 1, sizeof(int), hipMemcpyDeviceToHost,
                           g->stream));
    RSB_TRY(hipMemcpyAsync(&last[1], cnt + K - 1, sizeof(int), hipMemcpyDeviceToHost,
                           g->stream));
    RSB_TRY(hipStreamSynchronize(g->stream));
    int K2 = last[0] + last[1];
    if ((size_t)K2 > cap) RSB_FAIL(2, "RCB node table overflow (%d > %zu)", K2, cap);
    k_children<<<nblocks(K, 256), 256, 0, g->stream>>>(K, A, first, B, split);
    RSB_LAUNCH_CHECK();
    k_nodeof_next<<<nblocks(nel, 256), 256, 0, g->stream>>>(nel, first, split, nodeof);
    RSB_LAUNCH_CHECK();
    K = K2;
    std::swap(A, B);
  }
  hipFree(buf), hipFree(ext), hipFree(key);
  return 0;
}

// rsb_pre == 0: stable sort of each active segment by global id (= input
// index), rsb.c:284-287.  Inactive segments keep their order.
__global__ void k_gid_keys(const int *__restrict__ order, const int *__restrict__ segof,
                           const int *__restrict__ seg, int nseg, int nel,
                           uint64_t *__restrict__ key) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int e = order[i], s = segof[e];
  int active = seg[2 * nseg + s] > 1;
  key[i] = ((uint64_t)s << 32) | (uint32_t)(active ? e : i - seg[s]);
}

static int run_gid(rsbgpu *g) {
  const int nel = g->nel;
  uint64_t *key = nullptr;
  RSB_TRY(hipMalloc((void **)&key, sizeof(uint64_t) * ((size_t)nel + 1)));
  k_gid_keys<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, g->d_segof, g->d_seg,
                                                       g->nseg, nel, key);
  RSB_LAUNCH_CHECK();
  int sb = 1;
  while ((1ll << sb) < g->nseg) sb++;
  int rc = rsb_sort_pairs_u64(g, key, g->d_order, nel, 0, 32 + sb);
  hipFree(key);
  return rc;
}

extern "C" int rsbgpu_prepartition(rsbgpu *g, int rsb_pre) {
  RSB_TRY(hipSetDevice(g->device));
  if (g->nel == 0) return 0;
  rsb_phase_begin(g);
  int rc = 0;
  if (rsb_pre == 1) {
    if (!g->has_coord) RSB_FAIL(3, "RCB needs coordinates");
    rc = run_rcb(g);
  } else if (rsb_pre == 0) {
    rc = run_gid(g);
  } else {
    RSB_FAIL(3, "rsb_pre=%d is not supported by this build", rsb_pre);
  }
  rsb_phase_end(g, PH_PRE);
  return rc;
}
