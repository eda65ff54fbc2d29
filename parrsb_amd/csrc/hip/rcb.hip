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
#include <cfloat>

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
  RSB_TRY(dev_alloc(g, g->d_seg, h.size()));
  RSB_TRY(hipMemcpyAsync(g->d_seg, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice,
                         g->stream));
  RSB_TRY(dev_alloc(g, g->d_tile, g->h_tile.size()));
  if (!g->h_tile.empty())
    RSB_TRY(hipMemcpyAsync(g->d_tile, g->h_tile.data(), sizeof(int) * g->h_tile.size(),
                           hipMemcpyHostToDevice, g->stream));
  RSB_TRY(dev_alloc(g, g->d_seg_tiles, g->h_seg_tiles.size()));
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

// Each workgroup covers NE_PER * 256 consecutive positions.  A workgroup
// inside one node (large nodes) reduces in registers, shuffles and LDS and
// issues one atomic per extent: on the first rounds every workgroup hits the
// same node, and one atomic per 256 positions serialised ~400k atomics on
// six addresses (0.57 ms a round at 16.7M).  Otherwise runs of equal node in
// a wave issue one atomic each.
#define NE_PER 8
__global__ void __launch_bounds__(256) k_node_extent(const int *__restrict__ order,
                                                     const int *__restrict__ nodeof,
                                                     const int *__restrict__ frozen,
                                                     const double *__restrict__ coord, int nel,
                                                     int ndim, int K,
                                                     unsigned long long *__restrict__ ext) {
  __shared__ uint64_t smin[4][3], smax[4][3];
  __shared__ int block_node;
  const int64_t base = (int64_t)blockIdx.x * blockDim.x * NE_PER;
  if (threadIdx.x == 0) {
    const int64_t last = min(base + (int64_t)blockDim.x * NE_PER, (int64_t)nel) - 1;
    const int nf = nodeof[base], nl = nodeof[last];
    block_node = (nf == nl && !frozen[nf]) ? nf : -1;
  }
  __syncthreads();
  const int lane = threadIdx.x & (RSB_WAVE - 1), wid = threadIdx.x / RSB_WAVE;
  const int bn = block_node;
  if (bn >= 0) {
    // every position of the workgroup lies in node bn (nodes are ranges)
    int e[NE_PER];
#pragma unroll
    for (int u = 0; u < NE_PER; u++) {
      const int64_t i = base + threadIdx.x + (int64_t)u * blockDim.x;
      e[u] = i < nel ? order[i] : -1;
    }
    for (int d = 0; d < ndim; d++) {
      uint64_t mn = ~0ull, mx = 0ull;
#pragma unroll
      for (int u = 0; u < NE_PER; u++)
        if (e[u] >= 0) {
          const uint64_t kv = dkey(coord[(size_t)d * nel + e[u]]);
          mn = kv < mn ? kv : mn;
          mx = kv > mx ? kv : mx;
        }
      for (int o = RSB_WAVE / 2; o > 0; o >>= 1) {
        uint64_t a = __shfl_xor(mn, o, RSB_WAVE), b = __shfl_xor(mx, o, RSB_WAVE);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
      }
      if (lane == 0) smin[wid][d] = mn, smax[wid][d] = mx;
    }
    __syncthreads();
    if (threadIdx.x < ndim) {
      const int d = threadIdx.x;
      uint64_t mn = smin[0][d], mx = smax[0][d];
      for (int w = 1; w < (int)blockDim.x / RSB_WAVE; w++) {
        mn = smin[w][d] < mn ? smin[w][d] : mn;
        mx = smax[w][d] > mx ? smax[w][d] : mx;
      }
      atomicMin(&ext[(size_t)(2 * d) * K + bn], (unsigned long long)mn);
      atomicMax(&ext[(size_t)(2 * d + 1) * K + bn], (unsigned long long)mx);
    }
    return;
  }
  for (int u = 0; u < NE_PER; u++) {
    const int64_t i = base + threadIdx.x + (int64_t)u * blockDim.x;
    bool act = i < nel;
    const int node = act ? nodeof[i] : -1;
    if (act && frozen[node]) act = false;
    const int e = act ? order[i] : 0;
    for (int d = 0; d < ndim; d++) {
      const uint64_t kv = act ? dkey(coord[(size_t)d * nel + e]) : 0;
      uint64_t mn = act ? kv : ~0ull, mx = act ? kv : 0ull;
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
                           int *__restrict__ split, int *__restrict__ stat) {
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
  // stat[0]: longest child that still splits; stat[1]: some such child has
  // more ranks than elements (then empty nodes can appear below it)
  if (stat) {
    if (nl > 1) {
      atomicMax(&stat[0], nl);
      if (pl > nl) atomicOr(&stat[1], 1);
    }
    if (n - nl > 1) {
      atomicMax(&stat[0], n - nl);
      if (pu > n - nl) atomicOr(&stat[1], 1);
    }
  }
}

__global__ void k_nodeof_next(int nel, const int *__restrict__ first,
                              const int *__restrict__ split, int *__restrict__ nodeof) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int k = nodeof[i];
  nodeof[i] = first[k] + (i >= split[k] ? 1 : 0);
}

// Tail of the RCB tree: once no node holds more than RCB_TAIL elements, one
// workgroup finishes each node in LDS, depth by depth, with the same rules as
// the global rounds: per sub-node extents (coordinate min/max), axis of the
// largest extent (ties to the lower axis), a stable sort along it (bitonic
// sort of (sub-node, dense rank, position) keys, so ties keep their order),
// and the split after the lower ranks' share (p > 1) or at n / 2 (p == 1).
#define RCB_TAIL 1024
#define RCB_TAIL_T 256
// depths where no splitting sub-node is larger than this rank by counting
// instead of the 55-stage bitonic network over the whole node (environment
// PARRSB_AMD_RCB_RANKMAX overrides; 0 always sorts)
#define RCB_RANKMAX 64

// Exclusive prefix sum of v[0..n) (n <= RCB_TAIL) in place by the block;
// returns the total.  Each thread owns RCB_TAIL / RCB_TAIL_T consecutive
// items.
__device__ inline int tail_scan(int *v, int n, int *wsum) {
  constexpr int PER = RCB_TAIL / RCB_TAIL_T;
  const int t = threadIdx.x, lane = t & (RSB_WAVE - 1), wid = t / RSB_WAVE;
  int loc[PER], run = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = t * PER + k;
    loc[k] = i < n ? v[i] : 0;
    run += loc[k];
  }
  // inclusive scan of the thread totals inside the wave
  int x = run;
  for (int o = 1; o < RSB_WAVE; o <<= 1) {
    const int y = __shfl_up(x, o, RSB_WAVE);
    if (lane >= o) x += y;
  }
  if (lane == RSB_WAVE - 1) wsum[wid] = x;
  __syncthreads();
  int before = 0, total = 0;
  for (int w = 0; w < RCB_TAIL_T / RSB_WAVE; w++) {
    if (w < wid) before += wsum[w];
    total += wsum[w];
  }
  int acc = before + x - run;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = t * PER + k;
    if (i < n) v[i] = acc;
    acc += loc[k];
  }
  __syncthreads();
  return total;
}

__global__ void __launch_bounds__(RCB_TAIL_T) k_rcb_tail(nodes_t nd, int K, int *__restrict__ order,
                                                        const double *__restrict__ coord,
                                                        const uint32_t *__restrict__ rank,
                                                        int nel, int ndim, int rankmax) {
  constexpr int PER = RCB_TAIL / RCB_TAIL_T;
  const int k = blockIdx.x;
  if (k >= K || nd.frozen[k]) return;
  const int start = nd.start[k], n = nd.len[k];
  __shared__ int el[RCB_TAIL];
  __shared__ uint64_t key[RCB_TAIL];
  __shared__ short sub[RCB_TAIL];
  // sub-node tables: start, length, ranks; scratch for the prefix sums
  __shared__ int s_start[RCB_TAIL], s_len[RCB_TAIL], s_p[RCB_TAIL], s_sc[RCB_TAIL];
  __shared__ unsigned long long s_ext[6][RCB_TAIL / 2];
  __shared__ signed char s_axis[RCB_TAIL];
  __shared__ int wsum[RCB_TAIL_T / RSB_WAVE];
  __shared__ int nsub, s_max;
  for (int i = threadIdx.x; i < n; i += RCB_TAIL_T) el[i] = order[start + i], sub[i] = 0;
  if (threadIdx.x == 0) s_start[0] = 0, s_len[0] = n, s_p[0] = nd.p[k], nsub = 1;
  __syncthreads();
  int npow = 1;
  while (npow < n) npow <<= 1;
  for (int depth = 0; depth < 64; depth++) {
    const int ns = nsub;
    // compact index of the sub-nodes that still split (>= 2 elements: at
    // most n / 2 of them)
    for (int t = threadIdx.x; t < ns; t += RCB_TAIL_T) s_sc[t] = s_len[t] >= 2;
    __syncthreads();
    const int nsplit = tail_scan(s_sc, ns, wsum);
    if (nsplit == 0) break;
    {
      // the widest sub-node that splits (picks the sort below)
      int m = 0;
      for (int t = threadIdx.x; t < ns; t += RCB_TAIL_T) m = max(m, s_len[t]);
      for (int o = RSB_WAVE / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, RSB_WAVE));
      if (threadIdx.x == 0) s_max = 0;
      __syncthreads();
      if ((threadIdx.x & (RSB_WAVE - 1)) == 0) atomicMax(&s_max, m);
    }
    for (int j = threadIdx.x; j < nsplit; j += RCB_TAIL_T)
      for (int d = 0; d < 3; d++) s_ext[2 * d][j] = ~0ull, s_ext[2 * d + 1][j] = 0ull;
    __syncthreads();
    {
      // each thread takes PER consecutive elements; runs of one sub-node are
      // reduced in registers, a wave whose elements all lie in one sub-node
      // reduces by shuffles, so the LDS atomics stay few (no contention on
      // the shallow depths, where a sub-node spans many waves)
      const int i0 = threadIdx.x * PER, lane = threadIdx.x & (RSB_WAVE - 1);
      const int tf = i0 < n ? sub[i0] : -1;
      const int tl = i0 + PER - 1 < n ? sub[i0 + PER - 1] : -2;
      const int wfirst = __shfl(tf, 0, RSB_WAVE), wlast = __shfl(tl, RSB_WAVE - 1, RSB_WAVE);
      const bool whole = wfirst >= 0 && wfirst == wlast && s_len[wfirst] >= 2;
      for (int d = 0; d < ndim; d++) {
        unsigned long long mn = ~0ull, mx = 0ull;
        int cur = -1;
#pragma unroll
        for (int q = 0; q < PER; q++) {
          const int i = i0 + q;
          const int t = i < n ? sub[i] : -1;
          if (t != cur) {
            if (!whole && cur >= 0 && s_len[cur] >= 2) {
              atomicMin(&s_ext[2 * d][s_sc[cur]], mn);
              atomicMax(&s_ext[2 * d + 1][s_sc[cur]], mx);
            }
            cur = t, mn = ~0ull, mx = 0ull;
          }
          if (t >= 0) {
            const unsigned long long v = dkey(coord[(size_t)d * nel + el[i]]);
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
          }
        }
        if (whole) {
          for (int o = RSB_WAVE / 2; o > 0; o >>= 1) {
            const unsigned long long a = __shfl_xor(mn, o, RSB_WAVE), b = __shfl_xor(mx, o, RSB_WAVE);
            mn = a < mn ? a : mn;
            mx = b > mx ? b : mx;
          }
          if (lane == 0) {
            atomicMin(&s_ext[2 * d][s_sc[wfirst]], mn);
            atomicMax(&s_ext[2 * d + 1][s_sc[wfirst]], mx);
          }
        } else if (cur >= 0 && s_len[cur] >= 2) {
          atomicMin(&s_ext[2 * d][s_sc[cur]], mn);
          atomicMax(&s_ext[2 * d + 1][s_sc[cur]], mx);
        }
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ns; t += RCB_TAIL_T) {
      int a = -1;
      if (s_len[t] >= 2) {
        const int j = s_sc[t];
        double len[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < ndim; d++)
          len[d] = dkey_inv(s_ext[2 * d + 1][j]) - dkey_inv(s_ext[2 * d][j]);
        a = 0;
        if (len[1] > len[0]) a = 1;
        if (ndim == 3 && len[2] > len[a]) a = 2;
      }
      s_axis[t] = (signed char)a;
    }
    __syncthreads();
    if (s_max <= rankmax) {
      // every sub-node that splits is small: an element's place in its
      // sub-node is the number of (dense rank, position) keys below its own
      // there, which is the order the stable sort gives
      for (int i = threadIdx.x; i < n; i += RCB_TAIL_T) {
        const int a = s_axis[sub[i]];
        const uint64_t r = a < 0 ? 0ull : (uint64_t)rank[(size_t)a * nel + el[i]];
        key[i] = (r << 10) | (uint64_t)i;
      }
      __syncthreads();
      int dst[PER], mv[PER];
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int i = threadIdx.x + q * RCB_TAIL_T;
        dst[q] = i, mv[q] = 0;
        if (i >= n) continue;
        mv[q] = el[i];
        const int t = sub[i], ln = s_len[t];
        if (ln < 2) continue;
        const int st = s_start[t];
        const uint64_t ki = key[i];
        int c = 0;
        for (int j = st; j < st + ln; j++) c += key[j] < ki;
        dst[q] = st + c;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PER; q++)
        if (threadIdx.x + q * RCB_TAIL_T < n) el[dst[q]] = mv[q];
    } else {
    // keys: sub-node (11 bits) | dense rank (32 bits) | position (10 bits)
    for (int i = threadIdx.x; i < npow; i += RCB_TAIL_T) {
      if (i < n) {
        const int t = sub[i], a = s_axis[t];
        const uint64_t r = a < 0 ? 0ull : (uint64_t)rank[(size_t)a * nel + el[i]];
        key[i] = ((uint64_t)t << 42) | (r << 10) | (uint64_t)i;
      } else {
        key[i] = ~0ull;
      }
    }
    __syncthreads();
    // bitonic network: thread work is indexed by compare pair (npow / 2 of
    // them per stage), so no lane idles on the l < i half
    for (int kk = 2; kk <= npow; kk <<= 1)
      for (int j = kk >> 1; j > 0; j >>= 1) {
        for (int pr = threadIdx.x; pr < (npow >> 1); pr += RCB_TAIL_T) {
          const int i = ((pr & ~(j - 1)) << 1) | (pr & (j - 1)), l = i + j;
          const uint64_t a = key[i], b = key[l];
          const bool up = (i & kk) == 0;
          if ((a > b) == up) key[i] = b, key[l] = a;
        }
        __syncthreads();
      }
    {
      int mv[PER];
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int i = threadIdx.x + q * RCB_TAIL_T;
        mv[q] = i < n ? el[key[i] & 1023u] : 0;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int i = threadIdx.x + q * RCB_TAIL_T;
        if (i < n) el[i] = mv[q];
      }
    }
    }
    // children: 2 per sub-node that splits, 1 otherwise
    for (int t = threadIdx.x; t < ns; t += RCB_TAIL_T) s_sc[t] = s_len[t] >= 2 ? 2 : 1;
    __syncthreads();
    const int ns2 = tail_scan(s_sc, ns, wsum);  // s_sc[t] = first child of t
    int cs[PER][2], cl[PER][2], cp[PER][2], cf[PER], two[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int t = threadIdx.x + q * RCB_TAIL_T;
      two[q] = 0;
      if (t >= ns) continue;
      const int st = s_start[t], ln = s_len[t], pp = s_p[t];
      cf[q] = s_sc[t];
      if (ln >= 2) {
        const int nl = pp > 1 ? (int)lower_count_dev(ln, pp) : ln / 2;
        cs[q][0] = st, cl[q][0] = nl, cp[q][0] = pp > 1 ? (pp + 1) / 2 : 1;
        cs[q][1] = st + nl, cl[q][1] = ln - nl, cp[q][1] = pp > 1 ? pp / 2 : 1;
        two[q] = 2;
      } else {
        cs[q][0] = st, cl[q][0] = ln, cp[q][0] = pp;
        two[q] = 1;
      }
    }
    // element -> child sub-node (reads the old tables)
    for (int i = threadIdx.x; i < n; i += RCB_TAIL_T) {
      const int t = sub[i], ln = s_len[t];
      int c = s_sc[t];
      if (ln >= 2) {
        const int pp = s_p[t];
        const int nl = pp > 1 ? (int)lower_count_dev(ln, pp) : ln / 2;
        if (i >= s_start[t] + nl) c++;
      }
      sub[i] = (short)c;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++)
      for (int h = 0; h < two[q]; h++)
        s_start[cf[q] + h] = cs[q][h], s_len[cf[q] + h] = cl[q][h], s_p[cf[q] + h] = cp[q][h];
    if (threadIdx.x == 0) nsub = ns2;
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += RCB_TAIL_T) order[start + i] = el[i];
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
  RSB_TRY(RSB_SCRATCH(buf, sizeof(int) * cap * 13 + sizeof(int) * (size_t)nel));
  nodes_t A{buf, buf + cap, buf + 2 * cap, buf + 3 * cap};
  nodes_t B{buf + 4 * cap, buf + 5 * cap, buf + 6 * cap, buf + 7 * cap};
  int *cnt = buf + 8 * cap, *first = buf + 9 * cap, *split = buf + 10 * cap;
  int *axis = buf + 11 * cap, *nodeof = buf + 13 * cap;
  uint64_t *ext = nullptr, *key = nullptr;
  RSB_TRY(RSB_SCRATCH(ext, sizeof(uint64_t) * 6 * cap));
  RSB_TRY(RSB_SCRATCH(key, sizeof(uint64_t) * ((size_t)nel + 1)));
  int K = nseg;
  int *stat = nullptr;
  RSB_TRY(RSB_SCRATCH(stat, sizeof(int) * 2));
  // the tail kernel takes over once every splitting node is small and has
  // no more ranks than elements
  const char *tv = getenv("PARRSB_AMD_RCB_TAIL");
  bool tail_ok = !(tv && atoi(tv) == 0);
  int maxlen = 0, pbig = 0;
  for (int s = 0; s < nseg; s++)
    if (g->h_p[s] > 1 && g->h_n[s] > 1) {
      maxlen = std::max(maxlen, g->h_n[s]);
      pbig |= g->h_p[s] > g->h_n[s];
    }
  k_nodes_from_segments<<<nblocks(nseg, 256), 256, 0, g->stream>>>(g->d_seg, nseg, A);
  RSB_LAUNCH_CHECK();
  k_nodeof_init<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_seg, nseg, nel, nodeof);
  RSB_LAUNCH_CHECK();
  for (int r = 0; r < rounds; r++) {
    if (tail_ok && maxlen <= RCB_TAIL && !pbig) {
      const char *rm = getenv("PARRSB_AMD_RCB_RANKMAX");
      k_rcb_tail<<<K, RCB_TAIL_T, 0, g->stream>>>(A, K, g->d_order, g->d_coord, g->d_rank, nel,
                                                  g->ndim, rm ? atoi(rm) : RCB_RANKMAX);
      RSB_LAUNCH_CHECK();
      break;
    }
    k_node_extent_init<<<nblocks(K, 256), 256, 0, g->stream>>>(K, ext);
    RSB_LAUNCH_CHECK();
    k_node_extent<<<nblocks(nel, 256 * NE_PER), 256, 0, g->stream>>>(
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
    RSB_TRY(hipMemsetAsync(stat, 0, sizeof(int) * 2, g->stream));
    k_children<<<nblocks(K, 256), 256, 0, g->stream>>>(K, A, first, B, split, stat);
    RSB_LAUNCH_CHECK();
    int hstat[2];
    RSB_TRY(hipMemcpyAsync(hstat, stat, sizeof(int) * 2, hipMemcpyDeviceToHost, g->stream));
    k_nodeof_next<<<nblocks(nel, 256), 256, 0, g->stream>>>(nel, first, split, nodeof);
    RSB_LAUNCH_CHECK();
    RSB_TRY(hipStreamSynchronize(g->stream));
    maxlen = hstat[0], pbig = hstat[1];
    K = K2;
    std::swap(A, B);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// RIB rounds (rib.c:6-111): the same node tree as RCB, but every node is
// sorted by the projection of its centroids onto the dominant eigenvector of
// their centred covariance.  The reference finds that vector with <= 100
// power steps from a random start (fiedler.c:38-73), so its sign and, to
// ~1e-6, its direction vary from run to run; here it is the exact dominant
// eigenvector (cyclic Jacobi), signed so that its largest component is
// positive.  Sums are compensated (two-sum), so node means and covariances
// are nearly independent of the summation order.  In 2D the covariance is
// 2 x 2 (the reference passes its 3 x 3 array with N = 2, rib.c:47).
#pragma clang fp contract(off)
#define RIB_SMALL 256  // nodes up to this size: one thread per node
#define RIB_CHUNK 2048 // larger nodes: chunks of this many positions

struct dsum {
  double s = 0.0, c = 0.0;
  __device__ inline void add(double x) {
    double t = s + x, bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  __device__ inline void merge(double os, double oc) {
    add(os);
    c += oc;
  }
  __device__ inline double value() const { return s + c; }
};

// Block-wide merge of NQ compensated sums, fixed tree; result in thread 0.
template <int NQ>
__device__ inline void block_dsum(dsum (&v)[NQ], double *red) {
  const int lane = threadIdx.x & (RSB_WAVE - 1), wid = threadIdx.x / RSB_WAVE;
  const int nw = blockDim.x / RSB_WAVE;
  for (int o = RSB_WAVE / 2; o > 0; o >>= 1)
    for (int q = 0; q < NQ; q++) {
      double os = __shfl_xor(v[q].s, o, RSB_WAVE), oc = __shfl_xor(v[q].c, o, RSB_WAVE);
      v[q].merge(os, oc);
    }
  if (lane == 0)
    for (int q = 0; q < NQ; q++) red[(wid * NQ + q) * 2] = v[q].s, red[(wid * NQ + q) * 2 + 1] = v[q].c;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < NQ; q++) {
      dsum t;
      for (int w = 0; w < nw; w++) t.merge(red[(w * NQ + q) * 2], red[(w * NQ + q) * 2 + 1]);
      v[q] = t;
    }
  __syncthreads();
}

// Dominant eigenvector of the symmetric n x n (n = 2, 3) matrix A (row-major
// 3 x 3) by cyclic Jacobi rotations (Golub & Van Loan, Matrix Computations,
// Alg. 8.5.1 and 8.5.3); ties between eigenvalues go to the lower index.
__device__ inline void rib_dominant(double (&A)[3][3], int n, double *ev) {
  double V[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  for (int sweep = 0; sweep < 32; sweep++) {
    bool rotated = false;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p][q];
        if (apq == 0.0) continue;
        const double scale = fabs(A[p][p]) + fabs(A[q][q]);
        if (fabs(apq) <= DBL_EPSILON * DBL_EPSILON * scale) continue;
        const double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
        const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                    : -1.0 / (-tau + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - sn * akq;
          A[k][q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - sn * aqk;
          A[q][k] = sn * apk + c * aqk;
        }
        A[p][q] = A[q][p] = 0.0;
        for (int k = 0; k < n; k++) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - sn * vkq;
          V[k][q] = sn * vkp + c * vkq;
        }
        rotated = true;
      }
    if (!rotated) break;
  }
  int im = 0;
  for (int i = 1; i < n; i++)
    if (A[i][i] > A[im][im]) im = i;
  int dm = 0;
  for (int d = 1; d < n; d++)
    if (fabs(V[d][im]) > fabs(V[dm][im])) dm = d;
  const double sg = V[dm][im] < 0.0 ? -1.0 : 1.0;
  for (int d = 0; d < 3; d++) ev[d] = d < n ? sg * V[d][im] : 0.0;
}

__device__ inline void rib_axis_from_cov(const double *cov6, int ndim, double *ev) {
  // cov6: xx, xy, xz, yy, yz, zz
  double A[3][3] = {{cov6[0], cov6[1], cov6[2]},
                    {cov6[1], cov6[3], cov6[4]},
                    {cov6[2], cov6[4], cov6[5]}};
  rib_dominant(A, ndim, ev);
}

// Small nodes: one thread computes mean, covariance and axis of its node.
// ax[6k..6k+2] = mean, ax[6k+3..6k+5] = axis.
__global__ void k_rib_small(int K, const nodes_t nd, const int *__restrict__ order,
                            const double *__restrict__ coord, int nel, int ndim,
                            double *__restrict__ ax) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K || nd.frozen[k] || nd.len[k] > RIB_SMALL) return;
  const int st = nd.start[k], n = nd.len[k];
  dsum m[3];
  for (int i = 0; i < n; i++) {
    const int e = order[st + i];
    for (int d = 0; d < ndim; d++) m[d].add(coord[(size_t)d * nel + e]);
  }
  double mean[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < ndim; d++) mean[d] = m[d].value() / n;
  dsum cv[6];
  for (int i = 0; i < n; i++) {
    const int e = order[st + i];
    double x[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ndim; d++) x[d] = coord[(size_t)d * nel + e] - mean[d];
    cv[0].add(x[0] * x[0]), cv[1].add(x[0] * x[1]), cv[2].add(x[0] * x[2]);
    cv[3].add(x[1] * x[1]), cv[4].add(x[1] * x[2]), cv[5].add(x[2] * x[2]);
  }
  double c6[6];
  for (int q = 0; q < 6; q++) c6[q] = cv[q].value();
  double ev[3];
  rib_axis_from_cov(c6, ndim, ev);
  for (int d = 0; d < 3; d++) ax[6 * (size_t)k + d] = mean[d], ax[6 * (size_t)k + 3 + d] = ev[d];
}

__global__ void k_rib_chunk_count(int K, const nodes_t nd, int *__restrict__ cnt) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const int n = nd.len[k];
  cnt[k] = (!nd.frozen[k] && n > RIB_SMALL) ? (n + RIB_CHUNK - 1) / RIB_CHUNK : 0;
}

__global__ void k_rib_chunk_map(int K, const int *__restrict__ cnt,
                                const int *__restrict__ cfirst, int *__restrict__ cnode) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  for (int c = 0; c < cnt[k]; c++) cnode[cfirst[k] + c] = k;
}

// Partial sums of one chunk of a large node: PH 0 the coordinates, PH 1 the
// centred products (xx, xy, xz, yy, yz, zz).  part[(chunk * 6 + q) * 2 + {0,1}].
template <int PH>
__global__ void __launch_bounds__(256) k_rib_chunk_sums(const int *__restrict__ cnode,
                                                        const int *__restrict__ cfirst,
                                                        const nodes_t nd,
                                                        const int *__restrict__ order,
                                                        const double *__restrict__ coord,
                                                        int nel, int ndim,
                                                        const double *__restrict__ ax,
                                                        double *__restrict__ part) {
  constexpr int NQ = PH == 0 ? 3 : 6;
  __shared__ double red[(256 / RSB_WAVE) * NQ * 2];
  const int c = blockIdx.x, k = cnode[c];
  const int st = nd.start[k], n = nd.len[k];
  const int i0 = (c - cfirst[k]) * RIB_CHUNK, i1 = min(n, i0 + RIB_CHUNK);
  double mean[3] = {0.0, 0.0, 0.0};
  if (PH == 1)
    for (int d = 0; d < ndim; d++) mean[d] = ax[6 * (size_t)k + d];
  dsum v[NQ];
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int e = order[st + i];
    double x[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ndim; d++) x[d] = coord[(size_t)d * nel + e] - mean[d];
    if (PH == 0) {
      for (int d = 0; d < 3; d++) v[d].add(x[d]);
    } else {
      v[0].add(x[0] * x[0]), v[1].add(x[0] * x[1]), v[2].add(x[0] * x[2]);
      v[3 % NQ].add(x[1] * x[1]), v[4 % NQ].add(x[1] * x[2]), v[5 % NQ].add(x[2] * x[2]);
    }
  }
  block_dsum<NQ>(v, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < NQ; q++) {
      part[((size_t)c * 6 + q) * 2] = v[q].s;
      part[((size_t)c * 6 + q) * 2 + 1] = v[q].c;
    }
}

// Large nodes: merge the chunk partials in chunk order; PH 0 -> mean,
// PH 1 -> axis.
template <int PH>
__global__ void k_rib_large_fin(int K, const nodes_t nd, const int *__restrict__ cnt,
                                const int *__restrict__ cfirst, int ndim,
                                const double *__restrict__ part, double *__restrict__ ax) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K || cnt[k] == 0) return;
  constexpr int NQ = PH == 0 ? 3 : 6;
  dsum v[NQ];
  for (int c = cfirst[k]; c < cfirst[k] + cnt[k]; c++)
    for (int q = 0; q < NQ; q++) v[q].merge(part[((size_t)c * 6 + q) * 2], part[((size_t)c * 6 + q) * 2 + 1]);
  if (PH == 0) {
    for (int d = 0; d < 3; d++) ax[6 * (size_t)k + d] = d < ndim ? v[d].value() / nd.len[k] : 0.0;
  } else {
    double c6[6];
    for (int q = 0; q < 6; q++) c6[q] = v[q % NQ].value();
    double ev[3];
    rib_axis_from_cov(c6, ndim, ev);
    for (int d = 0; d < 3; d++) ax[6 * (size_t)k + 3 + d] = ev[d];
  }
}

// Projection keys; frozen nodes get key 0 (their order is kept by the stable
// sorts).  val = current position.
__global__ void k_rib_keys(int nel, int ndim, const int *__restrict__ order,
                           const int *__restrict__ nodeof, const int *__restrict__ frozen,
                           const double *__restrict__ coord, const double *__restrict__ ax,
                           uint64_t *__restrict__ key, int *__restrict__ val) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  const int k = nodeof[i], e = order[i];
  double p = 0.0;
  if (!frozen[k])
    for (int d = 0; d < ndim; d++)
      p += (coord[(size_t)d * nel + e] - ax[6 * (size_t)k + d]) * ax[6 * (size_t)k + 3 + d];
  key[i] = frozen[k] ? 0ull : dkey(p);
  val[i] = i;
}

// After the stable sort by projection: node of each moved position, and its
// element.
__global__ void k_rib_node_keys(int nel, const int *__restrict__ pos,
                                const int *__restrict__ nodeof, const int *__restrict__ order,
                                uint32_t *__restrict__ key2, int *__restrict__ val2) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nel) return;
  const int i = pos[j];
  key2[j] = (uint32_t)nodeof[i];
  val2[j] = order[i];
}
#pragma clang fp contract(fast)

static int run_rib(rsbgpu *g) {
  const int nel = g->nel, nseg = g->nseg;
  int rounds = 0;
  for (int s = 0; s < nseg; s++)
    if (g->h_p[s] > 1) rounds = std::max(rounds, rcb_depth(g->h_n[s], g->h_p[s]));
  if (rounds == 0) return 0;
  size_t cap = (size_t)nseg + (size_t)nel * (size_t)std::min(rounds, 4) / 2 + (size_t)nel + 16;
  size_t ccap = (size_t)nel / RIB_SMALL + (size_t)nel / RIB_CHUNK + (size_t)nseg + 16;
  // node tables (two generations) and per-node counts, per-position arrays,
  // and the chunk tables of the large nodes
  int *buf = nullptr;
  RSB_TRY(RSB_SCRATCH(buf, sizeof(int) * (cap * 13 + ccap + 3 * (size_t)nel)));
  nodes_t A{buf, buf + cap, buf + 2 * cap, buf + 3 * cap};
  nodes_t B{buf + 4 * cap, buf + 5 * cap, buf + 6 * cap, buf + 7 * cap};
  int *cnt = buf + 8 * cap, *first = buf + 9 * cap, *split = buf + 10 * cap;
  int *ccnt = buf + 11 * cap, *cfirst = buf + 12 * cap, *cnode = buf + 13 * cap;
  int *nodeof = cnode + ccap, *pos = nodeof + nel, *val2 = pos + nel;
  double *ax = nullptr, *part = nullptr;
  uint64_t *key = nullptr;
  RSB_TRY(RSB_SCRATCH(ax, sizeof(double) * 6 * cap));
  RSB_TRY(RSB_SCRATCH(part, sizeof(double) * 12 * ccap));
  RSB_TRY(RSB_SCRATCH(key, sizeof(uint64_t) * ((size_t)nel + 1)));
  uint32_t *key2 = (uint32_t *)key;  // the second sort reuses the key array
  hipStream_t st = g->stream;
  int K = nseg;
  k_nodes_from_segments<<<nblocks(nseg, 256), 256, 0, st>>>(g->d_seg, nseg, A);
  RSB_LAUNCH_CHECK();
  k_nodeof_init<<<nblocks(nel, 256), 256, 0, st>>>(g->d_seg, nseg, nel, nodeof);
  RSB_LAUNCH_CHECK();
  for (int r = 0; r < rounds; r++) {
    // axis of every node
    k_rib_small<<<nblocks(K, 256), 256, 0, st>>>(K, A, g->d_order, g->d_coord, nel, g->ndim,
                                                 ax);
    k_rib_chunk_count<<<nblocks(K, 256), 256, 0, st>>>(K, A, ccnt);
    RSB_LAUNCH_CHECK();
    int rc = rsb_exclusive_scan_i32(g, ccnt, cfirst, K);
    if (rc) return rc;
    int tail[2];
    RSB_TRY(hipMemcpyAsync(&tail[0], cfirst + K - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipMemcpyAsync(&tail[1], ccnt + K - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipStreamSynchronize(st));
    const int C = tail[0] + tail[1];
    if ((size_t)C > ccap) RSB_FAIL(2, "RIB chunk table overflow (%d > %zu)", C, ccap);
    if (C > 0) {
      k_rib_chunk_map<<<nblocks(K, 256), 256, 0, st>>>(K, ccnt, cfirst, cnode);
      k_rib_chunk_sums<0><<<C, 256, 0, st>>>(cnode, cfirst, A, g->d_order, g->d_coord, nel,
                                             g->ndim, ax, part);
      k_rib_large_fin<0><<<nblocks(K, 256), 256, 0, st>>>(K, A, ccnt, cfirst, g->ndim, part, ax);
      k_rib_chunk_sums<1><<<C, 256, 0, st>>>(cnode, cfirst, A, g->d_order, g->d_coord, nel,
                                             g->ndim, ax, part);
      k_rib_large_fin<1><<<nblocks(K, 256), 256, 0, st>>>(K, A, ccnt, cfirst, g->ndim, part, ax);
      RSB_LAUNCH_CHECK();
    }
    // stable sort by projection, then stable sort by node
    k_rib_keys<<<nblocks(nel, 256), 256, 0, st>>>(nel, g->ndim, g->d_order, nodeof, A.frozen,
                                                  g->d_coord, ax, key, pos);
    RSB_LAUNCH_CHECK();
    if ((rc = rsb_sort_pairs_u64(g, key, pos, nel, 0, 64))) return rc;
    k_rib_node_keys<<<nblocks(nel, 256), 256, 0, st>>>(nel, pos, nodeof, g->d_order, key2,
                                                       val2);
    RSB_LAUNCH_CHECK();
    int node_bits = 1;
    while ((1ll << node_bits) < K) node_bits++;
    if ((rc = rsb_sort_pairs_u32(g, key2, val2, nel, 0, node_bits))) return rc;
    RSB_TRY(hipMemcpyAsync(g->d_order, val2, sizeof(int) * (size_t)nel,
                           hipMemcpyDeviceToDevice, st));
    // children, as RCB
    k_child_count<<<nblocks(K, 256), 256, 0, st>>>(K, A, cnt);
    RSB_LAUNCH_CHECK();
    if ((rc = rsb_exclusive_scan_i32(g, cnt, first, K))) return rc;
    RSB_TRY(hipMemcpyAsync(&tail[0], first + K - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipMemcpyAsync(&tail[1], cnt + K - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    RSB_TRY(hipStreamSynchronize(st));
    const int K2 = tail[0] + tail[1];
    if ((size_t)K2 > cap) RSB_FAIL(2, "RIB node table overflow (%d > %zu)", K2, cap);
    k_children<<<nblocks(K, 256), 256, 0, st>>>(K, A, first, B, split, nullptr);
    RSB_LAUNCH_CHECK();
    k_nodeof_next<<<nblocks(nel, 256), 256, 0, st>>>(nel, first, split, nodeof);
    RSB_LAUNCH_CHECK();
    K = K2;
    std::swap(A, B);
  }
  RSB_TRY(hipStreamSynchronize(st));
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
  RSB_TRY(RSB_SCRATCH(key, sizeof(uint64_t) * ((size_t)nel + 1)));
  k_gid_keys<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, g->d_segof, g->d_seg,
                                                       g->nseg, nel, key);
  RSB_LAUNCH_CHECK();
  int sb = 1;
  while ((1ll << sb) < g->nseg) sb++;
  return rsb_sort_pairs_u64(g, key, g->d_order, nel, 0, 32 + sb);
}

extern "C" int rsbgpu_prepartition(rsbgpu *g, int rsb_pre) {
  RSB_TRY(hipSetDevice(g->device));
  if (g->nel == 0) return 0;
  rsb_phase_begin(g);
  int rc = 0;
  if (rsb_pre == 1) {
    if (!g->has_coord) RSB_FAIL(3, "RCB needs coordinates");
    rc = run_rcb(g);
  } else if (rsb_pre == 2) {
    if (!g->has_coord) RSB_FAIL(3, "RIB needs coordinates");
    rc = run_rib(g);
  } else if (rsb_pre == 0) {
    rc = run_gid(g);
  } else {
    RSB_FAIL(3, "rsb_pre=%d is not supported by this build", rsb_pre);
  }
  rsb_phase_end(g, PH_PRE);
  return rc;
}
