// LDS-tiled form of the level Laplacian, used by the Lanczos apply kernel.
//
// A tile is the block of up to RSB_TILE consecutive rows one workgroup
// owns.  Its "halo" is the set of rows outside the tile that its rows
// reference; the apply kernel loads tile + halo values into LDS once and
// reads every neighbour from LDS (window slot i < nrows: tile row i; then
// the halo rows in ascending row order).  Entries are re-encoded relative to
// that LDS window as 16-bit words: local index (12 bits) | weight (4 bits), so a
// tile plus its halo may hold at most 4096 rows; a tile whose halo does not
// fit is marked (nhalo = -1) and its rows use the global 32-bit form.
//
// Layout of the 16-bit entries: per 64-row slice, its width w rounded up to
// an even number; the first 8 * (w / 8) entries in groups of 8, entry j of
// lane l at (j / 8) * 512 + l * 8 + (j % 8), so one 16-byte load per lane
// fetches 8 consecutive entries of its row; then a group of 4, if w % 8 >= 4,
// at (w / 8) * 512 + l * 4 + (j % 4), one 8-byte load per lane; then a group
// of 2, if w % 4 == 2, after it (l * 2 + (j % 2)), one 4-byte load per lane.
// (A hex row has 26 neighbours: 3 groups of 8 and one of 2, no padding.)
//
// The same entries in 12 bits: on a conforming mesh every weight (shared
// vertices) is 4 (face), 2 (edge) or 1 (corner), so a row's entries ordered
// by weight class (4s, 2s, 1s; column order within a class) need only their
// local index.  Each class has the same number of slots in every row of a
// slice (the largest count among its 64 rows; per slice n4 | n2 << 8 |
// n1 << 16), so the weight of a slot is known to the whole wave.  Per slice
// of S = n4 + n2 + n1 slots: S / 21 blocks of 21 indices in 32 bytes per
// lane (index k of a block at bit 12 k of its eight words; two 16-byte words
// per lane, row-interleaved like the 16-bit groups), then a tail of 5
// indices in 8 bytes or 10 in 16 bytes (more than 10 left over take another
// block); unused slots hold the row's own index, which adds zero.  A hex
// row: one block and a tail of 5, 40 bytes against 52.  A level where some
// weight is not 1, 2 or 4 or a slice needs more than 42 slots keeps only
// the 16-bit form (e12_ok = 0).
#include <limits.h>
#include <stdlib.h>

#include <vector>

#include "state.hpp"

#define TB_HASH 8192  // LDS hash slots per tile (power of two)
#define E12_BLK 21    // 12-bit indices per 32-byte block
#define E12_MAXW 42   // widest slice with the 12-bit form
#define TB_LOCAL 4096 // local index range (12 bits)

__device__ inline unsigned tb_hash(unsigned c) { return (c * 2654435761u) & (TB_HASH - 1); }

// Insert c into the LDS set; returns false if the set is full.
__device__ inline bool tb_insert(int *keys, int c) {
  unsigned h = tb_hash((unsigned)c);
  for (int probe = 0; probe < TB_HASH; probe++) {
    int old = atomicCAS(&keys[h], -1, c);
    if (old == -1 || old == c) return true;
    h = (h + 1) & (TB_HASH - 1);
  }
  return false;
}

__device__ inline int tb_find(const int *keys, int c) {
  unsigned h = tb_hash((unsigned)c);
  for (int probe = 0; probe < TB_HASH; probe++) {
    int k = keys[h];
    if (k == c) return (int)h;
    if (k == -1) return -1;
    h = (h + 1) & (TB_HASH - 1);
  }
  return -1;
}

// Collect the external columns of tile blockIdx.x into the LDS set.
__device__ inline bool tb_collect(int *keys, int row0, int row1,
                                  const int64_t *__restrict__ slice_off,
                                  const int *__restrict__ slice_w,
                                  const uint32_t *__restrict__ ell) {
  for (int i = threadIdx.x; i < TB_HASH; i += blockDim.x) keys[i] = -1;
  __syncthreads();
  bool ok = true;
  for (int row = row0 + threadIdx.x; row < row1; row += blockDim.x) {
    const int sl = row >> 6;
    const uint32_t *e = ell + slice_off[sl] + (row & 63);
    for (int j = 0; j < slice_w[sl]; j++) {
      uint32_t q = e[(size_t)j * RSB_WAVE];
      int c = (int)(q & 0x0fffffffu);
      if ((q >> 28) == 0 || (c >= row0 && c < row1)) continue;
      if (!tb_insert(keys, c)) ok = false;
    }
  }
  __shared__ int any_fail;
  if (threadIdx.x == 0) any_fail = 0;
  __syncthreads();
  if (!ok) atomicOr(&any_fail, 1);
  __syncthreads();
  return any_fail == 0;
}

// Block-wide exclusive scan of per-thread counts (deterministic).
__device__ inline int tb_scan(int v, int *total) {
  __shared__ int s[RSB_BLOCK];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < (int)blockDim.x; o <<= 1) {
    int t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int incl = s[threadIdx.x];
  *total = s[blockDim.x - 1];
  __syncthreads();
  return incl - v;
}

// Pass A: number of halo rows per tile (-1 if it does not fit).
__global__ void __launch_bounds__(RSB_BLOCK) k_tile_count(const int *__restrict__ tile,
                                                          const int64_t *__restrict__ slice_off,
                                                          const int *__restrict__ slice_w,
                                                          const uint32_t *__restrict__ ell,
                                                          int *__restrict__ nhalo) {
  __shared__ int keys[TB_HASH];
  const int t = blockIdx.x, row0 = tile[3 * t + 1], row1 = tile[3 * t + 2];
  bool ok = tb_collect(keys, row0, row1, slice_off, slice_w, ell);
  int cnt = 0;
  const int per = TB_HASH / RSB_BLOCK;
  for (int k = 0; k < per; k++) cnt += keys[threadIdx.x * per + k] >= 0;
  int total;
  tb_scan(cnt, &total);
  if (threadIdx.x == 0) nhalo[t] = (ok && (row1 - row0) + total <= TB_LOCAL) ? total : -1;
}

// Position of entry j of lane `lane` in its slice's 16-bit entries: groups
// of 8 (row-interleaved 16-byte words), then a group of 4 (8-byte words)
// and a group of 2 (4-byte words) where the width w8 has them.
__device__ inline int e16_pos(int j, int lane, int w8) {
  const int full = w8 & ~7, k = j - full, t4 = (w8 - full) & 4;
  return j < full ? (j >> 3) * 512 + lane * 8 + (j & 7)
                  : (k < t4 ? full * 64 + lane * 4 + k : full * 64 + t4 * 64 + lane * 2 + (k - t4));
}

// 21 slots (stride RSB_BLOCK in LDS) into the eight words of a block:
// index k at bit 12 k.  Unrolled, so the words stay in registers.
__device__ inline void pack21(const uint16_t *my, uint32_t (&pw)[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) pw[i] = 0u;
#pragma unroll
  for (int k = 0; k < E12_BLK; k++) {
    const uint32_t v = my[k * RSB_BLOCK];
    const int word = (12 * k) / 32, sh = (12 * k) % 32;
    pw[word] |= v << sh;
    if (sh > 20) pw[word + 1] |= v >> (32 - sh);
  }
}

// Ascending bitonic sort of a[0, n) in LDS, n a power of two.
__device__ inline void lds_bitonic(int *a, int n) {
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const int x = a[i], y = a[l];
          if ((x > y) == ((i & k) == 0)) a[i] = y, a[l] = x;
        }
      }
      __syncthreads();
    }
}

// Pass B: halo lists and the 16-bit entries.  A tile's halo rows are listed
// in ascending row order (local index nrows + rank), so the apply's staging
// gathers them at increasing addresses and neighbouring lanes share cache
// lines.
__global__ void __launch_bounds__(RSB_BLOCK) k_tile_fill(
    const int *__restrict__ tile, const int64_t *__restrict__ slice_off,
    const int *__restrict__ slice_w, const uint32_t *__restrict__ ell,
    const int *__restrict__ nhalo, const int64_t *__restrict__ halo_off,
    const int64_t *__restrict__ off16, const int *__restrict__ w16, int *__restrict__ halo,
    uint16_t *__restrict__ e16) {
  __shared__ int keys[TB_HASH];
  __shared__ short local[TB_HASH];
  __shared__ int srt[TB_LOCAL];
  const int t = blockIdx.x, row0 = tile[3 * t + 1], row1 = tile[3 * t + 2];
  const int nrows = row1 - row0;
  const bool fits = nhalo[t] >= 0;  // then nrows + halo rows <= TB_LOCAL
  if (fits) {
    tb_collect(keys, row0, row1, slice_off, slice_w, ell);
    const int per = TB_HASH / RSB_BLOCK;
    int cnt = 0;
    for (int k = 0; k < per; k++) cnt += keys[threadIdx.x * per + k] >= 0;
    int total;
    int base = tb_scan(cnt, &total);
    for (int k = 0; k < per; k++) {
      const int slot = threadIdx.x * per + k;
      if (keys[slot] >= 0) srt[base++] = keys[slot];
    }
    int n2 = 1;
    while (n2 < total) n2 <<= 1;
    for (int i = total + threadIdx.x; i < n2; i += blockDim.x) srt[i] = INT_MAX;
    __syncthreads();
    lds_bitonic(srt, n2);
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
      const int c = srt[i];
      local[tb_find(keys, c)] = (short)(nrows + i);
      halo[halo_off[t] + i] = c;
    }
    __syncthreads();
  }
  for (int row = row0 + threadIdx.x; row < row1; row += blockDim.x) {
    const int sl = row >> 6, lane = row & 63;
    const uint32_t *e = ell + slice_off[sl] + lane;
    uint16_t *o = e16 + off16[sl];
    const int w = slice_w[sl], w8 = w16[sl];
    const uint16_t self = (uint16_t)(row - row0);
    for (int j = 0; j < w8; j++) {
      uint16_t v = self;  // padding: own row, weight 0
      if (fits && j < w) {
        uint32_t q = e[(size_t)j * RSB_WAVE];
        int c = (int)(q & 0x0fffffffu);
        uint32_t wt = q >> 28;
        if (wt) {
          int li = c - row0;
          if (c < row0 || c >= row1) {
            int slot = tb_find(keys, c);
            li = slot >= 0 ? local[slot] : -1;
          }
          // every external column was inserted above, so li >= 0
          if (li >= 0) v = (uint16_t)(li | (wt << 12));
        }
      }
      o[e16_pos(j, lane, w8)] = v;
    }
  }
}

// Pass C: the 12-bit form from a tile's 16-bit entries, a thread per row:
// the row's local indices class by class (4s, 2s, 1s, each class in its
// region of the slice's size for it, the rest of the region set to the
// row's own index) into its LDS slots, then packed 21 to a block.
#define E12_SLOTS (3 * E12_BLK)
__global__ void __launch_bounds__(RSB_BLOCK) k_tile_e12(
    const int *__restrict__ tile, const int *__restrict__ slice_w,
    const int64_t *__restrict__ off16, const int *__restrict__ w16,
    const uint16_t *__restrict__ e16, const int64_t *__restrict__ off12,
    const int *__restrict__ w12, const int *__restrict__ c12, uint32_t *__restrict__ e12,
    int *__restrict__ bad) {
  __shared__ uint16_t lst[E12_SLOTS * RSB_BLOCK];
  uint16_t *my = lst + threadIdx.x;  // slot k at my[k * RSB_BLOCK]
  const int t = blockIdx.x, row0 = tile[3 * t + 1], row1 = tile[3 * t + 2];
  for (int row = row0 + threadIdx.x; row < row1; row += blockDim.x) {
    const int sl = row >> 6, lane = row & 63;
    const int info = w12[sl], nb = info & 0xff, tw = info >> 8;
    if (nb > E12_MAXW / E12_BLK) continue;  // too many slots (k_w12 flagged it)
    const uint16_t *o = e16 + off16[sl];
    const int w = slice_w[sl], w8 = w16[sl], cw = c12[sl];
    const uint16_t self = (uint16_t)(row - row0);
    int s0 = 0, over = 0;
#pragma unroll
    for (int cls = 0; cls < 3; cls++) {
      const unsigned want = 4u >> cls;
      const int end = s0 + ((cw >> (8 * cls)) & 0xff);
      int s = s0;
      for (int j = 0; j < w; j++) {
        const uint16_t e = o[e16_pos(j, lane, w8)];
        if ((unsigned)(e >> 12) != want) continue;
        if (s < end && s < E12_SLOTS) my[s * RSB_BLOCK] = e & 0xfffu;
        s++;
      }
      over |= s > end;
      for (; s < end && s < E12_SLOTS; s++) my[s * RSB_BLOCK] = self;
      s0 = end;
    }
    if (over) atomicOr(bad, 1);
    const int upto = (nb + 1) * E12_BLK;  // the tail packs from a whole group of slots
    for (int k = s0; k < upto && k < E12_SLOTS; k++) my[k * RSB_BLOCK] = self;
    uint32_t *base = e12 + off12[sl];
    for (int b = 0; b < nb; b++) {
      uint32_t pw[8];
      pack21(my + b * E12_BLK * RSB_BLOCK, pw);
      *(uint4 *)(base + (size_t)(2 * b) * 256 + lane * 4) = make_uint4(pw[0], pw[1], pw[2], pw[3]);
      *(uint4 *)(base + (size_t)(2 * b + 1) * 256 + lane * 4) =
          make_uint4(pw[4], pw[5], pw[6], pw[7]);
    }
    if (tw) {
      uint32_t pw[8];
      pack21(my + nb * E12_BLK * RSB_BLOCK, pw);
      uint32_t *tail = base + (size_t)nb * 512;
      if (tw == 2) *(uint2 *)(tail + lane * 2) = make_uint2(pw[0], pw[1]);
      else *(uint4 *)(tail + lane * 4) = make_uint4(pw[0], pw[1], pw[2], pw[3]);
    }
  }
}

__global__ void k_w16(int64_t nslices, const int *__restrict__ w, int *__restrict__ w8,
                      int64_t *__restrict__ bytes) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslices) {
    int v = (w[i] + 1) / 2 * 2;  // even: groups of 8, then of 4 and 2
    w8[i] = v;
    bytes[i] = (int64_t)v * RSB_WAVE;
  }
  if (i == nslices) bytes[i] = 0;
}

// Per slice, a wave of 64 threads (one per row): the largest count of
// weight-4, weight-2 and weight-1 entries over its rows, c12 = n4 | n2 << 8 |
// n1 << 16 (each class gets that many slots in every row of the slice, so
// the class bounds are uniform in a wave), and the layout of its 12-bit form
// for n4 + n2 + n1 slots: blocks and tail words (info = nb | tw << 8), words
// for all 64 lanes.  A weight other than 0, 1, 2 or 4, or more than
// E12_MAXW slots, marks the level.
__global__ void __launch_bounds__(256) k_w12(int64_t nslices, const int64_t *__restrict__ slice_off,
                                             const int *__restrict__ slice_w,
                                             const uint32_t *__restrict__ ell,
                                             int *__restrict__ c12, int *__restrict__ info,
                                             int64_t *__restrict__ words, int *__restrict__ bad) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t sl = row >> 6;  // uniform in a wave
  const int lane = (int)(row & 63);
  if (sl > nslices) return;
  if (sl == nslices) {
    if (lane == 0) words[sl] = 0;
    return;
  }
  const uint32_t *e = ell + slice_off[sl] + lane;
  const int w = slice_w[sl];
  int c4 = 0, c2 = 0, c1 = 0, odd = 0;
  for (int j = 0; j < w; j++) {
    const unsigned wt = e[(size_t)j * RSB_WAVE] >> 28;
    c4 += wt == 4, c2 += wt == 2, c1 += wt == 1;
    odd |= wt != 0 && wt != 1 && wt != 2 && wt != 4;
  }
  for (int o = 32; o > 0; o >>= 1) {
    c4 = max(c4, __shfl_xor(c4, o));
    c2 = max(c2, __shfl_xor(c2, o));
    c1 = max(c1, __shfl_xor(c1, o));
    odd = max(odd, __shfl_xor(odd, o));
  }
  if (lane == 0) {
    const int slots = c4 + c2 + c1;
    int nb = slots / E12_BLK, rem = slots % E12_BLK;
    if (rem > 10) nb++, rem = 0;
    const int tw = rem == 0 ? 0 : (rem <= 5 ? 2 : 4);
    info[sl] = nb | (tw << 8);
    words[sl] = (int64_t)(8 * nb + tw) * RSB_WAVE;
    c12[sl] = c4 | (c2 << 8) | (c1 << 16);
    if (odd || slots > E12_MAXW) atomicOr(bad, 1);
  }
}

__global__ void k_zero_int(int *__restrict__ p) { *p = 0; }

__global__ void k_all_wide(int n, int *__restrict__ nh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) nh[i] = -1;
}

__global__ void k_nhalo64(int n, const int *__restrict__ nh, int64_t *__restrict__ o) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = nh[i] > 0 ? nh[i] : 0;
  if (i == n) o[i] = 0;
}

int rsb_build_tiles(rsbgpu *g) {
  const int nt = g->ntiles;
  const int64_t ns = g->nslices;
  if (nt == 0) return 0;
  RSB_TRY(dev_alloc(g, g->d_tnhalo, (size_t)nt + 1));
  RSB_TRY(dev_alloc(g, g->d_thalo_off, (size_t)nt + 1));
  RSB_TRY(dev_alloc(g, g->d_w16, (size_t)ns + 1));
  RSB_TRY(dev_alloc(g, g->d_off16, (size_t)ns + 1));
  int64_t *tmp = nullptr;
  RSB_TRY(RSB_SCRATCH(tmp, sizeof(int64_t) * (std::max<int64_t>(ns, nt) + 1)));
  k_tile_count<<<nt, RSB_BLOCK, 0, g->stream>>>(g->d_tile, g->d_slice_off, g->d_slice_w,
                                                 g->d_ell, g->d_tnhalo);
  RSB_LAUNCH_CHECK();
  const char *off = getenv("PARRSB_AMD_NO_TILES");  // every tile on the global path
  if (off && atoi(off) == 1) {
    k_all_wide<<<nblocks(nt, 256), 256, 0, g->stream>>>(nt, g->d_tnhalo);
    RSB_LAUNCH_CHECK();
  }
  k_nhalo64<<<nblocks(nt + 1, 256), 256, 0, g->stream>>>(nt, g->d_tnhalo, tmp);
  RSB_LAUNCH_CHECK();
  int rc = rsb_exclusive_scan_i64(g, tmp, g->d_thalo_off, nt + 1);
  if (rc) return rc;
  k_w16<<<nblocks(ns + 1, 256), 256, 0, g->stream>>>(ns, g->d_slice_w, g->d_w16, tmp);
  RSB_LAUNCH_CHECK();
  rc = rsb_exclusive_scan_i64(g, tmp, g->d_off16, ns + 1);
  if (rc) return rc;
  int64_t nh = 0, n16 = 0;
  RSB_TRY(hipMemcpyAsync(&nh, g->d_thalo_off + nt, sizeof(int64_t), hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipMemcpyAsync(&n16, g->d_off16 + ns, sizeof(int64_t), hipMemcpyDeviceToHost,
                         g->stream));
  std::vector<int> nhv(nt);
  RSB_TRY(hipMemcpyAsync(nhv.data(), g->d_tnhalo, sizeof(int) * nt, hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  g->nhalo_total = nh;
  g->e16_count = n16;
  g->nwide = 0;
  for (int v : nhv) g->nwide += v < 0;
  RSB_TRY(dev_alloc(g, g->d_halo, (size_t)nh + 1));
  RSB_TRY(dev_alloc(g, g->d_e16, (size_t)n16 + 8));
  // the 12-bit form (unless PARRSB_AMD_E12=0; =2 lets the FP64 solve read
  // it as well)
  const char *ev = getenv("PARRSB_AMD_E12");
  const bool want12 = !(ev && atoi(ev) == 0);
  g->e12_ok = 0;
  g->e12_f64 = ev && atoi(ev) == 2;
  if (want12) {
    RSB_TRY(dev_alloc(g, g->d_w12, (size_t)ns + 1));
    RSB_TRY(dev_alloc(g, g->d_c12, (size_t)ns + 1));
    RSB_TRY(dev_alloc(g, g->d_off12, (size_t)ns + 1));
    RSB_TRY(dev_alloc(g, g->d_e12_bad, 1));
    k_zero_int<<<1, 1, 0, g->stream>>>(g->d_e12_bad);
    k_w12<<<nblocks((ns + 1) * RSB_WAVE, 256), 256, 0, g->stream>>>(
        ns, g->d_slice_off, g->d_slice_w, g->d_ell, g->d_c12, g->d_w12, tmp, g->d_e12_bad);
    RSB_LAUNCH_CHECK();
    rc = rsb_exclusive_scan_i64(g, tmp, g->d_off12, ns + 1);
    if (rc) return rc;
    int64_t n12 = 0;
    RSB_TRY(hipMemcpyAsync(&n12, g->d_off12 + ns, sizeof(int64_t), hipMemcpyDeviceToHost,
                           g->stream));
    RSB_TRY(hipStreamSynchronize(g->stream));
    RSB_TRY(dev_alloc(g, g->d_e12, (size_t)n12 + 8));
  }
  k_tile_fill<<<nt, RSB_BLOCK, 0, g->stream>>>(g->d_tile, g->d_slice_off, g->d_slice_w, g->d_ell,
                                                g->d_tnhalo, g->d_thalo_off, g->d_off16,
                                                g->d_w16, g->d_halo, g->d_e16);
  RSB_LAUNCH_CHECK();
  if (want12) {
    k_tile_e12<<<nt, RSB_BLOCK, 0, g->stream>>>(g->d_tile, g->d_slice_w, g->d_off16, g->d_w16,
                                                 g->d_e16, g->d_off12, g->d_w12, g->d_c12,
                                                 g->d_e12, g->d_e12_bad);
    RSB_LAUNCH_CHECK();
    int bad = 1;
    RSB_TRY(hipMemcpyAsync(&bad, g->d_e12_bad, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    RSB_TRY(hipStreamSynchronize(g->stream));
    g->e12_ok = bad == 0;
  }
  RSB_TRY(hipStreamSynchronize(g->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// Vertex form of the tiles: the Laplacian as the reference applies it, by
// gather-scatter over vertices (laplacian.c:159-205):
//   (L q)_e = sum over the slots s of e of (mult(v_s) q_e - S(v_s)),
//   S(v) = sum of q over the slots of the segment's elements at vertex v,
//   mult(v) = the number of those slots.
// Every element of the segment that touches a vertex of a tile row is a
// neighbour of that row, so it lies in the tile's window (rows + halo): per
// tile, the distinct vertex ids of its rows (at most VT_MAXV) get local
// indices; each has a list of exactly nv entries, the window indices of the
// slots at it in ascending order, padded with VT_MAXW (the index of an LDS
// zero); each row holds the local indices of its nv vertices (0xffff: vertex
// id 0, which is never shared, as in the reference's gather-scatter).  A
// tile whose rows touch more than VT_MAXV vertices, or with more than nv
// slots at a vertex (a non-conforming mesh), has no vertex form (vt_n = -1)
// and keeps its neighbour entries.  Per row this reads one 16-byte index
// word plus one list of nv 16-bit words per vertex of the tile (about 1.4
// per row on a hex mesh), one aligned load each, against 26-28 16-bit
// neighbour entries.
#define VT_HASH 4096
#define VT_NONE 0xffffu

__device__ inline unsigned vt_hash(unsigned long long v) {
  return (unsigned)((v * 0x9E3779B97F4A7C15ull) >> 52) & (VT_HASH - 1);
}

__device__ inline bool vt_insert(unsigned long long *keys, unsigned long long v) {
  unsigned h = vt_hash(v);
  for (int probe = 0; probe < VT_HASH; probe++) {
    unsigned long long old = atomicCAS(&keys[h], 0ull, v);
    if (old == 0ull || old == v) return true;
    h = (h + 1) & (VT_HASH - 1);
  }
  return false;
}

__device__ inline int vt_find(const unsigned long long *keys, unsigned long long v) {
  unsigned h = vt_hash(v);
  for (int probe = 0; probe < VT_HASH; probe++) {
    const unsigned long long k = keys[h];
    if (k == v) return (int)h;
    if (k == 0ull) return -1;
    h = (h + 1) & (VT_HASH - 1);
  }
  return -1;
}

struct vt_in {
  const int *tile, *nhalo, *halo, *row_elem;
  const int64_t *halo_off;
  const long long *vtx;
  int nv;
};

// The window row of local index i (tile rows first, then the halo).
__device__ inline int vt_window_row(const vt_in &in, int t, int row0, int nrows, int i) {
  return i < nrows ? row0 + i : in.halo[in.halo_off[t] + (i - nrows)];
}

// Insert the vertex ids of the tile's rows; false if the table overflowed.
__device__ inline bool vt_collect(const vt_in &in, unsigned long long *keys, int row0, int row1) {
  __shared__ int fail;
  for (int i = threadIdx.x; i < VT_HASH; i += blockDim.x) keys[i] = 0ull;
  if (threadIdx.x == 0) fail = 0;
  __syncthreads();
  bool ok = true;
  for (int row = row0 + threadIdx.x; row < row1; row += blockDim.x) {
    const int e = in.row_elem[row];
    if (e < 0) continue;
    for (int j = 0; j < in.nv; j++) {
      const unsigned long long v = (unsigned long long)in.vtx[(int64_t)e * in.nv + j];
      if (v && !vt_insert(keys, v)) ok = false;
    }
  }
  if (!ok) atomicOr(&fail, 1);
  __syncthreads();
  return fail == 0;
}

// Local indices in table order (deterministic for a given table); returns
// the number of vertices.
__device__ inline int vt_number(const unsigned long long *keys, short *lid) {
  const int per = VT_HASH / RSB_BLOCK;
  int c = 0;
  for (int k = 0; k < per; k++) c += keys[threadIdx.x * per + k] != 0ull;
  int nvert = 0, base = tb_scan(c, &nvert);
  for (int k = 0; k < per; k++) {
    const int slot = threadIdx.x * per + k;
    if (keys[slot] != 0ull) lid[slot] = (short)base++;
  }
  __syncthreads();
  return nvert;
}

// Slots of the window at each local vertex, counted into vcnt (LDS); with
// list != nullptr also written there at list[v * nv + (running count)].
__device__ inline void vt_slots(const vt_in &in, int t, int row0, int nrows, int nh,
                                const unsigned long long *keys, const short *lid, int *vcnt,
                                uint16_t *list) {
  for (int i = threadIdx.x; i < nrows + nh; i += blockDim.x) {
    const int e = in.row_elem[vt_window_row(in, t, row0, nrows, i)];
    if (e < 0) continue;
    for (int j = 0; j < in.nv; j++) {
      const unsigned long long v = (unsigned long long)in.vtx[(int64_t)e * in.nv + j];
      const int slot = v ? vt_find(keys, v) : -1;
      if (slot < 0) continue;
      const int lv = lid[slot];
      const int k = atomicAdd(&vcnt[lv], 1);
      if (list && k < in.nv) list[lv * in.nv + k] = (uint16_t)i;
    }
  }
  __syncthreads();
}

// Pass A: local vertices per tile (vt_n = -1: no vertex form).
__global__ void __launch_bounds__(RSB_BLOCK) k_vt_count(vt_in in, int *__restrict__ vt_n,
                                                        int64_t *__restrict__ nlist) {
  __shared__ unsigned long long keys[VT_HASH];
  __shared__ short lid[VT_HASH];
  __shared__ int vcnt[VT_MAXV];
  __shared__ int over;
  const int t = blockIdx.x, row0 = in.tile[3 * t + 1], row1 = in.tile[3 * t + 2];
  const int nrows = row1 - row0, nh = in.nhalo[t];
  if (nh < 0 || nrows + nh > VT_MAXW) {
    if (threadIdx.x == 0) vt_n[t] = -1, nlist[t] = 0;
    return;
  }
  const bool ok = vt_collect(in, keys, row0, row1);
  const int nvert = vt_number(keys, lid);
  if (!ok || nvert > VT_MAXV) {
    if (threadIdx.x == 0) vt_n[t] = -1, nlist[t] = 0;
    return;
  }
  for (int v = threadIdx.x; v < nvert; v += blockDim.x) vcnt[v] = 0;
  if (threadIdx.x == 0) over = 0;
  __syncthreads();
  vt_slots(in, t, row0, nrows, nh, keys, lid, vcnt, nullptr);
  for (int v = threadIdx.x; v < nvert; v += blockDim.x)
    if (vcnt[v] > in.nv) atomicOr(&over, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    vt_n[t] = over ? -1 : nvert;
    nlist[t] = over ? 0 : (int64_t)nvert * in.nv;
  }
}

// Pass B: the lists (window indices ascending, padded with VT_MAXW) and the
// rows' local vertex indices.
__global__ void __launch_bounds__(RSB_BLOCK) k_vt_fill(vt_in in, const int *__restrict__ vt_n,
                                                       const int64_t *__restrict__ vt_loff,
                                                       uint16_t *__restrict__ list,
                                                       uint16_t *__restrict__ rowv) {
  __shared__ unsigned long long keys[VT_HASH];
  __shared__ short lid[VT_HASH];
  __shared__ int vcnt[VT_MAXV];
  const int t = blockIdx.x, row0 = in.tile[3 * t + 1], row1 = in.tile[3 * t + 2];
  const int nrows = row1 - row0;
  if (vt_n[t] < 0) return;
  const int nh = in.nhalo[t];
  vt_collect(in, keys, row0, row1);
  const int nvert = vt_number(keys, lid);
  for (int v = threadIdx.x; v < nvert; v += blockDim.x) vcnt[v] = 0;
  __syncthreads();
  uint16_t *Ls = list + vt_loff[t];
  vt_slots(in, t, row0, nrows, nh, keys, lid, vcnt, Ls);
  // each list ascending (its order of summation), then padded
  const int W = in.nv;
  for (int v = threadIdx.x; v < nvert; v += blockDim.x) {
    uint16_t *L = Ls + v * W;
    const int n = vcnt[v];
    for (int x = 1; x < n; x++) {
      const uint16_t key = L[x];
      int y = x - 1;
      while (y >= 0 && L[y] > key) {
        L[y + 1] = L[y];
        y--;
      }
      L[y + 1] = key;
    }
    for (int x = n; x < W; x++) L[x] = (uint16_t)VT_MAXW;
  }
  for (int row = row0 + threadIdx.x; row < row1; row += blockDim.x) {
    const int e = in.row_elem[row];
    uint16_t ids[8];
    for (int j = 0; j < 8; j++) {
      ids[j] = (uint16_t)VT_NONE;
      if (e >= 0 && j < in.nv) {
        const unsigned long long v = (unsigned long long)in.vtx[(int64_t)e * in.nv + j];
        const int slot = v ? vt_find(keys, v) : -1;
        if (slot >= 0) ids[j] = (uint16_t)lid[slot];
      }
    }
    uint4 pk;
    pk.x = ids[0] | ((unsigned)ids[1] << 16), pk.y = ids[2] | ((unsigned)ids[3] << 16);
    pk.z = ids[4] | ((unsigned)ids[5] << 16), pk.w = ids[6] | ((unsigned)ids[7] << 16);
    *(uint4 *)(rowv + (int64_t)row * 8) = pk;
  }
}

int rsb_build_vtiles(rsbgpu *g) {
  const int nt = g->ntiles;
  g->vt_ok = 0;
  // opt-in: PARRSB_AMD_VTILES=1 (FP32 Lanczos only)
  const char *on = getenv("PARRSB_AMD_VTILES");
  if (nt == 0 || !g->d_vtx || (g->nv != 4 && g->nv != 8) || g->nwide > 0 || !on ||
      atoi(on) != 1)
    return 0;
  RSB_TRY(dev_alloc(g, g->d_vt_n, (size_t)nt + 1));
  RSB_TRY(dev_alloc(g, g->d_vt_loff, (size_t)nt + 1));
  int64_t *nl = nullptr;
  RSB_TRY(RSB_SCRATCH(nl, sizeof(int64_t) * ((size_t)nt + 1)));
  vt_in in{g->d_tile, g->d_tnhalo, g->d_halo, g->d_row_elem, g->d_thalo_off, g->d_vtx, g->nv};
  RSB_TRY(hipMemsetAsync(nl, 0, sizeof(int64_t) * ((size_t)nt + 1), g->stream));
  k_vt_count<<<nt, RSB_BLOCK, 0, g->stream>>>(in, g->d_vt_n, nl);
  RSB_LAUNCH_CHECK();
  int rc = rsb_exclusive_scan_i64(g, nl, g->d_vt_loff, nt + 1);
  if (rc) return rc;
  int64_t tl = 0;
  RSB_TRY(hipMemcpyAsync(&tl, g->d_vt_loff + nt, sizeof(int64_t), hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  RSB_TRY(dev_alloc(g, g->d_vt_list, (size_t)tl + 8));
  RSB_TRY(dev_alloc(g, g->d_vt_rowv, (size_t)g->mpad * 8 + 8));
  k_vt_fill<<<nt, RSB_BLOCK, 0, g->stream>>>(in, g->d_vt_n, g->d_vt_loff, g->d_vt_list,
                                              g->d_vt_rowv);
  RSB_LAUNCH_CHECK();
  g->vt_ok = 1;
  return 0;
}

extern "C" int rsbgpu_level_form(rsbgpu *g, int *wide_tiles, int *vertex_tiles) {
  RSB_TRY(hipSetDevice(g->device));
  *wide_tiles = g->nwide;
  *vertex_tiles = 0;
  if (!g->vt_ok || g->ntiles == 0) return 0;
  std::vector<int> vn(g->ntiles);
  RSB_TRY(hipMemcpyAsync(vn.data(), g->d_vt_n, sizeof(int) * g->ntiles, hipMemcpyDeviceToHost,
                         g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  for (int v : vn) *vertex_tiles += v >= 0;
  return 0;
}

// 1 if the level's tiles carry the 12-bit form (every weight 1, 2 or 4),
// else 0
extern "C" int rsbgpu_level_e12(rsbgpu *g) { return g->e12_ok; }
