// Level Laplacian: the shared-vertex Laplacian of every segment with p > 1,
// restricted to the segment (edges to other segments vanish, as the
// reference's gather-scatter over the segment's communicator has them,
// laplacian.c:159-205).  Stored as SELL-64: rows in groups of 64 (one
// wavefront), each group padded to its longest row, entries column-major
// within the group so a wavefront's loads are contiguous.  An entry packs
// the neighbour's row (28 bits) and the shared-vertex count (4 bits); the
// diagonal is the row sum of the weights and is never stored.
#include <vector>

#include "state.hpp"

__global__ void k_row_maps(const int *__restrict__ order, const int *__restrict__ segof,
                           const int *__restrict__ seg, int nseg, int nel,
                           int *__restrict__ row_elem, int *__restrict__ elem_row) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  int e = order[i], s = segof[e];
  int pstart = seg[4 * nseg + s];
  if (pstart < 0) {
    elem_row[e] = -1;
    return;
  }
  int row = pstart + (i - seg[s]);
  row_elem[row] = e;
  elem_row[e] = row;
}

// Graph entries per batch in the level kernels: all their loads are issued
// before any is used (one dependent gather at a time left the 16.7M-row
// levels latency-bound).
#define LV_U 8

__global__ void k_fill_i32(int *__restrict__ a, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_level_rowlen(int64_t mpad, const int *__restrict__ row_elem,
                               const int *__restrict__ segof, const int64_t *__restrict__ growptr,
                               const uint32_t *__restrict__ gadj, int *__restrict__ rowlen,
                               int *__restrict__ slice_w, unsigned long long *__restrict__ nnz,
                               int64_t *__restrict__ seg_nnz) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int len = 0;
  if (row < mpad) {
    int e = row_elem[row];
    if (e >= 0) {
      const int s = segof[e];
      const int64_t k1 = growptr[e + 1];
      // LV_U graph entries, then their segments, in flight at once
      for (int64_t k0 = growptr[e]; k0 < k1; k0 += LV_U) {
        int c[LV_U];
#pragma unroll
        for (int u = 0; u < LV_U; u++)
          c[u] = k0 + u < k1 ? (int)(gadj[k0 + u] & 0x0fffffffu) : e;
#pragma unroll
        for (int u = 0; u < LV_U; u++) len += (k0 + u < k1) && segof[c[u]] == s;
      }
    }
    rowlen[row] = len;
  }
  // slice width = max over the 64 rows of the wavefront (blockDim multiple of 64)
  int m = len;
  for (int o = RSB_WAVE / 2; o > 0; o >>= 1) {
    int t = __shfl_xor(m, o, RSB_WAVE);
    m = t > m ? t : m;
  }
  unsigned long long s = wave_sum((unsigned long long)len);
  if ((threadIdx.x & (RSB_WAVE - 1)) == 0 && row < mpad) {
    slice_w[row / RSB_WAVE] = m;
    atomicAdd(nnz, s);
    // a segment's rows start 64-aligned, so the first row of a wavefront is
    // a real row of the wavefront's segment
    atomicAdd((unsigned long long *)&seg_nnz[segof[row_elem[row]]], s);
  }
}

__global__ void k_slice_bytes(int64_t nslices, const int *__restrict__ w, int64_t *__restrict__ o) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslices) o[i] = (int64_t)w[i] * RSB_WAVE;
  if (i == nslices) o[i] = 0;
}

__global__ void k_level_fill(int64_t mpad, const int *__restrict__ row_elem,
                             const int *__restrict__ segof, const int *__restrict__ elem_row,
                             const int64_t *__restrict__ growptr,
                             const uint32_t *__restrict__ gadj,
                             const int64_t *__restrict__ slice_off,
                             const int *__restrict__ slice_w, uint32_t *__restrict__ ell) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= mpad) return;
  int64_t sl = row / RSB_WAVE;
  int lane = (int)(row % RSB_WAVE);
  uint32_t *out = ell + slice_off[sl] + lane;
  int w = slice_w[sl], j = 0;
  int e = row_elem[row];
  if (e >= 0) {
    const int s = segof[e];
    const int64_t k1 = growptr[e + 1];
    for (int64_t k0 = growptr[e]; k0 < k1; k0 += LV_U) {
      uint32_t a[LV_U];
      int t[LV_U], r[LV_U];
#pragma unroll
      for (int u = 0; u < LV_U; u++) a[u] = k0 + u < k1 ? gadj[k0 + u] : (uint32_t)e;
#pragma unroll
      for (int u = 0; u < LV_U; u++) {
        const int c = (int)(a[u] & 0x0fffffffu);
        t[u] = segof[c];
        r[u] = elem_row[c];
      }
#pragma unroll
      for (int u = 0; u < LV_U; u++)
        if (k0 + u < k1 && t[u] == s) {
          out[(int64_t)j * RSB_WAVE] = (uint32_t)r[u] | (a[u] & 0xf0000000u);
          j++;
        }
    }
  }
  for (; j < w; j++) out[(int64_t)j * RSB_WAVE] = (uint32_t)row;  // weight 0
}

extern "C" int rsbgpu_level_laplacian(rsbgpu *g) {
  RSB_TRY(hipSetDevice(g->device));
  rsb_phase_begin(g);
  const int64_t mpad = g->mpad;
  const int nel = g->nel;
  g->nslices = mpad / RSB_WAVE;
  RSB_TRY(dev_alloc(g, g->d_row_elem, (size_t)mpad));
  RSB_TRY(dev_alloc(g, g->d_elem_row, (size_t)nel));
  RSB_TRY(dev_alloc(g, g->d_rowlen, (size_t)mpad));
  RSB_TRY(dev_alloc(g, g->d_slice_w, (size_t)g->nslices + 1));
  RSB_TRY(dev_alloc(g, g->d_slice_off, (size_t)g->nslices + 1));
  if (mpad == 0) {
    g->lnnz = g->lnnz_pad = 0;
    g->h_seg_nnz.assign(g->nseg, 0);
    rsb_phase_end(g, PH_LAP);
    return 0;
  }
  k_fill_i32<<<nblocks(mpad, 256), 256, 0, g->stream>>>(g->d_row_elem, mpad, -1);
  RSB_LAUNCH_CHECK();
  k_row_maps<<<nblocks(nel, 256), 256, 0, g->stream>>>(g->d_order, g->d_segof, g->d_seg, g->nseg,
                                                       nel, g->d_row_elem, g->d_elem_row);
  RSB_LAUNCH_CHECK();
  unsigned long long *dnnz = nullptr;
  RSB_TRY(RSB_SCRATCH(dnnz, sizeof(unsigned long long)));
  RSB_TRY(hipMemsetAsync(dnnz, 0, sizeof(unsigned long long), g->stream));
  RSB_TRY(dev_alloc(g, g->d_seg_nnz, (size_t)g->nseg));
  RSB_TRY(hipMemsetAsync(g->d_seg_nnz, 0, sizeof(int64_t) * g->nseg, g->stream));
  k_level_rowlen<<<nblocks(mpad, 256), 256, 0, g->stream>>>(
      mpad, g->d_row_elem, g->d_segof, g->d_growptr, g->d_gadj, g->d_rowlen, g->d_slice_w, dnnz,
      g->d_seg_nnz);
  RSB_LAUNCH_CHECK();
  int64_t *bytes = nullptr;
  RSB_TRY(RSB_SCRATCH(bytes, sizeof(int64_t) * (g->nslices + 1)));
  k_slice_bytes<<<nblocks(g->nslices + 1, 256), 256, 0, g->stream>>>(g->nslices, g->d_slice_w,
                                                                     bytes);
  RSB_LAUNCH_CHECK();
  int rc = rsb_exclusive_scan_i64(g, bytes, g->d_slice_off, g->nslices + 1);
  if (rc) return rc;
  unsigned long long hn = 0;
  RSB_TRY(hipMemcpyAsync(&g->lnnz_pad, g->d_slice_off + g->nslices, sizeof(int64_t),
                         hipMemcpyDeviceToHost, g->stream));
  RSB_TRY(hipMemcpyAsync(&hn, dnnz, sizeof(hn), hipMemcpyDeviceToHost, g->stream));
  g->h_seg_nnz.assign(g->nseg, 0);
  RSB_TRY(hipMemcpyAsync(g->h_seg_nnz.data(), g->d_seg_nnz, sizeof(int64_t) * g->nseg,
                         hipMemcpyDeviceToHost, g->stream));
  RSB_TRY(hipStreamSynchronize(g->stream));
  g->lnnz = (int64_t)hn;
  RSB_TRY(dev_alloc(g, g->d_ell, (size_t)g->lnnz_pad));
  k_level_fill<<<nblocks(mpad, 256), 256, 0, g->stream>>>(mpad, g->d_row_elem, g->d_segof,
                                                          g->d_elem_row, g->d_growptr,
                                                          g->d_gadj, g->d_slice_off,
                                                          g->d_slice_w, g->d_ell);
  RSB_LAUNCH_CHECK();
  rc = rsb_build_tiles(g);
  if (rc) return rc;
  rc = rsb_build_vtiles(g);
  if (rc) return rc;
  rsb_phase_end(g, PH_LAP);
  return 0;
}

extern "C" int rsbgpu_get_level_csr(rsbgpu *g, int64_t *mpad, int64_t *nnz, int64_t *rowptr,
                                    int *col, int *wgt, int *row_elem) {
  RSB_TRY(hipSetDevice(g->device));
  *mpad = g->mpad;
  *nnz = g->lnnz;
  if (!rowptr) return 0;
  std::vector<int> rl(g->mpad), sw(g->nslices + 1), re(g->mpad);
  std::vector<int64_t> so(g->nslices + 1);
  std::vector<uint32_t> ell(g->lnnz_pad);
  if (g->mpad) {
    RSB_TRY(hipMemcpy(rl.data(), g->d_rowlen, sizeof(int) * g->mpad, hipMemcpyDeviceToHost));
    RSB_TRY(hipMemcpy(re.data(), g->d_row_elem, sizeof(int) * g->mpad, hipMemcpyDeviceToHost));
    RSB_TRY(hipMemcpy(so.data(), g->d_slice_off, sizeof(int64_t) * (g->nslices + 1),
                      hipMemcpyDeviceToHost));
    RSB_TRY(hipMemcpy(ell.data(), g->d_ell, sizeof(uint32_t) * g->lnnz_pad,
                      hipMemcpyDeviceToHost));
  }
  int64_t k = 0;
  for (int64_t r = 0; r < g->mpad; r++) {
    rowptr[r] = k;
    int64_t sl = r / RSB_WAVE, lane = r % RSB_WAVE;
    for (int j = 0; j < rl[r]; j++) {
      uint32_t a = ell[so[sl] + (int64_t)j * RSB_WAVE + lane];
      col[k] = (int)(a & 0x0fffffffu);
      wgt[k] = (int)(a >> 28);
      k++;
    }
    if (row_elem) row_elem[r] = re[r];
  }
  rowptr[g->mpad] = k;
  return 0;
}
