// Connectivity from corner coordinates (parrsb_conn_mesh, con.c:189-295):
// every corner of every element is a point; points closer than a tolerance
// relative to their element's edge lengths get one global vertex id.
//
// The segment refinement of con-unique-vertices.c:445-505 on one GPU: all
// points start in one segment; ndim times over the ndim axes, every segment
// is sorted by that coordinate (one device-wide stable sort by the
// coordinate, then one by the segment index, so segments stay in place and
// ties keep their order) and split wherever consecutive points differ by
// more than tol^2 * min(dx) in squared distance, dx being the point's
// smallest squared distance to the corners joined to it by an element edge
// (con.c:70-101).  Segments are numbered in layout order after every pass; a
// point's id is its final segment (con.c:120-145), 1-based in the output.
// The element and face checks of con-check.c follow.  Corner orders: input
// in the preprocessor order, output in the symmetric order (con.c:3-8).
#include <cfloat>
#include <cstdio>
#include <vector>

#include "state.hpp"

__constant__ int CON_PRE_TO_SYM[8] = {0, 1, 3, 2, 4, 5, 7, 6};
__constant__ int CON_NEIGHBOR[8][3] = {{1, 2, 4}, {0, 3, 5}, {0, 3, 6}, {1, 2, 7},
                                       {0, 5, 6}, {1, 4, 7}, {2, 4, 7}, {3, 5, 6}};
__constant__ int CON_FACES3D[6][4] = {{1, 5, 7, 3}, {2, 4, 8, 6}, {1, 2, 6, 5},
                                      {3, 7, 8, 4}, {1, 3, 4, 2}, {5, 6, 8, 7}};

#pragma clang fp contract(off)
__global__ void k_con_points(const double *__restrict__ coord, int64_t nel, int ndim, int nv,
                             double *__restrict__ x, double *__restrict__ dx) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t np = nel * nv;
  if (p >= np) return;
  const int64_t e = p / nv;
  const int k = (int)(p % nv);
  double a[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < ndim; d++) a[d] = coord[(e * nv + CON_PRE_TO_SYM[k]) * ndim + d];
  for (int d = 0; d < 3; d++) x[(size_t)d * np + p] = a[d];
  double m = DBL_MAX;
  for (int q = 0; q < ndim; q++) {
    const int j = CON_PRE_TO_SYM[CON_NEIGHBOR[k][q]];
    double s = 0.0;
    for (int d = 0; d < ndim; d++) {
      const double b = coord[(e * nv + j) * ndim + d];
      s += (a[d] - b) * (a[d] - b);
    }
    m = s < m ? s : m;
  }
  dx[p] = m;
}

// layout order -> sort keys along axis d
__global__ void k_con_axis_keys(int64_t np, const int *__restrict__ lay,
                                const double *__restrict__ xd, uint64_t *__restrict__ key,
                                int *__restrict__ val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int p = lay[i];
  key[i] = dkey(xd[p]);
  val[i] = p;
}

__global__ void k_con_seg_keys(int64_t np, const int *__restrict__ val,
                               const int *__restrict__ segp, uint32_t *__restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) key[i] = (uint32_t)segp[val[i]];
}

// 1 where a new segment starts at layout position i (i > 0)
__global__ void k_con_split(int64_t np, const int *__restrict__ lay,
                            const int *__restrict__ segp, const double *__restrict__ xd,
                            const double *__restrict__ dx, double tol2, int *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int f = 0;
  if (i > 0) {
    const int a = lay[i - 1], b = lay[i];
    const double df = xd[b] - xd[a];
    const double lim = (dx[b] < dx[a] ? dx[b] : dx[a]) * tol2;
    f = (segp[a] != segp[b]) || (df * df > lim);
  }
  flag[i] = f;
}
#pragma clang fp contract(fast)

__global__ void k_con_renumber(int64_t np, const int *__restrict__ lay,
                               const int *__restrict__ flag, const int *__restrict__ excl,
                               int *__restrict__ segp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) segp[lay[i]] = excl[i] + flag[i];
}

__global__ void k_con_out(int64_t np, const int *__restrict__ segp, long long *__restrict__ vtx) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) vtx[p] = (long long)segp[p] + 1;
}

// con-check.c:262-288: no repeated id within an element
__global__ void k_con_element_check(int64_t nel, int nv, const int *__restrict__ segp,
                                    int *__restrict__ err) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  for (int a = 0; a < nv; a++)
    for (int b = a + 1; b < nv; b++)
      if (segp[e * nv + a] == segp[e * nv + b]) atomicOr(err, 1);
}

// vertex -> elements holding it (CSR over vertex ids)
__global__ void k_con_vcount(int64_t np, const int *__restrict__ segp, int *__restrict__ cnt) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) atomicAdd(&cnt[segp[p]], 1);
}

__global__ void k_con_vfill(int64_t np, int nv, const int *__restrict__ segp,
                            const int64_t *__restrict__ off, int *__restrict__ fill,
                            int *__restrict__ els) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int v = segp[p];
  els[off[v] + atomicAdd(&fill[v], 1)] = (int)(p / nv);
}

// con-check.c:197-260: no element holds exactly 3 of a hex face's 4 corners
__global__ void k_con_face_check(int64_t nel, const int *__restrict__ segp,
                                 const int64_t *__restrict__ off, const int *__restrict__ els,
                                 int *__restrict__ err) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nel * 6) return;
  const int64_t e = t / 6;
  const int f = (int)(t % 6);
  int v[4];
  for (int c = 0; c < 4; c++) v[c] = segp[e * 8 + CON_FACES3D[f][c] - 1];
  for (int c = 0; c < 4; c++)
    for (int64_t l = off[v[c]]; l < off[v[c] + 1]; l++) {
      const int el = els[l];
      int m = 0;
      for (int c2 = 0; c2 < 4; c2++)
        for (int64_t l2 = off[v[c2]]; l2 < off[v[c2] + 1]; l2++) m += els[l2] == el;
      if (m == 3) atomicOr(err, 1);
    }
}

__global__ void k_con_widen(int64_t n, const int *__restrict__ in, int64_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// vtx_out[nel*nv], coord[nel*nv*ndim]: host or device arrays (on_device).
// Returns 0, 1 if the element or face check fails, 2 on a device error,
// 3 on bad input.
extern "C" int rsbgpu_conn_mesh(int device, long long *vtx_out, const double *coord, int nel,
                                int ndim, double tol, int on_device, char *msg, int msg_len) {
  if (ndim != 2 && ndim != 3) return 3;
  if (nel < 0) return 3;
  const int nv = ndim == 3 ? 8 : 4;
  const int64_t np = (int64_t)nel * nv;
  if (np >= ((int64_t)1 << 31)) return 3;
  if (np == 0) return 0;
  rsbgpu *g = nullptr;
  if (rsbgpu_open(&g, device)) return 2;
  int rc = 0;
  hipStream_t st = g->stream;
  double *dcoord = nullptr, *x = nullptr, *dx = nullptr;
  int *lay = nullptr, *val = nullptr, *segp = nullptr, *flag = nullptr, *excl = nullptr,
      *err = nullptr, *cnt = nullptr, *fill = nullptr;
  int64_t *off = nullptr, *wide = nullptr;
  uint64_t *key = nullptr;
  long long *dvtx = nullptr;
  const size_t cb = sizeof(double) * (size_t)np * ndim;
#define CON_TRY(x)                   \
  do {                               \
    if ((x) != hipSuccess) {         \
      rc = 2;                        \
      goto done;                     \
    }                                \
  } while (0)
  if (on_device) {
    dcoord = (double *)coord;
  } else {
    CON_TRY(hipMalloc((void **)&dcoord, cb));
    CON_TRY(hipMemcpyAsync(dcoord, coord, cb, hipMemcpyHostToDevice, st));
  }
  CON_TRY(hipMalloc((void **)&x, sizeof(double) * 3 * (size_t)np));
  CON_TRY(hipMalloc((void **)&dx, sizeof(double) * (size_t)np));
  CON_TRY(hipMalloc((void **)&lay, sizeof(int) * (size_t)np));
  CON_TRY(hipMalloc((void **)&val, sizeof(int) * (size_t)np));
  CON_TRY(hipMalloc((void **)&segp, sizeof(int) * (size_t)np));
  CON_TRY(hipMalloc((void **)&flag, sizeof(int) * (size_t)np));
  CON_TRY(hipMalloc((void **)&excl, sizeof(int) * (size_t)np));
  CON_TRY(hipMalloc((void **)&key, sizeof(uint64_t) * (size_t)np));
  CON_TRY(hipMalloc((void **)&err, sizeof(int)));
  {
    const unsigned nb = nblocks(np, 256);
    k_con_points<<<nb, 256, 0, st>>>(dcoord, nel, ndim, nv, x, dx);
    CON_TRY(hipGetLastError());
    {
      std::vector<int> iota((size_t)np);
      for (int64_t i = 0; i < np; i++) iota[i] = (int)i;
      CON_TRY(hipMemcpyAsync(lay, iota.data(), sizeof(int) * (size_t)np, hipMemcpyHostToDevice, st));
      CON_TRY(hipMemsetAsync(segp, 0, sizeof(int) * (size_t)np, st));
      CON_TRY(hipStreamSynchronize(st));
    }
    int nseg = 1;
    const double tol2 = tol * tol;
    for (int t = 0; t < ndim; t++)
      for (int d = 0; d < ndim; d++) {
        const double *xd = x + (size_t)d * np;
        k_con_axis_keys<<<nb, 256, 0, st>>>(np, lay, xd, key, val);
        CON_TRY(hipGetLastError());
        if (rsb_sort_pairs_u64(g, key, val, np, 0, 64)) { rc = 2; goto done; }
        k_con_seg_keys<<<nb, 256, 0, st>>>(np, val, segp, (uint32_t *)key);
        CON_TRY(hipGetLastError());
        int sb = 1;
        while ((1ll << sb) < nseg) sb++;
        if (rsb_sort_pairs_u32(g, (uint32_t *)key, val, np, 0, sb)) { rc = 2; goto done; }
        CON_TRY(hipMemcpyAsync(lay, val, sizeof(int) * (size_t)np, hipMemcpyDeviceToDevice, st));
        k_con_split<<<nb, 256, 0, st>>>(np, lay, segp, xd, dx, tol2, flag);
        CON_TRY(hipGetLastError());
        if (rsb_exclusive_scan_i32(g, flag, excl, np)) { rc = 2; goto done; }
        k_con_renumber<<<nb, 256, 0, st>>>(np, lay, flag, excl, segp);
        CON_TRY(hipGetLastError());
        int last[2];
        CON_TRY(hipMemcpyAsync(&last[0], excl + np - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        CON_TRY(hipMemcpyAsync(&last[1], flag + np - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        CON_TRY(hipStreamSynchronize(st));
        nseg = last[0] + last[1] + 1;
      }
    // output
    if (on_device) {
      dvtx = vtx_out;
    } else {
      CON_TRY(hipMalloc((void **)&dvtx, sizeof(long long) * (size_t)np));
    }
    k_con_out<<<nb, 256, 0, st>>>(np, segp, dvtx);
    CON_TRY(hipGetLastError());
    if (!on_device) {
      CON_TRY(hipMemcpyAsync(vtx_out, dvtx, sizeof(long long) * (size_t)np,
                             hipMemcpyDeviceToHost, st));
    }
    // checks
    CON_TRY(hipMemsetAsync(err, 0, sizeof(int), st));
    k_con_element_check<<<nblocks(nel, 256), 256, 0, st>>>(nel, nv, segp, err);
    CON_TRY(hipGetLastError());
    int herr = 0;
    CON_TRY(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    CON_TRY(hipStreamSynchronize(st));
    if (!herr && ndim == 3) {
      int *els = val;  // np entries, one per corner
      CON_TRY(hipMalloc((void **)&cnt, sizeof(int) * ((size_t)nseg + 1)));
      CON_TRY(hipMalloc((void **)&fill, sizeof(int) * ((size_t)nseg + 1)));
      CON_TRY(hipMalloc((void **)&off, sizeof(int64_t) * ((size_t)nseg + 1)));
      CON_TRY(hipMalloc((void **)&wide, sizeof(int64_t) * ((size_t)nseg + 1)));
      CON_TRY(hipMemsetAsync(cnt, 0, sizeof(int) * ((size_t)nseg + 1), st));
      CON_TRY(hipMemsetAsync(fill, 0, sizeof(int) * ((size_t)nseg + 1), st));
      k_con_vcount<<<nb, 256, 0, st>>>(np, segp, cnt);
      k_con_widen<<<nblocks(nseg + 1, 256), 256, 0, st>>>(nseg + 1, cnt, wide);
      CON_TRY(hipGetLastError());
      if (rsb_exclusive_scan_i64(g, wide, off, nseg + 1)) { rc = 2; goto done; }
      k_con_vfill<<<nb, 256, 0, st>>>(np, nv, segp, off, fill, els);
      k_con_face_check<<<nblocks((int64_t)nel * 6, 256), 256, 0, st>>>(nel, segp, off, els, err);
      CON_TRY(hipGetLastError());
      CON_TRY(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
      CON_TRY(hipStreamSynchronize(st));
    }
    if (herr) rc = 1;
  }
done:
#undef CON_TRY
  if (!on_device) {
    if (dcoord) hipFree(dcoord);
    if (dvtx) hipFree(dvtx);
  }
  hipFree(x), hipFree(dx), hipFree(lay), hipFree(val), hipFree(segp), hipFree(flag),
      hipFree(excl), hipFree(key), hipFree(err), hipFree(cnt), hipFree(fill), hipFree(off),
      hipFree(wide);
  if (rc == 2 && msg && msg_len > 0) snprintf(msg, (size_t)msg_len, "device failure in connectivity");
  rsbgpu_close(g);
  return rc;
}
