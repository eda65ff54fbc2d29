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
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <vector>

#include "state.hpp"

__constant__ int CON_PRE_TO_SYM[8] = {0, 1, 3, 2, 4, 5, 7, 6};
__constant__ int CON_NEIGHBOR[8][3] = {{1, 2, 4}, {0, 3, 5}, {0, 3, 6}, {1, 2, 7},
                                       {0, 5, 6}, {1, 4, 7}, {2, 4, 7}, {3, 5, 6}};
__constant__ int CON_FACES3D[6][4] = {{1, 5, 7, 3}, {2, 4, 8, 6}, {1, 2, 6, 5},
                                      {3, 7, 8, 4}, {1, 3, 4, 2}, {5, 6, 8, 7}};
__constant__ int CON_FACES2D[4][2] = {{3, 1}, {2, 4}, {1, 2}, {4, 3}};
// preprocessor face number - 1 -> symmetric face (con.c:4)
static const int CON_FACE_PRE_TO_SYM[6] = {2, 1, 3, 0, 4, 5};

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

// Periodic faces (con-periodic.c): corners of each record's face, in the
// face's corner order, with their current ids.  sface[i] is the record's
// symmetric face number, checked on the host.
__global__ void k_con_pface(int nrec, const int *__restrict__ elem, const int *__restrict__ sface,
                            int nv, int ndim, int64_t np, const double *__restrict__ x,
                            const int *__restrict__ segp, double *__restrict__ fx,
                            int *__restrict__ fid) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nvf = nv / 2;
  if (t >= nrec * nvf) return;
  const int i = t / nvf, j = t % nvf;
  const int s = (ndim == 3 ? CON_FACES3D[sface[i]][j] : CON_FACES2D[sface[i]][j]) - 1;
  const int64_t p = (int64_t)elem[i] * nv + s;
  for (int d = 0; d < 3; d++) fx[((int64_t)i * 4 + j) * 3 + d] = x[(size_t)d * np + p];
  fid[(int64_t)i * 4 + j] = segp[p];
}

// Renumbering after periodic matching: an id in the sorted list `gone` takes
// its class's smallest id root[k]; every id then moves down by the number of
// gone ids below it (the dense renumbering in ascending order).
__global__ void k_con_relabel(int64_t np, int ngone, const int *__restrict__ gone,
                              const int *__restrict__ root, int *__restrict__ segp) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int v = segp[p];
  int lo = 0, hi = ngone;  // first index with gone[idx] >= v
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (gone[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  if (lo < ngone && gone[lo] == v) {
    v = root[lo];
    int a = 0, b = lo;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (gone[mid] < v) a = mid + 1;
      else b = mid;
    }
    lo = a;
  }
  segp[p] = v - lo;
}

// Host part of the periodic matching, on the corners gathered by
// k_con_pface (con-periodic.c:115-188): returns 1 if a face pair does not
// match, else fills the sorted list of ids that merge into a smaller one and
// their class minima.
static int periodic_classes(int nrec, const int *elem, const int *sface, const int *pelem,
                            const int *psface, int ndim, const std::vector<double> &fx,
                            const std::vector<int> &fid, std::vector<int> &gone,
                            std::vector<int> &root, char *msg, int msg_len) {
  const int nvf = ndim == 3 ? 4 : 2;
  // records by (element, face); a record j pairs with every earlier record i
  // whose (element, face) is j's partner
  std::vector<std::pair<int64_t, int>> byface((size_t)nrec);
  for (int i = 0; i < nrec; i++) byface[i] = {(int64_t)elem[i] * 8 + sface[i], i};
  std::sort(byface.begin(), byface.end());
  std::vector<std::pair<int, int>> pairs;
  for (int j = 0; j < nrec; j++) {
    const int64_t want = (int64_t)pelem[j] * 8 + psface[j];
    auto it = std::lower_bound(byface.begin(), byface.end(), std::make_pair(want, -1));
    for (; it != byface.end() && it->first == want && it->second < j; ++it) {
      const int i = it->second;
      double F[4][3], G[4][3], fmax = 0.0, gmax = 0.0;
      for (int d = 0; d < ndim; d++) {
        double mf = 0.0, mg = 0.0;
        for (int c = 0; c < nvf; c++) {
          F[c][d] = fx[((size_t)i * 4 + c) * 3 + d], G[c][d] = fx[((size_t)j * 4 + c) * 3 + d];
          fmax = std::max(fmax, std::fabs(F[c][d])), gmax = std::max(gmax, std::fabs(G[c][d]));
          mf += F[c][d], mg += G[c][d];
        }
        for (int c = 0; c < nvf; c++) F[c][d] -= mf / nvf, G[c][d] -= mg / nvf;
      }
      int shift = 0;
      double best = 1e20;
      for (int sh = 0; sh < nvf; sh++) {
        double d2 = 0.0;
        for (int c = 0; c < nvf; c++) {
          const int k = nvf - 1 - (c + sh) % nvf;
          for (int d = 0; d < ndim; d++) d2 += (F[c][d] - G[k][d]) * (F[c][d] - G[k][d]);
        }
        if (d2 < best) best = d2, shift = sh;
      }
      if (std::sqrt(best) > 1e-3 * std::max(fmax, gmax)) {
        if (msg && msg_len > 0)
          snprintf(msg, (size_t)msg_len, "periodic faces did not match: records %d and %d", i + 1,
                   j + 1);
        return 1;
      }
      for (int c = 0; c < nvf; c++) {
        const int k = nvf - 1 - (c + shift) % nvf;
        pairs.push_back({fid[(size_t)i * 4 + c], fid[(size_t)j * 4 + k]});
      }
    }
  }
  // classes of the ids that appear in pairs; the root of a class is its
  // smallest id
  std::vector<int> ids;
  ids.reserve(2 * pairs.size());
  for (auto &q : pairs) ids.push_back(q.first), ids.push_back(q.second);
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  std::vector<int> par(ids.size());
  for (size_t k = 0; k < par.size(); k++) par[k] = (int)k;
  auto find = [&](int a) {
    while (par[a] != a) a = par[a] = par[par[a]];
    return a;
  };
  auto index = [&](int v) { return (int)(std::lower_bound(ids.begin(), ids.end(), v) - ids.begin()); };
  for (auto &q : pairs) {
    const int a = find(index(q.first)), b = find(index(q.second));
    if (a != b) par[std::max(a, b)] = std::min(a, b);  // ids sorted: smaller index = smaller id
  }
  gone.clear(), root.clear();
  for (size_t k = 0; k < ids.size(); k++) {
    const int r = find((int)k);
    if (r != (int)k) gone.push_back(ids[k]), root.push_back(ids[r]);
  }
  return 0;
}

__global__ void k_con_widen(int64_t n, const int *__restrict__ in, int64_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// vtx_out[nel*nv], coord[nel*nv*ndim]: host or device arrays (on_device).
// Returns 0, 1 if the element or face check fails, 2 on a device error,
// 3 on bad input.
extern "C" int rsbgpu_conn_mesh(int device, long long *vtx_out, const double *coord, int nel,
                                int ndim, double tol, const long long *pinfo, int npinfo,
                                int on_device, char *msg, int msg_len) {
  if (ndim != 2 && ndim != 3) return 3;
  if (nel < 0 || npinfo < 0 || (npinfo > 0 && !pinfo)) return 3;
  const int nv = ndim == 3 ? 8 : 4;
  const int64_t np = (int64_t)nel * nv;
  if (np >= ((int64_t)1 << 31)) return 3;
  // periodic records: 1-based (element, face, element, face), faces in the
  // preprocessor numbering (con.c:41-47)
  std::vector<int> pel((size_t)npinfo), psf((size_t)npinfo), ppel((size_t)npinfo),
      ppsf((size_t)npinfo);
  for (int i = 0; i < npinfo; i++) {
    const long long *r = pinfo + 4 * (int64_t)i;
    if (r[0] < 1 || r[0] > nel || r[2] < 1 || r[2] > nel || r[1] < 1 || r[1] > 2 * ndim ||
        r[3] < 1 || r[3] > 2 * ndim) {
      if (msg && msg_len > 0) snprintf(msg, (size_t)msg_len, "periodic record %d out of range", i + 1);
      return 3;
    }
    pel[i] = (int)(r[0] - 1), psf[i] = CON_FACE_PRE_TO_SYM[r[1] - 1];
    ppel[i] = (int)(r[2] - 1), ppsf[i] = CON_FACE_PRE_TO_SYM[r[3] - 1];
  }
  if (np == 0) return 0;
  rsbgpu *g = nullptr;
  if (rsbgpu_open(&g, device)) return 2;
  int rc = 0;
  hipStream_t st = g->stream;
  double *dcoord = nullptr, *x = nullptr, *dx = nullptr;
  int *lay = nullptr, *val = nullptr, *segp = nullptr, *flag = nullptr, *excl = nullptr,
      *err = nullptr, *cnt = nullptr, *fill = nullptr;
  int64_t *off = nullptr, *wide = nullptr;
  int *pdev = nullptr, *pfid = nullptr, *pgone = nullptr;
  double *pfx = nullptr;
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
    CON_TRY(rsb_dev_malloc((void **)&dcoord, cb));
    CON_TRY(hipMemcpyAsync(dcoord, coord, cb, hipMemcpyHostToDevice, st));
  }
  CON_TRY(rsb_dev_malloc((void **)&x, sizeof(double) * 3 * (size_t)np));
  CON_TRY(rsb_dev_malloc((void **)&dx, sizeof(double) * (size_t)np));
  CON_TRY(rsb_dev_malloc((void **)&lay, sizeof(int) * (size_t)np));
  CON_TRY(rsb_dev_malloc((void **)&val, sizeof(int) * (size_t)np));
  CON_TRY(rsb_dev_malloc((void **)&segp, sizeof(int) * (size_t)np));
  CON_TRY(rsb_dev_malloc((void **)&flag, sizeof(int) * (size_t)np));
  CON_TRY(rsb_dev_malloc((void **)&excl, sizeof(int) * (size_t)np));
  CON_TRY(rsb_dev_malloc((void **)&key, sizeof(uint64_t) * (size_t)np));
  CON_TRY(rsb_dev_malloc((void **)&err, sizeof(int)));
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
      CON_TRY(rsb_dev_malloc((void **)&dvtx, sizeof(long long) * (size_t)np));
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
      CON_TRY(rsb_dev_malloc((void **)&cnt, sizeof(int) * ((size_t)nseg + 1)));
      CON_TRY(rsb_dev_malloc((void **)&fill, sizeof(int) * ((size_t)nseg + 1)));
      CON_TRY(rsb_dev_malloc((void **)&off, sizeof(int64_t) * ((size_t)nseg + 1)));
      CON_TRY(rsb_dev_malloc((void **)&wide, sizeof(int64_t) * ((size_t)nseg + 1)));
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
    if (!herr && npinfo > 0) {
      // periodic faces (con-periodic.c:258-288): gather the faces' corners,
      // match them on the host, relabel every point on the device
      const int nvf = nv / 2;
      CON_TRY(rsb_dev_malloc((void **)&pdev, sizeof(int) * 2 * (size_t)npinfo));
      CON_TRY(rsb_dev_malloc((void **)&pfx, sizeof(double) * 12 * (size_t)npinfo));
      CON_TRY(rsb_dev_malloc((void **)&pfid, sizeof(int) * 4 * (size_t)npinfo));
      CON_TRY(hipMemcpyAsync(pdev, pel.data(), sizeof(int) * (size_t)npinfo, hipMemcpyHostToDevice, st));
      CON_TRY(hipMemcpyAsync(pdev + npinfo, psf.data(), sizeof(int) * (size_t)npinfo,
                             hipMemcpyHostToDevice, st));
      k_con_pface<<<nblocks((int64_t)npinfo * nvf, 256), 256, 0, st>>>(
          npinfo, pdev, pdev + npinfo, nv, ndim, np, x, segp, pfx, pfid);
      CON_TRY(hipGetLastError());
      std::vector<double> hfx((size_t)npinfo * 12);
      std::vector<int> hfid((size_t)npinfo * 4), gone, root;
      CON_TRY(hipMemcpyAsync(hfx.data(), pfx, sizeof(double) * hfx.size(), hipMemcpyDeviceToHost, st));
      CON_TRY(hipMemcpyAsync(hfid.data(), pfid, sizeof(int) * hfid.size(), hipMemcpyDeviceToHost, st));
      CON_TRY(hipStreamSynchronize(st));
      if (periodic_classes(npinfo, pel.data(), psf.data(), ppel.data(), ppsf.data(), ndim, hfx,
                           hfid, gone, root, msg, msg_len)) {
        rc = 1;
      } else if (!gone.empty()) {
        const int ng = (int)gone.size();
        CON_TRY(rsb_dev_malloc((void **)&pgone, sizeof(int) * 2 * (size_t)ng));
        CON_TRY(hipMemcpyAsync(pgone, gone.data(), sizeof(int) * ng, hipMemcpyHostToDevice, st));
        CON_TRY(hipMemcpyAsync(pgone + ng, root.data(), sizeof(int) * ng, hipMemcpyHostToDevice, st));
        k_con_relabel<<<nb, 256, 0, st>>>(np, ng, pgone, pgone + ng, segp);
        k_con_out<<<nb, 256, 0, st>>>(np, segp, dvtx);
        CON_TRY(hipGetLastError());
        if (!on_device) {
          CON_TRY(hipMemcpyAsync(vtx_out, dvtx, sizeof(long long) * (size_t)np,
                                 hipMemcpyDeviceToHost, st));
        }
        CON_TRY(hipStreamSynchronize(st));
      }
    }
  }
done:
#undef CON_TRY
  if (!on_device) {
    if (dcoord) rsb_dev_free(dcoord);
    if (dvtx) rsb_dev_free(dvtx);
  }
  rsb_dev_free(x), rsb_dev_free(dx), rsb_dev_free(lay), rsb_dev_free(val), rsb_dev_free(segp), rsb_dev_free(flag),
      rsb_dev_free(excl), rsb_dev_free(key), rsb_dev_free(err), rsb_dev_free(cnt), rsb_dev_free(fill), rsb_dev_free(off),
      rsb_dev_free(wide), rsb_dev_free(pdev), rsb_dev_free(pfx), rsb_dev_free(pfid), rsb_dev_free(pgone);
  if (rc == 2 && msg && msg_len > 0) snprintf(msg, (size_t)msg_len, "device failure in connectivity");
  rsbgpu_close(g);
  return rc;
}
