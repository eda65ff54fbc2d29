// Device-wide radix sorts and scans (rocPRIM building blocks), in place on
// the caller's arrays.  LSD radix sorting is stable, which every sort of the
// partitioner relies on: equal keys keep the current order.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "state.hpp"

static int ensure_tmp(rsbgpu *g, size_t bytes) {
  if (bytes <= g->sort_tmp_bytes && g->d_sort_tmp) return 0;
  if (g->d_sort_tmp) RSB_TRY(rsb_dev_free(g->d_sort_tmp));
  g->d_sort_tmp = nullptr;
  RSB_TRY(rsb_dev_malloc(&g->d_sort_tmp, bytes + 256));
  g->sort_tmp_bytes = bytes + 256;
  return 0;
}

static int ensure_kv(rsbgpu *g, size_t n) {
  if (n <= g->sort_cap && g->d_k1) return 0;
  if (g->d_k1) RSB_TRY(rsb_dev_free(g->d_k1));
  if (g->d_v1) RSB_TRY(rsb_dev_free(g->d_v1));
  g->d_k1 = nullptr, g->d_v1 = nullptr;
  RSB_TRY(rsb_dev_malloc((void **)&g->d_k1, n * sizeof(uint64_t)));
  RSB_TRY(rsb_dev_malloc((void **)&g->d_v1, n * sizeof(int)));
  g->sort_cap = n;
  return 0;
}

int rsb_sort_pairs_u64(rsbgpu *g, uint64_t *keys, int *vals, int64_t n, int bb, int eb) {
  if (n <= 1) return 0;
  int rc = ensure_kv(g, (size_t)n);
  if (rc) return rc;
  size_t bytes = 0;
  RSB_TRY(rocprim::radix_sort_pairs(nullptr, bytes, keys, g->d_k1, vals, g->d_v1, (size_t)n,
                                    bb, eb, g->stream));
  if ((rc = ensure_tmp(g, bytes))) return rc;
  bytes = g->sort_tmp_bytes;
  RSB_TRY(rocprim::radix_sort_pairs(g->d_sort_tmp, bytes, keys, g->d_k1, vals, g->d_v1,
                                    (size_t)n, bb, eb, g->stream));
  RSB_TRY(hipMemcpyAsync(keys, g->d_k1, n * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                         g->stream));
  RSB_TRY(hipMemcpyAsync(vals, g->d_v1, n * sizeof(int), hipMemcpyDeviceToDevice, g->stream));
  return 0;
}

int rsb_sort_pairs_u32(rsbgpu *g, uint32_t *keys, int *vals, int64_t n, int bb, int eb) {
  if (n <= 1) return 0;
  int rc = ensure_kv(g, (size_t)n);
  if (rc) return rc;
  uint32_t *k1 = (uint32_t *)g->d_k1;
  size_t bytes = 0;
  RSB_TRY(rocprim::radix_sort_pairs(nullptr, bytes, keys, k1, vals, g->d_v1, (size_t)n, bb,
                                    eb, g->stream));
  if ((rc = ensure_tmp(g, bytes))) return rc;
  bytes = g->sort_tmp_bytes;
  RSB_TRY(rocprim::radix_sort_pairs(g->d_sort_tmp, bytes, keys, k1, vals, g->d_v1, (size_t)n,
                                    bb, eb, g->stream));
  RSB_TRY(hipMemcpyAsync(keys, k1, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, g->stream));
  RSB_TRY(hipMemcpyAsync(vals, g->d_v1, n * sizeof(int), hipMemcpyDeviceToDevice, g->stream));
  return 0;
}

int rsb_exclusive_scan_i64(rsbgpu *g, const int64_t *in, int64_t *out, int64_t n) {
  if (n <= 0) return 0;
  size_t bytes = 0;
  RSB_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, (size_t)n,
                                  rocprim::plus<int64_t>(), g->stream));
  int rc = ensure_tmp(g, bytes);
  if (rc) return rc;
  bytes = g->sort_tmp_bytes;
  RSB_TRY(rocprim::exclusive_scan(g->d_sort_tmp, bytes, in, out, (int64_t)0, (size_t)n,
                                  rocprim::plus<int64_t>(), g->stream));
  return 0;
}

int rsb_exclusive_scan_i32(rsbgpu *g, const int *in, int *out, int64_t n) {
  if (n <= 0) return 0;
  size_t bytes = 0;
  RSB_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(),
                                  g->stream));
  int rc = ensure_tmp(g, bytes);
  if (rc) return rc;
  bytes = g->sort_tmp_bytes;
  RSB_TRY(rocprim::exclusive_scan(g->d_sort_tmp, bytes, in, out, 0, (size_t)n,
                                  rocprim::plus<int>(), g->stream));
  return 0;
}
