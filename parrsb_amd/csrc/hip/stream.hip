// Streaming kernels that measure what HBM delivers to plain, fully
// coalesced access on this device: read (a sum), write (a fill) and copy,
// 16-byte vector accesses, every workgroup striding over the whole array
// with four accesses in flight per lane.  They give the apply kernel's
// roofline its practical ceiling: the best rate a streaming kernel reaches,
// beside the published 8 TB/s.
#include "state.hpp"

typedef float st_f32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ inline st_f32x4 st_load(const st_f32x4 *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT>
__device__ inline void st_store(st_f32x4 v, st_f32x4 *p) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <bool NT>
__global__ void __launch_bounds__(256) k_stream_read(const st_f32x4 *__restrict__ x, size_t n4,
                                                     float *__restrict__ out) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  st_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const st_f32x4 a = st_load<NT>(x + i), b = st_load<NT>(x + i + stride),
                   c = st_load<NT>(x + i + 2 * stride), d = st_load<NT>(x + i + 3 * stride);
    acc += (a + b) + (c + d);
  }
  for (; i < n4; i += stride) acc += st_load<NT>(x + i);
  const float s = (acc.x + acc.y) + (acc.z + acc.w);
  if (s == -1.0f) out[0] = s;  // the array holds ones: never true, keeps the loads
}

template <bool NT>
__global__ void __launch_bounds__(256) k_stream_write(st_f32x4 *__restrict__ y, size_t n4,
                                                      float v) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const st_f32x4 w = {v, v, v, v};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    st_store<NT>(w, y + i);
}

template <bool NT>
__global__ void __launch_bounds__(256) k_stream_copy(const st_f32x4 *__restrict__ x,
                                                     st_f32x4 *__restrict__ y, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const st_f32x4 a = st_load<NT>(x + i), b = st_load<NT>(x + i + stride),
                   c = st_load<NT>(x + i + 2 * stride), d = st_load<NT>(x + i + 3 * stride);
    st_store<NT>(a, y + i), st_store<NT>(b, y + i + stride);
    st_store<NT>(c, y + i + 2 * stride), st_store<NT>(d, y + i + 3 * stride);
  }
  for (; i < n4; i += stride) st_store<NT>(st_load<NT>(x + i), y + i);
}

// Best rate in GB/s (1e9 bytes/s) of each kernel over grids of 1024 ..
// 16384 workgroups, plain and non-temporal accesses, reps timed launches
// each (after one untimed): gbs[0] read, gbs[1] write, gbs[2] copy (bytes
// read + bytes written).  bytes is rounded down to a multiple of 16.
extern "C" int rsbgpu_stream_test(int device, size_t bytes, int reps, double *gbs) {
  if (bytes < 16 || reps < 1 || !gbs) return 3;
  const size_t n4 = bytes / 16;
  bytes = n4 * 16;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return 2;
  st_f32x4 *x = nullptr, *y = nullptr;
  float *out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = rsb_dev_malloc((void **)&x, bytes);
  if (e == hipSuccess) e = rsb_dev_malloc((void **)&y, bytes);
  if (e == hipSuccess) e = rsb_dev_malloc((void **)&out, sizeof(float));
  if (e == hipSuccess) e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    k_stream_write<false><<<4096, 256, 0, st>>>(x, n4, 1.0f);
    e = hipGetLastError();
  }
  gbs[0] = gbs[1] = gbs[2] = 0.0;
  for (int grid = 1024; e == hipSuccess && grid <= 16384; grid *= 2)
    for (int kind = 0; e == hipSuccess && kind < 3; kind++)
      for (int nt = 0; e == hipSuccess && nt < 2; nt++) {
        auto launch = [&]() {
          if (kind == 0) {
            if (nt) k_stream_read<true><<<grid, 256, 0, st>>>(x, n4, out);
            else k_stream_read<false><<<grid, 256, 0, st>>>(x, n4, out);
          } else if (kind == 1) {
            if (nt) k_stream_write<true><<<grid, 256, 0, st>>>(y, n4, 2.0f);
            else k_stream_write<false><<<grid, 256, 0, st>>>(y, n4, 2.0f);
          } else {
            if (nt) k_stream_copy<true><<<grid, 256, 0, st>>>(x, y, n4);
            else k_stream_copy<false><<<grid, 256, 0, st>>>(x, y, n4);
          }
        };
        launch();
        e = hipEventRecord(e0, st);
        for (int r = 0; e == hipSuccess && r < reps; r++) launch();
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0.f) {
          const double moved = (kind == 2 ? 2.0 : 1.0) * (double)bytes * reps;
          const double rate = moved / (ms * 1e-3) / 1e9;
          if (rate > gbs[kind]) gbs[kind] = rate;
        }
      }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (x) (void)rsb_dev_free(x);
  if (y) (void)rsb_dev_free(y);
  if (out) (void)rsb_dev_free(out);
  return e == hipSuccess ? 0 : 2;
}
