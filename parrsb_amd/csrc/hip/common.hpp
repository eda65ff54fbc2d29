// Shared device helpers for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RSB_WAVE 64

// Order-preserving map of an IEEE double to uint64 (total order equal to `<`
// on non-NaN values, with -0.0 mapped onto +0.0 so they compare equal, as
// C's `<` has them).
__host__ __device__ inline uint64_t dkey(double x) {
  if (x == 0.0) x = 0.0;
  uint64_t u;
  __builtin_memcpy(&u, &x, 8);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double dkey_inv(uint64_t k) {
  uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  double x;
  __builtin_memcpy(&x, &u, 8);
  return x;
}
__host__ __device__ inline uint32_t fkey(float x) {
  if (x == 0.0f) x = 0.0f;
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = RSB_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, RSB_WAVE);
  return v;
}

// Block-wide sum of NV doubles with a fixed reduction tree (deterministic).
// `red` must hold (blockDim.x / 64) * NV doubles.  Result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double *red) {
  const int lane = threadIdx.x & (RSB_WAVE - 1), wid = threadIdx.x / RSB_WAVE;
  const int nw = blockDim.x / RSB_WAVE;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) red[wid * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = red[k];
      for (int w = 1; w < nw; w++) s += red[w * NV + k];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Error-free transformation a + b = s + e (Knuth two-sum), for compensated
// FP32 accumulation.
__device__ inline void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Compensated accumulator: value = hi + lo.
struct fsum {
  float hi, lo;
  __device__ fsum() : hi(0.f), lo(0.f) {}
  __device__ inline void add(float x) {
    float s, e;
    two_sum(hi, x, s, e);
    hi = s;
    lo += e;
  }
  __device__ inline double value() const { return (double)hi + (double)lo; }
};

// Agent-scope load that bypasses the non-coherent per-CU cache; used when a
// block reads values that other blocks published in the same launch.
template <typename T>
__device__ inline T load_acquire(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// "Last block of a group" handshake: every block of a group of `count`
// blocks calls this once after publishing its partial results; it returns
// true in exactly one block (the last to arrive), after which that block may
// read every partial of the group.  The counter is re-armed for the next
// use.  Must be called by all threads of the block.
__device__ inline bool last_block_of(unsigned *counter, unsigned count) {
  __shared__ bool is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    is_last = (prev == count - 1);
    if (is_last) {
      __threadfence();
      __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last;
}
