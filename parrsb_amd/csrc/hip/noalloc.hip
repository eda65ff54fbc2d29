// Allocation-free partitioning: every device buffer of one call is carved
// from a fixed device array (`scratch`, one copy per GPU, held until the
// process ends) by a bump allocator, instead of being hipMalloc'd.  For
// callers that cannot allocate device memory: code around HIP graph
// capture, tight serving loops, harnesses that forbid it.  No counterpart in
// the reference.
//
// rsb_dev_malloc / rsb_dev_free are the library's only device allocator:
// outside such a call they are hipMalloc / hipFree.  Inside one, blocks come
// first-fit (lowest address) from a free list over the array.  A freed block
// is not reused at once: work queued on the call's stream may still read it,
// and some copies of the library are synchronous on the null stream, which a
// non-blocking stream does not order against.  Freed blocks wait in a
// pending list, and only when the free list cannot serve a request does the
// allocator wait for the stream to drain and return them to the free list
// (merged with their neighbours).  The entry point bounds the element count
// so that a call stays inside the array.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

// the array's size and the element limit are part of the public header; this
// file uses no MPI type from it
#ifndef PARRSB_NO_MPI
#define PARRSB_NO_MPI
#endif
#include "parrsb_amd.h"
#include "state.hpp"

static __device__ __align__(256) unsigned char scratch[PARRSB_AMD_NOALLOC_BYTES];

namespace {

constexpr int kMaxDevices = 64;

// Set up once per device, with that device current: the device address of
// `scratch` (host code never uses the array's name for device memory), and
// the stream and four timing events every allocation-free call borrows.
struct dev_setup {
  bool ready = false;
  unsigned char *ptr = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};
dev_setup setups[kMaxDevices];
std::mutex setup_mu;

// Blocks of the array by offset; every size a multiple of 256 bytes.
struct arena {
  int device = -1;
  unsigned char *base = nullptr;
  hipStream_t stream = nullptr;
  size_t cap = 0, high = 0, live = 0, peak = 0;
  std::map<size_t, size_t> free_;             // offset -> size, merged neighbours
  std::unordered_map<size_t, size_t> used;    // offset -> size
  std::vector<std::pair<size_t, size_t>> pending;  // freed, stream not yet drained

  void reset(int dev, unsigned char *b, hipStream_t st, size_t c) {
    device = dev, base = b, stream = st, cap = c, high = 0, live = 0, peak = 0;
    free_.clear(), used.clear(), pending.clear();
    free_[0] = c;
  }
  void give_back(size_t off, size_t size) {
    auto it = free_.emplace(off, size).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  bool take(size_t need, size_t *off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      *off = it->first;
      const size_t rest = it->second - need;
      free_.erase(it);
      if (rest) free_[*off + need] = rest;
      used[*off] = need;
      high = std::max(high, *off + need);
      live += need;
      peak = std::max(peak, live);
      return true;
    }
    return false;
  }
};
thread_local arena cur_arena;
thread_local bool cur_active = false;

}  // namespace

hipError_t rsb_dev_malloc(void **p, size_t bytes) {
  if (!cur_active) return hipMalloc(p, bytes);
  arena &a = cur_arena;
  const size_t need = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
  size_t off = 0;
  bool ok = a.take(need, &off);
  if (!ok && !a.pending.empty()) {
    // every block freed so far becomes reusable once the stream has drained
    hipError_t e = hipStreamSynchronize(a.stream);
    if (e != hipSuccess) return e;
    for (auto &b : a.pending) a.give_back(b.first, b.second);
    a.pending.clear();
    ok = a.take(need, &off);
  }
  if (!ok) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  *p = a.base + off;
  return hipSuccess;
}

hipError_t rsb_dev_free(const void *p) {
  if (!p) return hipSuccess;
  if (cur_active) {
    const unsigned char *q = (const unsigned char *)p;
    if (q >= cur_arena.base && q < cur_arena.base + cur_arena.cap) {
      auto it = cur_arena.used.find((size_t)(q - cur_arena.base));
      if (it == cur_arena.used.end()) return hipErrorInvalidValue;
      cur_arena.pending.emplace_back(it->first, it->second);
      cur_arena.live -= it->second;
      cur_arena.used.erase(it);
      return hipSuccess;
    }
  }
  return hipFree((void *)p);
}

bool rsb_arena_stream(int device, hipStream_t *stream, hipEvent_t *ev) {
  if (!cur_active || cur_arena.device != device) return false;
  const dev_setup &d = setups[device];
  *stream = d.stream;
  for (int i = 0; i < 4; i++) ev[i] = d.ev[i];
  return true;
}

extern "C" int rsbgpu_noalloc_begin(int device, char *msg, int msglen) {
  auto fail = [&](int rc, const char *what, hipError_t e) {
    if (msg && msglen > 0)
      snprintf(msg, (size_t)msglen, "%s%s%s", what, e != hipSuccess ? ": " : "",
               e != hipSuccess ? hipGetErrorString(e) : "");
    return rc;
  };
  if (cur_active) return fail(3, "an allocation-free call is already running on this thread",
                              hipSuccess);
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(2, "no usable GPU", e);
  if (device < 0 || device >= ndev || device >= kMaxDevices)
    return fail(2, "no such device", hipSuccess);
  std::lock_guard<std::mutex> lock(setup_mu);
  dev_setup &d = setups[device];
  if (!d.ready) {
    if ((e = hipSetDevice(device)) != hipSuccess) return fail(2, "hipSetDevice", e);
    void *ptr = nullptr;
    if ((e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch))) != hipSuccess || !ptr)
      return fail(2, "hipGetSymbolAddress(scratch)", e);
    if ((e = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking)) != hipSuccess)
      return fail(2, "hipStreamCreateWithFlags", e);
    for (int i = 0; i < 4; i++)
      if ((e = hipEventCreate(&d.ev[i])) != hipSuccess) return fail(2, "hipEventCreate", e);
    d.ptr = (unsigned char *)ptr;
    d.ready = true;
  }
  // PARRSB_AMD_NOALLOC_LIMIT (bytes) uses only the start of the array: a
  // smaller region forces the reuse of freed blocks (the tests set it)
  size_t cap = PARRSB_AMD_NOALLOC_BYTES;
  const char *lim = getenv("PARRSB_AMD_NOALLOC_LIMIT");
  if (lim && *lim) {
    const long long v = atoll(lim);
    if (v > 0 && (size_t)v < cap) cap = (size_t)v & ~(size_t)255;
  }
  cur_arena.reset(device, d.ptr, d.stream, cap);
  cur_active = true;
  return 0;
}

extern "C" long long rsbgpu_noalloc_end(long long *high) {
  if (!cur_active) return 0;
  // the call's work is queued on the borrowed stream: drain it before the
  // region may be handed to the next call
  (void)hipStreamSynchronize(setups[cur_arena.device].stream);
  cur_active = false;
  if (high) *high = (long long)cur_arena.high;
  return (long long)cur_arena.peak;
}
