// pf_mem.hip -- the library's device block cache and the per-thread construction streams.
//
// Operators, solvers and their construction temporaries take their device memory from here.  A freed
// block goes back to a per-device cache (sizes rounded to 512 B below 1 MiB, to 2 MiB above) and the next
// request of at most its size, and at least four fifths of it, takes it again; nothing here waits for the
// device, unlike hipFree, which synchronises the whole device and so would hold a request that releases its
// operator behind every other request's queued work.  Above PF_CACHE_MAX cached bytes per device the largest
// cached blocks are released, and a failed hipMalloc releases the device's cache and tries once more.
//
// The construction stream is a non-blocking stream per host thread and device: two threads building
// operators at once run their kernels and copies side by side, and neither waits for work another thread
// queued on its own streams.
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {
namespace {

constexpr int MAXDEV = 64;

struct Block {
  int dev;
  size_t size;
  bool cached;
};

struct Cache {
  std::mutex mu;
  std::multimap<size_t, void*> free_blocks[MAXDEV];
  size_t cached[MAXDEV] = {};
  size_t live = 0;
  std::unordered_map<void*, Block> blocks;  // every block the cache owns, in use or cached
};

// never destroyed: handles released during static destruction still find it
Cache& cache() {
  static Cache* c = new Cache();
  return *c;
}

size_t round_size(size_t b) {
  if (b < 512) b = 512;
  if (b <= (size_t(1) << 20)) return (b + 511) & ~size_t(511);
  const size_t g = size_t(2) << 20;
  return (b + g - 1) & ~(g - 1);
}

int current_device() {
  int dev = 0;
  PF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAXDEV) throw Error(PF_ERR_UNSUPPORTED, "device index out of range");
  return dev;
}

// Takes cached blocks of device dev out of the cache (largest first) until at most keep bytes stay; the
// caller releases them with hipFree outside the lock.
std::vector<void*> take_cached(Cache& c, int dev, size_t keep) {
  std::vector<void*> out;
  auto& fl = c.free_blocks[dev];
  while (c.cached[dev] > keep && !fl.empty()) {
    auto it = std::prev(fl.end());
    c.cached[dev] -= it->first;
    c.blocks.erase(it->second);
    out.push_back(it->second);
    fl.erase(it);
  }
  return out;
}

void release(const std::vector<void*>& v) {
  for (void* p : v) (void)hipFree(p);
}

struct ThreadStreams {
  hipStream_t s[MAXDEV] = {};
  ~ThreadStreams() {
    for (hipStream_t x : s)
      if (x) (void)hipStreamDestroy(x);
  }
};
thread_local ThreadStreams t_streams;

}  // namespace

void* dev_alloc(size_t bytes) {
  const int dev = current_device();
  const size_t r = round_size(bytes);
  Cache& c = cache();
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto& fl = c.free_blocks[dev];
    auto it = fl.lower_bound(r);
    if (it != fl.end() && it->first <= r + r / 4) {
      void* p = it->second;
      c.cached[dev] -= it->first;
      c.live += it->first;
      c.blocks[p].cached = false;
      fl.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, r);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    std::vector<void*> v;
    {
      std::lock_guard<std::mutex> g(c.mu);
      v = take_cached(c, dev, 0);
    }
    release(v);
    e = hipMalloc(&p, r);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(PF_ERR_NOMEM, "hipMalloc of " + std::to_string(r) + " bytes: " + hipGetErrorString(e));
    }
  }
  std::lock_guard<std::mutex> g(c.mu);
  c.blocks[p] = Block{dev, r, false};
  c.live += r;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  Cache& c = cache();
  std::vector<void*> v;
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.blocks.find(p);
    if (it == c.blocks.end() || it->second.cached) return;  // not a cache block, or freed twice
    Block& b = it->second;
    b.cached = true;
    c.live -= b.size;
    c.cached[b.dev] += b.size;
    c.free_blocks[b.dev].emplace(b.size, p);
    v = take_cached(c, b.dev, PF_CACHE_MAX);
  }
  release(v);
}

size_t dev_release_cached() {
  Cache& c = cache();
  std::vector<void*> v;
  size_t n = 0;
  {
    std::lock_guard<std::mutex> g(c.mu);
    for (int d = 0; d < MAXDEV; ++d) {
      n += c.cached[d];
      auto part = take_cached(c, d, 0);
      v.insert(v.end(), part.begin(), part.end());
    }
  }
  release(v);
  return n;
}

void dev_cache_stats(int64_t& live_bytes, int64_t& cached_bytes) {
  Cache& c = cache();
  std::lock_guard<std::mutex> g(c.mu);
  live_bytes = (int64_t)c.live;
  size_t n = 0;
  for (int d = 0; d < MAXDEV; ++d) n += c.cached[d];
  cached_bytes = (int64_t)n;
}

Recorder*& recorder() {
  static thread_local Recorder* r = nullptr;
  return r;
}

hipStream_t cstream() {
  const int dev = current_device();
  hipStream_t& s = t_streams.s[dev];
  if (!s) PF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

void cstream_sync() noexcept {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV) return;
  if (t_streams.s[dev]) (void)hipStreamSynchronize(t_streams.s[dev]);
}

void h2d(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  hipStream_t s = cstream();
  PF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  PF_HIP(hipStreamSynchronize(s));
}

void d2h(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  hipStream_t s = cstream();
  PF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  PF_HIP(hipStreamSynchronize(s));
}

}  // namespace pf

extern "C" {

int pf_cache_stats(int64_t* live_bytes, int64_t* cached_bytes) {
  if (!live_bytes || !cached_bytes) return pf::fail(PF_ERR_ARG, "null argument");
  pf::dev_cache_stats(*live_bytes, *cached_bytes);
  return PF_OK;
}

int pf_cache_release(int64_t* released_bytes) {
  return pf::guarded([&] {
    const size_t n = pf::dev_release_cached();
    if (released_bytes) *released_bytes = (int64_t)n;
  });
}

}  // extern "C"
