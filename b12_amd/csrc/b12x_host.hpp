// Host-side helpers shared by the C ABI translation units: error state, HIP error checking,
// device allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/b12x_c_api.h"

namespace b12x {

inline std::string& lastError() {
  static thread_local std::string e;
  return e;
}

inline int setErr(int code, const std::string& msg) {
  lastError() = msg;
  return code;
}

struct HipFail {
  int code;
  std::string msg;
  HipFail(int c, std::string m) : code(c), msg(std::move(m)) {}
};

#define HIPCHK(call)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      throw ::b12x::HipFail(e_ == hipErrorOutOfMemory ? B12X_E_NOMEM : B12X_E_HIP,              \
                            std::string(#call) + " failed: " + hipGetErrorString(e_));          \
    }                                                                                           \
  } while (0)

template <class F>
int guarded(F f) {
  try {
    return f();
  } catch (const HipFail& e) {
    return setErr(e.code, e.msg);
  } catch (const std::exception& e) {
    return setErr(B12X_E_ARG, e.what());
  }
}

inline int ensureDevice() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return setErr(B12X_E_NODEVICE, "no HIP device available");
  return B12X_OK;
}

// Device memory cache. Fresh hipMalloc of a large block costs hundreds of microseconds (the
// pages are mapped on allocation), so freed blocks are kept and handed out again. A freed block
// may still be read or written by work queued on any of the library's streams, so it is reused
// only after a fence: one event recorded on every registered stream after the free, all of them
// complete. Fences are recorded when an allocation finds no idle block (one fence covers every
// block freed since the last one) and are never recorded on a stream that is being captured into
// a graph. A stream is registered by the object that creates it and retired (synchronised, then
// no longer fenced) before that object frees its memory and destroys the stream. When hipMalloc
// runs out of memory the cache waits for its fences, returns every idle block and tries again.
// One cache per device, never destroyed (the HIP runtime may be gone at process exit).
class DeviceCache {
 public:
  static DeviceCache& get() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
    static DeviceCache* caches[kMaxDevices] = {};
    static std::mutex m;
    std::lock_guard<std::mutex> lock(m);
    if (!caches[dev]) caches[dev] = new DeviceCache();
    return *caches[dev];
  }

  void* alloc(size_t bytes) {
    const size_t want = roundUp(bytes);
    std::lock_guard<std::mutex> lock(mu_);
    harvest(false);
    void* p = takeReady(want);
    if (!p && !pending_.empty()) {
      fence();
      harvest(false);
      p = takeReady(want);
    }
    if (!p) {
      hipError_t e = hipMalloc(&p, want);
      if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        if (!pending_.empty()) fence();
        harvest(true);
        trim(0);
        p = takeReady(want);
        if (!p) e = hipMalloc(&p, want);
        else e = hipSuccess;
      }
      if (e != hipSuccess)
        throw HipFail(e == hipErrorOutOfMemory ? B12X_E_NOMEM : B12X_E_HIP,
                      std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    live_[p] = want;
    return p;
  }

  void release(void* p) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) {  // not from this cache (or another device): plain free
      (void)hipFree(p);
      return;
    }
    pending_.push_back({p, it->second});
    cached_ += it->second;
    live_.erase(it);
  }

  void addStream(hipStream_t s) {
    std::lock_guard<std::mutex> lock(mu_);
    streams_.push_back(s);
  }

  // synchronises s and stops fencing it (its work is complete, so the blocks it used are idle)
  void retireStream(hipStream_t s) {
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lock(mu_);
    streams_.erase(std::remove(streams_.begin(), streams_.end(), s), streams_.end());
  }

 private:
  static constexpr int kMaxDevices = 16;
  static constexpr size_t kMaxCached = size_t(32) << 30;  // idle bytes kept before trimming

  static size_t roundUp(size_t b) {
    if (b == 0) b = 1;
    const size_t g = b >= (size_t(1) << 20) ? (size_t(1) << 16) : 256;
    return (b + g - 1) / g * g;
  }

  // the smallest idle block that holds `want` bytes without wasting more than half of it
  void* takeReady(size_t want) {
    auto it = ready_.lower_bound(want);
    if (it == ready_.end() || it->first > 2 * want + (size_t(1) << 16)) return nullptr;
    void* p = it->second;
    cached_ -= it->first;
    ready_.erase(it);
    return p;
  }

  void fence() {
    Fence f;
    for (hipStream_t s : streams_) {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    }
    for (hipStream_t s : streams_) {
      hipEvent_t e = nullptr;
      if (!pool_.empty()) {
        e = pool_.back();
        pool_.pop_back();
      } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        for (hipEvent_t x : f.ev) pool_.push_back(x);
        return;
      }
      if (hipEventRecord(e, s) != hipSuccess) {
        (void)hipGetLastError();
        pool_.push_back(e);
        for (hipEvent_t x : f.ev) pool_.push_back(x);
        return;
      }
      f.ev.push_back(e);
    }
    f.blocks.swap(pending_);
    fences_.push_back(std::move(f));
  }

  // fences whose events have all completed release their blocks (wait: block on each fence)
  void harvest(bool wait) {
    while (!fences_.empty()) {
      Fence& f = fences_.front();
      bool done = true;
      for (hipEvent_t e : f.ev) {
        if (wait) {
          if (hipEventSynchronize(e) != hipSuccess) (void)hipGetLastError();
        } else if (hipEventQuery(e) != hipSuccess) {
          (void)hipGetLastError();
          done = false;
          break;
        }
      }
      if (!done) break;
      for (hipEvent_t e : f.ev) pool_.push_back(e);
      for (auto& b : f.blocks) ready_.insert({b.second, b.first});
      fences_.pop_front();
    }
    if (cached_ > kMaxCached) trim(kMaxCached / 2);
  }

  // frees idle blocks, largest first, until at most `keep` cached bytes remain
  void trim(size_t keep) {
    while (cached_ > keep && !ready_.empty()) {
      auto it = std::prev(ready_.end());
      (void)hipFree(it->second);
      cached_ -= it->first;
      ready_.erase(it);
    }
  }

  struct Fence {
    std::vector<hipEvent_t> ev;
    std::vector<std::pair<void*, size_t>> blocks;
  };
  std::mutex mu_;
  std::multimap<size_t, void*> ready_;
  std::vector<std::pair<void*, size_t>> pending_;
  std::deque<Fence> fences_;
  std::vector<hipStream_t> streams_;
  std::unordered_map<void*, size_t> live_;
  std::vector<hipEvent_t> pool_;
  size_t cached_ = 0;  // bytes in pending, fenced and idle blocks
};

template <class T>
T* dalloc(size_t count) {
  return static_cast<T*>(DeviceCache::get().alloc((count ? count : 1) * sizeof(T)));
}

inline void dfree(void* p) {
  if (p) DeviceCache::get().release(p);
}

// Device blocks owned by a scope: returned to the cache when the scope ends, also when an error
// is thrown part-way (work still queued on them is covered by the cache's fences).
class DevScope {
 public:
  DevScope() = default;
  DevScope(const DevScope&) = delete;
  DevScope& operator=(const DevScope&) = delete;
  ~DevScope() {
    for (void* p : p_) dfree(p);
  }
  template <class T>
  T* alloc(size_t count) {
    p_.reserve(p_.size() + 1);  // the push below cannot fail after the block is taken
    T* q = dalloc<T>(count);
    p_.push_back(q);
    return q;
  }
  // takes ownership of a block from dalloc (nullptr is ignored)
  void adopt(void* q) {
    if (!q) return;
    try {
      p_.push_back(q);
    } catch (...) {
      dfree(q);
      throw;
    }
  }

 private:
  std::vector<void*> p_;
};

// Streams, pinned host blocks and instantiated graphs of released objects, kept per device for
// the next object:
// creating and destroying them costs tens of microseconds each, which a tree and a matrix per
// job would pay every job. A pooled stream is idle (its object retired it); a pooled block is
// no longer read by the device.
class ObjectPool {
 public:
  static ObjectPool& get() {
    static ObjectPool* p = new ObjectPool();  // lives to the end of the process
    return *p;
  }
  hipStream_t takeStream() {
    const int dev = device();
    {
      std::lock_guard<std::mutex> lock(mu_);
      for (size_t i = streams_.size(); i-- > 0;)
        if (streams_[i].first == dev) {
          const hipStream_t s = streams_[i].second;
          streams_.erase(streams_.begin() + i);
          return s;
        }
    }
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreate(&s));
    return s;
  }
  void giveStream(hipStream_t s) {
    const int dev = device();
    std::lock_guard<std::mutex> lock(mu_);
    if (streams_.size() < kKeep) streams_.push_back({dev, s});
    else (void)hipStreamDestroy(s);
  }
  void* takePinned(size_t bytes) {
    {
      std::lock_guard<std::mutex> lock(mu_);
      for (size_t i = pinned_.size(); i-- > 0;)
        if (pinned_[i].first == bytes) {
          void* q = pinned_[i].second;
          pinned_.erase(pinned_.begin() + i);
          return q;
        }
    }
    void* q = nullptr;
    HIPCHK(hipHostMalloc(&q, bytes, hipHostMallocDefault));
    return q;
  }
  void givePinned(void* q, size_t bytes) {
    std::lock_guard<std::mutex> lock(mu_);
    if (pinned_.size() < kKeep) pinned_.push_back({bytes, q});
    else (void)hipHostFree(q);
  }
  // instantiated graphs of a kind (the caller's tag) released with their object, for the next
  // object to update in place (hipGraphExecUpdate) instead of instantiating anew
  hipGraphExec_t takeExec(int kind) {
    const int dev = device();
    std::lock_guard<std::mutex> lock(mu_);
    for (size_t i = execs_.size(); i-- > 0;)
      if (execs_[i].dev == dev && execs_[i].kind == kind) {
        const hipGraphExec_t e = execs_[i].exec;
        execs_.erase(execs_.begin() + i);
        return e;
      }
    return nullptr;
  }
  void giveExec(int kind, hipGraphExec_t e) {
    const int dev = device();
    std::lock_guard<std::mutex> lock(mu_);
    if (execs_.size() < kKeep) execs_.push_back({dev, kind, e});
    else (void)hipGraphExecDestroy(e);
  }

 private:
  static constexpr size_t kKeep = 64;
  static int device() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) {
      (void)hipGetLastError();
      d = 0;
    }
    return d;
  }
  struct PooledExec {
    int dev, kind;
    hipGraphExec_t exec;
  };
  std::mutex mu_;
  std::vector<std::pair<int, hipStream_t>> streams_;
  std::vector<std::pair<size_t, void*>> pinned_;
  std::vector<PooledExec> execs_;
};

// a stream of one of the library's objects: registered for the cache's fences on creation,
// synchronised and retired before the object frees its memory, then released to the pool
inline void streamCreate(hipStream_t* s) {
  *s = ObjectPool::get().takeStream();
  DeviceCache::get().addStream(*s);
}
inline void streamRetire(hipStream_t s) {
  if (s) DeviceCache::get().retireStream(s);
}
inline void streamRelease(hipStream_t s) {
  if (s) ObjectPool::get().giveStream(s);
}

inline uint32_t ceilDiv(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

}  // namespace b12x
