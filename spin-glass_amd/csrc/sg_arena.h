// sg_arena.h — device-memory arena of the batch/serving engine.
//
// One hipMalloc at set-up; jobs take and return blocks with no device call,
// so admitting and retiring a job never synchronises the device (hipFree
// does).  First fit over an offset-ordered free list, neighbours coalesced on
// release.  Not thread-safe: the engine owns it and runs on one host thread.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace sg {

class DevArena {
 public:
  static constexpr size_t kAlign = 256;

  DevArena() = default;
  DevArena(const DevArena &) = delete;
  DevArena &operator=(const DevArena &) = delete;
  ~DevArena() {
    if (base_) (void)hipFree(base_);
  }

  hipError_t init(size_t bytes) {
    bytes = round_up(bytes);
    hipError_t e = hipMalloc((void **)&base_, bytes);
    if (e != hipSuccess) {
      base_ = nullptr;
      return e;
    }
    cap_ = bytes;
    free_.clear();
    used_.clear();
    free_[0] = bytes;
    in_use_ = 0;
    return hipSuccess;
  }

  static size_t round_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

  // nullptr when no free block is large enough
  void *alloc(size_t bytes) {
    bytes = round_up(bytes ? bytes : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, sz = it->second;
      free_.erase(it);
      if (sz > bytes) free_[off + bytes] = sz - bytes;
      used_[off] = bytes;
      in_use_ += bytes;
      return base_ + off;
    }
    return nullptr;
  }

  // returns false for a pointer this arena did not hand out
  bool release(void *p) {
    if (!p) return true;
    const size_t off = (size_t)((char *)p - base_);
    auto u = used_.find(off);
    if (u == used_.end()) return false;
    size_t start = off, sz = u->second;
    in_use_ -= sz;
    used_.erase(u);
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first == start + sz) {
      sz += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        sz += prev->second;
        free_.erase(prev);
      }
    }
    free_[start] = sz;
    return true;
  }

  size_t capacity() const { return cap_; }
  size_t in_use() const { return in_use_; }
  size_t largest_free() const {
    size_t m = 0;
    for (auto &f : free_) m = f.second > m ? f.second : m;
    return m;
  }
  bool valid() const { return base_ != nullptr; }

 private:
  char *base_ = nullptr;
  size_t cap_ = 0, in_use_ = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  std::map<size_t, size_t> used_;  // offset -> bytes
};

}  // namespace sg
