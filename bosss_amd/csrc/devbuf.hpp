// Per-call device copies of host arrays, for the host-pointer entry points.
//
// A host that holds its data in managed arrays pins them for one call only
// (DGField_Evaluation.cs:78-110 pins with `fixed`), so a host-pointer entry point
// copies its inputs to the device, runs the device entry point on its own stream,
// copies the outputs back and frees the copies before it returns.  DevArrays
// collects those copies: in() allocates and uploads, out() allocates (zeroed),
// back() queues a download, finish() waits for the stream; the destructor waits
// and frees whatever was allocated, on every return path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace bosss {

class DevArrays {
 public:
  explicit DevArrays(hipStream_t s) : s_(s) {}
  DevArrays(const DevArrays&) = delete;
  DevArrays& operator=(const DevArrays&) = delete;
  ~DevArrays() {
    if (!ptrs_.empty()) (void)hipStreamSynchronize(s_);
    for (void* p : ptrs_) (void)hipFree(p);
  }
  // device copy of host[0..n) (n may be 0: a valid one-element allocation is returned)
  template <class T>
  T* in(const T* host, size_t n) {
    T* d = alloc<T>(n);
    if (d && n && host && err_ == hipSuccess) err_ = hipMemcpyAsync(d, host, n * sizeof(T), hipMemcpyHostToDevice, s_);
    return d;
  }
  // zero-filled device array of n elements
  template <class T>
  T* out(size_t n) {
    T* d = alloc<T>(n);
    if (d && err_ == hipSuccess) err_ = hipMemsetAsync(d, 0, (n ? n : 1) * sizeof(T), s_);
    return d;
  }
  template <class T>
  void back(T* host, const T* dev, size_t n) {
    if (err_ == hipSuccess && n && host && dev) err_ = hipMemcpyAsync(host, dev, n * sizeof(T), hipMemcpyDeviceToHost, s_);
  }
  hipError_t finish() {
    if (err_ == hipSuccess) err_ = hipStreamSynchronize(s_);
    return err_;
  }
  hipError_t error() const { return err_; }
  hipStream_t stream() const { return s_; }

 private:
  template <class T>
  T* alloc(size_t n) {
    if (err_ != hipSuccess) return nullptr;
    void* p = nullptr;
    err_ = hipMalloc(&p, (n ? n : 1) * sizeof(T));
    if (err_ != hipSuccess) return nullptr;
    ptrs_.push_back(p);
    return static_cast<T*>(p);
  }
  hipStream_t s_;
  hipError_t err_ = hipSuccess;
  std::vector<void*> ptrs_;
};

// Extent of a batch of blocks: max over b of off[b] + size(b), or 0 for an empty batch.
template <class F>
inline int64_t batch_extent(int nblk, const int64_t* off, F size) {
  int64_t m = 0;
  for (int b = 0; b < nblk; b++) {
    int64_t e = off[b] + size(b);
    if (e > m) m = e;
  }
  return m;
}

}  // namespace bosss
