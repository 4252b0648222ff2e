// SyncedMemory: the reference's host/device buffer handle
// (caffe/interface.hpp:7-29, caffe/syncedmem.hpp:47-86), on HIP.
//
// A value-type handle; copies share one buffer pair (shared_ptr).  Host
// memory is pinned (hipHostMalloc).  Both sides are allocated on first use
// and zero-filled when nothing has been written yet.  The head records where
// the current data lives:
//   UNINITIALIZED -> first access allocates and zero-fills that side
//   HEAD_AT_CPU / HEAD_AT_GPU -> the other side is stale; reading it copies
//   SYNCED -> both sides equal
// mutable_* returns a writable pointer and moves the head to that side;
// *_async variants copy with hipMemcpyAsync on the given stream (the caller
// synchronizes before touching the data on the other side).  HIP failures
// throw pl::HipError (the reference exits the process instead).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>

namespace pl {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& what) : std::runtime_error(what) {}
};

void hip_check(hipError_t e, const char* what);

class SyncedMemoryKernel {
 public:
  enum SyncedHead { UNINITIALIZED, HEAD_AT_CPU, HEAD_AT_GPU, SYNCED };
  explicit SyncedMemoryKernel(size_t size);
  ~SyncedMemoryKernel();
  SyncedMemoryKernel(const SyncedMemoryKernel&) = delete;
  SyncedMemoryKernel& operator=(const SyncedMemoryKernel&) = delete;

  const void* cpu_data();
  const void* gpu_data();
  void set_cpu_data(void* data);
  void set_gpu_data(void* data);
  void* mutable_cpu_data();
  void* mutable_cpu_data_async(hipStream_t stream);
  void* mutable_gpu_data();
  void* mutable_gpu_data_async(hipStream_t stream);
  void async_gpu_push(hipStream_t stream);
  SyncedHead head() const { return head_; }
  size_t size() const { return size_; }

 private:
  void to_cpu(bool async, hipStream_t stream);
  void to_gpu(bool async, hipStream_t stream);
  void alloc_cpu();
  void alloc_gpu();
  void* cpu_ptr_ = nullptr;
  void* gpu_ptr_ = nullptr;
  size_t size_ = 0;
  SyncedHead head_ = UNINITIALIZED;
  bool own_cpu_ = false;
  bool own_gpu_ = false;
};

}  // namespace pl

class SyncedMemory {
 public:
  std::shared_ptr<pl::SyncedMemoryKernel> kernel;
  bool type_flag;  // true: Fr elements (the reference's Fr/Fq selector)
  enum SyncedHead { UNINITIALIZED, HEAD_AT_CPU, HEAD_AT_GPU, SYNCED };

  SyncedMemory();
  explicit SyncedMemory(size_t size);
  SyncedMemory(size_t size, bool flag);

  const void* cpu_data();
  void set_cpu_data(void* data);
  const void* gpu_data();
  void set_gpu_data(void* data);
  void* mutable_cpu_data();
  void* mutable_cpu_data_async(hipStream_t stream);
  void* mutable_gpu_data();
  void* mutable_gpu_data_async(hipStream_t stream);
  void async_gpu_push(const hipStream_t& stream);
  SyncedHead head();
  size_t size();
};
