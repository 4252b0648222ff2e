// SyncedMemory on HIP: pinned host + device buffers with a head-state machine
// (semantics of caffe/syncedmem.cpp:37-145 and caffe/interface.cpp).
#include <hip/hip_runtime.h>

#include <cstring>

#include "interface.hpp"

namespace pl {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

SyncedMemoryKernel::SyncedMemoryKernel(size_t size) : size_(size) {}

SyncedMemoryKernel::~SyncedMemoryKernel() {
  if (cpu_ptr_ && own_cpu_) (void)hipHostFree(cpu_ptr_);
  if (gpu_ptr_ && own_gpu_) (void)hipFree(gpu_ptr_);
}

void SyncedMemoryKernel::alloc_cpu() {
  if (cpu_ptr_) return;
  hip_check(hipHostMalloc(&cpu_ptr_, size_ ? size_ : 1, hipHostMallocDefault), "hipHostMalloc");
  own_cpu_ = true;
}

void SyncedMemoryKernel::alloc_gpu() {
  if (gpu_ptr_) return;
  hip_check(hipMalloc(&gpu_ptr_, size_ ? size_ : 1), "hipMalloc");
  own_gpu_ = true;
}

void SyncedMemoryKernel::to_cpu(bool async, hipStream_t stream) {
  switch (head_) {
    case UNINITIALIZED:
      alloc_cpu();
      std::memset(cpu_ptr_, 0, size_);
      head_ = HEAD_AT_CPU;
      break;
    case HEAD_AT_GPU:
      alloc_cpu();
      if (size_) {
        if (async)
          hip_check(hipMemcpyAsync(cpu_ptr_, gpu_ptr_, size_, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync D2H");
        else
          hip_check(hipMemcpy(cpu_ptr_, gpu_ptr_, size_, hipMemcpyDeviceToHost), "hipMemcpy D2H");
      }
      head_ = SYNCED;
      break;
    case HEAD_AT_CPU:
    case SYNCED:
      break;
  }
}

void SyncedMemoryKernel::to_gpu(bool async, hipStream_t stream) {
  switch (head_) {
    case UNINITIALIZED:
      alloc_gpu();
      if (async)
        hip_check(hipMemsetAsync(gpu_ptr_, 0, size_, stream), "hipMemsetAsync");
      else
        hip_check(hipMemset(gpu_ptr_, 0, size_), "hipMemset");
      head_ = HEAD_AT_GPU;
      break;
    case HEAD_AT_CPU:
      alloc_gpu();
      if (size_) {
        if (async)
          hip_check(hipMemcpyAsync(gpu_ptr_, cpu_ptr_, size_, hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
        else
          hip_check(hipMemcpy(gpu_ptr_, cpu_ptr_, size_, hipMemcpyHostToDevice), "hipMemcpy H2D");
      }
      head_ = SYNCED;
      break;
    case HEAD_AT_GPU:
    case SYNCED:
      break;
  }
}

const void* SyncedMemoryKernel::cpu_data() {
  to_cpu(false, nullptr);
  return cpu_ptr_;
}

const void* SyncedMemoryKernel::gpu_data() {
  to_gpu(false, nullptr);
  return gpu_ptr_;
}

void SyncedMemoryKernel::set_cpu_data(void* data) {
  if (cpu_ptr_ && own_cpu_) (void)hipHostFree(cpu_ptr_);
  cpu_ptr_ = data;
  own_cpu_ = false;
  head_ = HEAD_AT_CPU;
}

void SyncedMemoryKernel::set_gpu_data(void* data) {
  if (gpu_ptr_ && own_gpu_) (void)hipFree(gpu_ptr_);
  gpu_ptr_ = data;
  own_gpu_ = false;
  head_ = HEAD_AT_GPU;
}

void* SyncedMemoryKernel::mutable_cpu_data() {
  to_cpu(false, nullptr);
  head_ = HEAD_AT_CPU;
  return cpu_ptr_;
}

void* SyncedMemoryKernel::mutable_cpu_data_async(hipStream_t stream) {
  to_cpu(true, stream);
  head_ = HEAD_AT_CPU;
  return cpu_ptr_;
}

void* SyncedMemoryKernel::mutable_gpu_data() {
  to_gpu(false, nullptr);
  head_ = HEAD_AT_GPU;
  return gpu_ptr_;
}

void* SyncedMemoryKernel::mutable_gpu_data_async(hipStream_t stream) {
  to_gpu(true, stream);
  head_ = HEAD_AT_GPU;
  return gpu_ptr_;
}

void SyncedMemoryKernel::async_gpu_push(hipStream_t stream) {
  alloc_gpu();
  if (!cpu_ptr_) {
    alloc_cpu();
    std::memset(cpu_ptr_, 0, size_);
  }
  hip_check(hipMemcpyAsync(gpu_ptr_, cpu_ptr_, size_, hipMemcpyHostToDevice, stream), "hipMemcpyAsync push");
  head_ = SYNCED;
}

}  // namespace pl

SyncedMemory::SyncedMemory() : kernel(nullptr), type_flag(true) {}
SyncedMemory::SyncedMemory(size_t size) : kernel(std::make_shared<pl::SyncedMemoryKernel>(size)), type_flag(true) {}
SyncedMemory::SyncedMemory(size_t size, bool flag)
    : kernel(std::make_shared<pl::SyncedMemoryKernel>(size)), type_flag(flag) {}

static pl::SyncedMemoryKernel& k_of(std::shared_ptr<pl::SyncedMemoryKernel>& k) {
  if (!k) throw std::logic_error("SyncedMemory: empty handle");
  return *k;
}

const void* SyncedMemory::cpu_data() { return k_of(kernel).cpu_data(); }
void SyncedMemory::set_cpu_data(void* data) { k_of(kernel).set_cpu_data(data); }
const void* SyncedMemory::gpu_data() { return k_of(kernel).gpu_data(); }
void SyncedMemory::set_gpu_data(void* data) { k_of(kernel).set_gpu_data(data); }
void* SyncedMemory::mutable_cpu_data() { return k_of(kernel).mutable_cpu_data(); }
void* SyncedMemory::mutable_cpu_data_async(hipStream_t s) { return k_of(kernel).mutable_cpu_data_async(s); }
void* SyncedMemory::mutable_gpu_data() { return k_of(kernel).mutable_gpu_data(); }
void* SyncedMemory::mutable_gpu_data_async(hipStream_t s) { return k_of(kernel).mutable_gpu_data_async(s); }
void SyncedMemory::async_gpu_push(const hipStream_t& s) { k_of(kernel).async_gpu_push(s); }
SyncedMemory::SyncedHead SyncedMemory::head() { return (SyncedHead)k_of(kernel).head(); }
size_t SyncedMemory::size() { return kernel ? kernel->size() : 0; }
