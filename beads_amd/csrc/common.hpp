// Core utilities: error type carrying a C-ABI status, HIP error checks and a
// grow-only device buffer. Counterpart of the reference's
// native/include/beads/core/{types.hpp,cuda_check.cuh,device_buffer.cuh}.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "beads_amd.h"
#include "real.hpp"

namespace beads_amd {

using type_id_t = int32_t;
using index_t = uint32_t;
using image_t = int32_t;
using runstep_t = uint64_t;

// An exception that remembers which C-ABI status it maps to.
class Error : public std::runtime_error {
 public:
  Error(beads_status_t status, const std::string& message)
      : std::runtime_error(message), status_(status) {}
  beads_status_t status() const noexcept { return status_; }

 private:
  beads_status_t status_;
};

[[noreturn]] inline void fail(beads_status_t status, const std::string& message) {
  throw Error(status, message);
}

inline void require(bool condition, const std::string& message,
                    beads_status_t status = BEADS_E_VALUE) {
  if (!condition) {
    fail(status, message);
  }
}

inline void hip_check(hipError_t err, const char* what, const char* file, int line) {
  if (err != hipSuccess) {
    fail(BEADS_E_DEVICE, std::string("HIP error in ") + what + " (" + file + ":" +
                             std::to_string(line) + "): " + hipGetErrorString(err));
  }
}

#define BEADS_HIP(expr) ::beads_amd::hip_check((expr), #expr, __FILE__, __LINE__)

// Device allocation owned by RAII; resize() never shrinks (device_buffer.cuh).
template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t count) { resize(count); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), size_(o.size_), cap_(o.cap_) {
    o.ptr_ = nullptr;
    o.size_ = o.cap_ = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      size_ = o.size_;
      cap_ = o.cap_;
      o.ptr_ = nullptr;
      o.size_ = o.cap_ = 0;
    }
    return *this;
  }

  void resize(size_t count) {
    if (count > cap_) {
      release();
      if (count > 0) {
        void* p = nullptr;
        BEADS_HIP(hipMalloc(&p, count * sizeof(T)));
        ptr_ = static_cast<T*>(p);
        cap_ = count;
      }
    }
    size_ = count;
  }
  void zero(hipStream_t stream) {
    if (size_ > 0) {
      BEADS_HIP(hipMemsetAsync(ptr_, 0, size_ * sizeof(T), stream));
    }
  }
  void upload(const T* host, size_t count, hipStream_t stream) {
    resize(count);
    if (count > 0) {
      BEADS_HIP(hipMemcpyAsync(ptr_, host, count * sizeof(T), hipMemcpyHostToDevice, stream));
    }
  }
  void download(T* host, size_t count, hipStream_t stream) const {
    if (count > 0) {
      BEADS_HIP(hipMemcpyAsync(host, ptr_, count * sizeof(T), hipMemcpyDeviceToHost, stream));
    }
  }
  T* data() noexcept { return ptr_; }
  const T* data() const noexcept { return ptr_; }
  size_t size() const noexcept { return size_; }

 private:
  void release() noexcept {
    if (ptr_ != nullptr) {
      (void)hipFree(ptr_);
      ptr_ = nullptr;
    }
    size_ = cap_ = 0;
  }
  T* ptr_ = nullptr;
  size_t size_ = 0;
  size_t cap_ = 0;
};

// Pinned host allocation (hipHostMalloc) for staged device-to-host copies.
template <typename T>
class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  ~PinnedBuffer() {
    if (ptr_ != nullptr) {
      (void)hipHostFree(ptr_);
    }
  }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  void resize(size_t count) {
    if (count > cap_) {
      if (ptr_ != nullptr) {
        (void)hipHostFree(ptr_);
        ptr_ = nullptr;
      }
      void* p = nullptr;
      BEADS_HIP(hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault));
      ptr_ = static_cast<T*>(p);
      cap_ = count;
    }
  }
  T* data() noexcept { return ptr_; }
  const T* data() const noexcept { return ptr_; }

 private:
  T* ptr_ = nullptr;
  size_t cap_ = 0;
};

}  // namespace beads_amd
