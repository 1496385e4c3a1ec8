// Device memory and status checks over the C ABI (include/gpj_capi.h).
// Replaces the reference's CUDA runtime wrapper: DeviceArray, check,
// device_available, synchronize_and_check (include/gpjson/cuda/cuda.hpp:10-45)
// and CudaError (include/gpjson/cuda/error.hpp).
#pragma once

#include "gpjson/error/error.hpp"

#include <cstddef>
#include <cstdint>

namespace gpjson::error::device {

class DeviceError : public gpjson::error::GpJSONError {
public:
  using GpJSONError::GpJSONError;
};

} // namespace gpjson::error::device

namespace gpjson::device {

/* RAII owner of one device allocation; move-only. */
class DeviceArray {
public:
  DeviceArray() = default;
  explicit DeviceArray(size_t bytes);
  ~DeviceArray();

  DeviceArray(DeviceArray &&other) noexcept;
  DeviceArray &operator=(DeviceArray &&other) noexcept;
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;

  /* Take ownership of an allocation made through the C ABI. */
  static DeviceArray adopt(void *ptr, size_t bytes);

  void *data() { return ptr_; }
  const void *data() const { return ptr_; }
  size_t size_bytes() const { return size_bytes_; }
  bool empty() const { return size_bytes_ == 0; }

  template <typename T> T *as() { return static_cast<T *>(ptr_); }
  template <typename T> const T *as() const { return static_cast<const T *>(ptr_); }

  void copy_from_host(const void *src, size_t bytes);
  void copy_to_host(void *dst, size_t bytes) const;
  void memset(int value);

private:
  void *ptr_{nullptr};
  size_t size_bytes_{0};
};

/* Host bytes headed for the device: page-locked when the runtime provides
 * them (copies then run at full link speed and overlap kernels), pageable
 * otherwise (no device, or the pinned allocation fails).  Uninitialized;
 * move-only. */
class HostBuffer {
public:
  HostBuffer() = default;
  explicit HostBuffer(size_t bytes);
  ~HostBuffer();

  HostBuffer(HostBuffer &&other) noexcept;
  HostBuffer &operator=(HostBuffer &&other) noexcept;
  HostBuffer(const HostBuffer &) = delete;
  HostBuffer &operator=(const HostBuffer &) = delete;

  std::byte *data() { return ptr_; }
  const std::byte *data() const { return ptr_; }
  size_t size_bytes() const { return size_bytes_; }
  bool pinned() const { return pinned_; }

private:
  void release();
  std::byte *ptr_{nullptr};
  size_t size_bytes_{0};
  bool pinned_{false};
};

/* A HIP stream of its own (non-blocking, default priority), created on first
 * use; move-only. */
class Stream {
public:
  Stream() = default;
  ~Stream();
  Stream(const Stream &) = delete;
  Stream &operator=(const Stream &) = delete;
  void *get();

private:
  void *stream_{nullptr};
};

/* A host-to-device copy in flight on a caller's stream: waits for it (and
 * releases its event) when completed or destroyed. */
class PendingCopy {
public:
  PendingCopy(void *d_dst, const void *h_src, size_t bytes, void *stream);
  ~PendingCopy();
  PendingCopy(const PendingCopy &) = delete;
  PendingCopy &operator=(const PendingCopy &) = delete;
  void wait();

private:
  void *event_{nullptr};
};

/* Throws DeviceError (or the matching GpJSONError subclass) for a non-zero
 * gpj_status, with `context` and the library's message. */
void check(int32_t status, const char *context);
bool device_available();
void synchronize_and_check();

} // namespace gpjson::device
