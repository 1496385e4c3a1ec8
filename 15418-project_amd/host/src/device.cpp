// DeviceArray and status checks over the C ABI.
#include "gpjson/device/device.hpp"

#include "gpj_capi.h"
#include "gpjson/error/common.hpp"
#include "gpjson/query/error.hpp"

#include <stdexcept>
#include <string>

namespace gpjson::device {

void check(int32_t status, const char *context) {
  if (status == GPJ_OK) return;
  const std::string msg = std::string(context) + ": " + gpj_last_error();
  switch (status) {
  case GPJ_ERR_QUERY_EXECUTION: throw error::query::QueryExecutionError(msg);
  case GPJ_ERR_MALFORMED_QUERY: throw error::query::MalformedQueryError(msg);
  case GPJ_ERR_UNSUPPORTED_QUERY: throw error::query::UnsupportedQueryError(msg);
  case GPJ_ERR_INVALID_ARGUMENT: throw error::common::ImplementationError(msg);
  default: throw error::device::DeviceError(msg);
  }
}

bool device_available() { return gpj_device_available() != 0; }

void synchronize_and_check() { check(gpj_device_sync(), "device synchronize"); }

DeviceArray::DeviceArray(size_t bytes) : size_bytes_(bytes) {
  if (bytes == 0) return;
  check(gpj_alloc(&ptr_, bytes), "device allocation");
}

DeviceArray::~DeviceArray() {
  if (ptr_) gpj_free(ptr_);
}

DeviceArray::DeviceArray(DeviceArray &&other) noexcept
    : ptr_(other.ptr_), size_bytes_(other.size_bytes_) {
  other.ptr_ = nullptr;
  other.size_bytes_ = 0;
}

DeviceArray &DeviceArray::operator=(DeviceArray &&other) noexcept {
  if (this == &other) return *this;
  if (ptr_) gpj_free(ptr_);
  ptr_ = other.ptr_;
  size_bytes_ = other.size_bytes_;
  other.ptr_ = nullptr;
  other.size_bytes_ = 0;
  return *this;
}

DeviceArray DeviceArray::adopt(void *ptr, size_t bytes) {
  DeviceArray a;
  a.ptr_ = ptr;
  a.size_bytes_ = bytes;
  return a;
}

void DeviceArray::copy_from_host(const void *src, size_t bytes) {
  if (bytes > size_bytes_)
    throw error::device::DeviceError("copy host to device: requested byte count exceeds allocation");
  if (bytes > 0 && src == nullptr)
    throw error::device::DeviceError("copy host to device: host pointer is null");
  if (bytes == 0) return;
  check(gpj_h2d(ptr_, src, bytes, nullptr), "copy host to device");
  check(gpj_stream_sync(nullptr), "copy host to device");
}

void DeviceArray::copy_to_host(void *dst, size_t bytes) const {
  if (bytes > size_bytes_)
    throw error::device::DeviceError("copy device to host: requested byte count exceeds allocation");
  if (bytes > 0 && dst == nullptr)
    throw error::device::DeviceError("copy device to host: host pointer is null");
  if (bytes == 0) return;
  check(gpj_d2h(dst, ptr_, bytes, nullptr), "copy device to host");
}

void DeviceArray::memset(int value) {
  if (size_bytes_ == 0) return;
  std::string host(size_bytes_, static_cast<char>(value));
  copy_from_host(host.data(), size_bytes_);
}

} // namespace gpjson::device
