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

HostBuffer::HostBuffer(size_t bytes) : size_bytes_(bytes) {
  if (bytes == 0) return;
  void *p = nullptr;
  if (gpj_host_alloc(&p, bytes) == GPJ_OK && p) {
    ptr_ = static_cast<std::byte *>(p);
    pinned_ = true;
  } else {
    ptr_ = new std::byte[bytes];
  }
}

void HostBuffer::release() {
  if (!ptr_) return;
  if (pinned_) gpj_host_free(ptr_);
  else delete[] ptr_;
  ptr_ = nullptr;
  size_bytes_ = 0;
  pinned_ = false;
}

HostBuffer::~HostBuffer() { release(); }

HostBuffer::HostBuffer(HostBuffer &&other) noexcept
    : ptr_(other.ptr_), size_bytes_(other.size_bytes_), pinned_(other.pinned_) {
  other.ptr_ = nullptr;
  other.size_bytes_ = 0;
  other.pinned_ = false;
}

HostBuffer &HostBuffer::operator=(HostBuffer &&other) noexcept {
  if (this == &other) return *this;
  release();
  ptr_ = other.ptr_;
  size_bytes_ = other.size_bytes_;
  pinned_ = other.pinned_;
  other.ptr_ = nullptr;
  other.size_bytes_ = 0;
  other.pinned_ = false;
  return *this;
}

Stream::~Stream() {
  if (stream_) gpj_stream_destroy(stream_);
}

void *Stream::get() {
  if (!stream_) check(gpj_stream_create(&stream_), "stream");
  return stream_;
}

PendingCopy::PendingCopy(void *d_dst, const void *h_src, size_t bytes, void *stream) {
  check(gpj_event_create(&event_), "copy event");
  const int32_t st = gpj_h2d(d_dst, h_src, bytes, stream);
  if (st == GPJ_OK) {
    const int32_t rec = gpj_event_record(event_, stream);
    if (rec == GPJ_OK) return;
    gpj_stream_sync(stream);
    gpj_event_destroy(event_);
    event_ = nullptr;
    check(rec, "copy event");
  }
  gpj_event_destroy(event_);
  event_ = nullptr;
  check(st, "copy host to device");
}

void PendingCopy::wait() {
  if (!event_) return;
  const int32_t st = gpj_event_sync(event_);
  gpj_event_destroy(event_);
  event_ = nullptr;
  check(st, "copy host to device");
}

PendingCopy::~PendingCopy() {
  if (!event_) return;
  gpj_event_sync(event_);
  gpj_event_destroy(event_);
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
