// One newline-aligned slice of a file: refers to host bytes owned by the
// FileReader and owns its device copy (reference: include/gpjson/file/file.hpp:19-48).
#pragma once

#include "gpjson/device/device.hpp"

#include <cstddef>
#include <memory>
#include <string>

namespace gpjson::file {

class FilePartition {
public:
  FilePartition() = default;
  FilePartition(size_t partition_id, size_t global_start_offset,
                size_t global_end_offset, const void *data);

  FilePartition(const FilePartition &) = delete;
  FilePartition &operator=(const FilePartition &) = delete;
  FilePartition(FilePartition &&) noexcept = default;
  FilePartition &operator=(FilePartition &&) noexcept = default;

  size_t partition_id() const { return partition_id_; }
  size_t global_start_offset() const { return global_start_offset_; }
  size_t global_end_offset() const { return global_end_offset_; }
  size_t size_bytes() const { return global_end_offset_ - global_start_offset_; }
  const void *host_bytes() const { return host_data_; }
  const void *device_bytes() const { return device_data_.data(); }
  bool device_loaded() const;
  /* Copy the partition to the device; idempotent.  Completes a copy started
   * by prefetch_to_device. */
  void load_to_device();
  /* Start the copy on `stream` and return (no counterpart in the reference:
   * the engine overlaps the next partition's copy with the current one's
   * index and queries).  device_bytes() is usable after load_to_device(). */
  void prefetch_to_device(void *stream);

private:
  size_t partition_id_{0};
  size_t global_start_offset_{0};
  size_t global_end_offset_{0};
  const void *host_data_{nullptr};
  device::DeviceArray device_data_;
  std::unique_ptr<device::PendingCopy> pending_; // destroyed (waited for) first
};

struct FileMetadata {
  std::string file_path;
  size_t file_size_bytes{0};
  size_t num_partitions{0};
};

} // namespace gpjson::file
