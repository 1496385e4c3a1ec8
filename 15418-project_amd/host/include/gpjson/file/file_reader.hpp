// File readers and newline-aligned partitioning
// (reference: include/gpjson/file/file_reader.hpp:39-95).
#pragma once

#include "gpjson/file/file.hpp"

#include <cstddef>
#include <ostream>
#include <string>
#include <type_traits>
#include <vector>

namespace gpjson::file {

enum class FileReaderType { COPIED = 0, MMAP };

inline std::ostream &operator<<(std::ostream &os, FileReaderType type) {
  switch (type) {
  case FileReaderType::COPIED: return os << "COPIED";
  case FileReaderType::MMAP: return os << "MMAP";
  }
  return os << "UNKNOWN(" << static_cast<std::underlying_type_t<FileReaderType>>(type) << ")";
}

struct FileReaderOptions {
  FileReaderType file_reader_type{FileReaderType::COPIED};
};

inline std::ostream &operator<<(std::ostream &os, const FileReaderOptions &options) {
  return os << "FileReaderOptions{file_reader_type=" << options.file_reader_type << "}";
}

class FileReader {
public:
  virtual ~FileReader() = default;
  virtual void create_partitions(size_t partition_size_bytes) = 0;
  virtual std::vector<FilePartition> &get_partitions() = 0;
  virtual const std::vector<FilePartition> &get_partitions() const = 0;
  virtual const FileMetadata &metadata() const = 0;

protected:
  FileReader() = default;
};

/* Reads the whole file into host memory. */
class CopiedFileReader final : public FileReader {
public:
  explicit CopiedFileReader(std::string file_path);
  void create_partitions(size_t partition_size_bytes) override;
  std::vector<FilePartition> &get_partitions() override { return partitions_; }
  const std::vector<FilePartition> &get_partitions() const override { return partitions_; }
  const FileMetadata &metadata() const override { return metadata_; }

private:
  std::string file_path_;
  FileMetadata metadata_;
  device::HostBuffer bytes_;              // the file, page-locked when possible
  std::vector<FilePartition> partitions_; // destroyed first: waits for their copies
};

/* Maps the file read-only. */
class MmapFileReader final : public FileReader {
public:
  explicit MmapFileReader(std::string file_path);
  ~MmapFileReader() override;
  MmapFileReader(const MmapFileReader &) = delete;
  MmapFileReader &operator=(const MmapFileReader &) = delete;
  void create_partitions(size_t partition_size_bytes) override;
  std::vector<FilePartition> &get_partitions() override { return partitions_; }
  const std::vector<FilePartition> &get_partitions() const override { return partitions_; }
  const FileMetadata &metadata() const override { return metadata_; }

private:
  void map();
  void unmap() noexcept;
  std::string file_path_;
  FileMetadata metadata_;
  const std::byte *mapped_{nullptr};
  size_t mapped_size_{0};
  int fd_{-1};
  std::vector<FilePartition> partitions_;
};

/* Cut points of create_partitions (file_reader.cpp:41-77): each partition
 * ends right before the last '\n' at or below start+size; the newline at a
 * cut belongs to no partition.  Throws PartitionError when a cut finds no
 * newline. */
std::vector<FilePartition> partition_bytes(const std::string &file_path,
                                           const std::byte *base, size_t size,
                                           size_t partition_size_bytes);

} // namespace gpjson::file
