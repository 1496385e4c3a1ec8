// File readers and partitioning (reference: src/gpjson/file/file_reader.cpp,
// src/gpjson/file/file.cpp).
#include "gpjson/file/error.hpp"
#include "gpjson/file/file.hpp"
#include "gpjson/file/file_reader.hpp"

#include <cerrno>
#include <cstring>
#include <fcntl.h>
#include <filesystem>
#include <algorithm>
#include <atomic>
#include <sstream>
#include <sys/mman.h>
#include <thread>
#include <unistd.h>
#include <vector>

namespace gpjson::file {

FilePartition::FilePartition(size_t partition_id, size_t global_start_offset,
                             size_t global_end_offset, const void *data)
    : partition_id_(partition_id), global_start_offset_(global_start_offset),
      global_end_offset_(global_end_offset), host_data_(data) {}

bool FilePartition::device_loaded() const {
  return size_bytes() == 0 || (device_data_.data() != nullptr && !pending_);
}

void FilePartition::prefetch_to_device(void *stream) {
  if (size_bytes() == 0 || device_data_.data() != nullptr) return;
  device::DeviceArray d(size_bytes());
  pending_ = std::make_unique<device::PendingCopy>(d.data(), host_data_, size_bytes(), stream);
  device_data_ = std::move(d);
}

void FilePartition::load_to_device() {
  if (pending_) {
    pending_->wait();
    pending_.reset();
    return;
  }
  if (device_loaded()) return;
  device::DeviceArray d(size_bytes());
  d.copy_from_host(host_data_, size_bytes());
  device_data_ = std::move(d);
}

std::vector<FilePartition> partition_bytes(const std::string &file_path,
                                           const std::byte *base, size_t size,
                                           size_t partition_size_bytes) {
  std::vector<FilePartition> parts;
  if (partition_size_bytes == 0 || size == 0 || size <= partition_size_bytes) {
    parts.emplace_back(0, 0, size, base);
    return parts;
  }
  size_t start = 0, id = 0;
  while (size - start > partition_size_bytes) {
    // scan back from start+size for the last newline of this partition
    const size_t target = start + partition_size_bytes;
    size_t cut = target;
    bool found = false;
    for (;;) {
      if (base[cut] == std::byte{'\n'}) {
        found = true;
        break;
      }
      if (cut == start) break;
      --cut;
    }
    if (!found) {
      std::ostringstream m;
      m << "Cannot partition file '" << file_path << "' using partition size "
        << partition_size_bytes << ": no newline found before offset " << target << ".";
      throw error::file::PartitionError(m.str());
    }
    const size_t next = cut + 1;
    if (next >= size) break;
    parts.emplace_back(id++, start, next - 1, base + start);
    start = next;
  }
  parts.emplace_back(id, start, size, base + start);
  return parts;
}

namespace {
size_t file_size_or_throw(const std::string &path) {
  std::error_code ec;
  const auto sz = std::filesystem::file_size(std::filesystem::path(path), ec);
  if (ec)
    throw error::file::FileOpenError("Failed to get size for file '" + path + "': " + ec.message());
  return static_cast<size_t>(sz);
}

// Reads bytes [0, size) of the open file into dst: 64 MiB pieces on up to 16
// threads (a cached file is then read at memory speed, not one core's copy
// speed).  False if any piece comes up short.
bool read_all(int fd, std::byte *dst, size_t size) {
  constexpr size_t kPiece = size_t{64} << 20;
  const size_t pieces = (size + kPiece - 1) / kPiece;
  const size_t workers =
      std::min<size_t>({std::max(1u, std::thread::hardware_concurrency()), 16, pieces});
  std::atomic<size_t> next{0};
  std::atomic<bool> ok{true};
  auto work = [&] {
    for (size_t k; ok.load(std::memory_order_relaxed) && (k = next++) < pieces;) {
      size_t off = k * kPiece;
      const size_t end = std::min(size, off + kPiece);
      while (off < end) {
        const ssize_t got = pread(fd, dst + off, end - off, static_cast<off_t>(off));
        if (got < 0 && errno == EINTR) continue;
        if (got <= 0) {
          ok = false;
          return;
        }
        off += static_cast<size_t>(got);
      }
    }
  };
  std::vector<std::thread> pool;
  try {
    for (size_t t = 1; t < workers; ++t) pool.emplace_back(work);
  } catch (...) { // fewer threads: the ones running and this one finish the pieces
  }
  work();
  for (auto &th : pool) th.join();
  return ok.load();
}
} // namespace

CopiedFileReader::CopiedFileReader(std::string file_path) : file_path_(std::move(file_path)) {
  metadata_.file_path = file_path_;
}

void CopiedFileReader::create_partitions(size_t partition_size_bytes) {
  metadata_.file_size_bytes = file_size_or_throw(file_path_);
  partitions_.clear();
  bytes_ = device::HostBuffer(metadata_.file_size_bytes); // pinned when possible
  const int fd = open(file_path_.c_str(), O_RDONLY);
  if (fd == -1) throw error::file::FileOpenError("Failed to open file '" + file_path_ + "' for reading.");
  const bool ok = metadata_.file_size_bytes == 0 ||
                  read_all(fd, bytes_.data(), metadata_.file_size_bytes);
  close(fd);
  if (!ok)
    throw error::file::FileReadError("Failed to read all bytes from file '" + file_path_ + "'.");
  partitions_ = partition_bytes(file_path_, bytes_.size_bytes() == 0 ? nullptr : bytes_.data(),
                                metadata_.file_size_bytes, partition_size_bytes);
  metadata_.num_partitions = partitions_.size();
}

MmapFileReader::MmapFileReader(std::string file_path) : file_path_(std::move(file_path)) {
  metadata_.file_path = file_path_;
}

MmapFileReader::~MmapFileReader() {
  partitions_.clear(); // waits for copies still reading the mapping
  unmap();
}

void MmapFileReader::map() {
  if (fd_ != -1 || mapped_size_ != 0) return;
  metadata_.file_size_bytes = file_size_or_throw(file_path_);
  fd_ = open(file_path_.c_str(), O_RDONLY);
  if (fd_ == -1)
    throw error::file::FileOpenError("Failed to open file '" + file_path_ +
                                     "' for reading: " + std::strerror(errno) + ".");
  if (metadata_.file_size_bytes == 0) return;
  void *m = mmap(nullptr, metadata_.file_size_bytes, PROT_READ, MAP_PRIVATE, fd_, 0);
  if (m == MAP_FAILED) {
    const std::string e = std::strerror(errno);
    unmap();
    throw error::file::FileReadError("Failed to mmap file '" + file_path_ + "': " + e + ".");
  }
  mapped_ = static_cast<const std::byte *>(m);
  mapped_size_ = metadata_.file_size_bytes;
}

void MmapFileReader::unmap() noexcept {
  if (mapped_ && mapped_size_) munmap(const_cast<std::byte *>(mapped_), mapped_size_);
  mapped_ = nullptr;
  mapped_size_ = 0;
  if (fd_ != -1) {
    close(fd_);
    fd_ = -1;
  }
}

void MmapFileReader::create_partitions(size_t partition_size_bytes) {
  map();
  partitions_ = partition_bytes(file_path_, mapped_, metadata_.file_size_bytes, partition_size_bytes);
  metadata_.num_partitions = partitions_.size();
}

} // namespace gpjson::file
