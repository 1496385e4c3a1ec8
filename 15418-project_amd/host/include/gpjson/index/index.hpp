// Device-resident indices of one partition
// (reference: include/gpjson/index/index.hpp:10-109).
//   NewlineIndex        int64[num_lines]: 0, then each '\n' offset
//   StringIndex         uint64[ceil(n/64)]: opening quote and body set
//   LeveledBitmapIndex  uint64[num_levels][ceil(n/64)], level-major
#pragma once

#include "gpjson/device/device.hpp"

#include <cstdint>
#include <utility>

namespace gpjson::index {

class Index {
public:
  Index() = default;
  virtual ~Index() = default;
  Index(Index &) = delete;
  Index &operator=(Index &) = delete;
  Index(Index &&) noexcept = default;
  Index &operator=(Index &&) noexcept = default;

  void *data() { return data_.data(); }
  const void *data() const { return data_.data(); }
  size_t size_bytes() const { return data_.size_bytes(); }

protected:
  explicit Index(device::DeviceArray data) : data_(std::move(data)) {}
  device::DeviceArray data_;
};

class NewlineIndex final : public Index {
public:
  NewlineIndex() = default;
  NewlineIndex(device::DeviceArray data, int64_t num_lines)
      : Index(std::move(data)), num_lines_(num_lines) {}
  NewlineIndex(device::DeviceArray data, int64_t num_lines, device::DeviceArray device_count)
      : Index(std::move(data)), num_lines_(num_lines), device_count_(std::move(device_count)) {}
  NewlineIndex(NewlineIndex &&) noexcept = default;
  NewlineIndex &operator=(NewlineIndex &&) noexcept = default;

  int64_t num_lines() const { return num_lines_; }
  /* Device int64 holding num_lines (allocated on first use if absent). */
  const int64_t *device_num_lines() const;

private:
  int64_t num_lines_{0};
  mutable device::DeviceArray device_count_;
};

class StringIndex final : public Index {
public:
  StringIndex() = default;
  explicit StringIndex(device::DeviceArray data) : Index(std::move(data)) {}
  StringIndex(StringIndex &&) noexcept = default;
  StringIndex &operator=(StringIndex &&) noexcept = default;
};

class LeveledBitmapIndex final : public Index {
public:
  LeveledBitmapIndex() = default;
  LeveledBitmapIndex(device::DeviceArray data, int num_levels)
      : Index(std::move(data)), num_levels_(num_levels) {}
  LeveledBitmapIndex(LeveledBitmapIndex &&) noexcept = default;
  LeveledBitmapIndex &operator=(LeveledBitmapIndex &&) noexcept = default;
  int num_levels() const { return num_levels_; }

private:
  int num_levels_{0};
};

class BuiltIndices {
public:
  BuiltIndices() = default;
  BuiltIndices(NewlineIndex newline_index, StringIndex string_index,
               LeveledBitmapIndex leveled_bitmap_index)
      : newline_index_(std::move(newline_index)), string_index_(std::move(string_index)),
        leveled_bitmap_index_(std::move(leveled_bitmap_index)) {}
  BuiltIndices(BuiltIndices &&) noexcept = default;
  BuiltIndices &operator=(BuiltIndices &&) noexcept = default;
  BuiltIndices(const BuiltIndices &) = delete;
  BuiltIndices &operator=(const BuiltIndices &) = delete;

  const NewlineIndex &get_newline_index() const { return newline_index_; }
  const StringIndex &get_string_index() const { return string_index_; }
  const LeveledBitmapIndex &get_leveled_bitmap_index() const { return leveled_bitmap_index_; }

private:
  NewlineIndex newline_index_;
  StringIndex string_index_;
  LeveledBitmapIndex leveled_bitmap_index_;
};

} // namespace gpjson::index
