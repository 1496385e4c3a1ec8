// IndexBuilder::build on the HIP pipeline (gpj_index_build of include/gpj_capi.h).
// Replaces the reference's four builders (src/gpjson/index/orig_index_builder.cu,
// fuse_index_builder.cu, sharemem_index_builder.cu); the builder type and the
// launch geometry are recorded in the profiler segment name only.
#include "gpjson/index/index_builder.hpp"

#include "gpj_capi.h"
#include "gpjson/error/common.hpp"

#include <string>

namespace gpjson::index {

const int64_t *NewlineIndex::device_num_lines() const {
  if (device_count_.data() == nullptr) {
    device::DeviceArray d(sizeof(int64_t));
    d.copy_from_host(&num_lines_, sizeof(int64_t));
    device_count_ = std::move(d);
  }
  return device_count_.as<int64_t>();
}

namespace detail {

BuiltIndices build_hip(const file::FilePartition &partition, size_t max_depth,
                       const IndexBuilderOptions &options, const gpjson::RunContext &run_context,
                       const char *builder_name) {
  if (partition.size_bytes() > 0 && partition.device_bytes() == nullptr)
    throw error::common::ImplementationError(
        "Index building requires a GPU-resident partition buffer");
  if (max_depth > static_cast<size_t>(GPJ_MAX_INDEX_LEVELS))
    throw error::common::ImplementationError(
        "Index builder supports at most " + std::to_string(GPJ_MAX_INDEX_LEVELS) + " levels");
  profiler::Profiler &profiler = run_context.profiler;
  const auto seg = profiler.beginf("%s hip_index_build (grid=%d block=%d)", builder_name,
                                   options.grid_size, options.block_size);
  gpj_index_opts opts{options.grid_size, options.block_size, options.reduction_grid_size,
                      options.reduction_block_size, 0};
  gpj_index raw{};
  device::check(gpj_index_build(static_cast<const uint8_t *>(partition.device_bytes()),
                                partition.size_bytes(), static_cast<int32_t>(max_depth), &opts,
                                &raw, nullptr),
                "index build");
  profiler.end(seg);
  // take over the allocations
  const size_t words = static_cast<size_t>(raw.level_size);
  NewlineIndex newline(device::DeviceArray::adopt(raw.d_newline, raw.num_lines * sizeof(int64_t)),
                       raw.num_lines,
                       device::DeviceArray::adopt(raw.d_num_lines, sizeof(int64_t)));
  StringIndex string(device::DeviceArray::adopt(raw.d_string, words * sizeof(uint64_t)));
  LeveledBitmapIndex leveled(
      device::DeviceArray::adopt(raw.d_leveled, words * sizeof(uint64_t) * max_depth),
      static_cast<int>(max_depth));
  return BuiltIndices(std::move(newline), std::move(string), std::move(leveled));
}

} // namespace detail
} // namespace gpjson::index
