// Index builders (reference: include/gpjson/index/index_builder.hpp:15-111).
// The four builder types of the reference are kept as names and options; on
// MI355X they all run the same single-pass HIP pipeline and return identical
// indices, as the reference's tests require of its four variants.
#pragma once

#include "gpjson/file/file.hpp"
#include "gpjson/file/file_reader.hpp"
#include "gpjson/index/index.hpp"
#include "gpjson/run_context.hpp"

#include <cstddef>
#include <ostream>
#include <type_traits>

namespace gpjson::index {

enum class IndexBuilderType { UNCOMBINED = 0, COMBINED, FUSED, SHAREMEM };

inline std::ostream &operator<<(std::ostream &os, IndexBuilderType type) {
  switch (type) {
  case IndexBuilderType::UNCOMBINED: return os << "UNCOMBINED";
  case IndexBuilderType::COMBINED: return os << "COMBINED";
  case IndexBuilderType::FUSED: return os << "FUSED";
  case IndexBuilderType::SHAREMEM: return os << "SHAREMEM";
  }
  return os << "UNKNOWN(" << static_cast<std::underlying_type_t<IndexBuilderType>>(type) << ")";
}

struct IndexBuilderOptions {
  IndexBuilderType index_builder_type{IndexBuilderType::UNCOMBINED};
  size_t file_partition_size{0};
  /* Launch geometry of the reference; accepted and recorded, never changes
   * results. */
  int grid_size{0};
  int block_size{0};
  int reduction_grid_size{0};
  int reduction_block_size{0};
};

inline std::ostream &operator<<(std::ostream &os, const IndexBuilderOptions &options) {
  return os << "IndexBuilderOptions{index_builder_type=" << options.index_builder_type
            << ", file_partition_size=" << options.file_partition_size
            << ", grid_size=" << options.grid_size << ", block_size=" << options.block_size
            << ", reduction_grid_size=" << options.reduction_grid_size
            << ", reduction_block_size=" << options.reduction_block_size << "}";
}

class IndexBuilder {
public:
  virtual ~IndexBuilder() = default;
  virtual BuiltIndices build(const file::FilePartition &partition, size_t max_depth,
                             const IndexBuilderOptions &options,
                             const gpjson::RunContext &run_context) const = 0;
};

namespace detail {
BuiltIndices build_hip(const file::FilePartition &partition, size_t max_depth,
                       const IndexBuilderOptions &options, const gpjson::RunContext &run_context,
                       const char *builder_name);
}

#define GPJSON_DECLARE_BUILDER(Name)                                                       \
  class Name final : public IndexBuilder {                                                 \
  public:                                                                                  \
    explicit Name(const file::FileReader &file_reader) : file_reader_(&file_reader) {}     \
    BuiltIndices build(const file::FilePartition &partition, size_t max_depth,             \
                       const IndexBuilderOptions &options,                                 \
                       const gpjson::RunContext &run_context) const override {             \
      return detail::build_hip(partition, max_depth, options, run_context, #Name);         \
    }                                                                                      \
    const file::FileReader &file_reader() const { return *file_reader_; }                  \
                                                                                           \
  private:                                                                                 \
    const file::FileReader *file_reader_;                                                  \
  };

GPJSON_DECLARE_BUILDER(UncombinedIndexBuilder)
GPJSON_DECLARE_BUILDER(CombinedIndexBuilder)
GPJSON_DECLARE_BUILDER(FusedIndexBuilder)
GPJSON_DECLARE_BUILDER(SharememIndexBuilder)
#undef GPJSON_DECLARE_BUILDER

} // namespace gpjson::index
