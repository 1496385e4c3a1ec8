// Engine::query (reference: src/gpjson/engine.cpp:17-175): compile the
// queries, partition the file at newlines, and per partition copy it to the
// device, build the indices, run the queries, materialize and append.  The
// next partition's copy runs on a stream of its own while the current one is
// indexed and queried (the reference copies each partition synchronously).
#include "gpjson/engine.hpp"

#include "gpjson/error/common.hpp"
#include "gpjson/profiler/profiler.hpp"
#include "gpjson/query/query_compiler.hpp"
#include "gpjson/run_context.hpp"

#include <memory>

namespace gpjson {

namespace {

std::unique_ptr<file::FileReader> make_reader(const std::string &path,
                                              const file::FileReaderOptions &o) {
  switch (o.file_reader_type) {
  case file::FileReaderType::COPIED: return std::make_unique<file::CopiedFileReader>(path);
  case file::FileReaderType::MMAP: return std::make_unique<file::MmapFileReader>(path);
  }
  throw error::common::ImplementationError("Undefined file reader type.");
}

std::unique_ptr<index::IndexBuilder> make_builder(const file::FileReader &reader,
                                                  const index::IndexBuilderOptions &o) {
  switch (o.index_builder_type) {
  case index::IndexBuilderType::UNCOMBINED:
    return std::make_unique<index::UncombinedIndexBuilder>(reader);
  case index::IndexBuilderType::COMBINED:
    return std::make_unique<index::CombinedIndexBuilder>(reader);
  case index::IndexBuilderType::FUSED: return std::make_unique<index::FusedIndexBuilder>(reader);
  case index::IndexBuilderType::SHAREMEM:
    return std::make_unique<index::SharememIndexBuilder>(reader);
  }
  throw error::common::ImplementationError("Undefined index builder type.");
}

} // namespace

query::MaterializedBatchResult Engine::query(const std::string &file_path,
                                             const std::vector<std::string> &queries_src,
                                             const EngineOptions &options) const {
  if (queries_src.empty()) return {};
  profiler::Profiler profiler("Engine::query profiler");
  RunContext ctx{profiler};
  const auto total = profiler.begin("Engine::query");

  query::QueryCompiler compiler(options);
  query::QueryExecutor executor(options);
  const auto comp = profiler.begin("query_compilation");
  query::BatchCompiledQuery compiled;
  for (const auto &q : queries_src) compiled.add(compiler.compile(q, options));
  profiler.end(comp);

  device::Stream copy_stream; // outlives the reader's partitions and their copies
  auto reader = make_reader(file_path, options.file_reader_options);
  const auto part_seg = profiler.begin("file_reader.create_partitions");
  reader->create_partitions(options.index_builder_options.file_partition_size);
  profiler.end(part_seg);
  auto builder = make_builder(*reader, options.index_builder_options);

  std::vector<query::MaterializedQueryResult> merged;
  merged.reserve(compiled.size());
  for (const auto &q : compiled.queries()) merged.emplace_back(q.query_text());

  auto &partitions = reader->get_partitions();
  for (size_t i = 0; i < partitions.size(); ++i) {
    auto &partition = partitions[i];
    const size_t pid = partition.partition_id();
    const auto pseg = profiler.begin_nestedf("partition %zu total", pid);
    const auto lseg = profiler.begin_nestedf("partition %zu load_to_device", pid);
    partition.load_to_device();
    if (i + 1 < partitions.size()) partitions[i + 1].prefetch_to_device(copy_stream.get());
    profiler.end(lseg);
    const auto bseg = profiler.begin_nestedf("partition %zu build_index", pid);
    index::BuiltIndices built = builder->build(partition, compiled.get_max_depth(),
                                               options.index_builder_options, ctx);
    profiler.end(bseg);
    const auto qseg = profiler.begin_nestedf("partition %zu execute_query", pid);
    query::BatchQueryResult batch = executor.execute_batch(compiled, partition, built);
    profiler.end(qseg);
    const auto mseg = profiler.begin_nestedf("partition %zu materialize_results", pid);
    query::MaterializedBatchResult part = std::move(batch).materialize();
    profiler.end(mseg);
    if (part.queries().size() != merged.size())
      throw error::common::ImplementationError(
          "Partition query results do not match the compiled query batch.");
    for (size_t q = 0; q < merged.size(); ++q) merged[q].append(part.take_query(q));
    profiler.end(pseg);
  }

  query::MaterializedBatchResult result;
  for (auto &q : merged) result.add_query_result(std::move(q));
  profiler.end(total);
  return result;
}

} // namespace gpjson
