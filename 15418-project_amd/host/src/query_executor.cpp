// QueryExecutor::execute_batch over gpj_query_run
// (reference: src/gpjson/query/kernels/orig/query_executor_host.cu:80-241).
#include "gpjson/query/query_executor.hpp"

#include "gpj_capi.h"
#include "gpjson/engine.hpp"
#include "gpjson/query/error.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

namespace gpjson::query {

namespace {

void validate(const CompiledQuery &q, const index::BuiltIndices &idx) {
  if (q.max_depth() > static_cast<size_t>(idx.get_leveled_bitmap_index().num_levels()))
    throw error::query::QueryExecutionError(
        "Query requires more bitmap levels than the index provides");
  if (q.max_depth() >= static_cast<size_t>(kMaxQueryLevels))
    throw error::query::QueryExecutionError("Query exceeds the maximum executor depth");
  if (q.ir_bytes().size() > kMaxQueryIrBytes)
    throw error::query::QueryExecutionError("Compiled query IR exceeds the 4096-byte IR buffer");
}

} // namespace

QueryExecutor::QueryExecutor(const gpjson::EngineOptions &options)
    : grid_size_(options.query_executor_options.grid_size),
      block_size_(options.query_executor_options.block_size) {}

BatchQueryResult QueryExecutor::execute_batch(const BatchCompiledQuery &compiled,
                                              const file::FilePartition &partition,
                                              const index::BuiltIndices &idx) const {
  BatchQueryResult batch(compiled.size());
  for (size_t q = 0; q < compiled.size(); ++q)
    batch.set_query_text(q, compiled.queries()[q].query_text());
  if (compiled.size() == 0 || partition.size_bytes() == 0 || partition.host_bytes() == nullptr)
    return batch;
  if (partition.device_bytes() == nullptr)
    throw error::query::QueryExecutionError(
        "Query execution requires a GPU-resident partition buffer");
  if (!device::device_available())
    throw error::query::QueryExecutionError("HIP device unavailable for query execution");
  const int64_t num_lines = idx.get_newline_index().num_lines();
  if (num_lines <= 0) return batch;
  if (!idx.get_newline_index().data() || !idx.get_string_index().data() ||
      !idx.get_leveled_bitmap_index().data())
    throw error::query::QueryExecutionError(
        "Query execution requires GPU-resident newline, string, and bitmap indices");

  gpj_index view{};
  view.d_newline = const_cast<int64_t *>(static_cast<const int64_t *>(idx.get_newline_index().data()));
  view.d_num_lines = const_cast<int64_t *>(idx.get_newline_index().device_num_lines());
  view.num_lines = num_lines;
  view.d_string = const_cast<uint64_t *>(static_cast<const uint64_t *>(idx.get_string_index().data()));
  view.d_leveled = const_cast<uint64_t *>(
      static_cast<const uint64_t *>(idx.get_leveled_bitmap_index().data()));
  view.level_size = static_cast<int64_t>((partition.size_bytes() + 63) / 64);
  view.num_levels = idx.get_leveled_bitmap_index().num_levels();

  const char *host = static_cast<const char *>(partition.host_bytes());
  const size_t base = partition.global_start_offset();
  // Queries run GPJ_MAX_BATCH_QUERIES at a time in one launch each (the
  // record is read once per group); results are the per-query results.
  for (size_t first = 0; first < compiled.size(); first += GPJ_MAX_BATCH_QUERIES) {
    const size_t last = std::min(compiled.size(), first + GPJ_MAX_BATCH_QUERIES);
    std::vector<device::DeviceArray> irs, spans;
    std::vector<gpj_query_desc> descs;
    uint32_t ir_total = 0;
    for (size_t q = first; q < last; ++q) {
      const CompiledQuery &cq = compiled.queries()[q];
      validate(cq, idx);
      ir_total += static_cast<uint32_t>(cq.ir_bytes().size());
      irs.emplace_back(cq.ir_bytes().size());
      irs.back().copy_from_host(cq.ir_bytes().data(), cq.ir_bytes().size());
      const int slots = static_cast<int>(cq.num_result_slots());
      spans.emplace_back(static_cast<size_t>(num_lines) * slots * 2 * sizeof(int64_t));
      descs.push_back(gpj_query_desc{irs.back().as<uint8_t>(),
                                     static_cast<uint32_t>(cq.ir_bytes().size()), slots, 0,
                                     static_cast<int32_t>(cq.max_depth()),
                                     spans.back().as<int64_t>(), nullptr});
    }
    if (ir_total > 8192) {
      // oversized bytecode: one launch per query
      for (size_t i = 0; i < descs.size(); ++i)
        device::check(gpj_query_run_batch(static_cast<const uint8_t *>(partition.device_bytes()),
                                          partition.size_bytes(), &view, &descs[i], 1, num_lines,
                                          nullptr),
                      "execute_query_kernel launch");
    } else {
      device::check(gpj_query_run_batch(static_cast<const uint8_t *>(partition.device_bytes()),
                                        partition.size_bytes(), &view, descs.data(),
                                        static_cast<int32_t>(descs.size()), num_lines, nullptr),
                    "execute_query_kernel launch");
    }
    for (size_t q = first; q < last; ++q) {
      const int slots = descs[q - first].capacity;
      const size_t lines_n = static_cast<size_t>(num_lines);
      const size_t count = lines_n * slots * 2;
      CompactLines out;
      out.base = base;
      out.lines = lines_n;
      out.slots = slots;
      out.spans.assign(count, -1);
      spans[q - first].copy_to_host(out.spans.data(), count * sizeof(int64_t));
      // the values' texts, in span order, in one buffer (no string per value:
      // those are made only when a line's results are read or materialized);
      // ranges of lines on up to 16 threads, each at its prefix of the bytes
      const size_t hw = std::max(1u, std::thread::hardware_concurrency());
      const size_t workers = std::min({hw, size_t{16}, std::max<size_t>(1, (lines_n + 65535) / 65536)});
      const size_t per = (lines_n + workers - 1) / std::max<size_t>(workers, 1);
      std::vector<size_t> bytes(workers + 1, 0);
      std::atomic<bool> bad_range{false};
      auto measure = [&](size_t t) {
        size_t sum = 0;
        for (size_t i = t * per * slots; i < std::min(lines_n, (t + 1) * per) * slots; ++i) {
          const int64_t a = out.spans[2 * i], b = out.spans[2 * i + 1];
          if (a < 0 || b < 0) continue;
          if (a > b || static_cast<size_t>(b) > partition.size_bytes()) bad_range = true;
          else sum += static_cast<size_t>(b - a);
        }
        bytes[t + 1] = sum;
      };
      auto copy = [&](size_t t) {
        char *dst = out.texts.data() + bytes[t];
        for (size_t i = t * per * slots; i < std::min(lines_n, (t + 1) * per) * slots; ++i) {
          const int64_t a = out.spans[2 * i], b = out.spans[2 * i + 1];
          if (a < 0 || b < 0) continue;
          std::memcpy(dst, host + a, static_cast<size_t>(b - a));
          dst += b - a;
        }
      };
      auto run = [&](auto &&fn) {
        if (workers <= 1) {
          fn(0);
          return;
        }
        std::vector<std::thread> pool;
        try {
          for (size_t t = 0; t < workers; ++t) pool.emplace_back([&, t] { fn(t); });
        } catch (...) { // a thread could not start: join the others, then fail
          for (auto &th : pool) th.join();
          throw;
        }
        for (auto &th : pool) th.join();
      };
      run(measure);
      if (bad_range)
        throw error::query::QueryExecutionError("Query kernel produced an invalid result range");
      for (size_t t = 0; t < workers; ++t) bytes[t + 1] += bytes[t];
      out.texts.resize(bytes[workers]);
      run(copy);
      batch.add_compact(q, std::move(out));
    }
  }
  return batch;
}

} // namespace gpjson::query
