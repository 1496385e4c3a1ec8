// Compiled queries and result containers
// (reference: include/gpjson/query/query.hpp:10-135).
#pragma once

#include <cstddef>
#include <optional>
#include <string>
#include <vector>

namespace gpjson::query {

enum class QueryOpcode : unsigned char {
  END = 0,
  STORE_RESULT = 1,
  MOVE_UP = 2,
  MOVE_DOWN = 3,
  MOVE_TO_KEY = 4,
  MOVE_TO_INDEX = 5,
  MOVE_TO_INDEX_REVERSE = 6,
  MARK_POSITION = 7,
  RESET_POSITION = 8,
  EXPRESSION_STRING_EQUALS = 9,
};

class CompiledQuery {
public:
  CompiledQuery() = default;
  CompiledQuery(std::string query_text, std::vector<std::byte> ir_bytes, size_t max_depth,
                size_t num_result_slots)
      : query_text_(std::move(query_text)), ir_bytes_(std::move(ir_bytes)),
        max_depth_(max_depth), num_result_slots_(num_result_slots) {}
  const std::string &query_text() const { return query_text_; }
  const std::vector<std::byte> &ir_bytes() const { return ir_bytes_; }
  size_t max_depth() const { return max_depth_; }
  size_t num_result_slots() const { return num_result_slots_; }

private:
  std::string query_text_;
  std::vector<std::byte> ir_bytes_;
  size_t max_depth_{1};
  size_t num_result_slots_{1};
};

class BatchCompiledQuery {
public:
  void add(CompiledQuery query);
  size_t size() const { return queries_.size(); }
  size_t get_max_depth() const { return max_depth_; }
  const std::vector<CompiledQuery> &queries() const { return queries_; }

private:
  std::vector<CompiledQuery> queries_;
  size_t max_depth_{1};
};

struct QueryOffset {
  std::optional<size_t> start;
  std::optional<size_t> end;
  std::optional<std::string> json_text;
};

class LineQueryResult {
public:
  const std::vector<QueryOffset> &offsets() const { return offsets_; }
  void add_offset(QueryOffset offset) { offsets_.push_back(std::move(offset)); }

private:
  friend class BatchQueryResult; // materialize() && moves the texts out
  std::vector<QueryOffset> offsets_;
};

class QueryResult {
public:
  void add_line_result(LineQueryResult line_result) { lines_.push_back(std::move(line_result)); }
  const std::vector<LineQueryResult> &lines() const { return lines_; }

private:
  friend class BatchQueryResult;
  std::vector<LineQueryResult> lines_;
};

class MaterializedBatchResult;

/* Results of one query over the lines of one partition in compact form (what
 * the executor produces): spans [lines][slots][2], partition-local {start,
 * end} (-1 = no value), global offsets = base + local, and the texts of the
 * values in one buffer, in span order.  No string per value is made until
 * the results are read per line (BatchQueryResult::queries) or materialized,
 * and then only for values a line holds (extension; the reference builds a
 * string per match in the executor, query_executor_host.cu:111-114). */
struct CompactLines {
  size_t base = 0;
  size_t lines = 0;
  int slots = 1;
  std::vector<int64_t> spans;
  std::string texts;
};

class BatchQueryResult {
public:
  explicit BatchQueryResult(size_t num_queries)
      : queries_(num_queries), pending_(num_queries), query_texts_(num_queries) {}
  void add_line_result(size_t query_index, LineQueryResult line_result);
  /* Appends a partition's results of one query in compact form. */
  void add_compact(size_t query_index, CompactLines lines);
  void merge(const BatchQueryResult &other);
  void set_query_text(size_t query_index, std::string query_text);
  /* Drops unmatched slots and lines without any value (query.cpp:87-109).
   * The rvalue form moves the value texts instead of copying them. */
  MaterializedBatchResult materialize() const &;
  MaterializedBatchResult materialize() &&;
  size_t num_queries() const { return queries_.size(); }
  /* Per-line results (QueryOffset with start, end and json_text); compact
   * results are turned into this form on the first call. */
  const std::vector<QueryResult> &queries() const;

private:
  void realize() const;
  mutable std::vector<QueryResult> queries_;
  mutable std::vector<std::vector<CompactLines>> pending_; // after queries_, in order
  std::vector<std::string> query_texts_;
};

class MaterializedValue {
public:
  explicit MaterializedValue(std::string json_text) : json_text_(std::move(json_text)) {}
  const std::string &json_text() const { return json_text_; }

private:
  std::string json_text_;
};

class MaterializedLineResult {
public:
  const std::vector<MaterializedValue> &values() const { return values_; }
  void add_value(MaterializedValue value) { values_.push_back(std::move(value)); }

private:
  std::vector<MaterializedValue> values_;
};

class MaterializedQueryResult {
public:
  explicit MaterializedQueryResult(std::string query_text) : query_text_(std::move(query_text)) {}
  const std::string &query_text() const { return query_text_; }
  const std::vector<MaterializedLineResult> &lines() const { return lines_; }
  void add_line_result(MaterializedLineResult line) { lines_.push_back(std::move(line)); }
  /* Moves the lines of `other` to the end of this result. */
  void append(MaterializedQueryResult &&other) {
    if (lines_.empty()) {
      lines_ = std::move(other.lines_);
    } else {
      lines_.reserve(lines_.size() + other.lines_.size());
      for (auto &l : other.lines_) lines_.push_back(std::move(l));
    }
    other.lines_.clear();
  }

private:
  std::string query_text_;
  std::vector<MaterializedLineResult> lines_;
};

class MaterializedBatchResult {
public:
  const std::vector<MaterializedQueryResult> &queries() const { return queries_; }
  void add_query_result(MaterializedQueryResult result) { queries_.push_back(std::move(result)); }
  /* The results of query q, to move them out (e.g. MaterializedQueryResult::append). */
  MaterializedQueryResult &&take_query(size_t q) { return std::move(queries_.at(q)); }

private:
  std::vector<MaterializedQueryResult> queries_;
};

class QueryOptions {};

} // namespace gpjson::query
