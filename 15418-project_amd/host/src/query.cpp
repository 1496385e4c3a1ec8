// Result containers and the bytecode reader
// (reference: src/gpjson/query/query.cpp, src/gpjson/query/ir_input_buffer.cpp).
#include "gpjson/query/ir_input_buffer.hpp"
#include "gpjson/query/query.hpp"

#include <algorithm>

namespace gpjson::query {

void BatchCompiledQuery::add(CompiledQuery query) {
  max_depth_ = std::max(max_depth_, query.max_depth());
  queries_.push_back(std::move(query));
}

void BatchQueryResult::add_line_result(size_t query_index, LineQueryResult line_result) {
  if (query_index >= queries_.size()) return;
  realize(); // keeps the order: compact results came first
  queries_[query_index].add_line_result(std::move(line_result));
}

void BatchQueryResult::add_compact(size_t query_index, CompactLines lines) {
  if (query_index < pending_.size()) pending_[query_index].push_back(std::move(lines));
}

namespace {

// Calls fn(line, slot, start, end, text, length) for every value of a compact
// chunk in order (global offsets; text points into the chunk's buffer).
template <typename Fn> void for_each_value(const CompactLines &c, Fn fn) {
  size_t t = 0;
  for (size_t line = 0; line < c.lines; ++line)
    for (int s = 0; s < c.slots; ++s) {
      const int64_t a = c.spans[(line * c.slots + s) * 2];
      const int64_t b = c.spans[(line * c.slots + s) * 2 + 1];
      if (a < 0 || b < 0) {
        fn(line, s, -1, -1, nullptr, 0);
        continue;
      }
      const size_t len = static_cast<size_t>(b - a);
      fn(line, s, static_cast<int64_t>(c.base) + a, static_cast<int64_t>(c.base) + b,
         c.texts.data() + t, len);
      t += len;
    }
}

// The per-line form of a compact chunk.
void append_lines(const CompactLines &c, QueryResult &out) {
  std::vector<LineQueryResult> lines(c.lines);
  for_each_value(c, [&](size_t line, int, int64_t a, int64_t b, const char *text, size_t len) {
    QueryOffset off;
    if (a >= 0) {
      off.start = static_cast<size_t>(a);
      off.end = static_cast<size_t>(b);
      off.json_text = std::string(text, len);
    }
    lines[line].add_offset(std::move(off));
  });
  for (auto &l : lines) out.add_line_result(std::move(l));
}

// The materialized form of a compact chunk: values of lines holding any.
void materialize_lines(const CompactLines &c, MaterializedQueryResult &out) {
  MaterializedLineResult m;
  size_t current = 0;
  for_each_value(c, [&](size_t line, int, int64_t a, int64_t, const char *text, size_t len) {
    if (line != current) {
      if (!m.values().empty()) out.add_line_result(std::move(m));
      m = MaterializedLineResult();
      current = line;
    }
    if (a >= 0) m.add_value(MaterializedValue(std::string(text, len)));
  });
  if (!m.values().empty()) out.add_line_result(std::move(m));
}

} // namespace

void BatchQueryResult::realize() const {
  for (size_t q = 0; q < pending_.size(); ++q) {
    for (const auto &c : pending_[q]) append_lines(c, queries_[q]);
    pending_[q].clear();
  }
}

const std::vector<QueryResult> &BatchQueryResult::queries() const {
  realize();
  return queries_;
}

void BatchQueryResult::merge(const BatchQueryResult &other) {
  if (queries_.size() != other.queries_.size()) return;
  bool other_realized = false;
  for (const auto &q : other.queries_) other_realized |= !q.lines().empty();
  if (other_realized) realize(); // its per-line results follow this one's
  for (size_t q = 0; q < queries_.size(); ++q) {
    if (query_texts_[q].empty()) query_texts_[q] = other.query_texts_[q];
    for (const auto &line : other.queries_[q].lines()) queries_[q].add_line_result(line);
    for (const auto &c : other.pending_[q]) pending_[q].push_back(c);
  }
}

void BatchQueryResult::set_query_text(size_t query_index, std::string query_text) {
  if (query_index < query_texts_.size()) query_texts_[query_index] = std::move(query_text);
}

MaterializedBatchResult BatchQueryResult::materialize() && {
  MaterializedBatchResult batch;
  for (size_t q = 0; q < queries_.size(); ++q) {
    MaterializedQueryResult result(std::move(query_texts_[q]));
    for (auto &line : queries_[q].lines_) {
      MaterializedLineResult m;
      for (auto &off : line.offsets_)
        if (off.json_text.has_value()) m.add_value(MaterializedValue(std::move(*off.json_text)));
      if (!m.values().empty()) result.add_line_result(std::move(m));
    }
    for (const auto &c : pending_[q]) materialize_lines(c, result);
    batch.add_query_result(std::move(result));
  }
  queries_.clear();
  pending_.clear();
  query_texts_.clear();
  return batch;
}

MaterializedBatchResult BatchQueryResult::materialize() const & {
  MaterializedBatchResult batch;
  for (size_t q = 0; q < queries_.size(); ++q) {
    MaterializedQueryResult result(query_texts_[q]);
    for (const auto &line : queries_[q].lines()) {
      MaterializedLineResult m;
      for (const auto &off : line.offsets())
        if (off.json_text.has_value()) m.add_value(MaterializedValue(*off.json_text));
      if (!m.values().empty()) result.add_line_result(std::move(m));
    }
    for (const auto &c : pending_[q]) materialize_lines(c, result);
    batch.add_query_result(std::move(result));
  }
  return batch;
}

std::byte IRByteInputBuffer::read_byte() {
  if (position_ >= size_) throw error::query::QueryExecutionError("Unexpected end of query IR");
  return bytes_[position_++];
}

size_t IRByteInputBuffer::read_varint() {
  size_t value = 0, shift = 0;
  for (;;) {
    const auto b = static_cast<unsigned char>(read_byte());
    value |= static_cast<size_t>(b & 0x7F) << shift;
    if (!(b & 0x80)) return value;
    shift += 7;
  }
}

std::string IRByteInputBuffer::read_string() {
  const size_t n = read_varint();
  std::string s;
  s.reserve(n);
  for (size_t i = 0; i < n; ++i) s.push_back(static_cast<char>(read_byte()));
  return s;
}

} // namespace gpjson::query
