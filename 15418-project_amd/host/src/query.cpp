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
  if (query_index < queries_.size()) queries_[query_index].add_line_result(std::move(line_result));
}

void BatchQueryResult::merge(const BatchQueryResult &other) {
  if (queries_.size() != other.queries_.size()) return;
  for (size_t q = 0; q < queries_.size(); ++q) {
    if (query_texts_[q].empty()) query_texts_[q] = other.query_texts_[q];
    for (const auto &line : other.queries_[q].lines()) queries_[q].add_line_result(line);
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
    batch.add_query_result(std::move(result));
  }
  queries_.clear();
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
