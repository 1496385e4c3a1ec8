// QueryCompiler over gpj_query_compile (include/gpj_capi.h).
#include "gpjson/query/query_compiler.hpp"

#include "gpj_capi.h"

#include <stdexcept>
#include <vector>

namespace gpjson::query {

QueryCompiler::QueryCompiler(const gpjson::EngineOptions &options) { (void)options; }

CompiledQuery QueryCompiler::compile(const std::string &query_src,
                                     const gpjson::EngineOptions &options) const {
  (void)options;
  std::vector<uint8_t> buf(1 << 16);
  uint32_t len = 0, depth = 0, slots = 0;
  const gpj_status s = gpj_query_compile(query_src.c_str(), buf.data(),
                                         static_cast<uint32_t>(buf.size()), &len, &depth, &slots, 0);
  switch (s) {
  case GPJ_OK: break;
  case GPJ_ERR_MALFORMED_QUERY: throw error::query::MalformedQueryError(gpj_last_error());
  case GPJ_ERR_UNSUPPORTED_QUERY: throw error::query::UnsupportedQueryError(gpj_last_error());
  case GPJ_ERR_OUT_OF_RANGE: throw std::out_of_range(gpj_last_error());
  case GPJ_ERR_INVALID_ARGUMENT: throw std::invalid_argument(gpj_last_error());
  default: throw error::query::MalformedQueryError(gpj_last_error());
  }
  std::vector<std::byte> ir(len);
  for (uint32_t i = 0; i < len; ++i) ir[i] = static_cast<std::byte>(buf[i]);
  return CompiledQuery(query_src, std::move(ir), depth, slots);
}

} // namespace gpjson::query
