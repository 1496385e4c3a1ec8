// JSONPath -> bytecode compiler (reference: include/gpjson/query/query_compiler.hpp:110-116).
// Errors: MalformedQueryError, UnsupportedQueryError, and std::invalid_argument /
// std::out_of_range for integer literals, as the reference's std::stoull does.
#pragma once

#include "gpjson/query/error.hpp"
#include "gpjson/query/query.hpp"

#include <string>

namespace gpjson {
struct EngineOptions;
}

namespace gpjson::query {

class QueryCompiler {
public:
  explicit QueryCompiler(const gpjson::EngineOptions &options);
  CompiledQuery compile(const std::string &query_src, const gpjson::EngineOptions &options) const;
};

} // namespace gpjson::query
