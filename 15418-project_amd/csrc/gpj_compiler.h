// gpj_compiler.h -- internal interface of the JSONPath bytecode compiler.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace gpj {

enum class CompileStatus {
  kOk = 0,
  kMalformed,       // MalformedQueryError
  kUnsupported,     // UnsupportedQueryError
  kInvalidArgument, // std::invalid_argument from std::stoull
  kOutOfRange,      // std::out_of_range from std::stoull
};

struct CompiledBytecode {
  std::vector<uint8_t> bytes;
  size_t max_depth = 0;
  size_t num_slots = 0;
};

CompileStatus compile_query(const std::string &query, bool extensions,
                            CompiledBytecode *out, std::string *error);

} // namespace gpj
