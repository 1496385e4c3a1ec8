// C entry point over the reference's query compiler, compiled together with
// the reference's own src/gpjson/query/{query_compiler,query,ir_input_buffer}.cpp
// by oracle/Makefile into oracle/_ref/.  Test infrastructure: used only to
// write and check the golden IR vectors under tests/golden/.
//
// EngineOptions is only forward-declared by the compiler and never read; the
// reference's compiler test defines it empty the same way
// (tests/query/query_compiler_test.cpp:20-22).
#include "gpjson/query/error.hpp"
#include "gpjson/query/query_compiler.hpp"

#include <cstring>
#include <string>

namespace gpjson {
struct EngineOptions {};
} // namespace gpjson

extern "C" {
// Returns the IR length (>= 0) or -1 malformed, -2 unsupported, -3 other
// error, -4 buffer too small.  max_depth / slots are written on success.
long gpj_ref_compile(const char *query, unsigned char *ir, long cap,
                     long *max_depth, long *slots, char *err, long err_cap) {
  gpjson::EngineOptions options{};
  gpjson::query::QueryCompiler compiler(options);
  auto put_err = [&](const char *m) {
    if (err && err_cap > 0) {
      std::strncpy(err, m, static_cast<size_t>(err_cap - 1));
      err[err_cap - 1] = 0;
    }
  };
  try {
    auto q = compiler.compile(std::string(query), options);
    const auto &bytes = q.ir_bytes();
    if (static_cast<long>(bytes.size()) > cap) return -4;
    std::memcpy(ir, bytes.data(), bytes.size());
    *max_depth = static_cast<long>(q.max_depth());
    *slots = static_cast<long>(q.num_result_slots());
    return static_cast<long>(bytes.size());
  } catch (const gpjson::error::query::MalformedQueryError &e) {
    put_err(e.what());
    return -1;
  } catch (const gpjson::error::query::UnsupportedQueryError &e) {
    put_err(e.what());
    return -2;
  } catch (const std::exception &e) {
    put_err(e.what());
    return -3;
  }
}
}
