// Per-run context handed to IndexBuilder::build (reference: include/gpjson/run_context.hpp).
#pragma once

#include "gpjson/profiler/profiler.hpp"

namespace gpjson {

struct RunContext {
  profiler::Profiler &profiler;
};

} // namespace gpjson
