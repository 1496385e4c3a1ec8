// Host wall-clock profiler (same output format as the reference's
// src/gpjson/profiler/profiler.cpp:103-113).
#include "gpjson/profiler/profiler.hpp"

#include <cstdio>
#include <vector>

namespace gpjson::profiler {

Profiler::Profiler(std::string name, bool print) : name_(std::move(name)), print_(print) {}

Profiler::SegmentId Profiler::open(std::string name, bool nested) {
  Record r;
  r.name = std::move(name);
  r.start = Clock::now();
  r.indent = indent_;
  r.nested = nested;
  records_.push_back(std::move(r));
  if (nested) ++indent_;
  return SegmentId(records_.size() - 1);
}

Profiler::SegmentId Profiler::begin(std::string_view name) { return open(std::string(name), false); }
Profiler::SegmentId Profiler::begin_nested(std::string_view name) {
  return open(std::string(name), true);
}

static std::string vfmt(const char *fmt, va_list ap) {
  va_list cp;
  va_copy(cp, ap);
  const int n = std::vsnprintf(nullptr, 0, fmt, cp);
  va_end(cp);
  if (n < 0) return {};
  std::vector<char> buf(static_cast<size_t>(n) + 1);
  std::vsnprintf(buf.data(), buf.size(), fmt, ap);
  return std::string(buf.data(), static_cast<size_t>(n));
}

Profiler::SegmentId Profiler::beginf(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::string s = vfmt(fmt, ap);
  va_end(ap);
  return open(std::move(s), false);
}

Profiler::SegmentId Profiler::begin_nestedf(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::string s = vfmt(fmt, ap);
  va_end(ap);
  return open(std::move(s), true);
}

void Profiler::end(SegmentId id) {
  if (!id.valid_ || id.index_ >= records_.size()) return;
  Record &r = records_[id.index_];
  if (r.done) return;
  if (r.nested && indent_ > 0) --indent_;
  r.ms = std::chrono::duration<double, std::milli>(Clock::now() - r.start).count();
  r.done = true;
  if (print_) {
    std::fprintf(stdout, "%s: %*s%s: %.3f ms\n", name_.c_str(), static_cast<int>(r.indent * 2), "",
                 r.name.c_str(), r.ms);
    std::fflush(stdout);
  }
}

std::vector<Profiler::Segment> Profiler::completed() const {
  std::vector<Segment> out;
  for (const auto &r : records_)
    if (r.done) out.push_back({r.name, r.indent, r.ms});
  return out;
}

} // namespace gpjson::profiler
