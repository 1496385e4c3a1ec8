// Host wall-clock segments, printed as "<name>: <indent><segment>: <ms> ms"
// when a segment ends (reference: include/gpjson/profiler/profiler.hpp).
#pragma once

#include <chrono>
#include <cstdarg>
#include <cstddef>
#include <string>
#include <string_view>
#include <vector>

namespace gpjson::profiler {

class Profiler {
public:
  explicit Profiler(std::string name = "Profiler", bool print = true);

  class SegmentId {
  public:
    SegmentId() = default;

  private:
    explicit SegmentId(size_t index) : index_(index), valid_(true) {}
    size_t index_{0};
    bool valid_{false};
    friend class Profiler;
  };

  SegmentId begin(std::string_view name);
  SegmentId beginf(const char *fmt, ...);
  SegmentId begin_nested(std::string_view name);
  SegmentId begin_nestedf(const char *fmt, ...);
  void end(SegmentId segment_id);

  /* Completed segments in start order: (name, indent, milliseconds). */
  struct Segment {
    std::string name;
    size_t indent;
    double ms;
  };
  std::vector<Segment> completed() const;

private:
  using Clock = std::chrono::steady_clock;
  struct Record {
    std::string name;
    Clock::time_point start{};
    double ms{0};
    size_t indent{0};
    bool nested{false};
    bool done{false};
  };
  SegmentId open(std::string name, bool nested);

  std::string name_;
  bool print_;
  std::vector<Record> records_;
  size_t indent_{0};
};

} // namespace gpjson::profiler
