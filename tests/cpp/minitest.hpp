// Minimal test registry for the C++ API tests (no third-party framework is
// available offline).  TEST(group, name) registers a CPU test, GPU_TEST a test
// that needs a HIP device.  Run: gpjson_cpp_tests [--cpu|--gpu] [filter].
#pragma once

#include <cstdio>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace minitest {

struct Case {
  std::string name;
  bool gpu;
  std::function<void()> fn;
};

inline std::vector<Case> &registry() {
  static std::vector<Case> r;
  return r;
}

struct Registrar {
  Registrar(const char *name, bool gpu, std::function<void()> fn) {
    registry().push_back({name, gpu, std::move(fn)});
  }
};

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class A, class B>
void expect_eq(const A &a, const B &b, const char *ea, const char *eb, const char *file, int line) {
  if (!(a == b)) {
    std::ostringstream s;
    s << file << ":" << line << ": expected " << ea << " == " << eb;
    throw Failure(s.str());
  }
}

} // namespace minitest

#define MT_CAT2(a, b) a##b
#define MT_CAT(a, b) MT_CAT2(a, b)
#define MT_DEFINE(group, name, gpu)                                                         \
  static void group##_##name();                                                             \
  static minitest::Registrar MT_CAT(reg_, MT_CAT(group##_##name, __LINE__))(#group "." #name, \
                                                                            gpu,            \
                                                                            group##_##name); \
  static void group##_##name()
#define TEST(group, name) MT_DEFINE(group, name, false)
#define GPU_TEST(group, name) MT_DEFINE(group, name, true)

#define EXPECT_EQ(a, b) minitest::expect_eq((a), (b), #a, #b, __FILE__, __LINE__)
#define EXPECT_TRUE(a) minitest::expect_eq(static_cast<bool>(a), true, #a, "true", __FILE__, __LINE__)
#define EXPECT_FALSE(a) minitest::expect_eq(static_cast<bool>(a), false, #a, "false", __FILE__, __LINE__)
#define EXPECT_THROW(stmt, exc)                                                             \
  do {                                                                                      \
    bool caught_ = false;                                                                   \
    try {                                                                                   \
      stmt;                                                                                 \
    } catch (const exc &) {                                                                 \
      caught_ = true;                                                                       \
    }                                                                                       \
    if (!caught_)                                                                           \
      throw minitest::Failure(std::string(__FILE__) + ":" + std::to_string(__LINE__) +      \
                              ": expected " #exc " from " #stmt);                           \
  } while (0)
