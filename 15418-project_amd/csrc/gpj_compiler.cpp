// gpj_compiler.cpp -- JSONPath -> bytecode compiler behind gpj_query_compile.
//
// Accepts the reference's JSONPath subset and emits byte-identical bytecode
// (grammar and emission order of src/gpjson/query/query_compiler.cpp:170-465;
// opcodes of include/gpjson/query/query.hpp:10-21; LEB128 varints and
// length-prefixed strings of query_compiler.cpp:14-33).  Supported: `.key`,
// `['key']`, `[n]`, `[a:b]`, `[:b]`, `[-n]`, `[-n:]`, `[i,j,...]`,
// `[?(@... == "s")]`.  Rejected as unsupported: `..`, `[*]`, `[--n]`, other
// filter operators.  Error texts follow the reference's "<message> at <pos>
// ('<char>')" form.
#include <cctype>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpj_compiler.h"

namespace gpj {

namespace {

enum Op : uint8_t {
  kEnd = 0, kStoreResult = 1, kMoveUp = 2, kMoveDown = 3, kMoveToKey = 4,
  kMoveToIndex = 5, kMoveToIndexReverse = 6, kMarkPosition = 7,
  kResetPosition = 8, kExpressionStringEquals = 9
};

struct CompileFailure {
  CompileStatus status;
  std::string message;
};

class Emitter {
public:
  void op(Op o) { bytes.push_back(o); }
  void varint(uint64_t v) {
    while (v & ~uint64_t{0x7F}) {
      bytes.push_back(static_cast<uint8_t>((v & 0x7F) | 0x80));
      v >>= 7;
    }
    bytes.push_back(static_cast<uint8_t>(v));
  }
  void str(const std::string &s) {
    varint(s.size());
    bytes.insert(bytes.end(), s.begin(), s.end());
  }
  // Level bookkeeping of the bytecode (never below zero).
  void down() {
    op(kMoveDown);
    ++level;
  }
  void up() {
    op(kMoveUp);
    if (level > 0) --level;
  }
  void store() {
    op(kStoreResult);
    ++stores;
  }
  std::vector<uint8_t> bytes;
  size_t level = 0;
  size_t stores = 0;
};

class Parser {
public:
  Parser(const std::string &q, bool extensions) : q_(q), ext_(extensions) {}

  CompiledBytecode run() {
    expect('$');
    expression();
    out_.store();
    out_.op(kEnd);
    return {out_.bytes, depth_, out_.stores};
  }

private:
  // Cursor sits on the last consumed character; -1 before the first.
  bool more() const { return cur_ + 1 < static_cast<long>(q_.size()); }
  char look() const {
    if (!more()) fail_malformed("Expected character, got EOF");
    return q_[cur_ + 1];
  }
  char take() {
    if (!more()) fail_malformed("Expected character, got EOF");
    return q_[++cur_];
  }
  void expect(char c) {
    const char got = take();
    if (got != c)
      fail_malformed(std::string("Expected character '") + c + "', got '" + got + "'");
  }
  std::string text_between(long from, long to) const {
    if (to < from) return {};
    return q_.substr(static_cast<size_t>(from + 1), static_cast<size_t>(to - from));
  }
  std::string where(const std::string &m) const {
    std::string s = m + " at " + std::to_string(cur_ + 1);
    if (more()) s += std::string(" ('") + q_[cur_ + 1] + "')";
    return s;
  }
  [[noreturn]] void fail_malformed(const std::string &m) const {
    throw CompileFailure{CompileStatus::kMalformed, where(m)};
  }
  [[noreturn]] void fail_unsupported(const std::string &m) const {
    throw CompileFailure{CompileStatus::kUnsupported, where(m)};
  }

  void expression() {
    if (!more()) return;
    switch (look()) {
    case '.': dot(); break;
    case '[': bracket(); break;
    case ' ': return;
    default: fail_unsupported("Unsupported expression type");
    }
  }

  void dot() {
    expect('.');
    if (look() == '.') fail_unsupported("Unsupported recursive descent");
    const long from = cur_;
    while (more()) {
      const char c = look();
      if (c == '.' || c == '[' || c == ' ') break;
      take();
    }
    const std::string name = text_between(from, cur_);
    if (name.empty()) fail_malformed("Unexpected empty property");
    key(name);
    if (more()) expression();
  }

  void bracket() {
    expect('[');
    const char c = look();
    if (c == '\'' || c == '"') {
      const std::string name = quoted();
      if (name.empty()) fail_malformed("Unexpected empty property");
      key(name);
    } else if (std::isdigit(static_cast<unsigned char>(c))) {
      const uint64_t first = integer(":],");
      switch (look()) {
      case ':': {
        expect(':');
        const uint64_t last = integer("]");
        expect(']');
        range(first, last, false);
        return;
      }
      case ',': {
        std::vector<uint64_t> list{first};
        while (look() == ',') {
          expect(',');
          if (!std::isdigit(static_cast<unsigned char>(look())))
            fail_malformed(std::string("Expected digit, got '") + look() + "'");
          list.push_back(integer(",]"));
        }
        expect(']');
        branches(list);
        return;
      }
      case ']':
        index(first, false);
        break;
      default:
        fail_malformed("Unexpected character in index expression");
      }
    } else if (c == ':') {
      expect(':');
      if (!std::isdigit(static_cast<unsigned char>(look())))
        fail_malformed("Unexpected character in index, expected an integer");
      const uint64_t last = integer("]");
      expect(']');
      range(0, last, false);
      return;
    } else if (c == '-') {
      expect('-');
      if (std::isdigit(static_cast<unsigned char>(look()))) {
        const uint64_t k = integer("]:");
        if (k == 0) fail_malformed("Invalid reverse index 0");
        if (look() == ']') {
          index(k, true);
        } else {
          expect(':');
          expect(']');
          range(0, k, true);
          return;
        }
      } else if (look() == '-') {
        fail_unsupported("Unsupported last n elements of the array query");
      } else {
        fail_malformed("Unexpected character in index, expected an integer");
      }
    } else if (c == '*') {
      if (!ext_) fail_unsupported("Unsupported wildcard expression");
      fail_unsupported("Unsupported wildcard expression");
    } else if (c == '?') {
      filter();
    } else {
      fail_malformed("Unexpected character in index, expected a quoted string or integer");
    }
    expect(']');
    if (more()) expression();
  }

  // One array branch per listed index; each branch compiles the rest of the
  // query, stores, and climbs back (query_compiler.cpp:302-364).
  void branches(const std::vector<uint64_t> &list) {
    size_t deepest = depth_;
    for (size_t i = 0; i < list.size(); ++i)
      deepest = branch(list[i], i + 1 == list.size(), deepest, false);
    depth_ = deepest + 1;
  }
  void range(uint64_t from, uint64_t to, bool reverse) {
    size_t deepest = depth_;
    if (reverse) {
      for (uint64_t i = to; i > from; --i) deepest = branch(i, i == from + 1, deepest, true);
    } else {
      for (uint64_t i = from; i < to; ++i) deepest = branch(i, i + 1 == to, deepest, false);
    }
    depth_ = deepest + 1;
  }
  size_t branch(uint64_t i, bool last, size_t deepest, bool reverse) {
    const size_t base_level = out_.level;
    out_.op(reverse ? kMoveToIndexReverse : kMoveToIndex);
    out_.varint(i);
    out_.down();
    marks_.push_back(cur_);
    const size_t saved = depth_;
    if (more()) expression();
    if (depth_ > deepest) deepest = depth_;
    depth_ = saved;
    if (!last) {
      out_.store();
      cur_ = marks_.back();
      marks_.pop_back();
      const size_t top = out_.level;
      for (size_t l = base_level; l < top; ++l) out_.up();
    }
    return deepest;
  }

  void filter() {
    expect('?');
    expect('(');
    expect('@');
    out_.op(kMarkPosition);
    expression();
    while (more() && look() == ' ') ++cur_;
    if (look() != '=') fail_unsupported("Unsupported character for expression");
    expect('=');
    expect('=');
    while (more() && look() == ' ') ++cur_;
    const std::string value = quoted();
    out_.op(kExpressionStringEquals);
    out_.str("\"" + value + "\"");
    out_.op(kResetPosition);
    expect(')');
  }

  void key(const std::string &name) {
    out_.op(kMoveToKey);
    out_.str(name);
    out_.down();
    ++depth_;
  }
  void index(uint64_t i, bool reverse) {
    out_.op(reverse ? kMoveToIndexReverse : kMoveToIndex);
    out_.varint(i);
    out_.down();
    ++depth_;
  }

  std::string quoted() {
    const char open = take();
    if (open != '\'' && open != '"') fail_malformed("Invalid quoted string");
    const long from = cur_;
    bool escaped = false;
    while (more()) {
      const char c = look();
      if (escaped) escaped = false;
      else if (c == '\\') escaped = true;
      else if (c == open) break;
      take();
    }
    const long to = cur_;
    expect(open);
    return text_between(from, to);
  }

  uint64_t integer(const char *terminators) {
    const long from = cur_;
    while (more()) {
      const char c = look();
      if (std::isdigit(static_cast<unsigned char>(c))) {
        take();
        continue;
      }
      if (std::strchr(terminators, c) != nullptr) break;
      fail_malformed("Invalid integer");
    }
    const std::string digits = text_between(from, cur_);
    // std::stoull semantics of the reference (query_compiler.cpp:462-464)
    if (digits.empty())
      throw CompileFailure{CompileStatus::kInvalidArgument, "stoull"};
    uint64_t v = 0;
    for (char d : digits) {
      const uint64_t digit = static_cast<uint64_t>(d - '0');
      if (v > (UINT64_MAX - digit) / 10)
        throw CompileFailure{CompileStatus::kOutOfRange, "stoull"};
      v = v * 10 + digit;
    }
    return v;
  }

  const std::string &q_;
  bool ext_;
  long cur_ = -1;
  std::vector<long> marks_;
  size_t depth_ = 0;
  Emitter out_;
};

} // namespace

CompileStatus compile_query(const std::string &query, bool extensions,
                            CompiledBytecode *out, std::string *error) {
  try {
    Parser p(query, extensions);
    *out = p.run();
    return CompileStatus::kOk;
  } catch (const CompileFailure &f) {
    if (error) *error = f.message;
    return f.status;
  }
}

} // namespace gpj
