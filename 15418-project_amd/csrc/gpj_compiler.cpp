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
  kResetPosition = 8, kExpressionStringEquals = 9,
  // extension opcodes (emitted only with GPJ_COMPILE_EXTENSIONS)
  kMoveToEach = 10,      // iterate the elements of the current array
  kJumpIfNotArray = 11   // <4-byte varint forward offset>
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
  // Fixed-width (4-byte) LEB128 so a forward jump can be patched later.
  size_t jump_placeholder() {
    op(kJumpIfNotArray);
    const size_t at = bytes.size();
    bytes.insert(bytes.end(), {0x80, 0x80, 0x80, 0x00});
    return at;
  }
  void patch_jump(size_t at) {
    const uint32_t off = static_cast<uint32_t>(bytes.size() - (at + 4));
    bytes[at] = 0x80 | (off & 0x7F);
    bytes[at + 1] = 0x80 | ((off >> 7) & 0x7F);
    bytes[at + 2] = 0x80 | ((off >> 14) & 0x7F);
    bytes[at + 3] = (off >> 21) & 0x7F;
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
    // Extension bytecode reports the leveled rows it actually reads; the
    // reference-mode figure keeps the reference's own count.
    const size_t depth = ext_ ? rows_needed(out_.bytes) : depth_;
    return {out_.bytes, depth, out_.stores};
  }

  // Highest leveled row any path of the bytecode reads, plus one.
  static size_t rows_needed(const std::vector<uint8_t> &ir) {
    size_t need = 0;
    walk(ir, 0, 0, std::vector<size_t>{}, &need);
    return need;
  }
  static uint64_t read_varint(const std::vector<uint8_t> &ir, size_t &pos) {
    uint64_t v = 0;
    int shift = 0;
    while (pos < ir.size()) {
      const uint8_t b = ir[pos++];
      v |= static_cast<uint64_t>(b & 0x7F) << shift;
      if (!(b & 0x80)) break;
      shift += 7;
    }
    return v;
  }
  static void walk(const std::vector<uint8_t> &ir, size_t pos, size_t lvl,
                   std::vector<size_t> marks, size_t *need) {
    auto touch = [&](size_t row) {
      if (row + 1 > *need) *need = row + 1;
    };
    while (pos < ir.size()) {
      switch (ir[pos++]) {
      case kEnd: return;
      case kStoreResult: break;
      case kMoveUp: if (lvl > 0) --lvl; break;
      case kMoveDown: touch(lvl); ++lvl; break;
      case kMoveToKey: {
        const uint64_t n = read_varint(ir, pos);
        pos += n;
        touch(lvl);
        break;
      }
      case kMoveToIndex:
      case kMoveToIndexReverse: read_varint(ir, pos); touch(lvl); break;
      case kMarkPosition: marks.push_back(lvl); break;
      case kResetPosition:
        if (!marks.empty()) { lvl = marks.back(); marks.pop_back(); }
        break;
      case kExpressionStringEquals: {
        const uint64_t n = read_varint(ir, pos);
        pos += n;
        break;
      }
      case kMoveToEach: touch(lvl); break;
      case kJumpIfNotArray: {
        const uint64_t off = read_varint(ir, pos);
        walk(ir, pos + off, lvl, marks, need);
        break;
      }
      default: return;
      }
    }
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
      expect('*');
      out_.op(kMoveToEach);
      out_.down();
      ++depth_;
    } else if (c == '?') {
      if (ext_) {
        element_filter();
        return;
      }
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

  // Extension form of [?(...)]: on an array, filter its elements; on any
  // other value, test the value itself as the reference does.  Emitted as
  //   JUMP_IF_NOT_ARRAY L; EACH; DOWN; <filter>; <rest>; STORE; END
  //   L: <filter>; <rest>            (run() appends STORE; END)
  void element_filter() {
    const long cursor = cur_;
    const size_t saved_depth = depth_;
    const size_t saved_level = out_.level;
    const size_t jump = out_.jump_placeholder();
    out_.op(kMoveToEach);
    out_.down();
    ++depth_;
    filter();
    expect(']');
    if (more()) expression();
    out_.store();
    out_.op(kEnd);
    const size_t array_depth = depth_;
    out_.patch_jump(jump);
    cur_ = cursor;
    depth_ = saved_depth;
    out_.level = saved_level;
    filter();
    expect(']');
    if (more()) expression();
    if (array_depth > depth_) depth_ = array_depth;
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
