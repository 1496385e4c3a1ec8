// Driver configuration: a TOML subset (tables, dotted keys, strings,
// integers, string arrays, comments) and a small option parser, so the driver
// needs no third-party library.  Precedence: defaults < TOML < command line.
#include "driver_config.hpp"

#include <cctype>
#include <chrono>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>
#include <variant>

namespace gpjson::driver {

namespace {

using TomlValue = std::variant<int64_t, std::string, std::vector<std::string>>;

class TomlSubset {
public:
  static std::map<std::string, TomlValue> parse_file(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open TOML config '" + path + "'.");
    std::stringstream ss;
    ss << in.rdbuf();
    TomlSubset p(ss.str(), path);
    return p.parse();
  }

private:
  TomlSubset(std::string text, std::string path) : s_(std::move(text)), path_(std::move(path)) {}

  [[noreturn]] void fail(const std::string &m) const {
    throw std::runtime_error(path_ + ":" + std::to_string(line_) + ": " + m);
  }
  bool eof() const { return i_ >= s_.size(); }
  char cur() const { return eof() ? '\0' : s_[i_]; }
  void skip_ws() {
    while (!eof() && (cur() == ' ' || cur() == '\t' || cur() == '\r')) ++i_;
  }
  void skip_ws_nl() {
    for (;;) {
      skip_ws();
      if (cur() == '#') {
        while (!eof() && cur() != '\n') ++i_;
      }
      if (cur() == '\n') {
        ++line_;
        ++i_;
        continue;
      }
      return;
    }
  }
  std::string bare_key() {
    std::string k;
    while (!eof() && (std::isalnum(static_cast<unsigned char>(cur())) || cur() == '_' ||
                      cur() == '-'))
      k.push_back(s_[i_++]);
    if (k.empty()) fail("expected a key");
    return k;
  }
  std::string key() {
    std::string k;
    for (;;) {
      skip_ws();
      k += (cur() == '"') ? basic_string() : bare_key();
      skip_ws();
      if (cur() != '.') return k;
      ++i_;
      k.push_back('.');
    }
  }
  std::string basic_string() {
    const char q = cur();
    ++i_;
    std::string out;
    while (!eof() && cur() != q) {
      char c = s_[i_++];
      if (q == '"' && c == '\\') {
        const char e = s_[i_++];
        switch (e) {
        case 'n': out.push_back('\n'); break;
        case 't': out.push_back('\t'); break;
        case 'r': out.push_back('\r'); break;
        case '"': out.push_back('"'); break;
        case '\\': out.push_back('\\'); break;
        default: fail(std::string("unsupported escape \\") + e);
        }
      } else {
        if (c == '\n') fail("newline in string");
        out.push_back(c);
      }
    }
    if (eof()) fail("unterminated string");
    ++i_;
    return out;
  }
  TomlValue value() {
    skip_ws();
    if (cur() == '"' || cur() == '\'') return basic_string();
    if (cur() == '[') {
      ++i_;
      std::vector<std::string> arr;
      for (;;) {
        skip_ws_nl();
        if (cur() == ']') {
          ++i_;
          return arr;
        }
        if (cur() != '"' && cur() != '\'') fail("array values must be strings");
        arr.push_back(basic_string());
        skip_ws_nl();
        if (cur() == ',') {
          ++i_;
          continue;
        }
        if (cur() == ']') {
          ++i_;
          return arr;
        }
        fail("expected ',' or ']'");
      }
    }
    std::string num;
    if (cur() == '+' || cur() == '-') num.push_back(s_[i_++]);
    while (!eof() && (std::isdigit(static_cast<unsigned char>(cur())) || cur() == '_'))
      if (s_[i_++] != '_') num.push_back(s_[i_ - 1]);
    if (num.empty() || num == "+" || num == "-") fail("expected a value");
    try {
      return static_cast<int64_t>(std::stoll(num));
    } catch (const std::exception &) {
      fail("integer out of range");
    }
  }
  std::map<std::string, TomlValue> parse() {
    std::map<std::string, TomlValue> out;
    std::string table;
    for (;;) {
      skip_ws_nl();
      if (eof()) return out;
      if (cur() == '[') {
        ++i_;
        table = key();
        if (cur() != ']') fail("expected ']'");
        ++i_;
        continue;
      }
      const std::string k = key();
      if (cur() != '=') fail("expected '='");
      ++i_;
      const std::string full = table.empty() ? k : table + "." + k;
      if (out.count(full)) fail("duplicate key '" + full + "'");
      out[full] = value();
      skip_ws();
      if (cur() == '#') {
        while (!eof() && cur() != '\n') ++i_;
      }
      if (!eof() && cur() != '\n') fail("trailing characters after value");
    }
  }

  std::string s_, path_;
  size_t i_ = 0;
  int line_ = 1;
};

std::string json_escape(const std::string &t) {
  std::string e;
  for (const char c : t) {
    switch (c) {
    case '"': e += "\\\""; break;
    case '\\': e += "\\\\"; break;
    case '\b': e += "\\b"; break;
    case '\f': e += "\\f"; break;
    case '\n': e += "\\n"; break;
    case '\r': e += "\\r"; break;
    case '\t': e += "\\t"; break;
    default:
      if (static_cast<unsigned char>(c) < 0x20) {
        std::ostringstream s;
        s << "\\u" << std::hex << std::setw(4) << std::setfill('0')
          << static_cast<int>(static_cast<unsigned char>(c));
        e += s.str();
      } else {
        e.push_back(c);
      }
    }
  }
  return e;
}
std::string jstr(const std::string &t) { return "\"" + json_escape(t) + "\""; }

template <class T> T narrow(int64_t v, const std::string &key) {
  if constexpr (std::is_unsigned_v<T>) {
    if (v < 0) throw std::runtime_error("TOML value for '" + key + "' must be non-negative.");
  } else {
    if (v < static_cast<int64_t>(std::numeric_limits<T>::lowest()) ||
        v > static_cast<int64_t>(std::numeric_limits<T>::max()))
      throw std::runtime_error("TOML value for '" + key + "' is out of range.");
  }
  return static_cast<T>(v);
}

// One binding table serves TOML and the command line.
struct Binding {
  std::string key;
  std::vector<std::string> aliases;
  enum Kind { kString, kStringList, kInt, kSize, kReaderEnum, kBuilderEnum } kind;
  void *field;
};

std::vector<Binding> bindings(DriverInputs &in) {
  auto &fr = in.engine_options.file_reader_options;
  auto &ib = in.engine_options.index_builder_options;
  auto &qe = in.engine_options.query_executor_options;
  return {
      {"run_name", {}, Binding::kString, &in.run_name},
      {"dataset_path", {"dataset"}, Binding::kString, &in.dataset_path},
      {"queries", {"query", "q"}, Binding::kStringList, &in.queries},
      {"file_reader.type", {}, Binding::kReaderEnum, &fr.file_reader_type},
      {"index_builder.type", {}, Binding::kBuilderEnum, &ib.index_builder_type},
      {"index_builder.file_partition_size", {}, Binding::kSize, &ib.file_partition_size},
      {"index_builder.grid_size", {}, Binding::kInt, &ib.grid_size},
      {"index_builder.block_size", {}, Binding::kInt, &ib.block_size},
      {"index_builder.reduction_grid_size", {}, Binding::kInt, &ib.reduction_grid_size},
      {"index_builder.reduction_block_size", {}, Binding::kInt, &ib.reduction_block_size},
      {"query_executor.grid_size", {}, Binding::kInt, &qe.grid_size},
      {"query_executor.block_size", {}, Binding::kInt, &qe.block_size},
  };
}

void assign_toml(const Binding &b, const TomlValue &v) {
  auto need_int = [&]() -> int64_t {
    if (!std::holds_alternative<int64_t>(v))
      throw std::runtime_error("TOML value for '" + b.key + "' must be an integer.");
    return std::get<int64_t>(v);
  };
  auto need_str = [&]() -> std::string {
    if (!std::holds_alternative<std::string>(v))
      throw std::runtime_error("TOML value for '" + b.key + "' has the wrong type.");
    return std::get<std::string>(v);
  };
  switch (b.kind) {
  case Binding::kString: *static_cast<std::string *>(b.field) = need_str(); break;
  case Binding::kStringList:
    if (!std::holds_alternative<std::vector<std::string>>(v))
      throw std::runtime_error("TOML value for '" + b.key + "' must be an array of strings.");
    *static_cast<std::vector<std::string> *>(b.field) = std::get<std::vector<std::string>>(v);
    break;
  case Binding::kInt: *static_cast<int *>(b.field) = narrow<int>(need_int(), b.key); break;
  case Binding::kSize: *static_cast<size_t *>(b.field) = narrow<size_t>(need_int(), b.key); break;
  case Binding::kReaderEnum: {
    const std::string s = need_str();
    if (!s.empty())
      *static_cast<file::FileReaderType *>(b.field) = parse_enum_value<file::FileReaderType>(s);
    break;
  }
  case Binding::kBuilderEnum: {
    const std::string s = need_str();
    if (!s.empty())
      *static_cast<index::IndexBuilderType *>(b.field) = parse_enum_value<index::IndexBuilderType>(s);
    break;
  }
  }
}

std::string cli_name(const std::string &key) {
  std::string k = key;
  for (char &c : k)
    if (c == '_') c = '-';
  return k.size() == 1 ? "-" + k : "--" + k;
}

int64_t parse_cli_int(const std::string &opt, const std::string &text) {
  try {
    size_t used = 0;
    const long long v = std::stoll(text, &used);
    if (used != text.size()) throw std::invalid_argument(text);
    return v;
  } catch (const std::exception &) {
    std::cerr << opt << ": Value " << text << " could not be converted\n";
    throw CliExit(106);
  }
}

void print_help(const std::vector<Binding> &bs) {
  std::cout << "Run gpjson queries against a JSON-lines dataset\nOptions:\n"
            << "  -h,--help\n  -c,--config TEXT    TOML config file path\n";
  for (const auto &b : bs) {
    std::cout << "  " << cli_name(b.key);
    for (const auto &a : b.aliases) std::cout << "," << cli_name(a);
    std::cout << "\n";
  }
}

} // namespace

template <> file::FileReaderType parse_enum_value<file::FileReaderType>(const std::string &text) {
  if (text == "COPIED") return file::FileReaderType::COPIED;
  if (text == "MMAP") return file::FileReaderType::MMAP;
  throw std::runtime_error("Invalid enum value '" + text + "'. Expected one of: COPIED, MMAP");
}

template <>
index::IndexBuilderType parse_enum_value<index::IndexBuilderType>(const std::string &text) {
  if (text == "UNCOMBINED") return index::IndexBuilderType::UNCOMBINED;
  if (text == "COMBINED") return index::IndexBuilderType::COMBINED;
  if (text == "FUSED") return index::IndexBuilderType::FUSED;
  if (text == "SHAREMEM") return index::IndexBuilderType::SHAREMEM;
  throw std::runtime_error("Invalid enum value '" + text +
                           "'. Expected one of: UNCOMBINED, COMBINED, FUSED, SHAREMEM");
}

DriverInputs default_driver_inputs() {
  DriverInputs in;
  const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                      std::chrono::system_clock::now().time_since_epoch())
                      .count();
  in.run_name = "run_" + std::to_string(ms);
  return in;
}

DriverInputs parse_driver_inputs(int argc, char **argv) {
  DriverInputs in = default_driver_inputs();
  std::vector<std::string> args(argv + 1, argv + argc);
  // pre-pass: the config file
  std::string config;
  for (size_t i = 0; i < args.size(); ++i) {
    const std::string &a = args[i];
    if ((a == "-c" || a == "--config") && i + 1 < args.size()) config = args[i + 1];
    else if (a.rfind("--config=", 0) == 0) config = a.substr(9);
  }
  auto bs = bindings(in);
  if (!config.empty()) {
    const auto table = TomlSubset::parse_file(config);
    for (const auto &b : bs) {
      auto it = table.find(b.key);
      if (it != table.end()) assign_toml(b, it->second);
    }
  }
  // command line overrides
  std::vector<std::string> cli_queries;
  bool cli_queries_seen = false;
  for (size_t i = 0; i < args.size(); ++i) {
    std::string opt = args[i], val;
    bool has_inline = false;
    if (opt == "-h" || opt == "--help") {
      print_help(bs);
      throw CliExit(0);
    }
    const size_t eq = opt.find('=');
    if (opt.rfind("--", 0) == 0 && eq != std::string::npos) {
      val = opt.substr(eq + 1);
      opt = opt.substr(0, eq);
      has_inline = true;
    }
    if (opt == "-c" || opt == "--config") {
      if (!has_inline) ++i;
      continue;
    }
    const Binding *match = nullptr;
    for (const auto &b : bs) {
      if (cli_name(b.key) == opt) match = &b;
      for (const auto &a : b.aliases)
        if (cli_name(a) == opt) match = &b;
    }
    if (!match) {
      std::cerr << "The following argument was not expected: " << opt << "\n";
      throw CliExit(109);
    }
    if (!has_inline) {
      if (i + 1 >= args.size()) {
        std::cerr << opt << ": 1 required TEXT missing\n";
        throw CliExit(107);
      }
      val = args[++i];
    }
    switch (match->kind) {
    case Binding::kString: *static_cast<std::string *>(match->field) = val; break;
    case Binding::kStringList:
      cli_queries_seen = true;
      cli_queries.push_back(val);
      break;
    case Binding::kInt:
      *static_cast<int *>(match->field) = static_cast<int>(parse_cli_int(opt, val));
      break;
    case Binding::kSize: {
      const int64_t v = parse_cli_int(opt, val);
      if (v < 0) {
        std::cerr << opt << ": Value " << val << " could not be converted\n";
        throw CliExit(106);
      }
      *static_cast<size_t *>(match->field) = static_cast<size_t>(v);
      break;
    }
    case Binding::kReaderEnum:
      *static_cast<file::FileReaderType *>(match->field) = parse_enum_value<file::FileReaderType>(val);
      break;
    case Binding::kBuilderEnum:
      *static_cast<index::IndexBuilderType *>(match->field) =
          parse_enum_value<index::IndexBuilderType>(val);
      break;
    }
  }
  if (cli_queries_seen) in.queries = cli_queries;
  validate_driver_inputs(in);
  return in;
}

void validate_driver_inputs(const DriverInputs &in) {
  if (in.run_name.empty()) throw std::runtime_error("run_name must not be empty.");
  if (in.dataset_path.empty())
    throw std::runtime_error(
        "dataset_path is required. Use --dataset or set dataset_path in the TOML config.");
  if (in.queries.empty())
    throw std::runtime_error(
        "At least one query is required. Use --query or set queries in the TOML config.");
  const auto &o = in.engine_options.index_builder_options;
  const auto &q = in.engine_options.query_executor_options;
  if (o.grid_size < 0 || o.block_size < 0 || o.reduction_grid_size < 0 ||
      o.reduction_block_size < 0)
    throw std::runtime_error("Index builder grid and block sizes must be non-negative.");
  if (q.grid_size < 0 || q.block_size < 0)
    throw std::runtime_error("Query executor grid and block sizes must be non-negative.");
}

std::string engine_options_to_json(const gpjson::EngineOptions &o) {
  std::ostringstream s;
  std::ostringstream fr, ib;
  fr << o.file_reader_options.file_reader_type;
  ib << o.index_builder_options.index_builder_type;
  const auto &b = o.index_builder_options;
  s << "{\"file_reader\":{\"type\":" << jstr(fr.str()) << "},\"index_builder\":{\"type\":"
    << jstr(ib.str()) << ",\"file_partition_size\":" << b.file_partition_size
    << ",\"grid_size\":" << b.grid_size << ",\"block_size\":" << b.block_size
    << ",\"reduction_grid_size\":" << b.reduction_grid_size
    << ",\"reduction_block_size\":" << b.reduction_block_size
    << "},\"query_executor\":{\"grid_size\":" << o.query_executor_options.grid_size
    << ",\"block_size\":" << o.query_executor_options.block_size << "}}";
  return s.str();
}

std::string driver_inputs_to_json(const DriverInputs &in) {
  std::ostringstream s;
  s << "{\"run_name\":" << jstr(in.run_name) << ",\"dataset_path\":" << jstr(in.dataset_path)
    << ",\"queries\":[";
  for (size_t i = 0; i < in.queries.size(); ++i) s << (i ? "," : "") << jstr(in.queries[i]);
  s << "],\"engine_options\":" << engine_options_to_json(in.engine_options) << "}";
  return s.str();
}

} // namespace gpjson::driver
