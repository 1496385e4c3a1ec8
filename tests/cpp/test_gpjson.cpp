// C++ API tests: the reference's gtest cases re-expressed against the
// MI355X engine (tests/app, tests/file, tests/query, tests/index,
// tests/engine_test.cpp, tests/cuda of the reference).
#include "minitest.hpp"

#include "../../15418-project_amd/host/apps/driver_config.hpp"
#include "gpj_capi.h"
#include "gpjson/device/device.hpp"
#include "gpjson/engine.hpp"
#include "gpjson/file/error.hpp"
#include "gpjson/file/file_reader.hpp"
#include "gpjson/index/index_builder.hpp"
#include "gpjson/profiler/profiler.hpp"
#include "gpjson/query/ir_input_buffer.hpp"
#include "gpjson/query/query_compiler.hpp"
#include "gpjson/query/query_executor.hpp"

#include <array>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>

using namespace gpjson;

namespace {

std::filesystem::path temp_file(const std::string &name, const std::string &contents) {
  const auto now = std::chrono::steady_clock::now().time_since_epoch().count();
  auto p = std::filesystem::temp_directory_path() / (std::to_string(now) + "_" + name);
  std::ofstream o(p, std::ios::binary);
  o << contents;
  return p;
}

struct TempFile {
  std::filesystem::path path;
  TempFile(const std::string &name, const std::string &contents) : path(temp_file(name, contents)) {}
  ~TempFile() {
    std::error_code ec;
    std::filesystem::remove(path, ec);
  }
};

std::vector<char *> argv_of(std::vector<std::string> &args) {
  std::vector<char *> v;
  for (auto &a : args) v.push_back(a.data());
  return v;
}

driver::DriverInputs parse(std::vector<std::string> args) {
  auto v = argv_of(args);
  return driver::parse_driver_inputs(static_cast<int>(v.size()), v.data());
}

std::unique_ptr<file::FileReader> reader_of(file::FileReaderType t, const std::string &path) {
  if (t == file::FileReaderType::COPIED) return std::make_unique<file::CopiedFileReader>(path);
  return std::make_unique<file::MmapFileReader>(path);
}

std::string bytes_of(const file::FilePartition &p) {
  if (!p.host_bytes() || !p.size_bytes()) return {};
  return std::string(static_cast<const char *>(p.host_bytes()), p.size_bytes());
}

const std::array<file::FileReaderType, 2> kReaders{file::FileReaderType::COPIED,
                                                   file::FileReaderType::MMAP};
const std::array<index::IndexBuilderType, 4> kBuilders{
    index::IndexBuilderType::UNCOMBINED, index::IndexBuilderType::COMBINED,
    index::IndexBuilderType::FUSED, index::IndexBuilderType::SHAREMEM};

} // namespace

// ---- driver config (tests/app/driver_config_test.cpp:49-217) --------------
TEST(DriverConfig, DefaultInputsUseOptionStructDefaults) {
  const auto in = driver::default_driver_inputs();
  EXPECT_FALSE(in.run_name.empty());
  EXPECT_TRUE(in.dataset_path.empty());
  EXPECT_TRUE(in.queries.empty());
  EXPECT_TRUE(in.engine_options.file_reader_options.file_reader_type == file::FileReaderType::COPIED);
  const auto &ib = in.engine_options.index_builder_options;
  EXPECT_TRUE(ib.index_builder_type == index::IndexBuilderType::UNCOMBINED);
  EXPECT_EQ(ib.file_partition_size, 0u);
  EXPECT_EQ(ib.grid_size, 0);
  EXPECT_EQ(ib.block_size, 0);
  EXPECT_EQ(ib.reduction_grid_size, 0);
  EXPECT_EQ(ib.reduction_block_size, 0);
  EXPECT_EQ(in.engine_options.query_executor_options.grid_size, 512);
  EXPECT_EQ(in.engine_options.query_executor_options.block_size, 1024);
}

TEST(DriverConfig, ParsesTomlConfig) {
  TempFile f("cfg.toml", R"toml(
dataset_path = "dataset/input.json"
run_name = "toml_run"
queries = ["$.user.lang", "$.user.id"]

[file_reader]
type = "MMAP"

[index_builder]
type = "COMBINED"
file_partition_size = 65536
grid_size = 16384
block_size = 1024
reduction_grid_size = 32
reduction_block_size = 16

[query_executor]
grid_size = 99
block_size = 128
)toml");
  const auto in = parse({"gpjson_driver", "--config", f.path.string()});
  EXPECT_EQ(in.run_name, std::string("toml_run"));
  EXPECT_EQ(in.dataset_path, std::string("dataset/input.json"));
  EXPECT_EQ(in.queries.size(), 2u);
  EXPECT_EQ(in.queries[0], std::string("$.user.lang"));
  EXPECT_EQ(in.queries[1], std::string("$.user.id"));
  EXPECT_TRUE(in.engine_options.file_reader_options.file_reader_type == file::FileReaderType::MMAP);
  const auto &ib = in.engine_options.index_builder_options;
  EXPECT_TRUE(ib.index_builder_type == index::IndexBuilderType::COMBINED);
  EXPECT_EQ(ib.file_partition_size, 65536u);
  EXPECT_EQ(ib.grid_size, 16384);
  EXPECT_EQ(ib.block_size, 1024);
  EXPECT_EQ(ib.reduction_grid_size, 32);
  EXPECT_EQ(ib.reduction_block_size, 16);
  EXPECT_EQ(in.engine_options.query_executor_options.grid_size, 99);
  EXPECT_EQ(in.engine_options.query_executor_options.block_size, 128);
}

TEST(DriverConfig, CliOverridesTomlConfig) {
  TempFile f("cfg.toml", R"toml(
dataset_path = "dataset/from_toml.json"
run_name = "toml_run"
queries = ["$.toml"]

[file_reader]
type = "COPIED"

[index_builder]
type = "UNCOMBINED"
grid_size = 1
block_size = 2

[query_executor]
grid_size = 3
block_size = 4
)toml");
  const auto in = parse({"gpjson_driver", "--config", f.path.string(), "--dataset",
                         "dataset/from_cli.json", "--query", "$.cli.one", "--query", "$.cli.two",
                         "--run-name", "cli_run", "--index-builder.type", "COMBINED",
                         "--file-reader.type", "MMAP", "--index-builder.grid-size", "16384",
                         "--query-executor.grid-size", "55", "--query-executor.block-size", "64"});
  EXPECT_EQ(in.run_name, std::string("cli_run"));
  EXPECT_EQ(in.dataset_path, std::string("dataset/from_cli.json"));
  EXPECT_EQ(in.queries.size(), 2u);
  EXPECT_EQ(in.queries[0], std::string("$.cli.one"));
  EXPECT_EQ(in.queries[1], std::string("$.cli.two"));
  EXPECT_TRUE(in.engine_options.file_reader_options.file_reader_type == file::FileReaderType::MMAP);
  const auto &ib = in.engine_options.index_builder_options;
  EXPECT_TRUE(ib.index_builder_type == index::IndexBuilderType::COMBINED);
  EXPECT_EQ(ib.grid_size, 16384);
  EXPECT_EQ(ib.block_size, 2);
  EXPECT_EQ(in.engine_options.query_executor_options.grid_size, 55);
  EXPECT_EQ(in.engine_options.query_executor_options.block_size, 64);
}

TEST(DriverConfig, RejectsNegativeQueryExecutorLaunchSizes) {
  auto in = driver::default_driver_inputs();
  in.dataset_path = "dataset/input.json";
  in.queries = {"$.user.lang"};
  in.engine_options.query_executor_options.grid_size = -1;
  EXPECT_THROW(driver::validate_driver_inputs(in), std::runtime_error);
}

TEST(DriverConfig, ParsesEnumNames) {
  EXPECT_TRUE(driver::parse_enum_value<file::FileReaderType>("COPIED") == file::FileReaderType::COPIED);
  EXPECT_TRUE(driver::parse_enum_value<file::FileReaderType>("MMAP") == file::FileReaderType::MMAP);
  EXPECT_TRUE(driver::parse_enum_value<index::IndexBuilderType>("COMBINED") ==
              index::IndexBuilderType::COMBINED);
  EXPECT_TRUE(driver::parse_enum_value<index::IndexBuilderType>("FUSED") ==
              index::IndexBuilderType::FUSED);
}

TEST(DriverConfig, RejectsInvalidEnum) {
  EXPECT_THROW(driver::parse_enum_value<file::FileReaderType>("BAD_READER"), std::runtime_error);
  EXPECT_THROW(driver::parse_enum_value<index::IndexBuilderType>("BAD_BUILDER"), std::runtime_error);
}

TEST(DriverConfig, RunConfigJson) {
  auto in = driver::default_driver_inputs();
  in.run_name = "r\"1";
  in.dataset_path = "d";
  in.queries = {"$.a"};
  EXPECT_EQ(driver::driver_inputs_to_json(in),
            std::string("{\"run_name\":\"r\\\"1\",\"dataset_path\":\"d\",\"queries\":[\"$.a\"],"
                        "\"engine_options\":{\"file_reader\":{\"type\":\"COPIED\"},"
                        "\"index_builder\":{\"type\":\"UNCOMBINED\",\"file_partition_size\":0,"
                        "\"grid_size\":0,\"block_size\":0,\"reduction_grid_size\":0,"
                        "\"reduction_block_size\":0},\"query_executor\":{\"grid_size\":512,"
                        "\"block_size\":1024}}}"));
}

// ---- file readers (tests/file/file_reader_test.cpp:89-155) ----------------
TEST(FileReader, CreatesSinglePartitionForWholeFile) {
  for (auto t : kReaders) {
    const std::string c = "{\"id\":1}\n{\"id\":2}\n";
    TempFile f("single.json", c);
    auto r = reader_of(t, f.path.string());
    r->create_partitions(0);
    EXPECT_EQ(r->metadata().file_path, f.path.string());
    EXPECT_EQ(r->metadata().file_size_bytes, c.size());
    EXPECT_EQ(r->metadata().num_partitions, 1u);
    EXPECT_EQ(r->get_partitions().size(), 1u);
    EXPECT_EQ(bytes_of(r->get_partitions()[0]), c);
  }
}

TEST(FileReader, SplitsPartitionsAtNewlineBoundaries) {
  for (auto t : kReaders) {
    const std::string c = "alpha\nbeta\ngamma";
    TempFile f("split.txt", c);
    auto r = reader_of(t, f.path.string());
    r->create_partitions(7);
    const auto &p = r->get_partitions();
    EXPECT_EQ(p.size(), 3u);
    EXPECT_EQ(p[0].global_start_offset(), 0u);
    EXPECT_EQ(p[0].global_end_offset(), 5u);
    EXPECT_EQ(bytes_of(p[0]), std::string("alpha"));
    EXPECT_EQ(p[1].global_start_offset(), 6u);
    EXPECT_EQ(p[1].global_end_offset(), 10u);
    EXPECT_EQ(bytes_of(p[1]), std::string("beta"));
    EXPECT_EQ(p[2].partition_id(), 2u);
    EXPECT_EQ(p[2].global_start_offset(), 11u);
    EXPECT_EQ(bytes_of(p[2]), std::string("gamma"));
  }
}

TEST(FileReader, ThrowsWhenNoNewlineExistsBeforeBoundary) {
  for (auto t : kReaders) {
    TempFile f("noline.txt", "abcdefghij");
    auto r = reader_of(t, f.path.string());
    EXPECT_THROW(r->create_partitions(4), error::file::PartitionError);
  }
}

TEST(FileReader, CreatesSingleEmptyPartitionForEmptyFile) {
  for (auto t : kReaders) {
    TempFile f("empty.txt", "");
    auto r = reader_of(t, f.path.string());
    r->create_partitions(64);
    EXPECT_EQ(r->metadata().file_size_bytes, 0u);
    EXPECT_EQ(r->get_partitions().size(), 1u);
    EXPECT_EQ(r->get_partitions()[0].size_bytes(), 0u);
  }
}

TEST(FileReader, MissingFileRaisesFileOpenError) {
  for (auto t : kReaders) {
    auto r = reader_of(t, "/nonexistent/gpjson/input.json");
    EXPECT_THROW(r->create_partitions(0), error::file::FileOpenError);
  }
}

// ---- query compiler (tests/query/query_compiler_test.cpp:58-147) ----------
namespace {
struct Ins {
  query::QueryOpcode op;
  std::string s;
  size_t n;
};
std::vector<Ins> decode(const query::CompiledQuery &q) {
  std::vector<Ins> out;
  query::IRByteInputBuffer in(q.ir_bytes().data(), q.ir_bytes().size());
  while (in.has_next()) {
    Ins i{in.read_opcode(), {}, 0};
    if (i.op == query::QueryOpcode::MOVE_TO_KEY || i.op == query::QueryOpcode::EXPRESSION_STRING_EQUALS)
      i.s = in.read_string();
    else if (i.op == query::QueryOpcode::MOVE_TO_INDEX || i.op == query::QueryOpcode::MOVE_TO_INDEX_REVERSE)
      i.n = in.read_varint();
    out.push_back(i);
    if (i.op == query::QueryOpcode::END) break;
  }
  return out;
}
} // namespace

TEST(QueryCompiler, CompilesSimplePropertyQuery) {
  EngineOptions o{};
  query::QueryCompiler c(o);
  const auto q = c.compile("$.name", o);
  const auto ins = decode(q);
  EXPECT_EQ(q.query_text(), std::string("$.name"));
  EXPECT_EQ(q.max_depth(), 1u);
  EXPECT_EQ(q.num_result_slots(), 1u);
  EXPECT_EQ(ins.size(), 4u);
  EXPECT_TRUE(ins[0].op == query::QueryOpcode::MOVE_TO_KEY);
  EXPECT_EQ(ins[0].s, std::string("name"));
  EXPECT_TRUE(ins[1].op == query::QueryOpcode::MOVE_DOWN);
  EXPECT_TRUE(ins[2].op == query::QueryOpcode::STORE_RESULT);
  EXPECT_TRUE(ins[3].op == query::QueryOpcode::END);
}

TEST(QueryCompiler, CompilesArrayRangeIntoTwoResultSlots) {
  EngineOptions o{};
  const auto q = query::QueryCompiler(o).compile("$.arr[0:2]", o);
  const auto ins = decode(q);
  EXPECT_EQ(q.max_depth(), 2u);
  EXPECT_EQ(q.num_result_slots(), 2u);
  EXPECT_EQ(ins.size(), 10u);
  EXPECT_TRUE(ins[2].op == query::QueryOpcode::MOVE_TO_INDEX);
  EXPECT_EQ(ins[2].n, 0u);
  EXPECT_TRUE(ins[5].op == query::QueryOpcode::MOVE_UP);
  EXPECT_TRUE(ins[6].op == query::QueryOpcode::MOVE_TO_INDEX);
  EXPECT_EQ(ins[6].n, 1u);
}

TEST(QueryCompiler, CompilesFilterExpression) {
  EngineOptions o{};
  const auto q = query::QueryCompiler(o).compile("$.user.lang[?(@ == \"en\")]", o);
  const auto ins = decode(q);
  EXPECT_EQ(q.max_depth(), 2u);
  EXPECT_EQ(q.num_result_slots(), 1u);
  EXPECT_EQ(ins.size(), 9u);
  EXPECT_TRUE(ins[4].op == query::QueryOpcode::MARK_POSITION);
  EXPECT_TRUE(ins[5].op == query::QueryOpcode::EXPRESSION_STRING_EQUALS);
  EXPECT_EQ(ins[5].s, std::string("\"en\""));
  EXPECT_TRUE(ins[6].op == query::QueryOpcode::RESET_POSITION);
}

TEST(QueryCompiler, RejectsMalformedAndUnsupportedQueries) {
  EngineOptions o{};
  query::QueryCompiler c(o);
  EXPECT_THROW(c.compile("$.", o), error::query::MalformedQueryError);
  EXPECT_THROW(c.compile("$.arr[-0]", o), error::query::MalformedQueryError);
  EXPECT_THROW(c.compile("$..name", o), error::query::UnsupportedQueryError);
  EXPECT_THROW(c.compile("$.arr[*]", o), error::query::UnsupportedQueryError);
  EXPECT_THROW(c.compile("$.a[1:]", o), std::invalid_argument);
}

TEST(Profiler, RecordsNestedSegments) {
  profiler::Profiler p("t", false);
  const auto a = p.begin_nested("outer");
  const auto b = p.beginf("inner %d", 7);
  p.end(b);
  p.end(a);
  const auto segs = p.completed();
  EXPECT_EQ(segs.size(), 2u);
  EXPECT_EQ(segs[0].name, std::string("outer"));
  EXPECT_EQ(segs[1].name, std::string("inner 7"));
  EXPECT_EQ(segs[1].indent, 1u);
}

// ---- device (tests/cuda/test_cuda.cpp) ------------------------------------
// ---- compact executor results (texts made only when read) -----------------
namespace {
query::CompactLines sample_chunk(size_t base) {
  // "abcdefghij": line 0 holds [1,3) and nothing; line 1 nothing; line 2
  // holds [4,7) and [0,1)
  query::CompactLines c;
  c.base = base;
  c.lines = 3;
  c.slots = 2;
  c.spans = {1, 3, -1, -1, -1, -1, -1, -1, 4, 7, 0, 1};
  c.texts = "bc" "efg" "a";
  return c;
}
} // namespace

TEST(BatchQueryResult, CompactResultsReadPerLineAndMaterialized) {
  query::BatchQueryResult r(1);
  r.set_query_text(0, "$.x");
  r.add_compact(0, sample_chunk(100));
  const auto &lines = r.queries()[0].lines();
  EXPECT_EQ(lines.size(), 3u);
  EXPECT_EQ(*lines[0].offsets()[0].start, size_t{101});
  EXPECT_EQ(*lines[0].offsets()[0].end, size_t{103});
  EXPECT_EQ(*lines[0].offsets()[0].json_text, std::string("bc"));
  EXPECT_TRUE(!lines[0].offsets()[1].json_text.has_value());
  EXPECT_TRUE(!lines[1].offsets()[0].start.has_value());
  EXPECT_EQ(*lines[2].offsets()[0].json_text, std::string("efg"));
  EXPECT_EQ(*lines[2].offsets()[1].json_text, std::string("a"));
  const auto m = r.materialize();
  EXPECT_EQ(m.queries()[0].lines().size(), 2u);
  EXPECT_EQ(m.queries()[0].lines()[1].values()[1].json_text(), std::string("a"));
}

TEST(BatchQueryResult, CompactMaterializeWithoutPerLineForm) {
  // materialized straight from the compact form, in order with eager lines
  query::BatchQueryResult r(1);
  query::LineQueryResult eager;
  query::QueryOffset off;
  off.start = 0;
  off.end = 1;
  off.json_text = "z";
  eager.add_offset(off);
  r.add_line_result(0, eager);
  r.add_compact(0, sample_chunk(0));
  query::BatchQueryResult other(1);
  other.add_compact(0, sample_chunk(50));
  r.merge(other);
  const auto m = std::move(r).materialize();
  const auto &l = m.queries()[0].lines();
  EXPECT_EQ(l.size(), 5u); // "z", then 2 lines of each chunk
  EXPECT_EQ(l[0].values()[0].json_text(), std::string("z"));
  EXPECT_EQ(l[1].values()[0].json_text(), std::string("bc"));
  EXPECT_EQ(l[4].values()[0].json_text(), std::string("efg"));
}

GPU_TEST(DeviceArray, MoveAndOversizeCopy) {
  device::DeviceArray a(64);
  void *p = a.data();
  device::DeviceArray b(std::move(a));
  EXPECT_TRUE(a.data() == nullptr);
  EXPECT_TRUE(b.data() == p);
  std::array<char, 128> host{};
  EXPECT_THROW(b.copy_from_host(host.data(), 128), error::device::DeviceError);
  b.copy_from_host(host.data(), 64);
  std::array<char, 64> back{};
  b.copy_to_host(back.data(), 64);
  EXPECT_TRUE(std::memcmp(back.data(), host.data(), 64) == 0);
}

GPU_TEST(FilePartition, LoadToDeviceIsIdempotent) {
  const std::string t = R"({"id":1})";
  file::FilePartition p(0, 0, t.size(), t.data());
  EXPECT_FALSE(p.device_loaded());
  p.load_to_device();
  const void *d = p.device_bytes();
  EXPECT_TRUE(d != nullptr);
  p.load_to_device();
  EXPECT_TRUE(p.device_bytes() == d);
}

// ---- index builders vs closed forms (tests/index/*_test.cpp) --------------
namespace {
const std::string kFixture =
    "{\"id\":1,\"name\":\"Ada\",\"tags\":[\"gpu\",\"json\"],\"active\":true}\n"
    "{\"id\":2,\"name\":\"Bob \\\"The Builder\\\"\",\"note\":\"literal { braces } and [brackets] "
    "stay strings\"}\n"
    "{\"id\":3,\"nested\":{\"arr\":[1,{\"x\":\"y,z\"}],\"empty\":{}},\"path\":\"C:\\\\tmp\\\\file\"}";

bool escaped_quote(const std::string &d, size_t i) {
  size_t k = 0;
  while (i > k && d[i - k - 1] == '\\') ++k;
  return k % 2 == 1;
}
std::vector<uint64_t> expected_string(const std::string &d) {
  std::vector<uint64_t> w((d.size() + 63) / 64, 0);
  bool in = false;
  for (size_t i = 0; i < d.size(); ++i) {
    if (d[i] == '"' && !escaped_quote(d, i)) {
      if (!in) w[i / 64] |= 1ull << (i % 64);
      in = !in;
    } else if (in) {
      w[i / 64] |= 1ull << (i % 64);
    }
  }
  return w;
}
std::vector<uint64_t> expected_leveled(const std::string &d, int levels) {
  const size_t ls = (d.size() + 63) / 64;
  const auto s = expected_string(d);
  std::vector<uint64_t> out(ls * levels, 0);
  int level = -1;
  for (size_t i = 0; i < d.size(); ++i) {
    if (s[i / 64] >> (i % 64) & 1) continue;
    const char c = d[i];
    const uint64_t bit = 1ull << (i % 64);
    if (c == '{' || c == '[') {
      ++level;
      if (level >= 0 && level < levels) out[ls * level + i / 64] |= bit;
    } else if (c == '}' || c == ']') {
      if (level >= 0 && level < levels) out[ls * level + i / 64] |= bit;
      --level;
    } else if (c == ':' || c == ',') {
      if (level >= 0 && level < levels) out[ls * level + i / 64] |= bit;
    }
  }
  return out;
}
template <class T> std::vector<T> copy_words(const index::Index &idx) {
  std::vector<T> v(idx.size_bytes() / sizeof(T));
  if (!v.empty() && gpj_d2h(v.data(), idx.data(), v.size() * sizeof(T), nullptr) != 0)
    throw std::runtime_error(gpj_last_error());
  return v;
}
} // namespace

GPU_TEST(IndexBuilder, AllBuildersMatchClosedForms) {
  constexpr int kMaxDepth = 8;
  file::FilePartition p(0, 0, kFixture.size(), kFixture.data());
  p.load_to_device();
  file::CopiedFileReader reader{"in-memory.ldjson"};
  profiler::Profiler prof("IndexBuilderTest", false);
  RunContext ctx{prof};
  for (auto t : kBuilders) {
    index::IndexBuilderOptions o{t, 0, 1, 512, 1, 64};
    std::unique_ptr<index::IndexBuilder> b;
    switch (t) {
    case index::IndexBuilderType::UNCOMBINED: b = std::make_unique<index::UncombinedIndexBuilder>(reader); break;
    case index::IndexBuilderType::COMBINED: b = std::make_unique<index::CombinedIndexBuilder>(reader); break;
    case index::IndexBuilderType::FUSED: b = std::make_unique<index::FusedIndexBuilder>(reader); break;
    case index::IndexBuilderType::SHAREMEM: b = std::make_unique<index::SharememIndexBuilder>(reader); break;
    }
    const auto built = b->build(p, kMaxDepth, o, ctx);
    std::vector<int64_t> nl{0};
    for (size_t i = 0; i < kFixture.size(); ++i)
      if (kFixture[i] == '\n') nl.push_back(static_cast<int64_t>(i));
    EXPECT_EQ(built.get_newline_index().num_lines(), static_cast<int64_t>(nl.size()));
    EXPECT_EQ(copy_words<int64_t>(built.get_newline_index()), nl);
    EXPECT_EQ(copy_words<uint64_t>(built.get_string_index()), expected_string(kFixture));
    EXPECT_EQ(built.get_leveled_bitmap_index().num_levels(), kMaxDepth);
    EXPECT_EQ(copy_words<uint64_t>(built.get_leveled_bitmap_index()),
              expected_leveled(kFixture, kMaxDepth));
  }
}

// ---- query executor (tests/query/query_executor_test.cpp:155-245) ---------
namespace {
const std::string kBookstore =
    R"({"store":{"book":[{"category":"reference","author":"Nigel Rees","title":"Sayings of the Century","price":"8.95"},{"category":"fiction","author":"Herman Melville","title":"Moby Dick","isbn":"0-553-21311-3","price":"8.99"},{"category":"fiction","author":"J.R.R. Tolkien","title":"The Lord of the Rings","isbn":"0-395-19395-8","price":"22.99"}],"bicycle":{"color":"red","price":"19.95"}},"expensive":"10"})";

// Indices built on the CPU the way the reference test does: the string index
// marks only the last byte of each string, which the key search needs.
index::BuiltIndices handmade_indices(const std::string &t) {
  const size_t ls = (t.size() + 63) / 64;
  std::vector<int64_t> nl{0};
  std::vector<uint64_t> str(ls, 0);
  std::vector<std::vector<size_t>> by_level;
  std::vector<char> stack;
  bool in = false, esc = false;
  for (size_t i = 0; i < t.size(); ++i) {
    const char c = t[i];
    if (in) {
      if (esc) { esc = false; continue; }
      if (c == '\\') { esc = true; continue; }
      if (c == '"') {
        if (i > 0) str[(i - 1) / 64] |= 1ull << ((i - 1) % 64);
        in = false;
      }
      continue;
    }
    if (c == '"') { in = true; continue; }
    if (c == '{' || c == '[') {
      stack.push_back(c);
      if (by_level.size() < stack.size()) by_level.resize(stack.size());
      continue;
    }
    if (stack.empty()) continue;
    if (c == ':' || c == ',') { by_level[stack.size() - 1].push_back(i); continue; }
    if (c == '}' || c == ']') { by_level[stack.size() - 1].push_back(i); stack.pop_back(); }
  }
  std::vector<uint64_t> lv(ls * by_level.size(), 0);
  for (size_t l = 0; l < by_level.size(); ++l)
    for (size_t p : by_level[l]) lv[l * ls + p / 64] |= 1ull << (p % 64);
  device::DeviceArray dn(nl.size() * 8), ds(str.size() * 8), dl(lv.size() * 8);
  dn.copy_from_host(nl.data(), nl.size() * 8);
  ds.copy_from_host(str.data(), str.size() * 8);
  dl.copy_from_host(lv.data(), lv.size() * 8);
  return index::BuiltIndices(index::NewlineIndex(std::move(dn), static_cast<int64_t>(nl.size())),
                             index::StringIndex(std::move(ds)),
                             index::LeveledBitmapIndex(std::move(dl), static_cast<int>(by_level.size())));
}
} // namespace

GPU_TEST(QueryExecutor, ExecutesCompiledQueriesWithHandcraftedIndices) {
  file::FilePartition p(0, 0, kBookstore.size(), kBookstore.data());
  p.load_to_device();
  auto built = handmade_indices(kBookstore);
  EngineOptions o{};
  o.query_executor_options.grid_size = 7;
  o.query_executor_options.block_size = 32;
  query::QueryCompiler c(o);
  query::BatchCompiledQuery qs;
  qs.add(c.compile("$.expensive", o));
  qs.add(c.compile("$.store.book[1].author", o));
  const auto r = query::QueryExecutor(o).execute_batch(qs, p, built);
  const auto m = r.materialize();
  EXPECT_EQ(r.num_queries(), 2u);
  EXPECT_EQ(r.queries()[0].lines().size(), 1u);
  EXPECT_EQ(m.queries()[0].query_text(), std::string("$.expensive"));
  EXPECT_EQ(m.queries()[0].lines()[0].values()[0].json_text(), std::string("\"10\""));
  EXPECT_EQ(m.queries()[1].lines()[0].values()[0].json_text(), std::string("\"Herman Melville\""));
}

GPU_TEST(QueryExecutor, MaterializedResultsSkipUnmatchedLines) {
  file::FilePartition p(0, 0, kBookstore.size(), kBookstore.data());
  p.load_to_device();
  auto built = handmade_indices(kBookstore);
  EngineOptions o{};
  query::QueryCompiler c(o);
  query::BatchCompiledQuery qs;
  qs.add(c.compile("$.store.magazine", o));
  const auto r = query::QueryExecutor(o).execute_batch(qs, p, built);
  EXPECT_EQ(r.queries()[0].lines().size(), 1u);
  EXPECT_TRUE(r.materialize().queries()[0].lines().empty());
}

GPU_TEST(QueryExecutor, RejectsQueriesDeeperThanTheIndex) {
  file::FilePartition p(0, 0, kBookstore.size(), kBookstore.data());
  p.load_to_device();
  file::CopiedFileReader reader{"x"};
  profiler::Profiler prof("t", false);
  RunContext ctx{prof};
  const auto built = index::CombinedIndexBuilder(reader).build(p, 1, {}, ctx);
  EngineOptions o{};
  query::BatchCompiledQuery qs;
  qs.add(query::QueryCompiler(o).compile("$.store.book", o));
  EXPECT_THROW(query::QueryExecutor(o).execute_batch(qs, p, built),
               error::query::QueryExecutionError);
}

// ---- engine (tests/engine_test.cpp:41-129) --------------------------------
GPU_TEST(Engine, QueriesPartitionedLdJsonFile) {
  TempFile f("engine.ldjson", "{\"name\":\"Ada\",\"scores\":[1,2],\"lang\":\"en\"}\n"
                              "{\"name\":\"Bob\",\"scores\":[3,4],\"lang\":\"fr\"}\n");
  for (auto reader : kReaders) {
    for (auto builder : {index::IndexBuilderType::UNCOMBINED, index::IndexBuilderType::COMBINED}) {
      EngineOptions o{{reader}, {builder, 43, 1, 64, 1, 64}, {512, 1024}};
      const auto r = Engine{}.query(f.path.string(), {"$.name", "$.scores[1]"}, o);
      EXPECT_EQ(r.queries().size(), 2u);
      EXPECT_EQ(r.queries()[0].query_text(), std::string("$.name"));
      EXPECT_EQ(r.queries()[0].lines().size(), 2u);
      EXPECT_EQ(r.queries()[0].lines()[0].values()[0].json_text(), std::string("\"Ada\""));
      EXPECT_EQ(r.queries()[0].lines()[1].values()[0].json_text(), std::string("\"Bob\""));
      EXPECT_EQ(r.queries()[1].lines()[0].values()[0].json_text(), std::string("2"));
      EXPECT_EQ(r.queries()[1].lines()[1].values()[0].json_text(), std::string("4"));
    }
  }
}

int main(int argc, char **argv) {
  bool want_cpu = true, want_gpu = true;
  std::string filter;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--cpu") want_gpu = false;
    else if (a == "--gpu") want_cpu = false;
    else filter = a;
  }
  if (want_gpu && !device::device_available()) {
    if (!want_cpu) {
      std::cerr << "no HIP device for --gpu tests\n";
      return 2;
    }
    want_gpu = false;
  }
  int run = 0, failed = 0;
  for (const auto &c : minitest::registry()) {
    if ((c.gpu && !want_gpu) || (!c.gpu && !want_cpu)) continue;
    if (!filter.empty() && c.name.find(filter) == std::string::npos) continue;
    ++run;
    try {
      c.fn();
      std::cout << "[ OK ] " << c.name << "\n";
    } catch (const std::exception &e) {
      ++failed;
      std::cout << "[FAIL] " << c.name << ": " << e.what() << "\n";
    }
    std::cout.flush();
  }
  std::cout << run - failed << "/" << run << " passed\n";
  return failed ? 1 : (run ? 0 : 3);
}
