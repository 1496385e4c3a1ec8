#pragma once

// Request server over the MI355X engine (no reference counterpart: the
// reference engine runs one file at a time and rebuilds its indices on every
// call, src/gpjson/engine.cpp:131-166).
//
// A request is (document, 1..8 compiled queries).  Requests become visible at
// their arrival times; each waits in FIFO order for a free slot (a HIP stream
// of its own) and for room in the HBM budget, then runs stream-ordered on its
// slot: pinned host-to-device copy of the document, index build (and, for
// element queries, the bounds of the top-level array's elements), one batched
// query launch with dense results, copy-back of the results.  Work of different slots overlaps on
// the device.  Documents and their indices stay resident after their request
// (least recently used first out when the budget needs room), so a request
// for a resident document runs the query launch only.  A request's results do
// not depend on what else runs: run_alone() serves one request with nothing
// shared, and the two agree bit for bit.

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace gpjson::serve {

// One compiled query of the server's query table.  An element query runs
// over the elements of a document whose top-level value is an array (its
// text is written `$[*]...` and compiled without the leading `$[*]`); the
// others run over the lines of an NDJSON document.
struct QuerySpec {
  std::string text;
  std::vector<uint8_t> ir;
  int max_depth = 0;
  int slots = 1;
  bool variable = false;
  bool elements = false;
};

// Compiles JSONPath text (extension syntax allowed) into a QuerySpec; text
// that starts with `$[*]` becomes an element query.
QuerySpec compile_query_spec(const std::string &text);

struct Request {
  uint64_t id = 0;
  uint64_t doc_id = 0;          // equal ids name the same document
  const uint8_t *doc = nullptr; // host bytes; pinned memory lets copies overlap
  uint64_t doc_bytes = 0;
  std::vector<int> queries;     // indices into the query table, 1..8 of them,
                                // all element queries or none
  double arrival_s = 0;         // seconds after the start of the run
};

// Results of one query over one document, dense: record r's {start, end}
// spans (document offsets) are spans[2 * offsets[r] .. 2 * offsets[r + 1]),
// offsets has records + 1 entries.
struct QueryOutput {
  std::vector<int64_t> offsets;
  std::vector<int64_t> spans;
};

struct RequestResult {
  uint64_t id = 0;
  double arrival_s = 0; // when it became visible
  double start_s = 0;   // when its work was enqueued
  double done_s = 0;    // when its results were on the host
  bool resident = false;
  int64_t num_records = 0; // lines, or elements for element queries
  std::vector<QueryOutput> outputs; // one per requested query, in order
};

struct ServerOptions {
  uint64_t hbm_budget_bytes = 80ull << 30;
  int slots = 8;            // requests in flight at once
  int spans_per_record = 1; // first guess of an extension query's spans per
                            // record (the output is sized from it and from
                            // the most seen since; a larger total reruns the
                            // request with the exact size)
  bool keep_resident = true;
  int index_levels = 0;     // 0: the deepest the table's queries need
  // First guesses of a document's line and array-element sizes (the
  // newline index and the element bounds are sized from them, and later from
  // the densest document seen; a document with more lines or elements is
  // rebuilt with the exact count).
  uint64_t bytes_per_line = 1024;
  uint64_t bytes_per_element = 512;
};

struct ServerStats {
  uint64_t admitted = 0;
  uint64_t resident_hits = 0;
  uint64_t evictions = 0;
  uint64_t retries = 0;     // a capacity guess (lines, elements, spans) was
                            // exceeded: the request reran with exact sizes
  uint64_t peak_bytes = 0;  // largest budget use seen
  uint64_t device_bytes = 0; // budget use at the end of the run
};

class Server {
public:
  Server(std::vector<QuerySpec> queries, ServerOptions options);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  // Replays `requests` against the wall clock (arrival_s relative to the
  // call) and returns their results in the order given.
  std::vector<RequestResult> run(const std::vector<Request> &requests);

  // Serves one request with nothing shared and nothing kept.
  RequestResult run_alone(const Request &request);

  // Drops every resident document.
  void clear_resident();

  const ServerStats &stats() const;

private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

} // namespace gpjson::serve
