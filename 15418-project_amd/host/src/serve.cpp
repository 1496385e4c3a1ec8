// Request server: slot admission under an HBM budget, resident documents
// and indices, overlapping per-slot streams, dense results.  See
// gpjson/serve/server.hpp.
#include "gpjson/serve/server.hpp"

#include "gpj_capi.h"
#include "gpjson/device/device.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>

namespace gpjson::serve {

namespace {

void ok(gpj_status s, const char *what) {
  if (s < 0) throw std::runtime_error(std::string(what) + ": " + gpj_last_error());
}

uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
uint64_t round256(uint64_t a) { return div_up(a, 256) * 256; }

// Capacities assumed before a count is known: `density` items per byte
// (the densest document served so far, or the configured first guess) with a
// quarter of slack.  A document with more lines or elements is rebuilt with
// its exact count (counted as a retry).
uint64_t density_hint(uint64_t n, double density, uint64_t extra) {
  return (uint64_t)((double)n * density * 1.25) + extra;
}

// True when the bytecode holds an extension opcode (MOVE_TO_EACH or
// JUMP_IF_NOT_ARRAY), i.e. it runs in variable-length mode.
bool uses_extension_opcodes(const std::vector<uint8_t> &ir) {
  size_t p = 0;
  auto varint = [&]() {
    uint32_t v = 0, shift = 0;
    while (p < ir.size()) {
      const uint8_t b = ir[p++];
      v |= (uint32_t)(b & 0x7F) << shift;
      if (!(b & 0x80)) break;
      shift += 7;
    }
    return v;
  };
  while (p < ir.size()) {
    const uint8_t op = ir[p++];
    if (op == 10 || op == 11) return true;
    if (op == 4 || op == 9) p += varint();
    else if (op == 5 || op == 6) (void)varint();
  }
  return false;
}

// A line count of -1 marks an index build whose bounded carry wait gave up
// (gpj_index_build_into); its indices and results are not valid.
void check_lines(int64_t lines) {
  if (lines < 0)
    throw error::device::DeviceError("index build: a tile's wait for its carry timed out");
}

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

QuerySpec compile_query_spec(const std::string &text) {
  QuerySpec q;
  q.text = text;
  std::string src = text;
  const size_t lead = text.find_first_not_of(" \t");
  if (lead != std::string::npos && text.compare(lead, 4, "$[*]") == 0) {
    q.elements = true; // runs as `$<rest>` on each element of a top-level array
    src = "$" + text.substr(lead + 4);
  }
  std::vector<uint8_t> buf(GPJ_MAX_QUERY_IR_BYTES * 4);
  uint32_t len = 0, depth = 0, slots = 0;
  ok(gpj_query_compile(src.c_str(), buf.data(), (uint32_t)buf.size(), &len, &depth, &slots,
                       GPJ_COMPILE_EXTENSIONS),
     "query compile");
  q.ir.assign(buf.begin(), buf.begin() + len);
  q.max_depth = (int)depth;
  q.slots = (int)slots;
  q.variable = uses_extension_opcodes(q.ir);
  return q;
}

struct Server::Impl {
  struct Pinned {
    uint8_t *p = nullptr;
    uint64_t bytes = 0;
  };
  // A document on the device with its index (one allocation) and, once an
  // element query needs them, the bounds of its top-level array's elements.
  struct Doc {
    uint64_t id = 0, bytes = 0, words = 0, nl_cap = 0, footprint = 0;
    uint8_t *block = nullptr;
    uint8_t *d_in = nullptr;
    uint64_t *d_string = nullptr, *d_leveled = nullptr;
    int64_t *d_newline = nullptr, *d_lines = nullptr;
    void *ready = nullptr; // recorded after the index build
    uint8_t *bounds_block = nullptr;
    int64_t *d_bounds = nullptr, *d_nbounds = nullptr;
    uint64_t bounds_cap = 0, bounds_bytes = 0;
    void *bounds_ready = nullptr;
    int users = 0;
    uint64_t last_use = 0;
  };
  // Layout of one request's device results: per query, CSR offsets
  // [max_records + 1] and spans [caps[q]][2]; then the dense scratch.
  struct Layout {
    int64_t max_records = 0;
    std::vector<uint64_t> caps;
    uint64_t outputs = 0; // bytes of the offsets and spans (they come first)
    uint64_t bytes = 0;
  };
  struct Slot {
    void *stream = nullptr, *ev_a = nullptr, *ev_b = nullptr;
    void *scratch = nullptr; // index scratch
    uint64_t scratch_bytes = 0;
    int stage = 0; // 0 idle, 1 computing, 2 copying results back
    size_t req = 0;
    Doc *doc = nullptr;
    uint8_t *d_results = nullptr; // kept across requests, grown when needed
    uint64_t results_cap = 0;
    Layout lay;
    uint64_t reserved = 0; // budget reserved for d_results and new bounds
    int64_t *h_small = nullptr; // pinned: line count, bound count, totals[8]
    int64_t records = 0;
    Pinned h_results;
    double start_s = 0;
    bool resident = false;
  };

  std::vector<QuerySpec> queries;
  std::vector<void *> d_irs;
  ServerOptions opt;
  int levels = 1;
  ServerStats stats;
  std::unordered_map<uint64_t, Doc *> docs;
  std::unordered_map<uint64_t, uint64_t> exact_caps;   // doc id -> line count
  std::unordered_map<uint64_t, uint64_t> exact_bounds; // doc id -> bound count
  std::map<std::pair<uint64_t, int>, uint64_t> exact_spans; // (doc, query) -> spans
  std::vector<Slot> slots;
  uint64_t used = 0, tick = 0;
  // learned sizes: lines and array elements per document byte, spans per
  // record of each query (the largest seen; they only size first guesses)
  double line_density = 0, element_density = 0;
  std::vector<double> spans_ratio;

  // Pinned result buffers: a completed request's buffer goes to a worker
  // that unpacks it into the caller's vectors and then back to the pool, so
  // the scheduler thread only enqueues work and polls events.
  std::mutex pool_mu;
  std::vector<Pinned> pool;
  std::mutex task_mu;
  std::condition_variable task_cv, idle_cv;
  std::deque<std::function<void()>> tasks;
  int running = 0;
  bool stopping = false;
  std::vector<std::thread> workers;

  Pinned take_pinned(uint64_t bytes) {
    {
      // the smallest pooled buffer that holds `bytes`
      std::lock_guard<std::mutex> g(pool_mu);
      size_t best = pool.size();
      for (size_t i = 0; i < pool.size(); ++i)
        if (pool[i].bytes >= bytes && (best == pool.size() || pool[i].bytes < pool[best].bytes))
          best = i;
      if (best < pool.size()) {
        Pinned b = pool[best];
        pool.erase(pool.begin() + (long)best);
        return b;
      }
    }
    Pinned b;
    b.bytes = 1 << 20; // powers of two: few distinct buffers in the pool
    while (b.bytes < bytes) b.bytes <<= 1;
    ok(gpj_host_alloc(reinterpret_cast<void **>(&b.p), b.bytes), "pinned alloc");
    return b;
  }
  void give_pinned(Pinned b) {
    std::lock_guard<std::mutex> g(pool_mu);
    pool.push_back(b);
  }
  void post(std::function<void()> fn) {
    {
      std::lock_guard<std::mutex> g(task_mu);
      tasks.push_back(std::move(fn));
    }
    task_cv.notify_one();
  }
  void drain() {
    std::unique_lock<std::mutex> g(task_mu);
    idle_cv.wait(g, [&] { return tasks.empty() && running == 0; });
  }
  void worker_loop() {
    for (;;) {
      std::function<void()> fn;
      {
        std::unique_lock<std::mutex> g(task_mu);
        task_cv.wait(g, [&] { return stopping || !tasks.empty(); });
        if (tasks.empty()) return;
        fn = std::move(tasks.front());
        tasks.pop_front();
        ++running;
      }
      fn();
      {
        std::lock_guard<std::mutex> g(task_mu);
        --running;
      }
      idle_cv.notify_all();
    }
  }

  Impl(std::vector<QuerySpec> q, ServerOptions o) : queries(std::move(q)), opt(o) {
    if (opt.slots < 1) opt.slots = 1;
    if (opt.spans_per_record < 1) opt.spans_per_record = 1;
    line_density = 1.0 / (double)std::max<uint64_t>(opt.bytes_per_line, 1);
    element_density = 1.0 / (double)std::max<uint64_t>(opt.bytes_per_element, 1);
    spans_ratio.assign(queries.size(), (double)opt.spans_per_record);
    int deepest = 1;
    for (const auto &qs : queries) deepest = std::max(deepest, qs.max_depth + (qs.elements ? 1 : 0));
    levels = opt.index_levels > 0 ? opt.index_levels : deepest;
    for (const auto &qs : queries) {
      void *d = nullptr;
      ok(gpj_alloc(&d, qs.ir.size()), "IR alloc");
      ok(gpj_h2d(d, qs.ir.data(), qs.ir.size(), nullptr), "IR upload");
      d_irs.push_back(d);
    }
    ok(gpj_stream_sync(nullptr), "IR upload");
    slots.resize(opt.slots);
    for (auto &s : slots) {
      ok(gpj_stream_create(&s.stream), "stream");
      ok(gpj_event_create(&s.ev_a), "event");
      ok(gpj_event_create(&s.ev_b), "event");
      ok(gpj_host_alloc(reinterpret_cast<void **>(&s.h_small), 8 * (2 + GPJ_MAX_BATCH_QUERIES)),
         "pinned alloc");
    }
    // page-locked result buffers for the slots' first requests (allocating
    // them later would stall the scheduler thread)
    std::vector<Pinned> warm;
    for (int i = 0; i < 2 * opt.slots; ++i) warm.push_back(take_pinned(4u << 20));
    for (auto &b : warm) give_pinned(b);
    for (int i = 0; i < 4; ++i) workers.emplace_back([this] { worker_loop(); });
  }

  ~Impl() {
    {
      std::lock_guard<std::mutex> g(task_mu);
      stopping = true;
    }
    task_cv.notify_all();
    for (auto &w : workers) w.join();
    for (auto &b : pool) (void)gpj_host_free(b.p);
    (void)gpj_device_sync();
    for (auto &kv : docs) release_doc(kv.second, nullptr);
    docs.clear();
    for (auto &s : slots) {
      (void)gpj_free(s.scratch);
      (void)gpj_free(s.d_results);
      (void)gpj_host_free(s.h_small);
      (void)gpj_event_destroy(s.ev_a);
      (void)gpj_event_destroy(s.ev_b);
      (void)gpj_stream_destroy(s.stream);
    }
    for (void *d : d_irs) (void)gpj_free(d);
    (void)gpj_device_sync();
  }

  bool elements_of(const Request &r) const {
    bool el = false, line = false;
    for (int qi : r.queries) (queries[qi].elements ? el : line) = true;
    if (el && line)
      throw std::invalid_argument("a request mixes element and line queries");
    return el;
  }
  uint64_t cap_of(const Request &r) const {
    auto it = exact_caps.find(r.doc_id);
    return it != exact_caps.end() ? (uint64_t)it->second
                                  : density_hint(r.doc_bytes, line_density, 2);
  }
  uint64_t bounds_cap_of(const Request &r) const {
    auto it = exact_bounds.find(r.doc_id);
    return it != exact_bounds.end() ? std::max<uint64_t>(it->second, 1)
                                    : density_hint(r.doc_bytes, element_density, 3);
  }
  uint64_t doc_footprint(uint64_t bytes, uint64_t nl_cap) const {
    const uint64_t w = div_up(bytes, 64);
    return round256(bytes) + round256(w * 8) + round256(w * 8 * (uint64_t)levels) +
           round256(nl_cap * 8) + 256;
  }
  uint64_t bounds_footprint(uint64_t doc_bytes, uint64_t cap) const {
    return round256(cap * 8) + 256 +
           round256(gpj_row_bounds_scratch_bytes((int64_t)div_up(doc_bytes, 64)));
  }
  // The results layout of request r over max_records records.
  Layout layout(const Request &r, int64_t max_records) const {
    Layout l;
    l.max_records = max_records;
    uint64_t total = 0;
    for (int qi : r.queries) {
      const QuerySpec &q = queries[qi];
      uint64_t cap;
      auto it = exact_spans.find({r.doc_id, qi});
      if (it != exact_spans.end()) cap = it->second;
      else if (!q.variable) cap = (uint64_t)max_records * (uint64_t)q.slots; // at most slots each
      else cap = (uint64_t)((double)max_records * spans_ratio[qi] * 1.25) + 16;
      cap = std::max<uint64_t>(cap, 1);
      l.caps.push_back(cap);
      l.bytes += round256(((uint64_t)max_records + 1) * 8) + round256(cap * 16);
      total += cap;
    }
    l.outputs = l.bytes;
    l.bytes += round256(gpj_query_dense_scratch_bytes(max_records, (int32_t)r.queries.size(),
                                                      (int64_t)total));
    l.bytes += 256; // the gathered counts (last 256 bytes)
    return l;
  }

  void release_bounds(Doc *d, void *stream) {
    if (d->bounds_block) (void)gpj_free_async(d->bounds_block, stream);
    if (d->bounds_ready) (void)gpj_event_destroy(d->bounds_ready);
    used -= d->bounds_bytes;
    d->bounds_block = nullptr;
    d->d_bounds = d->d_nbounds = nullptr;
    d->bounds_ready = nullptr;
    d->bounds_bytes = d->bounds_cap = 0;
  }
  void release_doc(Doc *d, void *stream) {
    release_bounds(d, stream);
    if (d->block) (void)gpj_free_async(d->block, stream);
    if (d->ready) (void)gpj_event_destroy(d->ready);
    used -= d->footprint;
    delete d;
  }

  // Frees least recently used idle documents until `need` more bytes fit.
  bool make_room(uint64_t need, void *stream) {
    while (used + need > opt.hbm_budget_bytes) {
      Doc *victim = nullptr;
      for (auto &kv : docs)
        if (kv.second->users == 0 && (!victim || kv.second->last_use < victim->last_use))
          victim = kv.second;
      if (!victim) return false;
      docs.erase(victim->id);
      release_doc(victim, stream);
      ++stats.evictions;
    }
    return true;
  }

  // Enqueues on `stream`: the document upload and its index build.
  Doc *build_doc(const Request &r, void *stream, void **scratch, uint64_t *scratch_bytes) {
    Doc *d = new Doc;
    d->id = r.doc_id;
    d->bytes = r.doc_bytes;
    d->words = div_up(r.doc_bytes, 64);
    d->nl_cap = cap_of(r);
    d->footprint = doc_footprint(r.doc_bytes, d->nl_cap);
    ok(gpj_alloc_async(reinterpret_cast<void **>(&d->block), d->footprint, stream), "doc alloc");
    uint8_t *p = d->block;
    d->d_in = p;
    p += round256(d->bytes);
    d->d_string = reinterpret_cast<uint64_t *>(p);
    p += round256(d->words * 8);
    d->d_leveled = reinterpret_cast<uint64_t *>(p);
    p += round256(d->words * 8 * (uint64_t)levels);
    d->d_newline = reinterpret_cast<int64_t *>(p);
    p += round256(d->nl_cap * 8);
    d->d_lines = reinterpret_cast<int64_t *>(p);
    used += d->footprint;
    ok(gpj_h2d(d->d_in, r.doc, r.doc_bytes, stream), "doc upload");
    // index scratch of the slot: sized for 1 GiB documents at first (64 KiB
    // per 4 GiB), grown stream-ordered for larger ones
    const uint64_t sb = std::max<uint64_t>(gpj_index_scratch_bytes(r.doc_bytes),
                                           gpj_index_scratch_bytes(1ull << 30));
    if (sb > *scratch_bytes) {
      if (*scratch) ok(gpj_free_async(*scratch, stream), "scratch free");
      ok(gpj_alloc_async(scratch, sb, stream), "scratch alloc");
      *scratch_bytes = sb;
    }
    ok(gpj_index_build_into(d->d_in, d->bytes, levels, d->d_string, d->d_leveled, d->d_newline,
                            (int64_t)d->nl_cap, d->d_lines, *scratch, stream),
       "index build");
    ok(gpj_event_create(&d->ready), "event");
    ok(gpj_event_record(d->ready, stream), "event record");
    return d;
  }

  gpj_index index_view(const Doc *d) const {
    return gpj_index{d->d_newline, d->d_lines, (int64_t)d->nl_cap, d->d_string, d->d_leveled,
                     (int64_t)d->words, levels, 0};
  }

  // Enqueues on `stream` (after the index): the bounds of the document's
  // top-level array elements, cap entries at most.
  void build_bounds(Doc *d, uint64_t cap, void *stream) {
    d->bounds_cap = cap;
    d->bounds_bytes = bounds_footprint(d->bytes, cap);
    ok(gpj_alloc_async(reinterpret_cast<void **>(&d->bounds_block), d->bounds_bytes, stream),
       "bounds alloc");
    used += d->bounds_bytes;
    uint8_t *p = d->bounds_block;
    d->d_bounds = reinterpret_cast<int64_t *>(p);
    p += round256(cap * 8);
    d->d_nbounds = reinterpret_cast<int64_t *>(p);
    p += 256;
    const gpj_index idx = index_view(d);
    ok(gpj_row_bounds(&idx, d->d_bounds, (int64_t)cap, d->d_nbounds, p, stream), "row bounds");
    ok(gpj_event_create(&d->bounds_ready), "event");
    ok(gpj_event_record(d->bounds_ready, stream), "event record");
  }

  // Enqueues the dense query launch of request r over doc d into d_results,
  // then the copies of the line count, the bound count and each query's total.
  void enqueue_queries(const Request &r, Doc *d, const Layout &lay, uint8_t *d_results,
                       int64_t *h_small, void *stream) {
    std::vector<gpj_query_desc> descs;
    std::vector<gpj_dense_result> outs;
    uint8_t *p = d_results;
    for (size_t k = 0; k < r.queries.size(); ++k) {
      const QuerySpec &q = queries[r.queries[k]];
      gpj_query_desc desc{};
      desc.d_ir = static_cast<const uint8_t *>(d_irs[r.queries[k]]);
      desc.ir_len = (uint32_t)q.ir.size();
      desc.capacity = q.variable ? 1 : q.slots;
      desc.variable = q.variable ? 1 : 0;
      desc.max_level = q.max_depth;
      descs.push_back(desc);
      gpj_dense_result o{};
      o.d_offsets = reinterpret_cast<int64_t *>(p);
      p += round256(((uint64_t)lay.max_records + 1) * 8);
      o.d_spans = reinterpret_cast<int64_t *>(p);
      p += round256(lay.caps[k] * 16);
      o.capacity = (int64_t)lay.caps[k];
      outs.push_back(o);
    }
    const gpj_index idx = index_view(d);
    if (elements_of(r))
      ok(gpj_query_run_records_dense(d->d_in, d->bytes, &idx, d->d_bounds, d->d_nbounds,
                                     descs.data(), (int32_t)descs.size(), lay.max_records,
                                     outs.data(), p, stream),
         "element query launch");
    else
      ok(gpj_query_run_dense(d->d_in, d->bytes, &idx, descs.data(), (int32_t)descs.size(),
                             lay.max_records, outs.data(), p, stream),
         "query launch");
    // line count, bound count and each query's total in one copy
    const uint64_t *srcs[2 + GPJ_MAX_BATCH_QUERIES] = {};
    srcs[0] = reinterpret_cast<const uint64_t *>(d->d_lines);
    srcs[1] = reinterpret_cast<const uint64_t *>(d->d_nbounds);
    for (size_t k = 0; k < outs.size(); ++k)
      srcs[2 + k] = reinterpret_cast<const uint64_t *>(outs[k].d_offsets + lay.max_records);
    uint64_t *d_small = reinterpret_cast<uint64_t *>(d_results + lay.bytes - 256);
    ok(gpj_gather_words(d_small, srcs, (int32_t)(2 + outs.size()), stream), "count gather");
    ok(gpj_d2h_async(h_small, d_small, 8 * (2 + outs.size()), stream), "count copy");
  }

  // Reads the small copies of a finished query launch: the record count, or
  // -1 when a capacity guess was too small (the exact size is noted for the
  // rerun).  A failed index build throws.
  int64_t settle(const Request &r, const Doc *d, const Layout &lay, const int64_t *h_small,
                 bool *drop_doc) {
    const int64_t lines = h_small[0];
    check_lines(lines);
    *drop_doc = false;
    // the next documents' first guesses follow the densest seen so far
    line_density = std::max(line_density, (double)lines / (double)r.doc_bytes);
    if (elements_of(r))
      element_density = std::max(element_density, (double)h_small[1] / (double)r.doc_bytes);
    if ((uint64_t)lines > d->nl_cap) { // rebuild the index with the exact count
      exact_caps[r.doc_id] = (uint64_t)lines;
      *drop_doc = true;
      return -1;
    }
    int64_t records = lines;
    if (elements_of(r)) {
      const int64_t nb = h_small[1];
      if ((uint64_t)nb > d->bounds_cap) { // rebuild the bounds with the exact count
        exact_bounds[r.doc_id] = (uint64_t)nb;
        *drop_doc = true;
        return -1;
      }
      records = std::max<int64_t>(nb - 1, 0);
    }
    bool fits = true;
    for (size_t k = 0; k < r.queries.size(); ++k)
      if ((uint64_t)h_small[2 + k] > lay.caps[k]) {
        exact_spans[{r.doc_id, r.queries[k]}] = (uint64_t)h_small[2 + k];
        fits = false;
      }
    for (size_t k = 0; k < r.queries.size(); ++k)
      if (records > 0 && queries[r.queries[k]].variable)
        spans_ratio[r.queries[k]] = std::max(spans_ratio[r.queries[k]],
                                             (double)h_small[2 + k] / (double)records);
    return fits ? records : -1;
  }

  // One copy brings every query's outputs (offsets and spans, as laid out
  // on the device) to the host.
  void enqueue_copy_back(const Layout &lay, const uint8_t *d_results, uint8_t *h, void *stream) {
    ok(gpj_d2h_async(h, d_results, lay.outputs, stream), "results copy");
  }

  void unpack(const Request &r, const Layout &lay, int64_t records, const int64_t *totals,
              const uint8_t *h, RequestResult &out) const {
    out.num_records = records;
    out.outputs.clear();
    for (size_t k = 0; k < r.queries.size(); ++k) {
      QueryOutput o;
      o.offsets.resize((size_t)records + 1);
      o.spans.resize((size_t)totals[k] * 2);
      std::memcpy(o.offsets.data(), h, o.offsets.size() * 8);
      h += round256(((uint64_t)lay.max_records + 1) * 8);
      std::memcpy(o.spans.data(), h, o.spans.size() * 8);
      h += round256(lay.caps[k] * 16);
      out.outputs.push_back(std::move(o));
    }
  }

  // The slot's device results buffer, grown to `bytes` when smaller (the
  // slot is idle: its stream has no work pending on the old buffer).
  void slot_results(Slot &s, uint64_t bytes) {
    if (s.results_cap >= bytes) return;
    if (s.d_results) (void)gpj_free_async(s.d_results, s.stream);
    const uint64_t grown = std::max<uint64_t>(bytes, s.results_cap + s.results_cap / 2);
    ok(gpj_alloc_async(reinterpret_cast<void **>(&s.d_results), grown, s.stream),
       "results alloc");
    s.results_cap = grown;
  }

  std::vector<RequestResult> run(const std::vector<Request> &reqs) {
    std::vector<RequestResult> out(reqs.size());
    std::vector<size_t> order(reqs.size());
    for (size_t i = 0; i < reqs.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    for (size_t i = 0; i < reqs.size(); ++i) {
      out[i].id = reqs[i].id;
      out[i].arrival_s = reqs[i].arrival_s;
      if (reqs[i].queries.empty() || reqs[i].queries.size() > GPJ_MAX_BATCH_QUERIES ||
          reqs[i].doc_bytes == 0)
        throw std::invalid_argument("request without 1..8 queries or document bytes");
      (void)elements_of(reqs[i]);
    }
    std::deque<size_t> waiting;
    size_t next = 0, done = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while (done < reqs.size()) {
      bool progress = false;
      const double t = now_s(t0);
      while (next < order.size() && reqs[order[next]].arrival_s <= t) {
        waiting.push_back(order[next++]);
        progress = true;
      }
      // admit in arrival order while a slot is free and the budget has room
      while (!waiting.empty()) {
        Slot *s = nullptr;
        for (auto &sl : slots)
          if (sl.stage == 0) {
            s = &sl;
            break;
          }
        if (!s) break;
        const Request &r = reqs[waiting.front()];
        const bool el = elements_of(r);
        auto it = docs.find(r.doc_id);
        Doc *d = it != docs.end() ? it->second : nullptr;
        const uint64_t nl_cap = d ? d->nl_cap : cap_of(r);
        const bool new_bounds = el && (!d || !d->d_bounds);
        const uint64_t b_cap = el ? (new_bounds ? bounds_cap_of(r) : d->bounds_cap) : 0;
        const int64_t max_records = el ? (int64_t)b_cap - 1 : (int64_t)nl_cap;
        Layout lay = layout(r, max_records);
        const uint64_t extra = lay.bytes + (new_bounds ? bounds_footprint(r.doc_bytes, b_cap) : 0);
        const uint64_t need = extra + (d ? 0 : doc_footprint(r.doc_bytes, nl_cap));
        if (d) ++d->users; // pin before making room
        const bool fits = make_room(need, s->stream);
        if (!fits) {
          if (d) --d->users;
          bool busy = false;
          for (auto &sl : slots) busy |= sl.stage != 0;
          if (!busy) throw std::runtime_error("request does not fit the HBM budget");
          break;
        }
        waiting.pop_front();
        s->req = &r - reqs.data();
        s->resident = d != nullptr;
        if (d) {
          ++stats.resident_hits;
          ok(gpj_stream_wait_event(s->stream, d->ready), "wait for index");
          if (d->bounds_ready) ok(gpj_stream_wait_event(s->stream, d->bounds_ready), "wait");
        } else {
          d = build_doc(r, s->stream, &s->scratch, &s->scratch_bytes);
          d->users = 1;
          if (opt.keep_resident) docs[d->id] = d;
        }
        if (new_bounds) build_bounds(d, b_cap, s->stream);
        s->doc = d;
        d->last_use = ++tick;
        slot_results(*s, lay.bytes);
        s->lay = std::move(lay);
        s->reserved = s->lay.bytes;
        used += s->reserved;
        stats.peak_bytes = std::max(stats.peak_bytes, used);
        enqueue_queries(r, d, s->lay, s->d_results, s->h_small, s->stream);
        ok(gpj_event_record(s->ev_a, s->stream), "event record");
        s->stage = 1;
        s->start_s = now_s(t0);
        out[s->req].start_s = s->start_s;
        ++stats.admitted;
        progress = true;
      }
      // completions
      for (auto &s : slots) {
        if (s.stage == 1) {
          const gpj_status q = gpj_event_query(s.ev_a);
          ok(q, "event query");
          if (q == 0) continue;
          progress = true;
          const Request &r = reqs[s.req];
          bool drop_doc = false;
          const int64_t records = settle(r, s.doc, s.lay, s.h_small, &drop_doc);
          if (records < 0) { // a capacity guess was too small: rerun with exact sizes
            ++stats.retries;
            finish_slot(s, drop_doc);
            waiting.push_front(&r - reqs.data());
            continue;
          }
          s.records = records;
          s.h_results = take_pinned(s.lay.outputs);
          enqueue_copy_back(s.lay, s.d_results, s.h_results.p, s.stream);
          ok(gpj_event_record(s.ev_b, s.stream), "event record");
          s.stage = 2;
        } else if (s.stage == 2) {
          const gpj_status q = gpj_event_query(s.ev_b);
          ok(q, "event query");
          if (q == 0) continue;
          progress = true;
          const Request &r = reqs[s.req];
          RequestResult &res = out[s.req];
          res.resident = s.resident;
          res.done_s = now_s(t0); // results are in host memory
          const int64_t records = s.records;
          const Pinned buf = s.h_results;
          std::vector<int64_t> totals(s.h_small + 2, s.h_small + 2 + r.queries.size());
          s.h_results = Pinned{};
          post([this, &r, &res, records, totals, buf, lay = s.lay] {
            unpack(r, lay, records, totals.data(), buf.p, res);
            give_pinned(buf);
          });
          finish_slot(s, /*drop_doc=*/false);
          ++done;
        }
      }
      if (!progress) std::this_thread::yield();
    }
    drain();
    stats.device_bytes = used;
    return out;
  }

  void finish_slot(Slot &s, bool drop_doc) {
    used -= s.reserved; // the buffer stays with the slot for its next request
    s.reserved = 0;
    Doc *d = s.doc;
    --d->users;
    if (drop_doc) {
      auto it = docs.find(d->id);
      if (it != docs.end() && it->second == d) docs.erase(it);
      if (d->users == 0) release_doc(d, s.stream);
    } else if (!opt.keep_resident && d->users == 0) {
      release_doc(d, s.stream);
    }
    s.doc = nullptr;
    s.stage = 0;
  }

  RequestResult run_alone(const Request &r) {
    RequestResult res;
    res.id = r.id;
    if (r.queries.empty() || r.queries.size() > GPJ_MAX_BATCH_QUERIES || r.doc_bytes == 0)
      throw std::invalid_argument("request without 1..8 queries or document bytes");
    const bool el = elements_of(r);
    void *stream = nullptr, *scratch = nullptr;
    uint64_t scratch_bytes = 0;
    ok(gpj_stream_create(&stream), "stream");
    // nothing shared: sizes start from the hints and are learned here
    std::unordered_map<uint64_t, uint64_t> caps_keep, bounds_keep;
    std::map<std::pair<uint64_t, int>, uint64_t> spans_keep;
    std::swap(caps_keep, exact_caps);
    std::swap(bounds_keep, exact_bounds);
    std::swap(spans_keep, exact_spans);
    int64_t failed_lines = 0;
    std::vector<int64_t> h_small(2 + GPJ_MAX_BATCH_QUERIES, 0);
    for (;;) {
      Doc *d = build_doc(r, stream, &scratch, &scratch_bytes);
      if (el) build_bounds(d, bounds_cap_of(r), stream);
      const Layout lay = layout(r, el ? (int64_t)d->bounds_cap - 1 : (int64_t)d->nl_cap);
      uint8_t *d_results = nullptr;
      ok(gpj_alloc_async(reinterpret_cast<void **>(&d_results), lay.bytes, stream), "alloc");
      std::fill(h_small.begin(), h_small.end(), 0);
      enqueue_queries(r, d, lay, d_results, h_small.data(), stream);
      ok(gpj_stream_sync(stream), "query");
      bool drop_doc = false;
      int64_t records = -1;
      if (h_small[0] < 0) failed_lines = h_small[0];
      else records = settle(r, d, lay, h_small.data(), &drop_doc);
      if (records >= 0) {
        std::vector<uint8_t> h(lay.outputs);
        enqueue_copy_back(lay, d_results, h.data(), stream);
        ok(gpj_stream_sync(stream), "copy back");
        unpack(r, lay, records, h_small.data() + 2, h.data(), res);
      }
      (void)gpj_free_async(d_results, stream);
      release_doc(d, stream);
      ok(gpj_stream_sync(stream), "release");
      if (records >= 0 || failed_lines < 0) break;
    }
    std::swap(caps_keep, exact_caps);
    std::swap(bounds_keep, exact_bounds);
    std::swap(spans_keep, exact_spans);
    (void)gpj_free(scratch);
    (void)gpj_stream_destroy(stream);
    check_lines(failed_lines);
    return res;
  }

  void clear_resident() {
    (void)gpj_device_sync();
    for (auto it = docs.begin(); it != docs.end();) {
      if (it->second->users == 0) {
        release_doc(it->second, nullptr);
        it = docs.erase(it);
      } else {
        ++it;
      }
    }
    (void)gpj_device_sync();
  }
};

Server::Server(std::vector<QuerySpec> queries, ServerOptions options)
    : impl_(std::make_unique<Impl>(std::move(queries), options)) {}
Server::~Server() = default;
std::vector<RequestResult> Server::run(const std::vector<Request> &requests) {
  return impl_->run(requests);
}
RequestResult Server::run_alone(const Request &request) { return impl_->run_alone(request); }
void Server::clear_resident() { impl_->clear_resident(); }
const ServerStats &Server::stats() const { return impl_->stats; }

} // namespace gpjson::serve
