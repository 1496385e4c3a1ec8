// Request server: slot admission under an HBM budget, resident documents,
// overlapping per-slot streams.  See gpjson/serve/server.hpp.
#include "gpjson/serve/server.hpp"

#include "gpj_capi.h"
#include "gpjson/device/device.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
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

// Newline-index capacity assumed before the count is known: records of 256
// bytes and more on average.  A document with more lines is rebuilt with its
// exact count (counted as a retry).
uint64_t newline_hint(uint64_t n) { return n / 256 + 2; }

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
  std::vector<uint8_t> buf(GPJ_MAX_QUERY_IR_BYTES * 4);
  uint32_t len = 0, depth = 0, slots = 0;
  ok(gpj_query_compile(text.c_str(), buf.data(), (uint32_t)buf.size(), &len, &depth, &slots,
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
  // A document on the device with its index (one allocation).
  struct Doc {
    uint64_t id = 0, bytes = 0, words = 0, nl_cap = 0, footprint = 0;
    uint8_t *block = nullptr;
    uint8_t *d_in = nullptr;
    uint64_t *d_string = nullptr, *d_leveled = nullptr;
    int64_t *d_newline = nullptr, *d_lines = nullptr;
    void *ready = nullptr; // recorded after the index build
    int users = 0;
    uint64_t last_use = 0;
  };
  struct Slot {
    void *stream = nullptr, *ev_a = nullptr, *ev_b = nullptr;
    void *scratch = nullptr;
    uint64_t scratch_bytes = 0;
    int stage = 0; // 0 idle, 1 computing, 2 copying results back
    size_t req = 0;
    Doc *doc = nullptr;
    uint8_t *d_results = nullptr;
    uint64_t results_bytes = 0; // budget reserved for d_results
    int64_t *h_lines = nullptr; // pinned
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
  std::unordered_map<uint64_t, uint64_t> exact_caps; // doc id -> line count
  std::vector<Slot> slots;
  uint64_t used = 0, tick = 0;

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
      std::lock_guard<std::mutex> g(pool_mu);
      for (size_t i = 0; i < pool.size(); ++i)
        if (pool[i].bytes >= bytes) {
          Pinned b = pool[i];
          pool.erase(pool.begin() + (long)i);
          return b;
        }
    }
    Pinned b;
    b.bytes = std::max<uint64_t>(bytes, 1 << 20);
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
    if (opt.result_capacity < 1) opt.result_capacity = 1;
    int deepest = 1;
    for (const auto &qs : queries) deepest = std::max(deepest, qs.max_depth);
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
      ok(gpj_host_alloc(reinterpret_cast<void **>(&s.h_lines), 8), "pinned alloc");
    }
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
      (void)gpj_host_free(s.h_lines);
      (void)gpj_event_destroy(s.ev_a);
      (void)gpj_event_destroy(s.ev_b);
      (void)gpj_stream_destroy(s.stream);
    }
    for (void *d : d_irs) (void)gpj_free(d);
    (void)gpj_device_sync();
  }

  uint64_t cap_of(const Request &r) const {
    auto it = exact_caps.find(r.doc_id);
    return it != exact_caps.end() ? (uint64_t)it->second : newline_hint(r.doc_bytes);
  }
  uint64_t doc_footprint(uint64_t bytes, uint64_t nl_cap) const {
    const uint64_t w = div_up(bytes, 64);
    return round256(bytes) + round256(w * 8) + round256(w * 8 * (uint64_t)levels) +
           round256(nl_cap * 8) + 256;
  }
  int cap_q(const QuerySpec &q) const { return q.variable ? opt.result_capacity : q.slots; }
  uint64_t results_footprint(const Request &r, uint64_t nl_cap) const {
    uint64_t b = 0;
    for (int qi : r.queries) b += round256(nl_cap * 16 * (uint64_t)cap_q(queries[qi])) +
                                  round256(nl_cap * 8);
    return b;
  }

  void release_doc(Doc *d, void *stream) {
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
    const uint64_t sb = gpj_index_scratch_bytes(r.doc_bytes);
    if (sb > *scratch_bytes) {
      ok(gpj_stream_sync(stream), "scratch");
      ok(gpj_free(*scratch), "scratch free");
      ok(gpj_alloc(scratch, sb), "scratch alloc");
      *scratch_bytes = sb;
    }
    ok(gpj_index_build_into(d->d_in, d->bytes, levels, d->d_string, d->d_leveled, d->d_newline,
                            (int64_t)d->nl_cap, d->d_lines, *scratch, stream),
       "index build");
    ok(gpj_event_create(&d->ready), "event");
    ok(gpj_event_record(d->ready, stream), "event record");
    return d;
  }

  // Enqueues the query launch of request r over doc d into d_results.
  void enqueue_queries(const Request &r, Doc *d, uint8_t *d_results, void *stream) {
    std::vector<gpj_query_desc> descs;
    uint8_t *p = d_results;
    for (int qi : r.queries) {
      const QuerySpec &q = queries[qi];
      const int cap = cap_q(q);
      gpj_query_desc desc{};
      desc.d_ir = static_cast<const uint8_t *>(d_irs[qi]);
      desc.ir_len = (uint32_t)q.ir.size();
      desc.capacity = cap;
      desc.variable = q.variable ? 1 : 0;
      desc.max_level = q.max_depth;
      desc.d_spans = reinterpret_cast<int64_t *>(p);
      p += round256(d->nl_cap * 16 * (uint64_t)cap);
      desc.d_counts = reinterpret_cast<int64_t *>(p);
      p += round256(d->nl_cap * 8);
      descs.push_back(desc);
    }
    gpj_index idx{d->d_newline, d->d_lines, (int64_t)d->nl_cap, d->d_string, d->d_leveled,
                  (int64_t)d->words, levels, 0};
    // up to 8 queries per launch
    for (size_t b = 0; b < descs.size(); b += GPJ_MAX_BATCH_QUERIES) {
      const int cnt = (int)std::min<size_t>(GPJ_MAX_BATCH_QUERIES, descs.size() - b);
      ok(gpj_query_run_batch(d->d_in, d->bytes, &idx, descs.data() + b, cnt,
                             (int64_t)d->nl_cap, stream),
         "query launch");
    }
  }

  uint64_t host_results_bytes(const Request &r, int64_t lines) const {
    uint64_t b = 0;
    for (int qi : r.queries) b += (uint64_t)lines * 16 * (uint64_t)cap_q(queries[qi]) +
                                  (uint64_t)lines * 8;
    return b;
  }

  // Enqueues the copy-back of the first `lines` rows of every query's results.
  void enqueue_copy_back(const Request &r, Doc *d, const uint8_t *d_results, int64_t lines,
                         uint8_t *h, void *stream) {
    const uint8_t *p = d_results;
    for (int qi : r.queries) {
      const int cap = cap_q(queries[qi]);
      const uint64_t sp = (uint64_t)lines * 16 * (uint64_t)cap, ct = (uint64_t)lines * 8;
      ok(gpj_d2h_async(h, p, sp, stream), "results copy");
      h += sp;
      p += round256(d->nl_cap * 16 * (uint64_t)cap);
      ok(gpj_d2h_async(h, p, ct, stream), "counts copy");
      h += ct;
      p += round256(d->nl_cap * 8);
    }
  }

  void unpack(const Request &r, int64_t lines, const uint8_t *h, RequestResult &out) const {
    out.num_lines = lines;
    out.outputs.clear();
    for (int qi : r.queries) {
      const int cap = cap_q(queries[qi]);
      QueryOutput o;
      o.spans.resize((size_t)lines * cap * 2);
      o.counts.resize((size_t)lines);
      std::memcpy(o.spans.data(), h, o.spans.size() * 8);
      h += o.spans.size() * 8;
      std::memcpy(o.counts.data(), h, o.counts.size() * 8);
      h += o.counts.size() * 8;
      if (queries[qi].variable) // entries past a line's count are unspecified
        for (int64_t l = 0; l < lines; ++l)
          for (int64_t k = std::max<int64_t>(0, o.counts[l]); k < cap; ++k)
            o.spans[(size_t)(l * cap + k) * 2] = o.spans[(size_t)(l * cap + k) * 2 + 1] = -1;
      out.outputs.push_back(std::move(o));
    }
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
      if (reqs[i].queries.empty() || reqs[i].doc_bytes == 0)
        throw std::invalid_argument("request without queries or document bytes");
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
        auto it = docs.find(r.doc_id);
        Doc *d = it != docs.end() ? it->second : nullptr;
        const uint64_t cap = d ? d->nl_cap : cap_of(r);
        const uint64_t res_bytes = results_footprint(r, cap);
        const uint64_t need = res_bytes + (d ? 0 : doc_footprint(r.doc_bytes, cap));
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
        } else {
          d = build_doc(r, s->stream, &s->scratch, &s->scratch_bytes);
          d->users = 1;
          if (opt.keep_resident) docs[d->id] = d;
        }
        s->doc = d;
        d->last_use = ++tick;
        ok(gpj_alloc_async(reinterpret_cast<void **>(&s->d_results), res_bytes, s->stream),
           "results alloc");
        s->results_bytes = res_bytes;
        used += res_bytes;
        stats.peak_bytes = std::max(stats.peak_bytes, used);
        enqueue_queries(r, d, s->d_results, s->stream);
        ok(gpj_d2h_async(s->h_lines, d->d_lines, 8, s->stream), "line count copy");
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
          const int64_t lines = *s.h_lines;
          check_lines(lines);
          if ((uint64_t)lines > s.doc->nl_cap) {
            // capacity hint too small: rebuild this document exactly, rerun
            ++stats.retries;
            exact_caps[r.doc_id] = (uint64_t)lines;
            finish_slot(s, /*drop_doc=*/true);
            waiting.push_front(&r - reqs.data());
            continue;
          }
          s.h_results = take_pinned(host_results_bytes(r, lines));
          enqueue_copy_back(r, s.doc, s.d_results, lines, s.h_results.p, s.stream);
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
          const int64_t lines = *s.h_lines;
          const Pinned buf = s.h_results;
          s.h_results = Pinned{};
          post([this, &r, &res, lines, buf] {
            unpack(r, lines, buf.p, res);
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
    (void)gpj_free_async(s.d_results, s.stream);
    used -= s.results_bytes;
    s.d_results = nullptr;
    s.results_bytes = 0;
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
    void *stream = nullptr, *scratch = nullptr;
    uint64_t scratch_bytes = 0;
    ok(gpj_stream_create(&stream), "stream");
    uint64_t cap = newline_hint(r.doc_bytes);
    int64_t counted = 0;
    for (;;) {
      Request rr = r;
      exact_caps[r.doc_id] = cap;
      Doc *d = build_doc(rr, stream, &scratch, &scratch_bytes);
      exact_caps.erase(r.doc_id);
      const uint64_t res_bytes = results_footprint(r, d->nl_cap);
      uint8_t *d_results = nullptr;
      ok(gpj_alloc_async(reinterpret_cast<void **>(&d_results), res_bytes, stream), "alloc");
      enqueue_queries(r, d, d_results, stream);
      int64_t lines = 0;
      ok(gpj_d2h(&lines, d->d_lines, 8, stream), "line count");
      counted = lines;
      if (lines >= 0 && (uint64_t)lines <= d->nl_cap) {
        std::vector<uint8_t> h(host_results_bytes(r, lines));
        enqueue_copy_back(r, d, d_results, lines, h.data(), stream);
        ok(gpj_stream_sync(stream), "copy back");
        unpack(r, lines, h.data(), res);
      }
      (void)gpj_free_async(d_results, stream);
      release_doc(d, stream);
      ok(gpj_stream_sync(stream), "release");
      if (lines < 0 || (uint64_t)lines <= cap) break;
      cap = (uint64_t)lines;
    }
    (void)gpj_free(scratch);
    (void)gpj_stream_destroy(stream);
    check_lines(counted);
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
