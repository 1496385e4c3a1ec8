// gpjson_serve: Poisson request replay against gpjson::serve::Server.
//
// Requests: documents of 4 KiB .. 64 MiB (log-uniform) from the generator,
// one of three record schemas with equal odds (NDJSON records of the 64 MiB
// plain configuration, one top-level array of the 1 GiB single-document
// configuration, NDJSON records of the 8 GiB throughput configuration), 1..8
// distinct queries each (at most the pool's size) drawn from the query sets
// of the configurations: the five line queries of the 64 MiB and 8 GiB
// configurations for NDJSON documents, the element query of the 1 GiB
// configuration for array documents, a share of
// requests (default 30%) re-querying an earlier document.  For every arrival
// rate it prints one JSON line: p50/p99 latency (arrival to results on the
// host), sustained GB/s, resident hits, evictions, retries, and whether every
// verified request's results equal the same request served alone (every
// request at every rate unless --verify limits it).  --digest-out writes the
// document manifest and one digest line per request and query, from which a
// checker can regenerate the documents and compare every result.
#include "gpjson/serve/server.hpp"

#include "gpj_capi.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" {
uint64_t gpj_gen_files(int32_t num_files, const int32_t *schemas, const uint64_t *targets,
                       uint64_t seed, int threads, uint64_t align, uint64_t *offsets_out,
                       uint64_t *sizes_out, uint8_t **buffer_out);
void gpj_gen_free(uint8_t *buffer);
}

namespace {

using gpjson::serve::Request;
using gpjson::serve::RequestResult;

struct Args {
  int requests = 1000;
  std::vector<double> rates{500, 1000, 2000, 5000};
  uint64_t seed = 0x15418004;
  uint64_t min_doc = 4096, max_doc = 64ull << 20;
  double requery = 0.3;
  double budget_gib = 80;
  int slots = 8;
  int spans_per_record = 1;
  uint64_t bytes_per_line = 1024, bytes_per_element = 512;
  int verify = -1; // requests per rate checked against run_alone (-1 = all)
  int threads = 16;
  std::string digest_out;
  int64_t wait_limit = -1; // index tiles' carry-wait polls (-1: default)
};

std::vector<double> parse_list(const std::string &s) {
  std::vector<double> v;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ',')) v.push_back(std::stod(tok));
  return v;
}

uint64_t parse_size(const std::string &s) {
  char *end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  std::string u(end);
  if (u == "KiB" || u == "K") v *= 1024;
  else if (u == "MiB" || u == "M") v *= 1024 * 1024;
  else if (u == "GiB" || u == "G") v *= 1024.0 * 1024 * 1024;
  return (uint64_t)v;
}

Args parse(int argc, char **argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    std::string k = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) throw std::invalid_argument("missing value for " + k);
      return argv[++i];
    };
    if (k == "--requests") a.requests = std::stoi(val());
    else if (k == "--rates") a.rates = parse_list(val());
    else if (k == "--seed") a.seed = std::stoull(val(), nullptr, 0);
    else if (k == "--min-doc") a.min_doc = parse_size(val());
    else if (k == "--max-doc") a.max_doc = parse_size(val());
    else if (k == "--requery") a.requery = std::stod(val());
    else if (k == "--budget-gib") a.budget_gib = std::stod(val());
    else if (k == "--slots") a.slots = std::stoi(val());
    else if (k == "--spans-per-record") a.spans_per_record = std::stoi(val());
    else if (k == "--digest-out") a.digest_out = val();
    else if (k == "--bytes-per-line") a.bytes_per_line = std::stoull(val());
    else if (k == "--bytes-per-element") a.bytes_per_element = std::stoull(val());
    else if (k == "--verify") a.verify = std::stoi(val());
    else if (k == "--threads") a.threads = std::stoi(val());
    else if (k == "--debug-wait-limit") a.wait_limit = std::stoll(val());
    else throw std::invalid_argument("unknown option " + k);
  }
  return a;
}

// query sets of the 64 MiB plain (schema 0), 1 GiB single-document (schema
// 1, element queries) and 8 GiB throughput (schema 2) configurations;
// indices into the server's table
const std::vector<std::string> kQueries = {"$.a.b", "$[*].items[*].sku", "$.user.lang",
                                           "$.e[0].f", "$.h[*].t", "$.h[?(@.k == \"v\")].id"};
// NDJSON documents (either schema) draw from the line queries of both
// NDJSON configurations (a query of the other configuration finds nothing
// there); array documents run the single-document configuration's query.
const std::vector<int> kLineQueries = {0, 2, 3, 4, 5};
const std::vector<int> kElementQueries = {1};
const uint64_t kAlign = 64;

// Digest of a result array: sum of value * odd weight (SplitMix64 of the
// index), modulo 2^64.
uint64_t weight(uint64_t i) {
  uint64_t z = i + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (z ^ (z >> 31)) | 1ull;
}
uint64_t digest(const std::vector<int64_t> &v) {
  uint64_t h = 0;
  for (size_t i = 0; i < v.size(); ++i) h += (uint64_t)v[i] * weight(i);
  return h;
}

template <typename T> std::string json_list(const std::vector<T> &v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + "]";
}

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const double r = p / 100.0 * (double)(v.size() - 1);
  const size_t lo = (size_t)r, hi = std::min(v.size() - 1, lo + 1);
  return v[lo] + (v[hi] - v[lo]) * (r - (double)lo);
}

bool same(const RequestResult &a, const RequestResult &b) {
  if (a.num_records != b.num_records || a.outputs.size() != b.outputs.size()) return false;
  for (size_t q = 0; q < a.outputs.size(); ++q)
    if (a.outputs[q].offsets != b.outputs[q].offsets || a.outputs[q].spans != b.outputs[q].spans)
      return false;
  return true;
}

} // namespace

int main(int argc, char **argv) {
  try {
    const Args a = parse(argc, argv);
    if (!gpj_device_available()) throw std::runtime_error("no HIP device");
    if (gpj_debug_set_wait_limit(a.wait_limit) < 0) throw std::runtime_error(gpj_last_error());
    std::vector<gpjson::serve::QuerySpec> table;
    for (const auto &q : kQueries) table.push_back(gpjson::serve::compile_query_spec(q));

    // the request shapes (documents and query sets) are shared by all rates
    std::mt19937_64 rng(a.seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    struct Shape {
      int doc;
      std::vector<int> queries;
    };
    std::vector<Shape> shapes;
    std::vector<int32_t> schemas;
    std::vector<uint64_t> targets;
    for (int i = 0; i < a.requests; ++i) {
      Shape s;
      if (!schemas.empty() && U(rng) < a.requery) {
        s.doc = (int)(U(rng) * (double)schemas.size()) % (int)schemas.size();
      } else {
        s.doc = (int)schemas.size();
        const double pick = U(rng);
        schemas.push_back(pick < 1.0 / 3 ? 0 : (pick < 2.0 / 3 ? 1 : 2));
        const double lg = std::log((double)a.min_doc) +
                          U(rng) * (std::log((double)a.max_doc) - std::log((double)a.min_doc));
        targets.push_back((uint64_t)std::exp(lg));
      }
      // 1..8 distinct queries (at most the pool's size), in random order
      std::vector<int> pool = schemas[s.doc] == 1 ? kElementQueries : kLineQueries;
      const int k = std::min<int>(1 + (int)(U(rng) * 8) % 8, (int)pool.size());
      for (int j = 0; j < k; ++j) {
        const size_t pick = j + (size_t)(U(rng) * (double)(pool.size() - j)) % (pool.size() - j);
        std::swap(pool[j], pool[pick]);
        s.queries.push_back(pool[j]);
      }
      shapes.push_back(std::move(s));
    }
    std::vector<uint64_t> off(schemas.size()), sz(schemas.size());
    uint8_t *gen = nullptr;
    const uint64_t total = gpj_gen_files((int32_t)schemas.size(), schemas.data(), targets.data(),
                                         a.seed, a.threads, kAlign, off.data(), sz.data(),
                                         &gen);
    if (!total) throw std::runtime_error("document generation failed");
    uint8_t *docs = nullptr; // pinned, so uploads overlap other slots' kernels
    if (gpj_host_alloc(reinterpret_cast<void **>(&docs), total) < 0)
      throw std::runtime_error(gpj_last_error());
    std::memcpy(docs, gen, total);
    gpj_gen_free(gen);
    uint64_t doc_bytes_total = 0;
    for (auto v : sz) doc_bytes_total += v;
    std::fprintf(stderr, "[serve] %d requests over %zu documents (%llu bytes)\n", a.requests,
                 schemas.size(), (unsigned long long)doc_bytes_total);
    FILE *dig = nullptr;
    if (!a.digest_out.empty()) {
      dig = std::fopen(a.digest_out.c_str(), "w");
      if (!dig) throw std::runtime_error("cannot write " + a.digest_out);
      std::string qs = "[";
      for (size_t i = 0; i < kQueries.size(); ++i) {
        std::string esc;
        for (char c : kQueries[i]) esc += (c == '"' ? std::string("\\\"") : std::string(1, c));
        qs += (i ? ",\"" : "\"") + esc + "\"";
      }
      qs += "]";
      std::fprintf(dig,
                   "{\"manifest\": {\"seed\": %llu, \"align\": %llu, \"levels\": %d, "
                   "\"queries\": %s, \"schemas\": %s, \"targets\": %s, \"offsets\": %s, "
                   "\"sizes\": %s}}\n",
                   (unsigned long long)a.seed, (unsigned long long)kAlign,
                   [&] {
                     int lv = 1;
                     for (const auto &q : table) lv = std::max(lv, q.max_depth + (q.elements ? 1 : 0));
                     return lv;
                   }(),
                   qs.c_str(), json_list(schemas).c_str(), json_list(targets).c_str(),
                   json_list(off).c_str(), json_list(sz).c_str());
    }

    for (size_t ri = 0; ri < a.rates.size(); ++ri) {
      const double rate = a.rates[ri];
      std::mt19937_64 arng(a.seed ^ (uint64_t)(rate * 1000));
      std::exponential_distribution<double> gap(rate);
      std::vector<Request> reqs(a.requests);
      double t = 0;
      uint64_t req_bytes = 0, new_bytes = 0;
      std::vector<char> seen(schemas.size(), 0);
      for (int i = 0; i < a.requests; ++i) {
        t += gap(arng);
        Request &r = reqs[i];
        r.id = (uint64_t)i;
        r.doc_id = (uint64_t)shapes[i].doc;
        r.doc = docs + off[shapes[i].doc];
        r.doc_bytes = sz[shapes[i].doc];
        r.queries = shapes[i].queries;
        r.arrival_s = t;
        req_bytes += r.doc_bytes;
        if (!seen[shapes[i].doc]) new_bytes += r.doc_bytes;
        seen[shapes[i].doc] = 1;
      }
      gpjson::serve::ServerOptions so;
      so.hbm_budget_bytes = (uint64_t)(a.budget_gib * (double)(1ull << 30));
      so.slots = a.slots;
      so.spans_per_record = a.spans_per_record;
      so.bytes_per_line = a.bytes_per_line;
      so.bytes_per_element = a.bytes_per_element;
      gpjson::serve::Server server(table, so);
      const std::vector<RequestResult> res = server.run(reqs);
      std::vector<double> lat;
      double first = 1e300, last = 0;
      uint64_t resident = 0;
      for (const auto &r : res) {
        lat.push_back((r.done_s - r.arrival_s) * 1e3);
        first = std::min(first, r.arrival_s);
        last = std::max(last, r.done_s);
        resident += r.resident ? 1 : 0;
      }
      const double span = last - first;
      int checked = 0, identical = 1;
      const int want = a.verify < 0 ? a.requests : std::min(a.verify, a.requests);
      server.clear_resident();
      if (dig)
        for (int i = 0; i < a.requests; ++i) {
          std::vector<uint64_t> d;
          for (const auto &o : res[i].outputs) {
            d.push_back(digest(o.offsets));
            d.push_back(digest(o.spans));
          }
          std::fprintf(dig,
                       "{\"rate\": %.0f, \"request\": %d, \"doc\": %d, \"queries\": %s, "
                       "\"records\": %lld, \"digests\": %s}\n",
                       rate, i, shapes[i].doc, json_list(shapes[i].queries).c_str(),
                       (long long)res[i].num_records, json_list(d).c_str());
        }
      for (int i = 0; i < want; ++i) {
        const RequestResult alone = server.run_alone(reqs[i]);
        if (!same(alone, res[i])) {
          identical = 0;
          std::fprintf(stderr, "[serve] request %d differs from its solo run\n", i);
        }
        ++checked;
      }
      const auto &st = server.stats();
      std::printf(
          "{\"metric\": \"serving p50/p99 request latency and sustained JSON GB/s\", "
          "\"rate_per_s\": %.0f, \"requests\": %d, \"documents\": %zu, \"p50_ms\": %.3f, "
          "\"p99_ms\": %.3f, \"mean_ms\": %.3f, \"sustained_GBps\": %.3f, "
          "\"new_document_GBps\": %.3f, \"makespan_s\": %.4f, \"resident_hits\": %llu, "
          "\"evictions\": %llu, \"retries\": %llu, \"peak_device_bytes\": %llu, "
          "\"slots\": %d, \"hbm_budget_bytes\": %llu, \"verified\": %d, \"identical\": %s}\n",
          rate, a.requests, schemas.size(), pct(lat, 50), pct(lat, 99),
          [&] { double m = 0; for (double x : lat) m += x; return lat.empty() ? 0 : m / lat.size(); }(),
          (double)req_bytes / span / 1e9, (double)new_bytes / span / 1e9, span,
          (unsigned long long)resident, (unsigned long long)st.evictions,
          (unsigned long long)st.retries, (unsigned long long)st.peak_bytes, a.slots,
          (unsigned long long)so.hbm_budget_bytes, checked, identical ? "true" : "false");
      std::fflush(stdout);
      if (!identical) return 3;
    }
    if (dig) std::fclose(dig);
    gpj_host_free(docs);
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "gpjson_serve: %s\n", e.what());
    return 2;
  }
}
