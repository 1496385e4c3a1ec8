/*
 * gpj_gen.c -- seeded synthetic JSON generator for the benchmark workloads.
 *
 * Deterministic for a given (schema, target size, seed): the output is split
 * into kStreams independent record streams, each seeded from the global seed
 * with SplitMix64 and driven by xoshiro256**, generated in parallel and
 * concatenated in stream order.  The host thread count never changes the
 * bytes.  Records are compact ASCII JSON; NDJSON schemas end every record with
 * '\n' and never exceed the target size.
 *
 * Schemas:
 *   0  NDJSON, 2-4 KiB records of 10-30 keys ([a-z]{3,12}); values 40% strings
 *      (8-200 printable chars, 2% with 1-3 escape runs of 1-5 escapes from
 *      \" and \\), 25% numbers (half int64, half decimal), 5% true/false/null,
 *      30% nested objects/arrays (depth <= 6); half the records carry
 *      "a":{..."b":v...}.
 *   1  one document: a top-level array of objects
 *      {"id":N,"items":[0-20 x {"sku":int,"name":str (3% escaped)}],
 *       "meta":<filler nested to depth 10>} with no newlines.
 *   2  NDJSON, ~1 KiB records of depth <= 4 with "user":{"lang":<30 codes>},
 *      "e":[{"f":..}..], "h":[{"t":..,"k":..,"id":..}..] where k=="v" in 5% of
 *      the elements, plus filler keys.
 *   3  NDJSON, twitter-like records (user object, entities arrays, text).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { kStreams = 64 };

typedef struct {
  uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t next_u64(rng_t *r) {
  uint64_t *s = r->s;
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}
static void seed_rng(rng_t *r, uint64_t seed, uint64_t stream) {
  uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}
/* uniform in [lo, hi] */
static int64_t rint_(rng_t *r, int64_t lo, int64_t hi) {
  return lo + (int64_t)(next_u64(r) % (uint64_t)(hi - lo + 1));
}
static double runit(rng_t *r) { return (double)(next_u64(r) >> 11) * (1.0 / 9007199254740992.0); }

typedef struct {
  uint8_t *p;
  uint64_t len, cap;
} buf_t;

static void put(buf_t *b, const char *s, uint64_t n) {
  if (b->len + n > b->cap) {
    uint64_t nc = b->cap ? b->cap * 2 : 1 << 16;
    while (nc < b->len + n) nc *= 2;
    b->p = (uint8_t *)realloc(b->p, nc);
    b->cap = nc;
  }
  memcpy(b->p + b->len, s, n);
  b->len += n;
}
static void putc_(buf_t *b, char c) { put(b, &c, 1); }
static void puts_(buf_t *b, const char *s) { put(b, s, strlen(s)); }
static void put_int(buf_t *b, int64_t v) {
  char t[32];
  int n = snprintf(t, sizeof t, "%lld", (long long)v);
  put(b, t, (uint64_t)n);
}

static void put_key_name(buf_t *b, rng_t *r, char *name_out) {
  int len = (int)rint_(r, 3, 12);
  for (int i = 0; i < len; ++i) name_out[i] = (char)('a' + rint_(r, 0, 25));
  name_out[len] = 0;
  (void)b;
}

/* printable ASCII excluding '"' and '\\' */
static char plain_char(rng_t *r) {
  for (;;) {
    char c = (char)rint_(r, 0x20, 0x7E);
    if (c != '"' && c != '\\') return c;
  }
}

static void put_string(buf_t *b, rng_t *r, int minlen, int maxlen, double escape_p) {
  int len = (int)rint_(r, minlen, maxlen);
  putc_(b, '"');
  int runs = (runit(r) < escape_p) ? (int)rint_(r, 1, 3) : 0;
  int run_at[3] = {-1, -1, -1};
  for (int i = 0; i < runs; ++i) run_at[i] = (int)rint_(r, 0, len > 0 ? len - 1 : 0);
  for (int i = 0; i < len; ++i) {
    for (int k = 0; k < runs; ++k) {
      if (run_at[k] == i) {
        int n = (int)rint_(r, 1, 5);
        for (int j = 0; j < n; ++j) puts_(b, (next_u64(r) & 1) ? "\\\"" : "\\\\");
      }
    }
    putc_(b, plain_char(r));
  }
  putc_(b, '"');
}

static void put_number(buf_t *b, rng_t *r) {
  if (next_u64(r) & 1) {
    put_int(b, (int64_t)(next_u64(r) >> (int)rint_(r, 1, 63)) * ((next_u64(r) & 1) ? 1 : -1));
  } else {
    char t[48];
    int n = snprintf(t, sizeof t, "%.*f", (int)rint_(r, 1, 6), (runit(r) - 0.5) * 2e6);
    put(b, t, (uint64_t)n);
  }
}

static void put_value0(buf_t *b, rng_t *r, int depth, int max_depth);

static void put_object0(buf_t *b, rng_t *r, int depth, int max_depth, int min_keys,
                        int max_keys) {
  int nk = (int)rint_(r, min_keys, max_keys);
  char names[32][13];
  putc_(b, '{');
  int emitted = 0;
  for (int i = 0; i < nk && i < 32; ++i) {
    char nm[13];
    put_key_name(b, r, nm);
    int dup = 0;
    for (int j = 0; j < emitted; ++j)
      if (!strcmp(names[j], nm)) dup = 1;
    if (dup) continue;
    strcpy(names[emitted++], nm);
    if (emitted > 1) putc_(b, ',');
    putc_(b, '"');
    puts_(b, nm);
    puts_(b, "\":");
    put_value0(b, r, depth + 1, max_depth);
  }
  putc_(b, '}');
}

static void put_value0(buf_t *b, rng_t *r, int depth, int max_depth) {
  double x = runit(r);
  if (x < 0.40 || (x >= 0.70 && depth >= max_depth)) {
    put_string(b, r, 8, 200, 0.02);
  } else if (x < 0.65) {
    put_number(b, r);
  } else if (x < 0.70) {
    const char *lits[3] = {"true", "false", "null"};
    puts_(b, lits[rint_(r, 0, 2)]);
  } else if (next_u64(r) & 1) {
    put_object0(b, r, depth, max_depth, 1, 4);
  } else {
    int n = (int)rint_(r, 0, 5);
    putc_(b, '[');
    for (int i = 0; i < n; ++i) {
      if (i) putc_(b, ',');
      put_value0(b, r, depth + 1, max_depth);
    }
    putc_(b, ']');
  }
}

/* schema 0 record */
static void record0(buf_t *b, rng_t *r) {
  uint64_t start = b->len;
  int64_t target = rint_(r, 2048, 4096);
  int nk_max = (int)rint_(r, 10, 30);
  char names[31][13];
  int emitted = 0;
  int plant = (int)(next_u64(r) & 1);
  int plant_at = plant ? (int)rint_(r, 0, 9) : -1;
  putc_(b, '{');
  for (int i = 0; i < 30; ++i) {
    if (emitted >= 10 && (emitted >= nk_max || (int64_t)(b->len - start) >= target)) break;
    if (i == plant_at) {
      if (emitted) putc_(b, ',');
      puts_(b, "\"a\":{");
      int m = (int)rint_(r, 0, 3), bpos = (int)rint_(r, 0, m);
      for (int k = 0; k <= m; ++k) {
        if (k) putc_(b, ',');
        if (k == bpos) {
          puts_(b, "\"b\":");
        } else {
          char nm[13];
          put_key_name(b, r, nm);
          putc_(b, '"');
          puts_(b, nm);
          puts_(b, "\":");
        }
        put_value0(b, r, 2, 6);
      }
      putc_(b, '}');
      strcpy(names[emitted++], "a");
      continue;
    }
    char nm[13];
    put_key_name(b, r, nm);
    int dup = 0;
    for (int j = 0; j < emitted; ++j)
      if (!strcmp(names[j], nm)) dup = 1;
    if (dup) continue;
    if (emitted) putc_(b, ',');
    strcpy(names[emitted++], nm);
    putc_(b, '"');
    puts_(b, nm);
    puts_(b, "\":");
    put_value0(b, r, 1, 6);
  }
  putc_(b, '}');
  putc_(b, '\n');
}

static const char *kLangs[30] = {"en", "nl", "fr", "de", "es", "it", "pt", "ru", "ja", "ko",
                                 "zh", "ar", "tr", "pl", "sv", "da", "fi", "no", "cs", "el",
                                 "he", "hi", "id", "th", "vi", "uk", "hu", "ro", "fa", "und"};

/* schema 2 record, ~1 KiB, depth <= 4 */
static void record2(buf_t *b, rng_t *r) {
  uint64_t start = b->len;
  int64_t target = rint_(r, 768, 1280);
  puts_(b, "{\"id\":");
  put_int(b, (int64_t)(next_u64(r) >> 12));
  puts_(b, ",\"user\":{\"id\":");
  put_int(b, rint_(r, 1, 1000000000));
  puts_(b, ",\"lang\":\"");
  puts_(b, kLangs[rint_(r, 0, 29)]);
  puts_(b, "\",\"name\":");
  put_string(b, r, 5, 24, 0.02);
  puts_(b, "},\"e\":[");
  int ne = (int)rint_(r, 1, 4);
  for (int i = 0; i < ne; ++i) {
    if (i) putc_(b, ',');
    puts_(b, "{\"f\":");
    if (next_u64(r) & 1) put_number(b, r);
    else put_string(b, r, 3, 30, 0.02);
    puts_(b, ",\"g\":[");
    put_int(b, rint_(r, 0, 99));
    puts_(b, "]}");
  }
  puts_(b, "],\"h\":[");
  int nh = (int)rint_(r, 1, 6);
  for (int i = 0; i < nh; ++i) {
    if (i) putc_(b, ',');
    puts_(b, "{\"t\":");
    put_string(b, r, 4, 40, 0.02);
    puts_(b, ",\"k\":\"");
    if (runit(r) < 0.05) putc_(b, 'v');
    else {
      int n = (int)rint_(r, 1, 6);
      for (int j = 0; j < n; ++j) putc_(b, (char)('a' + rint_(r, 0, 20)));
    }
    puts_(b, "\",\"id\":");
    put_int(b, rint_(r, 0, 1 << 30));
    putc_(b, '}');
  }
  putc_(b, ']');
  int fk = 0;
  while ((int64_t)(b->len - start) < target - 40 && fk < 40) {
    char nm[13];
    put_key_name(b, r, nm);
    puts_(b, ",\"x");
    puts_(b, nm);
    puts_(b, "\":");
    double x = runit(r);
    if (x < 0.55) put_string(b, r, 8, 120, 0.02);
    else if (x < 0.85) put_number(b, r);
    else if (x < 0.92) puts_(b, (next_u64(r) & 1) ? "true" : "null");
    else {
      puts_(b, "{\"p\":");
      put_number(b, r);
      puts_(b, ",\"q\":[");
      put_int(b, rint_(r, 0, 9));
      puts_(b, "]}");
    }
    ++fk;
  }
  puts_(b, "}\n");
}

/* schema 3: twitter-like record */
static void record3(buf_t *b, rng_t *r) {
  puts_(b, "{\"created_at\":\"Thu Apr 06 15:24:15 +0000 2017\",\"id\":");
  put_int(b, (int64_t)(next_u64(r) >> 4));
  puts_(b, ",\"text\":");
  put_string(b, r, 20, 280, 0.05);
  puts_(b, ",\"user\":{\"id\":");
  put_int(b, rint_(r, 1, 1 << 30));
  puts_(b, ",\"name\":");
  put_string(b, r, 3, 20, 0.02);
  puts_(b, ",\"lang\":\"");
  puts_(b, kLangs[rint_(r, 0, 29)]);
  puts_(b, "\",\"followers_count\":");
  put_int(b, rint_(r, 0, 5000000));
  puts_(b, ",\"description\":");
  put_string(b, r, 0, 160, 0.03);
  puts_(b, "},\"entities\":{\"hashtags\":[");
  int nh = (int)rint_(r, 0, 4);
  for (int i = 0; i < nh; ++i) {
    if (i) putc_(b, ',');
    puts_(b, "{\"text\":");
    put_string(b, r, 2, 15, 0.0);
    puts_(b, ",\"indices\":[");
    put_int(b, rint_(r, 0, 100));
    putc_(b, ',');
    put_int(b, rint_(r, 100, 140));
    puts_(b, "]}");
  }
  puts_(b, "],\"urls\":[]},\"lang\":\"");
  puts_(b, kLangs[rint_(r, 0, 29)]);
  puts_(b, "\",\"retweet_count\":");
  put_int(b, rint_(r, 0, 100000));
  puts_(b, "}\n");
}

/* schema 1: one array element (no trailing separator) */
static void put_filler(buf_t *b, rng_t *r, int depth, int max_depth) {
  if (depth >= max_depth) {
    put_int(b, rint_(r, 0, 999));
    return;
  }
  if (next_u64(r) & 1) {
    puts_(b, "{\"m\":");
    put_filler(b, r, depth + 1, max_depth);
    putc_(b, '}');
  } else {
    putc_(b, '[');
    put_filler(b, r, depth + 1, max_depth);
    putc_(b, ']');
  }
}
static void element1(buf_t *b, rng_t *r, int64_t id) {
  puts_(b, "{\"id\":");
  put_int(b, id);
  puts_(b, ",\"items\":[");
  int n = (int)rint_(r, 0, 20);
  for (int i = 0; i < n; ++i) {
    if (i) putc_(b, ',');
    puts_(b, "{\"sku\":");
    put_int(b, rint_(r, 0, 1000000000));
    puts_(b, ",\"name\":");
    put_string(b, r, 4, 24, 0.03);
    putc_(b, '}');
  }
  puts_(b, "],\"meta\":");
  put_filler(b, r, 2, 9); /* element object = depth 2; filler reaches depth 10 */
  putc_(b, '}');
}

typedef struct {
  int schema;
  uint64_t seed, stream, share;
  int64_t first_id, count; /* schema 1: element ids */
  buf_t out;
} job_t;

static void *gen_worker(void *arg) {
  job_t *j = (job_t *)arg;
  rng_t r;
  seed_rng(&r, j->seed, j->stream);
  buf_t tmp = {0, 0, 0};
  if (j->schema == 1) {
    for (int64_t i = 0; i < j->count; ++i) {
      tmp.len = 0;
      element1(&tmp, &r, j->first_id + i);
      if (j->first_id + i > 0) putc_(&j->out, ',');
      put(&j->out, (const char *)tmp.p, tmp.len);
    }
  } else {
    for (;;) {
      tmp.len = 0;
      if (j->schema == 0) record0(&tmp, &r);
      else if (j->schema == 2) record2(&tmp, &r);
      else record3(&tmp, &r);
      if (j->out.len + tmp.len > j->share) break;
      put(&j->out, (const char *)tmp.p, tmp.len);
    }
  }
  free(tmp.p);
  return NULL;
}

/* Generates into `out` (capacity cap).  For NDJSON schemas `target` is the
 * size bound in bytes; for schema 1 it is the number of array elements.
 * Returns bytes written, or 0 if cap is too small. */
uint64_t gpj_gen(int schema, uint64_t target, uint64_t seed, int threads, uint8_t *out,
                 uint64_t cap) {
  job_t jobs[kStreams];
  memset(jobs, 0, sizeof jobs);
  for (int s = 0; s < kStreams; ++s) {
    jobs[s].schema = schema;
    jobs[s].seed = seed;
    jobs[s].stream = (uint64_t)s;
    if (schema == 1) {
      int64_t per = (int64_t)(target / kStreams), extra = (int64_t)(target % kStreams);
      jobs[s].first_id = s * per + (s < extra ? s : extra);
      jobs[s].count = per + (s < extra ? 1 : 0);
    } else {
      jobs[s].share = target / kStreams;
    }
  }
  if (threads < 1) threads = 1;
  for (int base = 0; base < kStreams; base += threads) {
    pthread_t tid[kStreams];
    int k = 0;
    for (int s = base; s < base + threads && s < kStreams; ++s, ++k)
      pthread_create(&tid[k], NULL, gen_worker, &jobs[s]);
    for (int i = 0; i < k; ++i) pthread_join(tid[i], NULL);
  }
  uint64_t total = schema == 1 ? 2 : 0;
  for (int s = 0; s < kStreams; ++s) total += jobs[s].out.len;
  uint64_t off = 0;
  if (total <= cap) {
    if (schema == 1) out[off++] = '[';
    for (int s = 0; s < kStreams; ++s) {
      memcpy(out + off, jobs[s].out.p, jobs[s].out.len);
      off += jobs[s].out.len;
    }
    if (schema == 1) out[off++] = ']';
  }
  for (int s = 0; s < kStreams; ++s) free(jobs[s].out.p);
  return total <= cap ? total : 0;
}

/* ---- many files (the batched and serving configurations) -----------------
 * File f: NDJSON of schema schemas[f] (0, 2 or 3) from its own generator
 * stream (seed, stream 1000 + f), records appended while the file stays
 * within targets[f] bytes, but at least one record; or, for schema 1, one
 * top-level array document of schema-1 elements within targets[f] bytes (at
 * least one element).  The files are laid out
 * in one zero-filled buffer, each at an offset that is a multiple of
 * `align`; offsets_out / sizes_out receive the layout.  Returns the buffer
 * size (0 on allocation failure); free the buffer with gpj_gen_free. */
typedef struct {
  const int32_t *schemas;
  const uint64_t *targets;
  uint64_t seed;
  int32_t num_files;
  int32_t next; /* work queue */
  pthread_mutex_t mu;
  buf_t *bufs;
} files_job_t;

static void *files_worker(void *arg) {
  files_job_t *j = (files_job_t *)arg;
  buf_t tmp = {0, 0, 0};
  for (;;) {
    pthread_mutex_lock(&j->mu);
    const int32_t f = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (f >= j->num_files) break;
    rng_t r;
    seed_rng(&r, j->seed, 1000u + (uint64_t)f);
    buf_t *out = &j->bufs[f];
    if (j->schemas[f] == 1) {
      /* one top-level array document: '[' elements ']', no newline */
      putc_(out, '[');
      for (int64_t id = 0;; ++id) {
        tmp.len = 0;
        element1(&tmp, &r, id);
        if (id > 0 && out->len + 1 + tmp.len + 1 > j->targets[f]) break;
        if (id) putc_(out, ',');
        put(out, (const char *)tmp.p, tmp.len);
      }
      putc_(out, ']');
      continue;
    }
    for (;;) {
      tmp.len = 0;
      if (j->schemas[f] == 0) record0(&tmp, &r);
      else if (j->schemas[f] == 2) record2(&tmp, &r);
      else record3(&tmp, &r);
      if (out->len > 0 && out->len + tmp.len > j->targets[f]) break;
      put(out, (const char *)tmp.p, tmp.len);
    }
  }
  free(tmp.p);
  return NULL;
}

uint64_t gpj_gen_files(int32_t num_files, const int32_t *schemas, const uint64_t *targets,
                       uint64_t seed, int threads, uint64_t align, uint64_t *offsets_out,
                       uint64_t *sizes_out, uint8_t **buffer_out) {
  *buffer_out = NULL;
  if (num_files <= 0 || align == 0) return 0;
  files_job_t j;
  j.schemas = schemas;
  j.targets = targets;
  j.seed = seed;
  j.num_files = num_files;
  j.next = 0;
  pthread_mutex_init(&j.mu, NULL);
  j.bufs = (buf_t *)calloc((size_t)num_files, sizeof(buf_t));
  if (!j.bufs) return 0;
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t tid[256];
  for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, files_worker, &j);
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  pthread_mutex_destroy(&j.mu);
  uint64_t total = 0;
  for (int32_t f = 0; f < num_files; ++f) {
    offsets_out[f] = total;
    sizes_out[f] = j.bufs[f].len;
    total += (j.bufs[f].len + align - 1) / align * align;
  }
  uint8_t *buf = (uint8_t *)calloc(total ? total : 1, 1);
  if (buf) {
    for (int32_t f = 0; f < num_files; ++f)
      memcpy(buf + offsets_out[f], j.bufs[f].p, j.bufs[f].len);
  }
  for (int32_t f = 0; f < num_files; ++f) free(j.bufs[f].p);
  free(j.bufs);
  *buffer_out = buf;
  return buf ? total : 0;
}

void gpj_gen_free(uint8_t *buffer) { free(buffer); }

