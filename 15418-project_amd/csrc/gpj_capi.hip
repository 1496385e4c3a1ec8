// gpj_capi.hip -- implementation of the C ABI declared in include/gpj_capi.h.
//
// Host side of the HIP path: allocation, stream-ordered launches and status
// mapping.  No C++ exception crosses the ABI; failures return a negative
// gpj_status and leave a thread-local message for gpj_last_error().
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "gpj_capi.h"
#include "gpj_compiler.h"
#include "gpj_kernels.h"

namespace {

thread_local std::string g_error;

gpj_status fail(gpj_status s, const std::string &msg) {
  g_error = msg;
  return s;
}

gpj_status hip_fail(hipError_t e, const char *ctx) {
  return fail(e == hipErrorOutOfMemory ? GPJ_ERR_OUT_OF_MEMORY : GPJ_ERR_DEVICE,
              std::string(ctx) + ": " + hipGetErrorString(e));
}

#define GPJ_HIP(call, ctx)                                                     \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) return hip_fail(e_, ctx);                            \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

uint64_t tiles_for(uint64_t n) { return div_up(div_up(n, 64), gpj::kTileWords); }

// Scratch layout of the index build: tile aggregates (multi-file: tile
// descriptors), tile inclusive states (one uint64 per tile each), then a
// spare word and the failure flag.
uint64_t scratch_bytes(uint64_t n) {
  return tiles_for(n) * 16 + 256;
}
uint32_t *scratch_counter(void *scratch, uint64_t n) {
  return reinterpret_cast<uint32_t *>(static_cast<uint64_t *>(scratch) + 2 * tiles_for(n));
}

// Polls a waiting index tile makes before it gives up (tests lower it).
std::atomic<uint32_t> g_wait_limit{gpj::kDefaultWaitLimit};
// Single inputs of at least this many tiles take two per workgroup (tests
// set it to drive either form on any input).
std::atomic<uint64_t> g_pair_min_tiles{gpj::kPairMinTiles};

__global__ void empty_index_kernel(int64_t *newline, int64_t cap, int64_t *num_lines) {
  if (cap > 0) newline[0] = 0;
  *num_lines = 1;
}

template <int LMAX, bool SEG, int NH>
gpj_status launch_index(const uint8_t *d_in, uint64_t n, uint64_t words, int levels,
                        uint64_t *d_string, uint64_t *d_leveled, int64_t *d_newline,
                        int64_t cap, int64_t *d_num_lines, uint64_t *desc,
                        uint32_t *counter, uint64_t tiles, hipStream_t st,
                        gpj::BatchSeg seg = gpj::BatchSeg{nullptr, nullptr, nullptr},
                        uint64_t init = gpj::index_start_state(0, 0, -1)) {
  const int aligned = (reinterpret_cast<uintptr_t>(d_in) & 15u) == 0;
  // a workgroup per NH tiles, plus one (block 0) that runs the prefix scan
  hipLaunchKernelGGL((gpj::index_kernel<LMAX, SEG, NH>),
                     dim3((uint32_t)(div_up(tiles, NH) + 1)),
                     dim3(gpj::kTileWords), 0, st, d_in, n, words, levels, aligned, d_string,
                     d_leveled, d_newline, cap, d_num_lines, desc, counter, tiles, seg,
                     g_wait_limit.load(), init);
  GPJ_HIP(hipGetLastError(), "index_kernel launch");
  if (!SEG) { // a failed build reports -1 lines
    hipLaunchKernelGGL(gpj::index_status_kernel, dim3(1), dim3(64), 0, st, counter + 1,
                       d_num_lines);
    GPJ_HIP(hipGetLastError(), "index_status_kernel launch");
  }
  return GPJ_OK;
}

__global__ void fill_i64_kernel(int64_t *p, int64_t count, int64_t v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    p[i] = v;
}

// Up to 16 single words from anywhere to one place (so a host needs one copy
// for a handful of device counts).
struct WordSources {
  const uint64_t *src[16];
};
__global__ void gather_words_kernel(WordSources s, int32_t count, uint64_t *dst) {
  const int i = (int)threadIdx.x;
  if (i < count) dst[i] = s.src[i] ? *s.src[i] : 0;
}

// Plain 16-byte-per-lane grid-stride copy; the calibration kernel for HBM
// traffic counters.
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

} // namespace

extern "C" {

const char *gpj_last_error(void) { return g_error.c_str(); }

int gpj_device_available(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return count > 0;
}

gpj_status gpj_alloc(void **d_ptr, uint64_t bytes) {
  if (!d_ptr) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_alloc: null output pointer");
  *d_ptr = nullptr;
  if (bytes == 0) return GPJ_OK;
  GPJ_HIP(hipMalloc(d_ptr, bytes), "hipMalloc");
  return GPJ_OK;
}

gpj_status gpj_free(void *d_ptr) {
  if (d_ptr) GPJ_HIP(hipFree(d_ptr), "hipFree");
  return GPJ_OK;
}

gpj_status gpj_h2d(void *d_dst, const void *h_src, uint64_t bytes, void *stream) {
  if (bytes == 0) return GPJ_OK;
  if (!d_dst || !h_src) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_h2d: null pointer");
  GPJ_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, as_stream(stream)),
          "hipMemcpyAsync H2D");
  return GPJ_OK;
}

gpj_status gpj_d2h(void *h_dst, const void *d_src, uint64_t bytes, void *stream) {
  if (bytes == 0) return GPJ_OK;
  if (!h_dst || !d_src) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_d2h: null pointer");
  GPJ_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, as_stream(stream)),
          "hipMemcpyAsync D2H");
  GPJ_HIP(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
  return GPJ_OK;
}

gpj_status gpj_host_alloc(void **h_ptr, uint64_t bytes) {
  if (!h_ptr) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_host_alloc: null output pointer");
  *h_ptr = nullptr;
  if (bytes == 0) return GPJ_OK;
  GPJ_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
  return GPJ_OK;
}

gpj_status gpj_host_free(void *h_ptr) {
  if (h_ptr) GPJ_HIP(hipHostFree(h_ptr), "hipHostFree");
  return GPJ_OK;
}

gpj_status gpj_alloc_async(void **d_ptr, uint64_t bytes, void *stream) {
  if (!d_ptr) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_alloc_async: null output pointer");
  *d_ptr = nullptr;
  if (bytes == 0) return GPJ_OK;
  static bool pool_ready = false;
  if (!pool_ready) {
    // keep freed blocks in the pool instead of returning them at each sync
    int dev = 0;
    hipMemPool_t pool = nullptr;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t keep = ~0ull;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
    pool_ready = true;
  }
  GPJ_HIP(hipMallocAsync(d_ptr, bytes, as_stream(stream)), "hipMallocAsync");
  return GPJ_OK;
}

gpj_status gpj_free_async(void *d_ptr, void *stream) {
  if (d_ptr) GPJ_HIP(hipFreeAsync(d_ptr, as_stream(stream)), "hipFreeAsync");
  return GPJ_OK;
}

gpj_status gpj_d2h_async(void *h_dst, const void *d_src, uint64_t bytes, void *stream) {
  if (bytes == 0) return GPJ_OK;
  if (!h_dst || !d_src) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_d2h_async: null pointer");
  GPJ_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, as_stream(stream)),
          "hipMemcpyAsync D2H");
  return GPJ_OK;
}

gpj_status gpj_gather_words(uint64_t *d_dst, const uint64_t *const *d_srcs, int32_t count,
                            void *stream) {
  if (count < 0 || count > 16 || (count > 0 && (!d_dst || !d_srcs)))
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_gather_words: 0..16 words, non-null pointers");
  if (count == 0) return GPJ_OK;
  WordSources w{};
  for (int32_t i = 0; i < count; ++i) w.src[i] = d_srcs[i];
  hipLaunchKernelGGL(gather_words_kernel, dim3(1), dim3(64), 0, as_stream(stream), w, count,
                     d_dst);
  GPJ_HIP(hipGetLastError(), "gather_words_kernel launch");
  return GPJ_OK;
}

gpj_status gpj_event_create(void **event) {
  if (!event) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_event_create: null");
  hipEvent_t e = nullptr;
  GPJ_HIP(hipEventCreate(&e), "hipEventCreate");
  *event = e;
  return GPJ_OK;
}

gpj_status gpj_event_record(void *event, void *stream) {
  GPJ_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)),
          "hipEventRecord");
  return GPJ_OK;
}

gpj_status gpj_event_query(void *event) {
  const hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(event));
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return hip_fail(e, "hipEventQuery");
}

gpj_status gpj_event_sync(void *event) {
  GPJ_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)), "hipEventSynchronize");
  return GPJ_OK;
}

gpj_status gpj_event_elapsed_ms(void *start, void *end, float *ms) {
  if (!ms) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_event_elapsed_ms: null");
  GPJ_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                              reinterpret_cast<hipEvent_t>(end)),
          "hipEventElapsedTime");
  return GPJ_OK;
}

gpj_status gpj_event_destroy(void *event) {
  if (event) GPJ_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)), "hipEventDestroy");
  return GPJ_OK;
}

gpj_status gpj_stream_wait_event(void *stream, void *event) {
  GPJ_HIP(hipStreamWaitEvent(as_stream(stream), reinterpret_cast<hipEvent_t>(event), 0),
          "hipStreamWaitEvent");
  return GPJ_OK;
}

gpj_status gpj_stream_sync(void *stream) {
  GPJ_HIP(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
  return GPJ_OK;
}

gpj_status gpj_device_sync(void) {
  GPJ_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return GPJ_OK;
}

uint64_t gpj_index_scratch_bytes(uint64_t n) { return scratch_bytes(n); }

gpj_status gpj_index_status(const void *d_scratch, uint64_t n, int32_t *failed, void *stream) {
  if (!failed) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_index_status: null output");
  *failed = 0;
  if (n == 0) return GPJ_OK; // an empty input builds without scratch
  if (!d_scratch) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_index_status: null scratch");
  uint32_t flag = 0;
  GPJ_HIP(hipMemcpyAsync(&flag, scratch_counter(const_cast<void *>(d_scratch), n) + 1, 4,
                         hipMemcpyDeviceToHost, as_stream(stream)),
          "copy index status");
  GPJ_HIP(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
  *failed = flag != 0;
  return GPJ_OK;
}

gpj_status gpj_debug_set_wait_limit(int64_t polls) {
  if (polls < 0) polls = gpj::kDefaultWaitLimit;
  if (polls > (int64_t)gpj::kDefaultWaitLimit)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "wait limit: negative (default) or 0 .. 2^24 polls");
  g_wait_limit.store((uint32_t)polls);
  return GPJ_OK;
}

gpj_status gpj_debug_copy(void *d_dst, const void *d_src, uint64_t bytes, void *stream) {
  if (bytes % 16 != 0 || (reinterpret_cast<uintptr_t>(d_dst) & 15u) ||
      (reinterpret_cast<uintptr_t>(d_src) & 15u))
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_debug_copy: 16-byte aligned sizes and pointers");
  if (bytes == 0) return GPJ_OK;
  const uint64_t n16 = bytes / 16;
  const uint64_t blocks = std::min<uint64_t>(div_up(n16, 256), 8192);
  hipLaunchKernelGGL(copy_kernel, dim3((uint32_t)blocks), dim3(256), 0, as_stream(stream),
                     static_cast<const uint4 *>(d_src), static_cast<uint4 *>(d_dst), n16);
  GPJ_HIP(hipGetLastError(), "copy_kernel launch");
  return GPJ_OK;
}

} // extern "C"

namespace {

gpj_status build_into(const uint8_t *d_in, uint64_t n, int32_t levels, uint64_t *d_string,
                      uint64_t *d_leveled, int64_t *d_newline, int64_t newline_capacity,
                      int64_t *d_num_lines, void *d_scratch, void *stream, uint64_t init) {
  if (levels < 0 || levels > GPJ_MAX_INDEX_LEVELS)
    return fail(GPJ_ERR_INVALID_ARGUMENT,
                "index levels must be in [0, " + std::to_string(GPJ_MAX_INDEX_LEVELS) + "]");
  if (!d_num_lines || (newline_capacity > 0 && !d_newline))
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_index_build_into: null output");
  hipStream_t st = as_stream(stream);
  if (n == 0) {
    hipLaunchKernelGGL(empty_index_kernel, dim3(1), dim3(1), 0, st, d_newline,
                       newline_capacity, d_num_lines);
    GPJ_HIP(hipGetLastError(), "empty_index_kernel launch");
    return GPJ_OK;
  }
  if (!d_in || !d_string || (levels > 0 && !d_leveled) || !d_scratch)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_index_build_into: null buffer");
  const uint64_t words = div_up(n, 64);
  const uint64_t tiles = tiles_for(n);
  if (tiles >= 0xFFFFFFFFull) // the grid holds tiles + 1 workgroups
    return fail(GPJ_ERR_INVALID_ARGUMENT, "input too large for one index build");
  uint64_t *desc = static_cast<uint64_t *>(d_scratch);
  uint32_t *counter = reinterpret_cast<uint32_t *>(desc + 2 * tiles);
  const uint64_t clear_blocks = std::min<uint64_t>(div_up(tiles, 256), 2048);
  hipLaunchKernelGGL(gpj::index_clear_kernel, dim3((uint32_t)clear_blocks), dim3(256), 0, st,
                     desc, 2 * tiles, counter);
  GPJ_HIP(hipGetLastError(), "index_clear_kernel launch");
#define GPJ_DISPATCH(L)                                                                   \
  return tiles >= g_pair_min_tiles.load()                                                  \
             ? launch_index<L, false, gpj::kTileHalves>(                                   \
                   d_in, n, words, levels, d_string, d_leveled, d_newline, newline_capacity, \
                   d_num_lines, desc, counter, tiles, st,                                  \
                   gpj::BatchSeg{nullptr, nullptr, nullptr}, init)                         \
             : launch_index<L, false, 1>(d_in, n, words, levels, d_string, d_leveled,      \
                                         d_newline, newline_capacity, d_num_lines, desc,   \
                                         counter, tiles, st,                               \
                                         gpj::BatchSeg{nullptr, nullptr, nullptr}, init)
  if (levels <= 2) GPJ_DISPATCH(2);
  if (levels <= 4) GPJ_DISPATCH(4);
  if (levels <= 8) GPJ_DISPATCH(8);
  if (levels <= 16) GPJ_DISPATCH(16);
  GPJ_DISPATCH(22);
#undef GPJ_DISPATCH
}

} // namespace

extern "C" {

gpj_status gpj_index_build_into(const uint8_t *d_in, uint64_t n, int32_t levels,
                                uint64_t *d_string, uint64_t *d_leveled,
                                int64_t *d_newline, int64_t newline_capacity,
                                int64_t *d_num_lines, void *d_scratch,
                                void *stream) {
  return build_into(d_in, n, levels, d_string, d_leveled, d_newline, newline_capacity,
                    d_num_lines, d_scratch, stream, gpj::index_start_state(0, 0, -1));
}

gpj_status gpj_index_build_shard(const uint8_t *d_in, uint64_t n, int32_t levels,
                                 uint64_t *d_string, uint64_t *d_leveled,
                                 int64_t *d_newline, int64_t newline_capacity,
                                 int64_t *d_num_lines, void *d_scratch,
                                 const gpj_carry *carry_in, void *stream) {
  if (!carry_in) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_index_build_shard: null carry");
  if ((carry_in->escape != 0 && carry_in->escape != 1) ||
      (carry_in->in_string != 0 && carry_in->in_string != 1) || carry_in->depth < -128 ||
      carry_in->depth > 127)
    return fail(GPJ_ERR_INVALID_ARGUMENT,
                "gpj_index_build_shard: escape and in_string are 0 or 1, depth in [-128, 127]");
  return build_into(d_in, n, levels, d_string, d_leveled, d_newline, newline_capacity,
                    d_num_lines, d_scratch, stream,
                    gpj::index_start_state((uint32_t)carry_in->escape,
                                           (uint32_t)carry_in->in_string, carry_in->depth));
}

gpj_status gpj_index_shard_summary(const void *d_scratch, uint64_t n,
                                   gpj_shard_summary *d_summary, void *stream) {
  if (!d_scratch || !d_summary)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_index_shard_summary: null scratch or output");
  if (n == 0) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_index_shard_summary: empty shard");
  hipLaunchKernelGGL(gpj::shard_summary_kernel, dim3(1), dim3(256), 0, as_stream(stream),
                     static_cast<const uint64_t *>(d_scratch), tiles_for(n),
                     reinterpret_cast<uint64_t *>(d_summary));
  GPJ_HIP(hipGetLastError(), "shard_summary_kernel launch");
  return GPJ_OK;
}

gpj_status gpj_shard_carries(const gpj_shard_summary *summaries, int32_t count,
                             gpj_carry *carries) {
  if (count < 0 || (count > 0 && !summaries) || !carries)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_shard_carries: null summaries or carries");
  gpj_carry c{0, 0, -1, 0, 0};
  carries[0] = c;
  for (int32_t k = 0; k < count; ++k) {
    const uint32_t s = 8u * ((uint32_t)c.escape | ((uint32_t)c.in_string << 1));
    const uint32_t st = (uint32_t)(summaries[k].table >> s) & 3u;
    const uint32_t dd = (uint32_t)(summaries[k].table >> (32 + s)) & 0xFFu;
    c.escape = (int32_t)(st & 1u);
    c.in_string = (int32_t)(st >> 1);
    c.depth = (int32_t)(int8_t)(uint8_t)((uint32_t)c.depth + dd);
    c.newlines += (int64_t)summaries[k].newlines;
    carries[k + 1] = c;
  }
  return GPJ_OK;
}

} // extern "C"

// ---- index build without allocation --------------------------------------
// The index scratch of gpj_index_build_static (tile descriptors, the spare
// word and the failure flag) for inputs up to GPJ_STATIC_MAX_BYTES: every GPU
// the library runs on holds its own copy until the process ends.  In host code
// this name is not a device address; the address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
static __device__ __align__(256) uint8_t scratch[GPJ_STATIC_SCRATCH_BYTES];
static_assert(GPJ_STATIC_SCRATCH_BYTES >= (GPJ_STATIC_MAX_BYTES / gpj::kTileBytes) * 16 + 256,
              "static scratch covers the largest input");

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
void *g_static_scratch[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up on the first call on
// that device, with the device current).
gpj_status static_scratch(void **out) {
  int dev = 0;
  GPJ_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(GPJ_ERR_DEVICE, "gpj_index_build_static: device id beyond 63");
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_scratch[dev]) {
    void *p = nullptr;
    GPJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(GPJ_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
    g_static_scratch[dev] = p;
  }
  *out = g_static_scratch[dev];
  return GPJ_OK;
}
} // namespace

extern "C" {

gpj_status gpj_index_build_static(const uint8_t *d_in, uint64_t n, int32_t levels,
                                  uint64_t *d_string, uint64_t *d_leveled,
                                  int64_t *d_newline, int64_t newline_capacity,
                                  int64_t *d_num_lines, void *stream) {
  if (n > GPJ_STATIC_MAX_BYTES)
    return fail(GPJ_ERR_CAPACITY, "gpj_index_build_static: input of " + std::to_string(n) +
                                      " bytes is above the static scratch's limit of " +
                                      std::to_string(GPJ_STATIC_MAX_BYTES) +
                                      " bytes; use gpj_index_build_into");
  void *ptr = nullptr;
  const gpj_status st = static_scratch(&ptr);
  if (st != GPJ_OK) return st;
  // the same kernels as gpj_index_build_into; index_clear_kernel (a fill
  // kernel) resets the descriptors and flags at the start of every call
  return gpj_index_build_into(d_in, n, levels, d_string, d_leveled, d_newline,
                              newline_capacity, d_num_lines, ptr, stream);
}

gpj_status gpj_index_free(gpj_index *index) {
  if (!index) return GPJ_OK;
  if (index->owns_memory) {
    (void)hipFree(index->d_newline);
    (void)hipFree(index->d_num_lines);
    (void)hipFree(index->d_string);
    (void)hipFree(index->d_leveled);
  }
  std::memset(index, 0, sizeof(*index));
  return GPJ_OK;
}

gpj_status gpj_index_build(const uint8_t *d_in, uint64_t n, int32_t levels,
                           const gpj_index_opts *opts, gpj_index *out,
                           void *stream) {
  if (!out) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_index_build: null output");
  std::memset(out, 0, sizeof(*out));
  if (levels < 0 || levels > GPJ_MAX_INDEX_LEVELS)
    return fail(GPJ_ERR_INVALID_ARGUMENT,
                "index levels must be in [0, " + std::to_string(GPJ_MAX_INDEX_LEVELS) + "]");
  if (n > 0 && !d_in)
    return fail(GPJ_ERR_INVALID_ARGUMENT,
                "Index building requires a GPU-resident partition buffer");
  hipStream_t st = as_stream(stream);
  const uint64_t words = div_up(n, 64);
  int64_t cap = (opts && opts->newline_capacity > 0) ? opts->newline_capacity
                                                     : (int64_t)(n / 64 + 2);
  void *scratch = nullptr;
  gpj_index r;
  std::memset(&r, 0, sizeof(r));
  r.owns_memory = 1;
  r.level_size = (int64_t)words;
  r.num_levels = levels;
  auto cleanup = [&]() {
    (void)hipFree(scratch);
    gpj_index_free(&r);
  };
  hipError_t e;
#define GPJ_TRY(call, ctx)                                                     \
  do {                                                                         \
    e = (call);                                                                \
    if (e != hipSuccess) {                                                     \
      gpj_status s_ = hip_fail(e, ctx);                                        \
      cleanup();                                                               \
      return s_;                                                               \
    }                                                                          \
  } while (0)
  GPJ_TRY(hipMalloc(&r.d_num_lines, 8), "hipMalloc num_lines");
  GPJ_TRY(hipMalloc(&r.d_newline, (size_t)cap * 8), "hipMalloc newline index");
  GPJ_TRY(hipMalloc(&r.d_string, words ? words * 8 : 8), "hipMalloc string index");
  GPJ_TRY(hipMalloc(&r.d_leveled, words && levels ? words * 8 * (uint64_t)levels : 8),
          "hipMalloc leveled bitmap index");
  GPJ_TRY(hipMalloc(&scratch, scratch_bytes(n)), "hipMalloc index scratch");
  gpj_status s = gpj_index_build_into(d_in, n, levels, r.d_string, r.d_leveled,
                                      r.d_newline, cap, r.d_num_lines, scratch, stream);
  if (s != GPJ_OK) {
    cleanup();
    return s;
  }
  int64_t num_lines = 0;
  GPJ_TRY(hipMemcpyAsync(&num_lines, r.d_num_lines, 8, hipMemcpyDeviceToHost, st),
          "copy num_lines");
  GPJ_TRY(hipStreamSynchronize(st), "index build");
  if (n > 0) {
    uint32_t failed = 0;
    GPJ_TRY(hipMemcpy(&failed, scratch_counter(scratch, n) + 1, 4, hipMemcpyDeviceToHost),
            "copy index status");
    if (failed) {
      cleanup();
      return fail(GPJ_ERR_DEVICE, "index build: a tile's wait for its carry timed out");
    }
  }
  if (num_lines > cap) {
    // exact second pass: count per tile, scan, write
    const uint64_t tiles = tiles_for(n);
    const int aligned = (reinterpret_cast<uintptr_t>(d_in) & 15u) == 0;
    (void)hipFree(r.d_newline);
    r.d_newline = nullptr;
    GPJ_TRY(hipMalloc(&r.d_newline, (size_t)num_lines * 8), "hipMalloc newline index");
    uint64_t *counts = nullptr;
    GPJ_TRY(hipMalloc(&counts, tiles * 8), "hipMalloc newline scratch");
    hipLaunchKernelGGL(gpj::newline_count_kernel, dim3((uint32_t)tiles), dim3(gpj::kTileWords),
                       0, st, d_in, n, aligned, counts);
    hipLaunchKernelGGL(gpj::newline_scan_kernel, dim3(1), dim3(1024), 0, st, counts, tiles,
                       r.d_num_lines, (int64_t)1);
    hipLaunchKernelGGL(gpj::newline_write_kernel, dim3((uint32_t)tiles), dim3(gpj::kTileWords),
                       0, st, d_in, n, aligned, counts, r.d_newline);
    e = hipStreamSynchronize(st);
    (void)hipFree(counts);
    GPJ_TRY(e, "newline index pass");
  }
#undef GPJ_TRY
  (void)hipFree(scratch);
  r.num_lines = num_lines;
  *out = r;
  return GPJ_OK;
}

gpj_status gpj_debug_escape_quote(const uint8_t *d_in, uint64_t n, uint64_t *d_escape,
                                  uint64_t *d_quote, void *stream) {
  if (n == 0) return GPJ_OK;
  if (!d_in || !d_escape || !d_quote)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_debug_escape_quote: null pointer");
  const uint64_t words = div_up(n, 64);
  hipLaunchKernelGGL(gpj::escape_quote_kernel, dim3((uint32_t)div_up(words, 256)), dim3(256),
                     0, as_stream(stream), d_in, n, words, d_escape, d_quote);
  GPJ_HIP(hipGetLastError(), "escape_quote_kernel launch");
  return GPJ_OK;
}

namespace {

// Record extent the narrow query kernels take (gpj::PosMode); tests lower it
// through gpj_debug_set_narrow_span to drive the wide kernel on small inputs.
std::atomic<int64_t> g_narrow_span{gpj::kNarrowSpan};
// Query walk (GPJ_WALK_*).
std::atomic<int32_t> g_walk{GPJ_WALK_AUTO};

// One record per wave (query_wave_kernel) for this launch?
bool wave_walk(const uint8_t *d_in, uint64_t n, int64_t max_records) {
  // (its positions are 32-bit: inputs up to kNarrowSpan bytes)
  if ((reinterpret_cast<uintptr_t>(d_in) & 15) != 0 || n > (uint64_t)gpj::kNarrowSpan) return false;
  const int32_t w = g_walk.load();
  if (w == GPJ_WALK_WAVE) return true;
  if (w == GPJ_WALK_LANE) return false;
  return max_records <= GPJ_WAVE_WALK_RECORDS ||
         (max_records <= GPJ_WAVE_WALK_LONG_RECORDS &&
          n >= (uint64_t)GPJ_WAVE_WALK_LONG_BYTES * (uint64_t)max_records);
}

// Deepest interpreter level of a query batch (0 = unknown counts as 15).
int deepest_level(const gpj_query_desc *queries, int32_t num_queries) {
  int deepest = 0;
  for (int32_t q = 0; q < num_queries; ++q) {
    const int lv = queries[q].max_level > 0 ? queries[q].max_level : 15;
    deepest = std::max(deepest, std::min(lv, 15));
  }
  return deepest;
}

// Shared by the line and the top-level-array record entry points.
// Scratch layout of dense results (see gpj_query_dense_scratch_bytes): region
// cursors, tmp_off per query, scan chunk sums per query, staging regions.
struct DenseLayout {
  uint64_t tmp_off, sums, tmp, bytes;
  int64_t blocks, region;
};
DenseLayout dense_layout(int64_t max_records, int32_t num_queries, int64_t total_capacity) {
  DenseLayout l{};
  const uint64_t rec = (uint64_t)std::max<int64_t>(max_records, 0);
  const uint64_t nq = (uint64_t)std::max(num_queries, 0);
  const uint64_t cap = (uint64_t)std::max<int64_t>(total_capacity, 0);
  l.blocks = (int64_t)div_up(rec + 1, gpj::kScanChunk);
  // room for the total plus an eighth, and a full stage per region
  l.region = (int64_t)div_up(cap + cap / 8 + gpj::kDenseRegions * 64ull * gpj::kDenseStage,
                             gpj::kDenseRegions);
  auto r256 = [](uint64_t b) { return div_up(b, 256) * 256; };
  l.tmp_off = (uint64_t)gpj::kDenseRegions * gpj::kCursorStride * 8; // the cursors
  l.sums = l.tmp_off + r256(nq * rec * 8);
  l.tmp = l.sums + r256(nq * ((uint64_t)l.blocks + 1) * 8);
  l.bytes = l.tmp + (uint64_t)gpj::kDenseRegions * (uint64_t)l.region * 16;
  return l;
}

gpj_status run_batch(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                     const int64_t *d_starts, const int64_t *d_num_starts,
                     gpj::RecordMode mode, const gpj_query_desc *queries,
                     int32_t num_queries, int64_t max_records, void *stream,
                     const gpj_dense_result *dense = nullptr, void *d_scratch = nullptr) {
  if (num_queries < 0 || num_queries > GPJ_MAX_BATCH_QUERIES || (num_queries > 0 && !queries))
    return fail(GPJ_ERR_INVALID_ARGUMENT,
                "query batch holds 0.." + std::to_string(GPJ_MAX_BATCH_QUERIES) + " queries");
  gpj::QueryBatch batch{};
  uint32_t total_ir = 0;
  int64_t total_cap = 0;
  for (int32_t q = 0; q < num_queries; ++q) {
    const gpj_query_desc &d = queries[q];
    if (d.ir_len == 0 || d.ir_len > GPJ_MAX_QUERY_IR_BYTES)
      return fail(GPJ_ERR_QUERY_EXECUTION, "Compiled query IR exceeds the 4096-byte IR buffer");
    if (!dense && d.capacity < 1)
      return fail(GPJ_ERR_QUERY_EXECUTION, "query needs at least one result slot");
    if (!d.d_ir || (!dense && (!d.d_spans || (d.variable && !d.d_counts))))
      return fail(GPJ_ERR_INVALID_ARGUMENT, "query: null bytecode or result buffer");
    total_ir += d.ir_len;
    if (dense) {
      const gpj_dense_result &r = dense[q];
      if (!d.variable && d.capacity < 1)
        return fail(GPJ_ERR_QUERY_EXECUTION, "fixed-slot query needs at least one result slot");
      if (!r.d_offsets || r.capacity < 0 || (r.capacity > 0 && !r.d_spans))
        return fail(GPJ_ERR_INVALID_ARGUMENT, "dense query: null offsets or spans");
      batch.q[q] = gpj::QuerySlot{d.d_ir, d.ir_len, d.variable ? 1 : d.capacity,
                                  d.variable ? 1 : 0, r.d_spans, r.d_offsets, 1, nullptr,
                                  r.capacity};
      total_cap += r.capacity;
    } else {
      batch.q[q] = gpj::QuerySlot{d.d_ir, d.ir_len, d.capacity, d.variable ? 1 : 0, d.d_spans,
                                  d.d_counts};
    }
  }
  if (total_ir > gpj::kBatchIrBytes)
    return fail(GPJ_ERR_QUERY_EXECUTION, "query batch bytecode exceeds 8192 bytes; split it");
  batch.count = num_queries;
  const int deepest = deepest_level(queries, num_queries);
  if (dense && num_queries > 0 && max_records >= 0) {
    // every offset array gets its entries even when no record runs
    if (!d_scratch) return fail(GPJ_ERR_INVALID_ARGUMENT, "dense queries: null scratch");
    if (!d_num_starts) return fail(GPJ_ERR_QUERY_EXECUTION, "dense queries: null record count");
    const DenseLayout lay = dense_layout(max_records, num_queries, total_cap);
    uint8_t *sc = static_cast<uint8_t *>(d_scratch);
    batch.ds = gpj::DenseStage{reinterpret_cast<int64_t *>(sc + lay.tmp),
                               reinterpret_cast<uint64_t *>(sc), lay.region};
    for (int32_t q = 0; q < num_queries; ++q)
      batch.q[q].tmp_off = reinterpret_cast<int64_t *>(sc + lay.tmp_off) + (int64_t)q * max_records;
  }
  if (max_records <= 0 || n == 0 || num_queries == 0) {
    if (dense && num_queries > 0 && max_records >= 0) {
      // no record: every offset is 0
      for (int32_t q = 0; q < num_queries; ++q) {
        hipLaunchKernelGGL(fill_i64_kernel, dim3((uint32_t)std::min<uint64_t>(div_up((uint64_t)max_records + 1, 256), 1024)),
                           dim3(256), 0, as_stream(stream), dense[q].d_offsets,
                           max_records + 1, (int64_t)0);
      }
      GPJ_HIP(hipGetLastError(), "fill launch");
    }
    return GPJ_OK;
  }
  if (n >= (1ull << 35) ||
      (uint64_t)std::max(index->num_levels, 1) * (uint64_t)index->level_size >= 0xFFFFFFF0ull)
    return fail(GPJ_ERR_QUERY_EXECUTION,
                "query input above 32 GiB per launch; split it into partitions");
  if (!d_in || !d_starts || !d_num_starts || !index->d_string || !index->d_leveled)
    return fail(GPJ_ERR_QUERY_EXECUTION,
                "Query execution requires GPU-resident newline, string, and bitmap indices");
  const bool wave = wave_walk(d_in, n, max_records);
  // one wave per block (the kernel keeps each wave's interpreter arrays in
  // its block's LDS).  Few records: fewer per wave (down to 8), so that more
  // waves share the SIMDs and a wave's step waits for fewer records.
  int32_t per_wave = 64;
  while (per_wave > 8 && (uint64_t)max_records / (uint64_t)per_wave < 4096) per_wave /= 2;
  mode.per_wave = per_wave;
  const uint64_t blocks =
      std::min<uint64_t>(div_up((uint64_t)max_records, (uint64_t)per_wave), 1u << 22);
  // the wave walk: one block (one wave) per record, up to 2^20 blocks
  const uint32_t wave_blocks = (uint32_t)std::min<uint64_t>((uint64_t)max_records, 1u << 20);
  // 32-bit positions; records the narrow kernel marks (only possible when
  // the input is longer than the span) are rerun by the wide kernel
  const int64_t span = g_narrow_span.load();
  const bool fixup = !wave && (int64_t)n > span;
  hipStream_t st = as_stream(stream);
  if (dense) { // empty staging regions
    hipLaunchKernelGGL(gpj::fill_u64_kernel, dim3(16), dim3(256), 0, st, batch.ds.cursors,
                       (int64_t)gpj::kDenseRegions * gpj::kCursorStride, (uint64_t)0);
    GPJ_HIP(hipGetLastError(), "fill launch");
  }
#define GPJ_QLAUNCH(LV, NW, FIX)                                                          \
  do {                                                                                    \
    if (dense)                                                                            \
      hipLaunchKernelGGL((gpj::query_batch_kernel<LV, NW, true>), dim3((uint32_t)blocks),  \
                         dim3(64), 0, st, d_in, (int64_t)n, d_starts, d_num_starts,          \
                         max_records, index->d_string, index->d_leveled,                    \
                         index->level_size, index->num_levels, mode,                        \
                         gpj::PosMode{span, FIX}, batch);                                    \
    else                                                                                  \
      hipLaunchKernelGGL((gpj::query_batch_kernel<LV, NW, false>), dim3((uint32_t)blocks), \
                         dim3(64), 0, st, d_in, (int64_t)n, d_starts, d_num_starts,          \
                         max_records, index->d_string, index->d_leveled,                    \
                         index->level_size, index->num_levels, mode,                        \
                         gpj::PosMode{span, FIX}, batch);                                    \
  } while (0)
#define GPJ_QPAIR(LV)                                                                     \
  do {                                                                                    \
    if (wave) {                                                                           \
      hipLaunchKernelGGL((gpj::query_wave_kernel<LV>), dim3(wave_blocks), dim3(64), 0, st,  \
                         d_in, (int64_t)n, d_starts, d_num_starts, max_records,            \
                         index->d_string, index->d_leveled, index->level_size,             \
                         index->num_levels, mode, batch);                                  \
      break;                                                                              \
    }                                                                                     \
    GPJ_QLAUNCH(LV, true, 0);                                                             \
    if (fixup) GPJ_QLAUNCH(LV, false, 1);                                                 \
  } while (0)
#define GPJ_QDISPATCH(M)                                                                  \
  do {                                                                                    \
    if (deepest + 1 <= 4) M(4);                                                           \
    else if (deepest + 1 <= 6) M(6);                                                      \
    else if (deepest + 1 <= 8) M(8);                                                      \
    else M(16);                                                                           \
  } while (0)
  GPJ_QDISPATCH(GPJ_QPAIR);
  GPJ_HIP(hipGetLastError(), "query_batch_kernel launch");
  if (dense) {
    // counts -> CSR offsets (chunk sums, their prefix, per-chunk scan fused
    // with the copy of staged spans into CSR order), then the records left
    // to the fixup pass
    const DenseLayout lay = dense_layout(max_records, num_queries, total_cap);
    uint8_t *sc = static_cast<uint8_t *>(d_scratch);
    gpj::DenseSet ds{};
    for (int32_t q = 0; q < num_queries; ++q) {
      ds.counts[q] = dense[q].d_offsets;
      ds.tmp_off[q] = batch.q[q].tmp_off;
      ds.spans[q] = dense[q].d_spans;
      ds.out_cap[q] = dense[q].capacity;
    }
    ds.block_sums = reinterpret_cast<uint64_t *>(sc + lay.sums);
    ds.blocks = lay.blocks;
    ds.num_starts = d_num_starts;
    ds.starts_cap = mode.starts_cap;
    ds.drop_last = mode.drop_last;
    ds.max_lines = max_records;
    ds.tmp = batch.ds.tmp;
    const dim3 chunks((uint32_t)lay.blocks, (uint32_t)num_queries);
    hipLaunchKernelGGL(gpj::dense_reduce_kernel, chunks, dim3(256), 0, st, ds);
    hipLaunchKernelGGL(gpj::dense_sums_kernel, dim3((uint32_t)num_queries), dim3(1024), 0, st,
                       ds);
    hipLaunchKernelGGL(gpj::dense_scan_kernel, chunks, dim3(256), 0, st, ds);
    GPJ_HIP(hipGetLastError(), "dense scan launch");
#define GPJ_QFIX(LV) GPJ_QLAUNCH(LV, false, 2)
    GPJ_QDISPATCH(GPJ_QFIX);
#undef GPJ_QFIX
    GPJ_HIP(hipGetLastError(), "dense fixup launch");
  }
#undef GPJ_QDISPATCH
#undef GPJ_QPAIR
#undef GPJ_QLAUNCH
  return GPJ_OK;
}

} // namespace

gpj_status gpj_debug_set_index_pairing(int64_t min_tiles) {
  if (min_tiles < 0)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "index pairing: 0 (default) or a tile count >= 1");
  g_pair_min_tiles.store(min_tiles == 0 ? gpj::kPairMinTiles : (uint64_t)min_tiles);
  return GPJ_OK;
}

gpj_status gpj_debug_set_narrow_span(int64_t bytes) {
  if (bytes == 0) bytes = gpj::kNarrowSpan;
  if (bytes < 4096 || bytes > gpj::kNarrowSpan)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "narrow span: 0 or 4096 .. 2^31 - 2^20 bytes");
  g_narrow_span.store(bytes);
  return GPJ_OK;
}

gpj_status gpj_query_set_walk(int32_t walk) {
  if (walk != GPJ_WALK_AUTO && walk != GPJ_WALK_LANE && walk != GPJ_WALK_WAVE)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "query walk: GPJ_WALK_AUTO, _LANE or _WAVE");
  g_walk.store(walk);
  return GPJ_OK;
}

gpj_status gpj_query_run_batch(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                               const gpj_query_desc *queries, int32_t num_queries,
                               int64_t max_lines, void *stream) {
  if (!index) return fail(GPJ_ERR_INVALID_ARGUMENT, "query: null index");
  // index->num_lines: the entries the newline buffer holds
  return run_batch(d_in, n, index, index->d_newline, index->d_num_lines,
                   gpj::RecordMode{0, 0, 64, index->num_lines}, queries, num_queries, max_lines,
                   stream);
}

uint64_t gpj_query_dense_scratch_bytes(int64_t max_records, int32_t num_queries,
                                       int64_t total_capacity) {
  return dense_layout(max_records, num_queries, total_capacity).bytes;
}

gpj_status gpj_query_run_dense(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                               const gpj_query_desc *queries, int32_t num_queries,
                               int64_t max_lines, const gpj_dense_result *results,
                               void *d_scratch, void *stream) {
  if (!index) return fail(GPJ_ERR_INVALID_ARGUMENT, "query: null index");
  if (num_queries > 0 && !results)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "dense queries: null results");
  if (max_lines < 0) return fail(GPJ_ERR_INVALID_ARGUMENT, "dense queries: negative line count");
  return run_batch(d_in, n, index, index->d_newline, index->d_num_lines,
                   gpj::RecordMode{0, 0, 64, index->num_lines}, queries, num_queries, max_lines,
                   stream, results, d_scratch);
}

uint64_t gpj_row_bounds_scratch_bytes(int64_t level_size) {
  return level_size > 0 ? div_up((uint64_t)level_size, gpj::kTileWords) * 8 : 8;
}

gpj_status gpj_row_bounds(const gpj_index *index, int64_t *d_bounds, int64_t capacity,
                          int64_t *d_num_bounds, void *d_scratch, void *stream) {
  if (!index || !d_num_bounds || (capacity > 0 && !d_bounds) || !d_scratch)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_row_bounds: null pointer");
  if (!index->d_leveled || index->num_levels < 1 || index->level_size < 0)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_row_bounds: the index has no leveled rows");
  hipStream_t st = as_stream(stream);
  const uint64_t words = (uint64_t)index->level_size;
  const uint64_t tiles = div_up(words, gpj::kTileWords);
  uint64_t *counts = static_cast<uint64_t *>(d_scratch);
  if (tiles == 0) {
    hipLaunchKernelGGL(gpj::newline_scan_kernel, dim3(1), dim3(1024), 0, st, counts,
                       (uint64_t)0, d_num_bounds, (int64_t)0);
  } else {
    hipLaunchKernelGGL(gpj::row_count_kernel, dim3((uint32_t)tiles), dim3(gpj::kTileWords), 0,
                       st, index->d_leveled, words, counts);
    hipLaunchKernelGGL(gpj::newline_scan_kernel, dim3(1), dim3(1024), 0, st, counts, tiles,
                       d_num_bounds, (int64_t)0);
    hipLaunchKernelGGL(gpj::row_write_kernel, dim3((uint32_t)tiles), dim3(gpj::kTileWords), 0,
                       st, index->d_leveled, words, counts, d_bounds, capacity);
  }
  GPJ_HIP(hipGetLastError(), "row bounds launch");
  return GPJ_OK;
}

gpj_status gpj_query_run_records(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                                 const int64_t *d_bounds, const int64_t *d_num_bounds,
                                 const gpj_query_desc *queries, int32_t num_queries,
                                 int64_t max_records, void *stream) {
  if (!index) return fail(GPJ_ERR_INVALID_ARGUMENT, "query: null index");
  if (index->num_levels < 2)
    return fail(GPJ_ERR_QUERY_EXECUTION,
                "array records need an index with at least 2 leveled rows");
  // the bounds buffer holds at least max_records + 1 entries
  return run_batch(d_in, n, index, d_bounds, d_num_bounds,
                   gpj::RecordMode{1, 1, 64, max_records + 1}, queries, num_queries, max_records,
                   stream);
}

gpj_status gpj_query_run_records_dense(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                                       const int64_t *d_bounds, const int64_t *d_num_bounds,
                                       const gpj_query_desc *queries, int32_t num_queries,
                                       int64_t max_records, const gpj_dense_result *results,
                                       void *d_scratch, void *stream) {
  if (!index) return fail(GPJ_ERR_INVALID_ARGUMENT, "query: null index");
  if (index->num_levels < 2)
    return fail(GPJ_ERR_QUERY_EXECUTION,
                "array records need an index with at least 2 leveled rows");
  if (num_queries > 0 && !results)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "dense queries: null results");
  if (max_records < 0)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "dense queries: negative record count");
  return run_batch(d_in, n, index, d_bounds, d_num_bounds,
                   gpj::RecordMode{1, 1, 64, max_records + 1}, queries, num_queries, max_records,
                   stream, results, d_scratch);
}

gpj_status gpj_query_run(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                         const uint8_t *d_ir, uint32_t ir_len, int32_t num_results,
                         int64_t max_lines, int64_t *d_spans, void *stream) {
  const gpj_query_desc d{d_ir, ir_len, num_results, 0, 0, d_spans, nullptr};
  return gpj_query_run_batch(d_in, n, index, &d, 1, max_lines, stream);
}

gpj_status gpj_query_run_ext(const uint8_t *d_in, uint64_t n, const gpj_index *index,
                             const uint8_t *d_ir, uint32_t ir_len, int32_t capacity,
                             int64_t max_lines, int64_t *d_spans, int64_t *d_counts,
                             void *stream) {
  const gpj_query_desc d{d_ir, ir_len, capacity, 1, 0, d_spans, d_counts};
  return gpj_query_run_batch(d_in, n, index, &d, 1, max_lines, stream);
}

static_assert(sizeof(gpj_batch_file) == sizeof(gpj::BatchFile), "batch file layout");
static_assert(GPJ_BATCH_ALIGN == gpj::kTileBytes, "files start on index tiles");
static_assert(GPJ_BATCH_SLOTS == gpj::kBatchSlots, "result slots");

uint64_t gpj_batch_scratch_bytes(uint64_t total_bytes) { return scratch_bytes(total_bytes); }

gpj_status gpj_batch_tile_map(const gpj_batch_file *files, int32_t num_files,
                              uint64_t total_bytes, uint32_t *tile_file) {
  if (num_files <= 0 || !files || !tile_file)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_tile_map: empty batch or null pointer");
  const uint64_t tiles = div_up(total_bytes, gpj::kTileBytes);
  uint64_t next_free = 0;
  for (int32_t f = 0; f < num_files; ++f) {
    const gpj_batch_file &b = files[f];
    if (b.bytes == 0)
      return fail(GPJ_ERR_INVALID_ARGUMENT, "batch file " + std::to_string(f) + " is empty");
    if (b.offset % gpj::kTileBytes != 0 || b.offset < next_free ||
        b.offset + b.bytes > total_bytes)
      return fail(GPJ_ERR_INVALID_ARGUMENT,
                  "batch file " + std::to_string(f) +
                      ": offsets must be GPJ_BATCH_ALIGN multiples, increasing, inside the "
                      "buffer and not overlapping");
    if (b.newline_capacity < 1)
      return fail(GPJ_ERR_INVALID_ARGUMENT, "batch file newline capacity must be >= 1");
    if (__builtin_popcount(b.query_mask) > GPJ_BATCH_SLOTS)
      return fail(GPJ_ERR_INVALID_ARGUMENT, "a batch file runs at most 2 queries");
    next_free = div_up(b.offset + b.bytes, gpj::kTileBytes) * gpj::kTileBytes;
  }
  int32_t f = 0;
  for (uint64_t t = 0; t < tiles; ++t) {
    while (f + 1 < num_files && files[f + 1].offset / gpj::kTileBytes <= t) ++f;
    tile_file[t] = (uint32_t)f;
  }
  return GPJ_OK;
}

gpj_status gpj_batch_line_counts(const uint8_t *d_in, uint64_t total_bytes,
                                 const gpj_batch_file *d_files, int32_t num_files,
                                 const uint32_t *d_tile_file, int64_t *d_line_prefix,
                                 void *d_scratch, void *stream) {
  if (num_files <= 0 || total_bytes == 0 || !d_in || !d_files || !d_tile_file ||
      !d_line_prefix || !d_scratch)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_line_counts: empty batch or null pointer");
  if (reinterpret_cast<uintptr_t>(d_in) % 64 != 0)
    return fail(GPJ_ERR_INVALID_ARGUMENT,
                "gpj_batch_line_counts: batch buffer must be 64-byte aligned");
  const uint64_t tiles = tiles_for(total_bytes);
  if (tiles > 0xFFFFFFFFull)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "batch too large for one count launch");
  hipStream_t st = as_stream(stream);
  uint64_t *counts = static_cast<uint64_t *>(d_scratch);
  const auto *files = reinterpret_cast<const gpj::BatchFile *>(d_files);
  hipLaunchKernelGGL(gpj::batch_count_kernel, dim3((uint32_t)tiles), dim3(gpj::kTileWords), 0, st,
                     d_in, files, d_tile_file, counts);
  hipLaunchKernelGGL(gpj::batch_file_lines_kernel, dim3((uint32_t)div_up((uint64_t)num_files, 256)),
                     dim3(256), 0, st, files, num_files, counts, d_line_prefix);
  // line counts -> exclusive prefix, total in entry num_files
  hipLaunchKernelGGL(gpj::newline_scan_kernel, dim3(1), dim3(1024), 0, st,
                     reinterpret_cast<uint64_t *>(d_line_prefix), (uint64_t)num_files,
                     d_line_prefix + num_files, (int64_t)0);
  GPJ_HIP(hipGetLastError(), "batch line count launch");
  return GPJ_OK;
}

gpj_status gpj_batch_index(const uint8_t *d_in, uint64_t total_bytes,
                           const gpj_batch_file *d_files, int32_t num_files,
                           const uint32_t *d_tile_file, int32_t levels, uint64_t *d_string,
                           uint64_t *d_leveled, int64_t *d_newline, int64_t *d_line_prefix,
                           void *d_scratch, void *stream) {
  if (levels < 1 || levels > GPJ_MAX_INDEX_LEVELS)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "batch index levels must be in [1, 22]");
  if (num_files <= 0 || total_bytes == 0 || !d_in || !d_files || !d_tile_file || !d_string ||
      !d_leveled || !d_newline || !d_line_prefix || !d_scratch)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_index: empty batch or null pointer");
  if (reinterpret_cast<uintptr_t>(d_in) % 64 != 0)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_index: batch buffer must be 64-byte aligned");
  hipStream_t st = as_stream(stream);
  const uint64_t words = div_up(total_bytes, 64);
  const uint64_t tiles = tiles_for(total_bytes);
  if (tiles > 0xFFFFFFFFull)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "batch too large for one index launch");
  uint64_t *desc = static_cast<uint64_t *>(d_scratch);
  uint32_t *counter = reinterpret_cast<uint32_t *>(desc + 2 * tiles);
  const uint64_t clear_blocks = std::min<uint64_t>(div_up(tiles, 256), 2048);
  hipLaunchKernelGGL(gpj::index_clear_kernel, dim3((uint32_t)clear_blocks), dim3(256), 0, st,
                     desc, 2 * tiles, counter);
  GPJ_HIP(hipGetLastError(), "index_clear_kernel launch");
  const gpj::BatchSeg seg{d_tile_file, reinterpret_cast<const gpj::BatchFile *>(d_files),
                          d_line_prefix};
  gpj_status r;
#define GPJ_SEG(L)                                                                         \
  r = launch_index<L, true, gpj::kTileHalves>(d_in, total_bytes, words, levels, d_string, d_leveled, d_newline, \
                            0, nullptr, desc, counter, tiles, st, seg)
  if (levels <= 2) GPJ_SEG(2);
  else if (levels <= 4) GPJ_SEG(4);
  else if (levels <= 8) GPJ_SEG(8);
  else if (levels <= 16) GPJ_SEG(16);
  else GPJ_SEG(22);
#undef GPJ_SEG
  if (r != GPJ_OK) return r;
  // line counts -> exclusive prefix, total in entry num_files
  hipLaunchKernelGGL(gpj::newline_scan_kernel, dim3(1), dim3(1024), 0, st,
                     reinterpret_cast<uint64_t *>(d_line_prefix), (uint64_t)num_files,
                     d_line_prefix + num_files, (int64_t)0);
  GPJ_HIP(hipGetLastError(), "line prefix launch");
  // a failed build reports -1 total lines (its queries then run no record)
  hipLaunchKernelGGL(gpj::index_status_kernel, dim3(1), dim3(64), 0, st, counter + 1,
                     d_line_prefix + num_files);
  GPJ_HIP(hipGetLastError(), "index_status_kernel launch");
  return GPJ_OK;
}

gpj_status gpj_batch_query(const uint8_t *d_in, uint64_t total_bytes,
                           const gpj_batch_file *d_files, int32_t num_files, int32_t levels,
                           const uint64_t *d_string, const uint64_t *d_leveled,
                           const int64_t *d_newline, const int64_t *d_line_prefix,
                           int64_t max_records, const gpj_query_desc *queries,
                           int32_t num_queries, const gpj_batch_results *results,
                           void *stream) {
  return gpj_batch_query_shaped(d_in, total_bytes, d_files, num_files, levels, d_string,
                                d_leveled, d_newline, d_line_prefix, max_records, 64, queries,
                                num_queries, results, stream);
}

gpj_status gpj_batch_query_shaped(const uint8_t *d_in, uint64_t total_bytes,
                                  const gpj_batch_file *d_files, int32_t num_files,
                                  int32_t levels, const uint64_t *d_string,
                                  const uint64_t *d_leveled, const int64_t *d_newline,
                                  const int64_t *d_line_prefix, int64_t max_records,
                                  int32_t records_per_wave, const gpj_query_desc *queries,
                                  int32_t num_queries, const gpj_batch_results *results,
                                  void *stream) {
  if (records_per_wave < 1 || records_per_wave > 64)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_query: records per wave 1..64");
  if (num_files <= 0 || !d_in || !d_files || !d_string || !d_leveled || !d_newline ||
      !d_line_prefix || !results)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_query: empty batch or null pointer");
  if (num_queries < 1 || num_queries > GPJ_MAX_BATCH_QUERIES || !queries)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_query: 1..8 queries");
  if (results->capacity < 1)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_query: result capacity must be >= 1");
  for (int k = 0; k < GPJ_BATCH_SLOTS; ++k)
    if (!results->d_spans[k] || !results->d_counts[k])
      return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_query: null result slot");
  gpj::QueryBatch batch{};
  uint32_t total_ir = 0;
  for (int32_t q = 0; q < num_queries; ++q) {
    const gpj_query_desc &d = queries[q];
    if (d.ir_len == 0 || d.ir_len > GPJ_MAX_QUERY_IR_BYTES || !d.d_ir)
      return fail(GPJ_ERR_QUERY_EXECUTION, "Compiled query IR exceeds the 4096-byte IR buffer");
    total_ir += d.ir_len;
    batch.q[q] = gpj::QuerySlot{d.d_ir, d.ir_len, results->capacity, d.variable ? 1 : 0,
                                nullptr, nullptr};
  }
  if (total_ir > gpj::kBatchIrBytes)
    return fail(GPJ_ERR_QUERY_EXECUTION, "query batch bytecode exceeds 8192 bytes; split it");
  batch.count = num_queries;
  if (max_records <= 0) return GPJ_OK;
  if (total_bytes >= (1ull << 35) ||
      (uint64_t)levels * div_up(total_bytes, 64) >= 0xFFFFFFF0ull)
    return fail(GPJ_ERR_QUERY_EXECUTION,
                "batch above 32 GiB per launch; split it into several batches");
  const gpj::FilesMode fm{reinterpret_cast<const gpj::BatchFile *>(d_files), num_files,
                          results->capacity, d_line_prefix,
                          {results->d_spans[0], results->d_spans[1]},
                          {results->d_counts[0], results->d_counts[1]}, records_per_wave};
  const uint64_t blocks =
      std::min<uint64_t>(div_up((uint64_t)max_records, (uint64_t)records_per_wave), 1u << 22);
  const int64_t stride = (int64_t)div_up(total_bytes, 64);
  const int deepest = deepest_level(queries, num_queries);
  // every file is at most total_bytes long (see run_batch for the two kernels)
  const int64_t span = g_narrow_span.load();
  const bool fixup = (int64_t)total_bytes > span;
#define GPJ_FLAUNCH(LV, NW, FIX)                                                          \
  hipLaunchKernelGGL((gpj::query_files_kernel<LV, NW>), dim3((uint32_t)blocks), dim3(64), 0, \
                     as_stream(stream), d_in, d_newline, d_string, d_leveled, stride,     \
                     levels, fm, gpj::PosMode{span, FIX}, batch)
#define GPJ_FPAIR(LV)                                                                     \
  do {                                                                                    \
    GPJ_FLAUNCH(LV, true, 0);                                                             \
    if (fixup) GPJ_FLAUNCH(LV, false, 1);                                                 \
  } while (0)
  if (deepest + 1 <= 4) GPJ_FPAIR(4);
  else if (deepest + 1 <= 6) GPJ_FPAIR(6);
  else if (deepest + 1 <= 8) GPJ_FPAIR(8);
  else GPJ_FPAIR(16);
#undef GPJ_FPAIR
#undef GPJ_FLAUNCH
  GPJ_HIP(hipGetLastError(), "query_files_kernel launch");
  return GPJ_OK;
}

gpj_status gpj_batch_per_file(const uint8_t *d_in, const gpj_batch_file *files,
                              int32_t num_files, int32_t levels, uint64_t *d_string,
                              uint64_t *d_leveled_per_file, int64_t *d_newline,
                              int64_t *d_line_counts, const int64_t *line_prefix,
                              const gpj_query_desc *queries, int32_t num_queries,
                              const gpj_batch_results *results, void *d_scratch,
                              void *stream) {
  if (num_files <= 0 || !d_in || !files || !d_string || !d_leveled_per_file || !d_newline ||
      !d_line_counts || !line_prefix || !queries || !results || !d_scratch)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_per_file: empty batch or null pointer");
  if (num_queries < 1 || num_queries > GPJ_MAX_BATCH_QUERIES)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_batch_per_file: 1..8 queries");
  hipStream_t st = as_stream(stream);
  const int32_t cap = results->capacity;
  uint64_t lv_words = 0;
  for (int32_t f = 0; f < num_files; ++f) {
    const gpj_batch_file &b = files[f];
    const uint64_t w = div_up(b.bytes, 64);
    uint64_t *lv = d_leveled_per_file + lv_words * (uint64_t)levels;
    lv_words += w;
    int64_t *nl = d_newline + b.newline_offset;
    gpj_status r = gpj_index_build_into(d_in + b.offset, b.bytes, levels,
                                        d_string + b.offset / 64, lv, nl,
                                        (int64_t)b.newline_capacity, d_line_counts + f,
                                        d_scratch, stream);
    if (r != GPJ_OK) return r;
    gpj_query_desc descs[GPJ_BATCH_SLOTS];
    int k = 0;
    for (int32_t q = 0; q < num_queries && k < GPJ_BATCH_SLOTS; ++q) {
      if (!((b.query_mask >> q) & 1u)) continue;
      gpj_query_desc d = queries[q];
      d.capacity = cap;
      d.d_spans = results->d_spans[k] + line_prefix[f] * 2 * (int64_t)cap;
      d.d_counts = results->d_counts[k] + line_prefix[f];
      descs[k++] = d;
    }
    const int64_t lines = line_prefix[f + 1] - line_prefix[f];
    for (int kk = k; kk < GPJ_BATCH_SLOTS; ++kk) {
      hipLaunchKernelGGL(fill_i64_kernel, dim3((uint32_t)std::min<uint64_t>(div_up(lines, 256), 1024)),
                         dim3(256), 0, st, results->d_counts[kk] + line_prefix[f], lines,
                         (int64_t)0);
      GPJ_HIP(hipGetLastError(), "fill launch");
    }
    gpj_index idx{nl, d_line_counts + f, (int64_t)b.newline_capacity, d_string + b.offset / 64,
                  lv, (int64_t)w, levels, 0};
    if (k > 0) {
      r = gpj_query_run_batch(d_in + b.offset, b.bytes, &idx, descs, k, lines, stream);
      if (r != GPJ_OK) return r;
    }
  }
  return GPJ_OK;
}

gpj_status gpj_stream_create(void **stream) {
  if (!stream) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_stream_create: null");
  hipStream_t s = nullptr;
  GPJ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  *stream = s;
  return GPJ_OK;
}

gpj_status gpj_stream_destroy(void *stream) {
  if (stream) GPJ_HIP(hipStreamDestroy(as_stream(stream)), "hipStreamDestroy");
  return GPJ_OK;
}

gpj_status gpj_graph_begin(void *stream) {
  if (!stream) return fail(GPJ_ERR_INVALID_ARGUMENT, "graph capture needs a non-default stream");
  GPJ_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal),
          "hipStreamBeginCapture");
  return GPJ_OK;
}

gpj_status gpj_graph_end(void *stream, void **graph_exec) {
  if (!stream || !graph_exec) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_graph_end: null");
  hipGraph_t g = nullptr;
  GPJ_HIP(hipStreamEndCapture(as_stream(stream), &g), "hipStreamEndCapture");
  hipGraphExec_t e = nullptr;
  const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  GPJ_HIP(err, "hipGraphInstantiate");
  *graph_exec = e;
  return GPJ_OK;
}

gpj_status gpj_graph_launch(void *graph_exec, void *stream) {
  if (!graph_exec) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_graph_launch: null graph");
  GPJ_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph_exec), as_stream(stream)),
          "hipGraphLaunch");
  return GPJ_OK;
}

gpj_status gpj_graph_destroy(void *graph_exec) {
  if (graph_exec)
    GPJ_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph_exec)),
            "hipGraphExecDestroy");
  return GPJ_OK;
}

gpj_status gpj_query_compile(const char *query, uint8_t *ir, uint32_t ir_cap,
                             uint32_t *ir_len, uint32_t *max_depth, uint32_t *num_slots,
                             uint32_t flags) {
  if (!query || !ir_len || !max_depth || !num_slots)
    return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_query_compile: null pointer");
  gpj::CompiledBytecode out;
  std::string err;
  const gpj::CompileStatus s =
      gpj::compile_query(query, (flags & GPJ_COMPILE_EXTENSIONS) != 0, &out, &err);
  switch (s) {
  case gpj::CompileStatus::kOk: break;
  case gpj::CompileStatus::kMalformed: return fail(GPJ_ERR_MALFORMED_QUERY, err);
  case gpj::CompileStatus::kUnsupported: return fail(GPJ_ERR_UNSUPPORTED_QUERY, err);
  case gpj::CompileStatus::kInvalidArgument: return fail(GPJ_ERR_INVALID_ARGUMENT, err);
  case gpj::CompileStatus::kOutOfRange: return fail(GPJ_ERR_OUT_OF_RANGE, err);
  }
  *ir_len = (uint32_t)out.bytes.size();
  *max_depth = (uint32_t)out.max_depth;
  *num_slots = (uint32_t)out.num_slots;
  if (out.bytes.size() > ir_cap)
    return fail(GPJ_ERR_CAPACITY, "compiled query does not fit the IR buffer");
  if (!out.bytes.empty()) {
    if (!ir) return fail(GPJ_ERR_INVALID_ARGUMENT, "gpj_query_compile: null IR buffer");
    std::memcpy(ir, out.bytes.data(), out.bytes.size());
  }
  return GPJ_OK;
}

} // extern "C"
