// runtime.cpp -- library initialisation, the single-process communicator,
// contiguous maps, contexts, scratch memory and the HBM STREAM benchmarks.
//
// Reference behaviour: src/kernels/phist_kernels.h:79-185 (interface),
// src/kernels/builtin/kernels.cpp:192-230 (init: thread setup + RNG init via
// kernels_common_impl.cpp:23-88), src/kernels/builtin/map_module.F90:41-66,
// 289-320 (contiguous distribution; ilower/iupper are 0-based).
#include "amd_common.hpp"
#include <chrono>
#include <cstdarg>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>
#include <cstdlib>
#include <algorithm>

namespace phx {

void fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

Runtime& rt() {
  static Runtime r;
  return r;
}

// Device scratch of the allocation-free entry points (reduction partials of
// one call).  Host code never names it after initialisation: kernels and
// copies get the address hipGetSymbolAddress returned for the device.
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

static void resolve_noalloc_scratch(Runtime& r) {
  PHX_REQUIRE(r.device >= 0 && r.device < Runtime::kMaxDevices, PHIST_FUNCTIONAL_ERROR,
              "device %d beyond the %d devices the fixed scratch table holds", r.device, Runtime::kMaxDevices);
  if (r.noalloc_scratch[r.device]) return;
  void* p = nullptr;
  PHX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  PHX_REQUIRE(p != nullptr, PHIST_FUNCTIONAL_ERROR, "no device address for the fixed scratch array");
  r.noalloc_scratch[r.device] = p;
}

void* noalloc_scratch() {
  Runtime& r = rt();
  PHX_REQUIRE(r.initialized, PHIST_FUNCTIONAL_ERROR, "library not initialised");
  void* p = r.noalloc_scratch[r.device];
  PHX_REQUIRE(p != nullptr, PHIST_FUNCTIONAL_ERROR, "fixed scratch of device %d not resolved", r.device);
  return p;
}

NoAllocScope::NoAllocScope() {
  Runtime& r = rt();
  PHX_REQUIRE(!r.capturing, PHIST_FUNCTIONAL_ERROR, "allocation-free call inside the library's own graph capture");
  r.fixed_work = noalloc_scratch();
  r.fixed_work_bytes = kNoAllocScratchBytes;
}
NoAllocScope::~NoAllocScope() {
  Runtime& r = rt();
  r.fixed_work = nullptr;
  r.fixed_work_bytes = 0;
}

void ensure_init() {
  Runtime& r = rt();
  if (r.initialized) return;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  PHX_REQUIRE(e == hipSuccess && ndev > 0, PHIST_FUNCTIONAL_ERROR,
              "no HIP device available (the kernel library runs on the GPU only)");
  PHX_HIP(hipGetDevice(&r.device));
  hipDeviceProp_t prop;
  PHX_HIP(hipGetDeviceProperties(&prop, r.device));
  r.num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (prop.sharedMemPerBlock > 0) r.lds_per_block = (int64_t)prop.sharedMemPerBlock;
  PHX_HIP(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
  PHX_HIP(hipEventCreate(&r.ev_start));
  PHX_HIP(hipEventCreate(&r.ev_stop));
  PHX_HIP(hipEventCreate(&r.kt_start));
  PHX_HIP(hipEventCreate(&r.kt_stop));
  if (const char* s = getenv("PHIST_AMD_SELL_C")) r.sell_C = atoi(s);
  if (const char* s = getenv("PHIST_AMD_SELL_SIGMA")) r.sell_sigma = atoi(s);
  if (const char* s = getenv("PHIST_AMD_GRID_BLOCKS")) r.grid_blocks = atoi(s);
  if (const char* s = getenv("PHIST_AMD_SPMV_BLOCKS_PER_CU")) r.spmv_blocks_per_cu = atoi(s);
  if (const char* s = getenv("PHIST_AMD_SPMV_LANE_BYTES")) r.spmv_lane_bytes = atoi(s);
  if (const char* s = getenv("PHIST_AMD_SPMV_DEPTH")) r.spmv_depth = atoi(s);
  if (const char* s = getenv("PHIST_AMD_SPMV_COL_BLOCK")) r.spmv_col_block = std::max(-1, atoi(s));
  if (const char* s = getenv("PHIST_AMD_SPMV_STAGE")) r.spmv_stage = std::max(-1, std::min(1, atoi(s)));
  if (const char* s = getenv("PHIST_AMD_FAST_SYNC")) r.fast_sync = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_SPMV_NT_Y")) r.spmv_nt_y = atoi(s) < 0 ? -1 : (atoi(s) != 0 ? 1 : 0);
  if (const char* s = getenv("PHIST_AMD_GRAM_MFMA")) r.gram_mfma = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_GRAM_MFMA_WIDE")) r.gram_mfma_wide = std::max(1, atoi(s));
  if (const char* s = getenv("PHIST_AMD_MIX_MFMA_MIN_M")) r.mix_mfma_min_m = std::max(1, atoi(s));
  if (const char* s = getenv("PHIST_AMD_GRAM_BLOCKS_PER_CU")) r.gram_blocks_per_cu = atoi(s);
  if (const char* s = getenv("PHIST_AMD_MIX_MFMA")) r.mix_mfma = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_MINRES_DEVICE")) r.minres_device = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_GMRES_CGS")) r.gmres_cgs = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_GMRES_DEVICE")) r.gmres_device = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_SERVE_MAX_REPS")) r.serve_max_reps = std::max(1, atoi(s));
  if (const char* s = getenv("PHIST_AMD_GRAM_ROWS")) r.gram_rows = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_GRAM_SETS")) r.gram_sets = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_MINRES_GRAPH")) r.minres_graphs = atoi(s) != 0;
  if (const char* s = getenv("PHIST_AMD_MINRES_ITERS_PER_CALL")) r.minres_iters_per_call = std::max(0, atoi(s));
  if (const char* s = getenv("PHIST_AMD_MINRES_GRAPH_LEN")) r.minres_graph_len = std::max(1, atoi(s));
  PHX_HIP(hipSetDevice(r.device));
  resolve_noalloc_scratch(r);
  kiss_reset();
  r.initialized = true;
}

void* device_work(size_t bytes) {
  Runtime& r = rt();
  if (r.fixed_work) {
    PHX_REQUIRE(bytes <= r.fixed_work_bytes, PHIST_INVALID_INPUT,
                "the call needs %zu bytes of device scratch, the allocation-free entry point has %zu", bytes,
                r.fixed_work_bytes);
    return r.fixed_work;
  }
  if (r.capturing) {
    r.work_need = std::max(r.work_need, bytes);
    return r.work_arena;
  }
  if (bytes > r.d_work_bytes) {
    if (r.d_work) {
      PHX_HIP(hipStreamSynchronize(r.stream));
      PHX_HIP(hipFree(r.d_work));
      r.d_work = nullptr;
    }
    size_t nb = std::max<size_t>(bytes, 1 << 20);
    PHX_HIP(hipMalloc(&r.d_work, nb));
    r.d_work_bytes = nb;
  }
  return r.d_work;
}

void* host_work(size_t bytes) {
  Runtime& r = rt();
  PHX_REQUIRE(!r.fixed_work, PHIST_FUNCTIONAL_ERROR, "host-side reduction inside an allocation-free call");
  PHX_REQUIRE(!r.capturing, PHIST_FUNCTIONAL_ERROR, "a host-side reduction cannot be part of a captured graph");
  if (bytes > r.h_work_bytes) {
    if (r.h_work) {
      PHX_HIP(hipStreamSynchronize(r.stream));
      PHX_HIP(hipHostFree(r.h_work));
      r.h_work = nullptr;
    }
    size_t nb = std::max<size_t>(bytes, 1 << 16);
    PHX_HIP(hipHostMalloc(&r.h_work, nb, hipHostMallocDefault));
    r.h_work_bytes = nb;
  }
  return r.h_work;
}

bool signal_buffer(void** host, void** dev) {
  Runtime& r = rt();
  if (!r.fast_sync || r.fixed_work || r.capturing) return false;
  if (!r.h_sig) {
    void* h = nullptr;
    if (hipHostMalloc(&h, Runtime::kSignalBytes, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess || !h) {
      (void)hipGetLastError();  // not fatal: the copy + synchronisation path serves instead
      r.fast_sync = 0;
      return false;
    }
    void* d = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
    if (e != hipSuccess || !d) {
      (void)hipGetLastError();
      PHX_HIP(hipHostFree(h));
      r.fast_sync = 0;
      return false;
    }
    memset(h, 0, Runtime::kSignalBytes);
    r.h_sig = h;
    r.d_sig = d;
  }
  *host = r.h_sig;
  *dev = r.d_sig;
  return true;
}

void wait_signal(uint32_t seq) {
  Runtime& r = rt();
  const uint32_t* word = static_cast<const uint32_t*>(r.h_sig);
  const auto t0 = std::chrono::steady_clock::now();
  // the numbers grow in stream order: a later one means this one was written too
  auto reached = [&] { return (int32_t)(__atomic_load_n(word, __ATOMIC_ACQUIRE) - seq) >= 0; };
  for (int spin = 1;; spin++) {
    if (reached()) return;
    if ((spin & 63) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) break;
    __builtin_ia32_pause();
  }
  // long waits (much queued work) block in the runtime instead of spinning
  PHX_HIP(hipStreamSynchronize(r.stream));
  PHX_REQUIRE(reached(), PHIST_FUNCTIONAL_ERROR, "published reduction %u not seen after the stream finished", seq);
}

namespace {
// one wave copies nwords words to the published-result buffer, eight per lane
// in flight, then stores the sequence number (system-scope release)
template <typename W>
__global__ void __launch_bounds__(64) k_publish_words(const W* __restrict__ src, int64_t nwords, W* pub, uint32_t* word,
                                                      uint32_t seq) {
  constexpr int U = 8;
  for (int64_t b = 0; b < nwords; b += 64 * U) {
    W v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = b + u * 64 + threadIdx.x;
      if (i < nwords) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = b + u * 64 + threadIdx.x;
      if (i < nwords) pub[i] = v[u];
    }
  }
  if (threadIdx.x == 0) __hip_atomic_store(word, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
}  // namespace

namespace {
// enqueues the publication of bytes from dev at byte offset off of the buffer; the sequence number, or 0
uint32_t publish_at(const void* dev, size_t bytes, size_t off, void** host_area) {
  void *hs = nullptr, *ds = nullptr;
  const uintptr_t al = reinterpret_cast<uintptr_t>(dev) | bytes;
  if (bytes == 0 || al % 4 != 0 || !signal_buffer(&hs, &ds)) return 0;
  Runtime& r = rt();
  uint32_t seq = ++r.sig_seq;
  if (seq == 0) seq = ++r.sig_seq;  // 0 means "not published"
  char* pub = static_cast<char*>(ds) + off;
  uint32_t* word = static_cast<uint32_t*>(ds);
  typedef unsigned int w4_t __attribute__((ext_vector_type(4)));
  if (al % 16 == 0)
    hipLaunchKernelGGL(k_publish_words<w4_t>, dim3(1), dim3(64), 0, r.stream, static_cast<const w4_t*>(dev),
                       (int64_t)(bytes / 16), reinterpret_cast<w4_t*>(pub), word, seq);
  else
    hipLaunchKernelGGL(k_publish_words<uint32_t>, dim3(1), dim3(64), 0, r.stream, static_cast<const uint32_t*>(dev),
                       (int64_t)(bytes / 4), reinterpret_cast<uint32_t*>(pub), word, seq);
  PHX_HIP(hipGetLastError());
  *host_area = static_cast<char*>(hs) + off;
  return seq;
}
}  // namespace

// the buffer: sequence word, then slot 0 (fetch_to_host, up to the end minus two slots), then slots 1, 2
bool fetch_to_host(void* host, const void* dev, size_t bytes) {
  if (bytes > Runtime::kSignalBytes - Runtime::kSignalData - 2 * kSignalSlotBytes) return false;
  void* area = nullptr;
  const uint32_t seq = publish_at(dev, bytes, Runtime::kSignalData, &area);
  if (!seq) return false;
  wait_signal(seq);
  memcpy(host, area, bytes);
  return true;
}

uint32_t publish_async(const void* dev, size_t bytes, int slot) {
  PHX_REQUIRE(slot == 1 || slot == 2, PHIST_INVALID_INPUT, "published slot %d (1 or 2)", slot);
  if (bytes > kSignalSlotBytes) return 0;
  void* area = nullptr;
  return publish_at(dev, bytes, Runtime::kSignalBytes - (size_t)(3 - slot) * kSignalSlotBytes, &area);
}

const void* published_data(uint32_t seq, int slot) {
  PHX_REQUIRE(seq != 0 && (slot == 1 || slot == 2), PHIST_INVALID_INPUT, "published slot %d, number %u", slot, seq);
  wait_signal(seq);
  return static_cast<const char*>(rt().h_sig) + Runtime::kSignalBytes - (size_t)(3 - slot) * kSignalSlotBytes;
}

void sync_stream() { PHX_HIP(hipStreamSynchronize(rt().stream)); }

// The occupancy query costs microseconds of host time per call, as much as a
// launch; small solves launch thousands of kernels, so the answers are kept.
int occupancy_per_cu(const void* kernel, int threads, size_t lds_bytes) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, int, size_t>, int> known;
  const auto key = std::make_tuple(rt().device, kernel, threads, lds_bytes);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(key);
    if (it != known.end()) return it->second;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds_bytes) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  std::lock_guard<std::mutex> lk(mu);
  known[key] = per_cu;
  return per_cu;
}

bool stage_upload_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t row_bytes, size_t cols) {
  Runtime& r = rt();
  if (cols == 0 || row_bytes == 0) return true;
  const size_t need = spitch * (cols - 1) + row_bytes;
  if (need > Runtime::kStageSlotBytes || r.capturing) return false;
  if (!r.stage_base) {
    void* p = nullptr;
    PHX_HIP(hipHostMalloc(&p, Runtime::kStageSlots * Runtime::kStageSlotBytes, hipHostMallocDefault));
    r.stage_base = static_cast<char*>(p);
    for (int s = 0; s < Runtime::kStageSlots; s++)
      PHX_HIP(hipEventCreateWithFlags(&r.stage_ev[s], hipEventDisableTiming));
  }
  const int s = r.stage_next;
  r.stage_next = (s + 1) % Runtime::kStageSlots;
  if (r.stage_rec[s]) PHX_HIP(hipEventSynchronize(r.stage_ev[s]));  // normally long done
  char* slot = r.stage_base + (size_t)s * Runtime::kStageSlotBytes;
  const char* in = static_cast<const char*>(src);
  if (row_bytes == spitch && spitch == dpitch) {
    memcpy(slot, in, need);
    PHX_HIP(hipMemcpyAsync(dst, slot, need, hipMemcpyHostToDevice, r.stream));
  } else {
    for (size_t j = 0; j < cols; j++) memcpy(slot + j * spitch, in + j * spitch, row_bytes);
    PHX_HIP(hipMemcpy2DAsync(dst, dpitch, slot, spitch, row_bytes, cols, hipMemcpyHostToDevice, r.stream));
  }
  PHX_HIP(hipEventRecord(r.stage_ev[s], r.stream));
  r.stage_rec[s] = true;
  return true;
}

namespace {
// size classes: powers of two up to 2 MiB, then multiples of 2 MiB
size_t size_class(size_t bytes) {
  constexpr size_t kBig = size_t(2) << 20;
  if (bytes <= 4096) return 4096;
  if (bytes >= kBig) return (bytes + kBig - 1) / kBig * kBig;
  size_t c = 4096;
  while (c < bytes) c <<= 1;
  return c;
}

struct BlockCache {
  std::map<size_t, std::vector<void*>> free;
  size_t cached = 0, cap;
  explicit BlockCache(size_t c) : cap(c) {}
  void* take(size_t cls) {
    auto it = free.find(cls);
    if (it == free.end() || it->second.empty()) return nullptr;
    void* p = it->second.back();
    it->second.pop_back();
    cached -= cls;
    return p;
  }
};

BlockCache& dev_cache() {
  static BlockCache c(size_t(32) << 30);
  return c;
}
BlockCache& pin_cache() {
  static BlockCache c(size_t(1) << 30);
  return c;
}
std::mutex& cache_mutex() {
  static std::mutex m;
  return m;
}

void trim(BlockCache& c, bool device) {
  if (c.cached == 0) return;
  PHX_HIP(hipStreamSynchronize(rt().stream));
  for (auto& kv : c.free)
    for (void* p : kv.second) {
      if (device)
        PHX_HIP(hipFree(p));
      else
        PHX_HIP(hipHostFree(p));
    }
  c.free.clear();
  c.cached = 0;
}
}  // namespace

void* dev_alloc(size_t bytes) {
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  std::lock_guard<std::mutex> lk(cache_mutex());
  if (void* p = dev_cache().take(cls)) return p;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, cls);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim(dev_cache(), true);
    e = hipMalloc(&p, cls);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(PHIST_MEM_ALLOC_FAILED, "hipMalloc of %zu bytes failed", cls);
  }
  return p;
}

void dev_release(void* p, size_t bytes) {
  if (!p) return;
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  std::lock_guard<std::mutex> lk(cache_mutex());
  BlockCache& c = dev_cache();
  if (c.cached + cls > c.cap) {
    PHX_HIP(hipStreamSynchronize(rt().stream));
    PHX_HIP(hipFree(p));
    return;
  }
  c.free[cls].push_back(p);
  c.cached += cls;
}

void* pinned_alloc(size_t bytes) {
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  std::lock_guard<std::mutex> lk(cache_mutex());
  if (void* p = pin_cache().take(cls)) return p;
  void* p = nullptr;
  hipError_t e = hipHostMalloc(&p, cls, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim(pin_cache(), false);
    e = hipHostMalloc(&p, cls, hipHostMallocDefault);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(PHIST_MEM_ALLOC_FAILED, "hipHostMalloc of %zu bytes failed", cls);
  }
  return p;
}

void pinned_release(void* p, size_t bytes) {
  if (!p) return;
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  std::lock_guard<std::mutex> lk(cache_mutex());
  PHX_HIP(hipStreamSynchronize(rt().stream));
  BlockCache& c = pin_cache();
  if (c.cached + cls > c.cap) {
    PHX_HIP(hipHostFree(p));
    return;
  }
  c.free[cls].push_back(p);
  c.cached += cls;
}

void pool_trim() {
  std::lock_guard<std::mutex> lk(cache_mutex());
  trim(dev_cache(), true);
  trim(pin_cache(), false);
}

void ktimer_begin() {
  Runtime& r = rt();
  if (r.ktimer_on) PHX_HIP(hipEventRecord(r.kt_start, r.stream));
}

void ktimer_end() {
  Runtime& r = rt();
  if (!r.ktimer_on) return;
  PHX_HIP(hipEventRecord(r.kt_stop, r.stream));
  PHX_HIP(hipEventSynchronize(r.kt_stop));
  float ms = 0;
  PHX_HIP(hipEventElapsedTime(&ms, r.kt_start, r.kt_stop));
  r.ktimer_ms += ms;
  r.ktimer_count++;
}

Comm* default_comm() {
  static Comm c;
  return &c;
}

// ---- STREAM kernels on HBM ----
namespace {
__global__ void k_stream_load(const double* __restrict__ a, int64_t n, double* out) {
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += a[i];
  if (s == 12345.678) out[0] = s;  // keeps the loads alive
}
__global__ void k_stream_store(double* __restrict__ a, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}
__global__ void k_stream_copy(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
__global__ void k_stream_triad(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c,
                               int64_t n, double s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c[i] = a[i] + s * b[i];
}

// Known-size gather: every seg-byte segment of an n-double array is read
// exactly once, LPR = seg/16 lanes per segment (one 16-byte load each), in
// the scrambled order s(i) = (i * P) mod nseg (P odd, nseg a power of two: a
// bijection), so consecutive groups of a wave touch unrelated cache lines
// as the SpMV's x gathers do, and the bytes requested are exactly 8n.
template <int LPR>
__global__ void __launch_bounds__(256) k_gather_once(const double* __restrict__ a, int64_t nseg, double* out) {
  const int q = threadIdx.x % LPR;
  const int64_t g0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / LPR;
  const int64_t gs = (int64_t)gridDim.x * blockDim.x / LPR;
  const uint64_t P = 0x9E3779B97F4A7C15ull;
  double s0 = 0, s1 = 0;
  for (int64_t g = g0; g < nseg; g += gs) {
    const int64_t seg = (int64_t)(((uint64_t)g * P) & (uint64_t)(nseg - 1));
    const double2 v = *reinterpret_cast<const double2*>(a + seg * (2 * LPR) + 2 * q);
    s0 += v.x;
    s1 += v.y;
  }
  if (s0 + s1 == 12345.678) out[0] = s0;  // keeps the loads alive
}

// Device-ceiling kernels: 16-byte accesses, four independent ones in flight
// per thread; each workgroup walks its own contiguous part of the array in
// steps of 4 x 256 vectors (16 KB), the form a bandwidth-bound kernel of this
// library can at best take.  kind 0 reads n doubles, kind 1 copies them.
template <int KIND>
__global__ void __launch_bounds__(256) k_hbm_ceiling(const double2* __restrict__ a, double2* __restrict__ b, int64_t nv,
                                                     double* out) {
  const int64_t per = (nv + gridDim.x - 1) / gridDim.x;
  const int64_t lo = blockIdx.x * per, hi = lo + per < nv ? lo + per : nv;
  double s = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += 4 * 256) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t j = i + u * 256;
      v[u] = j < hi ? a[j] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t j = i + u * 256;
      if (KIND == 0) {
        s += v[u].x + v[u].y;
      } else if (j < hi) {
        b[j] = v[u];
      }
    }
  }
  if (KIND == 0 && s == 12345.678) out[0] = s;  // keeps the loads alive
}

// kind 0 read / 1 copy of 1 GiB with k_hbm_ceiling; mean/max GB/s over reps
void run_ceiling(int kind, double* mean_bw, double* max_bw) {
  ensure_init();
  Runtime& r = rt();
  PHX_REQUIRE(kind == 0 || kind == 1, PHIST_INVALID_INPUT, "kind %d: 0 (read) or 1 (copy)", kind);
  const int64_t n = int64_t(1) << 27;  // 1 GiB
  double *a = nullptr, *b = nullptr, *c = nullptr;
  PHX_HIP(hipMalloc(&a, n * 8));
  PHX_HIP(hipMalloc(&b, n * 8));
  PHX_HIP(hipMalloc(&c, 64));
  PHX_HIP(hipMemsetAsync(a, 0, n * 8, r.stream));
  const int blocks = r.num_cu * 8;
  const int reps = 20;
  double sum = 0, best = 0;
  for (int it = -2; it < reps; it++) {
    PHX_HIP(hipEventRecord(r.ev_start, r.stream));
    if (kind == 0)
      hipLaunchKernelGGL(k_hbm_ceiling<0>, dim3(blocks), dim3(256), 0, r.stream, (const double2*)a, (double2*)b,
                         n / 2, c);
    else
      hipLaunchKernelGGL(k_hbm_ceiling<1>, dim3(blocks), dim3(256), 0, r.stream, (const double2*)a, (double2*)b,
                         n / 2, c);
    PHX_HIP(hipGetLastError());
    PHX_HIP(hipEventRecord(r.ev_stop, r.stream));
    PHX_HIP(hipEventSynchronize(r.ev_stop));
    float ms = 0;
    PHX_HIP(hipEventElapsedTime(&ms, r.ev_start, r.ev_stop));
    if (it < 0) continue;
    const double gbs = (kind == 0 ? 8.0 : 16.0) * n / (ms * 1e-3) / 1e9;
    sum += gbs;
    best = std::max(best, gbs);
  }
  PHX_HIP(hipFree(a));
  PHX_HIP(hipFree(b));
  PHX_HIP(hipFree(c));
  *mean_bw = sum / reps;
  *max_bw = best;
}

// one launch of k_gather_once over 1 GiB with seg_bytes-byte segments
// (16..256), timed over reps after two untimed launches
void run_gather(int seg_bytes, double* mean_bw, double* max_bw) {
  ensure_init();
  Runtime& r = rt();
  PHX_REQUIRE(seg_bytes >= 16 && seg_bytes <= 256 && (seg_bytes & (seg_bytes - 1)) == 0, PHIST_INVALID_INPUT,
              "segment size %d: a power of two in 16..256", seg_bytes);
  const int64_t n = int64_t(1) << 27;  // 1 GiB
  const int64_t nseg = n * 8 / seg_bytes;
  double *a = nullptr, *c = nullptr;
  PHX_HIP(hipMalloc(&a, n * 8));
  PHX_HIP(hipMalloc(&c, 64));
  PHX_HIP(hipMemsetAsync(a, 0, n * 8, r.stream));
  const int blocks = r.num_cu * 8;
  const int reps = 20;
  double sum = 0, best = 0;
  for (int it = -2; it < reps; it++) {
    PHX_HIP(hipEventRecord(r.ev_start, r.stream));
    switch (seg_bytes) {
      case 16: hipLaunchKernelGGL(k_gather_once<1>, dim3(blocks), dim3(256), 0, r.stream, a, nseg, c); break;
      case 32: hipLaunchKernelGGL(k_gather_once<2>, dim3(blocks), dim3(256), 0, r.stream, a, nseg, c); break;
      case 64: hipLaunchKernelGGL(k_gather_once<4>, dim3(blocks), dim3(256), 0, r.stream, a, nseg, c); break;
      case 128: hipLaunchKernelGGL(k_gather_once<8>, dim3(blocks), dim3(256), 0, r.stream, a, nseg, c); break;
      default: hipLaunchKernelGGL(k_gather_once<16>, dim3(blocks), dim3(256), 0, r.stream, a, nseg, c); break;
    }
    PHX_HIP(hipGetLastError());
    PHX_HIP(hipEventRecord(r.ev_stop, r.stream));
    PHX_HIP(hipEventSynchronize(r.ev_stop));
    float ms = 0;
    PHX_HIP(hipEventElapsedTime(&ms, r.ev_start, r.ev_stop));
    if (it < 0) continue;
    const double gbs = 8.0 * n / (ms * 1e-3) / 1e9;
    sum += gbs;
    best = std::max(best, gbs);
  }
  PHX_HIP(hipFree(a));
  PHX_HIP(hipFree(c));
  *mean_bw = sum / reps;
  *max_bw = best;
}

// kind: 0 load, 1 store, 2 copy, 3 triad; returns mean/max GB/s over reps
void run_stream(int kind, double* mean_bw, double* max_bw) {
  ensure_init();
  Runtime& r = rt();
  const int64_t n = int64_t(1) << 27;  // 1 GiB per array
  double *a = nullptr, *b = nullptr, *c = nullptr;
  PHX_HIP(hipMalloc(&a, n * 8));
  PHX_HIP(hipMalloc(&b, n * 8));
  PHX_HIP(hipMalloc(&c, n * 8));
  PHX_HIP(hipMemsetAsync(a, 0, n * 8, r.stream));
  PHX_HIP(hipMemsetAsync(b, 0, n * 8, r.stream));
  const int blocks = r.num_cu * 8, threads = 256;
  const double bytes[4] = {8.0 * n, 8.0 * n, 16.0 * n, 24.0 * n};
  const int reps = 20;
  double sum = 0, best = 0;
  for (int it = -2; it < reps; it++) {
    PHX_HIP(hipEventRecord(r.ev_start, r.stream));
    switch (kind) {
      case 0: hipLaunchKernelGGL(k_stream_load, dim3(blocks), dim3(threads), 0, r.stream, a, n, c); break;
      case 1: hipLaunchKernelGGL(k_stream_store, dim3(blocks), dim3(threads), 0, r.stream, a, n, 1.0); break;
      case 2: hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(threads), 0, r.stream, a, b, n); break;
      default: hipLaunchKernelGGL(k_stream_triad, dim3(blocks), dim3(threads), 0, r.stream, a, b, c, n, 3.0); break;
    }
    PHX_HIP(hipGetLastError());
    PHX_HIP(hipEventRecord(r.ev_stop, r.stream));
    PHX_HIP(hipEventSynchronize(r.ev_stop));
    float ms = 0;
    PHX_HIP(hipEventElapsedTime(&ms, r.ev_start, r.ev_stop));
    if (it < 0) continue;
    double gbs = bytes[kind] / (ms * 1e-3) / 1e9;
    sum += gbs;
    best = std::max(best, gbs);
  }
  PHX_HIP(hipFree(a));
  PHX_HIP(hipFree(b));
  PHX_HIP(hipFree(c));
  *mean_bw = sum / reps;
  *max_bw = best;
}
}  // namespace

}  // namespace phx

using namespace phx;

extern "C" {

void phist_kernels_init(int* argc, char*** argv, int* iflag) {
  (void)argc;
  (void)argv;
  guard(iflag, __func__, [&] {
    ensure_init();
    *iflag = 0;
  });
}

void phist_kernels_finalize(int* iflag) {
  guard(iflag, __func__, [&] {
    Runtime& r = rt();
    if (r.initialized) {
      PHX_HIP(hipStreamSynchronize(r.stream));
      pool_trim();
      if (r.d_work) PHX_HIP(hipFree(r.d_work));
      if (r.h_work) PHX_HIP(hipHostFree(r.h_work));
      if (r.stage_base) {
        for (int s = 0; s < Runtime::kStageSlots; s++) PHX_HIP(hipEventDestroy(r.stage_ev[s]));
        PHX_HIP(hipHostFree(r.stage_base));
      }
      PHX_HIP(hipEventDestroy(r.ev_start));
      PHX_HIP(hipEventDestroy(r.ev_stop));
      PHX_HIP(hipEventDestroy(r.kt_start));
      PHX_HIP(hipEventDestroy(r.kt_stop));
      PHX_HIP(hipStreamDestroy(r.stream));
      r = Runtime();
    }
    *iflag = 0;
  });
}

void phist_comm_create(phist_comm_ptr* comm, int* iflag) {
  guard(iflag, __func__, [&] {
    *comm = new Comm();
    *iflag = 0;
  });
}
void phist_comm_delete(phist_comm_ptr comm, int* iflag) {
  guard(iflag, __func__, [&] {
    if (comm != default_comm()) delete (Comm*)comm;
    *iflag = 0;
  });
}
void phist_comm_get_rank(phist_const_comm_ptr comm, int* rank, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(comm, PHIST_INVALID_INPUT, "NULL comm");
    *rank = ((const Comm*)comm)->rank;
    *iflag = 0;
  });
}
void phist_comm_get_size(phist_const_comm_ptr comm, int* size, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(comm, PHIST_INVALID_INPUT, "NULL comm");
    *size = ((const Comm*)comm)->size;
    *iflag = 0;
  });
}
void phist_comm_get_mpi_comm(phist_const_comm_ptr comm, MPI_Comm* mpiComm, int* iflag) {
  guard(iflag, __func__, [&] {
    (void)comm;
    *mpiComm = 0;
    *iflag = 0;
  });
}

void phist_map_create(phist_map_ptr* map, phist_const_comm_ptr comm, phist_gidx nglob, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(nglob >= 0 && nglob <= INT32_MAX, PHIST_INVALID_INPUT, "map size %lld out of range", (long long)nglob);
    Map* m = new Map();
    m->comm = comm ? (const Comm*)comm : default_comm();
    m->nglob = nglob;
    m->nloc = (int32_t)nglob;
    m->ilower = 0;
    *map = m;
    *iflag = 0;
  });
}
void phist_map_delete(phist_map_ptr map, int* iflag) {
  guard(iflag, __func__, [&] {
    delete (Map*)map;
    *iflag = 0;
  });
}
void phist_map_get_comm(phist_const_map_ptr map, phist_const_comm_ptr* comm, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(map, PHIST_INVALID_INPUT, "NULL map");
    *comm = ((const Map*)map)->comm;
    *iflag = 0;
  });
}
void phist_map_get_local_length(phist_const_map_ptr map, phist_lidx* nloc, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(map, PHIST_INVALID_INPUT, "NULL map");
    *nloc = ((const Map*)map)->nloc;
    *iflag = 0;
  });
}
void phist_map_get_global_length(phist_const_map_ptr map, phist_gidx* nglob, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(map, PHIST_INVALID_INPUT, "NULL map");
    *nglob = ((const Map*)map)->nglob;
    *iflag = 0;
  });
}
void phist_map_get_ilower(phist_const_map_ptr map, phist_gidx* ilower, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(map, PHIST_INVALID_INPUT, "NULL map");
    *ilower = ((const Map*)map)->ilower;
    *iflag = 0;
  });
}
void phist_map_get_iupper(phist_const_map_ptr map, phist_gidx* iupper, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(map, PHIST_INVALID_INPUT, "NULL map");
    const Map* m = (const Map*)map;
    *iupper = m->ilower + m->nloc - 1;
    *iflag = 0;
  });
}
void phist_maps_compatible(phist_const_map_ptr map1, phist_const_map_ptr map2, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(map1 && map2, PHIST_INVALID_INPUT, "NULL map");
    const Map* a = (const Map*)map1;
    const Map* b = (const Map*)map2;
    *iflag = (a == b || (a->nglob == b->nglob && a->nloc == b->nloc && a->ilower == b->ilower)) ? 0 : -1;
  });
}

void phist_context_create(phist_context_ptr* ctx, phist_const_map_ptr row_map, phist_const_map_ptr col_map,
                          phist_const_map_ptr range_map, phist_const_map_ptr domain_map, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(row_map, PHIST_INVALID_INPUT, "context needs a row map");
    Context* c = new Context();
    c->row_map = (const Map*)row_map;
    c->col_map = (const Map*)col_map;
    c->range_map = range_map ? (const Map*)range_map : c->row_map;
    c->domain_map = domain_map ? (const Map*)domain_map : c->row_map;
    *ctx = c;
    *iflag = 0;
  });
}
void phist_context_delete(phist_context_ptr vctx, int* iflag) {
  guard(iflag, __func__, [&] {
    delete (Context*)vctx;
    *iflag = 0;
  });
}

void phist_bench_stream_load(double* mean_bw, double* max_bw, int* iflag) {
  guard(iflag, __func__, [&] { run_stream(0, mean_bw, max_bw); *iflag = 0; });
}
void phist_bench_stream_store(double* mean_bw, double* max_bw, int* iflag) {
  guard(iflag, __func__, [&] { run_stream(1, mean_bw, max_bw); *iflag = 0; });
}
void phist_bench_stream_copy(double* mean_bw, double* max_bw, int* iflag) {
  guard(iflag, __func__, [&] { run_stream(2, mean_bw, max_bw); *iflag = 0; });
}
void phist_bench_stream_triad(double* mean_bw, double* max_bw, int* iflag) {
  guard(iflag, __func__, [&] { run_stream(3, mean_bw, max_bw); *iflag = 0; });
}
void phist_amd_bench_gather(int seg_bytes, double* mean_bw, double* max_bw, int* iflag) {
  guard(iflag, __func__, [&] { run_gather(seg_bytes, mean_bw, max_bw); *iflag = 0; });
}
void phist_amd_bench_hbm(int kind, double* mean_bw, double* max_bw, int* iflag) {
  guard(iflag, __func__, [&] { run_ceiling(kind, mean_bw, max_bw); *iflag = 0; });
}

// ---- extension hooks (phist_amd_ext.h) ----
void phist_amd_sparseMat_set_format(int C, int sigma, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(C >= 0 && sigma >= 0, PHIST_INVALID_INPUT, "negative SELL parameters");
    PHX_REQUIRE(C == 0 || (C & (C - 1)) == 0, PHIST_INVALID_INPUT, "SELL chunk height must be a power of two");
    rt().sell_C = C;
    rt().sell_sigma = sigma;
    *iflag = 0;
  });
}
void phist_amd_set_grid_blocks(int nblocks, int* iflag) {
  guard(iflag, __func__, [&] {
    PHX_REQUIRE(nblocks >= 0, PHIST_INVALID_INPUT, "negative grid size");
    rt().grid_blocks = nblocks;
    *iflag = 0;
  });
}
void phist_amd_set_spmv_blocks_per_cu(int n, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    rt().spmv_blocks_per_cu = n < 0 ? -1 : n;
    *iflag = 0;
  });
}
void phist_amd_set_spmv_lane_bytes(int bytes, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_REQUIRE(bytes == 0 || bytes == 4 || bytes == 8 || bytes == 16 || bytes == 32 || bytes == 64,
                PHIST_INVALID_INPUT, "lane width %d bytes: 0, 4, 8, 16, 32 or 64", bytes);
    rt().spmv_lane_bytes = bytes;
    *iflag = 0;
  });
}
void phist_amd_set_spmv_depth(int slots, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_REQUIRE(slots == 0 || slots == 8 || slots == 16 || slots == 32, PHIST_INVALID_INPUT,
                "depth %d: 0, 8, 16 or 32", slots);
    rt().spmv_depth = slots == 0 ? -1 : slots;
    *iflag = 0;
  });
}
void phist_amd_set_spmv_col_block(int ncols, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_REQUIRE(ncols >= -1, PHIST_INVALID_INPUT, "column block %d: -1, 0 or a positive width", ncols);
    rt().spmv_col_block = ncols;
    *iflag = 0;
  });
}
void phist_amd_set_spmv_stage(int mode, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_REQUIRE(mode >= -1 && mode <= 1, PHIST_INVALID_INPUT, "staging mode %d: -1, 0 or 1", mode);
    rt().spmv_stage = mode;
    *iflag = 0;
  });
}
void phist_amd_set_gmres_device(int enable, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    rt().gmres_device = enable != 0;
    *iflag = 0;
  });
}
void phist_amd_set_fast_sync(int enable, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    rt().fast_sync = enable != 0;
    *iflag = 0;
  });
}
void phist_amd_spmv_staged_launches(int64_t* count, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_REQUIRE(count, PHIST_INVALID_INPUT, "NULL count");
    *count = rt().spmv_staged_launches;
    *iflag = 0;
  });
}
void phist_amd_set_minres_graphs(int enable, int iters_per_call, int graph_len, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_REQUIRE(iters_per_call >= 0 && graph_len >= 1, PHIST_INVALID_INPUT, "%d iterations per call, %d per graph",
                iters_per_call, graph_len);
    rt().minres_graphs = enable != 0;
    rt().minres_iters_per_call = iters_per_call;
    rt().minres_graph_len = graph_len;
    *iflag = 0;
  });
}
void phist_amd_minres_graph_stats(int64_t* captures, int64_t* launches, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    if (captures) *captures = rt().graph_captures;
    if (launches) *launches = rt().graph_launches;
    *iflag = 0;
  });
}
void phist_amd_get_stream(void** stream, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    *stream = (void*)rt().stream;
    *iflag = 0;
  });
}
void phist_amd_synchronize(int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    sync_stream();
    *iflag = 0;
  });
}
void phist_amd_timer_start(int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_HIP(hipEventRecord(rt().ev_start, rt().stream));
    *iflag = 0;
  });
}
void phist_amd_timer_stop(float* ms, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_HIP(hipEventRecord(rt().ev_stop, rt().stream));
    PHX_HIP(hipEventSynchronize(rt().ev_stop));
    PHX_HIP(hipEventElapsedTime(ms, rt().ev_start, rt().ev_stop));
    *iflag = 0;
  });
}
void phist_amd_kernel_timer(int enable, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    PHX_REQUIRE(enable >= 0 && enable <= 3, PHIST_INVALID_INPUT, "kernel timer mode %d (0..3)", enable);
    if (enable >= 2) {  // pause / resume: the totals stay
      rt().ktimer_on = enable == 3;
    } else {
      rt().ktimer_on = enable != 0;
      rt().ktimer_ms = 0.0;
      rt().ktimer_count = 0;
    }
    *iflag = 0;
  });
}
void phist_amd_kernel_timer_read(double* total_ms, int64_t* count, int* iflag) {
  guard(iflag, __func__, [&] {
    *total_ms = rt().ktimer_ms;
    *count = rt().ktimer_count;
    *iflag = 0;
  });
}
void phist_amd_device_info(char* name, int name_len, int* num_cu, int* iflag) {
  guard(iflag, __func__, [&] {
    ensure_init();
    hipDeviceProp_t prop;
    PHX_HIP(hipGetDeviceProperties(&prop, rt().device));
    if (name && name_len > 0) {
      snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cu) *num_cu = prop.multiProcessorCount;
    *iflag = 0;
  });
}

}  // extern "C"
