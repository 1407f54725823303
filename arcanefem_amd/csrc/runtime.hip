// runtime.hip -- library runtime: device/stream set-up, error reporting,
// device-memory plumbing, per-kernel event timing and a device-wide scan.
#include "common.h"
#include "device_utils.h"

#include <algorithm>
#include <map>
#include <mutex>

namespace afem {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

void fail(int status, const std::string& msg) { throw Error(status, msg); }

Runtime& rt()
{
  static Runtime r;
  return r;
}

static std::mutex g_init_mutex;

static void init_device(int device)
{
  std::lock_guard<std::mutex> lock(g_init_mutex);
  Runtime& r = rt();
  if (r.stream && r.device == device)
    return;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0)
    fail(AFEM_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= count)
    fail(AFEM_ERR_ARGUMENT, "device index out of range");
  AFEM_HIP(hipSetDevice(device));
  if (r.stream) {
    (void)hipStreamDestroy(r.stream);
    r.stream = nullptr;
  }
  AFEM_HIP(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  AFEM_HIP(hipGetDeviceProperties(&prop, device));
  r.num_cu = prop.multiProcessorCount;
  r.device = device;
}

void ensure_init()
{
  if (!rt().stream)
    init_device(0);
}

// ---- profiling ----------------------------------------------------------------

struct ProfRecord
{
  std::string name;
  hipEvent_t start, stop;
};
static std::vector<ProfRecord> g_records;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t take_event()
{
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  AFEM_HIP(hipEventCreate(&e));
  return e;
}

ProfScope::ProfScope(const char* kernel_name)
: name(kernel_name)
{
  if (rt().profiling) {
    start = take_event();
    AFEM_HIP(hipEventRecord(start, stream()));
  }
}

ProfScope::~ProfScope()
{
  if (start) {
    hipEvent_t stop = nullptr;
    if (hipEventCreate(&stop) == hipSuccess && hipEventRecord(stop, rt().stream) == hipSuccess)
      g_records.push_back({ name, start, stop });
  }
}

static void profile_reset()
{
  for (auto& r : g_records) {
    g_event_pool.push_back(r.start);
    (void)hipEventDestroy(r.stop);
  }
  g_records.clear();
}

// ---- device-wide exclusive scan (int32) --------------------------------------------
// Three phases: per-tile sums, a scan of the tile sums (recursive), and a
// per-tile scan that adds the tile offset.  Tile = 256 threads x 8 items.

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ void __launch_bounds__(kScanBlock)
k_scan_tile_sums(const int32_t* __restrict__ in, int64_t n, int32_t* __restrict__ tile_sums)
{
  __shared__ int32_t red[kScanBlock / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  int32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    if (i < n)
      s += in[i];
  }
  s = wave_reduce_sum(s);
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0;
    for (int w = 0; w < kScanBlock / 64; ++w)
      t += red[w];
    tile_sums[blockIdx.x] = t;
  }
}

// Exclusive scan of one tile in LDS, plus the tile offset; writes out[0..n).
__global__ void __launch_bounds__(kScanBlock)
k_scan_tile_apply(const int32_t* __restrict__ in, int64_t n, const int32_t* __restrict__ tile_offsets,
                  int32_t* __restrict__ out)
{
  __shared__ int32_t tile[kScanTile];
  __shared__ int32_t thread_sums[kScanBlock];
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int k = 0; k < kScanItems; ++k) {
    int idx = k * kScanBlock + threadIdx.x;
    int64_t i = base + idx;
    tile[idx] = (i < n) ? in[i] : 0;
  }
  __syncthreads();
  // each thread scans its 8 consecutive items
  int32_t local[kScanItems];
  int32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    local[k] = s;
    s += tile[threadIdx.x * kScanItems + k];
  }
  thread_sums[threadIdx.x] = s;
  __syncthreads();
  // Hillis-Steele scan over the 256 thread sums
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int32_t v = (threadIdx.x >= (unsigned)off) ? thread_sums[threadIdx.x - off] : 0;
    __syncthreads();
    thread_sums[threadIdx.x] += v;
    __syncthreads();
  }
  int32_t prefix = tile_offsets[blockIdx.x] + thread_sums[threadIdx.x] - s;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)threadIdx.x * kScanItems + k;
    if (i < n)
      out[i] = prefix + local[k];
  }
}

__global__ void k_store_total(const int32_t* __restrict__ in, const int32_t* __restrict__ out, int64_t n,
                              int32_t* __restrict__ total_slot)
{
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *total_slot = (n > 0) ? out[n - 1] + in[n - 1] : 0;
}

static void scan_rec(const int32_t* in, int32_t* out, int64_t n)
{
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles < 1)
    tiles = 1;
  DBuf<int32_t> sums((size_t)tiles);
  DBuf<int32_t> offs((size_t)tiles + 1);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)tiles), dim3(kScanBlock), 0, stream(), in, n, sums.p);
  check_launch("scan_tile_sums");
  if (tiles == 1) {
    AFEM_HIP(hipMemsetAsync(offs.p, 0, sizeof(int32_t), stream()));
  }
  else {
    scan_rec(sums.p, offs.p, tiles);
  }
  hipLaunchKernelGGL(k_scan_tile_apply, dim3((unsigned)tiles), dim3(kScanBlock), 0, stream(), in, n, offs.p, out);
  check_launch("scan_tile_apply");
  // keep the temporaries alive until the kernels have consumed them
  AFEM_HIP(hipStreamSynchronize(stream()));
}

int64_t exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n)
{
  scan_rec(in, out, n);
  hipLaunchKernelGGL(k_store_total, dim3(1), dim3(64), 0, stream(), in, out, n, out + n);
  check_launch("scan_store_total");
  int32_t total = 0;
  AFEM_HIP(hipMemcpyAsync(&total, out + n, sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
  return total;
}

// ---- streaming copy ----------------------------------------------------------------
// Each workgroup copies one contiguous 16 KB chunk (4 x 16 bytes per lane,
// all loads in flight before the stores), one workgroup per chunk: the
// device-to-device copy the roofline figures are measured against.  On
// MI355X this form measured 5.8 TB/s for 1 GiB against 4.4-5.3 TB/s for
// grid-stride loops over 4-32 workgroups per CU and 5.1 TB/s for
// hipMemcpyDeviceToDevice.
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int kCopyUnroll = 4;

__global__ void __launch_bounds__(256)
k_stream_copy(const v4i* __restrict__ src, v4i* __restrict__ dst, int64_t n16)
{
  const int64_t b = (int64_t)blockIdx.x * (256 * kCopyUnroll) + threadIdx.x;
  v4i a[kCopyUnroll];
#pragma unroll
  for (int u = 0; u < kCopyUnroll; ++u)
    if (b + 256 * u < n16)
      a[u] = src[b + 256 * u];
#pragma unroll
  for (int u = 0; u < kCopyUnroll; ++u)
    if (b + 256 * u < n16)
      dst[b + 256 * u] = a[u];
}

} // namespace afem

// ---- C ABI: runtime -----------------------------------------------------------------

using namespace afem;



extern "C" {

AFEM_API const char* afem_last_error(void) { return g_last_error.c_str(); }

AFEM_API int afem_version(void) { return 100; }

AFEM_API int afem_init(int device)
{
  AFEM_GUARD_BEGIN
  init_device(device);
  AFEM_GUARD_END
}

AFEM_API int afem_finalize(void)
{
  AFEM_GUARD_BEGIN
  Runtime& r = rt();
  if (r.stream) {
    AFEM_HIP(hipStreamSynchronize(r.stream));
    profile_reset();
    for (hipEvent_t e : g_event_pool)
      (void)hipEventDestroy(e);
    g_event_pool.clear();
    (void)hipStreamDestroy(r.stream);
    r.stream = nullptr;
    r.device = -1;
  }
  AFEM_GUARD_END
}

AFEM_API int afem_synchronize(void)
{
  AFEM_GUARD_BEGIN
  AFEM_HIP(hipStreamSynchronize(stream()));
  AFEM_GUARD_END
}

AFEM_API int afem_malloc(void** dptr, size_t bytes)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(dptr, AFEM_ERR_ARGUMENT, "null output pointer");
  ensure_init();
  *dptr = nullptr;
  if (bytes)
    AFEM_HIP(hipMalloc(dptr, bytes));
  AFEM_GUARD_END
}

AFEM_API int afem_free(void* dptr)
{
  AFEM_GUARD_BEGIN
  if (dptr) {
    AFEM_HIP(hipStreamSynchronize(stream()));
    AFEM_HIP(hipFree(dptr));
  }
  AFEM_GUARD_END
}

AFEM_API int afem_memcpy_htod(void* dst, const void* src, size_t bytes)
{
  AFEM_GUARD_BEGIN
  if (bytes) {
    AFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    AFEM_HIP(hipStreamSynchronize(stream()));
  }
  AFEM_GUARD_END
}

AFEM_API int afem_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
  AFEM_GUARD_BEGIN
  if (bytes) {
    AFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    AFEM_HIP(hipStreamSynchronize(stream()));
  }
  AFEM_GUARD_END
}

AFEM_API int afem_memcpy_dtod(void* dst, const void* src, size_t bytes)
{
  AFEM_GUARD_BEGIN
  if (bytes)
    AFEM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
  AFEM_GUARD_END
}

AFEM_API int afem_stream_copy(void* dst, const void* src, size_t bytes)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(dst && src, AFEM_ERR_ARGUMENT, "null pointer");
  AFEM_REQUIRE(bytes % 16 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0 &&
                 reinterpret_cast<uintptr_t>(src) % 16 == 0,
               AFEM_ERR_ARGUMENT, "stream copy needs 16-byte aligned pointers and a multiple of 16 bytes");
  const int64_t n16 = (int64_t)(bytes / 16);
  if (n16 == 0)
    return AFEM_OK;
  const int64_t chunks = (n16 + 256 * kCopyUnroll - 1) / (256 * kCopyUnroll);
  AFEM_REQUIRE(chunks <= 0x7fffffff, AFEM_ERR_ARGUMENT, "stream copy too large for one launch");
  const unsigned grid = (unsigned)chunks;
  ProfScope prof("stream_copy");
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, stream(), static_cast<const v4i*>(src),
                     static_cast<v4i*>(dst), n16);
  check_launch("stream_copy");
  AFEM_GUARD_END
}

AFEM_API int afem_memset(void* dptr, int value, size_t bytes)
{
  AFEM_GUARD_BEGIN
  if (bytes)
    AFEM_HIP(hipMemsetAsync(dptr, value, bytes, stream()));
  AFEM_GUARD_END
}

AFEM_API int afem_profile_enable(int on)
{
  AFEM_GUARD_BEGIN
  ensure_init();
  rt().profiling = (on != 0);
  AFEM_GUARD_END
}

AFEM_API int afem_profile_reset(void)
{
  AFEM_GUARD_BEGIN
  AFEM_HIP(hipStreamSynchronize(stream()));
  profile_reset();
  AFEM_GUARD_END
}

AFEM_API int afem_profile_query(const char* kernel, int64_t* launches, double* total_ms)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(kernel && launches && total_ms, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_HIP(hipStreamSynchronize(stream()));
  int64_t n = 0;
  double t = 0.0;
  for (auto& r : g_records)
    if (r.name == kernel) {
      float ms = 0.f;
      AFEM_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
      t += ms;
      ++n;
    }
  *launches = n;
  *total_ms = t;
  AFEM_GUARD_END
}

} // extern "C"
