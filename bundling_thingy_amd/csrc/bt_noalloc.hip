// Endpoint dedup that allocates no device memory (bt_read_endpoints_noalloc).
//
// The same result as bt_read_endpoints (endpoints.wgsl:9-49: the distinct
// positions among the first and last samples of the paths with two or more
// samples, each position keeping its smallest sample index, ascending), for
// callers that cannot allocate. Its working memory is one file-scope device
// array, `scratch`, of BT_NOALLOC_SCRATCH_BYTES (12 MiB) per GPU:
//   [0, 256)   control words (the index count)
//   table      BT_NOALLOC_MAX_SLOTS 64-bit slots: (index << 32) | (sorted slot + 1)
//   bitmap     BT_NOALLOC_MAX_SAMPLES bits: sample index kept
//   list       2 * BT_NOALLOC_MAX_PATHS indices, ascending
// The host resolves the array's address once per device with
// hipGetSymbolAddress (bt::noalloc_setup, called when a context is created on
// that device) and passes that address to the kernels and copies; every call
// on a GPU shares that GPU's array, so calls that overlap on one GPU race.
// Everything a call reads is cleared by a kernel first (no memset node, so
// the launches can also be captured).
//
// Unlike bt_read_endpoints, which unpermutes the samples into a staging
// buffer first, the insert kernel reads the current records in sorted order:
// a record's id carries its Layout index and the endpoint flag, and the
// endpoints are exactly the samples with that flag.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "bt_common.h"
#include "bt_device.h"
#include "bt_noalloc.h"

namespace {

constexpr size_t kNxCtlBytes = 256;
constexpr size_t kNxTableOff = kNxCtlBytes;
constexpr size_t kNxBitmapOff = kNxTableOff + (size_t)BT_NOALLOC_MAX_SLOTS * 8u;
constexpr size_t kNxListOff = kNxBitmapOff + (size_t)BT_NOALLOC_MAX_SAMPLES / 8u;
constexpr size_t kNxScratchBytes = kNxListOff + (size_t)BT_NOALLOC_MAX_PATHS * 2u * 4u;
static_assert(kNxScratchBytes == BT_NOALLOC_SCRATCH_BYTES, "scratch layout");
static_assert(kNxBitmapOff % 256 == 0 && kNxListOff % 256 == 0, "scratch alignment");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNxScratchBytes];

#define NX_LANE() (threadIdx.x & 63u)

// Zeroes the control words, `slots` table entries and `words` bitmap words.
extern "C" __global__ __launch_bounds__(256) void bt_nx_clear(unsigned char* base, uint32_t slots,
                                                              uint32_t words) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  unsigned long long* table = reinterpret_cast<unsigned long long*>(base + kNxTableOff);
  uint32_t* bitmap = reinterpret_cast<uint32_t*>(base + kNxBitmapOff);
  if (i < slots) table[i] = 0ull;
  if (i < words) bitmap[i] = 0u;
  if (i < kNxCtlBytes / 4u) reinterpret_cast<uint32_t*>(base)[i] = 0u;
}

// One thread per sorted record: an endpoint sample (BT_FIXED_BIT) goes into
// an open-addressing table keyed by its position (-0 and 0 alike, compared
// with ==; linear probing, load <= 1/2). An entry holds the Layout index in
// its upper word, so atomicMin keeps each position's smallest index, and the
// sorted slot + 1 in its lower word, which locates the entry's position.
extern "C" __global__ __launch_bounds__(256) void bt_nx_insert(const float* __restrict__ rec,
                                                               uint32_t n, unsigned char* base,
                                                               uint32_t mask) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t id = __float_as_uint(rec[3u * i + 2u]);
  if (!(id & BT_FIXED_BIT)) return;
  unsigned long long* table = reinterpret_cast<unsigned long long*>(base + kNxTableOff);
  const float x = rec[3u * i], y = rec[3u * i + 1u];
  const unsigned long long mine = ((unsigned long long)(id & ~BT_FIXED_BIT) << 32) | (i + 1u);
  const uint32_t bx = __float_as_uint(x == 0.0f ? 0.0f : x);
  const uint32_t by = __float_as_uint(y == 0.0f ? 0.0f : y);
  uint32_t h = bx * 0x9e3779b1u ^ (by + 0x7f4a7c15u + (bx << 6) + (bx >> 2));
  h ^= h >> 15;
  h *= 0x2c1b3c6du;
  h ^= h >> 12;
  h *= 0x297a2d39u;
  h ^= h >> 15;
  uint32_t slot = h & mask;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    unsigned long long old =
        __hip_atomic_load(&table[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0ull) {
      old = atomicCAS(&table[slot], 0ull, mine);
      if (old == 0ull) return;  // inserted
    }
    const uint32_t os = (uint32_t)(old & 0xffffffffull) - 1u;
    if (rec[3u * os] == x && rec[3u * os + 1u] == y) {
      atomicMin(&table[slot], mine);
      return;
    }
    slot = (slot + 1u) & mask;
  }
}

// Sets the bitmap bit of the index each occupied slot kept.
extern "C" __global__ __launch_bounds__(256) void bt_nx_mark(unsigned char* base, uint32_t slots) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= slots) return;
  const unsigned long long e =
      reinterpret_cast<const unsigned long long*>(base + kNxTableOff)[i];
  if (e == 0ull) return;
  const uint32_t index = (uint32_t)(e >> 32);
  atomicOr(reinterpret_cast<uint32_t*>(base + kNxBitmapOff) + (index >> 5), 1u << (index & 31u));
}

// One workgroup of 1024 threads lists the set bits of `words` bitmap words
// in ascending order: per chunk of 1024 words, each thread's popcount, a
// workgroup exclusive scan, then each thread writes its word's indices at
// its offset. The total goes to control word 0.
extern "C" __global__ __launch_bounds__(1024) void bt_nx_compact(unsigned char* base,
                                                                 uint32_t words) {
  __shared__ uint32_t wave_tot[16];
  __shared__ uint32_t carry_s;
  const uint32_t* bitmap = reinterpret_cast<const uint32_t*>(base + kNxBitmapOff);
  uint32_t* list = reinterpret_cast<uint32_t*>(base + kNxListOff);
  const uint32_t lane = NX_LANE(), w = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_s = 0u;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < words; c0 += 1024u) {
    const uint32_t wi = c0 + threadIdx.x;
    uint32_t bits = wi < words ? bitmap[wi] : 0u;
    const uint32_t pc = (uint32_t)__popc(bits);
    uint32_t incl = pc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(incl, d, 64);
      if ((int)lane >= d) incl += t;
    }
    if (lane == 63u) wave_tot[w] = incl;
    __syncthreads();
    uint32_t before = carry_s;
    for (uint32_t k = 0; k < w; ++k) before += wave_tot[k];
    uint32_t pos = before + incl - pc;
    while (bits) {
      const uint32_t b = (uint32_t)__builtin_ctz(bits);
      if (pos < 2u * BT_NOALLOC_MAX_PATHS) list[pos] = wi * 32u + b;
      ++pos;
      bits &= bits - 1u;
    }
    __syncthreads();  // every thread has read carry_s and wave_tot
    if (threadIdx.x == 0) {
      uint32_t t = carry_s;
      for (uint32_t k = 0; k < 16u; ++k) t += wave_tot[k];
      carry_s = t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) reinterpret_cast<uint32_t*>(base)[0] = carry_s;
}

namespace bt {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) fail(BT_E_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device]) return;
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  if (cur != device) fail(BT_E_DEVICE, "noalloc set-up: the device is not current");
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    fail(BT_E_DEVICE, std::string("noalloc set-up: hipGetSymbolAddress: ") + hipGetErrorString(e));
  g_ptr[device] = p;
}

void endpoints_noalloc(hipStream_t s, int device, const float* rec, uint32_t n, uint32_t n_paths,
                       uint32_t* indices, size_t capacity, uint32_t* count) {
  *count = 0;
  if (n_paths > BT_NOALLOC_MAX_PATHS)
    fail(BT_E_CAPACITY, "graph has " + std::to_string(n_paths) + " paths; the allocation-free "
                        "endpoint dedup takes at most " + std::to_string(BT_NOALLOC_MAX_PATHS));
  if (n > BT_NOALLOC_MAX_SAMPLES)
    fail(BT_E_CAPACITY, "graph has " + std::to_string(n) + " samples; the allocation-free "
                        "endpoint dedup takes at most " + std::to_string(BT_NOALLOC_MAX_SAMPLES));
  void* base_v = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (device >= 0 && device < kMaxDevices) base_v = g_ptr[device];
  }
  if (!base_v) fail(BT_E_STATE, "allocation-free scratch is not set up on this device");
  unsigned char* base = static_cast<unsigned char*>(base_v);
  if (n == 0 || n_paths == 0) return;
  uint32_t slots = 64;
  while (slots < 4u * n_paths) slots <<= 1;  // load <= 1/2 with two ends per path
  const uint32_t words = (n + 31u) / 32u;
  const uint32_t clear_items = slots > words ? slots : words;
  hipLaunchKernelGGL(bt_nx_clear, dim3((clear_items + 255u) / 256u), dim3(256), 0, s, base,
                     slots, words);
  hipLaunchKernelGGL(bt_nx_insert, dim3((n + 255u) / 256u), dim3(256), 0, s, rec, n, base,
                     slots - 1u);
  hipLaunchKernelGGL(bt_nx_mark, dim3((slots + 255u) / 256u), dim3(256), 0, s, base, slots);
  hipLaunchKernelGGL(bt_nx_compact, dim3(1), dim3(1024), 0, s, base, words);
  HIPCHK(hipGetLastError());
  uint32_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, base, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *count = total;
  const size_t take = std::min<size_t>(capacity, total);
  if (take) {
    HIPCHK(hipMemcpyAsync(indices, base + kNxListOff, take * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
}

}  // namespace bt
