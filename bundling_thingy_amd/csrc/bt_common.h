// Host-side helpers shared by the runtime: errors, device buffers, the
// per-iteration parameter model and the reference's host-side algorithms.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/bundling.h"

namespace bt {

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);
#define HIPCHK(x) ::bt::hip_check((x), #x)

// Per-engine cache of device blocks: a freed block is kept and handed out
// again (same stream, so reuse is ordered after every earlier use) instead of
// going back through hipFree, which waits for the device. Blocks of at most
// kSlabBlockMax bytes are carved from kSlabBytes slabs, so an upload of many
// small jobs makes a few hipMalloc calls rather than one per buffer (each
// costs about 10 us). When hipMalloc fails, the cached blocks of their own
// and the slabs whose blocks are all cached return to HIP and the call is
// tried once more; everything else returns when the pool is destroyed.
class DevPool {
 public:
  static constexpr size_t kSlabBytes = 64u << 20;
  static constexpr size_t kSlabBlockMax = 4u << 20;
  DevPool() = default;
  DevPool(const DevPool&) = delete;
  DevPool& operator=(const DevPool&) = delete;
  ~DevPool();
  // Bytes a request of `bytes` holds: a power of two (>= 256) up to 1 MiB,
  // whole MiB above.
  static size_t block_size(size_t bytes);
  void* get(size_t bytes, size_t* cap);
  void put(void* p, size_t cap);
  size_t cached() const { return cached_; }
  size_t slabs() const { return slabs_.size(); }

 private:
  struct Slab {
    char* base;
    size_t carved;  // bytes handed out from the slab so far
  };
  void* device_alloc(size_t bytes);  // hipMalloc, retried once after release_cached()
  void release_cached();
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> own_;      // blocks with a hipMalloc of their own
  std::vector<Slab> slabs_;          // the last one is being carved
  size_t cached_ = 0;
};

// Owning device allocation (from `pool` when set, else hipMalloc).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  size_t cap = 0;
  DevPool* pool = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t b);
  // Grows (never shrinks); returns true if the address changed.
  bool reserve(size_t b);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Options with the reference defaults filled in (main.rs:25-90, gpu.rs:21).
bt_options resolve_options(const bt_options* in);

bool has_radius(const bt_options& o);

// Grid rows per radius when none is given: rows about 4 texels high, at
// most 8 per radius (the autotuner refines it per iteration).
uint32_t default_split(float radius);

// Grid rows per radius from the sample density once the sample count is
// known: rows about 2.9 / sqrt(samples per texel^2) texels high (the set-up
// cost per task grows with the rows of its window, the tested band with the
// row height times the density), 1 to kMaxSplit rows per radius.
uint32_t density_split(float radius, uint64_t samples, float res_x, float res_y);
constexpr uint32_t kMaxSplit = 28;  // windows stay within 64 rows

// Bins per row height along x when bt_options::x_split is 0.
constexpr uint32_t kDefaultXSplit = 16;
// Upper bound on one job's grid bins (4 bytes each); x_split is halved until
// the grid fits.
constexpr uint64_t kMaxGridCells = 1ull << 22;

// Parameters of iteration `it`; split = grid rows per radius (0 = default).
void iter_params(const bt_options& o, uint32_t it, uint32_t split, bt_iter_params* p);

// Number of distinct parameter sets of the schedule (reference mode: 1).
uint32_t schedule_len(const bt_options& o);

uint32_t blur_weights(float sigma, float* w, uint32_t max_taps);

// graph.rs:79-91
bool normalization(const float* bbox, float* out);

// The count pass of resampling.wgsl:82-159 (count_paths) on the host, in the
// f32 arithmetic the emit kernel uses (no contraction, correctly rounded
// sqrtf and division): each path's sample count in closed form per segment,
// duplicates of the previous sample dropped, the closed-path special case,
// counts over `cap` stopped at cap + 1. path_scan (n_paths + 1 entries)
// receives the exclusive prefix sums of (count << 32 | 1) over the paths
// with two or more samples (0 for the others); returns the total, i.e.
// samples << 32 | drawable paths. norm = {min_x, min_y, scale, tr_x, tr_y}.
uint64_t resample_counts(const float* xy, const uint32_t* starts, uint32_t n_paths,
                         const float* norm, float ts_x, float ts_y, uint32_t cap,
                         uint64_t* path_scan);

// resampling.rs:72-89,321-397; returns BT_OK or an error code with message.
int capacity_model(const bt_options& o, uint32_t n_points, uint32_t n_starts, uint64_t* capacity,
                   uint64_t* reserved, std::string* err);

// graph.rs:35-77
int parse_trl(const char* text, size_t len, std::vector<float>& xy, std::vector<uint32_t>& starts,
              float* bbox, std::string* err);

// distortion.rs:118-216 (+ geometry.rs:28-73 layout rules)
int distortion(const float* original, const float* current, const uint32_t* rec3,
               uint32_t n_paths, uint32_t n_samples, uint32_t map_size, double* out,
               std::string* err);

}  // namespace bt
