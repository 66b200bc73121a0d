// Hand-written gfx950 kernels for the edge-bundling hot path.
//
// Reference (lpstankus/bundling_thingy) kernels replaced here, per stage:
//   resample_emit                    <- resampling.wgsl:161-183 (emit_paths; the count pass,
//                                      count_paths :138-159, runs on the host: resample_counts)
//   bin / cell scan / tasks / scatter <- spatial_grid.wgsl:72-138 (count_cells,
//                                      reserve_cells, scatter_points, prepare_dispatch)
//   neighbor_mean                    <- neighbor_mean.wgsl:75-203 (tiled + sparse)
//   blur_rows / blur_cols (+ peak),  <- blur.wgsl:13-83
//   or both in blur_fused
//   normalize                        <- normalization.wgsl:5-17
//
// Arithmetic contract (identical to the CPU oracle): FP32 without contraction
// (built with -ffp-contract=off; fmaf only where written), correctly rounded
// '/' and sqrtf, rintf for WGSL round(). Neighbor sums are exact 2^-24
// fixed-point int64 sums, so results do not depend on visiting order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bt_device.h"

#define LANE() (threadIdx.x & 63u)

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const f4v cf4v;
typedef __attribute__((address_space(4))) const uint32_t cu32;
typedef __attribute__((address_space(4))) const float cf32;

// Loads through the constant address space are eligible for scalar (SMEM)
// loads when the address is wave-uniform; the buffers read this way are never
// written by the kernel that reads them.
__device__ __forceinline__ float4 ld_uniform(const float4* p, uint32_t i) {
  const f4v v = ((cf4v*)(uintptr_t)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t ld_uniform(const uint32_t* p, uint32_t i) {
  return ((cu32*)(uintptr_t)p)[i];
}
__device__ __forceinline__ float ld_uniform(const float* p, uint32_t i) {
  return ((cf32*)(uintptr_t)p)[i];
}

__device__ __forceinline__ int find_job(const uint32_t* __restrict__ base, int njobs, uint32_t b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ld_uniform(base, (uint32_t)mid) <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// spatial_grid.wgsl:59-61
__device__ __forceinline__ float texel(float p, float res) { return (p + 1.0f) * 0.5f * res; }

// spatial_grid.wgsl:63-65, one axis; clamps before converting.
__device__ __forceinline__ uint32_t cell_axis(float t, float cell, uint32_t g) {
  float v = fmaxf(t / cell, 0.0f);
  return v >= (float)(g - 1u) ? g - 1u : (uint32_t)v;
}

// neighbor_mean.wgsl:66-73 / resampling.wgsl:69-75 pixel of a sample (y flipped).
__device__ __forceinline__ uint32_t splat_index(float x, float y, uint32_t w, uint32_t h) {
  float ux = x * 0.5f + 0.5f, uy = 0.5f - y * 0.5f;
  float fx = rintf(ux * (float)(w - 1u)), fy = rintf(uy * (float)(h - 1u));
  fx = fminf(fmaxf(fx, 0.0f), (float)(w - 1u));
  fy = fminf(fmaxf(fy, 0.0f), (float)(h - 1u));
  return (uint32_t)fy * w + (uint32_t)fx;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const uint32_t lane = LANE();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

// Exclusive scan across a 256-thread block; returns the exclusive prefix and
// the block total. lds must hold 4 values.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* lds, T* total) {
  const uint32_t lane = LANE(), wave = threadIdx.x >> 6;
  T inc = wave_incl_scan(v);
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  T w0 = lds[0], w1 = lds[1], w2 = lds[2], w3 = lds[3];
  T before = (wave > 0 ? w0 : T(0)) + (wave > 1 ? w1 : T(0)) + (wave > 2 ? w2 : T(0));
  *total = w0 + w1 + w2 + w3;
  __syncthreads();
  return before + inc - v;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
  return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ long long wave_sum64(long long v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

#define JOB_PROLOGUE(KIND)                                                   \
  const int j = find_job(base, njobs, blockIdx.x);                          \
  const uint32_t lb = blockIdx.x - ld_uniform(base, (uint32_t)j);           \
  const BtJob& J = jobs[j];

// ------------------------------------------------------------------ clear
// mode bit 0: zeroes the cell counters (n_cells + 1 words) of each job and
// the launch's control block; bit 1: the density counts and the peak of each
// job (apart, the density clear runs beside the next iteration's binning).
extern "C" __global__ __launch_bounds__(256) void bt_clear(const BtJob* __restrict__ jobs,
                                                           const uint32_t* __restrict__ base,
                                                           int njobs, BtControl* ctl, int mode) {
  JOB_PROLOGUE(BT_K_CLEAR);
  const uint32_t ncell = (mode & 1) ? J.n_cells + 1u : 0u;
  const uint32_t npix = (mode & 2) && J.dens_cnt ? J.dw * J.dh : 0u;
  uint32_t* cells = J.cells;
  uint32_t* dens = J.dens_cnt;
  const uint32_t i0 = lb * BT_CLEAR_ITEMS + threadIdx.x;
#pragma unroll 4
  for (uint32_t k = 0; k < BT_CLEAR_ITEMS / 256; ++k) {
    uint32_t i = i0 + k * 256u;
    if (i < ncell) cells[i] = 0u;
    if (i < npix) dens[i] = 0u;
  }
  if (lb == 0 && threadIdx.x == 0) {
    if ((mode & 2) && J.peak) J.peak[0] = 0u;  // no peak buffer when density is off
    if ((mode & 1) && j == 0) {
      ctl->task_total = 0u;
      ctl->heavy_count = 0u;
      ctl->item_count = 0u;
      for (int x = 0; x < BT_XCDS; ++x) {
        ctl->xq[x * BT_CTR_STRIDE] = 0u;
        ctl->iq[x * BT_CTR_STRIDE] = 0u;
      }
    }
  }
}

// -------------------------------------------------------------------- bin
// spatial_grid.wgsl:72-81 (count_cells) fused with the rank half of
// scatter_points (:127-138): each sample gets its cell and its rank among the
// cell's samples. The input is in the previous iteration's sorted order, so
// equal cells mostly come in runs of consecutive samples. Each wave takes
// BT_RANK_CHUNKS chunks of 64 consecutive samples; the first sample of each
// run (across chunk boundaries) adds the run's length to the cell counter,
// all these atomics issued together, and the run's samples take consecutive
// ranks from the returned base. A collapsed bundle (one cell, 10^5 samples)
// costs one atomic per 512 samples instead of one per sample or per wave. A
// cell split over several runs gets one disjoint range per run. Ranks depend
// on scheduling, which is harmless: every consumer is order-independent.
extern "C" __global__ __launch_bounds__(256) void bt_bin(const BtJob* __restrict__ jobs,
                                                         const uint32_t* __restrict__ base,
                                                         int njobs) {
  JOB_PROLOGUE(BT_K_RANK);
  constexpr uint32_t C = BT_RANK_CHUNKS;
  const uint32_t lane = LANE();
  const uint32_t w0 = lb * BT_RANK_BLOCK + (threadIdx.x >> 6) * (C * 64u);
  const uint32_t nv = w0 < J.n ? min(C * 64u, J.n - w0) : 0u;  // valid samples (a prefix)
  const float* rec = J.rec[J.parity];
  uint32_t key[C];
  uint64_t lead[C];
#pragma unroll
  for (uint32_t c = 0; c < C; ++c) {
    const uint32_t p = c * 64u + lane;
    key[c] = 0xffffffffu;
    if (p < nv) {
      const uint32_t i = w0 + p;
      const float x = rec[3u * i], y = rec[3u * i + 1u];
      const uint32_t cx = cell_axis(texel(x, J.res_x), J.cell_w, J.gx);
      const uint32_t cy = cell_axis(texel(y, J.res_y), J.cell, J.gy);
      key[c] = cy * J.gx + cx;
    }
  }
#pragma unroll
  for (uint32_t c = 0; c < C; ++c) {
    // the previous sample in sequence: lane - 1, or lane 63 of the last chunk
    uint32_t prev = __shfl(key[c], (int)(lane == 0u ? 0u : lane - 1u), 64);
    if (lane == 0u)
      prev = c ? (uint32_t)__builtin_amdgcn_readlane(key[c - 1u], 63) : 0xfffffffeu;
    lead[c] = ballot(c * 64u + lane < nv && key[c] != prev);
  }
  const uint64_t upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
  // run lengths: from a leader to the next leader (or the end of the samples)
  uint32_t b[C];
  uint32_t first_after = nv;
#pragma unroll
  for (int c = (int)C - 1; c >= 0; --c) {
    const uint64_t above = lead[c] & ~upto;
    const uint32_t next = above ? (uint32_t)c * 64u + (uint32_t)__builtin_ctzll(above) : first_after;
    b[c] = 0u;
    if ((lead[c] >> lane) & 1ull) b[c] = atomicAdd(&J.cells[key[c]], next - ((uint32_t)c * 64u + lane));
    if (lead[c]) first_after = (uint32_t)c * 64u + (uint32_t)__builtin_ctzll(lead[c]);
  }
  // ranks: base of the run's leader plus the distance to it
  uint32_t carry_pos = 0u, carry_base = 0u;
#pragma unroll
  for (uint32_t c = 0; c < C; ++c) {
    const uint32_t p = c * 64u + lane;
    const uint64_t below = lead[c] & upto;
    const uint32_t hl = below ? 63u - (uint32_t)__builtin_clzll(below) : 0u;
    const uint32_t in_base = __shfl(b[c], (int)hl, 64);
    if (p < nv) {
      const uint32_t head = below ? c * 64u + hl : carry_pos;
      const uint32_t rb = below ? in_base : carry_base;
      J.cellrank[w0 + p] = make_uint2(key[c], rb + (p - head));
    }
    if (lead[c]) {
      const uint32_t L = 63u - (uint32_t)__builtin_clzll(lead[c]);
      carry_pos = c * 64u + L;
      carry_base = (uint32_t)__builtin_amdgcn_readlane(b[c], L);
    }
  }
}

// -------------------------------------------------------------- cell scan
// Two-phase exclusive scan of each job's n_cells + 1 counters (the last one
// is zero, so it becomes the total): block totals, then each block's local
// scan offset by the sum of the totals before it. Replaces reserve_cells
// (spatial_grid.wgsl:86-125), whose atomically ordered block bases are not
// deterministic.
template <typename T>
__device__ __forceinline__ void load_block(const T* src, uint32_t n, uint32_t first, T* lds) {
  for (uint32_t k = threadIdx.x; k < BT_SCAN_ITEMS; k += 256u) {
    uint32_t i = first + k;
    lds[k] = i < n ? src[i] : T(0);
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void scan_reduce_body(const T* src, uint32_t n, uint32_t lb, T* bsum) {
  __shared__ T lds[BT_SCAN_ITEMS];
  __shared__ T part[4];
  load_block(src, n, lb * BT_SCAN_ITEMS, lds);
  T s = T(0);
#pragma unroll
  for (int k = 0; k < 16; ++k) s += lds[threadIdx.x * 16 + k];
  T total;
  (void)block_excl_scan(s, part, &total);
  if (threadIdx.x == 0) bsum[lb] = total;
}

// Sum of the block totals bsum[0 .. lb) over a 256-thread block: the
// block's offset (the totals are few, so each block adds its predecessors'
// instead of a separate scan of the totals).
template <typename T>
__device__ __forceinline__ T block_offset(const T* bsum, uint32_t lb, T* part) {
  T s = T(0);
  for (uint32_t k = threadIdx.x; k < lb; k += 256u) s += bsum[k];
  T total;
  (void)block_excl_scan(s, part, &total);
  return total;
}

template <typename T>
__device__ __forceinline__ void scan_apply_body(T* data, uint32_t n, uint32_t lb, const T* bsum) {
  __shared__ T lds[BT_SCAN_ITEMS];
  __shared__ T part[4];
  __shared__ T opart[4];
  const uint32_t first = lb * BT_SCAN_ITEMS;
  const T offset = block_offset(bsum, lb, opart);
  load_block(data, n, first, lds);
  T v[16];
  T s = T(0);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = s;
    s += lds[threadIdx.x * 16 + k];
  }
  T total;
  T before = block_excl_scan(s, part, &total) + offset;
#pragma unroll
  for (int k = 0; k < 16; ++k) lds[threadIdx.x * 16 + k] = before + v[k];
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < BT_SCAN_ITEMS; k += 256u) {
    uint32_t i = first + k;
    if (i < n) data[i] = lds[k];
  }
}

extern "C" __global__ __launch_bounds__(256) void bt_cell_scan_reduce(
    const BtJob* __restrict__ jobs, const uint32_t* __restrict__ base, int njobs) {
  JOB_PROLOGUE(BT_K_SCAN);
  scan_reduce_body<uint32_t>(J.cells, J.n_cells + 1u, lb, J.bsum);
}

extern "C" __global__ __launch_bounds__(256) void bt_cell_scan_apply(
    const BtJob* __restrict__ jobs, const uint32_t* __restrict__ base, int njobs) {
  JOB_PROLOGUE(BT_K_SCAN);
  scan_apply_body<uint32_t>(J.cells, J.n_cells + 1u, lb, J.bsum);
}

// ------------------------------------------------------------------ tasks
// Neighbor-mean tasks (the reference's Task, spatial_grid.wgsl:117-123, one
// per <= 128 samples of a cell). A row's bins are grouped into blocks of
// task_bins consecutive bins, whose samples are contiguous in sorted order;
// each block is cut into tasks of up to 64 consecutive samples. Slots are
// reserved with one atomic per workgroup and written by all its threads (one
// block of a collapsed bundle can hold thousands of tasks). Task order is
// irrelevant to the results.
extern "C" __global__ __launch_bounds__(256) void bt_tasks(const BtJob* __restrict__ jobs,
                                                           const uint32_t* __restrict__ base,
                                                           int njobs, BtControl* ctl,
                                                           BtTask* __restrict__ tasks) {
  JOB_PROLOGUE(BT_K_TASK);
  __shared__ uint32_t part[4];
  __shared__ uint32_t block_base;
  __shared__ uint32_t first[BT_TASK_BLOCKS];
  __shared__ uint32_t bstart[BT_TASK_BLOCKS];
  __shared__ uint32_t bend[BT_TASK_BLOCKS];
  constexpr uint32_t PER = BT_TASK_BLOCKS / 256u;  // task blocks per thread
  const uint32_t nblocks = J.task_bx * J.gy;
  const uint32_t b0 = lb * BT_TASK_BLOCKS + threadIdx.x * PER;
  const uint32_t* cells = J.cells;
  uint32_t st[PER], en[PER], nt[PER];
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    const uint32_t b = b0 + (uint32_t)k;
    st[k] = en[k] = 0u;
    if (b < nblocks) {
      const uint32_t row = b / J.task_bx, bx = b - row * J.task_bx;
      const uint32_t x0 = bx * J.task_bins, x1 = min(x0 + J.task_bins, J.gx);
      st[k] = cells[row * J.gx + x0];
      en[k] = cells[row * J.gx + x1];
    }
    nt[k] = (en[k] - st[k] + BT_TASK_POINTS - 1u) / BT_TASK_POINTS;
    mine += nt[k];
  }
  uint32_t total;
  uint32_t before = block_excl_scan(mine, part, &total);
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    const uint32_t l = threadIdx.x * PER + k;
    first[l] = before;
    bstart[l] = st[k];
    bend[l] = en[k];
    before += nt[k];
  }
  if (threadIdx.x == 0) block_base = total ? atomicAdd(&ctl->task_total, total) : 0u;
  __syncthreads();
  const uint32_t pos0 = block_base;
  for (uint32_t i = threadIdx.x; i < total; i += 256u) {
    // the block holding task i: the last one whose first task is <= i
    uint32_t lo = 0, hi = BT_TASK_BLOCKS - 1u;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1u) >> 1;
      if (first[mid] <= i) lo = mid; else hi = mid - 1u;
    }
    const uint32_t s = bstart[lo] + (i - first[lo]) * BT_TASK_POINTS;
    BtTask task;
    task.job = (uint32_t)j;
    task.start = s;
    task.end = min(s + BT_TASK_POINTS, bend[lo]);
    task.pad = 0u;
    tasks[pos0 + i] = task;
  }
}

// ---------------------------------------------------------------- scatter
// Counting-sort scatter (spatial_grid.wgsl:127-138) with the rank from bt_bin.
// Also quantizes each position to the 2^-24 fixed-point grid once.
extern "C" __global__ __launch_bounds__(256) void bt_scatter(const BtJob* __restrict__ jobs,
                                                             const uint32_t* __restrict__ base,
                                                             int njobs) {
  JOB_PROLOGUE(BT_K_BIN);
  const uint32_t i = lb * BT_BIN_BLOCK + threadIdx.x;
  if (i >= J.n) return;
  const float* rec = J.rec[J.parity];
  float x = rec[3u * i], y = rec[3u * i + 1u];
  uint32_t id = __float_as_uint(rec[3u * i + 2u]);
  uint2 cr = J.cellrank[i];
  uint32_t slot = J.cells[cr.x] + cr.y;
  int fx = (int)rintf(x * BT_FIX_ONE), fy = (int)rintf(y * BT_FIX_ONE);
  J.sorted[slot] = make_float4(x, y, __int_as_float(fx), __int_as_float(fy));
  J.sid[slot] = id;
}

// Exclusive prefix sums of the sorted fixed-point coordinates, both axes in
// one pass (two's complement int64 in u64 lanes), so any run of sorted
// samples sums in O(1). The coordinates are read from the sorted records
// (their upper 8 bytes); entry n is the total. Two phases over blocks of
// BT_SCAN_ITEMS samples: block totals (x, y interleaved in bsum), then the
// block-local scan plus the sum of the totals before the block.
__device__ __forceinline__ int2 qf_of(const BtJob& J, uint32_t i) {
  if (i >= J.n) return make_int2(0, 0);
  const int2* q = reinterpret_cast<const int2*>(J.sorted);
  return q[2u * i + 1u];
}

extern "C" __global__ __launch_bounds__(256) void bt_pre_scan_reduce(
    const BtJob* __restrict__ jobs, const uint32_t* __restrict__ base, int njobs) {
  JOB_PROLOGUE(BT_K_PTS);
  __shared__ long long part[8];
  const uint32_t first = lb * BT_SCAN_ITEMS;
  long long sx = 0, sy = 0;
#pragma unroll 4
  for (uint32_t k = threadIdx.x; k < BT_SCAN_ITEMS; k += 256u) {
    const int2 q = qf_of(J, first + k);
    sx += q.x;
    sy += q.y;
  }
  sx = wave_sum64(sx);
  sy = wave_sum64(sy);
  if (LANE() == 0) {
    part[threadIdx.x >> 6] = sx;
    part[4 + (threadIdx.x >> 6)] = sy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long* bs = reinterpret_cast<unsigned long long*>(J.bsum);
    bs[2u * lb] = (unsigned long long)(part[0] + part[1] + part[2] + part[3]);
    bs[2u * lb + 1u] = (unsigned long long)(part[4] + part[5] + part[6] + part[7]);
  }
}

// 1024 threads per block: each of the 16 waves scans 256 consecutive samples
// as 4 chunks of 64 (lane l holds samples c * 64 + l), so loads and the int64
// stores are coalesced and enough waves are resident to keep HBM busy; the
// waves' totals and the block's offset from bsum are added last.
constexpr uint32_t BT_PRE_APPLY_THREADS = 1024;
extern "C" __global__ __launch_bounds__(1024) void bt_pre_scan_apply(
    const BtJob* __restrict__ jobs, const uint32_t* __restrict__ base, int njobs) {
  JOB_PROLOGUE(BT_K_PTS);
  constexpr uint32_t WAVES = BT_PRE_APPLY_THREADS / 64u;
  constexpr uint32_t PER_WAVE = BT_SCAN_ITEMS / WAVES, CHUNKS = PER_WAVE / 64u;
  __shared__ long long part[2 * WAVES];
  __shared__ long long opart[2 * WAVES];
  const uint32_t lane = LANE(), wave = threadIdx.x >> 6;
  const uint32_t w0 = lb * BT_SCAN_ITEMS + wave * PER_WAVE;
  // the block's offset: the totals of the blocks before it
  const unsigned long long* bs = reinterpret_cast<const unsigned long long*>(J.bsum);
  long long sx = 0, sy = 0;
  for (uint32_t k = threadIdx.x; k < lb; k += BT_PRE_APPLY_THREADS) {
    sx += (long long)bs[2u * k];
    sy += (long long)bs[2u * k + 1u];
  }
  sx = wave_sum64(sx);
  sy = wave_sum64(sy);
  if (lane == 0) {
    opart[wave] = sx;
    opart[WAVES + wave] = sy;
  }
  long long ex[CHUNKS], ey[CHUNKS];
  long long rx = 0, ry = 0;
#pragma unroll
  for (uint32_t c = 0; c < CHUNKS; ++c) {
    const int2 q = qf_of(J, w0 + c * 64u + lane);
    const long long ix = wave_incl_scan((long long)q.x);
    const long long iy = wave_incl_scan((long long)q.y);
    ex[c] = rx + ix - q.x;
    ey[c] = ry + iy - q.y;
    rx += __shfl(ix, 63, 64);
    ry += __shfl(iy, 63, 64);
  }
  if (lane == 0) {
    part[wave] = rx;
    part[WAVES + wave] = ry;
  }
  __syncthreads();
  long long ox = 0, oy = 0;
  for (uint32_t w = 0; w < WAVES; ++w) {
    ox += opart[w];
    oy += opart[WAVES + w];
  }
  for (uint32_t w = 0; w < wave; ++w) {
    ox += part[w];
    oy += part[WAVES + w];
  }
  const uint32_t n1 = J.n + 1u;
#pragma unroll
  for (uint32_t c = 0; c < CHUNKS; ++c) {
    const uint32_t i = w0 + c * 64u + lane;
    if (i < n1) {
      J.pre_x[i] = (uint64_t)(ox + ex[c]);
      J.pre_y[i] = (uint64_t)(oy + ey[c]);
    }
  }
}

// ---------------------------------------------------------- neighbor mean
// neighbor_mean.wgsl:75-203. One wave per task (<= 64 query samples that are
// consecutive in sorted order, from neighbouring bins of one row). The grid
// rows of the task's window are set up in parallel, one row per lane (up to
// NM_ROWS per pass): each row's bins split (uniformly for the wave) into
//   - bins fully inside the disk of every query sample (conservative, with
//     an eps margin): all their samples pass the neighbor test, so their
//     fixed-point sums come in O(1) from the prefix sums of the sorted
//     coordinates (the role of the reference's CellSummary, :56-59);
//   - boundary bins, whose samples must be tested;
//   - bins beyond every disk, skipped.
// Within a row the bins are contiguous in sorted order, so a row yields at
// most two boundary ranges; bins narrower than the row height (x_split) end
// those ranges within a fraction of a row height of the disk's edge. The
// ranges form a candidate stream, fetched in chunks with coalesced loads into
// a wave-private LDS tile (the next chunk is in flight while the current one
// is tested) and broadcast to the lanes, which test their own sample:
//   t = (q - p) * texel_scale; accept when fmaf(tx, tx, ty*ty) <= r^2.
// Long streams are set up again per lane group (nm_group_rows) and the
// longest are split across waves (bt_neighbor_mean_heavy). All sums are
// exact int64, so the result equals the all-pairs test whatever the split.
// Output is written in sorted order (the next iteration's input order), and
// every sample, fixed or not, is splatted into the density counts.

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Cell index of a texel coordinate for window bounds (multiplies by 1/cell;
// the eps margins absorb the difference from the binning's division).
__device__ __forceinline__ int cell_of_fast(float t, float inv_cell, uint32_t g) {
  float v = fminf(fmaxf(t * inv_cell, 0.0f), (float)(g - 1u));
  return (int)v;
}


// Work distribution: a work list (tasks, or parts of heavy tasks) is cut into
// BT_XCDS contiguous parts (tasks are emitted in roughly spatial order, so a
// part is a spatially coherent region); workgroup b runs on XCD b % BT_XCDS
// and first drains that XCD's part, keeping its candidate ranges in one L2,
// then takes work from the other parts. The cursors sit on separate cache
// lines; each only grows, so every wave ends after at most one failed claim
// per part.
__device__ __forceinline__ uint32_t claim_work(uint32_t* cursors, uint32_t total, uint32_t& part,
                                               uint32_t& tried, uint32_t batch = 1u,
                                               uint32_t* end = nullptr) {
  while (tried < BT_XCDS) {
    const uint32_t lo = (uint32_t)(((uint64_t)total * part) / BT_XCDS);
    const uint32_t hi = (uint32_t)(((uint64_t)total * (part + 1u)) / BT_XCDS);
    if (lo < hi) {
      uint32_t k = 0;
      if (LANE() == 0) k = atomicAdd(&cursors[part * BT_CTR_STRIDE], batch);
      k = __builtin_amdgcn_readfirstlane(k);
      if (lo + k < hi) {
        if (end) *end = min(lo + k + batch, hi);
        return lo + k;
      }
    }
    part = (part + 1u) % BT_XCDS;
    ++tried;
  }
  return 0xffffffffu;
}

#define NM_ROWS 64          // window rows set up per pass (one per lane)
#define NM_RANGES (2 * NM_ROWS)  // ranges per row pass (two per row)
#define NM_GCHUNK 64        // candidates per group per chunk of the grouped stream
#define NM_GROUP_MIN 2048u  // candidates above which a task streams per lane group
#define NM_MF_ROWS 9        // B columns of the matrix-instruction sums: 4 + 4 digits, count
#define NM_BATCH_CAND 256.0 // expected candidates per task below which tasks are claimed 4 at a time
#define NM_BATCH8_CAND 64.0 // ... and below which 8 at a time

// Kernel shape, a template parameter the autotuner chooses per iteration:
// lanes per group of the grouped stream (16: 4 groups, 32: 2 groups) and
// waves per workgroup. The grouped streams add their accepted candidates
// with matrix instructions; whole-wave streams (short tasks) with VALU adds
// (as matrix products they spilled the main kernel's registers further and
// were slower). Results do not depend on the shape (exact sums).
template <int GL, int W>
struct NmCfg {
  static constexpr uint32_t GLANES = GL;
  static constexpr uint32_t GROUPS = 64 / GL;
  static constexpr uint32_t WAVES = W;
  // per wave: the whole-wave stream's two chunks of 64 float4 candidates, or
  // the grouped stream's digit planes of B (GROUPS x NM_MF_ROWS x 64 bytes)
  // and candidate positions (GROUPS x 64 float2), and (over both, between
  // chunks) the 4 x 16 x 16 int32 results of the matrix products
  static constexpr uint32_t TILE_BYTES = 2u * 64u * 16u;
  static constexpr uint32_t MF_BYTES = GROUPS * (NM_MF_ROWS * NM_GCHUNK + NM_GCHUNK * 8u);
  static constexpr uint32_t MF_MIN = MF_BYTES > 4096u ? MF_BYTES : 4096u;
  static constexpr uint32_t WAVE_BYTES = TILE_BYTES > MF_MIN ? TILE_BYTES : MF_MIN;

  // record of a heavy task's set-up (BT_REC_WORDS words): per lane group its
  // ranges (start[128], pref[129]), then stream lengths m[GROUPS], range
  // counts n[GROUPS], then per group the inside sums x, y, count (int64)
  static constexpr uint32_t REC_GROUP = NM_RANGES + NM_RANGES + 1u;
  static constexpr uint32_t REC_M = GROUPS * REC_GROUP;
  static constexpr uint32_t REC_N = REC_M + GROUPS;
  static constexpr uint32_t REC_SUMS = REC_N + GROUPS + ((REC_N + GROUPS) & 1u);
  static_assert(REC_SUMS + 6 * GROUPS <= BT_REC_WORDS, "heavy record layout");
  static_assert(REC_SUMS % 2 == 0, "int64 sums must be 8-byte aligned");
  static_assert(GL == 16 || GL == 32, "lane groups of 16 or 32");
};

// Per wave, `raw` holds either two chunks of 64 candidates (whole-wave
// stream, double buffered, as float4 records) or the grouped stream's
// matrix-instruction layout.
template <class C>
struct NmLds {
  alignas(16) unsigned char raw[C::WAVES][C::WAVE_BYTES];
  uint32_t start[C::WAVES][C::GROUPS][NM_RANGES];
  uint32_t pref[C::WAVES][C::GROUPS][NM_RANGES + 1];
  __device__ float4* tile(uint32_t w) { return reinterpret_cast<float4*>(raw[w]); }
};

// Per-wave state of one task.
struct NmTask {
  const BtJob* J;
  uint32_t slot, id;
  bool valid, movable;
  float px, py;
  int pfx, pfy;
  // window
  float tx0, tx1, ty0, ty1;
  uint32_t cy0, cy1;
};

__device__ __forceinline__ void nm_load(NmTask& T, const BtJob* jobs, const BtTask* tasks,
                                        uint32_t t) {
  const uint32_t lane = LANE();
  const uint32_t tj = uni(tasks[t].job);
  const uint32_t ts = uni(tasks[t].start);
  const uint32_t te = uni(tasks[t].end);
  T.J = &jobs[tj];
  const BtJob& J = *T.J;
  T.slot = ts + lane;
  T.valid = T.slot < te;
  float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
  T.id = 0;
  if (T.valid) {
    q = J.sorted[T.slot];
    T.id = J.sid[T.slot];
  }
  T.movable = T.valid && !(T.id & BT_FIXED_BIT);
  T.px = q.x;
  T.py = q.y;
  T.pfx = __float_as_int(q.z);
  T.pfy = __float_as_int(q.w);
  const bool mv = T.movable;
  // query bbox in texels (the same texel() as the binning; monotone)
  T.tx0 = __uint_as_float(uni(__float_as_uint(texel(wave_min(mv ? T.px : 3.0e38f), J.res_x))));
  T.tx1 = __uint_as_float(uni(__float_as_uint(texel(wave_max(mv ? T.px : -3.0e38f), J.res_x))));
  T.ty0 = __uint_as_float(uni(__float_as_uint(texel(wave_min(mv ? T.py : 3.0e38f), J.res_y))));
  T.ty1 = __uint_as_float(uni(__float_as_uint(texel(wave_max(mv ? T.py : -3.0e38f), J.res_y))));
  const float rr = J.radius + J.eps;
  T.cy0 = uni((uint32_t)cell_of_fast(T.ty0 - rr, J.inv_cell, J.gy));
  T.cy1 = uni((uint32_t)cell_of_fast(T.ty1 + rr, J.inv_cell, J.gy));
}

// Sets up rows [rb, min(rb + NM_ROWS, cy1 + 1)) of the window of the query box
// [tx0, tx1] x [ty0, ty1] (texels, wave-uniform): the boundary ranges go to
// start/pref (LDS), the inside-run sums to (sx, sy, sn) (wave-uniform).
// Returns the number of candidates m and of ranges.
__device__ __forceinline__ void nm_rows(const BtJob& J, float tx0, float tx1, float ty0, float ty1,
                                        uint32_t cy1, uint32_t rb, uint32_t* start,
                                        uint32_t* pref, long long& sx, long long& sy,
                                        long long& sn, uint32_t& m, uint32_t& nranges) {
  const uint32_t lane = LANE();
  const float c = J.cell, inv_w = J.inv_cell_w, eps = J.eps, r = J.radius;
  const float rr = r + eps, ri = r - eps;
  const uint32_t gx = J.gx, gy = J.gy;
  const uint32_t* cells = J.cells;
  const uint32_t cy = rb + lane;
  uint32_t r0b = 0, r0e = 0, r1b = 0, r1e = 0;
  long long lx = 0, ly = 0, ln = 0;
  if (lane < NM_ROWS && cy <= cy1) {
    const float Y0 = (float)cy * c, Y1 = Y0 + c;
    const float dyn = fmaxf(fmaxf(Y0 - eps - ty1, ty0 - (Y1 + eps)), 0.0f);
    if (dyn < rr) {
      const float wout = __builtin_sqrtf(rr * rr - dyn * dyn);
      const uint32_t xlo = (uint32_t)cell_of_fast(tx0 - wout, inv_w, gx);
      const uint32_t xhi = (uint32_t)cell_of_fast(tx1 + wout, inv_w, gx);
      int xa = 1, xb = 0;
      const float dyf = fmaxf(Y1 + eps - ty0, ty1 - (Y0 - eps));
      if (cy > 0u && cy + 1u < gy && dyf < ri) {
        const float win = __builtin_sqrtf(ri * ri - dyf * dyf);
        const float fa = ceilf((tx1 - win + eps) * inv_w);
        const float fb = floorf((tx0 + win - eps) * inv_w) - 1.0f;
        xa = (int)fminf(fmaxf(fa, fmaxf((float)xlo, 1.0f)), (float)gx);
        xb = (int)fmaxf(fminf(fb, fminf((float)xhi, (float)gx - 2.0f)), -1.0f);
      }
      const uint32_t row = cy * gx;
      const uint32_t s_lo = cells[row + xlo];
      const uint32_t s_hi = cells[row + xhi + 1u];
      if (xa <= xb) {
        const uint32_t s_a = cells[row + (uint32_t)xa];
        const uint32_t s_b = cells[row + (uint32_t)xb + 1u];
        lx = (long long)(J.pre_x[s_b] - J.pre_x[s_a]);
        ly = (long long)(J.pre_y[s_b] - J.pre_y[s_a]);
        ln = (long long)(s_b - s_a);
        r0b = s_lo;
        r0e = s_a;
        r1b = s_b;
        r1e = s_hi;
      } else {
        r0b = s_lo;
        r0e = s_hi;
      }
    }
  }
  sx = wave_sum64(lx);
  sy = wave_sum64(ly);
  sn = wave_sum64(ln);
  const uint32_t l0 = r0e - r0b, l1 = r1e - r1b;
  const uint32_t nr = (l0 ? 1u : 0u) + (l1 ? 1u : 0u);
  const uint32_t len = l0 + l1;
  const uint32_t rpos = wave_incl_scan(nr) - nr;
  const uint32_t lpos = wave_incl_scan(len) - len;
  nranges = uni(__shfl(rpos + nr, 63, 64));
  m = uni(__shfl(lpos + len, 63, 64));
  if (l0) {
    start[rpos] = r0b;
    pref[rpos] = lpos;
  }
  if (l1) {
    start[rpos + (l0 ? 1u : 0u)] = r1b;
    pref[rpos + (l0 ? 1u : 0u)] = lpos + l0;
  }
  if (lane == 0) pref[nranges] = m;
  __builtin_amdgcn_wave_barrier();
}

typedef __attribute__((address_space(1))) const f4v gf4v;

// A load from the global address space (its wait counter is separate from
// LDS waits; a generic pointer would make it a flat load, counted with both).
__device__ __forceinline__ float4 ld_global(const float4* p, uint32_t i) {
  const f4v v = ((gf4v*)(uintptr_t)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}

// Candidate v of a stream set up by nm_rows (v < m). `r` is the lane's range
// cursor: the range holding the lane's previous candidate (0 at the start);
// a lane's candidates only move forward, so the cursor advances past the
// ranges that end at or before v (usually none or one per call). Loads from
// the global address space, so the wait for them is separate from LDS waits.
__device__ __forceinline__ float4 nm_candidate(const float4* S, const uint32_t* start,
                                               const uint32_t* pref, uint32_t nranges,
                                               uint32_t v, uint32_t& r) {
  while (r + 1u < nranges && pref[r + 1u] <= v) ++r;
  return ld_global(S, start[r] + (v - pref[r]));
}

// The neighbor test of one candidate against the lane's sample, adding the
// candidate's fixed-point coordinates and 1 to the chunk sums when accepted.
#define NM_TEST(s)                                            \
  {                                                           \
    const float dx = ((s).x - px) * tsx;                      \
    const float dy = ((s).y - py) * tsy;                      \
    const bool ok = fmaf(dx, dx, dy * dy) <= r2;              \
    sx += ok ? __float_as_int((s).z) : 0;                     \
    sy += ok ? __float_as_int((s).w) : 0;                     \
    sc += ok ? 1 : 0;                                         \
  }

// Tests the candidate chunks first, first + step, ... of the stream set up in
// group 0's LDS ranges; adds the lane's accepted sums to (ax, ay, n).
template <class C>
__device__ __forceinline__ void nm_stream(const NmTask& T, NmLds<C>& L, uint32_t m, uint32_t nranges,
                                          uint32_t first, uint32_t step, long long& ax,
                                          long long& ay, long long& n, uint64_t& steps) {
  const uint32_t lane = LANE(), w = threadIdx.x >> 6;
  const BtJob& J = *T.J;
  const float tsx = J.ts_x, tsy = J.ts_y, r2 = J.r2, px = T.px, py = T.py;
  const float4* S = J.sorted;
  const uint32_t* start = L.start[w][0];
  const uint32_t* pref = L.pref[w][0];
  const uint32_t nch = (m + 63u) >> 6;
  uint32_t cur = 0;
  auto fetch = [&](uint32_t ch) -> float4 {
    const uint32_t v = (ch << 6) + lane;
    return v < m ? nm_candidate(S, start, pref, nranges, v, cur)
                 : make_float4(0.f, 0.f, 0.f, 0.f);
  };
  float4 nxt = first < nch ? fetch(first) : make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t b = 0;
  for (uint32_t ch = first; ch < nch; ch += step) {
    float4* tile = L.tile(w) + b * 64u;
    tile[lane] = nxt;
    if (ch + step < nch) nxt = fetch(ch + step);
    __builtin_amdgcn_wave_barrier();
    const uint32_t cnt = min(64u, m - (ch << 6));
    steps += cnt;
    int sx = 0, sy = 0, sc = 0;
    uint32_t k = 0;
    for (; k + 4u <= cnt; k += 4u) {
#pragma unroll
      for (uint32_t u = 0; u < 4u; ++u) {
        const float4 s = tile[k + u];
        NM_TEST(s)
      }
    }
    for (; k < cnt; ++k) {
      const float4 s = tile[k];
      NM_TEST(s)
    }
    ax += sx;
    ay += sy;
    n += sc;
    b ^= 1u;
    __builtin_amdgcn_wave_barrier();
  }
}

template <class C>
__device__ __forceinline__ float group_min(float v) {
#pragma unroll
  for (int d = C::GLANES / 2; d > 0; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
  return v;
}
template <class C>
__device__ __forceinline__ float group_max(float v) {
#pragma unroll
  for (int d = C::GLANES / 2; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
  return v;
}

// Grouped streams of a task whose window is one row group: the wave's lanes
// form 64 / GLANES groups of GLANES consecutive samples (neighbours along x),
// each with its own query box, inside sums and boundary ranges, which are
// narrower than the whole task's. Per step, each group's lanes test the next
// candidate of their group's stream (exhausted streams read candidates at
// infinity, which fail the test), so the task takes max over groups of the
// stream length instead of the length of the task's union stream.
struct NmGroups {
  uint32_t my_m, my_n;  // the lane's group: stream length and ranges
  uint32_t mmax;        // longest group stream (wave-uniform)
  long long sx, sy, sn; // inside sums of the lane's group
};

// Sets up every group's ranges (LDS) and inside sums.
template <class C>
__device__ __forceinline__ NmGroups nm_group_rows(const NmTask& T, NmLds<C>& L) {
  const uint32_t lane = LANE(), w = threadIdx.x >> 6, grp = lane / C::GLANES;
  const BtJob& J = *T.J;
  const bool mv = T.movable;
  const float gx0 = group_min<C>(mv ? T.px : 3.0e38f), gx1 = group_max<C>(mv ? T.px : -3.0e38f);
  const float gy0 = group_min<C>(mv ? T.py : 3.0e38f), gy1 = group_max<C>(mv ? T.py : -3.0e38f);
  const float rr = J.radius + J.eps;
  NmGroups G;
  G.my_m = G.my_n = G.mmax = 0u;
  G.sx = G.sy = G.sn = 0;
  const uint64_t movable = ballot(mv);
#pragma unroll
  for (uint32_t g = 0; g < C::GROUPS; ++g) {
    const uint32_t l = g * C::GLANES;
    if (!(movable & (((1ull << C::GLANES) - 1ull) << l))) continue;  // no movable sample
    const float tx0 = __uint_as_float(uni(__float_as_uint(texel(__shfl(gx0, (int)l, 64), J.res_x))));
    const float tx1 = __uint_as_float(uni(__float_as_uint(texel(__shfl(gx1, (int)l, 64), J.res_x))));
    const float ty0 = __uint_as_float(uni(__float_as_uint(texel(__shfl(gy0, (int)l, 64), J.res_y))));
    const float ty1 = __uint_as_float(uni(__float_as_uint(texel(__shfl(gy1, (int)l, 64), J.res_y))));
    const uint32_t c0 = uni((uint32_t)cell_of_fast(ty0 - rr, J.inv_cell, J.gy));
    const uint32_t c1 = uni((uint32_t)cell_of_fast(ty1 + rr, J.inv_cell, J.gy));
    long long sx, sy, sn;
    uint32_t m, nr;
    nm_rows(J, tx0, tx1, ty0, ty1, c1, c0, L.start[w][g], L.pref[w][g], sx, sy, sn, m, nr);
    if (grp == g) {
      G.sx = sx;
      G.sy = sy;
      G.sn = sn;
      G.my_m = m;
      G.my_n = nr;
    }
    G.mmax = max(G.mmax, m);
  }
  return G;
}

typedef int nm_v4i __attribute__((ext_vector_type(4)));

// Folds the four 16 x 16 matrix-product accumulators of the grouped MF
// stream into each lane's exact sums. Accumulator m (one per lane group)
// holds, in lane l = 16 q + c, register r, minus the column-c total of
// sample row 4 q + r of group m (the sample of lane 16 m + 4 q + r): columns
// 0-3 the 7-bit digits of (x fixed point + 2^24), 4-7 those of y, 8 the
// count. The products pass through LDS (red: 4 x 16 x 16 int32, over the
// stream's staging area, free between chunks).
__device__ __forceinline__ void nm_mf_fold(nm_v4i (&D)[4], int* red, long long& ax,
                                           long long& ay, long long& n) {
  const uint32_t lane = LANE(), q = lane >> 4, c = lane & 15u;
#pragma unroll
  for (uint32_t m = 0; m < 4u; ++m)
#pragma unroll
    for (uint32_t r = 0; r < 4u; ++r) red[(m * 16u + 4u * q + r) * 16u + c] = D[m][r];
  __builtin_amdgcn_wave_barrier();
  // A holds -1 for an accepted candidate: every total is the negated sum
  const int* row = red + lane * 16u;  // this lane's sample: group lane / 16, row lane % 16
  const long long sx = (long long)row[0] + (long long)row[1] * 128ll +
                       (long long)row[2] * 16384ll + (long long)row[3] * 2097152ll;
  const long long sy = (long long)row[4] + (long long)row[5] * 128ll +
                       (long long)row[6] * 16384ll + (long long)row[7] * 2097152ll;
  const long long cnt = row[8];  // minus the count
  ax += cnt * 16777216ll - sx;
  ay += cnt * 16777216ll - sy;
  n -= cnt;
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (uint32_t m = 0; m < 4u; ++m) D[m] = nm_v4i{0, 0, 0, 0};
}

// A candidate stream with the accepted sums as matrix products. The wave's
// 64 samples form 64 / GL lane groups of GL consecutive lanes, each group
// with its own stream (GL = 64: one stream for the whole wave). Per chunk,
// each group's next 64 candidates are staged as positions and as the 7-bit
// digits of their fixed-point coordinates plus 2^24 (B, 64 x 9: x digits, y
// digits, a column of ones; |coordinate| <= 1, so the biased value fits 25
// bits). Matrix product m (m = 0..3) covers the samples of lanes 16 m ..
// 16 m + 15, whose candidates are those of their group; lane l = 16 q + i
// tests candidates 16 q .. 16 q + 15 of that group against the sample of
// lane 16 m + i, giving the bytes of A (16 x 64: -1 accepted, 0 not), and
// v_mfma_i32_16x16x64_i8 adds A x B to accumulator m. Integer products: the
// sums are exact and equal the VALU stream's. Accumulators fold into the
// lane sums at the end (and every 2^16 chunks, far below int32 overflow: a
// chunk adds at most 64 x 127 per entry). Exhausted group streams stage
// candidates at infinity, which fail the test.
//
// The test is the reference's fmaf(dx, dx, dy * dy) <= r2, bit for bit.
// d2 = fmaf(...) is +0, positive or +inf, and non-negative floats order as
// their bit patterns, so d2 <= r2 exactly when bits(d2) - (bits(r2) + 1) <
// 0 as int32. v_perm_b32's sign-replicating selectors (9: bit 31 of the low
// source, 11: bit 31 of the high one) turn two such differences into two A
// bytes of 0xFF or 0: a subtraction per test and three instructions per
// four tests, where a compare, a select and half an OR took 2.5.
template <int GL, class C, class Fetch>
__device__ __forceinline__ void nm_mf_stream(const NmTask& T, NmLds<C>& L, uint32_t nch,
                                             uint32_t mmax, uint32_t first, uint32_t step,
                                             Fetch fetch, long long& ax, long long& ay,
                                             long long& n, uint64_t& steps) {
  static_assert(GL == 16 || GL == 32 || GL == 64, "lane groups of 16, 32 or 64");
  static_assert(C::WAVE_BYTES >= (64u / GL) * (NM_MF_ROWS * NM_GCHUNK + NM_GCHUNK * 8u) &&
                    C::WAVE_BYTES >= 4096u, "matrix-product stream layout");
  constexpr uint32_t G = 64u / GL;        // groups (streams)
  constexpr uint32_t PER = NM_GCHUNK / GL;  // candidates each lane stages per chunk
  constexpr uint32_t RB = GL / 16u;       // products per group
  const uint32_t lane = LANE(), w = threadIdx.x >> 6, grp = lane / GL, sub = lane % GL;
  const uint32_t q = lane >> 4, i16 = lane & 15u;
  const BtJob& J = *T.J;
  const float tsx = J.ts_x, tsy = J.ts_y, r2 = J.r2;
  // bits(r2) + 1; a NaN radius accepts nothing, as the float compare does
  const int rb = r2 == r2 ? (int)__float_as_uint(r2) + 1 : 0;
  unsigned char* bt = L.raw[w];                                             // [G][9][64]
  float2* xy = reinterpret_cast<float2*>(L.raw[w] + G * NM_MF_ROWS * NM_GCHUNK);  // [G][64]
  int* red = reinterpret_cast<int*>(L.raw[w]);
  auto ones = [&]() {  // the count column: row 8 of each group's B
    if (lane < 16u * G)
      reinterpret_cast<uint32_t*>(bt + ((lane >> 4) * NM_MF_ROWS + 8u) * NM_GCHUNK)[i16] =
          0x01010101u;
  };
  ones();
  // the samples of this lane's rows (row i16 of each product)
  float qx[4], qy[4];
#pragma unroll
  for (uint32_t m = 0; m < 4u; ++m) {
    qx[m] = __shfl(T.px, (int)(16u * m + i16), 64);
    qy[m] = __shfl(T.py, (int)(16u * m + i16), 64);
  }
  nm_v4i D[4];
#pragma unroll
  for (uint32_t m = 0; m < 4u; ++m) D[m] = nm_v4i{0, 0, 0, 0};
  float4 nxt[PER];
  if (first < nch) {
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) nxt[k] = fetch(first, k);
  }
  uint32_t since = 0;
  for (uint32_t ch = first; ch < nch; ch += step) {
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
      const uint32_t c = k * GL + sub;
      xy[grp * NM_GCHUNK + c] = make_float2(nxt[k].x, nxt[k].y);
      const uint32_t ux = (uint32_t)(__float_as_int(nxt[k].z) + (1 << 24));
      const uint32_t uy = (uint32_t)(__float_as_int(nxt[k].w) + (1 << 24));
      unsigned char* col = bt + grp * NM_MF_ROWS * NM_GCHUNK + c;
      col[0] = (unsigned char)(ux & 127u);
      col[NM_GCHUNK] = (unsigned char)((ux >> 7) & 127u);
      col[2 * NM_GCHUNK] = (unsigned char)((ux >> 14) & 127u);
      col[3 * NM_GCHUNK] = (unsigned char)(ux >> 21);
      col[4 * NM_GCHUNK] = (unsigned char)(uy & 127u);
      col[5 * NM_GCHUNK] = (unsigned char)((uy >> 7) & 127u);
      col[6 * NM_GCHUNK] = (unsigned char)((uy >> 14) & 127u);
      col[7 * NM_GCHUNK] = (unsigned char)(uy >> 21);
    }
    __builtin_amdgcn_wave_barrier();
    if (ch + step < nch) {
#pragma unroll
      for (uint32_t k = 0; k < PER; ++k) nxt[k] = fetch(ch + step, k);
    }
    steps += min((uint32_t)NM_GCHUNK, mmax - ch * (uint32_t)NM_GCHUNK);
#pragma unroll
    for (uint32_t m = 0; m < 4u; ++m) {
      const uint32_t g = m / RB;
      const float2* cand = xy + g * NM_GCHUNK + 16u * q;
      uint32_t d[16];
#pragma unroll
      for (uint32_t j = 0; j < 16u; ++j) {
        const float2 c = cand[j];
        const float dx = (c.x - qx[m]) * tsx;
        const float dy = (c.y - qy[m]) * tsy;
        d[j] = (uint32_t)(__float_as_int(fmaf(dx, dx, dy * dy)) - rb);
      }
      nm_v4i a;
#pragma unroll
      for (uint32_t t = 0; t < 4u; ++t)
        a[t] = (int)(__builtin_amdgcn_perm(d[4 * t + 1], d[4 * t], 0x0c0c0b09u) |
                     __builtin_amdgcn_perm(d[4 * t + 3], d[4 * t + 2], 0x0b090c0cu));
      const nm_v4i b = *reinterpret_cast<const nm_v4i*>(
          bt + (g * NM_MF_ROWS + min(i16, NM_MF_ROWS - 1u)) * NM_GCHUNK + 16u * q);
      D[m] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, D[m], 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
    if (++since == 65536u) {
      nm_mf_fold(D, red, ax, ay, n);
      ones();
      since = 0;
    }
  }
  nm_mf_fold(D, red, ax, ay, n);
}

// The grouped stream (C::GLANES-lane groups) as matrix products.
template <class C>
__device__ __forceinline__ void nm_group_stream_mf(const NmTask& T, NmLds<C>& L,
                                                   const NmGroups& G, uint32_t first,
                                                   uint32_t step, long long& ax, long long& ay,
                                                   long long& n, uint64_t& steps) {
  const uint32_t lane = LANE(), w = threadIdx.x >> 6, grp = lane / C::GLANES;
  const uint32_t sub = lane % C::GLANES;
  const float4* S = T.J->sorted;
  const uint32_t* start = L.start[w][grp];
  const uint32_t* pref = L.pref[w][grp];
  const float inf = __int_as_float(0x7f800000);
  uint32_t cur = 0;
  auto fetch = [&](uint32_t ch, uint32_t k) -> float4 {
    const uint32_t v = ch * NM_GCHUNK + k * C::GLANES + sub;
    return v < G.my_m ? nm_candidate(S, start, pref, G.my_n, v, cur)
                      : make_float4(inf, inf, 0.f, 0.f);
  };
  const uint32_t nch = (G.mmax + NM_GCHUNK - 1u) / NM_GCHUNK;
  nm_mf_stream<C::GLANES>(T, L, nch, G.mmax, first, step, fetch, ax, ay, n, steps);
}

// Tests the group chunks first, first + step, ... (NM_GCHUNK steps each);
// adds the lane's accepted sums to (ax, ay, n).
template <class C>
__device__ __forceinline__ void nm_group_stream(const NmTask& T, NmLds<C>& L, const NmGroups& G,
                                                uint32_t first, uint32_t step, long long& ax,
                                                long long& ay, long long& n, uint64_t& steps) {
  nm_group_stream_mf<C>(T, L, G, first, step, ax, ay, n, steps);
}

// Stores a heavy task's set-up as its record (NmCfg's REC_ layout).
template <class C>
__device__ __forceinline__ void nm_store_groups(const NmGroups& G, NmLds<C>& L, uint32_t* rec) {
  const uint32_t lane = LANE(), w = threadIdx.x >> 6, grp = lane / C::GLANES;
  for (uint32_t i = lane; i < C::REC_M; i += 64u) {
    const uint32_t g = i / C::REC_GROUP, o = i - g * C::REC_GROUP;
    rec[i] = o < NM_RANGES ? L.start[w][g][o] : L.pref[w][g][o - NM_RANGES];
  }
  if (lane % C::GLANES == 0) {
    rec[C::REC_M + grp] = G.my_m;
    rec[C::REC_N + grp] = G.my_n;
    long long* sums = (long long*)(rec + C::REC_SUMS) + 3u * grp;
    sums[0] = G.sx;
    sums[1] = G.sy;
    sums[2] = G.sn;
  }
}

template <class C>
__device__ __forceinline__ NmGroups nm_load_groups(NmLds<C>& L, const uint32_t* rec) {
  const uint32_t lane = LANE(), w = threadIdx.x >> 6, grp = lane / C::GLANES;
  for (uint32_t i = lane; i < C::REC_M; i += 64u) {
    const uint32_t g = i / C::REC_GROUP, o = i - g * C::REC_GROUP;
    const uint32_t v = rec[i];
    if (o < NM_RANGES) L.start[w][g][o] = v; else L.pref[w][g][o - NM_RANGES] = v;
  }
  NmGroups G;
  G.mmax = 0u;
#pragma unroll
  for (uint32_t g = 0; g < C::GROUPS; ++g) G.mmax = max(G.mmax, uni(rec[C::REC_M + g]));
  G.my_m = rec[C::REC_M + grp];
  G.my_n = rec[C::REC_N + grp];
  const long long* sums = (const long long*)(rec + C::REC_SUMS) + 3u * grp;
  G.sx = sums[0];
  G.sy = sums[1];
  G.sn = sums[2];
  __builtin_amdgcn_wave_barrier();
  return G;
}

// Mean of a movable sample from its exact sums, output record, density splat
// (one atomic per distinct pixel of the wave).
__device__ __forceinline__ uint64_t nm_finish(const NmTask& T, long long sx, long long sy,
                                              long long tn) {
  const uint32_t lane = LANE();
  const BtJob& J = *T.J;
  // neighbors of the wave's movable samples (executed-work report)
  const uint64_t nbr = wave_sum_u64(T.movable ? (unsigned long long)tn : 0ull);
  float mx = T.px, my = T.py;
  if (T.movable) {
    const long long dxs = sx - tn * (long long)T.pfx, dys = sy - tn * (long long)T.pfy;
    const double den = (double)tn * 16777216.0;
    mx = (float)((double)T.px + (double)dxs / den);
    my = (float)((double)T.py + (double)dys / den);
  }
  if (T.valid) {
    float* out = J.rec[J.parity ^ 1u];
    out[3u * T.slot] = mx;
    out[3u * T.slot + 1u] = my;
    out[3u * T.slot + 2u] = __uint_as_float(T.id);
  }
  if (J.dens_cnt) {
    const uint32_t pix = T.valid ? splat_index(mx, my, J.dw, J.dh) : 0u;
    uint64_t act = ballot(T.valid);
    while (act) {
      const int leader = __builtin_ctzll(act);
      const uint32_t key = __builtin_amdgcn_readlane(pix, leader);
      const uint64_t same = ballot(T.valid && pix == key);
      if (lane == (uint32_t)leader) atomicAdd(&J.dens_cnt[key], (uint32_t)__popcll(same));
      act &= ~same;
    }
  }
  return nbr;
}

// Per-workgroup totals of a launch (plain stores at exit, summed on the
// host): wave-steps executed (one candidate tested against each of the 64
// lanes' samples) and neighbors found.
template <class C>
__device__ __forceinline__ void nm_store_work(unsigned long long* steps_out,
                                              unsigned long long* nbr_out, uint64_t steps,
                                              uint64_t nbr, unsigned long long (*red)[2]) {
  if (LANE() == 0) {
    red[threadIdx.x >> 6][0] = steps;
    red[threadIdx.x >> 6][1] = nbr;
  }
  __syncthreads();
  if (threadIdx.x == 0 && blockIdx.x < BT_MAX_WG) {
    unsigned long long a = 0, b = 0;
    for (uint32_t w = 0; w < C::WAVES; ++w) {
      a += red[w][0];
      b += red[w][1];
    }
    steps_out[blockIdx.x] = a;
    if (nbr_out) nbr_out[blockIdx.x] = b;
  }
  __syncthreads();  // red is reused by a second call
}

// 4 waves per SIMD (16 per CU) whatever the workgroup size: at most 128
// VGPRs per lane.
template <int GL, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4, 4))) void
bt_neighbor_mean(
    const BtJob* __restrict__ jobs, const BtTask* __restrict__ tasks, BtControl* ctl,
    BtSplit split, unsigned long long samples) {
  typedef NmCfg<GL, W> C;
  __shared__ NmLds<C> L;
  __shared__ unsigned long long wave_cand[C::WAVES];
  const uint32_t lane = LANE(), w = threadIdx.x >> 6;
  const uint32_t total = uni(__atomic_load_n(&ctl->task_total, __ATOMIC_RELAXED));
  // split threshold from the previous launch's candidates per wave
  const unsigned long long prev = ctl->cand_prev, sprev = ctl->samples_prev;
  uint32_t heavy = BT_HEAVY_HI;
  if (prev) {
    const double expect = sprev ? (double)prev * ((double)samples / (double)sprev) : (double)prev;
    const unsigned long long per_wave =
        (unsigned long long)(expect / ((double)gridDim.x * C::WAVES));
    heavy = (uint32_t)min(max(per_wave / 2ull, (unsigned long long)BT_HEAVY_LO),
                          (unsigned long long)BT_HEAVY_HI);
  }
  const uint32_t part_steps = max(heavy / 2u, BT_PART_STEPS_MIN);
  // Tasks claimed per work-queue atomic: launches of short tasks (the
  // small-radius iterations, where each task is a chain of dependent loads
  // and the claims on the 8 cursors are a visible part of it) take 4 per
  // claim, 8 when a task expects under 64 candidates; longer tasks one, for
  // the balance at the end of the launch.
  uint32_t batch = 1u;
  if (prev && total) {
    const double expect = sprev ? (double)prev * ((double)samples / (double)sprev) : (double)prev;
    const double per_task = expect / (double)total;
    batch = per_task < NM_BATCH8_CAND ? 8u : per_task < NM_BATCH_CAND ? 4u : 1u;
  }
  unsigned long long cand = 0;
  uint64_t steps = 0, gsteps = 0, nbr = 0;
  uint32_t part = blockIdx.x % BT_XCDS, tried = 0;
  for (;;) {
    uint32_t tend = 0;
    const uint32_t t0 = uni(claim_work(ctl->xq, total, part, tried, batch, &tend));
    if (t0 >= total) break;
    tend = uni(tend);
    for (uint32_t t = t0; t < tend; ++t) {
    NmTask T;
    nm_load(T, jobs, tasks, t);
    long long ax = 0, ay = 0, n = 0, ux = 0, uy = 0, un = 0;
    if (ballot(T.movable)) {
      bool deferred = false;
      for (uint32_t rb = T.cy0; rb <= T.cy1; rb += NM_ROWS) {
        uint32_t m, nranges;
        long long sx, sy, sn;
        nm_rows(*T.J, T.tx0, T.tx1, T.ty0, T.ty1, T.cy1, rb, L.start[w][0], L.pref[w][0], sx, sy,
                sn, m, nranges);
        cand += m;
        const bool one_group = rb == T.cy0 && T.cy1 - T.cy0 < NM_ROWS;
        if (one_group && m > NM_GROUP_MIN) {
          // grouped streams; a long one is recorded for the split kernel
          const NmGroups G = nm_group_rows<C>(T, L);  // its own ranges and inside sums
          if (m > heavy) {
            const uint32_t parts = min(BT_MAX_PARTS, (G.mmax + part_steps - 1u) / part_steps);
            uint32_t h = 0, base = 0;
            if (lane == 0) {
              h = atomicAdd(&ctl->heavy_count, 1u);
              base = h < BT_MAX_HEAVY ? atomicAdd(&ctl->item_count, parts) : BT_MAX_ITEMS;
            }
            h = uni(h);
            base = uni(base);
            if (h < BT_MAX_HEAVY && base + parts <= BT_MAX_ITEMS) {
              nm_store_groups<C>(G, L, split.rec + (size_t)h * BT_REC_WORDS);
              for (uint32_t p = lane; p < parts; p += 64u) split.items[base + p] = make_uint2(h, p);
              long long* acc = split.acc + (size_t)h * 192u;
              acc[lane] = 0;
              acc[64u + lane] = 0;
              acc[128u + lane] = 0;
              if (lane == 0) {
                BtHeavy hv;
                hv.task = t;
                hv.parts = parts;
                hv.done = 0u;
                hv.pad = 0u;
                split.heavy[h] = hv;
              }
              deferred = true;
              break;
            }
            // No room: run it here; neutralize the part slots reserved in range.
            for (uint32_t p = lane; p < parts; p += 64u)
              if (base + p < BT_MAX_ITEMS) split.items[base + p] = make_uint2(0xffffffffu, 0u);
          }
          ax += G.sx;
          ay += G.sy;
          n += G.sn;
          nm_group_stream<C>(T, L, G, 0u, 1u, ax, ay, n, gsteps);
          break;
        }
        ux += sx;
        uy += sy;
        un += sn;
        nm_stream<C>(T, L, m, nranges, 0u, 1u, ax, ay, n, steps);
      }
      if (deferred) continue;
    }
    nbr += nm_finish(T, ax + ux, ay + uy, n + un);
    }
  }
  // this workgroup's candidates (every wave leaves the loop above)
  if (lane == 0) wave_cand[threadIdx.x >> 6] = cand;
  __syncthreads();
  if (threadIdx.x == 0 && blockIdx.x < BT_MAX_WG) {
    unsigned long long s = 0;
    for (uint32_t w = 0; w < C::WAVES; ++w) s += wave_cand[w];
    split.wg_cand[blockIdx.x] = s;
  }
  __shared__ unsigned long long work[C::WAVES][2];
  nm_store_work<C>(split.wg_steps, split.wg_nbr, steps + gsteps, nbr, work);
  nm_store_work<C>(split.wg_gsteps, nullptr, gsteps, 0, work);
}

// Runs the parts of heavy tasks: part p tests group chunks p, p + P, p + 2P,
// ... of the task's lane-group streams (set up by the main kernel and stored
// in the task's record) and adds its exact per-lane sums to the task's
// accumulators (part 0 also the inside-run sums). The part that finishes
// last (a counter; no wave waits) reads the accumulators and writes the
// means. Block 0 first sums the main launch's per-workgroup candidate counts
// into the control block, for the next launch's split threshold.
template <int GL, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) void
bt_neighbor_mean_heavy(
    const BtJob* __restrict__ jobs, const BtTask* __restrict__ tasks, BtControl* ctl,
    BtSplit split, unsigned long long samples) {
  typedef NmCfg<GL, W> C;
  __shared__ NmLds<C> L;
  const uint32_t lane = LANE();
  const uint32_t items = min(uni(__atomic_load_n(&ctl->item_count, __ATOMIC_RELAXED)),
                             BT_MAX_ITEMS);
  if (blockIdx.x == 0) {
    // total candidates of the main launch (same grid), for the next threshold
    __shared__ unsigned long long red[C::WAVES];
    unsigned long long s = 0;
    for (uint32_t b = threadIdx.x; b < min(gridDim.x, BT_MAX_WG); b += 64u * C::WAVES)
      s += split.wg_cand[b];
    s = wave_sum_u64(s);
    if (lane == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (uint32_t w = 0; w < C::WAVES; ++w) t += red[w];
      ctl->cand_prev = t;
      ctl->samples_prev = samples;
    }
  }
  uint32_t part = blockIdx.x % BT_XCDS, tried = 0;
  uint64_t steps = 0, nbr = 0;
  for (;;) {
    const uint32_t i = uni(claim_work(ctl->iq, items, part, tried));
    if (i >= items) break;
    const uint2 it = split.items[i];
    const uint32_t h = uni(it.x), p = uni(it.y);
    if (h >= BT_MAX_HEAVY) continue;  // a neutralized slot
    const uint32_t t = uni(split.heavy[h].task);
    const uint32_t parts = uni(split.heavy[h].parts);
    NmTask T;
    nm_load(T, jobs, tasks, t);
    long long ax = 0, ay = 0, n = 0, ux = 0, uy = 0, un = 0;
    // the set-up the main kernel stored; part p takes group chunks p,
    // p + parts, ...
    const NmGroups G = nm_load_groups<C>(L, split.rec + (size_t)h * BT_REC_WORDS);
    ux = G.sx;
    uy = G.sy;
    un = G.sn;
    nm_group_stream<C>(T, L, G, p, parts, ax, ay, n, steps);
    if (p == 0u) {
      ax += ux;
      ay += uy;
      n += un;
    }
    long long* acc = split.acc + (size_t)h * 192u;
    atomicAdd((unsigned long long*)&acc[lane], (unsigned long long)ax);
    atomicAdd((unsigned long long*)&acc[64u + lane], (unsigned long long)ay);
    atomicAdd((unsigned long long*)&acc[128u + lane], (unsigned long long)n);
    __threadfence();
    uint32_t prev = 0;
    if (lane == 0) prev = atomicAdd(&split.heavy[h].done, 1u);
    prev = uni(prev);
    if (prev + 1u == parts) {
      __threadfence();
      const long long sx = (long long)__hip_atomic_load((unsigned long long*)&acc[lane],
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const long long sy = (long long)__hip_atomic_load((unsigned long long*)&acc[64u + lane],
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const long long sn = (long long)__hip_atomic_load((unsigned long long*)&acc[128u + lane],
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      nbr += nm_finish(T, sx, sy, sn);
    }
  }
  __shared__ unsigned long long work[C::WAVES][2];
  nm_store_work<C>(split.wg_hsteps, split.wg_hnbr, steps, nbr, work);
}

// ------------------------------------------------------------------- blur
// blur.wgsl:13-65 generalized to any sigma: row pass from the u32 counts.
// A TW x TH tile (TW * TH = 256 threads, one output each) plus a taps-wide
// halo on each side sits in LDS; zero padding outside the image;
// v = t[c]*w0, then v += (t[c-k] + t[c+k]) * w[k]. The tile shape is a
// template parameter the autotuner picks per iteration.
template <int TW, int TH>
__global__ __launch_bounds__(256) void bt_blur_rows(const BtJob* __restrict__ jobs,
                                                    const uint32_t* __restrict__ base, int njobs) {
  static_assert(TW * TH == 256, "one output per thread");
  JOB_PROLOGUE(BT_K_ROW);
  __shared__ float tile[TH][TW + 2 * BT_MAX_TAPS];
  const uint32_t W = J.dw, H = J.dh, K = J.taps;
  const uint32_t tiles_x = (W + TW - 1u) / TW;
  const uint32_t x0 = (lb % tiles_x) * TW, y0 = (lb / tiles_x) * TH;
  const uint32_t tx = threadIdx.x % TW, ty = threadIdx.x / TW;
  const uint32_t y = y0 + ty;
  const uint32_t span = TW + 2u * K;
  const uint32_t* src = J.dens_cnt + (size_t)y * W;
  for (uint32_t e = tx; e < span; e += TW) {
    const int gx = (int)x0 - (int)K + (int)e;
    float v = 0.0f;
    if (y < H && gx >= 0 && gx < (int)W) v = (float)src[gx];
    tile[ty][e] = v;
  }
  __syncthreads();
  const uint32_t x = x0 + tx;
  if (x < W && y < H) {
    const float* wt = J.wt;
    const uint32_t c = K + tx;
    float v = tile[ty][c] * ld_uniform(wt, 0u);
    for (uint32_t k = 1; k <= K; ++k) v = v + (tile[ty][c - k] + tile[ty][c + k]) * ld_uniform(wt, k);
    J.dens_row[(size_t)y * W + x] = v;
  }
}

// The column pass over a tile of 64 columns in LDS whose row c holds image
// row y0 - K + c: thread (tx, ty) computes rows ty, ty + 4, ... of column
// tx. Its TH / 4 outputs advance together through the taps, so their LDS
// reads overlap; each output still adds its taps in the order
// v = t[c]*w0, v += (t[c-k] + t[c+k]) * w[k], k = 1..K. Returns the thread's
// largest in-image value (0 if none).
template <int TH, int ROWS, int COLS>
__device__ __forceinline__ float col_pass(float (&tile)[ROWS][COLS], const BtJob& J, uint32_t x,
                                          uint32_t y0, uint32_t tx, uint32_t ty) {
  constexpr uint32_t Q = TH / 4;
  const uint32_t W = J.dw, H = J.dh, K = J.taps;
  const float* wt = J.wt;
  float v[Q];
  const float w0 = ld_uniform(wt, 0u);
#pragma unroll
  for (uint32_t q = 0; q < Q; ++q) v[q] = tile[K + ty + 4u * q][tx] * w0;
  for (uint32_t k = 1; k <= K; ++k) {
    const float wk = ld_uniform(wt, k);
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q) {
      const uint32_t c = K + ty + 4u * q;
      v[q] = v[q] + (tile[c - k][tx] + tile[c + k][tx]) * wk;
    }
  }
  float peak = 0.0f;
#pragma unroll
  for (uint32_t q = 0; q < Q; ++q) {
    const uint32_t y = y0 + ty + 4u * q;
    if (x < W && y < H) {
      J.dens_col[(size_t)y * W + x] = v[q];
      peak = fmaxf(peak, v[q]);
    }
  }
  return wave_max(peak);
}

// blur.wgsl:13-83 column pass plus the peak (atomicMax on the f32 bits of a
// non-negative value, as the reference does). Tiles of 64 columns x TH rows
// (each thread TH / 4 outputs of one column).
template <int TH>
__global__ __launch_bounds__(256) void bt_blur_cols(const BtJob* __restrict__ jobs,
                                                    const uint32_t* __restrict__ base, int njobs) {
  JOB_PROLOGUE(BT_K_COL);
  constexpr uint32_t TW = 64;
  __shared__ float tile[TH + 2 * BT_MAX_TAPS][TW];
  __shared__ float wmax[4];
  const uint32_t W = J.dw, H = J.dh, K = J.taps;
  const uint32_t tiles_x = (W + TW - 1u) / TW;
  const uint32_t x0 = (lb % tiles_x) * TW, y0 = (lb / tiles_x) * TH;
  const uint32_t tx = threadIdx.x & 63u, ty = threadIdx.x >> 6;
  const uint32_t x = x0 + tx;
  const uint32_t span = TH + 2u * K;
  const float* src = J.dens_row;
  // eight independent loads in flight per thread before the LDS stores
  constexpr uint32_t B = 8;
  for (uint32_t e0 = ty; e0 < span; e0 += 4u * B) {
    float v[B];
#pragma unroll
    for (uint32_t b = 0; b < B; ++b) {
      const uint32_t e = e0 + 4u * b;
      const int gy = (int)y0 - (int)K + (int)e;
      v[b] = (e < span && x < W && gy >= 0 && gy < (int)H) ? src[(size_t)gy * W + x] : 0.0f;
    }
#pragma unroll
    for (uint32_t b = 0; b < B; ++b) {
      const uint32_t e = e0 + 4u * b;
      if (e < span) tile[e][tx] = v[b];
    }
  }
  __syncthreads();
  const float peak = col_pass<TH>(tile, J, x, y0, tx, ty);
  if ((threadIdx.x & 63u) == 0) wmax[ty] = peak;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    atomicMax(J.peak, __float_as_uint(m));
  }
}

// Both blur passes in one kernel, for blurs of at most BT_FUSED_TAPS taps:
// the counts of a (TH + 2K) x (64 + 2K) window go to LDS, its rows are
// blurred (each row value computed as bt_blur_rows computes it, so the row
// plane is the same), the tile's own rows are written to the row plane, and
// the column pass runs on the row values still in LDS. Halo rows outside the
// image hold zero counts and so blur to zero, the value the column pass of
// the two-kernel form reads there.
template <int TH>
__global__ __launch_bounds__(256) void bt_blur_fused(const BtJob* __restrict__ jobs,
                                                     const uint32_t* __restrict__ base, int njobs) {
  JOB_PROLOGUE(BT_K_COL);
  constexpr uint32_t TW = 64, KMAX = BT_FUSED_TAPS;
  constexpr uint32_t RMAX = TH + 2u * KMAX, PER = RMAX / 4u;
  __shared__ float tile[RMAX][TW + 2u * KMAX];
  __shared__ float wmax[4];
  const uint32_t W = J.dw, H = J.dh, K = J.taps;
  const uint32_t tiles_x = (W + TW - 1u) / TW;
  const uint32_t x0 = (lb % tiles_x) * TW, y0 = (lb / tiles_x) * TH;
  const uint32_t tx = threadIdx.x & 63u, ty = threadIdx.x >> 6;
  const uint32_t R = TH + 2u * K, S = TW + 2u * K;
  const uint32_t* src = J.dens_cnt;
  // four rows x two column halves: eight independent loads in flight per
  // thread before the LDS stores
  for (uint32_t r0 = ty; r0 < R; r0 += 16u) {
    float v[4][2];
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
      const uint32_t r = r0 + 4u * b;
      const int gy = (int)y0 - (int)K + (int)r;
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t c = tx + h * TW;
        const int gx = (int)x0 - (int)K + (int)c;
        v[b][h] = (r < R && c < S && gy >= 0 && gy < (int)H && gx >= 0 && gx < (int)W)
                      ? (float)src[(size_t)gy * W + gx] : 0.0f;
      }
    }
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
      const uint32_t r = r0 + 4u * b;
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t c = tx + h * TW;
        if (r < R && c < S) tile[r][c] = v[b][h];
      }
    }
  }
  __syncthreads();
  // row pass over rows ty, ty + 4, ... of the window, four rows at a time
  // advancing together through the taps (wave-uniform guard per group)
  const float* wt = J.wt;
  const uint32_t x = x0 + tx;
  const float w0 = ld_uniform(wt, 0u);
  float rv[PER];
#pragma unroll
  for (uint32_t q0 = 0; q0 < PER; q0 += 4u) {
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) rv[q0 + b] = 0.0f;
    if (ty + 4u * q0 < R) {
      const uint32_t c = K + tx;
      float v[4];
#pragma unroll
      for (uint32_t b = 0; b < 4; ++b) v[b] = tile[ty + 4u * (q0 + b)][c] * w0;
      for (uint32_t k = 1; k <= K; ++k) {
        const float wk = ld_uniform(wt, k);
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
          const uint32_t r = ty + 4u * (q0 + b);
          v[b] = v[b] + (tile[r][c - k] + tile[r][c + k]) * wk;
        }
      }
#pragma unroll
      for (uint32_t b = 0; b < 4; ++b) {
        const uint32_t r = ty + 4u * (q0 + b);
        rv[q0 + b] = v[b];
        const uint32_t y = y0 + r - K;
        if (r >= K && r < K + TH && x < W && y < H) J.dens_row[(size_t)y * W + x] = v[b];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t q = 0; q < PER; ++q) {
    const uint32_t r = ty + 4u * q;
    if (r < R) tile[r][tx] = rv[q];
  }
  __syncthreads();
  const float peak = col_pass<TH>(tile, J, x, y0, tx, ty);
  if ((threadIdx.x & 63u) == 0) wmax[ty] = peak;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    atomicMax(J.peak, __float_as_uint(m));
  }
}

// normalization.wgsl:5-17, f32 R channel.
extern "C" __global__ __launch_bounds__(256) void bt_normalize(const BtJob* __restrict__ jobs,
                                                               const uint32_t* __restrict__ base,
                                                               int njobs) {
  JOB_PROLOGUE(BT_K_NORM);
  const uint32_t npix = J.dw * J.dh;
  const float peak = __uint_as_float(*(volatile uint32_t*)J.peak);
  const uint32_t i0 = lb * BT_NORM_ITEMS + threadIdx.x;
#pragma unroll
  for (uint32_t k = 0; k < BT_NORM_ITEMS / 256u; ++k) {
    const uint32_t i = i0 + k * 256u;
    if (i < npix) {
      const float v = J.dens_col[i];
      J.dens_norm[i] = peak > 0.0f ? fminf(fmaxf(v / peak, 0.0f), 1.0f) : 0.0f;
    }
  }
}

// ------------------------------------------------- density on read
// The density of the current records outside an iteration
// (BT_FLAG_DENSITY_ON_READ): counts and peak cleared, every sample splatted
// as the neighbor-mean kernel splats its output (neighbor_mean.wgsl:66-73),
// then the usual blur and normalize. Integer adds commute, so the counts
// equal those an iteration writes.
extern "C" __global__ __launch_bounds__(256) void bt_density_clear(const BtJob* __restrict__ jobs,
                                                                   const uint32_t* __restrict__ base,
                                                                   int njobs) {
  JOB_PROLOGUE(BT_K_NORM);
  const uint32_t npix = J.dw * J.dh;
  const uint32_t i0 = lb * BT_NORM_ITEMS + threadIdx.x;
#pragma unroll
  for (uint32_t k = 0; k < BT_NORM_ITEMS / 256u; ++k) {
    const uint32_t i = i0 + k * 256u;
    if (i < npix) J.dens_cnt[i] = 0u;
  }
  if (lb == 0 && threadIdx.x == 0 && J.peak) J.peak[0] = 0u;
}

extern "C" __global__ __launch_bounds__(256) void bt_splat_points(const BtJob* __restrict__ jobs,
                                                                  const uint32_t* __restrict__ base,
                                                                  int njobs) {
  JOB_PROLOGUE(BT_K_BIN);
  const uint32_t i = lb * BT_BIN_BLOCK + threadIdx.x;
  const bool valid = i < J.n;
  const float* rec = J.rec[J.parity];
  const uint32_t pix = valid ? splat_index(rec[3u * i], rec[3u * i + 1u], J.dw, J.dh) : 0u;
  // records are in cell order: lanes of a wave mostly share pixels
  uint64_t act = ballot(valid);
  while (act) {
    const int leader = __builtin_ctzll(act);
    const uint32_t key = __builtin_amdgcn_readlane(pix, leader);
    const uint64_t same = ballot(valid && pix == key);
    if (LANE() == (uint32_t)leader) atomicAdd(&J.dens_cnt[key], (uint32_t)__popcll(same));
    act &= ~same;
  }
}

// ------------------------------------------------------------- resampling
// resampling.wgsl:45-48
__device__ __forceinline__ void normalized(const BtJob& J, uint32_t i, float& x, float& y) {
  const float* c = J.ctrl;
  x = ((c[2u * i] - J.norm_min_x) * J.norm_scale + J.norm_tr_x - 0.5f) * 2.0f;
  y = ((c[2u * i + 1u] - J.norm_min_y) * J.norm_scale + J.norm_tr_y - 0.5f) * 2.0f;
}

__device__ __forceinline__ uint32_t steps_of(float distance, float remaining) {
  float f = fmaxf(floorf(distance - remaining) + 1.0f, 0.0f);
  return f >= 4294967296.0f ? 0xffffffffu : (uint32_t)f;
}


// resampling.wgsl:50-136 write mode + emit_paths (:161-183): one wave per
// path. Within a segment the samples are independent (sample k sits at
// previous + delta * ((remaining + k) / distance)); lanes produce 64 at a
// time and the duplicate rule (a sample equal to the previous one is
// dropped, :51-53) is applied against the neighbouring lane's sample.
// Writes records {x, y, slot | fixed}, the path id of each sample, the
// Reservation triple of the path, and splats every written sample.
extern "C" __global__ __launch_bounds__(256) void bt_resample_emit(
    const BtJob* __restrict__ jobs, const uint32_t* __restrict__ base, int njobs) {
  JOB_PROLOGUE(BT_K_EMIT);
  const uint32_t lane = LANE();
  const uint32_t p = lb * BT_EMIT_PATHS + (threadIdx.x >> 6);
  if (p >= J.n_paths) return;
  const unsigned long long sc = J.path_scan[p];
  const unsigned long long nx_scan = J.path_scan[p + 1u];
  const uint32_t count = (uint32_t)((nx_scan >> 32) - (sc >> 32));
  const uint32_t off = (uint32_t)(sc >> 32);
  const uint32_t valid_before = (uint32_t)(sc & 0xffffffffu);
  if (lane == 0) {
    J.paths[3u * p] = count ? off : 0u;
    J.paths[3u * p + 1u] = count ? off - valid_before : 0u;
    J.paths[3u * p + 2u] = count;
  }
  if (count < 2u) return;
  float* rec = J.rec[0];
  uint32_t* edge_of = J.edge_of;
  uint32_t* dens = J.dens_cnt;
  const uint32_t W = J.dw, H = J.dh;
  bool bad = false;
  uint32_t emitted = 0;
  float lx, ly;
  // Writes one uniform sample (lane 0), applying the duplicate rule.
  auto push1 = [&](float x, float y) {
    if (emitted > 0u && x == lx && y == ly) return;
    if (emitted >= count) { bad = true; return; }
    if (lane == 0) {
      const uint32_t slot = off + emitted;
      const uint32_t fixed = (emitted == 0u || emitted == count - 1u) ? BT_FIXED_BIT : 0u;
      rec[3u * slot] = x;
      rec[3u * slot + 1u] = y;
      rec[3u * slot + 2u] = __uint_as_float(slot | fixed);
      edge_of[slot] = p;
      if (dens) atomicAdd(&dens[splat_index(x, y, W, H)], 1u);
    }
    lx = x;
    ly = y;
    ++emitted;
  };
  const uint32_t s = J.starts[p], e = J.starts[p + 1u];
  float fx, fy;
  normalized(J, s, fx, fy);
  lx = fx;
  ly = fy;
  push1(fx, fy);
  const float tsx = J.ts_x, tsy = J.ts_y;
  float px = fx, py = fy, remaining = 1.0f;
  for (uint32_t k = s + 1u; k < e && !bad; ++k) {
    float cx, cy;
    normalized(J, k, cx, cy);
    const float dx = cx - px, dy = cy - py;
    const float ex = dx * tsx, ey = dy * tsy;
    const float distance = sqrtf(fmaf(ex, ex, ey * ey));
    if (distance > 0.0f) {
      const uint32_t steps = steps_of(distance, remaining);
      for (uint32_t b0 = 0; b0 < steps && !bad; b0 += 64u) {
        const uint32_t st = b0 + lane;
        const bool active = st < steps;
        const float t = (remaining + (float)st) / distance;
        const float sx = px + dx * t, sy = py + dy * t;
        float qx = __shfl_up(sx, 1, 64), qy = __shfl_up(sy, 1, 64);
        if (lane == 0) { qx = lx; qy = ly; }
        const bool keep = active && !(sx == qx && sy == qy);
        const uint64_t km = ballot(keep);
        const uint32_t pos = emitted + mbcnt64(km);
        if (keep) {
          if (pos >= count) {
            bad = true;
          } else {
            const uint32_t slot = off + pos;
            const uint32_t fixed = (pos == 0u || pos == count - 1u) ? BT_FIXED_BIT : 0u;
            rec[3u * slot] = sx;
            rec[3u * slot + 1u] = sy;
            rec[3u * slot + 2u] = __uint_as_float(slot | fixed);
            edge_of[slot] = p;
            if (dens) atomicAdd(&dens[splat_index(sx, sy, W, H)], 1u);
          }
        }
        bad = ballot(bad) != 0ull;
        emitted += (uint32_t)__popcll(km);
        const uint64_t am = ballot(active);
        const int last = 63 - __builtin_clzll(am);
        lx = __shfl(sx, last, 64);
        ly = __shfl(sy, last, 64);
      }
      remaining = remaining + (float)steps - distance;
    }
    px = cx;
    py = cy;
  }
  if (!bad) push1(px, py);
  if (!bad && emitted == 1u) {
    for (uint32_t k = s + 1u; k < e; ++k) {
      float nx, ny;
      normalized(J, k, nx, ny);
      if (nx != fx || ny != fy) {
        push1(nx, ny);
        if (!bad) push1(px, py);
        break;
      }
    }
  }
  if ((bad || emitted != count) && lane == 0) atomicOr(&J.status[0], 2u);
}

// ------------------------------------------------------------ readbacks
// Undo the sorted order: record at slot -> original index, with the path id
// (the reference's Point {x, y, edge_id}, graph.wgsl:2-6).
extern "C" __global__ __launch_bounds__(256) void bt_unpermute(const BtJob* __restrict__ jobs,
                                                               const uint32_t* __restrict__ base,
                                                               int njobs, float* __restrict__ out,
                                                               uint32_t* __restrict__ cell_out,
                                                               uint32_t want_cells,
                                                               uint32_t split) {
  JOB_PROLOGUE(BT_K_BIN);
  const uint32_t i = lb * BT_BIN_BLOCK + threadIdx.x;
  if (i >= J.n) return;
  const float* rec = J.rec[J.parity];
  const uint32_t id = __float_as_uint(rec[3u * i + 2u]);
  const uint32_t orig = id & ~BT_FIXED_BIT;
  if (split) {
    // n positions (x, y), then n edge ids
    reinterpret_cast<float2*>(out)[orig] = make_float2(rec[3u * i], rec[3u * i + 1u]);
    reinterpret_cast<uint32_t*>(out)[2u * J.n + orig] = J.edge_of[orig];
  } else {
    out[3u * orig] = rec[3u * i];
    out[3u * orig + 1u] = rec[3u * i + 1u];
    out[3u * orig + 2u] = __uint_as_float(J.edge_of[orig]);
  }
  if (want_cells) {
    // cell of the sample as binned in the last iteration: sid/sorted are in
    // the order this iteration's input was binned from (rec of the other parity)
    const float* prev = J.rec[J.parity ^ 1u];
    const uint32_t pid = __float_as_uint(prev[3u * i + 2u]) & ~BT_FIXED_BIT;
    cell_out[pid] = J.cellrank[i].x;
  }
}

// The normalized density as the reference's Rgba16Float texture holds its R
// channel (density.rs:96, written by textureStore in normalization.wgsl:16):
// binary16 bits, rounded to nearest even (v_cvt_f16_f32 under the default
// round-to-nearest mode, f16 denormals kept). Two pixels per thread.
extern "C" __global__ __launch_bounds__(256) void bt_half_plane(const float* __restrict__ in,
                                                                uint16_t* __restrict__ out,
                                                                uint32_t n) {
  const uint32_t i = 2u * (blockIdx.x * 256u + threadIdx.x);
  if (i + 1u < n) {
    const float2 v = *reinterpret_cast<const float2*>(in + i);
    const _Float16 a = (_Float16)v.x, b = (_Float16)v.y;
    *reinterpret_cast<uint32_t*>(out + i) =
        (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
  } else if (i < n) {
    out[i] = __builtin_bit_cast(uint16_t, (_Float16)in[i]);
  }
}

// --------------------------------------------------- upload / rewind set-up
// Per-job set-up of an upload from the packed upload area (one H2D copy for
// all jobs): each job's blur weights go to its own buffer and its status
// words are cleared. area[2j], area[2j + 1]: word offset and length of job
// j's weights. One workgroup per job.
extern "C" __global__ __launch_bounds__(256) void bt_upload_setup(const BtJob* __restrict__ jobs,
                                                                  int njobs,
                                                                  const uint32_t* __restrict__ area) {
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const BtJob& J = jobs[j];
  const uint32_t off = area[2 * j], len = area[2 * j + 1];
  for (uint32_t i = threadIdx.x; i < len; i += 256u) J.wt[i] = __uint_as_float(area[off + i]);
  if (threadIdx.x < 4u) J.status[threadIdx.x] = 0u;
}

// One 64-bit word per job for the host: the job's first status word.
extern "C" __global__ __launch_bounds__(256) void bt_job_words(const BtJob* __restrict__ jobs,
                                                               int njobs,
                                                               unsigned long long* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= njobs) return;
  out[j] = (unsigned long long)jobs[j].status[0];
}

// The samples after upload: rec[0] -> rec_saved (restore = 0) or back.
extern "C" __global__ __launch_bounds__(256) void bt_save_points(const BtJob* __restrict__ jobs,
                                                                 const uint32_t* __restrict__ base,
                                                                 int njobs, int restore) {
  JOB_PROLOGUE(BT_K_BIN);
  const uint32_t i = lb * BT_BIN_BLOCK + threadIdx.x;
  if (i >= J.n || !J.rec_saved) return;
  float* live = J.rec[0] + 3u * i;
  float* saved = J.rec_saved + 3u * i;
  const float* src = restore ? saved : live;
  float* dst = restore ? live : saved;
  dst[0] = src[0];
  dst[1] = src[1];
  dst[2] = src[2];
}

// The four density planes after upload: counts, row, column and normalized
// planes -> dens_saved (restore = 0) or back.
extern "C" __global__ __launch_bounds__(256) void bt_save_density(const BtJob* __restrict__ jobs,
                                                                  const uint32_t* __restrict__ base,
                                                                  int njobs, int restore) {
  JOB_PROLOGUE(BT_K_NORM);
  if (!J.dens_saved) return;
  const uint32_t npix = J.dw * J.dh;
  uint32_t* planes[4] = {J.dens_cnt, reinterpret_cast<uint32_t*>(J.dens_row),
                         reinterpret_cast<uint32_t*>(J.dens_col),
                         reinterpret_cast<uint32_t*>(J.dens_norm)};
  uint32_t* saved = reinterpret_cast<uint32_t*>(J.dens_saved);
  const uint32_t i0 = lb * BT_NORM_ITEMS + threadIdx.x;
#pragma unroll
  for (uint32_t k = 0; k < BT_NORM_ITEMS / 256u; ++k) {
    const uint32_t i = i0 + k * 256u;
    if (i >= npix) continue;
#pragma unroll
    for (uint32_t q = 0; q < 4u; ++q) {
      uint32_t* live = planes[q] + i;
      uint32_t* copy = saved + (size_t)q * npix + i;
      if (restore)
        *live = *copy;
      else
        *copy = *live;
    }
  }
}

// ------------------------------------------------------------- endpoints
// endpoints.wgsl:9-49 (deduplicate, for --endpoint-dots): the distinct
// positions among the first and last samples of the paths with two or more
// samples. One thread per path inserts its two endpoints into an
// open-addressing hash table (linear probing, load <= 1/2) keyed by the
// position (-0 and 0 hash alike; positions compare with ==, as the
// reference). A slot holds sample index + 1; a position already present
// keeps the smallest index (atomicMin), so the table's content does not
// depend on scheduling. `pts` holds the current records in Layout order
// (endpoints never move).
__device__ __forceinline__ void endpoint_insert(const float* pts, uint32_t* table, uint32_t mask,
                                                uint32_t index) {
  const float x = pts[3u * index], y = pts[3u * index + 1u];
  const uint32_t bx = __float_as_uint(x == 0.0f ? 0.0f : x);
  const uint32_t by = __float_as_uint(y == 0.0f ? 0.0f : y);
  // 32-bit finalizer mix (MurmurHash3 fmix32) of the combined bits
  uint32_t h = bx * 0x85ebca6bu + by;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  uint32_t slot = h & mask;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    uint32_t old = __hip_atomic_load(&table[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0u) {
      old = atomicCAS(&table[slot], 0u, index + 1u);
      if (old == 0u) return;  // inserted
    }
    const float ox = pts[3u * (old - 1u)], oy = pts[3u * (old - 1u) + 1u];
    if (ox == x && oy == y) {
      atomicMin(&table[slot], index + 1u);
      return;
    }
    slot = (slot + 1u) & mask;
  }
}

extern "C" __global__ __launch_bounds__(256) void bt_endpoints(const BtJob* __restrict__ jobs,
                                                               const uint32_t* __restrict__ base,
                                                               int njobs, const float* pts,
                                                               uint32_t* table, uint32_t mask) {
  JOB_PROLOGUE(BT_K_PATH);
  const uint32_t i = lb * 256u + threadIdx.x;
  if (i >= J.n_paths) return;
  const uint32_t off = J.paths[3u * i], count = J.paths[3u * i + 2u];
  if (count < 2u) return;
  endpoint_insert(pts, table, mask, off);
  endpoint_insert(pts, table, mask, off + count - 1u);
}

// ------------------------------------------------------- launch wrappers
// Host-callable entry points used by the C++ runtime (bt_engine.cpp).
#include "bt_launch.h"

namespace btk {
#define L256(grid) dim3(grid), dim3(256), 0, s
void clear(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
           BtControl* ctl, int mode) {
  hipLaunchKernelGGL(bt_clear, L256(grid), jobs, base, nj, ctl, mode);
}
void bin(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj) {
  hipLaunchKernelGGL(bt_bin, L256(grid), jobs, base, nj);
}
void cell_scan(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj) {
  hipLaunchKernelGGL(bt_cell_scan_reduce, L256(grid), jobs, base, nj);
  hipLaunchKernelGGL(bt_cell_scan_apply, L256(grid), jobs, base, nj);
}
void pre_scan(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj) {
  hipLaunchKernelGGL(bt_pre_scan_reduce, L256(grid), jobs, base, nj);
  hipLaunchKernelGGL(bt_pre_scan_apply, dim3(grid), dim3(BT_PRE_APPLY_THREADS), 0, s, jobs, base,
                     nj);
}
void tasks(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
           BtControl* ctl, BtTask* tasks_out) {
  hipLaunchKernelGGL(bt_tasks, L256(grid), jobs, base, nj, ctl, tasks_out);
}
void scatter(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj) {
  hipLaunchKernelGGL(bt_scatter, L256(grid), jobs, base, nj);
}
template <int GL, int W>
static uint32_t nm_blocks_per_cu() {
  int a = 0, b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, bt_neighbor_mean<GL, W>, 64 * W, 0) !=
      hipSuccess)
    a = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, bt_neighbor_mean_heavy<GL, W>, 64 * W,
                                                   0) != hipSuccess)
    b = 1;
  return (uint32_t)std::max(1, std::min(a, b));
}
template <int GL, int W>
static void nm_launch(hipStream_t s, uint32_t grid, const BtJob* jobs, const BtTask* tasks_in,
                      BtControl* ctl, BtSplit split, unsigned long long samples) {
  hipLaunchKernelGGL((bt_neighbor_mean<GL, W>), dim3(grid), dim3(64 * W), 0, s, jobs, tasks_in,
                     ctl, split, samples);
  hipLaunchKernelGGL((bt_neighbor_mean_heavy<GL, W>), dim3(grid), dim3(64 * W), 0, s, jobs,
                     tasks_in, ctl, split, samples);
}
// Kernel shapes (lanes per group, waves per workgroup), by index.
#define BT_NM_SHAPES(X) X(0, 16, 4) X(1, 32, 4) X(2, 16, 2) X(3, 32, 2) X(4, 16, 8) X(5, 32, 8)
uint32_t neighbor_mean_blocks_per_cu(uint32_t shape) {
  switch (shape) {
#define BT_X(i, gl, w) \
  case i:              \
    return nm_blocks_per_cu<gl, w>();
    BT_NM_SHAPES(BT_X)
#undef BT_X
  }
  return 1;
}
uint32_t neighbor_mean_waves(uint32_t shape) {
  switch (shape) {
#define BT_X(i, gl, w) \
  case i:              \
    return w;
    BT_NM_SHAPES(BT_X)
#undef BT_X
  }
  return 4;
}
void neighbor_mean(hipStream_t s, uint32_t grid, const BtJob* jobs, const BtTask* tasks_in,
                   BtControl* ctl, BtSplit split, unsigned long long samples, uint32_t shape) {
  switch (shape) {
#define BT_X(i, gl, w)                                           \
  case i:                                                        \
    nm_launch<gl, w>(s, grid, jobs, tasks_in, ctl, split, samples); \
    return;
    BT_NM_SHAPES(BT_X)
#undef BT_X
  }
}
void row_tile(uint32_t shape, uint32_t* tw, uint32_t* th) {
  static const uint32_t w[BT_ROW_SHAPE_COUNT] = {64, 128, 32, 256};
  *tw = w[shape < BT_ROW_SHAPE_COUNT ? shape : 0];
  *th = 256u / *tw;
}
uint32_t col_tile_h(uint32_t shape) {
  static const uint32_t h[BT_COL_SHAPE_COUNT] = {32, 16, 64, 32, 16};
  return h[shape < BT_COL_SHAPE_COUNT ? shape : 0];
}
bool col_fused(uint32_t shape) { return shape == 3u || shape == 4u; }
void blur_rows(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
               uint32_t shape) {
  switch (shape) {
    case 1: hipLaunchKernelGGL((bt_blur_rows<128, 2>), L256(grid), jobs, base, nj); return;
    case 2: hipLaunchKernelGGL((bt_blur_rows<32, 8>), L256(grid), jobs, base, nj); return;
    case 3: hipLaunchKernelGGL((bt_blur_rows<256, 1>), L256(grid), jobs, base, nj); return;
    default: hipLaunchKernelGGL((bt_blur_rows<64, 4>), L256(grid), jobs, base, nj); return;
  }
}
void blur_cols(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
               uint32_t shape) {
  switch (shape) {
    case 1: hipLaunchKernelGGL((bt_blur_cols<16>), L256(grid), jobs, base, nj); return;
    case 2: hipLaunchKernelGGL((bt_blur_cols<64>), L256(grid), jobs, base, nj); return;
    case 3: hipLaunchKernelGGL((bt_blur_fused<32>), L256(grid), jobs, base, nj); return;
    case 4: hipLaunchKernelGGL((bt_blur_fused<16>), L256(grid), jobs, base, nj); return;
    default: hipLaunchKernelGGL((bt_blur_cols<32>), L256(grid), jobs, base, nj); return;
  }
}
void density_clear(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj) {
  if (grid) hipLaunchKernelGGL(bt_density_clear, L256(grid), jobs, base, nj);
}
void splat_points(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj) {
  if (grid) hipLaunchKernelGGL(bt_splat_points, L256(grid), jobs, base, nj);
}
void normalize(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj) {
  hipLaunchKernelGGL(bt_normalize, L256(grid), jobs, base, nj);
}
void resample_emit(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj) {
  hipLaunchKernelGGL(bt_resample_emit, L256(grid), jobs, base, nj);
}
void unpermute(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
               float* out, uint32_t* cell_out, uint32_t want_cells, uint32_t split) {
  hipLaunchKernelGGL(bt_unpermute, L256(grid), jobs, base, nj, out, cell_out, want_cells, split);
}
void half_plane(hipStream_t s, const float* in, uint16_t* out, uint32_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(bt_half_plane, dim3((n + 511u) / 512u), dim3(256), 0, s, in, out, n);
}
void upload_setup(hipStream_t s, const BtJob* jobs, int nj, const uint32_t* area) {
  hipLaunchKernelGGL(bt_upload_setup, dim3((unsigned)nj), dim3(256), 0, s, jobs, nj, area);
}
void job_words(hipStream_t s, const BtJob* jobs, int nj, unsigned long long* out) {
  hipLaunchKernelGGL(bt_job_words, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0, s, jobs, nj,
                     out);
}
void save_points(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
                 int restore) {
  hipLaunchKernelGGL(bt_save_points, L256(grid), jobs, base, nj, restore);
}
void save_density(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
                  int restore) {
  hipLaunchKernelGGL(bt_save_density, L256(grid), jobs, base, nj, restore);
}
void endpoints(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
               const float* pts, uint32_t* table, uint32_t mask) {
  hipLaunchKernelGGL(bt_endpoints, L256(grid), jobs, base, nj, pts, table, mask);
}
#undef L256
}  // namespace btk
