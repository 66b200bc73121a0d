// Device runtime. Mirrors the reference's Gpu::upload / Gpu::iterate
// (gpu.rs:492-553), Resampler::encode (resampling.rs:186-290),
// NeighborMean::encode (neighbor_mean.rs:200-289) and Density::encode
// (density.rs:155-166), restructured around batched launches over a job table.
#include "bt_engine.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>

#include "bt_launch.h"
#include "bt_noalloc.h"

namespace bt {

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
static inline size_t align_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

static const uint32_t kWt = BT_MAX_TAPS + 1;  // weights per iteration row

// Pinned host slots for plan uploads (the server queues up to 4 ticks ahead,
// each with its plan, besides admissions' plans), and device slots for
// one-job readback plans.
static const uint32_t kPinnedSlots = 16;
static const uint32_t kRetireSlots = 64;
static const size_t kRetireSlotBytes = 2048;

// Layout of the heavy-task split scratch (one per engine).
static const size_t kSplitItemsOff = (size_t)BT_MAX_HEAVY * sizeof(BtHeavy);
static const size_t kSplitAccOff = kSplitItemsOff + (size_t)BT_MAX_ITEMS * sizeof(uint2);
static const size_t kSplitCandOff = kSplitAccOff + (size_t)BT_MAX_HEAVY * 192 * sizeof(long long);
static const size_t kSplitRecOff = kSplitCandOff + (size_t)BT_MAX_WG * sizeof(unsigned long long);
static const size_t kSplitWorkOff = kSplitRecOff + (size_t)BT_MAX_HEAVY * BT_REC_WORDS * 4;
static const size_t kSplitBytes = kSplitWorkOff + 5 * (size_t)BT_MAX_WG * sizeof(unsigned long long);

BtJob Job::dev(uint32_t it, uint32_t parity_at, bool force_density) const {
  BtJob d;
  std::memset(&d, 0, sizeof d);
  d.rec[0] = rec0.as<float>();
  d.rec[1] = rec1.as<float>();
  d.sorted = sorted.as<float4>();
  d.sid = sid.as<uint32_t>();
  d.cellrank = cellrank.as<uint2>();
  d.cells = cells.as<uint32_t>();
  d.bsum = bsum.as<uint32_t>();
  d.pre_x = pre_x.as<uint64_t>();
  d.pre_y = pre_y.as<uint64_t>();
  d.edge_of = edge_of.as<uint32_t>();
  // density stages run for this iteration (kernels test dens_cnt, dw x dh, peak)
  const bool density = dens_cnt.p != nullptr && (force_density || it >= density_from);
  d.dens_cnt = density ? dens_cnt.as<uint32_t>() : nullptr;
  d.dens_row = dens_row.as<float>();
  d.dens_col = dens_col.as<float>();
  d.dens_norm = dens_norm.as<float>();
  d.peak = density ? peak.as<uint32_t>() : nullptr;
  d.status = status.as<uint32_t>();
  d.ctrl = in_ctrl;
  d.starts = in_starts;
  d.path_scan = nullptr;  // set by an upload's plan (the offsets sit in its input area)
  d.paths = paths.as<uint32_t>();
  d.n = n;
  d.n_paths = n_paths;
  d.dw = density ? dw : 0;
  d.dh = density ? dh : 0;
  const bt_iter_params& p = at(it);
  d.gx = p.gx;
  d.gy = p.gy;
  d.n_cells = p.gx * p.gy;
  d.taps = p.taps;
  d.res_x = p.res_x;
  d.res_y = p.res_y;
  d.ts_x = p.res_x * 0.5f;
  d.ts_y = p.res_y * 0.5f;
  d.cell = p.cell;
  d.inv_cell = p.cell > 0.0f ? 1.0f / p.cell : 0.0f;
  d.cell_w = p.cell_w;
  d.inv_cell_w = p.cell_w > 0.0f ? 1.0f / p.cell_w : 0.0f;
  d.task_bins = std::max(p.task_bins, 1u);
  d.task_bx = (p.gx + d.task_bins - 1u) / d.task_bins;
  d.radius = p.radius;
  d.eps = std::max(1e-3f * p.cell, (std::max(p.res_x, p.res_y) + 1.0f) * 4e-6f);
  d.r2 = p.radius * p.radius;
  const float pad_x = (p.res_x + 1.0f) * 2.4e-7f, pad_y = (p.res_y + 1.0f) * 2.4e-7f;
  d.reach_x = p.radius + pad_x * 4.0f;
  d.reach_y = p.radius + pad_y * 4.0f;
  d.norm_min_x = norm[0];
  d.norm_min_y = norm[1];
  d.norm_scale = norm[2];
  d.norm_tr_x = norm[3];
  d.norm_tr_y = norm[4];
  d.parity = parity_at;
  d.capacity = (uint32_t)std::min<uint64_t>(capacity, 0xfffffffeull);
  d.wt = wt.as<float>() ? wt.as<float>() + (size_t)std::min(it, sched - 1) * kWt : nullptr;
  d.rec_saved = rec_saved.as<float>();
  d.dens_saved = dens_saved.as<float>();
  return d;
}

bool Job::set_split(uint32_t e, uint32_t k) {
  splits[e] = k;
  iter_params(o, e, k, &params[e]);
  const uint32_t nc = params[e].gx * params[e].gy;
  if (nc <= max_cells && cells.p) return false;
  max_cells = std::max(max_cells, nc);
  if (max_cells + 1ull > (uint64_t)BT_SCAN_ITEMS * BT_SCAN_ITEMS)
    fail(BT_E_CAPACITY, "neighbor grid exceeds GPU dispatch limits");
  if (cells.p) {
    cells.alloc(((size_t)max_cells + 1) * 4);
    const size_t nb = cdiv(max_cells + 1ull, BT_SCAN_ITEMS);
    if (nb * 8 > bsum.bytes) bsum.alloc(nb * 8);
    return true;
  }
  return false;
}

void Job::set_pool(DevPool* pool) {
  DevBuf* all[] = {&paths, &bsum, &status, &rec0,
                   &rec1, &rec_saved, &sorted, &sid, &cellrank, &cells, &edge_of,
                   &pre_x, &pre_y, &dens_cnt, &dens_row, &dens_col, &dens_norm, &dens_saved,
                   &peak, &wt};
  for (DevBuf* b : all) b->pool = pool;
}

size_t Job::device_bytes() const {
  const DevBuf* all[] = {&paths, &bsum, &status, &rec0,
                         &rec1, &rec_saved, &sorted, &sid, &cellrank, &cells, &edge_of,
                         &pre_x, &pre_y,
                         &dens_cnt, &dens_row, &dens_col, &dens_norm, &dens_saved, &peak, &wt};
  size_t s = 0;
  for (const DevBuf* b : all) s += b->cap;
  return s;
}

uint32_t Job::max_cells_for(const bt_options& o, uint64_t n) {
  uint32_t cells = 1;
  for (uint32_t i = 0; i < schedule_len(o); ++i) {
    bt_iter_params q;
    iter_params(o, i, o.cell_split, &q);
    cells = std::max(cells, q.gx * q.gy);
    if (o.cell_split == 0 && q.radius > 0.0f) {
      iter_params(o, i, density_split(q.radius, n, q.res_x, q.res_y), &q);
      cells = std::max(cells, q.gx * q.gy);
    }
  }
  return cells;
}

size_t Job::footprint(const bt_options& o, uint64_t n, uint32_t n_points, uint32_t n_starts,
                      bool rewindable) {
  const uint32_t sched = schedule_len(o);
  const uint64_t paths = n_starts ? n_starts - 1ull : 0ull;
  const uint64_t cells = max_cells_for(o, n);
  bt_iter_params p0;
  iter_params(o, 0, 1, &p0);
  auto blk = [](uint64_t b) { return b ? DevPool::block_size((size_t)b) : (size_t)0; };
  const uint64_t nb = std::max<uint64_t>(
      std::max<uint64_t>(cdiv(paths + 1, BT_SCAN_ITEMS), cdiv(cells + 1, BT_SCAN_ITEMS)),
      2ull * cdiv(n + 1, BT_SCAN_ITEMS));
  size_t s = blk(12ull * std::max<uint64_t>(paths, 1)) + blk(16) + blk(8 * nb) +
             blk(4ull * sched * kWt);
  s += 3 * blk(12 * n) + blk(16 * n) + blk(4 * n) + blk(8 * n) + blk(4 * n) +
       blk(4 * (cells + 1)) + 2 * blk(8 * (n + 1));
  if (o.flags & BT_FLAG_DENSITY) {
    const uint64_t pb = 4ull * p0.density_w * p0.density_h;
    s += 4 * blk(pb) + blk(8) + (rewindable ? blk(4 * pb) : 0);
  }
  // upload area: control points, starts, weights and path offsets, 64-word aligned
  s += blk(4 * (align_up(2ull * n_points, 64) + align_up(n_starts, 64) +
                align_up((size_t)sched * kWt, 64) + align_up(2ull * n_starts, 64) + 64));
  return s;
}

Engine::Engine(const bt_options& o, uint32_t cu_mask) : opts_(o), pool_(new DevPool) {
  device_ = o.device;
  HIPCHK(hipSetDevice(device_));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device_));
  cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  for (uint32_t k = 0; k < BT_NM_SHAPE_COUNT; ++k)
    nm_per_cu_[k] = btk::neighbor_mean_blocks_per_cu(k);
  noalloc_setup(device_);  // the allocation-free entry point's scratch on this device
  if (cu_mask && cu_mask != 0xffffffffu) {
    // Every group of 32 CUs gets the same pattern, so every XCD keeps CUs
    // (workgroups are dealt to the XCDs round-robin); persistent grids are
    // sized to the CUs the stream may use.
    uint32_t xcds = 0;
    for (uint32_t b = 0; b < 32; ++b)
      if (cu_mask >> b & 1u) xcds |= 1u << (b % BT_XCDS);
    if (xcds != (1u << BT_XCDS) - 1u) fail(BT_E_INVALID, "CU mask leaves an XCD without CUs");
    const uint32_t words = ((uint32_t)cus_ + 31u) / 32u;
    std::vector<uint32_t> m(words, cu_mask);
    HIPCHK(hipExtStreamCreateWithCUMask(&streams_[0], words, m.data()));
    HIPCHK(hipStreamCreateWithFlags(&streams_[1], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&switch_ev_, hipEventDisableTiming));
    stream_cus_[1] = cus_;
    cus_ = std::max(1, (int)(words * (uint32_t)__builtin_popcount(cu_mask)));
  } else {
    HIPCHK(hipStreamCreateWithFlags(&streams_[0], hipStreamNonBlocking));
  }
  stream_cus_[0] = cus_;
  stream_ = streams_[0];
  ctl_.alloc(sizeof(BtControl));
  HIPCHK(hipMemset(ctl_.p, 0, sizeof(BtControl)));
  split_.alloc(kSplitBytes);
  HIPCHK(hipMemset(split_.p, 0, kSplitBytes));
  pinned_.resize(kPinnedSlots);
  for (Pinned& s : pinned_) HIPCHK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
  up_pinned_.resize(4);
  for (Pinned& s : up_pinned_) HIPCHK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
  retire_plans_.alloc(kRetireSlots * kRetireSlotBytes);
}

void Engine::set_wide(bool wide) {
  const int want = wide && streams_[1] ? 1 : 0;
  if (want == active_) return;
  HIPCHK(hipEventRecord(switch_ev_, streams_[active_]));
  HIPCHK(hipStreamWaitEvent(streams_[want], switch_ev_, 0));
  active_ = want;
  stream_ = streams_[active_];
  cus_ = stream_cus_[active_];
}

Engine::~Engine() {
  drop_graphs();
  // the active stream is ordered after everything queued on the other
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (hipStream_t& st : streams_)
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
      st = nullptr;
    }
  stream_ = nullptr;
  if (switch_ev_) (void)hipEventDestroy(switch_ev_);
  for (hipStream_t st : {side_, tasks_st_})
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  for (hipEvent_t e : ov_ev_) (void)hipEventDestroy(e);
  for (Pinned& s : pinned_) {
    if (s.ev) (void)hipEventDestroy(s.ev);
    if (s.p) (void)hipHostFree(s.p);
  }
  for (Pinned& s : up_pinned_) {
    if (s.ev) (void)hipEventDestroy(s.ev);
    if (s.p) (void)hipHostFree(s.p);
  }
  if (rd_pinned_.p) (void)hipHostFree(rd_pinned_.p);
}

void Engine::sync() { HIPCHK(hipStreamSynchronize(stream_)); }

void Engine::split_counts(uint32_t* tasks, uint32_t* parts) {
  uint32_t w[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(w, ctl_.p, sizeof w, hipMemcpyDeviceToHost, stream_));
  sync();
  *tasks = std::min(w[1], (uint32_t)BT_MAX_HEAVY);
  *parts = std::min(w[2], (uint32_t)BT_MAX_ITEMS);
}

void Engine::work_counts(uint64_t* steps, uint64_t* neighbors, uint64_t* candidates,
                         uint64_t* group_steps) {
  const size_t wg = std::min<size_t>(last_nm_grid_, BT_MAX_WG);
  if (wg == 0) {
    *steps = *neighbors = *candidates = *group_steps = 0;
    return;
  }
  std::vector<unsigned long long> h(6 * wg);
  const char* base = split_.as<char>();
  for (int k = 0; k < 5; ++k)
    HIPCHK(hipMemcpyAsync(h.data() + k * wg, base + kSplitWorkOff + (size_t)k * BT_MAX_WG * 8,
                          wg * 8, hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipMemcpyAsync(h.data() + 5 * wg, base + kSplitCandOff, wg * 8, hipMemcpyDeviceToHost,
                        stream_));
  sync();
  uint64_t a[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 6; ++k)
    for (size_t b = 0; b < wg; ++b) a[k] += h[k * wg + b];
  *steps = a[0] + a[2];
  *neighbors = a[1] + a[3];
  *candidates = a[5];
  *group_steps = a[4] + a[2];  // the heavy kernel's parts stream per lane group
}

void Engine::drop_graphs() {
  if (!graphs_.empty() && stream_) (void)hipStreamSynchronize(stream_);
  for (auto& kv : graphs_) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
  }
  graphs_.clear();
  graph_bufs_.clear();
}

Job& Engine::job(uint32_t i) {
  if (i >= jobs_.size()) fail(jobs_.empty() ? BT_E_NO_GRAPH : BT_E_INVALID,
                              jobs_.empty() ? "no uploaded graph" : "job index out of range");
  return *jobs_[i];
}

std::vector<Job*> Engine::all_jobs() {
  std::vector<Job*> v;
  for (auto& j : jobs_) v.push_back(j.get());
  return v;
}

size_t Engine::device_bytes() const {
  size_t s = ctl_.bytes + tasks_.bytes + plans_.bytes + staging_.bytes + staging2_.bytes;
  for (auto& j : jobs_) s += j->device_bytes();
  for (auto& g : graph_bufs_) s += g.second->bytes;
  return s;
}

// Workgroups each job needs for each kernel kind.
static uint32_t kind_blocks(int k, const Job& j, const BtJob& d, uint32_t kshape) {
  const uint64_t npix = (uint64_t)d.dw * d.dh;
  switch (k) {
    case BT_K_CLEAR: return cdiv(std::max<uint64_t>(d.n_cells + 1ull, npix), BT_CLEAR_ITEMS);
    case BT_K_BIN: return cdiv(j.n, BT_BIN_BLOCK);
    case BT_K_RANK: return cdiv(j.n, BT_RANK_BLOCK);
    case BT_K_SCAN: return cdiv(d.n_cells + 1ull, BT_SCAN_ITEMS);
    case BT_K_TASK: return cdiv((uint64_t)d.task_bx * d.gy, BT_TASK_BLOCKS);
    case BT_K_ROW: {
      uint32_t tw, th;
      btk::row_tile((kshape >> 8) & 0xffu, &tw, &th);
      return cdiv(d.dw, tw) * cdiv(d.dh, th);
    }
    case BT_K_COL: return cdiv(d.dw, 64) * cdiv(d.dh, btk::col_tile_h((kshape >> 16) & 0xffu));
    case BT_K_NORM: return cdiv(npix, BT_NORM_ITEMS);
    case BT_K_PATH: return cdiv(j.n_paths + 1ull, 256);
    case BT_K_EMIT: return cdiv(j.n_paths, BT_EMIT_PATHS);
    case BT_K_PSCAN: return cdiv(j.n_paths + 1ull, BT_SCAN_ITEMS);
    case BT_K_PTS: return cdiv(j.n + 1ull, BT_SCAN_ITEMS);
  }
  return 0;
}

Plan Engine::make_plan(const std::vector<Job*>& js, const std::vector<uint32_t>& iters) const {
  std::vector<BtJob> entries(js.size());
  for (size_t i = 0; i < js.size(); ++i) {
    const uint32_t it = iters[i];
    entries[i] = js[i]->dev(it, (js[i]->parity + (it - js[i]->iteration)) & 1u);
  }
  // one kernel shape per launch: the first job's choice for its iteration
  const uint32_t ks = js.empty() ? 0u : js[0]->kshape[js[0]->entry(iters[0])];
  return make_plan_of(js, entries, ks);
}

Plan Engine::make_plan_of(const std::vector<Job*>& js, const std::vector<BtJob>& entries,
                          uint32_t kshape) const {
  Plan p;
  const uint32_t nj = (uint32_t)js.size();
  if (btk::col_fused((kshape >> 16) & 0xffu)) {
    uint32_t taps = 0;
    for (const BtJob& e : entries) taps = std::max(taps, e.taps);
    if (taps > BT_FUSED_TAPS) kshape &= ~(0xffu << 16);  // the two-kernel blur
  }
  p.kshape = kshape;
  p.njobs = nj;
  p.base_off = align_up(nj * sizeof(BtJob), 256);
  p.host.assign(p.base_off + (size_t)BT_K_COUNT * (nj + 1) * 4, 0);
  BtJob* J = reinterpret_cast<BtJob*>(p.host.data());
  uint32_t* B = reinterpret_cast<uint32_t*>(p.host.data() + p.base_off);
  for (uint32_t i = 0; i < nj; ++i) {
    J[i] = entries[i];
    p.samples += entries[i].n;
  }
  for (int k = 0; k < BT_K_COUNT; ++k) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < nj; ++i) {
      B[k * (nj + 1) + i] = (uint32_t)acc;
      acc += kind_blocks(k, *js[i], J[i], kshape);
    }
    if (acc > 0x7fffffffull) fail(BT_E_CAPACITY, "launch exceeds the workgroup limit");
    B[k * (nj + 1) + nj] = (uint32_t)acc;
    p.grid[k] = (uint32_t)acc;
  }
  return p;
}

void Engine::put_plan(const Plan& p, void* dst, hipStream_t st) {
  if (!st) st = stream_;
  Pinned& s = pinned_[pinned_next_];
  pinned_next_ = (pinned_next_ + 1u) % (uint32_t)pinned_.size();
  if (s.pending) {
    HIPCHK(hipEventSynchronize(s.ev));
    s.pending = false;
  }
  if (s.bytes < p.host.size()) {
    if (s.p) HIPCHK(hipHostFree(s.p));
    s.p = nullptr;
    s.bytes = 0;
    const size_t b = align_up(std::max<size_t>(p.host.size(), 64 << 10), 64 << 10);
    HIPCHK(hipHostMalloc(&s.p, b, hipHostMallocDefault));
    s.bytes = b;
  }
  std::memcpy(s.p, p.host.data(), p.host.size());
  HIPCHK(hipMemcpyAsync(dst, s.p, p.host.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(s.ev, st));
  s.pending = true;
}

void Engine::queue_readback(Job& j, void* rec12, float* density, hipEvent_t done) {
  std::vector<Job*> js = {&j};
  Plan p = make_plan(js, {j.iteration});
  if (p.host.size() > kRetireSlotBytes) fail(BT_E_DEVICE, "readback plan exceeds its slot");
  unsigned char* dev = retire_plans_.as<unsigned char>() + (size_t)retire_next_ * kRetireSlotBytes;
  retire_next_ = (retire_next_ + 1u) % kRetireSlots;
  put_plan(p, dev);
  btk::unpermute(stream_, p.grid[BT_K_BIN], reinterpret_cast<const BtJob*>(dev),
                 reinterpret_cast<const uint32_t*>(dev + p.base_off) + BT_K_BIN * 2, 1,
                 j.rec_saved.as<float>(), nullptr, 0u);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(rec12, j.rec_saved.p, (size_t)j.n * 12, hipMemcpyDeviceToHost, stream_));
  size_t at = (size_t)j.n * 12;
  if (density && j.dens_norm.p) {
    HIPCHK(hipMemcpyAsync(density, j.dens_norm.p, (size_t)j.dw * j.dh * 4, hipMemcpyDeviceToHost,
                          stream_));
    at += (size_t)j.dw * j.dh * 4;
  }
  // the upload's status word (an asynchronous upload leaves it unchecked)
  HIPCHK(hipMemcpyAsync(static_cast<unsigned char*>(rec12) + at, j.status.p, 4,
                        hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipEventRecord(done, stream_));
}

void Engine::ensure_tasks(const std::vector<Job*>& js) {
  uint64_t t = 0;
  for (Job* j : js) t += j->n / BT_TASK_POINTS + 1 + std::min<uint64_t>(j->n, j->max_cells);
  if ((size_t)t * sizeof(BtTask) > tasks_.bytes) {
    drop_graphs();
    sync();
    tasks_.reserve((size_t)t * sizeof(BtTask) + (1 << 16));
  }
}

void Engine::launch_iteration(const Plan& p, const unsigned char* dev, bool density,
                              hipEvent_t* ev, const Overlap* ov) {
  const BtJob* J = reinterpret_cast<const BtJob*>(dev);
  const uint32_t nj = p.njobs;
  auto B = [&](int k) {
    return reinterpret_cast<const uint32_t*>(dev + p.base_off) + (size_t)k * (nj + 1);
  };
  BtControl* ctl = ctl_.as<BtControl>();
  auto mark = [&](int i) {
    if (ev) HIPCHK(hipEventRecord(ev[i], stream_));
  };
  // side: this iteration's density stages go to side_ (ov->side, set by the
  // capture for every iteration with density stages); prev: the previous
  // iteration's did, and side_ has cleared this iteration's density counts
  const bool side = ov && ov->side && density && p.grid[BT_K_ROW];
  const bool prev = ov && ov->prev;
  mark(0);
  btk::clear(stream_, p.grid[BT_K_CLEAR], J, B(BT_K_CLEAR), (int)nj, ctl, prev ? 1 : 3);
  mark(1);
  btk::bin(stream_, p.grid[BT_K_RANK], J, B(BT_K_RANK), (int)nj);
  mark(2);
  btk::cell_scan(stream_, p.grid[BT_K_SCAN], J, B(BT_K_SCAN), (int)nj);
  mark(3);
  if (ov) {
    // the task list (it reads only the cell offsets) beside the scatter and
    // the prefix scans
    HIPCHK(hipEventRecord(ov->scan_done, stream_));
    HIPCHK(hipStreamWaitEvent(tasks_st_, ov->scan_done, 0));
    btk::tasks(tasks_st_, p.grid[BT_K_TASK], J, B(BT_K_TASK), (int)nj, ctl, tasks_.as<BtTask>());
    HIPCHK(hipEventRecord(ov->tasks_done, tasks_st_));
  } else {
    btk::tasks(stream_, p.grid[BT_K_TASK], J, B(BT_K_TASK), (int)nj, ctl, tasks_.as<BtTask>());
  }
  mark(4);
  btk::scatter(stream_, p.grid[BT_K_BIN], J, B(BT_K_BIN), (int)nj);
  btk::pre_scan(stream_, p.grid[BT_K_PTS], J, B(BT_K_PTS), (int)nj);
  if (ov) HIPCHK(hipStreamWaitEvent(stream_, ov->tasks_done, 0));
  mark(5);
  // Persistent grid: 4 workgroups (16 waves) per CU, the kernel's occupancy
  // at its register count; a multiple of the XCD count.
  BtSplit split;
  split.heavy = split_.as<BtHeavy>();
  split.items = reinterpret_cast<uint2*>(split_.as<char>() + kSplitItemsOff);
  split.acc = reinterpret_cast<long long*>(split_.as<char>() + kSplitAccOff);
  split.wg_cand = reinterpret_cast<unsigned long long*>(split_.as<char>() + kSplitCandOff);
  split.rec = reinterpret_cast<uint32_t*>(split_.as<char>() + kSplitRecOff);
  unsigned long long* work = reinterpret_cast<unsigned long long*>(split_.as<char>() + kSplitWorkOff);
  split.wg_steps = work;
  split.wg_nbr = work + BT_MAX_WG;
  split.wg_hsteps = work + 2 * BT_MAX_WG;
  split.wg_hnbr = work + 3 * BT_MAX_WG;
  split.wg_gsteps = work + 4 * BT_MAX_WG;
  const uint32_t shape = p.kshape & 0xffu;
  last_nm_grid_ = (uint32_t)cus_ * nm_per_cu_[shape];
  if (prev) HIPCHK(hipStreamWaitEvent(stream_, ov->prev_done, 0));  // the counts are clear
  btk::neighbor_mean(stream_, last_nm_grid_, J, tasks_.as<BtTask>(), ctl, split, p.samples,
                     shape);
  mark(6);
  if (density && p.grid[BT_K_ROW]) {
    hipStream_t ds = stream_;
    if (side) {
      HIPCHK(hipEventRecord(ov->nm_done, stream_));
      HIPCHK(hipStreamWaitEvent(side_, ov->nm_done, 0));
      ds = side_;
    }
    const uint32_t col = (p.kshape >> 16) & 0xffu;
    if (!btk::col_fused(col))
      btk::blur_rows(ds, p.grid[BT_K_ROW], J, B(BT_K_ROW), (int)nj, (p.kshape >> 8) & 0xffu);
    mark(7);
    btk::blur_cols(ds, p.grid[BT_K_COL], J, B(BT_K_COL), (int)nj, col);
    mark(8);
    btk::normalize(ds, p.grid[BT_K_NORM], J, B(BT_K_NORM), (int)nj);
    mark(9);
    if (side) {
      // the next iteration's density counts and peak, from its own job table
      if (ov->next)
        btk::clear(side_, ov->next->grid[BT_K_CLEAR], reinterpret_cast<const BtJob*>(ov->next_dev),
                   reinterpret_cast<const uint32_t*>(ov->next_dev + ov->next->base_off) +
                       (size_t)BT_K_CLEAR * (ov->next->njobs + 1),
                   (int)ov->next->njobs, ctl, 2);
      HIPCHK(hipEventRecord(ov->side_done, side_));
    }
  } else if (ev) {
    mark(7);
    mark(8);
    mark(9);
  }
  HIPCHK(hipGetLastError());
}

void Engine::upload(const std::vector<GraphIn>& gs, std::vector<uint32_t>* samples) {
  sync();
  drop_graphs();
  jobs_.clear();
  ++generation_;
  std::vector<std::unique_ptr<Job>> out;
  upload_jobs(gs, &out);
  jobs_ = std::move(out);
  reset_split_estimate();
  if (samples) {
    samples->clear();
    for (auto& j : jobs_) samples->push_back(j->n);
  }
}

void* Engine::upload_staging(size_t bytes, Pinned** slot) {
  Pinned& s = up_pinned_[up_next_];
  up_next_ = (up_next_ + 1u) % (uint32_t)up_pinned_.size();
  if (s.pending) {  // its previous copy must be done before it is refilled
    HIPCHK(hipEventSynchronize(s.ev));
    s.pending = false;
  }
  if (s.bytes < bytes) {
    if (s.p) HIPCHK(hipHostFree(s.p));
    s.p = nullptr;
    s.bytes = 0;
    const size_t b = align_up(std::max<size_t>(bytes, 1 << 20), 1 << 20);
    HIPCHK(hipHostMalloc(&s.p, b, hipHostMallocDefault));
    s.bytes = b;
  }
  *slot = &s;
  return s.p;
}

void* Engine::readback_staging(size_t bytes) {
  if (rd_pinned_.bytes < bytes) {
    if (rd_pinned_.p) HIPCHK(hipHostFree(rd_pinned_.p));
    rd_pinned_.p = nullptr;
    rd_pinned_.bytes = 0;
    const size_t b = align_up(std::max<size_t>(bytes, 1 << 20), 1 << 20);
    HIPCHK(hipHostMalloc(&rd_pinned_.p, b, hipHostMallocDefault));
    rd_pinned_.bytes = b;
  }
  return rd_pinned_.p;
}

void Engine::upload_jobs(const std::vector<GraphIn>& gs, std::vector<std::unique_ptr<Job>>* out,
                         bool async, bool rewindable) {
  hipStream_t const st = stream_;
  BtControl* const ctl = ctl_.as<BtControl>();
  std::vector<std::unique_ptr<Job>> nj;
  std::vector<Job*> js;
  std::vector<std::vector<float>> wts;
  std::vector<std::vector<uint64_t>> scans;
  for (const GraphIn& g : gs) {
    auto j = std::make_unique<Job>();
    j->set_pool(pool_.get());
    j->o = resolve_options(&g.opts);
    if (!g.xy || !g.starts || g.n_starts < 2 || g.n_points < 2)
      fail(BT_E_INVALID, "graph has no drawable edges");
    for (uint32_t i = 0; i + 1 < g.n_starts; ++i)
      if (g.starts[i] > g.starts[i + 1]) fail(BT_E_INVALID, "path starts are not sorted");
    if (g.starts[g.n_starts - 1] > g.n_points) fail(BT_E_INVALID, "path starts exceed the points");
    float bbox[4] = {3.40282347e38f, 3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (uint32_t i = 0; i < g.n_points; ++i) {
      const float x = g.xy[2 * i], y = g.xy[2 * i + 1];
      if (!std::isfinite(x) || !std::isfinite(y)) fail(BT_E_INVALID, "non-finite TRL coordinate");
      bbox[0] = std::min(bbox[0], x);
      bbox[1] = std::min(bbox[1], y);
      bbox[2] = std::max(bbox[2], x);
      bbox[3] = std::max(bbox[3], y);
    }
    if (!normalization(bbox, j->norm)) fail(BT_E_INVALID, "graph has invalid extent");
    std::string e;
    int rc = capacity_model(j->o, g.n_points, g.n_starts, &j->capacity, &j->reserved, &e);
    if (rc != BT_OK) fail(rc, e);
    j->n_ctrl = g.n_points;
    j->n_paths = g.n_starts - 1;
    j->sched = schedule_len(j->o);
    j->params.resize(j->sched);
    j->splits.resize(j->sched);
    j->kshape.assign(j->sched, 0u);
    std::vector<float> wt((size_t)j->sched * kWt, 0.0f);
    for (uint32_t i = 0; i < j->sched; ++i) {
      iter_params(j->o, i, j->o.cell_split, &j->params[i]);
      j->splits[i] = j->params[i].radius > 0.0f
                         ? (uint32_t)std::lround(j->params[i].radius / j->params[i].cell)
                         : 1u;
      j->max_cells = std::max(j->max_cells, j->params[i].gx * j->params[i].gy);
      blur_weights(j->params[i].sigma, wt.data() + (size_t)i * kWt, BT_MAX_TAPS);
    }
    j->dw = j->params[0].density_w;
    j->dh = j->params[0].density_h;
    if (j->dw < 1 || j->dh < 1) fail(BT_E_INVALID, "density map dimensions must be nonzero");
    if (j->n_paths + 1ull > (uint64_t)BT_SCAN_ITEMS * BT_SCAN_ITEMS)
      fail(BT_E_CAPACITY, "paths exceed GPU dispatch limits");
    // Count pass (count_paths, resampling.wgsl:138-159) on the host, with
    // path offsets from a scan instead of atomics: the sample count is known
    // before anything is launched, so no device round trip is needed.
    std::vector<uint64_t> scan;
    if (g.path_scan) {
      scan.assign(g.path_scan, g.path_scan + g.n_starts);
    } else {
      scan.resize(g.n_starts);
      resample_counts(g.xy, g.starts, j->n_paths, j->norm, j->params[0].res_x * 0.5f,
                      j->params[0].res_y * 0.5f,
                      (uint32_t)std::min<uint64_t>(j->capacity, 0xfffffffeull), scan.data());
    }
    const uint64_t n = scan[j->n_paths] >> 32, valid = scan[j->n_paths] & 0xffffffffull;
    if (n > j->capacity)
      fail(BT_E_CAPACITY, "GPU resampling exceeds the " + std::to_string(j->capacity) +
                              "-sample buffer capacity; increase --max-graph-memory-mib or reduce "
                              "the graph or render dimensions. Device buffer limits also apply.");
    if (valid == 0 || n < 2) fail(BT_E_INVALID, "resampled graph has no drawable edges");
    if (n > 0x7fffffffull) fail(BT_E_CAPACITY, "sample count exceeds 2^31");
    if (n + 1 > (uint64_t)BT_SCAN_ITEMS * BT_SCAN_ITEMS) fail(BT_E_CAPACITY, "too many samples");
    j->n = (uint32_t)n;
    j->valid_paths = (uint32_t)valid;
    if (j->o.cell_split == 0)  // no split given: rows sized for the sample density
      for (uint32_t k = 0; k < j->sched; ++k)
        if (j->params[k].radius > 0.0f)
          j->set_split(k, density_split(j->params[k].radius, n, j->params[k].res_x,
                                        j->params[k].res_y));
    if (j->max_cells + 1ull > (uint64_t)BT_SCAN_ITEMS * BT_SCAN_ITEMS)
      fail(BT_E_CAPACITY, "neighbor grid exceeds GPU dispatch limits");
    j->paths.alloc((size_t)std::max(j->n_paths, 1u) * 12);
    j->status.alloc(16);
    // block sums: paths and cells (one per block), sample prefix sums (x, y per block)
    const size_t nb = std::max<size_t>(
        std::max<size_t>(cdiv(j->n_paths + 1ull, BT_SCAN_ITEMS), cdiv(j->max_cells + 1ull,
                                                                      BT_SCAN_ITEMS)),
        2 * (size_t)cdiv(n + 1, BT_SCAN_ITEMS));
    j->bsum.alloc(nb * 8);
    j->wt.alloc(wt.size() * 4);
    j->rec0.alloc((size_t)n * 12);
    j->rec1.alloc((size_t)n * 12);
    j->rec_saved.alloc((size_t)n * 12);
    j->sorted.alloc((size_t)n * 16);
    j->sid.alloc((size_t)n * 4);
    j->cellrank.alloc((size_t)n * 8);
    j->edge_of.alloc((size_t)n * 4);
    j->cells.alloc(((size_t)j->max_cells + 1) * 4);
    j->pre_x.alloc(((size_t)n + 1) * 8);
    j->pre_y.alloc(((size_t)n + 1) * 8);
    if (j->o.flags & BT_FLAG_DENSITY) {
      const size_t pb = (size_t)j->dw * j->dh * 4;
      j->dens_cnt.alloc(pb);
      j->dens_row.alloc(pb);
      j->dens_col.alloc(pb);
      j->dens_norm.alloc(pb);
      if (rewindable) j->dens_saved.alloc(pb * 4);
      j->peak.alloc(8);
    }
    wts.push_back(std::move(wt));
    scans.push_back(std::move(scan));
    js.push_back(j.get());
    nj.push_back(std::move(j));
  }
  if (js.empty()) fail(BT_E_INVALID, "no graphs");
  // Every job's inputs in one copy through pinned memory. The upload area
  // holds a (weights offset, weights length) word pair per job, then each
  // job's control points, path starts, blur weights and path offsets (u64),
  // 64-word aligned; the set-up kernel moves the weights to the jobs' own
  // buffers and clears their status words.
  const size_t njs = js.size();
  std::vector<size_t> off(4 * njs);
  size_t words = align_up(2 * njs, 64);
  for (size_t i = 0; i < njs; ++i) {
    off[4 * i] = words;
    words += align_up(2ull * gs[i].n_points, 64);
    off[4 * i + 1] = words;
    words += align_up(gs[i].n_starts, 64);
    off[4 * i + 2] = words;
    words += align_up(wts[i].size(), 64);
    off[4 * i + 3] = words;
    words += align_up(2ull * gs[i].n_starts, 64);
  }
  if (words > 0xffffffffull) fail(BT_E_CAPACITY, "upload exceeds 2^32 words");
  Pinned* slot = nullptr;
  uint32_t* host = static_cast<uint32_t*>(upload_staging(words * 4, &slot));
  for (size_t i = 0; i < njs; ++i) {
    host[2 * i] = (uint32_t)off[4 * i + 2];
    host[2 * i + 1] = (uint32_t)wts[i].size();
    std::memcpy(host + off[4 * i], gs[i].xy, (size_t)gs[i].n_points * 8);
    std::memcpy(host + off[4 * i + 1], gs[i].starts, (size_t)gs[i].n_starts * 4);
    std::memcpy(host + off[4 * i + 2], wts[i].data(), wts[i].size() * 4);
    std::memcpy(host + off[4 * i + 3], scans[i].data(), scans[i].size() * 8);
  }
  // Blocks of the pool are reused in stream order: the area, the plan and
  // the status words go back to it on return while the launches below may
  // still be queued, which is safe because every user is on this stream.
  DevBuf area;
  area.pool = pool_.get();
  area.alloc(words * 4);
  HIPCHK(hipMemcpyAsync(area.p, host, words * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(slot->ev, st));
  slot->pending = true;
  std::vector<BtJob> entries;
  for (size_t i = 0; i < njs; ++i) {
    Job& j = *js[i];
    j.in_ctrl = reinterpret_cast<const float*>(area.as<uint32_t>() + off[4 * i]);
    j.in_starts = area.as<uint32_t>() + off[4 * i + 1];
    BtJob d = j.dev(0, 0, true);
    d.path_scan = reinterpret_cast<uint64_t*>(area.as<uint32_t>() + off[4 * i + 3]);
    entries.push_back(d);
    j.in_ctrl = nullptr;  // the upload area is released on return
    j.in_starts = nullptr;
  }
  // Emit the samples (emit_paths, resampling.wgsl:161-183) with the initial
  // density splat, then the density chain (gpu.rs:512-514).
  const Plan p = make_plan_of(js, entries);
  DevBuf table, jw;
  table.pool = pool_.get();
  jw.pool = pool_.get();
  table.alloc(p.host.size());
  put_plan(p, table.p, st);
  {
    const unsigned char* dev = table.as<unsigned char>();
    const BtJob* J = reinterpret_cast<const BtJob*>(dev);
    const uint32_t n2 = p.njobs;
    auto B = [&](int k) {
      return reinterpret_cast<const uint32_t*>(dev + p.base_off) + (size_t)k * (n2 + 1);
    };
    btk::upload_setup(st, J, (int)n2, area.as<uint32_t>());
    btk::clear(st, p.grid[BT_K_CLEAR], J, B(BT_K_CLEAR), (int)n2, ctl, 3);
    btk::resample_emit(st, p.grid[BT_K_EMIT], J, B(BT_K_EMIT), (int)n2);
    if (p.grid[BT_K_ROW]) {
      btk::blur_rows(st, p.grid[BT_K_ROW], J, B(BT_K_ROW), (int)n2);
      btk::blur_cols(st, p.grid[BT_K_COL], J, B(BT_K_COL), (int)n2);
      btk::normalize(st, p.grid[BT_K_NORM], J, B(BT_K_NORM), (int)n2);
    }
    // the state after upload: rewinds return to it (the served form stages
    // its results in rec_saved instead)
    btk::save_points(st, p.grid[BT_K_BIN], J, B(BT_K_BIN), (int)n2, 0);
    if (rewindable && p.grid[BT_K_NORM])
      btk::save_density(st, p.grid[BT_K_NORM], J, B(BT_K_NORM), (int)n2, 0);
    if (!async) {
      jw.alloc(njs * 8);
      btk::job_words(st, J, (int)n2, jw.as<unsigned long long>());
    }
    HIPCHK(hipGetLastError());
  }
  if (!async) {
    // the emit pass must agree with the counts (resampling.rs:296-299); the
    // asynchronous form leaves the status word for its owner to check
    std::vector<uint64_t> status(njs);
    HIPCHK(hipMemcpyAsync(status.data(), jw.p, njs * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (size_t i = 0; i < njs; ++i)
      if (status[i] & 2u) fail(BT_E_MISMATCH, "GPU resampling count and write passes disagree");
  }
  for (auto& j : nj) {
    j->iteration = 0;
    j->parity = 0;
    j->density_at = 0;
    if (j->o.flags & BT_FLAG_DENSITY_ON_READ) j->density_from = 0xffffffffu;
    out->push_back(std::move(j));
  }
}

void Engine::rewind() {
  if (jobs_.empty()) fail(BT_E_NO_GRAPH, "no uploaded graph");
  std::vector<Job*> js = all_jobs();
  // entries with the density planes whatever the density policy: the planes
  // saved at upload come back too
  std::vector<BtJob> entries;
  for (Job* j : js) entries.push_back(j->dev(0, 0, true));
  Plan p = make_plan_of(js, entries);
  DevBuf table;
  table.pool = pool_.get();
  table.alloc(p.host.size());
  put_plan(p, table.p);
  const unsigned char* dev = table.as<unsigned char>();
  const BtJob* J = reinterpret_cast<const BtJob*>(dev);
  auto B = [&](int k) {
    return reinterpret_cast<const uint32_t*>(dev + p.base_off) + (size_t)k * (p.njobs + 1);
  };
  btk::save_points(stream_, p.grid[BT_K_BIN], J, B(BT_K_BIN), (int)p.njobs, 1);
  if (p.grid[BT_K_NORM])
    btk::save_density(stream_, p.grid[BT_K_NORM], J, B(BT_K_NORM), (int)p.njobs, 1);
  HIPCHK(hipGetLastError());
  for (Job* j : js) {
    j->iteration = 0;
    j->parity = 0;
    j->density_at = 0;
  }
  reset_split_estimate();
  sync();
}

void Engine::reset_split_estimate() {
  HIPCHK(hipMemsetAsync(ctl_.as<char>() + offsetof(BtControl, cand_prev), 0,
                        sizeof(unsigned long long), stream_));
}

static void check_iterable(const std::vector<Job*>& js) {
  if (js.empty()) fail(BT_E_NO_GRAPH, "no uploaded graph");
  for (Job* j : js)
    if (!has_radius(j->o)) fail(BT_E_STATE, "neighbor radius is not configured");
}

void Engine::iterate_jobs(const std::vector<Job*>& js, bool wait) {
  check_iterable(js);
  ensure_tasks(js);
  std::vector<uint32_t> it;
  bool density = false;
  for (Job* j : js) {
    it.push_back(j->iteration);
    density |= j->dens_cnt.p != nullptr;
  }
  Plan p = make_plan(js, it);
  const size_t slot = align_up(p.host.size(), 256);
  if (2 * slot > plans_.bytes) {
    sync();
    plans_.reserve(2 * slot + 4096);
  }
  slot_bytes_ = slot;
  eager_slot_ ^= 1u;
  unsigned char* dev = plans_.as<unsigned char>() + eager_slot_ * slot;
  put_plan(p, dev);
  launch_iteration(p, dev, density);
  for (Job* j : js) j->advanced(1);
  if (wait) sync();
}

void Engine::iterate(uint32_t n, bool use_graph) {
  std::vector<Job*> js = all_jobs();
  check_iterable(js);
  if (n == 0) return;
  ensure_tasks(js);
  bool density = false;
  for (Job* j : js) density |= j->dens_cnt.p != nullptr;
  if (!use_graph) {
    for (uint32_t i = 0; i < n; ++i) {
      std::vector<uint32_t> it;
      for (Job* j : js) it.push_back(j->iteration);
      const Plan p = make_plan(js, it);
      const size_t slot = align_up(p.host.size(), 256);
      if (2 * slot > plans_.bytes) {
        sync();
        plans_.reserve(2 * slot + 4096);
      }
      eager_slot_ ^= 1u;
      unsigned char* dev = plans_.as<unsigned char>() + eager_slot_ * slot;
      put_plan(p, dev);
      launch_iteration(p, dev, density);
      for (Job* j : js) j->advanced(1);
    }
    sync();
    return;
  }
  std::vector<uint64_t> key = {generation_, n};
  for (Job* j : js) key.push_back(((uint64_t)j->iteration << 1) | j->parity);
  auto found = graphs_.find(key);
  if (found == graphs_.end()) {
    std::vector<Plan> plans;
    size_t slot = 0;
    for (uint32_t i = 0; i < n; ++i) {
      std::vector<uint32_t> it;
      for (Job* j : js) it.push_back(j->iteration + i);
      plans.push_back(make_plan(js, it));
      slot = std::max(slot, align_up(plans.back().host.size(), 256));
    }
    auto buf = std::make_unique<DevBuf>();
    buf->alloc(slot * n);
    for (uint32_t i = 0; i < n; ++i)
      put_plan(plans[i], buf->as<unsigned char>() + i * slot);
    sync();
    GraphEntry g;
    // each iteration's density stages run on side_, beside the next
    // iteration's binning, scans and task list (which wait for nothing of
    // them); the next neighbor mean waits for them and for side_'s clear of
    // its density counts
    if (!side_) HIPCHK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    if (!tasks_st_) HIPCHK(hipStreamCreateWithFlags(&tasks_st_, hipStreamNonBlocking));
    while (ov_ev_.size() < 4 * (size_t)n) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ov_ev_.push_back(e);
    }
    HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeGlobal));
    // a launch that throws mid-capture must not leave the stream capturing
    struct CaptureGuard {
      hipStream_t s;
      bool open = true;
      ~CaptureGuard() {
        if (!open) return;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
      }
    } guard{stream_};
    bool prev = false;
    for (uint32_t i = 0; i < n; ++i) {
      Overlap ov;
      ov.side = true;
      ov.prev = prev;
      ov.prev_done = i ? ov_ev_[2 * i - 1] : nullptr;
      ov.nm_done = ov_ev_[2 * i];
      ov.side_done = ov_ev_[2 * i + 1];
      ov.scan_done = ov_ev_[2 * n + 2 * i];
      ov.tasks_done = ov_ev_[2 * n + 2 * i + 1];
      ov.next = i + 1 < n ? &plans[i + 1] : nullptr;
      ov.next_dev = buf->as<unsigned char>() + (i + 1) * slot;
      launch_iteration(plans[i], buf->as<unsigned char>() + i * slot, density, nullptr, &ov);
      prev = density && plans[i].grid[BT_K_ROW] && ov.next;
      if (density && plans[i].grid[BT_K_ROW] && !ov.next)
        HIPCHK(hipStreamWaitEvent(stream_, ov.side_done, 0));  // join
    }
    guard.open = false;
    HIPCHK(hipStreamEndCapture(stream_, &g.graph));
    const hipError_t ie = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g.graph);
      HIPCHK(ie);
    }
    g.last_nm_grid = last_nm_grid_;
    found = graphs_.emplace(key, g).first;
    graph_bufs_.emplace(key, std::move(buf));
  }
  HIPCHK(hipGraphLaunch(found->second.exec, stream_));
  last_nm_grid_ = found->second.last_nm_grid;
  for (Job* j : js) j->advanced(n);
  sync();
}

void Engine::profile(uint32_t n, double* stage_ms, uint32_t nst, double* total_ms) {
  std::vector<Job*> js = all_jobs();
  check_iterable(js);
  ensure_tasks(js);
  bool density = false;
  for (Job* j : js) density |= j->dens_cnt.p != nullptr;
  hipEvent_t ev[10];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  double acc[9] = {0};
  double tot = 0.0;
  try {
    for (uint32_t i = 0; i < n; ++i) {
      std::vector<uint32_t> it;
      for (Job* j : js) it.push_back(j->iteration);
      Plan p = make_plan(js, it);
      const size_t slot = align_up(p.host.size(), 256);
      if (2 * slot > plans_.bytes) {
        sync();
        plans_.reserve(2 * slot + 4096);
      }
      eager_slot_ ^= 1u;
      unsigned char* dev = plans_.as<unsigned char>() + eager_slot_ * slot;
      put_plan(p, dev);
      launch_iteration(p, dev, density, ev);
      sync();
      for (int s = 0; s < 9; ++s) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
        acc[s] += ms;
      }
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[9]));
      tot += ms;
      for (Job* j : js) j->advanced(1);
    }
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (uint32_t s = 0; s < nst && s < 9; ++s) stage_ms[s] = acc[s];
  if (total_ms) *total_ms = tot;
}

void Engine::autotune(uint32_t n, const uint32_t* cand, uint32_t ncand, uint32_t* chosen,
                      double* best_ms) {
  std::vector<Job*> js = all_jobs();
  check_iterable(js);
  if (!cand || ncand == 0) fail(BT_E_INVALID, "no autotune candidates");
  uint32_t sched = 1;
  for (Job* j : js) sched = std::max(sched, j->sched);
  n = std::min(n, sched);
  rewind();
  drop_graphs();
  std::vector<std::unique_ptr<DevBuf>> snap;
  for (Job* j : js) {
    snap.push_back(std::make_unique<DevBuf>());
    snap.back()->alloc((size_t)j->n * 12);
  }
  auto cur = [](Job* j) { return j->parity ? j->rec1.p : j->rec0.p; };
  auto restore = [&]() {
    for (size_t k = 0; k < js.size(); ++k)
      HIPCHK(hipMemcpyAsync(cur(js[k]), snap[k]->p, (size_t)js[k]->n * 12,
                            hipMemcpyDeviceToDevice, stream_));
  };
  bool density = false;
  for (Job* j : js) density |= j->dens_cnt.p != nullptr;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto run_one = [&](bool timed) -> double {
    ensure_tasks(js);
    std::vector<uint32_t> it;
    for (Job* j : js) it.push_back(j->iteration);
    Plan p = make_plan(js, it);
    const size_t slot = align_up(p.host.size(), 256);
    if (2 * slot > plans_.bytes) {
      sync();
      plans_.reserve(2 * slot + 4096);
    }
    eager_slot_ ^= 1u;
    unsigned char* dev = plans_.as<unsigned char>() + eager_slot_ * slot;
    put_plan(p, dev);
    HIPCHK(hipEventRecord(e0, stream_));
    launch_iteration(p, dev, density);
    HIPCHK(hipEventRecord(e1, stream_));
    sync();
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)timed;
    return ms;
  };
  try {
    for (uint32_t i = 0; i < n; ++i) {
      for (size_t k = 0; k < js.size(); ++k)
        HIPCHK(hipMemcpyAsync(snap[k]->p, cur(js[k]), (size_t)js[k]->n * 12,
                              hipMemcpyDeviceToDevice, stream_));
      double best = 1e300;
      uint32_t bk = cand[0];
      for (uint32_t c = 0; c < ncand; ++c) {
        if (cand[c] == 0) continue;
        restore();
        for (Job* j : js) j->set_split(j->entry(i), cand[c]);
        const double t = run_one(true);
        if (t < best) {
          best = t;
          bk = cand[c];
        }
      }
      restore();
      for (Job* j : js) j->set_split(j->entry(i), bk);
      (void)run_one(false);
      for (Job* j : js) j->advanced(1);
      if (chosen) chosen[i] = bk;
      if (best_ms) best_ms[i] = best;
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  rewind();
}

void Engine::set_shapes(uint32_t it, uint32_t nm, uint32_t row, uint32_t col) {
  if (nm >= BT_NM_SHAPE_COUNT || row >= BT_ROW_SHAPE_COUNT || col >= BT_COL_SHAPE_COUNT)
    fail(BT_E_INVALID, "unknown kernel shape");
  if (jobs_.empty()) fail(BT_E_NO_GRAPH, "no uploaded graph");
  const uint32_t v = nm | (row << 8) | (col << 16);
  for (auto& j : jobs_) {
    if (it == ~0u)
      std::fill(j->kshape.begin(), j->kshape.end(), v);
    else
      j->kshape[j->entry(it)] = v;
  }
  drop_graphs();  // captured launches hold the old shapes
}

void Engine::autotune_shapes(uint32_t n, uint32_t* shapes, double* stage_ms) {
  std::vector<Job*> js = all_jobs();
  check_iterable(js);
  uint32_t sched = 1;
  for (Job* j : js) sched = std::max(sched, j->sched);
  n = std::min(n, sched);
  rewind();
  drop_graphs();
  std::vector<std::unique_ptr<DevBuf>> snap;
  for (Job* j : js) {
    snap.push_back(std::make_unique<DevBuf>());
    snap.back()->alloc((size_t)j->n * 12);
  }
  auto cur = [](Job* j) { return j->parity ? j->rec1.p : j->rec0.p; };
  auto restore = [&]() {
    for (size_t k = 0; k < js.size(); ++k)
      HIPCHK(hipMemcpyAsync(cur(js[k]), snap[k]->p, (size_t)js[k]->n * 12,
                            hipMemcpyDeviceToDevice, stream_));
  };
  bool density = false;
  for (Job* j : js) density |= j->dens_cnt.p != nullptr;
  hipEvent_t ev[10];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  // one iteration from the current state; the times of the neighbor-mean,
  // row-blur and column-blur stages
  auto run_one = [&](double* t3) {
    ensure_tasks(js);
    std::vector<uint32_t> it;
    for (Job* j : js) it.push_back(j->iteration);
    Plan p = make_plan(js, it);
    const size_t slot = align_up(p.host.size(), 256);
    if (2 * slot > plans_.bytes) {
      sync();
      plans_.reserve(2 * slot + 4096);
    }
    eager_slot_ ^= 1u;
    unsigned char* dev = plans_.as<unsigned char>() + eager_slot_ * slot;
    put_plan(p, dev);
    launch_iteration(p, dev, density, ev);
    sync();
    float a = 0.0f, b = 0.0f, c = 0.0f;
    HIPCHK(hipEventElapsedTime(&a, ev[5], ev[6]));
    HIPCHK(hipEventElapsedTime(&b, ev[6], ev[7]));
    HIPCHK(hipEventElapsedTime(&c, ev[7], ev[8]));
    t3[0] = a;
    t3[1] = b;
    t3[2] = c;
  };
  try {
    for (uint32_t i = 0; i < n; ++i) {
      for (size_t k = 0; k < js.size(); ++k)
        HIPCHK(hipMemcpyAsync(snap[k]->p, cur(js[k]), (size_t)js[k]->n * 12,
                              hipMemcpyDeviceToDevice, stream_));
      uint32_t best[3] = {0, 0, 0};
      double best_t[3] = {1e300, 1e300, 1e300};
      const uint32_t counts[3] = {BT_NM_SHAPE_COUNT, BT_ROW_SHAPE_COUNT, BT_COL_SHAPE_COUNT};
      const uint32_t stages = density ? 3u : 1u;
      for (uint32_t st = 0; st < stages; ++st) {  // one stage at a time, the others fixed
        for (uint32_t c = 0; c < counts[st]; ++c) {
          uint32_t v[3] = {best[0], best[1], best[2]};
          v[st] = c;
          restore();
          for (Job* j : js) j->kshape[j->entry(i)] = v[0] | (v[1] << 8) | (v[2] << 16);
          double t3[3];
          run_one(t3);
          const double t = st == 2 ? t3[1] + t3[2] : t3[st];
          if (t < best_t[st]) {
            best_t[st] = t;
            best[st] = c;
          }
        }
      }
      restore();
      for (Job* j : js) j->kshape[j->entry(i)] = best[0] | (best[1] << 8) | (best[2] << 16);
      double t3[3];
      run_one(t3);
      for (Job* j : js) j->advanced(1);
      for (uint32_t k = 0; k < 3; ++k) {
        if (shapes) shapes[3 * i + k] = best[k];
        if (stage_ms) stage_ms[3 * i + k] = k == 0 ? best_t[0] : density ? t3[k] : 0.0;
      }
    }
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  rewind();
}

void Engine::read_paths(uint32_t job_i, uint32_t* rec3, size_t n) {
  Job& j = job(job_i);
  if (n != j.n_paths) fail(BT_E_INVALID, "path count mismatch");
  HIPCHK(hipMemcpyAsync(rec3, j.paths.p, (size_t)n * 12, hipMemcpyDeviceToHost, stream_));
  sync();
}

void Engine::unpermute_to_staging(Engine& e, Job& j, bool split) {
  e.staging_.reserve((size_t)j.n * 12);
  std::vector<Job*> js = {&j};
  Plan p = e.make_plan(js, {j.iteration});
  DevBuf table;
  table.pool = e.pool_.get();
  table.alloc(p.host.size());
  e.put_plan(p, table.p);
  const unsigned char* dev = table.as<unsigned char>();
  btk::unpermute(e.stream_, p.grid[BT_K_BIN], reinterpret_cast<const BtJob*>(dev),
                 reinterpret_cast<const uint32_t*>(dev + p.base_off) + BT_K_BIN * 2, 1,
                 e.staging_.as<float>(), nullptr, 0u, split ? 1u : 0u);
  HIPCHK(hipGetLastError());
}

void Engine::read_points_of(Engine& e, Job& j, void* rec12, size_t n) {
  if (n != j.n) fail(BT_E_INVALID, "sample count mismatch");
  unpermute_to_staging(e, j, false);
  HIPCHK(hipMemcpyAsync(rec12, e.staging_.p, (size_t)j.n * 12, hipMemcpyDeviceToHost, e.stream_));
  e.sync();
}

void Engine::read_points(uint32_t job_i, void* rec12, size_t n) {
  read_points_of(*this, job(job_i), rec12, n);
}

void Engine::read_positions(uint32_t job_i, float* xy, uint32_t* ids, size_t n) {
  Job& j = job(job_i);
  if (n != j.n) fail(BT_E_INVALID, "sample count mismatch");
  unpermute_to_staging(*this, j, true);
  HIPCHK(hipMemcpyAsync(xy, staging_.p, (size_t)j.n * 8, hipMemcpyDeviceToHost, stream_));
  if (ids)
    HIPCHK(hipMemcpyAsync(ids, staging_.as<unsigned char>() + (size_t)j.n * 8, (size_t)j.n * 4,
                          hipMemcpyDeviceToHost, stream_));
  sync();
}

void Engine::ensure_density(Job& j) {
  if (!j.dens_cnt.p || j.density_at == j.iteration) return;
  // the density of the current records, with the parameters of the
  // iteration that produced them (iteration k ran with entry k - 1)
  const uint32_t e = j.iteration ? j.iteration - 1u : 0u;
  std::vector<Job*> js = {&j};
  Plan p = make_plan_of(js, {j.dev(e, j.parity, true)});
  DevBuf table;
  table.pool = pool_.get();
  table.alloc(p.host.size());
  put_plan(p, table.p);
  const unsigned char* dev = table.as<unsigned char>();
  const BtJob* J = reinterpret_cast<const BtJob*>(dev);
  auto B = [&](int k) { return reinterpret_cast<const uint32_t*>(dev + p.base_off) + k * 2; };
  btk::density_clear(stream_, p.grid[BT_K_NORM], J, B(BT_K_NORM), 1);
  btk::splat_points(stream_, p.grid[BT_K_BIN], J, B(BT_K_BIN), 1);
  btk::blur_rows(stream_, p.grid[BT_K_ROW], J, B(BT_K_ROW), 1);
  btk::blur_cols(stream_, p.grid[BT_K_COL], J, B(BT_K_COL), 1);
  btk::normalize(stream_, p.grid[BT_K_NORM], J, B(BT_K_NORM), 1);
  HIPCHK(hipGetLastError());
  sync();
  j.density_at = j.iteration;
}

void Engine::read_density_of(Engine& e, Job& j, uint32_t kind, void* out, size_t n) {
  if (!j.dens_cnt.p) fail(BT_E_STATE, "density is disabled (color_by != density)");
  if (n != (size_t)j.dw * j.dh) fail(BT_E_INVALID, "pixel count mismatch");
  e.ensure_density(j);
  if (kind == 4) {  // binary16 R channel of the normalized density (the texture's bits)
    e.staging_.reserve(n * 2 + 4);
    btk::half_plane(e.stream_, j.dens_norm.as<float>(), e.staging_.as<uint16_t>(), (uint32_t)n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, e.staging_.p, n * 2, hipMemcpyDeviceToHost, e.stream_));
    e.sync();
    return;
  }
  const DevBuf* src = kind == 0 ? &j.dens_cnt : kind == 1 ? &j.dens_row
                    : kind == 2 ? &j.dens_col : kind == 3 ? &j.dens_norm : nullptr;
  if (!src) fail(BT_E_INVALID, "unknown density kind");
  HIPCHK(hipMemcpyAsync(out, src->p, n * 4, hipMemcpyDeviceToHost, e.stream_));
  e.sync();
}

void Engine::read_density(uint32_t job_i, uint32_t kind, void* out, size_t n) {
  read_density_of(*this, job(job_i), kind, out, n);
}

void Engine::read_grid(uint32_t job_i, uint32_t* cell_of, size_t n, uint32_t* counts,
                       size_t ncells) {
  Job& j = job(job_i);
  if (j.iteration == 0) fail(BT_E_STATE, "no iteration has binned the samples yet");
  if (n != j.n) fail(BT_E_INVALID, "sample count mismatch");
  const bt_iter_params& p = j.at(j.iteration - 1);
  const size_t nc = (size_t)p.gx * p.gy;
  if (ncells != nc) fail(BT_E_INVALID, "cell count mismatch");
  staging_.reserve((size_t)j.n * 12);
  staging2_.reserve((size_t)j.n * 4);
  std::vector<Job*> js = {&j};
  Plan pl = make_plan(js, {j.iteration});
  DevBuf table;
  table.pool = pool_.get();
  table.alloc(pl.host.size());
  put_plan(pl, table.p);
  const unsigned char* dev = table.as<unsigned char>();
  btk::unpermute(stream_, pl.grid[BT_K_BIN], reinterpret_cast<const BtJob*>(dev),
                 reinterpret_cast<const uint32_t*>(dev + pl.base_off) + BT_K_BIN * 2, 1,
                 staging_.as<float>(), staging2_.as<uint32_t>(), 1u);
  HIPCHK(hipGetLastError());
  std::vector<uint32_t> starts(nc + 1);
  HIPCHK(hipMemcpyAsync(cell_of, staging2_.p, (size_t)j.n * 4, hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipMemcpyAsync(starts.data(), j.cells.p, (nc + 1) * 4, hipMemcpyDeviceToHost, stream_));
  sync();
  for (size_t c = 0; c < nc; ++c) counts[c] = starts[c + 1] - starts[c];
}

// resampling.rs:137-149,240-249 and endpoints.wgsl:9-49: the distinct
// endpoint positions of the paths with two or more samples, one sample index
// per position. The reference keeps whichever index wins its insertion race
// and lists them in completion order; here each position keeps its smallest
// index and the list is sorted, so the result is the same on every run.
std::vector<uint32_t> Engine::read_endpoints(uint32_t job_i) {
  Job& j = job(job_i);
  std::vector<uint32_t> out;
  if (j.n_paths == 0 || j.n == 0) return out;
  size_t slots = 64;
  while (slots < (size_t)j.n_paths * 4) slots <<= 1;
  if (slots > 0x80000000ull) fail(BT_E_CAPACITY, "endpoint table exceeds 2^31 slots");
  staging_.reserve((size_t)j.n * 12);
  std::vector<Job*> js = {&j};
  Plan p = make_plan(js, {j.iteration});
  DevBuf plan, table;
  plan.pool = pool_.get();
  table.pool = pool_.get();
  plan.alloc(p.host.size());
  table.alloc(slots * 4);
  put_plan(p, plan.p);
  const unsigned char* dev = plan.as<unsigned char>();
  const BtJob* J = reinterpret_cast<const BtJob*>(dev);
  const uint32_t* base = reinterpret_cast<const uint32_t*>(dev + p.base_off);
  btk::unpermute(stream_, p.grid[BT_K_BIN], J, base + BT_K_BIN * 2, 1, staging_.as<float>(),
                 nullptr, 0u);
  HIPCHK(hipMemsetAsync(table.p, 0, slots * 4, stream_));
  btk::endpoints(stream_, p.grid[BT_K_PATH], J, base + BT_K_PATH * 2, 1, staging_.as<float>(),
                 table.as<uint32_t>(), (uint32_t)(slots - 1));
  HIPCHK(hipGetLastError());
  std::vector<uint32_t> host(slots);
  HIPCHK(hipMemcpyAsync(host.data(), table.p, slots * 4, hipMemcpyDeviceToHost, stream_));
  sync();
  for (uint32_t v : host)
    if (v) out.push_back(v - 1u);
  std::sort(out.begin(), out.end());
  return out;
}

void Engine::read_endpoints_noalloc(uint32_t job_i, uint32_t* indices, size_t capacity,
                                    uint32_t* count) {
  Job& j = job(job_i);
  const float* rec = (j.parity ? j.rec1 : j.rec0).as<float>();
  endpoints_noalloc(stream_, device_, rec, j.n, j.n_paths, indices, capacity, count);
}

}  // namespace bt
