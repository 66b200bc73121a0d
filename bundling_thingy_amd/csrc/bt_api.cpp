// The C ABI (include/bundling.h): argument checks, error capture, and the
// served form's scheduler.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>

#include "bt_common.h"
#include "bt_engine.h"
#include "bt_palette.h"

struct bt_ctx {
  std::unique_ptr<bt::Engine> engine;
  std::string err;
};

namespace {

template <class F>
int guarded(std::string* err, F&& f) {
  try {
    f();
    return BT_OK;
  } catch (const bt::Error& e) {
    if (err) *err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    if (err) *err = "host allocation failed";
    return BT_E_CAPACITY;
  } catch (const std::exception& e) {
    if (err) *err = e.what();
    return BT_E_DEVICE;
  }
}

}  // namespace

extern "C" {

void bt_options_default(bt_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->map_w = 1600;
  o->map_h = 1600;
  o->flags = BT_FLAG_DENSITY;
  o->max_graph_memory_bytes = 1024ull << 20;
}

int bt_create(const bt_options* opts, bt_ctx** out) {
  if (!out) return BT_E_INVALID;
  *out = nullptr;
  auto ctx = std::make_unique<bt_ctx>();
  std::string err;
  int rc = guarded(&err, [&] { ctx->engine = std::make_unique<bt::Engine>(bt::resolve_options(opts)); });
  if (rc != BT_OK) {
    // No context to hold the message: keep it in a static slot.
    static thread_local std::string last;
    last = err;
    return rc;
  }
  *out = ctx.release();
  return BT_OK;
}

void bt_destroy(bt_ctx* ctx) {
  if (!ctx) return;
  try {
    ctx->engine.reset();
  } catch (...) {
  }
  delete ctx;
}

const char* bt_last_error(const bt_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

#define CTX_CHECK()                      \
  if (!ctx || !ctx->engine) return BT_E_INVALID; \
  ctx->err.clear();

int bt_upload(bt_ctx* ctx, const float* xy, uint32_t n_points, const uint32_t* starts,
              uint32_t n_starts, uint32_t* samples) {
  CTX_CHECK();
  return guarded(&ctx->err, [&] {
    std::vector<bt::GraphIn> gs = {{xy, n_points, starts, n_starts, ctx->engine->options()}};
    std::vector<uint32_t> s;
    ctx->engine->upload(gs, &s);
    if (samples) *samples = s[0];
  });
}

int bt_upload_batch(bt_ctx* ctx, const bt_graph_desc* graphs, uint32_t n_graphs,
                    uint32_t* samples) {
  CTX_CHECK();
  if (!graphs || n_graphs == 0) return BT_E_INVALID;
  return guarded(&ctx->err, [&] {
    std::vector<bt::GraphIn> gs;
    for (uint32_t i = 0; i < n_graphs; ++i) {
      const bt_graph_desc& g = graphs[i];
      gs.push_back({g.xy, g.n_points, g.starts, g.n_starts,
                    g.opts ? *g.opts : ctx->engine->options()});
    }
    std::vector<uint32_t> s;
    ctx->engine->upload(gs, &s);
    if (samples) std::copy(s.begin(), s.end(), samples);
  });
}

int bt_iterate(bt_ctx* ctx) {
  CTX_CHECK();
  return guarded(&ctx->err, [&] { ctx->engine->iterate(1, false); });
}

int bt_iterate_n(bt_ctx* ctx, uint32_t n) {
  CTX_CHECK();
  return guarded(&ctx->err, [&] { ctx->engine->iterate(n, true); });
}

int bt_rewind(bt_ctx* ctx) {
  CTX_CHECK();
  return guarded(&ctx->err, [&] { ctx->engine->rewind(); });
}

int bt_read_paths(bt_ctx* ctx, uint32_t* rec3, size_t n_paths) {
  return bt_read_paths_job(ctx, 0, rec3, n_paths);
}

int bt_read_paths_job(bt_ctx* ctx, uint32_t job, uint32_t* rec3, size_t n_paths) {
  CTX_CHECK();
  if (!rec3) return BT_E_INVALID;
  return guarded(&ctx->err, [&] { ctx->engine->read_paths(job, rec3, n_paths); });
}

int bt_read_points(bt_ctx* ctx, void* rec12, size_t n_samples) {
  return bt_read_points_job(ctx, 0, rec12, n_samples);
}

int bt_read_points_job(bt_ctx* ctx, uint32_t job, void* rec12, size_t n_samples) {
  CTX_CHECK();
  if (!rec12) return BT_E_INVALID;
  return guarded(&ctx->err, [&] { ctx->engine->read_points(job, rec12, n_samples); });
}

int bt_read_positions_job(bt_ctx* ctx, uint32_t job, float* xy, uint32_t* edge_ids,
                          size_t n_samples) {
  CTX_CHECK();
  if (!xy) return BT_E_INVALID;
  return guarded(&ctx->err,
                 [&] { ctx->engine->read_positions(job, xy, edge_ids, n_samples); });
}

int bt_read_density(bt_ctx* ctx, uint32_t kind, void* out, size_t n_pixels) {
  return bt_read_density_job(ctx, 0, kind, out, n_pixels);
}

int bt_read_density_job(bt_ctx* ctx, uint32_t job, uint32_t kind, void* out, size_t n_pixels) {
  CTX_CHECK();
  if (!out) return BT_E_INVALID;
  return guarded(&ctx->err, [&] { ctx->engine->read_density(job, kind, out, n_pixels); });
}

int bt_read_grid(bt_ctx* ctx, uint32_t* cell_of, size_t n_samples, uint32_t* counts,
                 size_t n_cells) {
  CTX_CHECK();
  if (!cell_of || !counts) return BT_E_INVALID;
  return guarded(&ctx->err,
                 [&] { ctx->engine->read_grid(0, cell_of, n_samples, counts, n_cells); });
}

int bt_read_endpoints_noalloc(bt_ctx* ctx, uint32_t* indices, size_t capacity, uint32_t* count) {
  CTX_CHECK();
  if (!count || (capacity && !indices)) return BT_E_INVALID;
  *count = 0;
  return guarded(&ctx->err, [&] {
    ctx->engine->read_endpoints_noalloc(0, indices, capacity, count);
    if (capacity < *count) bt::fail(BT_E_CAPACITY, "endpoint buffer too small");
  });
}

int bt_read_endpoints(bt_ctx* ctx, uint32_t* indices, size_t capacity, uint32_t* count) {
  CTX_CHECK();
  if (!count || (capacity && !indices)) return BT_E_INVALID;
  *count = 0;
  return guarded(&ctx->err, [&] {
    const std::vector<uint32_t> e = ctx->engine->read_endpoints(0);
    *count = (uint32_t)e.size();
    std::copy(e.begin(), e.begin() + std::min(capacity, e.size()), indices);
    if (capacity < e.size()) bt::fail(BT_E_CAPACITY, "endpoint buffer too small");
  });
}

int bt_palette(uint32_t palette, float* colors48, float* rgb3) {
  static const float pal[4][16][3] = BT_PALETTE_TABLE;
  static const float ep[4][3] = BT_ENDPOINT_TABLE;
  if (palette > BT_PALETTE_BATLOW) return BT_E_INVALID;
  if (colors48) std::memcpy(colors48, pal[palette], sizeof pal[palette]);
  if (rgb3) std::memcpy(rgb3, ep[palette], sizeof ep[palette]);
  return BT_OK;
}

void bt_render_options_default(bt_render_options* ro) {
  if (!ro) return;
  std::memset(ro, 0, sizeof *ro);
  ro->palette = BT_PALETTE_VIRIDIS;
  ro->color_by = BT_COLOR_DENSITY;
  ro->msaa = 4;
}

int bt_render_size(bt_ctx* ctx, uint32_t* width, uint32_t* height) {
  CTX_CHECK();
  if (!width || !height) return BT_E_INVALID;
  return guarded(&ctx->err, [&] { ctx->engine->render_size(width, height); });
}

int bt_render(bt_ctx* ctx, const bt_render_options* ro, uint8_t* rgba, size_t n_bytes) {
  return bt_render_job(ctx, 0, ro, rgba, n_bytes);
}

int bt_render_job(bt_ctx* ctx, uint32_t job, const bt_render_options* ro, uint8_t* rgba,
                  size_t n_bytes) {
  CTX_CHECK();
  if (!rgba) return BT_E_INVALID;
  bt_render_options d;
  bt_render_options_default(&d);
  return guarded(&ctx->err, [&] { ctx->engine->render(ro ? *ro : d, rgba, n_bytes, job); });
}

int bt_render_size_job(bt_ctx* ctx, uint32_t job, uint32_t* width, uint32_t* height) {
  CTX_CHECK();
  if (!width || !height) return BT_E_INVALID;
  return guarded(&ctx->err, [&] { ctx->engine->render_size(width, height, job); });
}

int bt_job_samples(bt_ctx* ctx, uint32_t job, uint32_t* samples, uint32_t* iteration) {
  CTX_CHECK();
  return guarded(&ctx->err, [&] {
    bt::Job& j = ctx->engine->job(job);
    if (samples) *samples = j.n;
    if (iteration) *iteration = j.iteration;
  });
}

int bt_get_info(bt_ctx* ctx, bt_info* info) {
  CTX_CHECK();
  if (!info) return BT_E_INVALID;
  std::memset(info, 0, sizeof *info);
  info->jobs = (uint32_t)ctx->engine->njobs();
  info->device_bytes = ctx->engine->device_bytes();
  if (info->jobs) {
    bt::Job& j = ctx->engine->job(0);
    info->samples = j.n;
    info->paths = j.n_paths;
    info->iteration = j.iteration;
    info->capacity = j.capacity;
    info->reserved_bytes = j.reserved;
  }
  return guarded(&ctx->err,
                 [&] {
                   ctx->engine->split_counts(&info->split_tasks, &info->split_parts);
                   ctx->engine->work_counts(&info->test_steps, &info->neighbors,
                                            &info->candidates, &info->group_steps);
                 });
}

int bt_iter_params_for(const bt_options* opts, uint32_t iteration, bt_iter_params* out) {
  if (!out) return BT_E_INVALID;
  bt_options o = bt::resolve_options(opts);
  bt::iter_params(o, iteration, o.cell_split, out);
  return BT_OK;
}

int bt_profile(bt_ctx* ctx, uint32_t n, double* stage_ms, uint32_t n_stages, double* total_ms) {
  CTX_CHECK();
  if (!stage_ms && n_stages) return BT_E_INVALID;
  return guarded(&ctx->err, [&] { ctx->engine->profile(n, stage_ms, n_stages, total_ms); });
}

int bt_autotune(bt_ctx* ctx, uint32_t n, const uint32_t* candidates, uint32_t n_candidates,
                uint32_t* chosen, double* best_ms) {
  CTX_CHECK();
  return guarded(&ctx->err, [&] {
    ctx->engine->autotune(n, candidates, n_candidates, chosen, best_ms);
  });
}

int bt_autotune_shapes(bt_ctx* ctx, uint32_t n, uint32_t* shapes, double* stage_ms) {
  CTX_CHECK();
  return guarded(&ctx->err, [&] { ctx->engine->autotune_shapes(n, shapes, stage_ms); });
}

int bt_set_shapes(bt_ctx* ctx, uint32_t iteration, uint32_t nm_shape, uint32_t row_shape,
                  uint32_t col_shape) {
  CTX_CHECK();
  return guarded(&ctx->err,
                 [&] { ctx->engine->set_shapes(iteration, nm_shape, row_shape, col_shape); });
}

int bt_get_shapes(bt_ctx* ctx, uint32_t job, uint32_t iteration, uint32_t* shapes3) {
  CTX_CHECK();
  if (!shapes3) return BT_E_INVALID;
  return guarded(&ctx->err, [&] {
    const bt::Job& j = ctx->engine->job(job);
    const uint32_t v = j.kshape[j.entry(iteration)];
    shapes3[0] = v & 0xffu;
    shapes3[1] = (v >> 8) & 0xffu;
    shapes3[2] = (v >> 16) & 0xffu;
  });
}

int bt_copy_probe(bt_ctx* ctx, uint64_t bytes, double* ms) {
  CTX_CHECK();
  return guarded(&ctx->err, [&] {
    bt::DevBuf a, b;
    a.alloc(bytes);
    b.alloc(bytes);
    HIPCHK(hipMemsetAsync(a.p, 1, bytes, ctx->engine->stream()));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, ctx->engine->stream()));
    HIPCHK(hipMemcpyAsync(b.p, a.p, bytes, hipMemcpyDeviceToDevice, ctx->engine->stream()));
    HIPCHK(hipEventRecord(e1, ctx->engine->stream()));
    ctx->engine->sync();
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
  });
}

int bt_job_iter_params(bt_ctx* ctx, uint32_t job, uint32_t iteration, bt_iter_params* out) {
  CTX_CHECK();
  if (!out) return BT_E_INVALID;
  return guarded(&ctx->err, [&] { *out = ctx->engine->job(job).at(iteration); });
}

int bt_trl_parse(const char* text, size_t len, float** xy, uint32_t* n_points, uint32_t** starts,
                 uint32_t* n_starts, float* bbox, char* err, size_t errlen) {
  if (!text || !xy || !n_points || !starts || !n_starts || !bbox) return BT_E_INVALID;
  std::vector<float> p;
  std::vector<uint32_t> s;
  std::string e;
  int rc = bt::parse_trl(text, len, p, s, bbox, &e);
  if (rc != BT_OK) {
    if (err && errlen) {
      std::strncpy(err, e.c_str(), errlen - 1);
      err[errlen - 1] = 0;
    }
    return rc;
  }
  *xy = (float*)std::malloc(p.size() * sizeof(float));
  *starts = (uint32_t*)std::malloc(s.size() * sizeof(uint32_t));
  if (!*xy || !*starts) {
    std::free(*xy);
    std::free(*starts);
    return BT_E_CAPACITY;
  }
  std::memcpy(*xy, p.data(), p.size() * sizeof(float));
  std::memcpy(*starts, s.data(), s.size() * sizeof(uint32_t));
  *n_points = (uint32_t)(p.size() / 2);
  *n_starts = (uint32_t)s.size();
  return BT_OK;
}

void bt_free(void* p) { std::free(p); }

int bt_normalization(const float* bbox, float* out) {
  if (!bbox || !out) return BT_E_INVALID;
  return bt::normalization(bbox, out) ? BT_OK : BT_E_INVALID;
}

int bt_capacity(const bt_options* opts, uint32_t n_points, uint32_t n_starts, uint64_t* capacity,
                uint64_t* reserved_bytes) {
  if (!capacity || !reserved_bytes) return BT_E_INVALID;
  std::string e;
  return bt::capacity_model(bt::resolve_options(opts), n_points, n_starts, capacity,
                            reserved_bytes, &e);
}

uint32_t bt_blur_weights(float sigma, float* w, uint32_t max_taps) {
  if (!w) return 0;
  return bt::blur_weights(sigma, w, max_taps);
}

int bt_distortion(const float* original_xy, const float* current_xy, const uint32_t* rec3,
                  uint32_t n_paths, uint32_t n_samples, uint32_t map_size, double* out) {
  if (!original_xy || !current_xy || !rec3 || !out) return BT_E_INVALID;
  std::string e;
  return bt::distortion(original_xy, current_xy, rec3, n_paths, n_samples, map_size, out, &e);
}

}  // extern "C"

// ------------------------------------------------------------------ server

namespace {

struct ServedJob {
  uint64_t ticket = 0;
  uint32_t state = 0;  // 0 queued, 1 running, 2 done, 3 failed
  uint32_t lane = 0;   // the engine (stream) that runs it
  std::vector<float> xy;
  std::vector<uint32_t> starts;
  bt_options opts{};
  uint32_t iterations = 0;
  uint64_t estimate = 0;  // device bytes expected before admission
  uint64_t held = 0;      // device bytes its buffers hold while admitted
  uint64_t est_samples = 0;  // samples of the count pass (exact)
  std::vector<uint64_t> path_scan;  // the count pass's path offsets, for the upload
  double last_ms = 0;     // server clock of admission or of the last iteration
  // retirement: the results are copied into a pinned host buffer (`done`
  // marks the end of the copies), which holds them until release: records
  // (12 B each), then the normalized density (dw x dh floats), then the
  // upload's status word
  hipEvent_t done = nullptr;
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  std::unique_ptr<bt::Job> job;
  uint32_t samples = 0, dw = 0, dh = 0;
  double submit_ms = 0, admit_ms = 0, done_ms = 0;
  std::string err;
};

}  // namespace

static const double kStarveMs = 1000.0;       // SHORTEST: a job waiting longer goes first
static const size_t kTicksAhead = 4;          // ticks per lane queued on the GPU at most
static const uint64_t kTickSamples = 4ull << 20;  // SHORTEST: default sample-updates per tick
static const uint64_t kLaneSamples = 1ull << 20;  // LANES: default small-lane job limit

struct bt_server {
  bt_server_options o{};
  // one engine (its own stream and launch state) per lane: BT_POLICY_LANES
  // runs small and large jobs on two streams whose ticks overlap on the GPU
  std::vector<std::unique_ptr<bt::Engine>> engines;
  std::chrono::steady_clock::time_point t0;
  uint64_t next_ticket = 1;
  uint64_t used = 0;   // admitted jobs' footprints (estimate, or held if larger)
  uint64_t held = 0;   // device bytes the admitted jobs' buffers hold
  uint64_t held_peak = 0;
  std::deque<ServedJob*> queue;
  std::vector<ServedJob*> running;
  std::vector<ServedJob*> finishing;          // results being copied out
  std::multimap<size_t, void*> pinned_free;   // pinned result buffers for reuse
  std::vector<std::deque<hipEvent_t>> ticks;  // per lane: ends of its ticks still on the GPU
  std::deque<uint64_t> finished;              // tickets done or failed, not yet taken
  std::map<uint64_t, std::unique_ptr<ServedJob>> all;
  std::string err;
  double now() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
  }
};

namespace {

// The count pass on the host at submission (the exact sample count, and the
// path offsets the upload reuses), then the device bytes of the job's
// buffers as upload_jobs allocates them (pool blocks). Returns 0 when the
// graph cannot be normalized.
uint64_t estimate_bytes(ServedJob& s, uint64_t* samples_out) {
  bt_options o = bt::resolve_options(&s.opts);
  float bbox[4] = {3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f};
  const uint32_t np = (uint32_t)(s.xy.size() / 2);
  for (uint32_t i = 0; i < np; ++i) {
    bbox[0] = std::min(bbox[0], s.xy[2 * i]);
    bbox[1] = std::min(bbox[1], s.xy[2 * i + 1]);
    bbox[2] = std::max(bbox[2], s.xy[2 * i]);
    bbox[3] = std::max(bbox[3], s.xy[2 * i + 1]);
  }
  float norm[5];
  if (!bt::normalization(bbox, norm)) return 0;
  bt_iter_params p;
  bt::iter_params(o, 0, o.cell_split, &p);
  uint64_t cap = 0, reserved = 0;
  std::string err;
  if (bt::capacity_model(o, np, (uint32_t)s.starts.size(), &cap, &reserved, &err) != BT_OK)
    cap = 0xfffffffeull;  // the upload reports the budget error at admission
  const uint32_t n_paths = (uint32_t)s.starts.size() - 1u;
  s.path_scan.resize(s.starts.size());
  const uint64_t total = bt::resample_counts(s.xy.data(), s.starts.data(), n_paths, norm,
                                             p.res_x * 0.5f, p.res_y * 0.5f,
                                             (uint32_t)std::min<uint64_t>(cap, 0xfffffffeull),
                                             s.path_scan.data());
  const uint64_t n = std::max<uint64_t>(total >> 32, 2);
  *samples_out = n;
  return bt::Job::footprint(o, n, np, (uint32_t)s.starts.size(), false);
}

// Finishes retirements whose copies are done (all of them if `block`):
// the job's device state is released; its results stay in the pinned
// buffer until the job is released.
// Returns a job's result buffer to the free list.
void put_pinned(bt_server* srv, ServedJob& s) {
  if (!s.pinned) return;
  srv->pinned_free.emplace(s.pinned_bytes, s.pinned);
  s.pinned = nullptr;
}

// A job whose results could not be queued for copying: it fails alone and
// gives back its device state and its share of the budget.
void retire_failed(bt_server* srv, ServedJob& s, const std::string& msg) {
  if (s.done) (void)hipEventDestroy(s.done);
  s.done = nullptr;
  put_pinned(srv, s);
  s.job.reset();
  srv->used -= std::min(srv->used, s.estimate);
  srv->held -= std::min(srv->held, s.held);
  s.held = 0;
  s.state = 3;
  s.err = msg;
  s.done_ms = srv->now();
  srv->finished.push_back(s.ticket);
}

void reap(bt_server* srv, bool block) {
  // Each job leaves `finishing` before anything can throw: on an error the
  // list holds exactly the jobs not handled yet, so a later release of a
  // handled job cannot leave a dangling entry behind.
  std::vector<ServedJob*> todo;
  todo.swap(srv->finishing);
  for (size_t k = 0; k < todo.size(); ++k) {
    ServedJob* s = todo[k];
    const hipError_t q = block ? hipEventSynchronize(s->done) : hipEventQuery(s->done);
    if (q == hipErrorNotReady) {
      srv->finishing.push_back(s);
      continue;
    }
    if (q != hipSuccess) {
      srv->finishing.insert(srv->finishing.end(), todo.begin() + k, todo.end());
      bt::fail(BT_E_DEVICE, std::string("HIP: ") + hipGetErrorString(q));
    }
    const size_t pb = (size_t)s->samples * 12;
    const unsigned char* pin = (const unsigned char*)s->pinned;
    uint32_t status = 0;
    std::memcpy(&status, pin + pb + (size_t)s->dw * s->dh * 4, 4);
    if (status & 2u) {
      // the emit pass disagreed with the count pass (resampling.rs:296-299)
      s->state = 3;
      s->err = "GPU resampling count and write passes disagree";
      put_pinned(srv, *s);
    } else {
      s->state = 2;  // the results stay in the pinned buffer until release
    }
    (void)hipEventDestroy(s->done);
    s->done = nullptr;
    s->job.reset();
    srv->used -= std::min(srv->used, s->estimate);
    srv->held -= std::min(srv->held, s->held);
    s->held = 0;
    s->done_ms = srv->now();
    srv->finished.push_back(s->ticket);
  }
}

// A pinned host buffer of at least `bytes`: a free one of the same size
// class (powers of two from 1 MiB) when there is one, else a new one.
void* get_pinned(bt_server* srv, size_t bytes, size_t* got) {
  size_t b = 1 << 20;
  while (b < bytes) b <<= 1;
  auto it = srv->pinned_free.find(b);
  if (it != srv->pinned_free.end()) {
    void* p = it->second;
    *got = it->first;
    srv->pinned_free.erase(it);
    return p;
  }
  void* p = nullptr;
  HIPCHK(hipHostMalloc(&p, b, hipHostMallocDefault));
  *got = b;
  return p;
}

}  // namespace

extern "C" {

int bt_server_create(const bt_server_options* opts, bt_server** out) {
  if (!opts || !out) return BT_E_INVALID;
  *out = nullptr;
  auto s = std::make_unique<bt_server>();
  s->o = *opts;
  if (s->o.max_inflight == 0) s->o.max_inflight = 64;
  if (s->o.memory_budget == 0) s->o.memory_budget = 64ull << 30;
  // A CU mask must leave CUs on every XCD: the hardware deals workgroups to
  // the XCDs round-robin, and bit b of each 32-bit group selects a CU of XCD
  // b mod 8, so every residue mod 8 needs a set bit (workgroups dealt to an
  // XCD without CUs would never run).
  for (uint32_t m : {s->o.small_cu_mask, s->o.large_cu_mask}) {
    if (!m) continue;
    uint32_t seen = 0;
    for (uint32_t b = 0; b < 32; ++b)
      if (m >> b & 1u) seen |= 1u << (b % BT_XCDS);
    if (seen != (1u << BT_XCDS) - 1u) return BT_E_INVALID;
  }
  std::string err;
  int rc = guarded(&err, [&] {
    const int lanes = s->o.policy == BT_POLICY_LANES ? 2 : 1;
    for (int l = 0; l < lanes; ++l) {
      const uint32_t mask = lanes == 1 ? 0u : (l == 0 ? s->o.small_cu_mask : s->o.large_cu_mask);
      s->engines.push_back(
          std::make_unique<bt::Engine>(bt::resolve_options(&s->o.defaults), mask));
    }
    s->ticks.resize(lanes);
  });
  if (rc != BT_OK) return rc;
  s->t0 = std::chrono::steady_clock::now();
  *out = s.release();
  return BT_OK;
}

void bt_server_destroy(bt_server* srv) {
  if (!srv) return;
  try {
    for (auto& e : srv->engines) e->sync();
    for (auto& q : srv->ticks)
      for (hipEvent_t ev : q) (void)hipEventDestroy(ev);
    for (ServedJob* s : srv->finishing) {
      if (s->done) (void)hipEventDestroy(s->done);
      if (s->pinned) (void)hipHostFree(s->pinned);
    }
    for (auto& kv : srv->all)
      if (kv.second->state >= 2 && kv.second->pinned) (void)hipHostFree(kv.second->pinned);
    for (auto& kv : srv->pinned_free) (void)hipHostFree(kv.second);
    srv->all.clear();
    srv->engines.clear();
  } catch (...) {
  }
  delete srv;
}

const char* bt_server_last_error(const bt_server* srv) { return srv ? srv->err.c_str() : ""; }

double bt_server_clock_ms(const bt_server* srv) { return srv ? srv->now() : 0.0; }

int bt_server_submit(bt_server* srv, const float* xy, uint32_t n_points, const uint32_t* starts,
                     uint32_t n_starts, const bt_options* opts, uint32_t iterations,
                     uint64_t* ticket) {
  if (!srv || !xy || !starts || n_starts < 2 || iterations == 0 || !ticket) return BT_E_INVALID;
  // the footprint estimate walks the polylines: check them first (the same
  // rules as an upload)
  for (uint32_t i = 0; i + 1 < n_starts; ++i)
    if (starts[i] > starts[i + 1]) {
      srv->err = "path starts are not sorted";
      return BT_E_INVALID;
    }
  if (starts[n_starts - 1] > n_points || n_points < 2) {
    srv->err = "path starts exceed the points";
    return BT_E_INVALID;
  }
  auto s = std::make_unique<ServedJob>();
  s->ticket = srv->next_ticket++;
  s->xy.assign(xy, xy + 2ull * n_points);
  s->starts.assign(starts, starts + n_starts);
  s->opts = bt::resolve_options(opts ? opts : &srv->o.defaults);
  s->iterations = iterations;
  s->submit_ms = srv->now();
  s->estimate = estimate_bytes(*s, &s->est_samples);
  if (s->estimate == 0 || s->estimate > srv->o.memory_budget) {
    srv->err = "job footprint exceeds the server memory budget";
    return BT_E_CAPACITY;
  }
  if (srv->o.policy == BT_POLICY_LANES)
    s->lane = s->est_samples > (srv->o.lane_samples ? srv->o.lane_samples : kLaneSamples);
  *ticket = s->ticket;
  srv->queue.push_back(s.get());
  srv->all.emplace(s->ticket, std::move(s));
  return BT_OK;
}

int bt_server_step(bt_server* srv, uint32_t* running, uint32_t* queued) {
  if (!srv) return BT_E_INVALID;
  srv->err.clear();
  return guarded(&srv->err, [&] {
    // At most kTicksAhead ticks per lane queued on the GPU: the host prepares
    // the next tick while the previous one runs, and a loop of steps keeps
    // pace. Each lane is paced by its own ticks (a full lane skips this step);
    // the host blocks only when every lane with running jobs is full, on the
    // first such lane's oldest tick (lane 0 holds the small jobs).
    const uint32_t lanes = (uint32_t)srv->engines.size();
    auto pop_tick = [&](uint32_t lane, bool block) -> bool {
      std::deque<hipEvent_t>& q = srv->ticks[lane];
      if (q.empty()) return false;
      const hipError_t r = block ? hipEventSynchronize(q.front()) : hipEventQuery(q.front());
      if (r == hipErrorNotReady) return false;
      (void)hipEventDestroy(q.front());
      q.pop_front();
      HIPCHK(r);
      return true;
    };
    for (uint32_t lane = 0; lane < lanes; ++lane)
      while (pop_tick(lane, false)) {
      }
    {
      std::vector<bool> active(lanes, false);
      for (const ServedJob* s : srv->running) active[s->lane] = true;
      int wait_lane = -1;
      bool room = false;
      for (uint32_t lane = 0; lane < lanes; ++lane) {
        if (!active[lane]) continue;
        if (srv->ticks[lane].size() < kTicksAhead) room = true;
        else if (wait_lane < 0) wait_lane = (int)lane;
      }
      if (!room && wait_lane >= 0) pop_tick((uint32_t)wait_lane, true);
    }
    // nothing left to launch: wait for the results being copied out
    reap(srv, srv->running.empty() && srv->queue.empty());
    const bool shortest = srv->o.policy == BT_POLICY_SHORTEST;
    auto remaining = [](const ServedJob* j) -> uint64_t {
      const uint64_t n = j->job ? j->job->n : j->est_samples;
      const uint32_t done = j->job ? j->job->iteration : 0u;
      return n * (uint64_t)(j->iterations - std::min(done, j->iterations));
    };
    if (shortest) {
      // least work first, arrival order among equals; a job queued for more
      // than kStarveMs goes first (in arrival order), so continuous small
      // arrivals cannot hold a large job back indefinitely
      const double now = srv->now();
      std::stable_sort(srv->queue.begin(), srv->queue.end(),
                       [&](const ServedJob* a, const ServedJob* b) {
                         const bool sa = now - a->submit_ms > kStarveMs;
                         const bool sb = now - b->submit_ms > kStarveMs;
                         if (sa != sb) return sa;
                         if (sa) return a->ticket < b->ticket;
                         return remaining(a) < remaining(b);
                       });
    }
    // Admission while the footprint fits the budget and the in-flight cap.
    std::vector<ServedJob*> admit;
    uint64_t used = srv->used;
    while (!srv->queue.empty() && srv->running.size() + admit.size() < srv->o.max_inflight) {
      ServedJob* s = srv->queue.front();
      if (used + s->estimate > srv->o.memory_budget) break;
      used += s->estimate;
      admit.push_back(s);
      srv->queue.pop_front();
    }
    if (!admit.empty()) {
      // each lane's admitted jobs in one upload on its engine, queued behind
      // the ticks on the GPU (the host does not wait)
      std::vector<std::unique_ptr<bt::Job>> jobs(admit.size());
      for (uint32_t lane = 0; lane < srv->engines.size(); ++lane) {
        bt::Engine& e = *srv->engines[lane];
        std::vector<size_t> idx;
        std::vector<bt::GraphIn> gs;
        for (size_t i = 0; i < admit.size(); ++i) {
          ServedJob* s = admit[i];
          if (s->lane != lane) continue;
          idx.push_back(i);
          gs.push_back({s->xy.data(), (uint32_t)(s->xy.size() / 2), s->starts.data(),
                        (uint32_t)s->starts.size(), s->opts, s->path_scan.data()});
        }
        if (gs.empty()) continue;
        std::vector<std::unique_ptr<bt::Job>> up;
        try {
          e.upload_jobs(gs, &up, true, false);
          for (size_t k = 0; k < idx.size(); ++k) jobs[idx[k]] = std::move(up[k]);
        } catch (const bt::Error& err) {
          // Isolate failures: retry one by one so one bad graph fails alone.
          for (size_t k = 0; k < idx.size(); ++k) {
            std::vector<std::unique_ptr<bt::Job>> one;
            try {
              e.upload_jobs({gs[k]}, &one, true, false);
              jobs[idx[k]] = std::move(one[0]);
            } catch (const bt::Error& err1) {
              ServedJob* s = admit[idx[k]];
              s->state = 3;
              s->err = err1.msg;
              s->done_ms = srv->now();
              srv->finished.push_back(s->ticket);
            }
          }
          (void)err;
        }
      }
      const double t = srv->now();
      for (size_t i = 0; i < admit.size(); ++i) {
        if (!jobs[i]) continue;
        ServedJob* s = admit[i];
        s->job = std::move(jobs[i]);
        // The density the server returns is that of the last iteration. With
        // BT_FLAG_DENSITY_ON_READ the earlier ones, never read, are not
        // computed; without it every iteration computes its density, as
        // Gpu::iterate does (both give the last iteration's bits).
        s->job->density_from =
            (s->opts.flags & BT_FLAG_DENSITY_ON_READ) ? s->iterations - 1 : 0u;
        s->state = 1;
        s->admit_ms = t;
        s->last_ms = t;
        s->samples = s->job->n;
        s->held = s->job->device_bytes();
        srv->held += s->held;
        srv->held_peak = std::max(srv->held_peak, srv->held);
        s->estimate = std::max<uint64_t>(s->estimate, s->held);
        srv->used += s->estimate;
        srv->running.push_back(s);
      }
    }
    // Lanes with CU masks: a lane runs on every CU while the other lane has
    // no job running or queued (its masked CUs would idle), on its mask
    // otherwise.
    if (lanes == 2 && (srv->engines[0]->masked() || srv->engines[1]->masked())) {
      bool busy[2] = {false, false};
      for (const ServedJob* s : srv->running) busy[s->lane] = true;
      for (const ServedJob* s : srv->queue) busy[s->lane] = true;
      for (uint32_t lane = 0; lane < 2; ++lane) srv->engines[lane]->set_wide(!busy[1 - lane]);
    }
    // One iteration of the advanced jobs in one batched launch per stage
    // (per lane).
    if (!srv->running.empty()) {
      std::vector<ServedJob*> adv;
      // SHORTEST, and LANES with a tick budget while no job waits for
      // admission: each lane's tick takes its jobs with the least remaining
      // work first, up to the budget (with a backlog, LANES ticks every
      // running job: throughput first)
      const bool lane_budget = srv->o.policy == BT_POLICY_LANES && srv->o.tick_samples &&
                               srv->queue.empty();
      if (shortest || lane_budget) {
        const double now = srv->now();
        std::vector<ServedJob*> order(srv->running);
        std::stable_sort(order.begin(), order.end(), [&](const ServedJob* a, const ServedJob* b) {
          const bool sa = now - a->last_ms > kStarveMs, sb = now - b->last_ms > kStarveMs;
          if (sa != sb) return sa;
          return remaining(a) < remaining(b);
        });
        const uint64_t budget = srv->o.tick_samples ? srv->o.tick_samples : kTickSamples;
        std::vector<uint64_t> taken(lanes, 0);
        for (ServedJob* s : order) {
          if (taken[s->lane] && taken[s->lane] + s->job->n > budget) continue;
          adv.push_back(s);
          taken[s->lane] += s->job->n;
        }
      } else {
        adv = srv->running;
      }
      const double t = srv->now();
      for (uint32_t lane = 0; lane < lanes; ++lane) {
        if (srv->ticks[lane].size() >= kTicksAhead) continue;  // this lane waits
        bt::Engine& e = *srv->engines[lane];
        std::vector<bt::Job*> js;
        for (ServedJob* s : adv)
          if (s->lane == lane) js.push_back(s->job.get());
        if (js.empty()) continue;
        e.iterate_jobs(js, false);  // queued; the next tick's host work overlaps it
        for (ServedJob* s : adv)
          if (s->lane == lane) s->last_ms = t;
        hipEvent_t ev = nullptr;
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        const hipError_t rq = hipEventRecord(ev, e.stream());
        if (rq != hipSuccess) (void)hipEventDestroy(ev);
        HIPCHK(rq);
        srv->ticks[lane].push_back(ev);
      }
      // Retire the jobs that ran their last iteration: they leave `running`
      // first, then each either joins `finishing` (copies queued; reap()
      // completes them) or, if queuing fails, fails alone; the first error
      // is reported after every retiring job has been handled.
      std::vector<ServedJob*> still, retire;
      for (ServedJob* s : srv->running)
        (s->job->iteration < s->iterations ? still : retire).push_back(s);
      srv->running.swap(still);
      int err_code = BT_OK;
      std::string err_msg;
      for (ServedJob* s : retire) {
        try {
          const bool dens = s->job->dens_cnt.p != nullptr;
          s->dw = dens ? s->job->dw : 0;
          s->dh = dens ? s->job->dh : 0;
          const size_t pb = (size_t)s->samples * 12;
          // records, density, then the upload's status word
          s->pinned = get_pinned(srv, pb + (size_t)s->dw * s->dh * 4 + 4, &s->pinned_bytes);
          HIPCHK(hipEventCreateWithFlags(&s->done, hipEventDisableTiming));
          srv->engines[s->lane]->queue_readback(
              *s->job, s->pinned, dens ? (float*)((unsigned char*)s->pinned + pb) : nullptr,
              s->done);
          srv->finishing.push_back(s);
        } catch (const bt::Error& e) {
          // copies queued before the error must end before the buffer is reused
          (void)hipStreamSynchronize(srv->engines[s->lane]->stream());
          retire_failed(srv, *s, e.msg);
          if (err_code == BT_OK) {
            err_code = e.code;
            err_msg = e.msg;
          }
        }
      }
      if (err_code != BT_OK) bt::fail(err_code, err_msg);
    }
    if (running) *running = (uint32_t)(srv->running.size() + srv->finishing.size());
    if (queued) *queued = (uint32_t)srv->queue.size();
  });
}

int bt_server_take_finished(bt_server* srv, uint64_t* tickets, size_t capacity, size_t* n) {
  if (!srv || !n || (capacity && !tickets)) return BT_E_INVALID;
  int rc = guarded(&srv->err, [&] { reap(srv, false); });
  if (rc != BT_OK) return rc;
  size_t k = 0;
  while (k < capacity && !srv->finished.empty()) {
    tickets[k++] = srv->finished.front();
    srv->finished.pop_front();
  }
  *n = k;
  return BT_OK;
}

int bt_server_get_info(const bt_server* srv, bt_server_info* out) {
  if (!srv || !out) return BT_E_INVALID;
  std::memset(out, 0, sizeof *out);
  out->queued = (uint32_t)srv->queue.size();
  out->running = (uint32_t)srv->running.size();
  out->finishing = (uint32_t)srv->finishing.size();
  out->footprint_bytes = srv->used;
  out->held_bytes = srv->held;
  out->held_peak_bytes = srv->held_peak;
  out->memory_budget = srv->o.memory_budget;
  return BT_OK;
}

int bt_server_poll(bt_server* srv, uint64_t ticket, bt_job_result* out) {
  if (!srv || !out) return BT_E_INVALID;
  auto it = srv->all.find(ticket);
  if (it == srv->all.end()) return BT_E_INVALID;
  if (it->second->done) {
    int rc = guarded(&srv->err, [&] { reap(srv, false); });
    if (rc != BT_OK) return rc;
  }
  const ServedJob& s = *it->second;
  out->ticket = s.ticket;
  out->state = s.state;
  out->samples = s.samples;
  out->iterations_done = s.job ? s.job->iteration : (s.state == 2 ? s.iterations : 0);
  out->iterations_total = s.iterations;
  out->submit_ms = s.submit_ms;
  out->admit_ms = s.admit_ms;
  out->done_ms = s.done_ms;
  if (s.state == 3) srv->err = s.err;
  return BT_OK;
}

int bt_server_read_points(bt_server* srv, uint64_t ticket, void* rec12, size_t n_samples) {
  if (!srv || !rec12) return BT_E_INVALID;
  auto it = srv->all.find(ticket);
  if (it == srv->all.end()) return BT_E_INVALID;
  const ServedJob& s = *it->second;
  if (s.state != 2) return BT_E_BUSY;
  if (n_samples != s.samples) return BT_E_INVALID;
  std::memcpy(rec12, s.pinned, n_samples * 12);
  return BT_OK;
}

int bt_server_read_density(bt_server* srv, uint64_t ticket, uint32_t kind, void* out,
                           size_t n_pixels) {
  if (!srv || !out || kind != 3) return BT_E_INVALID;
  auto it = srv->all.find(ticket);
  if (it == srv->all.end()) return BT_E_INVALID;
  const ServedJob& s = *it->second;
  if (s.state != 2) return BT_E_BUSY;
  if (n_pixels != (size_t)s.dw * s.dh) return BT_E_INVALID;
  std::memcpy(out, (const unsigned char*)s.pinned + (size_t)s.samples * 12, n_pixels * 4);
  return BT_OK;
}

int bt_server_release(bt_server* srv, uint64_t ticket) {
  if (!srv) return BT_E_INVALID;
  auto it = srv->all.find(ticket);
  if (it == srv->all.end()) return BT_E_INVALID;
  if (it->second->state < 2) return BT_E_BUSY;
  put_pinned(srv, *it->second);
  srv->all.erase(it);
  return BT_OK;
}

}  // extern "C"
