/*
 * bundling.h - C ABI of the MI355X edge-bundling engine.
 *
 * Drop-in for the compute half of the reference's `Gpu` facade
 * (lpstankus/bundling_thingy, src/gpu.rs) and the host helpers around it.
 * All arrays are plain host pointers and sizes; the caller owns them and they
 * are copied during the call. A context owns its device memory and one HIP
 * stream; it is not thread-safe (the reference's `&mut self`); many contexts
 * may coexist. Every call blocks until its device work is complete, like the
 * reference's submit + poll(Wait) (gpu.rs:484-490,515-516,551-552).
 *
 * Return values: BT_OK or a BT_E_* code; bt_last_error() gives the message.
 */
#ifndef BUNDLING_H
#define BUNDLING_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BT_ABI_VERSION 1

#define BT_OK 0
#define BT_E_INVALID 1   /* bad argument or malformed input */
#define BT_E_NO_GRAPH 2  /* "no uploaded graph" (gpu.rs:533) */
#define BT_E_CAPACITY 3  /* graph memory budget / sample capacity (resampling.rs:86-89,300-304,346-372) */
#define BT_E_DEVICE 4    /* HIP runtime failure */
#define BT_E_STATE 5     /* "neighbor radius is not configured" (gpu.rs:529-532) */
#define BT_E_MISMATCH 6  /* "GPU resampling count and write passes disagree" (resampling.rs:296-299) */
#define BT_E_BUSY 7      /* server: not admitted yet / queue full */

/* bt_options.flags */
#define BT_FLAG_DENSITY 1u    /* ColorBy::Density (palette.rs:3-10, the reference default) */
#define BT_FLAG_DISTORTION 2u /* Options::evaluate_distortion: staging bytes count in the budget */
/* With BT_FLAG_DENSITY: iterations skip the splat, blur and normalize; the
 * density of the current positions is computed when it is read
 * (bt_read_density*, density colouring in bt_render*), with the bits an
 * iteration would have produced. On the server (bt_server_submit options) a
 * job computes only its last iteration's density, the one it returns;
 * without the flag it computes every iteration's, as Gpu::iterate does.
 * No counterpart in the reference. */
#define BT_FLAG_DENSITY_ON_READ 4u

/* Options of one context / job. Replaces the reference's `Options`
 * (main.rs:25-90) plus the map size (gpu.rs:21,99-102).
 * spacing == 0 selects reference mode: samples 1 texel apart on a map_w x
 * map_h map, integer radius_texels, a map-sized density and sigma = 6.
 * spacing > 0 selects generalized mode (canvas units = the unit square after
 * the reference's 95 % normalization): radius and sigma follow geometric
 * schedules over schedule_len iterations and the density is density_w x
 * density_h. The generalized formulas reduce exactly to reference mode when
 * spacing = 1/(S-1), radius = r/(S-1), density = S x S and sigma = 6. */
typedef struct bt_options {
  uint32_t map_w, map_h;           /* reference map size (SIZE = 1600) */
  uint32_t radius_texels;          /* Options::neighbor_radius; 0 = not configured */
  uint32_t flags;                  /* BT_FLAG_* */
  uint64_t max_graph_memory_bytes; /* Options::max_graph_memory_mib * MiB */
  int32_t device;                  /* HIP device ordinal */
  float spacing;                   /* generalized: sample spacing, canvas units */
  float radius0, radius1;          /* generalized: radius schedule, canvas units */
  float sigma0, sigma1;            /* generalized: blur sigma schedule, pixels */
  uint32_t schedule_len;           /* generalized: iterations the schedules span */
  uint32_t density_w, density_h;   /* generalized: density size */
  uint32_t cell_split;             /* grid rows per radius; 0 = autotuned */
  uint32_t x_split;                /* grid bins per row height along x; 0 = 16,
                                      1 = square cells as in the reference */
  uint32_t reserved[2];
} bt_options;

/* Exact parameters of one iteration (what the kernels use). */
typedef struct bt_iter_params {
  float res_x, res_y;   /* texel resolution: texel_scale = res * 0.5 */
  float radius;         /* texels */
  float cell;           /* grid row height, texels (the reference's cell side) */
  uint32_t gx, gy;      /* grid dims: bins per row, rows */
  float sigma;          /* pixels */
  uint32_t taps;        /* blur half-width = ceil(3 sigma) */
  uint32_t density_w, density_h;
  float cell_w;         /* grid bin width along x, texels (cell / x_split) */
  uint32_t task_bins;   /* bins per neighbor-mean task block along x */
} bt_iter_params;

typedef struct bt_info {
  uint32_t jobs;
  uint32_t samples;     /* job 0 */
  uint32_t paths;       /* job 0: path count (starts - 1) */
  uint32_t iteration;   /* job 0: iterations since upload */
  uint64_t capacity;    /* job 0: sample capacity of the budget model */
  uint64_t reserved_bytes;
  uint64_t device_bytes;/* device memory held by the context */
  uint32_t split_tasks; /* neighbor-mean tasks of the last iteration run that were
                           split across waves (long candidate streams) */
  uint32_t split_parts; /* their parts */
  /* executed work of that iteration's neighbor mean (all jobs): */
  uint64_t test_steps;  /* wave-steps: one candidate tested against a wave's 64 samples */
  uint64_t neighbors;   /* neighbors found for the movable samples, |N(p)| summed */
  uint64_t candidates;  /* candidates in the tasks' union streams (before lane groups) */
  uint64_t group_steps; /* of test_steps, those of lane-group streams (the rest stream
                           one candidate to all 64 lanes) */
} bt_info;

typedef struct bt_ctx bt_ctx;

/* Reference defaults: map 1600 x 1600, no radius, 1024 MiB, density on. */
void bt_options_default(bt_options* opts);

/* Gpu::new (gpu.rs:59-168). */
int bt_create(const bt_options* opts, bt_ctx** out);
void bt_destroy(bt_ctx* ctx);
const char* bt_last_error(const bt_ctx* ctx);

/* Gpu::upload (gpu.rs:492-525): resamples the graph (CSR polylines: xy holds
 * n_points (x, y) pairs, starts holds n_starts = paths + 1 offsets, as in
 * Graph::points()/starts(), graph.rs:93-99), computes the initial density and
 * drops any previous graph. *samples receives the sample count. */
int bt_upload(bt_ctx* ctx, const float* xy, uint32_t n_points, const uint32_t* starts,
              uint32_t n_starts, uint32_t* samples);

/* Gpu::iterate (gpu.rs:527-553): one neighbor-mean iteration, density splat,
 * blur and normalize. */
int bt_iterate(bt_ctx* ctx);

/* n iterations back to back, captured once as a hipGraph and replayed. */
int bt_iterate_n(bt_ctx* ctx, uint32_t n);

/* Restore the state right after the last upload (positions, density,
 * iteration 0) without resampling. */
int bt_rewind(bt_ctx* ctx);

/* Gpu::distortion_layout (gpu.rs:438-467): one Reservation triple
 * {point_offset, segment_offset, count} per path (count 0 = skipped path). */
int bt_read_paths(bt_ctx* ctx, uint32_t* rec3, size_t n_paths);

/* Gpu::distortion_snapshot (gpu.rs:469-482): n_samples little-endian 12-byte
 * records {f32 x, f32 y, u32 edge_id} in Layout order (geometry.rs:131-173). */
int bt_read_points(bt_ctx* ctx, void* rec12, size_t n_samples);

/* Density readback (density.rs:82-166). kind: 0 = u32 counts, 1 = row-blurred
 * f32, 2 = blurred f32, 3 = normalized f32 (the R channel of the reference's
 * rgba16f texture, before f16 rounding), 4 = that R channel as the texture
 * holds it (density.rs:96, normalization.wgsl:16): n_pixels binary16 values
 * (uint16_t bits, rounded to nearest even). */
int bt_read_density(bt_ctx* ctx, uint32_t kind, void* out, size_t n_pixels);

/* Grid of the last iteration: the cell of every sample (in Layout order) and
 * the per-cell counts (row-major cy * gx + cx, spatial_grid.wgsl:63-70). */
int bt_read_grid(bt_ctx* ctx, uint32_t* cell_of, size_t n_samples, uint32_t* counts,
                 size_t n_cells);

/* Endpoint dedup (endpoints.wgsl:9-49, run by resampling.rs:137-149,240-249
 * when --endpoint-dots is set): the sample indices (Layout order) of the
 * distinct positions among the first and last samples of the paths with two
 * or more samples. Each position keeps its smallest index and the list is
 * ascending (the reference keeps the first inserter, in completion order).
 * *count always receives the number found; indices receives min(capacity,
 * count) of them and BT_E_CAPACITY is returned when capacity < count. */
int bt_read_endpoints(bt_ctx* ctx, uint32_t* indices, size_t capacity, uint32_t* count);

/* bt_read_endpoints that allocates no device memory. No counterpart in the
 * reference: it is for callers that cannot allocate (code under hipGraph
 * capture of its own launches, serving loops, harnesses that forbid
 * hipMalloc). The result equals bt_read_endpoints'. Its hash table, index
 * bitmap and index list sit in one static device array of
 * BT_NOALLOC_SCRATCH_BYTES per GPU, set up when a context is created on that
 * GPU and held until the process ends; graphs of more than
 * BT_NOALLOC_MAX_PATHS paths or BT_NOALLOC_MAX_SAMPLES samples are refused
 * with BT_E_CAPACITY. Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (from contexts used by different threads) race. */
#define BT_NOALLOC_MAX_PATHS (1u << 18)
#define BT_NOALLOC_MAX_SAMPLES (1u << 24)
#define BT_NOALLOC_SCRATCH_BYTES (256u + 32u * BT_NOALLOC_MAX_PATHS + \
                                  BT_NOALLOC_MAX_SAMPLES / 8u + 8u * BT_NOALLOC_MAX_PATHS)
int bt_read_endpoints_noalloc(bt_ctx* ctx, uint32_t* indices, size_t capacity, uint32_t* count);

/* ---------------------------------------------------------------- render */

/* Palette (palette.rs:13-24; VIRIDIS is the reference default). */
#define BT_PALETTE_CIVIDIS 0u
#define BT_PALETTE_VIRIDIS 1u
#define BT_PALETTE_PLASMA 2u
#define BT_PALETTE_BATLOW 3u
/* ColorBy (palette.rs:3-10): the normalized density map, or the edge id. */
#define BT_COLOR_DENSITY 0u
#define BT_COLOR_EDGE_ID 1u

/* Options of the path rasterizer (render_paths.wgsl, gpu.rs:234-335,572-613;
 * main.rs:58-81 --palette, --color-by, --endpoint-dots, --msaa). */
typedef struct bt_render_options {
  uint32_t palette;        /* BT_PALETTE_* */
  uint32_t color_by;       /* BT_COLOR_*; DENSITY needs BT_FLAG_DENSITY */
  uint32_t msaa;           /* 2 or 4 samples per pixel */
  uint32_t endpoint_dots;  /* nonzero: dots at path ends (endpoints.wgsl) */
  uint32_t reserved[4];
} bt_render_options;

/* The 16 RGB stops of a palette (colors48 = 16 x {r, g, b}) and its endpoint
 * dot color (rgb3), palette.rs:26-130. */
int bt_palette(uint32_t palette, float* colors48, float* rgb3);

/* viridis, density, 4x MSAA, no dots (the reference's CLI defaults). */
void bt_render_options_default(bt_render_options* ro);

/* Render target size: the map (gpu.rs:99-102,133-134), i.e. the density
 * map's width and height. */
int bt_render_size(bt_ctx* ctx, uint32_t* width, uint32_t* height);

/* Gpu::render + Gpu::readback (gpu.rs:572-651): draws the current paths
 * (one quad of half-width 0.001 clip per segment, opacity 0.1, alpha
 * blending), then the endpoint outlines and dots, over the clear color
 * (0.025, 0.025, 0.035), resolves the MSAA samples and copies the
 * width x height RGBA8 image (row 0 = top) into rgba (n_bytes = w * h * 4).
 * With no uploaded graph it returns the blank frame that calibrates the
 * noisiness background (main.rs:117-123). */
int bt_render(bt_ctx* ctx, const bt_render_options* ro, uint8_t* rgba, size_t n_bytes);

/* The same for one job of a batched context (bt_upload_batch). */
int bt_render_job(bt_ctx* ctx, uint32_t job, const bt_render_options* ro, uint8_t* rgba,
                  size_t n_bytes);
int bt_render_size_job(bt_ctx* ctx, uint32_t job, uint32_t* width, uint32_t* height);

int bt_get_info(bt_ctx* ctx, bt_info* info);

/* Parameters iteration `iteration` of a job with these options runs with
 * (neighbor_mean.rs:19-46 grid sizes, blur.wgsl:41-60 taps). */
int bt_iter_params_for(const bt_options* opts, uint32_t iteration, bt_iter_params* out);

/* Per-stage device time of the last bt_profile call, in ms, by stage:
 * 0 clear, 1 bin, 2 scan, 3 tasks, 4 scatter, 5 neighbor_mean, 6 blur_rows,
 * 7 blur_cols, 8 normalize (the reference's GpuTimings stage keys,
 * neighbor_mean.rs:5-13, density.rs:5-9). Runs n iterations eagerly with
 * hipEvents around each stage; returns total ms. */
int bt_profile(bt_ctx* ctx, uint32_t n, double* stage_ms, uint32_t n_stages, double* total_ms);

/* Autotuner (no counterpart in the reference): for each of the first n
 * iterations, times one iteration under every candidate grid cell split
 * (cells per radius) from the same state, keeps the fastest in chosen[i]
 * (best_ms[i] its time) and uses it from then on, including in hipGraphs.
 * Results are bit-identical under any split. Leaves the jobs rewound. */
int bt_autotune(bt_ctx* ctx, uint32_t n, const uint32_t* candidates, uint32_t n_candidates,
                uint32_t* chosen, double* best_ms);

/* Kernel shapes (no counterpart in the reference: its workgroup sizes are
 * fixed, neighbor_mean.rs:4, blur.wgsl:13-35, normalization.wgsl:5). Per
 * schedule entry, three shape indices: the neighbor mean's lanes per group
 * x waves per workgroup (0 = 16 x 4, 1 = 32 x 4, 2 = 16 x 2, 3 = 32 x 2,
 * 4 = 16 x 8, 5 = 32 x 8), the row blur's tile (0 = 64 x 4, 1 = 128 x 2,
 * 2 = 32 x 8, 3 = 256 x 1 pixels) and the column blur's (0 = 64 x 32,
 * 1 = 64 x 16, 2 = 64 x 64; 3 = 64 x 32 and 4 = 64 x 16 run both blur
 * passes in one kernel, which ignores the row shape, for blurs of at most
 * 32 taps, sigma <= 10.67 px; wider blurs fall back to column shape 0).
 * Results are bit-identical under every shape. A launch over several jobs
 * runs the first job's shapes. */
#define BT_NM_SHAPES 6u
#define BT_ROW_SHAPES 4u
#define BT_COL_SHAPES 5u
/* Autotuner of the shapes: for each of the first n iterations, times each
 * stage under every shape of that stage from the same state (the other
 * stages at their current choice; the column shapes by the time of both
 * blur stages, as the fused shapes do both), keeps the fastest for every
 * job (shapes[3 i + k]; either output may be NULL) and uses it from then
 * on, including in hipGraphs. stage_ms[3 i + k] is the stage time of the
 * chosen shapes: neighbor mean, row blur (0 under a fused column shape),
 * column blur. Leaves the jobs rewound. Run after bt_autotune (the grid
 * split) when both are wanted. */
int bt_autotune_shapes(bt_ctx* ctx, uint32_t n, uint32_t* shapes, double* stage_ms);
/* Sets the shapes of schedule entry `iteration` of every job (all entries
 * when iteration is 0xffffffff). */
int bt_set_shapes(bt_ctx* ctx, uint32_t iteration, uint32_t nm_shape, uint32_t row_shape,
                  uint32_t col_shape);
/* The shapes job `job` runs iteration `iteration` with. */
int bt_get_shapes(bt_ctx* ctx, uint32_t job, uint32_t iteration, uint32_t* shapes3);

/* Measurement helper (no counterpart in the reference): copies `bytes` bytes
 * device-to-device on the context's stream (HIP's copy kernel), so profiler
 * byte counters can be scaled against a transfer of known size. *ms receives
 * the copy time from HIP events. */
int bt_copy_probe(bt_ctx* ctx, uint64_t bytes, double* ms);

/* Parameters job `job` runs iteration `iteration` with (after autotuning). */
int bt_job_iter_params(bt_ctx* ctx, uint32_t job, uint32_t iteration, bt_iter_params* out);

/* ---------------------------------------------------------------- batched */

/* One graph of a batch. */
typedef struct bt_graph_desc {
  const float* xy;
  uint32_t n_points;
  const uint32_t* starts;
  uint32_t n_starts;
  const bt_options* opts; /* per-graph options; NULL = the context's */
} bt_graph_desc;

/* Replaces the context's graphs with n_graphs independent jobs resampled in
 * one set of launches; samples[i] receives job i's sample count. Every later
 * bt_iterate / bt_iterate_n advances all jobs with one launch per stage. */
int bt_upload_batch(bt_ctx* ctx, const bt_graph_desc* graphs, uint32_t n_graphs,
                    uint32_t* samples);

int bt_read_paths_job(bt_ctx* ctx, uint32_t job, uint32_t* rec3, size_t n_paths);
int bt_read_points_job(bt_ctx* ctx, uint32_t job, void* rec12, size_t n_samples);
/* The samples of bt_read_points_job split into two arrays: n_samples
 * positions {f32 x, f32 y} and n_samples edge ids (edge_ids may be NULL).
 * No counterpart in the reference, whose Decoder reads the 12-byte records;
 * for hosts that keep positions and ids apart (no host-side decode). */
int bt_read_positions_job(bt_ctx* ctx, uint32_t job, float* xy, uint32_t* edge_ids,
                          size_t n_samples);
int bt_read_density_job(bt_ctx* ctx, uint32_t job, uint32_t kind, void* out, size_t n_pixels);
int bt_job_samples(bt_ctx* ctx, uint32_t job, uint32_t* samples, uint32_t* iteration);

/* ------------------------------------------------------------ host helpers */

/* Graph::parse (graph.rs:35-77). Arrays are malloc'ed; free with bt_free.
 * bbox = {min_x, min_y, max_x, max_y}. err receives up to errlen bytes. */
int bt_trl_parse(const char* text, size_t len, float** xy, uint32_t* n_points,
                 uint32_t** starts, uint32_t* n_starts, float* bbox, char* err,
                 size_t errlen);
void bt_free(void* p);

/* noisiness::background (noisiness.rs:28-42): the uniform RGB of an opaque
 * blank frame. err receives the reference's message on failure. */
int bt_noisiness_background(const uint8_t* rgba, uint32_t width, uint32_t height, uint8_t* rgb,
                            char* err, size_t errlen);

/* noisiness::measure (noisiness.rs:44-74): pixels whose RGB differs exactly
 * from rgb. ratio = colored / background, NaN when background == 0. */
typedef struct bt_noisiness {
  uint64_t colored_pixels, background_pixels, total_pixels;
  double noisiness, colored_to_background_ratio;
} bt_noisiness;
int bt_noisiness_measure(const uint8_t* rgba, uint32_t width, uint32_t height,
                         const uint8_t* rgb, bt_noisiness* out, char* err, size_t errlen);

/* write_image (main.rs:193-221): an 8-bit RGBA PNG (zlib deflate, filter 0). */
int bt_write_png(const char* path, const uint8_t* rgba, uint32_t width, uint32_t height,
                 char* err, size_t errlen);

/* Graph::normalization (graph.rs:79-91): out = {min_x, min_y, scale, tr_x, tr_y}. */
int bt_normalization(const float* bbox, float* out);

/* Allocation::new (resampling.rs:321-397) with Resampler::new's budget check
 * (resampling.rs:72-89): sample capacity and reserved bytes of a graph. */
int bt_capacity(const bt_options* opts, uint32_t n_points, uint32_t n_starts,
                uint64_t* capacity, uint64_t* reserved_bytes);

/* Blur weights w[0..taps] for sigma (blur.wgsl:42-60 at sigma = 6); returns taps. */
uint32_t bt_blur_weights(float sigma, float* w, uint32_t max_taps);

/* Aggregate distortion (distortion.rs:118-216) of a snapshot against the
 * baseline snapshot under Layout paths (rec3 triples). out = {distortion,
 * rms displacement (clip units), rms displacement in pixels at map size}. */
int bt_distortion(const float* original_xy, const float* current_xy, const uint32_t* rec3,
                  uint32_t n_paths, uint32_t n_samples, uint32_t map_size, double* out);

/* ----------------------------------------------------------------- served */

typedef struct bt_server bt_server;

#define BT_POLICY_FIFO 0      /* admit in arrival order; every running job advances each tick */
#define BT_POLICY_SHORTEST 1  /* admit and advance the least remaining work first, up to
                                 tick_samples sample-updates per tick */
#define BT_POLICY_LANES 2     /* admit in arrival order; jobs of at most lane_samples
                                 samples run in a small-job lane, the others in a large-job
                                 lane, each lane ticking on its own stream, so the two lanes'
                                 ticks overlap on the GPU and small jobs do not wait behind
                                 large jobs' iterations. With tick_samples = 0 each tick takes
                                 all of its lane's running jobs; otherwise, while no job waits
                                 for admission, it takes the lane's jobs with the least
                                 remaining work first up to tick_samples sample-updates, and
                                 all of them when a backlog forms (throughput first). With
                                 CU masks (small_cu_mask, large_cu_mask) the lanes run on
                                 disjoint CUs while both have work */

typedef struct bt_server_options {
  bt_options defaults;       /* device, flags */
  uint64_t memory_budget;    /* device bytes the in-flight jobs may hold */
  uint32_t max_inflight;     /* jobs holding device state at once */
  uint32_t policy;           /* BT_POLICY_* */
  uint64_t tick_samples;     /* BT_POLICY_SHORTEST: sample-updates per tick (0 = 4M);
                                BT_POLICY_LANES: per lane and tick while no job waits
                                (0 = every running job) */
  uint64_t lane_samples;     /* BT_POLICY_LANES: largest job (samples, from the count pass
                                at submission) of the small-job lane (0 = 1M) */
  uint32_t small_cu_mask;    /* BT_POLICY_LANES: CUs the small-job lane may use while the
                                large-job lane has work, one bit per CU of every group of
                                32 (0 = all). Bit b selects a CU of XCD b mod 8, so every
                                residue mod 8 needs a set bit (else BT_E_INVALID); on
                                MI355X each byte is one shader engine of every XCD */
  uint32_t large_cu_mask;    /* the same for the large-job lane. A lane runs on every CU
                                while the other lane has no job running or queued */
} bt_server_options;

typedef struct bt_job_result {
  uint64_t ticket;
  uint32_t state;            /* 0 queued, 1 running, 2 done, 3 failed */
  uint32_t samples;
  uint32_t iterations_done;
  uint32_t iterations_total;
  double submit_ms, admit_ms, done_ms;  /* server clock */
} bt_job_result;

int bt_server_create(const bt_server_options* opts, bt_server** out);
void bt_server_destroy(bt_server* srv);
const char* bt_server_last_error(const bt_server* srv);

/* Queues a job (graph copied): `iterations` iterations under `opts`. */
int bt_server_submit(bt_server* srv, const float* xy, uint32_t n_points, const uint32_t* starts,
                     uint32_t n_starts, const bt_options* opts, uint32_t iterations,
                     uint64_t* ticket);

/* One scheduler tick: admits queued jobs that fit the memory budget and the
 * in-flight cap (and resamples them in one batched launch), advances running
 * jobs by one iteration in one batched launch per stage, retires finished
 * jobs. FIFO admits in arrival order and advances every running job;
 * SHORTEST admits and advances the jobs with the least remaining work
 * (samples x iterations left) first, until tick_samples sample-updates are
 * taken (at least one job), except that a job queued, or running but not
 * advanced, for more than a second goes first. */
int bt_server_step(bt_server* srv, uint32_t* running, uint32_t* queued);

int bt_server_poll(bt_server* srv, uint64_t ticket, bt_job_result* out);
/* Tickets of the jobs that finished (done or failed) since the last call,
 * in completion order: up to `capacity` of them, *n receives how many. Picks
 * up completed retirements first, like bt_server_poll. A serving loop calls
 * this instead of polling every pending ticket. */
int bt_server_take_finished(bt_server* srv, uint64_t* tickets, size_t capacity, size_t* n);
/* A finished job's results: its samples after the last iteration, and the
 * normalized density of that iteration (the only density a job computes on
 * the server; bits equal to running the job alone through bt_iterate). */
int bt_server_read_points(bt_server* srv, uint64_t ticket, void* rec12, size_t n_samples);
int bt_server_read_density(bt_server* srv, uint64_t ticket, uint32_t kind, void* out,
                           size_t n_pixels);
/* Frees a finished job's device state. */
int bt_server_release(bt_server* srv, uint64_t ticket);
double bt_server_clock_ms(const bt_server* srv);

/* Admission accounting (no counterpart in the reference): footprint_bytes
 * is what admission charges against memory_budget (each admitted job's
 * estimate, or the bytes its buffers hold if larger); held_bytes is what the
 * admitted jobs' device buffers hold now (allocator blocks), held_peak_bytes
 * its maximum so far. Admission keeps held_bytes <= footprint_bytes <=
 * memory_budget. */
typedef struct bt_server_info {
  uint32_t queued, running, finishing, reserved;
  uint64_t footprint_bytes, held_bytes, held_peak_bytes, memory_budget;
} bt_server_info;
int bt_server_get_info(const bt_server* srv, bt_server_info* out);

#ifdef __cplusplus
}
#endif
#endif
