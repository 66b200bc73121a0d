/*
 * CPU oracle for the edge-bundling hot path. TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference pipeline (lpstankus/bundling_thingy):
 * TRL parsing, 95 % normalization, arc-length resampling, spatial-grid binning,
 * the fixed-radius neighbor mean with fixed path endpoints, the unit-weight
 * density splat, the separable Gaussian blur, the peak and normalize steps,
 * the f64 distortion metric, and the export image (segment quads, MSAA
 * coverage, 8-bit alpha blending, endpoint dots, resolve; a written
 * rasterization contract shared with the HIP rasterizer, since the
 * reference's coverage comes from its driver). Each function cites the reference lines it
 * follows. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as the checker or the timed CPU
 * baseline; the product never links it.
 *
 * Arithmetic contract shared with the HIP path (so results are bit-identical):
 *  - FP32 with no contraction (-ffp-contract=off) except where fmaf() is
 *    written explicitly; correctly rounded '/' and sqrtf; WGSL round() is
 *    rintf() (half to even); float->u32 conversions are clamped first.
 *  - Neighbor test: d2 = fmaf(tx, tx, ty*ty) with t = (q - p) * texel_scale,
 *    accepted when d2 <= r*r (neighbor_mean.wgsl:140-146,193-197).
 *  - Neighbor sums are exact: every coordinate is quantized once per
 *    iteration to a 2^-24 fixed-point integer qf = rint(q * 2^24) and
 *    D = sum(qf) - n*pf is accumulated in int64, so the result does not depend
 *    on the order in which neighbors are visited. The mean is
 *    (float)((double)p + (double)D / ((double)n * 2^24)).
 */
#ifndef BT_ORACLE_H
#define BT_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Graph::normalization result (graph.rs:14-18,79-91). */
typedef struct {
    float min_x, min_y;
    float scale;
    float tr_x, tr_y;
} or_norm;

/* Parameters of one neighbor-mean iteration in texel units. */
typedef struct {
    float res_x, res_y;   /* samples per clip half-extent * 2: reference map size - 1 */
    float radius;         /* neighbor radius in texels */
    float cell;           /* grid cell side (row height) in texels (reference: == radius) */
    uint32_t gx, gy;      /* grid dims: ceil((res_x + 1) / cell_w), ceil((res_y + 1) / cell) */
    float cell_w;         /* bin width along x in texels; 0 = cell (the reference's square cells) */
} or_grid;

/* Derived per-iteration constants; identical formulas on the device. */
typedef struct {
    float ts_x, ts_y;     /* texel scale = res * 0.5 */
    float r2;             /* radius * radius */
    float pad_x, pad_y;   /* (res + 1) * 2.4e-7 (neighbor_mean.wgsl:44) */
    float reach_x, reach_y; /* radius + 4 * pad: cell window covers the whole disk */
} or_derived;

int or_version(void);

/* graph.rs:35-77. Parses TRL text. Output arrays are malloc'ed (free with or_free).
 * Returns 0, or 1 for malformed input (message in err, up to 256 bytes). */
int or_parse_trl(const char* text, size_t len, float** xy, uint32_t* n_points,
                 uint32_t** starts, uint32_t* n_starts, float bbox[4], char* err);
void or_free(void* p);

/* graph.rs:79-91. Returns 0, or 1 for an invalid extent. */
int or_normalization(const float bbox[4], or_norm* out);

/* resampling.wgsl:45-48. */
void or_normalized(const or_norm* nm, const float* xy, uint32_t i, float out[2]);

/* resampling.wgsl:82-136 count mode + count_paths (:138-159). counts[i] is the
 * path's sample count (0 when < 2, the path is skipped). Returns the total. */
uint64_t or_resample_count(const float* xy, const uint32_t* starts, uint32_t n_paths,
                           const or_norm* nm, float res_x, float res_y, uint32_t capacity,
                           uint32_t* counts);

/* resampling.wgsl:50-136 write mode + emit_paths (:161-183). Paths are laid out
 * in path order: offsets[i] is the exclusive prefix of counts. Writes points
 * (x, y) and the owning path id. Returns 0, or 2 if any path emits a count
 * different from counts[i] (the reference's overflow == 2). */
int or_resample_emit(const float* xy, const uint32_t* starts, uint32_t n_paths,
                     const or_norm* nm, float res_x, float res_y,
                     const uint32_t* counts, const uint64_t* offsets,
                     float* pts, uint32_t* edge_id, uint32_t* density,
                     uint32_t dens_w, uint32_t dens_h);

void or_derive(const or_grid* g, or_derived* d);

/* spatial_grid.wgsl:59-81: cell index per point and per-cell counts. */
void or_grid_bin(const float* pts, uint32_t n, const or_grid* g,
                 uint32_t* cell_of, uint32_t* counts);

/* neighbor_mean.wgsl:56-203 semantics with exact fixed-point sums (see the
 * header). fixed[i] != 0 marks a path endpoint, which is copied. Optional
 * neighbors[i] receives |N(p)| (0 for fixed points). threads <= 0 means all. */
void or_neighbor_mean(const float* in, const uint8_t* fixed, uint32_t n,
                      const or_grid* g, float* out, uint32_t* neighbors, int threads);

/* Same, computing only every stride-th sample (the others are copied): a
 * bounded sample of the work for the timed CPU baseline. */
void or_neighbor_mean_strided(const float* in, const uint8_t* fixed, uint32_t n,
                              const or_grid* g, uint32_t stride, float* out, int threads);

/* The same means through the reference's tiled algorithm
 * (neighbor_mean.wgsl:40-59,75-161): per grid cell a summary (sample count,
 * coordinate sums, bounding box of its samples); a cell whose box lies
 * entirely within the radius (with the reference's padding of 2.4e-7 texels
 * per texel of map size added to its farthest corner) contributes its
 * summary, a cell entirely outside nothing, and the samples of the others
 * are tested one by one. Sums are the exact fixed-point sums, so the result
 * equals or_neighbor_mean's bit for bit; the classification is evaluated in
 * double precision. Computes every stride-th sample (the others are copied)
 * for the timed CPU baseline. */
void or_neighbor_mean_cells(const float* in, const uint8_t* fixed, uint32_t n,
                            const or_grid* g, uint32_t stride, float* out, int threads);

/* The reference's literal sparse form (neighbor_mean.wgsl:163-203): f32
 * sequential sums over the cells in [texel - reach, texel + reach] with
 * reach = r + pad, points visited in index order within each cell. Used only
 * as a tolerance cross-check of or_neighbor_mean. */
void or_neighbor_mean_f32(const float* in, const uint8_t* fixed, uint32_t n,
                          const or_grid* g, float* out);

/* Count candidate pairs in the 3x3 cells of side r around each point. */
uint64_t or_candidate_pairs(const float* pts, uint32_t n, const or_grid* g);

/* neighbor_mean.wgsl:66-73, resampling.wgsl:69-75 (u32 splat, y flipped). */
void or_splat(const float* pts, uint32_t n, uint32_t w, uint32_t h, uint32_t* counts);

/* Gaussian weights w[0..taps] for sigma; taps = ceil(3 sigma). Each weight is
 * exp(-k^2 / (2 sigma^2)) / Z rounded to 10 decimals, then to f32, which
 * reproduces blur.wgsl:42-60 exactly for sigma = 6. Returns taps. */
uint32_t or_blur_weights(float sigma, float* w, uint32_t max_taps);

/* blur.wgsl:13-65: horizontal pass from u32 counts (zero padding). */
void or_blur_rows(const uint32_t* in, uint32_t w, uint32_t h, const float* wt,
                  uint32_t taps, float* out);
/* blur.wgsl:13-83: vertical pass on f32 plus the peak. Returns the peak. */
float or_blur_cols(const float* in, uint32_t w, uint32_t h, const float* wt,
                   uint32_t taps, float* out);
/* normalization.wgsl:5-17 (R channel, f32). */
void or_normalize(const float* in, size_t n, float peak, float* out);

/* distortion.rs:118-216. Per-path point ranges (offset, count) into two
 * snapshots; returns 0 and writes distortion and rms, or 1 on invalid input. */
int or_distortion(const float* original, const float* current, const uint64_t* offsets,
                  const uint32_t* counts, uint32_t n_paths, double* distortion,
                  double* rms);

/* One whole iteration (the scope of Gpu::iterate, gpu.rs:527-553): mean,
 * splat, blur rows, blur cols + peak, normalize. Returns the peak. */
float or_iterate(const float* in, const uint8_t* fixed, uint32_t n, const or_grid* g,
                 float* out, uint32_t dens_w, uint32_t dens_h, float sigma,
                 uint32_t* counts, float* tmp, float* blurred, float* normalized,
                 int threads);

/* Export image (render_paths.wgsl:1-97, gpu.rs:572-651). pts: n x {x, y}
 * clip units in Layout order; edge: n edge ids; paths: n_paths x {offset,
 * segment offset, count}; density: dw x dh normalized map for density
 * colouring, NULL for edge-id colouring; palette 16 x 3; endpoint 3. */
typedef struct {
    const float* pts;
    const uint32_t* edge;
    uint32_t n, n_edges;
    const uint32_t* paths;
    uint32_t n_paths;
    const float* density;
    uint32_t dw, dh, w, h, msaa, dots;
    const float* palette;
    const float* endpoint;
} or_render_in;

/* w x h RGBA8, row 0 at the top. Returns nonzero on bad input. */
int or_render(const or_render_in* in, uint8_t* rgba);
/* Fragments per sample (w*h*msaa, sample k of pixel p at p*msaa+k). */
void or_coverage(const or_render_in* in, uint32_t* counts);
/* Nearest binary16 value (ties to even) of x, as an f32. */
float or_half_round(float x);

#ifdef __cplusplus
}
#endif
#endif
