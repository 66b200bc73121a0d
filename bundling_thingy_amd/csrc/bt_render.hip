// Path rasterizer for the export image (render_paths.wgsl:1-97, gpu.rs:572-651)
// and the endpoint dots, as compute kernels:
//   count    one thread per sample: the quad of the segment (i, i+1), its
//            MSAA coverage per pixel, and how many fragments it makes
//   scan     exclusive sum of the counts (fragment offsets in draw order):
//            workgroup sums, a scan of those, then each tile's local scan
//   emit     the same walk, writing {pixel | coverage << 24, segment}
//   sort     stable LSD radix sort on the pixel bits, 10 bits per pass:
//            per-tile digit histograms, their scan in digit-major order,
//            then a scatter that ranks equal digits with wave ballots, so
//            each pixel's fragments end up together, still in draw order
//   counts   density colouring (one colour per pixel): fragments per sample,
//            one atomic per run of a pixel's fragments in a wave
//   runs     edge colouring: first and last fragment of every pixel
//   dots     endpoint outline / dot coverage bits per sample (atomicOr)
//   resolve  one thread per pixel: blends its fragments into its samples in
//            order (8-bit storage after every blend, as an RGBA8 MSAA target
//            holds them; with density colouring, its count of blends of the
//            pixel's colour), applies the dots, averages the samples
//
// Rasterization contract (the hardware's MSAA coverage is not pinned by any
// reference output, so this is the definition the oracle shares):
// corners are computed in f32 as the vertex shader does, mapped to pixels by
// ((x * 0.5 + 0.5) * W, (0.5 - y * 0.5) * H) and snapped to 1/256 pixel; each
// quad is the strip's two triangles (c0 c1 c2) and (c2 c1 c3) with no
// culling; a sample is inside a triangle when all three integer edge
// functions are positive, or zero on a top or left edge. Sample positions are
// the standard 2x / 4x locations. Fragment shading happens once per pixel at
// its center; endpoint dots are shaded per sample (interpolate(sample)).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "bt_engine.h"
#include "bt_launch.h"
#include "bt_palette.h"

namespace bt {
namespace {

// Standard sample locations, 1/256 pixel units: {x, y} per sample.
__constant__ int kSample4[4][2] = {{96, 32}, {224, 96}, {32, 160}, {160, 224}};
__constant__ int kSample2[2][2] = {{192, 192}, {64, 64}};

__constant__ float kPalette[4][16][3] = BT_PALETTE_TABLE;
__constant__ float kEndpoint[4][3] = BT_ENDPOINT_TABLE;

constexpr float kOpacity = 0.1f;       // PATH_OPACITY
constexpr float kHalfWidth = 0.001f;   // quad half-width, clip units
constexpr float kDotRadius = 3.0f;     // ENDPOINT_RADIUS, pixels
constexpr float kDotOutline = 1.0f;    // ENDPOINT_OUTLINE_WIDTH, pixels
constexpr float kClear[3] = {0.025f, 0.025f, 0.035f};  // gpu.rs:586-591
constexpr uint32_t kMaxPixelBits = 24;  // fragment keys hold the pixel in bits 0..23

struct Frame {
  const float* pts;  // current samples, Layout order, {x, y, edge id}
  uint32_t n;        // samples
  uint32_t w, h;     // render target
  uint32_t msaa;     // 2 or 4
  uint32_t n_edges;  // paths (the reference's edges buffer length)
  const float* dens; // normalized density, dw x dh (density colouring)
  uint32_t dw, dh;
  uint32_t palette, color_by;
};

__device__ __forceinline__ void sample_pos(uint32_t msaa, uint32_t k, int& sx, int& sy) {
  if (msaa == 4u) {
    sx = kSample4[k][0];
    sy = kSample4[k][1];
  } else {
    sx = kSample2[k][0];
    sy = kSample2[k][1];
  }
}

// Unorm8 storage of a blend result.
__host__ __device__ __forceinline__ uint32_t unorm8(float x) {
  return (uint32_t)rintf(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f);
}

// BlendState::ALPHA_BLENDING with a constant source alpha: one channel.
__device__ __forceinline__ uint32_t blend8(uint32_t dst, float src) {
  return unorm8(src * kOpacity + ((float)dst * (1.0f / 255.0f)) * (1.0f - kOpacity));
}

// Fixed-point (1/256 px) corners of the quad of segment (i, i + 1): the
// vertex shader's corners (render_paths.wgsl:36-50) through the viewport
// transform. False when i is the last sample of its path, or a corner is not
// finite / out of the fixed-point range.
struct Quad {
  int x[4], y[4];
};

__device__ bool seg_quad(const Frame& f, uint32_t i, Quad& q) {
  if (i + 1u >= f.n) return false;
  const float* a = f.pts + 3ull * i;
  if (__float_as_uint(a[2]) != __float_as_uint(a[5])) return false;
  const float x0 = a[0], y0 = a[1], x1 = a[3], y1 = a[4];
  const float dx = x1 - x0, dy = y1 - y0;
  const float len = sqrtf(dx * dx + dy * dy);
  const float k = kHalfWidth / fmaxf(len, 1e-10f);
  const float nx = -dy * k, ny = dx * k;
  const float cx[4] = {x0 + nx, x0 - nx, x1 + nx, x1 - nx};
  const float cy[4] = {y0 + ny, y0 - ny, y1 + ny, y1 - ny};
  const float W = (float)f.w, H = (float)f.h;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float fx = (cx[c] * 0.5f + 0.5f) * W;
    const float fy = (0.5f - cy[c] * 0.5f) * H;
    if (!(fabsf(fx) < 1.0e6f && fabsf(fy) < 1.0e6f)) return false;
    q.x[c] = (int)rintf(fx * 256.0f);
    q.y[c] = (int)rintf(fy * 256.0f);
  }
  return true;
}

// One triangle's edge functions E = dx * (sy - ay) - dy * (sx - ax), oriented
// so the inside is E > 0; bias 0 on top / left edges (ties inside), else 1.
struct Tri {
  int dx[3], dy[3], ax[3], ay[3], bias[3];
  bool live;
};

__device__ void tri_setup(int ax, int ay, int bx, int by, int cx, int cy, Tri& t) {
  const long long area = (long long)(bx - ax) * (cy - ay) - (long long)(by - ay) * (cx - ax);
  t.live = area != 0;
  if (area < 0) {
    int tx = bx, ty = by;
    bx = cx;
    by = cy;
    cx = tx;
    cy = ty;
  }
  const int px[3] = {ax, bx, cx}, py[3] = {ay, by, cy};
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int n = e == 2 ? 0 : e + 1;
    t.ax[e] = px[e];
    t.ay[e] = py[e];
    t.dx[e] = px[n] - px[e];
    t.dy[e] = py[n] - py[e];
    const bool top_left = t.dy[e] < 0 || (t.dy[e] == 0 && t.dx[e] > 0);
    t.bias[e] = top_left ? 0 : 1;
  }
}

struct Span {
  int x0, x1, y0, y1;  // inclusive pixel range (empty when x0 > x1 or y0 > y1)
};

__device__ Span quad_span(const Frame& f, const Quad& q) {
  int mnx = q.x[0], mxx = q.x[0], mny = q.y[0], mxy = q.y[0];
#pragma unroll
  for (int c = 1; c < 4; ++c) {
    mnx = min(mnx, q.x[c]);
    mxx = max(mxx, q.x[c]);
    mny = min(mny, q.y[c]);
    mxy = max(mxy, q.y[c]);
  }
  // samples sit at 32..224 / 256 inside a pixel
  Span s;
  s.x0 = max(0, -((224 - mnx) >> 8));
  s.x1 = min((int)f.w - 1, (mxx - 32) >> 8);
  s.y0 = max(0, -((224 - mny) >> 8));
  s.y1 = min((int)f.h - 1, (mxy - 32) >> 8);
  return s;
}

// x-extent of edge (a, b) inside the horizontal band [yl, yh] (fixed point,
// double precision), merged into [lo, hi].
__device__ __forceinline__ void edge_band(const Quad& q, int a, int b, double yl, double yh,
                                          double& lo, double& hi) {
  const double x0 = q.x[a], y0 = q.y[a], x1 = q.x[b], y1 = q.y[b];
  const double ya = fmin(y0, y1), yb = fmax(y0, y1);
  if (yb < yl || ya > yh) return;
  if (y0 == y1) {
    lo = fmin(lo, fmin(x0, x1));
    hi = fmax(hi, fmax(x0, x1));
    return;
  }
  const double s = (x1 - x0) / (y1 - y0);
  const double xa = x0 + (fmax(ya, yl) - y0) * s, xb = x0 + (fmin(yb, yh) - y0) * s;
  lo = fmin(lo, fmin(xa, xb));
  hi = fmax(hi, fmax(xa, xb));
}

// A segment's quad, triangles and pixel span, set up once per work item.
struct SegSetup {
  Quad q;
  Tri t0, t1;
  Span s;
  bool ok;
};

__device__ __forceinline__ SegSetup seg_setup(const Frame& f, uint32_t i) {
  SegSetup g;
  g.ok = seg_quad(f, i, g.q);
  if (!g.ok) return g;
  tri_setup(g.q.x[0], g.q.y[0], g.q.x[1], g.q.y[1], g.q.x[2], g.q.y[2], g.t0);
  tri_setup(g.q.x[2], g.q.y[2], g.q.x[1], g.q.y[1], g.q.x[3], g.q.y[3], g.t1);
  g.s = quad_span(f, g.q);
  g.ok = g.s.x0 <= g.s.x1 && g.s.y0 <= g.s.y1;
  return g;
}

// Pixel rows of segment i's quad inside the image (0 if none).
__device__ __forceinline__ uint32_t seg_rows(const SegSetup& g) {
  return g.ok ? (uint32_t)(g.s.y1 - g.s.y0 + 1) : 0u;
}

// Walks the fragments of one pixel row py of a segment in draw order
// (left to right); a pixel that both triangles cover on a common sample gets
// two fragments (first triangle first), otherwise one with both masks. The
// row only visits the columns between the leftmost and rightmost points of
// the triangles' edges inside its sample band (a closed triangle meets a
// horizontal line between two of its edge points), widened by 4/256 px
// against rounding; the exact per-sample test decides coverage, so the bound
// changes work, not results.
template <class F>
__device__ void walk_row(const Frame& f, const SegSetup& g, int py, F&& emit) {
  const double yl = (double)py * 256.0 + 32.0, yh = (double)py * 256.0 + 224.0;
  double lo = 1.0e300, hi = -1.0e300;
  edge_band(g.q, 0, 1, yl, yh, lo, hi);  // the edges of (c0 c1 c2) and (c2 c1 c3)
  edge_band(g.q, 1, 2, yl, yh, lo, hi);
  edge_band(g.q, 2, 0, yl, yh, lo, hi);
  edge_band(g.q, 1, 3, yl, yh, lo, hi);
  edge_band(g.q, 3, 2, yl, yh, lo, hi);
  if (lo > hi) return;
  const int x0 = max(g.s.x0, (int)ceil((lo - 4.0 - 224.0) / 256.0));
  const int x1 = min(g.s.x1, (int)floor((hi + 4.0 - 32.0) / 256.0));
  if (x0 > x1) return;
  // Edge functions E = dx * (sy - ay) - dy * (sx - ax) of both triangles at
  // the (up to 4) samples of pixel x0, stepped one pixel right by
  // subtracting 256 * dy: exact integers, so a sample is inside (E >= bias
  // on all three edges of a live triangle) exactly when the direct
  // evaluation says so.
  long long E[2][3][4], step[2][3];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const Tri& T = t ? g.t1 : g.t0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      step[t][e] = 256ll * T.dy[e];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        int sx = 0, sy = 0;
        if ((uint32_t)k < f.msaa) sample_pos(f.msaa, (uint32_t)k, sx, sy);
        sx += x0 * 256;
        sy += py * 256;
        E[t][e][k] = (long long)T.dx[e] * (sy - T.ay[e]) - (long long)T.dy[e] * (sx - T.ax[e]);
      }
    }
  }
  for (int px = x0; px <= x1; ++px) {
    uint32_t m0 = 0u, m1 = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if ((uint32_t)k >= f.msaa) break;
      const bool in0 = g.t0.live && E[0][0][k] >= g.t0.bias[0] && E[0][1][k] >= g.t0.bias[1] &&
                       E[0][2][k] >= g.t0.bias[2];
      const bool in1 = g.t1.live && E[1][0][k] >= g.t1.bias[0] && E[1][1][k] >= g.t1.bias[1] &&
                       E[1][2][k] >= g.t1.bias[2];
      m0 |= in0 ? 1u << k : 0u;
      m1 |= in1 ? 1u << k : 0u;
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int k = 0; k < 4; ++k) E[t][e][k] -= step[t][e];
    const uint32_t pix = (uint32_t)py * f.w + (uint32_t)px;
    if (m0 & m1) {
      emit(pix, m0);
      emit(pix, m1);
    } else if (m0 | m1) {
      emit(pix, m0 | m1);
    }
  }
}

// Work items are (segment, pixel row) pairs, so a segment hundreds of pixels
// long (bundling leaves such segments next to fixed endpoints) is spread over
// as many threads as it has rows instead of holding one lane of a wave.
// rs_rows: rows per segment (entry n = 0, for the scan).
__global__ __launch_bounds__(256) void rs_rows(Frame f, uint32_t* rows) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i == 0) rows[f.n] = 0;
  if (i >= f.n) return;
  rows[i] = seg_rows(seg_setup(f, i));
}

// rs_row_seg: the segment of every row item (row_off = exclusive scan of rows).
__global__ __launch_bounds__(256) void rs_row_seg(Frame f, const uint32_t* rows,
                                                  const uint32_t* row_off, uint32_t* row_seg) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= f.n) return;
  const uint32_t o = row_off[i], k = rows[i];
  for (uint32_t r = 0; r < k; ++r) row_seg[o + r] = i;
}

// rs_count: fragments per row item (entry R = 0, for the scan).
__global__ __launch_bounds__(256) void rs_count(Frame f, uint32_t n_items, const uint32_t* row_seg,
                                                const uint32_t* row_off, uint32_t* cnt) {
  const uint32_t r = blockIdx.x * 256u + threadIdx.x;
  if (r == 0) cnt[n_items] = 0;
  if (r >= n_items) return;
  const uint32_t i = row_seg[r];
  const SegSetup g = seg_setup(f, i);
  uint32_t c = 0;
  walk_row(f, g, g.s.y0 + (int)(r - row_off[i]), [&](uint32_t, uint32_t) { ++c; });
  cnt[r] = c;
}

__global__ __launch_bounds__(256) void rs_emit(Frame f, uint32_t n_items, const uint32_t* row_seg,
                                               const uint32_t* row_off, const uint32_t* off,
                                               uint32_t* keys, uint32_t* vals) {
  const uint32_t r = blockIdx.x * 256u + threadIdx.x;
  if (r >= n_items) return;
  const uint32_t i = row_seg[r];
  const SegSetup g = seg_setup(f, i);
  uint32_t o = off[r];
  walk_row(f, g, g.s.y0 + (int)(r - row_off[i]), [&](uint32_t pix, uint32_t mask) {
    keys[o] = pix | (mask << kMaxPixelBits);
    vals[o] = i;
    ++o;
  });
}

__global__ __launch_bounds__(256) void rs_runs(const uint32_t* keys, uint32_t m, uint32_t* first,
                                               uint32_t* last) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= m) return;
  const uint32_t pm = (1u << kMaxPixelBits) - 1u;
  const uint32_t p = keys[j] & pm;
  if (j == 0 || (keys[j - 1] & pm) != p) first[p] = j;
  if (j + 1 == m || (keys[j + 1] & pm) != p) last[p] = j + 1;
}

// Endpoint outline (bit k) and dot (bit 4 + k) coverage of sample k:
// vs_endpoint's square of +-3 px around the endpoint, discarded outside the
// unit circle (outline) or the (1 - 1/3) circle (dot), render_paths.wgsl:73-97.
__global__ __launch_bounds__(256) void rs_dots(Frame f, const uint32_t* paths, uint32_t n_paths,
                                               uint32_t* dots) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t path = t >> 1;
  if (path >= n_paths) return;
  const uint32_t count = paths[3u * path + 2u];
  if (count < 2u) return;
  const uint32_t idx = paths[3u * path] + ((t & 1u) ? count - 1u : 0u);
  const float cx = (f.pts[3ull * idx] * 0.5f + 0.5f) * (float)f.w;
  const float cy = (0.5f - f.pts[3ull * idx + 1] * 0.5f) * (float)f.h;
  if (!(fabsf(cx) < 1.0e6f && fabsf(cy) < 1.0e6f)) return;
  const float inner = 1.0f - kDotOutline / kDotRadius;
  const float inner2 = inner * inner;
  const int x0 = max(0, (int)floorf(cx - kDotRadius) - 1);
  const int x1 = min((int)f.w - 1, (int)floorf(cx + kDotRadius) + 1);
  const int y0 = max(0, (int)floorf(cy - kDotRadius) - 1);
  const int y1 = min((int)f.h - 1, (int)floorf(cy + kDotRadius) + 1);
  for (int py = y0; py <= y1; ++py)
    for (int px = x0; px <= x1; ++px) {
      uint32_t bits = 0;
      for (uint32_t k = 0; k < f.msaa; ++k) {
        int sx, sy;
        sample_pos(f.msaa, k, sx, sy);
        const float ox = ((float)px + (float)sx * (1.0f / 256.0f) - cx) / kDotRadius;
        const float oy = ((float)py + (float)sy * (1.0f / 256.0f) - cy) / kDotRadius;
        if (!(ox >= -1.0f && ox < 1.0f && oy >= -1.0f && oy < 1.0f)) continue;
        const float d2 = ox * ox + oy * oy;
        if (d2 <= 1.0f) bits |= 1u << k;
        if (d2 <= inner2) bits |= 1u << (4u + k);
      }
      if (bits) atomicOr(&dots[(uint32_t)py * f.w + (uint32_t)px], bits);
    }
}

// fs_main's color (render_paths.wgsl:53-66): value -> 15 palette intervals.
__device__ __forceinline__ void palette_rgb(uint32_t pal, float value, float rgb[3]) {
  const float stop = value * 15.0f;
  const uint32_t idx = min((uint32_t)fmaxf(stop, 0.0f), 14u);
  const float t = stop - (float)idx;
#pragma unroll
  for (int c = 0; c < 3; ++c)
    rgb[c] = kPalette[pal][idx][c] * (1.0f - t) + kPalette[pal][idx + 1u][c] * t;
}

// The density texture (rgba16f, R channel) sampled bilinearly at a pixel's
// center with clamp-to-edge; weights quantized to 1/256 as texture units do.
__device__ float density_at(const Frame& f, uint32_t px, uint32_t py) {
  const float tx = ((float)px + 0.5f) / (float)f.w * (float)(f.dw - 1u);
  const float ty = ((float)py + 0.5f) / (float)f.h * (float)(f.dh - 1u);
  const float fx = floorf(tx), fy = floorf(ty);
  const float ax = rintf((tx - fx) * 256.0f) * (1.0f / 256.0f);
  const float ay = rintf((ty - fy) * 256.0f) * (1.0f / 256.0f);
  const uint32_t x0 = min((uint32_t)fx, f.dw - 1u), x1 = min(x0 + 1u, f.dw - 1u);
  const uint32_t y0 = min((uint32_t)fy, f.dh - 1u), y1 = min(y0 + 1u, f.dh - 1u);
  auto h = [&](uint32_t x, uint32_t y) {
    return __half2float(__float2half_rn(f.dens[(size_t)y * f.dw + x]));
  };
  const float top = h(x0, y0) * (1.0f - ax) + h(x1, y0) * ax;
  const float bot = h(x0, y1) * (1.0f - ax) + h(x1, y1) * ax;
  const float v = top * (1.0f - ay) + bot * ay;
  return fminf(fmaxf(v, 0.0f), 1.0f);
}

// Density colouring gives every fragment of a pixel the pixel's colour, so a
// sample's stored value after c blends is the blend applied c times to the
// clear colour, in whatever order: the resolve then needs only per-sample
// fragment counts. The fragments are sorted by pixel, so a wave's fragments
// of one pixel are adjacent lanes: the first lane of each such run adds the
// run's per-sample counts (one atomic per run and sample).
__global__ __launch_bounds__(256) void rs_sample_counts(const uint32_t* keys, uint32_t m,
                                                        uint32_t msaa, uint32_t* scount) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  const bool valid = j < m;
  const uint32_t key = valid ? keys[j] : 0u;
  const uint32_t pm = (1u << kMaxPixelBits) - 1u;
  const uint32_t p = key & pm, mask = key >> kMaxPixelBits;
  const uint32_t prev = __shfl_up(p, 1, 64);
  const bool head = valid && (lane == 0u || prev != p);
  const unsigned long long heads = __ballot(head);
  unsigned long long bits[4];
#pragma unroll
  for (uint32_t k = 0; k < 4u; ++k) bits[k] = __ballot(valid && ((mask >> k) & 1u));
  if (head) {
    const unsigned long long above = lane == 63u ? 0ull : heads & (~0ull << (lane + 1u));
    const unsigned long long upto = above ? (1ull << __builtin_ctzll(above)) - 1ull : ~0ull;
    const unsigned long long run = upto & (~0ull << lane);
    for (uint32_t k = 0; k < msaa; ++k) {
      const uint32_t c = (uint32_t)__popcll(bits[k] & run);
      if (c) atomicAdd(&scount[4u * p + k], c);
    }
  }
}

__global__ __launch_bounds__(256) void rs_resolve(Frame f, const uint32_t* keys,
                                                  const uint32_t* vals, const uint32_t* first,
                                                  const uint32_t* last, const uint32_t* scount,
                                                  const uint32_t* dots, uint32_t* out) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= f.w * f.h) return;
  const uint32_t px = p % f.w, py = p / f.w;
  uint32_t v[4][3];
  const uint32_t bg[3] = {unorm8(kClear[0]), unorm8(kClear[1]), unorm8(kClear[2])};
  for (int k = 0; k < 4; ++k)
    for (int c = 0; c < 3; ++c) v[k][c] = bg[c];
  if (scount) {
    // density colouring: c blends of the pixel's colour per sample; the
    // blend is monotone in the stored value, so it stops changing after at
    // most 255 applications and the loop ends there
    const uint4 cnt = reinterpret_cast<const uint4*>(scount)[p];
    const uint32_t ck[4] = {cnt.x, cnt.y, cnt.z, cnt.w};
    if (cnt.x | cnt.y | cnt.z | cnt.w) {
      float col[3];
      palette_rgb(f.palette, density_at(f, px, py), col);
      for (uint32_t k = 0; k < f.msaa; ++k)
        for (uint32_t t = 0; t < ck[k]; ++t) {
          uint32_t same = 1u;
          for (int c = 0; c < 3; ++c) {
            const uint32_t nv = blend8(v[k][c], col[c]);
            same &= nv == v[k][c];
            v[k][c] = nv;
          }
          if (same) break;
        }
    }
  }
  const uint32_t s0 = first ? first[p] : 0u, s1 = last ? last[p] : 0u;
  if (!scount && s1 > s0) {
    const bool by_edge = f.color_by == BT_COLOR_EDGE_ID;
    float col[3] = {0.0f, 0.0f, 0.0f};
    if (!by_edge) palette_rgb(f.palette, density_at(f, px, py), col);
    const float denom = (float)(max(f.n_edges, 2u) - 1u);
    uint32_t settled = 0;  // density colouring: samples a further blend leaves unchanged
    for (uint32_t j = s0; j < s1; ++j) {
      const uint32_t m = keys[j] >> kMaxPixelBits;
      if (by_edge) {
        const uint32_t id = __float_as_uint(f.pts[3ull * vals[j] + 2]);
        palette_rgb(f.palette, (float)id / denom, col);
      }
      for (uint32_t k = 0; k < f.msaa; ++k) {
        if (!((m >> k) & 1u)) continue;
        uint32_t same = 1u;
        for (int c = 0; c < 3; ++c) {
          const uint32_t nv = blend8(v[k][c], col[c]);
          same &= nv == v[k][c];
          v[k][c] = nv;
        }
        if (same) settled |= 1u << k;
      }
      // with one colour per pixel, a settled sample stays settled
      if (!by_edge && settled == (1u << f.msaa) - 1u) break;
    }
  }
  const uint32_t d = dots ? dots[p] : 0u;
  if (d) {
    uint32_t ring[3], dot[3];
    for (int c = 0; c < 3; ++c) {
      ring[c] = unorm8(kPalette[f.palette][0][c] * 0.1f);
      dot[c] = unorm8(kEndpoint[f.palette][c]);
    }
    for (uint32_t k = 0; k < f.msaa; ++k) {
      const uint32_t* src = ((d >> (4u + k)) & 1u) ? dot : ((d >> k) & 1u) ? ring : nullptr;
      if (src)
        for (int c = 0; c < 3; ++c) v[k][c] = src[c];
    }
  }
  uint32_t rgba = 0xff000000u;
  for (int c = 0; c < 3; ++c) {
    uint32_t sum = 0;
    for (uint32_t k = 0; k < f.msaa; ++k) sum += v[k][c];
    rgba |= ((sum + f.msaa / 2u) / f.msaa) << (8 * c);
  }
  out[p] = rgba;
}

uint32_t blocks(uint64_t n) { return (uint32_t)((n + 255u) / 256u); }

// ------------------------------------------------------------------ scan
// Exclusive scan of n u32 counts into u32 offsets, the total in 64 bits.
constexpr uint32_t kScanTile = 4096;  // items per workgroup (256 threads x 16)

__device__ __forceinline__ unsigned long long rs_wave_sum(unsigned long long v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ unsigned long long rs_wave_incl(unsigned long long v) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long t = __shfl_up(v, d, 64);
    if ((int)lane >= d) v += t;
  }
  return v;
}

__global__ __launch_bounds__(256) void rs_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                      unsigned long long* __restrict__ bsum) {
  __shared__ unsigned long long part[4];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  unsigned long long s = 0;
  for (uint32_t k = threadIdx.x; k < kScanTile; k += 256u)
    if (base + k < n) s += in[base + k];
  s = rs_wave_sum(s);
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) bsum[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// One workgroup: exclusive scan of the nb tile sums in place, the total to
// *total.
__global__ __launch_bounds__(1024) void rs_scan_bsum(unsigned long long* __restrict__ bsum,
                                                     uint32_t nb,
                                                     unsigned long long* __restrict__ total) {
  __shared__ unsigned long long wtot[16];
  __shared__ unsigned long long carry;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < nb; c0 += 1024u) {
    const uint32_t i = c0 + threadIdx.x;
    const unsigned long long v = i < nb ? bsum[i] : 0ull;
    const unsigned long long incl = rs_wave_incl(v);
    if (lane == 63u) wtot[w] = incl;
    __syncthreads();
    unsigned long long before = carry;
    for (uint32_t k = 0; k < w; ++k) before += wtot[k];
    if (i < nb) bsum[i] = before + incl - v;
    __syncthreads();  // every thread has read carry and wtot
    if (threadIdx.x == 0) {
      unsigned long long t = carry;
      for (uint32_t k = 0; k < 16u; ++k) t += wtot[k];
      carry = t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Each thread scans 16 consecutive items; the threads' sums are scanned
// across the workgroup and offset by the tile's scanned sum.
__global__ __launch_bounds__(256) void rs_scan_apply(const uint32_t* __restrict__ in,
                                                     uint32_t* __restrict__ out, uint64_t n,
                                                     const unsigned long long* __restrict__ bsum) {
  __shared__ unsigned long long wtot[4];
  constexpr uint32_t K = kScanTile / 256u;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint64_t first = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * K;
  uint32_t v[K];
  unsigned long long s = 0;
#pragma unroll
  for (uint32_t k = 0; k < K; ++k) {
    v[k] = first + k < n ? in[first + k] : 0u;
    s += v[k];
  }
  const unsigned long long incl = rs_wave_incl(s);
  if (lane == 63u) wtot[w] = incl;
  __syncthreads();
  unsigned long long at = bsum[blockIdx.x] + incl - s;
  for (uint32_t k = 0; k < w; ++k) at += wtot[k];
#pragma unroll
  for (uint32_t k = 0; k < K; ++k) {
    if (first + k < n) out[first + k] = (uint32_t)at;
    at += v[k];
  }
}

// ------------------------------------------------------------------ sort
constexpr uint32_t kSortTile = 4096;   // keys per workgroup per pass (256 threads x 16)
constexpr uint32_t kDigitBits = 10;    // bits per pass: 2 passes for a 2^20-pixel target
constexpr uint32_t kDigits = 1u << kDigitBits;

// Digit counts of each tile, written digit-major (hist[d * nb + tile]) so
// that one exclusive scan gives every (digit, tile) its output offset; the
// entry after the last is zeroed for the scan's total.
__global__ __launch_bounds__(256) void rs_sort_hist(const uint32_t* __restrict__ keys, uint32_t m,
                                                    uint32_t shift, uint32_t dmask, uint32_t nb,
                                                    uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kDigits];
  for (uint32_t d = threadIdx.x; d < kDigits; d += 256u) h[d] = 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (uint32_t k = threadIdx.x; k < kSortTile; k += 256u) {
    const uint32_t i = base + k;
    if (i < m) atomicAdd(&h[(keys[i] >> shift) & dmask], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kDigits; d += 256u) hist[d * nb + blockIdx.x] = h[d];
  if (blockIdx.x == 0 && threadIdx.x == 0) hist[kDigits * nb] = 0u;
}

// Stable scatter of one pass: the tile in rounds of 256 keys (wave w takes
// keys [64 w, 64 w + 64) of the round, in order). Within a wave the lanes
// holding the same digit are found with one ballot per digit bit and ranked
// by lane; the lowest of them (the leader) publishes the wave's count of
// that digit; a key's position is its digit's base plus the counts of the
// waves before its own plus its rank; then each leader advances its
// digit's base and clears its count.
__global__ __launch_bounds__(256) void rs_sort_scatter(const uint32_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin,
                                                       uint32_t* __restrict__ kout,
                                                       uint32_t* __restrict__ vout, uint32_t m,
                                                       uint32_t shift, uint32_t dmask,
                                                       uint32_t nb,
                                                       const uint32_t* __restrict__ offs) {
  __shared__ uint32_t base[kDigits];
  __shared__ uint32_t cnt[4][kDigits];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  for (uint32_t d = threadIdx.x; d < kDigits; d += 256u) {
    base[d] = offs[d * nb + blockIdx.x];
    cnt[0][d] = cnt[1][d] = cnt[2][d] = cnt[3][d] = 0u;
  }
  __syncthreads();
  const uint32_t tile = blockIdx.x * kSortTile;
  for (uint32_t r = 0; r < kSortTile / 256u; ++r) {
    const uint32_t i = tile + r * 256u + w * 64u + lane;
    const bool valid = i < m;
    const uint32_t key = valid ? kin[i] : 0u, val = valid ? vin[i] : 0u;
    const uint32_t d = (key >> shift) & dmask;
    uint64_t peers = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
    for (uint32_t b = 0; b < kDigitBits; ++b) {
      const uint64_t ones = __builtin_amdgcn_ballot_w64(((d >> b) & 1u) != 0u);
      peers &= ((d >> b) & 1u) ? ones : ~ones;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    const bool leader = valid && rank == 0u;
    if (leader) cnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = base[d] + rank;
      for (uint32_t k = 0; k < w; ++k) pos += cnt[k][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    if (leader) {
      atomicAdd(&base[d], (uint32_t)__popcll(peers));
      cnt[w][d] = 0u;
    }
    __syncthreads();
  }
}

}  // namespace

void Engine::render_size(uint32_t* w, uint32_t* h, uint32_t job_i) {
  if (!jobs_.empty()) {
    const Job& j = job(job_i);
    *w = j.dw;
    *h = j.dh;
    return;
  }
  bt_iter_params p;
  iter_params(opts_, 0, 1, &p);
  *w = p.density_w;
  *h = p.density_h;
}

void Engine::render(const bt_render_options& ro, uint8_t* rgba, size_t n_bytes, uint32_t job_i) {
  if (ro.msaa != 2u && ro.msaa != 4u) fail(BT_E_INVALID, "MSAA must be 2 or 4 samples");
  if (ro.palette > BT_PALETTE_BATLOW) fail(BT_E_INVALID, "unknown palette");
  if (ro.color_by > BT_COLOR_EDGE_ID) fail(BT_E_INVALID, "unknown color_by");
  Frame f{};
  render_size(&f.w, &f.h, job_i);
  if (f.w == 0 || f.h == 0) fail(BT_E_INVALID, "image dimensions must be nonzero");
  if ((uint64_t)f.w * f.h > (1ull << kMaxPixelBits))
    fail(BT_E_CAPACITY, "render target exceeds 2^24 pixels");
  if (n_bytes != (size_t)f.w * f.h * 4) fail(BT_E_INVALID, "image buffer size mismatch");
  f.msaa = ro.msaa;
  f.palette = ro.palette;
  f.color_by = ro.color_by;
  const size_t npix = (size_t)f.w * f.h;
  DevBuf image, plan, cnt, off, keys, vals, keys2, vals2, first, last, scount, dots, temp;
  for (DevBuf* b : {&image, &plan, &cnt, &off, &keys, &vals, &keys2, &vals2, &first, &last,
                    &scount, &dots, &temp})
    b->pool = pool_.get();
  image.alloc(npix * 4);
  const uint32_t *kkeys = nullptr, *kvals = nullptr, *kfirst = nullptr, *klast = nullptr,
                 *kscount = nullptr, *kdots = nullptr;
  if (!jobs_.empty() && job(job_i).n > 0) {
    Job& j = job(job_i);
    if (ro.color_by == BT_COLOR_DENSITY && !j.dens_norm.p)
      fail(BT_E_STATE, "density is disabled (color_by != density)");
    if (ro.color_by == BT_COLOR_DENSITY) ensure_density(j);
    f.n = j.n;
    f.n_edges = j.n_paths;
    f.dens = j.dens_norm.as<float>();
    f.dw = j.dw;
    f.dh = j.dh;
    // the current samples in Layout order
    staging_.reserve((size_t)j.n * 12);
    std::vector<Job*> js = {&j};
    Plan p = make_plan(js, {j.iteration});
    plan.alloc(p.host.size());
    put_plan(p, plan.p);
    const unsigned char* dev = plan.as<unsigned char>();
    btk::unpermute(stream_, p.grid[BT_K_BIN], reinterpret_cast<const BtJob*>(dev),
                   reinterpret_cast<const uint32_t*>(dev + p.base_off) + BT_K_BIN * 2, 1,
                   staging_.as<float>(), nullptr, 0u);
    HIPCHK(hipGetLastError());
    f.pts = staging_.as<float>();

    // Exclusive scan of `count` + 1 u32 entries (the last one 0) into out;
    // returns the total, summed in 64 bits: the 32-bit offsets (and the
    // sort's int count) hold at most 2^31 - 1 items.
    DevBuf total;
    total.pool = pool_.get();
    total.alloc(8);
    auto scan_launch = [&](const uint32_t* in, uint32_t* out, uint64_t items) {
      const uint64_t nb = (items + kScanTile - 1) / kScanTile;
      if (nb > 0xffffffffull) fail(BT_E_CAPACITY, "scan exceeds 2^32 tiles");
      temp.reserve(std::max<size_t>(nb * 8, 256));
      unsigned long long* bsum = temp.as<unsigned long long>();
      hipLaunchKernelGGL(rs_scan_reduce, dim3((uint32_t)nb), dim3(256), 0, stream_, in, items,
                         bsum);
      hipLaunchKernelGGL(rs_scan_bsum, dim3(1), dim3(1024), 0, stream_, bsum, (uint32_t)nb,
                         total.as<unsigned long long>());
      hipLaunchKernelGGL(rs_scan_apply, dim3((uint32_t)nb), dim3(256), 0, stream_, in, out,
                         items, bsum);
      HIPCHK(hipGetLastError());
    };
    auto scan = [&](const uint32_t* in, uint32_t* out, uint64_t count) -> uint64_t {
      scan_launch(in, out, count + 1);
      uint64_t t = 0;
      HIPCHK(hipMemcpyAsync(&t, total.p, 8, hipMemcpyDeviceToHost, stream_));
      sync();
      return t;
    };
    // rows per segment -> row items (segment order, then top to bottom)
    cnt.alloc(((size_t)f.n + 1) * 4);
    off.alloc(((size_t)f.n + 1) * 4);
    hipLaunchKernelGGL(rs_rows, dim3(blocks(f.n)), dim3(256), 0, stream_, f,
                       cnt.as<uint32_t>());
    HIPCHK(hipGetLastError());
    const uint64_t n_rows = scan(cnt.as<uint32_t>(), off.as<uint32_t>(), f.n);
    if (n_rows > 0x7fffffffull) fail(BT_E_CAPACITY, "frame exceeds 2^31 - 1 segment rows");
    const uint32_t R = (uint32_t)n_rows;
    DevBuf row_seg, fcnt, foff;
    for (DevBuf* b : {&row_seg, &fcnt, &foff}) b->pool = pool_.get();
    row_seg.alloc(std::max<size_t>(R, 1) * 4);
    fcnt.alloc(((size_t)R + 1) * 4);
    foff.alloc(((size_t)R + 1) * 4);
    hipLaunchKernelGGL(rs_row_seg, dim3(blocks(f.n)), dim3(256), 0, stream_, f,
                       cnt.as<uint32_t>(), off.as<uint32_t>(), row_seg.as<uint32_t>());
    HIPCHK(hipGetLastError());
    // fragments per row item -> offsets (draw order) -> total
    hipLaunchKernelGGL(rs_count, dim3(blocks(R + 1ull)), dim3(256), 0, stream_, f, R,
                       row_seg.as<uint32_t>(), off.as<uint32_t>(), fcnt.as<uint32_t>());
    HIPCHK(hipGetLastError());
    const uint64_t m64 = scan(fcnt.as<uint32_t>(), foff.as<uint32_t>(), R);
    if (m64 > 0x7fffffffull) fail(BT_E_CAPACITY, "frame exceeds 2^31 - 1 fragments");
    const uint32_t m = (uint32_t)m64;
    if (m > 0) {
      keys.alloc((size_t)m * 4);
      vals.alloc((size_t)m * 4);
      keys2.alloc((size_t)m * 4);
      vals2.alloc((size_t)m * 4);
      hipLaunchKernelGGL(rs_emit, dim3(blocks(R)), dim3(256), 0, stream_, f, R,
                         row_seg.as<uint32_t>(), off.as<uint32_t>(), foff.as<uint32_t>(),
                         keys.as<uint32_t>(), vals.as<uint32_t>());
      HIPCHK(hipGetLastError());
      // stable sort on the pixel bits: per pixel, fragments stay in draw order
      uint32_t bits = 1;
      while ((1ull << bits) < npix) ++bits;
      const uint32_t nb = (m + kSortTile - 1) / kSortTile;
      DevBuf hist, hoff;
      hist.pool = pool_.get();
      hoff.pool = pool_.get();
      hist.alloc(((size_t)nb * kDigits + 1) * 4);
      hoff.alloc(((size_t)nb * kDigits + 1) * 4);
      uint32_t *ka = keys.as<uint32_t>(), *va = vals.as<uint32_t>();
      uint32_t *kb = keys2.as<uint32_t>(), *vb = vals2.as<uint32_t>();
      for (uint32_t shift = 0; shift < bits; shift += kDigitBits) {
        // the last pass's digit stops at the pixel bits (the coverage mask above them is
        // not part of the key order)
        const uint32_t dmask = (1u << std::min(kDigitBits, bits - shift)) - 1u;
        hipLaunchKernelGGL(rs_sort_hist, dim3(nb), dim3(256), 0, stream_, ka, m, shift, dmask,
                           nb, hist.as<uint32_t>());
        scan_launch(hist.as<uint32_t>(), hoff.as<uint32_t>(), (uint64_t)nb * kDigits + 1);
        hipLaunchKernelGGL(rs_sort_scatter, dim3(nb), dim3(256), 0, stream_, ka, va, kb, vb, m,
                           shift, dmask, nb, hoff.as<uint32_t>());
        HIPCHK(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
      }
      kkeys = ka;
      kvals = va;
      if (ro.color_by == BT_COLOR_DENSITY) {
        scount.alloc(npix * 16);
        HIPCHK(hipMemsetAsync(scount.p, 0, npix * 16, stream_));
        hipLaunchKernelGGL(rs_sample_counts, dim3(blocks(m)), dim3(256), 0, stream_, kkeys, m,
                           f.msaa, scount.as<uint32_t>());
        HIPCHK(hipGetLastError());
        kscount = scount.as<uint32_t>();
      } else {
        first.alloc(npix * 4);
        last.alloc(npix * 4);
        HIPCHK(hipMemsetAsync(first.p, 0, npix * 4, stream_));
        HIPCHK(hipMemsetAsync(last.p, 0, npix * 4, stream_));
        hipLaunchKernelGGL(rs_runs, dim3(blocks(m)), dim3(256), 0, stream_, kkeys, m,
                           first.as<uint32_t>(), last.as<uint32_t>());
        HIPCHK(hipGetLastError());
        kfirst = first.as<uint32_t>();
        klast = last.as<uint32_t>();
      }
    }
    if (ro.endpoint_dots && j.n_paths > 0) {
      dots.alloc(npix * 4);
      HIPCHK(hipMemsetAsync(dots.p, 0, npix * 4, stream_));
      hipLaunchKernelGGL(rs_dots, dim3(blocks(2ull * j.n_paths)), dim3(256), 0, stream_, f,
                         j.paths.as<uint32_t>(), j.n_paths, dots.as<uint32_t>());
      HIPCHK(hipGetLastError());
      kdots = dots.as<uint32_t>();
    }
  }
  hipLaunchKernelGGL(rs_resolve, dim3(blocks(npix)), dim3(256), 0, stream_, f, kkeys, kvals,
                     kfirst, klast, kscount, kdots, image.as<uint32_t>());
  HIPCHK(hipGetLastError());
  // one DMA into pinned memory, then a host copy (a pageable destination
  // would be copied in staged chunks)
  void* pin = readback_staging(npix * 4);
  HIPCHK(hipMemcpyAsync(pin, image.p, npix * 4, hipMemcpyDeviceToHost, stream_));
  sync();
  std::memcpy(rgba, pin, npix * 4);
}

}  // namespace bt
