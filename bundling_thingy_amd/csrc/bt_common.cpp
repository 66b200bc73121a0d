// Host-side helpers: parameter model, TRL parsing, normalization, the
// reference's memory budget model and the f64 distortion metric.
#include "bt_common.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <strings.h>

#include "bt_device.h"

namespace bt {

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(BT_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

DevPool::~DevPool() {
  for (auto& kv : free_)
    if (own_.count(kv.second)) (void)hipFree(kv.second);
  for (const Slab& s : slabs_) (void)hipFree(s.base);
}

void DevPool::release_cached() {
  for (auto it = free_.begin(); it != free_.end();) {
    if (own_.erase(it->second)) {
      (void)hipFree(it->second);
      cached_ -= it->first;
      it = free_.erase(it);
    } else {
      ++it;
    }
  }
  // a slab whose carved bytes are all cached holds no live block
  std::vector<Slab> keep;
  for (size_t k = 0; k < slabs_.size(); ++k) {
    const Slab& s = slabs_[k];
    size_t in_cache = 0;
    for (const auto& kv : free_) {
      const char* q = static_cast<const char*>(kv.second);
      if (q >= s.base && q < s.base + kSlabBytes) in_cache += kv.first;
    }
    if (in_cache != s.carved || k + 1 == slabs_.size()) {
      keep.push_back(s);
      continue;
    }
    for (auto it = free_.begin(); it != free_.end();) {
      const char* q = static_cast<const char*>(it->second);
      if (q >= s.base && q < s.base + kSlabBytes) {
        cached_ -= it->first;
        it = free_.erase(it);
      } else {
        ++it;
      }
    }
    (void)hipFree(s.base);
  }
  slabs_.swap(keep);
}

void* DevPool::device_alloc(size_t bytes) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    // out of memory: return what the cache can give back and try once more
    (void)hipGetLastError();
    release_cached();
    e = hipMalloc(&q, bytes);
  }
  HIPCHK(e);
  return q;
}

size_t DevPool::block_size(size_t bytes) {
  size_t c = 256;
  if (bytes <= (1u << 20)) {
    while (c < bytes) c <<= 1;
  } else {
    c = (bytes + (1u << 20) - 1) >> 20 << 20;
  }
  return c;
}

void* DevPool::get(size_t bytes, size_t* cap) {
  const size_t c = block_size(bytes);
  auto it = free_.lower_bound(c);
  if (it != free_.end() && it->first <= c + c / 4) {
    void* q = it->second;
    *cap = it->first;
    cached_ -= it->first;
    free_.erase(it);
    return q;
  }
  if (c > kSlabBlockMax) {
    void* q = device_alloc(c);
    own_.emplace(q, c);
    *cap = c;
    return q;
  }
  if (slabs_.empty() || kSlabBytes - slabs_.back().carved < c) {
    if (!slabs_.empty()) {
      // the rest of the full slab goes to the cache as blocks of the size classes
      Slab& s = slabs_.back();
      while (kSlabBytes - s.carved >= 256) {
        const size_t left = kSlabBytes - s.carved;
        size_t piece = 256;
        if (left >= (1u << 20)) {
          piece = std::min(left >> 20 << 20, kSlabBlockMax);
        } else {
          while (piece * 2 <= left) piece <<= 1;
        }
        put(s.base + s.carved, piece);
        s.carved += piece;
      }
    }
    slabs_.push_back({static_cast<char*>(device_alloc(kSlabBytes)), 0});
  }
  Slab& s = slabs_.back();
  void* q = s.base + s.carved;
  s.carved += c;
  *cap = c;
  return q;
}

void DevPool::put(void* p, size_t cap) {
  free_.emplace(cap, p);
  cached_ += cap;
}

void DevBuf::alloc(size_t b) {
  release();
  if (b) {
    if (pool) {
      p = pool->get(b, &cap);
    } else {
      HIPCHK(hipMalloc(&p, b));
      cap = b;
    }
  }
  bytes = b;
}

bool DevBuf::reserve(size_t b) {
  if (b <= bytes && p) return false;
  alloc(std::max(b, (size_t)256));
  return true;
}

void DevBuf::release() {
  if (p) {
    if (pool) pool->put(p, cap);
    else (void)hipFree(p);
  }
  p = nullptr;
  bytes = 0;
  cap = 0;
}

bt_options resolve_options(const bt_options* in) {
  bt_options o;
  bt_options_default(&o);
  if (in) o = *in;
  if (o.map_w == 0) o.map_w = 1600;
  if (o.map_h == 0) o.map_h = o.map_w;
  if (o.max_graph_memory_bytes == 0) o.max_graph_memory_bytes = 1024ull << 20;
  if (o.spacing > 0.0f) {
    if (o.density_w == 0) o.density_w = 1024;
    if (o.density_h == 0) o.density_h = o.density_w;
    if (o.sigma0 <= 0.0f) o.sigma0 = 6.0f;
    if (o.sigma1 <= 0.0f) o.sigma1 = o.sigma0;
    if (o.radius1 <= 0.0f) o.radius1 = o.radius0;
    if (o.schedule_len == 0) o.schedule_len = 1;
  }
  return o;
}

bool has_radius(const bt_options& o) {
  return o.spacing > 0.0f ? o.radius0 > 0.0f : o.radius_texels > 0;
}

uint32_t schedule_len(const bt_options& o) {
  return o.spacing > 0.0f ? std::max(o.schedule_len, 1u) : 1u;
}

// Texel resolution of a generalized job: 1 / spacing, snapped to the nearest
// integer when within 1e-6 relative (so spacing = 1/(S-1) gives S-1 exactly).
static double resolution(float spacing) {
  double r = 1.0 / (double)spacing;
  double n = std::nearbyint(r);
  if (std::fabs(r - n) <= 1e-6 * r) r = n;
  return r;
}

static uint32_t grid_dim(float res, float cell) {
  double g = std::ceil(((double)res + 1.0) / (double)cell);
  if (g < 1.0) g = 1.0;
  if (g > 65535.0) g = 65535.0;
  return (uint32_t)g;
}

uint32_t default_split(float radius) {
  const float k = std::floor(radius / 4.0f);
  return (uint32_t)std::min(8.0f, std::max(1.0f, k));
}

uint32_t density_split(float radius, uint64_t samples, float res_x, float res_y) {
  const double area = ((double)res_x + 1.0) * ((double)res_y + 1.0);
  const double rho = std::max((double)samples / area, 1e-3);
  const double h = std::max(2.9 / std::sqrt(rho), 0.25);
  const double k = std::nearbyint((double)radius / h);
  return (uint32_t)std::min<double>((double)kMaxSplit, std::max(1.0, k));
}

void iter_params(const bt_options& o, uint32_t it, uint32_t split, bt_iter_params* p) {
  float sigma;
  if (o.spacing > 0.0f) {
    const double res = resolution(o.spacing);
    const uint32_t len = schedule_len(o);
    const uint32_t i = std::min(it, len - 1);
    const double f = len > 1 ? (double)i / (double)(len - 1) : 0.0;
    const double rho = (double)o.radius0 * std::pow((double)o.radius1 / (double)o.radius0, f);
    sigma = (float)((double)o.sigma0 * std::pow((double)o.sigma1 / (double)o.sigma0, f));
    p->res_x = p->res_y = (float)res;
    p->radius = (float)(rho * res);
    p->density_w = o.density_w;
    p->density_h = o.density_h;
  } else {
    sigma = 6.0f;
    p->res_x = (float)(o.map_w - 1);
    p->res_y = (float)(o.map_h - 1);
    p->radius = (float)o.radius_texels;
    p->density_w = o.map_w;
    p->density_h = o.map_h;
  }
  p->sigma = sigma;
  float w[BT_MAX_TAPS + 1];
  p->taps = blur_weights(sigma, w, BT_MAX_TAPS);
  if (p->radius > 0.0f) {
    if (split == 0) split = default_split(p->radius);
    // fewer rows per radius while even square cells exceed kMaxGridCells
    while (split > 1) {
      const float c = p->radius / (float)split;
      if ((uint64_t)grid_dim(p->res_x, c) * grid_dim(p->res_y, c) <= kMaxGridCells) break;
      --split;
    }
    p->cell = split == 1 ? p->radius : p->radius / (float)split;
    p->gy = grid_dim(p->res_y, p->cell);
    // Rows keep the cell height; bins along x are x_split times narrower, so
    // a row's candidate ranges end within a fraction of a cell of the disk.
    // Capped so a job's grid stays within kMaxGridCells bins.
    uint32_t a = std::min(o.x_split ? o.x_split : kDefaultXSplit, 64u);
    const uint64_t g1 = (uint64_t)grid_dim(p->res_x, p->cell) * p->gy;
    while (a > 1 && g1 * a > kMaxGridCells) a >>= 1;
    p->cell_w = a == 1 ? p->cell : p->cell / (float)a;
    p->gx = grid_dim(p->res_x, p->cell_w);
    // Task blocks: runs of bins one radius wide, cut into chunks of up to 64
    // consecutive samples (wider blocks fill more lanes; lane groups keep the
    // query boxes narrow).
    p->task_bins = a * split;
  } else {
    p->cell = p->cell_w = 0.0f;
    p->gx = p->gy = 1;
    p->task_bins = 1;
  }
}

uint32_t blur_weights(float sigma, float* w, uint32_t max_taps) {
  uint32_t taps = (uint32_t)std::ceil(3.0 * (double)sigma);
  if (taps > max_taps) taps = max_taps;
  const double s2 = 2.0 * (double)sigma * (double)sigma;
  double z = 0.0;
  for (int k = -(int)taps; k <= (int)taps; ++k) z += std::exp(-(double)k * k / s2);
  for (uint32_t k = 0; k <= taps; ++k) {
    double v = std::exp(-(double)k * k / s2) / z;
    w[k] = (float)(std::nearbyint(v * 1e10) / 1e10);
  }
  return taps;
}

bool normalization(const float* bbox, float* out) {
  const float rx = bbox[2] - bbox[0], ry = bbox[3] - bbox[1];
  const float span = std::max(rx, ry);
  if (!(span > 0.0f) || !std::isfinite(span)) return false;
  const float scale = 0.95f / span;
  if (!std::isfinite(scale)) return false;
  out[0] = bbox[0];
  out[1] = bbox[1];
  out[2] = scale;
  out[3] = (1.0f - scale * rx) / 2.0f;
  out[4] = (1.0f - scale * ry) / 2.0f;
  return true;
}

namespace {

// resampling.wgsl:45-48
inline void normalized_at(const float* xy, const float* nm, uint32_t i, float& x, float& y) {
  x = ((xy[2u * i] - nm[0]) * nm[2] + nm[3] - 0.5f) * 2.0f;
  y = ((xy[2u * i + 1u] - nm[1]) * nm[2] + nm[4] - 0.5f) * 2.0f;
}

inline uint32_t steps_for(float distance, float remaining) {
  const float f = std::max(std::floor(distance - remaining) + 1.0f, 0.0f);
  return f >= 4294967296.0f ? 0xffffffffu : (uint32_t)f;
}

// One path: the walk of resampling.wgsl:82-134 counting pushes. A segment
// of D texels takes steps = max(floor(D - remaining) + 1, 0) samples at
// prev + delta * ((remaining + k) / D); only its first can repeat the last
// sample pushed, and only its last is compared with what follows.
uint32_t count_path(const float* xy, uint32_t s, uint32_t e, const float* nm, float tsx,
                    float tsy, uint32_t cap) {
  float fx, fy;
  normalized_at(xy, nm, s, fx, fy);
  uint32_t count = 0;
  float lx = fx, ly = fy;
  auto push = [&](float x, float y) {
    if (count >= cap) {
      count = cap + 1u;
    } else {
      lx = x;
      ly = y;
      ++count;
    }
  };
  push(fx, fy);
  float px = fx, py = fy, remaining = 1.0f;
  bool over = count > cap;
  for (uint32_t k = s + 1u; k < e && !over; ++k) {
    float cx, cy;
    normalized_at(xy, nm, k, cx, cy);
    const float dx = cx - px, dy = cy - py;
    const float ex = dx * tsx, ey = dy * tsy;
    const float distance = std::sqrt(std::fma(ex, ex, ey * ey));
    if (distance > 0.0f) {
      const uint32_t steps = steps_for(distance, remaining);
      if (steps > 0u) {
        const float t0 = remaining / distance;
        const float t1 = (remaining + (float)(steps - 1u)) / distance;
        const float sx = px + dx * t0, sy = py + dy * t0;
        const float ux = px + dx * t1, uy = py + dy * t1;
        const uint64_t c = (uint64_t)count + steps - ((sx == lx && sy == ly) ? 1u : 0u);
        lx = ux;
        ly = uy;
        if (c > cap) {
          count = cap + 1u;
          over = true;
        } else {
          count = (uint32_t)c;
        }
      }
      remaining = remaining + (float)steps - distance;
    }
    px = cx;
    py = cy;
  }
  if (!over) {
    // the last control point (resampling.wgsl:117-119)
    if (!(count > 0u && px == lx && py == ly)) push(px, py);
    if (count == 1u) {
      // a closed path that collapsed to one sample keeps its first distinct
      // point and its end (resampling.wgsl:124-134)
      for (uint32_t k = s + 1u; k < e; ++k) {
        float nx, ny;
        normalized_at(xy, nm, k, nx, ny);
        if (nx != fx || ny != fy) {
          if (!(nx == lx && ny == ly)) push(nx, ny);
          if (!(count > 0u && px == lx && py == ly)) push(px, py);
          break;
        }
      }
    }
  }
  return count;
}

}  // namespace

uint64_t resample_counts(const float* xy, const uint32_t* starts, uint32_t n_paths,
                         const float* norm, float ts_x, float ts_y, uint32_t cap,
                         uint64_t* path_scan) {
  uint64_t acc = 0;
  for (uint32_t i = 0; i < n_paths; ++i) {
    path_scan[i] = acc;
    const uint32_t c = count_path(xy, starts[i], starts[i + 1u], norm, ts_x, ts_y, cap);
    if (c >= 2u) acc += ((uint64_t)c << 32) | 1ull;
  }
  path_scan[n_paths] = acc;
  return acc;
}

// NeighborMean::grid_bytes (neighbor_mean.rs:176-183) for a cell count.
static uint64_t grid_bytes_for(uint64_t cells) { return cells * 8 + 8 + cells * 24 + 52; }

int capacity_model(const bt_options& o, uint32_t n_points, uint32_t n_starts, uint64_t* capacity,
                   uint64_t* reserved, std::string* err) {
  const uint64_t budget = o.max_graph_memory_bytes;
  if (budget == 0) {
    *err = "graph memory budget must be nonzero";
    return BT_E_CAPACITY;
  }
  uint64_t grid = 0;
  if (has_radius(o)) {
    uint64_t cells = 0;
    if (o.spacing > 0.0f) {
      for (uint32_t i = 0; i < schedule_len(o); ++i) {
        bt_iter_params p;
        iter_params(o, i, 1, &p);
        cells = std::max(cells, (uint64_t)p.gx * p.gy);
      }
    } else {
      cells = (uint64_t)((o.map_w + o.radius_texels - 1) / o.radius_texels) *
              ((o.map_h + o.radius_texels - 1) / o.radius_texels);
    }
    grid = grid_bytes_for(cells);
  }
  const uint64_t staging = (o.flags & BT_FLAG_DISTORTION) ? 4194300ull : 0ull;
  if (grid + staging >= budget) {
    *err = "neighbor grid and optional distortion staging exceed --max-graph-memory-mib; "
           "increase the budget or --neighbor-radius";
    return BT_E_CAPACITY;
  }
  const uint64_t paths = n_starts ? n_starts - 1ull : 0ull;
  const uint64_t fixed = 8ull * n_points + 4ull * n_starts + 12 * paths + 16 * paths + 8 * paths +
                         32 + 16 + 48 + 16 + grid + staging;
  if (fixed > budget) {
    *err = "control points, path metadata, and neighbor grid exceed --max-graph-memory-mib";
    return BT_E_CAPACITY;
  }
  const uint64_t sample_bytes = 16 + (has_radius(o) ? 32 : 0);
  const uint64_t limit = 0xffffffffull / 12;
  uint64_t cap = std::min((budget - fixed) / sample_bytes, limit);
  if (cap < 3) {
    *err = "graph memory budget has no room for resampled geometry";
    return BT_E_CAPACITY;
  }
  *capacity = cap;
  *reserved = fixed + cap * sample_bytes;
  return BT_OK;
}

// ------------------------------------------------------------------- TRL

// Rust's f32::from_str grammar (sign, inf/infinity/nan, digits '.' digits,
// exponent); anything else is an invalid coordinate.
static bool rust_float(const char* s, size_t n, float* out) {
  size_t i = 0;
  if (n == 0 || n >= 128) return false;
  if (s[0] == '+' || s[0] == '-') i = 1;
  if (i == n) return false;
  const size_t rest = n - i;
  const char* r = s + i;
  bool special = (rest == 3 && strncasecmp(r, "inf", 3) == 0) ||
                 (rest == 8 && strncasecmp(r, "infinity", 8) == 0) ||
                 (rest == 3 && strncasecmp(r, "nan", 3) == 0);
  if (!special) {
    size_t digits = 0;
    while (i < n && std::isdigit((unsigned char)s[i])) ++i, ++digits;
    if (i < n && s[i] == '.') {
      ++i;
      while (i < n && std::isdigit((unsigned char)s[i])) ++i, ++digits;
    }
    if (digits == 0) return false;
    if (i < n && (s[i] == 'e' || s[i] == 'E')) {
      ++i;
      if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
      size_t ed = 0;
      while (i < n && std::isdigit((unsigned char)s[i])) ++i, ++ed;
      if (ed == 0) return false;
    }
    if (i != n) return false;
  }
  char buf[128];
  std::memcpy(buf, s, n);
  buf[n] = 0;
  *out = std::strtof(buf, nullptr);
  return true;
}

int parse_trl(const char* text, size_t len, std::vector<float>& xy, std::vector<uint32_t>& starts,
              float* bbox, std::string* err) {
  xy.clear();
  starts.clear();
  float mn[2] = {3.40282347e38f, 3.40282347e38f}, mx[2] = {-3.40282347e38f, -3.40282347e38f};
  size_t pos = 0, line = 0;
  while (pos < len) {
    size_t eol = pos;
    while (eol < len && text[eol] != '\n') ++eol;
    size_t end = eol;
    if (end > pos && text[end - 1] == '\r') --end;
    size_t cur = pos;
    while (cur < end && text[cur] != ' ') ++cur;  // token 0 is skipped
    const size_t start = xy.size() / 2;
    while (cur < end) {
      size_t a = cur + 1, b = a;
      while (b < end && text[b] != ' ') ++b;
      cur = b;
      if (b == a || text[a] == '(' || text[a] == '<') continue;
      float x, y;
      if (!rust_float(text + a, b - a, &x)) {
        *err = "line " + std::to_string(line + 1) + " x: invalid coordinate";
        return BT_E_INVALID;
      }
      if (cur >= end) {
        *err = "line " + std::to_string(line + 1) + " y: unexpected end of TRL line";
        return BT_E_INVALID;
      }
      size_t ya = cur + 1, yb = ya;
      while (yb < end && text[yb] != ' ') ++yb;
      cur = yb;
      if (!rust_float(text + ya, yb - ya, &y)) {
        *err = "line " + std::to_string(line + 1) + " y: invalid coordinate";
        return BT_E_INVALID;
      }
      if (!std::isfinite(x) || !std::isfinite(y)) {
        *err = "non-finite TRL coordinate";
        return BT_E_INVALID;
      }
      const size_t np = xy.size() / 2;
      if (np > start && xy[2 * np - 2] == x && xy[2 * np - 1] == y) continue;
      xy.push_back(x);
      xy.push_back(y);
    }
    const size_t np = xy.size() / 2;
    if (np - start < 2) {
      xy.resize(2 * start);
    } else {
      starts.push_back((uint32_t)start);
      for (size_t i = start; i < np; ++i) {
        mn[0] = std::min(mn[0], xy[2 * i]);
        mn[1] = std::min(mn[1], xy[2 * i + 1]);
        mx[0] = std::max(mx[0], xy[2 * i]);
        mx[1] = std::max(mx[1], xy[2 * i + 1]);
      }
    }
    ++line;
    pos = eol + 1;
  }
  starts.push_back((uint32_t)(xy.size() / 2));
  if (xy.empty()) {
    *err = "graph has no drawable edges";
    return BT_E_INVALID;
  }
  bbox[0] = mn[0];
  bbox[1] = mn[1];
  bbox[2] = mx[0];
  bbox[3] = mx[1];
  return BT_OK;
}

// ------------------------------------------------------------ distortion

namespace {

struct Curve {
  const float* p;
  uint32_t n;
  double length;
};

double seg_len(const float* a, const float* b) {
  return std::hypot((double)b[0] - (double)a[0], (double)b[1] - (double)a[1]);
}

bool curve_new(const float* p, uint32_t n, Curve* c) {
  if (n == 0) return false;
  for (uint32_t i = 0; i < 2 * n; ++i)
    if (!std::isfinite(p[i])) return false;
  double len = 0.0;
  for (uint32_t i = 0; i + 1 < n; ++i) len += seg_len(p + 2 * i, p + 2 * i + 2);
  if (!std::isfinite(len)) return false;
  *c = Curve{p, n, len};
  return true;
}

// distortion.rs:52-121: a cursor over a curve's segments in normalized arc length.
struct Cursor {
  const Curve* c;
  uint32_t next = 0;
  double total, traversed = 0.0, sx, sy, ex, ey, st = 0.0, et = 0.0;
  explicit Cursor(const Curve* cv) : c(cv), total(cv->length) {
    sx = ex = (double)cv->p[0];
    sy = ey = (double)cv->p[1];
  }
  void advance(double t) {
    while (et <= t && et < 1.0) {
      if (next + 1 >= c->n) {
        sx = ex;
        sy = ey;
        st = et;
        et = 1.0;
        return;
      }
      const float* a = c->p + 2 * next;
      const float* b = a + 2;
      ++next;
      const double len = seg_len(a, b);
      if (len == 0.0) continue;
      sx = a[0];
      sy = a[1];
      ex = b[0];
      ey = b[1];
      st = et;
      traversed += len;
      et = (next + 1 >= c->n) ? 1.0 : traversed / total;
    }
  }
  void pos(double t, double* x, double* y) const {
    if (t >= et) { *x = ex; *y = ey; return; }
    if (t <= st) { *x = sx; *y = sy; return; }
    const double f = (t - st) / (et - st);
    *x = std::fma(f, ex - sx, sx);
    *y = std::fma(f, ey - sy, sy);
  }
};

struct Kahan {
  double value = 0.0, correction = 0.0;
  void add(double term) {
    const double adjusted = term - correction;
    const double next = value + adjusted;
    correction = (next - value) - adjusted;
    value = next;
  }
};

bool squared_integral(const Curve& o, const Curve& c, double* out) {
  Cursor a(&o), b(&c);
  Kahan sum;
  double t = 0.0;
  while (t < 1.0) {
    a.advance(t);
    b.advance(t);
    const double next = std::min(a.et, b.et);
    if (!(next > t && next <= 1.0)) return false;
    double px, py, qx, qy, rx, ry, sx, sy;
    a.pos(t, &px, &py);
    b.pos(t, &qx, &qy);
    a.pos(next, &rx, &ry);
    b.pos(next, &sx, &sy);
    const double ax = px - qx, ay = py - qy, bx = rx - sx, by = ry - sy;
    const double smx = ax + bx, smy = ay + by, dlx = bx - ax, dly = by - ay;
    const double mid2 = std::fma(smx, smx, smy * smy);
    const double ch2 = std::fma(dlx, dlx, dly * dly);
    sum.add((next - t) * (mid2 / 4.0 + ch2 / 12.0));
    t = next;
  }
  if (!(std::isfinite(sum.value) && sum.value >= 0.0)) return false;
  *out = sum.value;
  return true;
}

}  // namespace

int distortion(const float* original, const float* current, const uint32_t* rec3,
               uint32_t n_paths, uint32_t n_samples, uint32_t map_size, double* out,
               std::string* err) {
  Kahan sum;
  uint32_t edges = 0;
  for (uint32_t i = 0; i < n_paths; ++i) {
    const uint32_t off = rec3[3 * i], count = rec3[3 * i + 2];
    if (count == 0) continue;
    if (count < 2) {
      *err = "edge " + std::to_string(i) + " has only one sample";
      return BT_E_INVALID;
    }
    if ((uint64_t)off + count > n_samples) {
      *err = "geometry ranges have a gap, overlap or exceed active count";
      return BT_E_INVALID;
    }
    Curve a, b;
    if (!curve_new(original + 2ull * off, count, &a) || !curve_new(current + 2ull * off, count, &b)) {
      *err = "distortion curve contains a nonfinite coordinate";
      return BT_E_INVALID;
    }
    double v;
    if (!squared_integral(a, b, &v)) {
      *err = "invalid distortion interval";
      return BT_E_INVALID;
    }
    sum.add(v);
    ++edges;
  }
  if (edges == 0) {
    *err = "distortion requires at least one edge";
    return BT_E_INVALID;
  }
  const double rms = std::sqrt(sum.value / (double)edges);
  out[0] = rms / (2.0 * 1.4142135623730951);
  out[1] = rms;
  out[2] = rms * ((double)map_size / 2.0);
  return BT_OK;
}

}  // namespace bt
