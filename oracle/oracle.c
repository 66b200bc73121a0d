/*
 * CPU oracle for the edge-bundling hot path. TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP path and the timed CPU baseline. The arithmetic
 * contract is stated in oracle.h. Reference citations are relative to the
 * lpstankus/bundling_thingy source tree.
 */
#include "oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FIX_ONE 16777216.0f /* 2^24 */

int or_version(void) { return 1; }

void or_free(void* p) { free(p); }

/* ---------------------------------------------------------------- parsing */

/* Rust's f32::from_str grammar: sign? (inf | infinity | nan | digits with an
 * optional '.', optional exponent). Anything else is "invalid coordinate". */
static int rust_float_token(const char* s, size_t n) {
    size_t i = 0;
    if (n == 0) return 0;
    if (s[0] == '+' || s[0] == '-') i = 1;
    if (i == n) return 0;
    {
        size_t rest = n - i;
        const char* r = s + i;
        if ((rest == 3 && strncasecmp(r, "inf", 3) == 0) ||
            (rest == 8 && strncasecmp(r, "infinity", 8) == 0) ||
            (rest == 3 && strncasecmp(r, "nan", 3) == 0))
            return 1;
    }
    size_t digits = 0;
    while (i < n && isdigit((unsigned char)s[i])) { i++; digits++; }
    if (i < n && s[i] == '.') {
        i++;
        while (i < n && isdigit((unsigned char)s[i])) { i++; digits++; }
    }
    if (digits == 0) return 0;
    if (i < n && (s[i] == 'e' || s[i] == 'E')) {
        i++;
        if (i < n && (s[i] == '+' || s[i] == '-')) i++;
        size_t ed = 0;
        while (i < n && isdigit((unsigned char)s[i])) { i++; ed++; }
        if (ed == 0) return 0;
    }
    return i == n;
}

static int parse_coord(const char* s, size_t n, float* out) {
    char buf[128];
    if (!rust_float_token(s, n) || n >= sizeof buf) return 0;
    memcpy(buf, s, n);
    buf[n] = 0;
    *out = strtof(buf, NULL);
    return 1;
}

/* graph.rs:35-77 */
int or_parse_trl(const char* text, size_t len, float** xy_out, uint32_t* n_points,
                 uint32_t** starts_out, uint32_t* n_starts, float bbox[4], char* err) {
    size_t cap = 1024, n = 0, scap = 256, ns = 0;
    float* xy = (float*)malloc(cap * 2 * sizeof(float));
    uint32_t* starts = (uint32_t*)malloc(scap * sizeof(uint32_t));
    float mnx = 3.40282347e38f, mny = 3.40282347e38f, mxx = -3.40282347e38f, mxy = -3.40282347e38f;
    size_t pos = 0, line_no = 0;
    while (pos < len) {
        size_t eol = pos;
        while (eol < len && text[eol] != '\n') eol++;
        size_t line_end = eol;
        if (line_end > pos && text[line_end - 1] == '\r') line_end--;
        /* tokens split on single spaces; the first token is skipped */
        size_t t = pos;
        while (t < line_end && text[t] != ' ') t++;
        size_t start = n;
        size_t cur = t; /* at ' ' or line_end */
        while (cur < line_end) {
            size_t a = cur + 1, b = a;
            while (b < line_end && text[b] != ' ') b++;
            cur = b;
            size_t tl = b - a;
            if (tl == 0 || text[a] == '(' || text[a] == '<') continue;
            float x, y;
            if (!parse_coord(text + a, tl, &x)) {
                snprintf(err, 256, "line %zu x: invalid coordinate", line_no + 1);
                goto fail;
            }
            if (cur >= line_end) {
                snprintf(err, 256, "line %zu y: unexpected end of TRL line", line_no + 1);
                goto fail;
            }
            size_t ya = cur + 1, yb = ya;
            while (yb < line_end && text[yb] != ' ') yb++;
            cur = yb;
            if (!parse_coord(text + ya, yb - ya, &y)) {
                snprintf(err, 256, "line %zu y: invalid coordinate", line_no + 1);
                goto fail;
            }
            if (!isfinite(x) || !isfinite(y)) {
                snprintf(err, 256, "non-finite TRL coordinate");
                goto fail;
            }
            if (n > start && xy[2 * (n - 1)] == x && xy[2 * (n - 1) + 1] == y) continue;
            if (n == cap) {
                cap *= 2;
                xy = (float*)realloc(xy, cap * 2 * sizeof(float));
            }
            xy[2 * n] = x;
            xy[2 * n + 1] = y;
            n++;
        }
        if (n - start < 2) {
            n = start;
        } else {
            if (ns + 2 > scap) {
                scap *= 2;
                starts = (uint32_t*)realloc(starts, scap * sizeof(uint32_t));
            }
            starts[ns++] = (uint32_t)start;
            for (size_t i = start; i < n; i++) {
                mnx = fminf(mnx, xy[2 * i]);
                mny = fminf(mny, xy[2 * i + 1]);
                mxx = fmaxf(mxx, xy[2 * i]);
                mxy = fmaxf(mxy, xy[2 * i + 1]);
            }
        }
        line_no++;
        pos = eol + 1;
    }
    if (ns + 1 > scap) starts = (uint32_t*)realloc(starts, (scap + 1) * sizeof(uint32_t));
    starts[ns++] = (uint32_t)n;
    if (n == 0) {
        snprintf(err, 256, "graph has no drawable edges");
        goto fail;
    }
    *xy_out = xy;
    *n_points = (uint32_t)n;
    *starts_out = starts;
    *n_starts = (uint32_t)ns;
    bbox[0] = mnx; bbox[1] = mny; bbox[2] = mxx; bbox[3] = mxy;
    return 0;
fail:
    free(xy);
    free(starts);
    return 1;
}

/* graph.rs:79-91 */
int or_normalization(const float bbox[4], or_norm* out) {
    float rx = bbox[2] - bbox[0], ry = bbox[3] - bbox[1];
    float span = fmaxf(rx, ry);
    if (!(span > 0.0f) || !isfinite(span)) return 1;
    float scale = 0.95f / span;
    if (!isfinite(scale)) return 1;
    out->min_x = bbox[0];
    out->min_y = bbox[1];
    out->scale = scale;
    out->tr_x = (1.0f - scale * rx) / 2.0f;
    out->tr_y = (1.0f - scale * ry) / 2.0f;
    return 0;
}

/* ------------------------------------------------------------- resampling */

/* resampling.wgsl:45-48 */
void or_normalized(const or_norm* nm, const float* xy, uint32_t i, float out[2]) {
    out[0] = ((xy[2 * i] - nm->min_x) * nm->scale + nm->tr_x - 0.5f) * 2.0f;
    out[1] = ((xy[2 * i + 1] - nm->min_y) * nm->scale + nm->tr_y - 0.5f) * 2.0f;
}

static uint32_t f2u_sat(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}

static uint32_t splat_index(float x, float y, uint32_t w, uint32_t h) {
    /* uv = (x * 0.5 + 0.5, 0.5 - y * 0.5); pixel = round(uv * (size - 1)) */
    float ux = x * 0.5f + 0.5f, uy = 0.5f - y * 0.5f;
    float fx = rintf(ux * (float)(w - 1)), fy = rintf(uy * (float)(h - 1));
    fx = fminf(fmaxf(fx, 0.0f), (float)(w - 1));
    fy = fminf(fmaxf(fy, 0.0f), (float)(h - 1));
    return (uint32_t)fy * w + (uint32_t)fx;
}

typedef struct {
    uint32_t count;
    float lx, ly;
} cursor_t;

/* resampling.wgsl:50-80 (write == false branch) */
static void push_count(cursor_t* c, float x, float y, uint32_t capacity) {
    if (c->count > 0 && x == c->lx && y == c->ly) return;
    if (c->count >= capacity) {
        c->count = capacity + 1;
        return;
    }
    c->lx = x;
    c->ly = y;
    c->count++;
}

/* resampling.wgsl:82-136, count mode */
static uint32_t resample_count_one(const float* xy, const uint32_t* starts, uint32_t i,
                                   const or_norm* nm, float res_x, float res_y,
                                   uint32_t capacity) {
    uint32_t s = starts[i], e = starts[i + 1];
    float first[2];
    or_normalized(nm, xy, s, first);
    float tsx = res_x * 0.5f, tsy = res_y * 0.5f;
    cursor_t c = {0, first[0], first[1]};
    push_count(&c, first[0], first[1], capacity);
    float px = first[0], py = first[1];
    float remaining = 1.0f;
    for (uint32_t k = s + 1; k < e; k++) {
        float cur[2];
        or_normalized(nm, xy, k, cur);
        float dx = cur[0] - px, dy = cur[1] - py;
        float ex = dx * tsx, ey = dy * tsy;
        float distance = sqrtf(fmaf(ex, ex, ey * ey));
        if (distance > 0.0f) {
            uint32_t steps = f2u_sat(fmaxf(floorf(distance - remaining) + 1.0f, 0.0f));
            if (steps > 0) {
                float t0 = remaining / distance;
                float t1 = (remaining + (float)(steps - 1)) / distance;
                float fx = px + dx * t0, fy = py + dy * t0;
                float lx = px + dx * t1, ly = py + dy * t1;
                uint32_t dup = (fx == c.lx && fy == c.ly) ? 1u : 0u;
                c.count += steps - dup;
                c.lx = lx;
                c.ly = ly;
                if (c.count > capacity) return capacity + 1;
            }
            remaining = remaining + (float)steps - distance;
        }
        px = cur[0];
        py = cur[1];
    }
    push_count(&c, px, py, capacity);
    if (c.count == 1) {
        for (uint32_t k = s + 1; k < e; k++) {
            float nx[2];
            or_normalized(nm, xy, k, nx);
            if (nx[0] != first[0] || nx[1] != first[1]) {
                push_count(&c, nx[0], nx[1], capacity);
                push_count(&c, px, py, capacity);
                break;
            }
        }
    }
    return c.count;
}

uint64_t or_resample_count(const float* xy, const uint32_t* starts, uint32_t n_paths,
                           const or_norm* nm, float res_x, float res_y, uint32_t capacity,
                           uint32_t* counts) {
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_paths; i++) {
        uint32_t c = resample_count_one(xy, starts, i, nm, res_x, res_y, capacity);
        counts[i] = c < 2 ? 0 : c;
        total += counts[i];
    }
    return total;
}

typedef struct {
    uint32_t count, expect, edge;
    float lx, ly;
    float* pts;
    uint32_t* edge_id;
    uint32_t* density;
    uint32_t dw, dh;
    int bad;
} writer_t;

/* resampling.wgsl:50-80 (write == true branch) */
static void push_write(writer_t* w, float x, float y) {
    if (w->count > 0 && x == w->lx && y == w->ly) return;
    if (w->count >= w->expect) {
        w->bad = 1;
        return;
    }
    w->pts[2 * w->count] = x;
    w->pts[2 * w->count + 1] = y;
    w->edge_id[w->count] = w->edge;
    if (w->density) w->density[splat_index(x, y, w->dw, w->dh)] += 1;
    w->lx = x;
    w->ly = y;
    w->count++;
}

int or_resample_emit(const float* xy, const uint32_t* starts, uint32_t n_paths,
                     const or_norm* nm, float res_x, float res_y,
                     const uint32_t* counts, const uint64_t* offsets,
                     float* pts, uint32_t* edge_id, uint32_t* density,
                     uint32_t dens_w, uint32_t dens_h) {
    int status = 0;
    float tsx = res_x * 0.5f, tsy = res_y * 0.5f;
    for (uint32_t i = 0; i < n_paths; i++) {
        if (counts[i] < 2) continue;
        uint32_t s = starts[i], e = starts[i + 1];
        float first[2];
        or_normalized(nm, xy, s, first);
        writer_t w = {0, counts[i], i, first[0], first[1], pts + 2 * offsets[i],
                      edge_id + offsets[i], density, dens_w, dens_h, 0};
        push_write(&w, first[0], first[1]);
        float px = first[0], py = first[1];
        float remaining = 1.0f;
        for (uint32_t k = s + 1; k < e && !w.bad; k++) {
            float cur[2];
            or_normalized(nm, xy, k, cur);
            float dx = cur[0] - px, dy = cur[1] - py;
            float ex = dx * tsx, ey = dy * tsy;
            float distance = sqrtf(fmaf(ex, ex, ey * ey));
            if (distance > 0.0f) {
                uint32_t steps = f2u_sat(fmaxf(floorf(distance - remaining) + 1.0f, 0.0f));
                for (uint32_t st = 0; st < steps && !w.bad; st++) {
                    float along = remaining + (float)st;
                    float t = along / distance;
                    push_write(&w, px + dx * t, py + dy * t);
                }
                remaining = remaining + (float)steps - distance;
            }
            px = cur[0];
            py = cur[1];
        }
        if (!w.bad) push_write(&w, px, py);
        if (!w.bad && w.count == 1) {
            for (uint32_t k = s + 1; k < e; k++) {
                float nx[2];
                or_normalized(nm, xy, k, nx);
                if (nx[0] != first[0] || nx[1] != first[1]) {
                    push_write(&w, nx[0], nx[1]);
                    push_write(&w, px, py);
                    break;
                }
            }
        }
        if (w.bad || w.count != counts[i]) status = 2;
    }
    return status;
}

/* ------------------------------------------------------------------- grid */

void or_derive(const or_grid* g, or_derived* d) {
    d->ts_x = g->res_x * 0.5f;
    d->ts_y = g->res_y * 0.5f;
    d->r2 = g->radius * g->radius;
    d->pad_x = (g->res_x + 1.0f) * 2.4e-7f;
    d->pad_y = (g->res_y + 1.0f) * 2.4e-7f;
    d->reach_x = g->radius + d->pad_x * 4.0f;
    d->reach_y = g->radius + d->pad_y * 4.0f;
}

/* spatial_grid.wgsl:59-61 */
static inline float texel(float p, float res) { return (p + 1.0f) * 0.5f * res; }

/* spatial_grid.wgsl:63-65, one axis */
static inline uint32_t cell_axis(float t, float cell, uint32_t g) {
    float v = fmaxf(t / cell, 0.0f);
    return v >= (float)(g - 1) ? g - 1 : (uint32_t)v;
}

/* Bin width along x: the reference's grid is square (cell_w = 0). Any bin
 * width gives the same neighbor sums; only the candidate windows change. */
static inline float bin_w(const or_grid* g) { return g->cell_w > 0.0f ? g->cell_w : g->cell; }

void or_grid_bin(const float* pts, uint32_t n, const or_grid* g, uint32_t* cell_of,
                 uint32_t* counts) {
    memset(counts, 0, (size_t)g->gx * g->gy * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) {
        uint32_t cx = cell_axis(texel(pts[2 * i], g->res_x), bin_w(g), g->gx);
        uint32_t cy = cell_axis(texel(pts[2 * i + 1], g->res_y), g->cell, g->gy);
        uint32_t c = cy * g->gx + cx;
        if (cell_of) cell_of[i] = c;
        counts[c]++;
    }
}

typedef struct {
    uint32_t* start; /* gx*gy + 1 */
    float* x;
    float* y;
    int32_t* fx;
    int32_t* fy;
} sorted_t;

/* Counting sort by cell, stable in index order. */
static void sort_points(const float* pts, uint32_t n, const or_grid* g, sorted_t* s) {
    size_t cells = (size_t)g->gx * g->gy;
    uint32_t* cell_of = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* counts = (uint32_t*)malloc(cells * sizeof(uint32_t));
    or_grid_bin(pts, n, g, cell_of, counts);
    s->start = (uint32_t*)malloc((cells + 1) * sizeof(uint32_t));
    uint32_t acc = 0;
    for (size_t c = 0; c < cells; c++) {
        s->start[c] = acc;
        acc += counts[c];
    }
    s->start[cells] = acc;
    s->x = (float*)malloc((size_t)n * sizeof(float));
    s->y = (float*)malloc((size_t)n * sizeof(float));
    s->fx = (int32_t*)malloc((size_t)n * sizeof(int32_t));
    s->fy = (int32_t*)malloc((size_t)n * sizeof(int32_t));
    memcpy(counts, s->start, cells * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) {
        uint32_t slot = counts[cell_of[i]]++;
        float x = pts[2 * i], y = pts[2 * i + 1];
        s->x[slot] = x;
        s->y[slot] = y;
        s->fx[slot] = (int32_t)rintf(x * FIX_ONE);
        s->fy[slot] = (int32_t)rintf(y * FIX_ONE);
    }
    free(cell_of);
    free(counts);
}

static void free_sorted(sorted_t* s) {
    free(s->start);
    free(s->x);
    free(s->y);
    free(s->fx);
    free(s->fy);
}

static void neighbor_mean_impl(const float* in, const uint8_t* fixed, uint32_t n,
                               const or_grid* g, float* out, uint32_t* neighbors, int threads,
                               uint32_t stride);

void or_neighbor_mean(const float* in, const uint8_t* fixed, uint32_t n, const or_grid* g,
                      float* out, uint32_t* neighbors, int threads) {
    neighbor_mean_impl(in, fixed, n, g, out, neighbors, threads, 1);
}

void or_neighbor_mean_strided(const float* in, const uint8_t* fixed, uint32_t n,
                              const or_grid* g, uint32_t stride, float* out, int threads) {
    neighbor_mean_impl(in, fixed, n, g, out, NULL, threads, stride ? stride : 1);
}

static void neighbor_mean_impl(const float* in, const uint8_t* fixed, uint32_t n,
                               const or_grid* g, float* out, uint32_t* neighbors, int threads,
                               uint32_t stride) {
    or_derived d;
    or_derive(g, &d);
    sorted_t s;
    sort_points(in, n, g, &s);
    const uint32_t gx = g->gx;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
    for (uint32_t i = 0; i < n; i++) {
        float px = in[2 * i], py = in[2 * i + 1];
        if (fixed[i] || i % stride != 0) {
            out[2 * i] = px;
            out[2 * i + 1] = py;
            if (neighbors) neighbors[i] = 0;
            continue;
        }
        float tx = texel(px, g->res_x), ty = texel(py, g->res_y);
        uint32_t x0 = cell_axis(tx - d.reach_x, bin_w(g), g->gx);
        uint32_t x1 = cell_axis(tx + d.reach_x, bin_w(g), g->gx);
        uint32_t y0 = cell_axis(ty - d.reach_y, g->cell, g->gy);
        uint32_t y1 = cell_axis(ty + d.reach_y, g->cell, g->gy);
        int64_t ax = 0, ay = 0;
        int64_t cnt = 0;
        for (uint32_t cy = y0; cy <= y1; cy++) {
            uint32_t b = s.start[cy * gx + x0], e = s.start[cy * gx + x1 + 1];
            int64_t sx = 0, sy = 0, c = 0;
            for (uint32_t j = b; j < e; j++) {
                float ddx = (s.x[j] - px) * d.ts_x;
                float ddy = (s.y[j] - py) * d.ts_y;
                float d2 = fmaf(ddx, ddx, ddy * ddy);
                int ok = d2 <= d.r2;
                sx += ok ? s.fx[j] : 0;
                sy += ok ? s.fy[j] : 0;
                c += ok;
            }
            ax += sx;
            ay += sy;
            cnt += c;
        }
        int64_t pfx = (int64_t)rintf(px * FIX_ONE), pfy = (int64_t)rintf(py * FIX_ONE);
        int64_t dx = ax - cnt * pfx, dy = ay - cnt * pfy;
        double den = (double)cnt * 16777216.0;
        out[2 * i] = (float)((double)px + (double)dx / den);
        out[2 * i + 1] = (float)((double)py + (double)dy / den);
        if (neighbors) neighbors[i] = (uint32_t)cnt;
    }
    free_sorted(&s);
}

/* neighbor_mean.wgsl:40-59 (classify_cell, cell_displacement) with
 * spatial_grid.wgsl:161-220 (summarize_cells): the reference's fully-inside
 * shortcut over per-cell summaries, in exact sums. */
void or_neighbor_mean_cells(const float* in, const uint8_t* fixed, uint32_t n,
                            const or_grid* g, uint32_t stride, float* out, int threads) {
    or_derived d;
    or_derive(g, &d);
    sorted_t s;
    sort_points(in, n, g, &s);
    const uint32_t gx = g->gx;
    const size_t cells = (size_t)g->gx * g->gy;
    int64_t* sum = (int64_t*)calloc(2 * cells, sizeof(int64_t));
    float* box = (float*)malloc(4 * cells * sizeof(float));
    for (size_t c = 0; c < cells; c++) {
        float x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
        int64_t sx = 0, sy = 0;
        for (uint32_t j = s.start[c]; j < s.start[c + 1]; j++) {
            x0 = fminf(x0, s.x[j]);
            y0 = fminf(y0, s.y[j]);
            x1 = fmaxf(x1, s.x[j]);
            y1 = fmaxf(y1, s.y[j]);
            sx += s.fx[j];
            sy += s.fy[j];
        }
        sum[2 * c] = sx;
        sum[2 * c + 1] = sy;
        box[4 * c] = x0;
        box[4 * c + 1] = y0;
        box[4 * c + 2] = x1;
        box[4 * c + 3] = y1;
    }
    const double r2 = (double)g->radius * (double)g->radius;
    const double tsx = d.ts_x, tsy = d.ts_y, padx = d.pad_x, pady = d.pad_y;
    if (stride == 0) stride = 1;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
    for (uint32_t i = 0; i < n; i++) {
        float px = in[2 * i], py = in[2 * i + 1];
        if (fixed[i] || i % stride != 0) {
            out[2 * i] = px;
            out[2 * i + 1] = py;
            continue;
        }
        float tx = texel(px, g->res_x), ty = texel(py, g->res_y);
        uint32_t x0 = cell_axis(tx - d.reach_x, bin_w(g), g->gx);
        uint32_t x1 = cell_axis(tx + d.reach_x, bin_w(g), g->gx);
        uint32_t y0 = cell_axis(ty - d.reach_y, g->cell, g->gy);
        uint32_t y1 = cell_axis(ty + d.reach_y, g->cell, g->gy);
        int64_t ax = 0, ay = 0, cnt = 0;
        for (uint32_t cy = y0; cy <= y1; cy++) {
            for (uint32_t cx = x0; cx <= x1; cx++) {
                const size_t c = (size_t)cy * gx + cx;
                const uint32_t b = s.start[c], e = s.start[c + 1];
                if (b == e) continue;
                const float* bx = box + 4 * c;
                const double nx = fmax(fmax((double)bx[0] - px, px - (double)bx[2]), 0.0);
                const double ny = fmax(fmax((double)bx[1] - py, py - (double)bx[3]), 0.0);
                const double ntx = fmax(nx * tsx - padx, 0.0), nty = fmax(ny * tsy - pady, 0.0);
                if (ntx * ntx + nty * nty > r2) continue;  /* outside */
                const double fx = fmax(fabs((double)bx[0] - px), fabs((double)bx[2] - px));
                const double fy = fmax(fabs((double)bx[1] - py), fabs((double)bx[3] - py));
                const double ftx = fx * tsx + padx, fty = fy * tsy + pady;
                if (ftx * ftx + fty * fty < r2) {  /* fully inside: the summary */
                    ax += sum[2 * c];
                    ay += sum[2 * c + 1];
                    cnt += e - b;
                    continue;
                }
                for (uint32_t j = b; j < e; j++) {
                    float ddx = (s.x[j] - px) * d.ts_x;
                    float ddy = (s.y[j] - py) * d.ts_y;
                    int ok = fmaf(ddx, ddx, ddy * ddy) <= d.r2;
                    ax += ok ? s.fx[j] : 0;
                    ay += ok ? s.fy[j] : 0;
                    cnt += ok;
                }
            }
        }
        int64_t pfx = (int64_t)rintf(px * FIX_ONE), pfy = (int64_t)rintf(py * FIX_ONE);
        int64_t dx = ax - cnt * pfx, dy = ay - cnt * pfy;
        double den = (double)cnt * 16777216.0;
        out[2 * i] = (float)((double)px + (double)dx / den);
        out[2 * i + 1] = (float)((double)py + (double)dy / den);
    }
    free(sum);
    free(box);
    free_sorted(&s);
}

void or_neighbor_mean_f32(const float* in, const uint8_t* fixed, uint32_t n,
                          const or_grid* g, float* out) {
    or_derived d;
    or_derive(g, &d);
    sorted_t s;
    sort_points(in, n, g, &s);
    float reach_x = g->radius + d.pad_x, reach_y = g->radius + d.pad_y;
    for (uint32_t i = 0; i < n; i++) {
        float px = in[2 * i], py = in[2 * i + 1];
        if (fixed[i]) {
            out[2 * i] = px;
            out[2 * i + 1] = py;
            continue;
        }
        float tx = texel(px, g->res_x), ty = texel(py, g->res_y);
        uint32_t x0 = cell_axis(tx - reach_x, bin_w(g), g->gx);
        uint32_t x1 = cell_axis(tx + reach_x, bin_w(g), g->gx);
        uint32_t y0 = cell_axis(ty - reach_y, g->cell, g->gy);
        uint32_t y1 = cell_axis(ty + reach_y, g->cell, g->gy);
        float sx = 0.0f, sy = 0.0f;
        uint32_t cnt = 0;
        for (uint32_t cy = y0; cy <= y1; cy++)
            for (uint32_t cx = x0; cx <= x1; cx++) {
                uint32_t c = cy * g->gx + cx;
                for (uint32_t j = s.start[c]; j < s.start[c + 1]; j++) {
                    float ddx = s.x[j] - px, ddy = s.y[j] - py;
                    float ex = ddx * d.ts_x, ey = ddy * d.ts_y;
                    if (ex * ex + ey * ey <= d.r2) {
                        sx += ddx;
                        sy += ddy;
                        cnt++;
                    }
                }
            }
        out[2 * i] = px + sx / (float)cnt;
        out[2 * i + 1] = py + sy / (float)cnt;
    }
    free_sorted(&s);
}

uint64_t or_candidate_pairs(const float* pts, uint32_t n, const or_grid* g) {
    size_t cells = (size_t)g->gx * g->gy;
    uint32_t* cell_of = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* counts = (uint32_t*)malloc(cells * sizeof(uint32_t));
    or_grid_bin(pts, n, g, cell_of, counts);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) {
        int cx = (int)(cell_of[i] % g->gx), cy = (int)(cell_of[i] / g->gx);
        for (int y = cy - 1; y <= cy + 1; y++)
            for (int x = cx - 1; x <= cx + 1; x++)
                if (x >= 0 && y >= 0 && x < (int)g->gx && y < (int)g->gy)
                    total += counts[(size_t)y * g->gx + x];
    }
    free(cell_of);
    free(counts);
    return total;
}

/* ---------------------------------------------------------------- density */

void or_splat(const float* pts, uint32_t n, uint32_t w, uint32_t h, uint32_t* counts) {
    for (uint32_t i = 0; i < n; i++) counts[splat_index(pts[2 * i], pts[2 * i + 1], w, h)] += 1;
}

uint32_t or_blur_weights(float sigma, float* w, uint32_t max_taps) {
    uint32_t taps = (uint32_t)ceil(3.0 * (double)sigma);
    if (taps > max_taps) taps = max_taps;
    double s2 = 2.0 * (double)sigma * (double)sigma;
    double z = 0.0;
    for (int k = -(int)taps; k <= (int)taps; k++) z += exp(-(double)k * k / s2);
    for (uint32_t k = 0; k <= taps; k++) {
        double v = exp(-(double)k * k / s2) / z;
        w[k] = (float)(nearbyint(v * 1e10) / 1e10);
    }
    return taps;
}

void or_blur_rows(const uint32_t* in, uint32_t w, uint32_t h, const float* wt, uint32_t taps,
                  float* out) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (uint32_t y = 0; y < h; y++) {
        const uint32_t* row = in + (size_t)y * w;
        for (uint32_t x = 0; x < w; x++) {
            float v = (float)row[x] * wt[0];
            for (uint32_t k = 1; k <= taps; k++) {
                float a = x >= k ? (float)row[x - k] : 0.0f;
                float b = x + k < w ? (float)row[x + k] : 0.0f;
                v = v + (a + b) * wt[k];
            }
            out[(size_t)y * w + x] = v;
        }
    }
}

float or_blur_cols(const float* in, uint32_t w, uint32_t h, const float* wt, uint32_t taps,
                   float* out) {
    float peak = 0.0f;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(max : peak)
#endif
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            float v = in[(size_t)y * w + x] * wt[0];
            for (uint32_t k = 1; k <= taps; k++) {
                float a = y >= k ? in[(size_t)(y - k) * w + x] : 0.0f;
                float b = y + k < h ? in[(size_t)(y + k) * w + x] : 0.0f;
                v = v + (a + b) * wt[k];
            }
            out[(size_t)y * w + x] = v;
            peak = fmaxf(peak, v);
        }
    }
    return peak;
}

void or_normalize(const float* in, size_t n, float peak, float* out) {
    for (size_t i = 0; i < n; i++)
        out[i] = peak > 0.0f ? fminf(fmaxf(in[i] / peak, 0.0f), 1.0f) : 0.0f;
}

float or_iterate(const float* in, const uint8_t* fixed, uint32_t n, const or_grid* g,
                 float* out, uint32_t dens_w, uint32_t dens_h, float sigma, uint32_t* counts,
                 float* tmp, float* blurred, float* normalized, int threads) {
    float wt[65];
    or_neighbor_mean(in, fixed, n, g, out, NULL, threads);
    if (!counts) return 0.0f;
    uint32_t taps = or_blur_weights(sigma, wt, 64);
    memset(counts, 0, (size_t)dens_w * dens_h * sizeof(uint32_t));
    or_splat(out, n, dens_w, dens_h, counts);
    or_blur_rows(counts, dens_w, dens_h, wt, taps, tmp);
    float peak = or_blur_cols(tmp, dens_w, dens_h, wt, taps, blurred);
    or_normalize(blurred, (size_t)dens_w * dens_h, peak, normalized);
    return peak;
}

/* ------------------------------------------------------------- distortion */

typedef struct {
    const float* p;
    uint32_t n;
    double length;
} curve_t;

static double seg_len(const float* a, const float* b) {
    return hypot((double)b[0] - (double)a[0], (double)b[1] - (double)a[1]);
}

/* distortion.rs:19-41 */
static int curve_new(const float* p, uint32_t n, curve_t* c) {
    if (n == 0) return 1;
    for (uint32_t i = 0; i < 2 * n; i++)
        if (!isfinite(p[i])) return 1;
    double len = 0.0;
    for (uint32_t i = 0; i + 1 < n; i++) len += seg_len(p + 2 * i, p + 2 * i + 2);
    if (!isfinite(len)) return 1;
    c->p = p;
    c->n = n;
    c->length = len;
    return 0;
}

typedef struct {
    const curve_t* c;
    uint32_t next; /* index of the next window's first point */
    double total, traversed, sx, sy, ex, ey, st, et;
} dcursor_t;

/* distortion.rs:63-76 */
static void dc_init(dcursor_t* k, const curve_t* c) {
    k->c = c;
    k->next = 0;
    k->total = c->length;
    k->traversed = 0.0;
    k->sx = k->ex = (double)c->p[0];
    k->sy = k->ey = (double)c->p[1];
    k->st = k->et = 0.0;
}

/* distortion.rs:80-105 */
static void dc_advance(dcursor_t* k, double t) {
    while (k->et <= t && k->et < 1.0) {
        if (k->next + 1 >= k->c->n) {
            k->sx = k->ex;
            k->sy = k->ey;
            k->st = k->et;
            k->et = 1.0;
            return;
        }
        const float* a = k->c->p + 2 * k->next;
        const float* b = a + 2;
        k->next++;
        double len = seg_len(a, b);
        if (len == 0.0) continue;
        k->sx = a[0];
        k->sy = a[1];
        k->ex = b[0];
        k->ey = b[1];
        k->st = k->et;
        k->traversed += len;
        k->et = (k->next + 1 >= k->c->n) ? 1.0 : k->traversed / k->total;
    }
}

/* distortion.rs:107-120 */
static void dc_pos(const dcursor_t* k, double t, double* x, double* y) {
    if (t >= k->et) { *x = k->ex; *y = k->ey; return; }
    if (t <= k->st) { *x = k->sx; *y = k->sy; return; }
    double f = (t - k->st) / (k->et - k->st);
    *x = fma(f, k->ex - k->sx, k->sx);
    *y = fma(f, k->ey - k->sy, k->sy);
}

typedef struct {
    double value, correction;
} kahan_t;

/* distortion.rs:129-135 */
static void kahan_add(kahan_t* s, double term) {
    double adjusted = term - s->correction;
    double next = s->value + adjusted;
    s->correction = (next - s->value) - adjusted;
    s->value = next;
}

/* distortion.rs:137-168 */
static int squared_integral(const curve_t* o, const curve_t* c, double* out) {
    dcursor_t a, b;
    dc_init(&a, o);
    dc_init(&b, c);
    kahan_t sum = {0.0, 0.0};
    double t = 0.0;
    for (;;) {
        if (t >= 1.0) break;
        dc_advance(&a, t);
        dc_advance(&b, t);
        double next = fmin(a.et, b.et);
        if (!(next > t && next <= 1.0)) return 1;
        double px, py, qx, qy, rx, ry, sx, sy;
        dc_pos(&a, t, &px, &py);
        dc_pos(&b, t, &qx, &qy);
        dc_pos(&a, next, &rx, &ry);
        dc_pos(&b, next, &sx, &sy);
        double ax = px - qx, ay = py - qy, bx = rx - sx, by = ry - sy;
        double smx = ax + bx, smy = ay + by, dlx = bx - ax, dly = by - ay;
        double mid2 = fma(smx, smx, smy * smy);
        double ch2 = fma(dlx, dlx, dly * dly);
        kahan_add(&sum, (next - t) * (mid2 / 4.0 + ch2 / 12.0));
        t = next;
    }
    if (!(isfinite(sum.value) && sum.value >= 0.0)) return 1;
    *out = sum.value;
    return 0;
}

/* distortion.rs:170-216 plus geometry.rs:28-73 (count 0 paths are skipped,
 * count 1 is an error). */
int or_distortion(const float* original, const float* current, const uint64_t* offsets,
                  const uint32_t* counts, uint32_t n_paths, double* distortion, double* rms) {
    kahan_t sum = {0.0, 0.0};
    uint32_t edges = 0;
    for (uint32_t i = 0; i < n_paths; i++) {
        if (counts[i] == 0) continue;
        if (counts[i] < 2) return 1;
        curve_t a, b;
        if (curve_new(original + 2 * offsets[i], counts[i], &a)) return 1;
        if (curve_new(current + 2 * offsets[i], counts[i], &b)) return 1;
        double v;
        if (squared_integral(&a, &b, &v)) return 1;
        kahan_add(&sum, v);
        if (!isfinite(sum.value)) return 1;
        edges++;
    }
    if (edges == 0) return 1;
    double r = sqrt(sum.value / (double)edges);
    *rms = r;
    *distortion = r / (2.0 * 1.4142135623730951);
    return 0;
}

/* ------------------------------------------------------------------------
 * Export image (render_paths.wgsl:1-97, gpu.rs:572-651, palette.rs).
 * Restated sequentially: every segment of every path in Layout order, its
 * first triangle then its second, blended sample by sample into 8-bit
 * storage; then the endpoint outlines and dots; then the box resolve.
 * Rasterization contract: corners in f32 as the vertex shader computes
 * them, viewport ((x*0.5+0.5)*W, (0.5-y*0.5)*H), snapped to 1/256 px;
 * integer edge functions, top-left fill rule, no culling; standard 2x / 4x
 * sample locations; fragment colour evaluated at the pixel center; dots
 * shaded per sample.
 * ---------------------------------------------------------------------- */

static const int OR_S4[4][2] = {{96, 32}, {224, 96}, {32, 160}, {160, 224}};
static const int OR_S2[2][2] = {{192, 192}, {64, 64}};

float or_half_round(float x) {
    /* nearest binary16 value, ties to even, as an f32 */
    uint32_t u;
    memcpy(&u, &x, 4);
    const uint32_t sign = u & 0x80000000u, a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return x;                 /* inf / nan */
    float r;
    if (a < 0x38800000u) {                          /* below 2^-14: quantum 2^-24 */
        float m;
        memcpy(&m, &a, 4);
        r = rintf(m * 16777216.0f) * (1.0f / 16777216.0f);
    } else {
        uint32_t b = (a + 0x0fffu + ((a >> 13) & 1u)) & ~0x1fffu;
        if (b >= 0x477ff000u) b = 0x7f800000u;      /* >= 65520 rounds to inf */
        memcpy(&r, &b, 4);
    }
    uint32_t ru;
    memcpy(&ru, &r, 4);
    ru |= sign;
    memcpy(&r, &ru, 4);
    return r;
}

static uint8_t or_unorm8(float x) {
    if (x < 0.0f) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return (uint8_t)rintf(x * 255.0f);
}

typedef struct {
    long long dx[3], dy[3], ax[3], ay[3], bias[3];
    int live;
} or_tri;

static void or_tri_make(const int* X, const int* Y, int a, int b, int c, or_tri* t) {
    long long px[3] = {X[a], X[b], X[c]}, py[3] = {Y[a], Y[b], Y[c]};
    long long area = (px[1] - px[0]) * (py[2] - py[0]) - (py[1] - py[0]) * (px[2] - px[0]);
    t->live = area != 0;
    if (area < 0) { /* make the inside the positive side of every edge */
        long long sx = px[1], sy = py[1];
        px[1] = px[2]; py[1] = py[2];
        px[2] = sx; py[2] = sy;
    }
    for (int e = 0; e < 3; e++) {
        int n = (e + 1) % 3;
        t->ax[e] = px[e];
        t->ay[e] = py[e];
        t->dx[e] = px[n] - px[e];
        t->dy[e] = py[n] - py[e];
        /* y grows downwards: a left edge runs up (dy < 0), a top edge runs
         * right (dy == 0, dx > 0) */
        int top_left = t->dy[e] < 0 || (t->dy[e] == 0 && t->dx[e] > 0);
        t->bias[e] = top_left ? 0 : 1;
    }
}

static int or_tri_covers(const or_tri* t, long long sx, long long sy) {
    if (!t->live) return 0;
    for (int e = 0; e < 3; e++) {
        long long E = t->dx[e] * (sy - t->ay[e]) - t->dy[e] * (sx - t->ax[e]);
        if (E < t->bias[e]) return 0;
    }
    return 1;
}

/* Corners of segment (p, q): 0 = p + n, 1 = p - n, 2 = q + n, 3 = q - n
 * (render_paths.wgsl:36-50, CORNERS), in 1/256 px. */
static int or_corners(const float* p, const float* q, uint32_t w, uint32_t h, int* X, int* Y) {
    float dx = q[0] - p[0], dy = q[1] - p[1];
    float len = sqrtf(dx * dx + dy * dy);
    float k = 0.001f / (len > 1e-10f ? len : 1e-10f);
    float nx = -dy * k, ny = dx * k;
    float cx[4] = {p[0] + nx, p[0] - nx, q[0] + nx, q[0] - nx};
    float cy[4] = {p[1] + ny, p[1] - ny, q[1] + ny, q[1] - ny};
    for (int c = 0; c < 4; c++) {
        float fx = (cx[c] * 0.5f + 0.5f) * (float)w;
        float fy = (0.5f - cy[c] * 0.5f) * (float)h;
        if (!(fabsf(fx) < 1.0e6f && fabsf(fy) < 1.0e6f)) return 0;
        X[c] = (int)rintf(fx * 256.0f);
        Y[c] = (int)rintf(fy * 256.0f);
    }
    return 1;
}

static void or_palette(const float* pal, float value, float* rgb) {
    float stop = value * 15.0f;
    uint32_t i = stop <= 0.0f ? 0u : (uint32_t)stop;
    if (i > 14u) i = 14u;
    float t = stop - (float)i;
    for (int c = 0; c < 3; c++) rgb[c] = pal[3 * i + c] * (1.0f - t) + pal[3 * (i + 1) + c] * t;
}

static float or_density_at(const or_render_in* in, uint32_t px, uint32_t py) {
    float tx = ((float)px + 0.5f) / (float)in->w * (float)(in->dw - 1);
    float ty = ((float)py + 0.5f) / (float)in->h * (float)(in->dh - 1);
    float fx = floorf(tx), fy = floorf(ty);
    float ax = rintf((tx - fx) * 256.0f) * (1.0f / 256.0f);
    float ay = rintf((ty - fy) * 256.0f) * (1.0f / 256.0f);
    uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy;
    if (x0 > in->dw - 1) x0 = in->dw - 1;
    if (y0 > in->dh - 1) y0 = in->dh - 1;
    uint32_t x1 = x0 + 1 < in->dw ? x0 + 1 : in->dw - 1;
    uint32_t y1 = y0 + 1 < in->dh ? y0 + 1 : in->dh - 1;
    const float* d = in->density;
    float a = or_half_round(d[(size_t)y0 * in->dw + x0]), b = or_half_round(d[(size_t)y0 * in->dw + x1]);
    float c = or_half_round(d[(size_t)y1 * in->dw + x0]), e = or_half_round(d[(size_t)y1 * in->dw + x1]);
    float top = a * (1.0f - ax) + b * ax;
    float bot = c * (1.0f - ax) + e * ax;
    float v = top * (1.0f - ay) + bot * ay;
    return v < 0.0f ? 0.0f : v > 1.0f ? 1.0f : v;
}

/* Visits every (segment, triangle) fragment sample in draw order. */
typedef void (*or_frag_fn)(void* ctx, uint32_t seg, uint32_t px, uint32_t py, uint32_t k);

static void or_raster(const or_render_in* in, or_frag_fn fn, void* ctx) {
    const int (*S)[2] = in->msaa == 4 ? OR_S4 : OR_S2;
    for (uint32_t p = 0; p < in->n_paths; p++) {
        uint32_t off = in->paths[3 * p], cnt = in->paths[3 * p + 2];
        for (uint32_t s = off; cnt >= 2 && s + 1 < off + cnt; s++) {
            int X[4], Y[4];
            if (!or_corners(in->pts + 2 * (size_t)s, in->pts + 2 * (size_t)(s + 1), in->w, in->h, X, Y))
                continue;
            or_tri tris[2];
            or_tri_make(X, Y, 0, 1, 2, &tris[0]); /* strip triangle 0 */
            or_tri_make(X, Y, 2, 1, 3, &tris[1]); /* strip triangle 1 */
            int mnx = X[0], mxx = X[0], mny = Y[0], mxy = Y[0];
            for (int c = 1; c < 4; c++) {
                if (X[c] < mnx) mnx = X[c];
                if (X[c] > mxx) mxx = X[c];
                if (Y[c] < mny) mny = Y[c];
                if (Y[c] > mxy) mxy = Y[c];
            }
            long long x0 = (long long)floor((mnx - 256.0) / 256.0), x1 = (long long)floor(mxx / 256.0);
            long long y0 = (long long)floor((mny - 256.0) / 256.0), y1 = (long long)floor(mxy / 256.0);
            if (x0 < 0) x0 = 0;
            if (y0 < 0) y0 = 0;
            if (x1 > (long long)in->w - 1) x1 = (long long)in->w - 1;
            if (y1 > (long long)in->h - 1) y1 = (long long)in->h - 1;
            for (int t = 0; t < 2; t++)
                for (long long py = y0; py <= y1; py++)
                    for (long long px = x0; px <= x1; px++)
                        for (uint32_t k = 0; k < in->msaa; k++)
                            if (or_tri_covers(&tris[t], px * 256 + S[k][0], py * 256 + S[k][1]))
                                fn(ctx, s, (uint32_t)px, (uint32_t)py, k);
        }
    }
}

static void or_count_frag(void* ctx, uint32_t seg, uint32_t px, uint32_t py, uint32_t k) {
    const or_render_in* in = ((const or_render_in**)ctx)[0];
    uint32_t* counts = ((uint32_t**)ctx)[1];
    counts[((size_t)py * in->w + px) * in->msaa + k]++;
}

void or_coverage(const or_render_in* in, uint32_t* counts) {
    memset(counts, 0, (size_t)in->w * in->h * in->msaa * 4);
    void* ctx[2] = {(void*)in, counts};
    or_raster(in, or_count_frag, ctx);
}

typedef struct {
    const or_render_in* in;
    uint8_t* s; /* w*h*msaa*3 */
} or_blend_ctx;

static void or_blend_frag(void* vctx, uint32_t seg, uint32_t px, uint32_t py, uint32_t k) {
    or_blend_ctx* c = (or_blend_ctx*)vctx;
    const or_render_in* in = c->in;
    float rgb[3];
    if (in->density) {
        or_palette(in->palette, or_density_at(in, px, py), rgb);
    } else {
        uint32_t e = in->n_edges > 2 ? in->n_edges : 2;
        or_palette(in->palette, (float)in->edge[seg] / (float)(e - 1), rgb);
    }
    uint8_t* v = c->s + (((size_t)py * in->w + px) * in->msaa + k) * 3;
    for (int ch = 0; ch < 3; ch++)
        v[ch] = or_unorm8(rgb[ch] * 0.1f + ((float)v[ch] * (1.0f / 255.0f)) * (1.0f - 0.1f));
}

int or_render(const or_render_in* in, uint8_t* rgba) {
    if (in->msaa != 2 && in->msaa != 4) return 1;
    const size_t npix = (size_t)in->w * in->h;
    uint8_t* s = (uint8_t*)malloc(npix * in->msaa * 3);
    uint8_t* dot = (uint8_t*)calloc(npix * in->msaa, 1); /* 1 outline, 2 dot */
    if (!s || !dot) {
        free(s);
        free(dot);
        return 1;
    }
    const float clear[3] = {0.025f, 0.025f, 0.035f};
    for (size_t i = 0; i < npix * in->msaa; i++)
        for (int ch = 0; ch < 3; ch++) s[3 * i + ch] = or_unorm8(clear[ch]);
    or_blend_ctx bc = {in, s};
    or_raster(in, or_blend_frag, &bc);
    if (in->dots) {
        const int (*S)[2] = in->msaa == 4 ? OR_S4 : OR_S2;
        const float inner = 1.0f - 1.0f / 3.0f;
        for (uint32_t p = 0; p < in->n_paths; p++) {
            uint32_t off = in->paths[3 * p], cnt = in->paths[3 * p + 2];
            if (cnt < 2) continue;
            for (int end = 0; end < 2; end++) {
                const float* q = in->pts + 2 * (size_t)(end ? off + cnt - 1 : off);
                float cx = (q[0] * 0.5f + 0.5f) * (float)in->w;
                float cy = (0.5f - q[1] * 0.5f) * (float)in->h;
                if (!(fabsf(cx) < 1.0e6f && fabsf(cy) < 1.0e6f)) continue;
                long long ya = (long long)floorf(cy) - 5, yb = (long long)floorf(cy) + 5;
                long long xa = (long long)floorf(cx) - 5, xb = (long long)floorf(cx) + 5;
                if (ya < 0) ya = 0;
                if (xa < 0) xa = 0;
                if (yb > (long long)in->h - 1) yb = (long long)in->h - 1;
                if (xb > (long long)in->w - 1) xb = (long long)in->w - 1;
                for (long long py = ya; py <= yb; py++) {
                    for (long long px = xa; px <= xb; px++) {
                        for (uint32_t k = 0; k < in->msaa; k++) {
                            float ox = ((float)px + (float)S[k][0] * (1.0f / 256.0f) - cx) / 3.0f;
                            float oy = ((float)py + (float)S[k][1] * (1.0f / 256.0f) - cy) / 3.0f;
                            if (!(ox >= -1.0f && ox < 1.0f && oy >= -1.0f && oy < 1.0f)) continue;
                            float d2 = ox * ox + oy * oy;
                            uint8_t* f = &dot[((size_t)py * in->w + px) * in->msaa + k];
                            if (d2 <= inner * inner) *f = 2;
                            else if (d2 <= 1.0f && *f == 0) *f = 1;
                        }
                    }
                }
            }
        }
        uint8_t ring[3], fill[3];
        for (int ch = 0; ch < 3; ch++) {
            ring[ch] = or_unorm8(in->palette[ch] * 0.1f);
            fill[ch] = or_unorm8(in->endpoint[ch]);
        }
        for (size_t i = 0; i < npix * in->msaa; i++)
            if (dot[i])
                for (int ch = 0; ch < 3; ch++) s[3 * i + ch] = dot[i] == 2 ? fill[ch] : ring[ch];
    }
    for (size_t p = 0; p < npix; p++) {
        for (int ch = 0; ch < 3; ch++) {
            uint32_t sum = 0;
            for (uint32_t k = 0; k < in->msaa; k++) sum += s[(p * in->msaa + k) * 3 + ch];
            rgba[4 * p + ch] = (uint8_t)((sum + in->msaa / 2) / in->msaa);
        }
        rgba[4 * p + 3] = 255;
    }
    free(s);
    free(dot);
    return 0;
}
