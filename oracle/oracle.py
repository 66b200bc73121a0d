"""ctypes wrapper of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker or the timed CPU baseline; never by the product."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "liboracle.so")


class OrNorm(C.Structure):
    _fields_ = [("min_x", C.c_float), ("min_y", C.c_float), ("scale", C.c_float),
                ("tr_x", C.c_float), ("tr_y", C.c_float)]


class OrGrid(C.Structure):
    _fields_ = [("res_x", C.c_float), ("res_y", C.c_float), ("radius", C.c_float),
                ("cell", C.c_float), ("gx", C.c_uint32), ("gy", C.c_uint32),
                ("cell_w", C.c_float)]


class OrDerived(C.Structure):
    _fields_ = [("ts_x", C.c_float), ("ts_y", C.c_float), ("r2", C.c_float),
                ("pad_x", C.c_float), ("pad_y", C.c_float), ("reach_x", C.c_float),
                ("reach_y", C.c_float)]


_PF = C.POINTER(C.c_float)
_PU32 = C.POINTER(C.c_uint32)
_PU64 = C.POINTER(C.c_uint64)
_PU8 = C.POINTER(C.c_uint8)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError(f"{_PATH} missing: run `make -C oracle`")
        L = C.CDLL(_PATH)
        sig = {
            "or_parse_trl": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(_PF), _PU32,
                                       C.POINTER(_PU32), _PU32, _PF, C.c_char_p]),
            "or_free": (None, [C.c_void_p]),
            "or_normalization": (C.c_int, [_PF, C.POINTER(OrNorm)]),
            "or_resample_count": (C.c_uint64, [_PF, _PU32, C.c_uint32, C.POINTER(OrNorm),
                                               C.c_float, C.c_float, C.c_uint32, _PU32]),
            "or_resample_emit": (C.c_int, [_PF, _PU32, C.c_uint32, C.POINTER(OrNorm), C.c_float,
                                           C.c_float, _PU32, _PU64, _PF, _PU32, _PU32,
                                           C.c_uint32, C.c_uint32]),
            "or_derive": (None, [C.POINTER(OrGrid), C.POINTER(OrDerived)]),
            "or_grid_bin": (None, [_PF, C.c_uint32, C.POINTER(OrGrid), _PU32, _PU32]),
            "or_neighbor_mean": (None, [_PF, _PU8, C.c_uint32, C.POINTER(OrGrid), _PF, _PU32,
                                        C.c_int]),
            "or_neighbor_mean_strided": (None, [_PF, _PU8, C.c_uint32, C.POINTER(OrGrid),
                                                C.c_uint32, _PF, C.c_int]),
            "or_neighbor_mean_f32": (None, [_PF, _PU8, C.c_uint32, C.POINTER(OrGrid), _PF]),
            "or_neighbor_mean_cells": (None, [_PF, _PU8, C.c_uint32, C.POINTER(OrGrid),
                                              C.c_uint32, _PF, C.c_int]),
            "or_candidate_pairs": (C.c_uint64, [_PF, C.c_uint32, C.POINTER(OrGrid)]),
            "or_splat": (None, [_PF, C.c_uint32, C.c_uint32, C.c_uint32, _PU32]),
            "or_blur_weights": (C.c_uint32, [C.c_float, _PF, C.c_uint32]),
            "or_blur_rows": (None, [_PU32, C.c_uint32, C.c_uint32, _PF, C.c_uint32, _PF]),
            "or_blur_cols": (C.c_float, [_PF, C.c_uint32, C.c_uint32, _PF, C.c_uint32, _PF]),
            "or_normalize": (None, [_PF, C.c_size_t, C.c_float, _PF]),
            "or_distortion": (C.c_int, [_PF, _PF, _PU64, _PU32, C.c_uint32,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "or_render": (C.c_int, [C.POINTER(OrRenderIn), _PU8]),
            "or_coverage": (None, [C.POINTER(OrRenderIn), _PU32]),
            "or_half_round": (C.c_float, [C.c_float]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(_PF)


def _u(a):
    return a.ctypes.data_as(_PU32)


def parse_trl(text):
    b = text.encode() if isinstance(text, str) else bytes(text)
    xy, st = _PF(), _PU32()
    n, ns = C.c_uint32(), C.c_uint32()
    bbox = (C.c_float * 4)()
    err = C.create_string_buffer(256)
    rc = lib().or_parse_trl(b, len(b), C.byref(xy), C.byref(n), C.byref(st), C.byref(ns), bbox,
                            err)
    if rc:
        raise ValueError(err.value.decode())
    pts = np.ctypeslib.as_array(xy, shape=(2 * n.value,)).copy().reshape(-1, 2)
    starts = np.ctypeslib.as_array(st, shape=(ns.value,)).copy()
    lib().or_free(C.cast(xy, C.c_void_p))
    lib().or_free(C.cast(st, C.c_void_p))
    return pts, starts, np.array(bbox[:], dtype=np.float32)


def normalization(bbox):
    nm = OrNorm()
    bb = np.ascontiguousarray(bbox, dtype=np.float32)
    if lib().or_normalization(_f(bb), C.byref(nm)):
        raise ValueError("graph has invalid extent")
    return nm


def grid(params):
    """OrGrid from a bt_iter_params-like object (res_x, res_y, radius, cell, gx,
    gy and, if present, the bin width cell_w)."""
    return OrGrid(params.res_x, params.res_y, params.radius, params.cell, params.gx, params.gy,
                  getattr(params, "cell_w", 0.0))


def make_grid(res, radius, cell=None, cell_w=None):
    """Square cells of side `cell` (default: the radius, as the reference), or
    rows of height `cell` split into bins of width `cell_w`."""
    cell = radius if cell is None else cell
    w = cell if cell_w is None else cell_w
    gy = int(np.ceil((np.float64(np.float32(res)) + 1.0) / np.float64(np.float32(cell))))
    gx = int(np.ceil((np.float64(np.float32(res)) + 1.0) / np.float64(np.float32(w))))
    return OrGrid(res, res, radius, cell, gx, gy, 0.0 if cell_w is None else cell_w)


def resample(points, starts, bbox, res_x, res_y, capacity=0xfffffffe, density=None):
    """Returns (counts per path, offsets, points (N,2), edge ids (N,), status)."""
    nm = normalization(bbox)
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    st = np.ascontiguousarray(starts, dtype=np.uint32)
    n_paths = len(st) - 1
    counts = np.zeros(n_paths, dtype=np.uint32)
    total = lib().or_resample_count(_f(pts), _u(st), n_paths, C.byref(nm), res_x, res_y,
                                    capacity, _u(counts))
    offsets = np.zeros(n_paths, dtype=np.uint64)
    offsets[1:] = np.cumsum(counts.astype(np.uint64))[:-1]
    out = np.zeros(2 * total, dtype=np.float32)
    eid = np.zeros(total, dtype=np.uint32)
    dptr, dw, dh = None, 0, 0
    if density is not None:
        dh, dw = density.shape
        dptr = _u(density)
    status = lib().or_resample_emit(_f(pts), _u(st), n_paths, C.byref(nm), res_x, res_y,
                                    _u(counts), offsets.ctypes.data_as(_PU64), _f(out), _u(eid),
                                    dptr, dw, dh)
    return counts, offsets, out.reshape(-1, 2), eid, status


def fixed_flags(counts, offsets, n):
    f = np.zeros(n, dtype=np.uint8)
    valid = counts >= 2
    f[offsets[valid].astype(np.int64)] = 1
    f[(offsets[valid] + counts[valid] - 1).astype(np.int64)] = 1
    return f


def grid_bin(points, g):
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    n = len(p) // 2
    cell = np.zeros(n, dtype=np.uint32)
    counts = np.zeros(g.gx * g.gy, dtype=np.uint32)
    lib().or_grid_bin(_f(p), n, C.byref(g), _u(cell), _u(counts))
    return cell, counts


def neighbor_mean(points, fixed, g, threads=0, neighbors=False):
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    n = len(p) // 2
    out = np.zeros_like(p)
    fx = np.ascontiguousarray(fixed, dtype=np.uint8)
    nb = np.zeros(n, dtype=np.uint32) if neighbors else None
    lib().or_neighbor_mean(_f(p), fx.ctypes.data_as(_PU8), n, C.byref(g), _f(out),
                           _u(nb) if nb is not None else None, threads)
    return (out.reshape(-1, 2), nb) if neighbors else out.reshape(-1, 2)


def neighbor_mean_strided(points, fixed, g, stride, threads=0):
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    out = np.zeros_like(p)
    fx = np.ascontiguousarray(fixed, dtype=np.uint8)
    lib().or_neighbor_mean_strided(_f(p), fx.ctypes.data_as(_PU8), len(p) // 2, C.byref(g),
                                   stride, _f(out), threads)
    return out.reshape(-1, 2)


def neighbor_mean_cells(points, fixed, g, stride=1, threads=0):
    """The reference's tiled algorithm (cell summaries with the fully-inside
    shortcut) in exact sums: the same bits as neighbor_mean."""
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    out = np.zeros_like(p)
    fx = np.ascontiguousarray(fixed, dtype=np.uint8)
    lib().or_neighbor_mean_cells(_f(p), fx.ctypes.data_as(_PU8), len(p) // 2, C.byref(g),
                                 stride, _f(out), threads)
    return out.reshape(-1, 2)


def neighbor_mean_f32(points, fixed, g):
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    out = np.zeros_like(p)
    fx = np.ascontiguousarray(fixed, dtype=np.uint8)
    lib().or_neighbor_mean_f32(_f(p), fx.ctypes.data_as(_PU8), len(p) // 2, C.byref(g), _f(out))
    return out.reshape(-1, 2)


def candidate_pairs(points, g):
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    return int(lib().or_candidate_pairs(_f(p), len(p) // 2, C.byref(g)))


def splat(points, w, h):
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1)
    c = np.zeros(w * h, dtype=np.uint32)
    lib().or_splat(_f(p), len(p) // 2, w, h, _u(c))
    return c.reshape(h, w)


def blur_weights(sigma, max_taps=64):
    w = np.zeros(max_taps + 1, dtype=np.float32)
    t = lib().or_blur_weights(sigma, _f(w), max_taps)
    return w[:t + 1]


def density(counts, sigma):
    """(row-blurred, blurred, peak, normalized) of u32 counts (H, W)."""
    h, w = counts.shape
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    wt = blur_weights(sigma)
    taps = len(wt) - 1
    rows = np.zeros(h * w, dtype=np.float32)
    cols = np.zeros(h * w, dtype=np.float32)
    lib().or_blur_rows(_u(c), w, h, _f(wt), taps, _f(rows))
    peak = lib().or_blur_cols(_f(rows), w, h, _f(wt), taps, _f(cols))
    norm = np.zeros(h * w, dtype=np.float32)
    lib().or_normalize(_f(cols), h * w, peak, _f(norm))
    return rows.reshape(h, w), cols.reshape(h, w), np.float32(peak), norm.reshape(h, w)


def distortion(original, current, offsets, counts):
    a = np.ascontiguousarray(original, dtype=np.float32).reshape(-1)
    b = np.ascontiguousarray(current, dtype=np.float32).reshape(-1)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    cnt = np.ascontiguousarray(counts, dtype=np.uint32)
    d, r = C.c_double(), C.c_double()
    if lib().or_distortion(_f(a), _f(b), off.ctypes.data_as(_PU64), _u(cnt), len(cnt),
                           C.byref(d), C.byref(r)):
        raise ValueError("invalid distortion input")
    return d.value, r.value


def endpoints(points, paths):
    """endpoints.wgsl:9-49 (deduplicate; run when --endpoint-dots is set,
    resampling.rs:137-149): for every path {point_offset, segment_offset,
    count} with count >= 2, its first and last sample; positions equal under
    float == (so -0 equals 0) count once. The reference keeps the index of
    whichever insertion wins and lists them in completion order; this
    canonical form keeps each position's smallest index and sorts them.
    points: (N, 2) float32 in Layout order; paths: (P, 3) uint32."""
    xy = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 2)
    pa = np.asarray(paths, dtype=np.int64).reshape(-1, 3)
    pa = pa[pa[:, 2] >= 2]
    cand = np.unique(np.concatenate([pa[:, 0], pa[:, 0] + pa[:, 2] - 1]))
    if len(cand) == 0:
        return np.zeros(0, dtype=np.uint32)
    pos = xy[cand] + np.float32(0.0)  # -0 + 0 = +0: one bit pattern per value
    key = pos.view(np.uint32).astype(np.uint64)
    key = (key[:, 0] << np.uint64(32)) | key[:, 1]
    _, first = np.unique(key, return_index=True)  # first occurrence = smallest index
    return np.sort(cand[first]).astype(np.uint32)


class OrRenderIn(C.Structure):
    _fields_ = [("pts", _PF), ("edge", _PU32), ("n", C.c_uint32), ("n_edges", C.c_uint32),
                ("paths", _PU32), ("n_paths", C.c_uint32), ("density", _PF),
                ("dw", C.c_uint32), ("dh", C.c_uint32), ("w", C.c_uint32), ("h", C.c_uint32),
                ("msaa", C.c_uint32), ("dots", C.c_uint32), ("palette", _PF),
                ("endpoint", _PF)]


def _render_in(points, edge_ids, paths, w, h, palette, endpoint, density=None, msaa=4,
               dots=False, n_edges=None):
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    xy = arr(np.asarray(points, dtype=np.float32).reshape(-1, 2), np.float32)
    ids = arr(np.asarray(edge_ids).reshape(-1), np.uint32)
    pa = arr(np.asarray(paths).reshape(-1, 3), np.uint32)
    pal = arr(np.asarray(palette).reshape(16, 3), np.float32)
    ep = arr(np.asarray(endpoint).reshape(3), np.float32)
    r = OrRenderIn()
    r.pts, r.edge, r.n = _f(xy), _u(ids), len(xy)
    r.n_edges = len(pa) if n_edges is None else n_edges
    r.paths, r.n_paths = _u(pa), len(pa)
    if density is not None:
        d = arr(density, np.float32)
        r.density, r.dh, r.dw = _f(d), d.shape[0], d.shape[1]
    r.w, r.h, r.msaa, r.dots = w, h, msaa, int(bool(dots))
    r.palette, r.endpoint = _f(pal), _f(ep)
    return r, keep


def render(points, edge_ids, paths, w, h, palette, endpoint, density=None, msaa=4, dots=False,
           n_edges=None):
    """The export image (render_paths.wgsl, gpu.rs:572-651) as (h, w, 4) uint8.
    density: the normalized (dh, dw) map for density colouring, None for
    edge-id colouring."""
    r, keep = _render_in(points, edge_ids, paths, w, h, palette, endpoint, density, msaa, dots,
                         n_edges)
    out = np.zeros((h, w, 4), dtype=np.uint8)
    if lib().or_render(C.byref(r), out.ctypes.data_as(_PU8)):
        raise ValueError("invalid render input")
    return out


def coverage(points, paths, w, h, msaa=4):
    """Fragments per sample, (h, w, msaa) uint32."""
    pal = np.zeros((16, 3), np.float32)
    r, keep = _render_in(points, np.zeros(len(points), np.uint32), paths, w, h, pal,
                         np.zeros(3, np.float32), None, msaa)
    out = np.zeros((h, w, msaa), dtype=np.uint32)
    lib().or_coverage(C.byref(r), _u(out))
    return out


def half_round(x):
    """Nearest binary16 values (ties to even) as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    return np.array([lib().or_half_round(float(v)) for v in x], dtype=np.float32)


class Pipeline:
    """The oracle's whole pipeline for one graph under given per-iteration
    parameters (from the product's bt_iter_params_for, or built by hand)."""

    def __init__(self, graph_points, starts, bbox, params0, with_density=True):
        self.p0 = params0
        d = None
        if with_density:
            d = np.zeros((params0.density_h, params0.density_w), dtype=np.uint32)
        self.counts, self.offsets, self.points, self.edge_id, self.status = resample(
            graph_points, starts, bbox, params0.res_x, params0.res_y, density=d)
        self.fixed = fixed_flags(self.counts, self.offsets, len(self.points))
        self.density_counts = d
        if with_density:
            self.rows, self.blurred, self.peak, self.normalized = density(d, params0.sigma)

    def iterate(self, params, threads=0, with_density=True):
        g = grid(params)
        self.points = neighbor_mean(self.points, self.fixed, g, threads)
        if with_density:
            self.density_counts = splat(self.points, params.density_w, params.density_h)
            self.rows, self.blurred, self.peak, self.normalized = density(
                self.density_counts, params.sigma)
        return self.points
