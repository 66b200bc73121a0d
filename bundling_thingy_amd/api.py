"""Python mirror of the reference's Graph / Gpu interface over the C ABI.

Names and behaviour follow lpstankus/bundling_thingy: ``Graph.parse`` /
``Graph.normalization`` (src/graph.rs:35-91), ``Gpu.upload`` / ``Gpu.iterate``
(src/gpu.rs:492-553), ``Gpu.distortion_layout`` / ``Gpu.distortion_snapshot``
(src/gpu.rs:438-482) returning a ``Layout`` / ``Snapshot``
(src/geometry.rs:7-224), and the distortion ``Aggregate``
(src/distortion.rs:170-216). Errors raise ``BtError`` with the reference's
messages. All device work runs in libbundling.so (HIP, gfx950); nothing here
computes on the CPU except the Layout/Decoder validation the reference also
does on the host.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbundling.so")

BT_OK, BT_E_INVALID, BT_E_NO_GRAPH, BT_E_CAPACITY = 0, 1, 2, 3
BT_E_DEVICE, BT_E_STATE, BT_E_MISMATCH, BT_E_BUSY = 4, 5, 6, 7
FLAG_DENSITY, FLAG_DISTORTION, FLAG_DENSITY_ON_READ = 1, 2, 4
DENSITY_COUNTS, DENSITY_ROWS, DENSITY_BLURRED, DENSITY_NORMALIZED = 0, 1, 2, 3
DENSITY_HALF = 4  # the normalized R channel as binary16 bits (the rgba16f texture)
PALETTE_CIVIDIS, PALETTE_VIRIDIS, PALETTE_PLASMA, PALETTE_BATLOW = 0, 1, 2, 3
PALETTES = {"cividis": 0, "viridis": 1, "plasma": 2, "batlow": 3}
COLOR_DENSITY, COLOR_EDGE_ID = 0, 1
STAGES = ["clear", "grid_count", "grid_scan", "grid_tasks", "grid_scatter", "neighbor_mean",
          "blur_horizontal", "blur_vertical_and_maximum", "normalize"]


class BtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


class bt_options(C.Structure):
    _fields_ = [("map_w", C.c_uint32), ("map_h", C.c_uint32), ("radius_texels", C.c_uint32),
                ("flags", C.c_uint32), ("max_graph_memory_bytes", C.c_uint64),
                ("device", C.c_int32), ("spacing", C.c_float), ("radius0", C.c_float),
                ("radius1", C.c_float), ("sigma0", C.c_float), ("sigma1", C.c_float),
                ("schedule_len", C.c_uint32), ("density_w", C.c_uint32),
                ("density_h", C.c_uint32), ("cell_split", C.c_uint32),
                ("x_split", C.c_uint32), ("reserved", C.c_uint32 * 2)]


class bt_iter_params(C.Structure):
    _fields_ = [("res_x", C.c_float), ("res_y", C.c_float), ("radius", C.c_float),
                ("cell", C.c_float), ("gx", C.c_uint32), ("gy", C.c_uint32),
                ("sigma", C.c_float), ("taps", C.c_uint32), ("density_w", C.c_uint32),
                ("density_h", C.c_uint32), ("cell_w", C.c_float), ("task_bins", C.c_uint32)]


class bt_info(C.Structure):
    _fields_ = [("jobs", C.c_uint32), ("samples", C.c_uint32), ("paths", C.c_uint32),
                ("iteration", C.c_uint32), ("capacity", C.c_uint64),
                ("reserved_bytes", C.c_uint64), ("device_bytes", C.c_uint64),
                ("split_tasks", C.c_uint32), ("split_parts", C.c_uint32),
                ("test_steps", C.c_uint64), ("neighbors", C.c_uint64),
                ("candidates", C.c_uint64), ("group_steps", C.c_uint64)]


class bt_render_options(C.Structure):
    _fields_ = [("palette", C.c_uint32), ("color_by", C.c_uint32), ("msaa", C.c_uint32),
                ("endpoint_dots", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class bt_noisiness(C.Structure):
    _fields_ = [("colored_pixels", C.c_uint64), ("background_pixels", C.c_uint64),
                ("total_pixels", C.c_uint64), ("noisiness", C.c_double),
                ("colored_to_background_ratio", C.c_double)]


class bt_graph_desc(C.Structure):
    _fields_ = [("xy", C.POINTER(C.c_float)), ("n_points", C.c_uint32),
                ("starts", C.POINTER(C.c_uint32)), ("n_starts", C.c_uint32),
                ("opts", C.POINTER(bt_options))]


class bt_server_options(C.Structure):
    _fields_ = [("defaults", bt_options), ("memory_budget", C.c_uint64),
                ("max_inflight", C.c_uint32), ("policy", C.c_uint32),
                ("tick_samples", C.c_uint64), ("lane_samples", C.c_uint64),
                ("small_cu_mask", C.c_uint32), ("large_cu_mask", C.c_uint32)]


class bt_server_info(C.Structure):
    _fields_ = [("queued", C.c_uint32), ("running", C.c_uint32), ("finishing", C.c_uint32),
                ("reserved", C.c_uint32), ("footprint_bytes", C.c_uint64),
                ("held_bytes", C.c_uint64), ("held_peak_bytes", C.c_uint64),
                ("memory_budget", C.c_uint64)]


class bt_job_result(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("state", C.c_uint32), ("samples", C.c_uint32),
                ("iterations_done", C.c_uint32), ("iterations_total", C.c_uint32),
                ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double)]


# Every symbol include/bundling.h declares, with its ctypes signature.
_P, _U32, _U64, _SZ, _F, _D = C.c_void_p, C.c_uint32, C.c_uint64, C.c_size_t, C.c_float, C.c_double
_PU32, _PU64, _PF, _PD = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_float), C.POINTER(C.c_double)
SIGNATURES = {
    "bt_options_default": (None, [C.POINTER(bt_options)]),
    "bt_create": (C.c_int, [C.POINTER(bt_options), C.POINTER(_P)]),
    "bt_destroy": (None, [_P]),
    "bt_last_error": (C.c_char_p, [_P]),
    "bt_upload": (C.c_int, [_P, _PF, _U32, _PU32, _U32, _PU32]),
    "bt_iterate": (C.c_int, [_P]),
    "bt_iterate_n": (C.c_int, [_P, _U32]),
    "bt_rewind": (C.c_int, [_P]),
    "bt_read_paths": (C.c_int, [_P, _PU32, _SZ]),
    "bt_read_points": (C.c_int, [_P, _P, _SZ]),
    "bt_read_density": (C.c_int, [_P, _U32, _P, _SZ]),
    "bt_read_grid": (C.c_int, [_P, _PU32, _SZ, _PU32, _SZ]),
    "bt_read_endpoints": (C.c_int, [_P, _PU32, _SZ, _PU32]),
    "bt_read_endpoints_noalloc": (C.c_int, [_P, _PU32, _SZ, _PU32]),
    "bt_palette": (C.c_int, [_U32, _PF, _PF]),
    "bt_render_options_default": (None, [C.POINTER(bt_render_options)]),
    "bt_render_size": (C.c_int, [_P, _PU32, _PU32]),
    "bt_render": (C.c_int, [_P, C.POINTER(bt_render_options), _P, _SZ]),
    "bt_render_job": (C.c_int, [_P, _U32, C.POINTER(bt_render_options), _P, _SZ]),
    "bt_render_size_job": (C.c_int, [_P, _U32, _PU32, _PU32]),
    "bt_get_info": (C.c_int, [_P, C.POINTER(bt_info)]),
    "bt_iter_params_for": (C.c_int, [C.POINTER(bt_options), _U32, C.POINTER(bt_iter_params)]),
    "bt_profile": (C.c_int, [_P, _U32, _PD, _U32, _PD]),
    "bt_autotune": (C.c_int, [_P, _U32, _PU32, _U32, _PU32, _PD]),
    "bt_job_iter_params": (C.c_int, [_P, _U32, _U32, C.POINTER(bt_iter_params)]),
    "bt_copy_probe": (C.c_int, [_P, _U64, _PD]),
    "bt_upload_batch": (C.c_int, [_P, C.POINTER(bt_graph_desc), _U32, _PU32]),
    "bt_read_paths_job": (C.c_int, [_P, _U32, _PU32, _SZ]),
    "bt_read_points_job": (C.c_int, [_P, _U32, _P, _SZ]),
    "bt_read_positions_job": (C.c_int, [_P, _U32, _PF, _PU32, _SZ]),
    "bt_read_density_job": (C.c_int, [_P, _U32, _U32, _P, _SZ]),
    "bt_job_samples": (C.c_int, [_P, _U32, _PU32, _PU32]),
    "bt_trl_parse": (C.c_int, [C.c_char_p, _SZ, C.POINTER(_PF), _PU32, C.POINTER(_PU32), _PU32,
                               _PF, C.c_char_p, _SZ]),
    "bt_free": (None, [_P]),
    "bt_noisiness_background": (C.c_int, [_P, _U32, _U32, _P, C.c_char_p, _SZ]),
    "bt_noisiness_measure": (C.c_int, [_P, _U32, _U32, _P, C.POINTER(bt_noisiness),
                                       C.c_char_p, _SZ]),
    "bt_write_png": (C.c_int, [C.c_char_p, _P, _U32, _U32, C.c_char_p, _SZ]),
    "bt_normalization": (C.c_int, [_PF, _PF]),
    "bt_capacity": (C.c_int, [C.POINTER(bt_options), _U32, _U32, _PU64, _PU64]),
    "bt_blur_weights": (_U32, [_F, _PF, _U32]),
    "bt_distortion": (C.c_int, [_PF, _PF, _PU32, _U32, _U32, _U32, _PD]),
    "bt_server_create": (C.c_int, [C.POINTER(bt_server_options), C.POINTER(_P)]),
    "bt_server_destroy": (None, [_P]),
    "bt_server_last_error": (C.c_char_p, [_P]),
    "bt_server_submit": (C.c_int, [_P, _PF, _U32, _PU32, _U32, C.POINTER(bt_options), _U32, _PU64]),
    "bt_server_step": (C.c_int, [_P, _PU32, _PU32]),
    "bt_server_poll": (C.c_int, [_P, _U64, C.POINTER(bt_job_result)]),
    "bt_server_read_points": (C.c_int, [_P, _U64, _P, _SZ]),
    "bt_server_read_density": (C.c_int, [_P, _U64, _U32, _P, _SZ]),
    "bt_server_release": (C.c_int, [_P, _U64]),
    "bt_server_clock_ms": (_D, [_P]),
    "bt_server_get_info": (C.c_int, [_P, C.POINTER(bt_server_info)]),
    "bt_autotune_shapes": (C.c_int, [_P, _U32, _PU32, _PD]),
    "bt_set_shapes": (C.c_int, [_P, _U32, _U32, _U32, _U32]),
    "bt_get_shapes": (C.c_int, [_P, _U32, _U32, _PU32]),
    "bt_server_take_finished": (C.c_int, [_P, _PU64, _SZ, C.POINTER(_SZ)]),
}

_lib = None


def lib():
    """Loads libbundling.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} is missing: build it with "
                              "`python -m bundling_thingy_amd.build`")
        L = C.CDLL(_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f32p(a):
    return a.ctypes.data_as(_PF)


def _u32p(a):
    return a.ctypes.data_as(_PU32)


# --------------------------------------------------------------------- options

class Options:
    """Reference Options (main.rs:25-90) plus generalized parameters.

    Reference mode (spacing=0): map_size x map_size map, samples 1 texel apart,
    integer ``neighbor_radius`` texels, density = map, sigma 6.
    Generalized mode (spacing>0): canvas-unit spacing and radius schedule
    radius0 -> radius1 (geometric over schedule_len iterations), density
    density_w x density_h with sigma0 -> sigma1.
    """

    def __init__(self, map_size=1600, neighbor_radius=None, max_graph_memory_mib=1024,
                 color_by_density=True, evaluate_distortion=False, device=0, spacing=0.0,
                 radius=None, sigma=None, schedule_len=1, density=None, cell_split=0,
                 map_h=None, x_split=0, density_on_read=False):
        o = bt_options()
        lib().bt_options_default(C.byref(o))
        o.map_w = map_size
        o.map_h = map_h or map_size
        o.radius_texels = neighbor_radius or 0
        o.max_graph_memory_bytes = int(max_graph_memory_mib * 1024 * 1024)
        o.flags = (FLAG_DENSITY if color_by_density else 0) | (
            FLAG_DISTORTION if evaluate_distortion else 0) | (
            FLAG_DENSITY_ON_READ if density_on_read else 0)
        o.device = device
        o.spacing = spacing
        if radius is not None:
            r0, r1 = radius if isinstance(radius, (tuple, list)) else (radius, radius)
            o.radius0, o.radius1 = r0, r1
        if sigma is not None:
            s0, s1 = sigma if isinstance(sigma, (tuple, list)) else (sigma, sigma)
            o.sigma0, o.sigma1 = s0, s1
        o.schedule_len = schedule_len
        if density is not None:
            dw, dh = density if isinstance(density, (tuple, list)) else (density, density)
            o.density_w, o.density_h = dw, dh
        o.cell_split = cell_split
        o.x_split = x_split
        self.c = o

    def iter_params(self, iteration):
        p = bt_iter_params()
        rc = lib().bt_iter_params_for(C.byref(self.c), iteration, C.byref(p))
        if rc:
            raise BtError(rc, "invalid options")
        return p


# ----------------------------------------------------------------------- graph

class Normalization:
    def __init__(self, minimum, scale, translation):
        self.minimum, self.scale, self.translation = minimum, scale, translation


class Graph:
    """CSR polylines: points (P, 2) float32, starts (E + 1,) uint32 (graph.rs:7-12)."""

    def __init__(self, points, starts, bbox=None):
        self.points = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 2)
        self.starts = np.ascontiguousarray(starts, dtype=np.uint32)
        if bbox is None:
            bbox = np.concatenate([self.points.min(0), self.points.max(0)])
        self.bbox = np.asarray(bbox, dtype=np.float32)

    @classmethod
    def parse(cls, text):
        """graph.rs:35-77."""
        b = text.encode() if isinstance(text, str) else bytes(text)
        xy, st = _PF(), _PU32()
        n, ns = C.c_uint32(), C.c_uint32()
        bbox = (C.c_float * 4)()
        err = C.create_string_buffer(256)
        rc = lib().bt_trl_parse(b, len(b), C.byref(xy), C.byref(n), C.byref(st), C.byref(ns),
                                bbox, err, 256)
        if rc:
            raise BtError(rc, err.value.decode())
        try:
            pts = np.ctypeslib.as_array(xy, shape=(n.value * 2,)).copy().reshape(-1, 2)
            starts = np.ctypeslib.as_array(st, shape=(ns.value,)).copy()
        finally:
            lib().bt_free(C.cast(xy, _P))
            lib().bt_free(C.cast(st, _P))
        return cls(pts, starts, np.array(bbox[:], dtype=np.float32))

    @classmethod
    def load(cls, path):
        with open(path, "rb") as f:
            return cls.parse(f.read())

    def normalization(self):
        """graph.rs:79-91."""
        out = (C.c_float * 5)()
        bb = np.ascontiguousarray(self.bbox, dtype=np.float32)
        rc = lib().bt_normalization(_f32p(bb), out)
        if rc:
            raise BtError(rc, "graph has invalid extent")
        return Normalization((out[0], out[1]), out[2], (out[3], out[4]))

    @property
    def edge_count(self):
        return len(self.starts) - 1


# ------------------------------------------------------------------- images

def _image_args(rgba):
    a = np.ascontiguousarray(rgba, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected a (height, width, 4) RGBA array")
    return a, a.shape[1], a.shape[0]


def palette(name):
    """(16 x 3 float32 stops, endpoint RGB) of a palette (palette.rs:26-130)."""
    cols = np.zeros((16, 3), dtype=np.float32)
    ep = np.zeros(3, dtype=np.float32)
    idx = PALETTES[name] if isinstance(name, str) else int(name)
    rc = lib().bt_palette(idx, _f32p(cols), _f32p(ep))
    if rc:
        raise BtError(rc, "unknown palette")
    return cols, ep


def noisiness_background(blank):
    """noisiness::background (noisiness.rs:28-42): RGB of an opaque uniform frame."""
    a, w, h = _image_args(blank)
    rgb = (C.c_uint8 * 3)()
    err = C.create_string_buffer(256)
    rc = lib().bt_noisiness_background(a.ctypes.data_as(_P), w, h, C.cast(rgb, _P), err, 256)
    if rc:
        raise BtError(rc, err.value.decode())
    return tuple(rgb)


def noisiness(rgba, rgb):
    """noisiness::measure (noisiness.rs:44-74) as a dict of the reference's fields."""
    a, w, h = _image_args(rgba)
    bg = (C.c_uint8 * 3)(*rgb)
    m = bt_noisiness()
    err = C.create_string_buffer(256)
    rc = lib().bt_noisiness_measure(a.ctypes.data_as(_P), w, h, C.cast(bg, _P), C.byref(m), err,
                                    256)
    if rc:
        raise BtError(rc, err.value.decode())
    ratio = m.colored_to_background_ratio
    return {"colored_pixels": m.colored_pixels, "background_pixels": m.background_pixels,
            "total_pixels": m.total_pixels, "noisiness": m.noisiness,
            "colored_to_background_ratio": None if ratio != ratio else ratio}


def write_png(path, rgba):
    """write_image (main.rs:193-221): 8-bit RGBA PNG."""
    a, w, h = _image_args(rgba)
    err = C.create_string_buffer(256)
    rc = lib().bt_write_png(os.fsencode(path), a.ctypes.data_as(_P), w, h, err, 256)
    if rc:
        raise BtError(rc, err.value.decode())


# ------------------------------------------------------------------- geometry

class Layout:
    """Validated path reservations (geometry.rs:28-73)."""

    def __init__(self, paths, samples):
        paths = np.asarray(paths, dtype=np.int64).reshape(-1, 3)
        self.paths = paths
        self.samples = int(samples)
        keep = paths[:, 2] != 0
        self.skipped = int((~keep).sum())
        if np.any(paths[keep, 2] < 2):
            i = int(np.nonzero(keep & (paths[:, 2] < 2))[0][0])
            raise BtError(BT_E_INVALID, f"edge {i} has only one sample")
        ids = np.nonzero(keep)[0]
        if len(ids) == 0:
            raise BtError(BT_E_INVALID, "no eligible edges")
        if self.samples < len(ids):
            raise BtError(BT_E_INVALID, "sample count below edge count")
        self.edges = ids[np.argsort(paths[ids, 0], kind="stable")]
        for col, expected, width in ((0, self.samples, 0), (1, self.samples - len(ids), 1)):
            order = np.argsort(paths[ids, col], kind="stable")
            st = paths[ids, col][order]
            en = st + paths[ids, 2][order] - width
            if st[0] != 0 or np.any(st[1:] != en[:-1]) or en[-1] != expected:
                raise BtError(BT_E_INVALID,
                              "geometry ranges have a gap, overlap or exceed active count")

    @property
    def edge_count(self):
        return len(self.edges)

    def decode(self, rec12):
        """Decoder::push + finish (geometry.rs:131-190): checks ids and finiteness."""
        rec = np.frombuffer(rec12, dtype=np.uint8)
        if rec.size % 12:
            raise BtError(BT_E_INVALID, "partial GPU point record")
        r = rec.view(np.dtype([("x", "<f4"), ("y", "<f4"), ("id", "<u4")]))
        if len(r) != self.samples:
            raise BtError(BT_E_INVALID, "missing GPU point records")
        owner = np.empty(self.samples, dtype=np.int64)
        for e in self.edges:
            o, _, c = self.paths[e]
            owner[o:o + c] = e
        bad = np.nonzero(r["id"] != owner)[0]
        if len(bad):
            raise BtError(BT_E_INVALID, f"point {int(bad[0])} has incorrect edge ID")
        xy = np.stack([r["x"], r["y"]], 1)
        if not np.all(np.isfinite(xy)):
            raise BtError(BT_E_INVALID, "nonfinite coordinates")
        return Snapshot(self, xy.astype(np.float32))


class Snapshot:
    def __init__(self, layout, positions):
        self.layout = layout
        self.positions = positions

    def edges(self):
        """(edge id, (count, 2) positions) per eligible edge (geometry.rs:197-206)."""
        for e in self.layout.edges:
            o, _, c = self.layout.paths[e]
            yield int(e), self.positions[o:o + c]


def distortion(original, current, map_size=1600):
    """Aggregate distortion of `current` against `original` (distortion.rs:170-216).

    Returns (distortion, rms_displacement, rms_displacement_pixels)."""
    L = original.layout
    rec3 = np.ascontiguousarray(L.paths, dtype=np.uint32)
    a = np.ascontiguousarray(original.positions, dtype=np.float32)
    b = np.ascontiguousarray(current.positions, dtype=np.float32)
    out = (C.c_double * 3)()
    rc = lib().bt_distortion(_f32p(a), _f32p(b), _u32p(rec3), len(rec3), L.samples, map_size, out)
    if rc:
        raise BtError(rc, "invalid distortion input")
    return out[0], out[1], out[2]


# ------------------------------------------------------------------------ gpu

class Gpu:
    """The compute half of the reference's Gpu (gpu.rs:59-553) on one MI355X."""

    def __init__(self, options=None):
        self.options = options or Options()
        h = _P()
        rc = lib().bt_create(C.byref(self.options.c), C.byref(h))
        if rc:
            raise BtError(rc, "device initialization failed")
        self._h = h
        self._graphs = []

    def close(self):
        if getattr(self, "_h", None):
            lib().bt_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc:
            raise BtError(rc, lib().bt_last_error(self._h).decode())

    # -- drop-in surface
    def upload(self, graph):
        s = C.c_uint32()
        self._check(lib().bt_upload(self._h, _f32p(graph.points), len(graph.points),
                                    _u32p(graph.starts), len(graph.starts), C.byref(s)))
        self._graphs = [graph]
        self._opts = [self.options]
        return s.value

    def upload_batch(self, graphs, options=None):
        """Uploads many graphs as independent jobs advanced together."""
        descs = (bt_graph_desc * len(graphs))()
        keep = []
        for i, g in enumerate(graphs):
            o = options[i].c if options else None
            if o is not None:
                keep.append(o)
            descs[i] = bt_graph_desc(_f32p(g.points), len(g.points), _u32p(g.starts),
                                     len(g.starts), C.pointer(o) if o is not None else None)
        samples = np.zeros(len(graphs), dtype=np.uint32)
        self._check(lib().bt_upload_batch(self._h, descs, len(graphs), _u32p(samples)))
        self._graphs = list(graphs)
        self._opts = list(options) if options else [self.options] * len(graphs)
        return samples

    def iterate(self, n=1, graph=False):
        if graph:
            self._check(lib().bt_iterate_n(self._h, n))
        else:
            for _ in range(n):
                self._check(lib().bt_iterate(self._h))

    def rewind(self):
        self._check(lib().bt_rewind(self._h))

    def info(self):
        i = bt_info()
        self._check(lib().bt_get_info(self._h, C.byref(i)))
        return i

    def job_samples(self, job):
        s, it = C.c_uint32(), C.c_uint32()
        self._check(lib().bt_job_samples(self._h, job, C.byref(s), C.byref(it)))
        return s.value, it.value

    def paths(self, job=0):
        n = len(self._graphs[job].starts) - 1
        out = np.zeros((n, 3), dtype=np.uint32)
        self._check(lib().bt_read_paths_job(self._h, job, _u32p(out), n))
        return out

    def points_raw(self, job=0, out=None):
        """The job's 12-byte sample records, into `out` (a contiguous buffer of
        at least 12 N bytes, reused by a serving loop) or a new array."""
        n, _ = self.job_samples(job)
        buf = np.empty(n * 12, dtype=np.uint8) if out is None else \
            np.frombuffer(out, dtype=np.uint8, count=n * 12)
        self._check(lib().bt_read_points_job(self._h, job, buf.ctypes.data_as(_P), n))
        return buf

    def points(self, job=0):
        """(positions (N, 2) float32, edge ids (N,) uint32) in Layout order:
        the fields of points_raw, split on the device."""
        n, _ = self.job_samples(job)
        xy = np.empty((n, 2), dtype=np.float32)
        ids = np.empty(n, dtype=np.uint32)
        self._check(lib().bt_read_positions_job(self._h, job, _f32p(xy), _u32p(ids), n))
        return xy, ids

    def distortion_layout(self, job=0):
        n, _ = self.job_samples(job)
        return Layout(self.paths(job), n)

    def distortion_snapshot(self, layout, job=0):
        return layout.decode(self.points_raw(job).tobytes())

    def density(self, kind=DENSITY_NORMALIZED, job=0, out=None):
        """The job's density plane `kind` (h, w), into `out` (a contiguous
        buffer of at least w * h values of the kind's type) or a new array."""
        p = self._opts[job].iter_params(0)
        w, h = p.density_w, p.density_h
        dt = {DENSITY_COUNTS: np.uint32, DENSITY_HALF: np.uint16}.get(kind, np.float32)
        buf = np.zeros(w * h, dtype=dt) if out is None else \
            np.frombuffer(out, dtype=dt, count=w * h)
        self._check(lib().bt_read_density_job(self._h, job, kind, buf.ctypes.data_as(_P), w * h))
        return buf.reshape(h, w)

    def copy_probe(self, nbytes):
        """Device-to-device copy of nbytes (profiler calibration); returns ms."""
        ms = C.c_double()
        self._check(lib().bt_copy_probe(self._h, nbytes, C.byref(ms)))
        return ms.value

    def job_params(self, iteration, job=0):
        """The exact parameters job `job` runs iteration `iteration` with."""
        p = bt_iter_params()
        self._check(lib().bt_job_iter_params(self._h, job, iteration, C.byref(p)))
        return p

    def autotune(self, iterations, candidates=(1, 2, 3, 4, 6, 8)):
        """Times every candidate grid cell split per iteration; keeps the best."""
        cand = np.ascontiguousarray(candidates, dtype=np.uint32)
        chosen = np.zeros(iterations, dtype=np.uint32)
        best = np.zeros(iterations, dtype=np.float64)
        self._check(lib().bt_autotune(self._h, iterations, _u32p(cand), len(cand),
                                      _u32p(chosen), best.ctypes.data_as(_PD)))
        return chosen, best

    def autotune_shapes(self, iterations):
        """Times every kernel shape of the neighbor mean, row blur and column
        blur per iteration; keeps the fastest. Returns (shapes (n, 3), stage
        ms (n, 3))."""
        shapes = np.zeros((iterations, 3), dtype=np.uint32)
        ms = np.zeros((iterations, 3), dtype=np.float64)
        self._check(lib().bt_autotune_shapes(self._h, iterations, _u32p(shapes),
                                             ms.ctypes.data_as(_PD)))
        return shapes, ms

    def set_shapes(self, nm=0, row=0, col=0, iteration=0xffffffff):
        self._check(lib().bt_set_shapes(self._h, iteration, nm, row, col))

    def shapes(self, iteration, job=0):
        out = np.zeros(3, dtype=np.uint32)
        self._check(lib().bt_get_shapes(self._h, job, iteration, _u32p(out)))
        return tuple(int(v) for v in out)

    def grid(self):
        """(cell of each sample in Layout order, per-cell counts) of the last iteration."""
        n, it = self.job_samples(0)
        p = self.job_params(max(it - 1, 0))
        cells = np.zeros(n, dtype=np.uint32)
        counts = np.zeros(p.gx * p.gy, dtype=np.uint32)
        self._check(lib().bt_read_grid(self._h, _u32p(cells), n, _u32p(counts), len(counts)))
        return cells, counts.reshape(p.gy, p.gx)

    def render_size(self, job=0):
        w, h = C.c_uint32(), C.c_uint32()
        self._check(lib().bt_render_size_job(self._h, job, C.byref(w), C.byref(h)))
        return w.value, h.value

    def render(self, palette="viridis", color_by=COLOR_DENSITY, msaa=4, endpoint_dots=False,
               job=0):
        """Gpu::render + Gpu::readback (gpu.rs:572-651): the export image as a
        (height, width, 4) uint8 RGBA array, row 0 at the top. Without an
        uploaded graph: the blank frame."""
        ro = bt_render_options()
        ro.palette = PALETTES[palette] if isinstance(palette, str) else int(palette)
        ro.color_by = int(color_by)
        ro.msaa = int(msaa)
        ro.endpoint_dots = int(bool(endpoint_dots))
        w, h = self.render_size(job)
        out = np.zeros((h, w, 4), dtype=np.uint8)
        self._check(lib().bt_render_job(self._h, job, C.byref(ro), out.ctypes.data_as(_P),
                                        out.nbytes))
        return out

    def endpoints(self):
        """Sorted sample indices (Layout order) of the distinct path endpoint
        positions: the deduplicate pass of endpoints.wgsl:9-49 that feeds the
        endpoint dots (resampling.rs:137-149)."""
        count = C.c_uint32()
        rc = lib().bt_read_endpoints(self._h, None, 0, C.byref(count))
        if rc != BT_E_CAPACITY:
            self._check(rc)
        out = np.zeros(count.value, dtype=np.uint32)
        if count.value:
            self._check(lib().bt_read_endpoints(self._h, _u32p(out), len(out), C.byref(count)))
        return out

    def profile(self, n):
        ms = (C.c_double * 9)()
        tot = C.c_double()
        self._check(lib().bt_profile(self._h, n, ms, 9, C.byref(tot)))
        return dict(zip(STAGES, ms[:])), tot.value


# --------------------------------------------------------------------- helpers

def capacity(options, n_points, n_starts):
    """Allocation::new (resampling.rs:321-397): (sample capacity, reserved bytes)."""
    cap, res = C.c_uint64(), C.c_uint64()
    rc = lib().bt_capacity(C.byref(options.c), n_points, n_starts, C.byref(cap), C.byref(res))
    if rc:
        raise BtError(rc, "graph memory budget")
    return cap.value, res.value


def blur_weights(sigma, max_taps=64):
    w = np.zeros(max_taps + 1, dtype=np.float32)
    taps = lib().bt_blur_weights(sigma, _f32p(w), max_taps)
    return w[:taps + 1]


# ---------------------------------------------------------------------- served

class Server:
    """Continuous-batching server: memory-based admission, one batched launch
    per stage per tick over the advanced jobs (policy FIFO: every in-flight
    job; SHORTEST: least remaining work first, tick_samples per tick; LANES:
    jobs of at most lane_samples samples tick on one stream, larger ones on
    another, the two overlapping on the GPU; with tick_samples, each lane's
    tick takes the least remaining work first up to tick_samples while no
    job waits for admission; small_cu_mask / large_cu_mask keep the lanes on
    the CUs they select (one bit per CU of every group of 32) while both
    lanes have work)."""

    FIFO, SHORTEST, LANES = 0, 1, 2

    def __init__(self, options=None, memory_budget=64 << 30, max_inflight=64, policy=0,
                 tick_samples=0, lane_samples=0, small_cu_mask=0, large_cu_mask=0):
        so = bt_server_options()
        so.defaults = (options or Options()).c
        so.memory_budget = memory_budget
        so.max_inflight = max_inflight
        so.policy = policy
        so.tick_samples = tick_samples
        so.lane_samples = lane_samples
        so.small_cu_mask = small_cu_mask
        so.large_cu_mask = large_cu_mask
        h = _P()
        rc = lib().bt_server_create(C.byref(so), C.byref(h))
        if rc:
            raise BtError(rc, "server creation failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().bt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise BtError(rc, lib().bt_server_last_error(self._h).decode())

    def submit(self, graph, options, iterations):
        t = C.c_uint64()
        self._check(lib().bt_server_submit(self._h, _f32p(graph.points), len(graph.points),
                                           _u32p(graph.starts), len(graph.starts),
                                           C.byref(options.c), iterations, C.byref(t)))
        return t.value

    def step(self):
        r, q = C.c_uint32(), C.c_uint32()
        self._check(lib().bt_server_step(self._h, C.byref(r), C.byref(q)))
        return r.value, q.value

    def poll(self, ticket):
        res = bt_job_result()
        self._check(lib().bt_server_poll(self._h, ticket, C.byref(res)))
        return res

    def points(self, ticket, samples):
        r = self.points_raw(ticket, samples)
        return np.stack([r["x"], r["y"]], 1), r["id"].copy()

    def points_raw(self, ticket, samples, out=None):
        """The job's 12-byte sample records {x, y, edge id}, copied into `out`
        (a contiguous buffer of at least samples * 12 bytes, reused by a
        serving loop) or a new array."""
        rec = np.dtype([("x", "<f4"), ("y", "<f4"), ("id", "<u4")])
        buf = np.empty(samples * 12, dtype=np.uint8) if out is None else \
            np.frombuffer(out, dtype=np.uint8, count=samples * 12)
        self._check(lib().bt_server_read_points(self._h, ticket, buf.ctypes.data_as(_P), samples))
        return buf.view(rec)

    def density(self, ticket, w, h, out=None):
        """The job's normalized density (h, w), copied into `out` (a
        contiguous float32 buffer of at least w * h values) or a new array."""
        buf = np.empty(w * h, dtype=np.float32) if out is None else \
            np.frombuffer(out, dtype=np.float32, count=w * h)
        self._check(lib().bt_server_read_density(self._h, ticket, DENSITY_NORMALIZED,
                                                 buf.ctypes.data_as(_P), w * h))
        return buf.reshape(h, w)

    def release(self, ticket):
        self._check(lib().bt_server_release(self._h, ticket))

    def clock_ms(self):
        return lib().bt_server_clock_ms(self._h)

    def take_finished(self, capacity=4096):
        """Tickets finished (done or failed) since the last call, in order."""
        buf = (C.c_uint64 * capacity)()
        n = _SZ()
        self._check(lib().bt_server_take_finished(self._h, buf, capacity, C.byref(n)))
        return list(buf[:n.value])

    def info(self):
        """Admission accounting: queued/running/finishing jobs, the footprint
        charged against the budget, the bytes the admitted jobs hold."""
        i = bt_server_info()
        self._check(lib().bt_server_get_info(self._h, C.byref(i)))
        return i
