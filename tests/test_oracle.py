"""Pins the CPU oracle: reference golden vectors (blur table), closed forms,
brute force on tiny inputs and invariants. No GPU needed."""
import json
import math
import os

import numpy as np
import pytest

from tests import brute

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
F32 = np.float32


# ------------------------------------------------------------------ parsing

TRL_CASES = [
    # (text, points, starts) - graph.rs:35-77 rules
    ("0 0 0 1 1\n", [[0, 0], [1, 1]], [0, 2]),
    ("e 0 0 (x <y 1 1 2 2\n", [[0, 0], [1, 1], [2, 2]], [0, 3]),            # ( and < tokens skipped
    ("e 0 0 0 0 1 1 1 1 2 0\n", [[0, 0], [1, 1], [2, 0]], [0, 3]),          # consecutive duplicates dropped
    ("a 5 5\nb 0 0 1 0\n", [[0, 0], [1, 0]], [0, 2]),                       # < 2 points dropped
    ("a 0 0  1 0\r\nb 2 2 3 3", [[0, 0], [1, 0], [2, 2], [3, 3]], [0, 2, 4]),  # double space, CRLF, no final newline
    ("a 0 0 1 0 0 0\n", [[0, 0], [1, 0], [0, 0]], [0, 3]),                  # non-consecutive repeat kept
    ("a 1e-1 +2 .5 3.\n", [[0.1, 2], [0.5, 3]], [0, 2]),                    # Rust float grammar
]

TRL_ERRORS = [
    "a 0 0 1\n",          # missing y
    "a 0 0 1 x\n",        # invalid y
    "a 0 0 inf 1\n",      # non-finite
    "a 0x10 0 1 1\n",     # hex is not a Rust float
    "a\n",                # no drawable edges
    "",
]


@pytest.mark.parametrize("text,pts,starts", TRL_CASES)
def test_parse_known_answers(orc, text, pts, starts):
    p, s, bbox = orc.parse_trl(text)
    np.testing.assert_array_equal(p, np.array(pts, dtype=np.float32))
    np.testing.assert_array_equal(s, np.array(starts, dtype=np.uint32))
    np.testing.assert_array_equal(bbox, np.concatenate([p.min(0), p.max(0)]))


@pytest.mark.parametrize("text", TRL_ERRORS)
def test_parse_errors(orc, text):
    with pytest.raises(ValueError):
        orc.parse_trl(text)


def test_normalization_closed_form(orc):
    nm = orc.normalization(np.array([1.0, 2.0, 5.0, 4.0], dtype=np.float32))
    assert nm.scale == F32(0.95) / F32(4.0)
    assert nm.tr_x == (F32(1) - F32(nm.scale) * F32(4)) / F32(2)
    assert math.isclose(nm.tr_y, (1 - 0.2375 * 2) / 2, rel_tol=1e-6)
    with pytest.raises(ValueError):
        orc.normalization(np.array([1, 1, 1, 1], dtype=np.float32))


# --------------------------------------------------------------- resampling

def _straight(p0, p1):
    return np.array([p0, p1], dtype=np.float32), np.array([0, 2], dtype=np.uint32)


def test_resample_straight_edge_closed_form(orc):
    # A horizontal edge spanning the whole normalized map: clip x from -0.95 to 0.95.
    res = 200.0
    pts, st = _straight([0, 0], [1, 0])
    bbox = np.array([0, 0, 1, 0], dtype=np.float32)
    counts, off, out, eid, status = orc.resample(pts, st, bbox, res, res)
    assert status == 0
    # normalized endpoints, resampling.wgsl:45-48 in float32
    nm = orc.normalization(bbox)
    x0 = F32((F32(F32(F32(0) - F32(nm.min_x)) * F32(nm.scale)) + F32(nm.tr_x) - F32(0.5)) * F32(2))
    x1 = F32((F32(F32(F32(1) - F32(nm.min_x)) * F32(nm.scale)) + F32(nm.tr_x) - F32(0.5)) * F32(2))
    ex = F32(F32(x1 - x0) * F32(100.0))
    D = float(np.sqrt(F32(ex * ex)))  # length in texels
    # samples at arc 0, 1, ..., floor(D), then the endpoint unless equal to the last
    last = F32(x0 + F32(F32(x1 - x0) * F32(F32(math.floor(D)) / F32(D))))
    expected = 1 + math.floor(D) + (0 if last == x1 else 1)
    assert counts[0] == len(out) == expected
    assert out[0, 0] == x0 and out[-1, 0] == x1
    # samples one texel apart (1/100 clip units) except the final remainder
    d = np.diff(out[:, 0].astype(np.float64)) * 100.0
    np.testing.assert_allclose(d[:-1], 1.0, atol=1e-4)
    assert 0 < d[-1] <= 1.0 + 1e-4
    assert np.all(eid == 0)


def test_resample_count_matches_emit_random(orc):
    rng = np.random.default_rng(7)
    pts = rng.uniform(0, 1, (300, 2)).astype(np.float32)
    st = np.array([0, 2, 5, 9, 10 + 2, 20, 40, 41 + 1, 100, 300], dtype=np.uint32)
    counts, off, out, eid, status = orc.resample(pts, st, np.concatenate([pts.min(0), pts.max(0)]),
                                                 117.0, 117.0)
    assert status == 0
    assert len(out) == counts.sum()
    for e in range(len(st) - 1):
        seg = out[off[e]:off[e] + counts[e]]
        assert np.all(np.any(seg[1:] != seg[:-1], axis=1)), "consecutive duplicates"
        assert np.all(eid[off[e]:off[e] + counts[e]] == e)


def test_resample_closed_path_special_case(orc):
    # A short closed path (A, B, A) shorter than a texel keeps its excursion
    # and returning endpoint (resampling.wgsl:124-134).
    pts = np.array([[0, 0], [1e-4, 0], [0, 0], [1, 1], [2, 2]], dtype=np.float32)
    st = np.array([0, 3, 5], dtype=np.uint32)
    counts, off, out, eid, status = orc.resample(pts, st, np.array([0, 0, 2, 2], np.float32),
                                                 100.0, 100.0)
    assert status == 0 and counts[0] == 3
    a = out[off[0]:off[0] + 3]
    assert np.all(a[0] == a[2]) and np.any(a[1] != a[0])


def _arc_positions(orc, pts, bbox, res, out):
    """Arc length (texels, float64) of each sample along the polyline `pts`
    after the reference's normalization; each sample is placed on the
    segment nearest to it."""
    nm = orc.normalization(bbox)
    c = ((pts.astype(np.float64) - [nm.min_x, nm.min_y]) * nm.scale + [nm.tr_x, nm.tr_y]
         - 0.5) * 2.0 * (res / 2.0)                     # control points in texels
    seg = np.diff(c, axis=0)
    lens = np.linalg.norm(seg, axis=1)
    cum = np.concatenate([[0.0], np.cumsum(lens)])
    q = out.astype(np.float64) * (res / 2.0)
    arc = np.empty(len(q))
    for i, p in enumerate(q):
        t = np.clip(np.einsum("ij,ij->i", p - c[:-1], seg) / lens ** 2, 0.0, 1.0)
        d = np.linalg.norm(c[:-1] + t[:, None] * seg - p, axis=1)
        k = int(np.argmin(d))
        assert d[k] < 1e-3, "sample off the polyline"
        arc[i] = cum[k] + t[k] * lens[k]
    return arc, cum


@pytest.mark.parametrize("shape", ["L", "zigzag"])
def test_resample_polyline_carries_remainder_across_joints(orc, shape):
    """resampling.wgsl:91-116: `remaining` carries over each joint, so the
    samples of a polyline lie exactly one texel of arc length apart across
    joints, not one texel from each segment start; the first sample is the
    first control point and the last the last one. Segment lengths are not
    whole texels, so resetting the remainder per segment breaks the spacing
    at every joint."""
    if shape == "L":
        pts = np.array([[0, 0], [0.52, 0], [0.52, 0.37]], dtype=np.float32)
    else:
        x = np.linspace(0.0, 1.0, 9)
        y = np.where(np.arange(9) % 2 == 0, 0.1, 0.43) + 0.021 * np.arange(9)
        pts = np.stack([x, y], 1).astype(np.float32)
    st = np.array([0, len(pts)], dtype=np.uint32)
    bbox = np.concatenate([pts.min(0), pts.max(0)]).astype(np.float32)
    res = 117.0
    counts, off, out, eid, status = orc.resample(pts, st, bbox, res, res)
    assert status == 0 and counts[0] == len(out)
    arc, cum = _arc_positions(orc, pts, bbox, res, out)
    frac = cum[1:-1] - np.floor(cum[1:-1])
    assert np.all((frac > 0.05) & (frac < 0.95)), "joints must fall between whole texels"
    gaps = np.diff(arc)
    np.testing.assert_allclose(gaps[:-1], 1.0, atol=2e-3)
    assert 0.0 < gaps[-1] <= 1.0 + 2e-3
    # one sample per whole texel of arc, then the endpoint
    assert counts[0] == 1 + int(np.floor(cum[-1])) + 1
    np.testing.assert_allclose(arc[:-1], np.arange(len(arc) - 1), atol=2e-3)
    assert arc[-1] == pytest.approx(cum[-1], abs=1e-3)


def test_polyline_arc_check_detects_reset_remainder(orc):
    """The check above fails for samples placed one texel from each
    segment start (the remainder reset per segment), built here from the
    same L-shaped path."""
    pts = np.array([[0, 0], [0.52, 0], [0.52, 0.37]], dtype=np.float32)
    bbox = np.concatenate([pts.min(0), pts.max(0)]).astype(np.float32)
    res = 117.0
    _, _, out, _, _ = orc.resample(pts, np.array([0, 3], np.uint32), bbox, res, res)
    arc, cum = _arc_positions(orc, pts, bbox, res, out)
    wrong = np.concatenate([np.arange(0.0, cum[1], 1.0), cum[1] + np.arange(0.0, cum[2] - cum[1], 1.0),
                            [cum[2]]])
    gaps = np.diff(wrong)
    assert not np.allclose(gaps[:-1], 1.0, atol=2e-3)


# --------------------------------------------------------------------- grid

def test_grid_bin_histogram_bruteforce(orc):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1, 1, (5000, 2)).astype(np.float32)
    g = orc.make_grid(200.0, 10.0)
    assert (g.gx, g.gy) == (21, 21)  # ceil(201 / 10), neighbor_mean.rs:27
    cell, counts = orc.grid_bin(pts, g)
    t = ((pts + F32(1)) * F32(0.5)) * F32(200.0)
    c = np.minimum((np.maximum(t / F32(10.0), 0)).astype(np.uint32), 20)
    np.testing.assert_array_equal(cell, c[:, 1] * 21 + c[:, 0])
    np.testing.assert_array_equal(counts, np.bincount(cell, minlength=441))
    assert counts.sum() == len(pts)


# ------------------------------------------------------------ neighbor mean

def _tiny(seed, n=160):
    rng = np.random.default_rng(seed)
    # two clusters and a few stragglers in clip space
    a = rng.normal([-0.3, 0.1], 0.05, (n // 2, 2))
    b = rng.normal([0.25, -0.2], 0.08, (n // 2 - 4, 2))
    c = rng.uniform(-0.9, 0.9, (4, 2))
    pts = np.concatenate([a, b, c]).astype(np.float32)
    fixed = (rng.random(len(pts)) < 0.1).astype(np.uint8)
    return pts, fixed


@pytest.mark.parametrize("seed,res,radius", [(0, 200.0, 10.0), (1, 500.0, 30.0), (2, 500.0, 5.3)])
def test_neighbor_mean_bitexact_vs_bruteforce(orc, seed, res, radius):
    pts, fixed = _tiny(seed)
    g = orc.make_grid(res, radius)
    got = orc.neighbor_mean(pts, fixed, g, threads=2)
    want = brute.neighbor_mean_bruteforce(pts, fixed, res, res, radius)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("cell,cell_w", [(10.0, None), (5.0, None), (2.5, None), (33.0, None),
                                         (10.0, 1.25), (2.5, 0.3125), (5.0, 0.7)])
def test_neighbor_mean_independent_of_cell_size(orc, cell, cell_w):
    pts, fixed = _tiny(5, 400)
    a = orc.neighbor_mean(pts, fixed, orc.make_grid(200.0, 10.0))
    b = orc.neighbor_mean(pts, fixed, orc.make_grid(200.0, 10.0, cell, cell_w))
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("seed,res,radius,cell", [(0, 200.0, 10.0, None), (1, 500.0, 30.0, None),
                                                 (2, 500.0, 5.3, 2.0), (3, 200.0, 40.0, None)])
def test_cell_summary_shortcut_bitexact_vs_bruteforce(orc, seed, res, radius, cell):
    """The reference's tiled algorithm (fully-inside cells contribute their
    summary, the others are tested point by point) gives the brute-force
    bits: tiny clusters, where many cells are fully inside, and every
    stride-th sample of the plain CPU case."""
    pts, fixed = _tiny(seed, 600)
    g = orc.make_grid(res, radius, cell)
    want = brute.neighbor_mean_bruteforce(pts, fixed, res, res, radius)
    got = orc.neighbor_mean_cells(pts, fixed, g, 1, threads=2)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_cell_summary_shortcut_equals_window_scan_on_graph(orc):
    from bundling_thingy_amd import workloads
    from bundling_thingy_amd.api import Options
    g = workloads.plain_cpu_graph()
    opts = Options(map_size=201, neighbor_radius=10)
    p = opts.iter_params(0)
    pl = orc.Pipeline(g.points, g.starts, g.bbox, p)
    pts = pl.points
    for _ in range(2):
        grid = orc.make_grid(p.res_x, p.radius)  # the reference's cells: side = radius
        a = orc.neighbor_mean(pts, pl.fixed, grid)
        b = orc.neighbor_mean_cells(pts, pl.fixed, grid, 1)
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        c = orc.neighbor_mean_cells(pts, pl.fixed, grid, 7)
        k = np.arange(len(pts))
        sel = (k % 7 == 0) & (pl.fixed == 0)
        np.testing.assert_array_equal(c[sel].view(np.uint32), a[sel].view(np.uint32))
        np.testing.assert_array_equal(c[~sel].view(np.uint32), pts[~sel].view(np.uint32))
        pts = a


def test_neighbor_mean_close_to_f64_and_reference_f32(orc):
    pts, fixed = _tiny(9, 600)
    g = orc.make_grid(200.0, 12.0)
    got = orc.neighbor_mean(pts, fixed, g).astype(np.float64)
    f64 = brute.neighbor_mean_f64(pts, fixed, 200.0, 200.0, 12.0)
    ref32 = orc.neighbor_mean_f32(pts, fixed, g).astype(np.float64)
    diag = np.linalg.norm(pts.max(0) - pts.min(0))
    assert np.abs(got - f64).max() <= 1e-6 * diag
    assert np.abs(got - ref32).max() <= 1e-5 * diag


def test_neighbor_mean_invariants(orc):
    g = orc.make_grid(200.0, 10.0)
    # two parallel horizontal lines 0.05 clip apart (5 texels), radius 10 texels
    x = np.linspace(-0.5, 0.5, 101, dtype=np.float32)
    top = np.stack([x, np.full_like(x, 0.025)], 1)
    bot = np.stack([x, np.full_like(x, -0.025)], 1)
    iso = np.array([[0.9, 0.9]], dtype=np.float32)
    pts = np.concatenate([top, bot, iso])
    fixed = np.zeros(len(pts), np.uint8)
    fixed[[0, 100, 101, 201]] = 1
    out = orc.neighbor_mean(pts, fixed, g)
    # fixed endpoints and the isolated point do not move
    np.testing.assert_array_equal(out[fixed == 1], pts[fixed == 1])
    np.testing.assert_array_equal(out[-1], pts[-1])
    # the lines move toward each other symmetrically
    mid = slice(20, 81)
    assert np.all(out[mid, 1] < 0.025) and np.all(out[101:202][mid, 1] > -0.025)
    np.testing.assert_allclose(out[mid, 1], -out[101:202][mid, 1], atol=2e-7)
    # means stay inside the bounding box of the inputs
    assert np.all(out >= pts.min(0)) and np.all(out <= pts.max(0))


# ------------------------------------------------------------------ density

def test_blur_weights_match_reference_table(orc):
    ref = json.load(open(os.path.join(GOLDEN, "blur_sigma6.json")))
    w = orc.blur_weights(6.0)
    assert len(w) == ref["taps"] + 1
    np.testing.assert_array_equal(w, np.array(ref["weights"], dtype=np.float32))


@pytest.mark.parametrize("sigma", [2.0, 4.0, 6.0, 8.0, 5.3])
def test_blur_weights_normalized_gaussian(orc, sigma):
    w = orc.blur_weights(sigma).astype(np.float64)
    assert len(w) - 1 == math.ceil(3 * sigma)
    assert abs(w[0] + 2 * w[1:].sum() - 1.0) < 1e-7  # f32 rounding of 2K+1 taps
    k = np.arange(len(w))
    np.testing.assert_allclose(w / w[0], np.exp(-k ** 2 / (2 * sigma ** 2)), rtol=2e-6, atol=2e-9)


def test_splat_half_to_even(orc):
    # S = 5: pixel = round(uv * 4); x = -0.75 -> u = 0.125 -> 0.5 -> 0 (even),
    # x = -0.25 -> 1.5 -> 2, y = 0.25 -> v = 0.375 -> 1.5 -> 2
    pts = np.array([[-0.75, 0.0], [-0.25, 0.25], [1.0, -1.0]], dtype=np.float32)
    c = orc.splat(pts, 5, 5)
    assert c[2, 0] == 1 and c[2, 2] == 1 and c[4, 4] == 1 and c.sum() == 3


def test_blur_impulse_is_outer_product(orc):
    c = np.zeros((64, 64), dtype=np.uint32)
    c[30, 33] = 1
    rows, blurred, peak, norm = orc.density(c, 4.0)
    w = orc.blur_weights(4.0)
    t = len(w) - 1
    for dy in range(-t, t + 1):
        for dx in range(-t, t + 1):
            assert blurred[30 + dy, 33 + dx] == F32(w[abs(dx)]) * F32(w[abs(dy)])
    assert blurred.sum() == pytest.approx(1.0, abs=1e-6)
    assert peak == blurred.max() and norm.max() == 1.0
    np.testing.assert_array_equal(norm, np.clip(blurred / peak, 0, 1).astype(np.float32))


def test_blur_zero_padding_and_mass(orc):
    rng = np.random.default_rng(0)
    c = rng.integers(0, 5, (40, 50)).astype(np.uint32)
    rows, blurred, peak, norm = orc.density(c, 2.0)
    # direct float64 convolution with zero padding
    w = orc.blur_weights(2.0).astype(np.float64)
    k = np.concatenate([w[::-1], w[1:]])
    r = np.array([np.convolve(row, k, mode="same") for row in c.astype(np.float64)])
    b = np.array([np.convolve(col, k, mode="same") for col in r.T]).T
    np.testing.assert_allclose(blurred, b, rtol=1e-5, atol=1e-6)
    zero = np.zeros((3, 3), np.uint32)
    _, _, pk, nz = orc.density(zero, 2.0)
    assert pk == 0 and not nz.any()


# --------------------------------------------------------------- distortion

def test_distortion_identical_and_translation(orc):
    rng = np.random.default_rng(1)
    pts = rng.uniform(-0.5, 0.5, (50, 2)).astype(np.float32)
    off = np.array([0, 10, 30], dtype=np.uint64)
    cnt = np.array([10, 20, 20], dtype=np.uint32)
    d, r = orc.distortion(pts, pts, off, cnt)
    assert d == 0.0 and r == 0.0
    v = np.array([0.03, -0.04], dtype=np.float32)
    moved = (pts + v).astype(np.float32)
    d, r = orc.distortion(pts, moved, off, cnt)
    assert d == pytest.approx(0.05 / (2 * math.sqrt(2)), rel=1e-5)


# ---------------------------------------------------------------- endpoints

def test_endpoints_closed_form(orc):
    # Layout: path 0 = samples 0..2, path 1 = 3..4, path 2 = one sample (5,
    # skipped: count < 2), path 3 = 6..8. Sample 3 repeats sample 2, sample 6
    # is (-0, 0) = sample 0, sample 8 repeats sample 4.
    pts = np.array([[0, 0], [0.5, 0], [1, 0], [1, 0], [2, 2], [7, 7],
                    [-0.0, 0], [1, 1], [2, 2]], dtype=np.float32)
    paths = np.array([[0, 0, 3], [3, 2, 2], [5, 3, 1], [6, 3, 3]], dtype=np.uint32)
    assert orc.endpoints(pts, paths).tolist() == [0, 2, 4]
    assert orc.endpoints(pts, np.zeros((0, 3), np.uint32)).tolist() == []
    # a closed path's two ends share a position: one endpoint
    assert orc.endpoints(pts[[0, 1, 0]], [[0, 0, 3]]).tolist() == [0]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_endpoints_bruteforce(orc, seed):
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, 6, 300)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]])
    n = int(counts.sum())
    # few distinct positions so that endpoints collide often; some -0 signs
    pts = rng.integers(-3, 4, (n, 2)).astype(np.float32) * np.float32(0.25)
    neg = rng.random((n, 2)) < 0.3
    pts[neg & (pts == 0)] = np.float32(-0.0)
    paths = np.stack([offsets, offsets, counts], 1).astype(np.uint32)
    seen = []  # (x, y, smallest index) by pairwise == over every endpoint
    for off, _, c in paths.tolist():
        if c < 2:
            continue
        for i in (off, off + c - 1):
            for k, (x, y, j) in enumerate(seen):
                if x == pts[i, 0] and y == pts[i, 1]:
                    seen[k] = (x, y, min(j, i))
                    break
            else:
                seen.append((pts[i, 0], pts[i, 1], i))
    want = sorted(j for _, _, j in seen)
    assert orc.endpoints(pts, paths).tolist() == want
