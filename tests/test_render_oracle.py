"""Pins the oracle's export-image restatement (render_paths.wgsl, gpu.rs:572-651,
palette.rs, noisiness.rs) and the host image helpers. No GPU needed.

The reference's MSAA coverage is set by the Vulkan driver and pinned by no
reference output; these tests pin the oracle's rasterization contract with
hand-derived closed forms, exact rational geometry away from edges, the
watertight split of a segment, and the published colormap values."""
import math
import struct
import zlib
from fractions import Fraction

import numpy as np
import pytest

W = H = 1600  # the reference's map / render size (gpu.rs:21)
S4 = [(0.375, 0.125), (0.875, 0.375), (0.125, 0.625), (0.625, 0.875)]
S2 = [(0.75, 0.75), (0.25, 0.25)]


@pytest.fixture(scope="module")
def bt():
    import bundling_thingy_amd as bt
    bt.lib()
    return bt


@pytest.fixture(scope="module")
def viridis(bt):
    return bt.api.palette("viridis")


def clip(fx, fy, w=W, h=H):
    """Pixel position -> clip position (exact for the values used here)."""
    return [np.float32(fx / w * 2 - 1), np.float32(1 - fy / h * 2)]


def one_path(*pix):
    pts = np.array([clip(*p) for p in pix], dtype=np.float32)
    return pts, np.array([[0, 0, len(pix)]], dtype=np.uint32)


# ------------------------------------------------------------------ palettes

def _lightness(rgb):
    c = np.where(rgb <= 0.04045, rgb / 12.92, ((rgb + 0.055) / 1.055) ** 2.4)
    y = c @ np.array([0.2126, 0.7152, 0.0722])
    return np.where(y > 216 / 24389, 116 * np.cbrt(y) - 16, y * 24389 / 27)


@pytest.mark.parametrize("name,first,last", [
    # matplotlib _cm_listed.py end points of the published colormaps
    ("viridis", (0.267004, 0.004874, 0.329415), (0.993248, 0.906157, 0.143936)),
    ("plasma", (0.050383, 0.029803, 0.527975), (0.940015, 0.975158, 0.131326)),
    ("cividis", (0.0, 0.135112, 0.304751), (0.995737, 0.909344, 0.217772)),
    ("batlow", None, None),
])
def test_palettes_are_the_published_colormaps(bt, name, first, last):
    cols, ep = bt.api.palette(name)
    assert cols.shape == (16, 3) and ep.shape == (3,)
    if first is not None:
        np.testing.assert_allclose(cols[0], first, atol=1e-6)
        np.testing.assert_allclose(cols[15], last, atol=1e-6)
    # perceptually uniform sequential maps: CIELAB lightness rises in even
    # steps (batlow's are less even at its dark end)
    steps = np.diff(_lightness(cols.astype(np.float64)))
    lo, hi = (3.0, 10.0) if name == "batlow" else (3.0, 7.0)
    assert steps.min() > lo and steps.max() < hi, steps
    assert ((cols >= 0) & (cols <= 1)).all()


# ------------------------------------------------------------------ coverage

def test_horizontal_segment_closed_form(orc):
    """A segment from pixel (10, 20) to (30, 20): half-width 0.001 clip =
    0.8 px, so the samples with |y - 20| < 0.8 and 10 <= x < 30 are covered:
    in row 19 the samples at y = 19.375, 19.625, 19.875 (4x samples 1, 2, 3),
    in row 20 those at 20.125, 20.375, 20.625 (samples 0, 1, 2)."""
    pts, paths = one_path((10.0, 20.0), (30.0, 20.0))
    cov = orc.coverage(pts, paths, W, H, 4)
    want = np.zeros_like(cov)
    want[19, 10:30, 1:4] = 1
    want[20, 10:30, 0:3] = 1
    np.testing.assert_array_equal(cov, want)
    cov2 = orc.coverage(pts, paths, W, H, 2)
    want2 = np.zeros_like(cov2)
    want2[19:21, 10:30, :] = 1   # y = 19.25, 19.75, 20.25, 20.75: all within 0.8
    np.testing.assert_array_equal(cov2, want2)


def test_vertical_segment_closed_form(orc):
    pts, paths = one_path((40.0, 5.0), (40.0, 12.0))
    cov = orc.coverage(pts, paths, W, H, 4)
    want = np.zeros_like(cov)
    # |x - 40| < 0.8: in column 39 samples at x = 39.375 / 39.625 / 39.875
    # (samples 0, 3, 1), in column 40 those at 40.125 / 40.375 / 40.625 (2, 0, 3)
    want[5:12, 39, [0, 1, 3]] = 1
    want[5:12, 40, [0, 2, 3]] = 1
    np.testing.assert_array_equal(cov, want)


def test_degenerate_segment_covers_nothing(orc):
    pts, paths = one_path((10.0, 10.0), (10.0, 10.0))
    assert orc.coverage(pts, paths, W, H, 4).sum() == 0


@pytest.mark.parametrize("a,m,b", [
    ((10.0, 20.0), (17.0, 20.0), (30.0, 20.0)),
    ((50.0, 4.0), (50.0, 9.5), (50.0, 30.0)),
    ((100.0, 100.0), (110.0, 110.0), (120.0, 120.0)),
])
def test_split_segment_is_watertight(orc, a, m, b):
    """Splitting a segment at an interior point of it covers every sample of
    the whole segment exactly once, the joint included (top-left rule)."""
    whole = orc.coverage(*one_path(a, b), W, H, 4)
    split = orc.coverage(*one_path(a, m, b), W, H, 4)
    assert whole.max() == 1
    np.testing.assert_array_equal(split, whole)


def _frac_pos(v):
    return Fraction(float(v))


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_coverage_matches_exact_quads_away_from_edges(orc, seed):
    """Random segments: a sample farther than 1/64 px from the quad's boundary
    is covered exactly once when it lies inside the exact (unsnapped) quad
    p +- n, q +- n with |n| = 0.001 (render_paths.wgsl:41-45), and not at all
    outside; checked in exact rational arithmetic."""
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(40):
        p = rng.uniform(-0.97, 0.97, 2).astype(np.float32)
        d = rng.normal(0, 0.004, 2).astype(np.float32)
        q = (p + d).astype(np.float32)
        pts = np.stack([p, q]).astype(np.float32)
        paths = np.array([[0, 0, 2]], np.uint32)
        cov = orc.coverage(pts, paths, W, H, 4)
        P = [_frac_pos(v) for v in p]
        Q = [_frac_pos(v) for v in q]
        D = [Q[0] - P[0], Q[1] - P[1]]
        L2 = D[0] ** 2 + D[1] ** 2
        L = math.sqrt(float(L2))

        def to_pix(x, y):
            return (x + 1) / 2 * W, (1 - y) / 2 * H

        ux, uy = to_pix(*P)
        vx, vy = to_pix(*Q)
        ex, ey = vx - ux, vy - uy                  # along, pixels
        lenp = math.hypot(float(ex), float(ey))
        hw = 0.001 * W / 2                          # half-width in pixels
        x0, x1 = int(min(ux, vx)) - 2, int(max(ux, vx)) + 3
        y0, y1 = int(min(uy, vy)) - 2, int(max(uy, vy)) + 3
        for py in range(max(y0, 0), min(y1, H)):
            for px in range(max(x0, 0), min(x1, W)):
                for k, (sx, sy) in enumerate(S4):
                    rx, ry = float(px + sx - ux), float(py + sy - uy)
                    along = (rx * float(ex) + ry * float(ey)) / lenp
                    across = abs(-rx * float(ey) + ry * float(ex)) / lenp
                    margin = min(along, lenp - along, hw - across)
                    if abs(margin) < 1 / 64:
                        continue
                    assert cov[py, px, k] == (1 if margin > 0 else 0), (seed, px, py, k)
                    checked += 1
        assert L > 0
    assert checked > 1000


# -------------------------------------------------------------------- colour

def test_single_fragment_colour_known_answer(orc, viridis):
    """Edge-id colouring of the only edge: value 0 -> viridis stop 0
    (0.267004, 0.004874, 0.329415). One 0.1-opacity blend over the clear
    colour (6, 6, 9) stores (12, 6, 17) in a sample; three of four samples
    covered resolve to (11, 6, 15)."""
    pal, ep = viridis
    pts, paths = one_path((10.0, 20.0), (30.0, 20.0))
    img = orc.render(pts, [0, 0], paths, W, H, pal, ep, density=None, msaa=4)
    np.testing.assert_array_equal(img[19, 10:30, :3], np.tile([11, 6, 15], (20, 1)))
    np.testing.assert_array_equal(img[20, 10:30, :3], np.tile([11, 6, 15], (20, 1)))
    mask = np.ones((H, W), bool)
    mask[19:21, 10:30] = False
    assert (img[mask][:, :3] == [6, 6, 9]).all() and (img[..., 3] == 255).all()
    # 2x: both samples of rows 19 and 20 covered -> the stored (12, 6, 17)
    img2 = orc.render(pts, [0, 0], paths, W, H, pal, ep, density=None, msaa=2)
    np.testing.assert_array_equal(img2[19:21, 10:30, :3], np.tile([12, 6, 17], (2, 20, 1)))


def test_density_colour_known_answer(orc, viridis):
    """A constant density of 0.5 samples to 0.5 anywhere (binary16 exact,
    bilinear weights sum to one): stop 7.5 = midway between viridis stops 7
    and 8 = (0.129856, 0.565117, 0.5494545); one blend stores (9, 20, 22)
    and three of four samples resolve to (8, 17, 19)."""
    pal, ep = viridis
    pts, paths = one_path((10.0, 20.0), (30.0, 20.0))
    dens = np.full((H, W), 0.5, np.float32)
    img = orc.render(pts, [0, 0], paths, W, H, pal, ep, density=dens, msaa=4)
    np.testing.assert_array_equal(img[19, 10:30, :3], np.tile([8, 17, 19], (20, 1)))


def test_repeated_blending_reaches_the_8bit_fixed_point(orc, viridis):
    """Many coincident segments: each sample converges to the value v with
    rint(255 * (0.1 c + 0.9 v / 255)) == v, the 8-bit fixed point nearest the
    clear colour, and stays there."""
    pal, ep = viridis
    n = 200
    pts = np.array([clip(10.0, 20.0), clip(30.0, 20.0)] * n, np.float32)
    paths = np.array([[2 * i, 0, 2] for i in range(n)], np.uint32)
    img = orc.render(pts, np.zeros(2 * n, np.uint32), paths, W, H, pal, ep, msaa=4, n_edges=1)
    c = pal[0].astype(np.float64)
    fixed = []
    for ch in range(3):
        v = [6, 6, 9][ch]
        for _ in range(1000):
            nv = int(np.rint(np.clip(c[ch] * 0.1 + v / 255 * 0.9, 0, 1) * 255))
            if nv == v:
                break
            v = nv
        fixed.append(v)
    bg = np.array([6, 6, 9])
    want = (3 * np.array(fixed) + bg + 2) // 4
    np.testing.assert_array_equal(img[19, 10:30, :3], np.tile(want, (20, 1)))
    assert (np.array(fixed) > bg).any()


def test_half_round_matches_numpy_binary16(orc):
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.random(3000), rng.random(500) * 1e-4, rng.random(200) * 1e-7,
                        [0.0, 1.0, 0.5, 2.0 ** -14, 2.0 ** -24, 3 * 2.0 ** -25, 1 + 2.0 ** -11,
                         1 + 3 * 2.0 ** -11, 65504.0, 65519.0, 65520.0]]).astype(np.float32)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(orc.half_round(x).view(np.uint32), want.view(np.uint32))


# ------------------------------------------------------------- endpoint dots

def test_endpoint_dots_match_exact_circles(orc, viridis):
    """Dots (render_paths.wgsl:73-97): samples within 2 px of an endpoint take
    the endpoint colour, within 3 px the outline colour colors[0] * 0.1,
    both opaque; checked per sample against exact distances."""
    pal, ep = viridis
    c0, c1 = (100.5, 200.5), (140.25, 220.75)
    pts, paths = one_path(c0, c1)
    img = orc.render(pts, [0, 0], paths, W, H, pal, ep, msaa=4, dots=True)
    path_cov = orc.coverage(pts, paths, W, H, 4).sum(2)
    ring = np.rint(pal[0].astype(np.float32) * np.float32(0.1) * 255).astype(int)
    fill = np.rint(ep * 255).astype(int)
    seen = [0, 0, 0]
    for cx, cy in (c0, c1):
        for py in range(int(cy) - 4, int(cy) + 5):
            for px in range(int(cx) - 4, int(cx) + 5):
                kinds = []
                for sx, sy in S4:
                    d2 = ((Fraction(px) + Fraction(sx) - Fraction(cx)) ** 2 +
                          (Fraction(py) + Fraction(sy) - Fraction(cy)) ** 2)
                    assert d2 not in (4, 9)
                    kinds.append(2 if d2 < 4 else 1 if d2 < 9 else 0)
                if min(kinds) == 2:
                    np.testing.assert_array_equal(img[py, px, :3], fill)
                elif path_cov[py, px] == 0:
                    # no path fragment here: the resolve averages dot, ring and
                    # clear-colour samples
                    cols = [np.array([6, 6, 9]), ring, fill]
                    want = (sum(cols[k] for k in kinds) + 2) // 4
                    np.testing.assert_array_equal(img[py, px, :3], want)
                for k in set(kinds):
                    seen[k] += 1
    assert min(seen) >= 8, seen


# --------------------------------------------------------------- noisiness

def test_blank_frame_and_noisiness(orc, bt, viridis):
    pal, ep = viridis
    img = orc.render(np.zeros((0, 2), np.float32), [], np.zeros((0, 3), np.uint32), 64, 48,
                     pal, ep)
    assert img.shape == (48, 64, 4) and (img == [6, 6, 9, 255]).all()
    rgb = bt.noisiness_background(img)
    assert rgb == (6, 6, 9)
    m = bt.noisiness(img, rgb)
    assert m["colored_pixels"] == 0 and m["total_pixels"] == 64 * 48
    assert m["noisiness"] == 0.0 and m["colored_to_background_ratio"] == 0.0
    img[3, 4, 0] = 7      # one channel differs
    img[5, 6, 3] = 0      # alpha is ignored by measure
    m = bt.noisiness(img, rgb)
    assert m["colored_pixels"] == 1 and m["background_pixels"] == 64 * 48 - 1
    assert m["noisiness"] == 1 / (64 * 48)
    assert m["colored_to_background_ratio"] == 1 / (64 * 48 - 1)
    full = np.zeros((2, 2, 4), np.uint8)
    assert bt.noisiness(full, (1, 1, 1))["colored_to_background_ratio"] is None


@pytest.mark.parametrize("mutate,msg", [
    (lambda a: a.__setitem__((0, 0, 3), 254), "blank frame must be opaque"),
    (lambda a: a.__setitem__((1, 1, 2), 10), "blank frame must have uniform RGB and opaque alpha"),
    (lambda a: a.__setitem__((1, 1, 3), 0), "blank frame must have uniform RGB and opaque alpha"),
])
def test_background_errors(bt, mutate, msg):
    img = np.full((4, 5, 4), [6, 6, 9, 255], np.uint8)
    mutate(img)
    with pytest.raises(bt.BtError) as e:
        bt.noisiness_background(img)
    assert e.value.msg == msg


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xffffffff
        if kind == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype, _, _, _ = ihdr
    assert (depth, ctype) == (8, 6)
    raw = zlib.decompress(idat)
    rows = [raw[y * (4 * w + 1):(y + 1) * (4 * w + 1)] for y in range(h)]
    assert all(r[0] == 0 for r in rows)
    return np.frombuffer(b"".join(r[1:] for r in rows), np.uint8).reshape(h, w, 4)


def test_png_round_trip(bt, tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    p = tmp_path / "x.png"
    bt.write_png(p, img)
    np.testing.assert_array_equal(_read_png(p), img)
    with pytest.raises(bt.BtError):
        bt.write_png(tmp_path / "missing" / "x.png", img)
