"""GPU parity of the export image (render_paths.wgsl, gpu.rs:572-651): the HIP
rasterizer through the C ABI against the oracle's sequential restatement,
byte for byte, and the noisiness of both (noisiness.rs)."""
import ctypes

import numpy as np
import pytest

from bundling_thingy_amd import api, workloads

pytestmark = pytest.mark.gpu


def _oracle_image(orc, bt, pl, w, h, palette, color_by, msaa, dots):
    pal, ep = bt.api.palette(palette)
    paths = np.stack([pl.offsets.astype(np.uint32), np.zeros_like(pl.counts), pl.counts], 1)
    dens = pl.normalized if color_by == bt.COLOR_DENSITY else None
    return orc.render(pl.points, pl.edge_id, paths, w, h, pal, ep, density=dens, msaa=msaa,
                      dots=dots)


def _check(gpu, orc, bt, pl, palette, color_by, msaa, dots, label):
    w, h = gpu.render_size()
    got = gpu.render(palette, color_by, msaa, dots)
    want = _oracle_image(orc, bt, pl, w, h, palette, color_by, msaa, dots)
    diff = np.argwhere((got != want).any(2))
    assert len(diff) == 0, f"{label}: {len(diff)} pixels differ, first {diff[:5].tolist()}"
    return got


def test_blank_frame(bt):
    with bt.Gpu(workloads.reference_cpu_options()) as gpu:
        img = gpu.render()
        assert img.shape == (201, 201, 4)
        assert bt.noisiness_background(img) == (6, 6, 9)
        assert bt.noisiness(img, (6, 6, 9))["colored_pixels"] == 0


@pytest.mark.parametrize("palette,color_by,msaa,dots", [
    ("viridis", 0, 4, False),
    ("viridis", 1, 4, False),
    ("plasma", 0, 2, True),
    ("batlow", 1, 2, False),
    ("cividis", 1, 4, True),
])
def test_reference_mode_image_bitexact(bt, orc, palette, color_by, msaa, dots):
    """Reference mode S = 201, r = 10 (the plain CPU case): after upload and
    after two iterations."""
    g = workloads.plain_cpu_graph()
    opts = workloads.reference_cpu_options()
    pl = orc.Pipeline(g.points, g.starts, g.bbox, opts.iter_params(0))
    with bt.Gpu(opts) as gpu:
        gpu.upload(g)
        _check(gpu, orc, bt, pl, palette, color_by, msaa, dots, "upload")
        for it in range(2):
            gpu.iterate()
            pl.iterate(opts.iter_params(it))
        img = _check(gpu, orc, bt, pl, palette, color_by, msaa, dots, "iteration 2")
        m = bt.noisiness(img, (6, 6, 9))
        assert 0 < m["noisiness"] < 1


def test_gpu_case_image_and_noisiness(bt, orc):
    """The plain GPU workload (9,800 edges, ~1.3M samples) on its 1024^2 map,
    density colouring, 4x MSAA: byte-identical images after upload and after
    one iteration at the schedule's smallest radius (0.01, sigma 2)."""
    g = workloads.plain_gpu_graph()
    opts = bt.Options(spacing=0.002, radius=0.01, sigma=2.0, density=1024)
    pl = orc.Pipeline(g.points, g.starts, g.bbox, opts.iter_params(0))
    with bt.Gpu(opts) as gpu:
        gpu.upload(g)
        img0 = _check(gpu, orc, bt, pl, "viridis", 0, 4, False, "upload")
        gpu.iterate()
        pl.iterate(opts.iter_params(0), threads=0)
        img1 = _check(gpu, orc, bt, pl, "viridis", 0, 4, False, "iteration 1")
    n0 = bt.noisiness(img0, (6, 6, 9))["noisiness"]
    n1 = bt.noisiness(img1, (6, 6, 9))["noisiness"]
    assert 0 < n1 < n0 < 1  # bundling concentrates the drawing


@pytest.mark.parametrize("spacing,msaa", [(0.15, 4), (0.05, 2)])
def test_long_segments_bitexact(bt, orc, spacing, msaa):
    """Segments of tens to hundreds of pixels in every direction (coarse
    spacing on a 1024^2 map): the row-bounded walk covers what the oracle's
    full bounding-box walk covers."""
    g = workloads.plain_cpu_graph()
    opts = bt.Options(spacing=spacing, radius=3 * spacing, sigma=2.0, density=1024)
    pl = orc.Pipeline(g.points, g.starts, g.bbox, opts.iter_params(0))
    with bt.Gpu(opts) as gpu:
        gpu.upload(g)
        _check(gpu, orc, bt, pl, "plasma", 1, msaa, True, "upload")
        gpu.iterate()
        pl.iterate(opts.iter_params(0))
        _check(gpu, orc, bt, pl, "viridis", 0, msaa, False, "iteration 1")


def test_batched_job_images_equal_alone(bt):
    """bt_render_job on a batched context: each job's image equals the image
    of the same graph uploaded alone (non-square 320 x 200 map included)."""
    graphs = workloads.batched_graphs(count=3, lo=50, hi=400)
    opts = [bt.Options(map_size=301, neighbor_radius=8),
            bt.Options(spacing=0.003, radius=0.03, sigma=3.0, density=(320, 200)),
            workloads.plain_cpu_options()]
    alone = []
    for g, o in zip(graphs, opts):
        with bt.Gpu(o) as gpu:
            gpu.upload(g)
            gpu.iterate(2)
            alone.append(gpu.render("plasma", 0, 4, True))
    with bt.Gpu(opts[0]) as gpu:
        gpu.upload_batch(graphs, opts)
        gpu.iterate(2)
        for j in range(len(graphs)):
            got = gpu.render("plasma", 0, 4, True, job=j)
            assert got.shape == alone[j].shape
            np.testing.assert_array_equal(got, alone[j])


def test_full_schedule_distortion_and_noisiness(bt, orc):
    """The acceptance criterion after a whole schedule: the plain CPU case's
    10 iterations (radius 0.05, 512^2, sigma 4). Distortion and noisiness
    must agree with the oracle's within 1 % (they agree exactly: positions,
    density and image are bit-identical)."""
    g = workloads.plain_cpu_graph()
    opts = workloads.plain_cpu_options()
    pl = orc.Pipeline(g.points, g.starts, g.bbox, opts.iter_params(0))
    base = pl.points.copy()
    with bt.Gpu(opts) as gpu:
        gpu.upload(g)
        layout = gpu.distortion_layout()
        b = gpu.distortion_snapshot(layout)
        gpu.iterate(10)
        d_gpu = bt.distortion(b, gpu.distortion_snapshot(layout), 512)[0]
        img = gpu.render()
    for it in range(10):
        pl.iterate(opts.iter_params(it))
    d_ref, _ = orc.distortion(base, pl.points, pl.offsets, pl.counts)
    assert d_ref > 0 and abs(d_gpu - d_ref) <= 0.01 * d_ref
    assert d_gpu == d_ref
    ref_img = _oracle_image(orc, bt, pl, 512, 512, "viridis", 0, 4, False)
    n_gpu = bt.noisiness(img, (6, 6, 9))["noisiness"]
    n_ref = bt.noisiness(ref_img, (6, 6, 9))["noisiness"]
    assert abs(n_gpu - n_ref) <= 0.01 * n_ref
    np.testing.assert_array_equal(img, ref_img)


def test_render_errors(bt):
    g = workloads.plain_cpu_graph()
    with bt.Gpu(bt.Options(map_size=201, neighbor_radius=10, color_by_density=False)) as gpu:
        gpu.upload(g)
        with pytest.raises(bt.BtError) as e:
            gpu.render(color_by=bt.COLOR_DENSITY)
        assert e.value.code == 5
        assert gpu.render(color_by=bt.COLOR_EDGE_ID).shape == (201, 201, 4)
        for kw in ({"msaa": 8}, {"palette": 7}, {"msaa": 1}):
            with pytest.raises(bt.BtError) as e:
                gpu.render(color_by=bt.COLOR_EDGE_ID, **kw)
            assert e.value.code == 1
        # a buffer of the wrong size is refused, not overrun
        ro = api.bt_render_options()
        api.lib().bt_render_options_default(ctypes.byref(ro))
        ro.color_by = bt.COLOR_EDGE_ID
        small = np.zeros(201 * 201 * 4 - 4, np.uint8)
        assert api.lib().bt_render(gpu._h, ctypes.byref(ro), small.ctypes.data_as(api._P),
                                   small.nbytes) == 1
        with pytest.raises(bt.BtError) as e:
            gpu.render(color_by=bt.COLOR_EDGE_ID, job=3)
        assert e.value.code == 1  # job index out of range
    with bt.Gpu(bt.Options(map_size=201, neighbor_radius=10)) as gpu:
        with pytest.raises(bt.BtError) as e:
            gpu.endpoints()
        assert e.value.code == 2  # no uploaded graph


def test_headless_export(bt, orc, tmp_path):
    """main.rs headless flow: original.png + one PNG per iteration, with the
    noisiness and distortion reports; PNGs hold exactly the rendered image and
    the distortion agrees with the oracle's."""
    import json

    from bundling_thingy_amd.headless import headless
    from tests.test_render_oracle import _read_png

    g = workloads.plain_cpu_graph()
    opts = workloads.reference_cpu_options()
    res = headless(g, str(tmp_path), opts, iterations=3, evaluate_noisiness=True,
                   evaluate_distortion=True, log=lambda *_: None)
    assert res["images"] == ["original.png", "iteration-01.png", "iteration-02.png",
                             "iteration-03.png"]
    noise = json.loads((tmp_path / "noisiness.json").read_text())
    dist = json.loads((tmp_path / "distortion.json").read_text())
    assert noise["render"]["background_rgb"] == [6, 6, 9]
    assert [i["iteration"] for i in noise["images"]] == [0, 1, 2, 3]
    assert noise["images"][0]["change_from_original"] == 0.0
    assert dist["geometry"]["edges_total"] == 2100
    assert dist["images"][0]["distortion"] == 0.0
    pl = orc.Pipeline(g.points, g.starts, g.bbox, opts.iter_params(0))
    base = pl.points.copy()
    for it in range(3):
        pl.iterate(opts.iter_params(it))
        d, _ = orc.distortion(base, pl.points, pl.offsets, pl.counts)
        assert dist["images"][it + 1]["distortion"] == pytest.approx(d, rel=1e-12)
    last = _read_png(tmp_path / "iteration-03.png")
    want = _oracle_image(orc, bt, pl, 201, 201, "viridis", 0, 4, False)
    np.testing.assert_array_equal(last, want)
    assert noise["images"][3]["noisiness"] == bt.noisiness(want, (6, 6, 9))["noisiness"]


def test_non_square_map_bitexact(bt, orc):
    """A 301 x 201 map in reference mode (texel scales differ per axis; the
    window's size after a resize, gpu.rs:653-697): positions, density and
    the image (quads, dots, bilinear density lookup) against the oracle."""
    g = workloads.plain_cpu_graph()
    opts = bt.Options(map_size=301, map_h=201, neighbor_radius=10)
    pl = orc.Pipeline(g.points, g.starts, g.bbox, opts.iter_params(0))
    with bt.Gpu(opts) as gpu:
        assert gpu.upload(g) == len(pl.points)
        for it in range(2):
            gpu.iterate()
            pl.iterate(opts.iter_params(it))
        got = gpu.points()[0]
        np.testing.assert_array_equal(got.view(np.uint32), pl.points.view(np.uint32))
        np.testing.assert_array_equal(gpu.density(3).view(np.uint32),
                                      pl.normalized.view(np.uint32))
        assert gpu.render_size() == (301, 201)
        for color_by in (0, 1):
            _check(gpu, orc, bt, pl, "viridis", color_by, 4, True, f"color_by {color_by}")
