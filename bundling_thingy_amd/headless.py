"""Headless export (main.rs:110-221): upload a graph, write original.png, then
iterate and write one PNG per iteration, with the optional noisiness
(noisiness.rs:84-137) and distortion (distortion_report.rs:6-32) reports.

The reports keep the reference's JSON field names and layout, so tools that
read its noisiness.json / distortion.json read these; the prose fields are
worded here. Writes go through a temporary file and a rename (report.rs)."""
import json
import math
import os

from . import api

_CLEAR_BLANK = "blank frame rendered before upload"


def _publish(directory, name, report):
    path = os.path.join(directory, name)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(report, f, indent=2)
    os.replace(tmp, path)


def _invalidate(directory, name):
    try:
        os.remove(os.path.join(directory, name))
    except FileNotFoundError:
        pass


def headless(graph, out, options, iterations=1, palette="viridis", color_by=None, msaa=4,
             endpoint_dots=False, evaluate_noisiness=False, evaluate_distortion=False,
             input_name="", log=print):
    """Runs the headless export of `graph` (an api.Graph) into directory `out`.
    `options` is an api.Options with a radius (reference or generalized
    mode); color_by defaults to density when the options compute it, else
    to edge ids. Returns {"samples", "images", "noisiness", "distortion"}."""
    os.makedirs(out, exist_ok=True)
    _invalidate(out, "distortion.json")
    _invalidate(out, "noisiness.json")
    if color_by is None:
        color_by = api.COLOR_DENSITY if options.c.flags & api.FLAG_DENSITY else api.COLOR_EDGE_ID
    rkw = dict(palette=palette, color_by=color_by, msaa=msaa, endpoint_dots=endpoint_dots)
    noise, dist_images, files = [], [], []
    with api.Gpu(options) as gpu:
        rgb = None
        if evaluate_noisiness:
            rgb = api.noisiness_background(gpu.render(**rkw))
        samples = gpu.upload(graph)
        w, h = gpu.render_size()
        layout = gpu.distortion_layout() if evaluate_distortion else None
        base = gpu.distortion_snapshot(layout) if layout is not None else None

        def write(name, iteration):
            img = gpu.render(**rkw)
            api.write_png(os.path.join(out, name), img)
            files.append(name)
            if rgb is not None:
                m = api.noisiness(img, rgb)
                first = noise[0]["noisiness"] if noise else m["noisiness"]
                m.update(file=name, iteration=iteration,
                         change_from_original=m["noisiness"] - first)
                noise.append(m)
                log(f"noisiness {name}: {m['noisiness']:.6f}, "
                    f"change {m['change_from_original']:+.6f}")
            log(f"wrote {os.path.join(out, name)}")

        write("original.png", 0)
        if base is not None:
            dist_images.append({"file": "original.png", "iteration": 0, "distortion": 0.0,
                                "rms_displacement_pixels": 0.0})
        for it in range(1, iterations + 1):
            gpu.iterate()
            name = "neighbor-mean.png" if iterations == 1 else f"iteration-{it:02d}.png"
            write(name, it)
            if base is not None:
                d, _, rms_px = api.distortion(base, gpu.distortion_snapshot(layout), max(w, h))
                dist_images.append({"file": name, "iteration": it, "distortion": d,
                                    "rms_displacement_pixels": rms_px})
                log(f"{name} distortion={d} rms_displacement_pixels={rms_px}")
    radius = options.c.radius_texels or None
    if evaluate_noisiness:
        _publish(out, "noisiness.json", {
            "schema_version": 1, "metric_version": 1, "metric": "noisiness",
            "definition": "share of exported pixels whose RGB is not exactly the background "
                          "RGB (alpha ignored)",
            "interpretation": "comparable only between runs with identical render settings",
            "input": input_name,
            "render": {
                "render_contract_version": 1, "width": w, "height": h, "format": "rgba8unorm",
                "background_rgb": list(rgb), "comparison": "rgb_exact_inequality",
                "palette": palette if isinstance(palette, str) else
                [k for k, v in api.PALETTES.items() if v == palette][0],
                "color_by": "density" if color_by == api.COLOR_DENSITY else "edge-id",
                "endpoint_dots": bool(endpoint_dots), "msaa": msaa,
                "path_style": {"half_width_clip": 0.001, "opacity": 0.1,
                               "endpoint_radius_pixels": 3, "endpoint_outline_width_pixels": 1},
                "normalization": "initial graph bounds, uniform 95% fit, centered; fixed "
                                 "through iterations"},
            "bundling": {"neighbor_radius": radius, "iterations": iterations},
            "images": noise,
        })
    if evaluate_distortion:
        L = layout
        _publish(out, "distortion.json", {
            "schema_version": 1, "metric_version": 1, "metric": "distortion",
            "definition": "RMS over edges (equal weights) of the integrated squared clip-space "
                          "displacement under normalized arc-length correspondence, over the "
                          "canvas diagonal",
            "interpretation": "lower means the geometry moved less from the resampled input",
            "input": input_name,
            "baseline": {"source": "gpu_resampled_iteration_0",
                         "correspondence": "directed_edge_normalized_arc_length",
                         "weighting": "equal_edges", "preprocessing_distortion_included": False},
            "geometry": {
                "geometry_contract_version": 1, "coordinate_space": "clip", "width": w,
                "height": h,
                "normalization": "initial graph bounds, centered uniform 95% fit, unchanged "
                                 "through iterations",
                "diagonal_clip_units": 2.0 * math.sqrt(2.0), "pixels_per_clip_unit": w / 2.0,
                "edges_total": len(L.edges) + L.skipped, "edges_evaluated": len(L.edges),
                "edges_skipped": L.skipped, "active_samples": L.samples},
            "bundling": {"neighbor_radius": radius, "iterations": iterations},
            "images": dist_images,
        })
    log(f"rendered {samples} samples to {out}")
    return {"samples": samples, "images": files, "noisiness": noise, "distortion": dist_images}
