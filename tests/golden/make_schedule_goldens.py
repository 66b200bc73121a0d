"""Writes the full-schedule golden fixtures of the plain GPU case and the large
case by running the CPU oracle (oracle/liboracle.so) over whole schedules.

The GPU tests cannot afford the oracle's full schedules on the GPU host
(minutes of CPU work), so the oracle's results are recorded here once:
for every iteration the SHA-256 of the sample positions (float32 bits, Layout
order: path by path, sample by sample) and the distortion against the
post-upload positions (distortion.rs:170-216), and after the last iteration
the SHA-256 of the normalized density plane, the noisiness of the export
image (density colouring, viridis, 4x MSAA; noisiness.rs:44-74 restated with
numpy) and the sample count.

Usage: python tests/golden/make_schedule_goldens.py [plain|large ...]
Threads: OMP-style, all host cores.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bundling_thingy_amd import api, workloads  # noqa: E402  (host-only calls)
from oracle import oracle as orc  # noqa: E402

BACKGROUND = (6, 6, 9)  # the blank frame's RGB (clear colour 0.025, 0.025, 0.035)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def noisiness(img, rgb=BACKGROUND):
    colored = int(np.any(img[:, :, :3] != np.array(rgb, np.uint8), axis=2).sum())
    return colored / float(img.shape[0] * img.shape[1])


def run(name, graph, opts, iterations, size):
    t0 = time.time()
    pl = orc.Pipeline(graph.points, graph.starts, graph.bbox, opts.iter_params(0))
    base = pl.points.copy()
    rows = []
    for it in range(iterations):
        t = time.time()
        pl.iterate(opts.iter_params(it), threads=0)
        d, _ = orc.distortion(base, pl.points, pl.offsets, pl.counts)
        rows.append({"iteration": it + 1, "positions_sha256": sha(pl.points), "distortion": d})
        print(f"{name}: iteration {it + 1} in {time.time() - t:.1f} s, distortion {d:.6g}",
              flush=True)
    pal, ep = api.palette("viridis")
    paths = np.stack([pl.offsets.astype(np.uint32), np.zeros_like(pl.counts), pl.counts], 1)
    img = orc.render(pl.points, pl.edge_id, paths, size, size, pal, ep, density=pl.normalized,
                     msaa=4, dots=False)
    out = {
        "workload": name,
        "generator": "tests/golden/make_schedule_goldens.py (CPU oracle)",
        "samples": int(len(pl.points)),
        "iterations": rows,
        "final": {"distortion": rows[-1]["distortion"], "noisiness": noisiness(img),
                  "normalized_density_sha256": sha(pl.normalized),
                  "image_sha256": sha(img)},
        "oracle_seconds": time.time() - t0,
    }
    path = os.path.join(HERE, f"{name}_schedule.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path} ({out['oracle_seconds']:.0f} s)", flush=True)


def main(which):
    if "plain" in which:
        run("plain_gpu", workloads.plain_gpu_graph(), workloads.plain_gpu_options(30), 30, 1024)
    if "large" in which:
        run("large", workloads.large_graph(), workloads.large_options(50), 50, 2048)


if __name__ == "__main__":
    main(sys.argv[1:] or ["plain", "large"])
