"""MI355X-native edge-bundling engine: the resample -> grid -> neighbor-mean ->
density hot path of lpstankus/bundling_thingy as hand-written gfx950 HIP
kernels behind a C ABI (include/bundling.h)."""
from .api import (COLOR_DENSITY, COLOR_EDGE_ID, BtError, Gpu, Graph, Layout, Options, Server,
                  Snapshot, blur_weights, capacity, distortion, lib, noisiness,
                  noisiness_background, write_png)

__all__ = ["COLOR_DENSITY", "COLOR_EDGE_ID", "BtError", "Gpu", "Graph", "Layout", "Options",
           "Server", "Snapshot", "blur_weights", "capacity", "distortion", "lib", "noisiness",
           "noisiness_background", "write_png"]
