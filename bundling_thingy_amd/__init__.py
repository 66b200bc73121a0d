"""MI355X-native edge-bundling engine: the resample -> grid -> neighbor-mean ->
density hot path of lpstankus/bundling_thingy as hand-written gfx950 HIP
kernels behind a C ABI (include/bundling.h)."""
from .api import (BtError, Gpu, Graph, Layout, Options, Server, Snapshot, blur_weights, capacity,
                  distortion, lib)

__all__ = ["BtError", "Gpu", "Graph", "Layout", "Options", "Server", "Snapshot", "blur_weights",
           "capacity", "distortion", "lib"]
