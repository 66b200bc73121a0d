"""Benchmark: sample-point updates/s of the edge-bundling hot path on one MI355X.

Workload (plain GPU case): 1,715 nodes from a 20-component Gaussian mixture,
9,800 straight edges (seed 2), resampled at spacing 0.002 (~1.33M samples),
30 iterations with the neighbor radius decaying 0.06 -> 0.01 geometrically,
1024 x 1024 FP32 density with sigma 8 -> 2. One step = the whole 30-iteration
schedule (the scope of 30 Gpu::iterate calls) replayed as one hipGraph from
the post-upload state; upload/resampling is timed separately.

Prints one JSON line (see the keys below). Extra objects: roofline (dominant
kernel vs HBM peak, from HIP events), valu (the same kernel's FP32 work vs the
vector peak), cpu_baseline (the CPU oracle on a bounded sample), latency.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
ITERS = 30
TUNE_CANDIDATES = (1, 2, 3, 4, 6, 8, 12, 16, 20, 24, 28)
COPY_PROBE_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def candidate_flops(counts2d):
    """8 FP32 ops per candidate pair; candidates of a sample = samples in the
    3 x 3 block of cells (side = radius) around its cell."""
    c = counts2d.astype(np.int64)
    p = np.pad(c, 1)
    s = sum(p[1 + dy:1 + dy + c.shape[0], 1 + dx:1 + dx + c.shape[1]]
            for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    return 8 * int((c * s).sum())


def cpu_baseline(graph, opts, target_s=15.0):
    """The CPU oracle (all host threads) on a bounded sample of the same
    schedule: neighbor means of every `stride`-th sample for all 30 radii at
    the initial positions, scaled by `stride`, plus the full density stages."""
    from oracle import oracle as orc
    p0 = opts.iter_params(0)
    pl = orc.Pipeline(graph.points, graph.starts, graph.bbox, p0)
    pts, fixed = pl.points, pl.fixed
    threads = os.cpu_count() or 1
    threads = max(1, min(threads, int(os.environ.get("OMP_NUM_THREADS", threads))))
    grids = [(orc.grid(opts.iter_params(i)), opts.iter_params(i)) for i in range(ITERS)]

    def schedule(stride, threads=threads):
        """Returns the scaled schedule time, and the sample's own mean and
        density times."""
        total = tmean = tdens = 0.0
        for g, p in grids:
            # per-call cost that does not scale with the sample (binning all
            # points, the pass over skipped ones): a call that means one sample
            t = time.perf_counter()
            orc.neighbor_mean_strided(pts, fixed, g, len(pts), threads)
            tf = time.perf_counter() - t
            t = time.perf_counter()
            out = orc.neighbor_mean_strided(pts, fixed, g, stride, threads)
            tm = max(time.perf_counter() - t - tf, 0.0)
            t = time.perf_counter()
            c = orc.splat(out, p.density_w, p.density_h)
            orc.density(c, p.sigma)
            td = time.perf_counter() - t
            total += tm * stride + tf + td
            tmean += tm
            tdens += td
        return total, tmean, tdens

    # a coarse pass over the whole schedule sets the stride of the timed pass
    coarse = 2048
    _, tm, td = schedule(coarse)
    stride = int(max(1, min(coarse, np.ceil(coarse * tm / max(target_s - td, 1.0)))))
    total, tm, td = schedule(stride)
    n = len(pts)
    # the same on one thread over a sample `threads` times sparser, so it takes
    # about as long as the timed pass
    stride1 = int(stride * threads)
    total1, _, _ = schedule(stride1, threads=1)
    return {"value": ITERS * n / total, "unit": "sample-point updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle neighbor means of one sample in {stride} for all {ITERS} "
                      f"radii at the initial positions (x{stride}), full splat/blur/normalize; "
                      f"est. {total:.1f} s per schedule",
            "value_1thread": ITERS * n / total1,
            "sample_1thread": f"one sample in {stride1}, 1 thread; est. {total1:.1f} s per "
                              f"schedule"}


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def batched_section(bt, workloads, n_graphs, iters=30):
    """256 independent graphs: the plain per-graph drop-in (upload, 30 eager
    iterates, readback, one graph at a time) against one batched job table
    (one launch per stage per iteration, the schedule replayed as a hipGraph)."""
    t = time.perf_counter()
    graphs = workloads.batched_graphs(n_graphs)
    opts = [workloads.batched_options(i, iters) for i in range(n_graphs)]
    log(f"batched: generated {n_graphs} graphs in {time.perf_counter() - t:.1f} s")
    gpu = bt.Gpu(opts[0])
    # warm the code paths once
    gpu.upload_batch(graphs[:2], opts[:2])
    gpu.iterate(2)
    # plain per-graph form
    lat, samples = [], []
    t0 = time.perf_counter()
    for g, o in zip(graphs, opts):
        ts = time.perf_counter()
        s = gpu.upload_batch([g], [o])
        gpu.iterate(iters)
        gpu.points(0)
        lat.append((time.perf_counter() - ts) * 1e3)
        samples.append(int(s[0]))
    plain_s = time.perf_counter() - t0
    updates = iters * sum(samples)
    log(f"batched: plain per-graph {plain_s:.2f} s")
    # batched form
    t0 = time.perf_counter()
    s = gpu.upload_batch(graphs, opts)
    up_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    gpu.iterate(iters, graph=True)  # first call captures + instantiates the hipGraph
    first_s = time.perf_counter() - t1
    gpu.rewind()
    t1 = time.perf_counter()
    gpu.iterate(iters, graph=True)
    run_s = time.perf_counter() - t1
    t1 = time.perf_counter()
    for j in range(n_graphs):
        gpu.points(j)
    read_s = time.perf_counter() - t1
    batch_s = up_s + run_s + read_s
    gpu.close()
    log(f"batched: upload {up_s:.3f} s, schedule {run_s:.3f} s (first {first_s:.3f} s), "
        f"readback {read_s:.3f} s")
    assert int(s.sum()) == sum(samples)
    return {"graphs": n_graphs, "samples": int(sum(samples)), "iterations": iters,
            "plain": {"updates_per_s": updates / plain_s, "seconds": plain_s,
                      "job_p50_ms": pct(lat, 50), "job_p99_ms": pct(lat, 99)},
            "batched": {"updates_per_s": updates / batch_s, "seconds": batch_s,
                        "upload_s": up_s, "schedule_s": run_s, "readback_s": read_s,
                        "graph_build_s": first_s - run_s,
                        "job_p50_ms": batch_s * 1e3, "job_p99_ms": batch_s * 1e3},
            "speedup": plain_s / batch_s}


def large_section(bt, workloads, steps=2, iters=50):
    """The scale case: 100,000 edges over 20,000 nodes (~5M samples), 50
    iterations radius 0.02 -> 0.005, 2048^2 density, sigma 8 -> 2; autotuned
    row splits, the schedule replayed as one hipGraph from the uploaded state."""
    t = time.perf_counter()
    graph = workloads.large_graph()
    log(f"large: generated {graph.edge_count} edges in {time.perf_counter() - t:.1f} s")
    gpu = bt.Gpu(workloads.large_options(iters))
    t = time.perf_counter()
    n = gpu.upload(graph)
    upload_ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    chosen, _ = gpu.autotune(iters, TUNE_CANDIDATES)
    tune_s = time.perf_counter() - t
    gpu.rewind()
    gpu.iterate(iters, graph=True)
    times = []
    for _ in range(steps):
        gpu.rewind()
        t = time.perf_counter()
        gpu.iterate(iters, graph=True)
        times.append(time.perf_counter() - t)
    gpu.close()
    sec = sum(times) / len(times)
    log(f"large: {n} samples, {sec * 1e3:.1f} ms per {iters}-iteration schedule")
    return {"edges": graph.edge_count, "samples": n, "iterations": iters,
            "density": [2048, 2048], "updates_per_s": iters * n / sec,
            "ms_per_schedule": sec * 1e3, "upload_ms": upload_ms, "autotune_s": tune_s,
            "splits": chosen.tolist()}


def render_section(bt, gpu, opts, reps=5):
    """Export image of the bundled graph (render_paths.wgsl, gpu.rs:572-651):
    density colouring, 4x MSAA, on the 1024^2 map; wall time of the whole
    call (rasterize, sort, resolve, 4 MiB readback) and of the noisiness
    measurement on the host."""
    with bt.Gpu(opts) as blank:
        rgb = bt.noisiness_background(blank.render())
    gpu.render()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        img = gpu.render()
        times.append(time.perf_counter() - t)
    t = time.perf_counter()
    m = bt.noisiness(img, rgb)
    noise_ms = (time.perf_counter() - t) * 1e3
    ms = float(np.median(times) * 1e3)
    log(f"render: {img.shape[1]}x{img.shape[0]} in {ms:.2f} ms, noisiness {m['noisiness']:.4f}")
    return {"size": [img.shape[1], img.shape[0]], "msaa": 4, "color_by": "density",
            "ms": ms, "noisiness": m["noisiness"], "noisiness_ms": noise_ms}


# FIFO for throughput; shortest-first with 1M sample-updates per tick for
# latency (shorter ticks let small jobs through sooner, at some throughput)
SERVER_POLICIES = (("server", "FIFO", 0), ("server_shortest", "SHORTEST", 1 << 20))


def serve(bt, jobs, policy, tick):
    """Feeds the arrival stream to a server in real time; returns (makespan s,
    [(job index, latency ms)], ticks, (start, end) on the monotonic clock)."""
    n_jobs = len(jobs)
    srv = bt.Server(bt.Options(), memory_budget=96 << 30, max_inflight=128, policy=policy,
                    tick_samples=tick)
    clk0 = srv.clock_ms()
    t0 = time.perf_counter()
    pending, slat, i, done = {}, [], 0, 0
    ticks = 0
    while done < n_jobs:
        now = time.perf_counter() - t0
        while i < n_jobs and jobs[i][0] <= now:
            ta, g, o, it = jobs[i]
            pending[srv.submit(g, o, it)] = (ta, i)
            i += 1
        running, queued = srv.step()
        ticks += 1
        for tk in list(pending):
            r = srv.poll(tk)
            if r.state >= 2:
                ta, k = pending.pop(tk)
                slat.append((k, r.done_ms - (clk0 + ta * 1e3)))
                srv.release(tk)
                done += 1
        if running == 0 and queued == 0 and i < n_jobs:
            time.sleep(max(0.0, jobs[i][0] - (time.perf_counter() - t0)))
    t1 = time.perf_counter()
    srv.close()
    return t1 - t0, slat, ticks, (t0, t1)


def profile_batched(bt, workloads, n_graphs=256, iters=30):
    """For rocprofv3 runs: upload the batched graphs, capture the schedule as
    a hipGraph, rewind, replay it; prints the replay's window."""
    graphs = workloads.batched_graphs(n_graphs)
    opts = [workloads.batched_options(i, iters) for i in range(n_graphs)]
    with bt.Gpu(opts[0]) as gpu:
        s = gpu.upload_batch(graphs, opts)
        gpu.iterate(iters, graph=True)
        gpu.rewind()
        t0 = time.perf_counter()
        gpu.iterate(iters, graph=True)
        t1 = time.perf_counter()
    log(f"batched replay: {int(s.sum())} samples, {(t1 - t0) * 1e3:.1f} ms, "
        f"window {t0:.4f} {t1:.4f}")


def profile_served(bt, workloads, n_jobs=100, rate_hz=200.0):
    """For rocprofv3 runs: the served stream under each server policy; prints
    each run's window."""
    jobs = workloads.served_jobs(n_jobs, rate_hz)
    serve(bt, jobs[:3], bt.Server.FIFO, 0)  # warm the code paths
    for name, policy, tick in SERVER_POLICIES:
        srv_s, _, ticks, (t0, t1) = serve(bt, jobs, getattr(bt.Server, policy), tick)
        log(f"served {name}: makespan {srv_s:.3f} s, {ticks} ticks, window {t0:.4f} {t1:.4f}")


def served_section(bt, workloads, n_jobs, rate_hz=200.0):
    """Poisson arrivals at rate_hz: the plain per-job form (FIFO, one job at a
    time, as the drop-in does) against the continuous-batching server."""
    t = time.perf_counter()
    jobs = workloads.served_jobs(n_jobs, rate_hz)
    log(f"served: generated {n_jobs} jobs in {time.perf_counter() - t:.1f} s")
    total_updates = 0
    # plain per-job FIFO
    gpu = bt.Gpu()
    gpu.upload_batch([jobs[0][1]], [jobs[0][2]])
    gpu.iterate(2)
    lat = []
    t0 = time.perf_counter()
    for ta, g, o, it in jobs:
        now = time.perf_counter() - t0
        if now < ta:
            time.sleep(ta - now)
        s = gpu.upload_batch([g], [o])
        gpu.iterate(it)
        gpu.points(0)
        lat.append((time.perf_counter() - t0 - ta) * 1e3)
        total_updates += it * int(s[0])
    plain_s = time.perf_counter() - t0
    gpu.close()
    log(f"served: plain FIFO makespan {plain_s:.2f} s")
    # continuous-batching server, both policies
    runs = {}
    for name, policy, tick in SERVER_POLICIES:
        srv_s, slat, ticks, _ = serve(bt, jobs, getattr(bt.Server, policy), tick)
        runs[name] = (srv_s, slat, ticks)
        log(f"served: {name} makespan {srv_s:.2f} s in {ticks} ticks")
    small = set(k for k, (_, g, _, _) in enumerate(jobs) if g.edge_count <= 20000)
    out = {"jobs": n_jobs, "rate_hz": rate_hz, "updates": total_updates,
           "plain_fifo": {"updates_per_s": total_updates / plain_s, "makespan_s": plain_s,
                          "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                          "p99_ms_le20k_edges": pct([lat[k] for k in sorted(small)], 99)}}
    for name, (srv_s, slat, ticks) in runs.items():
        out[name] = {"updates_per_s": total_updates / srv_s, "makespan_s": srv_s,
                     "p50_ms": pct([v for _, v in slat], 50),
                     "p99_ms": pct([v for _, v in slat], 99),
                     "p50_ms_le20k_edges": pct([v for k, v in slat if k in small], 50),
                     "p99_ms_le20k_edges": pct([v for k, v in slat if k in small], 99),
                     "ticks": ticks, "speedup": plain_s / srv_s}
    out["server_shortest"]["tick_samples"] = 1 << 20
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cell-split", type=int, default=0)
    ap.add_argument("--batched", type=int, default=256, help="graphs in the batched case (0: skip)")
    ap.add_argument("--served", type=int, default=100, help="jobs in the served case (0: skip)")
    ap.add_argument("--large", type=int, default=1, help="run the 100,000-edge case (0: skip)")
    ap.add_argument("--render", type=int, default=1,
                    help="time the export image of the bundled plain GPU case (0: skip)")
    ap.add_argument("--profile-only", action="store_true",
                    help="for rocprofv3 counter runs: a 256 MiB device copy (calibration), "
                         "then upload + autotune + one eager 30-iteration schedule")
    ap.add_argument("--profile-batched", action="store_true",
                    help="for rocprofv3 kernel traces: the batched schedule (capture, rewind, "
                         "replay), printing the replay's window on the monotonic clock")
    ap.add_argument("--profile-served", action="store_true",
                    help="for rocprofv3 kernel traces: the served stream under each server "
                         "policy, printing each run's window on the monotonic clock")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this "
              "single-GPU build (jobs shard across GPUs without exchange; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import bundling_thingy_amd as bt
    from bundling_thingy_amd import workloads

    if args.profile_batched or args.profile_served:
        if args.profile_batched:
            profile_batched(bt, workloads)
        if args.profile_served:
            profile_served(bt, workloads)
        return

    graph = workloads.plain_gpu_graph()
    opts = workloads.plain_gpu_options(ITERS, cell_split=args.cell_split)
    gpu = bt.Gpu(opts)
    if args.profile_only:
        ms = gpu.copy_probe(COPY_PROBE_BYTES)
        log(f"copy probe: {COPY_PROBE_BYTES} bytes in {ms:.3f} ms")
    t = time.perf_counter()
    n = gpu.upload(graph)
    upload_ms = (time.perf_counter() - t) * 1e3
    log(f"upload: {n} samples in {upload_ms:.1f} ms")
    if args.profile_only:
        if not args.cell_split:
            gpu.autotune(ITERS, TUNE_CANDIDATES)
        gpu.iterate(ITERS)
        log("profile-only schedule done")
        gpu.close()
        return

    # autotune the grid cell split per iteration (results are bit-identical
    # under any split; only speed changes)
    tune = None
    if not args.cell_split:
        t = time.perf_counter()
        chosen, best = gpu.autotune(ITERS, TUNE_CANDIDATES)
        tune = {"splits": chosen.tolist(), "seconds": time.perf_counter() - t}
        log(f"autotune: {chosen.tolist()} in {tune['seconds']:.2f} s")

    # warmup (also builds the hipGraph of the schedule)
    t = time.perf_counter()
    for _ in range(max(args.warmup, 1)):
        gpu.rewind()
        gpu.iterate(ITERS, graph=True)
    log(f"warmup done in {time.perf_counter() - t:.2f} s")

    times = []
    for s in range(args.steps):
        gpu.rewind()
        t = time.perf_counter()
        gpu.iterate(ITERS, graph=True)   # synchronizes the stream before returning
        times.append(time.perf_counter() - t)
        log(f"step {s}: {times[-1] * 1e3:.2f} ms")
    total = float(sum(times))
    ms = total / args.steps * 1e3
    value = ITERS * n * args.steps / total

    # Per-stage device time from HIP events on the engine's stream (eager run of
    # the same schedule), plus the grid counts for the FP32-work figure.
    gpu.rewind()
    stage_ms = {k: 0.0 for k in bt.api.STAGES}
    flops = 0
    per_iter = []
    for i in range(ITERS):
        st, _ = gpu.profile(1)
        for k, v in st.items():
            stage_ms[k] += v
        _, counts = gpu.grid()
        p = gpu.job_params(i)
        info = gpu.info()
        per_iter.append({"radius": round(p.radius, 3), "cell": round(p.cell, 3),
                         "mean_ms": round(st["neighbor_mean"], 3),
                         "split_tasks": info.split_tasks, "split_parts": info.split_parts})
        ky = max(1, int(round(p.radius / p.cell)))      # rows per radius
        kx = max(1, int(round(p.radius / p.cell_w)))    # bins per radius
        gy, gx = counts.shape
        pad = np.zeros(((gy + ky - 1) // ky * ky, (gx + kx - 1) // kx * kx), dtype=np.int64)
        pad[:gy, :gx] = counts
        coarse = pad.reshape(pad.shape[0] // ky, ky, pad.shape[1] // kx, kx).sum(axis=(1, 3))
        flops += candidate_flops(coarse)
    dom = max(stage_ms, key=stage_ms.get)
    mean_ms = stage_ms["neighbor_mean"] / ITERS
    P = 1024 * 1024
    alg_bytes = 24 * n + 4 * n  # read + write 12 B records, splat 4 B per sample
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "neighbor_mean_plain_gpu_case_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["bytes_per_dispatch"]
    roof = {"kernel": "bt_neighbor_mean+bt_neighbor_mean_heavy", "bound": "hbm",
            "achieved": alg_bytes / (mean_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "traffic": traffic, "algorithmic_bytes": alg_bytes,
            "avg_launch_ms": mean_ms}
    roof["frac"] = roof["achieved"] / roof["peak"]
    valu = {"kernel": "bt_neighbor_mean+bt_neighbor_mean_heavy", "bound": "valu",
            "achieved": flops / (stage_ms["neighbor_mean"] * 1e-3) / 1e12,
            "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "flops_per_schedule": flops}
    valu["frac"] = valu["achieved"] / valu["peak"]
    iter_bytes = 24 * n + 28 * P
    path = {"bytes_per_iteration": iter_bytes,
            "achieved_GBs": iter_bytes * ITERS / (total / args.steps) / 1e9}

    render = render_section(bt, gpu, opts) if args.render else None
    gpu.close()
    batched = batched_section(bt, workloads, args.batched) if args.batched else None
    served = served_section(bt, workloads, args.served) if args.served else None
    large = large_section(bt, workloads) if args.large else None

    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(graph, opts)
        log(f"cpu baseline: {cpu['value']:.3e} updates/s ({time.perf_counter() - t:.1f} s)")

    out = {
        "metric": "sample-point updates/s (plain, 1 GPU)",
        "value": value,
        "unit": "sample-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator; reference datasets are LFS stubs)",
        "config": {"workload": "plain GPU case: 1,715-node 20-component GMM, 9,800 edges, "
                               "spacing 0.002, 30 iterations radius 0.06->0.01, 1024^2 density "
                               "sigma 8->2",
                   "samples": n, "iterations": ITERS, "density": [1024, 1024]},
        "latency": {"p50_ms": float(np.percentile(np.array(times) * 1e3, 50)),
                    "p99_ms": float(np.percentile(np.array(times) * 1e3, 99)),
                    "upload_ms": upload_ms},
        "autotune": tune,
        "per_iteration": per_iter,
        "stage_ms_per_schedule": stage_ms,
        "dominant_stage": dom,
        "roofline": roof,
        "valu": valu,
        "path": path,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "large": large,
        "render": render,
    }
    if cpu:
        out["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
