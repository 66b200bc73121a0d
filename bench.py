"""Benchmark: sample-point updates/s of the edge-bundling hot path on one MI355X.

Workload (plain GPU case): 1,715 nodes from a 20-component Gaussian mixture,
9,800 straight edges (seed 2), resampled at spacing 0.002 (~1.33M samples),
30 iterations with the neighbor radius decaying 0.06 -> 0.01 geometrically,
1024 x 1024 FP32 density with sigma 8 -> 2. One step = the whole 30-iteration
schedule (the scope of 30 Gpu::iterate calls) replayed as one hipGraph from
the post-upload state; upload/resampling is timed separately.

Prints one JSON line (see the keys below). Extra objects: roofline (dominant
kernel vs HBM peak, from HIP events), executed (the same kernel's executed
test-loop instructions vs the SIMDs' issue capacity), cpu_baseline (the CPU
oracle on a bounded sample of the same schedule), latency.
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
# VALU issue capacity: 256 CUs x 4 SIMDs at 2.4 GHz, 16 lanes per SIMD, so a
# wave64 instruction every 4 cycles per SIMD (the published 157.3 TFLOP/s FP32
# = 256 x 4 x 16 lanes x 2 (packed FP32) x 2 (FMA) x 2.4e9)
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4
# VALU instructions per wave-step of the neighbor test loops (gfx950
# assembly). Whole-wave streams: 47 per 4 candidates (2 sub, 3 mul, 1 fma,
# 1 compare per test, 3 masked adds folded into add3/addc, one pointer
# increment per 4). Lane-group streams, whose sums are int8 matrix
# products: 499 per 64 (the same 6 arithmetic instructions per test and an
# integer subtraction whose sign is the test's outcome; per 4 tests two
# v_perm_b32 and an OR pack the A bytes; the 4 MFMAs per chunk run on the
# matrix units).
VALU_PER_STEP = 47 / 4
VALU_PER_STEP_MF = 499 / 64
ITERS = 30
TUNE_CANDIDATES = (1, 2, 3, 4, 6, 8, 12, 16, 20, 24, 28)
COPY_PROBE_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(graph, opts, positions, target_s=12.0):
    """The CPU oracle on all host threads, following the schedule: iteration
    i's neighbor means are computed from the positions the GPU run fed that
    iteration (`positions[i]`, read back after each eager iteration), with
    the radius of iteration i, on every stride-th sample, scaled by the
    stride; the per-call work that does not scale with the sample (sorting
    all samples into the grid, the cell summaries) is timed by a one-sample
    call and counted once. The density stages (splat, blur, normalize) run
    in full. Two forms of the mean: the reference's tiled algorithm (cells of
    side r, fully-inside cells by their summaries, neighbor_mean.wgsl:40-161)
    -- the baseline -- and the plain window scan of every candidate."""
    from oracle import oracle as orc
    p0 = opts.iter_params(0)
    pl = orc.Pipeline(graph.points, graph.starts, graph.bbox, p0, with_density=False)
    fixed = pl.fixed
    n = len(fixed)
    threads = os.cpu_count() or 1
    threads = max(1, min(threads, int(os.environ.get("OMP_NUM_THREADS", threads))))
    params = [opts.iter_params(i) for i in range(ITERS)]
    grids = [orc.make_grid(p.res_x, p.radius) for p in params]
    fns = {"cells": orc.neighbor_mean_cells, "window": orc.neighbor_mean_strided}

    def schedule(fn, stride, threads=threads):
        """Scaled schedule time, the sample's own mean time, and the time of
        the parts that do not scale with the sample."""
        total = tmean = tother = 0.0
        for pts, g, p in zip(positions, grids, params):
            t = time.perf_counter()
            fn(pts, fixed, g, n, threads)  # one sample: the per-call fixed work
            tf = time.perf_counter() - t
            t = time.perf_counter()
            out = fn(pts, fixed, g, stride, threads)
            tm = max(time.perf_counter() - t - tf, 0.0)
            t = time.perf_counter()
            orc.density(orc.splat(out, p.density_w, p.density_h), p.sigma)
            td = time.perf_counter() - t
            total += tm * stride + tf + td
            tmean += tm
            tother += tf + td
        return total, tmean, tother

    res = {}
    for name, fn in fns.items():
        # a coarse pass over the whole schedule sets the stride of the timed
        # pass (one sample in 64: enough that the means' time stands clear of
        # the per-call time subtracted from it)
        coarse = 64
        _, tm, other = schedule(fn, coarse)
        avail = max(target_s / 2.0 - other, 1.0)  # the timed pass takes about target_s / 2
        stride = int(max(1, min(coarse, np.ceil(coarse * tm / avail))))
        total, _, _ = schedule(fn, stride)
        res[name] = (total, stride)
        log(f"cpu baseline [{name}]: stride {stride}, est. {total:.1f} s per schedule")
    total, stride = res["cells"]
    stride1 = int(stride * threads)
    total1, _, _ = schedule(fns["cells"], stride1, threads=1)
    wtotal, wstride = res["window"]
    return {"value": ITERS * n / total, "unit": "sample-point updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle restatement of the reference's tiled neighbor mean (cells of side r, "
                      f"fully-inside cells by summary) on the positions each of the {ITERS} "
                      f"iterations of the GPU run started from, one sample in {stride} (x{stride}) "
                      f"plus full splat/blur/normalize; est. {total:.1f} s per schedule",
            "value_1thread": ITERS * n / total1,
            "sample_1thread": f"the same, one sample in {stride1}, 1 thread; est. {total1:.1f} s "
                              f"per schedule",
            "window_scan_value": ITERS * n / wtotal,
            "window_scan_sample": f"every candidate of the cells within reach tested, one sample "
                                  f"in {wstride}; est. {wtotal:.1f} s per schedule"}


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
    # batched form: kernel shapes autotuned per iteration on the whole batch
    # (outside the timed region, like the plain case's autotuning)
    gpu.upload_batch(graphs, opts)
    t = time.perf_counter()
    shapes, _ = gpu.autotune_shapes(iters)
    tune_s = time.perf_counter() - t
    log(f"batched: shapes {shapes.tolist()} in {tune_s:.2f} s")
    t0 = time.perf_counter()
    s = gpu.upload_batch(graphs, opts)
    up_s = time.perf_counter() - t0
    t = time.perf_counter()
    for i in range(iters):
        gpu.set_shapes(*(int(v) for v in shapes[i]), iteration=i)
    up_s += time.perf_counter() - t
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
                        "shapes": shapes.tolist(), "autotune_s": tune_s,
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
    shapes, _ = gpu.autotune_shapes(iters)
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
            "splits": chosen.tolist(), "shapes": shapes.tolist()}


def render_section(bt, gpu, opts, reps=5):
    """Export image of the bundled graph (render_paths.wgsl, gpu.rs:572-651):
    density colouring, 4x MSAA, on the 1024^2 map; wall time of the whole
    call (rasterize, sort, resolve, 4 MiB readback) and of the noisiness
    measurement on the host; the median of `reps` calls after three
    untimed ones."""
    with bt.Gpu(opts) as blank:
        rgb = bt.noisiness_background(blank.render())
    for _ in range(3):  # warm-up: first-call allocations, fresh host pages
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


# (name, policy, tick_samples, lane_samples, small_cu_mask, large_cu_mask).
# FIFO for throughput; shortest-first with 1M sample-updates per tick for
# latency (shorter ticks let small jobs through sooner, at some throughput);
# two lanes (jobs of at most 192k samples tick on their own stream, larger
# ones on another, overlapping on the GPU), each lane's tick taking the least
# remaining work first up to 1M sample-updates while no job waits for
# admission: on any CU, or with the small lane on shader engine 3 of every
# XCD and the large lane on shader engines 0-2 while both have work (each
# lane on every CU while the other has none).
SE3, SE012 = 0xFF000000, 0x00FFFFFF
SERVER_POLICIES = (("fifo", "FIFO", 0, 0, 0, 0), ("shortest_1M", "SHORTEST", 1 << 20, 0, 0, 0),
                   ("lanes_192k_1M", "LANES", 1 << 20, 3 << 16, 0, 0),
                   ("lanes_192k_1M_se", "LANES", 1 << 20, 3 << 16, SE3, SE012))
# Each server is warmed before its measured run with jobs of another seed (a
# long-running server has its allocator slabs and pinned result buffers in
# place); finished jobs are read into two reused host buffers.
WARM_JOBS = 64
_REC_BUF = None
_DENS_BUF = None
_WARM = {}
# A second, lighter arrival stream (the same jobs at half the rate) for the
# latency-load relation, under the final-only density policy.
LIGHT_RATE_FACTOR = 0.5
# Density policies, each applied to both forms: the density of every
# iteration (as Gpu::iterate computes it), or only the one that is returned,
# the last iteration's (BT_FLAG_DENSITY_ON_READ).
DENSITY_POLICIES = (("final_only", 4), ("every_iteration", 0))


def _flagged(bt, opts, flags):
    """A copy of job options with extra bt_options.flags set."""
    o = bt.Options()
    o.c = type(opts.c).from_buffer_copy(opts.c)
    o.c.flags |= flags
    return o


def _drive(bt, srv, jobs, flags, realtime):
    """Runs the jobs through the server (see serve); returns (makespan s,
    [(job index, latency ms)], ticks, (start, end))."""
    global _REC_BUF, _DENS_BUF
    if _REC_BUF is None:
        _REC_BUF = np.zeros(4 << 20, dtype=np.uint8)
        _DENS_BUF = np.zeros(2048 * 2048, dtype=np.float32)
    n_jobs = len(jobs)
    clk0 = srv.clock_ms()
    t0 = time.perf_counter()
    pending, slat, i, done = {}, [], 0, 0
    ticks = 0
    while done < n_jobs:
        now = time.perf_counter() - t0
        while i < n_jobs and (not realtime or jobs[i][0] <= now):
            ta, g, o, it = jobs[i]
            pending[srv.submit(g, _flagged(bt, o, flags) if flags else o, it)] = (
                ta if realtime else 0.0, i)
            i += 1
        running, queued = srv.step()
        ticks += 1
        for tk in srv.take_finished():
            r = srv.poll(tk)
            ta, k = pending.pop(tk)
            if r.state != 2:
                raise RuntimeError(f"served job {k} failed")
            slat.append((k, r.done_ms - (clk0 + ta * 1e3)))
            oc = jobs[k][2].c
            if 12 * r.samples > _REC_BUF.size:
                _REC_BUF = np.zeros(12 * r.samples, dtype=np.uint8)
            srv.points_raw(tk, r.samples, _REC_BUF)
            srv.density(tk, oc.density_w, oc.density_h, _DENS_BUF)
            srv.release(tk)
            done += 1
        if running == 0 and queued == 0 and i < n_jobs:
            time.sleep(max(0.0, jobs[i][0] - (time.perf_counter() - t0)))
    t1 = time.perf_counter()
    return t1 - t0, slat, ticks, (t0, t1)


def serve(bt, workloads, jobs, policy, tick, flags=0, realtime=True, lane=0, masks=(0, 0)):
    """Feeds the jobs to a server, warmed first with WARM_JOBS jobs of
    another seed (all at once, not measured): in real time at their arrival
    times (realtime) or all at once (saturated: the server's capacity). Each
    finished job's samples and density are read into host arrays, as the
    plain per-job form reads them. Returns (makespan s, [(job index, latency
    ms)], ticks, (start, end) on the monotonic clock); latency = completion -
    arrival (saturated: - start)."""
    srv = bt.Server(bt.Options(), memory_budget=96 << 30, max_inflight=128, policy=policy,
                    tick_samples=tick, lane_samples=lane, small_cu_mask=masks[0],
                    large_cu_mask=masks[1])
    if WARM_JOBS:
        if _WARM.get("jobs") is None:
            _WARM["jobs"] = workloads.served_jobs(WARM_JOBS, seed=77)
        _drive(bt, srv, _WARM["jobs"], flags, False)
    try:
        return _drive(bt, srv, jobs, flags, realtime)
    finally:
        srv.close()


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


def profile_served(bt, workloads, n_jobs=1000, rate_hz=200.0):
    """For rocprofv3 runs: the served stream (final-only density) under each
    server policy, in real time; prints each run's window."""
    jobs = workloads.served_jobs(n_jobs, rate_hz)
    serve(bt, workloads, jobs[:3], bt.Server.FIFO, 0, 4)  # warm the code paths
    for name, policy, tick, lane, m_small, m_large in SERVER_POLICIES:
        srv_s, _, ticks, (t0, t1) = serve(bt, workloads, jobs, getattr(bt.Server, policy), tick,
                                          4, lane=lane, masks=(m_small, m_large))
        log(f"served {name}: makespan {srv_s:.3f} s, {ticks} ticks, window {t0:.4f} {t1:.4f}")


def _plain_per_job(bt, jobs, flags):
    """The plain per-job drop-in over the jobs, one after another with no
    waiting (its capacity): upload, one blocking bt_iterate per iteration,
    read the samples and the density (into two reused host buffers, as the
    server's results are read). Returns (seconds, [service seconds],
    sample-updates)."""
    rec = np.zeros(12 * 4 << 20, dtype=np.uint8)
    dens = np.zeros(2048 * 2048, dtype=np.float32)
    gpu = bt.Gpu()
    gpu.upload_batch([jobs[0][1]], [_flagged(bt, jobs[0][2], flags)])
    gpu.iterate(2)
    gpu.density(3, 0)
    service, updates = [], 0
    t0 = time.perf_counter()
    for _, g, o, it in jobs:
        ts = time.perf_counter()
        n = gpu.upload_batch([g], [_flagged(bt, o, flags)])
        gpu.iterate(it)
        gpu.points_raw(0, rec)
        gpu.density(3, 0, dens)
        service.append(time.perf_counter() - ts)
        updates += it * int(n[0])
    total = time.perf_counter() - t0
    gpu.close()
    return total, service, updates


def _fifo_queue_latency(arrivals, service):
    """Latencies (ms) of a single FIFO server with these arrival times and
    service times: what the plain per-job form gives on the arrival stream
    (it is busy from one job's arrival until the job is done)."""
    out, free = [], 0.0
    for ta, d in zip(arrivals, service):
        free = max(free, ta) + d
        out.append((free - ta) * 1e3)
    return out


def served_section(bt, workloads, n_jobs, rate_hz=200.0):
    """Poisson arrivals at rate_hz (100-50,000 edges, 20-50 iterations,
    512^2-2048^2 density, spacing 0.01): the plain per-job drop-in against the
    continuous-batching server, both under the same density policy. For each
    form: capacity (every job available at once) and latency on the arrival
    stream (real time; the plain form's from a FIFO queue of its measured
    service times, since it cannot keep up with the stream)."""
    t = time.perf_counter()
    jobs = workloads.served_jobs(n_jobs, rate_hz)
    log(f"served: generated {n_jobs} jobs in {time.perf_counter() - t:.1f} s")
    small = set(k for k, (_, g, _, _) in enumerate(jobs) if g.edge_count <= 20000)
    arrivals = [j[0] for j in jobs]

    def lat_stats(lat):
        lat = dict(lat)
        return {"p50_ms": pct(list(lat.values()), 50), "p99_ms": pct(list(lat.values()), 99),
                "p50_ms_le20k_edges": pct([v for k, v in lat.items() if k in small], 50),
                "p99_ms_le20k_edges": pct([v for k, v in lat.items() if k in small], 99)}

    out = {"jobs": n_jobs, "rate_hz": rate_hz, "jobs_le20k_edges": len(small),
           "spacing": 0.01, "arrivals_s": arrivals[-1]}
    serve(bt, workloads, jobs[:3], bt.Server.FIFO, 0, 4)  # warm the code paths
    for dname, flags in DENSITY_POLICIES:
        plain_s, service, updates = _plain_per_job(bt, jobs, flags)
        out["updates"] = updates
        out["offered_updates_per_s"] = updates / arrivals[-1]
        plain_cap = updates / plain_s
        # each job's service time alone on an idle GPU: a floor under any
        # policy's latency for that job
        iso = [service[k] * 1e3 for k in sorted(small)]
        sec = {"plain": {"capacity_updates_per_s": plain_cap, "seconds": plain_s,
                         **lat_stats(enumerate(_fifo_queue_latency(arrivals, service))),
                         "isolated_p50_ms_le20k_edges": pct(iso, 50),
                         "isolated_p99_ms_le20k_edges": pct(iso, 99)}}
        log(f"served [{dname}]: plain per-job {plain_s:.2f} s, {plain_cap:.3e} updates/s")
        for name, policy, tick, lane, m_small, m_large in SERVER_POLICIES:
            pol = getattr(bt.Server, policy)
            kw = dict(lane=lane, masks=(m_small, m_large))
            cap_s, _, cap_ticks, _ = serve(bt, workloads, jobs, pol, tick, flags, realtime=False,
                                           **kw)
            rt_s, rt_lat, rt_ticks, _ = serve(bt, workloads, jobs, pol, tick, flags, **kw)
            sec[name] = {"capacity_updates_per_s": updates / cap_s, "capacity_seconds": cap_s,
                         "capacity_vs_plain": plain_s / cap_s, "capacity_ticks": cap_ticks,
                         "stream_updates_per_s": updates / rt_s, "stream_makespan_s": rt_s,
                         "stream_ticks": rt_ticks, **lat_stats(rt_lat)}
            if tick:
                sec[name]["tick_samples"] = tick
            if lane:
                sec[name]["lane_samples"] = lane
            if m_small or m_large:
                sec[name]["cu_masks"] = [hex(m_small), hex(m_large)]
            log(f"served [{dname}] {name}: capacity {updates / cap_s:.3e} updates/s "
                f"({plain_s / cap_s:.2f}x plain); stream p99 {sec[name]['p99_ms']:.1f} ms, "
                f"jobs <= 20k edges p99 {sec[name]['p99_ms_le20k_edges']:.1f} ms")
            if flags & 4:
                light = [(ta / LIGHT_RATE_FACTOR, g, o, it) for (ta, g, o, it) in jobs]
                l_s, l_lat, _, _ = serve(bt, workloads, light, pol, tick, flags, **kw)
                sec[name]["light_stream"] = {"rate_hz": rate_hz * LIGHT_RATE_FACTOR,
                                             "makespan_s": l_s, **lat_stats(l_lat)}
                log(f"served [{dname}] {name} at {rate_hz * LIGHT_RATE_FACTOR:.0f} jobs/s: "
                    f"p99 {sec[name]['light_stream']['p99_ms']:.1f} ms, jobs <= 20k edges p99 "
                    f"{sec[name]['light_stream']['p99_ms_le20k_edges']:.1f} ms")
        out[dname] = sec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cell-split", type=int, default=0)
    ap.add_argument("--batched", type=int, default=256, help="graphs in the batched case (0: skip)")
    ap.add_argument("--served", type=int, default=1000, help="jobs in the served case (0: skip)")
    ap.add_argument("--large", type=int, default=1, help="run the 100,000-edge case (0: skip)")
    ap.add_argument("--render", type=int, default=1,
                    help="time the export image of the bundled plain GPU case (0: skip)")
    ap.add_argument("--profile-only", action="store_true",
                    help="for rocprofv3 counter runs: a 256 MiB device copy (calibration), "
                         "then upload + autotune (splits, then shapes) + one eager "
                         "30-iteration schedule")
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
            gpu.autotune_shapes(ITERS)
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
        t_split = time.perf_counter() - t
        log(f"autotune: splits {chosen.tolist()} in {t_split:.2f} s")
        t = time.perf_counter()
        shapes, shape_ms = gpu.autotune_shapes(ITERS)
        tune = {"splits": chosen.tolist(), "split_seconds": t_split,
                "shapes": shapes.tolist(), "shape_seconds": time.perf_counter() - t,
                "shapes_key": "per iteration: [neighbor-mean shape (0 = 16 lanes x 4 waves, "
                              "1 = 32x4, 2 = 16x2, 3 = 32x2, 4 = 16x8, 5 = 32x8), row blur tile "
                              "(0 = 64x4, 1 = 128x2, 2 = 32x8, 3 = 256x1), column blur tile "
                              "(0 = 64x32, 1 = 64x16, 2 = 64x64; 3 = 64x32 and 4 = 64x16 fused "
                              "row + column kernel)]"}
        log(f"autotune: shapes {shapes.tolist()} in {tune['shape_seconds']:.2f} s")

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
        t1 = time.perf_counter()
        times.append(t1 - t)
        # the window on the monotonic clock, for matching profiler traces
        log(f"step {s}: {times[-1] * 1e3:.2f} ms (window {t:.4f} {t1:.4f})")
    total = float(sum(times))
    ms = total / args.steps * 1e3
    value = ITERS * n * args.steps / total

    # Per-stage device time from HIP events on the engine's stream (eager run of
    # the same schedule), plus the neighbor mean's executed work per iteration.
    gpu.rewind()
    stage_ms = {k: 0.0 for k in bt.api.STAGES}
    steps_total = 0
    valu_total = 0.0
    movable = n - 2 * int(np.count_nonzero(gpu.paths()[:, 2] >= 2))
    per_iter = []
    positions = []  # the input of every iteration, for the CPU baseline
    for i in range(ITERS):
        if not args.no_cpu:
            positions.append(gpu.points()[0])
        st, _ = gpu.profile(1)
        for k, v in st.items():
            stage_ms[k] += v
        p = gpu.job_params(i)
        info = gpu.info()
        steps_total += info.test_steps
        valu = (info.group_steps * VALU_PER_STEP_MF
                + (info.test_steps - info.group_steps) * VALU_PER_STEP)
        valu_total += valu
        nm_ms = st["neighbor_mean"]
        per_iter.append({"radius": round(p.radius, 3), "cell": round(p.cell, 3),
                         "mean_ms": round(nm_ms, 3),
                         "split_tasks": info.split_tasks, "split_parts": info.split_parts,
                         "wave_steps": info.test_steps,
                         "tests_per_sample": round(64 * info.test_steps / movable, 1),
                         "neighbors_per_sample": round(info.neighbors / movable, 1),
                         "group_steps": info.group_steps,
                         "valu_issue_frac": round(valu / (nm_ms * 1e-3) / VALU_ISSUE_PEAK, 3)})
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
    # executed work: wave-steps of the test loop x its VALU instructions,
    # against the SIMDs' issue capacity (the kernel's real bound)
    executed = {"kernel": "bt_neighbor_mean+bt_neighbor_mean_heavy", "bound": "valu",
                "wave_steps_per_schedule": steps_total,
                "lane_tests_per_schedule": 64 * steps_total,
                "valu_per_step": {"lane-group streams (int8 MFMA sums)": VALU_PER_STEP_MF,
                                  "whole-wave streams": VALU_PER_STEP},
                "valu_per_schedule": valu_total,
                "achieved": valu_total / (stage_ms["neighbor_mean"] * 1e-3),
                "peak": VALU_ISSUE_PEAK, "unit": "wave-instructions/s",
                "note": "test-loop VALU instructions only (staging, set-up, range walks and "
                        "the density splat are not counted); the int8 MFMAs of the "
                        "lane-group streams run on the matrix units"}
    executed["frac"] = executed["achieved"] / executed["peak"]
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
        cpu = cpu_baseline(graph, opts, positions)
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
        "executed": executed,
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
