"""GPU parity at the workloads' full sizes: the plain GPU case at its largest
radius and over its whole schedule, the 256-graph batched job table and the
served arrival stream, through the C ABI.

- Plain GPU case (9,800 edges, 1.33M samples): the first iteration (radius
  0.06 = 30 texels, where the heavy-task split runs) bit-exact against the
  oracle computed here; every iteration of the 30-iteration schedule against
  the oracle's position hashes and distortion recorded in
  tests/golden/plain_gpu_schedule.json (made by make_schedule_goldens.py),
  and the final noisiness within 1 % (north_star bar; it is in fact equal).
- Batched (256 graphs, 50-20,000 edges): after the 30-iteration hipGraph,
  every job's positions and normalized density equal that job run alone;
  the smallest, median and largest jobs equal the oracle after 2 iterations.
- Served (the 100-job Poisson stream, arrivals compressed into ticks): every
  job's final positions and density equal the job alone under both policies;
  two jobs equal the oracle over all their iterations.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from bundling_thingy_amd import workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BACKGROUND = (6, 6, 9)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _bits_equal(a, b, label):
    np.testing.assert_array_equal(np.ascontiguousarray(a).view(np.uint32),
                                  np.ascontiguousarray(b).view(np.uint32), err_msg=label)


def _oracle_run(orc, graph, opts, iterations, threads=0):
    pl = orc.Pipeline(graph.points, graph.starts, graph.bbox, opts.iter_params(0))
    for it in range(iterations):
        pl.iterate(opts.iter_params(it), threads=threads)
    return pl


def test_plain_gpu_case_largest_radius_vs_oracle(bt, orc):
    """Iteration 1 of the plain GPU case: radius 30 texels, sigma 8. Long
    candidate streams are split across waves (split_tasks > 0) and the
    result is still the oracle's, positions and density, bit for bit."""
    g = workloads.plain_gpu_graph()
    opts = workloads.plain_gpu_options(30)
    p0 = opts.iter_params(0)
    assert abs(p0.radius - 30.0) < 1e-4
    with bt.Gpu(opts) as gpu:
        n = gpu.upload(g)
        gpu.iterate()
        info = gpu.info()
        got = gpu.points()[0]
        dens = gpu.density(3)
        counts = gpu.density(0)
    assert info.split_tasks > 0 and info.split_parts >= info.split_tasks
    pl = _oracle_run(orc, g, opts, 1)
    assert n == len(pl.points)
    diag = float(np.linalg.norm(pl.points.max(0) - pl.points.min(0)))
    assert float(np.abs(got.astype(np.float64) - pl.points).max()) <= 1e-5 * diag
    _bits_equal(got, pl.points, "positions after iteration 1")
    np.testing.assert_array_equal(counts, pl.density_counts)
    _bits_equal(dens, pl.normalized, "normalized density after iteration 1")


@pytest.fixture(scope="module")
def plain_golden():
    with open(os.path.join(GOLDEN, "plain_gpu_schedule.json")) as f:
        return json.load(f)


def test_plain_gpu_full_schedule_vs_oracle(bt, plain_golden):
    """All 30 iterations through bt_iterate (the drop-in, one blocking call
    per iteration): each iteration's positions hash to the oracle's, the
    distortion after each equals the oracle's (within 1 % asserted), and the
    final normalized density, image and noisiness equal the oracle's. The
    same schedule replayed as one hipGraph after a rewind ends on the same
    bits."""
    gold = plain_golden
    g = workloads.plain_gpu_graph()
    opts = workloads.plain_gpu_options(30)
    with bt.Gpu(opts) as gpu:
        n = gpu.upload(g)
        assert n == gold["samples"]
        layout = gpu.distortion_layout()
        base = gpu.distortion_snapshot(layout)
        for row in gold["iterations"]:
            gpu.iterate()
            snap = gpu.distortion_snapshot(layout)
            assert _sha(snap.positions) == row["positions_sha256"], f"iteration {row['iteration']}"
            d = bt.distortion(base, snap, 1024)[0]
            assert abs(d - row["distortion"]) <= 0.01 * row["distortion"]
            assert d == pytest.approx(row["distortion"], rel=1e-12)
        final = gold["final"]
        assert _sha(gpu.density(3)) == final["normalized_density_sha256"]
        img = gpu.render("viridis", 0, 4, False)
        noise = bt.noisiness(img, BACKGROUND)["noisiness"]
        assert abs(noise - final["noisiness"]) <= 0.01 * final["noisiness"]
        assert _sha(img) == final["image_sha256"]
        gpu.rewind()
        gpu.iterate(30, graph=True)
        assert _sha(gpu.points()[0]) == gold["iterations"][-1]["positions_sha256"]


@pytest.fixture(scope="module")
def batched_case():
    graphs = workloads.batched_graphs(256)
    opts = [workloads.batched_options(i, 30) for i in range(256)]
    return graphs, opts


def test_batched_256_graphs_equal_alone(bt, batched_case):
    """The batched configuration at full size: 256 graphs of 50-20,000
    edges with per-graph radii, maps and sigmas, one launch per stage, the
    30-iteration schedule captured as a hipGraph. Each job's positions and
    normalized density equal the job uploaded and iterated alone."""
    graphs, opts = batched_case
    with bt.Gpu(opts[0]) as gpu:
        samples = gpu.upload_batch(graphs, opts)
        gpu.iterate(30, graph=True)
        batched = [(_sha(gpu.points(j)[0]), _sha(gpu.density(3, j))) for j in range(256)]
    assert len(samples) == 256 and int(samples.sum()) > 2e7
    with bt.Gpu(opts[0]) as gpu:
        for j, (gr, o) in enumerate(zip(graphs, opts)):
            gpu.upload_batch([gr], [o])
            gpu.iterate(30)
            assert (_sha(gpu.points(0)[0]), _sha(gpu.density(3, 0))) == batched[j], f"job {j}"


def test_batched_jobs_vs_oracle(bt, orc, batched_case):
    """The smallest, median and largest jobs of the 256-graph table after
    two batched iterations, against the oracle."""
    graphs, opts = batched_case
    with bt.Gpu(opts[0]) as gpu:
        samples = gpu.upload_batch(graphs, opts)
        order = np.argsort(samples, kind="stable")
        pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
        gpu.iterate(2, graph=True)
        got = {j: (gpu.points(j)[0], gpu.density(3, j), gpu.density(0, j)) for j in pick}
    for j in pick:
        pl = _oracle_run(orc, graphs[j], opts[j], 2)
        assert len(pl.points) == samples[j]
        _bits_equal(got[j][0], pl.points, f"job {j} positions")
        np.testing.assert_array_equal(got[j][2], pl.density_counts, err_msg=f"job {j} counts")
        _bits_equal(got[j][1], pl.normalized, f"job {j} density")


@pytest.fixture(scope="module")
def served_case(bt):
    """The served stream (100 Poisson arrivals at 200 jobs/s, 100-50,000
    edges, 20-50 iterations, 512^2-2048^2) and each job's results alone
    (drop-in: upload, then one bt_iterate per iteration), as hashes."""
    jobs = workloads.served_jobs(100)
    alone = []
    with bt.Gpu() as gpu:
        for _, g, o, it in jobs:
            gpu.upload_batch([g], [o])
            gpu.iterate(it)
            alone.append((_sha(gpu.points(0)[0]), _sha(gpu.density(3, 0))))
    return jobs, alone


def _serve_compressed(bt, jobs, policy, tick, per_step=3, keep=(), flags=0, lane=0,
                      masks=(0, 0)):
    """Submits `per_step` arrivals before each tick (about one tick's worth of
    the 200 jobs/s stream at full load), so admissions keep arriving while
    earlier jobs run. Returns {job index: (positions hash, density hash)} and
    {job index: (positions, density)} for the jobs in `keep`."""
    srv = bt.Server(bt.Options(), memory_budget=96 << 30, max_inflight=128, policy=policy,
                    tick_samples=tick, lane_samples=lane, small_cu_mask=masks[0],
                    large_cu_mask=masks[1])
    try:
        pending, out, raw, i = {}, {}, {}, 0
        for _ in range(100000):
            for _ in range(per_step):
                if i < len(jobs):
                    _, g, o, it = jobs[i]
                    if flags:
                        o2 = bt.Options()
                        o2.c = type(o.c).from_buffer_copy(o.c)
                        o2.c.flags |= flags
                        o = o2
                    pending[srv.submit(g, o, it)] = i
                    i += 1
            running, queued = srv.step()
            for t in list(pending):
                r = srv.poll(t)
                if r.state >= 2:
                    k = pending.pop(t)
                    assert r.state == 2, f"job {k} failed"
                    assert r.iterations_done == r.iterations_total == jobs[k][3]
                    o = jobs[k][2].iter_params(0)
                    pos = srv.points(t, r.samples)[0]
                    dens = srv.density(t, o.density_w, o.density_h)
                    out[k] = (_sha(pos), _sha(dens))
                    if k in keep:
                        raw[k] = (pos, dens)
                    srv.release(t)
            if i == len(jobs) and not pending:
                break
        assert len(out) == len(jobs)
        return out, raw
    finally:
        srv.close()


SE3, SE012 = 0xFF000000, 0x00FFFFFF  # shader engine 3 / 0-2 of every XCD


@pytest.mark.parametrize("policy,tick,lane,flags,masks", [
    (0, 0, 0, 0, (0, 0)), (1, 1 << 20, 0, 0, (0, 0)), (0, 0, 0, 4, (0, 0)),
    (1, 1 << 20, 0, 4, (0, 0)), (2, 0, 1 << 18, 0, (0, 0)), (2, 1 << 20, 1 << 18, 4, (0, 0)),
    (2, 1 << 20, 3 << 16, 4, (SE3, SE012)), (2, 0, 1 << 18, 0, (SE012, SE3))])
def test_served_stream_equals_alone(bt, served_case, policy, tick, lane, flags, masks):
    """Every job of the 100-job served stream, co-scheduled with the others
    under continuous batching (FIFO, shortest-first with 1M sample-updates
    per tick, and two lanes split at 256k or 192k samples, without and with a 1M
    per-lane tick budget, on any CU or with the lanes on disjoint shader
    engines while both have work), ends with the bits of the job alone,
    whether the jobs compute their density every iteration or only at the
    last one (flags 4 = BT_FLAG_DENSITY_ON_READ)."""
    jobs, alone = served_case
    out, _ = _serve_compressed(bt, jobs, policy, tick, flags=flags, lane=lane, masks=masks)
    for k in range(len(jobs)):
        assert out[k] == alone[k], f"job {k}"


def test_served_jobs_vs_oracle(bt, orc, served_case):
    """The smallest and the median job of the served stream, co-scheduled
    with the whole stream, against the oracle over all their iterations."""
    jobs, _ = served_case
    sizes = np.array([len(j[1].points) for j in jobs])
    order = np.argsort(sizes, kind="stable")
    pick = [int(order[0]), int(order[len(order) // 2])]
    _, out = _serve_compressed(bt, jobs, 0, 0, keep=pick)
    for k in pick:
        _, g, o, it = jobs[k]
        pl = _oracle_run(orc, g, o, it)
        _bits_equal(out[k][0], pl.points, f"job {k} positions")
        _bits_equal(out[k][1], pl.normalized, f"job {k} density")
