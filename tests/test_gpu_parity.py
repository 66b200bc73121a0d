"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs. Integer outputs (sample counts, path reservations, edge
ids, grid cells, per-cell counts, density counts) must be bit-exact; so must
the FP32 positions and density fields, because the kernels and the oracle
share one arithmetic contract. Tolerances from the reference's parity bar are
asserted too (positions within 1e-5 of the bbox diagonal, density relative
L-inf <= 1e-5) so the bar is visible in the test."""
import ctypes

import numpy as np
import pytest

from bundling_thingy_amd import api, workloads

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5   # of the bounding-box diagonal, after one iteration
DENS_TOL = 1e-5  # relative L-inf


def _oracle_pipeline(orc, graph, options):
    return orc.Pipeline(graph.points, graph.starts, graph.bbox, options.iter_params(0))


def _assert_points(gpu_pts, ref_pts, label):
    diag = float(np.linalg.norm(ref_pts.max(0) - ref_pts.min(0)))
    err = float(np.abs(gpu_pts.astype(np.float64) - ref_pts).max()) if len(ref_pts) else 0.0
    assert err <= POS_TOL * diag, f"{label}: max error {err} > {POS_TOL} * {diag}"
    np.testing.assert_array_equal(gpu_pts.view(np.uint32), ref_pts.view(np.uint32), err_msg=label)


def _assert_density(gpu, pl, label):
    np.testing.assert_array_equal(gpu.density(0), pl.density_counts, err_msg=label + " counts")
    for kind, ref in ((2, pl.blurred), (3, pl.normalized)):
        d = gpu.density(kind)
        scale = max(float(np.abs(ref).max()), 1e-30)
        assert float(np.abs(d - ref).max()) / scale <= DENS_TOL, label
        np.testing.assert_array_equal(d.view(np.uint32), ref.view(np.uint32), err_msg=label)


@pytest.fixture(scope="module")
def cpu_case():
    return workloads.plain_cpu_graph()


def test_reference_mode_upload_and_iterations(bt, orc, cpu_case):
    """Reference mode S=201, r=10, sigma 6 (the plain CPU case)."""
    opts = workloads.reference_cpu_options()
    pl = _oracle_pipeline(orc, cpu_case, opts)
    with bt.Gpu(opts) as gpu:
        n = gpu.upload(cpu_case)
        assert n == len(pl.points)
        paths = gpu.paths()
        np.testing.assert_array_equal(paths[:, 2], pl.counts)
        np.testing.assert_array_equal(paths[:, 0], pl.offsets.astype(np.uint32))
        pts, eid = gpu.points()
        np.testing.assert_array_equal(eid, pl.edge_id)
        _assert_points(pts, pl.points, "upload")
        _assert_density(gpu, pl, "upload")
        for it in range(3):
            gpu.iterate()
            p = opts.iter_params(it)
            ref = pl.iterate(p)
            pts, _ = gpu.points()
            _assert_points(pts, ref, f"iteration {it + 1}")
            _assert_density(gpu, pl, f"iteration {it + 1}")


def test_grid_cells_and_counts_bitexact(bt, orc, cpu_case):
    opts = workloads.reference_cpu_options()
    pl = _oracle_pipeline(orc, cpu_case, opts)
    with bt.Gpu(opts) as gpu:
        gpu.upload(cpu_case)
        before, _ = gpu.points()
        gpu.iterate()
        cells, counts = gpu.grid()
        ocell, ocounts = orc.grid_bin(before, orc.grid(gpu.job_params(0)))
        np.testing.assert_array_equal(cells, ocell)
        np.testing.assert_array_equal(counts.reshape(-1), ocounts)
        assert counts.sum() == len(before)


def test_generalized_cpu_case(bt, orc, cpu_case):
    opts = workloads.plain_cpu_options()
    pl = _oracle_pipeline(orc, cpu_case, opts)
    with bt.Gpu(opts) as gpu:
        assert gpu.upload(cpu_case) == len(pl.points)
        for it in range(2):
            gpu.iterate()
            ref = pl.iterate(opts.iter_params(it))
            _assert_points(gpu.points()[0], ref, f"iteration {it + 1}")
            _assert_density(gpu, pl, f"iteration {it + 1}")


def test_gpu_case_resample_and_last_radius(bt, orc):
    """The plain GPU workload (9,800 edges, ~1.3M samples): resampling and
    density bit-exact; one iteration at the schedule's smallest radius."""
    g = workloads.plain_gpu_graph()
    opts = bt.Options(spacing=0.002, radius=0.01, sigma=2.0, density=1024)
    pl = _oracle_pipeline(orc, g, opts)
    with bt.Gpu(opts) as gpu:
        assert gpu.upload(g) == len(pl.points)
        _assert_points(gpu.points()[0], pl.points, "upload")
        _assert_density(gpu, pl, "upload")
        gpu.iterate()
        ref = pl.iterate(opts.iter_params(0))
        _assert_points(gpu.points()[0], ref, "iteration 1")
        _assert_density(gpu, pl, "iteration 1")


@pytest.mark.parametrize("split,xs", [(1, 1), (1, 8), (2, 8), (3, 2), (5, 8), (8, 4), (16, 8),
                                      (16, 16), (1, 64), (40, 2)])
def test_cell_split_does_not_change_results(bt, orc, cpu_case, split, xs):
    """Any grid row height and bin width give the same bits (inside runs +
    scanned borders equal the all-pairs test), including the reference grid
    (square cells of side r: split 1, x_split 1) and windows taller than one
    row pass (split 40: 82 rows, set up in two passes, plain streams)."""
    base = workloads.plain_cpu_options()
    opts = bt.Options(spacing=0.005, radius=0.05, sigma=4.0, density=512, schedule_len=10,
                      cell_split=split, x_split=xs)
    pl = _oracle_pipeline(orc, cpu_case, base)
    with bt.Gpu(opts) as gpu:
        gpu.upload(cpu_case)
        for it in range(2):
            gpu.iterate()
            ref = pl.iterate(base.iter_params(it))
            _assert_points(gpu.points()[0], ref, f"split {split}/{xs} iteration {it + 1}")
        if split == 1 and xs == 1:
            cells, counts = gpu.grid()
            assert counts.shape == (21, 21)  # ceil(201 / 10): grid of cell = radius


def test_split_heavy_tasks_bitexact(bt, orc, cpu_case):
    """Square cells as wide as the radius leave long candidate streams; those
    tasks are split into parts run by many waves, whose exact partial sums
    must give the same bits as the oracle."""
    opts = bt.Options(spacing=0.005, radius=0.1, sigma=4.0, density=512, cell_split=1,
                      x_split=1)
    pl = _oracle_pipeline(orc, cpu_case, opts)
    with bt.Gpu(opts) as gpu:
        gpu.upload(cpu_case)
        for it in range(3):
            gpu.iterate()
            info = gpu.info()
            # after the first launch the threshold follows the candidates per
            # wave, which are few here: long streams are split
            if it > 0:
                assert info.split_tasks > 0 and info.split_parts > info.split_tasks
            ref = pl.iterate(opts.iter_params(it))
            _assert_points(gpu.points()[0], ref, f"iteration {it + 1}")
            _assert_density(gpu, pl, f"iteration {it + 1}")


def test_autotune_keeps_results(bt, cpu_case):
    opts = workloads.plain_cpu_options()
    with bt.Gpu(opts) as gpu:
        gpu.upload(cpu_case)
        gpu.iterate(4)
        ref = gpu.points()[0]
        chosen, best = gpu.autotune(4, (1, 2, 4, 8))
        assert len(chosen) == 4 and np.all(np.isin(chosen, [1, 2, 4, 8])) and np.all(best > 0)
        assert gpu.job_samples(0)[1] == 0  # rewound
        gpu.iterate(4, graph=True)
        np.testing.assert_array_equal(gpu.points()[0].view(np.uint32), ref.view(np.uint32))
        for it in range(4):
            p = gpu.job_params(it)
            assert abs(p.radius / p.cell - chosen[it]) < 1e-3


def test_deterministic_and_graph_replay(bt, cpu_case):
    """Two eager runs agree bit for bit, and a hipGraph of the schedule
    replayed three times (with rewinds and reads between) equals them."""
    opts = workloads.plain_cpu_options()
    with bt.Gpu(opts) as gpu:
        gpu.upload(cpu_case)
        gpu.iterate(5)
        a = gpu.points()[0]
        da = gpu.density(3)
        gpu.rewind()
        gpu.iterate(5)
        np.testing.assert_array_equal(gpu.points()[0].view(np.uint32), a.view(np.uint32))
        for _ in range(3):
            gpu.rewind()
            _ = gpu.points()
            gpu.iterate(5, graph=True)
            np.testing.assert_array_equal(gpu.points()[0].view(np.uint32), a.view(np.uint32))
            np.testing.assert_array_equal(gpu.density(3).view(np.uint32), da.view(np.uint32))


def test_batched_equals_alone(bt, cpu_case):
    """Jobs advanced together in one launch per stage give the same bits as
    each job alone."""
    graphs = workloads.batched_graphs(count=5, lo=50, hi=600)
    opts = [bt.Options(spacing=0.004, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=256,
                       schedule_len=4),
            bt.Options(map_size=301, neighbor_radius=8),
            workloads.plain_cpu_options(),
            bt.Options(spacing=0.003, radius=0.03, sigma=3.0, density=(320, 200)),
            bt.Options(spacing=0.006, radius=(0.08, 0.04), sigma=5.0, density=128,
                       schedule_len=3)]
    alone = []
    for g, o in zip(graphs, opts):
        with bt.Gpu(o) as gpu:
            gpu.upload(g)
            gpu.iterate(4)
            alone.append((gpu.points()[0], gpu.density(3)))
    with bt.Gpu(opts[0]) as gpu:
        gpu.upload_batch(graphs, opts)
        gpu.iterate(4, graph=True)
        for j in range(len(graphs)):
            p = gpu.points(j)[0]
            np.testing.assert_array_equal(p.view(np.uint32), alone[j][0].view(np.uint32))
        np.testing.assert_array_equal(gpu.density(3, job=0).view(np.uint32),
                                      alone[0][1].view(np.uint32))


def test_full_size_properties(bt):
    """At the full plain GPU workload: fixed endpoints never move, samples
    stay inside the initial bounding box, the grid partitions every sample."""
    g = workloads.plain_gpu_graph()
    opts = workloads.plain_gpu_options()
    with bt.Gpu(opts) as gpu:
        n = gpu.upload(g)
        layout = gpu.distortion_layout()
        base = gpu.distortion_snapshot(layout)
        gpu.iterate(3, graph=True)
        snap = gpu.distortion_snapshot(layout)
        ends = np.concatenate([layout.paths[layout.edges, 0],
                               layout.paths[layout.edges, 0] + layout.paths[layout.edges, 2] - 1])
        np.testing.assert_array_equal(snap.positions[ends], base.positions[ends])
        lo, hi = base.positions.min(0), base.positions.max(0)
        assert np.all(snap.positions >= lo) and np.all(snap.positions <= hi)
        cells, counts = gpu.grid()
        assert counts.sum() == n
        np.testing.assert_array_equal(np.bincount(cells, minlength=counts.size), counts.reshape(-1))
        d = bt.distortion(base, snap, 1024)
        assert 0 < d[0] < 0.1


def test_large_case_split_invariance(bt):
    """The 100,000-edge case (~5M samples, 2048^2 density): two grids (the
    density default, and 3 rows x 2 bins per radius) give the same bits after
    two iterations; endpoints stay fixed; the grid partitions every sample."""
    g = workloads.large_graph()
    base = workloads.large_options(50)
    other = bt.Options(spacing=0.0035, radius=(0.02, 0.005), sigma=(8.0, 2.0), density=2048,
                       schedule_len=50, cell_split=3, x_split=2)
    res = []
    for opts in (base, other):
        with bt.Gpu(opts) as gpu:
            n = gpu.upload(g)
            layout = gpu.distortion_layout()
            before = gpu.distortion_snapshot(layout)
            gpu.iterate(2)
            after = gpu.distortion_snapshot(layout)
            ends = np.concatenate([layout.paths[layout.edges, 0],
                                   layout.paths[layout.edges, 0] + layout.paths[layout.edges, 2]
                                   - 1])
            np.testing.assert_array_equal(after.positions[ends], before.positions[ends])
            cells, counts = gpu.grid()
            assert counts.sum() == n
            np.testing.assert_array_equal(np.bincount(cells, minlength=counts.size),
                                          counts.reshape(-1))
            res.append((gpu.points()[0], gpu.density(0)))
    assert n > 4_000_000
    np.testing.assert_array_equal(res[0][0].view(np.uint32), res[1][0].view(np.uint32))
    np.testing.assert_array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("case", ["cpu", "gpu"])
def test_endpoints_match_oracle(bt, orc, cpu_case, case):
    """Endpoint dedup (endpoints.wgsl): 2,100 / 9,800 edges between 235 /
    1,715 nodes, so most endpoints repeat. Exact index list, before and
    after iterations (endpoints are fixed samples)."""
    g = cpu_case if case == "cpu" else workloads.plain_gpu_graph()
    opts = (workloads.reference_cpu_options() if case == "cpu"
            else bt.Options(spacing=0.002, radius=0.01, sigma=2.0, density=1024))
    with bt.Gpu(opts) as gpu:
        gpu.upload(g)
        pts, _ = gpu.points()
        want = orc.endpoints(pts, gpu.paths())
        assert 0 < len(want) <= len(np.unique(g.points, axis=0))
        np.testing.assert_array_equal(gpu.endpoints(), want)
        gpu.iterate()
        gpu.iterate()
        got = gpu.endpoints()
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(gpu.points()[0][got].view(np.uint32),
                                      pts[want].view(np.uint32))
        count = ctypes.c_uint32()
        small = np.zeros(1, dtype=np.uint32)
        rc = api.lib().bt_read_endpoints(gpu._h, small.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                         1, ctypes.byref(count))
        assert rc == api.BT_E_CAPACITY and count.value == len(want) and small[0] == want[0]


def test_density_disabled_matches_enabled(bt, cpu_case):
    """color_by = edge id (no density buffers): the same positions, bit for
    bit, as with the density map on."""
    res = []
    for dens in (True, False):
        with bt.Gpu(bt.Options(map_size=201, neighbor_radius=10, color_by_density=dens)) as gpu:
            gpu.upload(cpu_case)
            gpu.iterate()
            gpu.iterate()
            res.append(gpu.points()[0])
            if not dens:
                with pytest.raises(bt.BtError) as e:
                    gpu.density(3)
                assert e.value.code == 5
    np.testing.assert_array_equal(res[0].view(np.uint32), res[1].view(np.uint32))


def test_rewind_mixed_density_batch(bt):
    """A batch mixing jobs with and without a density map: rewinding
    restores every job's samples (and density where it has one), so a
    second eager schedule and a graph replay give the first run's bits."""
    graphs = workloads.batched_graphs(count=4, lo=60, hi=400)
    opts = [bt.Options(spacing=0.004, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=256,
                       schedule_len=4, color_by_density=(i % 2 == 0)) for i in range(4)]
    with bt.Gpu(opts[0]) as gpu:
        gpu.upload_batch(graphs, opts)
        up = [gpu.points(j)[0] for j in range(4)]
        up_d = [gpu.density(3, j) for j in (0, 2)]
        gpu.iterate(4)
        first = [gpu.points(j)[0] for j in range(4)]
        first_d = [gpu.density(3, j) for j in (0, 2)]
        for graph in (False, True):
            gpu.rewind()
            for j in range(4):
                np.testing.assert_array_equal(gpu.points(j)[0].view(np.uint32),
                                              up[j].view(np.uint32))
            for k, j in enumerate((0, 2)):
                np.testing.assert_array_equal(gpu.density(3, j).view(np.uint32),
                                              up_d[k].view(np.uint32))
            gpu.iterate(4, graph=graph)
            for j in range(4):
                np.testing.assert_array_equal(gpu.points(j)[0].view(np.uint32),
                                              first[j].view(np.uint32))
            for k, j in enumerate((0, 2)):
                np.testing.assert_array_equal(gpu.density(3, j).view(np.uint32),
                                              first_d[k].view(np.uint32))


def test_error_paths(bt, cpu_case):
    with bt.Gpu(bt.Options(map_size=201)) as gpu:
        with pytest.raises(bt.BtError) as e:
            gpu.iterate()
        assert e.value.code == 2  # no uploaded graph
        gpu.upload(cpu_case)
        with pytest.raises(bt.BtError) as e:
            gpu.iterate()
        assert e.value.code == 5  # neighbor radius is not configured
    with bt.Gpu(bt.Options(map_size=1600, neighbor_radius=4, max_graph_memory_mib=8)) as gpu:
        with pytest.raises(bt.BtError) as e:
            gpu.upload(workloads.plain_gpu_graph())
        assert e.value.code == 3 and "sample buffer capacity" in e.value.msg


def test_split_positions_equal_records(bt):
    """bt_read_positions_job (positions and edge ids as two arrays) carries
    the bits of bt_read_points_job's 12-byte records, job by job."""
    from bundling_thingy_amd.api import lib, _f32p
    graphs = workloads.batched_graphs(count=3, lo=60, hi=400)
    opts = [workloads.batched_options(i, 4) for i in range(3)]
    rec = np.dtype([("x", "<f4"), ("y", "<f4"), ("id", "<u4")])
    with bt.Gpu(opts[0]) as gpu:
        gpu.upload_batch(graphs, opts)
        gpu.iterate(2)
        for j in range(3):
            r = gpu.points_raw(j).view(rec)
            xy, ids = gpu.points(j)
            np.testing.assert_array_equal(xy[:, 0].view(np.uint32), r["x"].view(np.uint32))
            np.testing.assert_array_equal(xy[:, 1].view(np.uint32), r["y"].view(np.uint32))
            np.testing.assert_array_equal(ids, r["id"])
            xy2 = np.empty_like(xy)
            assert lib().bt_read_positions_job(gpu._h, j, _f32p(xy2), None, len(xy2)) == 0
            np.testing.assert_array_equal(xy2.view(np.uint32), xy.view(np.uint32))
            assert lib().bt_read_positions_job(gpu._h, j, _f32p(xy2), None, len(xy2) - 1) == 1


@pytest.mark.parametrize("points,starts,msg", [
    ([[0.2, 0.3]], [0, 1], "no drawable edges"),
    ([[0, 0], [1, 1], [2, 0], [3, 1]], [0, 3, 2, 4], "not sorted"),
    ([[0, 0], [1, 1], [2, 0]], [0, 5], "exceed the points"),
    ([[0, 0], [np.nan, 1], [2, 0]], [0, 3], "non-finite"),
    ([[0.5, 0.5], [0.5, 0.5], [0.5, 0.5]], [0, 3], "invalid extent"),
])
def test_rejected_upload_drops_the_graph(bt, cpu_case, points, starts, msg):
    """gpu.rs:494: an upload drops the previous graph before anything can
    fail, so a rejected one leaves none (iterate: code 2). Malformed input is
    refused with the reference's messages (graph.rs:49-90); a valid upload
    afterwards runs as in a fresh context."""
    opts = workloads.reference_cpu_options()
    with bt.Gpu(opts) as fresh:
        fresh.upload(cpu_case)
        fresh.iterate()
        want = fresh.points()[0]
    with bt.Gpu(opts) as gpu:
        gpu.upload(cpu_case)
        gpu.iterate()
        with pytest.raises(bt.BtError) as e:
            gpu.upload(bt.Graph(np.array(points, np.float32), np.array(starts, np.uint32),
                                bbox=np.zeros(4, np.float32)))
        assert e.value.code == 1 and msg in e.value.msg, e.value.msg
        with pytest.raises(bt.BtError) as e:
            gpu.iterate()
        assert e.value.code == 2
        gpu.upload(cpu_case)
        gpu.iterate()
        np.testing.assert_array_equal(gpu.points()[0].view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("policy,tick,lane", [(0, 0, 0), (1, 0, 0), (1, 1, 0), (2, 0, 20000),
                                              (2, 1, 20000)])
def test_server_results_match_alone(bt, policy, tick, lane):
    """Co-scheduled jobs give the bits of each job alone, points and final
    density, under every policy (tick 1: the shortest-first policy advances
    one job per tick; policy 2 with jobs of at most 20,000 samples in the
    small-job lane: both lanes hold jobs; with tick 1 each lane advances one
    job per tick once no job waits for admission, all of them before)."""
    graphs = workloads.batched_graphs(count=4, lo=60, hi=400)
    opts = bt.Options(spacing=0.004, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=256,
                      schedule_len=6)
    iters = [6, 3, 5, 4]
    alone, alone_d = [], []
    for g, it in zip(graphs, iters):
        with bt.Gpu(opts) as gpu:
            gpu.upload(g)
            gpu.iterate(it)
            alone.append(gpu.points()[0])
            alone_d.append(gpu.density())
    srv = bt.Server(opts, memory_budget=1 << 30, max_inflight=3, policy=policy,
                    tick_samples=tick, lane_samples=lane)
    tickets = [srv.submit(g, opts, it) for g, it in zip(graphs, iters)]
    for _ in range(256):
        r, q = srv.step()
        if r == 0 and q == 0:
            break
    if policy == 2:
        small = [len(a) <= lane for a in alone]
        assert any(small) and not all(small), [len(a) for a in alone]
    for t, ref, ref_d in zip(tickets, alone, alone_d):
        res = srv.poll(t)
        assert res.state == 2 and res.iterations_done == res.iterations_total
        p, _ = srv.points(t, res.samples)
        np.testing.assert_array_equal(p.view(np.uint32), ref.view(np.uint32))
        # the server computes only the last iteration's density
        d = srv.density(t, ref_d.shape[1], ref_d.shape[0])
        assert d.max() == 1.0
        np.testing.assert_array_equal(d.view(np.uint32), ref_d.view(np.uint32))
    srv.close()


def test_server_isolates_a_failed_upload(bt):
    """A job whose resampled samples exceed its sample budget (2 MiB of
    graph memory: room for 15,932 samples against its 30,217) fails alone,
    state 3, when its admission batch's counts come back; the jobs admitted
    with it finish with the bits of each job alone."""
    graphs = workloads.batched_graphs(count=4, lo=60, hi=400)
    kw = dict(spacing=0.004, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=256, schedule_len=3)
    opts, tight = bt.Options(**kw), bt.Options(max_graph_memory_mib=2, **kw)
    alone = {}
    for k in (1, 2, 3):
        with bt.Gpu(opts) as gpu:
            gpu.upload(graphs[k])
            gpu.iterate(3)
            alone[k] = gpu.points()[0]
    with bt.Gpu(tight) as gpu:
        with pytest.raises(bt.BtError) as e:
            gpu.upload(graphs[0])
        assert e.value.code == 3 and "sample buffer capacity" in e.value.msg
    srv = bt.Server(opts, memory_budget=1 << 30, max_inflight=8, policy=0)
    tickets = [srv.submit(g, tight if k == 0 else opts, 3) for k, g in enumerate(graphs)]
    for _ in range(256):
        r, q = srv.step()
        if r == 0 and q == 0:
            break
    assert srv.poll(tickets[0]).state == 3
    for k in (1, 2, 3):
        res = srv.poll(tickets[k])
        assert res.state == 2
        p, _ = srv.points(tickets[k], res.samples)
        np.testing.assert_array_equal(p.view(np.uint32), alone[k].view(np.uint32))
    srv.close()


def _edge_case_graph(bt):
    """Polylines that stress the path: two identical co-located edges, a
    closed loop (first = last point), a zero-length edge (one repeated point:
    the closed-path special case), an edge shorter than one texel, a polyline
    with many short segments, and a lone long edge."""
    polys = [
        [(0.1, 0.5), (0.9, 0.5)],
        [(0.1, 0.5), (0.9, 0.5)],
        [(0.3, 0.3), (0.6, 0.3), (0.6, 0.6), (0.3, 0.3)],
        [(0.7, 0.8), (0.7, 0.8)],
        [(0.2, 0.2), (0.2005, 0.2001)],
        [(0.1 + 0.01 * k, 0.7 + 0.02 * ((-1) ** k)) for k in range(60)],
        [(0.05, 0.05), (0.95, 0.95)],
    ]
    pts = np.array([p for poly in polys for p in poly], np.float32)
    starts = np.cumsum([0] + [len(p) for p in polys]).astype(np.uint32)
    return bt.Graph(pts, starts)


@pytest.mark.parametrize("radius", [10, 400])
def test_edge_case_graph_bitexact(bt, orc, radius):
    """Reference mode S = 201: upload, three iterations and the image
    bit-identical to the oracle; radius 400 texels puts the whole map in one
    grid cell."""
    g = _edge_case_graph(bt)
    opts = bt.Options(map_size=201, neighbor_radius=radius)
    pl = _oracle_pipeline(orc, g, opts)
    with bt.Gpu(opts) as gpu:
        n = gpu.upload(g)
        assert n == len(pl.points)
        np.testing.assert_array_equal(gpu.paths()[:, 2], pl.counts)
        _assert_points(gpu.points()[0], pl.points, "upload")
        for it in range(3):
            gpu.iterate()
            _assert_points(gpu.points()[0], pl.iterate(opts.iter_params(it)), f"iteration {it + 1}")
            _assert_density(gpu, pl, f"iteration {it + 1}")
        pal, ep = bt.api.palette("viridis")
        paths = np.stack([pl.offsets.astype(np.uint32), np.zeros_like(pl.counts), pl.counts], 1)
        want = orc.render(pl.points, pl.edge_id, paths, 201, 201, pal, ep,
                          density=pl.normalized, msaa=4, dots=True)
        np.testing.assert_array_equal(gpu.render("viridis", 0, 4, True), want)
        np.testing.assert_array_equal(gpu.endpoints(), orc.endpoints(pl.points, paths))


def _random_polylines(bt, seed, n_paths=300):
    """Seeded random polylines: 2-12 control points each, steps from a
    hundredth of a texel to a third of the map, a fifth of them with a
    repeated point, some closed (last point = first)."""
    rng = np.random.default_rng(seed)
    polys = []
    for _ in range(n_paths):
        k = int(rng.integers(2, 13))
        p = [rng.uniform(0.0, 1.0, 2)]
        for _ in range(k - 1):
            step = 10.0 ** rng.uniform(-4.5, -0.5)
            p.append(p[-1] + rng.normal(0.0, 1.0, 2) * step)
        if rng.random() < 0.2:
            j = int(rng.integers(0, k))
            p.insert(j, p[j].copy())
        if rng.random() < 0.1:
            p.append(p[0].copy())
        polys.append(np.array(p, np.float32))
    pts = np.concatenate(polys).astype(np.float32)
    starts = np.cumsum([0] + [len(q) for q in polys]).astype(np.uint32)
    return bt.Graph(pts, starts)


@pytest.mark.parametrize("seed,mode", [(1, "reference"), (2, "reference"), (3, "generalized")])
def test_random_polylines_bitexact(bt, orc, seed, mode):
    """Seeded random polylines (joints carrying the resampling remainder,
    repeated points, closed paths, sub-texel segments): samples, path
    counts, two iterations and the density planes bit-identical to the
    oracle, in reference mode (S = 301, radius 12 texels) and with the
    generalized parameters (spacing 0.004, radius 0.03 -> 0.02, 256^2)."""
    g = _random_polylines(bt, seed)
    if mode == "reference":
        opts = bt.Options(map_size=301, neighbor_radius=12)
    else:
        opts = bt.Options(spacing=0.004, radius=(0.03, 0.02), sigma=(3.0, 2.0), density=256,
                          schedule_len=2)
    pl = _oracle_pipeline(orc, g, opts)
    with bt.Gpu(opts) as gpu:
        n = gpu.upload(g)
        assert n == len(pl.points)
        np.testing.assert_array_equal(gpu.paths()[:, 2], pl.counts)
        _assert_points(gpu.points()[0], pl.points, f"seed {seed}: upload")
        for it in range(2):
            gpu.iterate()
            _assert_points(gpu.points()[0], pl.iterate(opts.iter_params(it)),
                           f"seed {seed}: iteration {it + 1}")
            _assert_density(gpu, pl, f"seed {seed}: iteration {it + 1}")


def test_second_upload_replaces_the_graph(bt, cpu_case):
    """gpu.rs:494-497: a new upload drops the previous graph (and its
    iteration count); the results equal those of a fresh context."""
    other = workloads.gmm_graph(7, 80, 3, 500)
    opts = workloads.reference_cpu_options()
    with bt.Gpu(opts) as fresh:
        n_fresh = fresh.upload(other)
        fresh.iterate()
        fresh.iterate()
        want = fresh.points()[0]
        want_d = fresh.density(3)
    with bt.Gpu(opts) as gpu:
        gpu.upload(cpu_case)
        gpu.iterate()
        n = gpu.upload(other)
        assert n == n_fresh and gpu.job_samples(0) == (n, 0)
        assert len(gpu.paths()) == other.edge_count
        gpu.iterate()
        gpu.iterate()
        np.testing.assert_array_equal(gpu.points()[0].view(np.uint32), want.view(np.uint32))
        np.testing.assert_array_equal(gpu.density(3).view(np.uint32), want_d.view(np.uint32))


def test_server_rejects_malformed_starts(bt):
    """A submit validates the polylines before estimating their footprint."""
    srv = bt.Server(bt.Options(), memory_budget=1 << 30, max_inflight=4)
    try:
        pts = np.array([[0, 0], [1, 1], [2, 0]], np.float32)
        for starts in ([0, 3, 2], [0, 2, 9]):
            with pytest.raises(bt.BtError) as e:
                srv.submit(bt.Graph(pts, np.array(starts, np.uint32), [0, 0, 2, 1]),
                           bt.Options(map_size=201, neighbor_radius=10), 2)
            assert e.value.code == 1
    finally:
        srv.close()


@pytest.mark.parametrize("case", ["reference", "gpu"])
def test_density_binary16_readback(bt, orc, case):
    """Density kind 4: the normalized R channel as the reference's rgba16f
    texture holds it (density.rs:96, normalization.wgsl:16), bit-exact
    against the oracle's binary16 rounding (or_half_round, ties to even,
    subnormals kept) of the f32 plane, on the reference-mode map and on the
    plain GPU case's 1024^2 map after upload and after one iteration."""
    if case == "reference":
        g, opts = workloads.plain_cpu_graph(), workloads.reference_cpu_options()
    else:
        g = workloads.plain_gpu_graph()
        opts = bt.Options(spacing=0.002, radius=0.01, sigma=2.0, density=1024)
    with bt.Gpu(opts) as gpu:
        gpu.upload(g)
        for step in range(2):
            f32 = gpu.density(3)
            half = gpu.density(bt.api.DENSITY_HALF)
            assert half.dtype == np.uint16 and half.shape == f32.shape
            vals, inv = np.unique(f32.reshape(-1), return_inverse=True)
            want = orc.half_round(vals)[inv].astype(np.float16).view(np.uint16)
            np.testing.assert_array_equal(half.reshape(-1), want, err_msg=f"step {step}")
            np.testing.assert_array_equal(half, f32.astype(np.float16).view(np.uint16))
            assert np.any((half > 0) & (half < 0x0400)) or case == "reference"
            gpu.iterate()


def _run_shapes(bt, g, opts, shapes, iters=3, graph=False):
    with bt.Gpu(opts) as gpu:
        gpu.upload(g)
        if shapes is not None:
            gpu.set_shapes(*shapes)
        gpu.iterate(iters, graph=graph)
        info = gpu.info()
        return (gpu.points()[0], gpu.density(0), gpu.density(2), gpu.density(3),
                gpu.density(1), info.split_tasks)


@pytest.mark.parametrize("shape", [(nm, 0, 0) for nm in range(6)] +
                         [(0, row, 0) for row in range(1, 4)] +
                         [(0, 0, col) for col in (1, 2, 3, 4)])
def test_kernel_shapes_do_not_change_results(bt, cpu_case, shape):
    """Every neighbor-mean shape (lanes per group x waves per workgroup) and
    blur tile gives the default shapes' bits: positions, counts, blurred and
    normalized density and the row-blur plane, on the plain CPU case (plain
    and grouped streams) and on square cells as wide as the radius (tasks
    split across waves)."""
    cases = [workloads.plain_cpu_options(),
             bt.Options(spacing=0.005, radius=0.1, sigma=6.0, density=512, cell_split=1,
                        x_split=1)]
    for k, opts in enumerate(cases):
        want = _run_shapes(bt, cpu_case, opts, None)
        got = _run_shapes(bt, cpu_case, opts, shape)
        for a, b in zip(got[:5], want[:5]):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32),
                                          err_msg=f"shape {shape} case {k}")
        if k == 1:
            assert got[5] > 0  # the heavy-task kernel ran under this shape


@pytest.mark.parametrize("sigma", [10.5, 12.0])
def test_fused_blur_edges(bt, cpu_case, sigma):
    """The fused row + column blur at its widest blur (sigma 10.5: 32 taps,
    the halo of a 64 x 16 tile twice its height) and past it (sigma 12: 36
    taps, where the two-kernel blur runs instead) on a 200 x 200 map (tiles
    cut by both image edges) gives the default shapes' bits."""
    opts = bt.Options(spacing=0.005, radius=0.05, sigma=sigma, density=200)
    want = _run_shapes(bt, cpu_case, opts, None, iters=2)
    for col in (3, 4):
        got = _run_shapes(bt, cpu_case, opts, (0, 0, col), iters=2)
        for a, b in zip(got[:5], want[:5]):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32),
                                          err_msg=f"col shape {col}")


def test_autotune_shapes_keeps_results(bt, cpu_case):
    """The shape autotuner picks one shape per stage per iteration; the
    schedule replayed as a hipGraph under its choices gives the default
    bits."""
    opts = workloads.plain_cpu_options()
    want = _run_shapes(bt, cpu_case, opts, None, iters=4)
    with bt.Gpu(opts) as gpu:
        gpu.upload(cpu_case)
        shapes, ms = gpu.autotune_shapes(4)
        assert shapes.shape == (4, 3) and np.all(shapes[:, 0] < 6) and np.all(ms[:, 0] > 0)
        assert gpu.job_samples(0)[1] == 0  # rewound
        for i in range(4):
            assert gpu.shapes(i) == tuple(int(v) for v in shapes[i])
        gpu.iterate(4, graph=True)
        np.testing.assert_array_equal(gpu.points()[0].view(np.uint32), want[0].view(np.uint32))
        np.testing.assert_array_equal(gpu.density(3).view(np.uint32), want[3].view(np.uint32))
