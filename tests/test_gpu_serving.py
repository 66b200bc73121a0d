"""GPU tests of the density-on-read option and of the server's admission:
density computed when read equals the density every iteration computes;
server jobs under either density policy equal the job alone; admission
keeps the device bytes the admitted jobs hold within the memory budget; the
shortest-first policy admits a job queued for over a second even while
smaller jobs keep arriving."""
import numpy as np
import pytest

from bundling_thingy_amd import api, workloads

pytestmark = pytest.mark.gpu


def _with_flags(bt, opts, flags):
    o = bt.Options()
    o.c = type(opts.c).from_buffer_copy(opts.c)
    o.c.flags |= flags
    return o


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("case", ["reference", "generalized"])
def test_density_on_read_equals_every_iteration(bt, case):
    """Iterations that skip the density stages, with the density computed
    from the current positions when read: counts, blurred, normalized and
    binary16 planes and the density-coloured image equal those of a context
    that computes the density every iteration, after 0, 1 and 3 iterations,
    after a rewind, and after a hipGraph replay."""
    g = workloads.plain_cpu_graph()
    base = (workloads.reference_cpu_options() if case == "reference"
            else bt.Options(spacing=0.005, radius=(0.05, 0.03), sigma=(4.0, 2.0), density=512,
                            schedule_len=4))
    lazy = _with_flags(bt, base, api.FLAG_DENSITY_ON_READ)
    with bt.Gpu(base) as a, bt.Gpu(lazy) as b:
        a.upload(g)
        b.upload(g)

        def same(label):
            np.testing.assert_array_equal(_bits(a.points()[0]), _bits(b.points()[0]), label)
            for kind in (0, 1, 2, 3, api.DENSITY_HALF):
                np.testing.assert_array_equal(a.density(kind).view(np.uint8),
                                              b.density(kind).view(np.uint8), f"{label} {kind}")
            np.testing.assert_array_equal(a.render(), b.render(), label)

        same("upload")
        a.iterate()
        b.iterate()
        same("iteration 1")
        a.iterate(2)
        b.iterate(2)
        same("iteration 3")
        b.rewind()
        a.rewind()
        same("rewound")
        a.iterate(3, graph=True)
        b.iterate(3, graph=True)
        same("graph replay")


def test_server_density_policies_equal_alone(bt):
    """Jobs computing the density every iteration and jobs computing only
    the last one (BT_FLAG_DENSITY_ON_READ) share ticks; each returns the
    bits of the job alone, points and final density."""
    graphs = workloads.batched_graphs(count=6, lo=60, hi=800)
    kw = dict(spacing=0.004, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=256, schedule_len=6)
    opts = [bt.Options(density_on_read=(k % 2 == 1), **kw) for k in range(6)]
    iters = [6, 3, 5, 4, 6, 2]
    alone = []
    for g, o, it in zip(graphs, opts, iters):
        with bt.Gpu(o) as gpu:
            gpu.upload(g)
            gpu.iterate(it)
            alone.append((gpu.points()[0], gpu.density()))
    srv = bt.Server(opts[0], memory_budget=1 << 30, max_inflight=4, policy=0)
    try:
        tickets = [srv.submit(g, o, it) for g, o, it in zip(graphs, opts, iters)]
        for _ in range(512):
            r, q = srv.step()
            if r == 0 and q == 0:
                break
        for t, (p_ref, d_ref) in zip(tickets, alone):
            res = srv.poll(t)
            assert res.state == 2
            np.testing.assert_array_equal(_bits(srv.points(t, res.samples)[0]), _bits(p_ref))
            np.testing.assert_array_equal(_bits(srv.density(t, 256, 256)), _bits(d_ref))
    finally:
        srv.close()


def test_batched_graph_mixed_density_policies_equal_alone(bt):
    """One batch whose jobs compute the density every iteration or only when
    read, iterated as a captured schedule (the density stages and the next
    iteration's density clear on a second stream, the task list on a third)
    and replayed after a rewind: every job's points and density equal the
    job alone iterated eagerly, after each run."""
    graphs = workloads.batched_graphs(count=6, lo=60, hi=1500)
    kw = dict(spacing=0.004, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=256, schedule_len=5)
    opts = [bt.Options(density_on_read=(k % 2 == 1), **kw) for k in range(6)]
    alone = []
    for g, o in zip(graphs, opts):
        with bt.Gpu(o) as gpu:
            gpu.upload(g)
            gpu.iterate(5)
            alone.append((gpu.points()[0], gpu.density(3)))
    with bt.Gpu(opts[0]) as gpu:
        gpu.upload_batch(graphs, opts)
        for label in ("captured", "replayed"):
            gpu.iterate(5, graph=True)
            for j, (p_ref, d_ref) in enumerate(alone):
                np.testing.assert_array_equal(_bits(gpu.points(j)[0]), _bits(p_ref),
                                              f"{label} job {j} points")
                np.testing.assert_array_equal(_bits(gpu.density(3, j)), _bits(d_ref),
                                              f"{label} job {j} density")
            gpu.rewind()


def test_server_memory_budget_holds(bt):
    """A budget of under half the ten jobs' footprints: at every tick the bytes the
    admitted jobs' buffers hold stay within the footprint charged for them,
    and that within the budget; several jobs still run together, and every
    result equals the job alone."""
    graphs = workloads.batched_graphs(count=10, lo=300, hi=3000)
    opts = workloads.batched_options(0, 8)
    alone = []
    for g in graphs:
        with bt.Gpu(opts) as gpu:
            gpu.upload(g)
            gpu.iterate(8)
            alone.append(gpu.points()[0])
    # footprint of the three largest jobs alone, measured on a roomy server
    probe = bt.Server(opts, memory_budget=64 << 30, max_inflight=64)
    try:
        for g in graphs:
            probe.submit(g, opts, 8)
        foot = probe.info()  # nothing admitted yet
        assert foot.footprint_bytes == 0 and foot.queued == 10
        probe.step()
        big = probe.info().footprint_bytes
    finally:
        probe.close()
    budget = int(big * 0.45)
    srv = bt.Server(opts, memory_budget=budget, max_inflight=64, policy=0)
    try:
        tickets = [srv.submit(g, opts, 8) for g in graphs]
        most = 0
        for _ in range(4096):
            r, q = srv.step()
            i = srv.info()
            assert i.held_bytes <= i.footprint_bytes <= budget, (i.held_bytes, i.footprint_bytes)
            most = max(most, i.running)
            if r == 0 and q == 0:
                break
        assert srv.info().held_peak_bytes <= budget
        assert 2 <= most < 10
        for t, ref in zip(tickets, alone):
            res = srv.poll(t)
            assert res.state == 2
            np.testing.assert_array_equal(_bits(srv.points(t, res.samples)[0]), _bits(ref))
    finally:
        srv.close()


def test_shortest_first_admits_a_starving_queued_job(bt):
    """One in-flight slot, a large job queued behind small jobs that keep
    arriving: shortest-first would always admit the small ones, but a job
    queued for over a second goes first, so the large job is admitted
    within about a second of its submission."""
    small = workloads.batched_graphs(count=4, lo=40, hi=80)
    large = workloads.batched_graphs(count=1, lo=3000, hi=3001)[0]
    opts = bt.Options(spacing=0.004, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=256,
                      schedule_len=6)
    srv = bt.Server(opts, memory_budget=8 << 30, max_inflight=1, policy=1, tick_samples=1)
    try:
        srv.submit(small[0], opts, 3)
        srv.step()
        t_large = srv.submit(large, opts, 6)
        k = 0
        for _ in range(200000):
            if k % 2 == 0:
                srv.submit(small[k % 4], opts, 3)
            k += 1
            srv.step()
            if srv.poll(t_large).state >= 1:
                break
            assert srv.clock_ms() - srv.poll(t_large).submit_ms < 3000, "large job starved"
        res = srv.poll(t_large)
        assert res.state >= 1
        wait = res.admit_ms - res.submit_ms
        assert 1000.0 <= wait < 1600.0, wait
        for _ in range(200000):
            r, q = srv.step()
            if r == 0 and q == 0:
                break
        assert srv.poll(t_large).state == 2
    finally:
        srv.close()
