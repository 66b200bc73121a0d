"""Seeded synthetic graphs for the benchmark workloads.

The reference's datasets (data/*.trl) are Git-LFS stubs, so every workload is
generated. Choices the workload descriptions leave open are fixed here:
 - Gaussian-mixture nodes: component means ~ U[0.1, 0.9]^2, isotropic
   sigma_k ~ U[0.02, 0.08], equal weights, rejection outside [0, 1]^2;
 - edges between node pairs drawn with probability proportional to
   exp(-dist / 0.3) by rejection sampling, without repeats or self loops;
 - every edge is a straight 2-point polyline (a TRL line "id x0 y0 x1 y1").
Seeds: plain CPU case 1, plain GPU case 2, large case 3, batched 4 + index,
served 5.
"""
import numpy as np

from .api import Graph, Options


def gmm_nodes(rng, n, k):
    means = rng.uniform(0.1, 0.9, (k, 2))
    sig = rng.uniform(0.02, 0.08, k)
    out = np.empty((0, 2))
    while len(out) < n:
        m = 2 * (n - len(out)) + 16
        comp = rng.integers(k, size=m)
        p = means[comp] + rng.normal(size=(m, 2)) * sig[comp, None]
        p = p[np.all((p >= 0.0) & (p <= 1.0), axis=1)]
        out = np.concatenate([out, p])
    return out[:n]


def sample_pairs(rng, nodes, n_edges, scale=0.3):
    """Distinct unordered node pairs, each drawn with probability proportional
    to exp(-dist / scale) (rejection sampling), in draw order."""
    n = len(nodes)
    keys = np.empty(0, dtype=np.int64)
    while True:
        m = 4 * max(n_edges - len(keys), 16) + 64
        i = rng.integers(n, size=m)
        j = rng.integers(n, size=m)
        d = np.linalg.norm(nodes[i] - nodes[j], axis=1)
        acc = (rng.random(m) < np.exp(-d / scale)) & (i != j)
        a, b = np.minimum(i[acc], j[acc]), np.maximum(i[acc], j[acc])
        keys = np.concatenate([keys, a * n + b])
        _, first = np.unique(keys, return_index=True)
        keys = keys[np.sort(first)]
        if len(keys) >= n_edges:
            keys = keys[:n_edges]
            return np.stack([keys // n, keys % n], 1)


def straight_graph(nodes, pairs):
    pts = np.empty((2 * len(pairs), 2), dtype=np.float32)
    pts[0::2] = nodes[pairs[:, 0]]
    pts[1::2] = nodes[pairs[:, 1]]
    starts = np.arange(0, 2 * len(pairs) + 1, 2, dtype=np.uint32)
    return Graph(pts, starts)


def to_trl(graph):
    """TRL text of a graph: one line per path, "<id> x y x y ..."."""
    lines = []
    for e in range(len(graph.starts) - 1):
        seg = graph.points[graph.starts[e]:graph.starts[e + 1]]
        lines.append(str(e) + " " + " ".join(f"{x!r} {y!r}" for x, y in seg.tolist()))
    return "\n".join(lines) + "\n"


def gmm_graph(seed, n_nodes, n_comp, n_edges):
    rng = np.random.default_rng(seed)
    nodes = gmm_nodes(rng, n_nodes, n_comp)
    return straight_graph(nodes, sample_pairs(rng, nodes, n_edges))


def plain_cpu_graph():
    """235 nodes, 6-component mixture, 2,100 edges (seed 1)."""
    return gmm_graph(1, 235, 6, 2100)


def plain_gpu_graph():
    """1,715 nodes, 20-component mixture, 9,800 edges (seed 2)."""
    return gmm_graph(2, 1715, 20, 9800)


def large_graph():
    """20,000 nodes (15,000 uniform + 50 clusters of 100, sigma 0.01) and
    100,000 edges with lengths ~ LogNormal(-2, 0.7): each edge joins a random
    node to the node nearest a point at the drawn length (seed 3)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(3)
    uni = rng.uniform(0.0, 1.0, (15000, 2))
    centers = rng.uniform(0.1, 0.9, (50, 2))
    cl = (centers[:, None, :] + rng.normal(size=(50, 100, 2)) * 0.01).reshape(-1, 2)
    nodes = np.clip(np.concatenate([uni, cl]), 0.0, 1.0)
    tree = cKDTree(nodes)
    n = len(nodes)
    keys = np.empty(0, dtype=np.int64)
    while len(keys) < 100000:
        m = 120000
        a = rng.integers(n, size=m)
        length = rng.lognormal(-2.0, 0.7, m)
        th = rng.uniform(0, 2 * np.pi, m)
        tgt = nodes[a] + length[:, None] * np.stack([np.cos(th), np.sin(th)], 1)
        _, b = tree.query(tgt)
        ok = a != b
        k = np.minimum(a[ok], b[ok]) * n + np.maximum(a[ok], b[ok])
        keys = np.concatenate([keys, k])
        _, first = np.unique(keys, return_index=True)
        keys = keys[np.sort(first)]
    keys = keys[:100000]
    return straight_graph(nodes, np.stack([keys // n, keys % n], 1))


def batched_graphs(count=256, lo=50, hi=20000):
    """Graphs with edge counts log-uniform in [lo, hi], nodes = edges / 6."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(4 + i)
        e = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        n = max(8, e // 6)
        k = int(rng.integers(4, 21))
        nodes = gmm_nodes(rng, n, k)
        e = min(e, n * (n - 1) // 2)
        out.append(straight_graph(nodes, sample_pairs(rng, nodes, e)))
    return out


# --------------------------------------------------------------- parameters

def reference_cpu_options():
    """Plain CPU case in reference mode: S = 201 (spacing 0.005), r = 10 texels
    (0.05 canvas units), sigma 6, density 201 x 201."""
    return Options(map_size=201, neighbor_radius=10)


def plain_cpu_options():
    """Plain CPU case, generalized: spacing 0.005, radius 0.05, 512^2, sigma 4."""
    return Options(spacing=0.005, radius=0.05, sigma=4.0, density=512, schedule_len=10)


def plain_gpu_options(iterations=30, cell_split=0):
    """Plain GPU case: spacing 0.002, radius 0.06 -> 0.01 geometric over 30
    iterations, 1024^2 density, sigma 8 -> 2."""
    return Options(spacing=0.002, radius=(0.06, 0.01), sigma=(8.0, 2.0), density=1024,
                   schedule_len=iterations, cell_split=cell_split)


def batched_options(graph_index, iterations=30):
    """Per-graph parameters of the batched case (seeded by the graph index):
    spacing 0.01, radius 0.05 -> 0.02, sigma 4 -> 2, density 256^2 to 512^2."""
    rng = np.random.default_rng(1000 + graph_index)
    d = int(rng.choice([256, 384, 512]))
    return Options(spacing=0.01, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=d,
                   schedule_len=iterations)


def served_jobs(count, rate_hz=200.0, lo=100, hi=50000, seed=5):
    """Poisson arrival stream: (arrival seconds, graph, options, iterations).
    Edge counts log-uniform in [lo, hi], 20-50 iterations, density 512^2 to
    2048^2; spacing 0.01, radius 0.05 -> 0.02, sigma 4 -> 2."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate_hz)
        e = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        n = max(8, e // 6)
        k = int(rng.integers(4, 21))
        nodes = gmm_nodes(rng, n, k)
        e = min(e, n * (n - 1) // 2)
        g = straight_graph(nodes, sample_pairs(rng, nodes, e))
        iters = int(rng.integers(20, 51))
        d = int(rng.choice([512, 1024, 2048]))
        o = Options(spacing=0.01, radius=(0.05, 0.02), sigma=(4.0, 2.0), density=d,
                    schedule_len=iters)
        out.append((t, g, o, iters))
    return out


def large_options(iterations=50):
    """Large case: spacing 0.0035, radius 0.02 -> 0.005, 2048^2 density,
    sigma 8 -> 2 (blur half-width up to 24 px)."""
    return Options(spacing=0.0035, radius=(0.02, 0.005), sigma=(8.0, 2.0), density=2048,
                   schedule_len=iterations)
