"""Host-side logic of the product library (no GPU): the C ABI loads and
exports every declared symbol; TRL parsing, normalization, the capacity
model, blur weights, iteration parameters and distortion agree with the
oracle and the reference's rules."""
import math
import os
import re

import numpy as np
import pytest

from tests.test_oracle import TRL_CASES, TRL_ERRORS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(bt):
    import ctypes
    hdr = open(os.path.join(ROOT, "include", "bundling.h")).read()
    names = set(re.findall(r"^\s*(?:[a-z_0-9]+\s*\*?\s+)+\**(bt_[a-z_0-9]+)\(", hdr, re.M))
    assert len(names) >= 30
    L = ctypes.CDLL(os.path.join(ROOT, "bundling_thingy_amd", "libbundling.so"))
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    from bundling_thingy_amd.api import SIGNATURES
    assert names == set(SIGNATURES), names ^ set(SIGNATURES)


@pytest.mark.parametrize("text,pts,starts", TRL_CASES)
def test_parse_matches_oracle(bt, orc, text, pts, starts):
    g = bt.Graph.parse(text)
    p, s, bbox = orc.parse_trl(text)
    np.testing.assert_array_equal(g.points, p)
    np.testing.assert_array_equal(g.starts, s)
    np.testing.assert_array_equal(g.bbox, bbox)


@pytest.mark.parametrize("text", TRL_ERRORS)
def test_parse_errors(bt, text):
    with pytest.raises(bt.BtError):
        bt.Graph.parse(text)


def test_parse_generated_trl_roundtrip(bt):
    from bundling_thingy_amd import workloads
    g = workloads.plain_cpu_graph()
    h = bt.Graph.parse(workloads.to_trl(g))
    np.testing.assert_array_equal(g.points, h.points)
    np.testing.assert_array_equal(g.starts, h.starts)


def test_parse_large_trl_matches_oracle(bt, orc, tmp_path):
    """Loader at scale: the 100,000-edge graph (8 MB of TRL) through
    Graph.load, equal to the generator's arrays and to the oracle's parse."""
    from bundling_thingy_amd import workloads
    g = workloads.large_graph()
    path = tmp_path / "large.trl"
    path.write_text(workloads.to_trl(g))
    h = bt.Graph.load(path)
    np.testing.assert_array_equal(g.points, h.points)
    np.testing.assert_array_equal(g.starts, h.starts)
    p, st, bbox = orc.parse_trl(path.read_text())
    np.testing.assert_array_equal(p, h.points)
    np.testing.assert_array_equal(st, h.starts)
    np.testing.assert_array_equal(bbox, h.bbox)


def test_normalization_matches_oracle(bt, orc):
    rng = np.random.default_rng(2)
    for _ in range(20):
        bb = np.sort(rng.uniform(-5, 5, (2, 2)), axis=0).T.reshape(-1)[[0, 2, 1, 3]]
        bb = bb.astype(np.float32)
        n = bt.Graph(np.zeros((2, 2)), [0, 2], bb).normalization()
        o = orc.normalization(bb)
        assert (n.minimum, n.scale, n.translation) == ((o.min_x, o.min_y), o.scale, (o.tr_x, o.tr_y))


def test_blur_weights_match_oracle(bt, orc):
    for s in [2.0, 3.3, 4.0, 6.0, 8.0, 21.0]:
        np.testing.assert_array_equal(bt.blur_weights(s), orc.blur_weights(s))


def test_reference_mode_grid_dims(bt):
    # neighbor_mean.rs:27: grid dims ceil(S / r) per axis
    for S, r in [(1600, 10), (201, 10), (1600, 7), (512, 512), (100, 1)]:
        p = bt.Options(map_size=S, neighbor_radius=r, cell_split=1, x_split=1).iter_params(0)
        assert (p.gx, p.gy) == (math.ceil(S / r),) * 2
        assert p.res_x == S - 1 and p.radius == r and p.cell == r and p.cell_w == r
        assert p.sigma == 6.0 and p.taps == 18 and p.density_w == S


@pytest.mark.parametrize("S,r", [(201, 10), (1600, 10), (1025, 30)])
def test_generalized_reduces_to_reference_mode(bt, S, r):
    ref = bt.Options(map_size=S, neighbor_radius=r).iter_params(0)
    gen = bt.Options(spacing=1.0 / (S - 1), radius=r / (S - 1), sigma=6.0, density=S,
                     schedule_len=1).iter_params(0)
    for f in ["res_x", "res_y", "radius", "cell", "gx", "gy", "sigma", "taps", "density_w",
              "density_h", "cell_w", "task_bins"]:
        assert getattr(ref, f) == getattr(gen, f), f


@pytest.mark.parametrize("split,xs", [(1, 1), (1, 8), (4, 8), (16, 8), (6, 3), (2, 64)])
def test_x_split_bins(bt, split, xs):
    """Rows keep the cell height; bins along x are x_split times narrower; task
    blocks span one radius."""
    o = bt.Options(spacing=0.002, radius=0.06, sigma=8.0, density=1024, cell_split=split,
                   x_split=xs)
    p = o.iter_params(0)
    cell = np.float32(30.0) if split == 1 else np.float32(np.float32(30.0) / np.float32(split))
    w = cell if xs == 1 else np.float32(cell / np.float32(xs))
    assert p.cell == cell and p.cell_w == w
    assert p.gy == math.ceil(501.0 / float(cell)) and p.gx == math.ceil(501.0 / float(w))
    assert p.task_bins == xs * split


def test_row_split_capped_by_grid_size(bt):
    # 28 rows per radius of 1.43 texels on a 286-texel map would need 31M
    # square cells: the split drops until square cells fit 2^22
    o = bt.Options(spacing=0.0035, radius=0.005, sigma=2.0, density=2048, cell_split=28)
    p = o.iter_params(0)
    assert p.gx * p.gy <= 1 << 22 and p.cell == p.cell_w
    k = round(p.radius / p.cell)
    c = np.float32(p.radius / np.float32(k + 1))
    assert math.ceil((p.res_x + 1) / c) * math.ceil((p.res_y + 1) / c) > 1 << 22


def test_x_split_capped_by_grid_size(bt):
    # 2^22 bins at most: x_split halves until the grid fits
    o = bt.Options(spacing=0.001, radius=0.01, sigma=2.0, density=1024, cell_split=8,
                   x_split=64)
    p = o.iter_params(0)  # square grid 801 x 801; 64 -> 32 -> 16 -> 8 -> 4
    assert p.gy == 801 and round(p.cell / p.cell_w) == 4
    assert p.gx * p.gy <= 1 << 22 < p.gx * 2 * p.gy


def test_schedule_is_geometric(bt):
    o = bt.Options(spacing=0.002, radius=(0.06, 0.01), sigma=(8.0, 2.0), density=1024,
                   schedule_len=30)
    rad = [o.iter_params(i).radius for i in range(32)]
    np.testing.assert_allclose(rad[0], 30.0, rtol=1e-6)
    np.testing.assert_allclose(rad[29], 5.0, rtol=1e-6)
    assert rad[30] == rad[29] == rad[31]
    ratios = np.array(rad[1:30]) / np.array(rad[:29])
    np.testing.assert_allclose(ratios, (1 / 6) ** (1 / 29), rtol=1e-6)
    assert o.iter_params(0).taps == 24 and o.iter_params(29).taps == 6


def test_capacity_model(bt):
    o = bt.Options(neighbor_radius=10, max_graph_memory_mib=64)
    cap, res = bt.capacity(o, 1000, 501)
    assert res <= 64 << 20 and (64 << 20) - res < 48  # sample_bytes = 16 + 32
    cap2, _ = bt.capacity(bt.Options(max_graph_memory_mib=64), 1000, 501)
    assert cap2 > cap  # 16-byte samples without a radius
    with pytest.raises(bt.BtError) as e:
        bt.capacity(bt.Options(neighbor_radius=1, max_graph_memory_mib=1), 10, 3)
    assert e.value.code == 3
    with pytest.raises(bt.BtError):
        bt.capacity(bt.Options(neighbor_radius=100, max_graph_memory_mib=1), 200000, 3)


# Allocation::new (resampling.rs:321-397) term by term, for a graph of
# 19,600 control points and 9,801 starts (9,800 paths):
#   points 8 B each, starts 4 B each, Path 12 B, Edge 16 B and endpoint 8 B
#   per path, Parameters 32 B, map uniform 16 B (shader.rs:4), state 48 B,
#   statistics 16 B (resampling.rs:20-21), the neighbor grid
#   (neighbor_mean.rs:176-183: cells 8 C + 8, summaries 24 C, 52 B of
#   uniform and indirect dispatches) and the distortion staging 4,194,300 B
#   (resampling.rs:16); per sample 12 + 4 B, plus 32 B with a radius
#   (neighbor_mean.rs:15); capacity = min(available / sample_bytes,
#   limit / 12).
_PER_GRAPH = 8 * 19600 + 4 * 9801 + (12 + 16 + 8) * 9800 + 32 + 16 + 48 + 16  # 548,916
_GRID_1600_R4 = (8 * 400 * 400 + 8) + 24 * 400 * 400 + 52                    # 5,120,060
_MIB = 1 << 20


@pytest.mark.parametrize("kw,capacity,reserved", [
    # radius 4 on the 1600 map: grid of 400 x 400 cells, 48-byte samples
    (dict(neighbor_radius=4), 22251517, 548916 + 5120060 + 22251517 * 48),
    # the same with distortion staging
    (dict(neighbor_radius=4, evaluate_distortion=True), 22164136,
     548916 + 5120060 + 4194300 + 22164136 * 48),
    # no radius: no grid, 16-byte samples
    (dict(), 67074556, 548916 + 67074556 * 16),
    # a budget large enough for the 4 GiB binding limit to bind: 2^32 - 1 / 12
    (dict(max_graph_memory_mib=64 << 10), 357913941, 548916 + 357913941 * 16),
])
def test_capacity_model_known_answers(bt, kw, capacity, reserved):
    assert _PER_GRAPH == 548916 and _GRID_1600_R4 == 5120060
    o = bt.Options(**{"max_graph_memory_mib": 1024, **kw})
    cap, res = bt.capacity(o, 19600, 9801)
    assert (cap, res) == (capacity, reserved)
    budget = o.c.max_graph_memory_bytes
    fixed = res - cap * (48 if kw.get("neighbor_radius") else 16)
    if cap < 357913941:
        assert budget - res < (48 if kw.get("neighbor_radius") else 16)  # maximal
    assert fixed == _PER_GRAPH + (_GRID_1600_R4 if kw.get("neighbor_radius") else 0) + (
        4194300 if kw.get("evaluate_distortion") else 0)


def test_capacity_model_limits(bt):
    """The budget checks of Resampler::new and Allocation::new: grid plus
    staging must be below the budget; fixed bytes must fit; capacity >= 3."""
    # radius 1: grid of 1600^2 cells = 32 * 2,560,000 + 60 B > 64 MiB
    with pytest.raises(bt.BtError) as e:
        bt.capacity(bt.Options(neighbor_radius=1, max_graph_memory_mib=64), 19600, 9801)
    assert e.value.code == 3
    # exactly the grid + fixed bytes + 2 samples of budget: no room (capacity < 3)
    budget = 548916 + 5120060 + 2 * 48
    o = bt.Options(neighbor_radius=4, max_graph_memory_mib=budget / _MIB)
    assert o.c.max_graph_memory_bytes == budget
    with pytest.raises(bt.BtError) as e:
        bt.capacity(o, 19600, 9801)
    assert e.value.code == 3
    o = bt.Options(neighbor_radius=4, max_graph_memory_mib=(budget + 48) / _MIB)
    assert bt.capacity(o, 19600, 9801) == (3, budget + 48)
    # fixed bytes above the budget
    o = bt.Options(neighbor_radius=4, max_graph_memory_mib=(548916 + 5120060 - 1) / _MIB)
    with pytest.raises(bt.BtError) as e:
        bt.capacity(o, 19600, 9801)
    assert e.value.code == 3
    o = bt.Options(neighbor_radius=4, max_graph_memory_mib=(548916 + 5120060 + 3 * 48) / _MIB)
    assert bt.capacity(o, 19600, 9801)[0] == 3


def test_distortion_matches_oracle(bt, orc):
    rng = np.random.default_rng(4)
    n = 400
    a = rng.uniform(-0.9, 0.9, (n, 2)).astype(np.float32)
    b = (a + rng.normal(0, 0.01, (n, 2))).astype(np.float32)
    counts = np.array([0, 50, 150, 2, 198], dtype=np.uint32)
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint64)
    segs = np.concatenate([[0], np.cumsum(np.maximum(counts.astype(np.int64) - 1, 0))[:-1]])
    paths = np.stack([offs, segs, counts], 1).astype(np.uint32)
    paths[0] = 0
    L = bt.Layout(paths, n)
    d = bt.distortion(bt.Snapshot(L, a), bt.Snapshot(L, b), map_size=1600)
    od, orms = orc.distortion(a, b, offs, counts)
    assert d[0] == od and d[1] == orms and d[2] == orms * 800


def test_layout_rejects_bad_reservations(bt):
    ok = np.array([[0, 0, 3], [3, 2, 2], [0, 0, 0]], dtype=np.uint32)
    L = bt.Layout(ok, 5)
    assert L.edge_count == 2 and L.skipped == 1
    with pytest.raises(bt.BtError):
        bt.Layout(np.array([[0, 0, 3], [4, 2, 2]], dtype=np.uint32), 6)   # gap
    with pytest.raises(bt.BtError):
        bt.Layout(np.array([[0, 0, 1]], dtype=np.uint32), 1)              # one sample
    rec = np.zeros(5, dtype=[("x", "<f4"), ("y", "<f4"), ("id", "<u4")])
    rec["id"] = [0, 0, 0, 1, 1]
    L.decode(rec.tobytes())
    rec["id"][4] = 0
    with pytest.raises(bt.BtError):
        L.decode(rec.tobytes())


@pytest.mark.parametrize("small,large", [(0x0F000000, 0), (0, 0xF0F0F0F0), (0x01010101, 0)])
def test_server_refuses_cu_masks_that_leave_an_xcd_empty(bt, small, large):
    """A lane's CU mask selects, with bit b of every 32-bit group, a CU of
    XCD b mod 8 (the hardware deals workgroups to the 8 XCDs round-robin),
    so a mask without a set bit in some residue class mod 8 would leave
    that XCD's workgroups nowhere to run: creation refuses it before any
    device call."""
    from bundling_thingy_amd import api
    with pytest.raises(api.BtError) as e:
        bt.Server(policy=bt.Server.LANES, small_cu_mask=small, large_cu_mask=large)
    assert e.value.code == api.BT_E_INVALID


def test_sign_of_bit_difference_is_the_radius_test():
    """The neighbor-mean matrix-product streams decide fmaf(dx, dx, dy*dy) <= r2
    as the sign of bits(d2) - (bits(r2) + 1) in int32 (d2 is +0, positive or
    +inf). Checked against the float compare on random, near-tie, subnormal,
    zero and infinite squared distances and radii, ties included; shifting the
    bias by one (a plain bits(d2) - bits(r2)) fails on the ties."""
    rng = np.random.default_rng(5)
    f32 = np.float32
    r2 = np.concatenate([rng.uniform(0, 1e4, 2000), [0.0, 1e-45, 1.0, 900.0, 3.4e38]]).astype(f32)
    with np.errstate(over="ignore"):  # 2 x 3.4e38 -> +inf, a valid squared distance
        base = rng.uniform(0, 2, (r2.size, 64)).astype(f32) * r2[:, None]
    ties = np.repeat(r2[:, None], 4, 1)
    ulp_up = np.nextafter(ties, f32(np.inf))
    ulp_dn = np.nextafter(ties, f32(0))
    special = np.broadcast_to(np.array([0.0, 1e-45, 1e-40, np.inf], dtype=f32), (r2.size, 4))
    d2 = np.concatenate([base, ties, ulp_up, ulp_dn, special], axis=1).astype(f32)
    want = d2 <= r2[:, None]
    bits = d2.view(np.int32).astype(np.int64)
    rb = r2.view(np.int32).astype(np.int64)[:, None] + 1
    diff = ((bits - rb + 2**31) % 2**32 - 2**31)  # int32 wrap, as the kernel computes it
    np.testing.assert_array_equal(diff < 0, want)
    wrong = ((bits - (rb - 1) + 2**31) % 2**32 - 2**31) < 0
    assert np.any(wrong != want)
