"""Pins the CPU oracle against closed forms and brute force (CPU only).

The reference ships no tests, golden vectors or fixtures for this path, so
every oracle function is checked against an independent formulation in
tests/brute.py (set algebra / brute force) and against mathematical
invariants chosen so that plausible mistakes (key packing order, wrong link
slot, off-by-one subcell boxes, wrong apron neighbour) fail.
"""
import zlib

import numpy as np
import pytest

import brute
import oracle as O


def _topology_matches(ov, exp, levs):
    inf = ov.info()
    assert inf.depth == exp["depth"]
    assert inf.top_lev == exp["top_lev"]
    assert list(inf.root_pos) == list(exp["root"])
    np.testing.assert_array_equal(np.array(inf.pos_min), exp["pos_min"])
    np.testing.assert_array_equal(np.array(inf.pos_max), exp["pos_max"])
    for l in range(levs):
        assert inf.node_cnt[l] == exp["cnt"][l], f"level {l} count"
    for l in range(exp["depth"] + 1):
        nd = ov.nodes(l)
        e = exp["nodes"][l]
        np.testing.assert_array_equal(nd["mLev"], e["mLev"])
        np.testing.assert_array_equal(nd["mFlags"], e["mFlags"])
        np.testing.assert_array_equal(nd["mPos"], e["mPos"])
        np.testing.assert_array_equal(nd["mParent"], e["mParent"])
        np.testing.assert_array_equal(nd["mChildList"], e["mChildList"])
        if l == 0:
            np.testing.assert_array_equal(nd["mValue"], e["mValue"])
        if l >= 1:
            np.testing.assert_array_equal(ov.child_lists(l), exp["clists"][l])


CASES = [
    # (name, generator, radius, log2dims)
    ("box_r3", lambda r: (r.random((3000, 3)) * [70, 40, 55] + [9, 17, 30]), 3.0, (3, 3, 3, 3, 5)),
    ("box_r0.5_noextra", lambda r: (r.random((2000, 3)) * [60, 60, 60] + 5), 0.5, (3, 3, 3, 3, 5)),
    ("shell_r4", lambda r: _shell(r, 4000, 60, 40, 48), 4.0, (3, 3, 3, 3, 5)),
    ("clusters_r2", lambda r: np.concatenate([r.normal(c, 3.0, (500, 3)) for c in ([40, 40, 40], [90, 60, 30])]),
     2.0, (3, 3, 3, 3, 5)),
    ("single_l1_node", lambda r: (r.random((500, 3)) * 20 + 21), 3.0, (3, 3, 3, 3, 5)),
    ("top_below_root", lambda r: (r.random((600, 3)) * 60.9 + 2), 0.5, (3, 3, 3, 3, 5)),
    ("sample_cfg_32_bricks", lambda r: (r.random((3000, 3)) * [150, 90, 120] + 3), 3.0, (5, 3, 3, 3, 3)),
    ("small_bricks_4", lambda r: (r.random((800, 3)) * 30 + 2), 1.5, (2, 2, 2, 2, 3)),
]


def _shell(r, n, c, r0, r1):
    d = r.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rad = np.cbrt(r.uniform(r0 ** 3, r1 ** 3, n))
    return c + d * rad[:, None]


@pytest.mark.parametrize("name,gen,radius,log2dims", CASES, ids=[c[0] for c in CASES])
def test_oracle_topology_closed_form(name, gen, radius, log2dims):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    pts = np.asarray(gen(rng), dtype=np.float32)
    pts = np.abs(pts)
    ov = O.OracleVolume(log2dims)
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, radius)
    ov.update_atlas()
    exp = brute.expected_topology(pts, radius, (0, 0, 0), log2dims)
    _topology_matches(ov, exp, len(log2dims))
    if name == "top_below_root":
        assert exp["top_lev"] < exp["depth"]
    # atlas map: slot k -> leaf k, the rest undefined
    am = ov.atlas_map()
    n0 = exp["cnt"][0]
    np.testing.assert_array_equal(am["mLeafID"][:n0], np.arange(n0))
    np.testing.assert_array_equal(am["mPos"][:n0], exp["nodes"][0]["mPos"])
    assert (am["mLeafID"][n0:] == -1).all()


def test_oracle_invalid_points_dropped():
    rng = np.random.default_rng(3)
    pts = (rng.random((1000, 3)) * 50 + 10).astype(np.float32)
    bad = pts.copy()
    bad[::7, 2] = 1.0e10          # NOHIT
    bad[1::11, 0] = -3.0          # negative
    bad[2::13, 1] = np.nan        # non-finite
    keep = np.ones(len(pts), bool)
    keep[::7] = keep[1::11] = keep[2::13] = False
    a, b = O.OracleVolume(), O.OracleVolume()
    for v in (a, b):
        v.add_channel(0, 1)
    a.rebuild_topology(bad, 3.0)
    b.rebuild_topology(bad[keep], 3.0)
    for l in range(3):
        np.testing.assert_array_equal(a.nodes(l)["mPos"], b.nodes(l)["mPos"])


def test_oracle_translation_invariance():
    """Shifting points by a multiple of the top range shifts mPos only."""
    rng = np.random.default_rng(5)
    pts = (rng.random((1500, 3)) * 60 + 10).astype(np.float32)
    a, b = O.OracleVolume(), O.OracleVolume()
    a.rebuild_topology(pts, 3.0)
    b.rebuild_topology(pts, 3.0, orig=(512, 0, 512))
    ia, ib = a.info(), b.info()
    assert ia.depth == ib.depth and list(ia.node_cnt) == list(ib.node_cnt)
    for l in range(ia.depth):
        np.testing.assert_array_equal(a.nodes(l)["mPos"] + [512, 0, 512], b.nodes(l)["mPos"])
        np.testing.assert_array_equal(a.nodes(l)["mParent"], b.nodes(l)["mParent"])


def test_oracle_permutation_invariance():
    rng = np.random.default_rng(6)
    pts = (rng.random((2000, 3)) * 70 + 3).astype(np.float32)
    perm = rng.permutation(len(pts))
    a, b = O.OracleVolume(), O.OracleVolume()
    for v, p in ((a, pts), (b, pts[perm])):
        v.add_channel(0, 1)
        v.rebuild_topology(p, 3.0)
        v.update_atlas()
        ln = v.insert_points_subcell(4, p, 3.0)
        v.gather_level_set(1.5)
        v.update_apron(0, 3.0)
    assert a.info().sc_len == b.info().sc_len
    np.testing.assert_array_equal(a.bricks(), b.bricks())


@pytest.mark.parametrize("mode,R,r", [(1, 3.0, 1.5), (0, 4.0, 2.0), (1, 2.0, 1.0), (0, 1.0, 2.0)])
def test_oracle_gather_brute_force(mode, R, r):
    """Subcell-list membership + per-voxel min/max vs FP64 brute force (1e-5 abs; values are O(1))."""
    rng = np.random.default_rng(int(R * 10 + r))
    pts = (rng.random((1200, 3)) * [40, 30, 35] + [12, 9, 20]).astype(np.float32)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, R)
    ov.update_atlas()
    ov.insert_points_subcell(4, pts, R)
    if mode == 1:
        ov.gather_level_set(r)
    else:
        ov.gather_density(r)
    leaf_pos = ov.nodes(0)["mPos"]
    exp = brute.gather_f64(pts, (0, 0, 0), leaf_pos, 8, 4, R, r, mode)
    got = ov.bricks()[:, 1:9, 1:9, 1:9]
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-5)
    f64 = ov.gather_f64(mode, r)
    np.testing.assert_allclose(f64, exp, rtol=0, atol=1e-12)
    if mode == 0 and R >= r:
        # closed form (membership covers the whole support when R >= r)
        w = pts.astype(np.float64)
        k = 5
        p0 = leaf_pos[k]
        g = np.arange(8)
        vz, vy, vx = np.meshgrid(g + p0[2], g + p0[1], g + p0[0], indexing="ij")
        v = np.stack([vx.ravel(), vy.ravel(), vz.ravel()], 1)
        d = np.sqrt(((w[None] - v[:, None]) ** 2).sum(2))
        cf = np.maximum(0, (r - d).max(1)).reshape(8, 8, 8)
        np.testing.assert_allclose(got[k], cf, atol=1e-5)


def test_oracle_sc_len_closed_form():
    """List length = sum over points of the box subcells that lie in a leaf."""
    rng = np.random.default_rng(11)
    pts = (rng.random((700, 3)) * 45 + 6).astype(np.float32)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, 3.0)
    ov.update_atlas()
    ln = ov.insert_points_subcell(4, pts, 3.0)
    lut = brute.leaf_lookup(ov.nodes(0)["mPos"], 8)
    lo, hi = brute.subcell_box(pts, 3.0, 4)
    hi = hi - 3
    total = 0
    for i in range(len(pts)):
        for x in range(lo[i, 0], hi[i, 0] + 1, 4):
            for y in range(lo[i, 1], hi[i, 1] + 1, 4):
                for z in range(lo[i, 2], hi[i, 2] + 1, 4):
                    if min(x, y, z) >= 0 and (x // 8, y // 8, z // 8) in lut:
                        total += 1
    assert ln == total


@pytest.mark.parametrize("boundval", [3.0, 0.0])
def test_oracle_apron_neighbour_lookup(boundval):
    rng = np.random.default_rng(int(boundval) + 20)
    pts = (rng.random((1500, 3)) * [50, 35, 40] + 8).astype(np.float32)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, 3.0)
    ov.update_atlas()
    ov.insert_points_subcell(4, pts, 3.0)
    ov.gather_level_set(1.5)
    before = ov.bricks()
    ov.update_apron(0, boundval)
    after = ov.bricks()
    exp = brute.apron_expected(before, ov.nodes(0)["mPos"], 8, boundval)
    np.testing.assert_array_equal(after, exp)
    # interiors untouched
    np.testing.assert_array_equal(after[:, 1:9, 1:9, 1:9], before[:, 1:9, 1:9, 1:9])


def test_oracle_cubic_splat_brute_force():
    rng = np.random.default_rng(31)
    pts = (rng.random((900, 3)) * 40 + 10).astype(np.float32)
    wts = rng.uniform(0.1, 1.0, len(pts)).astype(np.float32)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, 4.0)
    ov.update_atlas()
    ov.splat_cubic(pts, 2.0, weights=wts)
    leaf_pos = ov.nodes(0)["mPos"]
    exp = brute.cubic_f64(pts, (0, 0, 0), leaf_pos, 8, 2.0, wts)
    got = ov.bricks()[:, 1:9, 1:9, 1:9]
    np.testing.assert_allclose(got, exp, rtol=1e-5, atol=1e-6 * exp.max())
    np.testing.assert_allclose(ov.splat_cubic_f64(pts, 2.0, weights=wts), exp, rtol=1e-12, atol=1e-12)
    # closed form: a lone point at a voxel centre gives k(0)=w at that voxel
    one = np.array([[20.0, 20.0, 20.0]], dtype=np.float32)
    o1 = O.OracleVolume()
    o1.add_channel(0, 1)
    o1.rebuild_topology(np.concatenate([one, one + 30]), 4.0)
    o1.update_atlas()
    o1.splat_cubic(one, 2.0, weights=np.array([0.5], np.float32))
    b = o1.bricks()
    k = [i for i, p in enumerate(o1.nodes(0)["mPos"]) if tuple(p) == (16, 16, 16)][0]
    assert b[k, 5, 5, 5] == np.float32(0.5)
    assert b[k, 5, 5, 6] == np.float32(0.5 * (1 - 1 / 4) ** 3)


def test_oracle_scatter_density_serial_sum():
    """ScatterDensity: unclamped cubic at the home voxel and 6 neighbours, summed serially."""
    pts = np.array([[20.25, 21.5, 22.75], [20.5, 21.25, 22.0], [40.0, 40.0, 40.0]], np.float32)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, 1.0)
    ov.update_atlas()
    ov.scatter_density(pts, 2.0, expand=True)
    b = ov.bricks()
    k = [i for i, p in enumerate(ov.nodes(0)["mPos"]) if tuple(p) == (16, 16, 16)][0]

    def df(p, q, r=2.0):
        c = np.float32(((np.float64(q) - p) ** 2).sum() / (r * r))
        return 1 + c * (-3 + c * (3 - c))

    p0 = pts[0].astype(np.float64) - 16
    p1 = pts[1].astype(np.float64) - 16
    home0, home1 = np.floor(p0), np.floor(p1)   # both (4,5,6)
    assert (home0 == home1).all()
    v = df(p0, home0) + df(p1, home1)
    h = home0.astype(int) + 1
    assert abs(b[k, h[2], h[1], h[0]] - v) < 1e-6
    nb = df(p0, home0 + [1, 0, 0]) + df(p1, home1 + [1, 0, 0])
    assert abs(b[k, h[2], h[1], h[0] + 1] - nb) < 1e-6


def test_oracle_insert_points_stable_rank():
    rng = np.random.default_rng(41)
    pts = (rng.random((800, 3)) * 30 + 9).astype(np.float32)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, 0.5)
    ov.update_atlas()
    pnode, poff = ov.insert_points(pts)
    lut = brute.leaf_lookup(ov.nodes(0)["mPos"], 8)
    cells = np.floor(pts / 8).astype(int)
    exp = np.array([lut[tuple(c)] for c in cells])
    np.testing.assert_array_equal(pnode, exp)
    seen = {}
    for i, k in enumerate(exp):
        assert poff[i] == seen.get(k, 0)
        seen[k] = seen.get(k, 0) + 1


def test_oracle_error_paths():
    ov = O.OracleVolume()
    with pytest.raises(O.OracleError) as e:
        ov.rebuild_topology(np.zeros((0, 3), np.float32), 1.0)
    assert e.value.code == 4
    with pytest.raises(O.OracleError) as e:
        ov.rebuild_topology(np.array([[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]], np.float32), 1.0)   # fits one brick
    assert e.value.code == 3
    with pytest.raises(O.OracleError):
        ov.insert_points_subcell(4, np.zeros((1, 3), np.float32), 3.0)
