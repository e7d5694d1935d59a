"""GPU parity of the sparse topology: cells found by a radix sort of 64-bit
cell keys and looked up through a hash table (GVDBX_TOPO_RADIX, and AUTO
beyond the dense-grid limit of 2^28 level-0 cells).

Every result is compared with the oracle as bit-exactly as the dense path
is: node records, child lists, atlas map, counts, top level and bounds;
subcell list length; level-set and density gathers; apron; InsertPoints and
ScatterDensity.  The cubic splat is checked against the FP64 evaluation
(1e-5 relative, floor 1e-6 * max|rho|).  The two-cluster cases span
billions of level-0 cells, which only the sparse grids can hold; the oracle
allocates per node, not per cell, so it runs them as they are.
"""
import sys
import zlib

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import TOPO_CASES, assert_same_topology

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build_pair(pts, radius, log2dims=(3, 3, 3, 3, 5), mode=None):
    import gvdbx as G
    dev = _dev()
    gv = G.VolumeGVDB(0)
    gv.configure(log2dims)
    gv.SetChannelDefault(16, 16, 1)
    gv.AddChannel(0, G.T_FLOAT, 1)
    if mode is not None:
        gv.SetTopologyMode(mode)
    tp = torch.from_numpy(np.ascontiguousarray(pts, dtype=np.float32)).to(dev)
    gv.SetPoints(tp)
    gv.RebuildTopology(len(pts), radius)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ov = O.OracleVolume(log2dims)
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, radius)
    ov.update_atlas()
    return gv, ov, tp


def _check_gathers(gv, ov, pts, R, r, sc=4):
    gl = gv.InsertPointsSubcell(sc, len(pts), R)
    ol = ov.insert_points_subcell(sc, pts, R)
    assert gl == ol
    gv.GatherLevelSet(sc, len(pts), r, (0, 0, 0), gl, 0)
    ov.gather_level_set(r)
    np.testing.assert_array_equal(gv.bricks()[:, 1:9, 1:9, 1:9], ov.bricks()[:, 1:9, 1:9, 1:9])
    gv.UpdateApron(0, 3.0)
    ov.update_apron(0, 3.0)
    np.testing.assert_array_equal(gv.bricks(), ov.bricks())
    gv.GatherDensity(sc, len(pts), r, (0, 0, 0), gl, 0)
    ov.gather_density(r)
    np.testing.assert_array_equal(gv.bricks()[:, 1:9, 1:9, 1:9], ov.bricks()[:, 1:9, 1:9, 1:9])
    gv.UpdateApron(0, 0.0)
    ov.update_apron(0, 0.0)
    np.testing.assert_array_equal(gv.bricks(), ov.bricks())


@pytest.mark.parametrize("name,gen,radius,log2dims", TOPO_CASES, ids=[c[0] for c in TOPO_CASES])
def test_radix_topology_parity(name, gen, radius, log2dims):
    """The dense-size topology cases built with the sparse grids forced."""
    import gvdbx as G
    _log(f"[gpu] radix topology {name}")
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    pts = np.asarray(gen(rng), dtype=np.float32)
    gv, ov, _ = build_pair(pts, radius, log2dims, mode=G.TOPO_RADIX)
    assert gv.info().topology_mode_used == G.TOPO_RADIX
    assert_same_topology(gv, ov)


@pytest.mark.parametrize("R,r,sc", [(3.0, 1.5, 4), (0.5, 0.5, 4), (6.0, 2.0, 2)])
def test_radix_gathers_parity(R, r, sc):
    """Subcell length, both gathers and the apron on a radix topology; R = 6
    takes the generic extra-brick boxes (2R beyond half a brick; subcells of
    2 keep R + sc within the brick), R = 0.5 makes no extra bricks."""
    import gvdbx as G
    _log(f"[gpu] radix gathers R={R}")
    rng = np.random.default_rng(int(R * 10))
    pts = (rng.random((30000, 3)) * [90, 70, 80] + [20, 12, 30]).astype(np.float32)
    pts[::19, 2] = 1.0e10
    gv, ov, _ = build_pair(pts, R, mode=G.TOPO_RADIX)
    assert_same_topology(gv, ov)
    _check_gathers(gv, ov, pts, R, r, sc)


def _two_clusters(seed, n, far):
    rng = np.random.default_rng(seed)
    a = rng.normal((100.0, 120.0, 90.0), 7.0, (n, 3))
    b = rng.normal(far, 7.0, (n, 3))
    c = rng.random((n // 4, 3)) * 40 + np.asarray(far) * 0.5   # a third, uniform clump half way
    return np.abs(np.concatenate([a, b, c])).astype(np.float32)


# (far cluster centre, tree): diagonal, 1.5e10 cells in the box, depth 4 of
# the default tree; x near the 16-bit key limit (cell 60012; 3.6e11 cells in
# the box), which needs a sixth level for the root to cover it
FAR = [((100.0 + 8 * 3000, 120.0 + 8 * 2500, 90.0 + 8 * 2000), (3, 3, 3, 3, 5)),
       ((100.0 + 8 * 60000, 120.0 + 8 * 3000, 95.0 + 8 * 2000), (3, 3, 3, 3, 3, 5))]


@pytest.mark.parametrize("far,log2dims", FAR, ids=["diagonal", "x_65k"])
def test_two_clusters_beyond_dense_limit(far, log2dims):
    """Clusters billions of level-0 cells apart: DENSE refuses the box, AUTO
    takes the sparse grids, and topology, subcell length, gathers and apron
    match the oracle bit for bit."""
    import gvdbx as G
    _log(f"[gpu] two clusters {far}")
    pts = _two_clusters(5, 12000, far)
    R, r = 3.0, 1.5
    gd = G.VolumeGVDB(0)
    gd.configure(log2dims)
    gd.AddChannel(0, G.T_FLOAT, 1)
    gd.SetTopologyMode(G.TOPO_DENSE)
    tp = torch.from_numpy(pts).to(_dev())
    gd.SetPoints(tp)
    with pytest.raises(G.GvdbxError):
        gd.RebuildTopology(len(pts), R)
    gv, ov, _ = build_pair(pts, R, log2dims)
    assert gv.info().topology_mode_used == G.TOPO_RADIX
    assert_same_topology(gv, ov)
    _check_gathers(gv, ov, pts, R, r)


def test_two_clusters_scatter_and_splat():
    """InsertPoints / ScatterDensity bit-exact and the cubic splat within the
    FP64 bound on a sparse topology."""
    import gvdbx as G
    _log("[gpu] two clusters scatter / splat")
    pts = _two_clusters(6, 8000, FAR[0][0])
    gv, ov, tp = build_pair(pts, 4.0)
    assert gv.info().topology_mode_used == G.TOPO_RADIX
    gv.InsertPoints(len(pts), (0, 0, 0))
    gp, go = gv.point_nodes(len(pts))
    op, oo = ov.insert_points(pts)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(go, oo)
    gv.ScatterDensity(len(pts), 2.0, 1.0, (0, 0, 0), True)
    ov.scatter_density(pts, 2.0, expand=True)
    np.testing.assert_array_equal(gv.bricks(), ov.bricks())
    rng = np.random.default_rng(3)
    w = rng.uniform(0.1, 1.0, len(pts)).astype(np.float32)
    gv.SetPointWeights(torch.from_numpy(w).to(tp.device))
    gv.SplatCubic(len(pts), 2.0)
    got = gv.bricks()[:, 1:9, 1:9, 1:9].astype(np.float64)
    ref = ov.splat_cubic_f64(pts, 2.0, weights=w)
    floor = 1e-6 * np.abs(ref).max()
    err = np.abs(got - ref)
    assert (err <= 1e-5 * np.abs(ref) + floor).all(), f"max err {err.max():.3e}"


def test_radix_incremental_rebuilds_in_full():
    """An incremental RebuildTopology of a sparse topology rebuilds it in
    full (the result equals a fresh rebuild); AccumulateTopology refuses."""
    import gvdbx as G
    _log("[gpu] radix incremental")
    pts = _two_clusters(7, 6000, FAR[0][0])
    gv, ov, tp = build_pair(pts, 3.0)
    with pytest.raises(G.GvdbxError):
        gv.AccumulateTopology(len(pts), 3.0)
    gv.RequestFullRebuild(False)
    pts2 = pts + np.float32(0.75)
    tp2 = torch.from_numpy(pts2).to(tp.device)
    gv.SetPoints(tp2)
    gv.RebuildTopology(len(pts2), 3.0)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ov2 = O.OracleVolume((3, 3, 3, 3, 5))
    ov2.add_channel(0, 1)
    ov2.rebuild_topology(pts2, 3.0)
    ov2.update_atlas()
    assert_same_topology(gv, ov2)


def test_radix_vbx_round_trip(tmp_path):
    """A sparse volume saved as VBX equals the oracle's file byte for byte;
    loading it builds sparse grids from the node pools (the same topology,
    bricks and file again), and a gather on the loaded volume equals the one
    on the rebuilt volume."""
    import gvdbx as G
    _log("[gpu] radix vbx round trip")
    pts = _two_clusters(8, 6000, FAR[0][0])
    gv, ov, tp = build_pair(pts, 3.0)
    assert gv.info().topology_mode_used == G.TOPO_RADIX
    ln = gv.InsertPointsSubcell(4, len(pts), 3.0)
    gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), ln, 0)
    gv.UpdateApron(0, 3.0)
    ov.insert_points_subcell(4, pts, 3.0)
    ov.gather_level_set(1.5)
    ov.update_apron(0, 3.0)
    fg, fo = tmp_path / "gpu.vbx", tmp_path / "oracle.vbx"
    gv.SaveVBX(fg)
    ov.save_vbx(fo)
    assert fg.read_bytes() == fo.read_bytes()
    lv = G.VolumeGVDB(0)
    lv.LoadVBX(fg)
    assert lv.info().topology_mode_used == G.TOPO_RADIX
    assert_same_topology(lv, ov, points_bbox=False)
    np.testing.assert_array_equal(lv.bricks(), ov.bricks())
    f2 = tmp_path / "again.vbx"
    lv.SaveVBX(f2)
    assert f2.read_bytes() == fg.read_bytes()
    lv.SetPoints(tp)
    l2 = lv.InsertPointsSubcell(4, len(pts), 3.0)
    assert l2 == ln
    lv.FillChannel(0, 0.0)
    lv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), l2, 0)
    lv.UpdateApron(0, 3.0)
    np.testing.assert_array_equal(lv.bricks(), gv.bricks())
