"""Oracle parity at the benchmark workloads' real sizes (GPU, through the C-ABI).

  - shell workload: 2^20 points uniform by volume in the shell r in [180, 200]
    about (256, 256, 256), topology and subcell radius 4, cubic splat r = 2
    (FP64 tolerance), GatherDensity r = 2 and UpdateApron(0) bit-exact;
  - dam-break level set at full size (2^24 points, frame 0 and one
    ABC-advected frame): the whole frame runs on the GPU exactly as the
    benchmark runs it, and every leaf within x <= 104 (bricks, aprons
    included) is compared bit for bit with the oracle run on the points
    x < 128, whose result there cannot differ (a voxel's subcell list and a
    leaf's existence only involve points less than 16 voxels away);
  - the binning sort's wide tiles, forced at small sizes, against the oracle.
"""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import assert_same_topology

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _G():
    import gvdbx
    return gvdbx


def test_shell_full_size():
    G = _G()
    dev = _dev()
    n = 1 << 20
    _log(f"[gpu] shell workload: {n} points")
    pts = G.gen_shell(20261203, n, (256.0, 256.0, 256.0), 180.0, 200.0)
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.AddChannel(0, G.T_FLOAT, 1)
    gv.AddChannel(1, G.T_FLOAT, 1)
    tp = torch.from_numpy(pts).to(dev)
    gv.SetPoints(tp)
    gv.RebuildTopology(n, 4.0)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ln = gv.InsertPointsSubcell(4, n, 4.0)
    gv.SplatCubic(n, 2.0, (0, 0, 0), 0)
    gv.UpdateApron(0, 0.0)
    gv.GatherDensity(4, n, 2.0, (0, 0, 0), ln, 1)
    gv.UpdateApron(1, 0.0)
    torch.cuda.synchronize()
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.add_channel(1, 1)
    ov.rebuild_topology(pts, 4.0)
    ov.update_atlas()
    assert_same_topology(gv, ov)
    inf = gv.info()
    assert inf.depth == 2 and 35000 < inf.node_cnt[0] < 50000, (inf.depth, inf.node_cnt[0])
    assert ln == ov.insert_points_subcell(4, pts, 4.0)
    _log(f"  leaves {inf.node_cnt[0]}, list length {ln}")
    ov.gather_density(2.0, chan=1)
    ov.update_apron(1, 0.0)
    np.testing.assert_array_equal(gv.bricks(1), ov.bricks(1))
    ref = ov.splat_cubic_f64(pts, 2.0)
    got = gv.bricks(0)
    floor = 1e-6 * np.abs(ref).max()
    err = np.abs(got[:, 1:9, 1:9, 1:9].astype(np.float64) - ref)
    assert (err <= 1e-5 * np.abs(ref) + floor).all(), f"cubic splat: max err {err.max():.3e}"
    assert np.isfinite(got).all()


def _keyed(nodes):
    return {tuple(int(c) for c in p): i for i, p in enumerate(nodes["mPos"])}


@pytest.mark.parametrize("frame", [0, 1])
def test_dambreak_full_frame_vs_oracle_slab(frame):
    G = _G()
    from gvdbx import workloads
    dev = _dev()
    n = 1 << 24
    _log(f"[gpu] dam-break full frame {frame}: {n} points")
    frames, _ = workloads.dambreak_frames(torch, G.gen_box, 20261203 + 1, n, frame + 1, dev)
    p = frames[frame]
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.AddChannel(0, G.T_FLOAT, 1)
    gv.SetPoints(p)
    gv.RebuildTopology(n, 3.0)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ln = gv.InsertPointsSubcell(4, n, 3.0)
    gv.GatherLevelSet(4, n, 1.5, (0, 0, 0), ln, 0)
    gv.UpdateApron(0, 3.0)
    torch.cuda.synchronize()
    nodes = gv.nodes(0)
    bricks = gv.bricks(0)
    host = p.cpu().numpy()
    del frames, p
    sub = host[host[:, 0] < 128.0]
    _log(f"  GPU leaves {len(nodes)}; oracle slab x < 128: {len(sub)} points")
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(sub, 3.0)
    ov.update_atlas()
    ov.insert_points_subcell(4, sub, 3.0)
    ov.gather_level_set(1.5)
    ov.update_apron(0, 3.0)
    onodes = ov.nodes(0)
    obricks = ov.bricks(0)
    gsel = np.nonzero(nodes["mPos"][:, 0] <= 104)[0]
    osel = np.nonzero(onodes["mPos"][:, 0] <= 104)[0]
    gk = {tuple(int(c) for c in nodes["mPos"][i]): i for i in gsel}
    ok = {tuple(int(c) for c in onodes["mPos"][i]): i for i in osel}
    assert len(gk) > 10000, len(gk)
    assert set(gk) == set(ok), "leaf sets differ in the compared region"
    keys = sorted(gk)
    gi = np.array([gk[k] for k in keys])
    oi = np.array([ok[k] for k in keys])
    np.testing.assert_array_equal(bricks[gi], obricks[oi])
    _log(f"  {len(keys)} leaves equal bit for bit (incl. aprons)")
    vals = bricks[gi][:, 1:9, 1:9, 1:9]
    assert vals.min() < 0.0 and vals.max() == 3.0


@pytest.mark.parametrize("tiles,records", [(1, 0), (2, 0), (0, 1)])
def test_binning_sort_tiles_bit_exact(tiles, records):
    """Both binning sorts (records moved through the passes; (cell, index)
    pairs then a gather, with 4096- and 8192-entry tiles forced) at 200k
    points with invalid points: topology, list length, level set, apron and
    InsertPoints equal the oracle."""
    G = _G()
    dev = _dev()
    _log(f"[gpu] sort tiles {tiles} records {records}")
    rng = np.random.default_rng(600 + tiles)
    n = 200000
    pts = (rng.random((n, 3)) * [200, 120, 180] + [10, 20, 30]).astype(np.float32)
    pts[::101, 2] = 1.0e10
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.AddChannel(0, G.T_FLOAT, 1)
    gv.SetTuning("sort_tiles", tiles)
    gv.SetTuning("bin_records", records)
    assert gv.GetTuning("sort_tiles") == tiles and gv.GetTuning("bin_records") == records
    tp = torch.from_numpy(pts).to(dev)
    gv.SetPoints(tp)
    gv.RebuildTopology(n, 3.0)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ln = gv.InsertPointsSubcell(4, n, 3.0)
    gv.GatherLevelSet(4, n, 1.5, (0, 0, 0), ln, 0)
    gv.UpdateApron(0, 3.0)
    gv.InsertPoints(n)
    pnode, poff = gv.point_nodes(n)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, 3.0)
    ov.update_atlas()
    assert_same_topology(gv, ov)
    assert ln == ov.insert_points_subcell(4, pts, 3.0)
    ov.gather_level_set(1.5)
    ov.update_apron(0, 3.0)
    np.testing.assert_array_equal(gv.bricks(0), ov.bricks(0))
    opn, opo = ov.insert_points(pts)
    np.testing.assert_array_equal(pnode, opn)
    np.testing.assert_array_equal(poff, opo)   # rank in ascending point order: the bins' stable order
    with pytest.raises(G.GvdbxError):
        gv.SetTuning("sort_tiles", 3)
    with pytest.raises(G.GvdbxError):
        gv.SetTuning("bin_records", 2)


def test_binning_records_translated_bins():
    """Bins made for a translation other than the topology's (InsertPoints
    with trans: some points leave the bin grid) are the same with both
    binning sorts."""
    G = _G()
    dev = _dev()
    _log("[gpu] binning with a translation")
    rng = np.random.default_rng(611)
    n = 150000
    pts = (rng.random((n, 3)) * [120, 90, 100] + [20, 30, 25]).astype(np.float32)
    tp = torch.from_numpy(pts).to(dev)
    outs = []
    for records in (0, 1):
        gv = G.VolumeGVDB(0)
        gv.configure([3, 3, 3, 3, 5])
        gv.AddChannel(0, G.T_FLOAT, 1)
        gv.SetTuning("bin_records", records)
        gv.SetPoints(tp)
        gv.RebuildTopology(n, 1.0)
        gv.FinishTopology(False, True)
        gv.UpdateAtlas()
        gv.InsertPoints(n, (17.5, -9.25, 3.0))
        outs.append(gv.point_nodes(n))
        gv.ScatterDensity(n, 2.0, 1.0, (17.5, -9.25, 3.0), True)
        outs.append(gv.bricks(0))
    np.testing.assert_array_equal(outs[0][0], outs[2][0])
    np.testing.assert_array_equal(outs[0][1], outs[2][1])
    np.testing.assert_array_equal(outs[1], outs[3])
    assert (outs[0][0] >= 0).sum() > n // 3 and (outs[0][0] < 0).sum() > 1000
