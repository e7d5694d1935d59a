"""VBX files on the GPU path (SaveVBX / LoadVBX, gvdb_volume_gvdb.cpp:
507-685, 1626-1767), through the C-ABI:
  - the file the GPU library writes equals the oracle's byte for byte
    (level set + colour channels, after a full rebuild and after an
    incremental update that deactivated leaves);
  - loading it into a fresh volume gives the oracle's topology, atlas map,
    bounds and active bricks exactly; saving again gives the same bytes;
  - the loaded volume renders the same image as the one saved;
  - truncated files, bitmask files and missing files are refused.
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


def _build(incremental):
    import gvdbx as G
    dev = _dev()
    rng = np.random.default_rng(19)
    p1 = (rng.random((60000, 3)) * [110, 80, 100] + [20, 30, 25]).astype(np.float32)
    clr = rng.integers(0, 256, (len(p1), 4), dtype=np.uint8)
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.AddChannel(0, G.T_FLOAT, 1)
    gv.AddChannel(1, G.T_UCHAR4, 1)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.add_channel(1, 1, dtype=O.T_UCHAR4)
    tp = torch.from_numpy(p1).to(dev)
    tc = torch.from_numpy(clr).to(dev)
    gv.SetPoints(tp, clr=tc)
    gv.RebuildTopology(len(p1), 3.0)
    ov.rebuild_topology(p1, 3.0)
    pts = p1
    if incremental:
        p2 = (rng.random((60000, 3)) * [100, 90, 80] + [90, 60, 70]).astype(np.float32)
        tp = torch.from_numpy(p2).to(dev)
        gv.SetPoints(tp, clr=tc)
        gv.RequestFullRebuild(False)
        gv.RebuildTopology(len(p2), 3.0)
        ov.incremental_topology(p2, 3.0, accum=False)
        pts = p2
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ov.update_atlas()
    ln = gv.InsertPointsSubcell(4, len(pts), 3.0)
    gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), ln, 0, chan_clr=1)
    gv.UpdateApron(0, 3.0)
    gv.UpdateApron(1, 0.0)
    ov.insert_points_subcell(4, pts, 3.0)
    ov.gather_level_set(1.5)
    ov.gather_color(clr, chan=1)
    ov.update_apron(0, 3.0)
    ov.update_apron(1, 0.0)
    torch.cuda.synchronize()
    return gv, ov, (tp, tc)


@pytest.mark.parametrize("incremental", [False, True])
def test_vbx_save_load_vs_oracle(tmp_path, incremental):
    import gvdbx as G
    _log(f"[gpu] vbx incremental={incremental}")
    gv, ov, keep = _build(incremental)
    assert_same_topology(gv, ov)
    fg, fo = tmp_path / "gpu.vbx", tmp_path / "oracle.vbx"
    gv.SaveVBX(fg)
    ov.save_vbx(fo)
    bg, bo = fg.read_bytes(), fo.read_bytes()
    assert len(bg) == len(bo) and bg == bo, "GPU and oracle VBX files differ"
    # load into a fresh volume (its own configuration comes from the file)
    lv = G.VolumeGVDB(0)
    lv.LoadVBX(fg)
    torch.cuda.synchronize()
    assert_same_topology(lv, ov, points_bbox=False)
    assert list(lv.info().pos_min) == list(lv.info().obj_min)
    act = ov.nodes(0)["mFlags"] == 1
    assert incremental == (not act.all())
    for c in (0, 1):
        np.testing.assert_array_equal(lv.bricks(c)[act].view(np.uint32), ov.bricks(c)[act].view(np.uint32))
    f2 = tmp_path / "again.vbx"
    lv.SaveVBX(f2)
    assert f2.read_bytes() == bg
    # the loaded volume renders as the saved one
    cam = G.camera_perspective(48, 40, (80.0, 70.0, -60.0), (-25.0, -20.0, 60.0), (50.0, 0, 0), (0, 40.0, 0))
    rp = G.render_params(step=0.5, thresh=0.0, light=(60, 200, -80), shadow_amt=0.8)
    a, a8 = gv.Render(G.SHADE_LEVELSET, 0, cam, rp)
    b, b8 = lv.Render(G.SHADE_LEVELSET, 0, cam, rp)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.equal(a8, b8)
    assert (a8[..., 0] != a8[0, 0, 0]).any()


def test_vbx_transform_and_errors(tmp_path):
    import gvdbx as G
    _log("[gpu] vbx errors")
    gv, ov, keep = _build(False)
    gv.SetTransform((1, 2, 3), (0.5, 0.5, 0.25), (10, 20, 30), (-4, 5, 6))
    f = tmp_path / "x.vbx"
    gv.SaveVBX(f)
    lv = G.VolumeGVDB(0)
    lv.LoadVBX(f)
    assert lv.GetTransform() == ((1, 2, 3), (0.5, 0.5, 0.25), (10, 20, 30), (-4, 5, 6))
    data = f.read_bytes()
    bad = tmp_path / "short.vbx"
    bad.write_bytes(data[:len(data) // 2])
    with pytest.raises(G.GvdbxError) as e:
        G.VolumeGVDB(0).LoadVBX(bad)
    assert e.value.code == 1
    masks = tmp_path / "masks.vbx"
    masks.write_bytes(bytes([2, 0]) + data[2:2 + 48 + 4] + bytes([1]) + data[2 + 48 + 4:])
    with pytest.raises(G.GvdbxError) as e:
        G.VolumeGVDB(0).LoadVBX(masks)
    assert e.value.code == 1
    with pytest.raises(G.GvdbxError):
        G.VolumeGVDB(0).LoadVBX(tmp_path / "missing.vbx")
    with pytest.raises(G.GvdbxError):
        G.VolumeGVDB(0).SaveVBX(tmp_path / "empty.vbx")   # nothing to save


def test_vbx_forced_version(tmp_path):
    """LoadVBX(fname, force_maj, force_min) (VG.cpp:551-554): a file whose
    version bytes are wrong loads as the version forced; without it the
    header is read as version 1.0 (bitmask node records) and refused."""
    import gvdbx as G
    _log("[gpu] vbx forced version")
    gv, ov, keep = _build(False)
    f = tmp_path / "x.vbx"
    gv.SaveVBX(f)
    data = f.read_bytes()
    assert data[:2] == bytes([1, 11])
    relabel = tmp_path / "relabel.vbx"
    relabel.write_bytes(bytes([1, 0]) + data[2:])   # 1.0: no transform, bitmask node records
    with pytest.raises(G.GvdbxError):
        G.VolumeGVDB(0).LoadVBX(relabel)
    lv = G.VolumeGVDB(0)
    lv.LoadVBX(relabel, 1, 11)
    torch.cuda.synchronize()
    assert_same_topology(lv, ov, points_bbox=False)
    for c in (0, 1):
        np.testing.assert_array_equal(lv.bricks(c).view(np.uint32), ov.bricks(c).view(np.uint32))
    f2 = tmp_path / "again.vbx"
    lv.SaveVBX(f2)
    assert f2.read_bytes() == data
    with pytest.raises(G.GvdbxError):
        G.VolumeGVDB(0).LoadVBX(f, 300, 0)   # not a version byte
