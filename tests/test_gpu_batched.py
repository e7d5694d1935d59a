"""GPU parity of the batched forms.

Depth fusion: back-projection of all frames in one launch is bit-exact with
the oracle's FP32 restatement; fusing all frames in one pass equals the
reference's per-frame sequence (GatherLevelSet with accumulate on the same
topology) bit for bit, since min is exact and order-free; and both equal the
oracle's level set of the union of the frames' points.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _vol(G):
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.SetChannelDefault(16, 16, 1)
    gv.AddChannel(0, G.T_FLOAT, 1)
    return gv


def test_depth_fusion_batched_equals_sequential_and_oracle():
    import gvdbx as G
    from gvdbx import workloads
    dev = _dev()
    F, W, H = 6, 96, 72
    _log(f"depth fusion {F} frames {W}x{H}")
    d, poses, intr = workloads.depth_scene(11, nframes=F, width=W, height=H, fx=78.75, fy=78.75)
    origin, vox = (-13.0, -13.0, -13.0), 0.05
    pts_ref = O.backproject_depth(d, intr, poses, vox, origin)
    dt = torch.from_numpy(d).to(dev)
    pt = torch.from_numpy(poses).to(dev)
    pts = G.backproject_depth(dt, intr, pt, vox, origin)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(pts.cpu().numpy(), pts_ref)
    n = pts.shape[0]
    # batched: every frame in one pass
    gb = _vol(G)
    gb.SetPoints(pts)
    gb.RebuildTopology(n, 2.0)
    gb.FinishTopology(False, True)
    gb.UpdateAtlas()
    ln = gb.InsertPointsSubcell(4, n, 2.0)
    gb.GatherLevelSet(4, n, 1.0, (0, 0, 0), ln, 0)
    gb.UpdateApron(0, 0.0)
    # per frame, accumulating on the same topology (main_point_fusion.cpp:588-590)
    gs = _vol(G)
    gs.SetPoints(pts)
    gs.RebuildTopology(n, 2.0)
    gs.FinishTopology(False, True)
    gs.UpdateAtlas()
    gs.FillChannel(0, 3.0)
    per = W * H
    for f in range(F):
        fp = pts[f * per:(f + 1) * per]
        gs.SetPoints(fp)
        lf = gs.InsertPointsSubcell(4, per, 2.0)
        gs.GatherLevelSet(4, per, 1.0, (0, 0, 0), lf, 0, accum=True)
    gs.UpdateApron(0, 0.0)
    torch.cuda.synchronize()
    bb, bs = gb.bricks(0), gs.bricks(0)
    assert bb.shape[0] > 50
    np.testing.assert_array_equal(bb, bs)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts_ref, 2.0)
    ov.update_atlas()
    ov.insert_points_subcell(4, pts_ref, 2.0)
    ov.gather_level_set(1.0)
    ov.update_apron(0, 0.0)
    np.testing.assert_array_equal(bb, ov.bricks(0))
