"""Incremental topology on the GPU (AccumulateTopology and RebuildTopology
after RequestFullRebuild(false); ActivateIncreBricksGPU,
gvdb_volume_gvdb.cpp:1020-1178) against the oracle, through the C-ABI:
every node record, child list, the atlas map and the bounds bit for bit
after each frame; then InsertPointsSubcell's length, GatherLevelSet,
UpdateApron and InsertPoints on the updated tree, compared on the active
leaves (a deactivated leaf owns no atlas brick)."""
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


def _box(seed, n, lo, hi):
    r = np.random.default_rng(seed)
    return (r.random((n, 3)) * (np.array(hi) - np.array(lo)) + np.array(lo)).astype(np.float32)


def _pair(n=60000):
    return (_box(11, n, (20, 30, 25), (150, 110, 140)), _box(12, n, (90, 60, 70), (230, 200, 180)),
            _box(13, n, (10, 20, 200), (120, 160, 300)))


class Pair:
    def __init__(self, dev):
        import gvdbx as G
        self.G = G
        self.dev = dev
        self.gv = G.VolumeGVDB(0)
        self.gv.configure([3, 3, 3, 3, 5])
        self.gv.AddChannel(0, G.T_FLOAT, 1)
        self.ov = O.OracleVolume()
        self.ov.add_channel(0, 1)
        self.tp = None

    def points(self, p):
        self.tp = torch.from_numpy(p).to(self.dev)
        self.gv.SetPoints(self.tp)

    def finish(self):
        self.gv.FinishTopology(False, True)
        self.gv.UpdateAtlas()
        self.ov.update_atlas()
        torch.cuda.synchronize()
        assert_same_topology(self.gv, self.ov)


def _active_bricks_equal(gv, ov):
    act = ov.nodes(0)["mFlags"] == 1
    np.testing.assert_array_equal(gv.nodes(0)["mFlags"] == 1, act)
    np.testing.assert_array_equal(gv.bricks(0)[act], ov.bricks(0)[act])
    return int(act.sum())


def _downstream(pr, p, radius):
    n = len(p)
    ln = pr.gv.InsertPointsSubcell(4, n, radius)
    assert ln == pr.ov.insert_points_subcell(4, p, radius)
    pr.gv.GatherLevelSet(4, n, radius / 2, (0, 0, 0), ln, 0)
    pr.gv.UpdateApron(0, 3.0)
    pr.ov.gather_level_set(radius / 2)
    pr.ov.update_apron(0, 3.0)
    torch.cuda.synchronize()
    act = _active_bricks_equal(pr.gv, pr.ov)
    pr.gv.InsertPoints(n)
    pnode, poff = pr.gv.point_nodes(n)
    opn, opo = pr.ov.insert_points(p)
    np.testing.assert_array_equal(pnode, opn)
    np.testing.assert_array_equal(poff, opo)
    return act


@pytest.mark.parametrize("radius", [0.0, 3.0])
def test_accumulate_frames_vs_oracle(radius):
    dev = _dev()
    frames = _pair()
    pr = Pair(dev)
    for f, p in enumerate(frames):
        _log(f"[gpu] accumulate r={radius} frame {f}")
        pr.points(p)
        pr.gv.AccumulateTopology(len(p), radius)
        if f == 0:
            pr.ov.rebuild_topology(p, radius)
        else:
            pr.ov.incremental_topology(p, radius, accum=True)
        pr.finish()
    if radius > 0:
        act = _downstream(pr, frames[-1], radius)
        _log(f"  {act} active leaves equal")


@pytest.mark.parametrize("radius", [1.5, 3.0])
def test_incremental_rebuild_frames_vs_oracle(radius):
    dev = _dev()
    p1, p2, p3 = _pair()
    pr = Pair(dev)
    pr.gv.RequestFullRebuild(False)
    seq = [p1, p2, p1, p3]
    for f, p in enumerate(seq):
        _log(f"[gpu] incremental rebuild r={radius} frame {f}")
        pr.points(p)
        pr.gv.RebuildTopology(len(p), radius)
        if f == 0:
            pr.ov.rebuild_topology(p, radius)
        else:
            pr.ov.incremental_topology(p, radius, accum=False)
        pr.finish()
        inactive = int((pr.ov.nodes(0)["mFlags"] == 0).sum())
        assert f < 1 or inactive > 0
        act = _downstream(pr, p, radius)
        _log(f"  {act} active / {inactive} inactive leaves equal")


def test_no_new_node_and_errors():
    G = __import__("gvdbx")
    dev = _dev()
    p1, _, _ = _pair()
    pr = Pair(dev)
    pr.points(p1)
    pr.gv.RebuildTopology(len(p1), 3.0)
    pr.ov.rebuild_topology(p1, 3.0)
    pr.finish()
    before = [pr.gv.nodes(l).copy() for l in range(3)]
    # every box of a radius-1 update lies inside the radius-3 topology: nothing changes
    sub = p1[::3].copy()
    pr.points(sub)
    pr.gv.RequestFullRebuild(False)
    pr.gv.RebuildTopology(len(sub), 1.0)
    pr.ov.incremental_topology(sub, 1.0, accum=False)
    pr.finish()
    for l in range(3):
        np.testing.assert_array_equal(pr.gv.nodes(l), before[l])
    # AccumulateTopology refuses points outside the root and changes nothing
    far = np.concatenate([sub[:1000], np.array([[700.0, 5.0, 5.0]], dtype=np.float32)])
    pr.points(far)
    with pytest.raises(G.GvdbxError) as e:
        pr.gv.AccumulateTopology(len(far), 1.0)
    assert e.value.code == 5   # GVDBX_ERR_RANGE
    with pytest.raises(O.OracleError) as e:
        pr.ov.incremental_topology(far, 1.0, accum=True)
    assert e.value.code == 5
    for l in range(3):
        np.testing.assert_array_equal(pr.gv.nodes(l), before[l])
    # RebuildTopology in incremental mode rebuilds fully there (the depth changes)
    pr.gv.RebuildTopology(len(far), 1.0)
    pr.ov.rebuild_topology(far, 1.0)
    pr.finish()
    assert pr.gv.info().depth == 3
    # after RequestFullRebuild(true) AccumulateTopology rebuilds
    pr.gv.RequestFullRebuild(True)
    pr.points(p1)
    pr.gv.AccumulateTopology(len(p1), 3.0)
    pr.ov.rebuild_topology(p1, 3.0)
    pr.finish()


def test_accumulate_large_frames():
    """Dam-break block points advected twice, accumulated: 2M points per frame."""
    G = __import__("gvdbx")
    from gvdbx import workloads
    dev = _dev()
    n = 1 << 21
    _log(f"[gpu] accumulate {n} points x 3 frames")
    frames, _ = workloads.dambreak_frames(torch, G.gen_box, 777, n, 3, dev)
    pr = Pair(dev)
    for f in range(3):
        p = frames[f]
        pr.gv.SetPoints(p)
        pr.gv.AccumulateTopology(n, 3.0)
        host = p.cpu().numpy()
        if f == 0:
            pr.ov.rebuild_topology(host, 3.0)
        else:
            pr.ov.incremental_topology(host, 3.0, accum=True)
        pr.finish()
        _log(f"  frame {f}: {pr.gv.info().node_cnt[0]} leaves equal")
