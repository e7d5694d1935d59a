"""The autotuner (gvdbx_autotune_gather / gvdbx_autotune_binning) on the
GPU: it keeps a configuration whose output is bit-identical to the one in
place (checked inside against a checksum), reports times, and the gather it
leaves set still equals the oracle bit for bit."""
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


@pytest.mark.parametrize("mode", [1, 0])
def test_autotune_keeps_results(mode):
    import gvdbx as G
    dev = _dev()
    _log(f"[gpu] autotune mode {mode}")
    rng = np.random.default_rng(404)
    n = 250000
    pts = (rng.random((n, 3)) * [150, 90, 130] + [20, 30, 25]).astype(np.float32)
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.AddChannel(0, G.T_FLOAT, 1)
    tp = torch.from_numpy(pts).to(dev)
    gv.SetPoints(tp)
    R = 3.0 if mode == 1 else 4.0
    r = 1.5 if mode == 1 else 2.0
    gv.RebuildTopology(n, R)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    tb = gv.AutotuneBinning(n, 4, R, reps=2)
    assert tb["candidates"] == 4 and tb["bin_records"] in (0, 1) and tb["sort_tiles"] in (1, 2)
    assert 0 < tb["best_ms"] <= tb["default_ms"]
    ln = gv.InsertPointsSubcell(4, n, R)
    tg = gv.AutotuneGather(mode, 4, r, chan=0, reps=2)
    _log(f"  binning {tb}  gather {tg}")
    assert tg["candidates"] == 22 and 0 < tg["best_ms"] <= tg["default_ms"]
    for k in ("gather_variant", "gather_classes", "gather_split", "xcd_remap"):
        assert gv.GetTuning(k) == tg[k]
    if mode == 1:
        gv.GatherLevelSet(4, n, r, (0, 0, 0), ln, 0)
        gv.UpdateApron(0, 3.0)
    else:
        gv.GatherDensity(4, n, r, (0, 0, 0), ln, 0)
        gv.UpdateApron(0, 0.0)
    torch.cuda.synchronize()
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, R)
    ov.update_atlas()
    assert ln == ov.insert_points_subcell(4, pts, R)
    if mode == 1:
        ov.gather_level_set(r)
        ov.update_apron(0, 3.0)
    else:
        ov.gather_density(r)
        ov.update_apron(0, 0.0)
    np.testing.assert_array_equal(gv.bricks(0), ov.bricks(0))
