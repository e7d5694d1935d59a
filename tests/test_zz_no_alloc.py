"""gvdbx_gather_level_set_noalloc (the GatherLevelSet that allocates nothing)
against gvdbx_gather_level_set through the C-ABI: the same interior bits,
and the same bricks after UpdateApron, on two calls in a row, the second on
a smaller point set, fresh and accumulated, and a clear error beyond the
scratch limit."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _volume(G, pts, R):
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.AddChannel(0, G.T_FLOAT, 1)
    tp = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).to("cuda:0")
    gv.SetPoints(tp)
    gv._pts = tp
    gv.RebuildTopology(len(pts), R)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    return gv


def test_noalloc_gather_matches_gather():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gvdbx as G
    L = G.lib()
    rng = np.random.default_rng(7070)
    for n, accum in ((120000, False), (30000, True)):   # the second call is smaller
        _log(f"[gpu] no-alloc gather: {n} points, accumulate {accum}")
        pts = (rng.random((n, 3)) * [150, 90, 120] + [10, 20, 15]).astype(np.float32)
        gv = _volume(G, pts, 3.0)
        ln = gv.InsertPointsSubcell(4, n, 3.0)
        gv.FillChannel(0, 2.0)
        gv.GatherLevelSet(4, n, 1.5, (0, 0, 0), ln, 0, -1, accum)
        ref = gv.bricks(0)
        gv.UpdateApron(0, 3.0)
        ref_apron = gv.bricks(0)
        gv.FillChannel(0, 2.0)
        code = L.gvdbx_gather_level_set_noalloc(gv._h, 4, n, C.c_float(1.5), C.c_float(0), C.c_float(0),
                                                 C.c_float(0), C.c_int64(ln), 0, -1, int(accum))
        assert code == 0, G.last_error()
        got = gv.bricks(0)
        # the gather defines the interior voxels (the default one also fills
        # the neighbours' aprons; the allocation-free one leaves them to
        # UpdateApron), and after UpdateApron the whole bricks agree
        np.testing.assert_array_equal(got[:, 1:9, 1:9, 1:9], ref[:, 1:9, 1:9, 1:9])
        gv.UpdateApron(0, 3.0)
        np.testing.assert_array_equal(gv.bricks(0), ref_apron)
        assert (ref[:, 1:9, 1:9, 1:9] < 3.0).any()
        gv.close()


def test_noalloc_gather_rejects_beyond_scratch():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gvdbx as G
    n = (64 << 20) // 4 + 4096   # one more page of points than the scratch holds
    _log(f"[gpu] no-alloc gather beyond the scratch: {n} points")
    rng = np.random.default_rng(7071)
    pts = (rng.random((n, 3)) * [500, 250, 500] + 64).astype(np.float32)
    gv = _volume(G, pts, 3.0)
    ln = gv.InsertPointsSubcell(4, n, 3.0)
    with pytest.raises(G.GvdbxError) as e:
        gv.GatherLevelSetNoAlloc(4, n, 1.5, (0, 0, 0), ln, 0)
    assert e.value.code == 5 and "scratch" in str(e.value)
    gv.close()
