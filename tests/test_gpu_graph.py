"""Frame graphs (a captured level-set frame replayed with hipGraphLaunch)
against the eager calls, through the C-ABI.  The graph replays over new
point data written into the captured point buffer; between replays the
test runs eager work and synchronises, and every replay's bricks must equal
the eager pipeline's (InsertPointsSubcell, GatherLevelSet, UpdateApron) bit
for bit, which the oracle parity tests tie to the reference."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _frames(n, k, seed=5):
    rng = np.random.default_rng(seed)
    base = rng.random((n, 3)) * [120, 70, 110] + [20, 30, 25]
    out = []
    for f in range(k):
        jit = rng.normal(scale=0.6, size=(n, 3)) + [2.0 * f, 0.5 * f, -1.0 * f]
        out.append((base + jit).astype(np.float32))
    return out


def _vol(G, union, radius):
    dev = _dev()
    gv = G.VolumeGVDB(0, stream=torch.cuda.current_stream())
    gv.configure([3, 3, 3, 3, 5])
    gv.AddChannel(0, G.T_FLOAT, 1)
    tu = torch.from_numpy(union).to(dev)
    gv.SetPoints(tu)
    gv.RebuildTopology(len(union), radius)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    return gv, tu


@pytest.mark.parametrize("accum", [False, True])
def test_frame_graph_replays_equal_eager(accum):
    import gvdbx as G
    dev = _dev()
    n, k = 120000, 4
    frames = _frames(n, k)
    union = np.concatenate(frames)
    _log(f"[gpu] frame graph accum={accum}")
    gg, tug = _vol(G, union, 3.0)   # replays into this one
    ge, tue = _vol(G, union, 3.0)   # the eager pipeline
    buf = torch.from_numpy(frames[0]).to(dev)
    gg.SetPoints(buf)
    if accum:
        gg.FillChannel(0, 3.0)
        ge.FillChannel(0, 3.0)
    graph = gg.FrameGraph(n, 4, 3.0, 1.5, chan=0, accum=accum, apron_chan=-1 if accum else 0, boundval=3.0)
    # the create call ran frame 0 eagerly; the eager volume does the same
    te = torch.from_numpy(frames[0]).to(dev)
    ge.SetPoints(te)
    ln = ge.InsertPointsSubcell(4, n, 3.0)
    ge.GatherLevelSet(4, n, 1.5, (0, 0, 0), ln, 0, accum=accum)
    if not accum:
        ge.UpdateApron(0, 3.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(gg.bricks(0), ge.bricks(0))
    for rep in range(2):
        for f in range(1, k):
            buf.copy_(torch.from_numpy(frames[f]).to(dev))
            graph.launch()
            te = torch.from_numpy(frames[f]).to(dev)
            ge.SetPoints(te)
            ln = ge.InsertPointsSubcell(4, n, 3.0)
            ge.GatherLevelSet(4, n, 1.5, (0, 0, 0), ln, 0, accum=accum)
            if not accum:
                ge.UpdateApron(0, 3.0)
            torch.cuda.synchronize()
            got, want = gg.bricks(0), ge.bricks(0)
            np.testing.assert_array_equal(got, want, err_msg=f"replay {rep}/{f}")
            _log(f"  replay {rep}/{f} equal")
    # eager calls on the graph's volume still see consistent bins
    ln = gg.InsertPointsSubcell(4, n, 3.0)
    assert ln == ge.InsertPointsSubcell(4, n, 3.0) or accum
    # a new topology invalidates the graph
    gg.RebuildTopology(n, 3.0)
    with pytest.raises(G.GvdbxError) as e:
        graph.launch()
    assert e.value.code == 2
    del graph


def test_frame_graph_refusals():
    import gvdbx as G
    dev = _dev()
    frames = _frames(20000, 1)
    gv, tu = _vol(G, frames[0], 3.0)
    with pytest.raises(G.GvdbxError):
        gv.FrameGraph(20000, 3, 3.0, 1.5)   # subcell size must divide the brick
    clr = torch.zeros((20000, 4), dtype=torch.uint8, device=dev)
    gv.SetPoints(tu, clr=clr)
    with pytest.raises(G.GvdbxError):
        gv.FrameGraph(20000, 4, 3.0, 1.5)   # colours are not captured


def test_frame_graph_outlives_its_volume():
    """Destroying the volume retires its graphs: launching one then fails
    cleanly and destroying it afterwards is safe."""
    import gvdbx as G
    dev = _dev()
    frames = _frames(20000, 1)
    gv, tu = _vol(G, frames[0], 3.0)
    graph = gv.FrameGraph(20000, 4, 3.0, 1.5)
    graph.launch()
    torch.cuda.synchronize()
    graph._vol = None
    del gv
    import gc
    gc.collect()
    with pytest.raises(G.GvdbxError) as e:
        graph.launch()
    assert e.value.code == 2
    del graph


def test_points_rewritten_in_place_are_read_again():
    """InsertPointsSubcell reads the point buffer at call time, as the
    reference does: rewriting the buffer in place (no SetPoints) between
    frames on a fixed topology gives the new frame's lists and level set."""
    import gvdbx as G
    dev = _dev()
    n = 50000
    frames = _frames(n, 2, seed=9)
    union = np.concatenate(frames)
    ga, _ = _vol(G, union, 3.0)
    gb, _ = _vol(G, union, 3.0)
    buf = torch.from_numpy(frames[0]).to(dev)
    ga.SetPoints(buf)
    la0 = ga.InsertPointsSubcell(4, n, 3.0)
    ga.GatherLevelSet(4, n, 1.5, (0, 0, 0), la0, 0)
    buf.copy_(torch.from_numpy(frames[1]).to(dev))   # in place
    la1 = ga.InsertPointsSubcell(4, n, 3.0)
    ga.GatherLevelSet(4, n, 1.5, (0, 0, 0), la1, 0)
    t1 = torch.from_numpy(frames[1]).to(dev)
    gb.SetPoints(t1)
    lb1 = gb.InsertPointsSubcell(4, n, 3.0)
    gb.GatherLevelSet(4, n, 1.5, (0, 0, 0), lb1, 0)
    torch.cuda.synchronize()
    assert la1 == lb1 and la0 != la1
    np.testing.assert_array_equal(ga.bricks(0), gb.bricks(0))
