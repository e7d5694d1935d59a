"""GPU parity of uchar4 colour channels (through the C-ABI) against the oracle.

The reference samples' call sequences with colour:
  - gFluidSurface (main_fluid_surface.cpp:92-94,325,354-356): float level set
    in channel 0, uchar4 colour in channel 1, GatherLevelSet(..., 0, 1),
    UpdateApron(0, 3.0), UpdateApron(1, 0.0);
  - gPointFusion (main_point_fusion.cpp:427-435,589): GatherLevelSet(..., 0, 1,
    accumulate);
  - gSprayDeposit (main_spray_deposit.cpp:150-151,290-298): GatherDensity(...,
    0, 1, accumulate).
Colour sums are exact integers on both sides, so every colour byte is
compared bit for bit, as are the float channels.
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


def _G():
    import gvdbx
    return gvdbx


def _pair(pts, clr, R, log2dims=(3, 3, 3, 3, 5), chans=(1,), pos_stride=12, clr_stride=4):
    """GPU volume with points (optionally interleaved with padding to test
    strides) and uchar4 colours, and the oracle volume with the same topology."""
    G = _G()
    dev = _dev()
    gv = G.VolumeGVDB(0)
    gv.configure(list(log2dims))
    gv.AddChannel(0, G.T_FLOAT, 1)
    for c in chans:
        gv.AddChannel(c, G.T_UCHAR4, 1)
    n = len(pts)
    pbuf = np.zeros((n, pos_stride // 4), np.float32)
    pbuf[:, :3] = pts
    cbuf = np.zeros((n, clr_stride), np.uint8)
    cbuf[:, :4] = clr
    tp = torch.from_numpy(pbuf).to(dev)
    tc = torch.from_numpy(cbuf).to(dev)
    gv.SetPoints(tp, None, tc, pos_stride=pos_stride, clr_stride=clr_stride)
    gv._keep_bufs = (tp, tc)
    gv.RebuildTopology(n, R)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ov = O.OracleVolume(log2dims)
    ov.add_channel(0, 1)
    for c in chans:
        ov.add_channel(c, 1, O.T_UCHAR4)
    ov.rebuild_topology(pts, R)
    ov.update_atlas()
    return gv, ov


def _data(seed, n, lo, span):
    rng = np.random.default_rng(seed)
    pts = (rng.random((n, 3)) * span + lo).astype(np.float32)
    clr = rng.integers(0, 256, (n, 4)).astype(np.uint8)
    return pts, clr


@pytest.mark.parametrize("pos_stride,clr_stride", [(12, 4), (16, 8)])
def test_fluid_surface_sequence(pos_stride, clr_stride):
    _log(f"[gpu] colour: fluid surface sequence strides {pos_stride}/{clr_stride}")
    pts, clr = _data(41, 30000, [20, 12, 30], [90, 70, 80])
    pts[::29, 2] = 1.0e10
    r = 1.0
    gv, ov = _pair(pts, clr, 2 * r, pos_stride=pos_stride, clr_stride=clr_stride)
    gl = gv.InsertPointsSubcell(4, len(pts), 2 * r)
    assert gl == ov.insert_points_subcell(4, pts, 2 * r)
    gv.GatherLevelSet(4, len(pts), r, (0, 0, 0), gl, 0, 1)
    ov.gather_level_set(r)
    ov.gather_color(clr, 1)
    gv.UpdateApron(0, 3.0)
    gv.UpdateApron(1, 0.0)
    ov.update_apron(0, 3.0)
    ov.update_apron(1, 0.0)
    np.testing.assert_array_equal(gv.bricks(0), ov.bricks(0))
    np.testing.assert_array_equal(gv.bricks_u8(1), ov.bricks_u8(1))


@pytest.mark.parametrize("mode", [1, 0])
def test_accumulate_sequences(mode):
    """Point fusion (level set) and spray deposit (density): accumulate twice
    onto filled channels, the second time with other points and colours."""
    _log(f"[gpu] colour: accumulate mode {mode}")
    pts, clr = _data(42 + mode, 20000, [10, 10, 10], [70, 60, 70])
    gv, ov = _pair(pts, clr, 3.0)
    gv.FillChannel(0, 2.5 if mode == 1 else 0.0)
    ov.fill_channel(0, 2.5 if mode == 1 else 0.0)
    gv.FillChannel(1, 0.3)
    ov.fill_channel(1, 0.3)
    np.testing.assert_array_equal(gv.bricks_u8(1), ov.bricks_u8(1))
    gl = gv.InsertPointsSubcell(4, len(pts), 3.0)
    assert gl == ov.insert_points_subcell(4, pts, 3.0)
    for rep in range(2):
        if mode == 1:
            gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), gl, 0, 1, True)
            ov.gather_level_set(1.5, accum=True)
        else:
            gv.GatherDensity(4, len(pts), 1.5, (0, 0, 0), gl, 0, 1, True)
            ov.gather_density(1.5, accum=True)
        ov.gather_color(clr, 1, accum=True)
        np.testing.assert_array_equal(gv.bricks(0)[:, 1:9, 1:9, 1:9], ov.bricks(0)[:, 1:9, 1:9, 1:9])
        np.testing.assert_array_equal(gv.bricks_u8(1)[:, 1:9, 1:9, 1:9], ov.bricks_u8(1)[:, 1:9, 1:9, 1:9],
                                      err_msg=f"accumulation {rep}")


@pytest.mark.parametrize("log2dims,sc,R", [((3, 3, 3, 3, 5), 2, 3.0), ((3, 3, 3, 3, 5), 1, 2.0),
                                           ((4, 3, 3, 3, 5), 1, 2.0)])
def test_gather_color_subcell_sizes(log2dims, sc, R):
    """Subcells of 2 and 1 voxels; 16^3 bricks at sc 1 hold 4096 subcells,
    which the kernel takes in several passes."""
    _log(f"[gpu] colour: subcells {log2dims} sc={sc}")
    pts, clr = _data(44 + sc, 8000, [6, 8, 5], [60, 50, 56])
    gv, ov = _pair(pts, clr, R, log2dims)
    gl = gv.InsertPointsSubcell(sc, len(pts), R)
    assert gl == ov.insert_points_subcell(sc, pts, R)
    gv.GatherLevelSet(sc, len(pts), 1.0, (0, 0, 0), gl, 0, 1)
    ov.gather_level_set(1.0)
    ov.gather_color(clr, 1)
    r0 = 1 << log2dims[0]
    np.testing.assert_array_equal(gv.bricks(0)[:, 1:r0 + 1, 1:r0 + 1, 1:r0 + 1],
                                  ov.bricks(0)[:, 1:r0 + 1, 1:r0 + 1, 1:r0 + 1])
    np.testing.assert_array_equal(gv.bricks_u8(1)[:, 1:r0 + 1, 1:r0 + 1, 1:r0 + 1],
                                  ov.bricks_u8(1)[:, 1:r0 + 1, 1:r0 + 1, 1:r0 + 1])


@pytest.mark.parametrize("avg", [True, False])
def test_scatter_color(avg):
    _log(f"[gpu] colour: scatter avg={avg}")
    pts, clr = _data(46, 20000, [5, 5, 5], [40, 30, 35])
    pts[::31, 1] = -1.0
    gv, ov = _pair(pts, clr, 0.5, chans=(1, 3))
    chan = 1 if avg else 3
    gv.FillChannel(chan, 0.2)
    ov.fill_channel(chan, 0.2)
    gv.InsertPoints(len(pts))
    gv.ScatterDensity(len(pts), 2.0, 1.0, (0, 0, 0), True, avg)
    ov.scatter_density(pts, 2.0, expand=True)
    ov.scatter_color(pts, clr, avg=avg, chan=chan)
    np.testing.assert_array_equal(gv.bricks(0), ov.bricks(0))
    np.testing.assert_array_equal(gv.bricks_u8(chan), ov.bricks_u8(chan))


def test_colour_error_paths():
    _log("[gpu] colour: error paths")
    G = _G()
    pts, clr = _data(47, 4000, [10, 10, 10], [40, 40, 40])
    gv, ov = _pair(pts, clr, 2.0, chans=(1,))
    gl = gv.InsertPointsSubcell(4, len(pts), 2.0)
    with pytest.raises(G.GvdbxError) as e:   # colour channel must be uchar4
        gv.GatherLevelSet(4, len(pts), 1.0, (0, 0, 0), gl, 0, 0)
    assert e.value.code == 1
    with pytest.raises(G.GvdbxError) as e:   # colour channel must exist
        gv.GatherLevelSet(4, len(pts), 1.0, (0, 0, 0), gl, 0, 5)
    assert e.value.code == 1
    with pytest.raises(G.GvdbxError) as e:   # the level set must be a float channel
        gv.GatherLevelSet(4, len(pts), 1.0, (0, 0, 0), gl, 1, 1)
    assert e.value.code == 1
    with pytest.raises(G.GvdbxError) as e:   # ScatterDensity without averaging writes channel 3
        gv.ScatterDensity(len(pts), 2.0, 1.0, (0, 0, 0), True, False)
    assert e.value.code == 1
    with pytest.raises(G.GvdbxError):
        gv.SplatCubic(len(pts), 2.0, (0, 0, 0), 1)
    with pytest.raises(G.GvdbxError):
        gv.AddChannel(2, 0, 1)   # T_UCHAR is not supported
    tc = torch.zeros(16, dtype=torch.uint8, device=_dev())
    tp = torch.zeros((3, 3), dtype=torch.float32, device=_dev())
    with pytest.raises(G.GvdbxError):   # misaligned colours
        gv.SetPoints(tp, None, tc[1:], clr_stride=4)
    # without colours, chan_clr is not used (the reference's colour list is null)
    gv2 = G.VolumeGVDB(0)
    gv2.configure([3, 3, 3, 3, 5])
    gv2.AddChannel(0, G.T_FLOAT, 1)
    tp2 = torch.from_numpy(pts).to(_dev())
    gv2.SetPoints(tp2)
    gv2.RebuildTopology(len(pts), 2.0)
    gv2.FinishTopology()
    gv2.UpdateAtlas()
    gl2 = gv2.InsertPointsSubcell(4, len(pts), 2.0)
    gv2.GatherLevelSet(4, len(pts), 1.0, (0, 0, 0), gl2, 0, 1)
