"""GPU parity of Render with a colour channel (SetColorChannel VG.cpp:686-690)
against the oracle, through the C-ABI.

A ball of points with random uchar4 colours is gathered into a density
channel (0) and its colour mean (1), both bit-exact with the oracle first;
every shading that reads the colour channel (voxel, trilinear, tricubic and
level-set surfaces, the volume, the 3D section) then renders with channel 1
set.  The float image agrees with the oracle's FP32 image to
|gpu - oracle| <= 1e-5 |oracle| + 1e-6 max|oracle| (both take the same
samples and the same colour bytes), the uchar4 image is the quantised float
image; a colour channel set back to none renders as before, and one that is
not uchar4 is refused."""
import sys

import numpy as np
import pytest

from test_gpu_color import _pair
from test_gpu_render import _ball, _cam_dict, _persp, _rp_dict
from test_gpu_render_modes import _close, _lut

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


MODES = {
    "voxel": dict(thresh=1.0, shadow_amt=0.8),
    "trilinear": dict(thresh=1.0, shadow_amt=0.8),
    "tricubic": dict(thresh=1.0, shadow_amt=0.8, finestep=0.1),
    "levelset": dict(thresh=1.0, shadow_amt=0.0),
    "volume": dict(thresh=0.05, vmin=0.0, vmax=2.0),
    "section3d": dict(thresh=1.0, vmin=0.0, vmax=2.0, slice_pnt=(60.0, 58.0, 57.25), slice_norm=(0.1, 0.2, 1.0)),
}


@pytest.fixture(scope="module")
def coloured():
    pts = _ball(61)
    rng = np.random.default_rng(62)
    clr = rng.integers(0, 256, (len(pts), 4)).astype(np.uint8)
    gv, ov = _pair(pts, clr, 4.0)
    gl = gv.InsertPointsSubcell(4, len(pts), 4.0)
    assert gl == ov.insert_points_subcell(4, pts, 4.0)
    gv.GatherDensity(4, len(pts), 2.0, (0, 0, 0), gl, 0, 1)
    ov.gather_density(2.0)
    ov.gather_color(clr, 1)
    for c, b in ((0, 0.0), (1, 0.0)):
        gv.UpdateApron(c, b)
        ov.update_apron(c, b)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(gv.bricks(0), ov.bricks(0))
    np.testing.assert_array_equal(gv.bricks_u8(1), ov.bricks_u8(1))
    return gv, ov


@pytest.mark.parametrize("mode", list(MODES))
def test_render_colour_vs_oracle(coloured, mode):
    import gvdbx as G
    gv, ov = coloured
    _log(f"[gpu] render colour {mode}")
    sh = {"voxel": G.SHADE_VOXEL, "trilinear": G.SHADE_TRILINEAR, "tricubic": G.SHADE_TRICUBIC,
          "levelset": G.SHADE_LEVELSET, "volume": G.SHADE_VOLUME, "section3d": G.SHADE_SECTION3D}[mode]
    lut = lut_dev = None
    if mode in ("volume", "section3d"):
        lut = _lut()
        lut_dev = torch.from_numpy(lut.reshape(-1)).to(torch.device("cuda:0"))
    kw = dict(step=0.5, light=(60, 200, -80), backclr=(0.1, 0.2, 0.3, 1.0))
    kw.update(MODES[mode])
    rp = G.render_params(lut=lut_dev, **kw)
    cam = _persp(G)
    gv.SetColorChannel(-1)
    plain, _ = gv.Render(sh, 0, cam, rp)
    plain = plain.cpu().numpy()
    gv.SetColorChannel(1)
    assert gv.GetColorChannel() == 1
    rgba, rgba8 = gv.Render(sh, 0, cam, rp)
    torch.cuda.synchronize()
    d = _rp_dict(rp, sh, lut)
    for k in ("finestep", "slice_pnt", "slice_norm"):
        if k in kw:
            d[k] = kw[k]
    d["clr_chan"] = 1
    r32, _ = ov.render(_cam_dict(cam), d)
    g = rgba.cpu().numpy()
    assert (np.abs(g - plain).sum(-1) > 1e-3).sum() > 100, "the colour changes too few pixels"
    _close(g, r32, f"render colour {mode}")
    q = np.clip(g * 255.0, 0, 255).astype(np.float32).astype(np.int32).astype(np.uint8)
    if mode == "section3d":
        q[..., 3] = 255
    np.testing.assert_array_equal(rgba8.cpu().numpy(), q)
    gv.SetColorChannel(255)   # CHAN_UNDEF
    again, _ = gv.Render(sh, 0, cam, rp)
    np.testing.assert_array_equal(again.cpu().numpy(), plain)


def test_colour_channel_must_be_uchar4(coloured):
    import gvdbx as G
    gv, _ = coloured
    _log("[gpu] render colour errors")
    cam = _persp(G)
    rp = G.render_params(step=0.5, thresh=1.0)
    gv.SetColorChannel(0)   # the float density
    with pytest.raises(G.GvdbxError):
        gv.Render(G.SHADE_TRILINEAR, 0, cam, rp)
    gv.SetColorChannel(5)   # no such channel
    with pytest.raises(G.GvdbxError):
        gv.Render(G.SHADE_TRILINEAR, 0, cam, rp)
    with pytest.raises(G.GvdbxError):
        gv.SetColorChannel(300)
    gv.SetColorChannel(-1)
    gv.Render(G.SHADE_TRILINEAR, 0, cam, rp)
