"""GPU parity of the remaining Render modes (SHADE_VOXEL, TRICUBIC,
EMPTYSKIP, SECTION2D, SECTION3D, OFF) and Raytrace against the oracle,
through the C-ABI, on fields built with the gathers (bricks bit-exact with
the oracle first).  The float image agrees with the oracle's FP32 image to
|gpu - oracle| <= 1e-5 |oracle| + 1e-6 max|oracle| (both march the same
samples; the normalisations may differ by an ulp); the uchar4 image is the
quantised float image (alpha 255 where the reference writes 255).  Raytrace
hits and normals agree to 1e-5 relative."""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_render import _ball, _cam_dict, _pair, _persp, _rp_dict

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


def _density_pair(seed):
    pts = _ball(seed)
    gv, ov, tp, _ = _pair(pts, 4.0)
    ln = gv.InsertPointsSubcell(4, len(pts), 4.0)
    gv.GatherDensity(4, len(pts), 2.0, (0, 0, 0), ln, 0)
    gv.UpdateApron(0, 0.0)
    ov.insert_points_subcell(4, pts, 4.0)
    ov.gather_density(2.0)
    ov.update_apron(0, 0.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(gv.bricks(0), ov.bricks(0))
    return gv, ov, tp


def _lut():
    t = np.linspace(0, 1, 16384, dtype=np.float32)
    return np.stack([t, t * t, 1 - t, 0.1 + 0.8 * t], 1).astype(np.float32)


def _close(gpu, ref, what):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    tol = 1e-5 * np.abs(ref) + 1e-6 * max(np.abs(ref).max(), 1e-30)
    bad = np.abs(gpu - ref) > tol
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} values off, max err {np.abs(gpu - ref).max():.3e}"


MODES = {
    "voxel": dict(thresh=1.0, shadow_amt=0.8),
    "tricubic": dict(thresh=1.0, shadow_amt=0.8, finestep=0.1),
    "emptyskip": dict(),
    "section2d": dict(thresh=0.05, vmin=0.0, vmax=2.0, slice_pnt=(60.0, 58.3, 62.0), slice_norm=(24.0, 1.0, 20.0)),
    "section3d": dict(thresh=1.0, vmin=0.0, vmax=2.0, slice_pnt=(60.0, 58.0, 57.25), slice_norm=(0.1, 0.2, 1.0)),
    "off": dict(),
}


@pytest.mark.parametrize("mode", list(MODES))
def test_render_mode_vs_oracle(mode):
    G = _G()
    _log(f"[gpu] render {mode}")
    gv, ov, tp = _density_pair(31 + len(mode))
    sh = {"voxel": G.SHADE_VOXEL, "tricubic": G.SHADE_TRICUBIC, "emptyskip": G.SHADE_EMPTYSKIP,
          "section2d": G.SHADE_SECTION2D, "section3d": G.SHADE_SECTION3D, "off": G.SHADE_OFF}[mode]
    lut = lut_dev = None
    if mode.startswith("section"):
        lut = _lut()
        lut_dev = torch.from_numpy(lut.reshape(-1)).to(_dev())
    kw = dict(step=0.5, light=(60, 200, -80), backclr=(0.1, 0.2, 0.3, 1.0))
    kw.update(MODES[mode])
    rp = G.render_params(lut=lut_dev, **kw)
    cam = _persp(G)
    rgba, rgba8 = gv.Render(sh, 0, cam, rp)
    torch.cuda.synchronize()
    d = _rp_dict(rp, sh, lut)
    for k in ("finestep", "slice_pnt", "slice_norm"):
        if k in kw:
            d[k] = kw[k]
    r32, _ = ov.render(_cam_dict(cam), d)
    g = rgba.cpu().numpy()
    if mode == "off":
        assert not g.any() and not rgba8.cpu().numpy().any()
        return
    assert (np.abs(r32 - np.float32(kw["backclr"])).sum(-1) > 1e-3).sum() > 100, "view misses the object"
    _close(g, r32, f"render {mode}")
    q = np.clip(g * 255.0, 0, 255).astype(np.float32).astype(np.int32).astype(np.uint8)
    if mode in ("emptyskip", "section2d", "section3d"):
        q[..., 3] = 255
    np.testing.assert_array_equal(rgba8.cpu().numpy(), q)


def test_raytrace_vs_oracle():
    G = _G()
    _log("[gpu] raytrace")
    gv, ov, tp = _density_pair(77)
    rng = np.random.default_rng(78)
    n = 4096
    rays = np.zeros(n, dtype=O.RAY_DTYPE)
    rays["orig"] = rng.uniform(20, 100, (n, 3)) * [1, 1, 0] + [0, 0, -20]
    d = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n), np.ones(n)], 1)
    rays["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["normal"] = (7, 8, 9)
    rp = G.render_params(step=0.5, thresh=1.0)
    dev_rays = torch.from_numpy(rays.view(np.uint8).copy()).to(_dev())
    gv.Raytrace(dev_rays, 0, rp, bias=0.25)
    torch.cuda.synchronize()
    got = dev_rays.cpu().numpy().view(O.RAY_DTYPE)
    ov.raytrace(rays, _rp_dict(rp), bias=0.25)
    hit = rays["hit"][:, 2] != np.float32(1e10)
    assert hit.sum() > 300 and (~hit).sum() > 300
    np.testing.assert_array_equal(got["hit"][:, 2] != np.float32(1e10), hit)
    np.testing.assert_allclose(got["hit"][hit], rays["hit"][hit], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(got["normal"][hit], rays["normal"][hit], rtol=1e-5, atol=1e-5)
    np.testing.assert_array_equal(got["normal"][~hit], rays["normal"][~hit])
    np.testing.assert_array_equal(got["orig"], rays["orig"])
