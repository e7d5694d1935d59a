"""Pins the oracle's colour channel in Render (SetColorChannel VG.cpp:686-690)
on the CPU.

The reference takes the colour of a uchar4 channel at a surface hit
(getColorF cuda_gvdb_raycast.cuh:206-209, each byte / 255 by make_float4(uchar4)
cuda_math.cuh:221-224) and shades it (performPhongShading cuda_gvdb_module.cu:
38-57: colour * (diffuse + ambient)); in SHADE_VOLUME every sample's
contribution is multiplied by the colour at the sample (rayDeepBrick
cuda_gvdb_raycast.cuh:519-522); SHADE_SECTION3D multiplies the slice colour
and the surface colour.  What fixes it independently of the oracle's code:
  - a white colour channel (bytes 255: exactly 1.0) leaves every image as it
    is without one, bit for bit;
  - a constant colour c scales each surface pixel to fl(c/255) * the plain
    pixel exactly, and the volume image (black background) to c/255 times it;
  - a colour channel that encodes each voxel's coordinates returns, at every
    hit, the voxel brute force finds: for SHADE_VOXEL the first voxel of the
    pixel's column above the threshold, for SHADE_TRILINEAR the voxel holding
    the hit Raytrace returns for the same ray.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_render import _phong, _surface_setup, column_leaves, ortho

BACK = (0.1, 0.2, 0.3, 0.4)
LIGHT = (48.0, 40.0, -60.0)


def _lut():
    t = np.linspace(0, 1, 16384, dtype=np.float32)
    return np.stack([t, t * t, 1 - t, 0.1 + 0.8 * t], 1).astype(np.float32)


MODES = {
    "voxel": dict(shade=O.SHADE_VOXEL, step=0.5, thresh=0.5, shadow_amt=0.8),
    "trilinear": dict(shade=O.SHADE_TRILINEAR, step=0.5, thresh=0.5, shadow_amt=0.8),
    "tricubic": dict(shade=O.SHADE_TRICUBIC, step=0.5, thresh=0.5, shadow_amt=0.8, finestep=0.1),
    "levelset": dict(shade=O.SHADE_LEVELSET, step=0.5, thresh=0.5, shadow_amt=0.8),
    "volume": dict(shade=O.SHADE_VOLUME, step=0.5, thresh=0.05, vmin=0.0, vmax=2.0, backclr=(0, 0, 0, 0)),
    "section3d": dict(shade=O.SHADE_SECTION3D, step=0.5, thresh=0.5, vmin=0.0, vmax=2.0,
                      slice_pnt=(48.0, 48.0, 47.25), slice_norm=(0.1, 0.2, 1.0)),
}
SURF = ("voxel", "trilinear", "tricubic", "levelset")


@pytest.fixture(scope="module")
def vol():
    pts = _surface_setup()
    v = O.OracleVolume()
    v.add_channel(0, 1)
    v.add_channel(1, 1, O.T_UCHAR4)
    v.rebuild_topology(pts, 4.0)
    v.update_atlas()
    v.splat_cubic(pts, 2.0)
    v.update_apron(0, 0.0)
    return v


def set_colour(v, fn):
    """Colour channel 1 := fn(X, Y, Z) (world voxel coordinates of every padded
    voxel) as uint8 (..., 4)."""
    nodes = v.nodes(0)
    br = 10
    idx = np.arange(br) - 1
    zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
    b = np.empty((len(nodes), br, br, br, 4), np.uint8)
    for k, m in enumerate(nodes["mPos"].astype(np.int64)):
        b[k] = fn(m[0] + xx, m[1] + yy, m[2] + zz)
    v.set_bricks(b.view(np.float32).reshape(b.shape[:4]), chan=1)


def _render(v, mode, clr):
    rp = dict(dict(light=LIGHT, backclr=BACK), **MODES[mode])
    if mode in ("volume", "section3d"):
        rp["lut"] = _lut()
    if clr:
        rp["clr_chan"] = 1
    return v.render(ortho(16, 16, 36.0, 36.0, 24, z0=-0.3), rp)[0]


def _const(c):
    return lambda X, Y, Z: np.broadcast_to(np.uint8(c), X.shape + (4,))


@pytest.mark.parametrize("mode", list(MODES))
def test_white_colour_changes_nothing(vol, mode):
    set_colour(vol, _const((255, 255, 255, 63)))
    plain, white = _render(vol, mode, False), _render(vol, mode, True)
    assert (plain[..., 3] != (0 if mode == "volume" else BACK[3])).sum() > 20, "view misses the object"
    k = 3 if mode == "section3d" else 4   # the surface colour's alpha enters the section's float alpha
    np.testing.assert_array_equal(white[..., :k], plain[..., :k])


@pytest.mark.parametrize("mode", SURF + ("volume",))
def test_constant_colour_scales(vol, mode):
    c = (51, 102, 204, 63)
    set_colour(vol, _const(c))
    plain, got = _render(vol, mode, False), _render(vol, mode, True)
    h = np.float32(c[:3]) / np.float32(255)
    if mode == "volume":
        ok = plain[..., :3] < 0.99 * plain[..., 3:]   # the pixel is alpha * colour: colour below the clamp
        np.testing.assert_allclose(got[..., :3][ok], (h * plain[..., :3])[ok], rtol=1e-5, atol=1e-7)
        np.testing.assert_array_equal(got[..., 3], plain[..., 3])
        return
    hit = plain[..., 3] == 1.0
    assert hit.sum() > 20
    np.testing.assert_array_equal(got[hit][:, :3], h * plain[hit][:, :3])
    np.testing.assert_array_equal(got[~hit], plain[~hit])
    assert (got[..., 3] == plain[..., 3]).all()


def _coords(X, Y, Z):
    return np.stack([X & 255, Y & 255, Z & 255, np.full_like(X, 63)], -1).astype(np.uint8)


def test_voxel_colour_is_the_hit_voxel(vol):
    set_colour(vol, _coords)
    br, nodes = vol.bricks(0), vol.nodes(0)
    rp = dict(shade=O.SHADE_VOXEL, step=0.5, thresh=0.5, light=LIGHT, shadow_amt=0.0, backclr=BACK, clr_chan=1)
    got = vol.render(ortho(8, 8, 36.0, 36.0, 24, z0=-0.3), rp)[0]
    hits = 0
    for px in range(64):
        x = 36 + ((px % 8) + 0.5) * 3
        y = 36 + ((px // 8) + 0.5) * 3
        vox = None
        for L in column_leaves(nodes, x, y):
            i, j = int(x - nodes["mPos"][L, 0]), int(y - nodes["mPos"][L, 1])
            for k in range(8):
                if br[L, k + 1, j + 1, i + 1] > 0.5:
                    vox = nodes["mPos"][L].astype(np.int64) + (i, j, k)
                    break
            if vox is not None:
                break
        if vox is None:
            continue
        hits += 1
        h = np.float32(vox & 255) / np.float32(255)
        sh = np.float32(_phong(np.array([0, 0, -1.0]), np.array([x, y, vox[2]], np.float64), LIGHT))
        np.testing.assert_allclose(got[px // 8, px % 8, :3], h * sh, rtol=1e-5)
    assert hits > 10


def test_trilinear_colour_is_the_raytrace_hit_voxel(vol):
    set_colour(vol, _coords)
    W = H = 16
    cam = ortho(W, H, 36.0, 36.0, 24, z0=-0.3)
    base = dict(shade=O.SHADE_TRILINEAR, step=0.5, thresh=0.5, light=LIGHT, shadow_amt=0.0, backclr=BACK)
    plain = vol.render(cam, base)[0]
    got = vol.render(cam, dict(base, clr_chan=1))[0]
    # the camera's rays (camera_ray's FP32 operations), traced with Raytrace
    f = np.float32
    xs, ys = np.meshgrid(np.arange(W, dtype=f), np.arange(H, dtype=f))
    u = (xs + f(0.5)) / f(W)
    w = (ys + f(0.5)) / f(H)
    rays = np.zeros(W * H, dtype=O.RAY_DTYPE)
    org, ou, ov = [np.asarray(cam[k], f) for k in ("org", "org_u", "org_v")]
    for a in range(3):
        rays["orig"][:, a] = ((u * ou[a] + w * ov[a]) + org[a]).reshape(-1)
    rays["dir"] = (0, 0, 1)
    vol.raytrace(rays, base, bias=0.0)
    hitm = rays["hit"][:, 2] != np.float32(1e10)
    assert hitm.sum() > 50
    np.testing.assert_array_equal(hitm, (plain[..., 3] == 1.0).reshape(-1))
    vox = np.floor(rays["hit"][hitm]).astype(np.int64)
    h = (vox & 255).astype(np.float32) / np.float32(255)
    np.testing.assert_array_equal(got.reshape(-1, 4)[hitm, :3], h * plain.reshape(-1, 4)[hitm, :3])


def test_colour_channel_must_be_uchar4(vol):
    rp = dict(MODES["trilinear"], clr_chan=0)   # a float channel
    with pytest.raises(Exception):
        vol.render(ortho(4, 4, 36.0, 36.0, 24), rp)
