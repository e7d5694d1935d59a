"""Pins the oracle's remaining Render modes and Raytrace (CPU only).

What fixes each one independently of the oracle's code:
  - SHADE_TRICUBIC: on a linear field the 27-tap filter of getTricubic
    (cuda_gvdb_raycast.cuh:32-98) is itself linear and, its taps sitting on
    integer points with weights ta^2, 2 ta tb, tb^2 (tb = frac/2 + 1/4),
    equals the field half a voxel back on every axis: sum w_i (floor(p) - 1
    + i) = p - 1/2.  The hit is then the exact plane crossing of that
    shifted field and the normal the negated plane normal (closed form);
  - SHADE_VOXEL: the first voxel of the pixel's column above the threshold,
    entered through its -z face with normal (0, 0, -1) (brute force over the
    leaf list);
  - SHADE_EMPTYSKIP: the entry point of the first leaf of the column, / 100;
  - SHADE_SECTION2D / 3D: the slice point's leaf found by brute force, the
    sample by scipy map_coordinates, the transfer colour by the index rule of
    transfer (cuda_gvdb_dda.cuh:20-23), and for 3D the first trilinear
    surface sample beyond the plane (brute force);
  - SHADE_OFF clears; Raytrace returns the trilinear brute-force hits pulled
    back by the bias and leaves a missed ray's normal as it was.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_render import _normal, _phong, _surface_setup, column_leaves, ortho, tri_scipy, vol_for

NOHIT = np.float32(1.0e10)


def _block_volume():
    rng = np.random.default_rng(3)
    pts = (rng.random((30000, 3)) * 72 + 12).astype(np.float32)   # leaves cover [8, 88)^3
    v = vol_for(pts, 0.0)
    return v


def _plant(v, G):
    """every leaf's padded voxel (i, j, k) := G(mPos + (i, j, k) - 1/2), so the
    trilinear sample at world point W is G(W)"""
    nodes = v.nodes(0)
    br = 10
    idx = np.arange(br, dtype=np.float64) - 0.5
    zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
    b = np.empty((len(nodes), br, br, br), np.float32)
    for k, m in enumerate(nodes["mPos"].astype(np.float64)):
        b[k] = G(m[0] + xx, m[1] + yy, m[2] + zz)
    v.set_bricks(b)
    return b


def test_tricubic_linear_field_closed_form():
    v = _block_volume()
    a = np.array([0.3, 0.2, 1.0])
    _plant(v, lambda x, y, z: a[0] * x + a[1] * y + a[2] * z - 60.0)
    cam = ortho(8, 8, 24.0, 24.0, 48, z0=-0.3)
    light = (48.0, 40.0, 10.0)
    rp = dict(shade=O.SHADE_TRICUBIC, step=0.5, finestep=0.1, thresh=0.0, light=light, shadow_amt=0.0,
              backclr=(0, 0, 0, 0))
    rgba32, _ = v.render(cam, rp)
    n = -a / np.linalg.norm(a)
    checked = 0
    for py in range(8):
        for px in range(8):
            x, y = 24 + (px + 0.5) * 6, 24 + (py + 0.5) * 6
            # G(W - 1/2) = 0 along the +z ray
            z = 60.0 - a[0] * (x - 0.5) - a[1] * (y - 0.5) + 0.5 * a[2]
            lx, ly, lz = x % 8, y % 8, z % 8
            if not (1.5 <= lx < 7.5 and 1.5 <= ly < 7.5 and 2 <= lz <= 6):
                continue   # taps (also the normal's, half a voxel off) would reach past the padded brick
            sv = _phong(n, np.array([x, y, z]), light)
            np.testing.assert_allclose(rgba32[py, px, :3], sv, rtol=2e-5)
            checked += 1
    assert checked >= 10
    # the unshifted crossing would shade visibly differently
    x, y = 24 + 0.5 * 6, 24 + 0.5 * 6
    z0 = 60.0 - a[0] * x - a[1] * y
    assert abs(_phong(n, np.array([x, y, z0]), light) - _phong(n, np.array([x, y, z0 + 0.75]), light)) > 1e-3


def test_voxel_surface_brute_force():
    pts = _surface_setup()
    v = vol_for(pts, 4.0)
    v.splat_cubic(pts, 2.0)
    v.update_apron(0, 0.0)
    br, nodes = v.bricks(0), v.nodes(0)
    cam = ortho(8, 8, 36.0, 36.0, 24, z0=-0.3)
    light = (48, 40, -60)
    rp = dict(shade=O.SHADE_VOXEL, step=0.5, thresh=0.5, light=light, shadow_amt=0.0, backclr=(0.1, 0.2, 0.3, 0.4))
    rgba32, _ = v.render(cam, rp)
    hits = 0
    for px in range(64):
        x = 36 + ((px % 8) + 0.5) * 3
        y = 36 + ((px // 8) + 0.5) * 3
        hit = None
        for L in column_leaves(nodes, x, y):
            i, j = int(x - nodes["mPos"][L, 0]), int(y - nodes["mPos"][L, 1])
            for k in range(8):
                if br[L, k + 1, j + 1, i + 1] > 0.5:
                    hit = np.array([x, y, nodes["mPos"][L, 2] + k], np.float64)
                    break
            if hit is not None:
                break
        got = rgba32[px // 8, px % 8]
        if hit is None:
            np.testing.assert_array_equal(got, np.float32([0.1, 0.2, 0.3, 0.4]))
            continue
        hits += 1
        np.testing.assert_allclose(got[:3], _phong(np.array([0, 0, -1.0]), hit, light), rtol=1e-5)
        assert got[3] == 1.0
    assert hits > 10


def test_empty_skip_first_leaf():
    pts = _surface_setup()
    v = vol_for(pts, 2.0)
    nodes = v.nodes(0)
    cam = ortho(8, 8, 30.0, 30.0, 36, z0=-0.3)
    rp = dict(shade=O.SHADE_EMPTYSKIP, backclr=(0.5, 0.25, 0.125, 0.0))
    rgba32, _ = v.render(cam, rp)
    hits = 0
    for px in range(64):
        x = 30 + ((px % 8) + 0.5) * 4.5
        y = 30 + ((px // 8) + 0.5) * 4.5
        col = column_leaves(nodes, x, y)
        got = rgba32[px // 8, px % 8]
        if len(col) == 0:
            np.testing.assert_array_equal(got, np.float32([0.5, 0.25, 0.125, 0.0]))
            continue
        hits += 1
        z = nodes["mPos"][col[0], 2]
        np.testing.assert_allclose(got, [x * 0.01, y * 0.01, z * 0.01, 1.0], rtol=0, atol=5e-5)
    assert hits > 10


def _lut():
    t = np.arange(16384, dtype=np.float64) / 16384
    return np.stack([t, 1 - t, 0.5 + 0 * t, 0.25 + 0.5 * t], 1).astype(np.float32)


def _transfer(lut, s, thresh, vmin, vmax):
    q = np.float32((np.float32(s) - np.float32(thresh)) / (np.float32(vmax) - np.float32(vmin)))
    return lut[int(min(1.0, max(0.0, float(q))) * 16300.0)]


def _leaf_at(nodes, w):
    m = nodes["mPos"]
    sel = np.nonzero(((m <= w) & (w < m + 8)).all(1))[0]
    return int(sel[0]) if len(sel) else None


def test_section2d_brute_force():
    pts = _surface_setup()
    v = vol_for(pts, 4.0)
    v.splat_cubic(pts, 2.0)
    v.update_apron(0, 0.0)
    br, nodes = v.bricks(0), v.nodes(0)
    lut = _lut()
    W = H = 16
    cam = dict(width=W, height=H, org=(0, 0, 0), org_u=(1, 0, 0), org_v=(0, 1, 0), dir=(0, 0, 1))
    rp = dict(shade=O.SHADE_SECTION2D, thresh=0.0, vmin=0.0, vmax=4.0, lut=lut, slice_pnt=(48.0, 47.3, 48.0),
              slice_norm=(14.0, 1.0, 14.0))
    rgba32, _ = v.render(cam, rp)
    inside = 0
    for y in range(H):
        for x in range(W):
            sx = np.float32(x * 2.0 / W - 1.0)
            sz = np.float32(y * 2.0 / H - 1.0)
            w = np.array([np.float32(48.0) + sx * np.float32(14.0), np.float32(47.3), np.float32(48.0) + sz * np.float32(14.0)],
                         np.float32)
            L = _leaf_at(nodes, w)
            got = rgba32[y, x]
            if L is None:
                np.testing.assert_array_equal(got, [0, 0, 0, 1])
                continue
            inside += 1
            s = tri_scipy(br[L], w.astype(np.float64) - nodes["mPos"][L])
            c = _transfer(lut, s, 0.0, 0.0, 4.0)
            np.testing.assert_allclose(got[:3], c[3] * c[:3], rtol=0, atol=3e-4)
    assert inside > 30


def test_section3d_brute_force():
    pts = _surface_setup()
    v = vol_for(pts, 4.0)
    v.splat_cubic(pts, 2.0)
    v.update_apron(0, 0.0)
    br, nodes = v.bricks(0), v.nodes(0)
    lut = _lut()
    cam = ortho(8, 8, 36.0, 36.0, 24, z0=-0.3)
    light = (48, 40, 100)
    thresh = 0.5
    rp = dict(shade=O.SHADE_SECTION3D, step=0.5, thresh=thresh, vmin=0.0, vmax=4.0, lut=lut, light=light,
              slice_pnt=(0.0, 0.0, 44.25), slice_norm=(0.0, 0.0, 1.0), backclr=(0.2, 0.3, 0.4, 1.0))
    rgba32, _ = v.render(cam, rp)
    surf = 0
    for px in range(64):
        x = 36 + ((px % 8) + 0.5) * 3
        y = 36 + ((px // 8) + 0.5) * 3
        w = np.array([x, y, 44.25])
        L = _leaf_at(nodes, w.astype(np.float32))
        clr = np.array([1.0, 1, 1, 0])
        t = 0.0
        if L is not None:
            t = tri_scipy(br[L], w - nodes["mPos"][L])
            clr = _transfer(lut, t, thresh, 0.0, 4.0).astype(np.float64)
        hit = None
        for z in 44.25 + 0.5 * np.arange(1, 80):
            L2 = _leaf_at(nodes, np.array([x, y, z], np.float32))
            if L2 is None:
                continue
            p = np.array([x, y, z]) - nodes["mPos"][L2]
            if tri_scipy(br[L2], p) >= thresh:
                hit = (L2, p)
                break
        got = rgba32[px // 8, px % 8]
        if hit is None:
            want = np.array([0.2, 0.3, 0.4, 1.0]) + clr[3] * (clr - np.array([0.2, 0.3, 0.4, 1.0]))
        else:
            surf += 1
            L2, p = hit
            hw = p + nodes["mPos"][L2]
            n = _normal(br, L2, p, -1)
            l = np.asarray(light, np.float64) - hw
            l /= np.linalg.norm(l)
            ds = 1.0 if t > thresh else 0.8 * max(0.0, float(n @ l))
            want = ds + clr[3] * (clr - ds)
        np.testing.assert_allclose(got, want, rtol=0, atol=5e-4)
    assert surf > 10


def test_shade_off_clears():
    pts = _surface_setup()
    v = vol_for(pts, 2.0)
    rgba32, rgba64 = v.render(ortho(4, 4, 30, 30, 40), dict(shade=O.SHADE_OFF))
    assert not rgba32.any() and not rgba64.any()
    with pytest.raises(O.OracleError):
        v.render(ortho(4, 4, 30, 30, 40), dict(shade=O.SHADE_TRICUBIC, finestep=0.0))
    with pytest.raises(O.OracleError):
        v.render(ortho(4, 4, 30, 30, 40), dict(shade=O.SHADE_SECTION2D))   # no transfer function


def test_raytrace_brute_force():
    pts = _surface_setup()
    v = vol_for(pts, 4.0)
    v.splat_cubic(pts, 2.0)
    v.update_apron(0, 0.0)
    br, nodes = v.bricks(0), v.nodes(0)
    xs = 36 + (np.arange(64) % 8 + 0.5) * 3
    ys = 36 + (np.arange(64) // 8 + 0.5) * 3
    rays = np.zeros(64, dtype=O.RAY_DTYPE)
    rays["orig"] = np.stack([xs, ys, np.full(64, -0.3)], 1)
    rays["dir"] = (0, 0, 1)
    rays["normal"] = (9, 9, 9)
    v.raytrace(rays, dict(step=0.5, thresh=0.5), bias=0.25)
    hits = 0
    for r in range(64):
        x, y = xs[r], ys[r]
        hit = None
        for L in column_leaves(nodes, x, y):
            for z in 0.2 + 0.5 * np.arange(16):
                p = (x - nodes["mPos"][L, 0], y - nodes["mPos"][L, 1], z)
                if tri_scipy(br[L], p) >= 0.5:
                    hit = (L, p)
                    break
            if hit:
                break
        if hit is None:
            assert (rays["hit"][r] == NOHIT).all() and (rays["normal"][r] == 9).all()
            continue
        hits += 1
        L, p = hit
        hw = np.array(p, np.float64) + nodes["mPos"][L]
        np.testing.assert_allclose(rays["hit"][r], hw - [0, 0, 0.25], rtol=0, atol=1e-4)
        np.testing.assert_allclose(rays["normal"][r], _normal(br, L, p, -1), rtol=0, atol=1e-5)
    assert hits > 10
