"""CPU pins of the batched-input restatements (back-projection, packing).

Back-projection is pinned by geometry rather than by its own formula: the
points of a frame that sees only a known plane lie on that plane, and
projecting a point back through the pose and intrinsics returns its pixel.
"""
import numpy as np

import oracle as O
from gvdbx import workloads


def _pose_looking_down_z(t=(0.0, 0.0, 0.0)):
    P = np.zeros(12, np.float32)
    P[:9] = np.eye(3, dtype=np.float32).reshape(-1)
    P[9:] = t
    return P


def test_backproject_plane_and_reprojection():
    W, H = 64, 48
    fx = fy = 100.0
    cx, cy = (W - 1) / 2.0, (H - 1) / 2.0
    # camera at (1, 2, 0) looking down +z at the plane z = 3 m: depth = 3 everywhere
    depth = np.full((1, H, W), 3.0, np.float32)
    depth[0, 5, 7] = np.nan
    depth[0, 6, 8] = 9.0          # beyond dmax
    pose = _pose_looking_down_z((1.0, 2.0, 0.0))[None]
    origin, vox = (-4.0, -4.0, -1.0), 0.02
    pts = O.backproject_depth(depth, (fx, fy, cx, cy), pose, vox, origin, 0.5, 8.0).reshape(H, W, 3)
    assert pts[5, 7, 2] == np.float32(1e10) and pts[6, 8, 2] == np.float32(1e10)
    ok = np.ones((H, W), bool)
    ok[5, 7] = ok[6, 8] = False
    world = pts.astype(np.float64) * vox + np.array(origin)
    np.testing.assert_allclose(world[ok][:, 2], 3.0, atol=1e-5)
    u = (world[..., 0] - 1.0) / 3.0 * fx + cx
    v = (world[..., 1] - 2.0) / 3.0 * fy + cy
    uu, vv = np.meshgrid(np.arange(W), np.arange(H))
    np.testing.assert_allclose(u[ok], uu[ok], atol=1e-3)
    np.testing.assert_allclose(v[ok], vv[ok], atol=1e-3)


def test_backproject_rotated_pose_reprojects():
    d, poses, intr = workloads.depth_scene(7, nframes=3, width=80, height=60, fx=65.625, fy=65.625)
    origin, vox = (-8.0, -8.0, -8.0), 0.02
    pts = O.backproject_depth(d, intr, poses, vox, origin).reshape(3, 60, 80, 3).astype(np.float64)
    fx, fy, cx, cy = intr
    for f in range(3):
        R = poses[f, :9].reshape(3, 3).astype(np.float64)
        t = poses[f, 9:].astype(np.float64)
        ok = np.isfinite(d[f]) & (d[f] >= 0.5) & (d[f] <= 8.0)
        assert ok.sum() > 100
        world = pts[f] * vox + np.array(origin)
        cam = (world - t) @ R          # R^T (w - t)
        np.testing.assert_allclose(cam[ok][:, 2], d[f][ok], rtol=2e-5, atol=2e-5)
        uu, vv = np.meshgrid(np.arange(80), np.arange(60))
        np.testing.assert_allclose(cam[ok][:, 0] / cam[ok][:, 2] * fx + cx, uu[ok], atol=2e-3)
        np.testing.assert_allclose(cam[ok][:, 1] / cam[ok][:, 2] * fy + cy, vv[ok], atol=2e-3)


def test_depth_scene_is_seeded():
    a = workloads.depth_scene(3, nframes=2, width=40, height=30, fx=32.8, fy=32.8)
    b = workloads.depth_scene(3, nframes=2, width=40, height=30, fx=32.8, fy=32.8)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
