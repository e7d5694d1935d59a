"""Seeded synthetic inputs for the batched workloads (host, numpy).

depth_scene(): frames of a room of random planes and spheres seen from
cameras on an orbit, as a depth sensor would report them:
  - 20 primitives: a floor (y = -1.5 m), a ceiling (y = 2.5 m), 4 walls with
    random horizontal normals at distance U(4.5,6) m from the centre (outside
    the camera orbit), 14 spheres with centres in the ball of radius 2.5 m and
    radii U(0.15,0.8) m (inside the orbit);
  - camera f on a 4 m orbit around the centre, looking at it, angle
    2*pi*f/F plus a seeded jitter, height 0.3 + 0.2 sin(3*angle);
  - pinhole fx = fy = 525, cx = (W-1)/2, cy = (H-1)/2; depth = distance along
    the optical axis of the first hit; no hit or depth outside [dmin, dmax]
    reads as invalid (NaN), and a seeded mask of U(5%,60%) of the pixels per
    frame is invalid as well.
poses are camera-to-world [R row-major (3x3) | t] with camera axes x right,
y down, z forward, the convention gvdbx_backproject_depth expects.

dambreak_frames(): the plain GPU workload, a dam-break block of points
advected by a divergence-free ABC flow, frame by frame on the device.
"""
import math

import numpy as np


def abc_velocity(torch, p, phases, amp=(1.0, 1.0, 1.0)):
    """ABC flow with k = 2 pi / 128: u = A sin(kz+f0) + C cos(ky+f1),
    v = B sin(kx+f2) + A cos(kz+f3), w = C sin(ky+f4) + B cos(kx+f5); with unit
    amplitudes every component is at most 2 voxels per frame."""
    k = 2.0 * math.pi / 128.0
    A, B, Cc = amp
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    u = A * torch.sin(k * z + phases[0]) + Cc * torch.cos(k * y + phases[1])
    v = B * torch.sin(k * x + phases[2]) + A * torch.cos(k * z + phases[3])
    w = Cc * torch.sin(k * y + phases[4]) + B * torch.cos(k * x + phases[5])
    return torch.stack([u, v, w], 1)


def dambreak_frames(torch, gen_box, seed, n, nframes, dev):
    """n points uniform in [64,576)x[64,320)x[64,576) with +-0.5 jitter
    (gen_box(seed, ...)), then nframes - 1 advection steps by the ABC flow
    with seeded phases, on dev.  Returns the frames (device tensors) and the
    frame-0 points on the host."""
    pts = gen_box(seed, n, (64, 64, 64), (576, 320, 576), 0.5)
    rng = np.random.default_rng(seed)
    phases = [float(x) for x in rng.uniform(0, 2 * math.pi, 6)]
    frames = [torch.from_numpy(pts).to(dev)]
    for _ in range(nframes - 1):
        p = frames[-1]
        frames.append((p + abc_velocity(torch, p, phases)).contiguous())
    return frames, pts


def _look_at(pos, target, up=(0.0, 1.0, 0.0)):
    f = np.asarray(target, np.float64) - pos
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    return np.stack([r, d, f], axis=1)   # columns: right, down, forward


def depth_scene(seed, nframes=64, width=640, height=480, fx=525.0, fy=525.0, dmin=0.5, dmax=8.0):
    rng = np.random.default_rng(seed)
    cx, cy = (width - 1) / 2.0, (height - 1) / 2.0
    planes = [(np.array([0.0, 1.0, 0.0]), 1.5), (np.array([0.0, -1.0, 0.0]), 2.5)]   # n.x + dist = 0
    for _ in range(4):
        a = rng.uniform(0, 2 * np.pi)
        planes.append((np.array([np.cos(a), 0.0, np.sin(a)]), rng.uniform(4.5, 6.0)))
    spheres = []
    for _ in range(14):
        c = rng.uniform(-2.5, 2.5, size=3)
        while np.linalg.norm(c) > 2.5:
            c = rng.uniform(-2.5, 2.5, size=3)
        spheres.append((c, rng.uniform(0.15, 0.8)))
    u = np.arange(width, dtype=np.float64)[None, :]
    v = np.arange(height, dtype=np.float64)[:, None]
    dirs_c = np.stack(np.broadcast_arrays((u - cx) / fx, (v - cy) / fy, np.ones((height, width))), -1)
    depth = np.empty((nframes, height, width), np.float32)
    poses = np.empty((nframes, 12), np.float32)
    for f in range(nframes):
        ang = 2 * np.pi * f / nframes + rng.uniform(-0.02, 0.02)
        pos = np.array([4.0 * np.cos(ang), 0.3 + 0.2 * np.sin(3 * ang), 4.0 * np.sin(ang)])
        R = _look_at(pos, (0.0, 0.0, 0.0))
        poses[f, :9] = R.reshape(-1)
        poses[f, 9:] = pos
        dw = dirs_c @ R.T                                      # world direction per unit depth
        best = np.full((height, width), np.inf)
        for n, dist in planes:
            den = dw @ n
            with np.errstate(divide="ignore", invalid="ignore"):
                t = -(pos @ n + dist) / den
            t[~(t > 0)] = np.inf
            best = np.minimum(best, t)
        for c, r in spheres:
            oc = pos - c
            a = np.einsum("hwk,hwk->hw", dw, dw)
            b = 2.0 * (dw @ oc)
            cc = oc @ oc - r * r
            disc = b * b - 4 * a * cc
            with np.errstate(invalid="ignore"):
                sq = np.sqrt(disc)
            t1 = (-b - sq) / (2 * a)
            t2 = (-b + sq) / (2 * a)
            t = np.where(t1 > 0, t1, t2)
            t[~(disc >= 0) | ~(t > 0)] = np.inf
            best = np.minimum(best, t)
        d = best.astype(np.float32)
        d[~((d >= dmin) & (d <= dmax))] = np.nan
        frac = rng.uniform(0.05, 0.60)
        d[rng.random((height, width)) < frac] = np.nan
        depth[f] = d
    return depth, poses, np.array([fx, fy, cx, cy], np.float32)


def cloud_batch(seed, nclouds=256, nmin=10_000, nmax=2_000_000):
    """Independent point clouds, each in its own index space (coordinates
    >= 8): size floor(exp(U(ln nmin, ln nmax))), shape a seeded choice of
      - shell: radius R ~ U(40,120), thickness 8, centred at R + 16;
      - box: sides U(40,200), corner (16,16,16);
      - clusters: 4..16 Gaussians, means U(40,200)^3, sigma U(3,10) per
        axis, truncated at 3 sigma;
    per-point weights U(0.1,1).  Returns a list of (points (n,3) float32,
    weights (n,) float32, shape name)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nclouds):
        n = int(np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        kind = ("shell", "box", "clusters")[int(rng.integers(0, 3))]
        if kind == "shell":
            R = rng.uniform(40.0, 120.0)
            d = rng.normal(size=(n, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            r = np.cbrt(rng.uniform((R - 8.0) ** 3, R ** 3, size=n))
            p = d * r[:, None] + (R + 16.0)
        elif kind == "box":
            side = rng.uniform(40.0, 200.0, size=3)
            p = rng.random((n, 3)) * side + 16.0
        else:
            k = int(rng.integers(4, 17))
            means = rng.uniform(40.0, 200.0, size=(k, 3))
            sig = rng.uniform(3.0, 10.0, size=(k, 3))
            lab = rng.integers(0, k, size=n)
            z = rng.standard_normal((n, 3))
            bad = np.abs(z) > 3
            while bad.any():
                z[bad] = rng.standard_normal(int(bad.sum()))
                bad = np.abs(z) > 3
            p = means[lab] + z * sig[lab]
        w = rng.uniform(0.1, 1.0, size=n).astype(np.float32)
        out.append((p.astype(np.float32), w, kind))
    return out


def cloud_camera_box(points):
    """xy rectangle (lo, hi) of a cloud, padded by 4 voxels, for an
    orthographic +z view."""
    lo = np.floor(points[:, :2].min(0)) - 4.0
    hi = np.ceil(points[:, :2].max(0)) + 4.0
    return lo, hi
