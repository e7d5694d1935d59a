"""GPU point gradients against FP64 central finite differences.

A config-3-shaped cluster set (64 anisotropic Gaussian clusters, weights
U(0.1, 1), cubic splat r = 2, UpdateApron(0), orthographic EA raycast with
step 0.5 and kappa 0.05) goes through the GPU forward and backward
(SplatCubic -> UpdateApron -> RaycastEA -> RaycastEABackward ->
ApronAdjoint -> SplatCubicBackward) for L = <gI, I>.  For 1000 sampled
points, dL/dp (three axes) and dL/dw are compared with central differences
of the FP64 chain (the oracle's splat formula, apron and EA raycast).

A perturbation of point i changes only the voxels within r of it (and
their apron copies), and a trilinear sample at s reads voxels
floor(s - 1/2) and floor(s - 1/2) + 1 per axis, so only rays whose x and y
lie in [p - r - 1/2, p + r + 3/2) change: an 8 x 8 pixel window (1 pixel per
voxel, one pixel of margin each side) holds all of them, and the loss
difference is taken over that window alone.  Window rays are the full
camera's rays exactly (dyadic origins and extents).

Tolerance per entry: |gpu - fd| <= 1e-5 |fd| + n 2^-24 M + 1e-7 max|fd|,
with M the oracle's magnitude adjoint and n the splat's rounding chain
(tests/tolerance.py: EA chain + apron + splat), the last term the finite differences' own accuracy,
which the test pins against the oracle's analytic FP64 adjoint.
"""
import sys

import numpy as np
import pytest

import oracle as O
from test_gpu_render import _cam_dict, _rp_dict
from tolerance import APRON_OPS, U32, ea_chain_ops, splat_ops

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

LO, EXT = 32, 128          # image covers [LO, LO + EXT)^2 in x, y at 1 pixel per voxel
R = 2.0                    # splat radius


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _clusters(seed, n):
    rng = np.random.default_rng(seed)
    k = 64
    means = rng.uniform(LO + 24, LO + EXT - 24, (k, 3))
    sig = rng.uniform(2.0, 10.0, (k, 3))
    c = rng.integers(0, k, n)
    pts = means[c] + np.clip(rng.normal(size=(n, 3)), -3, 3) * sig[c]
    return pts.astype(np.float32), rng.uniform(0.1, 1.0, n).astype(np.float32), rng


def _contrib(p, w, lm, pos):
    """The point's FP64 splat values (or_splat_cubic_f64's formula) at every
    padded-brick entry holding them: the voxel in its leaf and its apron
    copies in the neighbour leaves.  Returns (index tuple, values)."""
    lo = np.floor(p - R).astype(np.int64)
    hi = np.ceil(p + R).astype(np.int64)
    L, Z, Y, X, V = [], [], [], [], []
    for z in range(lo[2], hi[2] + 1):
        for y in range(lo[1], hi[1] + 1):
            for x in range(lo[0], hi[0] + 1):
                dx, dy, dz = x - p[0], y - p[1], z - p[2]
                d2 = dx * dx + dy * dy + dz * dz
                if not d2 < R * R:
                    continue
                c = (x >> 3, y >> 3, z >> 3)
                if c not in lm:
                    continue
                t = 1.0 - d2 / (R * R)
                val = w * t * t * t
                for oz in (-1, 0, 1):
                    for oy in (-1, 0, 1):
                        for ox in (-1, 0, 1):
                            nb = lm.get((c[0] + ox, c[1] + oy, c[2] + oz))
                            if nb is None:
                                continue
                            lx, ly, lz = x - pos[nb, 0], y - pos[nb, 1], z - pos[nb, 2]
                            if -1 <= lx <= 8 and -1 <= ly <= 8 and -1 <= lz <= 8:
                                L.append(nb), Z.append(lz + 1), Y.append(ly + 1), X.append(lx + 1), V.append(val)
    ix = tuple(np.array(a, dtype=np.int64) for a in (L, Z, Y, X))
    return ix, np.array(V, dtype=np.float64)


def test_point_gradients_vs_fp64_finite_differences():
    import gvdbx as G
    dev = torch.device("cuda:0")
    _log("[gpu] point gradients vs FP64 finite differences")
    pts, w, rng = _clusters(20261207, 1 << 17)
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.AddChannel(0, G.T_FLOAT, 1)
    tp = torch.from_numpy(pts).to(dev)
    tw = torch.from_numpy(w).to(dev)
    gv.SetPoints(tp)
    gv.SetPointWeights(tw)
    gv.RebuildTopology(len(pts), 2 * R)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    gv.SplatCubic(len(pts), R)
    gv.UpdateApron(0, 0.0)
    cam = G.camera_ortho(EXT, EXT, (float(LO), float(LO), 0.0), (float(EXT), 0, 0), (0, float(EXT), 0), (0, 0, 1))
    rp = G.ea_params(step=0.5, kappa=0.05)
    I, T = gv.RaycastEA(0, cam, rp, record=True)
    gI = rng.normal(size=(EXT, EXT)).astype(np.float32).astype(np.float64)
    grad = gv.RaycastEABackward(0, cam, rp, I, T, torch.from_numpy(gI.astype(np.float32)).to(dev))
    gv.ApronAdjoint(grad)
    dpos, dw = gv.SplatCubicBackward(len(pts), R, (0, 0, 0), grad)
    torch.cuda.synchronize()
    dpos, dw = dpos.cpu().numpy().astype(np.float64), dw.cpu().numpy().astype(np.float64)

    # FP64 chain of the oracle: bricks, analytic adjoint and magnitude adjoint
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, 2 * R)
    ov.update_atlas()
    p64, w64 = pts.astype(np.float64), w.astype(np.float64)
    b = ov.apron_f64(ov.splat_cubic_f64_padded(p64, R, w64), 0.0)
    ocam, EA = _cam_dict(cam), _rp_dict(rp)
    _, _, I64, _ = ov.raycast_ea(ocam, EA, bricks64=b)
    assert (I64 > 1e-3).mean() > 0.2
    g64 = ov.apron_adjoint_f64(ov.raycast_ea_backward(ocam, EA, gI, bricks64=b))
    m64 = ov.apron_adjoint_f64(ov.raycast_ea_backward(ocam, EA, gI, bricks64=b, mag=True))
    an_pos, an_w = ov.splat_cubic_backward_f64(p64, R, g64, w64)
    mg_pos, mg_w = ov.splat_cubic_backward_f64(p64, R, m64, w64, mag=True)

    nodes = ov.nodes(0)["mPos"].astype(np.int64)
    lm = {(int(x) >> 3, int(y) >> 3, int(z) >> 3): k for k, (x, y, z) in enumerate(nodes)}
    inside = ((p64[:, :2] >= LO + 8) & (p64[:, :2] < LO + EXT - 8)).all(1)
    sample = rng.choice(np.nonzero(inside)[0], 1000, replace=False)
    h = 1e-4
    fd = np.zeros((len(sample), 4))

    def window_loss(p, wt, x0, y0):
        """<gI, I> over the 8 x 8 window with point i at (p, wt)."""
        ix, vals = _contrib(p, wt, lm, nodes)
        saved = b[ix].copy()
        b[ix] += vals                     # entries are distinct: one per (voxel, leaf)
        wc = dict(ocam, width=8, height=8, org=(LO + x0, LO + y0, 0.0), org_u=(8.0, 0.0, 0.0), org_v=(0.0, 8.0, 0.0))
        _, _, Iw, _ = ov.raycast_ea(wc, EA, bricks64=b)
        b[ix] = saved
        return float(np.sum(gI[y0:y0 + 8, x0:x0 + 8] * Iw))

    for s, i in enumerate(sample):
        p0, w0 = p64[i], w64[i]
        # rays at x = LO + px + 1/2 in [p - 2.5, p + 3.5): pixels ceil(p - LO - 3) .. (at most 6), one margin each side
        x0 = int(np.ceil(p0[0] - LO - 3.0)) - 1
        y0 = int(np.ceil(p0[1] - LO - 3.0)) - 1
        idx, own = _contrib(p0, w0, lm, nodes)
        orig = b[idx].copy()
        b[idx] -= own                      # the bricks without point i
        for a in range(4):
            lv = []
            for sg in (1.0, -1.0):
                p, wt = p0.copy(), w0
                if a < 3:
                    p[a] += sg * h
                else:
                    wt = w0 + sg * h
                lv.append(window_loss(p, wt, x0, y0))
            fd[s, a] = (lv[0] - lv[1]) / (2 * h)
        b[idx] = orig
    gpu = np.concatenate([dpos[sample], dw[sample, None]], 1)
    an = np.concatenate([an_pos[sample], an_w[sample, None]], 1)
    mag = np.concatenate([mg_pos[sample], mg_w[sample, None]], 1)
    scale = np.abs(fd).max()
    assert scale > 1e-3
    # the finite differences themselves: the analytic FP64 adjoint of the same chain
    fd_err = np.abs(fd - an).max() / scale
    _log(f"  |fd - analytic FP64| / max|fd| = {fd_err:.2e}")
    assert fd_err <= 1e-7
    n = ea_chain_ops(gv.info().obj_min, gv.info().obj_max, 0.5) + APRON_OPS + splat_ops(R)
    bound = 1e-5 * np.abs(fd) + n * U32 * mag + 1e-7 * scale
    err = np.abs(gpu - fd)
    worst = float((err / bound).max())
    _log(f"  GPU vs finite differences: worst error / bound {worst:.3f} over {fd.size} entries")
    assert (err <= bound).all(), f"{int((err > bound).sum())} entries off, worst error/bound {worst:.3f}"
