"""GPU parity of the batched forms.

Depth fusion: back-projection of all frames in one launch is bit-exact with
the oracle's FP32 restatement; fusing all frames in one pass equals the
reference's per-frame sequence (GatherLevelSet with accumulate on the same
topology) bit for bit, since min is exact and order-free; and both equal the
oracle's level set of the union of the frames' points.
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


def _vol(G):
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])
    gv.SetChannelDefault(16, 16, 1)
    gv.AddChannel(0, G.T_FLOAT, 1)
    return gv


def test_depth_fusion_batched_equals_sequential_and_oracle():
    import gvdbx as G
    from gvdbx import workloads
    dev = _dev()
    F, W, H = 6, 96, 72
    _log(f"depth fusion {F} frames {W}x{H}")
    d, poses, intr = workloads.depth_scene(11, nframes=F, width=W, height=H, fx=78.75, fy=78.75)
    origin, vox = (-13.0, -13.0, -13.0), 0.05
    pts_ref = O.backproject_depth(d, intr, poses, vox, origin)
    dt = torch.from_numpy(d).to(dev)
    pt = torch.from_numpy(poses).to(dev)
    pts = G.backproject_depth(dt, intr, pt, vox, origin)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(pts.cpu().numpy(), pts_ref)
    n = pts.shape[0]
    # batched: every frame in one pass
    gb = _vol(G)
    gb.SetPoints(pts)
    gb.RebuildTopology(n, 2.0)
    gb.FinishTopology(False, True)
    gb.UpdateAtlas()
    ln = gb.InsertPointsSubcell(4, n, 2.0)
    gb.GatherLevelSet(4, n, 1.0, (0, 0, 0), ln, 0)
    gb.UpdateApron(0, 0.0)
    # per frame, accumulating on the same topology (main_point_fusion.cpp:588-590)
    gs = _vol(G)
    gs.SetPoints(pts)
    gs.RebuildTopology(n, 2.0)
    gs.FinishTopology(False, True)
    gs.UpdateAtlas()
    gs.FillChannel(0, 3.0)
    per = W * H
    for f in range(F):
        fp = pts[f * per:(f + 1) * per]
        gs.SetPoints(fp)
        lf = gs.InsertPointsSubcell(4, per, 2.0)
        gs.GatherLevelSet(4, per, 1.0, (0, 0, 0), lf, 0, accum=True)
    gs.UpdateApron(0, 0.0)
    torch.cuda.synchronize()
    bb, bs = gb.bricks(0), gs.bricks(0)
    assert bb.shape[0] > 50
    np.testing.assert_array_equal(bb, bs)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts_ref, 2.0)
    ov.update_atlas()
    ov.insert_points_subcell(4, pts_ref, 2.0)
    ov.gather_level_set(1.0)
    ov.update_apron(0, 0.0)
    np.testing.assert_array_equal(bb, ov.bricks(0))


def _cloud_cam(G, pts, W=64, H=64):
    from gvdbx import workloads
    lo, hi = workloads.cloud_camera_box(pts)
    # pixel centres a quarter voxel off the integer grid; rays start 0.3 below z = 0
    return G.camera_ortho(W, H, (lo[0] + 0.25, lo[1] + 0.25, -0.3), (hi[0] - lo[0], 0, 0), (0, hi[1] - lo[1], 0),
                          (0, 0, 1))


def _leaf_key(pos):
    return [tuple(int(x) for x in p) for p in pos]


def test_batched_clouds_equal_per_cloud():
    """Topology, splat, apron, EA raycast and every adjoint of a batch of
    clouds equal the same calls made for each cloud alone, bit for bit."""
    import gvdbx as G
    from gvdbx import workloads
    dev = _dev()
    clouds = workloads.cloud_batch(9, nclouds=6, nmin=3000, nmax=40000)
    _log("batched clouds: " + ", ".join(f"{len(p)} {k}" for p, w, k in clouds))
    rp = G.ea_params(step=0.5, kappa=0.05)
    rng = np.random.default_rng(4)
    gIs = [rng.normal(size=(64, 64)).astype(np.float32) for _ in clouds]
    singles = []
    for (p, w, k), gI in zip(clouds, gIs):
        gv = _vol(G)
        tp = torch.from_numpy(p).to(dev)
        tw = torch.from_numpy(w).to(dev)
        gv.SetPoints(tp)
        gv.SetPointWeights(tw)
        gv.RebuildTopology(len(p), 4.0)
        gv.FinishTopology(False, True)
        gv.UpdateAtlas()
        gv.SplatCubic(len(p), 2.0)
        gv.UpdateApron(0, 0.0)
        cam = _cloud_cam(G, p)
        I, T = gv.RaycastEA(0, cam, rp, record=True)
        grad = gv.RaycastEABackward(0, cam, rp, I, T, torch.from_numpy(gI).to(dev))
        gv.ApronAdjoint(grad)
        dpos, dw = gv.SplatCubicBackward(len(p), 2.0, (0, 0, 0), grad)
        torch.cuda.synchronize()
        inf = gv.info()
        singles.append(dict(nodes=gv.nodes(0), bricks=gv.bricks(0), I=I.cpu().numpy(), T=T.cpu().numpy(),
                            grad=grad.cpu().numpy(), dpos=dpos.cpu().numpy(), dw=dw.cpu().numpy(),
                            bounds=np.array(list(inf.obj_min) + list(inf.obj_max), np.float32), keep=(tp, tw)))
        gv.close()
    off = np.concatenate([[0], np.cumsum([len(p) for p, w, k in clouds])])
    P = torch.from_numpy(np.concatenate([p for p, w, k in clouds])).to(dev)
    Wt = torch.from_numpy(np.concatenate([w for p, w, k in clouds])).to(dev)
    gb = _vol(G)
    gb.SetPointsBatched(P, off)
    gb.SetPointWeights(Wt)
    gb.RebuildTopology(int(off[-1]), 4.0)
    gb.FinishTopology(False, True)
    gb.UpdateAtlas()
    gb.SplatCubic(int(off[-1]), 2.0)
    gb.UpdateApron(0, 0.0)
    cams = [_cloud_cam(G, p) for p, w, k in clouds]
    I, T = gb.RaycastEABatched(0, cams, rp, record=True)
    grad = gb.RaycastEABackwardBatched(0, cams, rp, I, T, torch.from_numpy(np.stack(gIs)).to(dev))
    gb.ApronAdjoint(grad)
    dpos, dw = gb.SplatCubicBackward(int(off[-1]), 2.0, (0, 0, 0), grad)
    torch.cuda.synchronize()
    cell_off, bounds = gb.cloud_layout()
    lc = gb.leaf_clouds()
    nodes, bricks, gradb = gb.nodes(0), gb.bricks(0), grad.cpu().numpy()
    Ib, Tb = I.cpu().numpy(), T.cpu().numpy()
    dpb, dwb = dpos.cpu().numpy(), dw.cpu().numpy()
    for c, s in enumerate(singles):
        sel = np.nonzero(lc == c)[0]
        local = nodes["mPos"][sel] - cell_off[c] * 8
        ref = {k: i for i, k in enumerate(_leaf_key(s["nodes"]["mPos"]))}
        keys = _leaf_key(local)
        assert sorted(keys) == sorted(ref), f"cloud {c}: leaf sets differ"
        order = np.array([ref[k] for k in keys])
        np.testing.assert_array_equal(bricks[sel], s["bricks"][order], err_msg=f"cloud {c} bricks")
        np.testing.assert_array_equal(bounds[c], s["bounds"], err_msg=f"cloud {c} bounds")
        np.testing.assert_array_equal(Ib[c], s["I"], err_msg=f"cloud {c} I")
        np.testing.assert_array_equal(Tb[c], s["T"], err_msg=f"cloud {c} T")
        inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
        np.testing.assert_array_equal(gradb[sel][inner], s["grad"][order][inner], err_msg=f"cloud {c} grad")
        np.testing.assert_array_equal(dpb[off[c]:off[c + 1]], s["dpos"], err_msg=f"cloud {c} dpos")
        np.testing.assert_array_equal(dwb[off[c]:off[c + 1]], s["dw"], err_msg=f"cloud {c} dw")
    assert (lc >= 0).all()


def test_batched_clouds_vs_oracle():
    """Every cloud of a batch against the oracle run on that cloud alone:
    leaf set bit-exact; padded bricks of the splat + apron within 1e-5 of the
    FP64 chain; EA images within 1e-5 (+1e-6 of the image maximum) of FP64;
    voxel gradients (after the apron adjoint) and point gradients entry by
    entry within 1e-5 |ref| + n 2^-24 M of the FP64 adjoints (M: the
    magnitude adjoint, tests/tolerance.py)."""
    import gvdbx as G
    from gvdbx import workloads
    from tolerance import APRON_OPS, check_entries, ea_chain_ops, splat_ops
    dev = _dev()
    clouds = workloads.cloud_batch(19, nclouds=5, nmin=3000, nmax=30000)
    _log("batched clouds vs oracle: " + ", ".join(f"{len(p)} {k}" for p, w, k in clouds))
    rp = G.ea_params(step=0.5, kappa=0.05)
    rng = np.random.default_rng(5)
    W = H = 48
    gIs = [rng.normal(size=(H, W)).astype(np.float32) for _ in clouds]
    off = np.concatenate([[0], np.cumsum([len(p) for p, w, k in clouds])])
    P = torch.from_numpy(np.concatenate([p for p, w, k in clouds])).to(dev)
    Wt = torch.from_numpy(np.concatenate([w for p, w, k in clouds])).to(dev)
    gb = _vol(G)
    gb.SetPointsBatched(P, off)
    gb.SetPointWeights(Wt)
    gb.RebuildTopology(int(off[-1]), 4.0)
    gb.FinishTopology(False, True)
    gb.UpdateAtlas()
    gb.SplatCubic(int(off[-1]), 2.0)
    gb.UpdateApron(0, 0.0)
    cams = [_cloud_cam(G, p, W, H) for p, w, k in clouds]
    I, T = gb.RaycastEABatched(0, cams, rp, record=True)
    grad = gb.RaycastEABackwardBatched(0, cams, rp, I, T, torch.from_numpy(np.stack(gIs)).to(dev))
    gb.ApronAdjoint(grad)
    dpos, dw = gb.SplatCubicBackward(int(off[-1]), 2.0, (0, 0, 0), grad)
    torch.cuda.synchronize()
    cell_off, bounds = gb.cloud_layout()
    lc = gb.leaf_clouds()
    nodes, bricks, gradb = gb.nodes(0), gb.bricks(0), grad.cpu().numpy()
    Ib, Tb = I.cpu().numpy(), T.cpu().numpy()
    dpb, dwb = dpos.cpu().numpy(), dw.cpu().numpy()
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    from test_gpu_render import _cam_dict, _img_close, _rp_dict
    for c, (p, w, k) in enumerate(clouds):
        ov = O.OracleVolume()
        ov.add_channel(0, 1)
        ov.rebuild_topology(p, 4.0)
        ov.update_atlas()
        onodes = ov.nodes(0)
        sel = np.nonzero(lc == c)[0]
        local = nodes["mPos"][sel] - cell_off[c] * 8
        ref = {kk: i for i, kk in enumerate(_leaf_key(onodes["mPos"]))}
        keys = _leaf_key(local)
        assert sorted(keys) == sorted(ref), f"cloud {c}: leaf set differs from the oracle"
        order = np.array([ref[kk] for kk in keys])
        oinf = ov.info()
        np.testing.assert_array_equal(bounds[c], np.array(list(oinf.obj_min) + list(oinf.obj_max), np.float32))
        p64, w64 = p.astype(np.float64), w.astype(np.float64)
        b64 = ov.apron_f64(ov.splat_cubic_f64_padded(p64, 2.0, w64), 0.0)
        got = bricks[sel].astype(np.float64)
        want = b64[order]
        floor = 1e-6 * np.abs(want).max()
        err = np.abs(got - want)
        assert (err <= 1e-5 * np.abs(want) + floor).all(), f"cloud {c}: splat max err {err.max():.3e}"
        cd, rd = _cam_dict(cams[c]), _rp_dict(rp)
        _, _, I64, T64 = ov.raycast_ea(cd, rd, bricks64=b64)
        assert (I64 > 0).sum() > 100, f"cloud {c}: the view misses the cloud"
        _img_close(Ib[c], I64, f"cloud {c} I")
        _img_close(Tb[c], T64, f"cloud {c} T")
        gI64 = gIs[c].astype(np.float64)
        g64 = ov.apron_adjoint_f64(ov.raycast_ea_backward(cd, rd, gI64, bricks64=b64))
        m64 = ov.apron_adjoint_f64(ov.raycast_ea_backward(cd, rd, gI64, bricks64=b64, mag=True))
        n_ea = ea_chain_ops(oinf.obj_min, oinf.obj_max, rp.step) + APRON_OPS
        check_entries(gradb[sel][inner], g64[order][inner], m64[order][inner], n_ea, f"cloud {c} voxel gradient")
        rdp, rdw = ov.splat_cubic_backward_f64(p64, 2.0, g64, w64)
        mdp, mdw = ov.splat_cubic_backward_f64(p64, 2.0, m64, w64, mag=True)
        n_pt = n_ea + splat_ops(2.0)
        check_entries(dpb[off[c]:off[c + 1]], rdp, mdp, n_pt, f"cloud {c} dL/dp")
        check_entries(dwb[off[c]:off[c + 1]], rdw, mdw, n_pt, f"cloud {c} dL/dw")
        _log(f"  cloud {c} ({k}, {len(p)} points, {len(sel)} leaves): oracle parity ok")
