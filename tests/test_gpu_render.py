"""GPU parity of Render, the emission-absorption raycast and the adjoints.

Fields are built with the gathers (bit-exact with the oracle) wherever a
threshold decides something (surface hits, transfer-function bins), so both
sides march the same samples.  Tolerances, per result:
  - images (Render float colour, EA I and T): |gpu - fp64| <= 1e-5 * |fp64|
    + 1e-6 * max|fp64| over the image (the FP64 evaluation of the FP32-defined
    sample set);
  - voxel gradient of the EA raycast, apron adjoint, point gradients, entry
    by entry: |gpu - fp64| <= 1e-5 * |fp64| + n * 2^-24 * M, where M is the
    oracle's magnitude adjoint (the same sums over absolute values of every
    term) and n bounds the dependent FP32 roundings (tests/tolerance.py);
  - the uchar4 image equals the quantised float image exactly.
"""
import sys

import numpy as np
import pytest

import oracle as O
from tolerance import APRON_OPS, check_entries, ea_chain_ops, splat_ops

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


def _pair(pts, topo_radius, weights=None):
    G = _G()
    dev = _dev()
    gv = G.VolumeGVDB(0)
    gv.configure([3, 3, 3, 3, 5])   # 8^3 bricks (level 0) ... 32^3 children at level 4
    gv.SetChannelDefault(16, 16, 1)
    gv.AddChannel(0, G.T_FLOAT, 1)
    tp = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).to(dev)
    gv.SetPoints(tp)
    tw = None
    if weights is not None:
        tw = torch.from_numpy(np.ascontiguousarray(weights, np.float32)).to(dev)
        gv.SetPointWeights(tw)
    gv.RebuildTopology(len(pts), topo_radius)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, topo_radius)
    ov.update_atlas()
    return gv, ov, tp, tw


def _ball(seed, n=20000, r=14.0, c=(60.0, 58.0, 62.0)):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (np.asarray(c) + d * np.cbrt(rng.random((n, 1))) * r).astype(np.float32)


def _cam_dict(c):
    return dict(width=c.width, height=c.height, org=tuple(c.org), org_u=tuple(c.org_u), org_v=tuple(c.org_v),
                dir=tuple(c.dir), dir_u=tuple(c.dir_u), dir_v=tuple(c.dir_v))


def _rp_dict(p, shade=None, lut=None):
    d = dict(step=p.step, extinct=p.extinct, albedo=p.albedo, minval=p.minval, alphacut=p.alphacut,
             thresh=p.thresh, vmin=p.vmin, vmax=p.vmax, epsilon=p.epsilon, max_iter=p.max_iter,
             backclr=tuple(p.backclr), light=tuple(p.light), shadow_amt=p.shadow_amt, shadow_bias=p.shadow_bias)
    if shade is not None:
        d["shade"] = shade
    if lut is not None:
        d["lut"] = lut
    return d


def _img_close(gpu, ref64, what):
    gpu = np.asarray(gpu, np.float64)
    tol = 1e-5 * np.abs(ref64) + 1e-6 * max(np.abs(ref64).max(), 1e-30)
    bad = np.abs(gpu - ref64) > tol
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} values off, max err {np.abs(gpu - ref64).max():.3e}"


def _persp(G, w=48, h=40):
    # camera at (60, 58, -40) looking at +z through a 40x34 window at z = 0
    return G.camera_perspective(w, h, (60.0, 58.3, -40.0), (-20.0, -17.0, 40.0), (40.0, 0, 0), (0, 34.0, 0))


@pytest.mark.parametrize("shade", ["trilinear", "levelset", "volume"])
def test_render_vs_oracle(shade):
    G = _G()
    _log(f"render {shade}")
    pts = _ball(5 + len(shade))
    if shade == "levelset":
        gv, ov, tp, _ = _pair(pts, 3.0)
        ln = gv.InsertPointsSubcell(4, len(pts), 3.0)
        gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), ln, 0)
        gv.UpdateApron(0, 3.0)
        ov.insert_points_subcell(4, pts, 3.0)
        ov.gather_level_set(1.5)
        ov.update_apron(0, 3.0)
        rp = G.render_params(step=0.5, thresh=0.0, light=(60, 200, -80), shadow_amt=0.8, shadow_bias=1.0,
                             backclr=(0.1, 0.2, 0.3, 1.0))
        sh, lut = G.SHADE_LEVELSET, None
    else:
        gv, ov, tp, _ = _pair(pts, 4.0)
        ln = gv.InsertPointsSubcell(4, len(pts), 4.0)
        gv.GatherDensity(4, len(pts), 2.0, (0, 0, 0), ln, 0)
        gv.UpdateApron(0, 0.0)
        ov.insert_points_subcell(4, pts, 4.0)
        ov.gather_density(2.0)
        ov.update_apron(0, 0.0)
        if shade == "trilinear":
            rp = G.render_params(step=0.5, thresh=1.0, light=(60, 200, -80), shadow_amt=0.8,
                                 backclr=(0.1, 0.2, 0.3, 1.0))
            sh, lut = G.SHADE_TRILINEAR, None
        else:
            t = np.linspace(0, 1, 16384, dtype=np.float32)
            lut = np.stack([t, t * t, 1 - t, 0.1 + t], 1).astype(np.float32)
            lut_dev = torch.from_numpy(lut.reshape(-1)).to(_dev())
            rp = G.render_params(step=0.5, thresh=0.05, vmin=0.0, vmax=2.0, extinct=-1.1, albedo=1.5,
                                 minval=0.005, alphacut=0.01, backclr=(0.1, 0.2, 0.3, 1.0), lut=lut_dev)
            sh = G.SHADE_VOLUME
    np.testing.assert_array_equal(gv.bricks(0), ov.bricks(0))
    cam = _persp(G)
    rgba, rgba8 = gv.Render(sh, 0, cam, rp)
    torch.cuda.synchronize()
    r32, r64 = ov.render(_cam_dict(cam), _rp_dict(rp, sh, lut))
    g = rgba.cpu().numpy()
    assert (r64[..., 3] > 0).sum() > 100, "view misses the object"
    _img_close(g, r64, f"render {shade}")
    q = np.clip(g * 255.0, 0, 255).astype(np.float32).astype(np.int32).astype(np.uint8)
    np.testing.assert_array_equal(rgba8.cpu().numpy(), q)


def _ea_setup(seed=21, n=30000):
    G = _G()
    rng = np.random.default_rng(seed)
    means = rng.uniform(40, 80, size=(6, 3))
    sig = rng.uniform(3, 7, size=(6, 3))
    lab = rng.integers(0, 6, n)
    pts = (means[lab] + rng.normal(size=(n, 3)) * sig[lab]).astype(np.float32)
    pts = np.clip(pts, 1.0, None)
    w = rng.uniform(0.1, 1.0, n).astype(np.float32)
    gv, ov, tp, tw = _pair(pts, 4.0, w)
    gv.SplatCubic(len(pts), 2.0)
    gv.UpdateApron(0, 0.0)
    b64 = ov.apron_f64(ov.splat_cubic_f64_padded(pts.astype(np.float64), 2.0, w.astype(np.float64)), 0.0)
    return gv, ov, tp, tw, pts, w, b64


@pytest.mark.parametrize("view", ["ortho", "perspective", "dense_perspective"])
def test_raycast_ea_and_backward_vs_oracle(view):
    """dense_perspective: 96x96 rays through a 12-voxel window (1/8 voxel
    apart), so rays two pixels apart, which share one wave of the backward,
    flush the same voxels: the claim decides plain and atomic adds there,
    and the result must still be the same bits on every run."""
    G = _G()
    _log(f"ea {view}")
    gv, ov, tp, tw, pts, w, b64 = _ea_setup()
    if view == "ortho":
        cam = G.camera_ortho(64, 64, (30.0, 30.0, 0.0), (64.0, 0, 0), (0, 64.0, 0), (0, 0, 1))
    elif view == "perspective":
        cam = _persp(G, 56, 56)
    else:
        cam = G.camera_perspective(96, 96, (55.0, 56.0, -30.0), (-6.0, -6.0, 30.0), (12.0, 0, 0), (0, 12.0, 0))
    rp = G.ea_params(step=0.5, kappa=0.05)
    I0, T0 = gv.RaycastEA(0, cam, rp, record=False)
    I, T = gv.RaycastEA(0, cam, rp, record=True)
    torch.cuda.synchronize()
    assert torch.equal(I0, I) and torch.equal(T0, T)
    _, _, I64, T64 = ov.raycast_ea(_cam_dict(cam), _rp_dict(rp), bricks64=b64)
    assert (I64 > 0).sum() > 500
    _img_close(I.cpu().numpy(), I64, "EA I")
    _img_close(T.cpu().numpy(), T64, "EA T")
    rng = np.random.default_rng(1)
    gI = torch.from_numpy(rng.normal(size=I.shape).astype(np.float32)).to(I.device)
    gT = torch.from_numpy(rng.normal(size=I.shape).astype(np.float32)).to(I.device)
    grad = gv.RaycastEABackward(0, cam, rp, I, T, gI, gT)
    for rep in range(3):
        grad2 = gv.RaycastEABackward(0, cam, rp, I, T, gI, gT)
        torch.cuda.synchronize()
        assert torch.equal(grad, grad2), f"backward is not deterministic (repeat {rep})"
    gI64, gT64 = gI.cpu().numpy().astype(np.float64), gT.cpu().numpy().astype(np.float64)
    g64 = ov.raycast_ea_backward(_cam_dict(cam), _rp_dict(rp), gI64, gT64, bricks64=b64)
    m64 = ov.raycast_ea_backward(_cam_dict(cam), _rp_dict(rp), gI64, gT64, bricks64=b64, mag=True)
    g = grad.cpu().numpy().astype(np.float64)
    assert np.abs(g64).max() > 0
    inf = gv.info()
    n = ea_chain_ops(inf.obj_min, inf.obj_max, rp.step)
    ratio = check_entries(g, g64, m64, n, f"EA voxel gradient ({view})")
    _log(f"  EA gradient: n = {n}, worst error / bound {ratio:.3f}")


def test_apron_adjoint_and_splat_backward_vs_oracle():
    G = _G()
    _log("apron adjoint + splat backward")
    gv, ov, tp, tw, pts, w, b64 = _ea_setup(seed=33, n=20000)
    inf = gv.info()
    B, br = inf.node_cnt[0], inf.brick_res
    rng = np.random.default_rng(2)
    gr = rng.normal(size=(B, br, br, br)).astype(np.float32)
    gt = torch.from_numpy(gr).to(tp.device)
    gv.ApronAdjoint(gt)
    torch.cuda.synchronize()
    ref = ov.apron_adjoint_f64(gr.astype(np.float64))
    mag = ov.apron_adjoint_f64(np.abs(gr.astype(np.float64)))
    got = gt.cpu().numpy().astype(np.float64)
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    check_entries(got[inner], ref[inner], mag[inner], APRON_OPS, "apron adjoint")
    dpos, dw = gv.SplatCubicBackward(len(pts), 2.0, (0, 0, 0), gt)
    torch.cuda.synchronize()
    p64, w64 = pts.astype(np.float64), w.astype(np.float64)
    rdpos, rdw = ov.splat_cubic_backward_f64(p64, 2.0, got, w64)
    mdpos, mdw = ov.splat_cubic_backward_f64(p64, 2.0, got, w64, mag=True)
    for a, b, m, name in ((dpos.cpu().numpy(), rdpos, mdpos, "dpos"), (dw.cpu().numpy(), rdw, mdw, "dw")):
        ratio = check_entries(a, b, m, splat_ops(2.0), name)
        _log(f"  {name}: worst error / bound {ratio:.3f}")


def test_autograd_splat_render_vs_oracle_chain():
    G = _G()
    from gvdbx.autograd import SplatRenderConfig, ortho_view, splat_render_ea
    _log("autograd chain")
    rng = np.random.default_rng(44)
    n = 20000
    means = rng.uniform(40, 80, size=(5, 3))
    pts = (means[rng.integers(0, 5, n)] + rng.normal(size=(n, 3)) * 4).astype(np.float32)
    w = rng.uniform(0.1, 1.0, n).astype(np.float32)
    dev = _dev()
    vol = G.VolumeGVDB(0)
    vol.configure([3, 3, 3, 3, 5])
    vol.AddChannel(0, G.T_FLOAT, 1)
    cam, rp = ortho_view((28, 28), (92, 92), 0.0, 64, 64)
    cfg = SplatRenderConfig(radius=2.0, topo_radius=4.0, cam=cam, rp=rp)
    P = torch.from_numpy(pts).to(dev).requires_grad_(True)
    W = torch.from_numpy(w).to(dev).requires_grad_(True)
    I, T = splat_render_ea(P, W, vol, cfg)
    target = torch.from_numpy(rng.normal(size=I.shape).astype(np.float32) * 0.1).to(dev)
    loss = 0.5 * ((I - target) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    gI = (I - target).detach().cpu().numpy().astype(np.float64)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, 4.0)
    ov.update_atlas()
    p64, w64 = pts.astype(np.float64), w.astype(np.float64)
    b64 = ov.apron_f64(ov.splat_cubic_f64_padded(p64, 2.0, w64), 0.0)
    _, _, I64, _ = ov.raycast_ea(_cam_dict(cam), _rp_dict(rp), bricks64=b64)
    _img_close(I.detach().cpu().numpy(), I64, "autograd forward I")
    g = ov.apron_adjoint_f64(ov.raycast_ea_backward(_cam_dict(cam), _rp_dict(rp), gI, bricks64=b64))
    gm = ov.apron_adjoint_f64(ov.raycast_ea_backward(_cam_dict(cam), _rp_dict(rp), gI, bricks64=b64, mag=True))
    rdpos, rdw = ov.splat_cubic_backward_f64(p64, 2.0, g, w64)
    mdpos, mdw = ov.splat_cubic_backward_f64(p64, 2.0, gm, w64, mag=True)
    inf = vol.info()
    n = ea_chain_ops(inf.obj_min, inf.obj_max, rp.step) + APRON_OPS + splat_ops(2.0)
    for a, b, m, name in ((P.grad.cpu().numpy(), rdpos, mdpos, "dL/dp"), (W.grad.cpu().numpy(), rdw, mdw, "dL/dw")):
        assert np.abs(b).max() > 0
        ratio = check_entries(a, b, m, n, name)
        _log(f"  {name}: n = {n}, worst error / bound {ratio:.3f}")


def test_render_error_paths():
    G = _G()
    _log("render errors")
    gv, ov, tp, tw, pts, w, b64 = _ea_setup(seed=5, n=5000)
    cam = G.camera_ortho(16, 16, (30.0, 30.0, 0.0), (64.0, 0, 0), (0, 64.0, 0), (0, 0, 1))
    rp = G.ea_params()
    I, T = gv.RaycastEA(0, cam, rp, record=False)
    with pytest.raises(G.GvdbxError) as e:
        gv.RaycastEABackward(0, cam, rp, I, T, torch.ones_like(I))
    assert e.value.code == 2
    with pytest.raises(G.GvdbxError) as e:
        gv.Render(9, 0, cam, rp)           # no such shading
    assert e.value.code == 1
    with pytest.raises(G.GvdbxError) as e:
        gv.Render(G.SHADE_TRICUBIC, 0, cam, G.render_params(finestep=0.0))
    assert e.value.code == 1
    with pytest.raises(G.GvdbxError) as e:
        gv.Render(G.SHADE_SECTION2D, 0, cam, rp)   # no transfer function
    assert e.value.code == 1
    with pytest.raises(G.GvdbxError) as e:
        gv.Render(G.SHADE_VOLUME, 0, cam, rp)   # no transfer function
    assert e.value.code == 1
    I, T = gv.RaycastEA(0, cam, rp, record=True)
    gv.RebuildTopology(len(pts), 4.0)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    with pytest.raises(G.GvdbxError) as e:        # the recorded view belongs to the old topology
        gv.RaycastEABackward(0, cam, rp, I, T, torch.ones_like(I))
    assert e.value.code == 2
