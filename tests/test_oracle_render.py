"""Pins the oracle's raycast and its adjoints (CPU only).

What fixes each function independently of the oracle's own code:
  - constant-density slab: T = a^n and I = c (1 - a^n) with a = exp(-kappa c step),
    n = 16 samples per 8-voxel brick crossed (closed form);
  - HDDA sample set: for +z rays whose origin sits 0.3 voxel off the brick
    grid, the samples are exactly the points z = Z + 0.2 + 0.5 j of every
    active leaf in the pixel's column (brute force over the leaf list);
  - trilinear sampling: scipy.ndimage.map_coordinates(order=1) on the padded
    brick (a library routine), feeding an independent numpy EA sum;
  - the SHADE_VOLUME transfer function: a constant LUT on a constant slab
    (closed form incl. the background blend of MOD.cu:75);
  - adjoints: FP64 central finite differences of the FP64 chain
    splat -> apron -> raycast, in positions, weights and voxel values;
  - surface shading: first sample below / above the threshold found by brute
    force, normal and Phong term recomputed with map_coordinates.
"""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

import oracle as O


def vol_for(points, radius):
    v = O.OracleVolume()
    v.add_channel(0, 1)
    v.rebuild_topology(points, radius)
    v.update_atlas()
    return v


def column_leaves(nodes, x, y):
    sel = (nodes["mPos"][:, 0] <= x) & (x < nodes["mPos"][:, 0] + 8) & (nodes["mPos"][:, 1] <= y) & \
          (y < nodes["mPos"][:, 1] + 8)
    idx = np.nonzero(sel)[0]
    return idx[np.argsort(nodes["mPos"][idx, 2])]


@pytest.fixture(scope="module")
def slab():
    rng = np.random.default_rng(1)
    pts = (rng.random((20000, 3)) * [32, 32, 64] + [64, 64, 64]).astype(np.float32)
    return vol_for(pts, 2.0)


EA = dict(step=0.5, extinct=-0.05, albedo=1.0, minval=-np.inf, alphacut=0.0)


def ortho(w, h, x0, y0, ext, z0=-0.3):
    return dict(width=w, height=h, org=(x0, y0, z0), org_u=(ext, 0, 0), org_v=(0, ext * h / w, 0), dir=(0, 0, 1))


def test_ea_constant_slab_closed_form(slab):
    c = np.float32(0.7)
    slab.fill_channel(0, c)
    cam = ortho(8, 8, 72, 72, 16)
    I32, T32, I64, T64 = slab.raycast_ea(cam, EA)
    nodes = slab.nodes(0)
    a = np.exp(float(np.float32(-0.05)) * float(c) * 0.5)   # the FP64 path uses the float parameters
    for px in range(64):
        x = 72 + ((px % 8) + 0.5) * 2
        y = 72 + ((px // 8) + 0.5) * 2
        n = 16 * len(column_leaves(nodes, x, y))
        assert n > 0
        np.testing.assert_allclose(T64.flat[px], a ** n, rtol=1e-12)
        np.testing.assert_allclose(I64.flat[px], float(c) * (1 - a ** n), rtol=1e-12)
        np.testing.assert_allclose(I32.flat[px], float(c) * (1 - a ** n), rtol=1e-5)


def test_ea_sample_set_brute_force():
    rng = np.random.default_rng(7)
    # two blobs with a gap along z so that rays cross empty level-1 cells
    a = rng.normal(size=(3000, 3)) * 5 + [40, 44, 30]
    b = rng.normal(size=(3000, 3)) * 5 + [40, 44, 110]
    v = vol_for(np.concatenate([a, b]).astype(np.float32), 2.0)
    v.splat_cubic(np.concatenate([a, b]).astype(np.float32), 2.0)
    cam = ortho(16, 16, 28, 32, 24)
    nodes = v.nodes(0)
    checked = 0
    for px in range(0, 256, 7):
        x = 28 + ((px % 16) + 0.5) * 1.5
        y = 32 + ((px // 16) + 0.5) * 1.5
        leaves = column_leaves(nodes, x, y)
        exp_leaf = np.repeat(leaves, 16)
        exp_z = np.tile(0.2 + 0.5 * np.arange(16), len(leaves))
        got_leaf, got_p = v.raycast_ea_samples(cam, EA, px)
        np.testing.assert_array_equal(got_leaf, exp_leaf)
        if len(leaves):
            np.testing.assert_allclose(got_p[:, 2], exp_z, atol=1e-5)   # FP32 rounding of z near 64-128
            np.testing.assert_allclose(got_p[:, 0], x - nodes["mPos"][got_leaf, 0], atol=0)
            np.testing.assert_allclose(got_p[:, 1], y - nodes["mPos"][got_leaf, 1], atol=0)
            checked += 1
    assert checked > 5


def tri_scipy(brick, p):
    """trilinear sample of a padded (z,y,x) brick at brick-local p via scipy."""
    coords = np.array([[p[2] + 0.5], [p[1] + 0.5], [p[0] + 0.5]], dtype=np.float64)
    return float(map_coordinates(brick.astype(np.float64), coords, order=1, mode="nearest")[0])


def test_ea_matches_library_trilinear_sum():
    rng = np.random.default_rng(11)
    pts = (rng.normal(size=(4000, 3)) * [6, 5, 7] + [50, 50, 50]).astype(np.float32)
    w = rng.uniform(0.1, 1, len(pts)).astype(np.float32)
    v = vol_for(pts, 4.0)
    v.splat_cubic(pts, 2.0, weights=w)
    v.update_apron(0, 0.0)
    cam = dict(width=10, height=10, org=(38.3, 39.1, 10.0), org_u=(24, 0, 0), org_v=(0, 24, 0),
               dir=(0.02, -0.01, 1.0), dir_u=(0.03, 0, 0), dir_v=(0, 0.02, 0))
    I32, T32, I64, T64 = v.raycast_ea(cam, EA)
    br = v.bricks(0)
    e = float(np.float32(-0.05)) * 0.5
    for px in range(0, 100, 9):
        leaf, p = v.raycast_ea_samples(cam, EA, px)
        I, T = 0.0, 1.0
        for k in range(len(leaf)):
            s = tri_scipy(br[leaf[k]], p[k])
            a = np.exp(e * s)
            I += s * T * (1 - a)
            T *= a
        np.testing.assert_allclose(I64.flat[px], I, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(T64.flat[px], T, rtol=1e-12)


def test_volume_shade_constant_lut(slab):
    c = np.float32(0.5)
    slab.fill_channel(0, c)
    lut = np.zeros((16384, 4), np.float32)
    lut[:] = (0.02, 0.04, 0.06, 0.8)
    back = (0.1, 0.2, 0.3, 1.0)
    rp = dict(shade=O.SHADE_VOLUME, step=0.5, extinct=-1.1, albedo=1.5, minval=0.005, alphacut=0.01, lut=lut,
              backclr=back)
    cam = ortho(4, 4, 72, 72, 16)
    rgba32, rgba64 = slab.render(cam, rp)
    nodes = slab.nodes(0)
    a = np.exp(float(np.float32(-1.1)) * float(np.float32(0.8)) * 0.5)
    for px in range(16):
        x = 72 + ((px % 4) + 0.5) * 4
        y = 72 + ((px // 4) + 0.5) * 4
        nb = len(column_leaves(nodes, x, y))
        # samples stop once T <= alphacut (checked before each sample)
        n = min(16 * nb, int(np.ceil(np.log(0.01) / np.log(a))))
        T = a ** n
        clr = np.minimum(np.array(lut[0, :3], np.float64) * 1.5 * (1 - T), 1.0)
        exp = np.array(back[:3]) + (1 - T) * (clr - np.array(back[:3]))
        np.testing.assert_allclose(rgba64[px // 4, px % 4, :3], exp, rtol=1e-7)
        np.testing.assert_allclose(rgba64[px // 4, px % 4, 3], 1 - T, rtol=1e-12)
        np.testing.assert_allclose(rgba32[px // 4, px % 4], rgba64[px // 4, px % 4], rtol=2e-6, atol=1e-7)


def _chain(v, pos64, w, r, cam, rp, gI):
    b = v.splat_cubic_f64_padded(pos64, r, w)
    b = v.apron_f64(b, 0.0)
    _, _, I64, _ = v.raycast_ea(cam, rp, bricks64=b)
    return float(np.dot(gI, I64.reshape(-1))), b


@pytest.fixture(scope="module")
def cluster():
    rng = np.random.default_rng(3)
    n = 1500
    pts = (rng.normal(size=(n, 3)) * [4, 5, 3] + [40, 40, 40]).astype(np.float32)
    w = rng.uniform(0.1, 1, n)
    v = vol_for(pts, 4.0)
    cam = dict(width=12, height=12, org=(28.1, 28.2, 0.0), org_u=(24, 0, 0), org_v=(0, 24, 0),
               dir=(0.05, -0.03, 1), dir_u=(0.01, 0, 0))
    gI = rng.normal(size=144)
    return v, pts.astype(np.float64), w, cam, gI, rng


def test_adjoint_chain_finite_differences(cluster):
    v, pos, w, cam, gI, rng = cluster
    r = 2.0
    L0, b0 = _chain(v, pos, w, r, cam, EA, gI)
    g = v.raycast_ea_backward(cam, EA, gI, bricks64=b0)
    g = v.apron_adjoint_f64(g)
    dpos, dw = v.splat_cubic_backward_f64(pos, r, g, w)
    h = 1e-5
    fd, an = [], []
    for i in rng.choice(len(pos), 12, replace=False):
        for a in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += h
            p2[i, a] -= h
            fd.append((_chain(v, p1, w, r, cam, EA, gI)[0] - _chain(v, p2, w, r, cam, EA, gI)[0]) / (2 * h))
            an.append(dpos[i, a])
        w1, w2 = w.copy(), w.copy()
        w1[i] += h
        w2[i] -= h
        fd.append((_chain(v, pos, w1, r, cam, EA, gI)[0] - _chain(v, pos, w2, r, cam, EA, gI)[0]) / (2 * h))
        an.append(dw[i])
    fd, an = np.array(fd), np.array(an)
    scale = np.abs(fd).max()
    assert scale > 1e-3
    assert np.abs(fd - an).max() <= 1e-7 * scale


def test_ea_backward_voxel_finite_differences(cluster):
    v, pos, w, cam, gI, rng = cluster
    b0 = v.apron_f64(v.splat_cubic_f64_padded(pos, 2.0, w), 0.0)
    gT = rng.normal(size=144)

    def loss(b):
        _, _, I, T = v.raycast_ea(cam, EA, bricks64=b)
        return float(np.dot(gI, I.reshape(-1)) + np.dot(gT, T.reshape(-1)))

    g = v.raycast_ea_backward(cam, EA, gI, gT, bricks64=b0)
    flat = np.argsort(-np.abs(g.reshape(-1)))[:10]
    h = 1e-6
    for f in flat:
        b1, b2 = b0.copy().reshape(-1), b0.copy().reshape(-1)
        b1[f] += h
        b2[f] -= h
        fd = (loss(b1.reshape(b0.shape)) - loss(b2.reshape(b0.shape))) / (2 * h)
        np.testing.assert_allclose(g.reshape(-1)[f], fd, rtol=1e-6)


def test_apron_adjoint_is_transpose(cluster):
    v, pos, w, cam, gI, rng = cluster
    B, br = v.info().node_cnt[0], v.info().brick_res
    x = rng.normal(size=(B, br, br, br))
    y = rng.normal(size=(B, br, br, br))
    ap = v.apron_f64(x, 0.0)           # A x (apron filled from interiors)
    interior = np.zeros_like(y)
    interior[:, 1:-1, 1:-1, 1:-1] = 1
    aty = v.apron_adjoint_f64(y)       # A^T y
    # <A x, y> over the whole padded array = <x, A^T y> over interiors
    lhs = float(np.sum(ap * y))
    rhs = float(np.sum(x * interior * aty))
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)


def _surface_setup():
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(6000, 3))
    pts = (pts / np.linalg.norm(pts, axis=1, keepdims=True) * np.cbrt(rng.random((6000, 1))) * 12 + 48)
    return pts.astype(np.float32)


def _normal(br, L, p, sign):
    g = np.array([tri_scipy(br[L], (p[0] + 0.5, p[1], p[2])) - tri_scipy(br[L], (p[0] - 0.5, p[1], p[2])),
                  tri_scipy(br[L], (p[0], p[1] + 0.5, p[2])) - tri_scipy(br[L], (p[0], p[1] - 0.5, p[2])),
                  tri_scipy(br[L], (p[0], p[1], p[2] + 0.5)) - tri_scipy(br[L], (p[0], p[1], p[2] - 0.5))])
    g *= sign
    return g / np.linalg.norm(g)


def _phong(g, hit, light):
    l = np.asarray(light, np.float64) - hit
    l /= np.linalg.norm(l)
    return 0.9 * max(0.0, float(g @ l)) + 0.1


def test_trilinear_surface_brute_force():
    """SHADE_TRILINEAR on a density bump: first sample >= thresh."""
    pts = _surface_setup()
    v = vol_for(pts, 4.0)
    v.splat_cubic(pts, 2.0)
    v.update_apron(0, 0.0)
    br, nodes = v.bricks(0), v.nodes(0)
    cam = ortho(8, 8, 36.0, 36.0, 24, z0=-0.3)
    light = (48, 48, -100)
    rp = dict(shade=O.SHADE_TRILINEAR, step=0.5, thresh=0.5, light=light, shadow_amt=0.0, backclr=(0, 0, 0, 0))
    rgba32, rgba64 = v.render(cam, rp)
    hits = 0
    for px in range(64):
        x = 36 + ((px % 8) + 0.5) * 3
        y = 36 + ((px // 8) + 0.5) * 3
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
            np.testing.assert_allclose(rgba64[px // 8, px % 8], (0, 0, 0, 0))
            continue
        hits += 1
        L, p = hit
        hw = np.array(p, np.float64) + nodes["mPos"][L]
        sv = _phong(_normal(br, L, p, -1), hw, light)
        # the oracle rounds c = p + 0.5 in FP32 (part of the defined sample
        # positions); scipy interpolates at the exact p + 0.5
        np.testing.assert_allclose(rgba64[px // 8, px % 8, :3], sv, rtol=1e-5)
        np.testing.assert_allclose(rgba32[px // 8, px % 8, :3], sv, rtol=2e-5)
    assert hits > 10


def test_levelset_surface_brute_force():
    """SHADE_LEVELSET on a gathered level set: first sample < thresh, the
    samples of a leaf starting at its (unsnapped) entry point."""
    pts = _surface_setup()
    v = vol_for(pts, 3.0)
    v.insert_points_subcell(4, pts, 3.0)
    v.gather_level_set(1.5)
    v.update_apron(0, 3.0)
    br, nodes = v.bricks(0), v.nodes(0)
    cam = ortho(8, 8, 36.0, 36.0, 24, z0=-0.3)
    light = (48, 48, -100)
    rp = dict(shade=O.SHADE_LEVELSET, step=0.5, thresh=0.0, light=light, shadow_amt=0.0, backclr=(0, 0, 0, 0))
    rgba32, rgba64 = v.render(cam, rp)
    hits = 0
    for px in range(64):
        x = 36 + ((px % 8) + 0.5) * 3
        y = 36 + ((px // 8) + 0.5) * 3
        hit = None
        for L in column_leaves(nodes, x, y):
            # entry at the brick face + a few epsilons (< 0.01); p.z <= 8 inclusive
            for z in 0.0 + 0.5 * np.arange(17):
                p = (x - nodes["mPos"][L, 0], y - nodes["mPos"][L, 1], z + 0.003)
                if tri_scipy(br[L], p) < 0.0:
                    hit = (L, p)
                    break
            if hit:
                break
        if hit is None:
            np.testing.assert_allclose(rgba64[px // 8, px % 8], (0, 0, 0, 0))
            continue
        hits += 1
        L, p = hit
        hw = np.array(p, np.float64) + nodes["mPos"][L]
        sv = _phong(_normal(br, L, p, +1), hw, light)
        np.testing.assert_allclose(rgba64[px // 8, px % 8, :3], sv, rtol=0, atol=2e-2)
        np.testing.assert_allclose(rgba32[px // 8, px % 8], rgba64[px // 8, px % 8], rtol=2e-5)
    assert hits > 10


def test_render_errors():
    v = O.OracleVolume()
    v.add_channel(0, 1)
    with pytest.raises(O.OracleError):
        v.raycast_ea(ortho(4, 4, 0, 0, 8), EA)   # no topology / atlas


def test_magnitude_adjoints_bound_the_adjoints(cluster):
    """The magnitude forms (every term's absolute value) bound the adjoints
    entry by entry, equal them where every term is non-negative (gT >= 0
    alone, weights and gradients >= 0), and are linear in |gI|, |gT|."""
    v, pos64, w, cam, _, rng = cluster
    r, rp = 2.0, EA
    b64 = v.apron_f64(v.splat_cubic_f64_padded(pos64, r, w), 0.0)
    gI = rng.normal(size=(cam["height"], cam["width"]))
    gT = rng.normal(size=(cam["height"], cam["width"]))
    g = v.raycast_ea_backward(cam, rp, gI, gT, bricks64=b64)
    m = v.raycast_ea_backward(cam, rp, gI, gT, bricks64=b64, mag=True)
    assert (m >= np.abs(g) - 1e-12 * np.abs(m).max()).all()
    assert (m > 0).any()
    m2 = v.raycast_ea_backward(cam, rp, 2 * np.abs(gI), -np.abs(gT), bricks64=b64, mag=True)
    m1 = v.raycast_ea_backward(cam, rp, np.abs(gI), np.abs(gT), bricks64=b64, mag=True)
    mI = v.raycast_ea_backward(cam, rp, np.abs(gI), None, bricks64=b64, mag=True)
    np.testing.assert_allclose(m2, m1 + mI, rtol=1e-12, atol=1e-300)
    # dT/drho has every term of one sign: with gT >= 0 alone, |adjoint| == magnitude
    gpos = np.abs(gT)
    np.testing.assert_allclose(np.abs(v.raycast_ea_backward(cam, rp, np.zeros_like(gI), gpos, bricks64=b64)),
                               v.raycast_ea_backward(cam, rp, np.zeros_like(gI), gpos, bricks64=b64, mag=True),
                               rtol=1e-12, atol=1e-300)
    gv = rng.normal(size=b64.shape)
    dp, dw = v.splat_cubic_backward_f64(pos64, r, gv, w)
    mp, mw = v.splat_cubic_backward_f64(pos64, r, gv, w, mag=True)
    assert (mp >= np.abs(dp) - 1e-12).all() and (mw >= np.abs(dw) - 1e-12).all()
    _, mw_pos = v.splat_cubic_backward_f64(pos64, r, np.abs(gv), w, mag=True)
    _, dw_pos = v.splat_cubic_backward_f64(pos64, r, np.abs(gv), w)
    np.testing.assert_allclose(mw_pos, dw_pos, rtol=1e-12, atol=1e-300)
