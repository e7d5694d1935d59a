"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Topology (node records, child lists, atlas map, counts, top level, bounds),
subcell list length, level-set / density gathers, apron, InsertPoints and
ScatterDensity are compared bit-exactly.  The cubic splat sums in a fixed
order that differs from the oracle's ascending point order, so it is checked
to 1e-5 relative error against the FP64 evaluation (absolute floor
1e-6 * max|rho|), and bit-exactly against itself across runs.
"""
import sys
import zlib

import numpy as np
import pytest

import oracle as O
from tolerance import check_entries, splat_ops

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _gvdbx():
    import gvdbx
    return gvdbx


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build_pair(pts, radius, log2dims=(3, 3, 3, 3, 5), orig=(0.0, 0.0, 0.0)):
    G = _gvdbx()
    dev = _dev()
    gv = G.VolumeGVDB(0)
    gv.configure(log2dims)
    gv.SetChannelDefault(16, 16, 1)
    gv.AddChannel(0, G.T_FLOAT, 1)
    tp = torch.from_numpy(np.ascontiguousarray(pts, dtype=np.float32)).to(dev)
    gv.SetPoints(tp)
    gv.RebuildTopology(len(pts), radius, orig)
    gv.FinishTopology(False, True)
    gv.UpdateAtlas()
    ov = O.OracleVolume(log2dims)
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, radius, orig)
    ov.update_atlas()
    return gv, ov, tp


def assert_same_topology(gv, ov, points_bbox=True):
    """points_bbox=False skips pos_min / pos_max (the points' box, which a
    loaded VBX file does not carry)."""
    gi, oi = gv.info(), ov.info()
    assert gi.depth == oi.depth
    assert gi.top_lev == oi.top_lev
    assert list(gi.root_pos) == list(oi.root_pos)
    assert list(gi.node_cnt) == list(oi.node_cnt)
    if points_bbox:
        assert list(gi.pos_min) == list(oi.pos_min) and list(gi.pos_max) == list(oi.pos_max)
    assert list(gi.obj_min) == list(oi.obj_min) and list(gi.obj_max) == list(oi.obj_max)
    assert list(gi.subdim) == list(oi.subdim)
    for l in range(gi.depth + 1):
        gn, on = gv.nodes(l), ov.nodes(l)
        for f in ("mLev", "mFlags", "mPos", "mValue", "mParent", "mChildList"):
            np.testing.assert_array_equal(gn[f], on[f], err_msg=f"level {l} field {f}")
        if l >= 1:
            np.testing.assert_array_equal(gv.child_lists(l), ov.child_lists(l), err_msg=f"child lists level {l}")
    np.testing.assert_array_equal(gv.atlas_map(), ov.atlas_map())


def _shell(rng, n, c, r0, r1):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rad = np.cbrt(rng.uniform(r0 ** 3, r1 ** 3, n))
    return (c + d * rad[:, None]).astype(np.float32)


TOPO_CASES = [
    ("box_r3", lambda r: r.random((20000, 3)) * [120, 64, 100] + [30, 40, 20], 3.0, (3, 3, 3, 3, 5)),
    ("box_r0.5", lambda r: r.random((20000, 3)) * [120, 64, 100] + [30, 40, 20], 0.5, (3, 3, 3, 3, 5)),
    ("shell_r4", lambda r: _shell(r, 30000, 120, 80, 96), 4.0, (3, 3, 3, 3, 5)),
    ("clusters_r2", lambda r: np.abs(np.concatenate([r.normal(c, 5.0, (4000, 3)) for c in
                                                    ([60, 60, 60], [150, 90, 40], [40, 200, 120])])), 2.0,
     (3, 3, 3, 3, 5)),
    ("top_below_root", lambda r: r.random((3000, 3)) * 60.9 + 2, 0.5, (3, 3, 3, 3, 5)),
    ("cfg2_shape_small", lambda r: r.random((60000, 3)) * [512, 256, 512] + 64 + (r.random((60000, 3)) - 0.5), 3.0,
     (3, 3, 3, 3, 5)),
    ("sample_cfg_32_bricks", lambda r: r.random((20000, 3)) * [150, 90, 120] + 3, 3.0, (5, 3, 3, 3, 3)),
    ("small_bricks_4", lambda r: r.random((5000, 3)) * 30 + 2, 1.5, (2, 2, 2, 2, 3)),
]


@pytest.mark.parametrize("name,gen,radius,log2dims", TOPO_CASES, ids=[c[0] for c in TOPO_CASES])
def test_topology_parity(name, gen, radius, log2dims):
    _log(f"[gpu] topology {name}")
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    pts = np.asarray(gen(rng), dtype=np.float32)
    gv, ov, _ = build_pair(pts, radius, log2dims)
    assert_same_topology(gv, ov)


def test_topology_invalid_points():
    _log("[gpu] topology invalid points")
    rng = np.random.default_rng(9)
    pts = (rng.random((20000, 3)) * 100 + 10).astype(np.float32)
    pts[::7, 2] = 1.0e10
    pts[1::11, 0] = -3.0
    pts[2::13, 1] = np.nan
    gv, ov, _ = build_pair(pts, 3.0)
    assert_same_topology(gv, ov)


@pytest.mark.parametrize("mode,R,r,boundval", [(1, 3.0, 1.5, 3.0), (0, 4.0, 2.0, 0.0), (1, 2.0, 1.0, 3.0),
                                               (0, 1.0, 2.0, 0.0)])
def test_gather_and_apron_parity(mode, R, r, boundval):
    _log(f"[gpu] gather mode={mode} R={R} r={r}")
    rng = np.random.default_rng(int(R * 100 + r * 10 + mode))
    pts = (rng.random((30000, 3)) * [90, 70, 80] + [20, 12, 30]).astype(np.float32)
    gv, ov, tp = build_pair(pts, R)
    gl = gv.InsertPointsSubcell(4, len(pts), R, (0, 0, 0))
    ol = ov.insert_points_subcell(4, pts, R)
    assert gl == ol
    if mode == 1:
        gv.GatherLevelSet(4, len(pts), r, (0, 0, 0), gl, 0)
        ov.gather_level_set(r)
    else:
        gv.GatherDensity(4, len(pts), r, (0, 0, 0), gl, 0)
        ov.gather_density(r)
    gb, ob = gv.bricks(), ov.bricks()
    np.testing.assert_array_equal(gb[:, 1:9, 1:9, 1:9], ob[:, 1:9, 1:9, 1:9])
    gv.UpdateApron(0, boundval)
    ov.update_apron(0, boundval)
    np.testing.assert_array_equal(gv.bricks(), ov.bricks())


def test_gather_accumulate_parity():
    _log("[gpu] gather accumulate")
    rng = np.random.default_rng(77)
    pts = (rng.random((20000, 3)) * 80 + 20).astype(np.float32)
    gv, ov, tp = build_pair(pts, 3.0)
    gl = gv.InsertPointsSubcell(4, len(pts), 3.0)
    ov.insert_points_subcell(4, pts, 3.0)
    gv.FillChannel(0, 2.5)
    ov.fill_channel(0, 2.5)
    gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), gl, 0, -1, True)
    ov.gather_level_set(1.5, accum=True)
    np.testing.assert_array_equal(gv.bricks()[:, 1:9, 1:9, 1:9], ov.bricks()[:, 1:9, 1:9, 1:9])


def test_levelset_fp64_tolerance():
    """Level set vs the FP64 evaluation: |err| <= 1e-5 * max(|ref|, r).

    phi = |p-v| - r cancels near the surface, so the absolute floor is the
    scale of the operands of that subtraction (r = 1.5), not of the result."""
    _log("[gpu] level set fp64")
    rng = np.random.default_rng(5)
    pts = (rng.random((20000, 3)) * 60 + 20).astype(np.float32)
    gv, ov, tp = build_pair(pts, 3.0)
    gl = gv.InsertPointsSubcell(4, len(pts), 3.0)
    ov.insert_points_subcell(4, pts, 3.0)
    gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), gl, 0)
    ref = ov.gather_f64(1, 1.5)
    got = gv.bricks()[:, 1:9, 1:9, 1:9].astype(np.float64)
    bound = 1e-5 * np.maximum(np.abs(ref), 1.5)
    err = np.abs(got - ref)
    assert (err <= bound).all(), f"max err {err.max():.3e}, worst ratio {(err / bound).max():.3f}"


def test_insert_points_parity():
    _log("[gpu] insert points")
    rng = np.random.default_rng(12)
    pts = (rng.random((20000, 3)) * 70 + 9).astype(np.float32)
    pts[::17, 2] = 1.0e10
    gv, ov, tp = build_pair(pts, 0.5)
    gv.InsertPoints(len(pts), (0, 0, 0))
    gp, go = gv.point_nodes(len(pts))
    op, oo = ov.insert_points(pts)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(go, oo)


def test_scatter_density_parity():
    _log("[gpu] scatter density")
    rng = np.random.default_rng(13)
    pts = (rng.random((20000, 3)) * 60 + 9).astype(np.float32)
    gv, ov, tp = build_pair(pts, 1.0)
    gv.InsertPoints(len(pts))
    gv.ScatterDensity(len(pts), 2.0, 1.0, (0, 0, 0), True)
    ov.scatter_density(pts, 2.0, expand=True)
    np.testing.assert_array_equal(gv.bricks(), ov.bricks())


def test_splat_cubic_fp64():
    _log("[gpu] cubic splat")
    rng = np.random.default_rng(14)
    pts = np.abs(np.concatenate([rng.normal(c, 6.0, (8000, 3)) for c in ([50, 50, 50], [90, 70, 40])])
                 ).astype(np.float32)
    w = rng.uniform(0.1, 1.0, len(pts)).astype(np.float32)
    gv, ov, tp = build_pair(pts, 4.0)
    tw = torch.from_numpy(w).to(tp.device)
    gv.SetPointWeights(tw)
    gv.SplatCubic(len(pts), 2.0)
    got = gv.bricks()[:, 1:9, 1:9, 1:9].astype(np.float64)
    ref = ov.splat_cubic_f64(pts, 2.0, weights=w)
    floor = 1e-6 * np.abs(ref).max()
    err = np.abs(got - ref)
    assert (err <= 1e-5 * np.abs(ref) + floor).all(), f"max err {err.max():.3e}"
    ov.splat_cubic(pts, 2.0, weights=w)
    np.testing.assert_allclose(got, ov.bricks()[:, 1:9, 1:9, 1:9], rtol=1e-5, atol=floor)
    gv.SplatCubic(len(pts), 2.0)
    np.testing.assert_array_equal(gv.bricks()[:, 1:9, 1:9, 1:9], got.astype(np.float32))


def test_splat_cubic_variants_and_runs():
    """Cell push (points in runs of `splat_chunk` per wavefront, per-run halo
    grids collected per leaf) and tile pull both stay within the FP64 bound
    for a dense cluster, for radii with and without a power-of-two r^2, for
    every run size; each is bit-identical on repeat and accumulates onto the
    channel."""
    _log("[gpu] cubic splat variants / runs")
    rng = np.random.default_rng(21)
    dense = rng.normal([40.0, 40.0, 40.0], 2.5, (40000, 3))
    sparse = rng.uniform(9.0, 80.0, (5000, 3))
    pts = np.concatenate([dense, sparse]).astype(np.float32)
    w = rng.uniform(0.1, 1.0, len(pts)).astype(np.float32)
    gv, ov, tp = build_pair(pts, 4.0)
    tw = torch.from_numpy(w).to(tp.device)
    gv.SetPointWeights(tw)
    for radius in (2.0, 1.5):
        ref = ov.splat_cubic_f64(pts, radius, weights=w)
        floor = 1e-6 * np.abs(ref).max()

        def check(got, expect):
            err = np.abs(got.astype(np.float64) - expect)
            assert (err <= 1e-5 * np.abs(expect) + floor).all(), f"r={radius}: max err {err.max():.3e}"

        for variant, chunk in ((1, 4096), (1, 64), (1, 1 << 20), (0, 4096)):
            gv.SetTuning("splat_variant", variant)
            gv.SetTuning("splat_chunk", chunk)
            assert gv.GetTuning("splat_chunk") == chunk and gv.GetTuning("splat_variant") == variant
            gv.SplatCubic(len(pts), radius)
            got = gv.bricks()[:, 1:9, 1:9, 1:9]
            check(got, ref)
            gv.SplatCubic(len(pts), radius)
            np.testing.assert_array_equal(gv.bricks()[:, 1:9, 1:9, 1:9], got)
            gv.FillChannel(0, 0.25)
            gv.SplatCubic(len(pts), radius, accum=True)
            check(gv.bricks()[:, 1:9, 1:9, 1:9], ref + 0.25)
    # the adjoint under both forms and several run lengths, against the FP64 adjoint
    inf = gv.info()
    rng = np.random.default_rng(22)
    gr = rng.normal(size=(inf.node_cnt[0], inf.brick_res, inf.brick_res, inf.brick_res)).astype(np.float32)
    gt = torch.from_numpy(gr).to(tp.device)
    rdpos, rdw = ov.splat_cubic_backward_f64(pts.astype(np.float64), 2.0, gr.astype(np.float64), w.astype(np.float64))
    mdpos, mdw = ov.splat_cubic_backward_f64(pts.astype(np.float64), 2.0, gr.astype(np.float64), w.astype(np.float64),
                                             mag=True)
    for variant, chunk in ((1, 1024), (1, 64), (0, 1024)):
        gv.SetTuning("splat_variant", variant)
        gv.SetTuning("splat_chunk", chunk)
        dpos, dw = gv.SplatCubicBackward(len(pts), 2.0, (0, 0, 0), gt)
        dpos2, dw2 = gv.SplatCubicBackward(len(pts), 2.0, (0, 0, 0), gt)
        torch.cuda.synchronize()
        assert torch.equal(dpos, dpos2) and torch.equal(dw, dw2)
        for got, ref, mag, name in ((dpos.cpu().numpy(), rdpos, mdpos, "dpos"), (dw.cpu().numpy(), rdw, mdw, "dw")):
            check_entries(got, ref, mag, splat_ops(2.0), f"{name} variant {variant} chunk {chunk}")
    gv.SetTuning("splat_variant", 1)
    gv.SetTuning("splat_chunk", 1024)
    with pytest.raises(Exception):
        gv.SetTuning("splat_chunk", 10)
    with pytest.raises(Exception):
        gv.SetTuning("splat_variant", 2)


def test_error_paths_leave_volume_unchanged():
    _log("[gpu] error paths")
    G = _gvdbx()
    rng = np.random.default_rng(15)
    pts = (rng.random((5000, 3)) * 60 + 9).astype(np.float32)
    gv, ov, tp = build_pair(pts, 3.0)
    before = gv.nodes(0)
    tiny = torch.tensor([[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]], device=tp.device)
    gv.SetPoints(tiny)
    with pytest.raises(G.GvdbxError) as e:
        gv.RebuildTopology(2, 3.0)
    assert e.value.code == 3
    bad = torch.full((4, 3), float("nan"), device=tp.device)
    gv.SetPoints(bad)
    with pytest.raises(G.GvdbxError) as e:
        gv.RebuildTopology(4, 3.0)
    assert e.value.code == 4
    np.testing.assert_array_equal(gv.nodes(0), before)
    with pytest.raises(G.GvdbxError):
        gv.GatherLevelSet(4, 4, 1.5, (0, 0, 0), 0, 0)   # no InsertPointsSubcell yet
    with pytest.raises(G.GvdbxError):
        gv.InsertPointsSubcell(4, 4, 9.0)                # box radius beyond the brick halo


def test_full_size_cfg2_properties():
    """16M-point frame: determinism, level-set bounds, topology size invariants."""
    _log("[gpu] full size config-2 frame")
    G = _gvdbx()
    dev = _dev()
    n = 1 << 24
    pts = G.gen_box(20261205, n, (64, 64, 64), (576, 320, 576), 0.5)
    tp = torch.from_numpy(pts).to(dev)
    gv = G.VolumeGVDB(0)
    gv.configure((3, 3, 3, 3, 5))
    gv.AddChannel(0, G.T_FLOAT, 1)
    gv.SetPoints(tp)
    outs = []
    for rep in range(2):
        gv.RebuildTopology(n, 3.0)
        gv.FinishTopology()
        gv.UpdateAtlas()
        ln = gv.InsertPointsSubcell(4, n, 3.0)
        gv.GatherLevelSet(4, n, 1.5, (0, 0, 0), ln, 0)
        gv.UpdateApron(0, 3.0)
        inf = gv.info()
        nb = inf.node_cnt[0]
        b = gv.bricks()
        outs.append((ln, nb, float(b.astype(np.float64).sum()), float(b.min()), float(b.max())))
        del b
    assert outs[0] == outs[1]
    ln, nb, s, mn, mx = outs[0]
    assert inf.depth == 3
    assert mx <= 3.0 and mn >= -1.5
    assert 130000 < nb < 170000
    # every leaf is the cell of some point or within the extra-brick box of one
    assert inf.node_cnt[3] == 1


@pytest.mark.parametrize("log2dims,sc,R", [((3, 3, 3, 3, 5), 4, 3.0), ((3, 3, 3, 3, 5), 2, 3.0),
                                           ((4, 3, 3, 3, 5), 4, 3.0), ((2, 2, 2, 2, 3), 2, 1.0)])
def test_gather_split_bit_identical(log2dims, sc, R):
    """Dense tiles cut into work items (gather_split), with and without the
    reach-class grouping of the bins (gather_classes), give the oracle's bits:
    a dense Gaussian cluster (~600 candidate chunks per tile, bins of
    thousands of points) beside a sparse box, level set and density, fresh
    and accumulated."""
    _log(f"[gpu] gather split {log2dims} sc={sc}")
    rng = np.random.default_rng(4242)
    a = rng.normal([40, 40, 40], 4.0, (40000, 3))
    b = rng.random((10000, 3)) * [90, 70, 80] + [10, 8, 12]
    pts = np.abs(np.concatenate([a, b])).astype(np.float32)
    gv, ov, tp = build_pair(pts, R, log2dims)
    gl = gv.InsertPointsSubcell(sc, len(pts), R)
    assert gl == ov.insert_points_subcell(sc, pts, R)
    assert gv.GetTuning("gather_split") >= 0
    for mode in (1, 0):
        gather = gv.GatherLevelSet if mode == 1 else gv.GatherDensity
        ogather = ov.gather_level_set if mode == 1 else ov.gather_density
        ogather(1.5)
        ref = ov.bricks()[:, 1:-1, 1:-1, 1:-1]
        ov.fill_channel(0, 0.25)
        ogather(1.5, accum=True)
        acc_ref = ov.bricks()[:, 1:-1, 1:-1, 1:-1]
        for classes in (1, 0):
            gv.SetTuning("gather_classes", classes)
            for split in (0, 1, 5, 64, 1 << 20):
                gv.SetTuning("gather_split", split)
                gv.FillChannel(0, -7.0)
                gather(sc, len(pts), 1.5, (0, 0, 0), gl, 0)
                np.testing.assert_array_equal(gv.bricks()[:, 1:-1, 1:-1, 1:-1], ref,
                                              err_msg=f"mode {mode} classes {classes} split {split}")
                gv.FillChannel(0, 0.25)
                gather(sc, len(pts), 1.5, (0, 0, 0), gl, 0, accum=True)
                np.testing.assert_array_equal(gv.bricks()[:, 1:-1, 1:-1, 1:-1], acc_ref,
                                              err_msg=f"accumulate, mode {mode} classes {classes} split {split}")
    assert gv.GetTuning("gather_classes") == 0
    with pytest.raises(Exception):
        gv.SetTuning("gather_split", -1)
    with pytest.raises(Exception):
        gv.SetTuning("gather_split", (1 << 20) + 1)
    with pytest.raises(Exception):
        gv.SetTuning("gather_classes", 2)


@pytest.mark.parametrize("sc", [1, 2])
def test_gather_classes_both_faces(sc):
    """8^3 bricks, radius 4: the membership range of a point 3..4 voxels into
    a bin reaches past both of its faces along an axis (the accumulating
    gathers, whose cutoff is unbounded, keep all of them; the level-set
    cutoff 4.5 a thin slice), so the reach-class gather must also take such
    points from the bin above.  Class path, bin path and oracle agree bit for
    bit, level set and density, fresh and accumulated."""
    _log(f"[gpu] gather classes both faces sc={sc}")
    R = 4.0
    rng = np.random.default_rng(77 + sc)
    pts = (rng.random((30000, 3)) * [70, 50, 60] + 5).astype(np.float32)
    # membership range of InsertPointsSubcell (trunc toward zero, subcell multiples)
    lo = np.floor(np.trunc(pts - R) / sc) * sc
    hi = np.floor(np.trunc(pts + 1 + R) / sc) * sc + sc - 1
    o = np.floor(pts / 8) * 8
    assert ((lo < o) & (hi >= o + 8)).any(), "no point reaches both faces: the case is not covered"
    gv, ov, tp = build_pair(pts, R)
    gl = gv.InsertPointsSubcell(sc, len(pts), R)
    assert gl == ov.insert_points_subcell(sc, pts, R)
    for mode in (1, 0):
        gather = gv.GatherLevelSet if mode == 1 else gv.GatherDensity
        ogather = ov.gather_level_set if mode == 1 else ov.gather_density
        ogather(1.5)
        ref = ov.bricks()[:, 1:-1, 1:-1, 1:-1]
        ov.fill_channel(0, 0.25)
        ogather(1.5, accum=True)
        acc_ref = ov.bricks()[:, 1:-1, 1:-1, 1:-1]
        for classes in (1, 0):
            gv.SetTuning("gather_classes", classes)
            gv.FillChannel(0, -7.0)
            gather(sc, len(pts), 1.5, (0, 0, 0), gl, 0)
            np.testing.assert_array_equal(gv.bricks()[:, 1:-1, 1:-1, 1:-1], ref, err_msg=f"mode {mode} classes {classes}")
            gv.FillChannel(0, 0.25)
            gather(sc, len(pts), 1.5, (0, 0, 0), gl, 0, accum=True)
            np.testing.assert_array_equal(gv.bricks()[:, 1:-1, 1:-1, 1:-1], acc_ref,
                                          err_msg=f"accumulate, mode {mode} classes {classes}")


@pytest.mark.parametrize("mode", [1, 0])
def test_gather_wide_box_parity(mode):
    """32^3 bricks, subcell 16, radius 14: ceil(R) + sc > 29, so the gather
    filters without the packed per-point ranges; oracle bits, fresh and
    accumulated."""
    _log(f"[gpu] gather wide box mode={mode}")
    rng = np.random.default_rng(515 + mode)
    pts = (rng.random((4000, 3)) * [150, 90, 120] + 3).astype(np.float32)
    gv, ov, tp = build_pair(pts, 14.0, (5, 3, 3, 3, 3))
    gl = gv.InsertPointsSubcell(16, len(pts), 14.0)
    assert gl == ov.insert_points_subcell(16, pts, 14.0)
    gather = gv.GatherLevelSet if mode == 1 else gv.GatherDensity
    ogather = ov.gather_level_set if mode == 1 else ov.gather_density
    ogather(2.0)
    gv.FillChannel(0, -7.0)
    gather(16, len(pts), 2.0, (0, 0, 0), gl, 0)
    np.testing.assert_array_equal(gv.bricks()[:, 1:-1, 1:-1, 1:-1], ov.bricks()[:, 1:-1, 1:-1, 1:-1])
    ov.fill_channel(0, 0.25)
    ogather(2.0, accum=True)
    gv.FillChannel(0, 0.25)
    gather(16, len(pts), 2.0, (0, 0, 0), gl, 0, accum=True)
    np.testing.assert_array_equal(gv.bricks()[:, 1:-1, 1:-1, 1:-1], ov.bricks()[:, 1:-1, 1:-1, 1:-1])


@pytest.mark.parametrize("mode", [1, 0])
def test_gather_variants_bit_identical(mode):
    """All gather kernel variants produce the oracle's bits."""
    _log(f"[gpu] gather variants mode={mode}")
    rng = np.random.default_rng(101 + mode)
    pts = (rng.random((40000, 3)) * [100, 60, 90] + [17, 25, 9]).astype(np.float32)
    gv, ov, tp = build_pair(pts, 3.0)
    gl = gv.InsertPointsSubcell(4, len(pts), 3.0)
    ov.insert_points_subcell(4, pts, 3.0)
    if mode == 1:
        ov.gather_level_set(1.5)
    else:
        ov.gather_density(1.5)
    ref = ov.bricks()[:, 1:9, 1:9, 1:9]
    acc_ref = None
    for var in (0, 1, 2):
        gv.SetTuning("gather_variant", var)
        gv.FillChannel(0, -7.0)
        if mode == 1:
            gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), gl, 0)
        else:
            gv.GatherDensity(4, len(pts), 1.5, (0, 0, 0), gl, 0)
        np.testing.assert_array_equal(gv.bricks()[:, 1:9, 1:9, 1:9], ref, err_msg=f"variant {var}")
        # accumulate mode (no cutoff pruning) on a non-trivial initial field
        gv.FillChannel(0, 0.25)
        if mode == 1:
            gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), gl, 0, accum=True)
        else:
            gv.GatherDensity(4, len(pts), 1.5, (0, 0, 0), gl, 0, accum=True)
        got = gv.bricks()[:, 1:9, 1:9, 1:9]
        if acc_ref is None:
            acc_ref = got
        np.testing.assert_array_equal(got, acc_ref, err_msg=f"accumulate, variant {var}")


@pytest.mark.parametrize("log2dims,sc,R,split", [((3, 3, 3, 3, 5), 4, 3.0, 16), ((3, 3, 3, 3, 5), 2, 2.0, 1),
                                                 ((4, 3, 3, 3, 5), 4, 3.0, 0)])
def test_fp16_subcell_parity(log2dims, sc, R, split):
    """InsertPointsSubcell_FP16 + GatherLevelSet_FP16 (positions read back
    from 16-bit fractions of the bounding box, membership from the inserted
    points) equal the oracle bit for bit, with and without split tiles; the
    FP16 and FP32 paths refuse each other's lists, and colours."""
    G = _gvdbx()
    _log(f"[gpu] fp16 subcells {log2dims} sc={sc} split={split}")
    rng = np.random.default_rng(808 + sc)
    a = rng.normal([40, 40, 40], 4.0, (30000, 3))
    b = rng.random((20000, 3)) * [110, 70, 90] + [10, 8, 12]
    pts = np.abs(np.concatenate([a, b])).astype(np.float32)
    gv, ov, tp = build_pair(pts, R, log2dims)
    gv.SetTuning("gather_split", split)
    gl = gv.InsertPointsSubcell_FP16(sc, len(pts), R)
    assert gl == ov.insert_points_subcell_fp16(sc, pts, R)
    with pytest.raises(G.GvdbxError) as e:
        gv.GatherLevelSet(sc, len(pts), 1.5, (0, 0, 0), gl, 0)
    assert e.value.code == 2
    gv.FillChannel(0, -7.0)
    gv.GatherLevelSet_FP16(sc, len(pts), 1.5, (0, 0, 0), gl, 0)
    ov.gather_level_set(1.5)
    r0 = 1 << log2dims[0]
    inner = (slice(None), slice(1, r0 + 1), slice(1, r0 + 1), slice(1, r0 + 1))
    np.testing.assert_array_equal(gv.bricks()[inner], ov.bricks()[inner])
    gv.UpdateApron(0, 3.0)
    ov.update_apron(0, 3.0)
    np.testing.assert_array_equal(gv.bricks(), ov.bricks())
    gl32 = gv.InsertPointsSubcell(sc, len(pts), R)
    assert gl32 == gl
    with pytest.raises(G.GvdbxError) as e:
        gv.GatherLevelSet_FP16(sc, len(pts), 1.5, (0, 0, 0), gl, 0)
    assert e.value.code == 2
    gv.AddChannel(1, G.T_UCHAR4, 1)
    tc = torch.zeros((len(pts), 4), dtype=torch.uint8, device=tp.device)
    gv.SetPoints(tp, None, tc)
    gv.InsertPointsSubcell_FP16(sc, len(pts), R)
    with pytest.raises(G.GvdbxError) as e:
        gv.GatherLevelSet_FP16(sc, len(pts), 1.5, (0, 0, 0), gl, 0, 1)
    assert e.value.code == 1


def test_apron_orders_bit_identical():
    """UpdateApron walking bricks in leaf order and in cell order (x
    fastest) writes the same words, for both boundary values."""
    _log("[gpu] apron orders")
    rng = np.random.default_rng(31)
    pts = (rng.random((30000, 3)) * [90, 70, 80] + [20, 12, 30]).astype(np.float32)
    gv, ov, tp = build_pair(pts, 3.0)
    gl = gv.InsertPointsSubcell(4, len(pts), 3.0)
    gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), gl, 0)
    for bv in (3.0, 0.0):
        out = []
        for order in (1, 0):
            gv.SetTuning("apron_order", order)
            gv.UpdateApron(0, bv)
            out.append(gv.bricks())
        np.testing.assert_array_equal(out[0], out[1])
    ov.insert_points_subcell(4, pts, 3.0)
    ov.gather_level_set(1.5)
    ov.update_apron(0, 0.0)
    np.testing.assert_array_equal(out[0], ov.bricks())


@pytest.mark.parametrize("case", ["levelset", "density", "accum", "split", "fp16", "radix", "unclassed"])
def test_apron_push_equals_full_apron(case):
    """A gather on 8^3 bricks also writes its boundary voxels into the
    neighbours' aprons, and UpdateApron then only writes the boundary value
    where no neighbour leaf is: after UpdateApron the bricks equal those of
    the full apron copy (apron_push 0) bit for bit, for every gather form,
    and the oracle's."""
    import gvdbx as G
    _log(f"[gpu] apron push {case}")
    rng = np.random.default_rng(41)
    pts = np.abs(np.concatenate([rng.normal(c, 7.0, (9000, 3)) for c in ([40, 40, 40], [90, 60, 50], [60, 110, 80])])
                 ).astype(np.float32)
    out = []
    for push in (1, 0):
        gv = G.VolumeGVDB(0)
        gv.configure((3, 3, 3, 3, 5))
        gv.AddChannel(0, G.T_FLOAT, 1)
        if case == "radix":
            gv.SetTopologyMode(G.TOPO_RADIX)
        tp = torch.from_numpy(pts).to(_dev())
        gv.SetPoints(tp)
        gv.RebuildTopology(len(pts), 3.0)
        gv.FinishTopology(False, True)
        gv.UpdateAtlas()
        gv.SetTuning("apron_push", push)
        if case == "split":
            gv.SetTuning("gather_split", 1)
        if case == "unclassed":
            gv.SetTuning("gather_classes", 0)
        if case == "fp16":
            ln = gv.InsertPointsSubcell_FP16(4, len(pts), 3.0)
            gv.GatherLevelSet_FP16(4, len(pts), 1.5, (0, 0, 0), ln, 0)
        else:
            ln = gv.InsertPointsSubcell(4, len(pts), 3.0)
            if case == "density":
                gv.GatherDensity(4, len(pts), 1.5, (0, 0, 0), ln, 0)
            elif case == "accum":
                gv.FillChannel(0, 0.75)
                gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), ln, 0, -1, True)
            else:
                gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), ln, 0)
        gv.UpdateApron(0, 3.0 if case != "density" else 0.0)
        b1 = gv.bricks()
        gv.UpdateApron(0, -2.0)   # another boundary value on the same pushed aprons
        out.append((b1, gv.bricks()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    if case == "levelset":
        ov = O.OracleVolume((3, 3, 3, 3, 5))
        ov.add_channel(0, 1)
        ov.rebuild_topology(pts, 3.0)
        ov.update_atlas()
        ov.insert_points_subcell(4, pts, 3.0)
        ov.gather_level_set(1.5)
        ov.update_apron(0, 3.0)
        np.testing.assert_array_equal(out[0][0], ov.bricks())


def test_apron_after_other_writes_copies_again():
    """A channel written after a pushing gather (splat, fill) gets the full
    apron copy from UpdateApron, as the reference does."""
    import gvdbx as G
    _log("[gpu] apron after other writes")
    rng = np.random.default_rng(42)
    pts = (rng.random((20000, 3)) * [80, 60, 70] + 20).astype(np.float32)
    w = rng.uniform(0.1, 1.0, len(pts)).astype(np.float32)
    gv, ov, tp = build_pair(pts, 3.0)
    gv.SetPointWeights(torch.from_numpy(w).to(tp.device))
    ln = gv.InsertPointsSubcell(4, len(pts), 3.0)
    gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), ln, 0)
    gv.SplatCubic(len(pts), 2.0, (0, 0, 0), 0, True)   # accumulates onto the level set
    gv.UpdateApron(0, 3.0)
    got = gv.bricks()
    gv.SetTuning("apron_push", 0)
    gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), ln, 0)
    gv.SplatCubic(len(pts), 2.0, (0, 0, 0), 0, True)
    gv.UpdateApron(0, 3.0)
    np.testing.assert_array_equal(got, gv.bricks())
    gv.SetTuning("apron_push", 1)
    gv.GatherLevelSet(4, len(pts), 1.5, (0, 0, 0), ln, 0)
    gv.FillChannel(0, 0.5)
    gv.UpdateApron(0, 3.0)
    ov.fill_channel(0, 0.5)
    ov.update_apron(0, 3.0)
    np.testing.assert_array_equal(gv.bricks(), ov.bricks())


@pytest.mark.parametrize("mode,R,r,sc,accum,split", [(1, 3.0, 1.5, 4, False, 16), (0, 4.0, 2.0, 4, False, 16),
                                                     (1, 3.0, 1.5, 4, True, 0), (1, 2.0, 1.0, 4, False, 1),
                                                     (0, 3.0, 0.5, 4, False, 0)])
def test_gather_quarters_bit_identical(mode, R, r, sc, accum, split):
    """z pruned per 2-voxel quarter instead of per 4-voxel half: the same bits
    (the oracle's) for both gathers, accumulate and split tiles."""
    _log(f"[gpu] gather quarters mode={mode} R={R} sc={sc} accum={accum} split={split}")
    rng = np.random.default_rng(int(R * 7 + sc + split))
    pts = np.abs(np.concatenate([rng.normal(c, 6.0, (8000, 3)) for c in ([40, 40, 40], [80, 60, 50], [50, 90, 70])])
                 ).astype(np.float32)
    gv, ov, tp = build_pair(pts, R)
    gl = gv.InsertPointsSubcell(sc, len(pts), R)
    ov.insert_points_subcell(sc, pts, R)
    gv.SetTuning("gather_split", split)
    out = []
    for q in (0, 1):
        gv.SetTuning("gather_quarters", q)
        gv.FillChannel(0, 1.25)
        if mode == 1:
            gv.GatherLevelSet(sc, len(pts), r, (0, 0, 0), gl, 0, -1, accum)
        else:
            gv.GatherDensity(sc, len(pts), r, (0, 0, 0), gl, 0, -1, accum)
        out.append(gv.bricks()[:, 1:9, 1:9, 1:9])
    np.testing.assert_array_equal(out[0], out[1])
    ov.fill_channel(0, 1.25)
    if mode == 1:
        ov.gather_level_set(r, accum=accum)
    else:
        ov.gather_density(r, accum=accum)
    np.testing.assert_array_equal(out[1], ov.bricks()[:, 1:9, 1:9, 1:9])
