"""Pins the oracle's colour channels (CPU only).

Colour means of GatherDensity / GatherLevelSet (cuda_gvdb_particles.cuh:
718-731,757-771) and the colour part of ScatterDensity
(cuda_gvdb_particles.cuh:554-572,610-642) against an independent numpy
formulation: subcell lists rebuilt by brute-force box membership
(tests/brute.py), exact integer sums, per-voxel grouping for the scatter; and
against closed forms (a uniform colour averages to itself within one step of
truncation, alpha saturates; an empty list gives (0,0,0,255)).
"""
import numpy as np

import brute
import oracle as O


def _setup(pts, R, sc, log2dims=(3, 3, 3, 3, 5)):
    ov = O.OracleVolume(log2dims)
    ov.add_channel(0, 1)
    ov.add_channel(1, 1, O.T_UCHAR4)
    ov.rebuild_topology(pts, R)
    ov.update_atlas()
    ln = ov.insert_points_subcell(sc, pts, R)
    return ov, ln


def _leaf_pos(ov):
    return ov.nodes(0)["mPos"].astype(np.int64)


def test_gather_color_counts_and_sums_brute():
    """Sparse points (lists of a few members, so alpha = trunc(255/n) shows
    every count) and random colours: every interior voxel equals the brute
    force over brute-force membership, for sc 4 and 2."""
    rng = np.random.default_rng(31)
    pts = (rng.random((400, 3)) * [40, 30, 36] + 6).astype(np.float32)
    pts[::37, 2] = 1.0e10   # invalid points have no colour in any list
    for sc in (4, 2):
        for alpha_zero in (True, False):
            clr = rng.integers(0, 256, (len(pts), 4)).astype(np.uint8)
            if alpha_zero:
                clr[:, 3] = 0
            ov, _ = _setup(pts, 2.0, sc)
            ov.gather_color(clr, 1)
            got = ov.bricks_u8(1)[:, 1:9, 1:9, 1:9]
            mem = brute.subcell_members(pts, (0, 0, 0), _leaf_pos(ov), 8, sc, 2.0)
            exp = brute.gather_color_expected(mem, clr, 8, sc)
            np.testing.assert_array_equal(got, exp, err_msg=f"sc {sc}")
            counts = np.array([len(m) for lm in mem for m in lm])
            assert (counts == 0).any() and (counts >= 2).any(), "membership cases not covered"
            if alpha_zero:
                # alpha = sat(trunc(255/n)); an empty list gives 255 (1/0)
                n = np.array([[len(m) for m in lm] for lm in mem])
                a = got[..., 3][:, ::sc, ::sc, ::sc].reshape(len(mem), -1)
                with np.errstate(divide="ignore"):
                    want = brute.sat_u8(np.float32(1) / n.astype(np.float32) * np.float32(255))
                np.testing.assert_array_equal(a, want)


def test_gather_color_uniform_closed_form():
    """Every point (c0, c1, c2, 255): a non-empty list averages to c or c-1
    (one truncation step), alpha saturates at 255; empty lists give
    (0, 0, 0, 255)."""
    rng = np.random.default_rng(32)
    pts = (rng.random((3000, 3)) * [50, 40, 45] + 5).astype(np.float32)
    c = np.array([200, 17, 99, 255], np.uint8)
    ov, _ = _setup(pts, 3.0, 4)
    ov.gather_color(np.tile(c, (len(pts), 1)), 1)
    got = ov.bricks_u8(1)[:, 1:9, 1:9, 1:9].reshape(-1, 4).astype(int)
    mem = brute.subcell_members(pts, (0, 0, 0), _leaf_pos(ov), 8, 4, 3.0)
    nonempty = np.array([[len(m) > 0 for m in lm] for lm in mem])
    vox = np.repeat(np.repeat(np.repeat(nonempty.reshape(-1, 2, 2, 2), 4, 1), 4, 2), 4, 3).reshape(-1)
    assert vox.any() and (~vox).any()
    for a in range(3):
        d = c[a] - got[vox, a]
        assert ((d == 0) | (d == 1)).all(), a
    assert (got[vox, 3] == 255).all()
    np.testing.assert_array_equal(got[~vox], np.tile([0, 0, 0, 255], ((~vox).sum(), 1)))


def test_gather_color_accumulate_brute():
    """Accumulate onto a filled channel: init is the stored colour / 255 and
    the divisor is count + 1."""
    rng = np.random.default_rng(33)
    pts = (rng.random((500, 3)) * [40, 30, 36] + 6).astype(np.float32)
    clr = rng.integers(0, 256, (len(pts), 4)).astype(np.uint8)
    ov, _ = _setup(pts, 2.0, 4)
    ov.fill_channel(1, 0.4)
    # FillChannel on uchar4: (0.4 * 255 -> 102) x3, alpha 255 (gvdbOpFillC4)
    init = ov.bricks_u8(1)[:, 1:9, 1:9, 1:9].copy()
    assert (init[..., :3] == 102).all() and (init[..., 3] == 255).all()
    ov.gather_color(clr, 1, accum=True)
    mem = brute.subcell_members(pts, (0, 0, 0), _leaf_pos(ov), 8, 4, 2.0)
    exp = brute.gather_color_expected(mem, clr, 8, 4, init=init)
    np.testing.assert_array_equal(ov.bricks_u8(1)[:, 1:9, 1:9, 1:9], exp)
    ov.fill_channel(1, 0.0)
    assert (ov.bricks_u8(1) == 0).all()


def _home_voxels(ov, pts):
    """Home leaf and voxel of every point by brute force: the leaf whose
    cell holds floor(p / 8); invalid points and points outside leaves -> -1."""
    ok, w = brute.valid_points(pts, (0, 0, 0))
    lp = _leaf_pos(ov)
    lut = brute.leaf_lookup(lp, 8)
    leaf = np.full(len(pts), -1)
    vox = np.zeros((len(pts), 3), np.int64)
    for i in np.nonzero(ok)[0]:
        c = np.floor(w[i] / 8).astype(np.int64)
        k = lut.get(tuple(c.tolist()))
        if k is not None:
            leaf[i] = k
            vox[i] = np.trunc(w[i].astype(np.float32) - lp[k].astype(np.float32)).astype(np.int64)
    return leaf, vox


def test_scatter_color_avg_and_last_brute():
    rng = np.random.default_rng(34)
    pts = (rng.random((6000, 3)) * [30, 24, 28] + 4).astype(np.float32)
    pts[::53, 0] = -2.0
    clr = rng.integers(0, 256, (len(pts), 4)).astype(np.uint8)
    for avg in (True, False):
        ov = O.OracleVolume()
        ov.add_channel(0, 1)
        ov.add_channel(1, 1, O.T_UCHAR4)
        ov.add_channel(3, 1, O.T_UCHAR4)
        ov.rebuild_topology(pts, 0.5)
        ov.update_atlas()
        chan = 1 if avg else 3
        ov.fill_channel(chan, 0.2)
        ov.scatter_color(pts, clr, avg=avg, chan=chan)
        got = ov.bricks_u8(chan)[:, 1:9, 1:9, 1:9]
        leaf, vox = _home_voxels(ov, pts)
        exp = np.zeros_like(got)
        exp[...] = [51, 51, 51, 255]
        sums = {}
        for i in np.nonzero(leaf >= 0)[0]:
            key = (leaf[i], vox[i, 2], vox[i, 1], vox[i, 0])
            if avg:
                s = sums.setdefault(key, np.zeros(4, np.int64))
                s[0] += 1
                s[1:] += clr[i, :3]
            else:
                exp[key] = clr[i]   # ascending i: the last point wins
        for key, s in sums.items():
            exp[key] = [s[1] // s[0], s[2] // s[0], s[3] // s[0], 255]
        assert len(sums) > 0 or not avg
        np.testing.assert_array_equal(got, exp, err_msg=f"avg {avg}")


def test_apron_color_channel_brute():
    """UpdateApron on a uchar4 channel copies whole words and stores
    make_uchar4(b, b, b, b) outside every leaf."""
    rng = np.random.default_rng(35)
    pts = (rng.random((2000, 3)) * [40, 30, 36] + 6).astype(np.float32)
    clr = rng.integers(0, 256, (len(pts), 4)).astype(np.uint8)
    ov, _ = _setup(pts, 2.0, 4)
    ov.gather_color(clr, 1)
    words = ov.bricks(1).view(np.uint32)
    for bv, byte in ((0.0, 0), (3.7, 3), (300.0, 255)):
        ov.update_apron(1, bv)
        got = ov.bricks(1).view(np.uint32)
        bound = np.uint32(byte * 0x01010101)
        exp = brute.apron_expected(words, _leaf_pos(ov), 8, bound)
        np.testing.assert_array_equal(got, exp, err_msg=f"boundval {bv}")
