"""Pins the oracle's FP16 subcell path (InsertPointsSubcell_FP16 +
GatherLevelSet_FP16, cuda_gvdb_particles.cuh:154-216,775-811) on the CPU.

  - brute force: membership from the full-precision points (the reference
    inserts by them), distances from the positions read back from 16-bit
    fractions of the bounding box, in FP64, against the oracle's FP32 result;
  - the list length equals the FP32 lists' (same membership);
  - the FP16 level set differs from the FP32 one by at most the read-back
    error, |dq| <= range / 65535 per axis.
"""
import numpy as np

import brute
import oracle as O


def _fp16_round_trip(w, mn, rg):
    f32 = np.float32
    t = ((w.astype(f32) - f32(mn)) / f32(rg)) * f32(65535.0)
    q = np.where(~(t > 0), 0, np.where(t >= 65535, 65535, np.trunc(np.nan_to_num(t)))).astype(np.uint32)
    return ((q.astype(f32) / f32(65535.0)) * f32(rg) + f32(mn)).astype(f32)


def test_fp16_levelset_brute_and_bound():
    rng = np.random.default_rng(77)
    pts = (rng.random((3000, 3)) * [70, 40, 55] + [9, 17, 30]).astype(np.float32)
    R, r = 3.0, 1.5
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.rebuild_topology(pts, R)
    ov.update_atlas()
    n32 = ov.insert_points_subcell(4, pts, R)
    ov.gather_level_set(r)
    ref32 = ov.bricks(0)[:, 1:9, 1:9, 1:9].astype(np.float64)
    n16 = ov.insert_points_subcell_fp16(4, pts, R)
    assert n16 == n32
    ov.fill_channel(0, -5.0)
    ov.gather_level_set(r)
    got = ov.bricks(0)[:, 1:9, 1:9, 1:9].astype(np.float64)
    inf = ov.info()
    mn = np.array(inf.pos_min, np.float32)
    rg = (np.array(inf.pos_max, np.float32) - mn).astype(np.float32)
    q = np.stack([_fp16_round_trip(pts[:, a], mn[a], rg[a]) for a in range(3)], 1)
    assert (q != pts).any(), "the read-back changes no position: the case is not covered"
    # membership from the original points, distances from the read-back ones
    lp = ov.nodes(0)["mPos"].astype(np.int64)
    lo, hi = brute.subcell_box(pts, R, 4)
    exp = np.full(got.shape, 3.0)
    g = np.arange(8)
    for k in range(len(lp)):
        p0 = lp[k]
        vz, vy, vx = np.meshgrid(g + p0[2], g + p0[1], g + p0[0], indexing="ij")
        v = np.stack([vx.ravel(), vy.ravel(), vz.ravel()], 1)
        s = v - v % 4
        near = ((hi >= p0) & (lo <= p0 + 7)).all(1)
        mem = ((lo[near][None] <= s[:, None]) & (s[:, None] <= hi[near][None])).all(2)
        d = np.sqrt(((q[near].astype(np.float64)[None] - v[:, None]) ** 2).sum(2))
        exp[k] = np.minimum(3.0, np.where(mem, d - r, np.inf).min(1, initial=np.inf)).reshape(8, 8, 8)
    bound = 1e-5 * np.maximum(np.abs(exp), r)
    assert (np.abs(got - exp) <= bound).all(), np.abs(got - exp).max()
    # the FP16 field moves by at most the read-back error
    dq = np.sqrt(3.0) * float(rg.max()) / 65535.0
    assert np.abs(got - ref32).max() <= dq + 1e-5
    assert (got != ref32).any()
