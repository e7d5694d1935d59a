"""Pins the oracle's incremental topology (ActivateIncreBricksGPU,
gvdb_volume_gvdb.cpp:1020-1155; AccumulateTopology 1158-1178) with set
algebra from brute.py, not with its own procedure:

  - accumulate with radius 0 reaches exactly the covering cells of the union
    of the frames, i.e. a full rebuild of the union (as node sets);
  - the non-accumulating update reaches exactly the cells the new frame's
    boxes touch, and leaves every other node unflagged, unlinked and
    childless;
  - with radius >= 1 the reached set is the old one united with the brute
    touched set (accumulate), or the touched set alone;
  - new nodes are appended after the old ones in key order, old nodes keep
    their records;
  - no missing node: nothing changes (the reference's early return);
  - UpdateAtlas gives active leaves atlas slots in leaf order, deactivated
    leaves keep their stale atlas coordinates;
  - GatherLevelSet over an incremental tree equals a full rebuild's on every
    leaf both trees hold.
"""
import numpy as np
import pytest

import oracle as O
from brute import expected_topology, incre_touched, lexsort_cells, reachable, reachable_cells, unique_cells

LOG2 = (3, 3, 3, 3, 5)
RNG = [8, 64, 512, 4096, 131072]
UNDEF64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _box(seed, n, lo, hi):
    r = np.random.default_rng(seed)
    return (r.random((n, 3)) * (np.array(hi) - np.array(lo)) + np.array(lo)).astype(np.float32)


def _covering(points, depth):
    w = points[np.isfinite(points).all(1)]
    return [unique_cells(np.floor(w / np.float32(RNG[l])).astype(np.int64)) for l in range(depth)]


def _vol(points, radius):
    ov = O.OracleVolume(LOG2)
    ov.add_channel(0, 1)
    ov.rebuild_topology(points, radius)
    return ov


def _rows(a):
    return {tuple(int(v) for v in x) for x in a}


def _snapshot(ov, depth):
    return [ov.nodes(l).copy() for l in range(depth + 1)], [ov.child_lists(l).copy() for l in range(1, depth + 1)]


def _check_invariants(ov, depth):
    """Every reached node is flagged and names the parent whose list holds
    it; every other node below the root is unflagged and unlinked, and an
    unreached inner node lists no child."""
    reach = reachable(ov, depth)
    for l in range(depth):
        nodes = ov.nodes(l)
        on = np.zeros(len(nodes), dtype=bool)
        on[reach[l]] = True
        assert (nodes["mFlags"][on] == 1).all(), f"level {l}"
        assert (nodes["mFlags"][~on] == 0).all(), f"level {l}"
        assert (nodes["mParent"][~on] == UNDEF64).all(), f"level {l}"
        assert len(np.unique(nodes["mPos"][on], axis=0)) == on.sum(), f"level {l}: one reached node per cell"
        par = nodes["mParent"][on]
        assert ((par >> np.uint64(8)) & np.uint64(0xFF) == np.uint64(l + 1)).all()
        pidx = (par >> np.uint64(16)).astype(np.int64)
        if l + 1 == depth:
            assert (pidx == 0).all()
        else:
            assert np.isin(pidx, reach[l + 1]).all()
        if l >= 1:
            cl = ov.child_lists(l)
            assert (cl[~on] == UNDEF64).all(), f"level {l}: unreached nodes list children"


P1 = dict(seed=1, n=6000, lo=(20, 30, 25), hi=(150, 110, 140))
P2 = dict(seed=2, n=6000, lo=(90, 60, 70), hi=(230, 200, 180))


def test_accumulate_radius0_is_the_union():
    p1, p2 = _box(**P1), _box(**P2)
    ov = _vol(p1, 0.0)
    depth = ov.info().depth
    assert depth == 2
    cnt1 = [len(ov.nodes(l)) for l in range(depth)]
    old_nodes, _ = _snapshot(ov, depth)
    ov.incremental_topology(p2, 0.0, accum=True)
    union = np.concatenate([p1, p2])
    exp = expected_topology(union, 0.0, (0, 0, 0), LOG2)
    got = reachable_cells(ov, depth, RNG)
    for l in range(depth):
        np.testing.assert_array_equal(got[l], lexsort_cells(exp["nodes"][l]["mPos"] // RNG[l]))
        nodes = ov.nodes(l)
        # old records untouched (positions, flags, parents, lists), new ones appended in key order
        for f in ("mPos", "mFlags", "mParent", "mChildList", "mValue"):
            np.testing.assert_array_equal(nodes[f][:cnt1[l]], old_nodes[l][f], err_msg=f"level {l} {f}")
        new = nodes["mPos"][cnt1[l]:] // RNG[l]
        np.testing.assert_array_equal(new, lexsort_cells(new))
        assert _rows(new) == _rows(got[l]) - _rows(old_nodes[l]["mPos"] // RNG[l])
    _check_invariants(ov, depth)
    # a full rebuild of the union holds the same node set
    full = _vol(union, 0.0)
    for l in range(depth):
        assert _rows(full.nodes(l)["mPos"]) == _rows(ov.nodes(l)["mPos"])


def test_rebuild_incremental_radius0_keeps_only_the_new_frame():
    p1, p2 = _box(**P1), _box(**P2)
    ov = _vol(p1, 0.0)
    depth = ov.info().depth
    ov.incremental_topology(p2, 0.0, accum=False)
    got = reachable_cells(ov, depth, RNG)
    cov2 = _covering(p2, depth)
    for l in range(depth):
        np.testing.assert_array_equal(got[l], cov2[l])
    _check_invariants(ov, depth)
    dead = (ov.nodes(0)["mFlags"] == 0).sum()
    assert dead > 100
    # a third frame back over the first region: the dead cells come back as new nodes
    cnt2 = [len(ov.nodes(l)) for l in range(depth)]
    ov.incremental_topology(p1, 0.0, accum=False)
    got = reachable_cells(ov, depth, RNG)
    cov1 = _covering(p1, depth)
    for l in range(depth):
        np.testing.assert_array_equal(got[l], cov1[l])
    assert len(ov.nodes(0)) > cnt2[0]
    pos = ov.nodes(0)["mPos"]
    assert len(np.unique(pos, axis=0)) < len(pos), "re-reached cells must be duplicated, not revived"
    _check_invariants(ov, depth)


@pytest.mark.parametrize("radius", [1.0, 1.5, 3.0])
@pytest.mark.parametrize("accum", [True, False])
def test_radius_box_reach(radius, accum):
    p1, p2 = _box(**P1), _box(**P2)
    ov = _vol(p1, radius)
    info = ov.info()
    depth, root = info.depth, np.array(info.root_pos)
    old = reachable_cells(ov, depth, RNG)
    ov.incremental_topology(p2, radius, accum=accum)
    touched = incre_touched(p2, radius, (0, 0, 0), RNG, depth, root)
    got = reachable_cells(ov, depth, RNG)
    for l in range(depth):
        want = _rows(touched[l]) | (_rows(old[l]) if accum else set())
        assert _rows(got[l]) == want, f"level {l}"
    _check_invariants(ov, depth)


def test_no_new_node_changes_nothing():
    p1 = _box(**P1)
    ov = _vol(p1, 3.0)
    info = ov.info()
    depth = info.depth
    sub = p1[::3]
    touched = incre_touched(sub, 1.0, (0, 0, 0), RNG, depth, np.array(info.root_pos))
    have = reachable_cells(ov, depth, RNG)
    assert all(_rows(touched[l]) <= _rows(have[l]) for l in range(depth))
    nodes, lists = _snapshot(ov, depth)
    ov.incremental_topology(sub, 1.0, accum=False)
    after_nodes, after_lists = _snapshot(ov, depth)
    for a, b in zip(nodes, after_nodes):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(lists, after_lists):
        np.testing.assert_array_equal(a, b)


def test_points_outside_the_root_are_refused():
    p1 = _box(**P1)
    ov = _vol(p1, 1.0)
    depth = ov.info().depth
    nodes, _ = _snapshot(ov, depth)
    far = np.array([[600.0, 10.0, 10.0]], dtype=np.float32)   # root covers [0, 512)
    with pytest.raises(O.OracleError) as e:
        ov.incremental_topology(np.concatenate([p1[:100], far]), 1.0)
    assert e.value.code == 5
    with pytest.raises(O.OracleError) as e:
        ov.incremental_topology(np.full((4, 3), np.nan, dtype=np.float32), 1.0)
    assert e.value.code == 4
    for a, b in zip(nodes, _snapshot(ov, depth)[0]):
        np.testing.assert_array_equal(a, b)


def test_atlas_after_deactivation():
    p1, p2 = _box(**P1), _box(**P2)
    ov = _vol(p1, 1.0)
    ov.update_atlas()
    depth = ov.info().depth
    ov.incremental_topology(p2, 1.0, accum=False)
    stale = ov.nodes(0)["mValue"].copy()
    ov.update_atlas()
    nodes = ov.nodes(0)
    act = nodes["mFlags"] == 1
    assert (~act).sum() > 50
    np.testing.assert_array_equal(nodes["mValue"][~act], stale[~act])
    k = np.arange(act.sum())
    sx, sy, br = 16, 16, 10
    z, rem = k // (sx * sy), k % (sx * sy)
    exp = np.stack([(rem % sx) * br + 1, (rem // sx) * br + 1, z * br + 1], 1)
    np.testing.assert_array_equal(nodes["mValue"][act], exp)
    # the atlas grows to the active leaves only (the pool's used count)
    assert list(ov.info().subdim) == [16, 16, max(1, -(-int(act.sum()) // 256))]
    # ComputeBounds: leaf 0 seeds both corners, then the active leaves widen them
    pos = nodes["mPos"].astype(np.int64)
    lo = np.minimum(pos[0], pos[act].min(0))
    hi = np.maximum(pos[0], (pos[act] + 8).max(0))
    assert list(ov.info().obj_min) == list(lo.astype(np.float32)) and list(ov.info().obj_max) == list(hi.astype(np.float32))
    amap = ov.atlas_map()
    np.testing.assert_array_equal(amap["mLeafID"][:act.sum()], np.nonzero(act)[0])
    assert (amap["mLeafID"][act.sum():] == -1).all()
    _check_invariants(ov, depth)


@pytest.mark.parametrize("accum", [True, False])
def test_level_set_on_common_leaves_equals_full_rebuild(accum):
    p1, p2 = _box(**P1), _box(**P2)
    ov = _vol(p1, 3.0)
    ov.incremental_topology(p2, 3.0, accum=accum)
    ov.update_atlas()
    ov.insert_points_subcell(4, p2, 3.0)
    ov.gather_level_set(1.5)
    full = _vol(p2, 3.0)
    full.update_atlas()
    full.insert_points_subcell(4, p2, 3.0)
    full.gather_level_set(1.5)
    depth = ov.info().depth
    ri = reachable(ov, depth)[0]
    rf = reachable(full, depth)[0]
    ki = {tuple(int(c) for c in ov.nodes(0)["mPos"][i]): i for i in ri}
    kf = {tuple(int(c) for c in full.nodes(0)["mPos"][i]): i for i in rf}
    common = sorted(set(ki) & set(kf))
    assert len(common) > 200
    bi, bf = ov.bricks(0), full.bricks(0)
    a = bi[[ki[k] for k in common]][:, 1:9, 1:9, 1:9]
    b = bf[[kf[k] for k in common]][:, 1:9, 1:9, 1:9]
    np.testing.assert_array_equal(a, b)
    assert a.min() < 0.0
