"""Pins the oracle's VBX writer (SaveVBX, gvdb_volume_gvdb.cpp:1626-1767)
with an independent reader (tests/vbx_reader.py, written from
GVDB_FILESPEC.txt and SaveVBX's field order): every header field, the level
table, the pools byte for byte against the oracle's node and child-list
arrays, and the atlas brick by brick against the leaves' atlas coordinates,
zeros where no active leaf owns a brick."""
import numpy as np
import pytest

import oracle as O
from vbx_reader import read_vbx


def _vol(incremental):
    rng = np.random.default_rng(9)
    p1 = (rng.random((8000, 3)) * [110, 80, 100] + [20, 30, 25]).astype(np.float32)
    ov = O.OracleVolume()
    ov.add_channel(0, 1)
    ov.add_channel(1, 1, dtype=O.T_UCHAR4)
    ov.rebuild_topology(p1, 3.0)
    if incremental:
        p2 = (rng.random((8000, 3)) * [100, 90, 80] + [90, 60, 70]).astype(np.float32)
        ov.incremental_topology(p2, 3.0, accum=False)
        p1 = p2
    ov.update_atlas()
    ov.insert_points_subcell(4, p1, 3.0)
    ov.gather_level_set(1.5)
    ov.update_apron(0, 3.0)
    ov.fill_channel(1, 0.0)
    return ov


@pytest.mark.parametrize("incremental", [False, True])
def test_oracle_vbx_fields(tmp_path, incremental):
    ov = _vol(incremental)
    path = tmp_path / "v.vbx"
    ov.save_vbx(path)
    f = read_vbx(path)
    assert f["consumed"] == f["size"]
    assert (f["major"], f["minor"]) == (1, 11)
    assert f["xform"] == (0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0)
    assert f["num_grids"] == 1 and f["grid_offs"] == (f["grids"][0]["offset"],) == (62,)
    g = f["grids"][0]
    info = ov.info()
    assert g["name"] == bytes(256)
    assert (g["dtype"], g["comps"], g["compress"], g["topotype"], g["reuse"], g["layout"]) == (b"f", 1, 0, 2, 0, 0)
    assert g["voxelsize"] == (1, 1, 1) and g["leafdim"] == (8, 8, 8) and g["apron"] == 1
    assert g["leafcnt"] == info.node_cnt[0] and g["num_chan"] == 2
    assert g["axiscnt"] == tuple(info.subdim) and g["axisres"] == tuple(10 * np.array(info.subdim))
    assert g["atlas_sz"] == 4 * int(np.prod(g["axisres"]))
    assert g["levels"] == 5 and g["root"] == info.depth << 8
    for l, L in enumerate(g["level_table"]):
        assert L["ld"] == 3 + 2 * (l == 4) and L["res"] == 1 << L["ld"]
        assert L["range"] == (int(np.prod([8, 8, 8, 8, 32][:l + 1])),) * 3
        assert L["cnt0"] == info.node_cnt[l] and L["w0"] == 64
        assert (L["cnt1"], L["w1"]) == ((0, 0) if l == 0 else (info.node_cnt[l], 8 * L["res"] ** 3))
        assert g["pool0"][l] == ov.nodes(l).tobytes()
        if l >= 1:
            assert g["pool1"][l] == ov.child_lists(l).tobytes()
    nodes = ov.nodes(0)
    act = nodes["mFlags"] == 1
    assert incremental == (not act.all())
    for c, (typ, stride, a) in enumerate(g["atlas"]):
        assert typ == (O.T_FLOAT if c == 0 else O.T_UCHAR4) and stride == 4
        ref = ov.bricks(c).view(np.uint32)
        owned = np.zeros(a.shape, dtype=bool)
        for k in np.nonzero(act)[0]:
            x, y, z = nodes["mValue"][k] - 1
            np.testing.assert_array_equal(a[z:z + 10, y:y + 10, x:x + 10], ref[k])
            owned[z:z + 10, y:y + 10, x:x + 10] = True
        assert not a[~owned].any()
