"""Independent VBX reader (GVDB_FILESPEC.txt, in the field order SaveVBX
writes, gvdb_volume_gvdb.cpp:1626-1767), used by the tests to check files
field by field."""
import struct

import numpy as np

NODE_BYTES = 64


class Reader:
    def __init__(self, data):
        self.d = data
        self.o = 0

    def take(self, fmt):
        n = struct.calcsize(fmt)
        out = struct.unpack_from(fmt, self.d, self.o)
        self.o += n
        return out

    def raw(self, n):
        b = self.d[self.o:self.o + n]
        assert len(b) == n, "truncated"
        self.o += n
        return b


def read_vbx(path):
    r = Reader(open(path, "rb").read())
    f = {}
    f["major"], f["minor"] = r.take("<BB")
    if (f["major"] == 1 and f["minor"] >= 11) or f["major"] > 1:
        f["xform"] = r.take("<12f")
    (ngrid,) = r.take("<i")
    f["num_grids"] = ngrid
    if f["major"] >= 2:
        (f["bitmasks"],) = r.take("<B")
    f["grid_offs"] = r.take("<%dQ" % ngrid)
    grids = []
    for _ in range(ngrid):
        g = {"offset": r.o}
        g["name"] = r.raw(256)
        g["dtype"], g["comps"], g["compress"] = r.take("<cBB")
        g["voxelsize"] = r.take("<3f")
        (g["leafcnt"],) = r.take("<i")
        g["leafdim"] = r.take("<3i")
        (g["apron"],) = r.take("<i")
        (g["num_chan"],) = r.take("<i")
        (g["atlas_sz"],) = r.take("<Q")
        (g["topotype"],) = r.take("<B")
        (g["reuse"],) = r.take("<i")
        (g["layout"],) = r.take("<B")
        g["axiscnt"] = r.take("<3i")
        g["axisres"] = r.take("<3i")
        (g["levels"],) = r.take("<i")
        (g["root"],) = r.take("<Q")
        lv = []
        for _ in range(g["levels"]):
            ld, res, rx, ry, rz, c0, w0, c1, w1 = r.take("<9i")
            lv.append(dict(ld=ld, res=res, range=(rx, ry, rz), cnt0=c0, w0=w0, cnt1=c1, w1=w1))
        g["level_table"] = lv
        g["pool0"] = [r.raw(L["cnt0"] * L["w0"]) for L in lv]
        g["pool1"] = [r.raw(L["cnt1"] * L["w1"]) for L in lv]
        ax = g["axisres"]
        g["atlas"] = []
        for _ in range(g["num_chan"]):
            typ, stride = r.take("<2i")
            a = np.frombuffer(r.raw(ax[0] * ax[1] * ax[2] * stride), dtype=np.uint32).reshape(ax[2], ax[1], ax[0])
            g["atlas"].append((typ, stride, a))
        grids.append(g)
    f["grids"] = grids
    f["size"] = len(r.d)
    f["consumed"] = r.o
    return f
