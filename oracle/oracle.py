"""ctypes view of the CPU oracle (oracle/_build/libgvdb_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never imports it.
The oracle restates the reference semantics documented in gvdb_oracle.c.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libgvdb_oracle.so")

OR_MAXLEV = 10


class OrNode(C.Structure):
    _fields_ = [("mLev", C.c_uint8), ("mFlags", C.c_uint8), ("mPriority", C.c_uint8), ("pad", C.c_uint8),
                ("mPos", C.c_int32 * 3), ("mValue", C.c_int32 * 3), ("mVRange", C.c_float * 3),
                ("mParent", C.c_uint64), ("mChildList", C.c_uint64), ("mMask", C.c_uint64)]


class OrInfo(C.Structure):
    _fields_ = [("levs", C.c_int), ("depth", C.c_int), ("top_lev", C.c_int), ("root_pos", C.c_int * 3),
                ("node_cnt", C.c_int64 * OR_MAXLEV), ("subdim", C.c_int * 3), ("brick_res", C.c_int),
                ("apron", C.c_int), ("pos_min", C.c_float * 3), ("pos_max", C.c_float * 3),
                ("obj_min", C.c_float * 3), ("obj_max", C.c_float * 3), ("sc_len", C.c_int64)]


class OrRenderParams(C.Structure):
    _fields_ = [("extinct", C.c_float), ("albedo", C.c_float), ("step", C.c_float), ("minval", C.c_float),
                ("alphacut", C.c_float), ("thresh", C.c_float), ("vmin", C.c_float), ("vmax", C.c_float),
                ("epsilon", C.c_float), ("max_iter", C.c_int), ("backclr", C.c_float * 4)]


NODE_DTYPE = np.dtype([("mLev", "u1"), ("mFlags", "u1"), ("mPriority", "u1"), ("pad", "u1"),
                       ("mPos", "<i4", 3), ("mValue", "<i4", 3), ("mVRange", "<f4", 3),
                       ("mParent", "<u8"), ("mChildList", "<u8"), ("mMask", "<u8")])
assert NODE_DTYPE.itemsize == 64
AMAP_DTYPE = np.dtype([("mPos", "<i4", 3), ("mLeafID", "<i4")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, f32, fp, dp = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.or_create.restype = vp
        L.or_destroy.argtypes = [vp]
        L.or_configure.argtypes = [vp, i32, C.POINTER(C.c_int)]
        L.or_set_channel_default.argtypes = [vp, i32, i32, i32]
        L.or_add_channel.argtypes = [vp, i32, i32]
        L.or_rebuild_topology.argtypes = [vp, vp, i64, i64, f32, fp]
        L.or_update_atlas.argtypes = [vp]
        L.or_insert_points_subcell.argtypes = [vp, i32, vp, i64, i64, f32, fp, C.POINTER(C.c_int64)]
        L.or_gather.argtypes = [vp, i32, f32, i32, i32]
        L.or_gather_f64.argtypes = [vp, i32, f32, i32, dp]
        L.or_update_apron.argtypes = [vp, i32, f32]
        L.or_insert_points.argtypes = [vp, vp, i64, i64, fp, vp, vp]
        L.or_scatter_density.argtypes = [vp, vp, i64, i64, f32, f32, fp, i32, i32]
        L.or_splat_cubic.argtypes = [vp, vp, i64, i64, vp, f32, fp, i32, i32]
        L.or_splat_cubic_f64.argtypes = [vp, vp, i64, i64, vp, f32, fp, dp]
        L.or_fill_channel.argtypes = [vp, i32, f32]
        L.or_get_info.argtypes = [vp, C.POINTER(OrInfo)]
        L.or_nodes.argtypes = [vp, i32]
        L.or_nodes.restype = vp
        L.or_child_lists.argtypes = [vp, i32]
        L.or_child_lists.restype = vp
        L.or_atlas_map.argtypes = [vp, C.POINTER(C.c_int64)]
        L.or_atlas_map.restype = vp
        L.or_atlas.argtypes = [vp, i32]
        L.or_atlas.restype = vp
        L.or_read_brick.argtypes = [vp, i32, i64, fp]
        L.or_raycast_ea.argtypes = [vp, i32, C.POINTER(OrRenderParams), i64, fp, fp, fp, dp]
        _lib = L
    return _lib


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"oracle {what} failed with status {code}")
        self.code = code


def _chk(code, what):
    if code != 0:
        raise OracleError(code, what)


def _pts(points):
    p = np.ascontiguousarray(points, dtype=np.float32)
    assert p.ndim == 2 and p.shape[1] == 3
    return p


class OracleVolume:
    """Mirror of the reference VolumeGVDB calls on the points path (CPU, serial)."""

    def __init__(self, log2dims=(3, 3, 3, 3, 5), channel_default=(16, 16, 1)):
        self._L = lib()
        self._h = self._L.or_create()
        arr = (C.c_int * len(log2dims))(*log2dims)
        _chk(self._L.or_configure(self._h, len(log2dims), arr), "configure")
        self._L.or_set_channel_default(self._h, *channel_default)
        self.log2dims = tuple(log2dims)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.or_destroy(h)
            self._h = None

    @property
    def res0(self):
        return 1 << self.log2dims[0]

    def add_channel(self, chan=0, apron=1):
        _chk(self._L.or_add_channel(self._h, chan, apron), "add_channel")

    def rebuild_topology(self, points, radius, orig=(0, 0, 0)):
        p = _pts(points)
        code = self._L.or_rebuild_topology(self._h, p.ctypes.data, p.shape[0], 12, float(radius), _f3(orig))
        _chk(code, "rebuild_topology")

    def update_atlas(self):
        _chk(self._L.or_update_atlas(self._h), "update_atlas")

    def insert_points_subcell(self, sc, points, radius, trans=(0, 0, 0)):
        p = _pts(points)
        out = C.c_int64(0)
        _chk(self._L.or_insert_points_subcell(self._h, sc, p.ctypes.data, p.shape[0], 12, float(radius),
                                              _f3(trans), C.byref(out)), "insert_points_subcell")
        return out.value

    def gather_density(self, radius, chan=0, accum=False):
        _chk(self._L.or_gather(self._h, 0, float(radius), chan, int(accum)), "gather_density")

    def gather_level_set(self, radius, chan=0, accum=False):
        _chk(self._L.or_gather(self._h, 1, float(radius), chan, int(accum)), "gather_level_set")

    def gather_f64(self, mode, radius):
        n0 = self.info().node_cnt[0]
        out = np.zeros(n0 * self.res0 ** 3, dtype=np.float64)
        _chk(self._L.or_gather_f64(self._h, mode, float(radius), 0,
                                   out.ctypes.data_as(C.POINTER(C.c_double))), "gather_f64")
        return out.reshape(n0, self.res0, self.res0, self.res0)

    def update_apron(self, chan=0, boundval=0.0):
        _chk(self._L.or_update_apron(self._h, chan, float(boundval)), "update_apron")

    def insert_points(self, points, trans=(0, 0, 0)):
        p = _pts(points)
        pnode = np.zeros(p.shape[0], dtype=np.int32)
        poff = np.zeros(p.shape[0], dtype=np.int32)
        _chk(self._L.or_insert_points(self._h, p.ctypes.data, p.shape[0], 12, _f3(trans), pnode.ctypes.data,
                                      poff.ctypes.data), "insert_points")
        return pnode, poff

    def scatter_density(self, points, radius, amp=1.0, trans=(0, 0, 0), expand=True, chan=0):
        p = _pts(points)
        _chk(self._L.or_scatter_density(self._h, p.ctypes.data, p.shape[0], 12, float(radius), float(amp),
                                        _f3(trans), int(expand), chan), "scatter_density")

    def splat_cubic(self, points, radius, weights=None, trans=(0, 0, 0), chan=0, accum=False):
        p = _pts(points)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        _chk(self._L.or_splat_cubic(self._h, p.ctypes.data, p.shape[0], 12,
                                    None if w is None else w.ctypes.data, float(radius), _f3(trans), chan,
                                    int(accum)), "splat_cubic")

    def splat_cubic_f64(self, points, radius, weights=None, trans=(0, 0, 0)):
        p = _pts(points)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        n0 = self.info().node_cnt[0]
        out = np.zeros(n0 * self.res0 ** 3, dtype=np.float64)
        _chk(self._L.or_splat_cubic_f64(self._h, p.ctypes.data, p.shape[0], 12,
                                        None if w is None else w.ctypes.data, float(radius), _f3(trans),
                                        out.ctypes.data_as(C.POINTER(C.c_double))), "splat_cubic_f64")
        return out.reshape(n0, self.res0, self.res0, self.res0)

    def fill_channel(self, chan, val):
        _chk(self._L.or_fill_channel(self._h, chan, float(val)), "fill_channel")

    # ---------------- readback ----------------
    def info(self):
        inf = OrInfo()
        _chk(self._L.or_get_info(self._h, C.byref(inf)), "get_info")
        return inf

    def nodes(self, lev):
        n = self.info().node_cnt[lev]
        if n == 0:
            return np.zeros(0, dtype=NODE_DTYPE)
        ptr = self._L.or_nodes(self._h, lev)
        buf = (C.c_char * (n * 64)).from_address(ptr)
        return np.frombuffer(bytes(buf), dtype=NODE_DTYPE).copy()

    def child_lists(self, lev):
        n = self.info().node_cnt[lev]
        r = 1 << self.log2dims[lev]
        if n == 0 or lev == 0:
            return np.zeros((0, r ** 3), dtype=np.uint64)
        ptr = self._L.or_child_lists(self._h, lev)
        buf = (C.c_char * (n * r ** 3 * 8)).from_address(ptr)
        return np.frombuffer(bytes(buf), dtype=np.uint64).reshape(n, r ** 3).copy()

    def atlas_map(self):
        cnt = C.c_int64(0)
        ptr = self._L.or_atlas_map(self._h, C.byref(cnt))
        if not ptr or cnt.value == 0:
            return np.zeros(0, dtype=AMAP_DTYPE)
        buf = (C.c_char * (cnt.value * 16)).from_address(ptr)
        return np.frombuffer(bytes(buf), dtype=AMAP_DTYPE).copy()

    def bricks(self, chan=0):
        """All leaf bricks incl. apron as (B, br, br, br) array indexed [z, y, x]."""
        inf = self.info()
        br = inf.brick_res
        B = inf.node_cnt[0]
        out = np.zeros((B, br, br, br), dtype=np.float32)
        tmp = np.zeros(br ** 3, dtype=np.float32)
        for k in range(B):
            _chk(self._L.or_read_brick(self._h, chan, k, tmp.ctypes.data_as(C.POINTER(C.c_float))), "read_brick")
            out[k] = tmp.reshape(br, br, br)
        return out
