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


class OrCamera(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("org", C.c_float * 3), ("org_u", C.c_float * 3),
                ("org_v", C.c_float * 3), ("dir", C.c_float * 3), ("dir_u", C.c_float * 3),
                ("dir_v", C.c_float * 3)]


class OrRenderParams(C.Structure):
    _fields_ = [("shade", C.c_int), ("step", C.c_float), ("extinct", C.c_float), ("albedo", C.c_float),
                ("minval", C.c_float), ("alphacut", C.c_float), ("thresh", C.c_float), ("vmin", C.c_float),
                ("vmax", C.c_float), ("epsilon", C.c_float), ("max_iter", C.c_int), ("backclr", C.c_float * 4),
                ("light", C.c_float * 3), ("shadow_amt", C.c_float), ("shadow_bias", C.c_float),
                ("lut", C.POINTER(C.c_float)), ("finestep", C.c_float), ("slice_pnt", C.c_float * 3),
                ("slice_norm", C.c_float * 3), ("use_clr", C.c_int), ("clr_chan", C.c_int)]


class OrScnRay(C.Structure):
    _fields_ = [("hit", C.c_float * 3), ("normal", C.c_float * 3), ("orig", C.c_float * 3), ("dir", C.c_float * 3),
                ("clr", C.c_uint32), ("pnode", C.c_uint32), ("pndx", C.c_uint32), ("pad", C.c_uint32)]


assert C.sizeof(OrScnRay) == 64
SHADE_VOXEL, SHADE_SECTION2D, SHADE_SECTION3D, SHADE_EMPTYSKIP = 0, 1, 2, 3   # gvdb_types.h:105-108
SHADE_TRILINEAR, SHADE_TRICUBIC, SHADE_LEVELSET, SHADE_VOLUME, SHADE_OFF = 4, 5, 6, 7, 100
T_UCHAR4, T_FLOAT = 2, 3   # gvdb_types.h:75-76


def make_camera(cam):
    """cam: dict with width, height, org, org_u, org_v, dir, dir_u, dir_v."""
    c = OrCamera()
    c.width, c.height = int(cam["width"]), int(cam["height"])
    for k in ("org", "org_u", "org_v", "dir", "dir_u", "dir_v"):
        setattr(c, k, (C.c_float * 3)(*[float(x) for x in cam.get(k, (0, 0, 0))]))
    return c


def make_params(rp):
    """rp: dict of scene parameters; missing keys take the Scene defaults
    (gvdb_scene.cpp:46-53, VG.cpp:71, RC.cuh:24)."""
    p = OrRenderParams()
    p.shade = int(rp.get("shade", SHADE_VOLUME))
    p.step = float(rp.get("step", 1.0))
    p.extinct = float(rp.get("extinct", -1.1))
    p.albedo = float(rp.get("albedo", 1.5))
    p.minval = float(rp.get("minval", 0.005))
    p.alphacut = float(rp.get("alphacut", 0.01))
    p.thresh = float(rp.get("thresh", 0.1))
    p.vmin = float(rp.get("vmin", 0.0))
    p.vmax = float(rp.get("vmax", 1.0))
    p.epsilon = float(rp.get("epsilon", 0.001))
    p.max_iter = int(rp.get("max_iter", 256))
    p.backclr = (C.c_float * 4)(*[float(x) for x in rp.get("backclr", (0, 0, 0, 0))])
    p.light = (C.c_float * 3)(*[float(x) for x in rp.get("light", (0, 0, 0))])
    p.shadow_amt = float(rp.get("shadow_amt", 0.8))
    p.shadow_bias = float(rp.get("shadow_bias", 1.0))
    p.finestep = float(rp.get("finestep", 0.1))
    p.slice_pnt = (C.c_float * 3)(*[float(x) for x in rp.get("slice_pnt", (0, 0, 0))])
    p.slice_norm = (C.c_float * 3)(*[float(x) for x in rp.get("slice_norm", (0, 0, 1))])
    clr = rp.get("clr_chan")   # SetColorChannel: a uchar4 channel colours or_render
    p.use_clr, p.clr_chan = (0, 0) if clr is None or int(clr) < 0 else (1, int(clr))
    lut = rp.get("lut")
    p._lut_keep = None
    if lut is not None:
        a = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1)
        assert a.size == 16384 * 4
        p._lut_keep = a
        p.lut = a.ctypes.data_as(C.POINTER(C.c_float))
    return p


NODE_DTYPE = np.dtype([("mLev", "u1"), ("mFlags", "u1"), ("mPriority", "u1"), ("pad", "u1"),
                       ("mPos", "<i4", 3), ("mValue", "<i4", 3), ("mVRange", "<f4", 3),
                       ("mParent", "<u8"), ("mChildList", "<u8"), ("mMask", "<u8")])
assert NODE_DTYPE.itemsize == 64
AMAP_DTYPE = np.dtype([("mPos", "<i4", 3), ("mLeafID", "<i4")])
RAY_DTYPE = np.dtype([("hit", "<f4", 3), ("normal", "<f4", 3), ("orig", "<f4", 3), ("dir", "<f4", 3),
                      ("clr", "<u4"), ("pnode", "<u4"), ("pndx", "<u4"), ("pad", "<u4")])
assert RAY_DTYPE.itemsize == 64

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
        L.or_add_channel_typed.argtypes = [vp, i32, i32, i32]
        L.or_gather_color.argtypes = [vp, vp, i64, i32, i32]
        L.or_scatter_color.argtypes = [vp, vp, i64, i64, vp, i64, fp, i32, i32]
        L.or_rebuild_topology.argtypes = [vp, vp, i64, i64, f32, fp]
        L.or_update_atlas.argtypes = [vp]
        L.or_incremental_topology.argtypes = [vp, vp, i64, i64, f32, fp, i32]
        L.or_insert_points_subcell.argtypes = [vp, i32, vp, i64, i64, f32, fp, C.POINTER(C.c_int64)]
        L.or_insert_points_subcell_fp16.argtypes = [vp, i32, vp, i64, i64, f32, fp, C.POINTER(C.c_int64)]
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
        L.or_write_brick.argtypes = [vp, i32, i64, fp]
        L.or_save_vbx.argtypes = [vp, C.c_char_p]
        L.or_raytrace.argtypes = [vp, i32, C.POINTER(OrRenderParams), C.POINTER(OrScnRay), i64, f32]
        cp, rpp = C.POINTER(OrCamera), C.POINTER(OrRenderParams)
        L.or_render.argtypes = [vp, i32, cp, rpp, dp, fp, dp]
        L.or_raycast_ea.argtypes = [vp, i32, cp, rpp, dp, fp, fp, dp, dp]
        L.or_raycast_ea_backward.argtypes = [vp, i32, cp, rpp, dp, dp, dp, dp]
        L.or_raycast_ea_backward_mag.argtypes = [vp, i32, cp, rpp, dp, dp, dp, dp]
        L.or_splat_cubic_backward_mag_f64.argtypes = [vp, dp, i64, dp, f32, dp, dp, dp]
        L.or_raycast_ea_samples.argtypes = [vp, i32, cp, rpp, i64, i64, C.POINTER(C.c_int64), fp,
                                            C.POINTER(C.c_int64)]
        L.or_apron_f64.argtypes = [vp, dp, C.c_double]
        L.or_apron_adjoint_f64.argtypes = [vp, dp]
        L.or_splat_cubic_f64_padded.argtypes = [vp, dp, i64, dp, f32, dp]
        L.or_splat_cubic_backward_f64.argtypes = [vp, dp, i64, dp, f32, dp, dp, dp]
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

    def add_channel(self, chan=0, apron=1, dtype=T_FLOAT):
        _chk(self._L.or_add_channel_typed(self._h, chan, apron, int(dtype)), "add_channel")

    def rebuild_topology(self, points, radius, orig=(0, 0, 0)):
        p = _pts(points)
        code = self._L.or_rebuild_topology(self._h, p.ctypes.data, p.shape[0], 12, float(radius), _f3(orig))
        _chk(code, "rebuild_topology")

    def incremental_topology(self, points, radius, orig=(0, 0, 0), accum=True):
        """ActivateIncreBricksGPU on the current tree (accum: AccumulateTopology)."""
        p = _pts(points)
        code = self._L.or_incremental_topology(self._h, p.ctypes.data, p.shape[0], 12, float(radius), _f3(orig),
                                               1 if accum else 0)
        _chk(code, "incremental_topology")

    def update_atlas(self):
        _chk(self._L.or_update_atlas(self._h), "update_atlas")

    def insert_points_subcell(self, sc, points, radius, trans=(0, 0, 0)):
        p = _pts(points)
        out = C.c_int64(0)
        _chk(self._L.or_insert_points_subcell(self._h, sc, p.ctypes.data, p.shape[0], 12, float(radius),
                                              _f3(trans), C.byref(out)), "insert_points_subcell")
        return out.value

    def insert_points_subcell_fp16(self, sc, points, radius, trans=(0, 0, 0)):
        p = _pts(points)
        out = C.c_int64(0)
        _chk(self._L.or_insert_points_subcell_fp16(self._h, sc, p.ctypes.data, p.shape[0], 12, float(radius),
                                                   _f3(trans), C.byref(out)), "insert_points_subcell_fp16")
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

    def gather_color(self, clr, chan=1, accum=False):
        """Colour mean of the gathers over the current subcell lists; clr:
        (N, 4) uint8 per point, in the order of the points given to
        insert_points_subcell."""
        c = np.ascontiguousarray(clr, dtype=np.uint8)
        assert c.ndim == 2 and c.shape[1] == 4
        _chk(self._L.or_gather_color(self._h, c.ctypes.data, 4, chan, int(accum)), "gather_color")

    def scatter_color(self, points, clr, avg=True, chan=1, trans=(0, 0, 0)):
        p = _pts(points)
        c = np.ascontiguousarray(clr, dtype=np.uint8)
        assert c.shape == (p.shape[0], 4)
        _chk(self._L.or_scatter_color(self._h, p.ctypes.data, p.shape[0], 12, c.ctypes.data, 4, _f3(trans),
                                      int(avg), chan), "scatter_color")

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

    def bricks_u8(self, chan=1):
        """A uchar4 channel's bricks as (B, br, br, br, 4) uint8 indexed [z, y, x, component]."""
        b = self.bricks(chan)
        return b.view(np.uint8).reshape(b.shape + (4,))

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

    def save_vbx(self, path):
        """SaveVBX (VG.cpp:1626-1767)."""
        _chk(self._L.or_save_vbx(self._h, str(path).encode()), "save_vbx")

    def set_bricks(self, bricks, chan=0):
        """Test input: overwrite every leaf's padded brick ((B, br, br, br), [z, y, x])."""
        b = np.ascontiguousarray(bricks, dtype=np.float32)
        for k in range(b.shape[0]):
            _chk(self._L.or_write_brick(self._h, chan, k, b[k].ctypes.data_as(C.POINTER(C.c_float))), "write_brick")

    def raytrace(self, rays, rp, bias=0.0, chan=0):
        """Raytrace (VG.cpp:4384-4408): rays is a structured array of RAY_DTYPE,
        updated in place (hit, normal)."""
        assert rays.dtype == RAY_DTYPE and rays.flags.c_contiguous
        p = make_params(rp)
        _chk(self._L.or_raytrace(self._h, chan, C.byref(p), rays.ctypes.data_as(C.POINTER(OrScnRay)), len(rays),
                                 float(bias)), "raytrace")
        return rays

    # ---------------- raycast and adjoints ----------------
    def _padded_count(self):
        inf = self.info()
        return inf.node_cnt[0], inf.brick_res

    @staticmethod
    def _dp(a):
        return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))

    def render(self, cam, rp, chan=0, bricks64=None):
        """Reference Render(shade): returns (rgba32, rgba64), each (H, W, 4)."""
        c, p = make_camera(cam), make_params(rp)
        n = c.width * c.height
        o32 = np.zeros(n * 4, dtype=np.float32)
        o64 = np.zeros(n * 4, dtype=np.float64)
        b = None if bricks64 is None else np.ascontiguousarray(bricks64, dtype=np.float64)
        _chk(self._L.or_render(self._h, chan, C.byref(c), C.byref(p), self._dp(b),
                               o32.ctypes.data_as(C.POINTER(C.c_float)), self._dp(o64)), "render")
        return o32.reshape(c.height, c.width, 4), o64.reshape(c.height, c.width, 4)

    def raycast_ea(self, cam, rp, chan=0, bricks64=None):
        """Linear-transfer emission-absorption raycast: (I32, T32, I64, T64), each (H, W)."""
        c, p = make_camera(cam), make_params(rp)
        n = c.width * c.height
        I32, T32 = np.zeros(n, np.float32), np.zeros(n, np.float32)
        I64, T64 = np.zeros(n, np.float64), np.zeros(n, np.float64)
        b = None if bricks64 is None else np.ascontiguousarray(bricks64, dtype=np.float64)
        fpp = C.POINTER(C.c_float)
        _chk(self._L.or_raycast_ea(self._h, chan, C.byref(c), C.byref(p), self._dp(b), I32.ctypes.data_as(fpp),
                                   T32.ctypes.data_as(fpp), self._dp(I64), self._dp(T64)), "raycast_ea")
        sh = (c.height, c.width)
        return I32.reshape(sh), T32.reshape(sh), I64.reshape(sh), T64.reshape(sh)

    def raycast_ea_backward(self, cam, rp, gI, gT=None, chan=0, bricks64=None, mag=False):
        """FP64 gradient over padded bricks (B, br, br, br) [z, y, x].  mag:
        the magnitude form (every term's absolute value), whose entries scaled
        by n * 2^-24 bound an FP32 evaluation's rounding error."""
        c, p = make_camera(cam), make_params(rp)
        B, br = self._padded_count()
        grad = np.zeros(B * br ** 3, dtype=np.float64)
        gi = np.ascontiguousarray(gI, dtype=np.float64).reshape(-1)
        gt = None if gT is None else np.ascontiguousarray(gT, dtype=np.float64).reshape(-1)
        b = None if bricks64 is None else np.ascontiguousarray(bricks64, dtype=np.float64)
        fn = self._L.or_raycast_ea_backward_mag if mag else self._L.or_raycast_ea_backward
        _chk(fn(self._h, chan, C.byref(c), C.byref(p), self._dp(b), self._dp(gi), self._dp(gt), self._dp(grad)),
             "raycast_ea_backward")
        return grad.reshape(B, br, br, br)

    def raycast_ea_samples(self, cam, rp, px, chan=0, max_out=4096):
        c, p = make_camera(cam), make_params(rp)
        leaf = np.zeros(max_out, dtype=np.int64)
        pp = np.zeros(max_out * 3, dtype=np.float32)
        cnt = C.c_int64(0)
        _chk(self._L.or_raycast_ea_samples(self._h, chan, C.byref(c), C.byref(p), int(px), max_out,
                                           leaf.ctypes.data_as(C.POINTER(C.c_int64)),
                                           pp.ctypes.data_as(C.POINTER(C.c_float)), C.byref(cnt)),
             "raycast_ea_samples")
        k = min(cnt.value, max_out)
        return leaf[:k], pp[:3 * k].reshape(k, 3)

    def apron_f64(self, bricks64, boundval=0.0):
        b = np.ascontiguousarray(bricks64, dtype=np.float64).copy()
        _chk(self._L.or_apron_f64(self._h, self._dp(b), float(boundval)), "apron_f64")
        return b

    def apron_adjoint_f64(self, grad):
        g = np.ascontiguousarray(grad, dtype=np.float64).copy()
        _chk(self._L.or_apron_adjoint_f64(self._h, self._dp(g)), "apron_adjoint_f64")
        return g

    def splat_cubic_f64_padded(self, pos64, radius, weights=None):
        B, br = self._padded_count()
        out = np.zeros((B, br, br, br), dtype=np.float64)
        p = np.ascontiguousarray(pos64, dtype=np.float64)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        _chk(self._L.or_splat_cubic_f64_padded(self._h, self._dp(p), p.shape[0], self._dp(w), float(radius),
                                               self._dp(out)), "splat_cubic_f64_padded")
        return out

    def splat_cubic_backward_f64(self, pos64, radius, grad, weights=None, mag=False):
        p = np.ascontiguousarray(pos64, dtype=np.float64)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        g = np.ascontiguousarray(grad, dtype=np.float64)
        dpos = np.zeros_like(p)
        dw = np.zeros(p.shape[0], dtype=np.float64)
        fn = self._L.or_splat_cubic_backward_mag_f64 if mag else self._L.or_splat_cubic_backward_f64
        _chk(fn(self._h, self._dp(p), p.shape[0], self._dp(w), float(radius), self._dp(g), self._dp(dpos),
                self._dp(dw)), "splat_cubic_backward_f64")
        return dpos, dw


def backproject_depth(depth, intr, poses, voxel, origin, dmin=0.5, dmax=8.0):
    """Depth frames -> index-space points, FP32 in the operand order of
    gvdbx_backproject_depth (numpy float32 arithmetic has no contraction):
    c = (d*((u-cx)/fx), d*((v-cy)/fy), d); w_i = ((R_i0 c0 + R_i1 c1) + R_i2 c2) + t_i;
    p = (w - origin) / voxel; pixels with d outside [dmin, dmax] -> (0, 0, 1e10)."""
    f32 = np.float32
    depth = np.asarray(depth, f32)
    poses = np.asarray(poses, f32)
    F, H, W = depth.shape
    fx, fy, cx, cy = (f32(x) for x in intr)
    ox, oy, oz = (f32(x) for x in origin)
    vox = f32(voxel)
    xn = (np.arange(W, dtype=f32)[None, :] - cx) / fx
    yn = (np.arange(H, dtype=f32)[:, None] - cy) / fy
    out = np.zeros((F, H, W, 3), f32)
    out[..., 2] = f32(1.0e10)
    for f in range(F):
        d = depth[f]
        with np.errstate(invalid="ignore"):
            ok = (d >= f32(dmin)) & (d <= f32(dmax))
        c0, c1, c2 = d * xn, d * yn, d
        P = poses[f]
        w0 = ((P[0] * c0 + P[1] * c1) + P[2] * c2) + P[9]
        w1 = ((P[3] * c0 + P[4] * c1) + P[5] * c2) + P[10]
        w2 = ((P[6] * c0 + P[7] * c1) + P[8] * c2) + P[11]
        p = np.stack([(w0 - ox) / vox, (w1 - oy) / vox, (w2 - oz) / vox], -1)
        out[f][ok] = p[ok]
    return out.reshape(-1, 3)
