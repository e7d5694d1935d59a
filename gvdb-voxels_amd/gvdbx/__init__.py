"""gvdbx: MI355X-native GVDB points -> sparse volume path.

`VolumeGVDB` mirrors the reference class nvdb::VolumeGVDB
(source/gvdb_library/src/gvdb_volume_gvdb.h:319-901) on the points path:
Configure, AddChannel, SetPoints, RebuildTopology, FinishTopology,
UpdateAtlas, InsertPointsSubcell, GatherDensity, GatherLevelSet,
InsertPoints, ScatterDensity, UpdateApron, Render, plus the extensions
(weighted cubic splat, emission-absorption raycast and their adjoints).
Every call goes through the C-ABI of libgvdbx.so; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from ._lib import (Camera, GvdbxError, Info, RenderParams, TOPO_AUTO, TOPO_DENSE, TOPO_RADIX,  # noqa: F401
                   call, exported_names, lib, last_error)

NODE_DTYPE = np.dtype([("mLev", "u1"), ("mFlags", "u1"), ("mPriority", "u1"), ("pad", "u1"),
                       ("mPos", "<i4", 3), ("mValue", "<i4", 3), ("mVRange", "<f4", 3),
                       ("mParent", "<u8"), ("mChildList", "<u8"), ("mMask", "<u8")])
AMAP_DTYPE = np.dtype([("mPos", "<i4", 3), ("mLeafID", "<i4")])
T_UCHAR4, T_FLOAT = 2, 3   # gvdb_types.h:75-76
F_LINEAR, F_CLAMP = 1, 1
SHADE_VOXEL, SHADE_SECTION2D, SHADE_SECTION3D, SHADE_EMPTYSKIP = 0, 1, 2, 3   # gvdb_types.h:105-108
SHADE_TRILINEAR, SHADE_TRICUBIC, SHADE_LEVELSET, SHADE_VOLUME, SHADE_OFF = 4, 5, 6, 7, 100   # gvdb_types.h:109-115


def _v3(x):
    return (C.c_float * 3)(*[float(a) for a in x])


def camera(width, height, org=(0, 0, 0), org_u=(0, 0, 0), org_v=(0, 0, 0), dir=(0, 0, 1), dir_u=(0, 0, 0),
           dir_v=(0, 0, 0)):
    """gvdbx_camera from its six vectors (index space)."""
    c = Camera()
    c.width, c.height = int(width), int(height)
    c.org, c.org_u, c.org_v = _v3(org), _v3(org_u), _v3(org_v)
    c.dir, c.dir_u, c.dir_v = _v3(dir), _v3(dir_u), _v3(dir_v)
    return c


def camera_ortho(width, height, corner, extent_u, extent_v, direction):
    """Orthographic camera: pixel centres span corner + [0,1)*extent_u + [0,1)*extent_v."""
    return camera(width, height, org=corner, org_u=extent_u, org_v=extent_v, dir=direction)


def camera_perspective(width, height, campos, cams, camu, camv):
    """PrepareRender's camera (VG.cpp:4260-4265): rays from campos through
    cams + u*camu + w*camv (top-left ray and the two image edges)."""
    return camera(width, height, org=campos, dir=cams, dir_u=camu, dir_v=camv)


def render_params(**kw):
    """gvdbx_render_params with the Scene defaults (gvdb_scene.cpp:46-53,
    VG.cpp:71, RC.cuh:24) for every field not given.  lut: a device float32
    tensor of 16384*4 entries (SHADE_VOLUME)."""
    p = RenderParams()
    p.step = float(kw.get("step", 1.0))
    p.extinct = float(kw.get("extinct", -1.1))
    p.albedo = float(kw.get("albedo", 1.5))
    p.minval = float(kw.get("minval", 0.005))
    p.alphacut = float(kw.get("alphacut", 0.01))
    p.thresh = float(kw.get("thresh", 0.1))
    p.vmin = float(kw.get("vmin", 0.0))
    p.vmax = float(kw.get("vmax", 1.0))
    p.epsilon = float(kw.get("epsilon", 0.001))
    p.max_iter = int(kw.get("max_iter", 256))
    p.backclr = (C.c_float * 4)(*[float(a) for a in kw.get("backclr", (0, 0, 0, 0))])
    p.light = _v3(kw.get("light", (0, 0, 0)))
    p.shadow_amt = float(kw.get("shadow_amt", 0.8))
    p.shadow_bias = float(kw.get("shadow_bias", 1.0))
    p.finestep = float(kw.get("finestep", 0.1))
    p.slice_pnt = _v3(kw.get("slice_pnt", (0, 0, 0)))
    p.slice_norm = _v3(kw.get("slice_norm", (0, 0, 1)))
    lut = kw.get("lut")
    p._lut = lut
    p.lut_dev = _ptr(lut) if lut is not None else None
    return p


def ea_params(step=0.5, kappa=0.05, albedo=1.0, **kw):
    """Parameters of the differentiable emission-absorption raycast: extinction
    kappa per unit density and length, no minval test, no alpha cut."""
    return render_params(step=step, extinct=-kappa, albedo=albedo, minval=float("-inf"), alphacut=0.0, **kw)


def _ptr(x):
    """Device address of a torch tensor (or an int address)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


class FrameGraph:
    """A captured level-set frame of a VolumeGVDB (gvdbx_frame_graph_*)."""

    def __init__(self, h, vol):
        self._h = h
        self._vol = vol   # the graph replays into this volume

    def launch(self):
        call("gvdbx_frame_graph_launch", self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gvdbx_frame_graph_destroy(self._h)
            self._h = None


class VolumeGVDB:
    """One sparse volume on one GPU (reference: VolumeGVDB)."""

    def __init__(self, device=0, stream=None):
        self._h = C.c_void_p()
        call("gvdbx_create", C.byref(self._h), int(device))
        self._keep = []
        if stream is not None:
            self.SetStream(stream)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gvdbx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ setup
    def SetStream(self, stream):
        """stream: a torch.cuda.Stream, or a raw hipStream_t address."""
        handle = getattr(stream, "cuda_stream", stream)
        call("gvdbx_set_stream", self._h, C.c_void_p(handle))

    def Synchronize(self):
        call("gvdbx_synchronize", self._h)

    def Configure(self, q4, q3, q2, q1, q0):
        """Configure(q4,q3,q2,q1,q0) (VG.h:387): level 0 (the brick) gets q0."""
        self.configure([q0, q1, q2, q3, q4])

    def configure(self, log2dims):
        arr = (C.c_int * len(log2dims))(*[int(x) for x in log2dims])
        call("gvdbx_configure", self._h, len(log2dims), arr)
        self.log2dims = tuple(int(x) for x in log2dims)

    def SetChannelDefault(self, sx, sy, sz):
        call("gvdbx_set_channel_default", self._h, sx, sy, sz)

    def AddChannel(self, chan, dtype=T_FLOAT, apron=1, filter=F_LINEAR, border=F_CLAMP):
        call("gvdbx_add_channel", self._h, chan, dtype, apron, filter, border)

    def FillChannel(self, chan, value):
        call("gvdbx_fill_channel", self._h, chan, float(value))

    def SetTopologyMode(self, mode):
        call("gvdbx_set_topology_mode", self._h, int(mode))

    def SetTuning(self, key, value):
        call("gvdbx_set_tuning", self._h, key.encode(), int(value))

    def GetTuning(self, key):
        out = C.c_int(0)
        call("gvdbx_get_tuning", self._h, key.encode(), C.byref(out))
        return out.value

    # ------------------------------------------------------------ points
    def SetPoints(self, pos, vel=None, clr=None, pos_off=0, pos_stride=12, clr_off=0, clr_stride=4):
        """pos: device float32 tensor (N,3) (or an address with offset/stride);
        clr: optional device uint8 tensor (N,4) of uchar4 colours."""
        self._keep = [pos, vel, clr]
        call("gvdbx_set_points", self._h, C.c_void_p(_ptr(pos)), pos_off, pos_stride,
             C.c_void_p(_ptr(vel)) if vel is not None else None, 0, 12,
             C.c_void_p(_ptr(clr)) if clr is not None else None, clr_off, clr_stride)

    def SetPointsBatched(self, pos, cloud_off):
        """Many clouds, each in its own index space: pos (N,3) device tensor,
        cloud_off: ncloud+1 point offsets (host sequence)."""
        self._keep = [pos]
        off = (C.c_int64 * len(cloud_off))(*[int(x) for x in cloud_off])
        call("gvdbx_set_points_batched", self._h, C.c_void_p(_ptr(pos)), 12, off, len(cloud_off) - 1)
        self.ncloud = len(cloud_off) - 1

    def cloud_layout(self):
        n = self.ncloud
        off = np.zeros((n, 3), np.int32)
        bnd = np.zeros((n, 6), np.float32)
        call("gvdbx_read_cloud_layout", self._h, off.ctypes.data, bnd.ctypes.data, n)
        return off, bnd

    def leaf_clouds(self):
        n = self.info().node_cnt[0]
        out = np.zeros(n, np.int32)
        if n:
            call("gvdbx_read_leaf_clouds", self._h, out.ctypes.data, n)
        return out

    def SetPointWeights(self, w, stride=4):
        self._keep_w = w
        call("gvdbx_set_point_weights", self._h, C.c_void_p(_ptr(w)) if w is not None else None, stride)

    # ------------------------------------------------------------ topology
    def RebuildTopology(self, num_pnts, radius, orig=(0.0, 0.0, 0.0)):
        call("gvdbx_rebuild_topology", self._h, int(num_pnts), float(radius), *[float(x) for x in orig])

    def AccumulateTopology(self, num_pnts, radius, orig=(0.0, 0.0, 0.0), iDepth=1):
        call("gvdbx_accumulate_topology", self._h, int(num_pnts), float(radius), *[float(x) for x in orig], int(iDepth))

    def RequestFullRebuild(self, tf):
        call("gvdbx_request_full_rebuild", self._h, 1 if tf else 0)

    def FrameGraph(self, num_pnts, subcell_size, sc_radius, radius, chan=0, accum=False, apron_chan=0,
                   boundval=3.0):
        """Captured level-set frame (gvdbx_frame_graph_create); returns a
        FrameGraph whose launch() replays it on the volume's stream."""
        h = C.c_void_p()
        call("gvdbx_frame_graph_create", self._h, int(num_pnts), int(subcell_size), float(sc_radius), float(radius),
             int(chan), int(accum), int(apron_chan), float(boundval), C.byref(h))
        return FrameGraph(h, self)

    def AutotuneGather(self, mode, subcell_size, radius, chan=0, reps=3):
        """gvdbx_autotune_gather: returns the chosen configuration and times."""
        from ._lib import TuneGather
        t = TuneGather()
        call("gvdbx_autotune_gather", self._h, int(mode), int(subcell_size), float(radius), int(chan), int(reps),
             C.byref(t))
        return {k: getattr(t, k) for k, _ in TuneGather._fields_}

    def AutotuneBinning(self, num_pnts, subcell_size, radius, reps=3):
        """gvdbx_autotune_binning: returns the chosen configuration and times."""
        from ._lib import TuneBinning
        t = TuneBinning()
        call("gvdbx_autotune_binning", self._h, int(num_pnts), int(subcell_size), float(radius), int(reps),
             C.byref(t))
        return {k: getattr(t, k) for k, _ in TuneBinning._fields_}

    def SaveVBX(self, fname):
        call("gvdbx_save_vbx", self._h, str(fname).encode())

    def LoadVBX(self, fname, force_maj=0, force_min=0):
        """LoadVBX (VG.cpp:507-685); force_maj / force_min > 0 read the file
        as that version whatever its header says."""
        call("gvdbx_load_vbx_version", self._h, str(fname).encode(), int(force_maj), int(force_min))
        inf = self.info()
        self.log2dims = tuple(int(inf.log2dim[l]) for l in range(inf.levs))
        return True

    def SetTransform(self, pretrans, scale, angs, trans):
        call("gvdbx_set_transform", self._h, _v3(pretrans), _v3(scale), _v3(angs), _v3(trans))

    def GetTransform(self):
        out = [(C.c_float * 3)() for _ in range(4)]
        call("gvdbx_get_transform", self._h, *out)
        return tuple(tuple(float(x) for x in a) for a in out)   # pretrans, scale, angs, trans

    def FinishTopology(self, commit_pool=False, compute_bounds=True):
        call("gvdbx_finish_topology", self._h, int(commit_pool), int(compute_bounds))

    def UpdateAtlas(self):
        call("gvdbx_update_atlas", self._h)

    # ------------------------------------------------------------ points -> voxels
    def InsertPointsSubcell(self, subcell_size, num_pnts, radius, trans=(0.0, 0.0, 0.0)):
        out = C.c_int64(0)
        call("gvdbx_insert_points_subcell", self._h, int(subcell_size), int(num_pnts), float(radius),
             *[float(x) for x in trans], C.byref(out))
        return out.value

    def InsertPointsSubcell_FP16(self, subcell_size, num_pnts, radius, trans=(0.0, 0.0, 0.0)):
        out = C.c_int64(0)
        call("gvdbx_insert_points_subcell_fp16", self._h, int(subcell_size), int(num_pnts), float(radius),
             *[float(x) for x in trans], C.byref(out))
        return out.value

    def GatherLevelSet_FP16(self, subcell_size, num_pnts, radius, trans, sc_len, chan_levelset=0, chan_clr=-1):
        call("gvdbx_gather_level_set_fp16", self._h, int(subcell_size), int(num_pnts), float(radius),
             *[float(x) for x in trans], int(sc_len), int(chan_levelset), int(chan_clr))

    def GatherDensity(self, subcell_size, num_pnts, radius, trans, sc_len, chan_density=0, chan_clr=-1, accum=False):
        call("gvdbx_gather_density", self._h, int(subcell_size), int(num_pnts), float(radius),
             *[float(x) for x in trans], int(sc_len), int(chan_density), int(chan_clr), int(accum))

    def GatherLevelSet(self, subcell_size, num_pnts, radius, trans, sc_len, chan_levelset=0, chan_clr=-1,
                       accum=False):
        call("gvdbx_gather_level_set", self._h, int(subcell_size), int(num_pnts), float(radius),
             *[float(x) for x in trans], int(sc_len), int(chan_levelset), int(chan_clr), int(accum))

    def GatherLevelSetNoAlloc(self, subcell_size, num_pnts, radius, trans, sc_len, chan_levelset=0, chan_clr=-1,
                              accum=False):
        """GatherLevelSet that allocates no device memory (gvdbx_gather_level_set_noalloc)."""
        call("gvdbx_gather_level_set_noalloc", self._h, int(subcell_size), int(num_pnts), float(radius),
             *[float(x) for x in trans], int(sc_len), int(chan_levelset), int(chan_clr), int(accum))

    def InsertPoints(self, num_pnts, trans=(0.0, 0.0, 0.0), prefix=True):
        call("gvdbx_insert_points", self._h, int(num_pnts), *[float(x) for x in trans], int(prefix))

    def ScatterDensity(self, num_pnts, radius, amp=1.0, trans=(0.0, 0.0, 0.0), expand=True, avg_color=False):
        call("gvdbx_scatter_density", self._h, int(num_pnts), float(radius), float(amp),
             *[float(x) for x in trans], int(expand), int(avg_color))

    def SplatCubic(self, num_pnts, radius, trans=(0.0, 0.0, 0.0), chan=0, accum=False):
        call("gvdbx_splat_cubic", self._h, int(num_pnts), float(radius), *[float(x) for x in trans], int(chan),
             int(accum))

    def UpdateApron(self, chan=0, boundval=0.0):
        call("gvdbx_update_apron", self._h, int(chan), float(boundval))

    # ------------------------------------------------------------ rendering
    def SetColorChannel(self, chan):
        """SetColorChannel(uchar chan) (VG.cpp:686-690): the uchar4 channel
        that colours Render; -1 or 255 for none."""
        call("gvdbx_set_color_channel", self._h, int(chan))

    def GetColorChannel(self):
        out = C.c_int(0)
        call("gvdbx_get_color_channel", self._h, C.byref(out))
        return out.value

    def Render(self, shade, chan, cam, rp, rgba=None, rgba8=None):
        """Render(shade, chan, buf) (VG.cpp:4336-4381).  Returns (rgba float32
        (H,W,4), rgba8 uint8 (H,W,4)) device tensors."""
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        if rgba is None:
            rgba = torch.empty((cam.height, cam.width, 4), dtype=torch.float32, device=dev)
        if rgba8 is None:
            rgba8 = torch.empty((cam.height, cam.width, 4), dtype=torch.uint8, device=dev)
        call("gvdbx_render", self._h, int(shade), int(chan), C.byref(cam), C.byref(rp), C.c_void_p(_ptr(rgba)),
             C.c_void_p(_ptr(rgba8)))
        return rgba, rgba8

    def Raytrace(self, rays, chan, rp, bias=0.0):
        """Raytrace (VG.cpp:4384-4408) over a device uint8 tensor of 64-byte
        ray records (gvdbx_scn_ray), updated in place."""
        assert rays.is_cuda and rays.numel() % 64 == 0
        call("gvdbx_raytrace", self._h, int(chan), C.byref(rp), C.c_void_p(_ptr(rays)), rays.numel() // 64, float(bias))
        return rays

    def RaycastEA(self, chan, cam, rp, I=None, T=None, record=False):
        """Linear-transfer emission-absorption raycast; (I, T) float32 (H,W)."""
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        if I is None:
            I = torch.empty((cam.height, cam.width), dtype=torch.float32, device=dev)
        if T is None:
            T = torch.empty((cam.height, cam.width), dtype=torch.float32, device=dev)
        call("gvdbx_raycast_ea", self._h, int(chan), C.byref(cam), C.byref(rp), C.c_void_p(_ptr(I)),
             C.c_void_p(_ptr(T)), int(record))
        return I, T

    def RaycastEABatched(self, chan, cams, rp, I=None, T=None, record=False):
        """One camera per batched cloud; (I, T) of shape (ncloud, H, W)."""
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        n = len(cams)
        arr = (Camera * n)(*cams)
        self._cams = arr
        if I is None:
            I = torch.empty((n, cams[0].height, cams[0].width), dtype=torch.float32, device=dev)
        if T is None:
            T = torch.empty((n, cams[0].height, cams[0].width), dtype=torch.float32, device=dev)
        call("gvdbx_raycast_ea_batched", self._h, int(chan), arr, n, C.byref(rp), C.c_void_p(_ptr(I)),
             C.c_void_p(_ptr(T)), int(record))
        return I, T

    def RaycastEABackwardBatched(self, chan, cams, rp, I, T, gI, gT=None, grad=None):
        import torch
        inf = self.info()
        br, B = inf.brick_res, inf.node_cnt[0]
        if grad is None:
            grad = torch.empty((B, br, br, br), dtype=torch.float32, device=I.device)
        n = len(cams)
        arr = (Camera * n)(*cams)
        call("gvdbx_raycast_ea_backward_batched", self._h, int(chan), arr, n, C.byref(rp), C.c_void_p(_ptr(I)),
             C.c_void_p(_ptr(T)), C.c_void_p(_ptr(gI)), C.c_void_p(_ptr(gT)) if gT is not None else None,
             C.c_void_p(_ptr(grad)))
        return grad

    def RaycastEABackward(self, chan, cam, rp, I, T, gI, gT=None, grad=None):
        """d(sum gI*I + gT*T)/d voxel over the padded bricks, (B, br, br, br)."""
        import torch
        inf = self.info()
        br, B = inf.brick_res, inf.node_cnt[0]
        if grad is None:
            grad = torch.empty((B, br, br, br), dtype=torch.float32, device=I.device)
        call("gvdbx_raycast_ea_backward", self._h, int(chan), C.byref(cam), C.byref(rp), C.c_void_p(_ptr(I)),
             C.c_void_p(_ptr(T)), C.c_void_p(_ptr(gI)), C.c_void_p(_ptr(gT)) if gT is not None else None,
             C.c_void_p(_ptr(grad)))
        return grad

    def ApronAdjoint(self, grad):
        call("gvdbx_apron_adjoint", self._h, C.c_void_p(_ptr(grad)))
        return grad

    def SplatCubicBackward(self, num_pnts, radius, trans, grad, dpos=None, dw=None, want_dw=True):
        import torch
        if dpos is None:
            dpos = torch.empty((num_pnts, 3), dtype=torch.float32, device=grad.device)
        if dw is None and want_dw:
            dw = torch.empty((num_pnts,), dtype=torch.float32, device=grad.device)
        call("gvdbx_splat_cubic_backward", self._h, int(num_pnts), float(radius), *[float(x) for x in trans],
             C.c_void_p(_ptr(grad)), C.c_void_p(_ptr(dpos)), C.c_void_p(_ptr(dw)) if dw is not None else None)
        return dpos, dw

    # ------------------------------------------------------------ readback
    def info(self):
        inf = Info()
        call("gvdbx_get_info", self._h, C.byref(inf))
        return inf

    def nodes(self, lev):
        n = self.info().node_cnt[lev]
        out = np.zeros(n, dtype=NODE_DTYPE)
        if n:
            call("gvdbx_read_nodes", self._h, lev, out.ctypes.data, n)
        return out

    def child_lists(self, lev):
        inf = self.info()
        r = 1 << self.log2dims[lev]
        n = inf.node_cnt[lev]
        out = np.zeros((n, r ** 3), dtype=np.uint64)
        if n:
            call("gvdbx_read_child_lists", self._h, lev, out.ctypes.data, out.size)
        return out

    def atlas_map(self):
        inf = self.info()
        n = inf.subdim[0] * inf.subdim[1] * inf.subdim[2]
        out = np.zeros(n, dtype=AMAP_DTYPE)
        call("gvdbx_read_atlas_map", self._h, out.ctypes.data, n)
        return out

    def bricks_u8(self, chan=1):
        """A uchar4 channel's bricks incl. apron, (B, br, br, br, 4) uint8."""
        b = self.bricks(chan)
        return b.view(np.uint8).reshape(b.shape + (4,))

    def bricks(self, chan=0):
        """All leaf bricks incl. apron, (B, br, br, br) indexed [k, z, y, x]."""
        inf = self.info()
        br = inf.brick_res
        B = inf.node_cnt[0]
        out = np.zeros((B, br, br, br), dtype=np.float32)
        call("gvdbx_read_bricks", self._h, chan, out.ctypes.data, out.size)
        return out

    def channel_ptr(self, chan=0):
        p = C.c_void_p()
        call("gvdbx_channel_device_ptr", self._h, chan, C.byref(p))
        return p.value

    def point_nodes(self, n):
        pnode = np.zeros(n, dtype=np.int32)
        poff = np.zeros(n, dtype=np.int32)
        call("gvdbx_read_point_nodes", self._h, pnode.ctypes.data, poff.ctypes.data, n)
        return pnode, poff


def backproject_depth(depth, intr, poses, voxel, origin, dmin=0.5, dmax=8.0, out=None, stream=None):
    """Index-space points of every pixel of a batch of depth frames (one launch).
    depth: (F,H,W) float32 device tensor; poses: (F,12) device tensor; intr:
    (fx, fy, cx, cy).  Returns (F*H*W, 3); invalid pixels are (0, 0, NOHIT)."""
    import torch
    F, H, W = depth.shape
    if out is None:
        out = torch.empty((F * H * W, 3), dtype=torch.float32, device=depth.device)
    fi = (C.c_float * 4)(*[float(x) for x in intr])
    fo = (C.c_float * 3)(*[float(x) for x in origin])
    handle = getattr(stream, "cuda_stream", stream) if stream is not None else torch.cuda.current_stream().cuda_stream
    call("gvdbx_backproject_depth", C.c_void_p(_ptr(depth.contiguous())), int(F), int(W), int(H), fi,
         C.c_void_p(_ptr(poses.contiguous())), float(voxel), fo, float(dmin), float(dmax), C.c_void_p(_ptr(out)),
         C.c_void_p(handle))
    return out


def gen_shell(seed, n, center=(256.0, 256.0, 256.0), r0=180.0, r1=200.0):
    """Seeded points uniform by volume in a spherical shell (host, float32 (n,3))."""
    out = np.zeros((n, 3), dtype=np.float32)
    call("gvdbx_gen_shell", C.c_uint64(seed), n, *[float(c) for c in center], float(r0), float(r1), out.ctypes.data)
    return out


def gen_box(seed, n, lo, hi, jitter=0.0):
    """Seeded points uniform in the box [lo, hi) plus per-axis jitter U(-j, j)."""
    out = np.zeros((n, 3), dtype=np.float32)
    flo = (C.c_float * 3)(*[float(x) for x in lo])
    fhi = (C.c_float * 3)(*[float(x) for x in hi])
    call("gvdbx_gen_box", C.c_uint64(seed), n, flo, fhi, float(jitter), out.ctypes.data)
    return out
