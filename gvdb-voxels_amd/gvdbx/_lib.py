"""ctypes binding of libgvdbx.so (the C-ABI declared in include/gvdbx/gvdbx_c.h)."""
import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libgvdbx.so")

OK = 0
STATUS_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_STATE", 3: "ERR_DEPTH", 4: "ERR_NOPOINTS", 5: "ERR_RANGE",
                6: "ERR_HIP", 7: "ERR_NOMEM"}

TOPO_AUTO, TOPO_DENSE, TOPO_RADIX = 0, 1, 2


class GvdbxError(RuntimeError):
    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed: {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code
        self.func = func


class Info(C.Structure):
    _fields_ = [("levs", C.c_int), ("depth", C.c_int), ("top_lev", C.c_int), ("root_pos", C.c_int * 3),
                ("node_cnt", C.c_int64 * 10), ("subdim", C.c_int * 3), ("brick_res", C.c_int), ("apron", C.c_int),
                ("pos_min", C.c_float * 3), ("pos_max", C.c_float * 3), ("obj_min", C.c_float * 3),
                ("obj_max", C.c_float * 3), ("sc_len", C.c_int64), ("topology_mode_used", C.c_int),
                ("log2dim", C.c_int * 10)]


class TuneGather(C.Structure):
    _fields_ = [("gather_variant", C.c_int), ("gather_classes", C.c_int), ("gather_split", C.c_int),
                ("xcd_remap", C.c_int), ("best_ms", C.c_float), ("default_ms", C.c_float), ("candidates", C.c_int)]


class TuneBinning(C.Structure):
    _fields_ = [("bin_records", C.c_int), ("sort_tiles", C.c_int), ("best_ms", C.c_float), ("default_ms", C.c_float),
                ("candidates", C.c_int)]


class Camera(C.Structure):
    """gvdbx_camera: pixel (x,y), u=(x+0.5)/width, w=(y+0.5)/height, ray
    origin (u*org_u + w*org_v) + org, direction normalize((u*dir_u + w*dir_v) + dir)."""
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("org", C.c_float * 3), ("org_u", C.c_float * 3),
                ("org_v", C.c_float * 3), ("dir", C.c_float * 3), ("dir_u", C.c_float * 3),
                ("dir_v", C.c_float * 3)]


class RenderParams(C.Structure):
    """gvdbx_render_params (ScnInfo fields of PrepareRender VG.cpp:4254-4306)."""
    _fields_ = [("step", C.c_float), ("extinct", C.c_float), ("albedo", C.c_float), ("minval", C.c_float),
                ("alphacut", C.c_float), ("thresh", C.c_float), ("vmin", C.c_float), ("vmax", C.c_float),
                ("epsilon", C.c_float), ("max_iter", C.c_int), ("backclr", C.c_float * 4), ("light", C.c_float * 3),
                ("shadow_amt", C.c_float), ("shadow_bias", C.c_float), ("lut_dev", C.c_void_p),
                ("finestep", C.c_float), ("slice_pnt", C.c_float * 3), ("slice_norm", C.c_float * 3)]


# name -> argtypes (all return int status unless listed in _RESTYPES)
_vp, _i, _i64, _f, _u64 = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_uint64
_SIGS = {
    "gvdbx_create": [C.POINTER(_vp), _i],
    "gvdbx_destroy": [_vp],
    "gvdbx_set_stream": [_vp, _vp],
    "gvdbx_synchronize": [_vp],
    "gvdbx_configure": [_vp, _i, C.POINTER(C.c_int)],
    "gvdbx_set_channel_default": [_vp, _i, _i, _i],
    "gvdbx_add_channel": [_vp, _i, _i, _i, _i, _i],
    "gvdbx_fill_channel": [_vp, _i, _f],
    "gvdbx_set_topology_mode": [_vp, _i],
    "gvdbx_set_tuning": [_vp, C.c_char_p, _i],
    "gvdbx_get_tuning": [_vp, C.c_char_p, C.POINTER(C.c_int)],
    "gvdbx_set_points": [_vp, _vp, _i, _i, _vp, _i, _i, _vp, _i, _i],
    "gvdbx_set_point_weights": [_vp, _vp, _i],
    "gvdbx_rebuild_topology": [_vp, _i, _f, _f, _f, _f],
    "gvdbx_raytrace": [_vp, _i, _vp, _vp, _i64, _f],
    "gvdbx_save_vbx": [_vp, C.c_char_p],
    "gvdbx_autotune_gather": [_vp, _i, _i, _f, _i, _i, _vp],
    "gvdbx_autotune_binning": [_vp, _i, _i, _f, _i, _vp],
    "gvdbx_frame_graph_create": [_vp, _i, _i, _f, _f, _i, _i, _i, _f, C.POINTER(_vp)],
    "gvdbx_frame_graph_launch": [_vp],
    "gvdbx_frame_graph_destroy": [_vp],
    "gvdbx_load_vbx": [_vp, C.c_char_p],
    "gvdbx_load_vbx_version": [_vp, C.c_char_p, _i, _i],
    "gvdbx_set_transform": [_vp, C.POINTER(_f), C.POINTER(_f), C.POINTER(_f), C.POINTER(_f)],
    "gvdbx_get_transform": [_vp, C.POINTER(_f), C.POINTER(_f), C.POINTER(_f), C.POINTER(_f)],
    "gvdbx_accumulate_topology": [_vp, _i, _f, _f, _f, _f, _i],
    "gvdbx_request_full_rebuild": [_vp, _i],
    "gvdbx_finish_topology": [_vp, _i, _i],
    "gvdbx_update_atlas": [_vp],
    "gvdbx_insert_points_subcell": [_vp, _i, _i, _f, _f, _f, _f, C.POINTER(C.c_int64)],
    "gvdbx_gather_density": [_vp, _i, _i, _f, _f, _f, _f, _i64, _i, _i, _i],
    "gvdbx_gather_level_set": [_vp, _i, _i, _f, _f, _f, _f, _i64, _i, _i, _i],
    "gvdbx_gather_level_set_noalloc": [_vp, _i, _i, _f, _f, _f, _f, _i64, _i, _i, _i],
    "gvdbx_insert_points_subcell_fp16": [_vp, _i, _i, _f, _f, _f, _f, C.POINTER(C.c_int64)],
    "gvdbx_gather_level_set_fp16": [_vp, _i, _i, _f, _f, _f, _f, _i64, _i, _i],
    "gvdbx_insert_points": [_vp, _i, _f, _f, _f, _i],
    "gvdbx_scatter_density": [_vp, _i, _f, _f, _f, _f, _f, _i, _i],
    "gvdbx_splat_cubic": [_vp, _i, _f, _f, _f, _f, _i, _i],
    "gvdbx_update_apron": [_vp, _i, _f],
    "gvdbx_render": [_vp, _i, _i, C.POINTER(Camera), C.POINTER(RenderParams), _vp, _vp],
    "gvdbx_set_color_channel": [_vp, _i],
    "gvdbx_get_color_channel": [_vp, C.POINTER(C.c_int)],
    "gvdbx_raycast_ea": [_vp, _i, C.POINTER(Camera), C.POINTER(RenderParams), _vp, _vp, _i],
    "gvdbx_raycast_ea_backward": [_vp, _i, C.POINTER(Camera), C.POINTER(RenderParams), _vp, _vp, _vp, _vp, _vp],
    "gvdbx_apron_adjoint": [_vp, _vp],
    "gvdbx_splat_cubic_backward": [_vp, _i, _f, _f, _f, _f, _vp, _vp, _vp],
    "gvdbx_backproject_depth": [_vp, _i, _i, _i, C.POINTER(C.c_float), _vp, _f, C.POINTER(C.c_float), _f, _f, _vp,
                                _vp],
    "gvdbx_set_points_batched": [_vp, _vp, _i, C.POINTER(C.c_int64), _i],
    "gvdbx_read_cloud_layout": [_vp, _vp, _vp, _i],
    "gvdbx_read_leaf_clouds": [_vp, _vp, _i64],
    "gvdbx_raycast_ea_batched": [_vp, _i, C.POINTER(Camera), _i, C.POINTER(RenderParams), _vp, _vp, _i],
    "gvdbx_raycast_ea_backward_batched": [_vp, _i, C.POINTER(Camera), _i, C.POINTER(RenderParams), _vp, _vp, _vp,
                                          _vp, _vp],
    "gvdbx_get_info": [_vp, C.POINTER(Info)],
    "gvdbx_read_nodes": [_vp, _i, _vp, _i64],
    "gvdbx_read_child_lists": [_vp, _i, _vp, _i64],
    "gvdbx_read_atlas_map": [_vp, _vp, _i64],
    "gvdbx_read_bricks": [_vp, _i, _vp, _i64],
    "gvdbx_channel_device_ptr": [_vp, _i, C.POINTER(_vp)],
    "gvdbx_read_point_nodes": [_vp, _vp, _vp, _i64],
    "gvdbx_gen_shell": [_u64, _i64, _f, _f, _f, _f, _f, _vp],
    "gvdbx_gen_box": [_u64, _i64, C.POINTER(C.c_float), C.POINTER(C.c_float), _f, _vp],
    "gvdbx_last_error": [],
    "gvdbx_version": [],
}
_RESTYPES = {"gvdbx_last_error": C.c_char_p, "gvdbx_version": C.c_char_p}

_lib = None


def exported_names():
    return sorted(_SIGS)


def lib():
    """Load libgvdbx.so; raise if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C gvdb-voxels_amd); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, C.c_int)
        _lib = L
    return _lib


def last_error():
    return lib().gvdbx_last_error().decode(errors="replace")


def call(name, *args):
    code = getattr(lib(), name)(*args)
    if code != OK:
        raise GvdbxError(code, name, last_error())
    return code
