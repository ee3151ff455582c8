"""ctypes binding of libgvdbx.so (declarations mirror include/gvdbx.h)."""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libgvdbx.so")

NODE_DTYPE = np.dtype([
    ("lev", np.uint8), ("flags", np.uint8), ("prio", np.uint8), ("pad", np.uint8),
    ("pos", np.int32, 3), ("value", np.int32, 3), ("vrange", np.float32, 3),
    ("parent", np.uint64), ("childlist", np.uint64), ("mask", np.uint64),
])
assert NODE_DTYPE.itemsize == 64

STATUS = {
    0: "GVDBX_OK", 1: "GVDBX_ERR_ARG", 2: "GVDBX_ERR_DEPTH0", 3: "GVDBX_ERR_DEPTH_MAX",
    4: "GVDBX_ERR_CAPACITY", 5: "GVDBX_ERR_HIP", 6: "GVDBX_ERR_STATE", 7: "GVDBX_ERR_NOMEM",
    8: "GVDBX_ERR_DEGENERATE",
}

VBX_MAXLEV = 10
VBX_MAXCHAN = 32


class VbxDesc(C.Structure):
    """gvdbx_vbx_desc (include/gvdbx.h): one grid of a VBX file on host memory."""
    _fields_ = [
        ("levels", C.c_int32), ("log2dim", C.c_int32 * VBX_MAXLEV), ("apron", C.c_int32),
        ("axis_count", C.c_int32 * 3), ("nchan", C.c_int32), ("chan_type", C.c_int32 * VBX_MAXCHAN),
        ("transform", C.c_float * 12), ("root", C.c_uint64), ("node_count", C.c_int64 * VBX_MAXLEV),
        ("clist_count", C.c_int64 * VBX_MAXLEV), ("slots", C.c_int64), ("nodes", C.c_void_p * VBX_MAXLEV),
        ("clists", C.c_void_p * VBX_MAXLEV), ("chan_data", C.c_void_p * VBX_MAXCHAN),
    ]


_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_F = C.c_float
_PF = C.POINTER(C.c_float)

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "gvdbx_ctx_create": [_I, _P, C.POINTER(_P)],
    "gvdbx_ctx_destroy": [_P],
    "gvdbx_status_string": [_I],
    "gvdbx_last_error": [_P, C.c_char_p, _I],
    "gvdbx_synchronize": [_P],
    "gvdbx_get_stream": [_P, C.POINTER(_P)],
    "gvdbx_configure": [_P, _I, _P],
    "gvdbx_set_channel_default": [_P, _I, _I, _I],
    "gvdbx_add_channel": [_P, _I, _I, _I],
    "gvdbx_fill_channel": [_P, _I, _F],
    "gvdbx_fill_channel4": [_P, _I, _P],
    "gvdbx_set_points": [_P, _I64, _P, _I, _I, _P, _I, _I, _P, _I, _I],
    "gvdbx_advance_points": [_P, _I64, _F],
    "gvdbx_batch_set_points": [_P, _I, _P, _P, _I, _I, _P, _I, _I, _I, _F],
    "gvdbx_batch_get_layout": [_P, C.POINTER(_I), C.POINTER(_I), _P],
    "gvdbx_rebuild_topology": [_P, _I64, _F, _PF],
    "gvdbx_accumulate_topology": [_P, _I64, _F, _PF, _I],
    "gvdbx_request_full_rebuild": [_P, _I],
    "gvdbx_finish_topology": [_P, _PF, _PF],
    "gvdbx_update_atlas": [_P],
    "gvdbx_insert_points_subcell": [_P, _I, _I64, _F, _PF, C.POINTER(_I64)],
    "gvdbx_gather_density": [_P, _I, _I64, _F, _PF, _I64, _I, _I, _I],
    "gvdbx_gather_levelset": [_P, _I, _I64, _F, _PF, _I64, _I, _I, _I],
    "gvdbx_insert_points_subcell_fp16": [_P, _I, _I64, _F, _PF, C.POINTER(_I64)],
    "gvdbx_gather_levelset_fp16": [_P, _I, _I64, _F, _PF, _I64, _I, _I],
    "gvdbx_get_min_max_vel": [_P, _I64, _PF],
    "gvdbx_gather_velocity": [_P, _I, _I, _I, _I],
    "gvdbx_set_subcell_options": [_P, _I],
    "gvdbx_scatter_reduce_levelset": [_P, _F, _I, _F],
    "gvdbx_update_apron": [_P, _I, _F],
    "gvdbx_update_apron_channels": [_P, _P, _P, _I],
    "gvdbx_sample_trilinear": [_P, _I, _P, _I64, _P],
    "gvdbx_splat_trilinear": [_P, _I, _P, _P, _I64, _I],
    "gvdbx_splat_trilinear_noalloc": [_P, _I, _P, _P, _I64, _I],
    "gvdbx_scatter_density": [_P, _I64, _F, _F, _P, _I, _I, _I, _I],
    "gvdbx_gather_density_velocity": [_P, _I, _I64, _F, _P, _I64, _I, _I, _I, _I],
    "gvdbx_splat_backward": [_P, _I, _P, _P, _I64, _P, _P],
    "gvdbx_sample_trilinear_backward": [_P, _I, _P, _P, _I64, _P, _I, _I],
    "gvdbx_get_depth": [_P, C.POINTER(_I), _P, _P, _P],
    "gvdbx_get_node_count": [_P, _I, C.POINTER(_I64), C.POINTER(_I64)],
    "gvdbx_copy_nodes": [_P, _I, _P, _I64],
    "gvdbx_copy_childlists": [_P, _I, _P, _I64],
    "gvdbx_get_num_slots": [_P, C.POINTER(_I64), C.POINTER(_I)],
    "gvdbx_copy_channel": [_P, _I, _P, _I64],
    "gvdbx_get_channel_ptr": [_P, _I, C.POINTER(_P), C.POINTER(_I64)],
    "gvdbx_get_subcell_info": [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I)],
    "gvdbx_copy_subcell_array": [_P, _I, _P, _I64],
    "gvdbx_save_vbx": [_P, C.c_char_p],
    "gvdbx_load_vbx": [_P, C.c_char_p],
    "gvdbx_vbx_write": [C.c_char_p, C.POINTER(VbxDesc)],
    "gvdbx_vbx_probe": [C.c_char_p, C.POINTER(VbxDesc)],
    "gvdbx_vbx_read": [C.c_char_p, C.POINTER(VbxDesc)],
    "gvdbx_dev_alloc": [_P, _I64, C.POINTER(_P)],
    "gvdbx_dev_free": [_P, _P],
    "gvdbx_commit": [_P, _P, _P, _I64],
    "gvdbx_retrieve": [_P, _P, _P, _I64],
    "gvdbx_set_launch": [_P, _I, _I],
    "gvdbx_synth_uniform": [C.c_uint64, _I64, _P, _P, _P],
    "gvdbx_synth_sphere": [C.c_uint64, _I64, _P, _F, _F, _P],
    "gvdbx_synth_lattice": [C.c_uint64, _P, _F, _F, _F, _P, _P],
    "gvdbx_graph_begin": [_P],
    "gvdbx_graph_end": [_P],
    "gvdbx_graph_launch": [_P],
    "gvdbx_graph_release": [_P],
    "gvdbx_kernel_timer": [_P, C.c_char_p, _I],
    "gvdbx_kernel_timer_read": [_P, C.POINTER(C.c_double), C.POINTER(_I64)],
}
_RESTYPE = {"gvdbx_status_string": C.c_char_p}

_lib = None


class GvdbxError(RuntimeError):
    def __init__(self, code, what, detail=""):
        super().__init__(f"{what}: {STATUS.get(code, code)} {detail}".strip())
        self.code = code


def load_library(path=None):
    """Load libgvdbx.so; raises when it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise GvdbxError(-1, "load_library", f"{p} not built (run __graft_entry__.build())")
    L = C.CDLL(p)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    if path is None:
        _lib = L
    return L


def fptr(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.ctypes.data_as(_PF)


class DeviceArray:
    """Device buffer owned through the library's allocation helpers (AllocData semantics)."""

    def __init__(self, vol, nbytes):
        self.vol = vol
        self.nbytes = int(nbytes)
        p = _P()
        vol._chk(vol.L.gvdbx_dev_alloc(vol.h, self.nbytes, C.byref(p)), "dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, vol, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(vol, arr.nbytes)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.vol._chk(self.vol.L.gvdbx_commit(self.vol.h, self.ptr, arr.ctypes.data, arr.nbytes), "commit")

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(n, dtype)
        if out.nbytes:
            self.vol._chk(self.vol.L.gvdbx_retrieve(self.vol.h, out.ctypes.data, self.ptr, out.nbytes), "retrieve")
        return out

    def free(self):
        if self.ptr and self.vol.h:
            self.vol.L.gvdbx_dev_free(self.vol.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
