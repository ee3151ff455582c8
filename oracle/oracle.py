"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NODE_DTYPE = np.dtype([
    ("lev", np.uint8), ("flags", np.uint8), ("prio", np.uint8), ("pad", np.uint8),
    ("pos", np.int32, 3), ("value", np.int32, 3), ("vrange", np.float32, 3),
    ("parent", np.uint64), ("childlist", np.uint64), ("mask", np.uint64),
])
assert NODE_DTYPE.itemsize == 64

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i64 = C.c_int64
        f32 = C.c_float
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_int, P, C.c_int, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [P]
        L.orc_rebuild_topology.argtypes = [P, P, i64, i64, f32, P]
        L.orc_accumulate_topology.argtypes = [P, P, i64, i64, f32, P, C.c_int]
        L.orc_request_full_rebuild.argtypes = [P, C.c_int]
        L.orc_update_atlas.argtypes = [P]
        for fn in ("orc_node_count", "orc_clist_count"):
            getattr(L, fn).restype = i64
            getattr(L, fn).argtypes = [P, C.c_int]
        for fn in ("orc_used0", "orc_num_slots", "orc_sc_num", "orc_channel_size"):
            getattr(L, fn).restype = i64
            getattr(L, fn).argtypes = [P]
        L.orc_nodes.restype = P
        L.orc_nodes.argtypes = [P, C.c_int]
        L.orc_clists.restype = P
        L.orc_clists.argtypes = [P, C.c_int]
        L.orc_depth.argtypes = [P]
        L.orc_rootpos.argtypes = [P, P]
        L.orc_posminmax.argtypes = [P, P]
        L.orc_leaf_slot.restype = P
        L.orc_leaf_slot.argtypes = [P]
        L.orc_insert_subcell.argtypes = [P, C.c_int, P, i64, i64, f32, P, P]
        for fn in ("orc_sc_map", "orc_sc_cnt", "orc_sc_off", "orc_sc_pos", "orc_sc_nid", "orc_sc_list",
                   "orc_sc_list_pos"):
            getattr(L, fn).restype = P
            getattr(L, fn).argtypes = [P]
        L.orc_channel_alloc.argtypes = [P, C.c_int]
        L.orc_channel.restype = P
        L.orc_channel.argtypes = [P, C.c_int]
        L.orc_fill_channel.argtypes = [P, C.c_int, C.c_double]
        L.orc_gather_density.argtypes = [P, f32, C.c_int, C.c_int]
        L.orc_gather_levelset.argtypes = [P, f32, C.c_int, C.c_int]
        L.orc_gather_velocity.argtypes = [P, P, i64, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_gather_colour.argtypes = [P, P, i64, C.c_int, C.c_int]
        L.orc_quantize_list_pos.argtypes = [P, P, P, P]
        L.orc_quantize_list_vel.argtypes = [P, P, i64, P, P, P]
        L.orc_gather_levelset_q16.argtypes = [P, f32, C.c_int, P, P, P]
        L.orc_scatter_colour.argtypes = [P, C.c_int, P, i64, i64, P, P, C.c_int]
        L.orc_update_apron.argtypes = [P, C.c_int, C.c_double]
        L.orc_scatter_reduce_levelset.argtypes = [P, f32, C.c_int, C.c_double]
        L.orc_sample_trilinear.argtypes = [P, C.c_int, P, i64, P]
        L.orc_splat_trilinear.argtypes = [P, C.c_int, P, P, i64]
        L.orc_scatter_density.argtypes = [P, C.c_int, P, i64, i64, C.c_float, P, C.c_int]
        L.orc_splat_backward.argtypes = [P, C.c_int, P, P, i64, P, P]
        L.orc_sample_backward.argtypes = [P, C.c_int, P, P, i64, P]
        L.orc_find_leaf.restype = i64
        L.orc_find_leaf.argtypes = [P, P]
        L.orc_cpu_activate_space.restype = i64
        L.orc_cpu_activate_space.argtypes = [C.c_int, P, P, i64, i64, P, P]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n):
    """Threads of the binning, gather and apron loops (1 = the serial loops; same results)."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _pts(pos):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    assert pos.ndim == 2 and pos.shape[1] >= 3
    return pos, pos.shape[1]


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed with status {code}")
        self.code = code


def _chk(code, what):
    if code != 0:
        raise OracleError(code, what)


class Tree:
    """One oracle volume (pools, atlas, subcells, channels)."""

    def __init__(self, log2dims=(3, 3, 3, 3, 5), atlas_cx=16, atlas_cy=16, apron=1):
        L = lib()
        ld = np.asarray(log2dims, dtype=np.int32)
        self.levs = len(ld)
        self.log2dims = tuple(int(x) for x in ld)
        self.apron = apron
        self.h = L.orc_create(self.levs, _ptr(ld), atlas_cx, atlas_cy, apron)
        assert self.h
        self.res0 = 1 << self.log2dims[0]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    # --- topology
    def rebuild_topology(self, pos, radius, trans=(0, 0, 0)):
        pos, st = _pts(pos)
        tr = np.asarray(trans, dtype=np.float32)
        _chk(lib().orc_rebuild_topology(self.h, _ptr(pos), len(pos), st, radius, _ptr(tr)), "rebuild_topology")

    def accumulate_topology(self, pos, radius, trans=(0, 0, 0), min_depth=0):
        pos, st = _pts(pos)
        tr = np.asarray(trans, dtype=np.float32)
        _chk(lib().orc_accumulate_topology(self.h, _ptr(pos), len(pos), st, radius, _ptr(tr), min_depth),
             "accumulate_topology")

    def request_full_rebuild(self, flag=True):
        lib().orc_request_full_rebuild(self.h, int(bool(flag)))

    def update_atlas(self):
        _chk(lib().orc_update_atlas(self.h), "update_atlas")

    def depth(self):
        return lib().orc_depth(self.h)

    def rootpos(self):
        o = np.zeros(3, np.int32)
        lib().orc_rootpos(self.h, _ptr(o))
        return o

    def posminmax(self):
        o = np.zeros(6, np.float32)
        lib().orc_posminmax(self.h, _ptr(o))
        return o

    def node_count(self, lev):
        return lib().orc_node_count(self.h, lev)

    def used0(self):
        return lib().orc_used0(self.h)

    def nodes(self, lev):
        n = self.node_count(lev)
        if n == 0:
            return np.zeros(0, NODE_DTYPE)
        p = lib().orc_nodes(self.h, lev)
        return np.frombuffer((C.c_char * (n * 64)).from_address(p), dtype=NODE_DTYPE).copy()

    def childlists(self, lev):
        n = lib().orc_clist_count(self.h, lev)
        per = 1 << (3 * self.log2dims[lev])
        if n == 0:
            return np.zeros((0, per), np.uint64)
        p = lib().orc_clists(self.h, lev)
        return np.frombuffer((C.c_char * (n * per * 8)).from_address(p), dtype=np.uint64).reshape(n, per).copy()

    def num_slots(self):
        return lib().orc_num_slots(self.h)

    def leaf_slot(self):
        n = self.node_count(0)
        p = lib().orc_leaf_slot(self.h)
        if n == 0 or not p:
            return np.zeros(0, np.int64)
        return np.frombuffer((C.c_char * (n * 8)).from_address(p), dtype=np.int64).copy()

    def find_leaf(self, p):
        a = np.asarray(p, np.int32)
        return lib().orc_find_leaf(self.h, _ptr(a))

    # --- subcells
    def insert_points_subcell(self, sc, pos, radius, trans=(0, 0, 0)):
        pos, st = _pts(pos)
        tr = np.asarray(trans, dtype=np.float32)
        ln = C.c_int64(0)
        _chk(lib().orc_insert_subcell(self.h, sc, _ptr(pos), len(pos), st, radius, _ptr(tr), C.byref(ln)),
             "insert_points_subcell")
        self.sc = sc
        return ln.value

    def _arr(self, fn, n, dtype, shape=None):
        if n == 0:
            return np.zeros(shape or (0,), dtype)
        p = getattr(lib(), fn)(self.h)
        a = np.frombuffer((C.c_char * (n * np.dtype(dtype).itemsize)).from_address(p), dtype=dtype).copy()
        return a.reshape(shape) if shape else a

    def subcells(self):
        nsc = lib().orc_sc_num(self.h)
        cnt = self._arr("orc_sc_cnt", nsc, np.int32)
        off = self._arr("orc_sc_off", nsc, np.int32)
        ln = int(cnt.sum())
        return dict(
            map=self._arr("orc_sc_map", self.node_count(0), np.int32),
            cnt=cnt, off=off,
            pos=self._arr("orc_sc_pos", nsc * 3, np.int32, (nsc, 3)),
            nid=self._arr("orc_sc_nid", nsc, np.int32),
            list=self._arr("orc_sc_list", ln, np.int64),
            list_pos=self._arr("orc_sc_list_pos", ln * 3, np.float32, (ln, 3)),
        )

    # --- channels
    def add_channel(self, chan):
        _chk(lib().orc_channel_alloc(self.h, chan), "channel_alloc")

    def fill_channel(self, chan, v):
        _chk(lib().orc_fill_channel(self.h, chan, v), "fill_channel")

    def channel(self, chan):
        """FP64 brick-linear atlas of shape (slots, br, br, br) indexed [slot, z, y, x]."""
        n = lib().orc_channel_size(self.h)
        br = self.res0 + 2 * self.apron
        if n == 0:
            return np.zeros((0, br, br, br))
        p = lib().orc_channel(self.h, chan)
        a = np.frombuffer((C.c_char * (n * 8)).from_address(p), dtype=np.float64).copy()
        return a.reshape(-1, br, br, br)

    def gather_density(self, radius, chan, accum=False):
        _chk(lib().orc_gather_density(self.h, radius, chan, int(accum)), "gather_density")

    def gather_levelset(self, radius, chan, accum=False):
        _chk(lib().orc_gather_levelset(self.h, radius, chan, int(accum)), "gather_levelset")

    def gather_velocity(self, vel, chans, accum=False):
        vel, st = _pts(vel)
        _chk(lib().orc_gather_velocity(self.h, _ptr(vel), st, chans[0], chans[1], chans[2], int(accum)),
             "gather_velocity")

    def gather_colour(self, clr, chan, accum=False):
        """Colour means of the gathers; clr: uint32 words (r | g << 8 | b << 16 | a << 24) per
        point, as listed at the insert.  The channel holds the words as doubles."""
        clr = np.ascontiguousarray(clr, dtype=np.uint32)
        _chk(lib().orc_gather_colour(self.h, _ptr(clr), 1, chan, int(accum)), "gather_colour")

    def scatter_colour(self, chan, pos, clr, avg=True, trans=(0, 0, 0)):
        """Colour part of ScatterDensity: averages (avg) or the largest-index point's colour per
        reached voxel; clr: uint32 words per point."""
        pos, st = _pts(pos)
        clr = np.ascontiguousarray(clr, dtype=np.uint32)
        tr = np.asarray(trans, np.float32)
        _chk(lib().orc_scatter_colour(self.h, chan, _ptr(pos), len(pos), st * 4, _ptr(tr), _ptr(clr), int(avg)),
             "scatter_colour")

    def quantize_list_pos(self, qmin, qrange):
        """The listed positions as FP16-path ushort3 (len x 3) over [qmin, qmin + qrange]."""
        ln = int(self.subcells()["cnt"].sum())
        out = np.zeros((ln, 3), np.uint16)
        lo, rg = np.asarray(qmin, np.float32), np.asarray(qrange, np.float32)
        _chk(lib().orc_quantize_list_pos(self.h, _ptr(lo), _ptr(rg), _ptr(out)), "quantize_list_pos")
        return out

    def quantize_list_vel(self, vel, vmin, vrange):
        ln = int(self.subcells()["cnt"].sum())
        out = np.zeros((ln, 3), np.uint16)
        vel, st = _pts(vel)
        lo, rg = np.asarray(vmin, np.float32), np.asarray(vrange, np.float32)
        _chk(lib().orc_quantize_list_vel(self.h, _ptr(vel), st, _ptr(lo), _ptr(rg), _ptr(out)), "quantize_list_vel")
        return out

    def gather_levelset_q16(self, radius, chan, q16, qmin, qrange):
        q16 = np.ascontiguousarray(q16, np.uint16)
        lo, rg = np.asarray(qmin, np.float32), np.asarray(qrange, np.float32)
        _chk(lib().orc_gather_levelset_q16(self.h, radius, chan, _ptr(q16), _ptr(lo), _ptr(rg)),
             "gather_levelset_q16")

    def update_apron(self, chan, boundval):
        _chk(lib().orc_update_apron(self.h, chan, boundval), "update_apron")

    def scatter_reduce_levelset(self, radius, chan, init=3.0):
        _chk(lib().orc_scatter_reduce_levelset(self.h, radius, chan, init), "scatter_reduce_levelset")

    def sample_trilinear(self, chan, q):
        q = np.ascontiguousarray(q, np.float32)
        out = np.zeros(len(q))
        _chk(lib().orc_sample_trilinear(self.h, chan, _ptr(q), len(q), _ptr(out)), "sample_trilinear")
        return out

    def splat_trilinear(self, chan, q, mass=None):
        q = np.ascontiguousarray(q, np.float32)
        m = None if mass is None else np.ascontiguousarray(mass, np.float32)
        _chk(lib().orc_splat_trilinear(self.h, chan, _ptr(q), _ptr(m), len(q)), "splat_trilinear")

    def scatter_density(self, chan, pos, radius, trans=(0, 0, 0), expand=True):
        pos = np.ascontiguousarray(pos, np.float32)
        tr = np.ascontiguousarray(trans, np.float32)
        _chk(lib().orc_scatter_density(self.h, chan, _ptr(pos), len(pos), 12, radius, _ptr(tr), int(expand)),
             "scatter_density")

    def splat_backward(self, chan_grad, q, mass=None):
        q = np.ascontiguousarray(q, np.float32)
        m = None if mass is None else np.ascontiguousarray(mass, np.float32)
        dq = np.zeros((len(q), 3))
        dm = np.zeros(len(q))
        _chk(lib().orc_splat_backward(self.h, chan_grad, _ptr(q), _ptr(m), len(q), _ptr(dq), _ptr(dm)),
             "splat_backward")
        return dq, dm

    def sample_backward(self, chan, q, dout):
        q = np.ascontiguousarray(q, np.float32)
        d = np.ascontiguousarray(dout, np.float64)
        dq = np.zeros((len(q), 3))
        _chk(lib().orc_sample_backward(self.h, chan, _ptr(q), _ptr(d), len(q), _ptr(dq)), "sample_backward")
        return dq


def cpu_activate_space(pos, log2dims=(3, 3, 3, 3, 5), trans=(0, 0, 0)):
    pos, st = _pts(pos)
    ld = np.asarray(log2dims, np.int32)
    tr = np.asarray(trans, np.float32)
    cnt = C.c_int64(0)
    r = lib().orc_cpu_activate_space(len(ld), _ptr(ld), _ptr(pos), len(pos), st, _ptr(tr), C.byref(cnt))
    return r
