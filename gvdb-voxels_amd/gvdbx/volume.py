"""VolumeGVDB: the reference's point-path methods (V.h:321-558) over the gvdbx C ABI.

Method names, argument order and meaning follow the reference VolumeGVDB
(source/gvdb_library/src/gvdb_volume_gvdb.h).  Failures raise GvdbxError where the
reference would print and exit (cudaCheck, A.cpp:1006-1034) or print and return
(RebuildTopology depth checks, V.cpp:1187-1194).
"""
import ctypes as C

import numpy as np

from .lib import NODE_DTYPE, DeviceArray, GvdbxError, fptr, load_library

T_UCHAR4 = 2  # src/gvdb_types.h:75-76
T_FLOAT = 3


class VolumeGVDB:
    def __init__(self, device=0, stream=None):
        self.L = load_library()
        h = C.c_void_p()
        code = self.L.gvdbx_ctx_create(device, stream, C.byref(h))
        if code != 0:
            raise GvdbxError(code, "gvdbx_ctx_create")
        self.h = h.value
        self.log2dims = None
        self.apron = None
        self._points = None
        self._channels = []
        self._types = {}

    # ------------------------------------------------------------ plumbing
    def _chk(self, code, what):
        if code != 0:
            buf = C.create_string_buffer(512)
            if self.h:
                self.L.gvdbx_last_error(self.h, buf, 512)
            raise GvdbxError(code, what, buf.value.decode(errors="replace"))

    def close(self):
        if getattr(self, "h", None):
            self.L.gvdbx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self._chk(self.L.gvdbx_synchronize(self.h), "synchronize")

    def alloc(self, nbytes):
        return DeviceArray(self, nbytes)

    def to_device(self, arr):
        return DeviceArray.from_numpy(self, arr)

    # ------------------------------------------------------------ configuration
    def Configure(self, q4, q3, q2, q1, q0):
        """Configure(q4..q0) as in V.cpp:2364: the last argument is the leaf level."""
        self.configure_levels([q0, q1, q2, q3, q4])

    def configure_levels(self, log2dims):
        ld = np.ascontiguousarray(log2dims, dtype=np.int32)
        self._chk(self.L.gvdbx_configure(self.h, len(ld), ld.ctypes.data), "configure")
        self.log2dims = tuple(int(max(x, 1)) for x in ld)

    def SetChannelDefault(self, cx, cy, cz):
        self._chk(self.L.gvdbx_set_channel_default(self.h, cx, cy, cz), "set_channel_default")

    def AddChannel(self, chan, dtype=T_FLOAT, apron=1):
        self._chk(self.L.gvdbx_add_channel(self.h, chan, dtype, apron), "add_channel")
        self.apron = apron
        if chan not in self._channels:
            self._channels.append(chan)
        self._types[chan] = dtype

    def FillChannel(self, chan, value):
        """FillChannel(chan, Vector4DF): value is a float (x) or an (x, y, z, w) sequence; a uchar4
        channel takes CLR2INT(x, y, z, 1), all zeros clear it."""
        v = [float(x) for x in value] if hasattr(value, "__len__") else [float(value), 0.0, 0.0, 0.0]
        v += [0.0] * (4 - len(v))
        self._chk(self.L.gvdbx_fill_channel4(self.h, chan, fptr(v[:4])), "fill_channel")

    def ClearChannel(self, chan):
        self.FillChannel(chan, 0.0)

    # ------------------------------------------------------------ points
    def SetPoints(self, pos, vel=None, clr=None, n=None, pos_off=0, pos_stride=12, vel_off=0, vel_stride=12,
                  clr_off=0, clr_stride=4):
        """SetPoints + SetDataGPU: device buffers (DeviceArray or raw pointer ints)."""
        ptr = lambda a: None if a is None else (a.ptr if isinstance(a, DeviceArray) else int(a))
        if n is None:
            n = pos.nbytes // pos_stride
        self._points = (pos, vel, clr)
        self.npts = n
        self._chk(self.L.gvdbx_set_points(self.h, n, ptr(pos), pos_off, pos_stride, ptr(vel), vel_off, vel_stride,
                                          ptr(clr), clr_off, clr_stride), "set_points")

    def AdvancePoints(self, n, dt):
        """q += v * dt over the first n bound points, in place on the device (FP32 product, then
        FP32 sum, as numpy's float32 arithmetic)."""
        self._chk(self.L.gvdbx_advance_points(self.h, n, float(dt)), "advance_points")

    # ------------------------------------------------------------ batched clouds
    def BatchSetPoints(self, pos, counts, vel=None, cell_log2=12, max_radius=1.0, pos_off=0, pos_stride=12,
                       vel_off=0, vel_stride=12):
        """Bind several clouds stored back to back in `pos` (counts[k] points each) as one batch:
        cloud k runs in its own grid cell (see batch_origin) through the usual calls."""
        ptr = lambda a: None if a is None else (a.ptr if isinstance(a, DeviceArray) else int(a))
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        self._points = (pos, vel, None)
        self.npts = int(cnt.sum())
        self._chk(self.L.gvdbx_batch_set_points(self.h, len(cnt), cnt.ctypes.data, ptr(pos), pos_off, pos_stride,
                                                ptr(vel), vel_off, vel_stride, cell_log2, float(max_radius)),
                  "batch_set_points")

    def batch_layout(self):
        ncloud, cell = C.c_int(0), C.c_int(0)
        grid = np.zeros(3, np.int32)
        self._chk(self.L.gvdbx_batch_get_layout(self.h, C.byref(ncloud), C.byref(cell), grid.ctypes.data),
                  "batch_get_layout")
        return ncloud.value, cell.value, grid

    def batch_origin(self, k):
        """Voxel origin of cloud k's cell."""
        _, cell, g = self.batch_layout()
        return np.array([k % g[0], (k // g[0]) % g[1], k // (g[0] * g[1])], np.int64) << cell

    # ------------------------------------------------------------ topology
    def RebuildTopology(self, n, radius, orig=(0, 0, 0)):
        self._chk(self.L.gvdbx_rebuild_topology(self.h, n, radius, fptr(orig)), "RebuildTopology")

    def AccumulateTopology(self, n, radius, orig=(0, 0, 0), min_depth=0):
        self._chk(self.L.gvdbx_accumulate_topology(self.h, n, radius, fptr(orig), min_depth), "AccumulateTopology")

    def RequestFullRebuild(self, flag=True):
        self._chk(self.L.gvdbx_request_full_rebuild(self.h, int(bool(flag))), "RequestFullRebuild")

    def FinishTopology(self):
        bmin = np.zeros(3, np.float32)
        bmax = np.zeros(3, np.float32)
        fp = C.POINTER(C.c_float)
        self._chk(self.L.gvdbx_finish_topology(self.h, bmin.ctypes.data_as(fp), bmax.ctypes.data_as(fp)),
                  "FinishTopology")
        return bmin, bmax

    def UpdateAtlas(self):
        self._chk(self.L.gvdbx_update_atlas(self.h), "UpdateAtlas")

    # ------------------------------------------------------------ subcells / gathers
    def InsertPointsSubcell(self, subcell, n, radius, trans=(0, 0, 0)):
        ln = C.c_int64(0)
        self._chk(self.L.gvdbx_insert_points_subcell(self.h, subcell, n, radius, fptr(trans), C.byref(ln)),
                  "InsertPointsSubcell")
        return ln.value

    def GatherDensity(self, subcell, n, radius, trans, sc_len, chan, chan_clr=-1, accum=False):
        self._chk(self.L.gvdbx_gather_density(self.h, subcell, n, radius, fptr(trans), sc_len, chan, chan_clr,
                                              int(bool(accum))), "GatherDensity")

    def GatherLevelSet(self, subcell, n, radius, trans, sc_len, chan, chan_clr=-1, accum=False):
        self._chk(self.L.gvdbx_gather_levelset(self.h, subcell, n, radius, fptr(trans), sc_len, chan, chan_clr,
                                               int(bool(accum))), "GatherLevelSet")

    def InsertPointsSubcell_FP16(self, subcell, n, radius, trans=(0, 0, 0)):
        """InsertPointsSubcell_FP16 (V.cpp:5094-5189): list positions as ushort3 over the rebuild's box."""
        ln = C.c_int64(0)
        self._chk(self.L.gvdbx_insert_points_subcell_fp16(self.h, subcell, n, radius, fptr(trans), C.byref(ln)),
                  "InsertPointsSubcell_FP16")
        return ln.value

    def GatherLevelSet_FP16(self, subcell, n, radius, trans, sc_len, chan, chan_clr=-1):
        self._chk(self.L.gvdbx_gather_levelset_fp16(self.h, subcell, n, radius, fptr(trans), sc_len, chan, chan_clr),
                  "GatherLevelSet_FP16")

    def GetMinMaxVel(self, n):
        """GetMinMaxVel (V.cpp:5470-5522): (min xyz, max xyz) of the bound velocities."""
        out = np.zeros(6, np.float32)
        self._chk(self.L.gvdbx_get_min_max_vel(self.h, n, out.ctypes.data_as(C.POINTER(C.c_float))),
                  "GetMinMaxVel")
        return out[:3].copy(), out[3:].copy()

    def GatherVelocity(self, chan_x, chan_y, chan_z, accum=False):
        self._chk(self.L.gvdbx_gather_velocity(self.h, chan_x, chan_y, chan_z, int(bool(accum))), "GatherVelocity")

    def SetSubcellOptions(self, store_velocity_lists=False):
        self._chk(self.L.gvdbx_set_subcell_options(self.h, int(bool(store_velocity_lists))), "set_subcell_options")

    def ScatterReduceLevelSet(self, radius, chan, init=3.0):
        self._chk(self.L.gvdbx_scatter_reduce_levelset(self.h, radius, chan, init), "ScatterReduceLevelSet")

    def UpdateApron(self, chan=None, boundval=0.0):
        """UpdateApron(chan, boundval); UpdateApron() covers every channel (V.cpp:4411-4415)."""
        if chan is None:
            self.UpdateApronChannels(self._channels, [boundval] * len(self._channels))
            return
        self._chk(self.L.gvdbx_update_apron(self.h, chan, float(boundval)), "UpdateApron")

    def UpdateApronChannels(self, chans, boundvals):
        ch = np.ascontiguousarray(chans, np.int32)
        bv = np.ascontiguousarray(boundvals, np.float32)
        assert len(ch) == len(bv)
        self._chk(self.L.gvdbx_update_apron_channels(self.h, ch.ctypes.data, bv.ctypes.data, len(ch)),
                  "UpdateApronChannels")

    def SampleTrilinear(self, chan, q_dev, n, out_dev):
        self._chk(self.L.gvdbx_sample_trilinear(self.h, chan, q_dev.ptr, n, out_dev.ptr), "SampleTrilinear")

    # ------------------------------------------------------------ splat and adjoints
    @staticmethod
    def _ptr(d):
        return None if d is None else d.ptr

    def GatherDensityVelocity(self, subcell, n, radius, trans, sc_len, chan, chans_vel):
        """GatherDensity into chan and the velocity means (GatherVelocity, no accumulation) into
        chans_vel in one pass over the bricks."""
        cx, cy, cz = chans_vel
        self._chk(self.L.gvdbx_gather_density_velocity(self.h, subcell, n, radius, fptr(trans), sc_len, chan, cx, cy, cz),
                  "GatherDensityVelocity")

    def ScatterDensity(self, num_pnts, radius, amp, trans=(0, 0, 0), expand=True, avgColor=False, chan=0,
                       chan_clr=None):
        """ScatterDensity (V.cpp:5625-5659): (1 - c)^3 of every bound point into its voxel (and
        its 6 face neighbours with expand) of chan; exact sums instead of the reference's race.
        With colours bound, the uchar4 channel chan_clr (default: the reference's channel, 1 with
        avgColor, else 3) takes the voxels' average colours (avgColor) or the colour of each
        voxel's highest-index point."""
        if chan_clr is None:
            chan_clr = (1 if avgColor else 3) if (self._points and self._points[2] is not None) else -1
        self._chk(self.L.gvdbx_scatter_density(self.h, num_pnts, radius, amp, fptr(trans), int(bool(expand)), chan,
                                               int(bool(avgColor)), chan_clr), "ScatterDensity")

    def SplatTrilinear(self, chan, q_dev, n, mass_dev=None, accum=False):
        """Trilinear mass splat (transpose of SampleTrilinear) into chan's interior."""
        self._chk(self.L.gvdbx_splat_trilinear(self.h, chan, q_dev.ptr, self._ptr(mass_dev), n, int(accum)),
                  "SplatTrilinear")

    def SplatTrilinearNoAlloc(self, chan, q_dev, n, mass_dev=None, accum=False):
        """SplatTrilinear with its scratch in the library's fixed per-GPU array (no allocation;
        at most 1,048,576 points per call)."""
        self._chk(self.L.gvdbx_splat_trilinear_noalloc(self.h, chan, q_dev.ptr, self._ptr(mass_dev), n, int(accum)),
                  "SplatTrilinearNoAlloc")

    def SplatBackward(self, chan_grad, q_dev, n, mass_dev=None, dq_dev=None, dmass_dev=None):
        """dL/dq and dL/dmass of the splat from dL/drho held in chan_grad."""
        self._chk(self.L.gvdbx_splat_backward(self.h, chan_grad, q_dev.ptr, self._ptr(mass_dev), n, self._ptr(dq_dev),
                                              self._ptr(dmass_dev)), "SplatBackward")

    def SampleTrilinearBackward(self, chan, q_dev, dout_dev, n, dq_dev=None, chan_grad=-1, accum_grad=False):
        """dL/dq of SampleTrilinear and (chan_grad >= 0) dL/dfield = splat of dout."""
        self._chk(self.L.gvdbx_sample_trilinear_backward(self.h, chan, q_dev.ptr, dout_dev.ptr, n, self._ptr(dq_dev),
                                                         chan_grad, int(accum_grad)), "SampleTrilinearBackward")

    def SaveVBX(self, path):
        """Write the pools and channels as a VBX file (SaveVBX, V.cpp:1626-1768)."""
        self._chk(self.L.gvdbx_save_vbx(self.h, str(path).encode()), "SaveVBX")

    def LoadVBX(self, path):
        """Replace the tree and channels by a VBX file's first grid (LoadVBX, V.cpp:507-685)."""
        from .lib import VbxDesc
        d = VbxDesc()
        self._chk(self.L.gvdbx_vbx_probe(str(path).encode(), C.byref(d)), "LoadVBX")
        self._chk(self.L.gvdbx_load_vbx(self.h, str(path).encode()), "LoadVBX")
        self.log2dims = tuple(int(d.log2dim[l]) for l in range(d.levels))
        self.apron = int(d.apron)
        self._channels = list(range(d.nchan))
        self._types = {c: int(d.chan_type[c]) for c in range(d.nchan)}

    def set_channel(self, chan, data):
        """Overwrite a channel (float[slots][br][br][br]) from host memory."""
        ptr, floats = C.c_void_p(), C.c_int64(0)
        self._chk(self.L.gvdbx_get_channel_ptr(self.h, chan, C.byref(ptr), C.byref(floats)), "get_channel_ptr")
        a = np.ascontiguousarray(data, np.float32).reshape(-1)
        if a.size > floats.value:
            raise ValueError(f"channel holds {floats.value} floats, got {a.size}")
        self._chk(self.L.gvdbx_commit(self.h, ptr.value, a.ctypes.data, a.nbytes), "commit")

    # ------------------------------------------------------------ introspection
    def depth(self):
        d = C.c_int(0)
        root = np.zeros(3, np.int32)
        pmin = np.zeros(3, np.float32)
        pmax = np.zeros(3, np.float32)
        self._chk(self.L.gvdbx_get_depth(self.h, C.byref(d), root.ctypes.data, pmin.ctypes.data, pmax.ctypes.data),
                  "get_depth")
        return d.value, root, pmin, pmax

    def node_count(self, lev):
        t = C.c_int64(0)
        u = C.c_int64(0)
        self._chk(self.L.gvdbx_get_node_count(self.h, lev, C.byref(t), C.byref(u)), "get_node_count")
        return t.value, u.value

    def getNumTotalNodes(self, lev):
        return self.node_count(lev)[0]

    def getNumUsedNodes(self, lev):
        return self.node_count(lev)[1]

    def get_nodes(self, lev):
        n = self.node_count(lev)[0]
        out = np.zeros(n, NODE_DTYPE)
        if n:
            self._chk(self.L.gvdbx_copy_nodes(self.h, lev, out.ctypes.data, n), "copy_nodes")
        return out

    def get_childlists(self, lev):
        n = self.node_count(lev)[0]
        per = 1 << (3 * self.log2dims[lev])
        out = np.zeros((n, per), np.uint64)
        if n:
            self._chk(self.L.gvdbx_copy_childlists(self.h, lev, out.ctypes.data, n * per), "copy_childlists")
        return out

    def num_slots(self):
        s = C.c_int64(0)
        br = C.c_int(0)
        self._chk(self.L.gvdbx_get_num_slots(self.h, C.byref(s), C.byref(br)), "get_num_slots")
        return s.value, br.value

    def get_channel(self, chan):
        """Brick-linear channel as (slots, br, br, br) indexed [slot, z, y, x]: float32, or for a
        uchar4 channel uint32 words (r | g << 8 | b << 16 | a << 24)."""
        s, br = self.num_slots()
        out = np.zeros(s * br ** 3, np.float32)
        if out.size:
            self._chk(self.L.gvdbx_copy_channel(self.h, chan, out.ctypes.data, out.size), "copy_channel")
        if self._types.get(chan) == T_UCHAR4:
            out = out.view(np.uint32)
        return out.reshape(s, br, br, br)

    def channel_ptr(self, chan):
        p = C.c_void_p()
        n = C.c_int64(0)
        self._chk(self.L.gvdbx_get_channel_ptr(self.h, chan, C.byref(p), C.byref(n)), "get_channel_ptr")
        return p.value, n.value

    def subcell_info(self):
        nsc = C.c_int64(0)
        ln = C.c_int64(0)
        sc = C.c_int(0)
        self._chk(self.L.gvdbx_get_subcell_info(self.h, C.byref(nsc), C.byref(ln), C.byref(sc)), "subcell_info")
        return nsc.value, ln.value, sc.value

    def get_subcells(self):
        nsc, ln, sc = self.subcell_info()
        total = self.node_count(0)[0]

        def grab(which, n, dtype, shape=None):
            a = np.zeros(n, dtype)
            if a.nbytes:
                self._chk(self.L.gvdbx_copy_subcell_array(self.h, which, a.ctypes.data, a.nbytes), "copy_subcell")
            return a.reshape(shape) if shape else a

        return dict(
            map=grab(0, total, np.int32), cnt=grab(1, nsc, np.int32), off=grab(2, nsc, np.int32),
            pos=grab(3, nsc * 3, np.int32, (nsc, 3)), nid=grab(4, nsc, np.int32),
            list_pos=grab(5, ln * 3, np.float32, (ln, 3)), list_vel=grab(6, ln * 3, np.float32, (ln, 3)),
            list_clr=grab(7, ln, np.uint32),
            list_q16=grab(8, ln * 3, np.uint16, (ln, 3)), list_vq16=grab(9, ln * 3, np.uint16, (ln, 3)),
        )

    # ------------------------------------------------------------ launch shape and graphs
    def SetLaunch(self, block_threads=0, persistent=False):
        """Threads per brick workgroup of the count, insert and density kernels (0 = chosen from
        the items per brick) and the persistent brick-queue form."""
        self._chk(self.L.gvdbx_set_launch(self.h, int(block_threads), int(bool(persistent))), "set_launch")

    def GraphBegin(self):
        """Record the following insert / gather / apron calls into one HIP graph (they run at
        GraphLaunch); run the same calls once before recording."""
        self._chk(self.L.gvdbx_graph_begin(self.h), "graph_begin")

    def GraphEnd(self):
        self._chk(self.L.gvdbx_graph_end(self.h), "graph_end")

    def GraphLaunch(self):
        self._chk(self.L.gvdbx_graph_launch(self.h), "graph_launch")

    def GraphRelease(self):
        self._chk(self.L.gvdbx_graph_release(self.h), "graph_release")

    # ------------------------------------------------------------ timing
    def kernel_timer(self, name, enable=True):
        self._chk(self.L.gvdbx_kernel_timer(self.h, name.encode() if name else None, int(enable)), "kernel_timer")

    def kernel_timer_read(self):
        ms = C.c_double(0)
        n = C.c_int64(0)
        self._chk(self.L.gvdbx_kernel_timer_read(self.h, C.byref(ms), C.byref(n)), "kernel_timer_read")
        return ms.value, n.value
