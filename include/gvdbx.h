/*
 * gvdbx.h -- C ABI of the MI355X (gfx950) point-to-sparse-grid library.
 *
 * Drop-in replacement for the point path of NVIDIA GVDB Voxels' VolumeGVDB
 * (V.cpp = source/gvdb_library/src/gvdb_volume_gvdb.cpp, V.h = .../gvdb_volume_gvdb.h).
 * Each entry point names the reference method it replaces.  The C++ facade in
 * gvdb-voxels_amd/gvdbx/volume_gvdb.hpp keeps the reference's method names and
 * signatures on top of these calls.
 *
 * ABI rules
 *   - every function returns int status (GVDBX_OK = 0); none exits or throws;
 *   - point buffers are caller-owned device pointers (SetDataGPU semantics, V.cpp:4872-4880);
 *     topology pools, atlas channels and subcell lists are library-owned;
 *   - work is enqueued on the context's HIP stream; functions that must return a
 *     count to the host (the reference's int& outputs and DtoH copies) synchronise that
 *     stream, as the reference does;
 *   - one context per host thread; calls on one context must not overlap.
 *
 * Topology layout (bit-identical to the reference pools):
 *   node pool of level l: gvdbx_node[count_l] (64 B, src/gvdb_node.h:27-40);
 *   child-list pool of level l >= 1: uint64[count_l][res_l^3], entries are
 *   Elem(0, l-1, index) = (index << 16) | ((l-1) << 8) or 0xFFFFFFFFFFFFFFFF (A.h:59-62).
 * Atlas layout (replaces the reference 3D CUarray atlas, A.cpp:528-560):
 *   channel c: float[slots][br][br][br], br = res0 + 2*apron, brick-linear, slot s is
 *   the s-th flagged leaf in pool order; leaf.value holds the reference atlas corner of s.
 */
#ifndef GVDBX_H
#define GVDBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GVDBX_ABI_VERSION 1

enum gvdbx_status {
  GVDBX_OK = 0,
  GVDBX_ERR_ARG = 1,          /* bad argument / unconfigured context */
  GVDBX_ERR_DEPTH0 = 2,       /* "Points extents smaller than 1 brick" (V.cpp:1187-1190) */
  GVDBX_ERR_DEPTH_MAX = 3,    /* "Point extents exceed 5 levels" (V.cpp:1191-1194) */
  GVDBX_ERR_CAPACITY = 4,     /* a size exceeds a 32-bit index or a fixed limit */
  GVDBX_ERR_HIP = 5,          /* a HIP runtime call failed */
  GVDBX_ERR_STATE = 6,        /* call order violated (e.g. gather before insert) */
  GVDBX_ERR_NOMEM = 7,        /* device allocation failed */
  GVDBX_ERR_DEGENERATE = 8    /* the whole topology is one brick at level 0 */
};

/* channel types: T_UCHAR4 and T_FLOAT (src/gvdb_types.h:75-76); both store one 32-bit word per voxel */
enum gvdbx_dtype { GVDBX_T_UCHAR4 = 2, GVDBX_T_FLOAT = 3 };

/* Node, same field order and size as the reference Node (src/gvdb_node.h:27-40). */
typedef struct gvdbx_node {
  uint8_t lev, flags, prio, pad;
  int32_t pos[3];
  int32_t value[3];
  float vrange[3];
  uint64_t parent;
  uint64_t childlist;
  uint64_t mask;
} gvdbx_node;

typedef struct gvdbx_ctx gvdbx_ctx;

/* ---- context: SetCudaDevice + Initialize (V.cpp:240-334, 2320-2361) */
int gvdbx_ctx_create(int device, void* hip_stream, gvdbx_ctx** out);
int gvdbx_ctx_destroy(gvdbx_ctx* ctx);
const char* gvdbx_status_string(int status);
int gvdbx_last_error(const gvdbx_ctx* ctx, char* buf, int buflen);
int gvdbx_synchronize(gvdbx_ctx* ctx);
int gvdbx_get_stream(const gvdbx_ctx* ctx, void** hip_stream);

/* ---- configuration: Configure (V.cpp:2364-2434), SetChannelDefault (V.h:390),
 *      AddChannel (V.cpp:2437-2453), FillChannel/ClearChannel (V.cpp:2456-2471, 1596-1603) */
int gvdbx_configure(gvdbx_ctx* ctx, int levs, const int* log2dims);
int gvdbx_set_channel_default(gvdbx_ctx* ctx, int cx, int cy, int cz);
int gvdbx_add_channel(gvdbx_ctx* ctx, int chan, int dtype, int apron);
int gvdbx_fill_channel(gvdbx_ctx* ctx, int chan, float value);   /* = gvdbx_fill_channel4 with (value, 0, 0, 0) */
/* FillChannel(chan, Vector4DF) (V.cpp:2456-2471): all-zero values clear the channel; a float
 * channel takes val[0]; a uchar4 channel takes CLR2INT(val[0], val[1], val[2], 1)
 * (gvdbOpFillC4, O.cuh:368-373), each component uchar(x * 255) saturated to [0, 255]. */
int gvdbx_fill_channel4(gvdbx_ctx* ctx, int chan, const float val[4]);

/* ---- points: SetPoints + SetDataGPU (V.cpp:4900-4908, 4872-4880); byte offsets and strides */
int gvdbx_set_points(gvdbx_ctx* ctx, int64_t n,
                     const void* pos, int pos_off, int pos_stride,
                     const void* vel, int vel_off, int vel_stride,
                     const void* clr, int clr_off, int clr_stride);

/* q += v * dt for the first n bound points, in place in the bound position buffer, on the
 * context's stream: an FP32 product, then an FP32 sum (no fused multiply-add), so the result
 * equals the same FP32 arithmetic on the host.  No reference counterpart: it keeps the SPH
 * step's displacement (the gFluidSurface frame moves its particles before RebuildTopology,
 * main_fluid_surface.cpp:320-356) on the device.  Needs bound velocities; not for batched clouds. */
int gvdbx_advance_points(gvdbx_ctx* ctx, int64_t n, float dt);

/* ---- batched clouds (no reference counterpart): ncloud independent clouds, each in its own
 *      frame, run through one RebuildTopology / UpdateAtlas / InsertPointsSubcell / Gather /
 *      UpdateApron sequence.  pos holds the clouds back to back (counts[k] points each, host
 *      array); cloud k is placed in grid cell k of 2^cell_log2 voxels, cell origin
 *      (k % g0, k / g0 % g1, k / (g0 g1)) << cell_log2 (gvdbx_batch_get_layout), where its
 *      bricks, subcells and channel values are those of the cloud run alone, shifted by the
 *      origin (positions stay in the cloud's frame, so every value is bit-identical).
 *      cell_log2 must be the range of an inner level; every valid point must stay below
 *      2^cell_log2 - res0 - 2 max_radius - 3 in each coordinate (GVDBX_ERR_CAPACITY otherwise),
 *      and later radii must not exceed max_radius (at most 1, which keeps every extra-pass box
 *      inside the cloud's own root).  The batch keeps one tree depth for all clouds: a cloud
 *      whose own depth would change between frames is not restarted.  gvdbx_set_points ends
 *      batched mode. */
int gvdbx_batch_set_points(gvdbx_ctx* ctx, int ncloud, const int64_t* counts, const void* pos, int pos_off,
                           int pos_stride, const void* vel, int vel_off, int vel_stride, int cell_log2,
                           float max_radius);
int gvdbx_batch_get_layout(const gvdbx_ctx* ctx, int* ncloud, int* cell_log2, int grid[3]);

/* ---- topology: RebuildTopology (V.cpp:1181-1218), AccumulateTopology (V.cpp:1158-1178),
 *      RequestFullRebuild (V.h:432), FinishTopology (V.cpp:1579-1593), UpdateAtlas (V.cpp:2630-2704) */
int gvdbx_rebuild_topology(gvdbx_ctx* ctx, int64_t n, float radius, const float orig[3]);
int gvdbx_accumulate_topology(gvdbx_ctx* ctx, int64_t n, float radius, const float orig[3], int min_depth);
int gvdbx_request_full_rebuild(gvdbx_ctx* ctx, int flag);
int gvdbx_finish_topology(gvdbx_ctx* ctx, float bmin[3], float bmax[3]);
int gvdbx_update_atlas(gvdbx_ctx* ctx);

/* ---- subcells and gathers: InsertPointsSubcell (V.cpp:4971-5092),
 *      GatherDensity (V.cpp:5318-5344), GatherLevelSet (V.cpp:5348-5373) */
int gvdbx_insert_points_subcell(gvdbx_ctx* ctx, int subcell, int64_t n, float radius, const float trans[3],
                                int64_t* sc_pnts_len);
/* Colour (P.cuh:713-731): when the points had colours bound (gvdbx_set_points clr) at the
 * insert and chan_clr >= 0, the uchar4 channel chan_clr receives per voxel, component by
 * component, (start + sum of the listed points' colours) / (cnt + accum) over the voxel's
 * subcell, start = the old value with accum, else (0, 0, 0, 255) (the reference's initial
 * colour (0, 0, 0, 1)).  Computed exactly: floor of the integer quotient, saturated at 255; an
 * empty subcell without accumulation gives 0 (the reference divides by zero there).  Without
 * accumulation the apron is written from the neighbour bricks (missing ones: 0). */
int gvdbx_gather_density(gvdbx_ctx* ctx, int subcell, int64_t n, float radius, const float trans[3],
                         int64_t sc_pnts_len, int chan, int chan_clr, int accum);
/* colour: as gvdbx_gather_density (the same rule, P.cuh:752-770) */
int gvdbx_gather_levelset(gvdbx_ctx* ctx, int subcell, int64_t n, float radius, const float trans[3],
                          int64_t sc_pnts_len, int chan, int chan_clr, int accum);
/* ---- FP16 subcell lists: InsertPointsSubcell_FP16 (V.cpp:5094-5189, kernel gvdbInsertSubcell_fp16
 *      P.cuh:154-216) stores each listed position as ushort3, ushort((q - posmin) / posrange *
 *      65535) over the last rebuild's bounding box (gvdbx_get_depth posmin / posmax; IEEE single
 *      operations in that order, saturated to [0, 65535]), and, when velocity lists are kept
 *      (gvdbx_set_subcell_options), each velocity the same way over GetMinMaxVel's box;
 *      counts, offsets and colours as gvdbx_insert_points_subcell.  Only
 *      gvdbx_gather_levelset_fp16 reads these lists (the float gathers refuse them);
 *      it decodes p = u / 65535 * posrange + posmin (unfused) and writes min(3, |p - v| - r)
 *      (P.cuh:775-811).  chan_clr must be < 0: the reference's FP16 colour accumulator is
 *      uninitialised.  GetMinMaxVel (V.cpp:5470-5522): out = min xyz, max xyz of the first n
 *      velocities. */
int gvdbx_insert_points_subcell_fp16(gvdbx_ctx* ctx, int subcell, int64_t n, float radius, const float trans[3],
                                     int64_t* sc_pnts_len);
int gvdbx_gather_levelset_fp16(gvdbx_ctx* ctx, int subcell, int64_t n, float radius, const float trans[3],
                               int64_t sc_pnts_len, int chan, int chan_clr);
int gvdbx_get_min_max_vel(gvdbx_ctx* ctx, int64_t n, float out[6]);
/* Velocity channels: the reference colour-mean rule (P.cuh:717-731) applied to the binned
 * float3 velocities (stored by gvdbInsertSubcell, P.cuh:266-267, read by no reference gather).
 * Sums are exact 64-bit fixed point (2^-24 units), so the result is order independent. */
int gvdbx_gather_velocity(gvdbx_ctx* ctx, int chan_x, int chan_y, int chan_z, int accum);
/* Whether InsertPointsSubcell also stores the binned velocities (AUX_SUBCELL_PNT_VEL,
 * P.cuh:266-267).  Off by default: no gather reads them; the velocity means come from the
 * per-subcell sums accumulated during the insert. */
int gvdbx_set_subcell_options(gvdbx_ctx* ctx, int store_velocity_lists);
/* ScatterReduceLevelSet: absent from the reference; defined as the GatherLevelSet field
 * (P.cuh:736-773) produced by scattering each (point, subcell) pair and min-reducing. */
int gvdbx_scatter_reduce_levelset(gvdbx_ctx* ctx, float radius, int chan, float init);

/* ---- apron: UpdateApron (V.cpp:4411-4457, kernel O.cuh:72-126); the channel-list form
 *      covers the all-channel UpdateApron() overload (V.cpp:4411-4415) in one launch. */
int gvdbx_update_apron(gvdbx_ctx* ctx, int chan, float boundval);
int gvdbx_update_apron_channels(gvdbx_ctx* ctx, const int* chans, const float* boundvals, int nchan);

/* ---- trilinear sampling: getTrilinear/getVolSampleF (R.cuh:104-126), voxel v at v+0.5,
 *      0 outside active bricks.  q: device float3[n]; out: device float[n]. */
int gvdbx_sample_trilinear(gvdbx_ctx* ctx, int chan, const void* q, int64_t n, void* out);

/* ---- GatherDensity with velocity channels bound (no reference counterpart as such: the
 *      float3 analogue of GatherDensity's colour channel, which the reference computes in the
 *      same kernel, P.cuh:717-731): chan as gvdbx_gather_density, (cx, cy, cz) as
 *      gvdbx_gather_velocity without accumulation, in one pass over the bricks.  The four
 *      channels must differ. */
int gvdbx_gather_density_velocity(gvdbx_ctx* ctx, int sc, int64_t n, float r, const float trans[3], int64_t len,
                                  int chan, int cx, int cy, int cz);

/* ---- ScatterDensity (V.h:555, V.cpp:5625-5659; kernel gvdbScatterPointDensity P.cuh:516-573):
 *      for every bound point with q = p + trans in a leaf with corner b, p' = q - b,
 *      pi = trunc(p'): voxel pi (and, with expand, its 6 face neighbours, which may lie in the
 *      brick's apron) of chan gains (1 - c)^3, c = |v - p'|^2 / radius^2.  amp is unused, as in
 *      the reference.  The reference's racy read-modify-write is replaced by exact 64-bit
 *      fixed-point sums (deterministic); inputs whose worst-case voxel sum would exceed 2^40
 *      (tiny radii at many points) return GVDBX_ERR_CAPACITY.  Not for batched clouds.
 *      Colour (P.cuh:553-572), when the points have colours bound and chan_clr >= 0 (a uchar4
 *      channel; the reference uses channel 1 with avgColor, else 3): avg_colour != 0 writes
 *      each voxel pi reached by points (r, g, b sums / count, 255) (gvdbScatterPointAvgCol,
 *      P.cuh:610-641); avg_colour == 0 writes the colour of one point of the voxel -- the
 *      reference's unordered last write, defined here as the point with the largest index.
 *      Voxels no point reached keep their word. */
int gvdbx_scatter_density(gvdbx_ctx* ctx, int64_t n, float radius, float amp, const float trans[3], int expand,
                          int chan, int avg_colour, int chan_clr);

/* ---- trilinear splat and adjoints (no reference counterpart: the transpose of the sampler
 *      above, R.cuh:104-126).  splat: chan interior (+)= sum_i m_i w_g(q_i) over the 8 voxels
 *      g of floor(q - 0.5) that lie in active bricks, points in no brick skipped; mass NULL =
 *      unit masses.  Deterministic: points sorted by (brick, region), exact 64-bit fixed-point
 *      sums per brick.  The apron is not written (UpdateApron afterwards). */
int gvdbx_splat_trilinear(gvdbx_ctx* ctx, int chan, const void* q, const void* mass, int64_t n, int accum);
/* gvdbx_splat_trilinear without any device allocation.  No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The splat's sort scratch lives in a fixed array of about 36 MiB per
 * GPU, allocated when the library is loaded on that GPU and held until the process ends; so
 * n <= 1,048,576 points per call and at most 38,836 active bricks (GVDBX_ERR_CAPACITY beyond;
 * splat larger clouds in several calls with accum = 1).  Same result as gvdbx_splat_trilinear,
 * bit for bit.  All calls on one GPU share the array: calls that overlap on one GPU (any
 * context or stream) race. */
int gvdbx_splat_trilinear_noalloc(gvdbx_ctx* ctx, int chan, const void* q, const void* mass, int64_t n, int accum);
/* Backward of the splat given G = dL/drho in chan_grad's interior: dmass_i = sum_g w_g(q_i) G(g),
 * dq_i = m_i sum_g grad w_g(q_i) G(g).  dq: device float3[n] or NULL; dmass: float[n] or NULL. */
int gvdbx_splat_backward(gvdbx_ctx* ctx, int chan_grad, const void* q, const void* mass, int64_t n, void* dq,
                         void* dmass);
/* Backward of gvdbx_sample_trilinear given dout = dL/dsample: dq_i = dout_i * grad sample(q_i)
 * (same padded-brick filter, dq NULL to skip) and, if chan_grad >= 0, chan_grad interior (+)=
 * splat of dout (the field gradient; exact when the apron holds neighbour copies and a zero
 * boundary value). */
int gvdbx_sample_trilinear_backward(gvdbx_ctx* ctx, int chan, const void* q, const void* dout, int64_t n, void* dq,
                                    int chan_grad, int accum_grad);

/* ---- introspection (getNumTotalNodes/getNumUsedNodes/getNode V.h:637-641, getAux) */
int gvdbx_get_depth(const gvdbx_ctx* ctx, int* depth, int32_t rootpos[3], float posmin[3], float posmax[3]);
int gvdbx_get_node_count(const gvdbx_ctx* ctx, int lev, int64_t* total, int64_t* used);
int gvdbx_copy_nodes(gvdbx_ctx* ctx, int lev, gvdbx_node* host_out, int64_t max_nodes);
int gvdbx_copy_childlists(gvdbx_ctx* ctx, int lev, uint64_t* host_out, int64_t max_entries);
int gvdbx_get_num_slots(const gvdbx_ctx* ctx, int64_t* slots, int* brick_res);
int gvdbx_copy_channel(gvdbx_ctx* ctx, int chan, float* host_out, int64_t max_floats);
int gvdbx_get_channel_ptr(gvdbx_ctx* ctx, int chan, void** dev_ptr, int64_t* floats);
/* list_len after a graph launch (gvdbx_graph_launch) is read back from the device (a host wait). */
int gvdbx_get_subcell_info(gvdbx_ctx* ctx, int64_t* num_subcells, int64_t* list_len, int* subcell);
/* which: 0 map(int32 per leaf) 1 cnt 2 off 3 pos(int32x3) 4 nid 5 list pos(float x3) 6 list vel(float x3)
 *        7 list colour (uint32 = uchar4) 8 FP16 list pos (uint16 x3) 9 FP16 list vel (uint16 x3) */
int gvdbx_copy_subcell_array(gvdbx_ctx* ctx, int which, void* host_out, int64_t max_bytes);

/* ---- VBX files (GVDB_FILESPEC.txt): SaveVBX (V.cpp:1626-1768), LoadVBX (V.cpp:507-685).
 *      gvdbx_save_vbx writes the context's pools and channels as the single-grid file SaveVBX
 *      writes (version 1.11, GVDB topology, atlas layout: channel c's bricks embedded in one
 *      (ax br) x (ay br) x (az br) volume, brick s at (s % ax, s / ax % ay, s / (ax ay)), the
 *      atlas depth az = the bricks per ax * ay layer that UpdateAtlas has needed so far).
 *      Channels 0 .. n-1 must all exist.  gvdbx_load_vbx replaces the context's tree and
 *      channels by a file's first grid as LoadVBX does (configure from the file's levels, pools
 *      as stored, one channel per stored channel with the file's apron and atlas layout,
 *      UpdateAtlas); the next RebuildTopology is a full one. */
int gvdbx_save_vbx(gvdbx_ctx* ctx, const char* path);
int gvdbx_load_vbx(gvdbx_ctx* ctx, const char* path);

/* The VBX codec on host memory (no GPU needed): the pools and the brick-linear channels of one
 * grid.  slots = bricks carrying data (the flagged leaves; atlas bricks 0 .. slots-1), each
 * channel slots * br^3 32-bit words (br = 2^log2dim[0] + 2 apron).  gvdbx_vbx_probe fills the
 * counts, types and layout of a file (pointers left NULL); the caller allocates the arrays and
 * gvdbx_vbx_read fills them.  Files with bitmask nodes, compression, brick layout or topology
 * reuse are refused (GVDBX_ERR_ARG), as LoadVBX cannot read them either. */
#define GVDBX_VBX_MAXLEV 10
#define GVDBX_VBX_MAXCHAN 32
typedef struct gvdbx_vbx_desc {
  int32_t levels;
  int32_t log2dim[GVDBX_VBX_MAXLEV];
  int32_t apron;
  int32_t axis_count[3];                 /* atlas bricks per axis */
  int32_t nchan;
  int32_t chan_type[GVDBX_VBX_MAXCHAN];  /* GVDBX_T_FLOAT or GVDBX_T_UCHAR4 */
  float transform[12];                   /* pretranslation, XYZ angles, scale, translation */
  uint64_t root;                         /* pool id of the root: Elem(0, depth, 0) */
  int64_t node_count[GVDBX_VBX_MAXLEV];  /* pool 0 rows */
  int64_t clist_count[GVDBX_VBX_MAXLEV]; /* pool 1 rows (0 at level 0) */
  int64_t slots;
  gvdbx_node* nodes[GVDBX_VBX_MAXLEV];
  uint64_t* clists[GVDBX_VBX_MAXLEV];    /* clist_count[l] * 8^log2dim[l] entries */
  uint32_t* chan_data[GVDBX_VBX_MAXCHAN];
} gvdbx_vbx_desc;
int gvdbx_vbx_write(const char* path, const gvdbx_vbx_desc* desc);
int gvdbx_vbx_probe(const char* path, gvdbx_vbx_desc* desc);
int gvdbx_vbx_read(const char* path, gvdbx_vbx_desc* desc);

/* ---- seeded synthetic point sets (host memory; no reference counterpart: the reference ships
 *      no point fixture).  The same bits as gvdbx/synth.py's shared_* functions: draw k of stream
 *      `seed` is the SplitMix64 finalizer of seed * 0xD1B54A32D192ED03 + (k + 1) *
 *      0x9E3779B97F4A7C15, a uniform its top 53 bits times 2^-53, a normal the in-order sum of 12
 *      uniforms minus 6; FP64 arithmetic without contraction, cast to FP32 at the end.
 *      gvdbx_synth_uniform: point i coordinate a = lo[a] + (hi[a] - lo[a]) u(3i + a).
 *      gvdbx_synth_sphere: point i = center + d (radius + sigma g3) / |d|, d = (g0, g1, g2), gj the
 *        normal at draw 48 i + 12 j, |d| = sqrt((d0 d0 + d1 d1) + d2 d2).
 *      gvdbx_synth_lattice: dims[0] dims[1] dims[2] points, i = (ix dims[1] + iy) dims[2] + iz, at
 *        (c + 0.5) spacing + (u(39 i + a) - 0.5) jitter; velocity a = vel_sigma * normal at draw
 *        39 i + 3 + 12 a (vel may be NULL). */
int gvdbx_synth_uniform(uint64_t seed, int64_t n, const float lo[3], const float hi[3], float* out);
int gvdbx_synth_sphere(uint64_t seed, int64_t n, const float center[3], float radius, float sigma, float* out);
int gvdbx_synth_lattice(uint64_t seed, const int dims[3], float spacing, float jitter, float vel_sigma, float* pos,
                        float* vel);

/* ---- device memory helpers: AllocData/CommitData/RetrieveData (V.cpp:4822-4860, A.cpp:263-298) */
int gvdbx_dev_alloc(gvdbx_ctx* ctx, int64_t bytes, void** dev_ptr);
int gvdbx_dev_free(gvdbx_ctx* ctx, void* dev_ptr);
int gvdbx_commit(gvdbx_ctx* ctx, void* dev_dst, const void* host_src, int64_t bytes);
int gvdbx_retrieve(gvdbx_ctx* ctx, void* host_dst, const void* dev_src, int64_t bytes);

/* ---- launch shape of the per-brick kernels (InsertPointsSubcell's count and insert passes, the
 *      density gather).  No reference counterpart (the reference launches fixed CUDA blocks).
 *      block_threads: 64, 128 or 256 threads per brick workgroup, or 0 to choose from the items
 *      per brick (one wave for sparse bricks, four otherwise).  persistent = 1: a grid of
 *      resident workgroups takes bricks from a device queue, four at a time, instead of one
 *      workgroup per brick.  Results are the same for every shape. */
int gvdbx_set_launch(gvdbx_ctx* ctx, int block_threads, int persistent);

/* ---- graphs: the calls made between gvdbx_graph_begin and gvdbx_graph_end are recorded on the
 *      context stream as one HIP graph instead of running; gvdbx_graph_launch runs them all
 *      with one launch.  No reference counterpart (the reference syncs inside its calls).
 *      Recordable: gvdbx_insert_points_subcell, the gathers (density, level set, velocity,
 *      density + velocity, colour), gvdbx_scatter_reduce_levelset, gvdbx_update_apron(_channels),
 *      gvdbx_fill_channel(4), gvdbx_advance_points.  A recorded insert sizes its lists for the
 *      largest possible length and leaves the length on the device (gvdbx_get_subcell_info
 *      reads it back after a launch).  The calls must have run once with the same arguments
 *      before recording, so that no buffer has to grow (GVDBX_ERR_STATE otherwise, as for any
 *      call that waits for the device or changes the topology or bindings; such a failure voids
 *      the recording, and gvdbx_graph_end returns it).  The host-side decisions of the recorded
 *      calls are made at recording; their effects take place at each launch.  A launch is
 *      refused (GVDBX_ERR_STATE) once the topology, atlas, bound points or any buffer the graph
 *      uses has changed since the recording; record again then.  One graph per context. */
int gvdbx_graph_begin(gvdbx_ctx* ctx);
int gvdbx_graph_end(gvdbx_ctx* ctx);
int gvdbx_graph_launch(gvdbx_ctx* ctx);
int gvdbx_graph_release(gvdbx_ctx* ctx);

/* ---- timing of one named kernel with HIP events on the context stream.
 *      kernel: "gather", "insert", "count", "apron", "keys", "sort" ... ; enable=0 stops. */
int gvdbx_kernel_timer(gvdbx_ctx* ctx, const char* kernel, int enable);
int gvdbx_kernel_timer_read(gvdbx_ctx* ctx, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
