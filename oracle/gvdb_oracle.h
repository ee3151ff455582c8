/*
 * gvdb_oracle.h -- CPU restatement of the GVDB point-to-sparse-grid path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it to compare against (or time beside) the
 * HIP library.  The product (libgvdbx.so) never links, loads or calls it.
 *
 * Integer / topology rules are restated exactly (same FP32 conversions as the
 * reference kernels); channel values are computed in FP64.  Every function cites
 * the reference file:line it restates (paths relative to the gvdb-voxels tree):
 *   V.cpp  = source/gvdb_library/src/gvdb_volume_gvdb.cpp
 *   P.cuh  = source/gvdb_library/kernels/cuda_gvdb_particles.cuh
 *   N.cuh  = source/gvdb_library/kernels/cuda_gvdb_nodes.cuh
 *   O.cuh  = source/gvdb_library/kernels/cuda_gvdb_operators.cuh
 *   A.cpp  = source/gvdb_library/src/gvdb_allocator.cpp
 *   R.cuh  = source/gvdb_library/kernels/cuda_gvdb_raycast.cuh
 *
 * Parity pinning (what checks each function independently of this file):
 *   topology (keys, unique, pools, links, extra/incremental passes): brute-force
 *     set enumeration + tree invariants in tests/test_oracle_topology.py;
 *   subcell binning: voxel-centric brute force membership in tests/test_oracle_subcell.py;
 *   gather density: closed form max(0, r - min_p |p-v|) over all points;
 *   gather level set: closed form for values <= 0, brute force elsewhere;
 *   apron: neighbour-interior / boundval invariant;
 *   trilinear sample / splat / adjoint: affine reproduction, partition of unity,
 *     adjoint identity, finite differences.
 */
#ifndef GVDB_ORACLE_H
#define GVDB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXLEV 10
#define ORC_UNDEFL 0xFFFFFFFFull
#define ORC_UNDEF64 0xFFFFFFFFFFFFFFFFull
#define ORC_NOHIT 1.0e10f

/* 64-byte node, same field order as the reference host Node (src/gvdb_node.h:27-40). */
typedef struct orc_node {
  uint8_t lev, flags, prio, pad;
  int32_t pos[3];
  int32_t value[3];
  float vrange[3];
  uint64_t parent;
  uint64_t childlist;
  uint64_t mask;
} orc_node;

typedef struct orc_tree orc_tree;

/* status codes mirror include/gvdbx.h */
enum { ORC_OK = 0, ORC_ERR_ARG = 1, ORC_ERR_DEPTH0 = 2, ORC_ERR_DEPTH_MAX = 3, ORC_ERR_STATE = 6, ORC_ERR_NOMEM = 7,
       ORC_ERR_DEGENERATE = 8 };

orc_tree* orc_create(int levs, const int* log2dims, int atlas_cx, int atlas_cy, int apron);
void orc_destroy(orc_tree* t);

/* RebuildTopology state machine (V.cpp:1181-1218): full, full, then incremental. */
int orc_rebuild_topology(orc_tree* t, const float* pos, int64_t n, int64_t stride, float radius, const float* trans);
/* AccumulateTopology (V.cpp:1158-1178): incremental with existing leaves kept. */
int orc_accumulate_topology(orc_tree* t, const float* pos, int64_t n, int64_t stride, float radius,
                            const float* trans, int min_depth);
void orc_request_full_rebuild(orc_tree* t, int flag);
int orc_update_atlas(orc_tree* t);

int64_t orc_node_count(const orc_tree* t, int lev);
int64_t orc_used0(const orc_tree* t);
int64_t orc_clist_count(const orc_tree* t, int lev);
const orc_node* orc_nodes(const orc_tree* t, int lev);
const uint64_t* orc_clists(const orc_tree* t, int lev);
int orc_depth(const orc_tree* t);
void orc_rootpos(const orc_tree* t, int32_t* out3);
void orc_posminmax(const orc_tree* t, float* out6);
int64_t orc_num_slots(const orc_tree* t);
const int64_t* orc_leaf_slot(const orc_tree* t);   /* per leaf: atlas slot or -1 */

/* InsertPointsSubcell (V.cpp:4971-5092, P.cuh:109-118, 218-315, 677-694).
 * Lists are in ascending point-index order (the build's deterministic order). */
/* Threads of the subcell binning, gathers and apron (1, the default, runs the serial loops); the
 * threaded forms return the same lists and fields.  orc_max_threads: OpenMP's thread limit. */
void orc_set_threads(int n);
int orc_max_threads(void);
int orc_insert_subcell(orc_tree* t, int sc, const float* pos, int64_t n, int64_t stride, float radius,
                       const float* trans, int64_t* len_out);
int64_t orc_sc_num(const orc_tree* t);
const int32_t* orc_sc_map(const orc_tree* t);   /* per leaf (total) */
const int32_t* orc_sc_cnt(const orc_tree* t);
const int32_t* orc_sc_off(const orc_tree* t);
const int32_t* orc_sc_pos(const orc_tree* t);   /* Nsc*3 */
const int32_t* orc_sc_nid(const orc_tree* t);
const int64_t* orc_sc_list(const orc_tree* t);  /* len point indices */
const float* orc_sc_list_pos(const orc_tree* t); /* len*3 FP32 positions (p + trans) */

/* Channels: FP64 brick-linear atlas, slot-major, (res0+2a)^3 voxels per slot. */
int orc_channel_alloc(orc_tree* t, int chan);
double* orc_channel(orc_tree* t, int chan);
int64_t orc_channel_size(const orc_tree* t);
int orc_fill_channel(orc_tree* t, int chan, double v);

int orc_gather_density(orc_tree* t, float radius, int chan, int accum);      /* P.cuh:696-733 */
int orc_gather_levelset(orc_tree* t, float radius, int chan, int accum);     /* P.cuh:736-773 */
int orc_gather_velocity(orc_tree* t, const float* vel, int64_t stride, int chan_x, int chan_y, int chan_z,
                        int accum);                                             /* colour rule P.cuh:717-731 */
/* colour means of the gathers (P.cuh:713-731): packed uchar4 words stored as doubles */
int orc_gather_colour(orc_tree* t, const uint32_t* clr, int64_t stride_words, int chan, int accum);
int orc_update_apron(orc_tree* t, int chan, double boundval);                 /* O.cuh:72-126 */
/* FP16 subcell lists (P.cuh:154-216) and their level set (P.cuh:775-811) */
int orc_quantize_list_pos(const orc_tree* t, const float* qmin, const float* qrange, uint16_t* out);
int orc_quantize_list_vel(const orc_tree* t, const float* vel, int64_t stride, const float* vmin, const float* vrange,
                          uint16_t* out);
int orc_gather_levelset_q16(orc_tree* t, float r, int chan, const uint16_t* q16, const float* qmin,
                            const float* qrange);
/* colour part of ScatterDensity (P.cuh:553-572, 610-641); clr: one uint32 word per point */
int orc_scatter_colour(orc_tree* t, int chan, const float* pos, int64_t n, int64_t stride, const float* trans,
                       const uint32_t* clr, int avg);

/* ScatterReduceLevelSet: same field as GatherLevelSet (a21). */
int orc_scatter_reduce_levelset(orc_tree* t, float radius, int chan, double init);

/* Trilinear sample at world points (R.cuh:104-126 convention: voxel v at v+0.5). */
int orc_sample_trilinear(const orc_tree* t, int chan, const float* q, int64_t n, double* out);
/* Trilinear mass splat (transpose of sampling) and its adjoint. */
int orc_splat_trilinear(orc_tree* t, int chan, const float* q, const float* mass, int64_t n);
int orc_scatter_density(orc_tree* t, int chan, const float* pos, int64_t n, int64_t stride, float r,
                        const float* trans, int expand);
int orc_splat_backward(const orc_tree* t, int chan_grad, const float* q, const float* mass, int64_t n,
                       double* dq, double* dmass);
/* Sampling backward with respect to the points: dq = dout * grad sample(q). */
int orc_sample_backward(const orc_tree* t, int chan, const float* q, const double* dout, int64_t n, double* dq);

/* Leaf lookup at integer voxel position (tree descent, -1 if none). */
int64_t orc_find_leaf(const orc_tree* t, const int32_t* p);

/* CPU baseline: per-point recursive activation (V.cpp:1220-1227, 2766-2855 restated). */
int64_t orc_cpu_activate_space(int levs, const int* log2dims, const float* pos, int64_t n, int64_t stride,
                               const float* trans, int64_t* leaf_count);

#ifdef __cplusplus
}
#endif
#endif
