// kernels.hpp -- launchers of the point-path kernels (device code in kernels.hip).
#pragma once
#include "common.hpp"

namespace gvdbx {

// Compact topology key: [lev | x_rel | y_rel | z_rel], coordinates relative to the root in
// units of the level's range, b bits each; invalid points use lev = depth.  Ascending
// compact order == ascending reference key order (lev<<48 | x<<32 | y<<16 | z, P.cuh:373-381).
struct KeyFmt {
  int b;       // bits per coordinate = lgR[depth] - lgR[0]
  int nbits;   // total live bits
};

struct PointsView {
  const char* pos; int pos_off, pos_stride;
  const char* vel; int vel_off, vel_stride;
  const char* clr; int clr_off, clr_stride;  // uchar4 colours (32-bit words), or null
  int64_t n;
  float tx, ty, tz;
  // Batched clouds: point i belongs to cloud cloud[i], whose cell origin is its grid cell
  // (cloud % grid_x, cloud / grid_x % grid_y, cloud / (grid_x * grid_y)) << cell_lg.  Positions
  // stay in the cloud's own frame; only integer brick / voxel coordinates carry the origin.
  // cloud == nullptr: one cloud at the origin.
  const uint32_t* cloud;
  int cell_lg, grid_x, grid_y;
};

__device__ __forceinline__ int3 cloud_origin(const PointsView& P, int64_t i) {
  if (!P.cloud) return make_int3(0, 0, 0);
  const uint32_t k = P.cloud[i];
  const uint32_t gx = (uint32_t)P.grid_x, gy = (uint32_t)P.grid_y;
  return make_int3((int)(k % gx) << P.cell_lg, (int)((k / gx) % gy) << P.cell_lg, (int)(k / (gx * gy)) << P.cell_lg);
}

// ---- keys
template <typename K> void brick_keys(const DevTree& T, const PointsView& P, KeyFmt f, int nlev, K* keys, hipStream_t s);
// up to this many topology keys are sorted and uniqued by one workgroup
constexpr int kSmallSortKeys = 4096;
template <typename K>
// mdev: read m from device memory instead (meta_unique = -1 when it is 0 or above the limit)
void sort_unique_small(const K* in, int64_t m, KeyFmt f, int depth, K* out, int* flags, int* rank, int* meta_unique,
                       int* meta_lev, int* meta_zero, const int* mdev, hipStream_t s);
// *out (zeroed by the caller) += flagged leaves among the first n, or with markers the leaves
// whose marker is set; writes no node
void count_leaf_flags(const DevTree& T, int64_t n, const int* markers, int* out, hipStream_t s);
template <typename K>
void expand_ancestors(const DevTree& T, const K* sorted, const int* flags, const int* rank, int64_t n, KeyFmt f, K* out,
                      hipStream_t s);
// Box-node keys of the extra / incremental passes: block b of 256 points writes its keys to
// region[b * cap_b ...] and their number to blockcnt[b]; compact_regions packs the slices at
// blockoff (exclusive scan of blockcnt).  cap_b must hold 256 * the per-point bound.
template <typename K>
// compact != null: the keys are also appended to compact[0, compact_cap) (any order), their count
// added to *counter
void box_keys(const DevTree& T, const PointsView& P, KeyFmt f, float r, float hi_mul, K* region, int64_t cap_b,
              int* blockcnt, int* markers, hipStream_t s, K* compact = nullptr, int compact_cap = 0,
              int* counter = nullptr);
template <typename K>
void compact_regions(const K* region, int64_t cap_b, const int* blockcnt, const int* blockoff, int64_t nblocks, K* out,
                     int64_t out_cap, hipStream_t s);
template <typename K> void unique_flags(const K* k, int64_t m, KeyFmt f, int depth, int* flags, hipStream_t s);
template <typename K>
void level_hist(const K* k, const int* flags, int64_t m, KeyFmt f, int* lev_cnt, hipStream_t s);

struct AllocPlan {
  int64_t base[kMaxLev];      // first new pool index per level
  int64_t levstart[kMaxLev];  // first unique rank of each level
};
template <typename K>
// link_leaves: the new nodes are all leaves under linked parents; they are linked here as well
void setup_nodes(const DevTree& T, const K* k, const int* flags, const int* rank, int64_t m, KeyFmt f,
                 const AllocPlan& plan, hipStream_t s, bool link_leaves = false);
void setup_root(const DevTree& T, hipStream_t s);

// Occupancy maps of a full rebuild: per level below the root, a box of root-relative node
// coordinates, one byte per node (x-major order = key order), each level starting on a group of
// 32 nodes ("word").
struct OmapLevels {
  int nlev;                     // levels below the root
  int lo[kMaxLev][3];           // the box's first node coordinates
  int n[kMaxLev][3];            // its size in nodes
  int64_t word0[kMaxLev + 1];   // first group of each level; word0[nlev] = all groups
};
struct OmapCounts {
  int64_t home[kMaxLev];  // home (covering) nodes per level
};
// marks the covering nodes of the valid points into home and, with box, their extra-pass box
// nodes into boxmap; *overflow = 1 when a node lies outside the map box
void omap_mark(const DevTree& T, const PointsView& P, const OmapLevels& M, float r, bool box, uint8_t* home,
               uint8_t* boxmap, int* overflow, hipStream_t s);
// per group: the home nodes (ch) and the box nodes that are not home nodes (cb); boxmap may be null
void omap_counts(const uint8_t* home, const uint8_t* boxmap, int64_t words, int* ch, int* cb, hipStream_t s);
// per level the node counts from the scans of ch / cb (words + 1 entries each)
void omap_level_totals(const OmapLevels& M, const int* sh, const int* sb, int* home_cnt, int* box_cnt,
                       hipStream_t s);
// writes every level's nodes (home nodes, then the other box nodes, each in key order) into the
// pools, which start empty
void omap_nodes(const DevTree& T, const OmapLevels& M, const uint8_t* home, const uint8_t* boxmap, const int* sh,
                const int* sb, const OmapCounts& C, hipStream_t s);
// gvdbLinkBricks over the level-lev nodes [first, first + count)
void link_level(const DevTree& T, int lev, int64_t first, int64_t count, hipStream_t s);
// gvdbDelinkLeafBricks over the first count leaves; with markers, the flags of the first nmark
// leaves are set to their markers first
void delink_leaves(const DevTree& T, int64_t count, const int* markers, int64_t nmark, hipStream_t s);
void delink_level(const DevTree& T, int lev, int64_t count, hipStream_t s);
// Leaf lookup table being built (DevTree::bmap): cells [lo, lo + n) of 2^lg0-voxel bricks, or none
struct BrickMapOut {
  int* bmap;
  int lg0;
  int lo[3], n[3];
};
// clear: when given, its table is also cleared to -1 by the same launch
void leaf_flags(const DevTree& T, int64_t count, int by_parent, int* out, hipStream_t s,
                const BrickMapOut* clear = nullptr);
void count_flags(const int* flags, int64_t n, int* out, hipStream_t s);
void copy_flags_to_markers(const DevTree& T, int64_t count, int* markers, hipStream_t s);

// ---- atlas
// table: when given, every flagged leaf is also written at its cell (cleared by leaf_flags)
void assign_slots(const DevTree& T, const int* flags, const int* rank, int64_t count, int apron, int br, int cx,
                  int cy, int* leaf_slot, int* slot_leaf, int4* slot_pos, hipStream_t s,
                  const BrickMapOut* table = nullptr);

// ---- subcells
struct SubcellCfg {
  int sc, lgsc, D, res0;
  float r;
  int cmask;  // position bits inside a batched cloud's cell (all bits for a single cloud)
};
// subcell corners and leaves of every linked leaf; also zeroes ocnt and vsum (when given) there
void sc_positions(const DevTree& T, int64_t leaves, SubcellCfg c, const int* map, int* sc_pos, int* sc_nid,
                  int* map_leaf, int* ocnt, unsigned long long* vsum, hipStream_t s);

// ---- gathers (mode 0 density, 1 level set starting from `init`)
void gather_field(SubcellCfg c, int64_t nsc, const int* sc_pos, const int* sc_nid, const int* cnt, const int* off,
                  const float* list_pos, const int* leaf_slot, int apron, float* chan, int accum, int mode, float init,
                  hipStream_t s);
// GatherLevelSet_FP16 (P.cuh:775-811): the level set from ushort3 list positions quantized to
// [qmin, qmin + qrange], starting from init, without accumulation
void levelset_field_q16(SubcellCfg c, int64_t nsc, const int* sc_pos, const int* sc_nid, const int* cnt,
                        const int* off, const uint16_t* list_q16, const float qmin[3], const float qrange[3],
                        const int* leaf_slot, int apron, float* chan, float init, hipStream_t s);
// level set (GatherLevelSet and ScatterReduceLevelSet): segmented min over the sorted pairs
void levelset_field(SubcellCfg c, int64_t nsc, const int* sc_pos, const int* sc_nid, const int* cnt, const int* off,
                    const float* list_pos, const int* leaf_slot, int apron, float* chan, int accum, float init,
                    hipStream_t s);

// ---- apron and sampling
void sample_trilinear(const DevTree& T, const int* leaf_slot, int res0, int apron, const float* chan,
                      const float* q, int64_t n, float* out, hipStream_t s);

// q += v * dt in place over n points of the given byte views (FP32 product, then FP32 sum)
// Counters to the host: host[0, n) = meta[0, n) (n <= 64), then host[n] = seq with a
// system-scope release, so a host thread that sees seq there reads the new counters.  host is
// the device address of host-coherent pinned memory.
void meta_export(const int* meta, int* host, int n, unsigned seq, hipStream_t s);
void advance_points(char* pos, int pos_off, int pos_stride, const char* vel, int vel_off, int vel_stride, int64_t n,
                    float dt, hipStream_t s);

}  // namespace gvdbx
