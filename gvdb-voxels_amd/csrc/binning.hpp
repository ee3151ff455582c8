// binning.hpp -- launchers of the home-brick subcell binning and its gathers (binning.hip).
#pragma once
#include <vector>
#include "kernels.hpp"

namespace gvdbx {

constexpr int kBBoxParts = 1024;

// two-pass bounding box: per-block partials, then one block writes meta[0..5] (ordered ints);
// the partial pass also clears clear0[0, n0) and clear1[0, n1) first
void bbox_two_pass(const PointsView& P, float* part, int* meta, hipStream_t s, uint32_t* clear0 = nullptr,
                   int64_t n0 = 0, uint32_t* clear1 = nullptr, int64_t n1 = 0);

void batch_ids(const PointsView& P, const int64_t* off, int ncloud, float limit, uint32_t* cloud, int* flag,
               hipStream_t s);
// home keys with the first radix pass's tile counts over `bits` key bits (radix_sort_u32_positions
// then sorts them; counts holds radix_count_ints(n) ints)
// slot_keys: map is leaf_slot (the keys are the points' slots, read from the lookup table)
void home_keys_counted(const DevTree& T, const PointsView& P, const int* map, bool slot_keys, int nmap,
                       uint32_t* keys, int bits, int* counts, hipStream_t s);
void key_ranges(const uint32_t* skeys, int64_t n, int nmap, int* start, hipStream_t s);
// the same start[0 .. nmax + 1] of sorted keys in [0, nmax], by one pass over the keys
void key_starts(const uint32_t* skeys, int64_t n, int nmax, int* start, hipStream_t s);

struct BinArgs {
  LaunchCfg L;   // launch shape of the count and insert kernels
  // non-null: the count kernel reads each brick's points through svals from the caller's arrays
  // and writes the sorted copies the insert reads (positions (q, index); velocities when
  // svel_out is given); null: k_gather_sorted wrote spos / svel before the count
  float4* spos_out;
  float4* svel_out;
  DevTree T;
  PointsView P;
  SubcellCfg c;
  int W;
  const int* map;
  const int* map_leaf;
  const uint32_t* svals;
  const int* start;
  int* cnt_in;   // per subcell: pairs placed through brick windows
  int* ocnt;     // per subcell: pairs placed one by one (outside windows, orphans)
  int* ocur;     // insert cursor of the one-by-one pairs
  int* wgb;      // [nmap][W^3]: this brick's offset inside cnt_in, or -1
  const int* off;
  const float4* spos;   // sorted positions (q = p + trans; w = point index bits)
  const float4* svel;   // sorted velocities, or null
  int* nbm;             // [nmap][27] neighbour subcell-map indices, written by the count kernel
  const int* nbr;       // the atlas neighbour table when the subcell map is the slot map, or null
  const int4* slot_pos; // with nbr: the slots' leaf corners (w: leaf), so brick m's corner is one load
  float* list_pos;
  float* list_vel;
  uint32_t* list_clr;   // the listed points' colours, or null
  // InsertPointsSubcell_FP16: positions (and velocities) stored as ushort3, quantized to the
  // box [qmin, qmin + qrange] (list_pos / list_vel are null then)
  uint16_t* list_q16;
  uint16_t* list_vq16;
  float qmin[3], qrange[3], vqmin[3], vqrange[3];
  unsigned long long* vsum;
  unsigned long long* wvs;  // [nmap][W^3][3]: the brick's window velocity sums (where its count > 0)
};

// the bricks' pairs, and those of the sorted points whose own leaf is missing ([start[nmap],
// start[nmap + 1]), placed one by one), in one launch each
void bin_count(const BinArgs& A, int nmap, hipStream_t s);
void bin_insert(const BinArgs& A, int nmap, hipStream_t s);
void sum_counts(const int* a, const int* b, int* out, int64_t n, hipStream_t s);
void gather_sorted(const PointsView& P, const uint32_t* svals, int64_t n, float4* spos, float4* svel, hipStream_t s);
// per subcell: cnt_in = sum of the brick-window counts of the (up to 8) bricks that reach it,
// cnt = cnt_in + ocnt, and each 256-subcell block's sum of cnt into tile[block]; then, once tile
// holds the exclusive scan of those sums (scan_tile_sums), off = exclusive scan of cnt and each
// window's slice base = off + prefix
void cell_totals(const BinArgs& A, int64_t nsc, int* cnt, int* tile, hipStream_t s);
void cell_bases(const BinArgs& A, int64_t nsc, int* cnt, int* tile, int* off, hipStream_t s);

constexpr int kMaxWindowCells = 1000;  // (D + 2)^3 for D <= 8
constexpr int kMaxPaddedVox = 1728;  // (8 + 2*2)^3: brick-window kernels handle res0 <= 8, apron <= 2
// nbr == nullptr: interior only; else apron voxels with a neighbour brick are written as well
// velocity means written by the density pass (GatherDensity with velocity channels bound)
struct VelOut {
  const int* cnt;
  const unsigned long long* vsum;
  float* ch[3];
};
// vel: null, or the velocity means' inputs and channels (needs nbr: the apron is computed)
// slot_pos: the slots' corners when the subcell map is the slot map (brick m = slot m), or null
void density_brick(SubcellCfg c, int nmap, int64_t nsc, int64_t len, const int* map_leaf, const int* off,
                   const float* list_pos, const int* leaf_slot, const gvdbx_node* nodes0, const int4* slot_pos,
                   int apron, const int* nbr, float* chan, const VelOut* vel, const LaunchCfg& L, hipStream_t s);
void velocity_brick(SubcellCfg c, int nmap, const int* map_leaf, const int* cnt, const unsigned long long* vsum,
                    const int* leaf_slot, int apron, const int* nbr, float* cx, float* cy, float* cz, int accum,
                    hipStream_t s);

// Colour means (the GatherDensity / GatherLevelSet colour rule, P.cuh:713-731): per subcell the
// exact component sums of its listed colours (csum, 4 per subcell) and the packed mean without
// accumulation (cmean); then every voxel of chan (a uchar4 channel) takes its subcell's mean, or
// with accum (old + sum) / (cnt + 1) from its own old value.  nbr != null (and !accum): the
// apron is written from the neighbour subcells, 0 where no neighbour brick exists.
void colour_sums(int64_t nsc, const int* off, const int* cnt, const uint32_t* list_clr, unsigned long long* csum,
                 uint32_t* cmean, hipStream_t s);
void colour_brick(SubcellCfg c, int nmap, const int* map_leaf, const int* cnt, const unsigned long long* csum,
                  const uint32_t* cmean, const int* leaf_slot, int apron, const int* nbr, uint32_t* chan, int accum,
                  hipStream_t s);

void velocity_subcell(SubcellCfg c, int64_t nsc, const int* sc_pos, const int* sc_nid, const int* cnt,
                      const unsigned long long* vsum, const int* leaf_slot, int apron, float* cx, float* cy,
                      float* cz, int accum, hipStream_t s);

constexpr int kMaxApronChannels = 8;
struct ApronChannels {
  int n;
  float* ptr[kMaxApronChannels];
  float bound[kMaxApronChannels];
};
// largest leaf lookup table (DevTree::bmap), in cells
constexpr int64_t kBrickMapMax = (int64_t)1 << 24;
void neighbor_table(const DevTree& T, int64_t nslots, const int4* slot_pos, const int* leaf_slot, int res0, int* nbr,
                    hipStream_t s);
void apron_multi(int64_t nslots, const int* nbr, int res0, int apron, const ApronChannels& ch, hipStream_t s);
// Apron voxels of a (res0 + 2 apron)^3 brick as (voxel index | neighbour direction << 27); the
// boundary fill reads it from device memory.
void apron_table(int res0, int apron, std::vector<uint32_t>& tab);
void apron_bound(int64_t nslots, const int* nbr, const uint32_t* tab, int ntab, int res0, int apron,
                 const ApronChannels& ch, hipStream_t s);

}  // namespace gvdbx
