// vbx.cpp -- the VBX file format (GVDB_FILESPEC.txt) for the pools and the brick-linear atlas.
//
// Host code only.  The writer produces the single-grid file of SaveVBX (V.cpp:1626-1768) at
// version 1.11: header with the grid transform, grid offset table, grid header, topology section
// (per-level dims, then pool 0 rows of every level, then pool 1 rows of every level) and the atlas
// section (per channel: type, stride, then the atlas volume slice by slice).  The reader accepts
// what LoadVBX (V.cpp:507-685) accepts for that layout: versions 1.0 to 2.x, the first grid,
// uncompressed GVDB topology in atlas layout, node rows of 64 bytes (no bitmasks).
//
// The atlas section stores one 3D volume per channel: axis_count bricks of br = res0 + 2 apron
// voxels per axis, x fastest, brick b at (b % ax, b / ax % ay, b / (ax ay)) (getAtlasPos,
// A.cpp:705-715).  The library keeps the same bricks brick-linear, slot s = atlas brick s, so
// the codec is a permutation of words; atlas bricks past the last slot are written as zeros.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/gvdbx.h"

namespace {

constexpr int kNodeBytes = 64;  // sizeof(Node), src/gvdb_node.h:27-40

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

template <typename T>
bool put(FILE* f, const T& v) {
  return fwrite(&v, sizeof(T), 1, f) == 1;
}
template <typename T>
bool get(FILE* f, T& v) {
  return fread(&v, sizeof(T), 1, f) == 1;
}
bool put_n(FILE* f, const void* p, size_t bytes) { return bytes == 0 || fwrite(p, 1, bytes, f) == bytes; }
bool get_n(FILE* f, void* p, size_t bytes) { return bytes == 0 || fread(p, 1, bytes, f) == bytes; }

int brick_res(const gvdbx_vbx_desc* d) { return (1 << d->log2dim[0]) + 2 * d->apron; }

// checks the fields a file needs; returns GVDBX_OK or GVDBX_ERR_ARG
int check_desc(const gvdbx_vbx_desc* d) {
  if (!d || d->levels < 1 || d->levels > GVDBX_VBX_MAXLEV) return GVDBX_ERR_ARG;
  int sum = 0;
  for (int l = 0; l < d->levels; ++l) {
    if (d->log2dim[l] < 1 || d->log2dim[l] > 8) return GVDBX_ERR_ARG;
    sum += d->log2dim[l];
    if (d->node_count[l] < 0 || d->node_count[l] > INT32_MAX) return GVDBX_ERR_ARG;
    if (d->clist_count[l] < 0 || d->clist_count[l] > INT32_MAX || (l == 0 && d->clist_count[l] != 0))
      return GVDBX_ERR_ARG;
  }
  if (sum > 30 || d->apron < 0 || d->apron > (1 << d->log2dim[0])) return GVDBX_ERR_ARG;
  if (d->nchan < 0 || d->nchan > GVDBX_VBX_MAXCHAN) return GVDBX_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (d->axis_count[a] < 1) return GVDBX_ERR_ARG;
  const int64_t bricks = (int64_t)d->axis_count[0] * d->axis_count[1] * d->axis_count[2];
  const int64_t br = brick_res(d);
  for (int a = 0; a < 3; ++a)
    if ((int64_t)d->axis_count[a] * br > INT32_MAX) return GVDBX_ERR_ARG;
  if (d->slots < 0 || d->slots > bricks) return GVDBX_ERR_ARG;
  for (int c = 0; c < d->nchan; ++c)
    if (d->chan_type[c] != GVDBX_T_FLOAT && d->chan_type[c] != GVDBX_T_UCHAR4) return GVDBX_ERR_ARG;
  return GVDBX_OK;
}

int bad_file() { return GVDBX_ERR_ARG; }

// Everything of the file up to the pools: the parsed desc (no pointers) and the file offsets of
// the pools and the atlas section.
struct Layout {
  long pools = 0, atlas = 0;
  int64_t width1[GVDBX_VBX_MAXLEV] = {};
};

int parse(FILE* f, gvdbx_vbx_desc* d, Layout* L) {
  memset(d, 0, sizeof(*d));
  uint8_t major = 0, minor = 0;
  if (!get(f, major) || !get(f, minor)) return bad_file();
  d->transform[6] = d->transform[7] = d->transform[8] = 1.0f;  // unit scale for files before 1.11
  if ((major == 1 && minor >= 11) || major > 1)
    for (int k = 0; k < 12; ++k)
      if (!get(f, d->transform[k])) return bad_file();
  int32_t num_grids = 0;
  if (!get(f, num_grids) || num_grids < 1) return bad_file();
  if (major >= 2) {
    uint8_t masks = 0;
    if (!get(f, masks)) return bad_file();
  }
  std::vector<uint64_t> offs((size_t)num_grids);
  if (!get_n(f, offs.data(), offs.size() * 8)) return bad_file();
  if (fseek(f, (long)offs[0], SEEK_SET) != 0) return bad_file();  // the first grid
  char name[256];
  uint8_t dtype, comps, compress, topotype, layout;
  float vox[3];
  int32_t leafcnt, leafdim[3], apron, nchan, reuse, axiscnt[3], axisres[3], levels;
  uint64_t atlas_sz;
  if (!get_n(f, name, 256) || !get(f, dtype) || !get(f, comps) || !get(f, compress) || !get_n(f, vox, 12) ||
      !get(f, leafcnt) || !get_n(f, leafdim, 12) || !get(f, apron) || !get(f, nchan) || !get(f, atlas_sz) ||
      !get(f, topotype) || !get(f, reuse) || !get(f, layout) || !get_n(f, axiscnt, 12) || !get_n(f, axisres, 12))
    return bad_file();
  if (compress != 0 || topotype != 2 || layout != 0) return GVDBX_ERR_ARG;  // only what LoadVBX reads
  if (!get(f, levels) || levels < 1 || levels > GVDBX_VBX_MAXLEV || !get(f, d->root)) return bad_file();
  d->levels = levels;
  d->apron = apron;
  d->nchan = nchan;
  for (int a = 0; a < 3; ++a) d->axis_count[a] = axiscnt[a];
  for (int l = 0; l < levels; ++l) {
    int32_t ld, res, range[3], cnt0, width0, cnt1, width1;
    if (!get(f, ld) || !get(f, res) || !get_n(f, range, 12) || !get(f, cnt0) || !get(f, width0) || !get(f, cnt1) ||
        !get(f, width1))
      return bad_file();
    if (width0 != kNodeBytes) return GVDBX_ERR_ARG;  // bitmask nodes (V.cpp:626-630)
    if (ld < 1 || ld > 8 || res != (1 << ld) || cnt0 < 0 || cnt1 < 0) return bad_file();
    const int64_t want1 = l > 0 ? (int64_t)8 << (3 * ld) : 0;
    if (cnt1 > 0 && width1 != want1) return bad_file();
    d->log2dim[l] = ld;
    d->node_count[l] = cnt0;
    d->clist_count[l] = l > 0 ? cnt1 : 0;
    L->width1[l] = width1;
  }
  if (leafdim[0] != (1 << d->log2dim[0]) || leafcnt != d->node_count[0]) return bad_file();
  const int br = brick_res(d);
  for (int a = 0; a < 3; ++a)
    if (axisres[a] != axiscnt[a] * br) return bad_file();
  L->pools = ftell(f);
  int64_t bytes = 0;
  for (int l = 0; l < levels; ++l) bytes += d->node_count[l] * kNodeBytes;
  for (int l = 0; l < levels; ++l) bytes += (l > 0 ? d->clist_count[l] : 0) * L->width1[l];
  L->atlas = L->pools + (long)bytes;
  // bricks with data: the flagged leaves (UpdateAtlas gives them atlas bricks 0, 1, ...)
  std::vector<gvdbx_node> leaves((size_t)d->node_count[0]);
  if (!get_n(f, leaves.data(), leaves.size() * kNodeBytes)) return bad_file();
  for (const gvdbx_node& n : leaves) d->slots += n.flags ? 1 : 0;
  if (fseek(f, L->atlas, SEEK_SET) != 0) return bad_file();
  for (int c = 0; c < nchan && c < GVDBX_VBX_MAXCHAN; ++c) {
    int32_t type = 0, stride = 0;
    if (!get(f, type) || !get(f, stride) || stride != 4) return GVDBX_ERR_ARG;
    d->chan_type[c] = type;
    const int64_t vol = (int64_t)axisres[0] * axisres[1] * axisres[2] * stride;
    if (fseek(f, (long)vol, SEEK_CUR) != 0) return bad_file();
  }
  return check_desc(d);
}

}  // namespace

extern "C" {

int gvdbx_vbx_write(const char* path, const gvdbx_vbx_desc* d) {
  if (!path) return GVDBX_ERR_ARG;
  const int chk = check_desc(d);
  if (chk != GVDBX_OK) return chk;
  for (int l = 0; l < d->levels; ++l)
    if ((d->node_count[l] && !d->nodes[l]) || (d->clist_count[l] && !d->clists[l])) return GVDBX_ERR_ARG;
  for (int c = 0; c < d->nchan; ++c)
    if (d->slots && !d->chan_data[c]) return GVDBX_ERR_ARG;
  File F;
  F.f = fopen(path, "wb");
  if (!F.f) return GVDBX_ERR_ARG;
  FILE* f = F.f;
  const int br = brick_res(d);
  const int32_t axisres[3] = {d->axis_count[0] * br, d->axis_count[1] * br, d->axis_count[2] * br};
  bool ok = put(f, (uint8_t)1) && put(f, (uint8_t)11);  // version 1.11 (V.cpp:31-32)
  ok = ok && put_n(f, d->transform, 48) && put(f, (int32_t)1);
  const long table = ftell(f);
  ok = ok && put(f, (uint64_t)0);
  const uint64_t grid_off = (uint64_t)ftell(f);
  char name[256];
  memset(name, 0, sizeof(name));
  const float vox[3] = {1.0f, 1.0f, 1.0f};
  const int32_t res0 = 1 << d->log2dim[0];
  const int32_t leafdim[3] = {res0, res0, res0};
  const uint64_t atlas_sz = d->nchan > 0 ? (uint64_t)axisres[0] * axisres[1] * axisres[2] * 4 : 0;
  ok = ok && put_n(f, name, 256) && put(f, (uint8_t)'f') && put(f, (uint8_t)1) && put(f, (uint8_t)0) &&
       put_n(f, vox, 12) && put(f, (int32_t)d->node_count[0]) && put_n(f, leafdim, 12) && put(f, (int32_t)d->apron) &&
       put(f, (int32_t)d->nchan) && put(f, atlas_sz) && put(f, (uint8_t)2) && put(f, (int32_t)0) &&
       put(f, (uint8_t)0) && put_n(f, d->axis_count, 12) && put_n(f, axisres, 12);
  ok = ok && put(f, (int32_t)d->levels) && put(f, d->root);
  int sum = 0;
  for (int l = 0; l < d->levels && ok; ++l) {
    sum += d->log2dim[l];
    const int32_t res = 1 << d->log2dim[l], rng = 1 << sum;
    const int32_t range[3] = {rng, rng, rng};
    const int32_t width1 = l > 0 ? (int32_t)(8 << (3 * d->log2dim[l])) : 0;
    ok = put(f, (int32_t)d->log2dim[l]) && put(f, res) && put_n(f, range, 12) && put(f, (int32_t)d->node_count[l]) &&
         put(f, (int32_t)kNodeBytes) && put(f, (int32_t)d->clist_count[l]) && put(f, width1);
  }
  for (int l = 0; l < d->levels && ok; ++l) ok = put_n(f, d->nodes[l], (size_t)d->node_count[l] * kNodeBytes);
  for (int l = 1; l < d->levels && ok; ++l)
    ok = put_n(f, d->clists[l], (size_t)d->clist_count[l] * ((size_t)8 << (3 * d->log2dim[l])));
  const int64_t bvox = (int64_t)br * br * br;
  std::vector<uint32_t> slice((size_t)axisres[0] * axisres[1]);
  for (int c = 0; c < d->nchan && ok; ++c) {
    ok = put(f, (int32_t)d->chan_type[c]) && put(f, (int32_t)4);
    for (int z = 0; z < axisres[2] && ok; ++z) {
      const int bz = z / br, lz = z % br;
      for (int y = 0; y < axisres[1]; ++y) {
        const int by = y / br, ly = y % br;
        uint32_t* row = slice.data() + (size_t)y * axisres[0];
        for (int bx = 0; bx < d->axis_count[0]; ++bx) {
          const int64_t b = ((int64_t)bz * d->axis_count[1] + by) * d->axis_count[0] + bx;
          if (b < d->slots)
            memcpy(row + (size_t)bx * br, d->chan_data[c] + b * bvox + ((int64_t)lz * br + ly) * br, (size_t)br * 4);
          else
            memset(row + (size_t)bx * br, 0, (size_t)br * 4);
        }
      }
      ok = put_n(f, slice.data(), slice.size() * 4);
    }
  }
  ok = ok && fseek(f, table, SEEK_SET) == 0 && put(f, grid_off);
  return ok ? GVDBX_OK : GVDBX_ERR_ARG;
}

int gvdbx_vbx_probe(const char* path, gvdbx_vbx_desc* d) {
  if (!path || !d) return GVDBX_ERR_ARG;
  File F;
  F.f = fopen(path, "rb");
  if (!F.f) return GVDBX_ERR_ARG;
  Layout L;
  return parse(F.f, d, &L);
}

int gvdbx_vbx_read(const char* path, gvdbx_vbx_desc* d) {
  if (!path || !d) return GVDBX_ERR_ARG;
  File F;
  F.f = fopen(path, "rb");
  if (!F.f) return GVDBX_ERR_ARG;
  gvdbx_vbx_desc h;
  Layout L;
  const int s = parse(F.f, &h, &L);
  if (s != GVDBX_OK) return s;
  // the caller's counts must be the probed ones (its arrays are sized by them)
  if (h.levels != d->levels || h.nchan != d->nchan || h.slots != d->slots || h.apron != d->apron) return GVDBX_ERR_ARG;
  for (int l = 0; l < h.levels; ++l)
    if (h.node_count[l] != d->node_count[l] || h.clist_count[l] != d->clist_count[l] ||
        h.log2dim[l] != d->log2dim[l])
      return GVDBX_ERR_ARG;
  for (int l = 0; l < h.levels; ++l)
    if ((h.node_count[l] && !d->nodes[l]) || (h.clist_count[l] && !d->clists[l])) return GVDBX_ERR_ARG;
  for (int c = 0; c < h.nchan; ++c)
    if (h.slots && !d->chan_data[c]) return GVDBX_ERR_ARG;
  FILE* f = F.f;
  if (fseek(f, L.pools, SEEK_SET) != 0) return GVDBX_ERR_ARG;
  for (int l = 0; l < h.levels; ++l)
    if (!get_n(f, d->nodes[l], (size_t)h.node_count[l] * kNodeBytes)) return GVDBX_ERR_ARG;
  for (int l = 0; l < h.levels; ++l) {
    const size_t bytes = (size_t)(l > 0 ? h.clist_count[l] : 0) * (size_t)L.width1[l];
    if (l == 0) {
      if (fseek(f, (long)bytes, SEEK_CUR) != 0) return GVDBX_ERR_ARG;
    } else if (!get_n(f, d->clists[l], bytes)) {
      return GVDBX_ERR_ARG;
    }
  }
  memcpy(d->transform, h.transform, sizeof(h.transform));
  d->root = h.root;
  for (int a = 0; a < 3; ++a) d->axis_count[a] = h.axis_count[a];
  for (int c = 0; c < h.nchan; ++c) d->chan_type[c] = h.chan_type[c];
  const int br = brick_res(&h);
  const int64_t bvox = (int64_t)br * br * br;
  const int32_t ax = h.axis_count[0] * br, ay = h.axis_count[1] * br, az = h.axis_count[2] * br;
  std::vector<uint32_t> slice((size_t)ax * ay);
  if (fseek(f, L.atlas, SEEK_SET) != 0) return GVDBX_ERR_ARG;
  for (int c = 0; c < h.nchan; ++c) {
    int32_t type = 0, stride = 0;
    if (!get(f, type) || !get(f, stride)) return GVDBX_ERR_ARG;
    for (int z = 0; z < az; ++z) {
      if (!get_n(f, slice.data(), slice.size() * 4)) return GVDBX_ERR_ARG;
      const int bz = z / br, lz = z % br;
      for (int y = 0; y < ay; ++y) {
        const int by = y / br, ly = y % br;
        const uint32_t* row = slice.data() + (size_t)y * ax;
        for (int bx = 0; bx < h.axis_count[0]; ++bx) {
          const int64_t b = ((int64_t)bz * h.axis_count[1] + by) * h.axis_count[0] + bx;
          if (b < h.slots)
            memcpy(d->chan_data[c] + b * bvox + ((int64_t)lz * br + ly) * br, row + (size_t)bx * br, (size_t)br * 4);
        }
      }
    }
  }
  return GVDBX_OK;
}

}  // extern "C"
