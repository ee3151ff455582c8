// api.hip -- host side of the C ABI (include/gvdbx.h): context, pools, the RebuildTopology
// state machine and the launch sequences of the point path.
//
// Host orchestration restates VolumeGVDB (V.cpp = source/gvdb_library/src/gvdb_volume_gvdb.cpp):
// the same passes, pool order and state transitions, with node allocation moved onto the
// device (pool index = rank of the sorted unique key) instead of the reference's host loop.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "binning.hpp"
#include "kernels.hpp"
#include "primitives.hpp"
#include "splat.hpp"

using namespace gvdbx;

namespace {

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

enum MetaSlot { M_BBOX = 0, M_UNIQUE = 8, M_COUNT = 9, M_USED = 10, M_TOTAL = 11, M_LEV = 16, M_WIDE = 32, M_INTS = 64 };

// Scratch of gvdbx_splat_trilinear_noalloc, for callers that cannot allocate device memory: the
// splat's sort (keys and values, both ping-pong halves, the radix tile counts, key starts and
// the sorted payload) sized for kNoAllocPoints points and kNoAllocKeys keys (27 per brick + 1).
constexpr int64_t kNoAllocPoints = (int64_t)1 << 20;
constexpr int64_t kNoAllocKeys = (int64_t)1 << 20;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kScrWord = align256((size_t)(kNoAllocPoints + 1) * 4);  // one of keys, keys_alt, vals, vals_alt
constexpr size_t kScrCounts = align256((size_t)(256 * ((kNoAllocPoints + kRadixTile - 1) / kRadixTile) + 256) * 4);
constexpr size_t kScrStart = align256((size_t)(kNoAllocKeys + 2) * 4);
constexpr size_t kScrPayload = align256((size_t)(kNoAllocPoints + 1) * 16);
constexpr size_t kScratchBytes = 4 * kScrWord + kScrCounts + kScrStart + kScrPayload;

}  // namespace

// One copy per GPU, held until the process ends.  Host code never uses this name as an address:
// each device's copy is reached through the pointer hipGetSymbolAddress returns for it
// (scratch_address), and every call on one GPU shares that GPU's copy.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// The device address of `scratch` on `device`, resolved once per device with that device
// current (the caller has made it current).  Returns the resolution's status.
hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) (void)hipGetLastError();
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    g_scratch_err[device] = e;
    g_scratch_done[device] = true;
  }
  *out = g_scratch_ptr[device];
  return g_scratch_err[device];
}

}  // namespace

struct gvdbx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;

  bool configured = false;
  int levs = 0;
  int ld[kMaxLev] = {};
  int lgR[kMaxLev] = {};

  Buf nodes[kMaxLev], clist[kMaxLev];
  int64_t cnt[kMaxLev] = {};
  int64_t used0 = 0;

  int rebuild = 1, cur_depth = -1, depth = -1;
  int root[3] = {0, 0, 0};
  float posmin[3] = {0, 0, 0}, posmax[3] = {0, 0, 0};

  int axis_cx = 16, axis_cy = 16, axis_cz = 1;
  // atlas bricks along z of the reference's 3D atlas (its AddChannel count, grown by UpdateAtlas
  // to hold the used leaves, A.cpp:622-650); only the VBX layout needs it
  int atlas_cz = 0;
  int apron = -1;
  bool chan_on[32] = {};
  int chan_type[32] = {};  // GVDBX_T_FLOAT or GVDBX_T_UCHAR4 (32-bit words either way)
  // apron voxels that have a neighbour brick hold current copies of its interior (set by the
  // gathers that write the apron and by full apron passes; cleared by every other writer)
  bool nbr_apron_ok[32] = {};
  // apron voxels that have no neighbour brick hold bound_val (set by writers that fill them:
  // FillChannel, the density gather with its apron (0), UpdateApron; cleared by the others)
  bool bound_ok[32] = {};
  float bound_val[32] = {};
  Buf chan[32];
  int64_t chan_slot_cap = 0;
  int64_t nslots = 0;
  Buf leaf_slot, slot_leaf, slot_pos, nbr;
  Buf apron_tab;                     // apron voxel table of the boundary fill
  std::vector<uint32_t> h_apron_tab;
  int apron_tab_res = -1, apron_tab_apron = -1;

  const char* pos = nullptr;
  int pos_off = 0, pos_stride = 12;
  const char* vel = nullptr;
  int vel_off = 0, vel_stride = 12;
  const char* clr = nullptr;
  int clr_off = 0, clr_stride = 4;
  int64_t npts = 0;

  // batched clouds (gvdbx_batch_set_points): cloud k occupies grid cell k of 2^cell_lg voxels
  bool batch = false;
  int ncloud = 0, cell_lg = 0, grid[3] = {0, 0, 0}, batch_depth = 0;
  float batch_rmax = 0.f;
  std::vector<int64_t> cloud_off;
  Buf cloud_id, cloud_off_dev;

  int sc = 0, scdim = 0;
  int64_t nsc = 0, sclen = 0;
  bool lists_have_vel = false;
  bool lists_have_clr = false;  // the insert stored the points' colours (AUX_SUBCELL_PNT_CLR)
  // InsertPointsSubcell_FP16: list positions (and velocities) as ushort3 quantized to the
  // rebuild's bounding box (and GetMinMaxVel's velocity box)
  bool lists_fp16 = false;
  Buf list_q16, list_vq16;
  float q_min[3] = {0, 0, 0}, q_range[3] = {1, 1, 1};
  bool vel_box = false;  // GetMinMaxVel ran
  float vel_min[3] = {0, 0, 0}, vel_range[3] = {1, 1, 1};
  bool store_vel_lists = false;
  bool have_vsum = false;
  Buf sc_map, sc_cnt, sc_off, sc_pos, sc_nid, list_pos, list_vel, map_leaf, vsum;
  bool sc_map_current = false;  // sc_map holds the scan map (the insert may have used the slots)
  bool spec_sorted = false;     // incre_launch also sorted the box keys (see there)
  bool spec_scanned = true;     // incre_launch's box launch scanned its block counts (else it appended its keys)
  Buf hkeys, hkeys_alt, hvals, hvals_alt, ranges, bbox_part, cnt_in, ocnt, ocur, wgb, nbm, spos, svel;
  Buf sp_keys, sp_keys_alt, sp_vals, sp_vals_alt, sp_counts, sp_start, sp_payload;  // splat: points sorted by (brick, region)
  Buf scatter_acc;  // ScatterDensity: 64-bit fixed-point sums per padded voxel
  Buf scatter_clr;  // ScatterDensity colour: per interior voxel sums (avgColor) or point index
  Buf wvs;          // insert: per-brick window velocity sums
  Buf om_home, om_box, om_ch, om_cb, om_sh, om_sb;  // full rebuild by occupancy maps: maps, group counts, scans
  Buf list_clr, csum, cmean;  // subcell list colours; per subcell colour sums and means
  // brick order of the last splat (a permutation of its n points), reused by the splat
  // backward on the same point buffer for locality
  const void* sp_order_q = nullptr;
  int64_t sp_order_n = 0;
  const uint32_t* sp_order = nullptr;

  Buf keys, keys_alt, flags, rank, radix_counts, tile_buf, markers, meta;
  Buf box_region, box_cnt, box_off;  // per-workgroup slices of the box-pass keys
  // host-coherent pinned copy of meta: [0, M_INTS) the counters, [M_INTS] the sequence number of
  // the last export (k_meta_export); d_meta is its device address
  int* h_meta = nullptr;
  int* d_meta = nullptr;
  unsigned meta_seq = 0;

  int64_t topo_version = 0, atlas_version = -1;
  int64_t new_nodes[kMaxLev] = {};  // per level: nodes the last allocation appended
  // leaf lookup table (DevTree::bmap), valid from UpdateAtlas until the tree next changes
  Buf bmap;
  bool bm_valid = false;
  int bm_lo[3] = {0, 0, 0}, bm_n[3] = {0, 0, 0};
  float last_r = 0.f;  // radius of the last rebuild

  // this device's copy of `scratch` (gvdbx_splat_trilinear_noalloc), resolved at creation
  void* scratch = nullptr;
  hipError_t scratch_err = hipSuccess;

  // launch shape of the per-brick kernels (gvdbx_set_launch)
  int launch_block = 0, launch_persist = 0, cus = 0;
  Buf queue;  // the persistent kernels' brick queue counter

  // list buffers are sized for the largest possible length (no read-back before the insert)
  // while that takes at most this many bytes per list
  size_t list_budget = (size_t)1 << 30;

  // graph recording (gvdbx_graph_begin / _end / _launch)
  bool recording = false;
  int record_err = GVDBX_OK;     // first failure of a call while recording
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int64_t graph_topo = 0, graph_atlas = 0;
  const char* graph_pos = nullptr;
  int64_t graph_npts = 0;
  Buf graph_len;                 // the list length a recorded insert leaves on the device
  bool sclen_dev = false;        // sclen is on the device (graph_len) until read back
  int64_t sclen_hint = 0;        // the last list length known on the host
  bool graph_insert = false;     // the recording holds an insert (its length is graph_len)
  uint64_t buf_gen = 0;          // counts reallocations: a graph refers to the buffers of its recording
  // buffers replaced by larger ones: kept until the context ends (or memory runs short), so a
  // reallocation needs no wait for the work still using them
  std::vector<void*> retired;
  size_t retired_bytes = 0;
  uint64_t graph_buf_gen = 0;

  // named-kernel timer: event pairs recorded on the stream, summed when read
  std::string timer_name;
  bool timer_on = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
};

namespace {

int fail(gvdbx_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) {
    c->err = buf;
    if (c->recording && c->record_err == GVDBX_OK) c->record_err = code;  // the recording is void
  }
  return code;
}

// Calls that need the host to wait for the device, or change the topology or bindings, cannot
// be recorded into a graph.
#define NOT_RECORDING(c)                                                                            \
  do {                                                                                              \
    if ((c) && (c)->recording) return fail((c), GVDBX_ERR_STATE, "%s cannot be recorded into a graph", __func__); \
  } while (0)

#define HIPCHK(c, expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail((c), GVDBX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define CHK(expr)            \
  do {                       \
    int s_ = (expr);         \
    if (s_ != GVDBX_OK) return s_; \
  } while (0)

// A kernel launch the runtime refuses (a grid or block it cannot dispatch) leaves its work
// undone without stopping the stream; this turns it into an error before anything reads the
// results back or relies on them.
int launch_status(gvdbx_ctx* c, const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, GVDBX_ERR_HIP, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
  return GVDBX_OK;
}

// Grow-only buffer; contents are not kept.  Work already enqueued may still read the old
// buffer, so the stream is drained before it is freed (growth is rare in steady state).
// Frees the retired buffers once the stream has finished the work that may use them.
int release_retired(gvdbx_ctx* c) {
  if (c->retired.empty()) return GVDBX_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (void* p : c->retired) (void)hipFree(p);
  c->retired.clear();
  c->retired_bytes = 0;
  return GVDBX_OK;
}

// hipMalloc, retrying once after freeing the retired buffers
hipError_t dev_malloc(gvdbx_ctx* c, void** p, size_t nb) {
  hipError_t e = hipMalloc(p, nb);
  if (e != hipSuccess && !c->retired.empty()) {
    (void)hipGetLastError();
    if (release_retired(c) == GVDBX_OK) e = hipMalloc(p, nb);
  }
  return e;
}

// The old buffer of a reallocation stays valid for the work already enqueued on the stream.
int retire(gvdbx_ctx* c, void* p, size_t bytes) {
  c->retired.push_back(p);
  c->retired_bytes += bytes;
  if (c->retired_bytes > c->list_budget) return release_retired(c);
  return GVDBX_OK;
}

// Grow-only buffer; contents are not kept.  A growing buffer takes twice what is asked (a quarter
// more above 256 MiB): the topology of a moving point set grows by a few per cent per step, and
// the several dozen buffers sized by it would otherwise reallocate every few steps, each
// reallocation a hipMalloc on the host's path between the rebuild's read-back and the launches
// that follow it.
int ensure(gvdbx_ctx* c, Buf& b, size_t bytes) {
  if (bytes <= b.cap) return GVDBX_OK;
  if (c->recording)
    return fail(c, GVDBX_ERR_STATE, "a buffer would grow while recording: run the same calls once before recording");
  const size_t big = (size_t)256 << 20;
  size_t nb = bytes <= big ? 2 * bytes : bytes + bytes / 4;
  nb = (nb + 255) & ~(size_t)255;
  if (b.p) CHK(retire(c, b.p, b.cap));
  b.p = nullptr;
  b.cap = 0;
  c->buf_gen++;
  hipError_t e = dev_malloc(c, &b.p, nb);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(c, GVDBX_ERR_NOMEM, "hipMalloc(%zu) failed: %s", nb, hipGetErrorString(e));
  }
  b.cap = nb;
  return GVDBX_OK;
}

// Grow-only buffer keeping the first `keep` bytes (pool growth, A.cpp:168-187).
int ensure_keep(gvdbx_ctx* c, Buf& b, size_t bytes, size_t keep) {
  if (bytes <= b.cap) return GVDBX_OK;
  if (c->recording) return fail(c, GVDBX_ERR_STATE, "a pool would grow while recording");
  size_t nb = b.cap ? b.cap : 256;
  while (nb < bytes) nb *= 2;
  void* p = nullptr;
  hipError_t e = dev_malloc(c, &p, nb);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, GVDBX_ERR_NOMEM, "hipMalloc(%zu) failed: %s", nb, hipGetErrorString(e));
  }
  c->buf_gen++;
  if (keep && b.p) HIPCHK(c, hipMemcpyAsync(p, b.p, keep, hipMemcpyDeviceToDevice, c->stream));
  void* old = b.p;
  const size_t old_cap = b.cap;
  b.p = p;
  b.cap = nb;
  if (old) CHK(retire(c, old, old_cap));  // the copy above still reads it
  return GVDBX_OK;
}

template <typename T>
T* P(Buf& b) {
  return (T*)b.p;
}

LaunchCfg launch_of(gvdbx_ctx* c) {
  return LaunchCfg{c->launch_block, c->launch_persist, (int*)c->queue.p, c->cus};
}

int res_of(const gvdbx_ctx* c, int l) { return 1 << c->ld[l]; }
int64_t clist_per(const gvdbx_ctx* c, int l) { return (int64_t)1 << (3 * c->ld[l]); }

// Leaves the per-leaf buffers are sized for: the leaf pool's capacity, which doubles when it
// grows, so those buffers grow with the pool rather than with every leaf an incremental rebuild
// appends (delinked leaves stay in the pool, so its count only grows).
int64_t leaf_room(const gvdbx_ctx* c) {
  return std::max<int64_t>(c->cnt[0], (int64_t)(c->nodes[0].cap / sizeof(gvdbx_node)));
}

DevTree tree_of(gvdbx_ctx* c) {
  DevTree T;
  memset(&T, 0, sizeof(T));
  T.depth = c->depth;
  T.levs = c->levs;
  for (int l = 0; l < kMaxLev; ++l) {
    T.ld[l] = c->ld[l];
    T.lgR[l] = c->lgR[l];
    T.nodes[l] = P<gvdbx_node>(c->nodes[l]);
    T.clist[l] = P<uint64_t>(c->clist[l]);
  }
  T.root[0] = c->root[0];
  T.root[1] = c->root[1];
  T.root[2] = c->root[2];
  T.bmap = c->bm_valid ? P<int>(c->bmap) : nullptr;
  for (int a = 0; a < 3; ++a) {
    T.bm_lo[a] = c->bm_lo[a];
    T.bm_n[a] = c->bm_n[a];
  }
  return T;
}

PointsView points_of(gvdbx_ctx* c, int64_t n, const float* tr) {
  PointsView v;
  v.pos = c->pos;
  v.pos_off = c->pos_off;
  v.pos_stride = c->pos_stride;
  v.vel = c->vel;
  v.vel_off = c->vel_off;
  v.vel_stride = c->vel_stride;
  v.clr = c->clr;
  v.clr_off = c->clr_off;
  v.clr_stride = c->clr_stride;
  v.n = n;
  v.tx = tr ? tr[0] : 0.f;
  v.ty = tr ? tr[1] : 0.f;
  v.tz = tr ? tr[2] : 0.f;
  v.cloud = c->batch ? P<uint32_t>(c->cloud_id) : nullptr;
  v.cell_lg = c->cell_lg;
  v.grid_x = c->grid[0] > 0 ? c->grid[0] : 1;
  v.grid_y = c->grid[1] > 0 ? c->grid[1] : 1;
  return v;
}

// Counter read-back in two halves: the export kernel is enqueued where the values are final and
// more work can be enqueued behind it; the host then waits for the export's sequence number,
// which k_meta_export stores into pinned memory after the counters with a system-scope release.
// The stream runs its work in order, so every launch before the export has finished by then.
// This needs no copy engine or runtime copy kernel, and the host sees the number within a
// microsecond or so by spinning; after kMetaSpinUs it waits on the stream instead (long work
// before the read-back, or a failed launch).
constexpr int64_t kMetaSpinUs = 2000;

int read_meta_begin(gvdbx_ctx* c) {
  NOT_RECORDING(c);
  CHK(launch_status(c, "read-back"));
  ++c->meta_seq;
  meta_export(P<int>(c->meta), c->d_meta, M_INTS, c->meta_seq, c->stream);
  return launch_status(c, "read-back");
}

int read_meta_end(gvdbx_ctx* c) {
  const volatile unsigned* seq = reinterpret_cast<const volatile unsigned*>(c->h_meta + M_INTS);
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned k = 1;; ++k) {
    if (*seq == c->meta_seq) {
      std::atomic_thread_fence(std::memory_order_acquire);
      return GVDBX_OK;
    }
    if ((k & 1023u) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(kMetaSpinUs))
      break;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (*seq != c->meta_seq) return fail(c, GVDBX_ERR_HIP, "the counter read-back did not arrive");
  std::atomic_thread_fence(std::memory_order_acquire);
  return GVDBX_OK;
}

int read_meta(gvdbx_ctx* c) {
  CHK(read_meta_begin(c));
  return read_meta_end(c);
}

int zero_meta(gvdbx_ctx* c) {
  fill_u32(P<uint32_t>(c->meta), M_INTS, 0u, c->stream);
  return GVDBX_OK;
}

// Named-kernel timer: brackets the launches of one kernel family with a pair of HIP events
// on the context stream (no host synchronisation until gvdbx_kernel_timer_read).
hipEvent_t next_event(gvdbx_ctx* c) {
  if (c->ev_used == c->ev.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev.push_back(e);
  }
  return c->ev[c->ev_used++];
}

struct TimedRegion {
  gvdbx_ctx* c;
  bool on;
  // allowed = false: never times (no event is created, for the entry point that allocates nothing)
  TimedRegion(gvdbx_ctx* ctx, const char* name, bool allowed = true)
      : c(ctx), on(allowed && !ctx->recording && ctx->timer_on && ctx->timer_name == name) {
    if (on) {
      hipEvent_t e = next_event(c);
      if (e) (void)hipEventRecord(e, c->stream);
      else on = false;
    }
  }
  ~TimedRegion() {
    if (!on) return;
    hipEvent_t e = next_event(c);
    if (e) (void)hipEventRecord(e, c->stream);
  }
};

// ---------------------------------------------------------------- topology

void clear_tree(gvdbx_ctx* c) {  // Clear (V.cpp:2487-2503)
  c->bm_valid = false;
  for (int l = 0; l < c->levs; ++l) c->cnt[l] = 0;
  c->used0 = 0;
  c->rebuild = 1;
}

// Capacity for `add` more nodes at level l (nodes + child lists keep their contents).
int reserve_level(gvdbx_ctx* c, int l, int64_t add) {
  const int64_t need = c->cnt[l] + add;
  CHK(ensure_keep(c, c->nodes[l], (size_t)need * sizeof(gvdbx_node), (size_t)c->cnt[l] * sizeof(gvdbx_node)));
  if (l > 0)
    CHK(ensure_keep(c, c->clist[l], (size_t)(need * clist_per(c, l)) * 8, (size_t)(c->cnt[l] * clist_per(c, l)) * 8));
  return GVDBX_OK;
}

// GetBoundingBox (V.cpp:5404-5467): launch; the result lands in meta[0..5].  The launch also
// clears the other meta counters, and the first nmark markers when nmark > 0.
int bounding_box_launch(gvdbx_ctx* c, int64_t n, const float* tr, int64_t nmark = 0) {
  CHK(ensure(c, c->bbox_part, (size_t)kBBoxParts * 6 * sizeof(float)));
  PointsView pv = points_of(c, n, tr);
  bbox_two_pass(pv, P<float>(c->bbox_part), P<int>(c->meta), c->stream, P<uint32_t>(c->meta), M_INTS,
                nmark > 0 ? P<uint32_t>(c->markers) : nullptr, nmark);
  return GVDBX_OK;
}

// posmin/posmax from the host copy of meta (min - 1 and max + 2, clamped at 0)
void bounding_box_finish(gvdbx_ctx* c) {
  for (int a = 0; a < 3; ++a) {
    float mn = ordered_to_float(c->h_meta[M_BBOX + a]);
    float mx = ordered_to_float(c->h_meta[M_BBOX + 3 + a]);
    c->posmin[a] = mn - 1.0f;
    c->posmax[a] = mx + 2.0f;
    if (0 > c->posmin[a]) c->posmin[a] = 0;
    if (0 > c->posmax[a]) c->posmax[a] = 0;
  }
}

int determine_depth(gvdbx_ctx* c, int* root) {  // DetermineDepth (V.cpp:821-837)
  for (int lev = 0; lev < c->levs; ++lev) {
    const int64_t R = (int64_t)1 << c->lgR[lev];
    bool same = true;
    int c0[3];
    for (int a = 0; a < 3; ++a) {
      int64_t pm = (int32_t)c->posmin[a], px = (int32_t)c->posmax[a];
      int64_t cm = pm / R * R, cx = px / R * R;
      if (pm < cm) cm -= R;
      if (px < cx) cx -= R;
      c0[a] = (int)cm;
      if (cm != cx) same = false;
    }
    if (same) {
      root[0] = c0[0];
      root[1] = c0[1];
      root[2] = c0[2];
      return lev;
    }
  }
  return -1;
}

KeyFmt key_fmt(const gvdbx_ctx* c, int d) {
  KeyFmt f;
  f.b = c->lgR[d] - c->lgR[0];
  int lb = 0;
  while ((1 << lb) <= d) ++lb;
  f.nbits = 3 * f.b + lb;
  return f;
}

template <typename K>
int alloc_unique(gvdbx_ctx* c, int64_t m, KeyFmt f, int d, const K* sorted, bool with_used);

// Sort m compact keys in c->keys, unique them, and append the nodes per level in sorted
// order (FindUniqueBrick V.cpp:843-873 + host allocation V.cpp:1308-1331 / 975-997).
// count_leaves >= 0: also count the flagged leaves among the first count_leaves into the used
// count, read back with the level counts (the incremental pass's used count, V.cpp:1144-1147).
template <typename K>
int sort_unique_alloc(gvdbx_ctx* c, int64_t m, KeyFmt f, int d, int64_t count_leaves = -1,
                      const int* markers = nullptr) {
  if (m <= 0) return GVDBX_OK;
  if (m > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "%lld topology keys exceed the 32-bit sort index", (long long)m);
  const int64_t ncnt = radix_count_ints(m);
  CHK(ensure(c, c->radix_counts, (size_t)ncnt * 4));
  CHK(ensure(c, c->tile_buf, (size_t)(scan_tile_count(ncnt > m ? ncnt : m) + 16) * 4));
  CHK(ensure(c, c->flags, (size_t)m * 4));
  CHK(ensure(c, c->rank, (size_t)m * 4));
  const K* sorted;
  if (m <= kSmallSortKeys) {  // one workgroup: sort + unique + ranks + level counts
    TimedRegion tr(c, "sort");
    sort_unique_small<K>(P<K>(c->keys), m, f, d, P<K>(c->keys_alt), P<int>(c->flags), P<int>(c->rank),
                         P<int>(c->meta) + M_UNIQUE, P<int>(c->meta) + M_LEV, P<int>(c->meta) + M_USED, nullptr,
                         c->stream);
    sorted = P<K>(c->keys_alt);
  } else {
    int par;
    {
      TimedRegion tr(c, "sort");
      if (sizeof(K) == 4)
        par = radix_sort_u32(P<uint32_t>(c->keys), P<uint32_t>(c->keys_alt), nullptr, nullptr, m, 0, f.nbits,
                             P<int>(c->radix_counts), P<int>(c->tile_buf), c->stream);
      else
        par = radix_sort_u64(P<uint64_t>(c->keys), P<uint64_t>(c->keys_alt), nullptr, nullptr, m, 0, f.nbits,
                             P<int>(c->radix_counts), P<int>(c->tile_buf), c->stream);
    }
    sorted = par ? P<K>(c->keys_alt) : P<K>(c->keys);
    CHK(zero_meta(c));
    unique_flags<K>(sorted, m, f, d, P<int>(c->flags), c->stream);
    exclusive_scan(P<int>(c->flags), P<int>(c->rank), m, P<int>(c->tile_buf), P<int>(c->meta) + M_UNIQUE,
                   c->stream);
    level_hist<K>(sorted, P<int>(c->flags), m, f, P<int>(c->meta) + M_LEV, c->stream);
  }
  if (count_leaves >= 0)
    count_leaf_flags(tree_of(c), count_leaves, markers, P<int>(c->meta) + M_USED, c->stream);
  CHK(read_meta(c));
  return alloc_unique<K>(c, m, f, d, sorted, count_leaves >= 0);
}

// The allocation half: per-level counts from the read-back meta, nodes appended in sorted order
// (with_used: the used count was read back too).
template <typename K>
int alloc_unique(gvdbx_ctx* c, int64_t m, KeyFmt f, int d, const K* sorted, bool with_used) {
  c->bm_valid = false;  // the tree changes from here on
  if (with_used) c->used0 = c->h_meta[M_USED];
  AllocPlan plan;
  memset(&plan, 0, sizeof(plan));
  int64_t acc = 0;
  int64_t levcnt[kMaxLev] = {};
  for (int l = 0; l < d; ++l) {
    levcnt[l] = c->h_meta[M_LEV + l];
    plan.levstart[l] = acc;
    plan.base[l] = c->cnt[l];
    acc += levcnt[l];
    CHK(reserve_level(c, l, levcnt[l]));
  }
  DevTree T = tree_of(c);
  // only new leaves (every inner node exists and is linked): the set-up launch links them too
  bool leaves_only = levcnt[0] > 0 && c->cnt[d] > 0;
  for (int l = 1; l < d; ++l) leaves_only = leaves_only && levcnt[l] == 0;
  setup_nodes<K>(T, sorted, P<int>(c->flags), P<int>(c->rank), m, f, plan, c->stream, leaves_only);
  for (int l = 1; l < d; ++l)
    if (levcnt[l])
      fill_u64(P<uint64_t>(c->clist[l]) + c->cnt[l] * clist_per(c, l), levcnt[l] * clist_per(c, l), kUndef64,
               c->stream);
  for (int l = 0; l < kMaxLev; ++l) c->new_nodes[l] = l < d ? levcnt[l] : 0;
  if (leaves_only) c->new_nodes[0] = 0;  // linked by the set-up launch
  for (int l = 0; l < d; ++l) c->cnt[l] += levcnt[l];
  c->used0 += levcnt[0];
  return GVDBX_OK;
}

// Per-level top-down gvdbLinkBricks launches over the nodes the last allocation appended: the
// older nodes are linked already (live ones since they were appended; dead ones are unflagged,
// and a delinked leaf is never found again, so it never comes back), and relinking them would
// write the same entries.
void link_all(gvdbx_ctx* c, int d) {
  DevTree T = tree_of(c);
  for (int l = d - 1; l >= 0; --l)
    if (c->new_nodes[l] > 0) link_level(T, l, c->cnt[l] - c->new_nodes[l], c->new_nodes[l], c->stream);
}

// full rebuilds of up to this many keys (points x levels) sort them all at once
constexpr int64_t kOneSortKeys = (int64_t)1 << 22;

template <typename K>
int full_pass(gvdbx_ctx* c, int64_t n, const float* tr, int d) {  // ActivateBricksGPU (V.cpp:1229-1413)
  // The sorted unique keys of every level (level in the top bits) are the unique leaf keys
  // followed by their ancestors, so only the n leaf keys are sorted at full size; each unique
  // leaf then expands into its d keys, and that short list goes through the usual sort /
  // unique / allocation.  The node set and order are those of sorting all n * d keys, which
  // small inputs still do in one sort (one host read-back fewer).
  KeyFmt f = key_fmt(c, d);
  c->bm_valid = false;
  if (n > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "point count exceeds 2^31");
  if (n * d <= kOneSortKeys) {
    CHK(ensure(c, c->keys, (size_t)(n * d) * sizeof(K)));
    CHK(ensure(c, c->keys_alt, (size_t)(n * d) * sizeof(K)));
    CHK(reserve_level(c, d, 1));
    c->cnt[d] = 1;
    DevTree T = tree_of(c);
    setup_root(T, c->stream);
    fill_u64(P<uint64_t>(c->clist[d]), clist_per(c, d), kUndef64, c->stream);
    {
      TimedRegion trg(c, "keys");
      brick_keys<K>(T, points_of(c, n, tr), f, d, P<K>(c->keys), c->stream);
    }
    CHK(sort_unique_alloc<K>(c, n * d, f, d));
    link_all(c, d);
    return GVDBX_OK;
  }
  CHK(ensure(c, c->keys, (size_t)n * sizeof(K)));
  CHK(ensure(c, c->keys_alt, (size_t)n * sizeof(K)));
  CHK(reserve_level(c, d, 1));
  c->cnt[d] = 1;
  DevTree T = tree_of(c);
  setup_root(T, c->stream);
  fill_u64(P<uint64_t>(c->clist[d]), clist_per(c, d), kUndef64, c->stream);
  {
    TimedRegion trg(c, "keys");
    brick_keys<K>(T, points_of(c, n, tr), f, 1, P<K>(c->keys), c->stream);
  }
  const int64_t ncnt = radix_count_ints(n);
  CHK(ensure(c, c->radix_counts, (size_t)ncnt * 4));
  CHK(ensure(c, c->tile_buf, (size_t)(scan_tile_count(ncnt > n ? ncnt : n) + 16) * 4));
  CHK(ensure(c, c->flags, (size_t)n * 4));
  CHK(ensure(c, c->rank, (size_t)n * 4));
  int par;
  {
    TimedRegion trs(c, "sort");
    if (sizeof(K) == 4)
      par = radix_sort_u32(P<uint32_t>(c->keys), P<uint32_t>(c->keys_alt), nullptr, nullptr, n, 0, f.nbits,
                           P<int>(c->radix_counts), P<int>(c->tile_buf), c->stream);
    else
      par = radix_sort_u64(P<uint64_t>(c->keys), P<uint64_t>(c->keys_alt), nullptr, nullptr, n, 0, f.nbits,
                           P<int>(c->radix_counts), P<int>(c->tile_buf), c->stream);
  }
  if (par) std::swap(c->keys, c->keys_alt);  // sorted leaf keys now in c->keys
  CHK(zero_meta(c));
  unique_flags<K>(P<K>(c->keys), n, f, d, P<int>(c->flags), c->stream);
  exclusive_scan(P<int>(c->flags), P<int>(c->rank), n, P<int>(c->tile_buf), P<int>(c->meta) + M_UNIQUE, c->stream);
  CHK(read_meta(c));
  const int64_t m = (int64_t)c->h_meta[M_UNIQUE] * d;
  CHK(ensure(c, c->keys_alt, (size_t)(m > 0 ? m : 1) * sizeof(K)));
  expand_ancestors<K>(T, P<K>(c->keys), P<int>(c->flags), P<int>(c->rank), n, f, P<K>(c->keys_alt), c->stream);
  std::swap(c->keys, c->keys_alt);  // expanded keys in c->keys for the sort / unique / allocation
  CHK(ensure(c, c->keys_alt, (size_t)(m > 0 ? m : 1) * sizeof(K)));
  CHK(sort_unique_alloc<K>(c, m, f, d));
  link_all(c, d);
  return GVDBX_OK;
}

// Box pass shared by the extra (hi = 2r) and incremental (hi = r) passes.  Keys go to
// per-workgroup slices sized by the per-point bound (per axis at most
// floor(((1 + hi_mul) r + 1) / res0) + 2 leaf nodes, each emitting at most `depth` keys), in
// launches of at most 4 GiB of slices; each launch is packed into c->keys after its count is read.
struct BoxGeom {
  int64_t per_point = 0, chunk = 0;
};

int box_geom(gvdbx_ctx* c, int64_t n, float r, float hi_mul, int d, size_t ksize, BoxGeom* g) {
  const double R0 = (double)(1 << c->lgR[0]);
  const int64_t per_axis = (int64_t)(((1.0 + hi_mul) * (double)r + 1.0) / R0) + 2;
  if (per_axis > 64) return fail(c, GVDBX_ERR_CAPACITY, "radius %g too large for the box pass", (double)r);
  g->per_point = per_axis * per_axis * per_axis * d;
  const int64_t budget = (int64_t)4 << 30;
  int64_t chunk = budget / (g->per_point * (int64_t)ksize) / 256 * 256;
  if (chunk < 256) chunk = 256;
  g->chunk = chunk < n ? chunk : n;
  return GVDBX_OK;
}

// Enqueue the box kernel for points [start, start + cnt) and the scan of its slice counts
// (total -> meta[M_COUNT]).
template <typename K>
// compact: the keys are also appended to c->keys (up to kSmallSortKeys; their count at M_COUNT)
// instead of the block counts being scanned (the caller scans them when it needs the slices)
int box_launch(gvdbx_ctx* c, int64_t start, int64_t cnt, const float* tr, float r, float hi_mul, int d, int* markers,
               const BoxGeom& g, bool compact = false) {
  const int64_t nb = (cnt + 255) / 256;
  CHK(ensure(c, c->box_region, (size_t)nb * 256 * g.per_point * sizeof(K)));
  CHK(ensure(c, c->box_cnt, (size_t)(nb + 1) * 4));
  CHK(ensure(c, c->box_off, (size_t)(nb + 1) * 4));
  CHK(ensure(c, c->tile_buf, (size_t)(scan_tile_count(nb + 1) + 16) * 4));
  TimedRegion trg(c, "box");
  PointsView pv = points_of(c, cnt, tr);
  pv.pos += start * (int64_t)pv.pos_stride;
  if (pv.cloud) pv.cloud += start;
  if (compact) {
    CHK(ensure(c, c->keys, (size_t)kSmallSortKeys * sizeof(K)));
    box_keys<K>(tree_of(c), pv, key_fmt(c, d), r, hi_mul, P<K>(c->box_region), 256 * g.per_point, P<int>(c->box_cnt),
                markers, c->stream, P<K>(c->keys), kSmallSortKeys, P<int>(c->meta) + M_COUNT);
  } else {
    box_keys<K>(tree_of(c), pv, key_fmt(c, d), r, hi_mul, P<K>(c->box_region), 256 * g.per_point, P<int>(c->box_cnt),
                markers, c->stream);
    exclusive_scan(P<int>(c->box_cnt), P<int>(c->box_off), nb, P<int>(c->tile_buf), P<int>(c->meta) + M_COUNT,
                   c->stream);
  }
  return GVDBX_OK;
}

// prelaunched: the launch of the first chunk was enqueued before the caller's meta read-back.
template <typename K>
int box_pass(gvdbx_ctx* c, int64_t n, const float* tr, float r, float hi_mul, int d, int* markers, int64_t* m_out,
             bool prelaunched = false) {
  BoxGeom g;
  CHK(box_geom(c, n, r, hi_mul, d, sizeof(K), &g));
  int64_t total = 0;
  for (int64_t start = 0; start < n; start += g.chunk) {
    const int64_t cnt = n - start < g.chunk ? n - start : g.chunk;
    if (!(prelaunched && start == 0)) {
      CHK(box_launch<K>(c, start, cnt, tr, r, hi_mul, d, markers, g));
      CHK(read_meta(c));
    } else if (!c->spec_scanned) {  // the speculative launch appended its keys instead: scan the slices
      const int64_t nb = (cnt + 255) / 256;
      exclusive_scan(P<int>(c->box_cnt), P<int>(c->box_off), nb, P<int>(c->tile_buf), nullptr, c->stream);
    }
    const int64_t m = (uint32_t)c->h_meta[M_COUNT];
    if (total + m > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "more than 2^31 box-pass keys");
    CHK(ensure_keep(c, c->keys, (size_t)(total + m + 1) * sizeof(K), (size_t)total * sizeof(K)));
    compact_regions<K>(P<K>(c->box_region), 256 * g.per_point, P<int>(c->box_cnt), P<int>(c->box_off),
                       (cnt + 255) / 256, P<K>(c->keys) + total, INT64_MAX, c->stream);
    total += m;
  }
  CHK(ensure(c, c->keys_alt, (size_t)(total > 0 ? total : 1) * sizeof(K)));
  *m_out = total;
  return GVDBX_OK;
}

template <typename K>
int extra_pass(gvdbx_ctx* c, int64_t n, const float* tr, float r, int d) {  // ActivateExtraBricksGPU (V.cpp:931-1016)
  int64_t m = 0;
  CHK(box_pass<K>(c, n, tr, r, 2.0f, d, nullptr, &m));
  if (m > 0) {
    CHK(sort_unique_alloc<K>(c, m, key_fmt(c, d), d));
    link_all(c, d);
  }
  return GVDBX_OK;
}

// Markers and the incremental box pass (hi = trunc(q) + r); launched before the read-back
// that also returns the bounding box.
template <typename K>
int incre_launch(gvdbx_ctx* c, int64_t n, const float* tr, float r, int d, bool accum, bool markers_clear = false) {
  const int64_t total0 = c->cnt[0];
  CHK(ensure(c, c->markers, (size_t)total0 * 4));
  DevTree T = tree_of(c);
  if (accum)
    copy_flags_to_markers(T, total0, P<int>(c->markers), c->stream);
  else if (!markers_clear)  // (otherwise the bounding-box launch cleared them)
    fill_u32(P<uint32_t>(c->markers), total0, 0u, c->stream);
  BoxGeom g;
  CHK(box_geom(c, n, r, 1.0f, d, sizeof(K), &g));
  c->spec_sorted = false;
  const bool small = n > 0 && g.chunk >= n;
  if (n > 0) CHK(box_launch<K>(c, 0, g.chunk, tr, r, 1.0f, d, P<int>(c->markers), g, small));
  c->spec_scanned = !small;
  if (small) {
    // One box launch covers the points: compact its keys, sort and unique them in one workgroup
    // and count the marked leaves, all reading the key count on the device, so the read-back
    // that returns it also returns the level and used counts.  A count above the one-workgroup
    // limit leaves this to the host path.  Nothing here writes a node: the call may still be
    // rejected (depth), and the markers are applied by the delink launch once it goes ahead.
    CHK(ensure(c, c->keys, (size_t)kSmallSortKeys * sizeof(K)));
    CHK(ensure(c, c->keys_alt, (size_t)kSmallSortKeys * sizeof(K)));
    CHK(ensure(c, c->flags, (size_t)kSmallSortKeys * 4));
    CHK(ensure(c, c->rank, (size_t)kSmallSortKeys * 4));
    // the box launch appended its keys to c->keys (count at M_COUNT); one workgroup sorts them
    sort_unique_small<K>(P<K>(c->keys), 0, key_fmt(c, d), d, P<K>(c->keys_alt), P<int>(c->flags), P<int>(c->rank),
                         P<int>(c->meta) + M_UNIQUE, P<int>(c->meta) + M_LEV, P<int>(c->meta) + M_USED,
                         P<int>(c->meta) + M_COUNT, c->stream);
    count_leaf_flags(T, total0, P<int>(c->markers), P<int>(c->meta) + M_USED, c->stream);
    c->spec_sorted = true;
  }
  return GVDBX_OK;
}

template <typename K>
int incre_pass(gvdbx_ctx* c, int64_t n, const float* tr, float r, int d, bool accum,
               bool prelaunched = false) {  // V.cpp:1020-1156
  const int64_t total0 = c->cnt[0];
  if (total0 == 0) return GVDBX_OK;
  if (!prelaunched) {
    CHK(zero_meta(c));
    CHK(incre_launch<K>(c, n, tr, r, d, accum));
    CHK(read_meta(c));
  }
  int64_t m = 0;
  if (c->spec_sorted) {
    c->spec_sorted = false;
    m = (uint32_t)c->h_meta[M_COUNT];
    if (m == 0) return GVDBX_OK;  // no new node: topology unchanged (V.cpp:1066)
    if (c->h_meta[M_UNIQUE] >= 0) {  // sorted, applied and counted on the device already
      CHK(alloc_unique<K>(c, m, key_fmt(c, d), d, P<K>(c->keys_alt), true));
      link_all(c, d);
      DevTree T = tree_of(c);
      delink_leaves(T, c->cnt[0], P<int>(c->markers), total0, c->stream);
      for (int l = 1; l < d; ++l) delink_level(T, l, c->cnt[l], c->stream);
      return GVDBX_OK;
    }
  }
  CHK(box_pass<K>(c, n, tr, r, 1.0f, d, P<int>(c->markers), &m, true));
  if (m == 0) return GVDBX_OK;  // no new node: topology unchanged (V.cpp:1066)
  // flags = markers, then the used count (V.cpp:1144-1147): the flagged leaves are the marked old
  // ones (counted from the markers, read back with the level counts) plus the new ones
  // (sort_unique_alloc adds those); the delink launch applies the markers, then unlinks the
  // leaves left unmarked
  CHK(sort_unique_alloc<K>(c, m, key_fmt(c, d), d, total0, P<int>(c->markers)));
  link_all(c, d);
  DevTree T = tree_of(c);
  delink_leaves(T, c->cnt[0], P<int>(c->markers), total0, c->stream);
  for (int l = 1; l < d; ++l) delink_level(T, l, c->cnt[l], c->stream);
  return GVDBX_OK;
}

// Full rebuild (full_pass, then extra_pass when r >= 1) by occupancy maps (kernels.hip): one
// marking launch over the points, a count and scan per map and one read-back, then the nodes are
// written in key order.  The maps cover the nodes the two passes can reach: the valid points lie
// in [posmin, posmax], their boxes reach r below and 2r above; clipped to the root.  *done stays
// false -- and nothing in the tree has changed -- when the maps would hold more than
// kOmapMaxNodes entries (points spread over a sparse box, batched clouds) or a node fell outside
// them; the caller then sorts keys.
constexpr int64_t kOmapMaxNodes = (int64_t)1 << 26;

int omap_rebuild(gvdbx_ctx* c, int64_t n, const float* tr, float r, int d, bool* done) {
  *done = false;
  if (c->batch || d < 1 || d >= kMaxLev) return GVDBX_OK;
  OmapLevels M;
  memset(&M, 0, sizeof(M));
  M.nlev = d;
  const int64_t Rr = (int64_t)1 << c->lgR[d];
  const int64_t reach = (int64_t)std::ceil((double)r) + 2;
  int64_t vlo[3], vhi[3];
  for (int a = 0; a < 3; ++a) {
    vlo[a] = std::max<int64_t>(c->root[a], (int64_t)std::floor((double)c->posmin[a]) - reach);
    vhi[a] = std::min<int64_t>(c->root[a] + Rr - 1, (int64_t)std::ceil((double)c->posmax[a]) + 2 * reach);
    if (vhi[a] < vlo[a]) return GVDBX_OK;
  }
  int64_t words = 0;  // groups of 32 map entries
  for (int l = 0; l < d; ++l) {
    const int sh = c->lgR[l];
    int64_t nodes = 1;
    for (int a = 0; a < 3; ++a) {
      const int64_t lo = (vlo[a] >> sh) - (c->root[a] >> sh), hi = (vhi[a] >> sh) - (c->root[a] >> sh);
      M.lo[l][a] = (int)lo;
      M.n[l][a] = (int)(hi - lo + 1);
      nodes *= hi - lo + 1;
      if (nodes > kOmapMaxNodes) return GVDBX_OK;
    }
    M.word0[l] = words;
    words += (nodes + 31) / 32;
    if (words * 32 > kOmapMaxNodes) return GVDBX_OK;
  }
  M.word0[d] = words;
  const bool box = r >= 1;
  CHK(ensure(c, c->om_home, (size_t)words * 32));
  if (box) CHK(ensure(c, c->om_box, (size_t)words * 32));
  CHK(ensure(c, c->om_ch, (size_t)(words + 1) * 4));
  CHK(ensure(c, c->om_cb, (size_t)(words + 1) * 4));
  CHK(ensure(c, c->om_sh, (size_t)(words + 1) * 4));
  CHK(ensure(c, c->om_sb, (size_t)(words + 1) * 4));
  CHK(ensure(c, c->tile_buf, (size_t)(scan_tile_count(words + 1) + 16) * 4));
  uint8_t* home = P<uint8_t>(c->om_home);
  uint8_t* boxmap = box ? P<uint8_t>(c->om_box) : nullptr;
  CHK(zero_meta(c));
  fill_u32((uint32_t*)home, words * 8, 0u, c->stream);
  if (box) fill_u32((uint32_t*)boxmap, words * 8, 0u, c->stream);
  {
    TimedRegion trg(c, "keys");
    omap_mark(tree_of(c), points_of(c, n, tr), M, r, box, home, boxmap, P<int>(c->meta) + M_COUNT, c->stream);
  }
  omap_counts(home, boxmap, words, P<int>(c->om_ch), P<int>(c->om_cb), c->stream);
  exclusive_scan(P<int>(c->om_ch), P<int>(c->om_sh), words, P<int>(c->tile_buf), P<int>(c->om_sh) + words, c->stream);
  exclusive_scan(P<int>(c->om_cb), P<int>(c->om_sb), words, P<int>(c->tile_buf), P<int>(c->om_sb) + words, c->stream);
  omap_level_totals(M, P<int>(c->om_sh), P<int>(c->om_sb), P<int>(c->meta) + M_LEV, P<int>(c->meta) + M_LEV + 8,
                    c->stream);
  CHK(read_meta(c));
  if (c->h_meta[M_COUNT]) return GVDBX_OK;  // a node outside the maps: sort instead
  // the tree changes from here on (each level holds at most 2 kOmapMaxNodes < 2^31 nodes)
  c->bm_valid = false;
  OmapCounts C;
  memset(&C, 0, sizeof(C));
  CHK(reserve_level(c, d, 1));
  c->cnt[d] = 1;
  for (int l = 0; l < d; ++l) {
    C.home[l] = (uint32_t)c->h_meta[M_LEV + l];
    const int64_t cnt = C.home[l] + (uint32_t)c->h_meta[M_LEV + 8 + l];
    CHK(reserve_level(c, l, cnt));
    c->cnt[l] = cnt;
    c->new_nodes[l] = cnt;
  }
  for (int l = d; l < kMaxLev; ++l) c->new_nodes[l] = 0;
  DevTree T = tree_of(c);
  setup_root(T, c->stream);
  fill_u64(P<uint64_t>(c->clist[d]), clist_per(c, d), kUndef64, c->stream);
  omap_nodes(T, M, home, boxmap, P<int>(c->om_sh), P<int>(c->om_sb), C, c->stream);
  for (int l = 1; l < d; ++l)
    if (c->cnt[l]) fill_u64(P<uint64_t>(c->clist[l]), c->cnt[l] * clist_per(c, l), kUndef64, c->stream);
  link_all(c, d);
  c->used0 = c->cnt[0];
  *done = true;
  return GVDBX_OK;
}

int check_points(gvdbx_ctx* c, int64_t n) {
  if (!c || !c->configured) return fail(c, GVDBX_ERR_ARG, "context not configured");
  if (!c->pos) return fail(c, GVDBX_ERR_STATE, "no points bound (gvdbx_set_points)");
  if (n <= 0 || n > c->npts) return fail(c, GVDBX_ERR_ARG, "point count %lld outside bound buffer (%lld)",
                                         (long long)n, (long long)c->npts);
  if (n > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "point count exceeds 2^31");
  return GVDBX_OK;
}

bool wide_keys(const gvdbx_ctx* c, int d) { return key_fmt(c, d).nbits > 32; }

int rebuild_impl(gvdbx_ctx* c, int64_t n, float r, const float* orig, bool accumulate, int min_depth) {
  NOT_RECORDING(c);
  CHK(check_points(c, n));
  // One read-back returns the bounding box and, when the tree is expected to stay at its
  // depth, the key count of the speculatively launched incremental box pass (it only writes
  // keys and markers, so it is simply dropped when a full rebuild follows).
  const bool spec = !c->rebuild && c->depth >= 1 && c->cnt[0] > 0;
  if (c->batch) {
    if (r > c->batch_rmax) return fail(c, GVDBX_ERR_ARG, "radius %g exceeds the batch guard radius %g", (double)r,
                                       (double)c->batch_rmax);
    if (orig && (orig[0] != 0.f || orig[1] != 0.f || orig[2] != 0.f))
      return fail(c, GVDBX_ERR_ARG, "batched clouds take no translation");
  }
  // the bounding-box launch clears the meta counters and, for an incremental pass without
  // accumulation, its markers
  const bool bbox_markers = spec && !c->batch && !accumulate;
  if (bbox_markers) CHK(ensure(c, c->markers, (size_t)c->cnt[0] * 4));
  if (!c->batch) CHK(bounding_box_launch(c, n, orig, bbox_markers ? c->cnt[0] : 0));
  else CHK(zero_meta(c));
  if (spec) {
    if (wide_keys(c, c->depth)) CHK(incre_launch<uint64_t>(c, n, orig, r, c->depth, accumulate, bbox_markers));
    else CHK(incre_launch<uint32_t>(c, n, orig, r, c->depth, accumulate, bbox_markers));
  }
  int root[3] = {0, 0, 0};
  int depth;
  if (c->batch) {
    // the merged root sits at the origin, one level above the grid of cloud cells
    if (spec) CHK(read_meta(c));
    depth = c->batch_depth;
    for (int a = 0; a < 3; ++a) {
      c->posmin[a] = 0.f;
      c->posmax[a] = (float)((int64_t)c->grid[a] << c->cell_lg);
    }
  } else {
    CHK(read_meta(c));
    bounding_box_finish(c);
    depth = determine_depth(c, root);
  }
  if (accumulate && depth >= 0 && depth < min_depth) depth = min_depth;
  if (depth == 0) return fail(c, GVDBX_ERR_DEPTH0, "Points extents smaller than 1 brick");
  if (depth < 0 || depth > 5 || depth >= c->levs)
    return fail(c, GVDBX_ERR_DEPTH_MAX, "Point extents exceed the configured levels (depth %d)", depth);
  // a rejected call returns above with the tree, atlas and subcells untouched (the speculative
  // pass writes scratch only); from here the topology may change
  c->topo_version++;
  c->sc = 0;
  c->last_r = r;
  if (accumulate && !c->rebuild) {  // AccumulateTopology (V.cpp:1158-1178)
    return wide_keys(c, c->depth) ? incre_pass<uint64_t>(c, n, orig, r, c->depth, true, spec)
                                  : incre_pass<uint32_t>(c, n, orig, r, c->depth, true, spec);
  }
  if (c->rebuild) {
    clear_tree(c);
  } else if (depth != c->cur_depth) {  // V.cpp:1198-1203
    clear_tree(c);
    c->rebuild = 1;
    c->cur_depth = depth;
  }
  if (c->rebuild) {
    c->depth = depth;
    for (int a = 0; a < 3; ++a) c->root[a] = root[a];
    bool done = false;
    CHK(omap_rebuild(c, n, orig, r, depth, &done));
    if (!done && wide_keys(c, depth)) {
      CHK(full_pass<uint64_t>(c, n, orig, depth));
      if (r >= 1) CHK(extra_pass<uint64_t>(c, n, orig, r, depth));
    } else if (!done) {
      CHK(full_pass<uint32_t>(c, n, orig, depth));
      if (r >= 1) CHK(extra_pass<uint32_t>(c, n, orig, r, depth));
    }
    c->rebuild = 0;
    return GVDBX_OK;
  }
  return wide_keys(c, c->depth) ? incre_pass<uint64_t>(c, n, orig, r, c->depth, false, spec)
                                : incre_pass<uint32_t>(c, n, orig, r, c->depth, false, spec);
}

int chan_any(gvdbx_ctx* c, int chan) {
  if (!c) return GVDBX_ERR_ARG;
  if (chan < 0 || chan >= 32 || !c->chan_on[chan]) return fail(c, GVDBX_ERR_ARG, "channel %d does not exist", chan);
  return GVDBX_OK;
}

// a float channel (every field operation but the colour means)
int chan_ok(gvdbx_ctx* c, int chan) {
  CHK(chan_any(c, chan));
  if (c->chan_type[chan] != GVDBX_T_FLOAT)
    return fail(c, GVDBX_ERR_ARG, "channel %d holds uchar4 colours, not floats", chan);
  return GVDBX_OK;
}

int chan_clr_ok(gvdbx_ctx* c, int chan) {
  CHK(chan_any(c, chan));
  if (c->chan_type[chan] != GVDBX_T_UCHAR4)
    return fail(c, GVDBX_ERR_ARG, "colour channel %d is not a uchar4 channel", chan);
  return GVDBX_OK;
}

// float -> uchar as the colour macros convert (uchar(x)): truncation, saturated to [0, 255]
// (NaN -> 0)
uint32_t sat_u8(float x) { return x >= 255.0f ? 255u : (x >= 0.0f ? (uint32_t)x : 0u); }

// The 32-bit word a boundary or fill value stands for in the channel: the float itself, or for
// uchar4 channels make_uchar4(v, v, v, v) (gvdbUpdateApronC4, O.cuh:143-146).
float bound_word(const gvdbx_ctx* c, int chan, float v) {
  if (c->chan_type[chan] != GVDBX_T_UCHAR4) return v;
  const uint32_t u = sat_u8(v) * 0x01010101u;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

int64_t brick_floats(const gvdbx_ctx* c) {
  const int64_t br = res_of(c, 0) + 2 * c->apron;
  return br * br * br;
}

// Channel storage for nslots bricks (AtlasResize keeps existing bricks, A.cpp:622-650).
int reserve_channels(gvdbx_ctx* c, int64_t slots) {
  if (slots <= c->chan_slot_cap) return GVDBX_OK;
  int64_t ns = c->chan_slot_cap ? c->chan_slot_cap : 64;
  while (ns < slots) ns *= 2;
  for (int ch = 0; ch < 32; ++ch) {
    if (!c->chan_on[ch]) continue;
    const size_t old = (size_t)(c->chan_slot_cap * brick_floats(c)) * 4;
    const size_t nb = (size_t)(ns * brick_floats(c)) * 4;
    CHK(ensure_keep(c, c->chan[ch], nb, old));
    fill_f32(P<float>(c->chan[ch]) + c->chan_slot_cap * brick_floats(c), (ns - c->chan_slot_cap) * brick_floats(c),
             0.f, c->stream);
  }
  c->chan_slot_cap = ns;
  return GVDBX_OK;
}

SubcellCfg subcell_cfg(const gvdbx_ctx* c, int sc, float r) {
  SubcellCfg s;
  s.sc = sc;
  s.lgsc = 0;
  while ((1 << s.lgsc) < sc) ++s.lgsc;
  s.res0 = res_of(c, 0);
  s.D = s.res0 / sc;
  s.r = r;
  s.cmask = c->batch ? (1 << c->cell_lg) - 1 : 0x7FFFFFFF;
  return s;
}

int atlas_current(gvdbx_ctx* c) {
  if (c->atlas_version != c->topo_version)
    return fail(c, GVDBX_ERR_STATE, "UpdateAtlas must run after the topology changed");
  return GVDBX_OK;
}

int gather_checks(gvdbx_ctx* c, int chan) {
  CHK(chan_ok(c, chan));
  if (c->sc <= 0) return fail(c, GVDBX_ERR_STATE, "InsertPointsSubcell must run before a gather");
  CHK(atlas_current(c));
  return GVDBX_OK;
}

// gathers that read the list positions need the float lists of InsertPointsSubcell
int float_lists(gvdbx_ctx* c) {
  if (c->lists_fp16)
    return fail(c, GVDBX_ERR_STATE, "the subcell lists hold FP16 positions (InsertPointsSubcell_FP16): "
                                    "only GatherLevelSet_FP16 reads them");
  return GVDBX_OK;
}

}  // namespace

// ======================================================================= C ABI

extern "C" {

const char* gvdbx_status_string(int s) {
  switch (s) {
    case GVDBX_OK: return "ok";
    case GVDBX_ERR_ARG: return "invalid argument";
    case GVDBX_ERR_DEPTH0: return "points extents smaller than 1 brick";
    case GVDBX_ERR_DEPTH_MAX: return "point extents exceed the configured levels";
    case GVDBX_ERR_CAPACITY: return "capacity exceeded";
    case GVDBX_ERR_HIP: return "HIP runtime error";
    case GVDBX_ERR_STATE: return "call order violated";
    case GVDBX_ERR_NOMEM: return "out of device memory";
    case GVDBX_ERR_DEGENERATE: return "degenerate single-brick topology";
    default: return "unknown status";
  }
}

int gvdbx_ctx_create(int device, void* hip_stream, gvdbx_ctx** out) {
  if (!out) return GVDBX_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    return GVDBX_ERR_HIP;
  }
  if (hipSetDevice(device) != hipSuccess) return GVDBX_ERR_HIP;
  gvdbx_ctx* c = new gvdbx_ctx();
  c->device = device;
  if (hip_stream) {
    c->stream = (hipStream_t)hip_stream;
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return GVDBX_ERR_HIP;
    }
    c->own_stream = true;
  }
  c->scratch_err = scratch_address(device, &c->scratch);
  if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) c->cus = 0;
  size_t mem_free = 0, mem_total = 0;
  if (hipMemGetInfo(&mem_free, &mem_total) == hipSuccess && mem_total / 8 > c->list_budget)
    c->list_budget = mem_total / 8;
  (void)hipGetLastError();
  if (ensure(c, c->meta, M_INTS * sizeof(int)) != GVDBX_OK || ensure(c, c->queue, 256) != GVDBX_OK ||
      ensure(c, c->graph_len, 256) != GVDBX_OK ||
      hipHostMalloc((void**)&c->h_meta, (M_INTS + 64) * sizeof(int), hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&c->d_meta, c->h_meta, 0) != hipSuccess) {
    gvdbx_ctx_destroy(c);
    return GVDBX_ERR_HIP;
  }
  memset(c->h_meta, 0, (M_INTS + 64) * sizeof(int));
  *out = c;
  return GVDBX_OK;
}

static void release_graph(gvdbx_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_insert = false;
}

int gvdbx_ctx_destroy(gvdbx_ctx* c) {
  if (!c) return GVDBX_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (c->recording) {  // an unfinished recording is dropped
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    c->recording = false;
    (void)hipGetLastError();
  }
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release_graph(c);
  Buf* bufs[] = {&c->leaf_slot, &c->slot_leaf, &c->slot_pos, &c->nbr, &c->apron_tab, &c->sc_map, &c->sc_cnt, &c->sc_off, &c->sc_pos,
                 &c->sc_nid, &c->list_pos, &c->list_vel, &c->map_leaf, &c->vsum, &c->hkeys, &c->hkeys_alt,
                 &c->hvals, &c->hvals_alt, &c->ranges, &c->bbox_part, &c->cnt_in, &c->ocnt, &c->ocur, &c->wgb, &c->nbm, &c->spos, &c->svel, &c->keys, &c->keys_alt, &c->flags,
                 &c->rank, &c->radix_counts, &c->tile_buf, &c->markers, &c->meta, &c->box_region, &c->box_cnt,
                 &c->box_off, &c->sp_keys, &c->sp_keys_alt, &c->sp_vals, &c->sp_vals_alt, &c->sp_counts, &c->sp_start, &c->sp_payload, &c->scatter_acc, &c->wvs,
                 &c->list_clr, &c->csum, &c->cmean, &c->scatter_clr, &c->bmap, &c->list_q16, &c->list_vq16,
                 &c->cloud_id, &c->om_home, &c->om_box, &c->om_ch, &c->om_cb, &c->om_sh, &c->om_sb,
                 &c->cloud_off_dev, &c->queue, &c->graph_len};
  for (Buf* b : bufs)
    if (b->p) (void)hipFree(b->p);
  for (void* p : c->retired) (void)hipFree(p);
  for (int l = 0; l < kMaxLev; ++l) {
    if (c->nodes[l].p) (void)hipFree(c->nodes[l].p);
    if (c->clist[l].p) (void)hipFree(c->clist[l].p);
  }
  for (int ch = 0; ch < 32; ++ch)
    if (c->chan[ch].p) (void)hipFree(c->chan[ch].p);
  if (c->h_meta) (void)hipHostFree(c->h_meta);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GVDBX_OK;
}

int gvdbx_last_error(const gvdbx_ctx* c, char* buf, int buflen) {
  if (!c || !buf || buflen <= 0) return GVDBX_ERR_ARG;
  snprintf(buf, (size_t)buflen, "%s", c->err.c_str());
  return GVDBX_OK;
}

int gvdbx_synchronize(gvdbx_ctx* c) {
  NOT_RECORDING(c);
  if (!c) return GVDBX_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GVDBX_OK;
}

int gvdbx_get_stream(const gvdbx_ctx* c, void** s) {
  if (!c || !s) return GVDBX_ERR_ARG;
  *s = (void*)c->stream;
  return GVDBX_OK;
}

int gvdbx_configure(gvdbx_ctx* c, int levs, const int* log2dims) {
  NOT_RECORDING(c);  // Configure (V.cpp:2364-2434)
  if (!c || !log2dims || levs < 2 || levs > kMaxLev) return fail(c, GVDBX_ERR_ARG, "bad level configuration");
  int sum = 0;
  for (int l = 0; l < levs; ++l) {
    int v = log2dims[l] == 0 ? 1 : log2dims[l];
    if (v < 1 || v > 8) return fail(c, GVDBX_ERR_ARG, "log2 dim %d out of range at level %d", v, l);
    c->ld[l] = v;
    sum += v;
    c->lgR[l] = sum;
  }
  if (sum > 30) return fail(c, GVDBX_ERR_ARG, "total tree range 2^%d exceeds 32-bit coordinates", sum);
  for (int l = levs; l < kMaxLev; ++l) {
    c->ld[l] = 0;
    c->lgR[l] = sum;
  }
  c->levs = levs;
  c->configured = true;
  clear_tree(c);
  c->depth = -1;
  c->cur_depth = -1;
  c->nslots = 0;
  c->sc = 0;
  return GVDBX_OK;
}

int gvdbx_set_channel_default(gvdbx_ctx* c, int cx, int cy, int cz) {
  NOT_RECORDING(c);  // SetChannelDefault (V.h:390)
  if (!c || cx <= 0 || cy <= 0 || cz <= 0) return fail(c, GVDBX_ERR_ARG, "bad channel default");
  c->axis_cx = cx;
  c->axis_cy = cy;
  c->axis_cz = cz;
  return GVDBX_OK;
}

int gvdbx_add_channel(gvdbx_ctx* c, int chan, int dtype, int apron) {
  NOT_RECORDING(c);  // AddChannel (V.cpp:2437-2453)
  if (!c || !c->configured) return fail(c, GVDBX_ERR_ARG, "context not configured");
  if (chan < 0 || chan >= 32 || (dtype != GVDBX_T_FLOAT && dtype != GVDBX_T_UCHAR4) || apron < 0 ||
      apron > res_of(c, 0))
    return fail(c, GVDBX_ERR_ARG, "unsupported channel (chan %d dtype %d apron %d)", chan, dtype, apron);
  if (c->chan_on[chan] && c->chan_type[chan] != dtype)
    return fail(c, GVDBX_ERR_ARG, "channel %d exists with another type", chan);
  if (c->apron >= 0 && apron != c->apron)
    return fail(c, GVDBX_ERR_ARG, "all channels share one apron width (%d)", c->apron);
  if (c->apron < 0) c->atlas_cz = c->axis_cz;
  c->apron = apron;
  if (!c->chan_on[chan]) {
    c->chan_on[chan] = true;
    c->chan_type[chan] = dtype;
    if (c->chan_slot_cap > 0) {
      const size_t nb = (size_t)(c->chan_slot_cap * brick_floats(c)) * 4;
      CHK(ensure(c, c->chan[chan], nb));
      fill_f32(P<float>(c->chan[chan]), c->chan_slot_cap * brick_floats(c), 0.f, c->stream);
    }
  }
  return GVDBX_OK;
}

// FillChannel (V.cpp:2456-2471): all-zero values clear the channel; a float channel takes
// val[0]; a uchar4 channel takes CLR2INT(val[0], val[1], val[2], 1) (gvdbOpFillC4, O.cuh:368-373).
int gvdbx_fill_channel4(gvdbx_ctx* c, int chan, const float val[4]) {
  CHK(chan_any(c, chan));
  if (!val) return fail(c, GVDBX_ERR_ARG, "fill value is null");
  float w = val[0];
  if (c->chan_type[chan] == GVDBX_T_UCHAR4) {
    const bool zero = val[0] == 0.0f && val[1] == 0.0f && val[2] == 0.0f && val[3] == 0.0f;
    const uint32_t u = zero ? 0u
                            : sat_u8(val[0] * 255.0f) | sat_u8(val[1] * 255.0f) << 8 |
                                  sat_u8(val[2] * 255.0f) << 16 | 255u << 24;
    memcpy(&w, &u, 4);
  }
  if (c->chan_slot_cap > 0) fill_f32(P<float>(c->chan[chan]), c->chan_slot_cap * brick_floats(c), w, c->stream);
  c->nbr_apron_ok[chan] = true;
  c->bound_ok[chan] = true;
  c->bound_val[chan] = w;
  return launch_status(c, "FillChannel");
}

int gvdbx_fill_channel(gvdbx_ctx* c, int chan, float v) {  // FillChannel / ClearChannel
  const float val[4] = {v, 0.0f, 0.0f, 0.0f};
  return gvdbx_fill_channel4(c, chan, val);
}

int gvdbx_set_points(gvdbx_ctx* c, int64_t n, const void* pos, int pos_off, int pos_stride, const void* vel,
                     int vel_off, int vel_stride, const void* clr, int clr_off, int clr_stride) {
  NOT_RECORDING(c);
  if (!c || !pos || n <= 0 || pos_stride < 12 || pos_off < 0 || (pos_off & 3) || (pos_stride & 3))
    return fail(c, GVDBX_ERR_ARG, "bad position buffer");
  if (vel && (vel_stride < 12 || vel_off < 0 || (vel_off & 3) || (vel_stride & 3)))
    return fail(c, GVDBX_ERR_ARG, "bad velocity buffer");
  if (clr && (clr_stride < 4 || clr_off < 0 || (clr_off & 3) || (clr_stride & 3)))
    return fail(c, GVDBX_ERR_ARG, "bad colour buffer");
  if (c->batch) {  // leaving the batched layout: the next rebuild starts over
    c->batch = false;
    c->ncloud = 0;
    clear_tree(c);
    c->cur_depth = -1;
  }
  c->npts = n;
  c->pos = (const char*)pos;
  c->pos_off = pos_off;
  c->pos_stride = pos_stride;
  c->vel = (const char*)vel;
  c->vel_off = vel_off;
  c->vel_stride = vel_stride;
  c->clr = (const char*)clr;
  c->clr_off = clr_off;
  c->clr_stride = clr_stride;
  return GVDBX_OK;
}

int gvdbx_advance_points(gvdbx_ctx* c, int64_t n, float dt) {
  CHK(check_points(c, n));
  if (c->batch) return fail(c, GVDBX_ERR_STATE, "batched clouds are moved by the caller and bound again");
  if (!c->vel) return fail(c, GVDBX_ERR_STATE, "no velocities bound (gvdbx_set_points vel)");
  if (!std::isfinite(dt)) return fail(c, GVDBX_ERR_ARG, "time step is not finite");
  advance_points((char*)c->pos, c->pos_off, c->pos_stride, c->vel, c->vel_off, c->vel_stride, n, dt, c->stream);
  HIPCHK(c, hipGetLastError());
  return GVDBX_OK;
}

int gvdbx_batch_set_points(gvdbx_ctx* c, int ncloud, const int64_t* counts, const void* pos, int pos_off,
                           int pos_stride, const void* vel, int vel_off, int vel_stride, int cell_log2,
                           float max_radius) {
  NOT_RECORDING(c);
  if (!c || !c->configured) return fail(c, GVDBX_ERR_ARG, "context not configured");
  if (ncloud < 1 || ncloud > 65536 || !counts) return fail(c, GVDBX_ERR_ARG, "bad cloud count %d", ncloud);
  // radii up to 1 keep every extra-pass box inside the cloud's own root (its box
  // [trunc(q) - r, trunc(q) + 2r] lies within the bounding box margins -1 / +2), so a batched
  // cloud gets exactly the bricks it gets alone
  if (!(max_radius >= 0.f) || max_radius > 1.f) return fail(c, GVDBX_ERR_ARG, "batch radius must lie in [0, 1]");
  int cell_lev = -1;
  for (int l = 1; l < c->levs - 1; ++l)
    if (c->lgR[l] == cell_log2) cell_lev = l;
  if (cell_lev < 0) return fail(c, GVDBX_ERR_ARG, "cell 2^%d is not the range of an inner level", cell_log2);
  std::vector<int64_t> off(ncloud + 1, 0);
  for (int k = 0; k < ncloud; ++k) {
    if (counts[k] < 0) return fail(c, GVDBX_ERR_ARG, "negative size of cloud %d", k);
    off[k + 1] = off[k] + counts[k];
  }
  const int64_t n = off[ncloud];
  if (n <= 0 || n > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "batched point count %lld", (long long)n);
  int g = 1;
  while ((int64_t)g * g * g < ncloud) ++g;
  const int grid[3] = {g, g, (int)((ncloud + (int64_t)g * g - 1) / ((int64_t)g * g))};
  const int64_t extent = (int64_t)(grid[0] > grid[2] ? grid[0] : grid[2]) << cell_log2;
  int depth = -1;
  for (int l = cell_lev + 1; l < c->levs && depth < 0; ++l)
    if (((int64_t)1 << c->lgR[l]) >= extent) depth = l;
  if (depth < 0) return fail(c, GVDBX_ERR_CAPACITY, "%d clouds of 2^%d voxels exceed the configured levels", ncloud,
                             cell_log2);
  const bool same = c->batch && c->ncloud == ncloud && c->cell_lg == cell_log2 && c->cloud_off == off;
  c->batch = false;  // rebinding the points; the layout check below decides whether the tree restarts
  CHK(gvdbx_set_points(c, n, pos, pos_off, pos_stride, vel, vel_off, vel_stride, nullptr, 0, 4));
  CHK(ensure(c, c->cloud_id, (size_t)(n + 1) * 4));
  CHK(ensure(c, c->cloud_off_dev, (size_t)(ncloud + 1) * 8));
  c->cloud_off = off;
  HIPCHK(c, hipMemcpyAsync(c->cloud_off_dev.p, c->cloud_off.data(), (ncloud + 1) * 8, hipMemcpyHostToDevice,
                           c->stream));
  // valid coordinates must stay below the cell's last brick row minus the reach of the box
  // passes (trunc(q) + 2r) and the bounding-box margin, so no brick, subcell list or apron of
  // one cloud can touch another cloud's cell
  const float limit = (float)((1 << cell_log2) - res_of(c, 0)) - 2.f * max_radius - 3.f;
  CHK(zero_meta(c));
  PointsView pv = points_of(c, n, nullptr);
  batch_ids(pv, P<int64_t>(c->cloud_off_dev), ncloud, limit, P<uint32_t>(c->cloud_id), P<int>(c->meta) + M_COUNT,
            c->stream);
  CHK(read_meta(c));
  if (c->h_meta[M_COUNT]) {
    c->ncloud = 0;
    clear_tree(c);
    c->cur_depth = -1;
    return fail(c, GVDBX_ERR_CAPACITY, "a cloud reaches beyond %g voxels of its 2^%d cell", (double)limit, cell_log2);
  }
  c->batch = true;
  c->ncloud = ncloud;
  c->cell_lg = cell_log2;
  for (int a = 0; a < 3; ++a) c->grid[a] = grid[a];
  c->batch_depth = depth;
  c->batch_rmax = max_radius;
  if (!same) {  // a new layout: the next rebuild starts over
    clear_tree(c);
    c->cur_depth = -1;
  }
  return GVDBX_OK;
}

int gvdbx_batch_get_layout(const gvdbx_ctx* c, int* ncloud, int* cell_log2, int grid[3]) {
  if (!c) return GVDBX_ERR_ARG;
  if (ncloud) *ncloud = c->batch ? c->ncloud : 0;
  if (cell_log2) *cell_log2 = c->cell_lg;
  for (int a = 0; a < 3 && grid; ++a) grid[a] = c->batch ? c->grid[a] : 0;
  return GVDBX_OK;
}

int gvdbx_rebuild_topology(gvdbx_ctx* c, int64_t n, float r, const float orig[3]) {
  if (!c) return GVDBX_ERR_ARG;
  CHK(rebuild_impl(c, n, r, orig, false, 0));
  return launch_status(c, "RebuildTopology");
}

int gvdbx_accumulate_topology(gvdbx_ctx* c, int64_t n, float r, const float orig[3], int min_depth) {
  if (!c) return GVDBX_ERR_ARG;
  CHK(rebuild_impl(c, n, r, orig, true, min_depth));
  return launch_status(c, "AccumulateTopology");
}

int gvdbx_request_full_rebuild(gvdbx_ctx* c, int flag) {
  if (!c) return GVDBX_ERR_ARG;
  c->rebuild = flag ? 1 : 0;
  return GVDBX_OK;
}

int gvdbx_finish_topology(gvdbx_ctx* c, float bmin[3], float bmax[3]) {  // ComputeBounds (V.cpp:1792-1816)
  if (!c || !c->configured) return GVDBX_ERR_ARG;
  const int64_t n = c->cnt[0];
  if (n == 0) return fail(c, GVDBX_ERR_STATE, "no leaves");
  gvdbx_node* h = (gvdbx_node*)malloc((size_t)n * sizeof(gvdbx_node));
  if (!h) return GVDBX_ERR_NOMEM;
  hipError_t e = hipMemcpyAsync(h, c->nodes[0].p, (size_t)n * sizeof(gvdbx_node), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    free(h);
    return fail(c, GVDBX_ERR_HIP, "%s", hipGetErrorString(e));
  }
  const int R = res_of(c, 0);
  float mn[3] = {(float)h[0].pos[0], (float)h[0].pos[1], (float)h[0].pos[2]};
  float mx[3] = {mn[0], mn[1], mn[2]};
  for (int64_t i = 0; i < n; ++i) {
    if (!h[i].flags) continue;
    for (int a = 0; a < 3; ++a) {
      if (h[i].pos[a] < mn[a]) mn[a] = (float)h[i].pos[a];
      if (h[i].pos[a] + R > mx[a]) mx[a] = (float)(h[i].pos[a] + R);
    }
  }
  free(h);
  for (int a = 0; a < 3; ++a) {
    if (bmin) bmin[a] = mn[a];
    if (bmax) bmax[a] = mx[a];
  }
  return GVDBX_OK;
}

// The leaf lookup table over the brick cells around the last rebuild's points (their bounding
// box widened by the reach of the box passes and two bricks), clipped to the root; leaves outside
// it are still found by descent.  Not built for batched clouds (cells far apart) or when the box
// holds more than kBrickMapMax cells.
static int plan_brick_map(gvdbx_ctx* c, BrickMapOut* out, bool* use) {
  c->bm_valid = false;
  *use = false;
  if (c->batch || c->depth < 1 || c->used0 == 0) return GVDBX_OK;
  const int s = c->lgR[0];
  const int64_t R0 = (int64_t)1 << s;
  const double margin = 2.0 * (double)c->last_r + 2.0 * (double)R0 + 4.0;
  int64_t lo[3], n[3], cells = 1;
  for (int a = 0; a < 3; ++a) {
    const int64_t rlo = (int64_t)c->root[a] >> s, rn = (int64_t)1 << (c->lgR[c->depth] - s);
    int64_t l = (int64_t)std::floor(((double)c->posmin[a] - margin) / (double)R0);
    int64_t h = (int64_t)std::floor(((double)c->posmax[a] + margin) / (double)R0);
    if (l < rlo) l = rlo;
    if (h > rlo + rn - 1) h = rlo + rn - 1;
    if (h < l) return GVDBX_OK;
    lo[a] = l;
    n[a] = h - l + 1;
    cells *= n[a];
    if (cells > kBrickMapMax) return GVDBX_OK;
  }
  CHK(ensure(c, c->bmap, (size_t)cells * 8));  // (leaf, slot) per cell
  out->bmap = P<int>(c->bmap);
  out->lg0 = s;
  for (int a = 0; a < 3; ++a) {
    out->lo[a] = (int)lo[a];
    out->n[a] = (int)n[a];
  }
  *use = true;
  return GVDBX_OK;
}

int gvdbx_update_atlas(gvdbx_ctx* c) {
  NOT_RECORDING(c);  // UpdateAtlas (V.cpp:2630-2704)
  if (!c || !c->configured) return GVDBX_ERR_ARG;
  if (c->apron < 0) return fail(c, GVDBX_ERR_STATE, "UpdateAtlas needs channel 0 (AddChannel)");
  const int64_t total = c->cnt[0], room = leaf_room(c);
  CHK(ensure(c, c->flags, (size_t)(room + 1) * 4));
  CHK(ensure(c, c->rank, (size_t)(room + 1) * 4));
  CHK(ensure(c, c->tile_buf, (size_t)(scan_tile_count(room + 1) + 16) * 4));
  CHK(ensure(c, c->leaf_slot, (size_t)(room + 1) * 4));
  CHK(ensure(c, c->slot_leaf, (size_t)(room + 1) * 4));
  CHK(ensure(c, c->slot_pos, (size_t)(room + 1) * 16));
  // the lookup table is cleared by the flag launch and filled by the slot launch
  BrickMapOut tab;
  bool use_tab = false;
  CHK(plan_brick_map(c, &tab, &use_tab));
  DevTree T = tree_of(c);
  const int br = res_of(c, 0) + 2 * c->apron;
  leaf_flags(T, total, 0, P<int>(c->flags), c->stream, use_tab ? &tab : nullptr);
  exclusive_scan(P<int>(c->flags), P<int>(c->rank), total, P<int>(c->tile_buf), nullptr, c->stream);
  assign_slots(T, P<int>(c->flags), P<int>(c->rank), total, c->apron, br, c->axis_cx, c->axis_cy,
               P<int>(c->leaf_slot), P<int>(c->slot_leaf), P<int4>(c->slot_pos), c->stream,
               use_tab ? &tab : nullptr);
  c->nslots = c->used0;
  {
    const int64_t layer = (int64_t)c->axis_cx * c->axis_cy;
    const int64_t need = (c->used0 + layer - 1) / layer;
    if (need > c->atlas_cz) c->atlas_cz = (int)need;
  }
  for (int ch = 0; ch < 32; ++ch) c->nbr_apron_ok[ch] = c->bound_ok[ch] = false;
  if (use_tab) {
    for (int a = 0; a < 3; ++a) {
      c->bm_lo[a] = tab.lo[a];
      c->bm_n[a] = tab.n[a];
    }
    c->bm_valid = true;
  }
  T = tree_of(c);
  CHK(ensure(c, c->nbr, (size_t)(room * 27 + 1) * 4));
  neighbor_table(T, c->nslots, P<int4>(c->slot_pos), P<int>(c->leaf_slot), res_of(c, 0), P<int>(c->nbr), c->stream);
  CHK(reserve_channels(c, c->nslots));
  c->atlas_version = c->topo_version;
  return launch_status(c, "UpdateAtlas");
}

// Largest number of subcell corners one point reaches per axis.  The corners are the multiples of
// sc in [trunc(q - r) / sc * sc, trunc(q + 1 + r) / sc * sc] (C division), a span below
// 2r + 1 + sc + e voxels, e = the FP32 rounding of q - r and q + 1 + r (at most 1.5 units in the
// last place of the largest coordinate that can reach a leaf), so there are at most
// ceil((2r + 1 + e) / sc) + 1 of them.
static int64_t subcells_per_axis(const gvdbx_ctx* c, float r, int sc) {
  double xmax = 1.0;
  if (c->batch) {
    xmax = std::ldexp(1.0, c->cell_lg);  // positions stay in their cloud's frame
  } else if (c->depth >= 0) {
    for (int a = 0; a < 3; ++a)
      xmax = std::max(xmax, std::fabs((double)c->root[a]) + std::ldexp(1.0, c->lgR[c->depth]));
  }
  xmax += (double)r + sc + 1.0;
  const double e = 2.0 * std::ldexp(1.0, (int)std::ceil(std::log2(2.0 * xmax)) - 23);
  return (int64_t)std::ceil((2.0 * (double)r + 1.0 + e) / (double)sc) + 1;
}

static int insert_impl(gvdbx_ctx* c, int sc, int64_t n, float r, const float trans[3], int64_t* len, bool fp16) {
  CHK(check_points(c, n));  // InsertPointsSubcell (V.cpp:4971-5092), _FP16 (V.cpp:5094-5189)
  if (fp16 && c->batch) return fail(c, GVDBX_ERR_ARG, "InsertPointsSubcell_FP16 is not for batched clouds");
  if (fp16 && c->recording) return fail(c, GVDBX_ERR_STATE, "InsertPointsSubcell_FP16 cannot be recorded");
  if (c->batch) {
    if (r > c->batch_rmax) return fail(c, GVDBX_ERR_ARG, "radius %g exceeds the batch guard radius %g", (double)r,
                                       (double)c->batch_rmax);
    if (trans && (trans[0] != 0.f || trans[1] != 0.f || trans[2] != 0.f))
      return fail(c, GVDBX_ERR_ARG, "batched clouds take no translation");
  }
  const int res0 = res_of(c, 0);
  if (sc <= 0 || (sc & (sc - 1)) || res0 % sc || sc * sc * sc > 512)
    return fail(c, GVDBX_ERR_ARG, "subcell size %d must be a power of two dividing %d", sc, res0);
  if (!(r >= 0.0f) || r > 1.0e6f) return fail(c, GVDBX_ERR_ARG, "radius %g out of range", (double)r);
  const bool with_vel = c->vel != nullptr;
  const bool keep_vel_lists = with_vel && c->store_vel_lists;
  if (fp16 && keep_vel_lists && !c->vel_box)
    return fail(c, GVDBX_ERR_STATE, "GetMinMaxVel must run before an FP16 insert that stores velocities");
  if (c->used0 > 0 && c->depth < 1) return fail(c, GVDBX_ERR_DEGENERATE, "no tree");
  if (len) *len = 0;
  // Every check that needs no device result is above.  From here the subcell arrays change, so
  // the gathers refuse (c->sc = 0) until this insert has completed.
  c->sc = 0;
  if (c->used0 == 0) {
    c->sc = sc;
    c->scdim = res_of(c, 0) / sc;
    c->nsc = 0;
    c->sclen = 0;
    c->sclen_dev = false;
    return GVDBX_OK;
  }
  const SubcellCfg cfg = subcell_cfg(c, sc, r);
  const int64_t per = (int64_t)cfg.D * cfg.D * cfg.D;
  const int64_t total = c->cnt[0];
  const int64_t nsc = per * c->used0;
  const int nmap = (int)c->used0;
  if (nsc > INT32_MAX || total > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "subcell count exceeds 2^31");
  // arrays sized for every pool leaf (an upper bound of the linked leaves indexed by map), up to
  // the pool's capacity
  const int64_t room = leaf_room(c);
  const int64_t nsc_cap = per * room;
  CHK(ensure(c, c->sc_map, (size_t)(room + 1) * 4));
  CHK(ensure(c, c->map_leaf, (size_t)(room + 1) * 4));
  CHK(ensure(c, c->flags, (size_t)(room + 1) * 4));
  CHK(ensure(c, c->sc_cnt, (size_t)(nsc_cap + 1) * 4));
  CHK(ensure(c, c->sc_off, (size_t)(nsc_cap + 1) * 4));
  CHK(ensure(c, c->sc_pos, (size_t)(nsc_cap + 1) * 12));
  CHK(ensure(c, c->sc_nid, (size_t)(nsc_cap + 1) * 4));
  if (with_vel) CHK(ensure(c, c->vsum, (size_t)(nsc_cap + 1) * 24));
  CHK(ensure(c, c->hkeys, (size_t)n * 4));
  CHK(ensure(c, c->hkeys_alt, (size_t)n * 4));
  CHK(ensure(c, c->hvals, (size_t)n * 4));
  CHK(ensure(c, c->hvals_alt, (size_t)n * 4));
  CHK(ensure(c, c->ranges, (size_t)(room + 2) * 4));
  const int64_t ncnt = radix_count_ints(n);
  CHK(ensure(c, c->radix_counts, (size_t)ncnt * 4));
  int64_t big = nsc_cap > room ? nsc_cap : room;
  if (ncnt > big) big = ncnt;
  // scan tiles of the map scan, and the 256-subcell block sums of the offsets scan
  int64_t tiles = scan_tile_count(big + 1);
  if ((nsc_cap + 255) / 256 > tiles) tiles = (nsc_cap + 255) / 256;
  CHK(ensure(c, c->tile_buf, (size_t)(tiles + 16) * 4));
  DevTree T = tree_of(c);
  PointsView pv = points_of(c, n, trans);

  const int W = cfg.D + 2;
  CHK(ensure(c, c->cnt_in, (size_t)(nsc_cap + 1) * 4));
  CHK(ensure(c, c->ocnt, (size_t)(nsc_cap + 1) * 4));
  CHK(ensure(c, c->ocur, (size_t)(nsc_cap + 1) * 4));
  CHK(ensure(c, c->wgb, (size_t)(room * W * W * W + 1) * 4));
  CHK(ensure(c, c->nbm, (size_t)(room * 27 + 1) * 4));
  CHK(ensure(c, c->spos, (size_t)(n + 1) * 16));
  if (with_vel) CHK(ensure(c, c->svel, (size_t)(n + 1) * 16));

  // subcell map (gvdbSetFlagSubcell + PrefixSum, P.cuh:109-118) and corners (P.cuh:677-694).
  // Leaves are flagged exactly when they are linked (linking takes every flagged node, delinking
  // every unflagged leaf), so on a current atlas the map of a linked leaf is its slot: the
  // kernels read leaf_slot, and the full map array (dead leaves included) is made on request.
  const bool slot_map = c->atlas_version == c->topo_version && c->nslots == c->used0 && c->leaf_slot.p != nullptr;
  const int* mapp = slot_map ? P<int>(c->leaf_slot) : P<int>(c->sc_map);
  if (!slot_map) {
    leaf_flags(T, total, 1, P<int>(c->flags), c->stream);
    exclusive_scan(P<int>(c->flags), P<int>(c->sc_map), total, P<int>(c->tile_buf), nullptr, c->stream);
    fill_u32(P<uint32_t>(c->map_leaf), total + 1, 0xFFFFFFFFu, c->stream);
  }
  c->sc_map_current = !slot_map;
  // the corner kernel also clears the one-by-one counts and the velocity sums of every subcell
  sc_positions(T, total, cfg, mapp, P<int>(c->sc_pos), P<int>(c->sc_nid), P<int>(c->map_leaf), P<int>(c->ocnt),
               with_vel ? P<unsigned long long>(c->vsum) : nullptr, c->stream);

  // sort points by home brick
  uint32_t* skeys;
  uint32_t* svals;
  {
    TimedRegion tr(c, "home");
    int bits = 1;  // keys run up to the sentinel nmap + 1
    while (((int64_t)1 << bits) <= (int64_t)nmap + 1) ++bits;
    // the key launch also counts the first radix digit; the sort generates the point positions
    home_keys_counted(T, pv, mapp, slot_map, nmap, P<uint32_t>(c->hkeys), bits, P<int>(c->radix_counts), c->stream);
    const int par = radix_sort_u32_positions(P<uint32_t>(c->hkeys), P<uint32_t>(c->hkeys_alt), P<uint32_t>(c->hvals),
                                             P<uint32_t>(c->hvals_alt), n, bits, P<int>(c->radix_counts), c->stream);
    skeys = par ? P<uint32_t>(c->hkeys_alt) : P<uint32_t>(c->hkeys);
    svals = par ? P<uint32_t>(c->hvals_alt) : P<uint32_t>(c->hvals);
    key_ranges(skeys, n, nmap, P<int>(c->ranges), c->stream);
  }
  // [ranges[nmap], ranges[nmap + 1]): the valid points whose own leaf is missing, which the count
  // and insert launches place one by one (invalid points sort after them)
  // Sorted copies of the points for the per-brick kernels: with bricks of 64 to 511 points on
  // average the count pass gathers them itself (its bricks' points, through the permutation);
  // sparse bricks leave the count pass too little work per point to hide that gather's latency,
  // and bricks of hundreds of points give each count thread several dependent gathers in turn
  // (16.7M level-set points, 787 per brick: 226 + 115 µs apart against 406 µs in the count
  // pass), so there k_gather_sorted makes the copies first.
  const bool gather_in_count = n >= (int64_t)64 * nmap && n < (int64_t)512 * nmap;
  if (!gather_in_count)
    gather_sorted(pv, svals, n, P<float4>(c->spos), with_vel ? P<float4>(c->svel) : nullptr, c->stream);
  BinArgs A;
  A.L = launch_of(c);
  A.spos_out = gather_in_count ? P<float4>(c->spos) : nullptr;
  A.svel_out = (gather_in_count && keep_vel_lists) ? P<float4>(c->svel) : nullptr;
  A.T = T;
  A.P = pv;
  A.c = cfg;
  A.W = W;
  A.map = mapp;
  A.map_leaf = P<int>(c->map_leaf);
  A.svals = svals;
  A.start = P<int>(c->ranges);
  A.cnt_in = P<int>(c->cnt_in);
  A.ocnt = P<int>(c->ocnt);
  A.ocur = P<int>(c->ocur);
  A.wgb = P<int>(c->wgb);
  A.off = P<int>(c->sc_off);
  A.spos = P<float4>(c->spos);
  A.svel = with_vel ? P<float4>(c->svel) : nullptr;
  A.nbm = P<int>(c->nbm);
  A.nbr = slot_map ? P<int>(c->nbr) : nullptr;
  A.slot_pos = slot_map ? P<int4>(c->slot_pos) : nullptr;
  A.list_pos = nullptr;
  A.list_vel = nullptr;
  A.list_clr = nullptr;
  A.list_q16 = nullptr;
  A.list_vq16 = nullptr;
  for (int a = 0; a < 3; ++a) {  // quantization boxes: the rebuild's bounding box, GetMinMaxVel's box
    A.qmin[a] = c->posmin[a];
    A.qrange[a] = c->posmax[a] - c->posmin[a];
    A.vqmin[a] = c->vel_min[a];
    A.vqrange[a] = c->vel_range[a];
  }
  A.vsum = with_vel ? P<unsigned long long>(c->vsum) : nullptr;
  if (with_vel) CHK(ensure(c, c->wvs, (size_t)(room * W * W * W * 3 + 1) * 8));
  A.wvs = with_vel ? P<unsigned long long>(c->wvs) : nullptr;

  {
    TimedRegion tr(c, "count");
    bin_count(A, nmap, c->stream);
  }
  // subcell offsets: block sums from the totals pass, their scan, the block scans in the bases pass
  const int64_t nblk = (nsc + 255) / 256;
  CHK(ensure(c, c->tile_buf, (size_t)(nblk + 16) * 4));
  cell_totals(A, nsc, P<int>(c->sc_cnt), P<int>(c->tile_buf), c->stream);
  scan_tile_sums(P<int>(c->tile_buf), nblk, P<int>(c->meta) + M_TOTAL, c->stream);  // stores the total
  cell_bases(A, nsc, P<int>(c->sc_cnt), P<int>(c->tile_buf), P<int>(c->sc_off), c->stream);
  const bool keep_clr = c->clr != nullptr;  // gvdbInsertSubcell stores colours when bound (P.cuh:268)
  Buf& lpos = fp16 ? c->list_q16 : c->list_pos;
  Buf& lvel = fp16 ? c->list_vq16 : c->list_vel;
  const size_t ebytes = fp16 ? 6 : 12;  // per list entry: float3 or ushort3
  // Lists sized for the largest possible length when that bound fits the list budget (and lies
  // below 2^31, so the 32-bit offsets cannot wrap), so the insert is enqueued before the single
  // length read-back (and a recorded insert needs none); otherwise the length is read and checked
  // before anything is inserted.  From a bound of 2^32 on, the 32-bit scan total can wrap, so the
  // length comes from a 64-bit sum.
  const int64_t per_axis = subcells_per_axis(c, r, sc);
  const int64_t bound = per_axis * per_axis * per_axis * n;
  const size_t list_room = lpos.cap > c->list_budget ? lpos.cap : c->list_budget;
  const bool early = bound <= INT32_MAX && (size_t)(bound + 1) * ebytes <= list_room;
  if (c->recording && !early)
    return fail(c, GVDBX_ERR_CAPACITY, "the list bound (%lld entries) does not fit a recorded insert",
                (long long)bound);
  int64_t ln = -1;
  if (early) {
    CHK(ensure(c, lpos, (size_t)(bound + 1) * ebytes));
    if (keep_vel_lists) CHK(ensure(c, lvel, (size_t)(bound + 1) * ebytes));
    if (keep_clr) CHK(ensure(c, c->list_clr, (size_t)(bound + 1) * 4));
  } else {
    const bool wide = bound >= ((int64_t)1 << 32);
    if (wide) {  // the 64-bit sum accumulates into a cleared counter
      fill_u32(P<uint32_t>(c->meta) + M_WIDE, 2, 0u, c->stream);
      sum_wide(P<int>(c->sc_cnt), nsc, (unsigned long long*)(P<int>(c->meta) + M_WIDE), c->stream);
    }
    CHK(read_meta(c));
    if (wide) {
      unsigned long long w;
      memcpy(&w, c->h_meta + M_WIDE, sizeof(w));
      ln = (int64_t)w;
    } else {
      ln = (uint32_t)c->h_meta[M_TOTAL];
    }
    if (ln > INT32_MAX || ln > bound) return fail(c, GVDBX_ERR_CAPACITY, "subcell list length %lld", (long long)ln);
    CHK(ensure(c, lpos, (size_t)(ln + 1) * ebytes));
    if (keep_vel_lists) CHK(ensure(c, lvel, (size_t)(ln + 1) * ebytes));
    if (keep_clr) CHK(ensure(c, c->list_clr, (size_t)(ln + 1) * 4));
  }
  A.list_pos = fp16 ? nullptr : P<float>(c->list_pos);
  A.list_vel = (keep_vel_lists && !fp16) ? P<float>(c->list_vel) : nullptr;
  A.list_q16 = fp16 ? P<uint16_t>(c->list_q16) : nullptr;
  A.list_vq16 = (keep_vel_lists && fp16) ? P<uint16_t>(c->list_vq16) : nullptr;
  A.list_clr = keep_clr ? P<uint32_t>(c->list_clr) : nullptr;
  if (c->recording) {  // the length stays on the device, where later recorded calls leave it alone
    copy_u32((const uint32_t*)(P<int>(c->meta) + M_TOTAL), P<uint32_t>(c->graph_len), 1, c->stream);
  } else if (early) {
    CHK(read_meta_begin(c));  // the length is final after the scan
  }
  {
    TimedRegion tr(c, "insert");
    bin_insert(A, nmap, c->stream);
  }
  if (early && !c->recording) {
    CHK(read_meta_end(c));  // returns while the insert kernels still run
    ln = (uint32_t)c->h_meta[M_TOTAL];
  }
  if (!c->recording && (ln > INT32_MAX || ln > bound))
    return fail(c, GVDBX_ERR_CAPACITY, "subcell list length %lld", (long long)ln);
  CHK(launch_status(c, "InsertPointsSubcell"));
  c->sc = sc;
  c->scdim = cfg.D;
  c->nsc = nsc;
  c->sclen = c->recording ? -1 : ln;
  c->sclen_dev = c->recording;
  if (c->recording) c->graph_insert = true;
  else c->sclen_hint = ln;
  c->lists_have_vel = keep_vel_lists;
  c->lists_have_clr = keep_clr;
  c->lists_fp16 = fp16;
  c->have_vsum = with_vel;
  for (int a = 0; a < 3; ++a) {
    c->q_min[a] = A.qmin[a];
    c->q_range[a] = A.qrange[a];
  }
  if (len) *len = ln;
  return GVDBX_OK;
}

int gvdbx_insert_points_subcell(gvdbx_ctx* c, int sc, int64_t n, float r, const float trans[3], int64_t* len) {
  return insert_impl(c, sc, n, r, trans, len, false);
}

int gvdbx_insert_points_subcell_fp16(gvdbx_ctx* c, int sc, int64_t n, float r, const float trans[3], int64_t* len) {
  return insert_impl(c, sc, n, r, trans, len, true);
}

// GetMinMaxVel (V.cpp:5470-5522): the velocity box of the first n bound points, kept for the
// FP16 insert's velocity quantization (min, and range = max - min); out[6] = min xyz, max xyz.
int gvdbx_get_min_max_vel(gvdbx_ctx* c, int64_t n, float out[6]) {
  CHK(check_points(c, n));
  if (!c->vel) return fail(c, GVDBX_ERR_STATE, "no velocities bound (gvdbx_set_points)");
  CHK(ensure(c, c->bbox_part, (size_t)kBBoxParts * 6 * sizeof(float)));
  PointsView pv = points_of(c, n, nullptr);
  pv.pos = c->vel;
  pv.pos_off = c->vel_off;
  pv.pos_stride = c->vel_stride;
  CHK(zero_meta(c));
  bbox_two_pass(pv, P<float>(c->bbox_part), P<int>(c->meta), c->stream);
  CHK(read_meta(c));
  for (int a = 0; a < 3; ++a) {
    const float mn = ordered_to_float(c->h_meta[M_BBOX + a]), mx = ordered_to_float(c->h_meta[M_BBOX + 3 + a]);
    c->vel_min[a] = mn;
    c->vel_range[a] = mx - mn;
    if (out) {
      out[a] = mn;
      out[3 + a] = mx;
    }
  }
  c->vel_box = true;
  return GVDBX_OK;
}

int gvdbx_set_subcell_options(gvdbx_ctx* c, int store_velocity_lists) {
  if (!c) return GVDBX_ERR_ARG;
  c->store_vel_lists = store_velocity_lists != 0;
  return GVDBX_OK;
}

static int gather_impl(gvdbx_ctx* c, float r, int chan, int accum, int mode) {
  CHK(gather_checks(c, chan));
  CHK(float_lists(c));
  if (c->nsc == 0) return GVDBX_OK;
  const SubcellCfg cfg = subcell_cfg(c, c->sc, r);
  TimedRegion tr(c, "gather");
  const int br = cfg.res0 + 2 * c->apron;
  if (mode == 0 && !accum && cfg.res0 <= 8 && br * br * br <= kMaxPaddedVox) {
    density_brick(cfg, (int)c->used0, c->nsc, c->sclen >= 0 ? c->sclen : c->sclen_hint, P<int>(c->map_leaf),
                  P<int>(c->sc_off), P<float>(c->list_pos), P<int>(c->leaf_slot), P<gvdbx_node>(c->nodes[0]),
                  c->sc_map_current ? nullptr : P<int4>(c->slot_pos), c->apron,
                  c->apron > 0 ? P<int>(c->nbr) : nullptr, P<float>(c->chan[chan]), nullptr, launch_of(c),
                  c->stream);
    c->nbr_apron_ok[chan] = c->bound_ok[chan] = c->apron > 0;
    c->bound_val[chan] = 0.0f;
  } else {
    if (mode == 1)
      levelset_field(cfg, c->nsc, P<int>(c->sc_pos), P<int>(c->sc_nid), P<int>(c->sc_cnt), P<int>(c->sc_off),
                     P<float>(c->list_pos), P<int>(c->leaf_slot), c->apron, P<float>(c->chan[chan]), accum, 3.0f,
                     c->stream);
    else
      gather_field(cfg, c->nsc, P<int>(c->sc_pos), P<int>(c->sc_nid), P<int>(c->sc_cnt), P<int>(c->sc_off),
                   P<float>(c->list_pos), P<int>(c->leaf_slot), c->apron, P<float>(c->chan[chan]), accum, mode, 0.0f,
                   c->stream);
    c->nbr_apron_ok[chan] = c->bound_ok[chan] = false;
  }
  return launch_status(c, "Gather");
}

// The colour part of GatherDensity / GatherLevelSet (P.cuh:713-731): runs when the insert stored
// colours (a colour attribute was bound) and a colour channel is given.
static int gather_colour(gvdbx_ctx* c, int chan_clr, int accum) {
  if (chan_clr < 0 || !c->lists_have_clr || c->nsc == 0) return GVDBX_OK;
  const SubcellCfg cfg = subcell_cfg(c, c->sc, 0.f);
  TimedRegion tr(c, "colour");
  CHK(ensure(c, c->csum, (size_t)(c->nsc * 4 + 1) * 8));
  CHK(ensure(c, c->cmean, (size_t)(c->nsc + 1) * 4));
  colour_sums(c->nsc, P<int>(c->sc_off), P<int>(c->sc_cnt), P<uint32_t>(c->list_clr),
              P<unsigned long long>(c->csum), P<uint32_t>(c->cmean), c->stream);
  const bool with_apron = c->apron > 0 && !accum;
  colour_brick(cfg, (int)c->used0, P<int>(c->map_leaf), P<int>(c->sc_cnt), P<unsigned long long>(c->csum),
               P<uint32_t>(c->cmean), P<int>(c->leaf_slot), c->apron, with_apron ? P<int>(c->nbr) : nullptr,
               P<uint32_t>(c->chan[chan_clr]), accum, c->stream);
  c->nbr_apron_ok[chan_clr] = c->bound_ok[chan_clr] = with_apron;  // 0 where no neighbour exists
  c->bound_val[chan_clr] = 0.0f;
  return launch_status(c, "Gather colour");
}

static int gather_with_colour(gvdbx_ctx* c, int sc, float r, int chan, int chan_clr, int accum, int mode) {
  if (!c) return GVDBX_ERR_ARG;
  if (sc != c->sc) return fail(c, GVDBX_ERR_ARG, "subcell size %d differs from the insert (%d)", sc, c->sc);
  if (chan_clr >= 0) {
    CHK(chan_clr_ok(c, chan_clr));
    if (chan_clr == chan) return fail(c, GVDBX_ERR_ARG, "field and colour channels must differ");
  }
  CHK(gather_impl(c, r, chan, accum, mode));
  return gather_colour(c, chan_clr, accum);
}

int gvdbx_gather_density(gvdbx_ctx* c, int sc, int64_t n, float r, const float trans[3], int64_t len, int chan,
                         int chan_clr, int accum) {
  (void)n; (void)trans; (void)len;
  return gather_with_colour(c, sc, r, chan, chan_clr, accum, 0);
}

int gvdbx_gather_levelset(gvdbx_ctx* c, int sc, int64_t n, float r, const float trans[3], int64_t len, int chan,
                          int chan_clr, int accum) {
  (void)n; (void)trans; (void)len;
  return gather_with_colour(c, sc, r, chan, chan_clr, accum, 1);
}

// GatherLevelSet_FP16 (V.cpp:5375-5401, kernel P.cuh:775-811): the level set from the FP16 lists,
// min(3, |p - v| - r) per voxel with p decoded from the list's ushort3; no accumulation.  The
// reference's colour part reads an uninitialised accumulator: not built (chan_clr must be < 0).
int gvdbx_gather_levelset_fp16(gvdbx_ctx* c, int sc, int64_t n, float r, const float trans[3], int64_t len, int chan,
                               int chan_clr) {
  (void)n; (void)trans; (void)len;
  if (!c) return GVDBX_ERR_ARG;
  if (sc != c->sc) return fail(c, GVDBX_ERR_ARG, "subcell size %d differs from the insert (%d)", sc, c->sc);
  CHK(gather_checks(c, chan));
  if (!c->lists_fp16) return fail(c, GVDBX_ERR_STATE, "GatherLevelSet_FP16 needs InsertPointsSubcell_FP16 lists");
  if (chan_clr >= 0) return fail(c, GVDBX_ERR_ARG, "the FP16 level set takes no colour channel");
  if (c->nsc == 0) return GVDBX_OK;
  const SubcellCfg cfg = subcell_cfg(c, c->sc, r);
  TimedRegion tr(c, "gather");
  levelset_field_q16(cfg, c->nsc, P<int>(c->sc_pos), P<int>(c->sc_nid), P<int>(c->sc_cnt), P<int>(c->sc_off),
                     P<uint16_t>(c->list_q16), c->q_min, c->q_range, P<int>(c->leaf_slot), c->apron,
                     P<float>(c->chan[chan]), 3.0f, c->stream);
  c->nbr_apron_ok[chan] = c->bound_ok[chan] = false;
  return launch_status(c, "GatherLevelSet_FP16");
}

int gvdbx_gather_velocity(gvdbx_ctx* c, int cx, int cy, int cz, int accum) {
  CHK(gather_checks(c, cx));
  CHK(chan_ok(c, cy));
  CHK(chan_ok(c, cz));
  if (!c->have_vsum) return fail(c, GVDBX_ERR_STATE, "no velocities were bound at InsertPointsSubcell");
  if (c->nsc == 0) return GVDBX_OK;
  const SubcellCfg cfg = subcell_cfg(c, c->sc, 0.f);
  TimedRegion tr(c, "velocity");
  bool apron_done = false;
  if ((int64_t)cfg.D * cfg.D * cfg.D <= 512) {
    const bool with_apron = c->apron > 0 && !accum;
    velocity_brick(cfg, (int)c->used0, P<int>(c->map_leaf), P<int>(c->sc_cnt), P<unsigned long long>(c->vsum),
                   P<int>(c->leaf_slot), c->apron, with_apron ? P<int>(c->nbr) : nullptr, P<float>(c->chan[cx]),
                   P<float>(c->chan[cy]), P<float>(c->chan[cz]), accum, c->stream);
    apron_done = with_apron;
  } else {
    velocity_subcell(cfg, c->nsc, P<int>(c->sc_pos), P<int>(c->sc_nid), P<int>(c->sc_cnt),
                     P<unsigned long long>(c->vsum), P<int>(c->leaf_slot), c->apron, P<float>(c->chan[cx]),
                     P<float>(c->chan[cy]), P<float>(c->chan[cz]), accum, c->stream);
  }
  c->nbr_apron_ok[cx] = c->nbr_apron_ok[cy] = c->nbr_apron_ok[cz] = apron_done;
  c->bound_ok[cx] = c->bound_ok[cy] = c->bound_ok[cz] = apron_done;  // velocity_brick writes 0 there
  c->bound_val[cx] = c->bound_val[cy] = c->bound_val[cz] = 0.0f;
  return launch_status(c, "GatherVelocity");
}

int gvdbx_gather_density_velocity(gvdbx_ctx* c, int sc, int64_t n, float r, const float trans[3], int64_t len,
                                  int chan, int cx, int cy, int cz) {
  (void)n; (void)trans; (void)len;
  if (!c) return GVDBX_ERR_ARG;
  if (sc != c->sc) return fail(c, GVDBX_ERR_ARG, "subcell size %d differs from the insert (%d)", sc, c->sc);
  CHK(gather_checks(c, chan));
  CHK(chan_ok(c, cx));
  CHK(chan_ok(c, cy));
  CHK(chan_ok(c, cz));
  if (chan == cx || chan == cy || chan == cz || cx == cy || cx == cz || cy == cz)
    return fail(c, GVDBX_ERR_ARG, "density and velocity channels must be four different channels");
  if (!c->have_vsum) return fail(c, GVDBX_ERR_STATE, "no velocities were bound at InsertPointsSubcell");
  CHK(float_lists(c));
  if (c->nsc == 0) return GVDBX_OK;
  const SubcellCfg cfg = subcell_cfg(c, c->sc, r);
  const int br = cfg.res0 + 2 * c->apron;
  const bool one_pass = c->apron > 0 && cfg.res0 <= 8 && br * br * br <= kMaxPaddedVox && cfg.D <= 8;
  if (!one_pass) {  // the two gathers one after the other give the same channels
    CHK(gather_impl(c, r, chan, 0, 0));
    return gvdbx_gather_velocity(c, cx, cy, cz, 0);
  }
  TimedRegion tr(c, "gather");
  const VelOut vo{P<int>(c->sc_cnt), P<unsigned long long>(c->vsum),
                  {P<float>(c->chan[cx]), P<float>(c->chan[cy]), P<float>(c->chan[cz])}};
  density_brick(cfg, (int)c->used0, c->nsc, c->sclen >= 0 ? c->sclen : c->sclen_hint, P<int>(c->map_leaf),
                P<int>(c->sc_off), P<float>(c->list_pos), P<int>(c->leaf_slot), P<gvdbx_node>(c->nodes[0]),
                c->sc_map_current ? nullptr : P<int4>(c->slot_pos), c->apron, P<int>(c->nbr),
                P<float>(c->chan[chan]), &vo, launch_of(c), c->stream);
  for (int ch : {chan, cx, cy, cz}) {
    c->nbr_apron_ok[ch] = c->bound_ok[ch] = true;
    c->bound_val[ch] = 0.0f;
  }
  return launch_status(c, "GatherDensityVelocity");
}

int gvdbx_scatter_reduce_levelset(gvdbx_ctx* c, float r, int chan, float init) {
  CHK(gather_checks(c, chan));
  CHK(float_lists(c));
  if (c->nsc == 0) return GVDBX_OK;
  const SubcellCfg cfg = subcell_cfg(c, c->sc, r);
  TimedRegion tr(c, "scatter");
  // the insert left the (point, subcell) pairs sorted by subcell; each voxel min-reduces the
  // segment of its subcell, starting from init
  levelset_field(cfg, c->nsc, P<int>(c->sc_pos), P<int>(c->sc_nid), P<int>(c->sc_cnt), P<int>(c->sc_off),
                 P<float>(c->list_pos), P<int>(c->leaf_slot), c->apron, P<float>(c->chan[chan]), 0, init, c->stream);
  c->nbr_apron_ok[chan] = c->bound_ok[chan] = false;
  return launch_status(c, "ScatterReduceLevelSet");
}

// The boundary fill's apron voxel table for the context's brick geometry, uploaded once.
static int apron_table_current(gvdbx_ctx* c) {
  const int res0 = res_of(c, 0);
  if (c->apron_tab_res == res0 && c->apron_tab_apron == c->apron) return GVDBX_OK;
  NOT_RECORDING(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));  // an earlier upload may still read the host table
  apron_table(res0, c->apron, c->h_apron_tab);
  CHK(ensure(c, c->apron_tab, c->h_apron_tab.size() * 4 + 4));
  HIPCHK(c, hipMemcpyAsync(c->apron_tab.p, c->h_apron_tab.data(), c->h_apron_tab.size() * 4, hipMemcpyHostToDevice,
                           c->stream));
  c->apron_tab_res = res0;
  c->apron_tab_apron = c->apron;
  return GVDBX_OK;
}

int gvdbx_update_apron_channels(gvdbx_ctx* c, const int* chans, const float* boundvals, int nchan) {
  if (!c || !chans || nchan < 0) return GVDBX_ERR_ARG;  // UpdateApron (V.cpp:4411-4457)
  for (int i = 0; i < nchan; ++i) CHK(chan_any(c, chans[i]));
  CHK(atlas_current(c));
  if (c->apron == 0 || c->nslots == 0) return GVDBX_OK;
  CHK(apron_table_current(c));
  const uint32_t* tab = P<uint32_t>(c->apron_tab);
  const int ntab = (int)c->h_apron_tab.size();
  TimedRegion tr(c, "apron");
  // channels whose neighbour copies are current only need the boundary value; the others
  // take the full neighbour pull (UpdateApron<T>, O.cuh:72-126)
  for (int pass = 0; pass < 2; ++pass) {
    ApronChannels ac;
    ac.n = 0;
    for (int i = 0; i < nchan; ++i) {
      if (c->nbr_apron_ok[chans[i]] != (pass == 0)) continue;
      const float bv = bound_word(c, chans[i], boundvals ? boundvals[i] : 0.0f);
      if (pass == 0 && c->bound_ok[chans[i]] && __builtin_memcmp(&bv, &c->bound_val[chans[i]], 4) == 0)
        continue;  // neighbour copies and boundary value both current
      ac.ptr[ac.n] = P<float>(c->chan[chans[i]]);
      ac.bound[ac.n] = bv;
      if (++ac.n == kMaxApronChannels || i == nchan - 1) {
        if (pass == 0) apron_bound(c->nslots, P<int>(c->nbr), tab, ntab, res_of(c, 0), c->apron, ac, c->stream);
        else apron_multi(c->nslots, P<int>(c->nbr), res_of(c, 0), c->apron, ac, c->stream);
        ac.n = 0;
      }
    }
    if (ac.n) {
      if (pass == 0) apron_bound(c->nslots, P<int>(c->nbr), tab, ntab, res_of(c, 0), c->apron, ac, c->stream);
      else apron_multi(c->nslots, P<int>(c->nbr), res_of(c, 0), c->apron, ac, c->stream);
    }
  }
  for (int i = 0; i < nchan; ++i) {
    c->nbr_apron_ok[chans[i]] = c->bound_ok[chans[i]] = true;
    c->bound_val[chans[i]] = bound_word(c, chans[i], boundvals ? boundvals[i] : 0.0f);
  }
  return launch_status(c, "UpdateApron");
}

int gvdbx_update_apron(gvdbx_ctx* c, int chan, float boundval) {
  return gvdbx_update_apron_channels(c, &chan, &boundval, 1);
}

int gvdbx_sample_trilinear(gvdbx_ctx* c, int chan, const void* q, int64_t n, void* out) {
  CHK(chan_ok(c, chan));
  if (c->batch) return fail(c, GVDBX_ERR_STATE, "sampling takes world points: not for batched clouds");
  if (!q || !out || n < 0) return fail(c, GVDBX_ERR_ARG, "bad sample buffers");
  if (n > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "more than 2^31 queries");
  if (c->apron < 1) return fail(c, GVDBX_ERR_STATE, "trilinear sampling needs an apron >= 1");
  CHK(atlas_current(c));
  if (c->depth < 0) return fail(c, GVDBX_ERR_STATE, "no topology");
  TimedRegion tr(c, "sample");
  sample_trilinear(tree_of(c), P<int>(c->leaf_slot), res_of(c, 0), c->apron, P<float>(c->chan[chan]),
                   (const float*)q, n, (float*)out, c->stream);
  return launch_status(c, "SampleTrilinear");
}

namespace {

int field_checks(gvdbx_ctx* c, int chan, const void* q, int64_t n) {
  CHK(chan_ok(c, chan));
  if (c->batch) return fail(c, GVDBX_ERR_STATE, "splat and sampling take world points: not for batched clouds");
  if (!q || n < 0) return fail(c, GVDBX_ERR_ARG, "bad point buffers");
  if (n > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "more than 2^31 points");
  CHK(atlas_current(c));
  if (c->depth < 0) return fail(c, GVDBX_ERR_STATE, "no topology");
  return GVDBX_OK;
}

// Scratch of one splat: the sort's keys and values (both ping-pong halves), the radix tile
// counts, the key starts and the sorted (q, m) payload.
struct SplatWork {
  uint32_t* keys;
  uint32_t* keys_alt;
  uint32_t* vals;
  uint32_t* vals_alt;
  int* counts;
  int* start;
  float4* payload;
};

int splat_limits(gvdbx_ctx* c) {
  if (!splat_brick_supported(c->ld[0]))
    return fail(c, GVDBX_ERR_CAPACITY, "splat supports leaf bricks of 2^1..2^4 voxels");
  if (c->nslots * 27 + 2 > INT32_MAX) return fail(c, GVDBX_ERR_CAPACITY, "too many bricks for the splat keys");
  return GVDBX_OK;
}

// Trilinear splat of (q, mass) into the interior of chan (sorted pull, see splat.hip), with the
// sort's scratch in w (sized for n points and nslots * 27 + 1 keys); *perm receives the sorted
// position -> point permutation (inside w).
int splat_run(gvdbx_ctx* c, int chan, const float* q, const float* mass, int64_t n, int accum, const SplatWork& w,
              bool timed, const uint32_t** perm) {
  const uint32_t sentinel = (uint32_t)(c->nslots * 27);
  *perm = nullptr;
  {
    TimedRegion tr(c, "splat_sort", timed);
    const uint32_t* skeys = w.keys;
    if (n > 0) {
      int bits = 1;  // keys run up to the sentinel
      while (((int64_t)1 << bits) <= (int64_t)sentinel) ++bits;
      // the key launch also counts the first radix digit; the sort generates the point positions
      splat_keys(tree_of(c), q, n, P<int>(c->leaf_slot), c->ld[0], sentinel, w.keys, bits, w.counts, c->stream);
      const int par = radix_sort_u32_positions(w.keys, w.keys_alt, w.vals, w.vals_alt, n, bits, w.counts, c->stream);
      skeys = par ? w.keys_alt : w.keys;
      *perm = par ? w.vals_alt : w.vals;
      // the other value buffer takes the inverse permutation
      splat_place(q, mass, n, *perm, par ? w.vals : w.vals_alt, w.payload, c->stream);
    }
    key_starts(skeys, n, (int)sentinel, w.start, c->stream);  // start[0 .. sentinel + 1]
  }
  {
    TimedRegion tr(c, "splat_brick", timed);
    splat_brick(c->nslots, w.start, w.payload, P<int>(c->nbr), P<int4>(c->slot_pos), c->ld[0], c->apron,
                P<float>(c->chan[chan]), accum, c->stream);
  }
  c->nbr_apron_ok[chan] = c->bound_ok[chan] = false;
  return launch_status(c, "SplatTrilinear");
}

int splat_impl(gvdbx_ctx* c, int chan, const float* q, const float* mass, int64_t n, int accum) {
  CHK(splat_limits(c));
  if (c->nslots == 0) return GVDBX_OK;
  const int64_t nkeys = c->nslots * 27 + 1;
  CHK(ensure(c, c->sp_keys, (size_t)(n + 1) * 4));
  CHK(ensure(c, c->sp_keys_alt, (size_t)(n + 1) * 4));
  CHK(ensure(c, c->sp_vals, (size_t)(n + 1) * 4));
  CHK(ensure(c, c->sp_vals_alt, (size_t)(n + 1) * 4));
  CHK(ensure(c, c->sp_counts, (size_t)radix_count_ints(n) * 4));
  CHK(ensure(c, c->sp_start, (size_t)(nkeys + 1) * 4));
  CHK(ensure(c, c->sp_payload, (size_t)(n + 1) * 16));
  const SplatWork w{P<uint32_t>(c->sp_keys), P<uint32_t>(c->sp_keys_alt), P<uint32_t>(c->sp_vals),
                    P<uint32_t>(c->sp_vals_alt), P<int>(c->sp_counts), P<int>(c->sp_start), P<float4>(c->sp_payload)};
  const uint32_t* perm = nullptr;
  CHK(splat_run(c, chan, q, mass, n, accum, w, true, &perm));
  c->sp_order_q = n > 0 ? q : nullptr;
  c->sp_order_n = n;
  c->sp_order = perm;
  return GVDBX_OK;
}

}  // namespace

int gvdbx_splat_trilinear(gvdbx_ctx* c, int chan, const void* q, const void* mass, int64_t n, int accum) {
  CHK(field_checks(c, chan, q, n));
  TimedRegion tr(c, "splat");
  return splat_impl(c, chan, (const float*)q, (const float*)mass, n, accum);
}

// gvdbx_splat_trilinear with its scratch in this device's copy of `scratch`: no allocation,
// sizes capped by the array (kNoAllocPoints points, kNoAllocKeys splat keys).  Calls that overlap
// on one GPU (any context, any stream) race on that copy.
int gvdbx_splat_trilinear_noalloc(gvdbx_ctx* c, int chan, const void* q, const void* mass, int64_t n, int accum) {
  CHK(field_checks(c, chan, q, n));
  CHK(splat_limits(c));
  if (n > kNoAllocPoints)
    return fail(c, GVDBX_ERR_CAPACITY, "SplatTrilinear without allocation takes at most %lld points per call (%lld given)",
                (long long)kNoAllocPoints, (long long)n);
  if (c->nslots * 27 + 1 > kNoAllocKeys)
    return fail(c, GVDBX_ERR_CAPACITY, "SplatTrilinear without allocation takes at most %lld bricks (%lld active)",
                (long long)((kNoAllocKeys - 1) / 27), (long long)c->nslots);
  if (!c->scratch)
    return fail(c, GVDBX_ERR_HIP, "no scratch array on device %d: %s", c->device, hipGetErrorString(c->scratch_err));
  if (c->nslots == 0) return GVDBX_OK;
  unsigned char* base = (unsigned char*)c->scratch;
  const SplatWork w{(uint32_t*)base, (uint32_t*)(base + kScrWord), (uint32_t*)(base + 2 * kScrWord),
                    (uint32_t*)(base + 3 * kScrWord), (int*)(base + 4 * kScrWord),
                    (int*)(base + 4 * kScrWord + kScrCounts),
                    (float4*)(base + 4 * kScrWord + kScrCounts + kScrStart)};
  const uint32_t* perm = nullptr;
  CHK(splat_run(c, chan, (const float*)q, (const float*)mass, n, accum, w, false, &perm));
  // the permutation lives in the shared scratch: the backward does not reuse it
  c->sp_order_q = nullptr;
  c->sp_order_n = 0;
  c->sp_order = nullptr;
  return GVDBX_OK;
}

int gvdbx_splat_backward(gvdbx_ctx* c, int chan_grad, const void* q, const void* mass, int64_t n, void* dq,
                         void* dmass) {
  CHK(field_checks(c, chan_grad, q, n));
  if (!splat_brick_supported(c->ld[0]))
    return fail(c, GVDBX_ERR_CAPACITY, "splat supports leaf bricks of 2^1..2^4 voxels");
  TimedRegion tr(c, "splat_backward");
  const uint32_t* order = (c->sp_order && c->sp_order_q == q && c->sp_order_n == n) ? c->sp_order : nullptr;
  splat_backward(tree_of(c), P<int>(c->leaf_slot), P<int>(c->nbr), c->ld[0], c->apron, P<float>(c->chan[chan_grad]),
                 (const float*)q, (const float*)mass, n, order, (float*)dq, (float*)dmass, c->stream);
  return launch_status(c, "SplatBackward");
}

int gvdbx_sample_trilinear_backward(gvdbx_ctx* c, int chan, const void* q, const void* dout, int64_t n, void* dq,
                                    int chan_grad, int accum_grad) {
  CHK(field_checks(c, chan, q, n));
  if (!dout) return fail(c, GVDBX_ERR_ARG, "dout is null");
  if (c->apron < 1) return fail(c, GVDBX_ERR_STATE, "trilinear sampling needs an apron >= 1");
  if (chan_grad >= 0) CHK(chan_ok(c, chan_grad));
  TimedRegion tr(c, "sample_backward");
  if (dq)
    sample_backward(tree_of(c), P<int>(c->leaf_slot), res_of(c, 0), c->apron, P<float>(c->chan[chan]),
                    (const float*)q, (const float*)dout, n, (float*)dq, c->stream);
  if (chan_grad >= 0) CHK(splat_impl(c, chan_grad, (const float*)q, (const float*)dout, n, accum_grad));
  return launch_status(c, "SampleTrilinearBackward");
}

int gvdbx_scatter_density(gvdbx_ctx* c, int64_t n, float radius, float amp, const float trans[3], int expand,
                          int chan, int avg_colour, int chan_clr) {
  (void)amp;  // unused by the reference kernel as well (P.cuh:525-573)
  CHK(chan_ok(c, chan));  // ScatterDensity (V.cpp:5625-5659)
  const int cmode = (c->clr && chan_clr >= 0) ? (avg_colour ? 1 : 2) : 0;
  if (cmode) {
    CHK(chan_clr_ok(c, chan_clr));
    if (chan_clr == chan) return fail(c, GVDBX_ERR_ARG, "density and colour channels must differ");
  }
  if (c->batch) return fail(c, GVDBX_ERR_STATE, "ScatterDensity takes world points: not for batched clouds");
  CHK(check_points(c, n));
  if (!(radius > 0.0f) || !(radius < 1.0e30f)) return fail(c, GVDBX_ERR_ARG, "radius %g out of range", (double)radius);
  CHK(atlas_current(c));
  if (c->depth < 0) return fail(c, GVDBX_ERR_STATE, "no topology");
  if (c->nslots == 0 || n == 0) return GVDBX_OK;
  // Exact sums need every voxel's total within 2^62 at the scale 2^shift: a voxel takes at most
  // one contribution per point, each |w| <= max(1, (cmax - 1)^3) with cmax = (expand ? 6 : 3)/r^2
  // (|v - p'|^2 < 3 for the point's voxel, < 6 for a face neighbour).  A bound above 2^40 would
  // leave less than 2^-22 of resolution: refused.
  const double cmax = (expand ? 6.0 : 3.0) / ((double)radius * (double)radius);
  const double wmax = cmax > 2.0 ? (cmax - 1.0) * (cmax - 1.0) * (cmax - 1.0) : 1.0;
  const double bound = (double)n * wmax * 1.01 + 1.0;
  if (bound > std::ldexp(1.0, 40))
    return fail(c, GVDBX_ERR_CAPACITY, "ScatterDensity: %lld points at radius %g exceed the exact-sum range",
                (long long)n, (double)radius);
  int shift = 62 - (int)std::ceil(std::log2(bound));
  if (shift > 48) shift = 48;
  const int br = res_of(c, 0) + 2 * c->apron;
  const int64_t nvox = c->nslots * (int64_t)br * br * br;
  CHK(ensure(c, c->scatter_acc, (size_t)nvox * 8));
  const int64_t nin = c->nslots << (3 * c->ld[0]);
  if (cmode) CHK(ensure(c, c->scatter_clr, (size_t)nin * (cmode == 1 ? 32 : 4) + 8));
  TimedRegion tr(c, "scatter_density");
  scatter_density(tree_of(c), points_of(c, n, trans), P<int>(c->leaf_slot), c->ld[0], c->apron, radius, expand,
                  shift, c->nslots, P<unsigned long long>(c->scatter_acc), P<float>(c->chan[chan]), cmode,
                  cmode ? c->scatter_clr.p : nullptr, cmode ? P<uint32_t>(c->chan[chan_clr]) : nullptr, c->stream);
  c->nbr_apron_ok[chan] = c->bound_ok[chan] = false;
  if (cmode) c->nbr_apron_ok[chan_clr] = c->bound_ok[chan_clr] = false;
  return launch_status(c, "ScatterDensity");
}

int gvdbx_get_depth(const gvdbx_ctx* c, int* depth, int32_t rootpos[3], float posmin[3], float posmax[3]) {
  if (!c) return GVDBX_ERR_ARG;
  if (depth) *depth = c->depth;
  for (int a = 0; a < 3; ++a) {
    if (rootpos) rootpos[a] = c->root[a];
    if (posmin) posmin[a] = c->posmin[a];
    if (posmax) posmax[a] = c->posmax[a];
  }
  return GVDBX_OK;
}

int gvdbx_get_node_count(const gvdbx_ctx* c, int lev, int64_t* total, int64_t* used) {
  if (!c || lev < 0 || lev >= kMaxLev) return GVDBX_ERR_ARG;
  if (total) *total = lev < c->levs ? c->cnt[lev] : 0;
  if (used) *used = lev == 0 ? c->used0 : (lev < c->levs ? c->cnt[lev] : 0);
  return GVDBX_OK;
}

static int copy_out(gvdbx_ctx* c, void* dst, const void* src, size_t bytes) {
  NOT_RECORDING(c);
  if (bytes == 0) return GVDBX_OK;
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GVDBX_OK;
}

int gvdbx_copy_nodes(gvdbx_ctx* c, int lev, gvdbx_node* out, int64_t max_nodes) {
  if (!c || !out || lev < 0 || lev >= c->levs) return GVDBX_ERR_ARG;
  const int64_t n = c->cnt[lev] < max_nodes ? c->cnt[lev] : max_nodes;
  return copy_out(c, out, c->nodes[lev].p, (size_t)n * sizeof(gvdbx_node));
}

int gvdbx_copy_childlists(gvdbx_ctx* c, int lev, uint64_t* out, int64_t max_entries) {
  if (!c || !out || lev < 1 || lev >= c->levs) return GVDBX_ERR_ARG;
  int64_t n = c->cnt[lev] * clist_per(c, lev);
  if (n > max_entries) n = max_entries;
  return copy_out(c, out, c->clist[lev].p, (size_t)n * 8);
}

int gvdbx_get_num_slots(const gvdbx_ctx* c, int64_t* slots, int* brick_res) {
  if (!c) return GVDBX_ERR_ARG;
  if (slots) *slots = c->nslots;
  if (brick_res) *brick_res = c->configured ? res_of(c, 0) + 2 * (c->apron < 0 ? 0 : c->apron) : 0;
  return GVDBX_OK;
}

int gvdbx_copy_channel(gvdbx_ctx* c, int chan, float* out, int64_t max_floats) {
  CHK(chan_any(c, chan));
  if (!out) return GVDBX_ERR_ARG;
  int64_t n = c->nslots * brick_floats(c);
  if (n > max_floats) n = max_floats;
  return copy_out(c, out, c->chan[chan].p, (size_t)n * 4);
}

int gvdbx_get_channel_ptr(gvdbx_ctx* c, int chan, void** p, int64_t* floats) {
  CHK(chan_any(c, chan));
  c->nbr_apron_ok[chan] = c->bound_ok[chan] = false;  // the caller may write through the pointer
  if (p) *p = c->chan[chan].p;
  if (floats) *floats = c->nslots * brick_floats(c);
  return GVDBX_OK;
}

// The list length of a launched graph's insert, read back on request.
static int resolve_sclen(gvdbx_ctx* c) {
  if (!c->sclen_dev) return GVDBX_OK;
  NOT_RECORDING(c);
  uint32_t v = 0;
  HIPCHK(c, hipMemcpyAsync(&v, c->graph_len.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sclen = (int64_t)v;
  c->sclen_hint = c->sclen;
  c->sclen_dev = false;
  return GVDBX_OK;
}

int gvdbx_get_subcell_info(gvdbx_ctx* c, int64_t* nsc, int64_t* len, int* sc) {
  if (!c) return GVDBX_ERR_ARG;
  if (len) CHK(resolve_sclen(c));
  if (nsc) *nsc = c->nsc;
  if (len) *len = c->sclen;
  if (sc) *sc = c->sc;
  return GVDBX_OK;
}

int gvdbx_copy_subcell_array(gvdbx_ctx* c, int which, void* out, int64_t max_bytes) {
  if (!c || !out) return GVDBX_ERR_ARG;
  CHK(resolve_sclen(c));
  const void* src = nullptr;
  int64_t bytes = 0;
  switch (which) {
    case 0:
      if (!c->sc_map_current && c->cnt[0] > 0) {  // the insert used the slots: make the full map
        CHK(ensure(c, c->flags, (size_t)(c->cnt[0] + 1) * 4));
        CHK(ensure(c, c->sc_map, (size_t)(c->cnt[0] + 1) * 4));
        CHK(ensure(c, c->tile_buf, (size_t)(scan_tile_count(c->cnt[0] + 1) + 16) * 4));
        leaf_flags(tree_of(c), c->cnt[0], 1, P<int>(c->flags), c->stream);
        exclusive_scan(P<int>(c->flags), P<int>(c->sc_map), c->cnt[0], P<int>(c->tile_buf), nullptr, c->stream);
        c->sc_map_current = true;
      }
      src = c->sc_map.p;
      bytes = c->cnt[0] * 4;
      break;
    case 1: src = c->sc_cnt.p; bytes = c->nsc * 4; break;
    case 2: src = c->sc_off.p; bytes = c->nsc * 4; break;
    case 3: src = c->sc_pos.p; bytes = c->nsc * 12; break;
    case 4: src = c->sc_nid.p; bytes = c->nsc * 4; break;
    case 5: src = c->lists_fp16 ? nullptr : c->list_pos.p; bytes = c->sclen * 12; break;
    case 6: src = (c->lists_have_vel && !c->lists_fp16) ? c->list_vel.p : nullptr; bytes = c->sclen * 12; break;
    case 7: src = c->lists_have_clr ? c->list_clr.p : nullptr; bytes = c->lists_have_clr ? c->sclen * 4 : 0; break;
    case 8: src = c->lists_fp16 ? c->list_q16.p : nullptr; bytes = c->sclen * 6; break;
    case 9: src = (c->lists_have_vel && c->lists_fp16) ? c->list_vq16.p : nullptr; bytes = c->sclen * 6; break;
    default: return fail(c, GVDBX_ERR_ARG, "unknown subcell array %d", which);
  }
  if (bytes > max_bytes) bytes = max_bytes;
  if (!src || bytes <= 0) return GVDBX_OK;
  return copy_out(c, out, src, (size_t)bytes);
}

// SaveVBX (V.cpp:1626-1768): pools and channels copied to the host and written by the codec
// (vbx.cpp).  The channels must be laid out for the current topology.
int gvdbx_save_vbx(gvdbx_ctx* c, const char* path) {
  if (!c || !c->configured || !path) return fail(c, GVDBX_ERR_ARG, "context not configured or no path");
  int nchan = 0;
  while (nchan < 32 && c->chan_on[nchan]) ++nchan;
  for (int k = nchan; k < 32; ++k)
    if (c->chan_on[k]) return fail(c, GVDBX_ERR_ARG, "VBX files hold channels 0 .. n-1; channel %d follows a gap", k);
  if (nchan > 0) CHK(atlas_current(c));
  gvdbx_vbx_desc d;
  memset(&d, 0, sizeof(d));
  d.levels = c->levs;
  for (int l = 0; l < c->levs; ++l) d.log2dim[l] = c->ld[l];
  d.apron = c->apron < 0 ? 0 : c->apron;
  d.axis_count[0] = c->axis_cx;
  d.axis_count[1] = c->axis_cy;
  d.axis_count[2] = c->atlas_cz > 0 ? c->atlas_cz : c->axis_cz;
  d.nchan = nchan;
  for (int k = 0; k < nchan; ++k) d.chan_type[k] = c->chan_type[k];
  d.transform[6] = d.transform[7] = d.transform[8] = 1.0f;  // identity grid transform (V.cpp:77)
  d.root = c->depth >= 0 && c->cnt[c->depth] > 0 ? (uint64_t)c->depth << 8 : kUndefL;  // Elem(0, depth, 0)
  d.slots = nchan > 0 ? c->nslots : 0;
  std::vector<std::vector<gvdbx_node>> nodes(c->levs);
  std::vector<std::vector<uint64_t>> lists(c->levs);
  std::vector<std::vector<uint32_t>> chans(nchan);
  for (int l = 0; l < c->levs; ++l) {
    d.node_count[l] = c->cnt[l];
    d.clist_count[l] = l > 0 ? c->cnt[l] : 0;
    nodes[l].resize((size_t)c->cnt[l]);
    d.nodes[l] = nodes[l].data();
    CHK(copy_out(c, nodes[l].data(), c->nodes[l].p, nodes[l].size() * sizeof(gvdbx_node)));
    if (l > 0) {
      lists[l].resize((size_t)(c->cnt[l] * clist_per(c, l)));
      d.clists[l] = lists[l].data();
      CHK(copy_out(c, lists[l].data(), c->clist[l].p, lists[l].size() * 8));
    }
  }
  for (int k = 0; k < nchan; ++k) {
    chans[k].resize((size_t)(d.slots * brick_floats(c)));
    d.chan_data[k] = chans[k].data();
    CHK(copy_out(c, chans[k].data(), c->chan[k].p, chans[k].size() * 4));
  }
  const int st = gvdbx_vbx_write(path, &d);
  if (st != GVDBX_OK) return fail(c, st, "SaveVBX: cannot write %s", path);
  return GVDBX_OK;
}

// LoadVBX (V.cpp:507-685): the file's first grid replaces the tree and the channels.
int gvdbx_load_vbx(gvdbx_ctx* c, const char* path) {
  if (!c || !path) return fail(c, GVDBX_ERR_ARG, "no context or path");
  gvdbx_vbx_desc d;
  int st = gvdbx_vbx_probe(path, &d);
  if (st != GVDBX_OK) return fail(c, st, "LoadVBX: %s is not a VBX file this library reads", path);
  if (d.levels < 2) return fail(c, GVDBX_ERR_ARG, "LoadVBX: %d levels (at least 2 needed)", d.levels);
  const int depth = d.root == kUndefL ? -1 : (int)((d.root >> 8) & 0xFF);
  if (depth >= d.levels || (depth >= 0 && d.node_count[depth] < 1))
    return fail(c, GVDBX_ERR_ARG, "LoadVBX: root id %llx does not name a stored node", (unsigned long long)d.root);
  for (int l = 1; l < d.levels; ++l)
    if (d.clist_count[l] != d.node_count[l])
      return fail(c, GVDBX_ERR_ARG, "LoadVBX: level %d holds %lld child lists for %lld nodes", l,
                  (long long)d.clist_count[l], (long long)d.node_count[l]);
  const int64_t br = ((int64_t)1 << d.log2dim[0]) + 2 * d.apron;
  std::vector<std::vector<gvdbx_node>> nodes(d.levels);
  std::vector<std::vector<uint64_t>> lists(d.levels);
  std::vector<std::vector<uint32_t>> chans(d.nchan);
  for (int l = 0; l < d.levels; ++l) {
    nodes[l].resize((size_t)d.node_count[l]);
    d.nodes[l] = nodes[l].data();
    if (l > 0) {
      lists[l].resize((size_t)(d.clist_count[l] << (3 * d.log2dim[l])));
      d.clists[l] = lists[l].data();
    }
  }
  for (int k = 0; k < d.nchan; ++k) {
    chans[k].resize((size_t)(d.slots * br * br * br));
    d.chan_data[k] = chans[k].data();
  }
  st = gvdbx_vbx_read(path, &d);
  if (st != GVDBX_OK) return fail(c, st, "LoadVBX: cannot read %s", path);
  // child list i must belong to node i (the layout every descent here assumes)
  for (int l = 1; l < d.levels; ++l)
    for (int64_t i = 0; i < d.node_count[l]; ++i)
      if ((nodes[l][i].childlist >> 16) != (uint64_t)i || ((nodes[l][i].childlist >> 8) & 0xFF) != (uint64_t)l)
        return fail(c, GVDBX_ERR_ARG, "LoadVBX: node %lld of level %d does not own child list %lld", (long long)i, l,
                    (long long)i);
  // DestroyChannels + Configure (Clear): the context starts over with the file's geometry
  for (int k = 0; k < 32; ++k) c->chan_on[k] = false;
  c->apron = -1;
  c->chan_slot_cap = 0;
  CHK(gvdbx_configure(c, d.levels, d.log2dim));
  c->axis_cx = d.axis_count[0];
  c->axis_cy = d.axis_count[1];
  c->axis_cz = 1;
  for (int k = 0; k < d.nchan; ++k) CHK(gvdbx_add_channel(c, k, d.chan_type[k], d.apron));
  c->atlas_cz = d.axis_count[2];
  for (int l = 0; l < d.levels; ++l) {
    CHK(reserve_level(c, l, d.node_count[l]));
    HIPCHK(c, hipMemcpyAsync(c->nodes[l].p, nodes[l].data(), nodes[l].size() * sizeof(gvdbx_node),
                             hipMemcpyHostToDevice, c->stream));
    if (l > 0)
      HIPCHK(c, hipMemcpyAsync(c->clist[l].p, lists[l].data(), lists[l].size() * 8, hipMemcpyHostToDevice, c->stream));
    c->cnt[l] = d.node_count[l];
  }
  c->depth = depth;
  c->used0 = d.slots;
  c->topo_version++;
  const int res0 = 1 << d.log2dim[0];
  bool any = false;
  for (int a = 0; a < 3; ++a) c->posmin[a] = c->posmax[a] = 0.f;
  if (depth >= 0)
    for (int a = 0; a < 3; ++a) c->root[a] = nodes[depth][0].pos[a];
  for (const gvdbx_node& n : nodes[0]) {  // FinishTopology's bounds of the flagged leaves
    if (!n.flags) continue;
    for (int a = 0; a < 3; ++a) {
      const float lo = (float)n.pos[a], hi = (float)(n.pos[a] + res0);
      c->posmin[a] = any ? std::fmin(c->posmin[a], lo) : lo;
      c->posmax[a] = any ? std::fmax(c->posmax[a], hi) : hi;
    }
    any = true;
  }
  if (d.nchan > 0) {
    CHK(gvdbx_update_atlas(c));
    if (c->nslots != d.slots) return fail(c, GVDBX_ERR_STATE, "LoadVBX: %lld slots for %lld stored bricks",
                                          (long long)c->nslots, (long long)d.slots);
    for (int k = 0; k < d.nchan; ++k)
      HIPCHK(c, hipMemcpyAsync(c->chan[k].p, chans[k].data(), chans[k].size() * 4, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));  // the host copies above are released on return
  return launch_status(c, "LoadVBX");
}

int gvdbx_dev_alloc(gvdbx_ctx* c, int64_t bytes, void** p) {
  NOT_RECORDING(c);
  if (!c || !p || bytes < 0) return GVDBX_ERR_ARG;
  hipError_t e = hipMalloc(p, (size_t)(bytes > 0 ? bytes : 1));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, GVDBX_ERR_NOMEM, "hipMalloc(%lld): %s", (long long)bytes, hipGetErrorString(e));
  }
  return GVDBX_OK;
}

int gvdbx_dev_free(gvdbx_ctx* c, void* p) {
  NOT_RECORDING(c);
  if (!c) return GVDBX_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (p) HIPCHK(c, hipFree(p));
  return GVDBX_OK;
}

int gvdbx_commit(gvdbx_ctx* c, void* dst, const void* src, int64_t bytes) {
  NOT_RECORDING(c);
  if (!c || bytes < 0) return GVDBX_ERR_ARG;
  if (bytes == 0) return GVDBX_OK;
  HIPCHK(c, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GVDBX_OK;
}

int gvdbx_retrieve(gvdbx_ctx* c, void* dst, const void* src, int64_t bytes) {
  if (!c || bytes < 0) return GVDBX_ERR_ARG;
  return copy_out(c, dst, src, (size_t)bytes);
}

int gvdbx_kernel_timer(gvdbx_ctx* c, const char* kernel, int enable) {
  NOT_RECORDING(c);
  if (!c) return GVDBX_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->timer_on = enable && kernel;
  c->timer_name = kernel ? kernel : "";
  c->ev_used = 0;
  return GVDBX_OK;
}

int gvdbx_kernel_timer_read(gvdbx_ctx* c, double* ms, int64_t* n) {
  NOT_RECORDING(c);
  if (!c) return GVDBX_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double tot = 0;
  for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[i], c->ev[i + 1]));
    tot += t;
  }
  if (ms) *ms = tot;
  if (n) *n = (int64_t)(c->ev_used / 2);
  c->ev_used = 0;
  return GVDBX_OK;
}

// ---------------------------------------------------------------- launch shape and graphs

int gvdbx_set_launch(gvdbx_ctx* c, int block_threads, int persistent) {
  if (!c) return GVDBX_ERR_ARG;
  NOT_RECORDING(c);
  if (block_threads != 0 && block_threads != 64 && block_threads != 128 && block_threads != 256)
    return fail(c, GVDBX_ERR_ARG, "block size %d: 0 (automatic), 64, 128 or 256", block_threads);
  if (persistent != 0 && persistent != 1) return fail(c, GVDBX_ERR_ARG, "persistent must be 0 or 1");
  c->launch_block = block_threads;
  c->launch_persist = persistent;
  return GVDBX_OK;
}

int gvdbx_graph_begin(gvdbx_ctx* c) {
  if (!c) return GVDBX_ERR_ARG;
  NOT_RECORDING(c);
  if (!c->configured) return fail(c, GVDBX_ERR_STATE, "context not configured");
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->recording = true;
  c->record_err = GVDBX_OK;
  c->graph_insert = false;
  return GVDBX_OK;
}

int gvdbx_graph_end(gvdbx_ctx* c) {
  if (!c) return GVDBX_ERR_ARG;
  if (!c->recording) return fail(c, GVDBX_ERR_STATE, "no recording in progress (gvdbx_graph_begin)");
  c->recording = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  const int rec = c->record_err;
  const bool had_insert = c->graph_insert;
  c->record_err = GVDBX_OK;
  if (e != hipSuccess || rec != GVDBX_OK || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    c->graph_insert = false;
    if (rec != GVDBX_OK) return rec;  // the failing call's message stays in place
    return fail(c, GVDBX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  release_graph(c);
  hipGraphExec_t x = nullptr;
  const hipError_t e2 = hipGraphInstantiateWithFlags(&x, g, 0);
  if (e2 != hipSuccess) {
    (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return fail(c, GVDBX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e2));
  }
  c->graph = g;
  c->graph_exec = x;
  c->graph_insert = had_insert;
  c->graph_topo = c->topo_version;
  c->graph_atlas = c->atlas_version;
  c->graph_pos = c->pos;
  c->graph_npts = c->npts;
  c->graph_buf_gen = c->buf_gen;
  return GVDBX_OK;
}

int gvdbx_graph_launch(gvdbx_ctx* c) {
  if (!c) return GVDBX_ERR_ARG;
  NOT_RECORDING(c);
  if (!c->graph_exec) return fail(c, GVDBX_ERR_STATE, "no graph recorded (gvdbx_graph_begin / _end)");
  if (c->topo_version != c->graph_topo || c->atlas_version != c->graph_atlas)
    return fail(c, GVDBX_ERR_STATE, "the topology or atlas changed since the graph was recorded");
  if (c->pos != c->graph_pos || c->npts != c->graph_npts)
    return fail(c, GVDBX_ERR_STATE, "the bound points changed since the graph was recorded");
  if (c->buf_gen != c->graph_buf_gen)
    return fail(c, GVDBX_ERR_STATE, "buffers the graph uses were reallocated since it was recorded");
  HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
  if (c->graph_insert) c->sclen_dev = true;
  return GVDBX_OK;
}

int gvdbx_graph_release(gvdbx_ctx* c) {
  if (!c) return GVDBX_ERR_ARG;
  NOT_RECORDING(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  release_graph(c);
  return GVDBX_OK;
}

}  // extern "C"
