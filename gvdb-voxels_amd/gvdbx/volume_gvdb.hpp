// volume_gvdb.hpp -- C++ drop-in for the point path of GVDB's VolumeGVDB.
//
// Header-only facade over the C ABI (include/gvdbx.h).  Method names, argument order and
// meaning follow source/gvdb_library/src/gvdb_volume_gvdb.h (V.h:321-558); failures behave
// like the reference: HIP/runtime errors print and exit (cudaCheck, gvdb_allocator.cpp:
// 1006-1034), topology depth errors print and return (gvdb_volume_gvdb.cpp:1187-1194).
// Device buffers are plain device pointers (the reference's CUdeviceptr in DataPtr).
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../../include/gvdbx.h"

namespace gvdbx {

typedef unsigned char uchar;
typedef uint64_t uint64;

struct Vector3DF {
  float x = 0, y = 0, z = 0;
  Vector3DF() {}
  Vector3DF(float a, float b, float c) : x(a), y(b), z(c) {}
};

struct Vector4DF {
  float x = 0, y = 0, z = 0, w = 0;
  Vector4DF() {}
  Vector4DF(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
};

// DataPtr subset used by the point path (gvdb_allocator.h:35-55): a caller-owned device
// buffer with element count, byte offset and byte stride.
struct DataPtr {
  void* gpu = nullptr;
  int64_t lastEle = 0;
  int offs = 0;
  int stride = 0;
};

enum { T_UCHAR4 = GVDBX_T_UCHAR4, T_FLOAT = GVDBX_T_FLOAT };

class VolumeGVDB {
 public:
  VolumeGVDB() {}
  ~VolumeGVDB() {
    if (ctx_) gvdbx_ctx_destroy(ctx_);
  }
  VolumeGVDB(const VolumeGVDB&) = delete;
  VolumeGVDB& operator=(const VolumeGVDB&) = delete;

  // SetCudaDevice (V.cpp:240-334) + Initialize (V.cpp:2320-2361)
  void SetCudaDevice(int devid = 0, void* stream = nullptr) {
    if (ctx_) gvdbx_ctx_destroy(ctx_);
    ctx_ = nullptr;
    check(gvdbx_ctx_create(devid, stream, &ctx_), "SetCudaDevice", true);
  }
  void Initialize() {
    if (!ctx_) SetCudaDevice(0);
  }

  // Configure(q4..q0): the last argument is the leaf level (V.cpp:2364-2378)
  void Configure(int q4, int q3, int q2, int q1, int q0) {
    int r[5] = {q0, q1, q2, q3, q4};
    Configure(5, r, nullptr);
  }
  void Configure(int levs, int* r, int* numcnt, bool use_masks = false) {
    (void)numcnt;
    (void)use_masks;
    check(gvdbx_configure(ctx_, levs, r), "Configure", true);
  }
  void SetChannelDefault(int cx, int cy, int cz) { check(gvdbx_set_channel_default(ctx_, cx, cy, cz), "SetChannelDefault", true); }
  void AddChannel(uchar chan, int dt, int apron, int filter = 0, int border = 0) {
    (void)filter;
    (void)border;
    check(gvdbx_add_channel(ctx_, chan, dt, apron), "AddChannel", true);
    channels_.push_back(chan);
  }
  void FillChannel(uchar chan, Vector4DF val) {
    const float v[4] = {val.x, val.y, val.z, val.w};
    check(gvdbx_fill_channel4(ctx_, chan, v), "FillChannel", true);
  }
  void FillChannel(uchar chan, float val) { FillChannel(chan, Vector4DF(val, 0.0f, 0.0f, 0.0f)); }
  void ClearChannel(uchar chan) { FillChannel(chan, Vector4DF(0.0f, 0.0f, 0.0f, 0.0f)); }

  // SetDataGPU (V.cpp:4872-4880) / SetPoints (V.cpp:4900-4908)
  void SetDataGPU(DataPtr& dat, int cnt, void* gptr, int offs, int stride) {
    dat.gpu = gptr;
    dat.lastEle = cnt;
    dat.offs = offs;
    dat.stride = stride;
  }
  void SetPoints(DataPtr& pntpos, DataPtr& pntvel, DataPtr& pntclr) {
    check(gvdbx_set_points(ctx_, pntpos.lastEle, pntpos.gpu, pntpos.offs, pntpos.stride, pntvel.gpu, pntvel.offs,
                           pntvel.stride, pntclr.gpu, pntclr.offs, pntclr.stride),
          "SetPoints", true);
    clr_bound_ = pntclr.gpu != nullptr;
  }

  // Topology (V.cpp:1158-1218, 1579-1593, 2630-2704)
  void RebuildTopology(int num_pnts, float radius, Vector3DF orig) {
    const float o[3] = {orig.x, orig.y, orig.z};
    check(gvdbx_rebuild_topology(ctx_, num_pnts, radius, o), "RebuildTopology", false);
  }
  void AccumulateTopology(int num_pnts, float radius, Vector3DF orig, int iDepth = 1) {
    const float o[3] = {orig.x, orig.y, orig.z};
    check(gvdbx_accumulate_topology(ctx_, num_pnts, radius, o, iDepth), "AccumulateTopology", false);
  }
  void RequestFullRebuild(bool tf) { check(gvdbx_request_full_rebuild(ctx_, tf ? 1 : 0), "RequestFullRebuild", true); }
  void FinishTopology(bool pCommitPool = true, bool pComputeBounds = true) {
    (void)pCommitPool;
    if (pComputeBounds) check(gvdbx_finish_topology(ctx_, bmin_, bmax_), "FinishTopology", false);
  }
  void UpdateAtlas() { check(gvdbx_update_atlas(ctx_), "UpdateAtlas", true); }

  // Points into subcells and gathers (V.cpp:4971-5092, 5318-5373)
  void InsertPointsSubcell(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength) {
    const float t[3] = {trans.x, trans.y, trans.z};
    int64_t len = 0;
    check(gvdbx_insert_points_subcell(ctx_, subcell_size, num_pnts, radius, t, &len), "InsertPointsSubcell", true);
    pSCPntsLength = (int)len;
  }
  void GatherDensity(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                     int chanDensity, int chanClr, bool bAccum = false) {
    if (pSCPntsLength == 0) return;  // V.cpp:5321
    const float t[3] = {trans.x, trans.y, trans.z};
    check(gvdbx_gather_density(ctx_, subcell_size, num_pnts, radius, t, pSCPntsLength, chanDensity, chanClr,
                               bAccum ? 1 : 0),
          "GatherDensity", true);
  }
  void GatherLevelSet(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                      int chanLevelset, int chanClr, bool bAccum = false) {
    const float t[3] = {trans.x, trans.y, trans.z};
    check(gvdbx_gather_levelset(ctx_, subcell_size, num_pnts, radius, t, pSCPntsLength, chanLevelset, chanClr,
                                bAccum ? 1 : 0),
          "GatherLevelSet", true);
  }
  // FP16 subcell lists and their level set (V.cpp:5094-5189, 5375-5401), GetMinMaxVel (V.cpp:5470-5522)
  void InsertPointsSubcell_FP16(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength) {
    const float t[3] = {trans.x, trans.y, trans.z};
    int64_t len = 0;
    check(gvdbx_insert_points_subcell_fp16(ctx_, subcell_size, num_pnts, radius, t, &len), "InsertPointsSubcell_FP16",
          true);
    pSCPntsLength = (int)len;
  }
  void GatherLevelSet_FP16(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                           int chanLevelSet, int chanClr) {
    const float t[3] = {trans.x, trans.y, trans.z};
    (void)chanClr;  // the reference's FP16 colour accumulator is uninitialised: no colour here
    check(gvdbx_gather_levelset_fp16(ctx_, subcell_size, num_pnts, radius, t, pSCPntsLength, chanLevelSet, -1),
          "GatherLevelSet_FP16", true);
  }
  void GetMinMaxVel(int num_pnts) { check(gvdbx_get_min_max_vel(ctx_, num_pnts, nullptr), "GetMinMaxVel", true); }
  // ScatterDensity (V.cpp:5625-5659) into the reference kernel's channels: density 0, colours
  // (when bound) 1 with avgColor, else 3.
  void ScatterDensity(int num_pnts, float radius, float amp, Vector3DF trans, bool expand = true,
                      bool avgColor = false) {
    const float t[3] = {trans.x, trans.y, trans.z};
    const int chan_clr = clr_bound_ ? (avgColor ? 1 : 3) : -1;
    check(gvdbx_scatter_density(ctx_, num_pnts, radius, amp, t, expand ? 1 : 0, 0, avgColor ? 1 : 0, chan_clr),
          "ScatterDensity", true);
  }
  // Additions of this library (no reference counterpart): velocity channel means and the
  // scatter-reduce level set.
  void GatherVelocity(int chanX, int chanY, int chanZ, bool bAccum = false) {
    check(gvdbx_gather_velocity(ctx_, chanX, chanY, chanZ, bAccum ? 1 : 0), "GatherVelocity", true);
  }
  // GatherDensity with the velocity means in the same pass (channels must differ)
  void GatherDensityVelocity(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                             int chanDensity, int chanX, int chanY, int chanZ) {
    if (pSCPntsLength == 0) return;
    const float t[3] = {trans.x, trans.y, trans.z};
    check(gvdbx_gather_density_velocity(ctx_, subcell_size, num_pnts, radius, t, pSCPntsLength, chanDensity, chanX,
                                        chanY, chanZ),
          "GatherDensityVelocity", true);
  }
  void ScatterReduceLevelSet(float radius, int chan, float init = 3.0f) {
    check(gvdbx_scatter_reduce_levelset(ctx_, radius, chan, init), "ScatterReduceLevelSet", true);
  }

  // UpdateApron (V.cpp:4411-4457)
  void UpdateApron() {
    std::vector<float> bv(channels_.size(), 0.0f);
    std::vector<int> ch(channels_.begin(), channels_.end());
    check(gvdbx_update_apron_channels(ctx_, ch.data(), bv.data(), (int)ch.size()), "UpdateApron", true);
  }
  void UpdateApron(uchar chan, float boundval = 0.0f, bool changeCtx = true) {
    (void)changeCtx;
    check(gvdbx_update_apron(ctx_, chan, boundval), "UpdateApron", true);
  }

  // Introspection (V.h:637-641)
  uint64 getNumTotalNodes(int lev) {
    int64_t t = 0, u = 0;
    gvdbx_get_node_count(ctx_, lev, &t, &u);
    return (uint64)t;
  }
  uint64 getNumUsedNodes(int lev) {
    int64_t t = 0, u = 0;
    gvdbx_get_node_count(ctx_, lev, &t, &u);
    return (uint64)u;
  }
  Vector3DF getWorldMin() const { return Vector3DF(bmin_[0], bmin_[1], bmin_[2]); }
  Vector3DF getWorldMax() const { return Vector3DF(bmax_[0], bmax_[1], bmax_[2]); }
  gvdbx_ctx* context() { return ctx_; }

 private:
  void check(int status, const char* where, bool fatal) {
    if (status == GVDBX_OK) return;
    char buf[512] = {0};
    if (ctx_) gvdbx_last_error(ctx_, buf, sizeof(buf));
    if (status == GVDBX_ERR_DEPTH0) {
      printf("ERROR: Points extents smaller than 1 brick. Not yet supported.\n");
      return;
    }
    if (status == GVDBX_ERR_DEPTH_MAX) {
      printf("ERROR: Point extents exceed 5 levels.\n");
      return;
    }
    printf("GVDB ERROR in %s: %s (%s)\n", where, gvdbx_status_string(status), buf);
    if (fatal || status == GVDBX_ERR_HIP) exit(-1);
  }

  gvdbx_ctx* ctx_ = nullptr;
  bool clr_bound_ = false;
  std::vector<int> channels_;
  float bmin_[3] = {0, 0, 0}, bmax_[3] = {0, 0, 0};
};

}  // namespace gvdbx
