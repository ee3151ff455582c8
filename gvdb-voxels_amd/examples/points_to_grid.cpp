// points_to_grid.cpp -- a C++ caller of the VolumeGVDB facade, following the per-frame call
// sequence of the reference's fluid-surface sample (source/gFluidSurface/main_fluid_surface.cpp:
// 320-356): SetPoints, RebuildTopology, FinishTopology, UpdateAtlas, InsertPointsSubcell,
// GatherDensity / GatherLevelSet, UpdateApron.
//
// usage: points_to_grid <points.f32 | synth:SEED> <n> <radius> <subcell> <levelset 0|1> <out.bin>
// points.f32 holds n float3 positions; synth:SEED takes n points of the shared seeded sphere
// (gvdbx_synth_sphere: radius 20 about (64, 64, 64), sigma 1.5; the points of the committed
// golden sphere fixture for SEED 0 and n 3000).  out.bin receives [nodes0 count][slots][brick res]
// (int64 x3), the level-0 node pool (64 B each) and channel 0 (float, brick-linear).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../gvdbx/volume_gvdb.hpp"

int main(int argc, char** argv) {
  if (argc < 7) {
    fprintf(stderr, "usage: %s points.f32 n radius subcell levelset out.bin\n", argv[0]);
    return 2;
  }
  const int n = atoi(argv[2]);
  const float radius = (float)atof(argv[3]);
  const int subcell = atoi(argv[4]);
  const bool levelset = atoi(argv[5]) != 0;
  std::vector<float> pts((size_t)n * 3);
  if (strncmp(argv[1], "synth:", 6) == 0) {
    const float center[3] = {64.f, 64.f, 64.f};
    if (gvdbx_synth_sphere(strtoull(argv[1] + 6, nullptr, 10), n, center, 20.f, 1.5f, pts.data()) != GVDBX_OK) {
      fprintf(stderr, "bad synthetic point count\n");
      return 2;
    }
  } else {
    FILE* f = fopen(argv[1], "rb");
    if (!f || fread(pts.data(), sizeof(float), pts.size(), f) != pts.size()) {
      fprintf(stderr, "cannot read %s\n", argv[1]);
      return 2;
    }
    fclose(f);
  }

  gvdbx::VolumeGVDB gvdb;
  gvdb.SetCudaDevice(0);
  gvdb.Initialize();
  gvdb.Configure(5, 3, 3, 3, 3);  // leaf bricks of 8^3
  gvdb.SetChannelDefault(16, 16, 1);
  gvdb.AddChannel(0, gvdbx::T_FLOAT, 1);

  void* dpos = nullptr;
  if (gvdbx_dev_alloc(gvdb.context(), (int64_t)pts.size() * 4, &dpos) ||
      gvdbx_commit(gvdb.context(), dpos, pts.data(), (int64_t)pts.size() * 4)) {
    fprintf(stderr, "device upload failed\n");
    return 1;
  }
  gvdbx::DataPtr pos, vel, clr;
  gvdb.SetDataGPU(pos, n, dpos, 0, 12);
  gvdb.SetPoints(pos, vel, clr);

  const gvdbx::Vector3DF orig(0, 0, 0);
  int len = 0;
  gvdb.RebuildTopology(n, radius, orig);
  gvdb.FinishTopology(false, true);
  gvdb.UpdateAtlas();
  gvdb.InsertPointsSubcell(subcell, n, radius, orig, len);
  if (levelset)
    gvdb.GatherLevelSet(subcell, n, radius, orig, len, 0, -1);
  else
    gvdb.GatherDensity(subcell, n, radius, orig, len, 0, -1);
  gvdb.UpdateApron(0, levelset ? 3.0f : 0.0f);

  int64_t total = 0, used = 0, slots = 0;
  int br = 0;
  gvdbx_get_node_count(gvdb.context(), 0, &total, &used);
  gvdbx_get_num_slots(gvdb.context(), &slots, &br);
  std::vector<gvdbx_node> nodes((size_t)total);
  std::vector<float> chan((size_t)slots * br * br * br);
  gvdbx_copy_nodes(gvdb.context(), 0, nodes.data(), total);
  gvdbx_copy_channel(gvdb.context(), 0, chan.data(), (int64_t)chan.size());
  FILE* o = fopen(argv[6], "wb");
  if (!o) return 1;
  const int64_t hdr[3] = {total, slots, br};
  fwrite(hdr, sizeof(hdr), 1, o);
  fwrite(nodes.data(), sizeof(gvdbx_node), nodes.size(), o);
  fwrite(chan.data(), sizeof(float), chan.size(), o);
  fclose(o);
  printf("points %d leaves %lld used %lld subcell-entries %d\n", n, (long long)total,
         (long long)gvdb.getNumUsedNodes(0), len);
  gvdbx_dev_free(gvdb.context(), dpos);
  return 0;
}
