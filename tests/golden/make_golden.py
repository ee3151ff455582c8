"""Writes the golden fixtures of tests/golden from the oracle (oracle/gvdb_oracle.c) on inputs of
the shared seeded generator (gvdbx/synth.py shared_*, the same bits as the library's
gvdbx_synth_*).  Run from the repository root:  python3 tests/golden/make_golden.py

sphere.npz: 3,000 points on a sphere of radius 20 about (64, 64, 64) (seed 0, sigma 1.5), tree
  <3,3,3,3,5>, apron 1: RebuildTopology(r=1), UpdateAtlas, InsertPointsSubcell(4, r=1),
  GatherDensity(r=1) + UpdateApron(0) into channel 0, GatherLevelSet(r=1) + UpdateApron(3) into
  channel 1.
sph.npz: the 24 x 10 x 24 jittered lattice (seed 1, spacing 1.2, jitter 1, N(0,1) velocities),
  three frames of q += 0.5 v (full, full, incremental rebuild with r=1; the topology of every
  frame), then in the last frame InsertPointsSubcell(4, r=1), GatherDensity(r=1) into channel 0
  and the velocity means into channels 1-3, UpdateApron of all four.
Channels are stored as FP32 roundings of the oracle's FP64 values.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "gvdb-voxels_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

LD = (3, 3, 3, 3, 5)
SPH_DIMS, SPH_SEED, SPH_H, DT = (24, 10, 24), 1, 1.2, np.float32(0.5)


def sphere_points():
    from gvdbx import synth
    return synth.shared_sphere(3000, 0, (64.0, 64.0, 64.0), 20.0, 1.5)


def sph_points():
    from gvdbx import synth
    return synth.shared_lattice(SPH_DIMS, SPH_SEED, SPH_H, 1.0, 1.0)


def sph_frames():
    p, v = sph_points()
    out = [p]
    for _ in range(2):
        out.append((out[-1] + v * DT).astype(np.float32))
    return out, v


def topology(t, prefix, out):
    out[prefix + "depth"] = np.int32(t.depth())
    out[prefix + "rootpos"] = t.rootpos()
    for l in range(len(LD)):
        out[f"{prefix}nodes{l}"] = t.nodes(l).view(np.uint8)
        if l:
            out[f"{prefix}clist{l}"] = t.childlists(l)


def make_sphere(O):
    p = sphere_points()
    t = O.Tree(LD)
    for c in (0, 1):
        t.add_channel(c)
    t.rebuild_topology(p, 1.0)
    t.update_atlas()
    ln = t.insert_points_subcell(4, p, 1.0)
    t.gather_density(1.0, 0)
    t.update_apron(0, 0.0)
    t.gather_levelset(1.0, 1)
    t.update_apron(1, 3.0)
    sc = t.subcells()
    out = {"points": p, "list_len": np.int64(ln), "slots": np.int64(t.num_slots()), "leaf_slot": t.leaf_slot(),
           "sc_map": sc["map"], "sc_cnt": sc["cnt"], "sc_off": sc["off"],
           "density": t.channel(0).astype(np.float32), "levelset": t.channel(1).astype(np.float32)}
    topology(t, "", out)
    return out


def make_sph(O):
    frames, v = sph_frames()
    t = O.Tree(LD)
    for c in range(4):
        t.add_channel(c)
    out = {"points0": frames[0], "velocities": v}
    for k, q in enumerate(frames):
        t.rebuild_topology(q, 1.0)
        t.update_atlas()
        topology(t, f"f{k}_", out)
    q = frames[-1]
    ln = t.insert_points_subcell(4, q, 1.0)
    t.gather_density(1.0, 0)
    t.gather_velocity(v, (1, 2, 3))
    for c in range(4):
        t.update_apron(c, 0.0)
    sc = t.subcells()
    out.update({"list_len": np.int64(ln), "slots": np.int64(t.num_slots()), "leaf_slot": t.leaf_slot(),
                "sc_map": sc["map"], "sc_cnt": sc["cnt"], "sc_off": sc["off"],
                "channels": np.stack([t.channel(c).astype(np.float32) for c in range(4)])})
    return out


def main():
    import oracle as O
    np.savez_compressed(os.path.join(HERE, "sphere.npz"), **make_sphere(O))
    np.savez_compressed(os.path.join(HERE, "sph.npz"), **make_sph(O))
    for f in ("sphere.npz", "sph.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
