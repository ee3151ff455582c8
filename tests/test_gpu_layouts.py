"""Point buffers inside records, and the samples' 32^3 bricks, against the oracle.

- gSprayDeposit binds its points inside the 64-byte ScnRay records (V.h:300-308): positions at
  offset 0 and colours at offset 48, stride sizeof(ScnRay) (main_spray_deposit.cpp:287-288),
  then inserts and gathers density with the colour channel, accumulating
  (main_spray_deposit.cpp:294-295).  Here velocities ride in the same record (offset 36, the
  ray direction) as well.  The oracle reads the same records with the same strides.
- gFluidSurface configures Configure(3, 3, 3, 3, 5), 32^3 leaf bricks (main_fluid_surface.cpp:87),
  and per frame rebuilds, inserts with radius 2r, gathers the level set with r and fills the
  apron with 3.0 (main_fluid_surface.cpp:320-356); the brick-window kernels handle bricks up to
  8^3, so 32^3 bricks run the general paths.  Checked over full, full and incremental frames.
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu

RAY = np.dtype([("hit", np.float32, 3), ("normal", np.float32, 3), ("orig", np.float32, 3), ("dir", np.float32, 3),
                ("clr", np.uint32), ("pnode", np.uint32), ("pndx", np.uint32), ("pad", np.uint32)])
assert RAY.itemsize == 64 and RAY.fields["clr"][1] == 48 and RAY.fields["dir"][1] == 36


def test_points_inside_64_byte_records(oracle_mod):
    U.log("case AoS records (ScnRay)")
    g = np.random.default_rng(41)
    n = 30_000
    rec = np.zeros(n, RAY)
    rec["hit"] = g.uniform(150.0, 250.0, (n, 3)).astype(np.float32)
    rec["hit"][:25] -= 400.0  # missed rays land at negative coordinates: skipped
    rec["normal"] = g.standard_normal((n, 3)).astype(np.float32)
    rec["orig"] = 999.0
    rec["dir"] = g.standard_normal((n, 3)).astype(np.float32)
    rec["clr"] = g.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    rec["pnode"], rec["pndx"] = 0xDEADBEEF, 0xFFFFFFFF
    from gvdbx.volume import T_UCHAR4
    v = U.new_volume(chans=(0,))
    for c in (2, 3, 4):
        v.AddChannel(c, apron=1)
    v.AddChannel(1, dtype=T_UCHAR4, apron=1)
    t = U.new_tree(oracle_mod, chans=(0, 1, 2, 3, 4))
    d = v.to_device(rec)
    v.SetPoints(d, d, d, n=n, pos_off=0, pos_stride=64, vel_off=36, vel_stride=64, clr_off=48, clr_stride=64)
    floats = rec.view(np.float32).reshape(n, 16)  # the oracle reads positions with a 16-float stride
    pos = floats[:, 0:3]
    vel = floats[:, 9:12]
    for frame in range(2):
        v.RebuildTopology(n, 1.0)
        t.rebuild_topology(floats, 1.0)
        v.UpdateAtlas()
        t.update_atlas()
        U.compare_topology(v, t)
        ln = v.InsertPointsSubcell(4, n, 1.0)
        assert ln == t.insert_points_subcell(4, floats, 1.0)
        U.compare_subcells(v, t)
        v.FillChannel(0, 0.0)
        v.FillChannel(1, (0.2, 0.4, 0.6, 1.0))
        t.fill_channel(0, 0.0)
        t.fill_channel(1, float(51 | 102 << 8 | 153 << 16 | 255 << 24))
        v.GatherDensity(4, n, 1.0, (0, 0, 0), ln, 0, 1, True)  # accumulate, as the sample does
        t.gather_density(1.0, 0, accum=True)
        t.gather_colour(rec["clr"], 1, accum=True)
        v.GatherVelocity(2, 3, 4)
        t.gather_velocity(vel, (2, 3, 4))
        U.compare_channel(v, t, 0, interior_only=True)
        for c in (2, 3, 4):
            U.compare_channel(v, t, c, interior_only=True)
        a = v.apron
        got = v.get_channel(1).view(np.uint32)[:, a:-a, a:-a, a:-a]
        want = t.channel(1)[:, a:-a, a:-a, a:-a].astype(np.uint32)
        assert np.array_equal(got, want), "colour words differ"
        U.log(f"  frame {frame}: {v.node_count(0)[1]} bricks, {ln} entries")
    assert np.array_equal(pos[:5], rec["hit"][:5])
    # the displacement writes the positions inside the records and nothing else
    v.AdvancePoints(n, 0.25)
    back = d.download(np.uint8).view(RAY)
    want = (rec["hit"] + rec["dir"] * np.float32(0.25)).astype(np.float32)
    assert back["hit"].tobytes() == want.tobytes()
    for f in ("normal", "orig", "dir", "clr", "pnode", "pndx", "pad"):
        assert back[f].tobytes() == rec[f].tobytes(), f
    v.close()


def test_brick32_fluid_surface_frames(oracle_mod):
    from gvdbx import synth
    U.log("case 32^3 bricks (Configure(3,3,3,3,5))")
    ld = (5, 3, 3, 3, 3)
    p, vel = synth.sph_block(n=150_000, seed=6, box=160.0)
    v, t = U.new_volume(ld, chans=(0, 1)), U.new_tree(oracle_mod, ld, chans=(0, 1))
    n = len(p)
    d = v.to_device(p)
    v.SetPoints(d, n=n)
    r = 1.0
    cur = p
    for frame in range(3):
        if frame:
            cur = (cur + vel * np.float32(0.5)).astype(np.float32)
            d.upload(cur)
        v.RebuildTopology(n, 2 * r)
        t.rebuild_topology(cur, 2 * r)
        v.UpdateAtlas()
        t.update_atlas()
        U.compare_topology(v, t)
        ln = v.InsertPointsSubcell(4, n, 2 * r)
        assert ln == t.insert_points_subcell(4, cur, 2 * r)
        U.compare_subcells(v, t)
        v.GatherLevelSet(4, n, r, (0, 0, 0), ln, 0)
        t.gather_levelset(r, 0)
        v.UpdateApron(0, 3.0)
        t.update_apron(0, 3.0)
        U.compare_channel(v, t, 0)
        v.GatherDensity(4, n, r, (0, 0, 0), ln, 1)
        t.gather_density(r, 1)
        v.UpdateApron(1, 0.0)
        t.update_apron(1, 0.0)
        U.compare_channel(v, t, 1)
        U.log(f"  frame {frame}: {v.node_count(0)[1]} bricks of 32^3, {ln} entries")
    v.close()


def test_advance_points_range_and_refusals():
    from gvdbx.lib import GvdbxError
    U.log("case AdvancePoints refusals and partial range")
    p = np.random.default_rng(5).uniform(10.0, 50.0, (1000, 3)).astype(np.float32)
    v = U.new_volume()
    d = v.to_device(p)
    v.SetPoints(d, n=1000)
    with pytest.raises(GvdbxError):
        v.AdvancePoints(1000, 0.5)  # no velocities bound
    dv = v.to_device(p)
    v.SetPoints(d, dv, n=1000)
    with pytest.raises(GvdbxError):
        v.AdvancePoints(1001, 0.5)  # beyond the bound buffer
    with pytest.raises(GvdbxError):
        v.AdvancePoints(1000, float("nan"))
    assert d.download(np.float32).tobytes() == p.tobytes(), "a refused call moved points"
    v.AdvancePoints(500, 1.0)
    got = d.download(np.float32).reshape(-1, 3)
    assert got[:500].tobytes() == (p[:500] + p[:500]).tobytes()
    assert got[500:].tobytes() == p[500:].tobytes()
    v.close()
