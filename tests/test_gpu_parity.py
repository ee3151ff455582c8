"""GPU (libgvdbx.so, gfx950) vs CPU oracle parity through the C ABI.

Bar: topology pools, child lists, slots, subcell maps/counts/offsets/positions bit-exact;
subcell list contents equal as multisets per subcell; FP32 channels within
|gpu - oracle| <= 1e-5 * max(|oracle|, 1) of the FP64 oracle.
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


def rand_points(n, seed, lo=20.0, hi=180.0):
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3)).astype(np.float32)


@pytest.mark.parametrize("seed,radius,sc", [(0, 1.0, 4), (1, 0.5, 4), (2, 2.0, 4), (3, 1.0, 2), (4, 1.0, 8)])
def test_small_random_pipeline(oracle_mod, seed, radius, sc):
    U.log(f"case small seed={seed} r={radius} sc={sc}")
    p = rand_points(3000, seed)
    p[:40] -= 250.0  # invalid points
    v, t = U.new_volume(), U.new_tree(oracle_mod)
    U.run_pipeline(v, t, p, radius, sc, levelset=False)


@pytest.mark.parametrize("radius", [0.5, 1.0])
def test_full_rebuild_sparse_extent_sorts(oracle_mod, radius):
    """Two clouds 6,000 voxels apart: the occupancy bitmaps of a full rebuild would cover over
    2^27 leaf bricks, so the rebuild sorts keys instead; the pools must still equal the oracle's."""
    U.log(f"case sparse extent r={radius}")
    g = np.random.default_rng(11)
    p = np.concatenate([g.uniform(10, 60, (1500, 3)), g.uniform(6000, 6050, (1500, 3))]).astype(np.float32)
    v, t = U.new_volume(), U.new_tree(oracle_mod)
    U.run_pipeline(v, t, p, radius, 4)


def test_levelset_small(oracle_mod):
    U.log("case levelset small")
    p = rand_points(4000, 7, 30, 120)
    v, t = U.new_volume(), U.new_tree(oracle_mod)
    U.run_pipeline(v, t, p, 2.0, 4, density=False, levelset=True, apron_bound=3.0)


def test_sphere_262k_density(oracle_mod):
    """The CPU-runnable workload: 262,144 points on a sphere, r = 1, sc = 4."""
    from gvdbx import synth
    U.log("case sphere 262k")
    p = synth.sphere_shell()
    v, t = U.new_volume(), U.new_tree(oracle_mod)
    U.run_pipeline(v, t, p, 1.0, 4)


def test_sph_steps_incremental_with_velocity(oracle_mod):
    """SPH-style steps on a reduced block: full, full, then incremental rebuilds."""
    from gvdbx import synth
    U.log("case sph incremental")
    p, vel = synth.sph_block(n=131_072, seed=1, box=96.0)

    def move(cur, step):
        return (cur + vel * np.float32(0.5)).astype(np.float32)

    v, t = U.new_volume(chans=(0, 1, 2, 3)), U.new_tree(oracle_mod, chans=(0, 1, 2, 3))
    U.run_pipeline(v, t, p, 1.0, 4, vel=vel, chans_vel=(1, 2, 3), steps=6, move=move)


def test_determinism_and_point_order(oracle_mod):
    """Two runs are bit-identical; shuffling the points leaves topology and fields unchanged."""
    from gvdbx import synth
    U.log("case determinism")
    p = synth.sphere_shell(n=65_536, seed=2)
    g = np.random.default_rng(0)
    vel = g.standard_normal(p.shape).astype(np.float32)
    outs = []
    for perm in (np.arange(len(p)), np.arange(len(p)), g.permutation(len(p))):
        v = U.new_volume(chans=(0, 1, 2, 3))
        dp = v.to_device(p[perm])
        dv = v.to_device(vel[perm])
        v.SetPoints(dp, dv, n=len(p))
        v.RebuildTopology(len(p), 1.0)
        v.UpdateAtlas()
        ln = v.InsertPointsSubcell(4, len(p), 1.0)
        v.GatherDensity(4, len(p), 1.0, (0, 0, 0), ln, 0)
        v.GatherVelocity(1, 2, 3)
        for c in range(4):
            v.UpdateApron(c, 0.0)
        outs.append((v.get_nodes(0).tobytes(), [v.get_channel(c).tobytes() for c in range(4)]))
        v.close()
    assert outs[0] == outs[1]
    assert outs[0][0] == outs[2][0]
    assert outs[0][1] == outs[2][1]


def test_trilinear_sample(oracle_mod):
    U.log("case trilinear")
    p = rand_points(20000, 8, 40, 100)
    v, t = U.new_volume(), U.new_tree(oracle_mod)
    U.run_pipeline(v, t, p, 1.0, 4)
    q = np.random.default_rng(5).uniform(30, 110, (20000, 3)).astype(np.float32)
    dq = v.to_device(q)
    dout = v.alloc(4 * len(q))
    v.SampleTrilinear(0, dq, len(q), dout)
    got = dout.download(np.float32).astype(np.float64)
    want = t.sample_trilinear(0, q)
    err = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    assert err.max() <= U.TOL, err.max()
    assert (want != 0).sum() > 1000


def test_scatter_reduce_levelset_equals_gather(oracle_mod):
    U.log("case scatter-reduce level set")
    p = rand_points(6000, 9, 30, 110)
    v, t = U.new_volume(chans=(0, 1)), U.new_tree(oracle_mod, chans=(0, 1))
    U.run_pipeline(v, t, p, 2.0, 4, density=False, levelset=True, apron_bound=3.0)
    v.ScatterReduceLevelSet(2.0, 1, 3.0)
    t.scatter_reduce_levelset(2.0, 1, 3.0)
    U.compare_channel(v, t, 1, interior_only=True)
    g0 = v.get_channel(0)[:, 1:-1, 1:-1, 1:-1]
    g1 = v.get_channel(1)[:, 1:-1, 1:-1, 1:-1]
    assert np.array_equal(g0, g1), "scatter-reduce and gather produce the same FP32 field"


def test_depth_errors_on_gpu():
    from gvdbx import GvdbxError
    U.log("case depth errors")
    v = U.new_volume()
    p = np.full((16, 3), 3.5, np.float32)
    d = v.to_device(p)
    v.SetPoints(d, n=16)
    with pytest.raises(GvdbxError) as e:
        v.RebuildTopology(16, 1.0)
    assert e.value.code == 2
    p = np.array([[0, 0, 0], [3.0e6, 1, 1]], np.float32)
    d2 = v.to_device(p)
    v.SetPoints(d2, n=2)
    with pytest.raises(GvdbxError) as e:
        v.RebuildTopology(2, 1.0)
    assert e.value.code == 3
    with pytest.raises(GvdbxError) as e:
        v.GatherDensity(4, 2, 1.0, (0, 0, 0), 0, 0)
    assert e.value.code in (1, 6)


def test_rejected_rebuild_keeps_state(oracle_mod):
    """A RebuildTopology that ends in a depth error (V.cpp:1187-1194 returns before touching the
    tree) leaves the pools, child lists, atlas and subcell lists as they were: a gather afterwards
    still matches, and the next (incremental) rebuild of the original points matches the
    oracle's, which never saw the rejected call.  The rejected call comes after two full
    rebuilds, so it runs the speculative incremental box pass, whose new keys (the 16 points
    reach a missing brick) must not reach the node flags."""
    from gvdbx import GvdbxError
    U.log("case rejected rebuild keeps state")
    p = rand_points(4000, 31, 30, 110)
    v, t = U.new_volume(), U.new_tree(oracle_mod)
    dpos, _ = U.run_pipeline(v, t, p, 1.0, 4, steps=2)
    before = v.depth()[0], v.node_count(0), v.subcell_info()
    pools = [v.get_nodes(l).tobytes() for l in range(len(v.log2dims))]
    lists = [v.get_childlists(l).tobytes() for l in range(1, before[0] + 1)]
    small = v.to_device(np.full((16, 3), 3.5, np.float32))
    v.SetPoints(small, n=16)
    with pytest.raises(GvdbxError) as e:
        v.RebuildTopology(16, 1.0)
    assert e.value.code == 2
    assert (v.depth()[0], v.node_count(0), v.subcell_info()) == before
    assert [v.get_nodes(l).tobytes() for l in range(len(v.log2dims))] == pools, "a rejected call changed a pool"
    assert [v.get_childlists(l).tobytes() for l in range(1, before[0] + 1)] == lists
    v.SetPoints(dpos, n=len(p))
    v.FillChannel(0, 0.0)
    t.fill_channel(0, 0.0)
    v.GatherDensity(4, len(p), 1.0, (0, 0, 0), before[2][1], 0)
    t.gather_density(1.0, 0)
    U.compare_channel(v, t, 0, interior_only=True)
    # the next call is the state machine's first incremental rebuild, on both sides
    v.RebuildTopology(len(p), 1.0)
    t.rebuild_topology(p, 1.0)
    v.UpdateAtlas()
    t.update_atlas()
    U.compare_topology(v, t)
    ln = v.InsertPointsSubcell(4, len(p), 1.0)
    assert ln == t.insert_points_subcell(4, p, 1.0)
    U.compare_subcells(v, t)
    v.GatherDensity(4, len(p), 1.0, (0, 0, 0), ln, 0)
    t.gather_density(1.0, 0)
    U.compare_channel(v, t, 0, interior_only=True)
    v.close()


@pytest.mark.parametrize("radius,sc", [(2.5, 2), (3.5, 4), (1.0, 1)])
def test_window_overflow_paths(oracle_mod, radius, sc):
    """r + 1 > sc: pairs outside the per-brick LDS window take the global path."""
    U.log(f"case window r={radius} sc={sc}")
    p = rand_points(3000, 21, 40, 120)
    v, t = U.new_volume(), U.new_tree(oracle_mod)
    U.run_pipeline(v, t, p, radius, sc, density=False, levelset=True, apron_bound=3.0)


def test_orphan_points_and_velocity_lists(oracle_mod):
    """Insert points whose own brick is missing (topology built from a subset) and keep the
    optional binned velocities."""
    U.log("case orphans")
    g = np.random.default_rng(4)
    a = rand_points(4000, 22, 40, 100)
    b = g.uniform(20, 130, (3000, 3)).astype(np.float32)
    allp = np.concatenate([a, b])
    vel = g.standard_normal(allp.shape).astype(np.float32)
    v, t = U.new_volume(chans=(0, 1, 2, 3)), U.new_tree(oracle_mod, chans=(0, 1, 2, 3))
    v.SetSubcellOptions(store_velocity_lists=True)
    dp = v.to_device(allp)
    dv = v.to_device(vel)
    v.SetPoints(dp, dv, n=len(a))
    v.RebuildTopology(len(a), 1.0)
    t.rebuild_topology(a, 1.0)
    v.UpdateAtlas()
    t.update_atlas()
    U.compare_topology(v, t)
    v.SetPoints(dp, dv, n=len(allp))
    ln = v.InsertPointsSubcell(4, len(allp), 1.0)
    assert ln == t.insert_points_subcell(4, allp, 1.0)
    gs = U.compare_subcells(v, t)
    o = t.subcells()
    seg = np.repeat(np.arange(len(o["cnt"])), o["cnt"])
    key_g = np.lexsort((gs["list_vel"][:, 2], gs["list_vel"][:, 1], gs["list_vel"][:, 0], seg))
    want_vel = vel[o["list"]]
    key_o = np.lexsort((want_vel[:, 2], want_vel[:, 1], want_vel[:, 0], seg))
    assert np.array_equal(gs["list_vel"][key_g], want_vel[key_o])
    v.GatherDensity(4, len(allp), 1.0, (0, 0, 0), ln, 0)
    t.gather_density(1.0, 0)
    v.GatherVelocity(1, 2, 3)
    t.gather_velocity(vel, (1, 2, 3))
    for c in range(4):
        U.compare_channel(v, t, c, interior_only=True)


def test_accumulate_gathers(oracle_mod):
    """accum = True reads the channel back (GatherDensity / GatherLevelSet / velocity means)."""
    U.log("case accumulate")
    p = rand_points(3000, 23, 40, 100)
    vel = np.random.default_rng(5).standard_normal(p.shape).astype(np.float32)
    v, t = U.new_volume(chans=(0, 1, 2, 3)), U.new_tree(oracle_mod, chans=(0, 1, 2, 3))
    U.run_pipeline(v, t, p, 1.0, 4, vel=vel, chans_vel=(1, 2, 3))
    ln = v.subcell_info()[1]
    v.GatherDensity(4, len(p), 1.5, (0, 0, 0), ln, 0, accum=True)
    t.gather_density(1.5, 0, accum=True)
    v.GatherVelocity(1, 2, 3, accum=True)
    t.gather_velocity(vel, (1, 2, 3), accum=True)
    for c in range(4):
        U.compare_channel(v, t, c, interior_only=True)


def test_apron_state_transitions(oracle_mod):
    """Fused apron (gather-written neighbour copies) + later boundary values, a second
    UpdateApron with another boundary value, and the full pull after a fill / level set."""
    U.log("case apron states")
    p = rand_points(5000, 31, 40, 110)
    vel = np.random.default_rng(6).standard_normal(p.shape).astype(np.float32)
    v, t = U.new_volume(chans=(0, 1, 2, 3)), U.new_tree(oracle_mod, chans=(0, 1, 2, 3))
    U.run_pipeline(v, t, p, 1.0, 4, vel=vel, chans_vel=(1, 2, 3))
    v.UpdateApron(0, 5.0)
    t.update_apron(0, 5.0)
    U.compare_channel(v, t, 0)
    v.UpdateApron()  # all channels, boundary 0
    for c in range(4):
        t.update_apron(c, 0.0)
        U.compare_channel(v, t, c)
    v.FillChannel(2, 1.5)
    t.fill_channel(2, 1.5)
    v.UpdateApron(2, -2.0)
    t.update_apron(2, -2.0)
    U.compare_channel(v, t, 2)
    ln = v.subcell_info()[1]
    v.GatherLevelSet(4, len(p), 1.0, (0, 0, 0), ln, 0)
    t.gather_levelset(1.0, 0)
    v.UpdateApron(0, 3.0)
    t.update_apron(0, 3.0)
    U.compare_channel(v, t, 0)


@pytest.mark.parametrize("sc,levelset", [(2, False), (4, False), (2, True), (1, False)])
def test_small_bricks_4cubed(oracle_mod, sc, levelset):
    """4^3 leaf bricks (log2 dims 2,3,3,3,5): the generic brick / subcell kernels."""
    U.log(f"case 4^3 bricks sc={sc} levelset={levelset}")
    ld = (2, 3, 3, 3, 5)
    p = rand_points(4000, 40 + sc, 20, 90)
    v, t = U.new_volume(ld), U.new_tree(oracle_mod, ld)
    U.run_pipeline(v, t, p, 1.0, sc, density=not levelset, levelset=levelset, apron_bound=3.0 if levelset else 0.0)


def test_bricks_4cubed_incremental_with_velocity(oracle_mod):
    from gvdbx import synth
    U.log("case 4^3 bricks incremental velocity")
    ld = (2, 3, 3, 3, 5)
    p, vel = synth.sph_block(n=32_768, seed=3, box=48.0)

    def move(cur, step):
        return (cur + vel * np.float32(0.5)).astype(np.float32)

    v, t = U.new_volume(ld, chans=(0, 1, 2, 3)), U.new_tree(oracle_mod, ld, chans=(0, 1, 2, 3))
    U.run_pipeline(v, t, p, 1.0, 2, vel=vel, chans_vel=(1, 2, 3), steps=4, move=move)


@pytest.mark.parametrize("ld", [(3, 3, 3, 3, 5), (2, 3, 3, 3, 5)])
def test_density_velocity_one_pass(oracle_mod, ld):
    """GatherDensityVelocity equals GatherDensity + GatherVelocity bit for bit (all four padded
    channels, before and after UpdateApron) and the oracle within the tolerance, over
    incremental SPH steps."""
    from gvdbx import synth
    U.log(f"case density+velocity one pass ld={ld}")
    p, vel = synth.sph_block(n=65_536, seed=4, box=80.0)
    t = U.new_tree(oracle_mod, ld, chans=(0, 1, 2, 3))
    vs = [U.new_volume(ld, chans=(0, 1, 2, 3)) for _ in range(2)]
    dvs = [(v.to_device(p), v.to_device(vel)) for v in vs]
    cur = p
    for step in range(4):
        outs = []
        for k, v in enumerate(vs):
            dp, dv = dvs[k]
            dp.upload(cur)
            v.SetPoints(dp, dv, n=len(cur))
            v.RebuildTopology(len(cur), 1.0)
            v.UpdateAtlas()
            ln = v.InsertPointsSubcell(4, len(cur), 1.0)
            if k == 0:
                v.GatherDensity(4, len(cur), 1.0, (0, 0, 0), ln, 0)
                v.GatherVelocity(1, 2, 3)
            else:
                v.GatherDensityVelocity(4, len(cur), 1.0, (0, 0, 0), ln, 0, (1, 2, 3))
            before = [v.get_channel(c).tobytes() for c in range(4)]
            v.UpdateApron()
            outs.append((before, [v.get_channel(c).tobytes() for c in range(4)]))
        assert outs[0] == outs[1], f"step {step}: one pass differs from the two gathers"
        t.rebuild_topology(cur, 1.0)
        t.update_atlas()
        t.insert_points_subcell(4, cur, 1.0)
        t.gather_density(1.0, 0)
        t.gather_velocity(vel, (1, 2, 3))
        for c in range(4):
            t.update_apron(c, 0.0)
            U.compare_channel(vs[1], t, c)
        cur = (cur + vel * np.float32(0.5)).astype(np.float32)
    for v in vs:
        v.close()


def test_small_velocities_absolute_resolution(oracle_mod):
    """Velocity means are exact sums of 2^-24 fixed-point terms: each term is rounded once, so a
    mean is within 2^-25 of the FP64 mean whatever the magnitudes (plus the FP32 cast).  With
    velocities of order 1e-3 that absolute bound is the statement (a 1e-5 relative bound would not
    hold there); the one-pass density + velocity gather meets the same bound."""
    from gvdbx import synth
    U.log("case small velocities")
    p, vel = synth.sph_block(n=40_000, seed=21, box=48.0)
    vel = (vel * np.float32(1e-3)).astype(np.float32)
    v, t = U.new_volume(chans=(0, 1, 2, 3)), U.new_tree(oracle_mod, chans=(0, 1, 2, 3))
    d, dv = v.to_device(p), v.to_device(vel)
    v.SetPoints(d, dv, n=len(p))
    v.RebuildTopology(len(p), 1.0)
    t.rebuild_topology(p, 1.0)
    v.UpdateAtlas()
    t.update_atlas()
    ln = v.InsertPointsSubcell(4, len(p), 1.0)
    t.insert_points_subcell(4, p, 1.0)
    t.gather_velocity(vel, (1, 2, 3))
    a = v.apron
    for one_pass in (False, True):
        if one_pass:
            v.GatherDensityVelocity(4, len(p), 1.0, (0, 0, 0), ln, 0, (1, 2, 3))
        else:
            v.GatherVelocity(1, 2, 3)
        for c in (1, 2, 3):
            g = v.get_channel(c)[:, a:-a, a:-a, a:-a].astype(np.float64)
            o = t.channel(c)[:, a:-a, a:-a, a:-a]
            err = np.abs(g - o)
            bound = 2.0 ** -25 + np.abs(o) * 2.0 ** -24
            assert (err <= bound).all(), (one_pass, c, float((err - bound).max()))
            assert np.abs(o).max() > 1e-4  # the means are of the small velocities' order
    v.close()


def test_dense_bricks_velocity_lists(oracle_mod):
    """Bricks of hundreds of points, where the count pass gathers the points itself and writes the
    sorted copies (velocities too, when the velocity lists are kept), with orphan points (their
    own brick missing) mixed in: lists with velocities as per-subcell multisets, density and
    velocity means against the oracle."""
    from gvdbx import synth
    U.log("case dense bricks, velocity lists")
    p, vel = synth.sph_block(n=120_000, seed=5, box=40.0)
    extra = np.random.default_rng(6).uniform(2.0, 60.0, (3000, 3)).astype(np.float32)
    allp = np.concatenate([p, extra])
    allv = np.concatenate([vel, np.random.default_rng(7).standard_normal(extra.shape).astype(np.float32)])
    v, t = U.new_volume(chans=(0, 1, 2, 3)), U.new_tree(oracle_mod, chans=(0, 1, 2, 3))
    v.SetSubcellOptions(store_velocity_lists=True)
    dp, dv = v.to_device(allp), v.to_device(allv)
    v.SetPoints(dp, dv, n=len(p))
    v.RebuildTopology(len(p), 1.0)
    t.rebuild_topology(p, 1.0)
    v.UpdateAtlas()
    t.update_atlas()
    U.compare_topology(v, t)
    assert len(allp) >= 64 * v.num_slots()[0], "the bricks are not dense"
    v.SetPoints(dp, dv, n=len(allp))
    ln = v.InsertPointsSubcell(4, len(allp), 1.0)
    assert ln == t.insert_points_subcell(4, allp, 1.0)
    gs = U.compare_subcells(v, t)
    o = t.subcells()
    seg = np.repeat(np.arange(len(o["cnt"])), o["cnt"])
    key_g = np.lexsort((gs["list_vel"][:, 2], gs["list_vel"][:, 1], gs["list_vel"][:, 0], seg))
    want_vel = allv[o["list"]]
    key_o = np.lexsort((want_vel[:, 2], want_vel[:, 1], want_vel[:, 0], seg))
    assert np.array_equal(gs["list_vel"][key_g], want_vel[key_o])
    v.GatherDensity(4, len(allp), 1.0, (0, 0, 0), ln, 0)
    t.gather_density(1.0, 0)
    v.GatherVelocity(1, 2, 3)
    t.gather_velocity(allv, (1, 2, 3))
    for c in range(4):
        U.compare_channel(v, t, c, interior_only=True)
    v.close()
