"""Colour channels of GatherDensity / GatherLevelSet (P.cuh:713-731, 752-770) on the GPU.

The points carry uchar4 colours (32-bit words); InsertPointsSubcell lists them beside the
positions and the gathers write, per voxel, the mean of its subcell's colours into a uchar4
channel.  The result is integer data, so it must equal the oracle's restatement (exact integer
sums, floored quotient, saturated at 255) bit for bit: interiors after the gather, whole padded
bricks after UpdateApron with a boundary value, with and without accumulation, for the density
and the level-set gathers.
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


def word(r, g, b, a):
    return r | g << 8 | b << 16 | a << 24


def colour_cloud(seed, n=30000):
    g = np.random.default_rng(seed)
    p = g.uniform(20, 110, (n, 3)).astype(np.float32)
    p[:30] -= 150.0  # invalid points: skipped by the insert
    clr = g.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    return p, clr


def interior(a, ap=1):
    return a[:, ap:-ap, ap:-ap, ap:-ap]


def colour_pair(oracle_mod, p, clr, sc, r):
    """GPU volume (float channel 0, uchar4 channel 1) and oracle tree after a full rebuild and
    the insert of the coloured points."""
    from gvdbx import T_UCHAR4
    v, t = U.new_volume(chans=(0,)), U.new_tree(oracle_mod, chans=(0, 1))
    v.AddChannel(1, T_UCHAR4, apron=1)
    dp, dc = v.to_device(p), v.to_device(clr)
    v.SetPoints(dp, clr=dc, n=len(p))
    v.RebuildTopology(len(p), r)
    v.UpdateAtlas()
    t.rebuild_topology(p, r)
    t.update_atlas()
    U.compare_topology(v, t)
    t.fill_channel(0, 0.0)
    v.FillChannel(0, 0.0)
    ln = v.InsertPointsSubcell(sc, len(p), r)
    assert ln == t.insert_points_subcell(sc, p, r)
    g = v.get_subcells()
    o = t.subcells()
    # listed colours are the listed points' colours (same multiset per subcell)
    seg = np.repeat(np.arange(len(o["cnt"])), o["cnt"])
    assert np.array_equal(np.sort(g["list_clr"] + (seg.astype(np.uint64) << 32)),
                          np.sort(clr[o["list"]] + (seg.astype(np.uint64) << 32)))
    return v, t, ln, (dp, dc)


def same_words(got, want):
    got = np.asarray(got, np.uint32)
    want = np.asarray(want)
    assert want.dtype == np.float64
    assert (want == np.floor(want)).all() and (want >= 0).all() and (want < 2.0 ** 32).all()
    bad = np.count_nonzero(got != want.astype(np.uint32))
    assert bad == 0, f"{bad} of {got.size} colour words differ"


@pytest.mark.parametrize("sc,r,seed", [(4, 1.0, 0), (2, 1.0, 1), (4, 2.0, 2)])
def test_colour_gather_matches_oracle(oracle_mod, sc, r, seed):
    U.log(f"case colour gather sc={sc} r={r}")
    p, clr = colour_cloud(seed)
    v, t, ln, _bufs = colour_pair(oracle_mod, p, clr, sc, r)
    n = len(p)
    # FillChannel(chan, Vector4DF) on a uchar4 channel: CLR2INT(x, y, z, 1)
    v.FillChannel(1, (0.5, 0.25, 1.0, 0.0))
    fill = word(127, 63, 255, 255)
    assert (v.get_channel(1) == fill).all()
    t.fill_channel(1, float(fill))
    # density + colour, no accumulation: interior, then the apron written from the neighbours
    v.GatherDensity(sc, n, r, (0, 0, 0), ln, 0, chan_clr=1)
    t.gather_density(r, 0)
    t.gather_colour(clr, 1)
    g1 = v.get_channel(1)
    same_words(interior(g1), interior(t.channel(1)[: len(g1)]))
    want = t.channel(1)[: len(g1)]
    assert len(np.unique(interior(want))) > 100  # many distinct means
    U.compare_channel(v, t, 0, interior_only=True)
    v.UpdateApron(1, 0.0)
    t.update_apron(1, 0.0)
    same_words(v.get_channel(1), t.channel(1)[: len(g1)])
    # accumulation on top of the previous words; the apron then needs UpdateApron
    v.GatherDensity(sc, n, r, (0, 0, 0), ln, 0, chan_clr=1, accum=True)
    t.gather_density(r, 0, accum=True)
    t.gather_colour(clr, 1, accum=True)
    same_words(interior(v.get_channel(1)), interior(t.channel(1)[: len(g1)]))
    v.UpdateApron(1, 7.0)  # make_uchar4(7, 7, 7, 7)
    t.update_apron(1, float(word(7, 7, 7, 7)))
    same_words(v.get_channel(1), t.channel(1)[: len(g1)])
    # the level-set gather writes colour by the same rule
    v.GatherLevelSet(sc, n, r, (0, 0, 0), ln, 0, chan_clr=1)
    t.gather_levelset(r, 0)
    t.gather_colour(clr, 1)
    same_words(interior(v.get_channel(1)), interior(t.channel(1)[: len(g1)]))
    U.compare_channel(v, t, 0, interior_only=True)
    v.close()


def test_colour_channel_rules(oracle_mod):
    from gvdbx import GvdbxError
    U.log("case colour channel rules")
    p, clr = colour_cloud(5, 8000)
    v, t, ln, bufs = colour_pair(oracle_mod, p, clr, 4, 1.0)
    n = len(p)
    for bad in (lambda: v.GatherDensity(4, n, 1.0, (0, 0, 0), ln, 0, chan_clr=0),   # float channel
                lambda: v.GatherDensity(4, n, 1.0, (0, 0, 0), ln, 1, chan_clr=-1),  # uchar4 field
                lambda: v.FillChannel(5, 1.0)):                                   # no channel 5
        with pytest.raises(GvdbxError) as e:
            bad()
        assert e.value.code == 1
    q = v.to_device(p[:100])
    out = v.alloc(400)
    with pytest.raises(GvdbxError):
        v.SampleTrilinear(1, q, 100, out)
    # all-zero FillChannel clears a uchar4 channel (ClearChannel)
    v.FillChannel(1, (0.0, 0.0, 0.0, 0.0))
    assert (v.get_channel(1) == 0).all()
    # without colours bound at the insert the colour channel is left alone (P.cuh:729)
    v.FillChannel(1, (1.0, 1.0, 1.0, 1.0))
    v.SetPoints(bufs[0], n=n)
    ln = v.InsertPointsSubcell(4, n, 1.0)
    v.GatherDensity(4, n, 1.0, (0, 0, 0), ln, 0, chan_clr=1)
    assert (v.get_channel(1) == 0xFFFFFFFF).all()
    v.close()


@pytest.mark.parametrize("avg", [True, False])
def test_scatter_density_colour_matches_oracle(oracle_mod, avg):
    """ScatterDensity with colours bound (P.cuh:553-572, 610-641): the uchar4 channel's words
    equal the oracle's bit for bit -- averages per reached voxel, or the colour of the voxel's
    highest-index point -- whatever order the points come in, while the density part still
    matches; voxels no point reached keep the fill word."""
    from gvdbx import T_UCHAR4
    U.log(f"case scatter colour avg={avg}")
    g = np.random.default_rng(11)
    p = g.uniform(20, 100, (40000, 3)).astype(np.float32)
    p[:20] -= 150.0
    clr = g.integers(0, 1 << 32, len(p), dtype=np.uint64).astype(np.uint32)
    v, t = U.new_volume(chans=(0,)), U.new_tree(oracle_mod, chans=(0, 1))
    v.AddChannel(1, T_UCHAR4, apron=1)
    dp = v.to_device(p)
    v.SetPoints(dp, n=len(p))
    v.RebuildTopology(len(p), 1.0)
    v.UpdateAtlas()
    t.rebuild_topology(p, 1.0)
    t.update_atlas()
    for perm in (np.arange(len(p)), g.permutation(len(p))):
        pp, cc = p[perm], clr[perm]
        dpp, dcc = v.to_device(pp), v.to_device(cc)
        v.SetPoints(dpp, clr=dcc, n=len(p))
        v.FillChannel(0, 0.0)
        v.FillChannel(1, (0.2, 0.4, 0.6, 1.0))
        v.ScatterDensity(len(p), 1.0, 1.0, (0, 0, 0), True, avgColor=avg, chan=0, chan_clr=1)
        t.fill_channel(0, 0.0)
        fill = word(51, 102, 153, 255)
        t.fill_channel(1, float(fill))
        t.scatter_density(0, pp, 1.0)
        t.scatter_colour(1, pp, cc, avg=avg)
        got = v.get_channel(1)
        want = t.channel(1)[: len(got)]
        same_words(got, want)
        reached = np.count_nonzero(want != fill)
        assert reached > 10000, reached
        U.compare_channel(v, t, 0)
    v.close()
