"""Launch shapes and graphs of the insert / gather / apron sequence.

- Every launch shape (64, 128 and 256 threads per brick workgroup, one workgroup per brick or
  persistent workgroups over a brick queue) gives the subcell counts, offsets and channels of
  the default shape bit for bit (counts are integer sums, the density a max, the velocity means
  come from exact fixed-point sums, the apron copies), on an SPH cloud and on a batch of clouds.
- A recorded graph of InsertPointsSubcell + GatherDensity(+ velocity means) + UpdateApron, launched
  several times with eager work (channel fills, synchronisations, reads) between the launches,
  equals the eager sequence after every launch, and its list length comes back from the device.
- Recording refuses what it cannot hold: topology changes, host waits, buffer growth; a launch
  after the topology or the bound points changed is refused.
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu

SHAPES = [(64, False), (128, False), (256, False), (64, True), (128, True), (256, True)]


def sph_volume(n=150_000, seed=12):
    from gvdbx import synth
    p, vel = synth.sph_block(n=n, seed=seed, box=96.0)
    v = U.new_volume(chans=(0, 1, 2, 3))
    dp, dv = v.to_device(p), v.to_device(vel)
    v.SetPoints(dp, dv, n=len(p))
    v.RebuildTopology(len(p), 1.0)
    v.UpdateAtlas()
    return v, len(p), (dp, dv)


def step(v, n, vel=True):
    ln = v.InsertPointsSubcell(4, n, 1.0)
    if vel:
        v.GatherDensityVelocity(4, n, 1.0, (0, 0, 0), ln, 0, (1, 2, 3))
        v.UpdateApron()
    else:
        v.GatherDensity(4, n, 1.0, (0, 0, 0), ln, 0)
        v.UpdateApron(0, 0.0)
    return ln


def snapshot(v, chans=(0, 1, 2, 3)):
    nsc, ln, _ = v.subcell_info()
    cnt = np.zeros(nsc, np.int32)
    off = np.zeros(nsc, np.int32)
    v._chk(v.L.gvdbx_copy_subcell_array(v.h, 1, cnt.ctypes.data, cnt.nbytes), "cnt")
    v._chk(v.L.gvdbx_copy_subcell_array(v.h, 2, off.ctypes.data, off.nbytes), "off")
    return ln, cnt, off, [v.get_channel(c).tobytes() for c in chans]


def same(a, b, what):
    assert a[0] == b[0], f"{what}: list length {a[0]} != {b[0]}"
    assert np.array_equal(a[1], b[1]), f"{what}: subcell counts"
    assert np.array_equal(a[2], b[2]), f"{what}: subcell offsets"
    for c, (x, y) in enumerate(zip(a[3], b[3])):
        assert x == y, f"{what}: channel {c}"


def test_launch_shapes_equal_default_sph():
    U.log("case launch shapes, SPH cloud")
    v, n, keep = sph_volume()
    step(v, n)
    ref = snapshot(v)
    for bs, pers in SHAPES:
        v.SetLaunch(bs, pers)
        for c in range(4):
            v.FillChannel(c, 7.0)
        step(v, n)
        same(snapshot(v), ref, f"block {bs} persistent {pers}")
        U.log(f"  block {bs} persistent {pers}: equal")
    v.close()


def test_graph_replays_equal_eager_sph():
    U.log("case graph replays, SPH cloud")
    v, n, keep = sph_volume(seed=13)
    step(v, n)
    ref = snapshot(v)
    for bs, pers in [(0, False), (128, True)]:
        v.SetLaunch(bs, pers)
        step(v, n)  # the calls run once with these arguments before recording
        v.GraphBegin()
        step(v, n)
        v.GraphEnd()
        for rep in range(3):
            for c in range(4):
                v.FillChannel(c, -1.0)  # eager work between launches
            v.synchronize()
            v.GraphLaunch()
            got = snapshot(v)
            same(got, ref, f"launch {rep} (block {bs}, persistent {pers})")
        U.log(f"  block {bs} persistent {pers}: 3 launches equal the eager sequence")
    v.GraphRelease()
    v.close()


def test_graph_batch_of_clouds():
    from test_gpu_batch import CELL, LD, clouds
    U.log("case graph + persistent, batch of clouds")
    cs = clouds(seed=21, count=8)
    counts = [len(c) for c in cs]
    v = U.new_volume(LD)
    d = v.to_device(np.concatenate(cs))
    v.BatchSetPoints(d, counts, cell_log2=CELL, max_radius=1.0)
    n = int(sum(counts))
    v.RebuildTopology(n, 1.0)
    v.UpdateAtlas()
    step(v, n, vel=False)
    ref = snapshot(v, chans=(0,))
    for bs, pers in [(64, True), (256, False)]:
        v.SetLaunch(bs, pers)
        step(v, n, vel=False)
        same(snapshot(v, chans=(0,)), ref, f"eager block {bs} persistent {pers}")
        v.GraphBegin()
        step(v, n, vel=False)
        v.GraphEnd()
        for rep in range(2):
            v.FillChannel(0, 3.0)
            v.synchronize()
            v.GraphLaunch()
            same(snapshot(v, chans=(0,)), ref, f"graph launch {rep} block {bs} persistent {pers}")
    U.log(f"  {len(cs)} clouds: graphs equal the eager batch")
    v.GraphRelease()
    v.close()


def test_graph_refusals():
    from gvdbx.lib import GvdbxError
    U.log("case graph refusals")
    v, n, (dp, dv) = sph_volume(n=40_000, seed=14)
    with pytest.raises(GvdbxError):
        v.SetLaunch(96, False)
    with pytest.raises(GvdbxError):
        v.GraphLaunch()  # nothing recorded
    step(v, n)
    ref = snapshot(v)
    # a rebuild cannot be recorded: the recording is void, GraphEnd reports it
    v.GraphBegin()
    with pytest.raises(GvdbxError):
        v.RebuildTopology(n, 1.0)
    with pytest.raises(GvdbxError):
        v.GraphEnd()
    # a host wait cannot be recorded either
    v.GraphBegin()
    with pytest.raises(GvdbxError):
        v.synchronize()
    with pytest.raises(GvdbxError):
        v.GraphEnd()
    # growth: more points than the buffers of the run before the recording hold (a growing
    # buffer reserves twice its request, so 2.5 times the points)
    from gvdbx import synth
    p2, vel2 = synth.sph_block(n=100_000, seed=14, box=96.0)
    w = U.new_volume(chans=(0, 1, 2, 3))
    dp2, dv2 = w.to_device(p2), w.to_device(vel2)
    w.SetPoints(dp2, dv2, n=40_000)
    w.RebuildTopology(40_000, 1.0)
    w.UpdateAtlas()
    step(w, 40_000)
    w.SetPoints(dp2, dv2, n=100_000)
    w.GraphBegin()
    with pytest.raises(GvdbxError):
        w.InsertPointsSubcell(4, 100_000, 1.0)
    with pytest.raises(GvdbxError):
        w.GraphEnd()
    w.close()
    # a good recording, then a rebuild: the launch is refused
    v.GraphBegin()
    step(v, n)
    v.GraphEnd()
    v.GraphLaunch()
    same(snapshot(v), ref, "launch before the rebuild")
    v.RequestFullRebuild(True)
    v.RebuildTopology(n, 1.0)
    with pytest.raises(GvdbxError):
        v.GraphLaunch()
    # the eager calls still work after all of it
    v.UpdateAtlas()
    step(v, n)
    same(snapshot(v), ref, "eager after the refusals")
    v.close()
