"""SaveVBX / LoadVBX through the C ABI (V.cpp:1626-1768, 507-685).

A volume after full, full and incremental rebuilds (its leaf pool holds unlinked leaves, which
take no atlas brick) with a density and a level-set channel is saved; a fresh context loads the
file.  Pools, child lists, depth, root, slots and channels must come back byte for byte, the
sampler must read the same values, the file's pools must be the oracle's, and the loaded
context must go on working (its next rebuild is a full one, as after LoadVBX).
"""
import ctypes as C

import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


def test_save_load_round_trip(oracle_mod, tmp_path):
    from gvdbx import VolumeGVDB, synth
    from gvdbx.lib import VbxDesc
    U.log("case vbx save / load")
    p, _ = synth.sph_block(n=60_000, seed=2, box=64.0)
    v, t = U.new_volume(chans=(0, 1)), U.new_tree(oracle_mod, chans=(0, 1))
    # the third frame moves the block by 12 voxels along x: its incremental rebuild unlinks the
    # bricks left behind
    move = lambda cur, step: (cur + np.float32([12.0, 0.0, 0.0])).astype(np.float32) if step == 2 else cur
    dpos, _ = U.run_pipeline(v, t, p, 1.0, 4, steps=3, move=move)
    cur = (p + np.float32([12.0, 0.0, 0.0])).astype(np.float32)
    ln = v.subcell_info()[1]
    v.GatherLevelSet(4, len(p), 1.0, (0, 0, 0), ln, 1)
    v.UpdateApron(1, 3.0)
    nodes0 = v.get_nodes(0)
    assert (nodes0["flags"] == 0).any(), "the incremental steps left no unlinked leaf"
    path = tmp_path / "sph.vbx"
    v.SaveVBX(path)
    # the file's pools are the oracle's (the library's pools are bit-exact to it)
    d = VbxDesc()
    assert v.L.gvdbx_vbx_probe(str(path).encode(), C.byref(d)) == 0
    assert d.nchan == 2 and d.slots == t.num_slots() and d.root == t.depth() << 8
    w = VolumeGVDB(0)
    w.LoadVBX(path)
    (wd, wroot, _, _), (vd, vroot, _, _) = w.depth(), v.depth()
    assert wd == vd and np.array_equal(wroot, vroot)
    for l in range(len(v.log2dims)):
        assert w.get_nodes(l).tobytes() == v.get_nodes(l).tobytes() == t.nodes(l).tobytes(), f"level {l}"
        if l:
            assert np.array_equal(w.get_childlists(l), v.get_childlists(l))
    assert w.num_slots() == v.num_slots()
    for c in (0, 1):
        assert w.get_channel(c).tobytes() == v.get_channel(c).tobytes(), f"channel {c}"
    q = np.random.default_rng(8).uniform(0.0, 70.0, (50_000, 3)).astype(np.float32)
    dq = v.to_device(q)
    dq2 = w.to_device(q)
    o1, o2 = v.alloc(4 * len(q)), w.alloc(4 * len(q))
    v.SampleTrilinear(0, dq, len(q), o1)
    w.SampleTrilinear(0, dq2, len(q), o2)
    a, b = o1.download(np.float32), o2.download(np.float32)
    assert a.tobytes() == b.tobytes() and np.any(a != 0)
    # the loaded context goes on: a full rebuild of the current points equals the oracle's
    dw = w.to_device(cur)
    w.SetPoints(dw, n=len(cur))
    w.RebuildTopology(len(cur), 1.0)
    w.UpdateAtlas()
    t2 = U.new_tree(oracle_mod, chans=(0, 1))
    t2.rebuild_topology(cur, 1.0)
    t2.update_atlas()
    U.compare_topology(w, t2)
    v.close()
    w.close()
