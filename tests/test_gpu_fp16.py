"""FP16 subcell lists on the GPU: InsertPointsSubcell_FP16 + GatherLevelSet_FP16 (V.cpp:5094-5189,
5375-5401; kernels P.cuh:154-216, 775-811) and GetMinMaxVel (V.cpp:5470-5522).

The quantized lists are integer data: per subcell they must equal the oracle's quantization of
the same entries as multisets, bit for bit (positions, and velocities when velocity lists are
kept).  The FP16 level set must match the oracle's decoding level set within 1e-5 (floored at 1)
and stay within the quantization bound of the float level set.  Float gathers refuse FP16 lists
and the FP16 gather refuses float lists.
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


def cloud(seed, n=20000):
    g = np.random.default_rng(seed)
    p = g.uniform(20, 110, (n, 3)).astype(np.float32)
    p[:15] -= 150.0  # invalid points
    v = g.standard_normal((n, 3)).astype(np.float32)
    return p, v


def keyed(seg, rows):
    """Per-subcell multiset of uint16 rows as one sortable uint64 per entry."""
    rows = rows.astype(np.uint64)
    k = (seg.astype(np.uint64) << 48) | (rows[:, 0] << 32) | (rows[:, 1] << 16) | rows[:, 2]
    return np.sort(k)


@pytest.mark.parametrize("sc,r,seed", [(4, 1.0, 0), (2, 1.0, 1), (4, 2.0, 2)])
def test_fp16_lists_and_levelset(oracle_mod, sc, r, seed):
    from gvdbx import GvdbxError
    U.log(f"case fp16 sc={sc} r={r}")
    p, vel = cloud(seed)
    v, t = U.new_volume(chans=(0, 1)), U.new_tree(oracle_mod, chans=(0, 1))
    dp, dv = v.to_device(p), v.to_device(vel)
    v.SetPoints(dp, dv, n=len(p))
    v.RebuildTopology(len(p), r)
    v.UpdateAtlas()
    t.rebuild_topology(p, r)
    t.update_atlas()
    U.compare_topology(v, t)
    _, _, pmin, pmax = v.depth()
    qmin = np.asarray(pmin, np.float32)
    qrange = (np.asarray(pmax, np.float32) - qmin).astype(np.float32)
    ln = v.InsertPointsSubcell_FP16(sc, len(p), r)
    assert ln == t.insert_points_subcell(sc, p, r)
    g, o = v.get_subcells(), t.subcells()
    for k in ("cnt", "off", "pos", "nid"):
        assert np.array_equal(g[k], o[k]), k
    seg = np.repeat(np.arange(len(o["cnt"])), o["cnt"])
    want_q = t.quantize_list_pos(qmin, qrange)
    assert np.array_equal(keyed(seg, g["list_q16"]), keyed(seg, want_q)), "FP16 list positions differ"
    assert len(np.unique(want_q[:, 0])) > 1000
    # the FP16 level set against the oracle's decoding one, and near the float level set
    v.FillChannel(0, 0.0)
    t.fill_channel(0, 0.0)
    v.GatherLevelSet_FP16(sc, len(p), r, (0, 0, 0), ln, 0)
    t.gather_levelset_q16(r, 0, want_q, qmin, qrange)
    U.compare_channel(v, t, 0, interior_only=True)
    t.gather_levelset(r, 1)
    a = v.apron
    d = np.abs(v.get_channel(0)[:, a:-a, a:-a, a:-a] - t.channel(1)[: v.num_slots()[0], a:-a, a:-a, a:-a])
    assert d.max() <= np.sqrt(3.0) * float(qrange.max()) / 65535.0 * 1.01 + 1e-5
    # float gathers refuse the FP16 lists; chan_clr is refused
    with pytest.raises(GvdbxError) as e:
        v.GatherDensity(sc, len(p), r, (0, 0, 0), ln, 1)
    assert e.value.code == 6
    with pytest.raises(GvdbxError) as e:
        v.GatherLevelSet_FP16(sc, len(p), r, (0, 0, 0), ln, 0, chan_clr=1)
    assert e.value.code == 1
    # velocity lists: quantized over GetMinMaxVel's box
    v.SetSubcellOptions(store_velocity_lists=True)
    vmin, vmax = v.GetMinMaxVel(len(p))
    assert np.array_equal(vmin, vel.min(0)) and np.array_equal(vmax, vel.max(0))
    ln2 = v.InsertPointsSubcell_FP16(sc, len(p), r)
    assert ln2 == ln
    g = v.get_subcells()
    vrange = (vmax - vmin).astype(np.float32)
    want_v = t.quantize_list_vel(vel, vmin, vrange)
    both_g = np.concatenate([g["list_q16"], g["list_vq16"]], axis=1)
    both_o = np.concatenate([want_q, want_v], axis=1)
    og = np.lexsort(tuple(both_g[:, c] for c in range(5, -1, -1)) + (seg,))
    oo = np.lexsort(tuple(both_o[:, c] for c in range(5, -1, -1)) + (seg,))
    assert np.array_equal(both_g[og], both_o[oo]), "FP16 list velocities differ"
    # back to float lists: the FP16 gather refuses them
    v.SetSubcellOptions(store_velocity_lists=False)
    ln3 = v.InsertPointsSubcell(sc, len(p), r)
    with pytest.raises(GvdbxError) as e:
        v.GatherLevelSet_FP16(sc, len(p), r, (0, 0, 0), ln3, 0)
    assert e.value.code == 6
    v.close()


def test_fp16_velocity_lists_need_min_max_vel(oracle_mod):
    """The refusal comes before anything is launched: the lists of the insert before stay as they
    were, so a gather after the refused call gives the same channel, bit for bit."""
    from gvdbx import GvdbxError
    U.log("case fp16 velocity box required")
    p, vel = cloud(5, 3000)
    v = U.new_volume(chans=(0,))
    dp, dv = v.to_device(p), v.to_device(vel)
    v.SetPoints(dp, dv, n=len(p))
    v.RebuildTopology(len(p), 1.0)
    v.UpdateAtlas()
    ln = v.InsertPointsSubcell(4, len(p), 1.0)
    v.GatherLevelSet(4, len(p), 1.0, (0, 0, 0), ln, 0)
    first = v.get_channel(0).tobytes()
    info = v.subcell_info()
    v.SetSubcellOptions(store_velocity_lists=True)
    with pytest.raises(GvdbxError) as e:
        v.InsertPointsSubcell_FP16(4, len(p), 1.0)
    assert e.value.code == 6
    assert v.subcell_info() == info
    v.FillChannel(0, 0.0)
    v.GatherLevelSet(4, len(p), 1.0, (0, 0, 0), ln, 0)
    assert v.get_channel(0).tobytes() == first
    v.close()
