"""Leaf lookups where the lookup table does not cover every leaf.

UpdateAtlas builds the leaf lookup table over the brick cells around the last rebuild's points;
leaves elsewhere (here: a first cloud kept by AccumulateTopology while the table covers the
second) must still be found by descent.  Topology, subcell lists, density, apron and trilinear
samples over both clouds are compared with the oracle, then a further accumulation runs its box
pass over both clouds (table hits for one, descents for the other).
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


def test_leaves_outside_the_table_box(oracle_mod):
    U.log("case lookup table: leaves outside its box")
    g = np.random.default_rng(21)
    a = g.uniform(30, 60, (4000, 3)).astype(np.float32)
    a[0] = (310.0, 310.0, 310.0)  # the first call is a full rebuild: this point makes its root span both clouds
    b = g.uniform(300, 330, (4000, 3)).astype(np.float32)
    v, t = U.new_volume(), U.new_tree(oracle_mod)
    for p in (a, b):  # depth 2 for both (one root), so the second call accumulates
        d = v.to_device(p)
        v.SetPoints(d, n=len(p))
        v.AccumulateTopology(len(p), 1.0)
        t.accumulate_topology(p, 1.0)
    assert v.depth()[0] == t.depth() == 2
    v.UpdateAtlas()
    t.update_atlas()
    U.compare_topology(v, t)
    ab = np.concatenate([a, b])
    dab = v.to_device(ab)
    v.SetPoints(dab, n=len(ab))
    ln = v.InsertPointsSubcell(4, len(ab), 1.0)
    assert ln == t.insert_points_subcell(4, ab, 1.0)
    U.compare_subcells(v, t)
    v.GatherDensity(4, len(ab), 1.0, (0, 0, 0), ln, 0)
    t.gather_density(1.0, 0)
    U.compare_channel(v, t, 0, interior_only=True)
    v.UpdateApron(0, 0.0)
    t.update_apron(0, 0.0)
    U.compare_channel(v, t, 0)
    q = np.concatenate([a[:2000], b[:2000]]) + np.float32(0.37)
    dq = v.to_device(q.astype(np.float32))
    out = v.alloc(4 * len(q))
    v.SampleTrilinear(0, dq, len(q), out)
    got = out.download(np.float32, len(q)).astype(np.float64)
    want = t.sample_trilinear(0, q)
    assert (np.abs(got - want) / np.maximum(np.abs(want), 1.0)).max() <= U.TOL
    assert (want[:2000] != 0).sum() > 500 and (want[2000:] != 0).sum() > 500  # both clouds sampled
    # a further accumulation: its box pass finds cloud b's leaves in the table, cloud a's by descent
    moved = (ab + g.uniform(-1.5, 1.5, ab.shape)).astype(np.float32)
    dm = v.to_device(moved)
    v.SetPoints(dm, n=len(moved))
    v.AccumulateTopology(len(moved), 1.0)
    t.accumulate_topology(moved, 1.0)
    v.UpdateAtlas()
    t.update_atlas()
    U.compare_topology(v, t)
    v.close()
