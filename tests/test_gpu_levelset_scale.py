"""The level-set workload at full size on the GPU: 16,777,216 points on the surfaces of 64
spheres in a 1024^3 domain, RebuildTopology + InsertPointsSubcell (sc 4) with r = 2,
ScatterReduceLevelSet(r = 2, init 3.0), UpdateApron(3.0), then 1,048,576 uniform trilinear
queries.

The oracle cannot run the whole field in test time, so it is pinned three ways:
  - topology: the oracle's full rebuild of all points, pools compared byte for byte;
  - 48 random bricks: the oracle pipeline on exactly the points that can reach them (every
    point whose brick is within one brick of a chosen one: a (point, subcell) pair needs
    |p - subcell| <= r + 1 and brick creation |p - brick| <= 2r + 1, both < 8), subcell counts
    equal and level-set values within 1e-5;
  - size-independent invariants over the whole field (bounds, every point's home voxel at
    most |p - v| - r) and the sampler against FP64 trilinear interpolation of the GPU field.
"""
import ctypes as C

import numpy as np
import pytest

import gpu_util as U

RADIUS = 2.0


def _subcell_counts(v):
    nsc, _, _ = v.subcell_info()
    total = v.node_count(0)[0]
    m = np.zeros(total, np.int32)
    cnt = np.zeros(nsc, np.int32)
    v._chk(v.L.gvdbx_copy_subcell_array(v.h, 0, m.ctypes.data, m.nbytes), "map")
    v._chk(v.L.gvdbx_copy_subcell_array(v.h, 1, cnt.ctypes.data, cnt.nbytes), "cnt")
    return m, cnt


def _brick_lookup(nodes):
    """(sorted linear keys of the flagged leaves, their slots, their pool indices)."""
    live = np.nonzero(nodes["flags"] == 1)[0]
    pos = nodes["pos"][live].astype(np.int64) // 8
    keys = (pos[:, 0] << 42) | (pos[:, 1] << 21) | pos[:, 2]
    order = np.argsort(keys)
    return keys[order], np.arange(len(live))[order], live[order]


def _find(lookup, bpos):
    keys, slots, pool = lookup
    k = (bpos[:, 0].astype(np.int64) << 42) | (bpos[:, 1].astype(np.int64) << 21) | bpos[:, 2].astype(np.int64)
    i = np.clip(np.searchsorted(keys, k), 0, len(keys) - 1)
    ok = keys[i] == k
    return np.where(ok, slots[i], -1), np.where(ok, pool[i], -1)


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_levelset_16m_points(oracle_mod):
    from gvdbx import synth
    U.log("case level set 16M: generate")
    p = synth.sphere_surfaces()
    n = len(p)
    v = U.new_volume(chans=(0,))
    d = v.to_device(p)
    v.SetPoints(d, n=n)
    v.RebuildTopology(n, RADIUS)
    v.UpdateAtlas()
    ln = v.InsertPointsSubcell(4, n, RADIUS)
    v.ScatterReduceLevelSet(RADIUS, 0, 3.0)
    v.UpdateApron(0, 3.0)
    qs = synth.uniform_queries()
    dq = v.to_device(qs)
    out = v.alloc(4 * len(qs))
    v.SampleTrilinear(0, dq, len(qs), out)
    samples = out.download(np.float32).astype(np.float64)
    ch = v.get_channel(0)
    nodes = v.get_nodes(0)
    gmap, gcnt = _subcell_counts(v)
    U.log(f"gpu done: bricks {len(ch)} subcell entries {ln}")
    assert ln > 150_000_000

    # ---- topology against the oracle's full rebuild
    t = U.new_tree(oracle_mod)
    t.rebuild_topology(p, RADIUS)
    t.update_atlas()
    U.compare_topology(v, t)
    del t
    U.log("topology equal")

    # ---- 48 random bricks against the oracle on the points that can reach them
    lookup = _brick_lookup(nodes)
    rng = np.random.default_rng(48)
    chosen = rng.choice(len(ch), 48, replace=False)
    live = np.nonzero(nodes["flags"] == 1)[0]
    cpos = nodes["pos"][live[chosen]].astype(np.int64) // 8
    near = set()
    for b in cpos:
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    near.add(((b[0] + dx) << 42) | ((b[1] + dy) << 21) | (b[2] + dz))
    pb = np.floor(p.astype(np.float64) / 8).astype(np.int64)
    pk = (pb[:, 0] << 42) | (pb[:, 1] << 21) | pb[:, 2]
    sub = p[np.isin(pk, np.fromiter(near, np.int64))]
    U.log(f"oracle subset: {len(sub)} points")
    t = U.new_tree(oracle_mod)
    t.rebuild_topology(sub, RADIUS)
    t.update_atlas()
    t.insert_points_subcell(4, sub, RADIUS)
    t.scatter_reduce_levelset(RADIUS, 0, 3.0)
    tn = t.nodes(0)
    tslot = t.leaf_slot()
    tch = t.channel(0)
    tsc = t.subcells()
    tl = _brick_lookup(tn)
    oslot, opool = _find(tl, cpos)
    assert (oslot >= 0).all()
    gslot, gpool = chosen, live[chosen]
    for i in range(48):
        assert tslot[opool[i]] == oslot[i]
        g = ch[gslot[i], 1:-1, 1:-1, 1:-1].astype(np.float64)
        o = tch[oslot[i], 1:-1, 1:-1, 1:-1]
        err = np.abs(g - o) / np.maximum(np.abs(o), 1.0)
        assert err.max() <= U.TOL, (i, err.max())
        gc = gcnt[gmap[gpool[i]] * 8: gmap[gpool[i]] * 8 + 8]
        oc = tsc["cnt"][tsc["map"][opool[i]] * 8: tsc["map"][opool[i]] * 8 + 8]
        assert np.array_equal(gc, oc), (i, gc, oc)
    del t
    U.log("48 bricks equal")

    # ---- invariants of the whole field
    interior = ch[:, 1:-1, 1:-1, 1:-1]
    assert interior.max() <= 3.0 and interior.min() >= -RADIUS
    valid = np.nonzero((p >= 0).all(1))[0]  # points with a negative coordinate are skipped
    assert len(valid) > 0.9 * n
    pick = p[rng.choice(valid, 200_000, replace=False)].astype(np.float64)
    hv = np.floor(pick).astype(np.int64)
    slot, _ = _find(lookup, hv // 8)
    assert (slot >= 0).all()
    loc = hv - (hv // 8) * 8 + 1
    val = ch[slot, loc[:, 2], loc[:, 1], loc[:, 0]].astype(np.float64)
    bound = np.sqrt(((pick - hv) ** 2).sum(1)) - RADIUS
    assert (val <= bound + 1e-5).all()

    # ---- sampler against FP64 trilinear interpolation of the GPU field (apron included)
    q = qs.astype(np.float64)
    hb = np.floor(q).astype(np.int64) // 8
    qslot, _ = _find(lookup, hb)
    want = np.zeros(len(q))
    hit = qslot >= 0
    x = q[hit] - hb[hit] * 8 + 1 - 0.5
    i0 = np.floor(x).astype(np.int64)
    a = x - i0
    s = qslot[hit]
    acc = np.zeros(hit.sum())
    for k in range(8):
        dx, dy, dz = k & 1, (k >> 1) & 1, k >> 2
        w = (a[:, 0] if dx else 1 - a[:, 0]) * (a[:, 1] if dy else 1 - a[:, 1]) * (a[:, 2] if dz else 1 - a[:, 2])
        acc += w * ch[s, i0[:, 2] + dz, i0[:, 1] + dy, i0[:, 0] + dx]
    want[hit] = acc
    err = np.abs(samples - want) / np.maximum(np.abs(want), 1.0)
    assert err.max() <= U.TOL, err.max()
    assert hit.sum() > 1000
    U.log(f"sampler equal on {hit.sum()} queries in bricks")
