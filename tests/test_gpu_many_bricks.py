"""More leaf bricks than one dispatch can hold at one workgroup each.

A dispatch carries at most 2^32 - 1 work-items per grid dimension, so kernels that give a
256-thread workgroup to every brick pass that limit beyond 16,777,215 bricks.  This runs the
pipeline on 4,000,000 uniform points in a 2048^3 domain with 4^3 leaf bricks (about 19.8M
leaves), subcell 2, r = 1: RebuildTopology, UpdateAtlas, InsertPointsSubcell,
GatherDensity, UpdateApron.  Pinned against the oracle three ways:
  - topology: the oracle's full rebuild of all points, pools compared byte for byte;
  - 400 bricks (300 at random points' home bricks, 100 at random): the oracle pipeline on the
    points that can reach them (every point whose brick is within one brick of a chosen one:
    pairs need |p - subcell| <= r + 1 = 2 and brick creation |p - brick| <= 2r + 1 = 3, both
    below the 4-voxel brick), subcell counts equal and density interiors within 1e-5;
  - the list length equals the sum of the subcell counts.
"""
import ctypes as C

import numpy as np
import pytest

import gpu_util as U

LD = (2, 3, 3, 3, 5)
RES = 4
RADIUS = 1.0
SC = 2
D3 = (RES // SC) ** 3


def _subcell_counts(v):
    nsc, _, _ = v.subcell_info()
    total = v.node_count(0)[0]
    m = np.zeros(total, np.int32)
    cnt = np.zeros(nsc, np.int32)
    v._chk(v.L.gvdbx_copy_subcell_array(v.h, 0, m.ctypes.data, m.nbytes), "map")
    v._chk(v.L.gvdbx_copy_subcell_array(v.h, 1, cnt.ctypes.data, cnt.nbytes), "cnt")
    return m, cnt


def _brick_keys(bpos):
    b = bpos.astype(np.int64)
    return (b[:, 0] << 42) | (b[:, 1] << 21) | b[:, 2]


def _brick_of_slot(v, chan, slot):
    """One channel brick (br^3 floats, [z, y, x]) read straight from the device."""
    ptr, _ = v.channel_ptr(chan)
    br = RES + 2 * v.apron
    out = np.empty(br ** 3, np.float32)
    v._chk(v.L.gvdbx_retrieve(v.h, out.ctypes.data, ptr + int(slot) * out.nbytes, out.nbytes), "retrieve")
    return out.reshape(br, br, br)


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_more_bricks_than_one_dispatch_row(oracle_mod):
    U.log("case 19.8M 4^3 bricks: generate")
    n = 4_000_000
    p = np.random.default_rng(11).uniform(0.0, 2048.0, (n, 3)).astype(np.float32)
    v = U.new_volume(LD, chans=(0,))
    d = v.to_device(p)
    v.SetPoints(d, n=n)
    v.RebuildTopology(n, RADIUS)
    v.UpdateAtlas()
    ln = v.InsertPointsSubcell(SC, n, RADIUS)
    v.GatherDensity(SC, n, RADIUS, (0, 0, 0), ln, 0)
    v.UpdateApron(0, 0.0)
    v.synchronize()
    nodes = v.get_nodes(0)
    live = np.nonzero(nodes["flags"] == 1)[0]
    U.log(f"gpu done: leaves {len(live)} list length {ln}")
    assert len(live) > (1 << 32) // 256

    gmap, gcnt = _subcell_counts(v)
    assert int(gcnt.astype(np.int64).sum()) == ln

    # ---- topology against the oracle's full rebuild
    t = U.new_tree(oracle_mod, LD, chans=())
    t.rebuild_topology(p, RADIUS)
    t.update_atlas()
    U.compare_topology(v, t)
    del t
    U.log("topology equal")

    # ---- 400 bricks against the oracle on the points that can reach them
    rng = np.random.default_rng(12)
    lkeys = _brick_keys(nodes["pos"][live] // RES)
    order = np.argsort(lkeys)
    home = np.floor(p[rng.choice(n, 300, replace=False)].astype(np.float64) / RES).astype(np.int64)
    hk = _brick_keys(home)
    i = np.clip(np.searchsorted(lkeys[order], hk), 0, len(order) - 1)
    assert (lkeys[order][i] == hk).all()
    chosen = np.unique(np.concatenate([order[i], rng.choice(len(live), 100, replace=False)]))
    cpos = nodes["pos"][live[chosen]].astype(np.int64) // RES
    near = set()
    for b in cpos:
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    near.add(((b[0] + dx) << 42) | ((b[1] + dy) << 21) | (b[2] + dz))
    pk = _brick_keys(np.floor(p.astype(np.float64) / RES).astype(np.int64))
    sub = p[np.isin(pk, np.fromiter(near, np.int64))]
    U.log(f"oracle subset: {len(sub)} points around {len(chosen)} bricks")
    t = U.new_tree(oracle_mod, LD, chans=(0,))
    t.rebuild_topology(sub, RADIUS)
    t.update_atlas()
    t.insert_points_subcell(SC, sub, RADIUS)
    t.gather_density(RADIUS, 0)
    tn = t.nodes(0)
    tslot = t.leaf_slot()
    tch = t.channel(0)
    tsc = t.subcells()
    tlive = np.nonzero(tn["flags"] == 1)[0]
    tkeys = _brick_keys(tn["pos"][tlive] // RES)
    torder = np.argsort(tkeys)
    ck = _brick_keys(cpos)
    j = np.clip(np.searchsorted(tkeys[torder], ck), 0, len(torder) - 1)
    assert (tkeys[torder][j] == ck).all()
    oslot = torder[j]
    opool = tlive[oslot]
    nonzero = 0
    for k in range(len(chosen)):
        assert tslot[opool[k]] == oslot[k]
        g = _brick_of_slot(v, 0, chosen[k])[1:-1, 1:-1, 1:-1].astype(np.float64)
        o = tch[oslot[k], 1:-1, 1:-1, 1:-1]
        err = np.abs(g - o) / np.maximum(np.abs(o), 1.0)
        assert err.max() <= U.TOL, (k, err.max())
        nonzero += int((o > 0).any())
        gp = live[chosen[k]]
        gc = gcnt[gmap[gp] * D3: gmap[gp] * D3 + D3]
        oc = tsc["cnt"][tsc["map"][opool[k]] * D3: tsc["map"][opool[k]] * D3 + D3]
        assert np.array_equal(gc, oc), (k, gc, oc)
    assert nonzero >= 250
    U.log(f"{len(chosen)} bricks equal ({nonzero} with density)")
    v.close()
