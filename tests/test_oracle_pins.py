"""Pins the oracle's accumulate rules, FP16 velocity words and the scatter-reduce level set.

Each check restates the rule independently of the oracle code:
  - AccumulateTopology (V.cpp:1158-1178, marker path V.cpp:1034-1042): the live leaves become the
    old live set united with the brick boxes [trunc(q) - r, trunc(q) + r] of the new points, old
    nodes keep their pool index and bytes, new nodes are appended per level in key order;
  - accumulating gathers (P.cuh:700-705, 740-745; velocity and colour means P.cuh:717-731):
    max(old, r - |p - v|) / min(old, |p - v| - r) over the subcell's members, and
    (old + sum v) / (cnt + 1), with the members from the voxel-centric membership rule;
  - FP16 velocity words (P.cuh:200-207): hand-computed words, the word of each entry belongs to
    that entry's point, and the decoding error bound;
  - ScatterReduceLevelSet (the build's definition from P.cuh:736-773): equal to GatherLevelSet bit
    for bit, and to brute force from any initial value.
"""
import numpy as np
import pytest

import helpers as H


def cloud(n, seed, lo, hi):
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3)).astype(np.float32)


def live_positions(t):
    nd = t.nodes(0)
    return {tuple(int(x) for x in p) for p in nd["pos"][nd["flags"] == 1]}


def key_sorted(cells):
    """Rows in ascending reference key order (x, then y, then z: P.cuh:373-381)."""
    cells = np.asarray(sorted(cells), np.int64).reshape(-1, 3)
    return cells


@pytest.mark.parametrize("r", [1.0, 2.0])
def test_accumulate_topology_is_union_append_in_key_order(oracle_mod, r):
    a = cloud(1500, 11, 30.0, 90.0)
    b = cloud(800, 12, 80.0, 140.0)
    t = oracle_mod.Tree()
    t.rebuild_topology(a, r)  # the first call is a full rebuild
    d, root = t.depth(), t.rootpos().astype(np.int64)
    old = [t.nodes(l) for l in range(d + 1)]
    old_live = live_positions(t)
    t.accumulate_topology(b, r)
    assert t.depth() == d and np.array_equal(t.rootpos(), root)
    qb, vb = H.valid_q(b)
    for l in range(d):
        box = {tuple(x) for x in H.box_nodes(qb, vb, l, r, 1.0, root, d).tolist()}
        now = t.nodes(l)
        n_old = len(old[l])
        # old nodes keep their index and their bytes (the marker path keeps every live leaf)
        assert now[:n_old].tobytes() == old[l].tobytes(), f"level {l}: an old node changed"
        if l == 0:
            existing = old_live
        else:
            existing = {tuple(int(x) for x in p) for p in old[l]["pos"]}
        want_new = key_sorted(box - existing)
        got_new = now["pos"][n_old:].astype(np.int64)
        assert np.array_equal(got_new, want_new), f"level {l}: appended nodes differ"
        assert (now["flags"][n_old:] == 1).all()
    assert live_positions(t) == old_live | {tuple(x) for x in H.box_nodes(qb, vb, 0, r, 1.0, root, d).tolist()}
    assert t.used0() == len(live_positions(t))
    nd = t.nodes(0)
    for i in np.nonzero(nd["flags"] == 1)[0]:
        assert t.find_leaf(nd["pos"][i]) == i  # every live leaf is linked


def test_accumulate_topology_without_new_nodes_changes_nothing(oracle_mod):
    """V.cpp:1066: no new key, no change (flags included)."""
    a = cloud(1500, 13, 30.0, 90.0)
    t = oracle_mod.Tree()
    t.rebuild_topology(a, 1.0)
    d = t.depth()
    before = [t.nodes(l).tobytes() for l in range(d + 1)]
    inner = a[(a > 40.0).all(1) & (a < 80.0).all(1)][:200]  # boxes inside the live set
    t.accumulate_topology(inner, 1.0)
    assert [t.nodes(l).tobytes() for l in range(d + 1)] == before


def setup_fields(oracle_mod, p, r, sc, chans):
    t = oracle_mod.Tree()
    t.rebuild_topology(p, r)
    for c in chans:
        t.add_channel(c)
    t.update_atlas()
    t.insert_points_subcell(sc, p, r)
    return t


def voxel_iter(t, sc, stride):
    """(subcell k, local offset, voxel v, atlas index tuple) for every stride-th subcell."""
    s = t.subcells()
    nodes = t.nodes(0)
    slot = t.leaf_slot()
    a = t.apron
    for k in range(0, len(s["cnt"]), stride):
        lf = nodes[s["nid"][k]]
        for u in range(sc ** 3):
            off = np.array([u % sc, (u // sc) % sc, u // (sc * sc)])
            v = s["pos"][k] + off
            loc = v - lf["pos"] + a
            yield k, v, (slot[s["nid"][k]], loc[2], loc[1], loc[0])


def test_accumulating_density_and_levelset_gathers(oracle_mod):
    """The accumulate forms start from the voxel's old value instead of 0 / 3.0: old values are
    taken from another field (a level set for the density, a density for the level set), so a
    rule that ignored them, or read another voxel's, would show."""
    p = cloud(1200, 21, 30.0, 80.0)
    r, sc = 1.5, 4
    t = setup_fields(oracle_mod, p, r, sc, (0,))
    q, valid = H.valid_q(p)
    qd = q.astype(np.float64)
    s = t.subcells()
    t.gather_levelset(r, 0)
    old = t.channel(0)
    t.gather_density(r, 0, accum=True)
    new = t.channel(0)
    checked = 0
    for k, v, ix in voxel_iter(t, sc, 5):
        mem = H.subcell_members(q, valid, s["pos"][k], sc, r)
        cand = (r - np.sqrt(((qd[mem] - v) ** 2).sum(1))).max() if len(mem) else -np.inf
        assert new[ix] == max(old[ix], cand)
        checked += old[ix] > cand
    assert checked > 0  # some voxels keep their old value
    t.gather_density(r, 0)
    old = t.channel(0)
    t.gather_levelset(r, 0, accum=True)
    new = t.channel(0)
    for k, v, ix in voxel_iter(t, sc, 5):
        mem = H.subcell_members(q, valid, s["pos"][k], sc, r)
        cand = (np.sqrt(((qd[mem] - v) ** 2).sum(1)) - r).min() if len(mem) else np.inf
        assert new[ix] == min(old[ix], cand)


def test_accumulating_velocity_means(oracle_mod):
    """(old + sum of the members' velocities) / (cnt + 1) per voxel (P.cuh:717-731 for float3)."""
    p = cloud(900, 22, 30.0, 80.0)
    g = np.random.default_rng(5)
    v1 = g.standard_normal(p.shape).astype(np.float32)
    v2 = g.standard_normal(p.shape).astype(np.float32)
    t = setup_fields(oracle_mod, p, 1.0, 4, (0, 1, 2))
    q, valid = H.valid_q(p)
    s = t.subcells()
    t.gather_velocity(v1, (0, 1, 2))
    old = [t.channel(c) for c in range(3)]
    t.gather_velocity(v2, (0, 1, 2), accum=True)
    new = [t.channel(c) for c in range(3)]
    for k, v, ix in voxel_iter(t, 4, 3):
        mem = H.subcell_members(q, valid, s["pos"][k], 4, 1.0)
        for c in range(3):
            want = (old[c][ix] + v2[mem, c].astype(np.float64).sum()) / (len(mem) + 1)
            assert abs(new[c][ix] - want) <= 1e-12 * max(1.0, abs(want))


def test_fp16_velocity_words_hand_computed_and_bounded(oracle_mod):
    """ushort((v - lo) / range * 65535) in single precision, for the velocity of the entry's own
    point.  Over lo = (-2, -2, -2), range = (5, 5, 5): v = (-1, 0, 2.5) -> 1/5, 2/5, 4.5/5 of
    65535 = 13107.0002, 26214.0004, 58981.498 -> (13107, 26214, 58981); v = lo -> 0; v beyond
    lo + range saturates at 65535; and the decoding u / 65535 * range + lo lies within
    range / 65535 of the velocity for every entry."""
    p = np.array([[10.5, 10.5, 10.5], [40.25, 30.5, 60.75], [150.0, 150.0, 150.0]], np.float32)
    vel = np.array([[-1.0, 0.0, 2.5], [-2.0, -2.0, -2.0], [9.0, 9.0, 9.0]], np.float32)
    t = setup_fields(oracle_mod, p, 1.0, 4, (0,))
    s = t.subcells()
    lo, rg = np.float32([-2.0, -2.0, -2.0]), np.float32([5.0, 5.0, 5.0])
    w = t.quantize_list_vel(vel, lo, rg)
    want = {0: (13107, 26214, 58981), 1: (0, 0, 0), 2: (65535, 65535, 65535)}
    for pt, word in want.items():
        rows = w[np.nonzero(s["list"] == pt)[0]]
        assert len(rows) > 0 and (rows == np.array(word, np.uint16)).all(), (pt, rows[:2])
    g = np.random.default_rng(6)
    pts = cloud(2000, 23, 20.0, 70.0)
    v = g.uniform(-3.0, 4.0, pts.shape).astype(np.float32)
    t = setup_fields(oracle_mod, pts, 1.0, 4, (0,))
    s = t.subcells()
    lo, rg = v.min(0), (v.max(0) - v.min(0)).astype(np.float32)
    w = t.quantize_list_vel(v, lo, rg)
    dec = w.astype(np.float64) / 65535.0 * rg + lo
    err = np.abs(dec - v[s["list"]])
    assert (err <= rg / 65535.0 * 1.0001 + 1e-6).all()
    assert len(np.unique(w[:, 0])) > 100  # the words carry the values, not a constant


@pytest.mark.parametrize("r,init", [(2.0, 3.0), (1.0, 0.75)])
def test_scatter_reduce_levelset_equals_gather_and_brute_force(oracle_mod, r, init):
    p = cloud(1000, 24, 30.0, 80.0)
    sc = 4
    t = setup_fields(oracle_mod, p, r, sc, (0, 1))
    t.fill_channel(0, 55.0)  # apron voxels keep it (neither gather writes them)
    t.fill_channel(1, 55.0)
    t.scatter_reduce_levelset(r, 0, init)
    if init == 3.0:
        t.gather_levelset(r, 1)
        assert t.channel(0).tobytes() == t.channel(1).tobytes()
    q, valid = H.valid_q(p)
    qd = q.astype(np.float64)
    s = t.subcells()
    ch = t.channel(0)
    for k, v, ix in voxel_iter(t, sc, 4):
        mem = H.subcell_members(q, valid, s["pos"][k], sc, r)
        cand = (np.sqrt(((qd[mem] - v) ** 2).sum(1)) - r).min() if len(mem) else np.inf
        assert ch[ix] == min(init, cand)
