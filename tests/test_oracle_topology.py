"""Pins the oracle's topology rules against brute-force set enumeration and tree invariants.

No reference-held vectors exist for this path (the reference has no tests and its points.dat
fixture is absent), so the oracle is checked against what the rules fix independently:
sorted unique key sets per level (P.cuh:352-382 + Thrust sort), the extra-brick box
(P.cuh:467-512), append order (V.cpp:975-997), links (N.cuh:470-505) and the incremental
mark/delink rules (P.cuh:420-465, N.cuh:401-467).
"""
import numpy as np
import pytest

import helpers as H

UNDEF64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def rand_points(n, seed, lo=20.0, hi=180.0):
    g = np.random.default_rng(seed)
    return g.uniform(lo, hi, (n, 3)).astype(np.float32)


def check_links(t, depth):
    """Every child-list entry points at a node inside its slot whose parent points back."""
    R = H.ranges(t.log2dims)
    for l in range(1, depth + 1):
        nodes = t.nodes(l)
        cl = t.childlists(l)
        child = t.nodes(l - 1)
        res = 1 << t.log2dims[l]
        for i, nd in enumerate(nodes):
            ci = int(nd["childlist"]) >> 16
            assert int(nd["childlist"]) & 0xFF == 1 and (int(nd["childlist"]) >> 8) & 0xFF == l
            for b in np.nonzero(cl[ci] != UNDEF64)[0]:
                e = int(cl[ci][b])
                assert (e >> 8) & 0xFF == l - 1 and e & 0xFF == 0
                c = child[e >> 16]
                bx, by, bz = b % res, (b // res) % res, b // (res * res)
                exp = nd["pos"] + np.array([bx, by, bz]) * R[l - 1]
                assert np.array_equal(c["pos"], exp)
                assert int(c["parent"]) == ((i << 16) | (l << 8))


@pytest.mark.parametrize("seed,radius", [(0, 0.5), (1, 1.0), (2, 2.0)])
def test_full_rebuild_matches_set_enumeration(oracle_mod, seed, radius):
    p = rand_points(3000, seed)
    p[:50] -= 200.0  # invalid (negative) points are dropped by every kernel
    t = oracle_mod.Tree()
    t.rebuild_topology(p, radius)
    q, valid = H.valid_q(p)
    d, root = H.depth_root(q)
    assert t.depth() == d and np.array_equal(t.rootpos(), root)
    mn, mx = H.posminmax(q)
    assert np.array_equal(t.posminmax(), np.concatenate([mn, mx]))
    for l in range(d):
        A = H.point_nodes(q, valid, l)
        pos = t.nodes(l)["pos"].astype(np.int64)
        assert np.array_equal(pos[: len(A)], A), f"first-pass pool order at level {l}"
        if radius >= 1:
            box = H.box_nodes(q, valid, l, radius, 2.0, root, d)
            B = box[~H.rows_in(box, A)]
            assert np.array_equal(pos[len(A):], B), f"extra-pass append order at level {l}"
        else:
            assert len(pos) == len(A)
    root_nd = t.nodes(d)
    assert len(root_nd) == 1 and np.array_equal(root_nd[0]["pos"], root)
    assert int(root_nd[0]["parent"]) == 0xFFFFFFFF
    leaves = t.nodes(0)
    assert (leaves["flags"] == 1).all() and (leaves["value"] == -1).all()
    assert all(int(c) == 0xFFFFFFFF for c in leaves["childlist"])
    assert t.used0() == len(leaves)
    check_links(t, d)


def test_state_machine_full_full_incremental(oracle_mod):
    """Calls 1 and 2 rebuild fully (mCurrDepth starts at -1), call 3 is incremental."""
    p = rand_points(2000, 5)
    t = oracle_mod.Tree()
    t.rebuild_topology(p, 1.0)
    n1 = [t.node_count(l) for l in range(5)]
    t.rebuild_topology(p, 1.0)
    n2 = [t.node_count(l) for l in range(5)]
    assert n1 == n2
    # incremental with the same points: every box node exists -> no new key -> unchanged
    before = t.nodes(0).copy()
    t.rebuild_topology(p, 1.0)
    assert np.array_equal(t.nodes(0), before)


def test_incremental_invariants(oracle_mod):
    g = np.random.default_rng(9)
    p = rand_points(4000, 7, 40, 150)
    t = oracle_mod.Tree()
    t.rebuild_topology(p, 1.0)
    t.rebuild_topology(p, 1.0)
    old_pos = t.nodes(0)["pos"].copy()
    p2 = (p + g.normal(0, 6.0, p.shape)).astype(np.float32)
    t.rebuild_topology(p2, 1.0)
    q, valid = H.valid_q(p2)
    d = t.depth()
    leaves = t.nodes(0)
    assert np.array_equal(leaves["pos"][: len(old_pos)], old_pos), "existing nodes keep their pool index"
    live = leaves[leaves["flags"] == 1]
    want = H.box_nodes(q, valid, 0, 1.0, 1.0, t.rootpos(), d)
    got = H.node_set(live["pos"])
    assert np.array_equal(got, want), "live leaves == bricks touched by the incremental boxes"
    assert len(np.unique(live["pos"], axis=0)) == len(live)
    assert t.used0() == len(live)
    # dead leaves are delinked and unreachable; live ones are reachable
    dead = leaves[leaves["flags"] == 0]
    assert all(int(x) == 0xFFFFFFFFFFFFFFFF for x in dead["parent"])
    for nd in live[:200]:
        assert t.find_leaf(nd["pos"]) >= 0
    new = leaves[len(old_pos):]
    assert np.array_equal(H.node_set(new["pos"]), new["pos"].astype(np.int64)), "new nodes appended sorted"
    # inner nodes: flagged iff linked and holding a child
    for l in range(1, d):
        nodes = t.nodes(l)
        cl = t.childlists(l)
        for i, nd in enumerate(nodes):
            has = (cl[int(nd["childlist"]) >> 16] != UNDEF64).any()
            linked = int(nd["parent"]) != 0xFFFFFFFFFFFFFFFF
            assert bool(nd["flags"]) == (linked and has)


def test_update_atlas_slots(oracle_mod):
    p = rand_points(3000, 11)
    t = oracle_mod.Tree(atlas_cx=16, atlas_cy=16, apron=1)
    t.rebuild_topology(p, 1.0)
    t.update_atlas()
    leaves = t.nodes(0)
    slot = t.leaf_slot()
    assert np.array_equal(slot, np.arange(len(leaves)))
    s = np.arange(len(leaves))
    br = 10
    exp = np.stack([(s % 16) * br + 1, ((s // 16) % 16) * br + 1, (s // 256) * br + 1], axis=1)
    assert np.array_equal(leaves["value"], exp)


def test_depth_errors(oracle_mod):
    t = oracle_mod.Tree()
    # all points inside one brick including the -1/+2 margin -> depth 0
    p = np.full((10, 3), 3.5, np.float32)
    with pytest.raises(oracle_mod.OracleError) as e:
        t.rebuild_topology(p, 1.0)
    assert e.value.code == 2
    # extents beyond the tree -> depth error
    p = np.array([[0, 0, 0], [3.0e6, 1, 1]], np.float32)
    with pytest.raises(oracle_mod.OracleError) as e:
        t.rebuild_topology(p, 1.0)
    assert e.value.code == 3


def test_cpu_activate_space_leaf_count(oracle_mod):
    p = rand_points(5000, 13)
    q, valid = H.valid_q(p)
    n = oracle_mod.cpu_activate_space(p)
    assert n == len(H.point_nodes(q, valid, 0))
