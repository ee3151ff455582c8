"""The SPH workload at full size on the GPU: 4,194,304 particles on the jittered lattice of the
lower 40 % of a 256^3 box, velocities N(0, 1), 60 steps of q += v dt (dt = 0.5, on the device
through gvdbx_advance_points; its positions must equal numpy's float32 q + v * dt bit for bit
every step), each step RebuildTopology(r = 1) (the reference state machine: full, full, then incremental) + UpdateAtlas +
InsertPointsSubcell(4) + GatherDensity with the three velocity means + UpdateApron of the four
channels -- the call sequence of gFluidSurface's frame (main_fluid_surface.cpp:320-356).

The oracle cannot follow 60 steps of 4M points in test time, so:
  - steps 0 to 5: the oracle follows the same state machine; pools and child lists byte for byte;
  - every step: every valid point's brick is live, the used count is the flagged leaves, the
    subcell counts sum to the list length;
  - steps 0, 1, 2, 30 and 59: 48 random live bricks against the oracle run on exactly the points
    that can reach them (a point reaches a brick's subcells only from that brick or a neighbour:
    |p - subcell| <= r + 1 < 8), subcell counts equal, density and velocity means within 1e-5;
  - every incremental step: when the boxes [trunc(q) - r, trunc(q) + r] of the valid points
    (gvdbCalcIncreExtraBrickId, P.cuh:420-465) touch a brick that was not live, the live bricks
    become exactly the bricks the boxes touch (the markers start at zero every step); when they
    touch none, the step finds no new node and leaves the topology as it was (V.cpp:1066).
"""
import ctypes as C

import numpy as np
import pytest

import gpu_util as U

R = 1.0
DT = np.float32(0.5)
STEPS = 60
ORACLE_TOPO_STEPS = 6
BRICK_STEPS = (0, 1, 2, 30, 59)


def key_of(b):
    b = np.asarray(b, np.int64)
    return (b[:, 0] << 42) | (b[:, 1] << 21) | b[:, 2]


def subcell_counts(v):
    nsc, _, _ = v.subcell_info()
    m = np.zeros(v.node_count(0)[0], np.int32)
    cnt = np.zeros(nsc, np.int32)
    v._chk(v.L.gvdbx_copy_subcell_array(v.h, 0, m.ctypes.data, m.nbytes), "map")
    v._chk(v.L.gvdbx_copy_subcell_array(v.h, 1, cnt.ctypes.data, cnt.nbytes), "cnt")
    return m, cnt


def incremental_live_set(q):
    """Keys of the leaf bricks touched by [trunc(trunc(q) - r), trunc(trunc(q) + r)] per valid q."""
    valid = (q >= 0).all(1)
    tq = q[valid].astype(np.int64).astype(np.float32)
    lo = (tq - np.float32(R)).astype(np.float32).astype(np.int64) // 8
    hi = (tq + np.float32(R)).astype(np.float32).astype(np.int64) // 8
    keys = []
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c = lo + np.array([dx, dy, dz])
                ok = (c <= hi).all(1) & (c >= 0).all(1)
                keys.append(key_of(c[ok]))
    return np.unique(np.concatenate(keys))


def check_bricks(oracle_mod, v, q, vel, step, rng):
    nodes = v.get_nodes(0)
    live = np.nonzero(nodes["flags"] == 1)[0]
    chosen = rng.choice(len(live), 48, replace=False)
    cpos = nodes["pos"][live[chosen]].astype(np.int64) // 8
    near = key_of(np.concatenate([cpos + np.array([dx, dy, dz]) for dx in (-1, 0, 1) for dy in (-1, 0, 1)
                                  for dz in (-1, 0, 1)]))
    pk = key_of(np.floor(q.astype(np.float64) / 8).astype(np.int64))
    sel = np.isin(pk, near) & (q >= 0).all(1)
    sub, subv = q[sel], vel[sel]
    t = U.new_tree(oracle_mod, chans=(0, 1, 2, 3))
    t.rebuild_topology(sub, R)
    t.update_atlas()
    t.insert_points_subcell(4, sub, R)
    t.gather_density(R, 0)
    t.gather_velocity(subv, (1, 2, 3))
    tn = t.nodes(0)
    tlive = np.nonzero(tn["flags"] == 1)[0]
    tkeys = key_of(tn["pos"][tlive].astype(np.int64) // 8)
    order = np.argsort(tkeys)
    idx = np.searchsorted(tkeys[order], key_of(cpos))
    opool = tlive[order][idx]
    assert np.array_equal(tkeys[order][idx], key_of(cpos)), "a chosen brick is missing from the oracle's tree"
    oslot = t.leaf_slot()[opool]
    tsc = t.subcells()
    gmap, gcnt = subcell_counts(v)
    for c in range(4):
        g = v.get_channel(c)[chosen, 1:-1, 1:-1, 1:-1].astype(np.float64)
        o = t.channel(c)[oslot, 1:-1, 1:-1, 1:-1]
        err = np.abs(g - o) / np.maximum(np.abs(o), 1.0)
        assert err.max() <= U.TOL, (step, c, err.max())
    for i in range(48):
        gc = gcnt[gmap[live[chosen[i]]] * 8: gmap[live[chosen[i]]] * 8 + 8]
        oc = tsc["cnt"][tsc["map"][opool[i]] * 8: tsc["map"][opool[i]] * 8 + 8]
        assert np.array_equal(gc, oc), (step, i, gc, oc)
    U.log(f"  step {step}: 48 bricks equal the oracle on {len(sub)} points")


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_sph_4m_60_steps(oracle_mod):
    from gvdbx import synth
    U.log("case SPH 4M x 60 steps: generate")
    p, vel = synth.sph_block(seed=1)
    n = len(p)
    assert n == 4_194_304
    v = U.new_volume(chans=(0, 1, 2, 3))
    dp, dv = v.to_device(p), v.to_device(vel)
    v.SetPoints(dp, dv, n=n)
    t = U.new_tree(oracle_mod)
    rng = np.random.default_rng(60)
    cur = p
    changed = 0
    for step in range(STEPS):
        if step:
            cur = (cur + vel * DT).astype(np.float32)
            v.AdvancePoints(n, DT)
            got = dp.download(np.float32, 3 * n)
            assert got.tobytes() == cur.tobytes(), f"step {step}: device displacement differs"
        v.RebuildTopology(n, R)
        v.UpdateAtlas()
        ln = v.InsertPointsSubcell(4, n, R)
        v.GatherDensityVelocity(4, n, R, (0, 0, 0), ln, 0, (1, 2, 3))
        v.UpdateApron()
        if step < ORACLE_TOPO_STEPS:
            t.rebuild_topology(cur, R)
            t.update_atlas()
            U.compare_topology(v, t)
        d, root, _, _ = v.depth()
        assert d == 2 and (root == 0).all()
        nodes = v.get_nodes(0)
        flagged = nodes["flags"] == 1
        assert v.node_count(0)[1] == flagged.sum() == v.num_slots()[0]
        _, cnt = subcell_counts(v)
        assert cnt.sum() == ln and (cnt >= 0).all()
        valid = (cur >= 0).all(1)
        lkeys = np.sort(key_of(nodes["pos"][flagged].astype(np.int64) // 8))
        hk = key_of(np.floor(cur[valid].astype(np.float64) / 8).astype(np.int64))
        i = np.clip(np.searchsorted(lkeys, hk), 0, len(lkeys) - 1)
        assert (lkeys[i] == hk).all(), f"step {step}: a valid point's brick is not live"
        if step >= 2:
            box = incremental_live_set(cur)
            grows = not np.isin(box, prev).all()
            want = box if grows else prev
            assert np.array_equal(lkeys, want), f"step {step}: live set ({'new bricks' if grows else 'unchanged'})"
            changed += grows
        prev = lkeys
        if step in BRICK_STEPS:
            check_bricks(oracle_mod, v, cur, vel, step, rng)
        U.log(f"  step {step}: {flagged.sum()} bricks ({len(nodes)} in the pool), {valid.sum()} valid points, "
              f"{ln} entries")
        if step == ORACLE_TOPO_STEPS - 1:
            del t
    assert changed > 0  # some steps did add bricks
    U.log(f"{changed} incremental steps added bricks")
    v.close()
