"""Shared GPU-vs-oracle pipeline drivers for the gpu-marked parity tests."""
import sys

import numpy as np

TOL = 1e-5  # |gpu - oracle| <= TOL * max(|oracle|, 1)  (FP32 channels vs FP64 oracle)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def new_volume(ld=(3, 3, 3, 3, 5), chans=(0,), apron=1):
    from gvdbx import VolumeGVDB
    v = VolumeGVDB(0)
    v.configure_levels(ld)
    v.SetChannelDefault(16, 16, 1)
    for c in chans:
        v.AddChannel(c, apron=apron)
    return v


def new_tree(oracle_mod, ld=(3, 3, 3, 3, 5), chans=(0,), apron=1):
    t = oracle_mod.Tree(ld, 16, 16, apron)
    for c in chans:
        t.add_channel(c)
    return t


def compare_topology(v, t):
    d, root, pmin, pmax = v.depth()
    assert d == t.depth(), (d, t.depth())
    assert np.array_equal(root, t.rootpos())
    assert np.array_equal(np.concatenate([pmin, pmax]), t.posminmax())
    for l in range(len(v.log2dims)):
        gn, on = v.get_nodes(l), t.nodes(l)
        assert len(gn) == len(on), f"level {l}: {len(gn)} vs {len(on)} nodes"
        if len(gn):
            bad = np.nonzero(gn.view(np.uint8).reshape(-1, 64) != on.view(np.uint8).reshape(-1, 64))
            assert gn.tobytes() == on.tobytes(), f"level {l} pool differs at nodes {np.unique(bad[0])[:10]}"
        if l >= 1 and len(gn):
            assert np.array_equal(v.get_childlists(l), t.childlists(l)), f"level {l} child lists differ"
    assert v.node_count(0)[1] == t.used0()


def compare_subcells(v, t):
    g, o = v.get_subcells(), t.subcells()
    for k in ("map", "cnt", "off", "pos", "nid"):
        assert np.array_equal(g[k], o[k]), f"subcell {k} differs"
    # list contents per subcell as multisets (the insert order within a subcell is free)
    seg = np.repeat(np.arange(len(o["cnt"])), o["cnt"])
    for lp in (g["list_pos"], o["list_pos"]):
        assert len(lp) == len(seg)
    og = np.lexsort((g["list_pos"][:, 2], g["list_pos"][:, 1], g["list_pos"][:, 0], seg))
    oo = np.lexsort((o["list_pos"][:, 2], o["list_pos"][:, 1], o["list_pos"][:, 0], seg))
    assert np.array_equal(g["list_pos"][og], o["list_pos"][oo]), "subcell list contents differ"
    return g


def compare_channel(v, t, chan, interior_only=False, tol=TOL):
    g = v.get_channel(chan).astype(np.float64)
    o = t.channel(chan)[: len(g)]
    assert g.shape == o.shape, (g.shape, o.shape)
    if interior_only:
        a = v.apron
        g = g[:, a:-a, a:-a, a:-a]
        o = o[:, a:-a, a:-a, a:-a]
    err = np.abs(g - o) / np.maximum(np.abs(o), 1.0)
    worst = float(err.max()) if err.size else 0.0
    assert worst <= tol, f"channel {chan}: max rel err {worst:.3e}"
    return worst


def run_pipeline(v, t, pos, radius=1.0, sc=4, vel=None, density=True, levelset=False, apron_bound=0.0,
                 chans_vel=None, compare=True, steps=1, move=None):
    """Drive RebuildTopology -> UpdateAtlas -> InsertPointsSubcell -> gather -> UpdateApron on
    both sides; returns the device buffers so callers can keep using them."""
    n = len(pos)
    dpos = v.to_device(pos)
    dvel = v.to_device(vel) if vel is not None else None
    v.SetPoints(dpos, dvel, n=n)
    cur = pos
    for step in range(steps):
        if step > 0 and move is not None:
            cur = move(cur, step)
            dpos.upload(cur)
        v.RebuildTopology(n, radius)
        t.rebuild_topology(cur, radius)
        v.UpdateAtlas()
        t.update_atlas()
        if compare:
            compare_topology(v, t)
        ln = v.InsertPointsSubcell(sc, n, radius)
        lo = t.insert_points_subcell(sc, cur, radius)
        assert ln == lo, (ln, lo)
        if compare:
            compare_subcells(v, t)
        if density:
            v.GatherDensity(sc, n, radius, (0, 0, 0), ln, 0)
            t.gather_density(radius, 0)
        if levelset:
            v.GatherLevelSet(sc, n, radius, (0, 0, 0), ln, 0)
            t.gather_levelset(radius, 0)
        if chans_vel is not None:
            v.GatherVelocity(*chans_vel)
            t.gather_velocity(vel, chans_vel)
        if compare:
            compare_channel(v, t, 0, interior_only=True)
        v.UpdateApron(0, apron_bound)
        t.update_apron(0, apron_bound)
        if chans_vel is not None:
            for c in chans_vel:
                v.UpdateApron(c, 0.0)
                t.update_apron(c, 0.0)
        if compare:
            compare_channel(v, t, 0)
            if chans_vel is not None:
                for c in chans_vel:
                    compare_channel(v, t, c)
        log(f"  step {step}: depth {t.depth()} leaves {t.node_count(0)} used {t.used0()} len {lo}")
    return dpos, dvel
