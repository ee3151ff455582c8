"""Autotuner of the batched-cloud path: picks the leaf brick size, the subcell size and the launch
shape of the per-brick kernels (threads per workgroup, grid or persistent form).

Every candidate gives the exact field of its own topology (the gathers are exact for any
brick / subcell size, and every launch shape gives the same bits), so the choice is purely one
of speed.  For each (brick, subcell) pair the batch's topology is built once; then, per launch
shape, InsertPointsSubcell + GatherDensity + UpdateApron is recorded as one graph and launched
a few times, and the fastest launch wins.  The cell of each cloud is the smallest inner-level
range that holds every cloud for that brick size (gvdbx_batch_set_points rejects a cell that is
too small).
"""
import time

from .lib import GvdbxError
from .volume import VolumeGVDB

# (leaf log2, subcell) candidates: 8^3, 16^3 and 4^3 bricks; inner levels stay at 3 (8^3
# children) and the root at 5
CANDIDATES = ((3, 4), (3, 2), (4, 4), (4, 8), (2, 2), (2, 4))
# (threads per brick workgroup, persistent); 0 = chosen from the items per brick
SHAPES = ((0, False), (64, False), (128, False), (256, False), (64, True), (128, True), (256, True))


def levels_for(leaf_log2):
    return (leaf_log2, 3, 3, 3, 5)


def cell_candidates(leaf_log2):
    """Ranges (log2) of the inner levels below the root for this brick size, smallest first."""
    ld = levels_for(leaf_log2)
    acc, out = 0, []
    for l in range(len(ld) - 1):
        acc += ld[l]
        if l >= 1:
            out.append(acc)
    return out


def run_batch(v, n, radius, subcell, chan=0):
    """One batched step: RebuildTopology + UpdateAtlas + InsertPointsSubcell + GatherDensity +
    UpdateApron; returns the subcell-list length."""
    v.RebuildTopology(n, radius)
    v.UpdateAtlas()
    return insert_gather_apron(v, n, radius, subcell, chan)


def insert_gather_apron(v, n, radius, subcell, chan=0):
    """InsertPointsSubcell + GatherDensity + UpdateApron on the current topology."""
    ln = v.InsertPointsSubcell(subcell, n, radius)
    v.GatherDensity(subcell, n, radius, (0, 0, 0), ln, chan)
    v.UpdateApron(chan, 0.0)
    return ln


def record_step(v, n, radius, subcell, chan=0):
    """Run insert + gather + apron once (sizing every buffer), then record it as the volume's
    graph (VolumeGVDB.GraphLaunch runs it)."""
    insert_gather_apron(v, n, radius, subcell, chan)
    v.GraphBegin()
    try:
        insert_gather_apron(v, n, radius, subcell, chan)
    finally:
        v.GraphEnd()


def time_launches(v, reps):
    """Best wall time (ms) of `reps` synchronised graph launches."""
    best = float("inf")
    for _ in range(reps):
        v.synchronize()
        t0 = time.perf_counter()
        v.GraphLaunch()
        v.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def new_batch_volume(leaf_log2, device=0):
    v = VolumeGVDB(device)
    v.configure_levels(levels_for(leaf_log2))
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, apron=1)
    return v


def cell_for(leaf_log2, max_coord, radius):
    """Smallest inner-level cell (log2) whose guard limit, 2^cell - brick - 2r - 3, lies above
    every valid coordinate of the clouds (max_coord)."""
    for cell in cell_candidates(leaf_log2):
        if max_coord < (1 << cell) - (1 << leaf_log2) - 2.0 * radius - 3.0:
            return cell
    raise GvdbxError(4, "autotune", f"clouds reach {max_coord}: no cell holds them")


def max_valid_coord(clouds):
    """Largest coordinate of any point with all coordinates >= 0, over host arrays."""
    m = 0.0
    for c in clouds:
        ok = (c >= 0).all(1)
        if ok.any():
            m = max(m, float(c[ok].max()))
    return m


def bind(v, leaf_log2, dpos, counts, radius, max_coord):
    """Bind the batch with the smallest cell that holds it; returns the cell log2."""
    cell = cell_for(leaf_log2, max_coord, radius)
    v.BatchSetPoints(dpos, counts, cell_log2=cell, max_radius=radius)
    return cell


def tune(upload, counts, max_coord, radius=1.0, candidates=CANDIDATES, shapes=SHAPES, reps=3, log=None):
    """upload(v) -> device buffer of the concatenated clouds on volume v; max_coord = the largest
    valid coordinate (max_valid_coord).  Returns {"leaf_log2", "subcell", "block", "persistent",
    "cell_log2", "ms"} of the fastest candidate (ms = one graph launch of insert + gather +
    apron) and every timing."""
    n = int(sum(int(c) for c in counts))
    timings = []
    for leaf, sc in candidates:
        if sc > (1 << leaf):
            continue
        v = new_batch_volume(leaf)
        try:
            d = upload(v)
            cell = bind(v, leaf, d, counts, radius, max_coord)
            v.RebuildTopology(n, radius)
            v.UpdateAtlas()
            for block, pers in shapes:
                v.SetLaunch(block, pers)
                record_step(v, n, radius, sc)
                ms = time_launches(v, reps)
                v.GraphRelease()
                timings.append({"leaf_log2": leaf, "subcell": sc, "block": block, "persistent": bool(pers),
                                "cell_log2": cell, "ms": round(ms, 4)})
                if log:
                    log(f"autotune leaf 2^{leaf} subcell {sc} block {block or 'auto'} "
                        f"{'persistent' if pers else 'grid'}: {ms:.3f} ms")
        finally:
            v.close()
    best = min(timings, key=lambda t: t["ms"])
    return dict(best, candidates=timings)
