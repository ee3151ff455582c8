"""Independent numpy restatements used to pin the oracle (brute force, closed forms)."""
import numpy as np

LD = (3, 3, 3, 3, 5)


def ranges(ld=LD):
    return [1 << int(sum(ld[: l + 1])) for l in range(len(ld))]


def valid_q(pos, trans=(0, 0, 0)):
    q = pos.astype(np.float32) + np.asarray(trans, np.float32)
    return q, (q >= 0).all(axis=1)


def posminmax(q):
    mn = (q.min(axis=0) - np.float32(1.0)).astype(np.float32)
    mx = (q.max(axis=0) + np.float32(2.0)).astype(np.float32)
    return np.maximum(mn, 0).astype(np.float32), np.maximum(mx, 0).astype(np.float32)


def depth_root(q, ld=LD):
    mn, mx = posminmax(q)
    imn, imx = mn.astype(np.int64), mx.astype(np.int64)
    for lev, R in enumerate(ranges(ld)):
        if np.array_equal(imn // R, imx // R):
            return lev, (imn // R) * R
    return -1, None


def node_set(cells):
    """Sorted unique rows (x-major, then y, then z) == ascending reference key order."""
    if len(cells) == 0:
        return np.zeros((0, 3), np.int64)
    u = np.unique(np.asarray(cells, np.int64), axis=0)
    return u


def point_nodes(q, valid, lev, ld=LD):
    R = ranges(ld)[lev]
    return node_set(np.floor(q[valid].astype(np.float64) / R).astype(np.int64) * R)


def box_nodes(q, valid, lev, r, hi_mul, root, depth, ld=LD):
    """Level-lev nodes intersecting [trunc(trunc(q)-r), trunc(trunc(q)+hi_mul*r)] inside the root."""
    R = ranges(ld)[lev]
    Rd = ranges(ld)[depth]
    tq = q[valid].astype(np.int64).astype(np.float32)
    bmin = (tq - np.float32(r)).astype(np.float32).astype(np.int64)
    bmax = (tq + np.float32(hi_mul) * np.float32(r)).astype(np.float32).astype(np.int64)
    lo = np.floor_divide(bmin, R)
    hi = np.floor_divide(bmax, R)
    out = []
    span = (hi - lo).max(axis=0) + 1 if len(lo) else np.zeros(3, np.int64)
    for dx in range(int(span[0])):
        for dy in range(int(span[1])):
            for dz in range(int(span[2])):
                c = lo + np.array([dx, dy, dz])
                ok = (c <= hi).all(axis=1)
                c = c[ok] * R
                ins = ((c >= root) & (c < root + Rd)).all(axis=1)
                out.append(c[ins])
    cells = np.concatenate(out) if out else np.zeros((0, 3), np.int64)
    return node_set(cells)


def rows_in(a, b):
    """Boolean mask of rows of a present in b."""
    if len(b) == 0:
        return np.zeros(len(a), bool)
    sb = {tuple(x) for x in b.tolist()}
    return np.array([tuple(x) in sb for x in a.tolist()], bool)


def subcell_members(q, valid, s_pos, sc, r):
    """Points whose subcell enumeration (P.cuh:275-315) contains subcell corner s_pos."""
    qq = q.astype(np.float32)
    lo = ((qq - np.float32(r)).astype(np.int64) // sc) * sc  # trunc toward zero for q-r > -1
    hq = (qq + np.float32(1.0)).astype(np.float32)
    hq = (hq + np.float32(r)).astype(np.float32)
    hi = (hq.astype(np.int64) // sc) * sc
    # C truncation of negatives in (-1, 0) gives 0; astype(int64) truncates toward zero too.
    ok = valid & (qq[:, 2] != np.float32(1e10)) & (lo <= s_pos).all(axis=1) & (hi >= s_pos).all(axis=1)
    return np.nonzero(ok)[0]
