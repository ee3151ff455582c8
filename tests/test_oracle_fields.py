"""Pins the oracle's subcell binning, gathers, apron and trilinear rules.

Checks are independent of the oracle code: a voxel-centric membership rule for the
subcell lists (P.cuh:275-315 enumerates point-centrically), closed forms for the
density field, brute force for the level set, a neighbour-lookup restatement of the
apron, and affine/partition-of-unity/adjoint/finite-difference checks for trilinear.
"""
import numpy as np
import pytest

import helpers as H


def small_cloud(n=1500, seed=3, lo=30.0, hi=90.0):
    g = np.random.default_rng(seed)
    return g.uniform(lo, hi, (n, 3)).astype(np.float32)


def build(oracle_mod, p, r=1.0, sc=4, chans=(0,)):
    t = oracle_mod.Tree()
    t.rebuild_topology(p, r)
    for c in chans:
        t.add_channel(c)
    t.update_atlas()
    ln = t.insert_points_subcell(sc, p, r)
    return t, ln


@pytest.mark.parametrize("r,sc", [(1.0, 4), (2.0, 4), (1.0, 2), (0.5, 8)])
def test_subcell_lists_match_membership(oracle_mod, r, sc):
    p = small_cloud()
    p[0, 2] = np.float32(1e10)  # NOHIT is skipped by the subcell kernels
    t = oracle_mod.Tree()
    t.rebuild_topology(p[1:], r)  # topology without the NOHIT point (it would blow the depth)
    t.add_channel(0)
    t.update_atlas()
    ln = t.insert_points_subcell(sc, p, r)
    s = t.subcells()
    q, valid = H.valid_q(p)
    assert ln == s["cnt"].sum() == len(s["list"])
    assert np.array_equal(s["off"], np.concatenate([[0], np.cumsum(s["cnt"])[:-1]]))
    leaves = t.nodes(0)
    for k in range(len(s["cnt"])):
        mem = H.subcell_members(q, valid, s["pos"][k], sc, r)
        lst = s["list"][s["off"][k]: s["off"][k] + s["cnt"][k]]
        assert np.array_equal(lst, mem), f"subcell {k}"
        lf = leaves[s["nid"][k]]
        assert ((s["pos"][k] - lf["pos"]) % sc == 0).all() and ((s["pos"][k] - lf["pos"]) < 8).all()


def test_gather_density_closed_form(oracle_mod):
    p = small_cloud(1200, 4)
    r = 1.0
    t, _ = build(oracle_mod, p, r)
    t.gather_density(r, 0)
    ch = t.channel(0)
    leaves = t.nodes(0)
    slot = t.leaf_slot()
    q = p.astype(np.float64)
    for i in range(0, len(leaves), 7):
        base = leaves[i]["pos"]
        vz, vy, vx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
        v = np.stack([vx, vy, vz], -1).reshape(-1, 3) + base
        dmin = np.sqrt(((v[:, None, :] - q[None]) ** 2).sum(-1)).min(1)
        want = np.maximum(0.0, r - dmin).reshape(8, 8, 8)
        got = ch[slot[i], 1:9, 1:9, 1:9]
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_gather_levelset_closed_form_and_bruteforce(oracle_mod):
    p = small_cloud(1200, 5)
    r, sc = 2.0, 4
    t, _ = build(oracle_mod, p, r, sc)
    t.gather_levelset(r, 0)
    ch = t.channel(0)
    s = t.subcells()
    q, valid = H.valid_q(p)
    qd = q.astype(np.float64)
    leaves = t.nodes(0)
    slot = t.leaf_slot()
    for k in range(0, len(s["cnt"]), 5):
        mem = H.subcell_members(q, valid, s["pos"][k], sc, r)
        lf = leaves[s["nid"][k]]
        for t_ in range(sc ** 3):
            off = np.array([t_ % sc, (t_ // sc) % sc, t_ // (sc * sc)])
            v = s["pos"][k] + off
            dist_all = np.sqrt(((qd[valid] - v) ** 2).sum(1))
            want = min(3.0, (np.sqrt(((qd[mem] - v) ** 2).sum(1)) - r).min()) if len(mem) else 3.0
            loc = v - lf["pos"] + 1
            got = ch[slot[s["nid"][k]], loc[2], loc[1], loc[0]]
            assert abs(got - want) < 1e-12
            if want <= 0:  # inside the band the value is the global nearest distance
                assert abs(got - (dist_all.min() - r)) < 1e-12


def test_velocity_mean(oracle_mod):
    p = small_cloud(800, 6)
    g = np.random.default_rng(1)
    vel = g.standard_normal((len(p), 3)).astype(np.float32)
    t, _ = build(oracle_mod, p, 1.0, 4, chans=(0, 1, 2))
    t.gather_velocity(vel, (0, 1, 2))
    s = t.subcells()
    q, valid = H.valid_q(p)
    chs = [t.channel(c) for c in range(3)]
    leaves = t.nodes(0)
    slot = t.leaf_slot()
    for k in range(0, len(s["cnt"]), 3):
        mem = H.subcell_members(q, valid, s["pos"][k], 4, 1.0)
        want = vel[mem].astype(np.float64).mean(0) if len(mem) else np.zeros(3)
        loc = s["pos"][k] - leaves[s["nid"][k]]["pos"] + 1
        for c in range(3):
            got = chs[c][slot[s["nid"][k]], loc[2]: loc[2] + 4, loc[1]: loc[1] + 4, loc[0]: loc[0] + 4]
            np.testing.assert_allclose(got, want[c], rtol=1e-12, atol=1e-12)


def test_apron_neighbour_invariant(oracle_mod):
    p = small_cloud(1500, 8)
    t, _ = build(oracle_mod, p, 1.0)
    t.gather_density(1.0, 0)
    t.update_apron(0, -7.0)
    ch = t.channel(0)
    leaves = t.nodes(0)
    slot = t.leaf_slot()
    where = {tuple(lf["pos"]): i for i, lf in enumerate(leaves) if lf["flags"]}
    for i in range(0, len(leaves), 3):
        base = leaves[i]["pos"]
        for z in range(10):
            for y in range(10):
                for x in range(10):
                    if 1 <= x <= 8 and 1 <= y <= 8 and 1 <= z <= 8:
                        continue
                    g = base + np.array([x, y, z]) - 1
                    nb = tuple((g // 8) * 8)
                    if nb in where:
                        j = where[nb]
                        l = g - leaves[j]["pos"] + 1
                        want = ch[slot[j], l[2], l[1], l[0]]
                    else:
                        want = -7.0
                    assert ch[slot[i], z, y, x] == want


def _affine_tree(oracle_mod):
    p = small_cloud(4000, 9, 40, 80)  # dense: interior bricks fully surrounded
    t, _ = build(oracle_mod, p, 1.0)
    t.fill_channel(0, 0.0)
    return t, p


def test_trilinear_reproduces_affine_fields(oracle_mod):
    t, p = _affine_tree(oracle_mod)
    c0, c = 0.25, np.array([0.5, -1.25, 2.0])
    # set interior voxel g to c0 + c.(g + 0.5) via the splat-free path: fill through the API
    leaves = t.nodes(0)
    slot = t.leaf_slot()
    ch = t.channel(0)
    vz, vy, vx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    for i, lf in enumerate(leaves):
        gx, gy, gz = vx + lf["pos"][0], vy + lf["pos"][1], vz + lf["pos"][2]
        ch[slot[i], 1:9, 1:9, 1:9] = c0 + c[0] * (gx + 0.5) + c[1] * (gy + 0.5) + c[2] * (gz + 0.5)
    import ctypes as C
    ptr = oracle_mod.lib().orc_channel(t.h, 0)
    C.memmove(ptr, ch.ctypes.data, ch.nbytes)
    t.update_apron(0, 0.0)
    g = np.random.default_rng(2)
    qs = g.uniform(52, 68, (500, 3)).astype(np.float32)
    got = t.sample_trilinear(0, qs)
    want = c0 + qs.astype(np.float64) @ c
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9)
    t.fill_channel(0, 1.0)
    np.testing.assert_allclose(t.sample_trilinear(0, qs), 1.0, atol=1e-12)  # partition of unity


def test_splat_adjoint_identity_and_mass(oracle_mod):
    t, p = _affine_tree(oracle_mod)
    g = np.random.default_rng(3)
    qs = g.uniform(50, 70, (300, 3)).astype(np.float32)
    m = g.uniform(0.5, 1.5, 300).astype(np.float32)
    t.splat_trilinear(0, qs, m)
    rho = t.channel(0)
    assert abs(rho[:, 1:9, 1:9, 1:9].sum() - m.astype(np.float64).sum()) < 1e-9  # mass conservation
    # <S^T m, f> == <m, S f> with f a random interior field
    import ctypes as C
    f = g.standard_normal(rho.shape)
    fi = np.zeros_like(f)
    fi[:, 1:9, 1:9, 1:9] = f[:, 1:9, 1:9, 1:9]
    t.fill_channel(0, 0.0)
    ptr = oracle_mod.lib().orc_channel(t.h, 0)
    C.memmove(ptr, fi.ctypes.data, fi.nbytes)
    t.update_apron(0, 0.0)
    lhs = (rho[:, 1:9, 1:9, 1:9] * fi[:, 1:9, 1:9, 1:9]).sum()
    rhs = (m.astype(np.float64) * t.sample_trilinear(0, qs)).sum()
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(lhs))


def test_splat_backward_matches_finite_differences(oracle_mod):
    import ctypes as C
    t, p = _affine_tree(oracle_mod)
    g = np.random.default_rng(4)
    qs = g.uniform(50, 70, (40, 3)).astype(np.float32)
    m = g.uniform(0.5, 1.5, 40).astype(np.float32)
    G = g.standard_normal(t.channel(0).shape)
    ptr = oracle_mod.lib().orc_channel(t.h, 0)
    C.memmove(ptr, G.ctypes.data, G.nbytes)
    dq, dm = t.splat_backward(0, qs, m)

    def loss(qv, mv):
        t.fill_channel(0, 0.0)
        t.splat_trilinear(0, qv, mv)
        return (t.channel(0)[:, 1:9, 1:9, 1:9] * G[:, 1:9, 1:9, 1:9]).sum()

    eps = 1e-3
    for i in range(0, 40, 4):
        for a in range(3):
            qp, qm = qs.copy(), qs.copy()
            qp[i, a] += eps
            qm[i, a] -= eps
            fd = (loss(qp, m) - loss(qm, m)) / (float(qp[i, a]) - float(qm[i, a]))
            assert abs(fd - dq[i, a]) <= 1e-3 * max(1.0, abs(fd))
        mp, mm = m.copy(), m.copy()
        mp[i] += np.float32(0.01)
        mm[i] -= np.float32(0.01)
        fd = (loss(qs, mp) - loss(qs, mm)) / (float(mp[i]) - float(mm[i]))
        assert abs(fd - dm[i]) <= 1e-6 * max(1.0, abs(fd))


def test_sample_backward_matches_finite_differences(oracle_mod):
    """dq of sampling = dout * grad of the sampled field: FP64 central differences of
    orc_sample_trilinear over a random field whose aprons hold neighbour copies and a
    non-zero boundary value."""
    import ctypes as C
    t, p = _affine_tree(oracle_mod)
    g = np.random.default_rng(6)
    f = g.standard_normal(t.channel(0).shape)
    ptr = oracle_mod.lib().orc_channel(t.h, 0)
    C.memmove(ptr, f.ctypes.data, f.nbytes)
    t.update_apron(0, 0.7)
    qs = g.uniform(35, 85, (400, 3)).astype(np.float32)
    dout = g.standard_normal(400)
    dq = t.sample_backward(0, qs, dout)
    eps = np.float32(1e-3)
    checked = 0
    for a in range(3):
        qp, qm = qs.copy(), qs.copy()
        qp[:, a] += eps
        qm[:, a] -= eps
        # skip points whose perturbation crosses a voxel-centre plane (the filter has a kink there)
        same = np.floor(qp[:, a].astype(np.float64) - 0.5) == np.floor(qm[:, a].astype(np.float64) - 0.5)
        same &= np.floor(qp[:, a]) == np.floor(qm[:, a])
        fd = (t.sample_trilinear(0, qp) - t.sample_trilinear(0, qm)) / (qp[:, a].astype(np.float64) - qm[:, a])
        want = dout * fd
        err = np.abs(want - dq[:, a])[same] / np.maximum(np.abs(want[same]), 1.0)
        assert err.max() < 1e-9, err.max()
        checked += same.sum()
    assert checked > 1000
    assert (np.abs(dq) > 0).sum() > 500
    outside = np.array([[1.0, 1.0, 1.0], [500.0, 3.0, 3.0]], np.float32)  # no leaf: zero gradient
    assert not t.sample_backward(0, outside, np.ones(2)).any()


def _scatter_tree(oracle_mod, topo, apron=1):
    t = oracle_mod.Tree((3, 3, 3, 3, 5), 16, 16, apron)
    t.rebuild_topology(topo, 1.0)
    t.add_channel(0)
    t.update_atlas()
    t.fill_channel(0, 0.0)
    return t


def _brick_of(t, corner):
    """(slot, padded brick array) of the leaf whose corner is `corner`."""
    nodes = t.nodes(0)
    live = np.nonzero(nodes["flags"] == 1)[0]
    k = np.nonzero((nodes["pos"][live] == np.asarray(corner)).all(1))[0]
    assert len(k) == 1
    slot = t.leaf_slot()[live[k[0]]]
    return slot, t.channel(0)[slot]


def test_scatter_density_closed_forms(oracle_mod):
    """ScatterDensity (P.cuh:516-573) at hand-computed points of the brick at (40, 40, 40):
    a point at voxel centre (42.5, 43.5, 44.5) with r = 1 puts (1 - 0.75)^3 = 0.015625 on voxel
    (2, 3, 4); its -x face neighbour is 1.5, 0.5, 0.5 away: c = 2.75, (1 - 2.75)^3 = -5.359375;
    its +x neighbour 0.015625.  A point exactly on voxel corner (45, 41, 42) gives 1 there and 0
    on the 6 neighbours (c = 1).  r = 2 scales c by 1/4: (1 - 0.1875)^3 = 0.536376953125."""
    topo = np.array([[44.0, 44.0, 44.0], [150.0, 150.0, 150.0]], np.float32)
    t = _scatter_tree(oracle_mod, topo)
    t.scatter_density(0, np.array([[42.5, 43.5, 44.5]], np.float32), 1.0)
    _, b = _brick_of(t, (40, 40, 40))
    a = 1  # apron: padded index = local + 1, array order [z, y, x]
    assert b[4 + a, 3 + a, 2 + a] == 0.015625
    assert b[4 + a, 3 + a, 1 + a] == -5.359375
    assert b[4 + a, 3 + a, 3 + a] == 0.015625
    assert b[4 + a, 2 + a, 2 + a] == -5.359375 and b[4 + a, 4 + a, 2 + a] == 0.015625
    assert b[3 + a, 3 + a, 2 + a] == -5.359375 and b[5 + a, 3 + a, 2 + a] == 0.015625
    assert np.count_nonzero(b) == 7
    t = _scatter_tree(oracle_mod, topo)
    t.scatter_density(0, np.array([[45.0, 41.0, 42.0]], np.float32), 1.0)
    _, b = _brick_of(t, (40, 40, 40))
    assert b[2 + a, 1 + a, 5 + a] == 1.0 and np.count_nonzero(b) == 1
    t = _scatter_tree(oracle_mod, topo)
    t.scatter_density(0, np.array([[42.5, 43.5, 44.5]], np.float32), 2.0, expand=False)
    _, b = _brick_of(t, (40, 40, 40))
    assert b[4 + a, 3 + a, 2 + a] == 0.536376953125 and np.count_nonzero(b) == 1


def test_scatter_density_apron_skips_and_sums(oracle_mod):
    """The -x neighbour of a point in local voxel 0 lands in the brick's own apron (no clamp);
    a point with a negative coordinate, one outside every brick and a NOHIT point add nothing;
    two calls add up (the contributions sum, so the reference's race is defined away); a
    translation shifts the point before the leaf lookup."""
    topo = np.array([[44.0, 44.0, 44.0], [150.0, 150.0, 150.0]], np.float32)
    t = _scatter_tree(oracle_mod, topo)
    pts = np.array([[40.5, 43.5, 44.5], [-1.0, 43.0, 43.0], [200.0, 200.0, 200.0], [43.0, 43.0, 1e10]], np.float32)
    t.scatter_density(0, pts, 1.0)
    _, b = _brick_of(t, (40, 40, 40))
    a = 1
    assert b[4 + a, 3 + a, 0] == -5.359375  # padded x = -1 + apron
    assert b[4 + a, 3 + a, 0 + a] == 0.015625
    assert np.count_nonzero(b) == 7
    t.scatter_density(0, pts, 1.0)
    _, b2 = _brick_of(t, (40, 40, 40))
    assert np.array_equal(b2, 2 * b)
    t = _scatter_tree(oracle_mod, topo)
    t.scatter_density(0, np.array([[40.0, 41.0, 42.0]], np.float32), 1.0, trans=(2.5, 2.5, 2.5), expand=False)
    _, b = _brick_of(t, (40, 40, 40))
    assert b[4 + a, 3 + a, 2 + a] == 0.015625 and np.count_nonzero(b) == 1


def _word(r, g, b, a):
    return r | g << 8 | b << 16 | a << 24


def test_colour_means_hand_computed(oracle_mod):
    """GatherDensity's colour rule (P.cuh:713-731) at hand-computed subcells (8^3 bricks, sc 4,
    r = 1).  Points at 9.x reach only the subcell at (8, 8, 8) (int(p - 1) and int(p + 2) both
    floor to 8), so its list holds exactly the first two points:
      colours (10, 20, 30, 40) and (11, 0, 255, 255), start (0, 0, 0, 255) without accumulation:
      (21 / 2, 20 / 2, 285 / 2, 550 / 2) -> (10, 10, 142, 255 saturated);
      the subcell at (12, 12, 12) of the same brick is empty -> 0;
      a lone point (200, 200, 200, 200) at 40.5 -> (200, 200, 200, 455 -> 255);
    with accumulation over the old word (100, 50, 0, 255):
      (121 / 3, 70 / 3, 285 / 3, 550 / 3) -> (40, 23, 95, 183); the empty subcell keeps the old
      word (old / 1)."""
    p = np.array([[9.2, 9.3, 9.4], [9.6, 9.1, 9.9], [40.5, 40.5, 40.5], [150.0, 150.0, 150.0]], np.float32)
    clr = np.array([_word(10, 20, 30, 40), _word(11, 0, 255, 255), _word(200, 200, 200, 200),
                    _word(1, 2, 3, 4)], np.uint32)
    t = oracle_mod.Tree((3, 3, 3, 3, 5), 16, 16, 1)
    t.rebuild_topology(p, 1.0)
    t.add_channel(1)
    t.update_atlas()
    t.insert_points_subcell(4, p, 1.0)
    t.fill_channel(1, 12345.0)
    t.gather_colour(clr, 1)

    def brick(corner):
        nodes = t.nodes(0)
        live = np.nonzero(nodes["flags"] == 1)[0]
        k = np.nonzero((nodes["pos"][live] == np.asarray(corner)).all(1))[0]
        assert len(k) == 1
        return t.channel(1)[t.leaf_slot()[live[k[0]]]]

    a = 1
    b = brick((8, 8, 8))
    assert b[1 + a, 1 + a, 1 + a] == _word(10, 10, 142, 255)   # voxel (9, 9, 9)
    assert b[3 + a, 0 + a, 2 + a] == _word(10, 10, 142, 255)   # voxel (10, 8, 11): same subcell
    assert b[5 + a, 6 + a, 4 + a] == 0.0                       # subcell (12, 12, 12): empty
    assert brick((40, 40, 40))[1 + a, 1 + a, 1 + a] == _word(200, 200, 200, 255)
    t.fill_channel(1, float(_word(100, 50, 0, 255)))
    t.gather_colour(clr, 1, accum=True)
    b = brick((8, 8, 8))
    assert b[1 + a, 1 + a, 1 + a] == _word(40, 23, 95, 183)
    assert b[5 + a, 6 + a, 4 + a] == _word(100, 50, 0, 255)


def test_scatter_colour_hand_computed(oracle_mod):
    """ScatterDensity's colour part (P.cuh:553-572, 610-641): points 0 and 1 share voxel
    (42, 43, 44) of the brick at (40, 40, 40), point 2 is alone in (45, 45, 45).
    avgColor: (21 / 2, 20 / 2, 285 / 2, 255) = (10, 10, 142, 255) and (200, 100, 50, 255);
    without it the voxel takes one point's word as it is -- the largest index, point 1 -- and
    the lone point's word keeps its alpha 0.  Voxels without points keep the fill word."""
    topo = np.array([[44.0, 44.0, 44.0], [150.0, 150.0, 150.0]], np.float32)
    pts = np.array([[42.5, 43.5, 44.5], [42.2, 43.7, 44.1], [45.5, 45.5, 45.5], [-1.0, 43.0, 43.0]], np.float32)
    clr = np.array([_word(10, 20, 30, 40), _word(11, 0, 255, 7), _word(200, 100, 50, 0), _word(9, 9, 9, 9)],
                   np.uint32)
    a = 1
    for avg, want0, want1 in ((True, _word(10, 10, 142, 255), _word(200, 100, 50, 255)),
                              (False, _word(11, 0, 255, 7), _word(200, 100, 50, 0))):
        t = _scatter_tree(oracle_mod, topo)
        t.fill_channel(0, 77.0)
        t.scatter_colour(0, pts, clr, avg=avg)
        _, b = _brick_of(t, (40, 40, 40))
        assert b[4 + a, 3 + a, 2 + a] == want0
        assert b[5 + a, 5 + a, 5 + a] == want1
        assert np.count_nonzero(b != 77.0) == 2


def test_fp16_lists_hand_computed_and_bounded(oracle_mod):
    """FP16 subcell lists (P.cuh:200-207): ushort((q - lo) / range * 65535) in single precision.
    Over lo = (2, 2, 2), range = (100, 100, 100): 10.5 -> 8.5 / 100 * 65535 = 5570.47 -> 5570;
    q = lo -> 0; q = lo + range -> 65535.  The FP16 level set (P.cuh:775-811) decodes
    u / 65535 * range + lo, at most range / 65535 per axis from the listed position, so it stays
    within sqrt(3) * 100 / 65535 (plus rounding) of the float level set on every voxel."""
    p = np.array([[10.5, 2.0, 102.0], [40.25, 30.5, 60.75], [150.0, 150.0, 150.0]], np.float32)
    t = oracle_mod.Tree((3, 3, 3, 3, 5), 16, 16, 1)
    t.rebuild_topology(p, 1.0)
    t.add_channel(0)
    t.add_channel(1)
    t.update_atlas()
    t.insert_points_subcell(4, p, 1.0)
    lo, rg = (2.0, 2.0, 2.0), (100.0, 100.0, 100.0)
    q = t.quantize_list_pos(lo, rg)
    s = t.subcells()
    first = q[np.nonzero(s["list"] == 0)[0]]
    assert len(first) > 0 and (first == np.array([5570, 0, 65535], np.uint16)).all()
    # saturation: a listed coordinate beyond the box maps to 65535
    assert (t.quantize_list_pos(lo, (50.0, 50.0, 50.0))[np.nonzero(s["list"] == 2)[0]] == 65535).all()
    g = np.random.default_rng(4).uniform(20, 90, (3000, 3)).astype(np.float32)
    t = oracle_mod.Tree((3, 3, 3, 3, 5), 16, 16, 1)
    t.rebuild_topology(g, 1.0)
    t.add_channel(0)
    t.add_channel(1)
    t.update_atlas()
    t.insert_points_subcell(4, g, 1.0)
    lo = np.float32([18.0, 18.0, 18.0])
    rg = np.float32([80.0, 80.0, 80.0])
    t.gather_levelset(1.0, 0)
    t.gather_levelset_q16(1.0, 1, t.quantize_list_pos(lo, rg), lo, rg)
    d = np.abs(t.channel(0) - t.channel(1))
    assert d.max() <= np.sqrt(3.0) * 80.0 / 65535.0 * 1.01 + 1e-6
    assert d.max() > 0  # the quantization shows
