"""GPU trilinear splat and the adjoints of splat and sampling, against the FP64 oracle.

Bar: FP32 fields, dL/dq and dL/dmass within |gpu - oracle| <= 1e-5 * max(|oracle|, 1); the
splat is bit-reproducible (same bits for repeated runs and shuffled point order); the
adjoint agrees with FP64 central finite differences of the oracle's forward to 1e-3.
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


def interior(a):
    return a[:, 1:-1, 1:-1, 1:-1]


def build_pair(oracle_mod, topo_pts, chans=(0, 1, 2), radius=1.0):
    """GPU volume and oracle tree with the topology of topo_pts (full rebuild, r = 1)."""
    v, t = U.new_volume(chans=chans), U.new_tree(oracle_mod, chans=chans)
    d = v.to_device(topo_pts)
    v.SetPoints(d, n=len(topo_pts))
    v.RebuildTopology(len(topo_pts), radius)
    v.UpdateAtlas()
    t.rebuild_topology(topo_pts, radius)
    t.update_atlas()
    U.compare_topology(v, t)
    for c in chans:
        t.fill_channel(c, 0.0)
    return v, t, d


def set_oracle_channel(oracle_mod, t, chan, data):
    import ctypes as C
    a = np.ascontiguousarray(data, np.float64)
    C.memmove(oracle_mod.lib().orc_channel(t.h, chan), a.ctypes.data, a.nbytes)


def rel_err(got, want):
    return float((np.abs(np.asarray(got, np.float64) - want) / np.maximum(np.abs(want), 1.0)).max())


@pytest.mark.parametrize("seed,use_mass", [(0, True), (1, False)])
def test_splat_matches_oracle(oracle_mod, seed, use_mass):
    U.log(f"case splat seed={seed} mass={use_mass}")
    g = np.random.default_rng(seed)
    p = g.uniform(20, 100, (20000, 3)).astype(np.float32)
    p[:25] -= 150.0  # invalid points are skipped
    q = np.concatenate([p, g.uniform(10, 110, (3000, 3)).astype(np.float32)])  # some outside every brick
    m = g.uniform(-1.0, 2.0, len(q)).astype(np.float32) if use_mass else None
    v, t, _ = build_pair(oracle_mod, p, chans=(0,))
    dq = v.to_device(q)
    dm = v.to_device(m) if use_mass else None
    v.SplatTrilinear(0, dq, len(q), dm)
    t.splat_trilinear(0, q, m)
    got, want = interior(v.get_channel(0)), interior(t.channel(0))
    assert rel_err(got, want) <= U.TOL
    assert (want != 0).mean() > 0.15
    # accumulate a second splat on top
    v.SplatTrilinear(0, dq, len(q), dm, accum=True)
    t.splat_trilinear(0, q, m)
    assert rel_err(interior(v.get_channel(0)), interior(t.channel(0))) <= U.TOL


def test_splat_mass_conservation_and_sampling_transpose(oracle_mod):
    """With r = 1 every corner of a valid point lies in a brick, so the splat conserves mass;
    <S^T m, f> == <m, S f> against the GPU sampler."""
    U.log("case splat conservation")
    g = np.random.default_rng(3)
    p = g.uniform(30, 90, (30000, 3)).astype(np.float32)
    m = g.uniform(0.5, 1.5, len(p)).astype(np.float32)
    v, t, dp = build_pair(oracle_mod, p, chans=(0, 1))
    dmass = v.to_device(m)
    v.SplatTrilinear(0, dp, len(p), dmass)
    rho = interior(v.get_channel(0)).astype(np.float64)
    assert abs(rho.sum() - m.astype(np.float64).sum()) <= 1e-5 * m.sum()
    f = g.standard_normal(v.get_channel(1).shape).astype(np.float32)
    v.set_channel(1, f)
    v.UpdateApron(1, 0.0)
    out = v.alloc(4 * len(p))
    v.SampleTrilinear(1, dp, len(p), out)
    s = out.download(np.float32).astype(np.float64)
    lhs = (rho * interior(f).astype(np.float64)).sum()
    rhs = (m.astype(np.float64) * s).sum()
    assert abs(lhs - rhs) <= 1e-5 * max(1.0, abs(lhs))


def test_splat_bit_reproducible(oracle_mod):
    U.log("case splat determinism")
    g = np.random.default_rng(5)
    p = g.uniform(40, 80, (60000, 3)).astype(np.float32)  # dense: many points per voxel
    m = g.uniform(0.5, 1.5, len(p)).astype(np.float32)
    outs = []
    for perm in (np.arange(len(p)), np.arange(len(p)), g.permutation(len(p))):
        v = U.new_volume(chans=(0,))
        d = v.to_device(p)
        v.SetPoints(d, n=len(p))
        v.RebuildTopology(len(p), 1.0)
        v.UpdateAtlas()
        dq, dm = v.to_device(p[perm]), v.to_device(m[perm])
        v.SplatTrilinear(0, dq, len(p), dm)
        outs.append(interior(v.get_channel(0)).tobytes())
        v.close()
    assert outs[0] == outs[1] == outs[2]


def test_splat_backward_matches_oracle(oracle_mod):
    U.log("case splat backward")
    g = np.random.default_rng(7)
    p = g.uniform(20, 100, (20000, 3)).astype(np.float32)
    q = np.concatenate([p, g.uniform(0, 130, (2000, 3)).astype(np.float32)])
    m = g.uniform(0.5, 1.5, len(q)).astype(np.float32)
    v, t, _ = build_pair(oracle_mod, p, chans=(0,))
    G = g.standard_normal(v.get_channel(0).shape).astype(np.float32)
    v.set_channel(0, G)
    set_oracle_channel(oracle_mod, t, 0, G.astype(np.float64))
    dqd, dmd = v.to_device(q), v.to_device(m)
    gq, gm = v.alloc(12 * len(q)), v.alloc(4 * len(q))
    v.SplatBackward(0, dqd, len(q), dmd, gq, gm)
    want_q, want_m = t.splat_backward(0, q, m)
    got_q, got_m = gq.download(np.float32), gm.download(np.float32)
    assert rel_err(got_q.reshape(-1, 3), want_q) <= U.TOL
    assert rel_err(got_m, want_m) <= U.TOL
    assert (want_m != 0).sum() > 15000
    # after a forward splat of the same buffer the backward visits the points in brick order:
    # per-point results are independent of the visiting order, so the bits are the same
    v.AddChannel(1, apron=1)
    v.SplatTrilinear(1, dqd, len(q), dmd)
    gq2, gm2 = v.alloc(12 * len(q)), v.alloc(4 * len(q))
    v.SplatBackward(0, dqd, len(q), dmd, gq2, gm2)
    assert gq2.download(np.float32).tobytes() == got_q.tobytes()
    assert gm2.download(np.float32).tobytes() == got_m.tobytes()


def test_sample_backward_matches_oracle(oracle_mod):
    """Sampling backward: dq (padded-brick filter, boundary value in the apron) and the field
    gradient (splat of dout) against the oracle."""
    U.log("case sample backward")
    g = np.random.default_rng(8)
    p = g.uniform(20, 100, (20000, 3)).astype(np.float32)
    q = np.concatenate([p, g.uniform(0, 130, (2000, 3)).astype(np.float32)])
    v, t, _ = build_pair(oracle_mod, p, chans=(0, 1))
    f = g.standard_normal(v.get_channel(0).shape).astype(np.float32)
    v.set_channel(0, f)
    set_oracle_channel(oracle_mod, t, 0, f.astype(np.float64))
    v.UpdateApron(0, 0.7)
    t.update_apron(0, 0.7)
    dout = g.standard_normal(len(q)).astype(np.float32)
    dq_dev, dout_dev = v.to_device(q), v.to_device(dout)
    gq = v.alloc(12 * len(q))
    v.SampleTrilinearBackward(0, dq_dev, dout_dev, len(q), gq, chan_grad=1)
    assert rel_err(gq.download(np.float32).reshape(-1, 3), t.sample_backward(0, q, dout.astype(np.float64))) <= U.TOL
    t.splat_trilinear(1, q, dout)
    assert rel_err(interior(v.get_channel(1)), interior(t.channel(1))) <= U.TOL


def _loss(rho, target):
    d = interior(rho).astype(np.float64) - interior(target).astype(np.float64)
    return (d * d).mean()


def test_adjoint_finite_differences_small(oracle_mod):
    """L = mean over active voxels of (rho - rho*)^2; GPU dL/dq, dL/dm against FP64 central
    differences of the oracle's own splat (eps = 1e-3 on positions)."""
    U.log("case adjoint FD small")
    g = np.random.default_rng(9)
    a = g.uniform(30, 62, (6000, 3)).astype(np.float32)
    b = g.uniform(30, 62, (6000, 3)).astype(np.float32)
    ma = g.uniform(0.5, 1.5, len(a)).astype(np.float32)
    mb = g.uniform(0.5, 1.5, len(b)).astype(np.float32)
    v, t, _ = build_pair(oracle_mod, np.concatenate([a, b]), chans=(0, 1, 2))
    da, dma, db, dmb = v.to_device(a), v.to_device(ma), v.to_device(b), v.to_device(mb)
    v.SplatTrilinear(0, da, len(a), dma)
    v.SplatTrilinear(1, db, len(b), dmb)
    rho, tgt = v.get_channel(0), v.get_channel(1)
    nvox = interior(rho).size
    G = np.zeros_like(rho)
    G[:, 1:-1, 1:-1, 1:-1] = 2.0 * (interior(rho).astype(np.float64) - interior(tgt)) / nvox
    v.set_channel(2, G)
    gq, gm = v.alloc(12 * len(a)), v.alloc(4 * len(a))
    v.SplatBackward(2, da, len(a), dma, gq, gm)
    dq = gq.download(np.float32).reshape(-1, 3)
    dm = gm.download(np.float32)

    t.splat_trilinear(1, b, mb)
    target = t.channel(1).copy()

    def loss(qv, mv):
        t.fill_channel(0, 0.0)
        t.splat_trilinear(0, qv, mv)
        return _loss(t.channel(0), target)

    idx = g.choice(len(a), 64, replace=False)
    eps = np.float32(1e-3)
    worst = 0.0
    for i in idx:
        for ax in range(3):
            qp, qm = a.copy(), a.copy()
            qp[i, ax] += eps
            qm[i, ax] -= eps
            if np.floor(float(qp[i, ax]) - 0.5) != np.floor(float(qm[i, ax]) - 0.5):
                continue  # the perturbation crosses a voxel-centre plane, where L has a kink
            fd = (loss(qp, ma) - loss(qm, ma)) / (float(qp[i, ax]) - float(qm[i, ax]))
            worst = max(worst, abs(fd - dq[i, ax]) / max(abs(fd), 1e-3 * np.abs(dq).max()))
        mp, mm = ma.copy(), ma.copy()
        mp[i] += np.float32(1e-2)
        mm[i] -= np.float32(1e-2)
        fd = (loss(a, mp) - loss(a, mm)) / (float(mp[i]) - float(mm[i]))
        worst = max(worst, abs(fd - dm[i]) / max(abs(fd), 1e-3 * np.abs(dm).max()))
    U.log(f"adjoint FD small: worst relative error {worst:.2e}")
    assert worst <= 1e-3


def test_adjoint_config_2m_points(oracle_mod):
    """The 2,097,152-point adjoint workload: union topology of the two draws, forward splats,
    MSE gradient; GPU gradients against the oracle for every particle and, on 64 random
    particles, against FP64 central differences (eps = 1e-3) of L through the oracle sampler
    (L(q_i +- e) differs by m_i * (S G)(q_i +- e) within a voxel-centre cell)."""
    from gvdbx import synth
    U.log("case adjoint 2M")
    a, ma = synth.uniform_massive(seed=4)
    b, mb = synth.uniform_massive(seed=4 + 1000)
    v, t, _ = build_pair(oracle_mod, np.concatenate([a, b]), chans=(0, 1, 2))
    da, dma, db, dmb = v.to_device(a), v.to_device(ma), v.to_device(b), v.to_device(mb)
    v.SplatTrilinear(0, da, len(a), dma)
    v.SplatTrilinear(1, db, len(b), dmb)
    rho, tgt = v.get_channel(0), v.get_channel(1)
    t.splat_trilinear(0, a, ma)
    assert rel_err(interior(rho), interior(t.channel(0))) <= U.TOL
    nvox = interior(rho).size
    G = np.zeros_like(rho)
    G[:, 1:-1, 1:-1, 1:-1] = 2.0 * (interior(rho).astype(np.float64) - interior(tgt)) / nvox
    v.set_channel(2, G)
    gq, gm = v.alloc(12 * len(a)), v.alloc(4 * len(a))
    v.SplatBackward(2, da, len(a), dma, gq, gm)
    dq = gq.download(np.float32).reshape(-1, 3)
    dm = gm.download(np.float32)
    set_oracle_channel(oracle_mod, t, 2, G.astype(np.float64))
    want_q, want_m = t.splat_backward(2, a, ma)
    scale_q, scale_m = np.abs(want_q).max(), np.abs(want_m).max()
    assert np.abs(dq - want_q).max() <= 1e-5 * scale_q
    assert np.abs(dm - want_m).max() <= 1e-5 * scale_m
    t.update_apron(2, 0.0)
    g = np.random.default_rng(64)
    idx = g.choice(len(a), 64, replace=False)
    eps = np.float32(1e-3)
    worst = 0.0
    for ax in range(3):
        qp, qm = a[idx].copy(), a[idx].copy()
        qp[:, ax] += eps
        qm[:, ax] -= eps
        h = qp[:, ax].astype(np.float64) - qm[:, ax]
        fd = ma[idx] * (t.sample_trilinear(2, qp) - t.sample_trilinear(2, qm)) / h
        same = np.floor(qp[:, ax].astype(np.float64) - 0.5) == np.floor(qm[:, ax].astype(np.float64) - 0.5)
        err = np.abs(fd - dq[idx, ax]) / np.maximum(np.abs(fd), 1e-3 * scale_q)
        worst = max(worst, float(err[same].max()))
    U.log(f"adjoint FD 2M: worst relative error {worst:.2e}")
    assert worst <= 1e-3
