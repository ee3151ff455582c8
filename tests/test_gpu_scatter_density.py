"""ScatterDensity (gvdbScatterPointDensity, P.cuh:516-573) on the GPU against the oracle.

The reference sums with a racy read-modify-write; the definition here is the race-free sum of
every contribution (the oracle sums in FP64), and the GPU's 64-bit fixed-point sums must match
it within 1e-5 relative (floored at 1) over whole padded bricks (the expand neighbours reach
the apron), be bitwise independent of the point order, and accumulate over calls.
"""
import numpy as np
import pytest

import gpu_util as U
from test_gpu_adjoint import build_pair, rel_err

pytestmark = pytest.mark.gpu


def scatter_points(seed, n=20000):
    g = np.random.default_rng(seed)
    p = g.uniform(20, 100, (n, 3)).astype(np.float32)
    p[:20] -= 150.0  # invalid: negative coordinates
    p[20, 2] = np.float32(1e10)  # NOHIT
    q = np.concatenate([p, g.uniform(5, 115, (2000, 3)).astype(np.float32)])  # some in no brick
    return p, q


@pytest.mark.parametrize("r,expand,seed", [(1.0, True, 0), (2.0, True, 1), (0.5, False, 2), (1.5, True, 3)])
def test_scatter_density_matches_oracle(oracle_mod, r, expand, seed):
    U.log(f"case scatter density r={r} expand={expand}")
    p, q = scatter_points(seed)
    topo = p[(p >= 0).all(1) & (p < 1e9).all(1)]
    v, t, _ = build_pair(oracle_mod, topo, chans=(0,))
    v.FillChannel(0, 0.25)
    t.fill_channel(0, 0.25)
    dq = v.to_device(q)
    v.SetPoints(dq, n=len(q))
    trans = (0.5, -0.25, 0.75)
    v.ScatterDensity(len(q), r, 1.0, trans, expand)
    t.scatter_density(0, q, r, trans, expand)
    got, want = v.get_channel(0), t.channel(0)[: v.num_slots()[0]]
    assert rel_err(got, want) <= U.TOL
    assert (np.abs(want - 0.25) > 1e-9).mean() > (0.05 if expand else 0.01)  # voxels the points reached
    v.ScatterDensity(len(q), r, 1.0, trans, expand)  # a second call adds the same contributions
    t.scatter_density(0, q, r, trans, expand)
    assert rel_err(v.get_channel(0), t.channel(0)[: v.num_slots()[0]]) <= U.TOL
    v.close()


def test_scatter_density_order_independent_and_limits(oracle_mod):
    from gvdbx import GvdbxError
    U.log("case scatter density determinism")
    p, q = scatter_points(7, 60000)
    topo = p[(p >= 0).all(1) & (p < 1e9).all(1)]
    v, t, _ = build_pair(oracle_mod, topo, chans=(0,))
    outs = []
    for perm in (np.arange(len(q)), np.random.default_rng(8).permutation(len(q))):
        v.FillChannel(0, 0.0)
        dq = v.to_device(q[perm])
        v.SetPoints(dq, n=len(q))
        v.ScatterDensity(len(q), 1.0, 1.0)
        outs.append(v.get_channel(0).tobytes())
    assert outs[0] == outs[1]
    with pytest.raises(GvdbxError) as e:
        v.ScatterDensity(len(q), 0.01, 1.0)  # (6 / 0.01^2)^3 per point: beyond exact sums
    assert e.value.code == 4
    # avgColor without colours bound: the density part alone (V.cpp:5637, 5649)
    v.FillChannel(0, 0.0)
    v.ScatterDensity(len(q), 1.0, 1.0, avgColor=True)
    assert v.get_channel(0).tobytes() == outs[1]
    v.close()
