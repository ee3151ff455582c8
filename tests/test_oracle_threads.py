"""The oracle's threaded binning, gathers and apron (the all-core CPU baseline) against its serial
loops: the subcell counts, offsets and lists (point order within every subcell) and every channel
voxel must be identical, for point counts above, equal to and below the thread count."""
import numpy as np
import pytest

from gvdbx import synth


def run(oracle_mod, p, vel, threads, r=1.0, sc=4):
    oracle_mod.set_threads(threads)
    try:
        t = oracle_mod.Tree()
        for c in range(5):
            t.add_channel(c)
        t.rebuild_topology(p, r)
        t.update_atlas()
        ln = t.insert_points_subcell(sc, p, r)
        t.gather_density(r, 0)
        t.gather_velocity(vel, (1, 2, 3))
        t.gather_levelset(r, 4)
        for c in range(5):
            t.update_apron(c, 0.5 * c)
        return ln, t.subcells(), [t.channel(c).copy() for c in range(5)]
    finally:
        oracle_mod.set_threads(1)


@pytest.mark.parametrize("n", [40_000, 7, 3])
def test_threaded_equals_serial(oracle_mod, n):
    p, vel = synth.sph_block(n=max(n, 64), seed=4, box=48.0)
    p, vel = p[:n], vel[:n]
    ln1, s1, c1 = run(oracle_mod, p, vel, 1)
    ln4, s4, c4 = run(oracle_mod, p, vel, 4)
    assert ln1 == ln4
    for k in ("map", "cnt", "off", "pos", "nid", "list", "list_pos"):
        assert np.array_equal(s1[k], s4[k]), k
    for c in range(5):
        assert c1[c].tobytes() == c4[c].tobytes(), c


def test_threaded_sphere_subcell_2(oracle_mod):
    p = synth.sphere_shell(n=30_000, seed=7)
    vel = np.random.default_rng(7).standard_normal(p.shape).astype(np.float32)
    a = run(oracle_mod, p, vel, 1, r=1.5, sc=2)
    b = run(oracle_mod, p, vel, 3, r=1.5, sc=2)
    assert a[0] == b[0]
    for k in ("cnt", "off", "list", "list_pos"):
        assert np.array_equal(a[1][k], b[1][k]), k
    for c in range(5):
        assert a[2][c].tobytes() == b[2][c].tobytes(), c
