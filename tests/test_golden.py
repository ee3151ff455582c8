"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from
the oracle on the shared seeded generator's points).

- CPU: the oracle reproduces them exactly, and the generator still yields their input points
  (a change to either shows here first).
- GPU: the HIP path reproduces them with no oracle in the loop: node pools and child lists byte
  for byte at every frame, the subcell map, counts and offsets exactly, and the channels within
  1e-5 relative of the stored values (FP32 roundings of the oracle's FP64 fields).
"""
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
LD = (3, 3, 3, 3, 5)
TOL = 1e-5


def load(name):
    with np.load(os.path.join(GOLD, name)) as z:
        return {k: z[k] for k in z.files}


def maker():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("name", ["sphere", "sph"])
def test_oracle_reproduces_fixture(oracle_mod, name):
    m = maker()
    want = load(name + ".npz")
    got = (m.make_sphere if name == "sphere" else m.make_sph)(oracle_mod)
    assert sorted(got) == sorted(want)
    for k in want:
        assert np.asarray(got[k]).tobytes() == want[k].tobytes(), k


def test_generator_yields_fixture_points():
    m = maker()
    assert m.sphere_points().tobytes() == load("sphere.npz")["points"].tobytes()
    p, v = m.sph_points()
    z = load("sph.npz")
    assert p.tobytes() == z["points0"].tobytes() and v.tobytes() == z["velocities"].tobytes()


# ---------------------------------------------------------------- GPU against the fixtures

def gpu_topology(v, z, prefix):
    d, root, _, _ = v.depth()
    assert d == int(z[prefix + "depth"]) and np.array_equal(root, z[prefix + "rootpos"])
    for l in range(len(LD)):
        assert v.get_nodes(l).view(np.uint8).tobytes() == z[f"{prefix}nodes{l}"].tobytes(), f"{prefix}level {l} pool"
        if l:
            assert np.array_equal(v.get_childlists(l), z[f"{prefix}clist{l}"]), f"{prefix}level {l} child lists"


def gpu_subcells(v, z):
    nsc, ln, _ = v.subcell_info()
    assert ln == int(z["list_len"]) and nsc == len(z["sc_cnt"])
    for which, key in ((0, "sc_map"), (1, "sc_cnt"), (2, "sc_off")):
        a = np.zeros_like(z[key])
        v._chk(v.L.gvdbx_copy_subcell_array(v.h, which, a.ctypes.data, a.nbytes), key)
        assert np.array_equal(a, z[key]), key


def close(got, want, what):
    err = np.abs(got.astype(np.float64) - want) / np.maximum(np.abs(want.astype(np.float64)), 1.0)
    assert err.max() <= TOL, (what, float(err.max()))


@pytest.mark.gpu
def test_gpu_sphere_fixture():
    import gpu_util as U
    z = load("sphere.npz")
    U.log("case golden sphere")
    v = U.new_volume(LD, chans=(0, 1))
    d = v.to_device(z["points"])
    n = len(z["points"])
    v.SetPoints(d, n=n)
    v.RebuildTopology(n, 1.0)
    v.UpdateAtlas()
    gpu_topology(v, z, "")
    assert v.num_slots()[0] == int(z["slots"])
    ln = v.InsertPointsSubcell(4, n, 1.0)
    gpu_subcells(v, z)
    v.GatherDensity(4, n, 1.0, (0, 0, 0), ln, 0)
    v.UpdateApron(0, 0.0)
    v.GatherLevelSet(4, n, 1.0, (0, 0, 0), ln, 1)
    v.UpdateApron(1, 3.0)
    close(v.get_channel(0), z["density"], "density")
    close(v.get_channel(1), z["levelset"], "level set")
    v.close()


@pytest.mark.gpu
def test_gpu_sph_fixture():
    import gpu_util as U
    z = load("sph.npz")
    U.log("case golden sph frames")
    v = U.new_volume(LD, chans=(0, 1, 2, 3))
    dp, dv = v.to_device(z["points0"]), v.to_device(z["velocities"])
    n = len(z["points0"])
    v.SetPoints(dp, dv, n=n)
    for k in range(3):
        if k:
            v.AdvancePoints(n, 0.5)
        v.RebuildTopology(n, 1.0)
        v.UpdateAtlas()
        gpu_topology(v, z, f"f{k}_")
    ln = v.InsertPointsSubcell(4, n, 1.0)
    gpu_subcells(v, z)
    v.GatherDensityVelocity(4, n, 1.0, (0, 0, 0), ln, 0, (1, 2, 3))
    v.UpdateApron()
    for c in range(4):
        close(v.get_channel(c), z["channels"][c], f"channel {c}")
    v.close()
