"""gvdbx_splat_trilinear_noalloc (scratch in the library's fixed per-GPU array) against
gvdbx_splat_trilinear (scratch the library allocates), through the C ABI.

The splat is bit-reproducible (exact 64-bit fixed-point sums per brick), so the two entry points
must write identical channels, bit for bit: on two calls in a row, the second with a smaller
input (the array keeps the first call's contents, so stale counts would show), then with
accumulation, and the cap on the point count is refused before anything runs.
"""
import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


def interior(v, chan):
    a = v.apron
    return v.get_channel(chan)[:, a:-a, a:-a, a:-a]


def test_splat_noalloc_matches_splat():
    U.log("case splat without allocation")
    g = np.random.default_rng(17)
    p = g.uniform(0.0, 96.0, (300_000, 3)).astype(np.float32)
    m = g.uniform(0.5, 1.5, len(p)).astype(np.float32)
    v = U.new_volume(chans=(0, 1))
    dp, dm = v.to_device(p), v.to_device(m)
    v.SetPoints(dp, n=len(p))
    v.RebuildTopology(len(p), 1.0)
    v.UpdateAtlas()
    for n in (300_000, 70_000):
        U.log(f"  n={n}")
        v.SplatTrilinear(0, dp, n, dm)
        want = interior(v, 0)
        v.SplatTrilinearNoAlloc(1, dp, n, dm)
        got = interior(v, 1)
        assert got.tobytes() == want.tobytes(), f"n={n}: the channels differ"
        assert np.any(got != 0)
    U.log("  accumulate")
    v.SplatTrilinear(0, dp, 50_000, None, accum=True)
    v.SplatTrilinearNoAlloc(1, dp, 50_000, None, accum=True)
    assert interior(v, 1).tobytes() == interior(v, 0).tobytes()
    st = v.L.gvdbx_splat_trilinear_noalloc(v.h, 1, dp.ptr, dm.ptr, (1 << 20) + 1, 0)
    assert st == 4  # GVDBX_ERR_CAPACITY, refused on the host
    v.close()
