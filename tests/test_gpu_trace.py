"""GPU parity: photon emission and bounce tracing vs the oracle, bit-exact.

Both sides follow the same path-ordered semantics (per-path Philox stream,
first n photons of each kind in (path, event) order), FP32 without
contraction and correctly rounded division / sqrt, so the photon maps must be
identical bit for bit.
"""
import numpy as np
import pytest

import pyoracle as O
from pm355.scenes import cornell_box

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[False, True], ids=["brute", "bvh"])
def dscene(ctx, request):
    return ctx.upload_scene(cornell_box(), accel=request.param)


def _np(t):
    return t.detach().cpu().numpy()


# (5, 2000, 9000, 10): the global map fills in the first batch and the
# caustic map only after it, in the caustic phase (emission filter, survivor
# queue, persistent continuation, one staging slot per path).
@pytest.mark.parametrize("seed,ng,nc,mb", [(1, 20000, 2000, 10), (7, 5000, 500, 8), (3, 3000, 50, 1),
                                           (5, 2000, 9000, 10)])
def test_trace_parity(ctx, dscene, seed, ng, nc, mb):
    g, c, st = ctx.trace_photons(dscene, seed, ng, nc, mb, max_paths=2_000_000)
    og, oc, ost = O.trace_photons(dscene.scene, seed, ng, nc, mb, max_paths=2_000_000)
    assert st[1:] == ost[1:]
    assert np.array_equal(_np(g).view(np.uint32), og.view(np.uint32))
    assert np.array_equal(_np(c)[:, : ost[2]].view(np.uint32), oc[:, : ost[2]].view(np.uint32))


def test_trace_max_paths(ctx, dscene):
    g, c, st = ctx.trace_photons(dscene, 2, 100000, 100000, 10, max_paths=1000)
    og, oc, ost = O.trace_photons(dscene.scene, 2, 100000, 100000, 10, max_paths=1000)
    assert st == ost
    assert np.array_equal(_np(g)[:, : st[1]].view(np.uint32), og[:, : st[1]].view(np.uint32))
    assert np.array_equal(_np(c)[:, : st[2]].view(np.uint32), oc[:, : st[2]].view(np.uint32))


def test_trace_full_size_prefix(ctx, dscene):
    """2M global + 500k caustic photons (the per-frame scene config): full
    maps, valid photons, and the first photons equal the oracle's (the maps
    are path-ordered, so a short oracle run is a prefix up to the power)."""
    from test_oracle_trace import _on_walls_or_light

    ng, nc = 2_000_000, 500_000
    g, c, st = ctx.trace_photons(dscene, 11, ng, nc, 10, max_paths=200_000_000)
    assert st[1:] == (ng, nc)
    gh, ch = _np(g), _np(c)
    for m in (gh, ch):
        assert np.all(m[3:6] == np.float32(1.0) / np.float32(ng))
        sel = np.random.default_rng(0).choice(m.shape[1], 200_000, replace=False)
        assert np.all(_on_walls_or_light(m[0:3, sel]))
    og, oc, _ = O.trace_photons(dscene.scene, 11, 20000, 3000, 10)
    assert np.array_equal(gh[0:3, :20000].view(np.uint32), og[0:3].view(np.uint32))
    assert np.array_equal(gh[6:9, :20000].view(np.uint32), og[6:9].view(np.uint32))
    assert np.array_equal(ch[0:3, :3000].view(np.uint32), oc[0:3].view(np.uint32))
