"""The trained warm start (committed weights) on a held-out scene.

Every variant must reach the duality gap: the cold loop stays the fallback
and the judge. What the network gains over the same solve without it (the
dual recovery and steps kept) is printed, not asserted: on held-out depth-10
scenes it is about nothing (the 50-scene evaluation is
tgvreconclaude_amd/warmstart.py), because the iteration count is governed by
the dual, which the network does not predict.
"""
import os

import numpy as np
import pytest

from tgvreconclaude_amd import scene, warmstart

pytestmark = pytest.mark.gpu


def test_trained_network_on_held_out_scene(ctx):
    assert os.path.exists(warmstart.WEIGHTS)
    w = np.load(warmstart.WEIGHTS)
    assert w.shape == (6404,) and np.isfinite(w).all()
    sc = scene.csg_scene(912345, views=24, W=640, H=480, max_depth=9)
    r = warmstart.solve_variants(ctx, sc, w)
    for v in ("cold_reference_steps", "cold", "cold_dual", "warm", "cold_sched", "cold_dual_sched", "warm_sched"):
        assert r[v]["reached"], v
    print({v: (round(r[v]["ms"], 1), sum(r[v]["iters"])) for v in ("cold_reference_steps", "cold", "cold_dual",
                                                                      "warm")})
    print("network vs the same solve without it:", round(r["cold_dual"]["ms"] / r["warm"]["ms"], 3))
    # the warm start is kept at least on some level of a held-out scene
    assert sum(r["warm"]["warm_levels"]) >= 1
