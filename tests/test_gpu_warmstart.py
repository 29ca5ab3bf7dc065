"""The trained warm start (committed weights) on a held-out depth-10 scene.

Every variant must reach the duality gap: the cold loop stays the fallback
and the judge. What the network gains over the same solve without it is
asserted in leaf-iterations, which do not depend on timing: with the same
steps the warm start needs fewer than the cold start (on all 50 held-out
scenes of the evaluation, 1.08x fewer in total). Against the cold start with
the dual recovered from its initial field the network does not gain (0.92x
in the evaluation); that ratio is printed. The 50-scene evaluation and the
bound on what any predictor could gain are tgvreconclaude_amd/warmstart.py
`eval` and `bound`.
"""
import os

import numpy as np
import pytest

from tgvreconclaude_amd import warmstart

pytestmark = pytest.mark.gpu


def test_trained_network_on_held_out_scene(ctx):
    assert os.path.exists(warmstart.WEIGHTS)
    w = np.load(warmstart.WEIGHTS)
    assert w.shape == (6404,) and np.isfinite(w).all()
    sc = warmstart.eval_scene(900000)  # held out: training used seeds 1000-2999
    r = warmstart.solve_variants(ctx, sc, w)
    for v in ("cold_reference_steps", "cold", "cold_dual", "warm", "cold_sched", "cold_dual_sched", "warm_sched"):
        assert r[v]["reached"], v
    print({v: (round(r[v]["ms"], 1), r[v]["leaf_iters"]) for v in ("cold", "cold_dual", "warm")})
    print("network vs the cold start with the recovered dual, leaf-iterations:",
          round(r["cold_dual"]["leaf_iters"] / r["warm"]["leaf_iters"], 3))
    # the network is kept on some levels and, with the same steps, saves work
    assert sum(r["warm"]["warm_levels"]) >= 1
    assert r["warm"]["leaf_iters"] < r["cold"]["leaf_iters"]
    assert r["warm_sched"]["leaf_iters"] < r["cold_sched"]["leaf_iters"]
