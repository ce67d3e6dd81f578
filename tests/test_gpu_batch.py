"""Batches of independent replicas: one context advancing replicas of
different boxes, chain counts and seeds must give, replica by replica, the
same bits as each replica run alone."""
import numpy as np
import pytest

from mfa_amd import configs, lib

pytestmark = pytest.mark.gpu


def _replicas():
    cases = configs.small_cases()
    out = []
    for name in ("brush_small", "brush_bend"):
        kind, inp = cases[name]
        out.append((kind, inp))
    # a third, smaller brush with its own seed and box
    out.append((1, configs.brush(6, 8, 150, 5.0, 5.0, 7.0, x_space=0.5, seed=21)))
    states = []
    for kind, inp in out:
        p, r, v, t = configs.build(inp, kind)
        states.append((p, r, v, t))
    return states


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_batch_matches_single_runs(precision):
    states = _replicas()
    print("batch of", len(states), "replicas, sizes", [p.n_part for p, *_ in states], flush=True)
    singles = []
    for p, r, v, t in states:
        with lib.Context(p, precision) as c:
            c.set_state(r, v, np.zeros_like(r), t)
            obs = c.run_steps(1, 40)
            singles.append((c.get_state(), obs))
    with lib.Batch([s[0] for s in states], precision) as b:
        for q, (p, r, v, t) in enumerate(states):
            b.set_replica(q, r, v, np.zeros_like(r), t)
        b.run_steps(1, 40)
        for q in range(len(states)):
            st = b.get_replica(q)
            ob = b.replica_obs(q)
            ref, robs = singles[q]
            for k in ("r", "v", "f", "mic"):
                assert np.array_equal(st[k], ref[k]), (q, k)
            assert ob["n_rebuilds"] == robs["n_rebuilds"]
            assert ob["v_fluid_fluid"] == robs["v_fluid_fluid"]
            assert ob["t_kin"] == robs["t_kin"]


def test_autotune_picks_a_supported_width():
    states = _replicas()
    with lib.Batch([s[0] for s in states], lib.MFA_MIXED) as b:
        for q, (p, r, v, t) in enumerate(states):
            b.set_replica(q, r, v, np.zeros_like(r), t)
        tpp, ms = b.autotune(1, 20)
        assert tpp in (1, 2, 4, 8, 16) and ms > 0.0
        b.run_steps(200, 10)
