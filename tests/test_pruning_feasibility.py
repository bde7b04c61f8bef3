"""Why the learned-pruning method is not built: on the synthetic workload the
exact marginals are diffuse, so the state-pruning rule (keep states with
marginal mass >= 1e-6, fall back when more than 1e-3 of the mass is
dropped) falls back on every part.  Measured on the float64 oracle."""
import numpy as np

import oracle
from partapp_amd.api import ExpParam, person_model, synth_joints, synth_unary


def test_keep_rule_drops_more_than_the_fallback_budget():
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(0)
    model = person_model(ep, joints)
    u = synth_unary(0, 10, 1, 24, 120, 136).astype(np.float64)
    _, post, _, _ = oracle.root_posterior_rot(oracle.exp_param_from(ep), model.is_detect, model.is_upright,
                                              model.root, joints, u)
    for p in range(10):
        q = oracle.normalized(post[p, 0]).ravel()
        dropped = q[q < 1e-6].sum()
        qs = np.sort(q)[::-1]
        n999 = int(np.searchsorted(np.cumsum(qs), 1 - 1e-3)) + 1
        assert dropped > 1e-3, (p, dropped)  # the instance falls back
        # holding 99.9% of the mass takes more than 5% of the states: under 20x fewer states
        assert n999 / q.size > 0.05, (p, n999 / q.size)
