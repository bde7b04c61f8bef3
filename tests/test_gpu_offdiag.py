"""Parity of the eigenbasis path (joints whose covariance has off-diagonal
terms: gaussFilter2dOffset, multi_array_filter.hpp:334-369) with the oracle:
single messages with the nearest-scatter (sparse, upward) and bilinear
(downward) rotations, and whole sweeps mixing full and diagonal covariances.
Same bar as the diagonal path: rel L1 of the normalised marginals <= 1e-5 vs
the float32 oracle, identical MAP, identical LOG_ZERO placement."""
import math
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import ExpParam, Joint, compute_rot_joint_marginal, person_model, synth_joints, synth_unary

from test_gpu_parity import check_parity, run_plan

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


OD_MESSAGES = {
    "tilted": dict(offset_c=(7.3, -11.2), offset_p=(-3.9, 8.8), C=(30.0, 12.0, 12.0, 50.0), rot_mean=0.3,
                   rot_sigma=0.25, scale=1.0),
    "negative_corr_scaled": dict(offset_c=(-5.0, 9.0), offset_p=(4.0, 2.5), C=(60.0, -25.0, -25.0, 20.0),
                                 rot_mean=-0.4, rot_sigma=0.3, scale=1.3),
    "thin": dict(offset_c=(2.0, 2.0), offset_p=(0.0, -3.0), C=(40.0, 39.0, 39.0, 40.0), rot_mean=0.0,
                 rot_sigma=0.0, scale=0.8),
    "asymmetric_entries": dict(offset_c=(1.0, -1.0), offset_p=(3.0, 3.0), C=(25.0, 8.0, 0.0, 16.0), rot_mean=0.2,
                               rot_sigma=0.5, scale=1.0),
}


@pytest.mark.parametrize("sparse", [True, False])
@pytest.mark.parametrize("case", sorted(OD_MESSAGES))
def test_offdiag_message_matches_oracle(gpu, case, sparse):
    import torch
    kw = OD_MESSAGES[case]
    R, H, W = 24, 47, 61
    ep = ExpParam(-180.0, 180.0, R, 1.0, 1.0, 1, 0.0)
    model = person_model(ep)
    rng = np.random.default_rng(sum(map(ord, case)) + int(sparse))
    child = rng.normal(-3, 1.0, (R, H, W)).astype(np.float32)
    say("off-diagonal message", case, "sparse" if sparse else "bilinear", kw)
    out = compute_rot_joint_marginal(model, torch.from_numpy(child).to(gpu), kw["offset_c"], kw["offset_p"], kw["C"],
                                     kw["rot_mean"], kw["rot_sigma"], kw["scale"], is_sparse=sparse).cpu().numpy()
    ref = oracle.rot_joint_marginal(oracle.exp_param_from(ep), child, is_sparse=sparse, **kw)
    np.testing.assert_array_equal(out <= -9e5, ref <= -9e5)
    live = ref > -9e5
    assert live.mean() > 0.3
    np.testing.assert_allclose(out[live], ref[live], rtol=1e-5, atol=5e-5)
    d = oracle.rel_l1(oracle.normalized(out), oracle.normalized(ref))
    say("  rel L1 %.3g, exact %.4f" % (d, np.mean(out == ref)))
    assert d <= 1e-6


def full_cov_joints(seed, every=1):
    """synth_joints with a correlation of +-0.45 added to every `every`-th joint."""
    out = []
    for i, j in enumerate(synth_joints(seed)):
        a, b = j.C[0], j.C[3]
        c = (0.45 if (i + seed) % 2 == 0 else -0.45) * math.sqrt(a * b) if i % every == 0 else 0.0
        out.append(Joint(j.child_idx, j.parent_idx, j.offset_c, j.offset_p, (a, c, c, b), j.rot_mean, j.rot_sigma))
    return out


@pytest.mark.parametrize("sparse", [True, False])
def test_offdiag_sweep_matches_oracle(gpu, sparse):
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = full_cov_joints(11)
    model = person_model(ep, joints, is_sparse=sparse)
    u = synth_unary(11, 10, 1, 24, 112, 128)
    say("off-diagonal sweep", "sparse" if sparse else "bilinear")
    out = run_plan(model, [u])[0]
    check_parity("off-diagonal sweep sparse=%s" % sparse, model, u, joints, out)


def test_offdiag_mixed_batch(gpu):
    # per-instance joints: one instance fully off-diagonal, one with every
    # other joint off-diagonal, one diagonal; two scales, masks on
    ep = ExpParam(-180.0, 180.0, 24, 0.8, 1.2, 2, 0.1)
    is_upright = [0] * 10
    is_upright[4] = is_upright[2] = 1
    joints = [full_cov_joints(40), full_cov_joints(41, every=2), synth_joints(42)]
    model = person_model(ep, joints[0], is_upright=is_upright, is_sparse=True)
    sizes = [(128, 144), (136, 120), (120, 128)]
    us = [synth_unary(50 + i, 10, 2, 24, h, w) for i, (h, w) in enumerate(sizes)]
    say("off-diagonal mixed batch", sizes)
    outs = run_plan(model, us, joints)
    for i in range(3):
        check_parity("mixed %d" % i, model, us[i], joints[i], outs[i], check_f64=(i == 0))
