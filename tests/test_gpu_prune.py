"""The pruned execution (ps_plan_execute_pruned) of the BF16 method.

* keep masks taken from exact marginals (per part, the 64 x 64 tiles of least
  mass, at most 1e-4 in all, dropped): the pruned sweep equals the FP64 oracle run
  on the pruned model (the non-leaf parts' unaries set to LOG_ZERO on the
  dropped tiles, the leaves' marginals restricted to their kept tiles) within
  the BF16 budget, writes LOG_ZERO on every dropped tile, and stays within
  twice the dropped mass (plus the budget) of the exact sweep;
* every tile kept, or every instance sent to the fallback: bit for bit the
  unpruned sweep;
* the trained net's masks (when the weights are present): the instances it
  does not send to the fallback stay within twice their true dropped mass.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from partapp_amd import prune
from partapp_amd.api import Plan, config_exp_param, decode_maps, person_model, synth_joints, synth_unary

pytestmark = pytest.mark.gpu
BF16_L1 = 1e-3


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(sizes, seeds):
    import torch
    ep = config_exp_param(1)
    joints = [synth_joints(s) for s in seeds]
    model = person_model(ep, joints[0])
    us = [synth_unary(s, 10, 1, 24, h, w) for s, (h, w) in zip(seeds, sizes)]
    plan = Plan(model, sizes, joints)
    plan.set_method("bf16mix")
    un = torch.from_numpy(np.concatenate([u.ravel() for u in us])).to("cuda:0")
    return ep, joints, model, us, plan, un


def run(plan, un, keep=None, fallback=None):
    import torch
    marg = torch.full_like(un, float("nan"))
    root = torch.full((plan.root_numel,), float("nan"), device=un.device)
    maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=un.device)
    if keep is None:
        plan.execute(un, marg, root, maps)
    else:
        plan.execute_pruned(un, marg, root, maps, keep, fallback)
    torch.cuda.synchronize()
    return marg, root, maps


def leaves(model):
    parents = {j.parent_idx for j in model.joints}
    return [p for p in range(model.num_parts) if p not in parents]


def test_exact_masks_match_the_pruned_model(gpu):
    import torch
    sizes = [(200, 232), (256, 256)]
    ep, joints, model, us, plan, un = setup(sizes, [200, 201])
    m_exact, _, _ = run(plan, un)
    pr = prune.Pruner(plan, gpu, net=object())
    masses = [[prune.cell_mass(plan.instance_view(m_exact, i), 0)] for i in range(2)]
    pr.update(un, masses=masses)
    say("exact masks: kept tiles %.3f, dropped mass %s" % (pr.kept_fraction(), pr.predicted_dropped.tolist()))
    m_pr, _, maps_pr = run(plan, un, pr.keep, None)
    offs, _ = plan.keep_layout()
    keep_h = pr.keep.cpu().numpy()
    leaf = leaves(model)
    for i, (H, W) in enumerate(sizes):
        ty, tx = (H + 63) // 64, (W + 63) // 64
        k = keep_h[offs[i]:offs[i] + 10 * 24 * ty * tx].reshape(10, 24, ty, tx)
        kfull = np.repeat(np.repeat(k, 64, axis=2), 64, axis=3)[:, :, :H, :W].astype(bool)  # [P, R, H, W]
        dropped = float(pr.predicted_dropped[i])
        mg = plan.instance_view(m_pr, i).cpu().numpy()[:, 0]
        me = plan.instance_view(m_exact, i).cpu().numpy()[:, 0]
        # the pruned model: non-leaf unaries at LOG_ZERO on their dropped tiles
        u = us[i].copy()
        for p in range(10):
            if p not in leaf:
                u[p, 0][~kfull[p]] = -1e6
        _, p64, _, _ = oracle.root_posterior_rot(oracle.exp_param_from(ep), model.is_detect, model.is_upright,
                                                 model.root, joints[i], u.astype(np.float64))
        worst_model = worst_exact = 0.0
        for p in range(10):
            if p != model.root:
                assert np.all(mg[p][~kfull[p]] == np.float32(-1e6)), (i, p)
            ref = p64[p, 0].copy()
            if p in leaf:
                ref[~kfull[p]] = -np.inf
            worst_model = max(worst_model, oracle.rel_l1(oracle.normalized(mg[p]), oracle.normalized(ref)))
            worst_exact = max(worst_exact, oracle.rel_l1(oracle.normalized(mg[p]), oracle.normalized(me[p])))
        say("  instance %d %s: dropped %.3g; vs the pruned model %.3g, vs the exact sweep %.3g" % (
            i, sizes[i], dropped, worst_model, worst_exact))
        assert worst_model <= BF16_L1
        assert worst_exact <= 2 * dropped + BF16_L1
    plan.close()


def test_all_kept_or_all_fallback_is_the_unpruned_sweep(gpu):
    import torch
    ep, joints, model, us, plan, un = setup([(96, 104), (88, 120)], [210, 211])
    ref = run(plan, un)
    _, total = plan.keep_layout()
    keep = torch.ones(total, dtype=torch.uint8, device=gpu)
    got = run(plan, un, keep, None)
    assert all(torch.equal(a, b) for a, b in zip(got, ref))
    zeros = torch.zeros(total, dtype=torch.uint8, device=gpu)
    fb = torch.ones(2, dtype=torch.int32, device=gpu)
    got = run(plan, un, zeros, fb)
    assert all(torch.equal(a, b) for a, b in zip(got, ref))
    plan.close()


@pytest.mark.skipif(not os.path.exists(prune.WEIGHTS), reason="no trained pruning net")
def test_net_masks_stay_within_the_dropped_mass(gpu):
    sizes = [(300, 300)] * 4
    ep, joints, model, us, plan, un = setup(sizes, [220, 221, 222, 223])
    m_exact, _, maps_e = run(plan, un)
    pr = prune.Pruner(plan, gpu)
    pr.update(un)
    fb = pr.fallback.cpu().numpy()
    m_pr, _, maps_p = run(plan, un, pr.keep, pr.fallback)
    offs, _ = plan.keep_layout()
    keep_h = pr.keep.cpu().numpy()
    say("net masks: kept tiles %.3f, predicted dropped %s, fallback %s" % (
        pr.kept_fraction(), [round(v, 6) for v in pr.predicted_dropped.tolist()], fb.tolist()))
    for i, (H, W) in enumerate(sizes):
        me = plan.instance_view(m_exact, i).cpu().numpy()[:, 0]
        mg = plan.instance_view(m_pr, i).cpu().numpy()[:, 0]
        if fb[i]:
            assert np.array_equal(me, mg)
            continue
        ty, tx = (H + 63) // 64, (W + 63) // 64
        k = keep_h[offs[i]:offs[i] + 10 * 24 * ty * tx].reshape(10, 24, ty, tx)
        true_drop = 0.0
        worst = 0.0
        for p in range(10):
            pe = oracle.normalized(me[p])
            kfull = np.repeat(np.repeat(k[p], 64, axis=1), 64, axis=2)[:, :H, :W].astype(bool)
            true_drop += float(pe[~kfull].sum())
            worst = max(worst, oracle.rel_l1(oracle.normalized(mg[p]), pe))
        say("  instance %d: true dropped mass %.3g, rel L1 vs exact %.3g" % (i, true_drop, worst))
        assert worst <= 2 * true_drop + BF16_L1
    plan.close()
