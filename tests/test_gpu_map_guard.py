"""The MAP guard of the BF16 methods (ps_plan_set_map_guard): instances
whose MAP certificate fails (some part's best log marginal beats its second
best by at most tau) are swept again in FP32 on the device.

* tau huge: every instance is re-evaluated, so every output is bit for bit
  the FP32 method's;
* tau tiny: none is (the random unaries leave no exact ties), so every
  output is bit for bit the unguarded BF16 method's;
* a batch with one instance of constant unaries (exact ties everywhere, gap
  0): only that instance is re-evaluated, and takes the FP32 outputs;
* graph replays of a guarded plan equal its eager execution;
* the error of the bf16mix log marginals near each part's maximum against
  the FP64 oracle, which sets the guard's threshold.
"""
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import ExpParam, Plan, decode_maps, person_model, synth_joints, synth_unary

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def execute(plan, un, method, tau=0.0, graph=False):
    import torch
    plan.set_method(method)
    plan.set_map_guard(tau)
    marg = torch.full_like(un, float("nan"))
    root = torch.full((plan.root_numel,), float("nan"), device=un.device)
    maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=un.device)
    if graph:
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            plan.execute(un, marg, root, maps, stream=s, graph=True)
        s.synchronize()
    else:
        plan.execute(un, marg, root, maps)
    torch.cuda.synchronize()
    n, flags = plan.read_map_guard()
    return marg, root, maps, n, flags


def same(a, b, i=None, plan=None, kind="unary"):
    import torch
    if i is None:
        return torch.equal(a, b)
    return torch.equal(plan.instance_view(a, i, kind), plan.instance_view(b, i, kind))


def make(sizes, seeds, const=()):
    import torch
    from partapp_amd.api import Joint
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = [synth_joints(s) for s in seeds]
    for k in const:  # flat unaries and centred joints: many exactly equal maxima
        joints[k] = [Joint(j.child_idx, j.parent_idx, (0.0, 0.0), (0.0, 0.0), (16.0, 0, 0, 16.0), 0.0, 0.0)
                     for j in joints[k]]
    model = person_model(ep, joints[0] if 0 not in const else joints[-1])
    us = []
    for k, ((h, w), s) in enumerate(zip(sizes, seeds)):
        us.append(np.full((10, 1, 24, h, w), -2.0, dtype=np.float32) if k in const else
                  synth_unary(s, 10, 1, 24, h, w))
    plan = Plan(model, sizes, joints)
    un = torch.from_numpy(np.concatenate([u.ravel() for u in us])).to("cuda:0")
    return plan, un, model, joints, us


def test_guard_all_and_none(gpu):
    plan, un, *_ = make([(96, 104), (88, 120)], [60, 61])
    say("guard: tau huge / tiny")
    m32, r32, p32, n0, _ = execute(plan, un, "fp32")
    mbf, rbf, pbf, _, _ = execute(plan, un, "bf16mix")
    assert n0 == 0
    ma, ra, pa, na, fa = execute(plan, un, "bf16mix", tau=1e30)
    assert na == 2 and fa == [1, 1]
    assert same(ma, m32) and same(ra, r32) and same(pa, p32)
    mn, rn, pn, nn, fn = execute(plan, un, "bf16mix", tau=1e-30)
    assert nn == 0 and fn == [0, 0]
    assert same(mn, mbf) and same(rn, rbf) and same(pn, pbf)
    plan.close()


def test_guard_reevaluates_only_the_tied_instance(gpu):
    plan, un, model, joints, us = make([(64, 72), (80, 96), (72, 64)], [62, 63, 64], const=(1,))
    say("guard: instance 1 has constant unaries")
    m32, r32, p32, _, _ = execute(plan, un, "fp32")
    mbf, rbf, pbf, _, _ = execute(plan, un, "bf16mix")
    mg, rg, pg, n, flags = execute(plan, un, "bf16mix", tau=1e-3)
    say("  re-evaluated %d: %s" % (n, flags))
    assert flags[1] == 1
    for i in range(3):
        ref_m, ref_r = (m32, r32) if flags[i] else (mbf, rbf)
        assert same(mg, ref_m, i, plan), i
        assert same(rg, ref_r, i, plan, "root"), i
    mg_d = decode_maps(pg, 3, 1, 10)
    for i in range(3):
        want = decode_maps(p32 if flags[i] else pbf, 3, 1, 10)[i]
        assert np.array_equal(mg_d[i], want), i
    # the tied instance now has the float32 oracle's MAPs (first index of the ties)
    _, _, b32, _ = oracle.root_posterior_rot(oracle.exp_param_from(model.exp_param), model.is_detect,
                                             model.is_upright, model.root, joints[1], us[1])
    for p in range(10):
        assert (int(mg_d[1, 0, p]["rotidx"]), int(mg_d[1, 0, p]["y"]), int(mg_d[1, 0, p]["x"])) == tuple(b32[0, p])
    plan.close()


def test_guard_graph_replays(gpu):
    plan, un, *_ = make([(64, 72), (80, 96)], [65, 66], const=(0,))
    ref = execute(plan, un, "bf16mix", tau=1e-3)
    for rep in range(3):
        got = execute(plan, un, "bf16mix", tau=1e-3, graph=True)
        _ = (un * 2).sum().item()  # eager work between replays
        assert same(got[0], ref[0]) and same(got[1], ref[1]) and same(got[2], ref[2]), rep
        assert got[3] == ref[3] and got[4] == ref[4]
    plan.close()


@pytest.mark.parametrize("method", ["bf16mix", "bf16x3"])
def test_log_error_spread_near_the_maximum(gpu, method):
    # max over states within 3 of each part's best FP64 value of
    # |(v_method - v_64) - (v_method - v_64)(FP64 argmax)|: a MAP can only move
    # to a state whose computed value is within this spread of the best, so
    # the guard's threshold must exceed it (the default keeps a 2x margin)
    from partapp_amd.api import config_exp_param, synth_unary_dev
    import torch
    cases = [("300x300", [(300, 300), (300, 300)], [0, 31], 1), ("uneven, 3 scales", [(232, 408)], [77], 2)]
    worst = 0.0
    for name, sizes, seeds, cfg in cases:
        ep = config_exp_param(cfg)
        joints = [synth_joints(s) for s in seeds]
        model = person_model(ep, joints[0])
        S = ep.num_scale_steps
        us = [synth_unary(s, 10, S, 24, h, w) for s, (h, w) in zip(seeds, sizes)]
        plan = Plan(model, sizes, joints)
        un = torch.from_numpy(np.concatenate([u.ravel() for u in us])).to("cuda:0")
        mx, _, _, _, _ = execute(plan, un, method)
        for i in range(len(sizes)):
            _, p64, _, _ = oracle.root_posterior_rot(oracle.exp_param_from(ep), model.is_detect, model.is_upright,
                                                     model.root, joints[i], us[i].astype(np.float64))
            m = plan.instance_view(mx, i).cpu().numpy()
            for s in range(S):
                for p in range(10):
                    v64 = p64[p, s].ravel()
                    vb = m[p, s].ravel().astype(np.float64)
                    j = int(np.argmax(v64))
                    near = v64 >= v64[j] - 3.0
                    worst = max(worst, float(np.abs((vb[near] - v64[near]) - (vb[j] - v64[j])).max()))
        plan.close()
        say("%s %s: max log-error spread within 3 of the maximum so far %.3g" % (method, name, worst))
    assert worst <= 0.5 * Plan.MAP_GUARD_TAU[method]
