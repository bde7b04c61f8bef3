"""Parity at the benchmark workloads' own sizes.

* The metric's batch: 32 instances of 300 x 300 positions x 24 rotations x 1
  scale, the 10-part person tree, per-image spatial models (seeds 0..31), one
  plan.  The plain FP32 sweep is held to the FP32 bar against the float32
  oracle (relative L1 of every part's normalised marginal <= 1e-5, identical
  MAP states) and the BF16 method (bf16mix) to <= 1e-3 relative L1 against the
  FP64 oracle with identical MAP states, on instances 0, 15 and 31.
* The uneven batch: sizes drawn like the 256-instance workload (H, W uniform
  in 160..480, multiples of 8) plus its extremes 480 x 480 and 160 x 160, 24
  rotations x 3 scales (0.8, 1.0, 1.2), per-image spatial models, one plan;
  both methods on every instance of the plan.

The unaries of the checked instances are read back from the device buffers
the sweep read, so the oracle sees the very same inputs.
"""
import concurrent.futures as cf
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import (Plan, config_exp_param, decode_maps, person_model, synth_joints, synth_unary,
                             synth_unary_dev, uneven_sizes)

pytestmark = pytest.mark.gpu

FP32_L1 = 1e-5
BF16_L1 = 1e-3


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def run_both(plan, unary):
    """The FP32 sweep and the bf16mix sweep of one plan on the same unaries;
    returns {method: (marginal tensor, maps [n][S][P])}."""
    import torch
    out = {}
    for method in ("fp32", "bf16mix"):
        levels, used = plan.set_method(method)
        if method != "fp32":
            assert used == levels, "every level of this workload runs on the matrix cores"
        marg = torch.full_like(unary, float("nan"))
        root = torch.full((plan.root_numel,), float("nan"), device=unary.device)
        maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=unary.device)
        plan.execute(unary, marg, root, maps)
        torch.cuda.synchronize()
        out[method] = (marg, decode_maps(maps, plan.n_instances, plan.S, plan.P))
    return out


def check_instance(plan, model, joints, unary, outs, i):
    """Errors of instance i: (FP32 vs f32 oracle, FP32 vs f64, bf16mix vs f64,
    MAP mismatches of each leg)."""
    ep = oracle.exp_param_from(model.exp_param)
    u = plan.instance_view(unary, i).cpu().numpy()
    m32 = plan.instance_view(outs["fp32"][0], i).cpu().numpy()
    mbf = plan.instance_view(outs["bf16mix"][0], i).cpu().numpy()

    def ref(dt):
        return oracle.root_posterior_rot(ep, model.is_detect, model.is_upright, model.root, joints[i], u.astype(dt))

    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f32, f64 = ex.submit(ref, np.float32), ex.submit(ref, np.float64)
        _, p32, b32, _ = f32.result()
        _, p64, b64, _ = f64.result()
    e32 = e32_64 = ebf = 0.0
    bad32, badbf = [], []
    for s in range(plan.S):
        for p in range(plan.P):
            n64 = oracle.normalized(p64[p, s])
            n32 = oracle.normalized(m32[p, s])
            e32 = max(e32, oracle.rel_l1(n32, oracle.normalized(p32[p, s])))
            e32_64 = max(e32_64, oracle.rel_l1(n32, n64))
            ebf = max(ebf, oracle.rel_l1(oracle.normalized(mbf[p, s]), n64))
            for mp, want, bad, refg in ((outs["fp32"][1], b32, bad32, p32), (outs["bf16mix"][1], b64, badbf, p64)):
                got = (int(mp[i, s, p]["rotidx"]), int(mp[i, s, p]["y"]), int(mp[i, s, p]["x"]))
                if got != tuple(want[s, p]):
                    bad.append((s, p, got, tuple(want[s, p]), oracle.top2_gap(refg[p, s])))
    return e32, e32_64, ebf, bad32, badbf


def assert_instance(tag, res):
    e32, e32_64, ebf, bad32, badbf = res
    say("%s: fp32 rel L1 %.3g vs f32 oracle (%.3g vs f64), bf16mix %.3g vs f64; MAP mismatches fp32 %s, bf16mix %s"
        % (tag, e32, e32_64, ebf, bad32, badbf))
    assert e32 <= FP32_L1, "%s: FP32 rel L1 %.3g" % (tag, e32)
    assert not bad32, "%s: FP32 MAP differs from the float32 oracle: %s" % (tag, bad32)
    assert ebf <= BF16_L1, "%s: bf16mix rel L1 %.3g vs FP64" % (tag, ebf)
    assert not badbf, "%s: bf16mix MAP differs from the FP64 oracle: %s" % (tag, badbf)


def test_metric_batch_32x300x300x24(gpu):
    import torch
    ep = config_exp_param(1)
    n, H, W = 32, 300, 300
    joints = [synth_joints(s) for s in range(n)]
    model = person_model(ep, joints[0])
    plan = Plan(model, [(H, W)] * n, joints)
    say("plan: 32 x 300x300x24, scratch %.2f GB" % (plan.scratch_bytes() / 1e9))
    unary = torch.empty((plan.unary_numel,), dtype=torch.float32, device=gpu)
    for i in range(n):
        o = plan.unary_offsets[i]
        unary[o:o + 10 * 24 * H * W].copy_(torch.from_numpy(synth_unary(i, 10, 1, 24, H, W, threads=16).ravel()))
    outs = run_both(plan, unary)
    for i in (0, 15, 31):
        assert_instance("instance %d" % i, check_instance(plan, model, joints, unary, outs, i))
    plan.close()


def test_uneven_batch_three_scales(gpu):
    import torch
    ep = config_exp_param(2)
    sizes = uneven_sizes(4) + [(480, 480), (160, 160)]
    n = len(sizes)
    joints = [synth_joints(1000 + i) for i in range(n)]
    model = person_model(ep, joints[0])
    plan = Plan(model, sizes, joints)
    say("plan: uneven %s x 24 rotations x 3 scales, scratch %.2f GB" % (sizes, plan.scratch_bytes() / 1e9))
    unary = torch.empty((plan.unary_numel,), dtype=torch.float32, device=gpu)
    for i, (h, w) in enumerate(sizes):
        o = plan.unary_offsets[i]
        synth_unary_dev(2000 + i, 10, 3, 24, h, w, unary[o:o + 10 * 3 * 24 * h * w])
    outs = run_both(plan, unary)
    for i in range(n):
        assert_instance("instance %d %s" % (i, sizes[i]), check_instance(plan, model, joints, unary, outs, i))
    plan.close()
