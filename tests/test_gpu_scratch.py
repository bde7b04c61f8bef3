"""The plan's scratch and the caller's marginal grids.

* Footprint: per (instance, scale) the plan holds K row-pass outputs, K message
  slots and one upward-belief grid per inner part (a part with a parent and a
  child), each [R][H][W'] with W' = W rounded up to 32 floats and each grid
  rounded up to 64 floats; log_from_root lives in the marginal grid of each
  chain's first part.  The 256-instance uneven batch only fits in HBM beside
  its unaries and marginals at this size, so it is pinned here.
* The marginal grids carry intermediate grids during a sweep: a second
  execution on the same buffers (holding the first one's results) gives the
  same outputs bit for bit, as does one on buffers filled with other values.
* Narrow widths, where a column strip is mostly the padding columns of the
  plan's grids (written as LOG_ZERO, never read as data): FP32 against the
  float32 oracle at its bar, bf16mix against FP64 at the method's bar (1e-3)
  with the MAP guard on (every MAP equal).
"""
import sys

import numpy as np
import pytest

from partapp_amd.api import ExpParam, Plan, person_model, synth_joints, synth_unary

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def r64(n):
    return (n + 63) & ~63


def test_scratch_footprint(gpu):
    ep = ExpParam(-180.0, 180.0, 24, 0.8, 1.2, 2, 0.0)
    model = person_model(ep, synth_joints(401))
    sizes = [(300, 300), (160, 488), (41, 57)]
    plan = Plan(model, sizes, [synth_joints(402 + i) for i in range(len(sizes))])
    K, inner, R, S = 5, 4, 24, 2  # person tree: five chains off the torso, four inner parts
    want = 0
    for h, w in sizes:
        grid = r64(R * h * ((w + 31) & ~31))
        want += S * (K + K + inner) * grid
    say("scratch %d bytes, expected %d" % (plan.scratch_bytes(), 4 * want))
    assert plan.scratch_bytes() == 4 * want


@pytest.mark.parametrize("method", ["fp32", "bf16mix"])
def test_marginal_buffers_start_state_does_not_matter(gpu, method):
    import torch
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    sizes = [(72, 100), (96, 84)]
    joints = [synth_joints(410 + i) for i in range(len(sizes))]
    model = person_model(ep, joints[0])
    plan = Plan(model, sizes, joints)
    plan.set_method(method)
    un = torch.from_numpy(np.concatenate([synth_unary(420 + i, 10, 1, 24, h, w).ravel()
                                          for i, (h, w) in enumerate(sizes)])).to(gpu)
    root = torch.zeros((plan.root_numel,), device=gpu)
    maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=gpu)
    say("%s: three executions on differently filled marginal buffers" % method)
    marg = torch.full_like(un, float("nan"))
    plan.execute(un, marg, root, maps)
    torch.cuda.synchronize()
    ref = (marg.clone(), root.clone(), maps.clone())
    plan.execute(un, marg, root, maps)  # the buffers hold the first execution's outputs
    torch.cuda.synchronize()
    for a, b in zip((marg, root, maps), ref):
        assert torch.equal(a, b)
    marg.fill_(123.0)
    root.fill_(-7.0)
    plan.execute(un, marg, root, maps)
    torch.cuda.synchronize()
    for a, b in zip((marg, root, maps), ref):
        assert torch.equal(a, b)
    assert torch.isfinite(marg).all()


# the last column strip holds 1, 33 and 1 columns of the grid and 31 padding
# columns (grids narrower than the joints' offsets of up to 40 px would move
# whole messages out of the grid)
NARROW = [(96, 65), (80, 97), (88, 129)]


def test_narrow_widths_fp32(gpu):
    from test_gpu_parity import check_parity, run_plan
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = [synth_joints(430 + i) for i in range(len(NARROW))]
    model = person_model(ep, joints[0])
    us = [synth_unary(440 + i, 10, 1, 24, h, w) for i, (h, w) in enumerate(NARROW)]
    say("FP32 on widths", [w for _, w in NARROW])
    outs = run_plan(model, us, joints)
    for i, u in enumerate(us):
        check_parity("%d x %d" % NARROW[i], model, u, joints[i], outs[i], check_f64=False)


def test_narrow_widths_bf16mix(gpu):
    from test_gpu_bf16 import BUDGET_L1, errors_vs_f64, run_method
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = [synth_joints(450 + i) for i in range(len(NARROW))]
    model = person_model(ep, joints[0])
    us = [synth_unary(460 + i, 10, 1, 24, h, w) for i, (h, w) in enumerate(NARROW)]
    say("bf16mix on widths", [w for _, w in NARROW])
    outs, _ = run_method(model, us, joints, "bf16mix")
    for i, u in enumerate(us):
        worst, map_bad = errors_vs_f64(model, u, joints[i], outs[i])
        say("  %d x %d: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (NARROW[i] + (worst, map_bad)))
        # the method's bar (these small grids reach 4.8e-4 / 8.4e-4 / 3.3e-4,
        # the same as with dense rows): padding read as data would be far off
        assert worst <= BUDGET_L1
        assert not map_bad
