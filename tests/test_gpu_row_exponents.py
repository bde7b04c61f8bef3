"""Row exponents of the fp16 row-pass output.  With the bf16, bf16x2 and
bf16mix methods a leaf's unary (a source whose maximum no kernel computes)
is exponentiated against a power-of-two exponent per row, read back by the
column pass; every other source keeps its grid maximum.

* leaf unaries shifted by large constants (row exponents near +-870 and
  -60), a chain node's unary shifted, an upright root (the torso's
  log_from_root grids are LOG_ZERO on every other rotation): the FP64
  oracle's marginals within the method's budget and identical MAPs (MAP
  guard on);
* the leaf max runs exactly when some level keeps an FP32 row pass: no
  leaf-max launch for the fp16 methods, one per sweep for bf16x3, for the
  FP32 method and for bf16mix with an off-diagonal joint (the eigenbasis
  path keeps the FP32 row pass).
"""
import sys

import numpy as np
import pytest

from partapp_amd.api import ExpParam, Plan, person_model, synth_joints, synth_unary

from test_gpu_bf16 import METHOD_L1, errors_vs_f64, run_method
from test_gpu_offdiag import full_cov_joints

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("method", ["bf16mix", "bf16x2"])
def test_shifted_unaries_and_upright_root(gpu, method):
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    is_upright = [0] * 10
    is_upright[4] = 1
    joints = [synth_joints(300), synth_joints(301)]
    model = person_model(ep, joints[0], is_upright=is_upright)
    sizes = [(104, 136), (96, 120)]
    us = [synth_unary(310 + i, 10, 1, 24, h, w) for i, (h, w) in enumerate(sizes)]
    # leaves 0 and 9 far above and below, leaf 6 just below zero, chain node 1 shifted
    for u in us:
        u[0] += 600.0
        u[9] -= 600.0
        u[6] -= 40.5
        u[1] -= 300.0
    say("shifted unaries, upright root:", method, sizes)
    outs, (levels, used) = run_method(model, us, joints, method)
    assert used == levels
    for i in range(len(sizes)):
        worst, map_bad = errors_vs_f64(model, us[i], joints[i], outs[i])
        say("  %s instance %d %s: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (method, i, sizes[i], worst,
                                                                                   map_bad))
        assert worst <= METHOD_L1[method]
        assert not map_bad, map_bad


def leaf_max_launches(joints, method):
    import torch
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    model = person_model(ep, joints)
    u = synth_unary(320, 10, 1, 24, 64, 72)
    plan = Plan(model, [(64, 72)], [joints])
    plan.set_method(method)
    dev = torch.device("cuda:0")
    un = torch.from_numpy(u.ravel()).to(dev)
    marg = torch.empty_like(un)
    plan.set_profiling(True, classes=("leaf_max", "rowpass"))
    plan.read_profile()
    plan.execute(un, marg)
    torch.cuda.synchronize()
    prof = plan.read_profile()
    plan.set_profiling(False)
    plan.close()
    assert prof["rowpass"][1] > 0
    return prof["leaf_max"][1]


def test_leaf_max_only_where_a_level_keeps_fp32(gpu):
    plain = synth_joints(321)
    for method in ("bf16mix", "bf16x2", "bf16"):
        assert leaf_max_launches(plain, method) == 0, method
    for method in ("bf16x3", "fp32"):
        assert leaf_max_launches(plain, method) == 1, method
    assert leaf_max_launches(full_cov_joints(322, every=3), "bf16mix") == 1
