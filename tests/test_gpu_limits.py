"""Shapes at the edges of the launch geometry, against the float32 oracle at
the FP32 bar (relative L1 <= 1e-5 per part, identical MAP states):

* two plans whose row passes need more than 64 KB of LDS, the wider one first:
  creating the narrower plan must not lower the wider plan's LDS limit;
* a y filter with half width >= 96 (sigma_y * scale >= 32 px), whose FP32
  column-pass tile needs more than 64 KB of LDS;
* the reference's default rotation grid, 48 steps (ExpParam.proto:195), on the
  48-slice FP32 row pass (its BF16 methods: tests/test_gpu_bf16.py).
"""
import sys

import numpy as np
import pytest

from partapp_amd.api import ExpParam, Joint, Plan, decode_maps, person_model, synth_joints, synth_unary

from test_gpu_parity import check_parity, run_plan

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def upload(plan, unaries, gpu):
    import torch
    un = torch.from_numpy(np.concatenate([u.ravel() for u in unaries])).to(gpu)
    marg = torch.full_like(un, float("nan"))
    root = torch.full((plan.root_numel,), float("nan"), device=gpu)
    maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=gpu)
    return un, marg, root, maps


def outputs(plan, marg, root, maps):
    import torch
    torch.cuda.synchronize()
    m = decode_maps(maps, plan.n_instances, plan.S, plan.P)
    return [(plan.instance_view(marg, i).cpu().numpy(), plan.instance_view(root, i, "root").cpu().numpy(), m[i])
            for i in range(plan.n_instances)]


def test_lds_limit_only_rises_across_plans(gpu):
    from partapp_amd._lib import lib
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(40)
    model = person_model(ep, joints)
    u_wide = synth_unary(41, 10, 1, 24, 96, 800)
    u_narrow = synth_unary(42, 10, 1, 24, 96, 700)
    wide = Plan(model, [(96, 800)], [joints])   # row pass ~83 KB of LDS
    narrow = Plan(model, [(96, 700)], [joints])  # ~73 KB, created second
    say("LDS: wide plan 96x800 then narrow 96x700, executing the wide one")
    un, marg, root, maps = upload(wide, [u_wide], gpu)
    wide.execute(un, marg, root, maps)
    out = outputs(wide, marg, root, maps)[0]
    check_parity("96x800 after a 96x700 plan", model, u_wide, joints, out, check_f64=False)
    un2, marg2, root2, maps2 = upload(narrow, [u_narrow], gpu)
    narrow.execute(un2, marg2, root2, maps2)
    out2 = outputs(narrow, marg2, root2, maps2)[0]
    check_parity("96x700", model, u_narrow, joints, out2, check_f64=False)
    wide.close()
    narrow.close()
    assert lib() is not None


def test_fp32_column_pass_wide_y_filter(gpu):
    # sigma_y = 32 px on two edges (half width 96: a 264-row column tile of
    # 64 columns, 67.6 KB of LDS), the other edges as generated
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    base = synth_joints(43)
    joints = [Joint(j.child_idx, j.parent_idx, j.offset_c, j.offset_p,
                    (j.C[0], 0.0, 0.0, 1024.0) if i in (1, 6) else j.C, j.rot_mean, j.rot_sigma)
              for i, j in enumerate(base)]
    model = person_model(ep, joints)
    u = synth_unary(44, 10, 1, 24, 104, 120)
    say("wide y filter: sigma_y 32 px on joints 1 and 6")
    out = run_plan(model, [u], [joints])[0]
    check_parity("ky 96", model, u, joints, out)


def test_default_rotation_grid_48(gpu):
    ep = ExpParam(-180.0, 180.0, 48, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(45)
    model = person_model(ep, joints)
    u = synth_unary(46, 10, 1, 48, 112, 128)
    say("48 rotations, 112x128")
    plan = Plan(model, [(112, 128)], [joints])
    un, marg, root, maps = upload(plan, [u], gpu)
    plan.set_method("fp32")
    plan.execute(un, marg, root, maps)
    ref = outputs(plan, marg, root, maps)[0]
    check_parity("R = 48", model, u, joints, ref)
    plan.close()


def test_launch_shape_knobs_do_not_change_results(gpu):
    # the FP32 row pass at 64..512 threads per block (a thread then covers
    # several columns), the FP32 column pass at 8 and 16 output rows per
    # lane, the BF16 x pass's rows per block, and per-class settings of each
    # knob: bit for bit the default launch shapes
    import torch
    from partapp_amd._lib import PsError
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = [synth_joints(47), synth_joints(48)]
    model = person_model(ep, joints[0])
    sizes = [(72, 136), (64, 200)]
    us = [synth_unary(49 + i, 10, 1, 24, h, w) for i, (h, w) in enumerate(sizes)]
    plan = Plan(model, sizes, joints)
    classes = plan.classes()
    say("launch classes:", classes)
    assert len(classes) >= 4 and all(c["jobs"] > 0 for c in classes)
    for method, knob, values in (("fp32", plan.TUNE_FP32_ROW_THREADS, (64, 128, 256, 512)),
                                 ("fp32", plan.TUNE_FP32_COL_ROWS, (8, 16)),
                                 ("bf16mix", plan.TUNE_MFMA_ROWS, (1, 3, 16))):
        plan.set_method(method)
        plan.tune(knob, 0 if method == "fp32" else 4)
        un, marg, root, maps = upload(plan, us, gpu)
        plan.execute(un, marg, root, maps)
        ref = [t.clone() for t in (marg, root, maps)]
        for v in values:
            plan.tune(knob, v)
            marg.fill_(float("nan"))
            plan.execute(un, marg, root, maps)
            torch.cuda.synchronize()
            assert torch.equal(marg, ref[0]) and torch.equal(root, ref[1]) and torch.equal(maps, ref[2]), (method, v)
        plan.tune(knob, 0 if method == "fp32" else 4)
        plan.tune(knob, values[0], cls=1)
        plan.tune(knob, values[-1], cls=2)
        marg.fill_(float("nan"))
        plan.execute(un, marg, root, maps)
        torch.cuda.synchronize()
        assert torch.equal(marg, ref[0]), method
    with pytest.raises(PsError):
        plan.tune(plan.TUNE_FP32_ROW_THREADS, 100)
    with pytest.raises(PsError):
        plan.tune(plan.TUNE_FP32_COL_ROWS, 12)
    with pytest.raises(PsError):
        plan.tune(plan.TUNE_MFMA_ROWS, 4, cls=len(classes))
    plan.close()
