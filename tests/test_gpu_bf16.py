"""The BF16 banded-Toeplitz spatial passes (PS_METHOD_BF16 / _BF16X3) against
the FP64 oracle: per-part relative L1 of the normalised marginals and MAP
agreement, on the same seeded inputs as the FP32 parity tests.  The FP32
method stays the drop-in; these methods carry their own error budget
(<= 1e-3 relative L1 against FP64, MAP identical)."""
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import ExpParam, Plan, config_exp_param, decode_maps, person_model, synth_joints, synth_unary

from test_gpu_offdiag import full_cov_joints
from test_gpu_parity import oracle_run

pytestmark = pytest.mark.gpu

BUDGET_L1 = 1e-3


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def run_method(model, unaries, joints, method, graph=False, guard=True):
    """With guard (the default) the plan's MAP guard runs at the method's
    threshold: instances whose MAP certificate fails are swept again in FP32,
    so every MAP must equal the FP64 oracle's."""
    import torch
    sizes = [u.shape[3:] for u in unaries]
    plan = Plan(model, sizes, joints)
    levels, used = plan.set_method(method)
    plan.set_map_guard(Plan.MAP_GUARD_TAU[method] if guard else 0.0)
    dev = torch.device("cuda:0")
    un = torch.from_numpy(np.concatenate([u.ravel() for u in unaries])).to(dev)
    marg = torch.full_like(un, float("nan"))
    root = torch.full((plan.root_numel,), float("nan"), device=dev)
    maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=dev)
    if graph:
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            plan.execute(un, marg, root, maps, stream=s, graph=True)
        s.synchronize()
    else:
        plan.execute(un, marg, root, maps)
        torch.cuda.synchronize()
    m = decode_maps(maps, plan.n_instances, plan.S, plan.P)
    out = [(plan.instance_view(marg, i).cpu().numpy(), plan.instance_view(root, i, "root").cpu().numpy(), m[i])
           for i in range(plan.n_instances)]
    plan.close()
    return out, (levels, used)


def errors_vs_f64(model, u, joints, out):
    marg, root, maps = out
    P, S = u.shape[0], u.shape[1]
    r64, p64, b64, s64 = oracle_run(model, u, joints, np.float64)
    worst, map_bad = 0.0, []
    for s in range(S):
        for p in range(P):
            worst = max(worst, oracle.rel_l1(oracle.normalized(marg[p, s]), oracle.normalized(p64[p, s])))
            got = (int(maps[s, p]["rotidx"]), int(maps[s, p]["y"]), int(maps[s, p]["x"]))
            if got != tuple(b64[s, p]):
                map_bad.append((s, p, got, tuple(b64[s, p]), oracle.top2_gap(p64[p, s])))
    return worst, map_bad


@pytest.mark.parametrize("method", ["bf16x3", "bf16x2", "bf16mix", "bf16"])
@pytest.mark.parametrize("seed", [0, 1])
def test_bf16_sweep_error_vs_f64(gpu, method, seed):
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(seed)
    model = person_model(ep, joints)
    u = synth_unary(seed, 10, 1, 24, 112, 128)
    say("bf16 sweep", method, "seed", seed)
    outs, (levels, used) = run_method(model, [u], [joints], method)
    assert used == levels == 4  # two upward and two downward levels
    worst, map_bad = errors_vs_f64(model, u, joints, outs[0])
    say("  %s seed %d: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (method, seed, worst, map_bad))
    assert worst <= min(BUDGET_L1, METHOD_L1[method])
    assert not map_bad


def test_bf16x3_cfg0_full_size(gpu):
    ep = config_exp_param(0)
    joints = synth_joints(0)
    model = person_model(ep, joints)
    u = synth_unary(0, 10, 1, 24, 200, 200)
    outs, _ = run_method(model, [u], [joints], "bf16x3")
    worst, map_bad = errors_vs_f64(model, u, joints, outs[0])
    say("bf16x3 cfg0: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (worst, map_bad))
    assert worst <= BUDGET_L1 and not map_bad


# per-method budgets of the tests: the method's own error level with
# margin, so that a change that loses precision (e.g. rounding a filter's taps
# to bf16 where they were split) fails even while it stays under 1e-3
METHOD_L1 = {"bf16x3": 1e-5, "bf16x2": 7e-4, "bf16mix": 7e-4, "bf16": 1e-3}


@pytest.mark.parametrize("method", ["bf16x3", "bf16mix"])
def test_uneven_scales_masks_and_offdiagonal(gpu, method):
    # partial tiles (W not a multiple of 32, H not of 64), three scales (the
    # filters and the fp16 scale differ per scale), masks (the root upright),
    # and an instance whose off-diagonal joints keep the FP32 eigenbasis path
    # inside the same levels
    ep = ExpParam(-180.0, 180.0, 24, 0.7, 1.3, 3, 0.1)
    is_upright = [0] * 10
    is_upright[4] = is_upright[5] = 1
    joints = [synth_joints(60), full_cov_joints(61, every=2)]
    model = person_model(ep, joints[0], is_upright=is_upright)
    sizes = [(152, 176), (136, 144)]
    us = [synth_unary(70 + i, 10, 3, 24, h, w) for i, (h, w) in enumerate(sizes)]
    outs, (levels, used) = run_method(model, us, joints, method)
    assert used == levels
    for i in range(2):
        worst, map_bad = errors_vs_f64(model, us[i], joints[i], outs[i])
        say("%s uneven %d: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (method, i, worst, map_bad))
        assert worst <= METHOD_L1[method]
        assert not map_bad, map_bad


@pytest.mark.parametrize("method", ["bf16mix", "bf16x2", "bf16"])
def test_fp16_intermediate_uneven_widths_and_rotation_copy(gpu, method):
    # the fp16 row-pass output with widths 72, 232 and 104 in one launch (waves
    # past a narrower job's width idle, odd tile counts), border masks, and
    # joints without rotation spread (rot_sigma = 0: the rotation GEMM's
    # identity table) next to ordinary ones
    from partapp_amd.api import Joint
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.1)
    is_upright = [0] * 10
    is_upright[4] = is_upright[5] = 1
    base = [synth_joints(80), synth_joints(81), synth_joints(82)]
    joints = []
    for k, js in enumerate(base):
        joints.append([Joint(j.child_idx, j.parent_idx, j.offset_c, j.offset_p, j.C, j.rot_mean,
                             0.0 if (i + k) % 4 == 0 else j.rot_sigma) for i, j in enumerate(js)])
    model = person_model(ep, joints[0], is_upright=is_upright)
    sizes = [(96, 72), (120, 232), (88, 104)]
    us = [synth_unary(90 + i, 10, 1, 24, h, w) for i, (h, w) in enumerate(sizes)]
    say("fp16 intermediate", method, sizes)
    outs, (levels, used) = run_method(model, us, joints, method)
    assert used == levels
    for i in range(len(sizes)):
        worst, map_bad = errors_vs_f64(model, us[i], joints[i], outs[i])
        say("  %s instance %d %s: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (method, i, sizes[i], worst,
                                                                                   map_bad))
        assert worst <= METHOD_L1[method]
        assert not map_bad, map_bad


# parts without a detector (is_detect = 0): a leaf whose source is the zero
# grid (row pass), a chain node's belief without its unary (up), a leaf
# marginal without one (down).  "mixed" puts them next to the ordinary kinds
# in one launch; with every chain node undetected the up level's launch holds
# only the kinds without a unary read.
DETECT_CASES = {"mixed": (7, 9), "no_mid_unaries": (1, 2, 7, 8)}


@pytest.mark.parametrize("method", ["bf16mix", "bf16x3"])
@pytest.mark.parametrize("case", sorted(DETECT_CASES))
def test_parts_without_detector(gpu, case, method):
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    is_detect = [1] * 10
    for p in DETECT_CASES[case]:
        is_detect[p] = 0
    joints = [synth_joints(100), synth_joints(101)]
    model = person_model(ep, joints[0], is_detect=is_detect)
    sizes = [(104, 120), (96, 136)]
    us = [synth_unary(110 + i, 10, 1, 24, h, w) for i, (h, w) in enumerate(sizes)]
    say("parts without detector", case, method, is_detect)
    outs, (levels, used) = run_method(model, us, joints, method)
    assert used == levels
    for i in range(len(sizes)):
        worst, map_bad = errors_vs_f64(model, us[i], joints[i], outs[i])
        say("  %s %s instance %d: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (case, method, i, worst, map_bad))
        assert worst <= METHOD_L1[method]
        assert not map_bad, map_bad


def test_fp16_widest_and_tallest_grids(gpu):
    # the extremes of the uneven batch: 480 columns (8-wave row-pass blocks,
    # 8 column strips) next to 480 rows, in one launch of each level
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = [synth_joints(120), synth_joints(121)]
    model = person_model(ep, joints[0])
    sizes = [(160, 480), (480, 168)]
    us = [synth_unary(130 + i, 10, 1, 24, h, w) for i, (h, w) in enumerate(sizes)]
    say("fp16 extremes", sizes)
    outs, (levels, used) = run_method(model, us, joints, "bf16mix")
    assert used == levels
    for i in range(len(sizes)):
        worst, map_bad = errors_vs_f64(model, us[i], joints[i], outs[i])
        say("  instance %d %s: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (i, sizes[i], worst, map_bad))
        assert worst <= METHOD_L1["bf16mix"]
        assert not map_bad, map_bad


@pytest.mark.parametrize("method", ["bf16mix", "bf16x3"])
def test_tiny_grids(gpu, method):
    # grids narrower and shorter than one MFMA tile (down to 1 x 1) in one
    # batch: every output finite, every MAP inside its grid, and the
    # non-degenerate 20 x 24 instance within the budget.  On the smaller ones
    # the joints shift every message out of the grid, so the normalised FP64
    # marginals are no reference there (the FP32 path differs from FP64 by
    # ~0.1 as well).
    ep = ExpParam(-180.0, 180.0, 24, 0.3, 0.3, 1, 0.0)
    sizes = [(1, 1), (3, 2), (9, 11), (20, 24)]
    joints = [synth_joints(40 + i) for i in range(len(sizes))]
    model = person_model(ep, joints[0])
    us = [synth_unary(50 + i, 10, 1, 24, h, w) for i, (h, w) in enumerate(sizes)]
    # unguarded: the degenerate instances have near ties everywhere, and the
    # guard's FP32 re-evaluation reproduces the reference's float32 overflow
    # there (+inf marginals on the 9 x 11 grid, as the float32 oracle)
    outs, (levels, used) = run_method(model, us, joints, method, guard=False)
    assert used == levels
    for i, (h, w) in enumerate(sizes):
        marg, root, maps = outs[i]
        assert np.isfinite(marg).all() and np.isfinite(root).all(), sizes[i]
        for p in range(10):
            m = maps[0, p]
            assert 0 <= int(m["rotidx"]) < 24 and 0 <= int(m["y"]) < h and 0 <= int(m["x"]) < w
    worst, map_bad = errors_vs_f64(model, us[3], joints[3], outs[3])
    say("%s tiny grids: 20x24 worst rel L1 vs f64 %.3g, MAP mismatches %s" % (method, worst, map_bad))
    assert worst <= BUDGET_L1 and not map_bad
    # guarded, the 20 x 24 instance keeps its MAPs and the 9 x 11 one (ties
    # everywhere) takes the float32 path's outputs
    outs_g, _ = run_method(model, us, joints, method, guard=True)
    outs_32, _ = run_method(model, us, joints, "fp32")
    assert np.array_equal(outs_g[2][0], outs_32[2][0], equal_nan=True)
    worst, map_bad = errors_vs_f64(model, us[3], joints[3], outs_g[3])
    assert worst <= BUDGET_L1 and not map_bad


def test_bf16_graph_equals_eager(gpu):
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(3)
    model = person_model(ep, joints)
    u = synth_unary(3, 10, 1, 24, 96, 104)
    eager, _ = run_method(model, [u], [joints], "bf16x3")
    graph, _ = run_method(model, [u], [joints], "bf16x3", graph=True)
    np.testing.assert_array_equal(eager[0][0], graph[0][0])
    np.testing.assert_array_equal(eager[0][1], graph[0][1])


def test_method_fallback_for_wide_filters(gpu):
    # a joint with sigma 40 px (half width 120) does not fit the BF16 passes:
    # its levels run FP32 and the result still matches the FP32 oracle
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(5)
    from partapp_amd.api import Joint
    j = joints[0]
    joints[0] = Joint(j.child_idx, j.parent_idx, j.offset_c, j.offset_p, (1600.0, 0, 0, 25.0), j.rot_mean, j.rot_sigma)
    model = person_model(ep, joints)
    u = synth_unary(5, 10, 1, 24, 112, 128)
    outs, (levels, used) = run_method(model, [u], [joints], "bf16x3")
    assert used < levels
    worst, map_bad = errors_vs_f64(model, u, joints, outs[0])
    assert worst <= BUDGET_L1 and not map_bad


def test_width_limit_of_the_mfma_passes(gpu):
    # 512 columns (8 full waves per row-pass block) is the widest grid the
    # MFMA passes take; a batch with a wider grid runs every level in FP32,
    # with the FP32 method's marginals bit for bit (the root posterior's
    # log-sum-exp stays the BF16 methods' shifted FP64 one: the same values
    # where the FP32 sum of expf does not underflow to LOG_ZERO, below the
    # underflow bound where it does)
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(140)
    model = person_model(ep, joints)
    u512 = synth_unary(141, 10, 1, 24, 96, 512)
    outs, (levels, used) = run_method(model, [u512], [joints], "bf16mix")
    assert used == levels == 4
    worst, map_bad = errors_vs_f64(model, u512, joints, outs[0])
    say("512 wide: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (worst, map_bad))
    assert worst <= METHOD_L1["bf16mix"]
    assert not map_bad, map_bad
    u520 = synth_unary(142, 10, 1, 24, 96, 520)
    outs_mix, (levels, used) = run_method(model, [u520], [joints], "bf16mix")
    assert used == 0 and levels == 4
    outs_fp32, _ = run_method(model, [u520], [joints], "fp32")
    assert np.array_equal(outs_mix[0][0], outs_fp32[0][0])
    r_mix, r_32 = outs_mix[0][1], outs_fp32[0][1]
    in_range = r_32 > -1e5
    np.testing.assert_allclose(r_mix[in_range], r_32[in_range], rtol=1e-5, atol=1e-5)
    assert (r_mix[~in_range] < -80.0).all()


def test_other_rotation_counts_run_fp32(gpu):
    # the MFMA passes hold 24 rotation slices per 32-row tile; with 16
    # rotations every level runs the FP32 passes, bit for bit
    ep = ExpParam(-180.0, 180.0, 16, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(150)
    model = person_model(ep, joints)
    u = synth_unary(151, 10, 1, 16, 72, 88)
    outs_mix, (levels, used) = run_method(model, [u], [joints], "bf16mix")
    assert used == 0 and levels == 4
    outs_fp32, _ = run_method(model, [u], [joints], "fp32")
    assert np.array_equal(outs_mix[0][0], outs_fp32[0][0])
    worst, map_bad = errors_vs_f64(model, u, joints, outs_mix[0])
    assert worst <= 1e-5 and not map_bad


def test_rows_per_block_knob_does_not_change_results(gpu):
    import torch
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(8)
    model = person_model(ep, joints)
    u = synth_unary(8, 10, 1, 24, 100, 116)  # H = 100: the last block of 3 / 16 rows is partial
    plan = Plan(model, [(100, 116)], [joints])
    plan.set_method("bf16mix")
    un = torch.from_numpy(u.ravel().copy()).to(gpu)
    outs = []
    for rows in (1, 3, 4, 16):
        plan.tune(plan.TUNE_MFMA_ROWS, rows)
        m = torch.full_like(un, float("nan"))
        r = torch.full((plan.root_numel,), float("nan"), device=gpu)
        maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=gpu)
        plan.execute(un, m, r, maps)
        torch.cuda.synchronize()
        outs.append((m.cpu(), r.cpu(), maps.cpu()))
    from partapp_amd._lib import PsError
    with pytest.raises(PsError):
        plan.tune(plan.TUNE_MFMA_ROWS, 0)
    plan.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert torch.equal(a, b)


@pytest.mark.parametrize("method", ["bf16mix", "bf16x3"])
def test_default_rotation_grid_48_on_the_matrix_cores(gpu, method):
    # the reference's default 48 rotations (ExpParam.proto:195): two 32-slice
    # tiles per x-filter GEMM and a 3-step rotation GEMM into two output
    # tiles; every level on the matrix cores, two widths, three scales
    ep = ExpParam(-180.0, 180.0, 48, 0.7, 1.3, 3, 0.0)
    joints = [synth_joints(160), synth_joints(161)]
    model = person_model(ep, joints[0])
    sizes = [(112, 128), (96, 200)]
    us = [synth_unary(162 + i, 10, 3, 48, h, w) for i, (h, w) in enumerate(sizes)]
    say("48 rotations", method, sizes)
    outs, (levels, used) = run_method(model, us, joints, method)
    assert used == levels == 4
    for i in range(len(sizes)):
        worst, map_bad = errors_vs_f64(model, us[i], joints[i], outs[i])
        say("  instance %d %s: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (i, sizes[i], worst, map_bad))
        assert worst <= METHOD_L1[method]
        assert not map_bad, map_bad
