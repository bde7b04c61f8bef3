"""Parity of the gfx950 sweep with the oracle (the CPU restatement of the
reference), on the same seeded inputs.

Bar (the FP32 path's tolerance): per part, the relative L1 distance of the
normalised marginals p = exp(b - LSE(b)) (computed in float64 from each side's
float32 log marginals) is <= 1e-5 against the float32 oracle, and the MAP
(rotation, y, x) of every part and scale is identical.  The distance to the
float64 oracle is reported alongside.
"""
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import (ExpParam, Joint, PartModel, Plan, compute_root_posterior_rot, compute_rot_joint_marginal,
                             config_exp_param, decode_maps, person_model, synth_joints, synth_unary)

pytestmark = pytest.mark.gpu

TOL_L1 = 1e-5


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def torch_mod():
    import torch
    return torch


def run_plan(model, unaries, joints=None, graph=False):
    """unaries: list of numpy [P][S][R][H][W]; returns per-instance marginals,
    root posteriors and MAP records (numpy)."""
    torch = torch_mod()
    sizes = [u.shape[3:] for u in unaries]
    plan = Plan(model, sizes, joints)
    dev = torch.device("cuda:0")
    flat = np.concatenate([u.ravel() for u in unaries])
    un = torch.from_numpy(flat).to(dev)
    marg = torch.full_like(un, float("nan"))
    root = torch.full((plan.root_numel,), float("nan"), device=dev)
    maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream() if graph else None
    if graph:
        with torch.cuda.stream(stream):
            plan.execute(un, marg, root, maps, stream=stream, graph=True)
        stream.synchronize()
    else:
        plan.execute(un, marg, root, maps)
        torch.cuda.synchronize()
    m = decode_maps(maps, plan.n_instances, plan.S, plan.P)
    out = [(plan.instance_view(marg, i).cpu().numpy(), plan.instance_view(root, i, "root").cpu().numpy(), m[i])
           for i in range(plan.n_instances)]
    plan.close()
    return out


def oracle_run(model, u, joints, dtype=np.float32):
    ep = oracle.exp_param_from(model.exp_param)
    return oracle.root_posterior_rot(ep, model.is_detect, model.is_upright, model.root, joints, u.astype(dtype),
                                     is_sparse=model.is_sparse)


def check_parity(name, model, u, joints, gpu_out, check_f64=True):
    marg, root, maps = gpu_out
    P, S = u.shape[0], u.shape[1]
    r32, p32, b32, s32 = oracle_run(model, u, joints)
    worst = 0.0
    for s in range(S):
        for p in range(P):
            d = oracle.rel_l1(oracle.normalized(marg[p, s]), oracle.normalized(p32[p, s]))
            worst = max(worst, d)
            assert d <= TOL_L1, "%s: part %d scale %d rel L1 %.3g" % (name, p, s, d)
            got = (int(maps[s, p]["rotidx"]), int(maps[s, p]["y"]), int(maps[s, p]["x"]))
            assert got == tuple(b32[s, p]), "%s: MAP of part %d scale %d: %s vs %s (top-2 gap %.3g)" % (
                name, p, s, got, tuple(b32[s, p]), oracle.top2_gap(p32[p, s]))
            assert np.float32(maps[s, p]["score"]) == np.float32(marg[p, s].ravel().max())
    # root posterior: same log-sum-exp over the valid root rotations
    finite = root > -1e5
    np.testing.assert_array_equal(finite, r32 > -1e5)
    np.testing.assert_allclose(root[finite], r32[finite], rtol=2e-5, atol=2e-5)
    msg = "%s: worst rel L1 vs f32 oracle %.3g" % (name, worst)
    if check_f64:
        r64, p64, b64, s64 = oracle_run(model, u, joints, np.float64)
        w64 = max(oracle.rel_l1(oracle.normalized(marg[p, s]), oracle.normalized(p64[p, s]))
                  for s in range(S) for p in range(P))
        msg += ", vs f64 oracle %.3g" % w64
    say(msg)
    return worst


# ----------------------------------------------------------------- messages

MESSAGE_CASES = {
    "generic": dict(offset_c=(7.3, -11.2), offset_p=(-3.9, 8.8), C=(30.0, 0, 0, 50.0), rot_mean=0.3,
                    rot_sigma=0.25, scale=1.0),
    "copy_rot": dict(offset_c=(0.0, 0.0), offset_p=(2.0, -1.0), C=(16.0, 0, 0, 9.0), rot_mean=0.0, rot_sigma=0.0,
                     scale=1.0),
    "zero_rot": dict(offset_c=(1.0, 1.0), offset_p=(0.0, 0.0), C=(16.0, 0, 0, 9.0), rot_mean=0.0, rot_sigma=-1.0,
                     scale=1.0),
    "big_offsets_scaled": dict(offset_c=(39.0, -38.0), offset_p=(-40.0, 37.5), C=(144.0, 0, 0, 16.0),
                               rot_mean=-0.5, rot_sigma=0.4, scale=1.2),
    "clipped_rot_kernel": dict(offset_c=(3.0, 4.0), offset_p=(-2.0, 1.0), C=(20.0, 0, 0, 20.0), rot_mean=0.1,
                               rot_sigma=3.0, scale=0.8),
}
# Offsets whose rotated value at slice 0 lies within an ulp of a half-integer:
# the nearest-neighbour warp's offset changes once along the axis (the double
# sum x + offset rounds differently in different binades), forward (offset_c)
# or inverse (offset_p), along x or along y (at 41 x 57, found by the oracle's
# per-pixel evaluation)
SPLIT_C = -3.5302013092419844
SPLIT_P = -3.5302013092419813
for _name, _oc, _op in (("split_fwd_x", (SPLIT_C, 0.0), (0.0, 0.0)), ("split_fwd_y", (0.0, SPLIT_C), (0.0, 0.0)),
                        ("split_inv_x", (0.0, 0.0), (SPLIT_P, 0.0)), ("split_inv_y", (0.0, 0.0), (0.0, SPLIT_P))):
    MESSAGE_CASES[_name] = dict(offset_c=_oc, offset_p=_op, C=(30.0, 0, 0, 40.0), rot_mean=0.0, rot_sigma=0.25,
                                scale=1.0)


@pytest.mark.parametrize("case", sorted(MESSAGE_CASES))
def test_message_matches_oracle(gpu, case):
    torch = torch_mod()
    kw = MESSAGE_CASES[case]
    R, H, W = 24, 41, 57
    ep = ExpParam(-180.0, 180.0, R, 1.0, 1.0, 1, 0.0)
    plan = oracle.message_plan(oracle.exp_param_from(ep), R, H, W, kw["offset_c"], kw["offset_p"], kw["rot_mean"],
                               kw["rot_sigma"], kw["scale"])
    assert plan["uniform"] == (not case.startswith("split_"))
    model = person_model(ep)
    rng = np.random.default_rng(sum(map(ord, case)))
    child = rng.normal(-3, 1.0, (R, H, W)).astype(np.float32)
    say("message case", case, kw)
    out = compute_rot_joint_marginal(model, torch.from_numpy(child).to(gpu), kw["offset_c"], kw["offset_p"], kw["C"],
                                     kw["rot_mean"], kw["rot_sigma"], kw["scale"]).cpu().numpy()
    ref = oracle.rot_joint_marginal(oracle.exp_param_from(ep), child, **kw)
    # the LOG_ZERO sentinels must sit at the same places
    np.testing.assert_array_equal(out <= -9e5, ref <= -9e5)
    live = ref > -9e5
    if live.any():
        np.testing.assert_allclose(out[live], ref[live], rtol=1e-5, atol=5e-5)
        d = oracle.rel_l1(oracle.normalized(out), oracle.normalized(ref))
        say("  rel L1 %.3g, exact %.4f" % (d, np.mean(out == ref)))
        assert d <= 1e-6
    else:
        np.testing.assert_array_equal(out, ref)


# ------------------------------------------------------------------- sweeps

# Grid sizes: the synthetic joints move a part by up to 40*sqrt(2)*scale px per
# warp, so grids of at least ~2x that keep every message away from the
# degenerate case where a warp leaves no pixel inside the grid (then only
# LOG_ZERO sentinels remain and float32 sums of them overflow, on both sides).

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sweep_small_matches_oracle(gpu, seed):
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(seed)
    model = person_model(ep, joints)
    u = synth_unary(seed, 10, 1, 24, 112, 128)
    say("sweep small seed", seed)
    out = run_plan(model, [u])[0]
    check_parity("small seed %d" % seed, model, u, joints, out)


@pytest.mark.parametrize("method", ["fp32", "bf16mix"])
def test_sweep_with_split_warps(gpu, method):
    # two joints whose warps change their offset once (forward along x, inverse
    # along y): the FP32 passes take their breaks; the BF16 methods run those
    # levels on the FP32 passes
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(5)
    j0, j5 = joints[0], joints[5]
    joints[0] = Joint(j0.child_idx, j0.parent_idx, (SPLIT_C, 0.0), j0.offset_p, j0.C, j0.rot_mean, j0.rot_sigma)
    joints[5] = Joint(j5.child_idx, j5.parent_idx, j5.offset_c, (0.0, SPLIT_P), j5.C, j5.rot_mean, j5.rot_sigma)
    model = person_model(ep, joints)
    u = synth_unary(5, 10, 1, 24, 96, 112)
    say("sweep with split warps", method)
    plan = Plan(model, [(96, 112)], [joints])
    levels, used = plan.set_method(method)
    plan.close()
    if method == "fp32":
        out = run_plan(model, [u])[0]
        check_parity("split warps", model, u, joints, out)
    else:
        assert used < levels  # the levels of the two split messages run in FP32
        from test_gpu_bf16 import METHOD_L1, errors_vs_f64, run_method
        outs, _ = run_method(model, [u], [joints], method)
        worst, map_bad = errors_vs_f64(model, u, joints, outs[0])
        say("  bf16mix: worst rel L1 vs f64 %.3g, MAP mismatches %s" % (worst, map_bad))
        assert worst <= METHOD_L1[method] and not map_bad


# trees other than the person tree, by parent of each part (-1 = root): the
# root's child count selects the root kernel's instantiation (1..8 unrolled,
# above 8 the count read at run time)
TREES = {
    "chain3": [-1, 0, 1],                      # one chain of two below the root
    "star3": [1, -1, 1, 1],                    # root = part 1, three leaves
    "mixed4": [-1, 0, 1, 0, 0, 4, 5, 0],       # chains of length 2, 1, 3, 1
    "star9": [-1, 0, 0, 0, 0, 0, 0, 0, 0, 0],  # nine leaves: the run-time count
}


@pytest.mark.parametrize("tree", sorted(TREES))
def test_root_child_counts(gpu, tree):
    parents = TREES[tree]
    P = len(parents)
    root = parents.index(-1)
    base = synth_joints(11) + synth_joints(12)
    joints = []
    for c, par in enumerate(parents):
        if par >= 0:
            j = base[len(joints)]
            joints.append(Joint(c, par, j.offset_c, j.offset_p, j.C, j.rot_mean, j.rot_sigma))
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    model = PartModel([1] * P, [0] * P, [1 if p == root else 0 for p in range(P)], joints, ep, True)
    u = synth_unary(13, P, 1, 24, 112, 128)
    say("tree", tree, parents)
    out = run_plan(model, [u])[0]
    check_parity("tree %s" % tree, model, u, joints, out, check_f64=False)


def test_sweep_cfg0_full_size(gpu):
    # one instance, 200 x 200 x 24 rotations, 10 parts, seed 0
    ep = config_exp_param(0)
    joints = synth_joints(0)
    model = person_model(ep, joints)
    u = synth_unary(0, 10, 1, 24, 200, 200)
    say("cfg0 full size")
    out = run_plan(model, [u])[0]
    check_parity("cfg0", model, u, joints, out)


def test_sweep_zero_rot_mean_equal_partition_functions(gpu):
    # with delta = 0 on every edge every part's unnormalised marginal has the
    # same log partition function.  The grid is large enough that no joint
    # offset (up to 40 px per axis, twice per edge) moves a grid entirely out
    # of bounds, where only LOG_ZERO sentinels would be left to sum.
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(7, zero_rot_mean=True)
    model = person_model(ep, joints)
    u = synth_unary(7, 10, 1, 24, 120, 136)
    marg, root, maps = run_plan(model, [u])[0]
    lse = []
    for p in range(10):
        b = marg[p, 0].astype(np.float64)
        lse.append(b.max() + np.log(np.exp(b - b.max()).sum()))
    assert max(lse) - min(lse) < 1e-3 * max(1.0, abs(lse[0])), lse
    check_parity("zero rot_mean", model, u, joints, (marg, root, maps), check_f64=False)


def test_uneven_batch_scales_masks_flags(gpu):
    # three instances of different sizes, three scales, per-instance joints,
    # upright parts (including the root), border strip, non-detect leaf and mid part
    ep = ExpParam(-180.0, 180.0, 24, 0.7, 1.3, 3, 0.1)
    is_detect = [1] * 10
    is_detect[9] = 0   # a leaf
    is_detect[7] = 0   # a chain node
    is_upright = [0] * 10
    is_upright[4] = is_upright[5] = is_upright[2] = 1
    joints = [synth_joints(20 + i) for i in range(3)]
    model = person_model(ep, joints[0], is_detect=is_detect, is_upright=is_upright)
    sizes = [(144, 160), (136, 176), (152, 144)]
    us = [synth_unary(30 + i, 10, 3, 24, h, w) for i, (h, w) in enumerate(sizes)]
    say("uneven batch", sizes)
    outs = run_plan(model, us, joints)
    for i in range(3):
        check_parity("uneven %d" % i, model, us[i], joints[i], outs[i], check_f64=(i == 0))


def test_graph_replays_match_eager(gpu):
    torch = torch_mod()
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    model = person_model(ep, synth_joints(3))
    u = synth_unary(3, 10, 1, 24, 96, 104)
    plan = Plan(model, [(96, 104)])
    dev = torch.device("cuda:0")
    un = torch.from_numpy(u.ravel().copy()).to(dev)
    ref_m = torch.empty_like(un)
    ref_r = torch.empty((plan.root_numel,), device=dev)
    ref_maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=dev)
    plan.execute(un, ref_m, ref_r, ref_maps)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    m = torch.empty_like(un)
    r = torch.empty_like(ref_r)
    maps = torch.zeros_like(ref_maps)
    for rep in range(4):
        m.fill_(float("nan"))
        r.fill_(float("nan"))
        maps.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.execute(un, m, r, maps, stream=s, graph=True)
        s.synchronize()
        _ = (un * 2).sum().item()  # eager work between replays
        assert torch.equal(m, ref_m), rep
        assert torch.equal(r, ref_r), rep
        assert torch.equal(maps, ref_maps), rep
    plan.close()


def test_reference_shaped_api_last_scale_map(gpu):
    torch = torch_mod()
    ep = ExpParam(-180.0, 180.0, 24, 0.7, 1.3, 3, 0.0)
    joints = synth_joints(5)
    model = person_model(ep, joints)
    u = synth_unary(5, 10, 3, 24, 144, 152)
    root, marg, best, maps = compute_root_posterior_rot(model, torch.from_numpy(u).to(gpu))
    r32, p32, b32, s32 = oracle_run(model, u, joints)
    # best_part_hyp keeps the last scale (objectdetect_findrot.cpp:257-259 inside :503)
    for p in range(10):
        assert (best[p].scaleidx, best[p].rotidx, best[p].y, best[p].x) == (2,) + tuple(b32[2, p])
        assert best[p].rot == pytest.approx(-172.5 + 15 * best[p].rotidx)
        assert best[p].scale == pytest.approx(1.2, abs=1e-6)
        assert best[p].to_vect().shape == (7,)
    for s in range(3):
        for p in range(10):
            assert oracle.rel_l1(oracle.normalized(marg[p, s].cpu().numpy()), oracle.normalized(p32[p, s])) <= TOL_L1


def test_constant_unaries_ties_take_the_first_index(gpu):
    # flat unaries make many exactly equal maxima; the MAP is the first in C order
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    base = synth_joints(9, zero_rot_mean=True)
    joints = [Joint(j.child_idx, j.parent_idx, (0.0, 0.0), (0.0, 0.0), (16.0, 0, 0, 16.0), 0.0, 0.0) for j in base]
    model = person_model(ep, joints)
    u = np.full((10, 1, 24, 30, 30), -2.0, dtype=np.float32)
    marg, root, maps = run_plan(model, [u])[0]
    r32, p32, b32, s32 = oracle_run(model, u, joints, np.float32)
    for p in range(10):
        got = (int(maps[0, p]["rotidx"]), int(maps[0, p]["y"]), int(maps[0, p]["x"]))
        flat = int(np.argmax(marg[p, 0].ravel()))  # numpy argmax is the first maximum
        assert got == (flat // 900, (flat % 900) // 30, flat % 30)
        assert got == tuple(b32[0, p])


def test_profiling_classes_bracket_only_their_launches(gpu):
    # per-kernel timing of every class, or of the column pass alone, leaves
    # the results unchanged; a mask beyond the six classes is an error
    import ctypes
    from partapp_amd._lib import PsError, check, lib
    torch = torch_mod()
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(4)
    model = person_model(ep, joints)
    u = synth_unary(4, 10, 1, 24, 64, 72)
    plan = Plan(model, [(64, 72)], [joints])
    dev = torch.device("cuda:0")
    un = torch.from_numpy(u.ravel()).to(dev)
    outs = []
    for classes in (None, Plan.KERNEL_CLASSES, ("colpass",)):
        marg = torch.full_like(un, float("nan"))
        root = torch.full((plan.root_numel,), float("nan"), device=dev)
        maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=dev)
        if classes is not None:
            plan.set_profiling(True, classes=classes)
            plan.read_profile()
        plan.execute(un, marg, root, maps)
        torch.cuda.synchronize()
        if classes is not None:
            prof = plan.read_profile()
            plan.set_profiling(False)
            for k, (ms, n) in prof.items():
                assert (n > 0) == (k in classes), (classes, k, n)
                assert ms >= 0.0
        outs.append(marg.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    # one column-pass launch per level and destination kind; the messages run
    # as late as the tree allows, so each of the four levels holds one kind
    assert prof["colpass"][1] == 4
    with pytest.raises(PsError):
        check(lib().ps_plan_set_profiling_classes(plan.handle, ctypes.c_uint(1 << len(Plan.KERNEL_CLASSES))),
              "ps_plan_set_profiling_classes")
    plan.close()


def test_plan_errors_on_gpu(gpu):
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    model = person_model(ep)
    bad = synth_joints(1)
    bad[2] = Joint(bad[2].child_idx, bad[2].parent_idx, bad[2].offset_c, bad[2].offset_p, (0.0, 0, 0, 1.0))
    from partapp_amd._lib import PsError
    with pytest.raises(PsError) as e:
        Plan(model, [(20, 20)], [bad])
    assert e.value.status == -3
    with pytest.raises(PsError):
        Plan(model, [(0, 20)])


def test_device_generator_matches_host(gpu):
    import torch
    from partapp_amd.api import synth_unary_dev
    u = synth_unary(13, 3, 2, 24, 37, 53)
    d = torch.empty(u.size, dtype=torch.float32, device=gpu)
    synth_unary_dev(13, 3, 2, 24, 37, 53, d)
    g = d.cpu().numpy().reshape(u.shape)
    # the device's double-precision log / sin / cos may differ in the last bit
    assert np.mean(g == u) > 0.999
    np.testing.assert_allclose(g, u, rtol=0, atol=1e-6)
