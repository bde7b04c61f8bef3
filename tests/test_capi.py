"""CPU tests of libpartapp_ps.so: it loads, exports every symbol of
include/partapp_ps.h, validates models like the reference's asserts, and its
host-side message plan equals the oracle's pixel-by-pixel evaluation of the
reference's coordinate arithmetic."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle
from partapp_amd import _lib
from partapp_amd.api import ExpParam, Joint, PartModel, config_exp_param, person_model, synth_joints, synth_unary

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"^[A-Za-z_][\w \*]*?\b(ps_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, "no ctypes signature for " + n
    assert set(_lib.SIGNATURES) == set(names)
    assert b"gfx950" in L.ps_version()


def _model(**kw):
    ep = kw.pop("exp_param", ExpParam(num_rotation_steps=24))
    return person_model(ep, **kw)


def test_model_create_person_tree():
    m = _model()
    assert m.root == 4


def _status(fn):
    with pytest.raises(_lib.PsError) as e:
        fn()
    return e.value.status


def test_model_validation_mirrors_the_asserts():
    joints = synth_joints(0)
    ep = ExpParam(num_rotation_steps=24)
    P = 10
    det, up = [1] * P, [0] * P
    root = [1 if p == 4 else 0 for p in range(P)]
    # no root (objectdetect_findrot.cpp:781-789)
    assert _status(lambda: PartModel(det, up, [0] * P, joints, ep)) == -2
    # two roots
    assert _status(lambda: PartModel(det, up, [1 if p in (4, 5) else 0 for p in range(P)], joints, ep)) == -2
    # joints != P - 1 (objectdetect_aux.cpp:131)
    assert _status(lambda: PartModel(det, up, root, joints[:-1], ep)) == -2
    # a non-root part with two children (objectdetect_findrot.cpp:210)
    bad = list(joints)
    bad[0] = Joint(0, 2, bad[0].offset_c, bad[0].offset_p, bad[0].C)  # 0 -> 2 while 3 -> 2 exists
    assert _status(lambda: PartModel(det, up, root, bad, ep)) == -2
    # a cycle / disconnected part
    bad = list(joints)
    bad[0] = Joint(0, 3, bad[0].offset_c, bad[0].offset_p, bad[0].C)
    bad[3] = Joint(3, 0, bad[3].offset_c, bad[3].offset_p, bad[3].C)
    assert _status(lambda: PartModel(det, up, root, bad, ep)) == -2
    # R above the 64-bit rotation masks
    assert _status(lambda: PartModel(det, up, root, joints, ExpParam(num_rotation_steps=65))) == -5
    # strip_border_detections >= 0.5 (objectdetect_findrot.cpp:528)
    assert _status(lambda: PartModel(det, up, root, joints, ExpParam(num_rotation_steps=24,
                                                                     strip_border_detections=0.5))) == -1


def _debug_plan(ep, H, W, oc, op, C, rot_mean, rot_sigma, scale):
    R = ep.num_rotation_steps
    ints = [(ctypes.c_int * R)() for _ in range(4)]
    delta, rot_len, kx, ky = (ctypes.c_int() for _ in range(4))
    taps = (ctypes.c_float * 1024)()
    st = _lib.lib().ps_debug_message_plan(
        ctypes.byref(ep.to_c()), H, W, (ctypes.c_double * 2)(*oc), (ctypes.c_double * 2)(*op),
        (ctypes.c_double * 4)(*C), rot_mean, rot_sigma, scale, ctypes.byref(delta), ints[0], ints[1], ints[2],
        ints[3], ctypes.byref(rot_len), taps, ctypes.byref(kx), ctypes.byref(ky))
    return st, dict(delta=delta.value, fwd_ox=list(ints[0]), fwd_oy=list(ints[1]), inv_ox=list(ints[2]),
                    inv_oy=list(ints[3]), rot_taps=list(taps[:rot_len.value]), kx=kx.value, ky=ky.value)


@pytest.mark.parametrize("seed", range(12))
def test_plan_matches_oracle_pixel_evaluation(seed):
    rng = np.random.default_rng(100 + seed)
    R = int(rng.choice([4, 5, 6, 24, 48]))
    S = 3
    ep = ExpParam(-180.0, 180.0, R, 0.7, 1.3, S, 0.0)
    H, W = int(rng.integers(8, 200)), int(rng.integers(8, 200))
    oc, op = rng.uniform(-40, 40, 2), rng.uniform(-40, 40, 2)
    sx, sy = rng.uniform(4, 12, 2)
    rot_mean = float(rng.uniform(-math.pi / 6, math.pi / 6))
    rot_sigma = float(rng.uniform(0.1, 0.4))
    for s in range(S):
        scale = ep.scale_from_index(s)
        st, mine = _debug_plan(ep, H, W, oc, op, [sx * sx, 0, 0, sy * sy], rot_mean, rot_sigma, scale)
        assert st == 0
        ref = oracle.message_plan(oracle.exp_param_from(ep), R, H, W, oc, op, rot_mean, rot_sigma, scale)
        assert ref["uniform"]
        for k in ("delta", "fwd_ox", "fwd_oy", "inv_ox", "inv_oy"):
            assert mine[k] == ref[k], k
        assert mine["rot_taps"] == [float(np.float32(t)) for t in ref["rot_taps"]]
        # spatial half widths: floor(3 * scale * sigma + 0.5)
        assert mine["kx"] == math.floor(3 * math.sqrt(scale * scale * sx * sx) + 0.5)
        assert mine["ky"] == math.floor(3 * math.sqrt(scale * scale * sy * sy) + 0.5)


def test_plan_error_codes():
    ep = ExpParam(num_rotation_steps=24)
    assert _debug_plan(ep, 10, 10, [0, 0], [0, 0], [0.0, 0, 0, 1.0], 0, 0.1, 1.0)[0] == -3   # C00 <= 0
    # off-diagonal C: the filter runs in its eigenbasis, which must be positive definite
    assert _debug_plan(ep, 10, 10, [0, 0], [0, 0], [1.0, 2.0, 2.0, 1.0], 0, 0.1, 1.0)[0] == -3
    st, p = _debug_plan(ep, 10, 10, [0, 0], [0, 0], [5.0, 2.0, 2.0, 3.0], 0, 0.1, 1.5)
    e_small, e_large = np.linalg.eigvalsh(np.array([[5.0, 2.0], [2.0, 3.0]]) * 1.5 ** 2)
    assert st == 0
    assert p["kx"] == math.floor(3 * math.sqrt(e_small) + 0.5) and p["ky"] == math.floor(3 * math.sqrt(e_large) + 0.5)
    assert _debug_plan(ep, 10, 10, [0, 0], [0, 0], [400.0 ** 2, 0, 0, 1.0], 0, 0.1, 1.0)[0] == -4  # >= 1000 taps
    assert _debug_plan(ExpParam(num_rotation_steps=1), 10, 10, [0, 0], [0, 0], [1.0, 0, 0, 1.0], 0, 0.1,
                       1.0)[0] == -4  # R = 1 cannot hold a clipped rotation kernel
    # rot_sigma == 0 copies, < 0 zeroes (objectdetect_findrot.cpp:377-420)
    st, p = _debug_plan(ep, 10, 10, [0, 0], [0, 0], [1.0, 0, 0, 1.0], 0, 0.0, 1.0)
    assert st == 0 and p["rot_taps"] == []


@pytest.mark.parametrize("oc,op", [((-3.5302013092419844, 0.0), (0.0, 0.0)), ((0.0, -3.5302013092419844), (0.0, 0.0)),
                                   ((0.0, 0.0), (-3.5302013092419813, 0.0)), ((0.0, 0.0), (0.0, -3.5302013092419813))])
def test_plan_takes_warps_whose_offset_changes_once(oc, op):
    # offsets within an ulp of a half-integer at slice 0: the oracle's per-pixel
    # warp changes its offset along one axis; the plan takes it (the offsets at
    # index 0 are the oracle's; the kernels take the break, GPU-tested)
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    ref = oracle.message_plan(oracle.exp_param_from(ep), 24, 41, 57, oc, op, 0.0, 0.25, 1.0)
    assert not ref["uniform"]
    st, mine = _debug_plan(ep, 41, 57, oc, op, [30.0, 0, 0, 40.0], 0.0, 0.25, 1.0)
    assert st == 0
    for k in ("delta", "fwd_ox", "fwd_oy", "inv_ox", "inv_oy"):
        assert mine[k] == ref[k], k


def test_plan_create_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    m = _model()
    from partapp_amd.api import Plan
    assert _status(lambda: Plan(m, [(20, 20)])) == -8


def test_synth_is_deterministic_and_shaped():
    a = synth_unary(3, 2, 1, 24, 30, 40, threads=4)
    b = synth_unary(3, 2, 1, 24, 30, 40, threads=1)
    np.testing.assert_array_equal(a, b)
    c = synth_unary(4, 2, 1, 24, 30, 40)
    assert not np.array_equal(a, c)
    assert a.dtype == np.float32 and a.shape == (2, 1, 24, 30, 40)
    big = synth_unary(5, 1, 1, 24, 200, 200)
    bg = big[big < -1.0]  # background N(-3, 0.5) away from the bumps
    assert abs(bg.mean() + 3.0) < 0.01 and abs(bg.std() - 0.5) < 0.01
    # three bumps of amplitude 4 per part, each in one rotation bin
    assert big.max() > 0.0 and (big > -0.5).sum() < 3 * 400


def test_synth_joints_ranges():
    for seed in range(5):
        js = synth_joints(seed)
        assert [(j.child_idx, j.parent_idx) for j in js] == [(0, 1), (1, 4), (2, 4), (3, 2), (5, 4), (6, 7),
                                                              (7, 4), (8, 4), (9, 8)]
        for j in js:
            assert all(-40 <= v <= 40 for v in j.offset_c + j.offset_p)
            assert 16 <= j.C[0] <= 144 and 16 <= j.C[3] <= 144 and j.C[1] == j.C[2] == 0
            assert 0.1 <= j.rot_sigma <= 0.4 and abs(j.rot_mean) <= math.pi / 6
    assert all(j.rot_mean == 0 for j in synth_joints(0, zero_rot_mean=True))
    assert synth_joints(0)[0].offset_c == synth_joints(0, zero_rot_mean=True)[0].offset_c


def test_config_exp_params():
    ep = config_exp_param(2)
    assert [ep.scale_from_index(s) for s in range(3)] == pytest.approx([0.8, 1.0, 1.2], abs=1e-6)
    assert config_exp_param(1).num_rotation_steps == 24
