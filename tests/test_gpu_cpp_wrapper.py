"""The C++ interface (include/partapp_ps.hpp) called the way partapp's
findObjectImageRotJoints calls computeRootPosteriorRot, with stand-ins of
partapp's PartApp / Joint / FloatGrid3 / PartHyp types, checked against the
oracle: marginals, root posterior, best_part_hyp of the last scale, the
in-place masking of the caller's score grids, best_conf and the root
HypothesisList."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import ExpParam, person_model, synth_joints, synth_unary
from partapp_amd.build import CPP_TEST_BIN

pytestmark = pytest.mark.gpu


def write_input(path, ep, model, joints, u, root):
    P, S, R, H, W = u.shape
    with open(path, "wb") as f:
        np.array([P, S, R, H, W, root], dtype=np.int32).tofile(f)
        np.array([ep.min_part_rotation, ep.max_part_rotation, ep.min_object_scale, ep.max_object_scale,
                  ep.strip_border_detections], dtype=np.float32).tofile(f)
        np.array(model.is_detect + model.is_upright + model.is_root, dtype=np.int32).tofile(f)
        for j in joints:
            np.array([j.child_idx, j.parent_idx], dtype=np.int32).tofile(f)
            np.array(list(j.offset_c) + list(j.offset_p) + list(j.C) + [j.rot_mean, j.rot_sigma],
                     dtype=np.float64).tofile(f)
        u.astype(np.float32).tofile(f)


def test_cpp_wrapper_matches_oracle(gpu, tmp_path):
    if not os.path.exists(CPP_TEST_BIN):
        pytest.skip("C++ driver not built (python -m partapp_amd.build)")
    ep = ExpParam(-180.0, 180.0, 24, 0.7, 1.3, 2, 0.1)
    joints = synth_joints(41)
    is_upright = [0] * 10
    is_upright[3] = is_upright[4] = 1
    model = person_model(ep, joints, is_upright=is_upright)
    P, S, R, H, W = 10, 2, 24, 136, 148
    u = synth_unary(41, P, S, R, H, W)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(inp, ep, model, joints, u, 4)
    print("cpp wrapper", CPP_TEST_BIN, file=sys.stderr, flush=True)
    res = subprocess.run([CPP_TEST_BIN, inp, out], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = open(out, "rb").read()
    o = 0
    n_root, n_grid = S * H * W, P * S * R * H * W
    root = np.frombuffer(raw, np.float32, n_root, o).reshape(S, H, W)
    o += 4 * n_root
    marg = np.frombuffer(raw, np.float32, n_grid, o).reshape(P, S, R, H, W)
    o += 4 * n_grid
    hyps = []
    for p in range(P):
        nh = int(np.frombuffer(raw, np.int32, 1, o)[0])
        o += 4
        hp = []
        for _ in range(nh):
            rec = np.frombuffer(raw, np.int32, 4, o)
            sc = np.frombuffer(raw, np.float32, 1, o + 16)[0]
            hp.append((tuple(int(v) for v in rec), float(sc)))
            o += 20
        hyps.append(hp)
    masked = np.frombuffer(raw, np.float32, n_grid, o).reshape(P, S, R, H, W)
    o += 4 * n_grid
    best_conf = np.frombuffer(raw, np.float32, 7 * P, o).reshape(P, 7)
    o += 28 * P
    nh = int(np.frombuffer(raw, np.int32, 1, o)[0])
    o += 4
    root_hyps = np.frombuffer(raw, np.float32, 5 * nh, o).reshape(nh, 5)
    assert o + 20 * nh == len(raw)

    r32, p32, b32, s32 = oracle.root_posterior_rot(oracle.exp_param_from(ep), model.is_detect, model.is_upright, 4,
                                                   joints, u)
    for s in range(S):
        for p in range(P):
            assert oracle.rel_l1(oracle.normalized(marg[p, s]), oracle.normalized(p32[p, s])) <= 1e-5
    for p in range(P):
        (scaleidx, rotidx, y, x), score = hyps[p][0]
        assert (scaleidx, rotidx, y, x) == (S - 1,) + tuple(b32[S - 1, p])
        # then the local maxima of the last scale's marginal, tagged scale 0
        # (objectdetect_findrot.cpp:280-283, objectdetect_aux.cpp:304); on the
        # float32 oracle's marginal where values agree to the last bit, else on ours
        lm = oracle.find_local_max(marg[p, S - 1], 1000)
        assert [h[0] for h in hyps[p][1:]] == [(0, int(d), int(y), int(x)) for d, x, y, v in lm]
        assert len(hyps[p]) - 1 == len(lm) > 0
    # best_conf rows = toVect() of best_part_hyp[p][0] (objectdetect_findrot.cpp:1005-1008)
    for p in range(P):
        (scaleidx, rotidx, y, x), score = hyps[p][0]
        scale = 0.7 + (1.3 - 0.7) / S * (0.5 + scaleidx)
        rot = -180.0 + 360.0 / R * (0.5 + rotidx)
        np.testing.assert_allclose(best_conf[p], [scaleidx, scale, rotidx, rot, x, y, score], rtol=1e-6)
    # root HypothesisList (:1037-1048): local maxima of the root posterior over
    # (scale, y, x), flip set, moved by the bbox offset (-12.7, 5.3) truncated to int
    lm = oracle.find_local_max(root, 1000)
    assert nh == len(lm) > 0
    for (d, x, y, v), h in zip(lm, root_hyps):
        # scale_from_index on the float proto fields (partapp_aux.hpp:44-55)
        mn, mx = float(np.float32(0.7)), float(np.float32(1.3))
        scale = np.float32(mn + (mx - mn) / S * (0.5 + d))
        assert (h[0], h[1]) == (int(x - 12.7), int(y + 5.3))
        assert h[2] == scale and h[3] == np.float32(v) and h[4] == 1.0
    finite = r32 > -1e5
    np.testing.assert_allclose(root[finite], r32[finite], rtol=2e-5, atol=2e-5)
    # masks written into the caller's grids as the reference does (:509-551)
    exp = u.copy()
    keep = np.zeros(R, bool)
    keep[[11, 12]] = True
    for p in (3, 4):
        exp[p][:, ~keep] = -1e6
    strip = int(np.float32(0.1) * np.float32(W))
    exp[4][..., :strip] = -1e6
    exp[4][..., W - strip:] = -1e6
    np.testing.assert_array_equal(masked, exp)
