"""Pins of the oracle's extra unaries (setTorsoPosPrior, addLoadDPMScore,
addDPMScore, computeLogGrid, addExtraUnary; objectdetect_icps.cpp:137-191,
445-548, multi_array_op.hpp:154-166) against values the formulas fix:
the prior's peak sits at (cx - mu_x, cy - mu_y) with value log(weight), one
standard deviation away it is log(weight) - 1/2, far away the 1e-4 clip;
raw DPM scores at or below 1e-4 add the unweighted log(1e-4) (a float 1e-4
is below the double 1e-4); slice selection by rotation count; LOG_ZERO for
zero scores; the plain weighted sum."""
import numpy as np
import pytest

import oracle


def ulps(a, b):
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


def test_torso_prior_peak_sigma_and_clip():
    H, W = 40, 60
    mu_x, mu_y, var_x, var_y, weight = 6.0, -4.0, 25.0, 16.0, 0.8
    g = np.zeros((3, H, W), dtype=np.float32)
    out = oracle.torso_pos_prior(g, (mu_x, mu_y, var_x, var_y), weight)
    # every slice gets the same addend
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    cx, cy = 0.5 * W, 0.5 * H
    px, py = int(cx - mu_x), int(cy - mu_y)  # ix_rel = cx - ix = mu_x
    assert np.unravel_index(np.argmax(out[0]), (H, W)) == (py, px)
    assert ulps(out[0, py, px], np.log(np.float32(weight))) <= 1
    # one standard deviation along x (5 px): log(weight * e^-1/2)
    assert out[0, py, px - 5] == pytest.approx(np.log(weight) - 0.5, abs=2e-6)
    assert out[0, py, px + 5] == pytest.approx(np.log(weight) - 0.5, abs=2e-6)
    # and along y (4 px)
    assert out[0, py + 4, px] == pytest.approx(np.log(weight) - 0.5, abs=2e-6)
    # far corner: the score clips to 1e-4 before the log
    assert ulps(out[0, 0, W - 1], np.log(np.float32(1e-4))) <= 1
    # the whole grid against a float64 evaluation of the formula
    x = np.arange(W, dtype=np.float64)
    y = np.arange(H, dtype=np.float64)
    sx = np.exp(-0.5 * (cx - x - mu_x) ** 2 / var_x)
    sy = np.exp(-0.5 * (cy - y - mu_y) ** 2 / var_y)
    ref = np.log(np.maximum(weight * sy[:, None] * sx[None, :], 1e-4))
    np.testing.assert_allclose(out[0], ref, rtol=0, atol=3e-6)


def test_torso_prior_adds_in_double():
    # float += double: the double addend is added before one rounding
    g = np.full((1, 4, 4), 1000.0, dtype=np.float32)
    out = oracle.torso_pos_prior(g, (100.0, 100.0, 1.0, 1.0), 1.0)  # clipped everywhere
    want = np.float32(np.float64(np.float32(1000.0)) + np.float64(np.log(np.float32(1e-4))))
    assert np.all(ulps(out, want) <= 1)


def test_dpm_raw_clip_and_weights():
    S, R, H, W = 2, 3, 2, 3
    dpm = np.array([[[0.5, 1e-5, np.float32(1e-4)], [2.0, 0.0, 1.5e-4]]], dtype=np.float32)
    g = np.zeros((S, R, H, W), dtype=np.float32)
    out = oracle.add_dpm_raw(g, dpm, 2.0)
    want = np.array([[2 * np.log(0.5), np.log(1e-4), np.log(1e-4)],
                     [2 * np.log(2.0), np.log(1e-4), 2 * np.log(1.5e-4)]])
    for s in range(S):
        for r in range(R):  # one DPM slice for every rotation
            np.testing.assert_allclose(out[s, r], want, rtol=0, atol=2e-6)
    # n_dpm == R: slice r goes to rotation r; any other count: slice 0
    d3 = np.stack([np.full((H, W), v, dtype=np.float32) for v in (0.25, 0.5, 1.0)])
    out3 = oracle.add_dpm_raw(g, d3, 1.0)
    for r, v in enumerate((0.25, 0.5, 1.0)):
        np.testing.assert_allclose(out3[:, r], np.log(v), atol=1e-6)
    out2 = oracle.add_dpm_raw(g, d3[:2], 1.0)
    np.testing.assert_allclose(out2, np.log(0.25), atol=1e-6)


def test_dpm_log_rotation_rule_and_log_grid():
    S, R, H, W = 1, 3, 2, 2
    g = np.ones((S, R, H, W), dtype=np.float32)
    one = np.full((1, H, W), -2.0, dtype=np.float32)
    np.testing.assert_array_equal(oracle.add_dpm_log(g, one, 0.5), np.float32(1.0) + np.float32(0.5) * np.float32(-2))
    per_rot = np.stack([np.full((H, W), -float(r), dtype=np.float32) for r in range(R)])
    out = oracle.add_dpm_log(g, per_rot, 2.0)
    for r in range(R):
        assert np.all(out[0, r] == np.float32(1.0 - 2.0 * r))
    with pytest.raises(oracle.OracleError):
        oracle.add_dpm_log(g, per_rot, 0.0)  # the reference asserts dpm_weight > 0
    lg = oracle.log_grid(np.array([0.0, 1.0, np.e, 0.5], dtype=np.float32))
    assert lg[0] == np.float32(-1e6) and lg[1] == 0.0
    assert ulps(lg[2], 1.0) <= 1 and ulps(lg[3], np.log(np.float32(0.5))) <= 1


def test_extra_unary_weighted_sum():
    rng = np.random.default_rng(3)
    g = rng.normal(size=(2, 3, 4, 5)).astype(np.float32)
    e = rng.normal(size=g.shape).astype(np.float32)
    out = oracle.add_extra_unary(g, e, 0.7)
    np.testing.assert_array_equal(out, g + np.float32(0.7) * e)
