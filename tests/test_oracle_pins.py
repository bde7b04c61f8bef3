"""Pins of the oracle (CPU).  The reference ships no golden vectors
(code_test.zip is a missing blob), so every function of the oracle is checked
against what mathematics or an independent library routine fixes:
closed forms, delta / constant responses, scipy's Gaussian filters, explicit
index arithmetic for the warps, and exact enumeration of a small joint.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle

LOG_ZERO = -1e6


# ------------------------------------------------------------ index helpers

def test_value_from_index_bin_centres():
    # partapp_aux.hpp:44-55: min + step * (idx + 0.5)
    for r in range(24):
        assert oracle.value_from_index(-180.0, 180.0, 24, r) == pytest.approx(-172.5 + 15.0 * r, abs=1e-12)
    assert oracle.value_from_index(1.0, 1.0, 1, 0) == 1.0  # min == max special case
    # scales of the uneven batch: float proto fields 0.7f / 1.3f, 3 steps
    lo, hi = float(np.float32(0.7)), float(np.float32(1.3))
    s = [oracle.value_from_index(lo, hi, 3, i) for i in range(3)]
    assert s == pytest.approx([0.8, 1.0, 1.2], abs=1e-6)
    # index_from_value inverts it, and the upright root keeps bins 11, 12 (:699-700)
    assert oracle.index_from_value(-180.0, 180.0, 24, -1e-6) == 11
    assert oracle.index_from_value(-180.0, 180.0, 24, 1e-6) == 12
    assert oracle.index_from_value(-180.0, 180.0, 24, 180.0) == -1  # out of range is an error


# ------------------------------------------------------- a8 gaussian taps

@pytest.mark.parametrize("sigma", [0.382, 1.0, 4.0, 12.0, 14.4])
def test_gaussian_filter_closed_form(sigma):
    f = oracle.gaussian_filter(sigma)
    k = (len(f) - 1) // 2
    assert k == math.floor(3 * sigma + 0.5)
    assert f[k] == 1.0
    np.testing.assert_array_equal(f, f[::-1])  # symmetric
    i = np.arange(1, k + 1)
    # log taps are exactly quadratic with curvature -1/sigma^2
    np.testing.assert_allclose(np.log(f[k + 1:]), -(i ** 2) / (2 * sigma ** 2), rtol=1e-13, atol=1e-300)
    if sigma >= 1.0:
        # unnormalised: sum of a +-3 sigma Gaussian of peak 1
        assert f.sum() == pytest.approx(sigma * math.sqrt(2 * math.pi) * math.erf(3 / math.sqrt(2)), rel=2e-2)


# ------------------------------------------------- a6 separable truncated filter

def _taps(sigma):
    return oracle.gaussian_filter(sigma)


def test_diag2d_delta_is_outer_product_truncated():
    H, W = 11, 17
    sx, sy = 2.2, 1.3
    g = np.zeros((H, W))
    y0, x0 = 2, 14  # near the top and right borders: truncation
    g[y0, x0] = 1.0
    out = oracle.gauss_filter_diag2d(g, [sx * sx, 0, 0, sy * sy])
    fx, fy = _taps(sx), _taps(sy)
    kx, ky = (len(fx) - 1) // 2, (len(fy) - 1) // 2
    exp = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            if abs(y - y0) <= ky and abs(x - x0) <= kx:
                exp[y, x] = fy[y - y0 + ky] * fx[x - x0 + kx]
    np.testing.assert_allclose(out, exp, rtol=0, atol=1e-15)


def test_diag2d_constant_gives_truncated_row_sums():
    H, W = 9, 13
    sx, sy = 1.7, 2.4
    out = oracle.gauss_filter_diag2d(np.full((H, W), 3.0), [sx * sx, 0, 0, sy * sy])
    fx, fy = _taps(sx), _taps(sy)
    kx, ky = (len(fx) - 1) // 2, (len(fy) - 1) // 2
    colsum = np.array([sum(fx[u - x + kx] for u in range(max(0, x - kx), min(W, x + kx + 1))) for x in range(W)])
    rowsum = np.array([sum(fy[v - y + ky] for v in range(max(0, y - ky), min(H, y + ky + 1))) for y in range(H)])
    np.testing.assert_allclose(out, 3.0 * np.outer(rowsum, colsum), rtol=1e-14)


def test_diag2d_matches_scipy_gaussian_filter():
    # scipy's kernel radius is int(truncate * sigma + 0.5) = floor(3 sigma + 0.5)
    # and it normalises; mode='constant' is zero padding without renormalisation
    rng = np.random.default_rng(1)
    g = rng.random((23, 31))
    sx, sy = 3.1, 2.2
    out = oracle.gauss_filter_diag2d(g, [sx * sx, 0, 0, sy * sy])
    ref = ndimage.gaussian_filter1d(g, sx, axis=1, mode="constant", truncate=3.0)
    ref = ndimage.gaussian_filter1d(ref, sy, axis=0, mode="constant", truncate=3.0)
    ref *= _taps(sx).sum() * _taps(sy).sum()
    np.testing.assert_allclose(out, ref, rtol=1e-12)


def test_diag2d_brute_force_2d():
    rng = np.random.default_rng(2)
    H, W = 7, 9
    g = rng.random((H, W))
    sx, sy = 0.9, 1.6
    out = oracle.gauss_filter_diag2d(g, [sx * sx, 0, 0, sy * sy])
    exp = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            for v in range(H):
                for u in range(W):
                    dx, dy = u - x, v - y
                    if abs(dx) <= math.floor(3 * sx + 0.5) and abs(dy) <= math.floor(3 * sy + 0.5):
                        exp[y, x] += g[v, u] * math.exp(-dx * dx / (2 * sx * sx)) * math.exp(-dy * dy / (2 * sy * sy))
    np.testing.assert_allclose(out, exp, rtol=1e-13)


def test_diag2d_float32_matches_float64():
    rng = np.random.default_rng(3)
    g = rng.random((40, 50)).astype(np.float32)
    C = [49.0, 0, 0, 25.0]
    o32 = oracle.gauss_filter_diag2d(g, C)
    o64 = oracle.gauss_filter_diag2d(g.astype(np.float64), C)
    np.testing.assert_allclose(o32, o64, rtol=2e-6)


def test_diag2d_rejects_bad_covariance():
    with pytest.raises(oracle.OracleError):
        oracle.gauss_filter_diag2d(np.ones((4, 4)), [0.0, 0, 0, 1.0])
    with pytest.raises(oracle.OracleError):
        oracle.gauss_filter_diag2d(np.ones((4, 4)), [1.0, 0.5, 0.5, 1.0])  # off-diagonal: another path


# ----------------------------------------------------- a5 circular filter

def test_wraparound_delta_gives_wrapped_kernel():
    R = 24
    f = _taps(1.0)  # 7 taps
    n = (len(f) - 1) // 2
    for j in (0, 3, 23):
        x = np.zeros(R)
        x[j] = 1.0
        out = oracle.filter_1d_wraparound(x, f)
        exp = np.zeros(R)
        for s in range(R):
            i = (j - s + n) % R
            if i < len(f):
                exp[s] = f[i]
        np.testing.assert_array_equal(out, exp)


def test_wraparound_matches_scipy_wrap():
    rng = np.random.default_rng(4)
    x = rng.random(24)
    sigma = 1.3
    out = oracle.filter_1d_wraparound(x, _taps(sigma))
    ref = ndimage.gaussian_filter1d(x, sigma, mode="wrap", truncate=3.0) * _taps(sigma).sum()
    np.testing.assert_allclose(out, ref, rtol=1e-13)


@pytest.mark.parametrize("R,expect_len", [(5, 3), (6, 5)])
def test_rotation_kernel_clipping(R, expect_len):
    # objectdetect_findrot.cpp:385-397: sigma_idx = 3 gives 19 taps >= R; keep
    # R-2 (odd R) or R-1 (even R) taps around the centre
    step = math.radians(float(np.float32(360.0) / np.float32(R)))
    ep = oracle.exp_param(R=R)
    plan = oracle.message_plan(ep, R, 4, 4, [0, 0], [0, 0], 0.0, 3.0 * step, 1.0)
    taps = np.array(plan["rot_taps"])
    assert len(taps) == expect_len
    h = (expect_len - 1) // 2
    i = np.arange(-h, h + 1)
    np.testing.assert_allclose(taps, np.exp(-(i ** 2) / (2 * 9.0)), rtol=1e-9)


# ------------------------------------------------------------- a4 warps

def _translation(dx, dy):
    return [[1, 0, dx], [0, 1, dy], [0, 0, 1]]


def test_warp_integer_shift_and_fill():
    rng = np.random.default_rng(5)
    g = rng.random((6, 8))
    out = oracle.transform_nearest(g, _translation(2.0, -1.0), -7.0)
    exp = np.full_like(g, -7.0)
    for y in range(6):
        for x in range(8):
            ys, xs = y + 1, x - 2  # out(y, x) = in(y - dy, x - dx)
            if 0 <= ys < 6 and 0 <= xs < 8:
                exp[y, x] = g[ys, xs]
    np.testing.assert_array_equal(out, exp)


@pytest.mark.parametrize("d,src_shift", [(0.5, 0), (-0.5, 1), (1.5, -1), (-1.5, 2), (0.49, 0), (0.51, -1)])
def test_warp_half_integer_direction(d, src_shift):
    # nearest = floor(x - d + 0.5): halves round up
    g = np.arange(10.0)[None, :].repeat(3, 0)
    out = oracle.transform_nearest(g, _translation(d, 0.0), -1.0)
    for x in range(10):
        xs = x + src_shift
        assert out[0, x] == (xs if 0 <= xs < 10 else -1.0)


def test_warp_rotation_quarter_turn():
    # T21 maps in -> out: x3 = -y1 + (n-1), y3 = x1, so out[y][x] = in[n-1-x][y]
    n = 5
    g = np.arange(n * n, dtype=np.float64).reshape(n, n)
    out = oracle.transform_nearest(g, [[0, -1, n - 1], [1, 0, 0], [0, 0, 1]], -1.0)
    for y in range(n):
        for x in range(n):
            assert out[y, x] == g[n - 1 - x, y]


# ------------------------------------------------------- a3 the message

def _message(child, **kw):
    args = dict(offset_c=[0, 0], offset_p=[0, 0], C=[4.0, 0, 0, 2.25], rot_mean=0.0, rot_sigma=0.0, scale=1.0)
    args.update(kw)
    ep = args.pop("ep", oracle.exp_param(R=child.shape[0]))
    return oracle.rot_joint_marginal(ep, child, **args)


def _textbook(b, sx, sy, sr=None):
    """log(G_y * G_x * G_r(exp(b - M))) + M with scipy's Gaussian filters."""
    M = b.max()
    e = np.exp(b - M)
    if sr is not None:
        e = ndimage.gaussian_filter1d(e, sr, axis=0, mode="wrap", truncate=3.0) * _taps(sr).sum()
    e = ndimage.gaussian_filter1d(e, sx, axis=2, mode="constant", truncate=3.0) * _taps(sx).sum()
    e = ndimage.gaussian_filter1d(e, sy, axis=1, mode="constant", truncate=3.0) * _taps(sy).sum()
    with np.errstate(divide="ignore"):
        return np.where(e == 0, LOG_ZERO, np.log(e)) + M


def test_message_special_case_is_a_gaussian_blur_in_prob_space():
    rng = np.random.default_rng(6)
    b = rng.normal(-3, 1, (24, 15, 19))
    out = _message(b)  # offsets 0, delta 0, rot_sigma 0
    np.testing.assert_allclose(out, _textbook(b, 2.0, 1.5), rtol=1e-11, atol=1e-12)


def test_message_rotation_filter_is_circular_blur():
    rng = np.random.default_rng(7)
    R = 24
    b = rng.normal(-3, 1, (R, 12, 14))
    step = math.radians(15.0)
    out = _message(b, rot_sigma=1.0 * step)
    np.testing.assert_allclose(out, _textbook(b, 2.0, 1.5, sr=1.0), rtol=1e-11, atol=1e-12)


def test_message_offsets_shift_forward_then_back():
    # R = 1: the only rotation bin is 0 degrees, so the joint offsets are exact
    # integer translations: J(y, x) = b(y - dy_c, x - dx_c), msg(y, x) = L(y + dy_p, x + dx_p)
    rng = np.random.default_rng(8)
    H, W = 20, 26
    b = rng.normal(-2, 1, (1, H, W))
    oc, op = (3.0, -2.0), (-4.0, 5.0)
    out = _message(b, offset_c=oc, offset_p=op)
    M = b.max()
    J = np.full((1, H, W), LOG_ZERO)
    for y in range(H):
        for x in range(W):
            ys, xs = y - int(oc[1]), x - int(oc[0])
            if 0 <= ys < H and 0 <= xs < W:
                J[0, y, x] = b[0, ys, xs]
    L = _textbook(J, 2.0, 1.5)
    L = L - L.max() + L.max()  # same M: J and b share the max unless it was cut off
    assert J.max() == M
    exp = np.full((1, H, W), LOG_ZERO)
    for y in range(H):
        for x in range(W):
            ys, xs = y + int(op[1]), x + int(op[0])
            if 0 <= ys < H and 0 <= xs < W:
                exp[0, y, x] = L[0, ys, xs]
    np.testing.assert_allclose(out, exp, rtol=1e-11, atol=1e-12)


def test_message_rot_mean_shifts_slices_without_wrapping():
    # objectdetect_findrot.cpp:323-347: delta = round(-rot_mean / step), output
    # slice r + delta receives source slice r; slices with no source are LOG_ZERO
    rng = np.random.default_rng(9)
    R = 24
    b = rng.normal(-2, 1, (R, 10, 12))
    step = math.radians(15.0)
    out = _message(b, rot_mean=-2.0 * step)  # delta = +2
    J = np.full_like(b, LOG_ZERO)
    J[2:] = b[:-2]
    exp = _textbook(J, 2.0, 1.5)
    # _textbook uses max(J); the message uses max(b) = M of the whole child
    M = b.max()
    E = np.exp(J - M)
    E = ndimage.gaussian_filter1d(E, 2.0, axis=2, mode="constant", truncate=3.0) * _taps(2.0).sum()
    E = ndimage.gaussian_filter1d(E, 1.5, axis=1, mode="constant", truncate=3.0) * _taps(1.5).sum()
    exp = np.where(E == 0, LOG_ZERO, np.log(np.where(E == 0, 1, E))) + M
    np.testing.assert_allclose(out, exp, rtol=1e-11, atol=1e-12)
    np.testing.assert_array_equal(out[:2], LOG_ZERO + M)


def test_message_sentinels():
    # a single finite entry: far pixels filter to exactly 0 -> LOG_ZERO + M,
    # pixels whose inverse warp leaves the grid -> LOG_ZERO
    R, H, W = 1, 30, 30
    b = np.full((R, H, W), LOG_ZERO)
    b[0, 15, 15] = 5.0
    out = _message(b, offset_p=(0.0, 12.0), C=[1.0, 0, 0, 1.0])
    assert out[0, 0, 0] == LOG_ZERO + 5.0          # inside, zero response
    assert out[0, 29, 5] == LOG_ZERO               # y + 12 >= H
    assert out[0, 3, 15] == pytest.approx(5.0)     # the impulse itself (3 + 12 = 15)


def test_message_float32_close_to_float64():
    rng = np.random.default_rng(10)
    b = rng.normal(-3, 0.5, (24, 30, 34))
    kw = dict(offset_c=[7.3, -11.2], offset_p=[-3.9, 8.8], C=[30.0, 0, 0, 50.0], rot_mean=0.3, rot_sigma=0.25)
    o64 = _message(b, **kw)
    o32 = _message(b.astype(np.float32), **kw)
    p64, p32 = oracle.normalized(o64), oracle.normalized(o32)
    assert oracle.rel_l1(p32, p64) < 1e-5


# ------------------------------------------- a1 + a2: the sweep, brute force

def _tiny_tree(seed, rot_mean=0.0):
    """Root 0 with children 1 and 2; 2 has child 3."""
    from dataclasses import dataclass

    @dataclass
    class J:
        child_idx: int
        parent_idx: int
        offset_c: tuple
        offset_p: tuple
        C: tuple
        rot_mean: float
        rot_sigma: float

    rng = np.random.default_rng(seed)
    step = math.radians(90.0)
    joints = []
    for c, p in ((1, 0), (2, 0), (3, 2)):
        sx, sy = rng.uniform(0.6, 1.0, 2)
        joints.append(J(c, p, tuple(rng.uniform(-1.3, 1.3, 2)), tuple(rng.uniform(-1.3, 1.3, 2)),
                        (sx * sx, 0.0, 0.0, sy * sy), rot_mean, 0.45 * step))
    return joints


def _pairwise_matrix(ep, jt, R, H, W):
    """K[s_parent, s_child] = exp(message of a one-hot child), float64."""
    n = R * H * W
    K = np.zeros((n, n))
    for i in range(n):
        b = np.full(n, LOG_ZERO)
        b[i] = 0.0
        m = oracle.rot_joint_marginal(ep, b.reshape(R, H, W), jt.offset_c, jt.offset_p, jt.C, jt.rot_mean,
                                      jt.rot_sigma, 1.0)
        K[:, i] = np.exp(m.ravel())
    return K


def test_sweep_equals_brute_force_enumeration():
    R, H, W = 4, 3, 4
    ep = oracle.exp_param(R=R)
    joints = _tiny_tree(11)
    # the half-integer-free offsets make the downward operator the transpose of the upward one
    rng = np.random.default_rng(12)
    u = rng.normal(0, 1, (4, 1, R, H, W))
    root_post, post, best, score = oracle.root_posterior_rot(ep, [1] * 4, [0] * 4, 0, joints, u)
    n = R * H * W
    K = {(j.child_idx, j.parent_idx): _pairwise_matrix(ep, j, R, H, W) for j in joints}
    # check the transpose property on which the enumeration relies
    for j in joints:
        down = _pairwise_matrix(ep, type(j)(j.parent_idx, j.child_idx, j.offset_p, j.offset_c, j.C, -j.rot_mean,
                                            j.rot_sigma), R, H, W)
        np.testing.assert_allclose(down, K[(j.child_idx, j.parent_idx)].T, rtol=1e-12, atol=1e-300)
    eu = [np.exp(u[p, 0].ravel()) for p in range(4)]
    joint = (eu[0][:, None, None, None] * eu[1][None, :, None, None] * eu[2][None, None, :, None]
             * eu[3][None, None, None, :])
    joint = joint * K[(1, 0)][:, :, None, None] * K[(2, 0)][:, None, :, None] * K[(3, 2)][None, None, :, :]
    Z = joint.sum()
    marg = [joint.sum(axis=tuple(a for a in range(4) if a != p)) / Z for p in range(4)]
    for p in range(4):
        np.testing.assert_allclose(oracle.normalized(post[p, 0]).ravel(), marg[p], rtol=1e-9, atol=1e-14)
        # MAP = first flat argmax
        flat = int(np.argmax(marg[p]))
        assert tuple(best[0, p]) == (flat // (H * W), (flat % (H * W)) // W, flat % W)
    # every unnormalised marginal sums to the same partition function
    lse = [np.log(np.exp(post[p, 0].astype(np.float64)).sum()) for p in range(4)]
    np.testing.assert_allclose(lse, np.log(Z), rtol=1e-12)
    # root posterior: log of the sum over all rotations of the root marginal
    np.testing.assert_allclose(root_post[0], np.log(np.exp(post[0, 0]).sum(axis=0)), rtol=1e-12)


def test_down_message_shifts_rotation_by_plus_rot_mean():
    # rot_mean is the rotation from parent to child (:320-323): a parent peak at
    # rotation r0 makes the child's marginal peak at r0 + rot_mean/step
    R, H, W = 24, 9, 9
    ep = oracle.exp_param(R=R)
    step = math.radians(15.0)

    class J:
        child_idx, parent_idx = 1, 0
        offset_c = offset_p = (0.0, 0.0)
        C = (1.0, 0.0, 0.0, 1.0)
        rot_mean = 3.0 * step
        rot_sigma = 0.0

    u = np.full((2, 1, R, H, W), -5.0)
    u[0, 0, 7] = 0.0  # parent prefers rotation 7
    root_post, post, best, score = oracle.root_posterior_rot(ep, [1, 1], [0, 0], 0, [J()], u)
    child_by_rot = np.exp(post[1, 0]).sum(axis=(1, 2))
    assert int(np.argmax(child_by_rot)) == 10
    assert best[0, 1][0] == 10


def test_upright_and_strip_masks():
    R, H, W = 24, 8, 20
    ep = oracle.exp_param(R=R, strip=0.2)

    class J:
        child_idx, parent_idx = 1, 0
        offset_c = offset_p = (0.0, 0.0)
        C = (1.0, 0.0, 0.0, 1.0)
        rot_mean = 0.0
        rot_sigma = 0.0

    rng = np.random.default_rng(13)
    u = rng.normal(0, 1, (2, 1, R, H, W))
    root_post, post, best, score = oracle.root_posterior_rot(ep, [1, 1], [1, 1], 0, [J()], u)
    # upright parts keep rotations with |rot| < 15 degrees: bins 11 and 12 at R = 24
    for p in range(2):
        keep = np.exp(post[p, 0]).sum(axis=(1, 2)) > 0
        assert list(np.nonzero(keep)[0]) == [11, 12]
    # the root loses int(0.2 * W) = 4 border columns on each side
    assert np.all(post[0, 0][:, :, :4] < -1e5) and np.all(post[0, 0][:, :, -4:] < -1e5)
    assert np.all(post[0, 0][11:13, :, 4:-4] > -1e5)
    # upright root: the posterior sums bins 11 and 12 only
    s = np.exp(post[0, 0, 11]) + np.exp(post[0, 0, 12])
    exp = np.where(s == 0, LOG_ZERO, np.log(np.where(s == 0, 1.0, s)))  # log of 0 is LOG_ZERO (:725)
    np.testing.assert_allclose(root_post[0], exp, rtol=1e-6)


# ----------------------------------------------------- a15 local maxima

def test_find_local_max_hand_built():
    g = np.zeros((3, 5, 6), dtype=np.float32)
    g[1, 2, 3] = 5.0           # strict maximum
    g[0, 2, 3] = 4.0           # below the slice above it -> not a maximum
    g[2, 0, 0] = 3.0           # corner maximum
    g[1, 4, 5] = 2.0
    out = oracle.find_local_max(g, 1000)
    got = {(int(d), int(x), int(y)): v for d, x, y, v in out if v > 0}
    assert got == {(1, 3, 2): 5.0, (2, 0, 0): 3.0, (1, 5, 4): 2.0}
    top = oracle.find_local_max(g, 2)
    assert [r[3] for r in top] == [5.0, 3.0]  # top-K by value
