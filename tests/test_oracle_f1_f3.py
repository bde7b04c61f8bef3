"""Pins of the oracle's off-diagonal covariance path (gaussFilter2dOffset,
multi_array_filter.hpp:334-369, eig2d boost_math.cpp:40-99) and of the
direct / bilinear warps and score ingestion it shares with loadScoreGrid
(multi_array_transform.hpp:122-239, partapp.cpp:866-899)."""
import math

import numpy as np
import pytest

import oracle

LOG_ZERO = -1e6


@pytest.mark.parametrize("C", [(5.0, 2.0, 2.0, 3.0), (9.0, -4.0, -4.0, 9.0), (2.0, 0.0, 0.0, 7.0),
                               (7.0, 0.0, 0.0, 2.0), (16.0, 7.9, 7.9, 4.0)])
def test_eig2d_reconstructs(C):
    V, E = oracle.eig2d(C)
    M = np.array(C).reshape(2, 2)
    np.testing.assert_allclose(V @ E @ V.T, M, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(V.T @ V, np.eye(2), atol=1e-12)
    assert E[0, 1] == E[1, 0] == 0 and E[0, 0] <= E[1, 1]
    np.testing.assert_allclose(np.diag(E), np.linalg.eigvalsh(M), rtol=1e-12)


def test_direct_warp_quarter_turn_and_last_writer():
    n = 6
    g = (np.arange(n * n, dtype=np.float32) + 1).reshape(n, n)
    # x3 = -y1 + (n-1), y3 = x1: out[y][x] = in[n-1-x][y]
    out = oracle.transform_direct(g, [[0, -1, n - 1], [1, 0, 0], [0, 0, 1]], (n, n))
    for y in range(n):
        for x in range(n):
            assert out[y, x] == g[n - 1 - x, y]
    # a 2x downscale maps pairs of cells onto one: the later cell of the scan
    # (x outer, y inner) wins
    h = np.arange(1, 17, dtype=np.float32).reshape(4, 4)
    out = oracle.transform_direct(h, [[0.5, 0, 0], [0, 0.5, 0], [0, 0, 1]], (2, 2))
    # x3 = floor(0.5 x1 + 0.5): x1 = 0 -> 0, 1, 2 -> 1, 3 -> 2 (dropped)
    assert out[0, 0] == h[0, 0]
    assert out[1, 1] == h[2, 2]
    assert out[0, 1] == h[0, 2] and out[1, 0] == h[2, 0]


def test_direct_warp_skips_default_values():
    g = np.array([[0.0, 5.0]], dtype=np.float32)
    # both cells land on output (0, 0): the zero (the default value) is not written
    out = oracle.transform_direct(g, [[0.0, 0, 0], [0, 1, 0], [0, 0, 1]], (1, 2), 0.0)
    assert out[0, 0] == 5.0


def test_bilinear_warp_half_pixel_and_borders():
    rng = np.random.default_rng(1)
    g = rng.random((5, 7)).astype(np.float32)
    # T21 translates by (+0.5, 0): out(x) samples in(x - 0.5)
    out = oracle.transform_bilinear(g, [[1, 0, 0.5], [0, 1, 0], [0, 0, 1]], (5, 7))
    for y in range(4):
        for x in range(1, 7):
            assert out[y, x] == pytest.approx(0.5 * g[y, x - 1] + 0.5 * g[y, x], rel=1e-6)
    assert np.all(out[:, 0] == 0)   # x - 0.5 < 0: outside
    assert np.all(out[4, 1:] == 0)  # last row: no neighbour below, not on a grid point
    # an integer translation lands on grid points: copied exactly
    out = oracle.transform_bilinear(g, [[1, 0, 2], [0, 1, -1], [0, 0, 1]], (5, 7))
    np.testing.assert_array_equal(out[:4, 2:], g[1:, :5])


def test_offset_filter_impulse_has_covariance_C():
    # far from the borders the response to an impulse is the anisotropic
    # Gaussian exp(-d' C^-1 d / 2) (up to interpolation): its normalised second
    # moments are C
    H = W = 81
    C = np.array([[16.0, 7.0], [7.0, 9.0]])
    g = np.zeros((H, W))
    g[40, 40] = 1.0
    for sparse in (False, True):
        out = oracle.gauss_filter2d(g, C.ravel(), sparse)
        yy, xx = np.mgrid[0:H, 0:W]
        m = out.sum()
        mx, my = (out * xx).sum() / m, (out * yy).sum() / m
        cxx = (out * (xx - mx) ** 2).sum() / m
        cyy = (out * (yy - my) ** 2).sum() / m
        cxy = (out * (xx - mx) * (yy - my)).sum() / m
        assert abs(mx - 40) < 0.6 and abs(my - 40) < 0.6
        np.testing.assert_allclose([cxx, cxy, cyy], [16.0, 7.0, 9.0], atol=1.6)


def test_offset_filter_diagonal_rotation_is_the_diagonal_filter():
    # C = R diag(a, b) R' with R a quarter turn is diag(b, a): filtering through
    # the eigenbasis of a slightly rotated C converges to the diagonal filter
    rng = np.random.default_rng(2)
    g = rng.random((40, 50))
    d = oracle.gauss_filter2d(g, [4.0, 0, 0, 9.0], False)
    o = oracle.gauss_filter2d(g, [4.0, 1e-9, 1e-9, 9.0], False)
    inner = (slice(5, -5), slice(5, -5))
    np.testing.assert_allclose(o[inner], d[inner], rtol=2e-2)


def test_message_with_full_covariance_runs_both_directions():
    rng = np.random.default_rng(3)
    b = rng.normal(-3, 1, (24, 30, 34))
    ep = oracle.exp_param(R=24)
    kw = dict(offset_c=[3.0, -4.0], offset_p=[2.0, 1.0], C=[20.0, 6.0, 6.0, 12.0], rot_mean=0.2, rot_sigma=0.25,
              scale=1.0)
    up = oracle.rot_joint_marginal(ep, b, is_sparse=True, **kw)
    dn = oracle.rot_joint_marginal(ep, b, is_sparse=False, **kw)
    assert np.isfinite(up).all() and np.isfinite(dn).all()
    # the sparse (direct) and bilinear rotations differ only by interpolation
    live = (up > -9e5) & (dn > -9e5)
    assert live.mean() > 0.5
    p_up, p_dn = oracle.normalized(np.where(live, up, -1e30)), oracle.normalized(np.where(live, dn, -1e30))
    assert oracle.rel_l1(p_up, p_dn) < 0.2


def test_ingest_scores_direct_and_bilinear():
    rng = np.random.default_rng(4)
    # a score grid sampled every 2 pixels, mapped to a 2x larger image
    sg = rng.normal(0.5, 1.0, (2, 6, 8)).astype(np.float32)
    T = np.array([[[2, 0, 1], [0, 2, 1], [0, 0, 1]]] * 2, dtype=np.float64)
    out = oracle.ingest_scores(sg, T, False, 14, 18)
    for s in range(2):
        for y in range(6):
            for x in range(8):
                v = sg[s, y, x]
                exp = math.log(np.float32(1e-4)) if v < 0 else math.log(v)
                assert out[s, 2 * y + 1, 2 * x + 1] == pytest.approx(exp, rel=1e-6)
        # unmapped pixels had NO_CLASS_VALUE = 0 -> LOG_ZERO
        assert out[s, 0, 0] == LOG_ZERO and out[s, 2, 2] == LOG_ZERO
    outb = oracle.ingest_scores(sg, T, True, 14, 18)

    def clog(v):
        return math.log(np.float32(1e-4)) if v < 0 else math.log(v)

    # bilinear: pixel (3, 3) is grid cell (1, 1) exactly; pixel (4, 4) sits
    # halfway between cells (1,1), (1,2), (2,1), (2,2)
    assert outb[0, 3, 3] == pytest.approx(clog(sg[0, 1, 1]), rel=1e-6)
    v = 0.25 * (float(sg[0, 1, 1]) + float(sg[0, 1, 2]) + float(sg[0, 2, 1]) + float(sg[0, 2, 2]))
    assert outb[0, 4, 4] == pytest.approx(clog(v), rel=1e-5, abs=1e-5)
