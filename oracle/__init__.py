"""The oracle: TEST INFRASTRUCTURE ONLY.

A CPU restatement of partapp's pictorial-structure inference path
(ps_oracle.cpp, built by oracle/Makefile into oracle/_build/libps_oracle.so),
with numpy wrappers.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, and only as the checker or the CPU baseline; the
product (partapp_amd) never imports it.

Pins: see tests/test_oracle_pins.py (closed forms, delta and constant
responses, brute-force convolution, integer-shift warps, the rot_sigma = 0 /
zero-offset special case against a textbook blur, exact enumeration of a
4-part tree, equal partition functions).  Unpinnable: agreement with the
reference binary's BLAS summation order and libm rounding (no reference build
exists: Boost, CBLAS, Qt4 and MATLAB's libmat are absent).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libps_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


class orc_exp_param(ctypes.Structure):
    _fields_ = [("min_part_rotation", ctypes.c_float), ("max_part_rotation", ctypes.c_float),
                ("num_rotation_steps", ctypes.c_uint), ("min_object_scale", ctypes.c_float),
                ("max_object_scale", ctypes.c_float), ("num_scale_steps", ctypes.c_uint),
                ("strip_border_detections", ctypes.c_float)]


class orc_joint(ctypes.Structure):
    _fields_ = [("child_idx", ctypes.c_int), ("parent_idx", ctypes.c_int), ("offset_c", ctypes.c_double * 2),
                ("offset_p", ctypes.c_double * 2), ("C", ctypes.c_double * 4), ("rot_mean", ctypes.c_double),
                ("rot_sigma", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_value_from_index.restype = ctypes.c_double
        L.orc_value_from_index.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_index_from_value.argtypes = [ctypes.c_double] * 4
        L.orc_gaussian_filter.argtypes = [ctypes.c_double, _dp, ctypes.c_int]
        for t, p in (("f32", _fp), ("f64", _dp)):
            getattr(L, "orc_filter_1d_wraparound_" + t).argtypes = [p, p, ctypes.c_int, p, ctypes.c_int]
            getattr(L, "orc_gauss_filter_diag2d_" + t).argtypes = [p, p, ctypes.c_int, ctypes.c_int, _dp]
            getattr(L, "orc_transform_nearest_" + t).argtypes = [p, p, ctypes.c_int, ctypes.c_int, _dp,
                                                                 ctypes.c_float if t == "f32" else ctypes.c_double]
            getattr(L, "orc_rot_joint_marginal_" + t).argtypes = [
                ctypes.POINTER(orc_exp_param), p, p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
            getattr(L, "orc_root_posterior_rot_" + t).argtypes = [
                ctypes.POINTER(orc_exp_param), ctypes.c_int, _ip, _ip, ctypes.c_int, ctypes.c_int,
                ctypes.POINTER(orc_joint), ctypes.c_int, ctypes.c_int, p, p, p, _ip, p, ctypes.c_int]
        L.orc_find_local_max_f32.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_message_plan.argtypes = [ctypes.POINTER(orc_exp_param), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _ip, _ip, _ip,
                                       _ip, _ip, _ip, _ip, _dp, _ip, _ip]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _chk(st, where):
    if st != 0:
        raise OracleError("%s returned %d" % (where, st))


def _ptr(a):
    return a.ctypes.data_as(_fp if a.dtype == np.float32 else _dp)


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def exp_param(min_rot=-180.0, max_rot=180.0, R=24, min_scale=1.0, max_scale=1.0, S=1, strip=0.0):
    return orc_exp_param(min_rot, max_rot, R, min_scale, max_scale, S, strip)


def exp_param_from(ep):
    """From any object with the ExpParam attribute names."""
    return orc_exp_param(ep.min_part_rotation, ep.max_part_rotation, ep.num_rotation_steps, ep.min_object_scale,
                         ep.max_object_scale, ep.num_scale_steps, ep.strip_border_detections)


def joints_from(joints):
    arr = (orc_joint * len(joints))()
    for i, j in enumerate(joints):
        arr[i].child_idx, arr[i].parent_idx = j.child_idx, j.parent_idx
        arr[i].offset_c[:] = list(j.offset_c)
        arr[i].offset_p[:] = list(j.offset_p)
        arr[i].C[:] = list(j.C)
        arr[i].rot_mean, arr[i].rot_sigma = j.rot_mean, j.rot_sigma
    return arr


def value_from_index(minval, maxval, num_steps, idx):
    return lib().orc_value_from_index(minval, maxval, num_steps, idx)


def index_from_value(minval, maxval, num_steps, val):
    return lib().orc_index_from_value(minval, maxval, num_steps, val)


def gaussian_filter(sigma):
    buf = np.zeros(4096, dtype=np.float64)
    n = lib().orc_gaussian_filter(sigma, _ptr(buf), buf.size)
    if n < 0:
        raise OracleError("filter too long")
    return buf[:n].copy()


def filter_1d_wraparound(x, f):
    x = np.ascontiguousarray(x)
    f = np.ascontiguousarray(f, dtype=x.dtype)
    out = np.empty_like(x)
    _chk(getattr(lib(), "orc_filter_1d_wraparound_" + _sfx(x.dtype))(_ptr(x), _ptr(out), x.size, _ptr(f), f.size),
         "filter_1d_wraparound")
    return out


def gauss_filter_diag2d(grid, C):
    g = np.ascontiguousarray(grid)
    out = np.empty_like(g)
    c = np.ascontiguousarray(np.asarray(C, dtype=np.float64).reshape(4))
    _chk(getattr(lib(), "orc_gauss_filter_diag2d_" + _sfx(g.dtype))(_ptr(g), _ptr(out), g.shape[0], g.shape[1],
                                                                     _ptr(c)), "gauss_filter_diag2d")
    return out


def transform_nearest(grid, T21, default_value):
    g = np.ascontiguousarray(grid)
    out = np.empty_like(g)
    T = np.ascontiguousarray(np.asarray(T21, dtype=np.float64).reshape(9))
    _chk(getattr(lib(), "orc_transform_nearest_" + _sfx(g.dtype))(_ptr(g), _ptr(out), g.shape[0], g.shape[1],
                                                                   _ptr(T), default_value), "transform_nearest")
    return out


def rot_joint_marginal(ep, child, offset_c, offset_p, C, rot_mean, rot_sigma, scale, is_sparse=False):
    """computeRotJointMarginal on a [R][H][W] grid (float32 or float64)."""
    c = np.ascontiguousarray(child)
    out = np.empty_like(c)
    oc = np.asarray(offset_c, dtype=np.float64)
    op = np.asarray(offset_p, dtype=np.float64)
    cc = np.ascontiguousarray(np.asarray(C, dtype=np.float64).reshape(4))
    R, H, W = c.shape
    _chk(getattr(lib(), "orc_rot_joint_marginal_" + _sfx(c.dtype))(
        ctypes.byref(ep), _ptr(c), _ptr(out), R, H, W, _ptr(oc), _ptr(op), _ptr(cc), rot_mean, rot_sigma, scale,
        int(is_sparse)), "rot_joint_marginal")
    return out


def root_posterior_rot(ep, is_detect, is_upright, root, joints, log_part_detections, is_sparse=True,
                       want_marginals=True):
    """computeRootPosteriorRot.  log_part_detections [P][S][R][H][W] (float32 or
    float64; a masked copy is made).  Returns (root_posterior [S][H][W],
    log_part_posterior [P][S][R][H][W] or None, best_idx [S][P][3] (rot, y, x),
    best_score [S][P])."""
    det = np.array(log_part_detections, copy=True, order="C")
    P, S, R, H, W = det.shape
    dt = det.dtype
    post = np.empty_like(det) if want_marginals else None
    rootp = np.empty((S, H, W), dtype=dt)
    best = np.empty((S, P, 3), dtype=np.int32)
    score = np.empty((S, P), dtype=dt)
    d = (ctypes.c_int * P)(*map(int, is_detect))
    u = (ctypes.c_int * P)(*map(int, is_upright))
    J = joints_from(joints)
    _chk(getattr(lib(), "orc_root_posterior_rot_" + _sfx(dt))(
        ctypes.byref(ep), P, d, u, root, len(joints), J, H, W, _ptr(det),
        _ptr(post) if post is not None else None, _ptr(rootp), best.ctypes.data_as(_ip), _ptr(score),
        int(is_sparse)), "root_posterior_rot")
    return rootp, post, best, score


def find_local_max(grid, max_n):
    g = np.ascontiguousarray(grid, dtype=np.float32)
    D, H, W = g.shape
    out = np.empty((D * H * W, 4), dtype=np.float64)
    n = lib().orc_find_local_max_f32(_ptr(g), D, H, W, max_n, _ptr(out))
    return out[:n].copy()


def message_plan(ep, R, H, W, offset_c, offset_p, rot_mean, rot_sigma, scale):
    ints = {k: (ctypes.c_int * R)() for k in ("fox", "foy", "iox", "ioy")}
    delta, uniform, rot_len, kx, ky = (ctypes.c_int() for _ in range(5))
    taps = (ctypes.c_double * 1024)()
    oc = (ctypes.c_double * 2)(*offset_c)
    op = (ctypes.c_double * 2)(*offset_p)
    _chk(lib().orc_message_plan(ctypes.byref(ep), R, H, W, oc, op, rot_mean, rot_sigma, scale, ctypes.byref(delta),
                                ints["fox"], ints["foy"], ints["iox"], ints["ioy"], ctypes.byref(uniform),
                                ctypes.byref(rot_len), taps, ctypes.byref(kx), ctypes.byref(ky)), "message_plan")
    return dict(delta=delta.value, fwd_ox=list(ints["fox"]), fwd_oy=list(ints["foy"]), inv_ox=list(ints["iox"]),
                inv_oy=list(ints["ioy"]), uniform=bool(uniform.value), rot_taps=list(taps[:rot_len.value]))


# ------------------------------------------------------------ parity metrics

def normalized(logb):
    """p = exp(b - LSE(b)) in float64 over the whole grid."""
    b = np.asarray(logb, dtype=np.float64)
    m = b.max()
    e = np.exp(b - m)
    return e / e.sum()


def rel_l1(p_impl, p_ref):
    return float(np.abs(p_impl - p_ref).sum() / np.abs(p_ref).sum())


def top2_gap(logb):
    """Gap between the largest and second largest value of a grid."""
    b = np.asarray(logb, dtype=np.float64).ravel()
    if b.size < 2:
        return float("inf")
    i = int(np.argmax(b))
    m = b[i]
    b2 = np.delete(b, i)
    return float(m - b2.max())


# ------------------------------------------------ off-diagonal covariance (f1)

def _bind_f1():
    L = lib()
    if getattr(L, "_f1_bound", False):
        return L
    for t, p in (("f32", _fp), ("f64", _dp)):
        getattr(L, "orc_gauss_filter2d_" + t).argtypes = [p, p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]
    L.orc_eig2d.argtypes = [_dp, _dp, _dp]
    L.orc_transform_direct_f32.argtypes = [_fp, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int, _dp,
                                           ctypes.c_float]
    L.orc_transform_bilinear_f32.argtypes = [_fp, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int, _dp,
                                             ctypes.c_float]
    L.orc_ingest_scores_f32.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _fp]
    L._f1_bound = True
    return L


def eig2d(C):
    L = _bind_f1()
    c = np.ascontiguousarray(np.asarray(C, dtype=np.float64).reshape(4))
    V = np.empty(4)
    E = np.empty(4)
    L.orc_eig2d(_ptr(c), _ptr(V), _ptr(E))
    return V.reshape(2, 2), E.reshape(2, 2)


def gauss_filter2d(grid, C, is_sparse):
    """gaussFilter2d: diagonal C, or gaussFilter2dOffset in C's eigenbasis."""
    L = _bind_f1()
    g = np.ascontiguousarray(grid)
    out = np.empty_like(g)
    c = np.ascontiguousarray(np.asarray(C, dtype=np.float64).reshape(4))
    _chk(getattr(L, "orc_gauss_filter2d_" + _sfx(g.dtype))(_ptr(g), _ptr(out), g.shape[0], g.shape[1], _ptr(c),
                                                            int(is_sparse)), "gauss_filter2d")
    return out


def transform_direct(grid, T31, out_shape, default_value=0.0):
    L = _bind_f1()
    g = np.ascontiguousarray(grid, dtype=np.float32)
    out = np.empty(out_shape, dtype=np.float32)
    T = np.ascontiguousarray(np.asarray(T31, dtype=np.float64).reshape(9))
    L.orc_transform_direct_f32(_ptr(g), g.shape[0], g.shape[1], _ptr(out), out_shape[0], out_shape[1], _ptr(T),
                               default_value)
    return out


def transform_bilinear(grid, T21, out_shape, default_value=0.0):
    L = _bind_f1()
    g = np.ascontiguousarray(grid, dtype=np.float32)
    out = np.empty(out_shape, dtype=np.float32)
    T = np.ascontiguousarray(np.asarray(T21, dtype=np.float64).reshape(9))
    L.orc_transform_bilinear_f32(_ptr(g), g.shape[0], g.shape[1], _ptr(out), out_shape[0], out_shape[1], _ptr(T),
                                 default_value)
    return out


def ingest_scores(score_grid, Tig, interpolate, H, W):
    """loadScoreGrid's grid-to-image mapping, then clip_scores_fill + log."""
    L = _bind_f1()
    g = np.ascontiguousarray(score_grid, dtype=np.float32)
    n, gh, gw = g.shape
    T = np.ascontiguousarray(np.asarray(Tig, dtype=np.float64).reshape(n, 9))
    out = np.empty((n, H, W), dtype=np.float32)
    L.orc_ingest_scores_f32(_ptr(g), n, gh, gw, _ptr(T), int(interpolate), H, W, _ptr(out))
    return out


# ------------------------------------------------------- extra unaries (f3)

def torso_pos_prior(grids, params, weight):
    """setTorsoPosPrior on a copy of grids [..][H][W] (float32)."""
    g = np.array(grids, dtype=np.float32, copy=True, order="C")
    H, W = g.shape[-2:]
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).ravel()[:4])
    _chk(lib().orc_torso_pos_prior_f32(_ptr(g), g.size // (H * W), H, W, _ptr(p), ctypes.c_float(weight)),
         "torso_pos_prior")
    return g


def add_dpm_raw(grids, dpm, weight):
    """addLoadDPMScore's accumulation on a copy of one part's [S][R][H][W]."""
    g = np.array(grids, dtype=np.float32, copy=True, order="C")
    d = np.ascontiguousarray(dpm, dtype=np.float32).reshape(-1, *g.shape[-2:])
    S, R, H, W = g.shape
    _chk(lib().orc_add_dpm_raw_f32(_ptr(g), S, R, H, W, _ptr(d), d.shape[0], ctypes.c_float(weight)), "add_dpm_raw")
    return g


def add_dpm_log(grids, log_dpm, weight):
    """addDPMScore on a copy of one part's [S][R][H][W]."""
    g = np.array(grids, dtype=np.float32, copy=True, order="C")
    d = np.ascontiguousarray(log_dpm, dtype=np.float32).reshape(-1, *g.shape[-2:])
    S, R, H, W = g.shape
    _chk(lib().orc_add_dpm_log_f32(_ptr(g), S, R, H, W, _ptr(d), d.shape[0], ctypes.c_float(weight)), "add_dpm_log")
    return g


def log_grid(grid):
    """computeLogGrid on a copy (0 -> LOG_ZERO)."""
    g = np.array(grid, dtype=np.float32, copy=True, order="C")
    _chk(lib().orc_log_grid_f32(_ptr(g), ctypes.c_longlong(g.size)), "log_grid")
    return g


def add_extra_unary(grids, extra, weight):
    g = np.array(grids, dtype=np.float32, copy=True, order="C")
    e = np.ascontiguousarray(extra, dtype=np.float32)
    assert e.size == g.size
    _chk(lib().orc_add_extra_unary_f32(_ptr(g), _ptr(e), ctypes.c_longlong(g.size), ctypes.c_float(weight)),
         "add_extra_unary")
    return g
