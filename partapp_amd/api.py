"""Host-side interface of the gfx950 pictorial-structure sweep.

Mirrors the reference's libPictStruct entry points for this path:

* ``compute_root_posterior_rot`` -- computeRootPosteriorRot
  (objectdetect_findrot.cpp:470-727, objectdetect.h:269-274): per-part log
  marginals, the root posterior and ``best_part_hyp`` (the MAP of every part,
  taken from the last scale as the reference does, :257-259 inside :503).
* ``compute_rot_joint_marginal`` -- computeRotJointMarginal (:292-456).
* ``Plan`` -- a batch of instances (uneven sizes, per-image spatial models)
  executed as one sweep; ``ps_plan_execute`` underneath.

Device memory is handed over as torch CUDA tensors (or raw device pointers);
torch is only the allocator and stream provider here, every computation runs
in libpartapp_ps.so.
"""
import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import check, lib, ps_exp_param, ps_instance, ps_joint, ps_map, ps_model_desc

LOG_ZERO = -1e6  # boost_math.h:23


@dataclass
class ExpParam:
    """ExpParam fields on the path (ExpParam.proto:189-195,229), stored as float32/uint32."""
    min_part_rotation: float = -180.0
    max_part_rotation: float = 180.0
    num_rotation_steps: int = 48
    min_object_scale: float = 1.0
    max_object_scale: float = 1.0
    num_scale_steps: int = 1
    strip_border_detections: float = 0.0

    def to_c(self):
        return ps_exp_param(self.min_part_rotation, self.max_part_rotation, self.num_rotation_steps,
                            self.min_object_scale, self.max_object_scale, self.num_scale_steps,
                            self.strip_border_detections)

    @staticmethod
    def _value_from_index(minval, maxval, num_steps, idx):  # partapp_aux.hpp:44-55
        minval, maxval = float(np.float32(minval)), float(np.float32(maxval))
        if minval == maxval:
            return minval
        return minval + (maxval - minval) / num_steps * (0.5 + idx)

    def rot_from_index(self, idx):
        return self._value_from_index(self.min_part_rotation, self.max_part_rotation, self.num_rotation_steps, idx)

    def scale_from_index(self, idx):
        return self._value_from_index(self.min_object_scale, self.max_object_scale, self.num_scale_steps, idx)


@dataclass
class Joint:
    """object_detect::Joint after loadJoints (objectdetect.h:54-86): 0-based ids, flipped."""
    child_idx: int
    parent_idx: int
    offset_c: Tuple[float, float]
    offset_p: Tuple[float, float]
    C: Tuple[float, float, float, float]  # row-major 2x2 covariance
    rot_mean: float = 0.0
    rot_sigma: float = 0.0

    def to_c(self):
        j = ps_joint()
        j.child_idx, j.parent_idx = self.child_idx, self.parent_idx
        j.offset_c[:] = list(self.offset_c)
        j.offset_p[:] = list(self.offset_p)
        j.C[:] = list(self.C)
        j.rot_mean, j.rot_sigma = self.rot_mean, self.rot_sigma
        return j

    @staticmethod
    def from_c(j):
        return Joint(j.child_idx, j.parent_idx, tuple(j.offset_c), tuple(j.offset_p), tuple(j.C), j.rot_mean,
                     j.rot_sigma)


def joints_to_c(joints: Sequence[Joint]):
    arr = (ps_joint * len(joints))()
    for i, j in enumerate(joints):
        arr[i] = j.to_c()
    return arr


@dataclass
class PartHyp:
    """objectdetect.h:88-195; to_vect() is the 7-float best_conf row (:139-159)."""
    scaleidx: int
    rotidx: int
    x: int
    y: int
    score: float
    scale: float = 0.0
    rot: float = 0.0

    def to_vect(self):
        return np.array([self.scaleidx, self.scale, self.rotidx, self.rot, self.x, self.y, self.score],
                        dtype=np.float32)


class PartModel:
    """PartConfig flags + joints + ExpParam: an immutable ps_model."""

    def __init__(self, is_detect, is_upright, is_root, joints: Sequence[Joint], exp_param: ExpParam,
                 is_sparse=True):
        self.num_parts = len(is_detect)
        self.is_detect = list(map(int, is_detect))
        self.is_upright = list(map(int, is_upright))
        self.is_root = list(map(int, is_root))
        self.joints = list(joints)
        self.exp_param = exp_param
        self.is_sparse = bool(is_sparse)
        P = self.num_parts
        self._det = (ctypes.c_int * P)(*self.is_detect)
        self._up = (ctypes.c_int * P)(*self.is_upright)
        self._root = (ctypes.c_int * P)(*self.is_root)
        self._joints_c = joints_to_c(self.joints)
        desc = ps_model_desc(P, self._det, self._up, self._root, len(self.joints), self._joints_c,
                             exp_param.to_c(), int(self.is_sparse))
        handle = ctypes.c_void_p()
        check(lib().ps_model_create(ctypes.byref(desc), ctypes.byref(handle)), "ps_model_create")
        self.handle = handle

    @property
    def root(self):
        return [p for p in range(self.num_parts) if self.is_detect[p] and self.is_root[p]][0]

    def close(self):
        if getattr(self, "handle", None):
            lib().ps_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(t):
    """Device address of a torch tensor or an int; None -> NULL."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream_handle(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Plan:
    """A batch of instances, packed back to back: unary and marginal buffers
    hold [inst][P][S][R][H_i][W_i] floats, the root posterior [inst][S][H_i][W_i]."""

    def __init__(self, model: PartModel, sizes: Sequence[Tuple[int, int]],
                 joints: Optional[Sequence[Optional[Sequence[Joint]]]] = None):
        self.model = model
        ep = model.exp_param
        self.P, self.S, self.R = model.num_parts, ep.num_scale_steps, ep.num_rotation_steps
        self.sizes = [(int(h), int(w)) for h, w in sizes]
        n = len(self.sizes)
        self._inst = (ps_instance * n)()
        self._joint_arrays = []
        self.unary_offsets, self.root_offsets = [], []
        u = r = 0
        for i, (h, w) in enumerate(self.sizes):
            inst = self._inst[i]
            inst.height, inst.width = h, w
            inst.unary_offset = inst.marginal_offset = u
            inst.root_offset = r
            self.unary_offsets.append(u)
            self.root_offsets.append(r)
            if joints is not None and joints[i] is not None:
                arr = joints_to_c(joints[i])
                self._joint_arrays.append(arr)
                inst.joints = ctypes.cast(arr, ctypes.POINTER(ps_joint))
            u += self.P * self.S * self.R * h * w
            r += self.S * h * w
        self.unary_numel = u
        self.root_numel = r
        self.map_numel = n * self.S * self.P
        handle = ctypes.c_void_p()
        check(lib().ps_plan_create(model.handle, self._inst, n, ctypes.byref(handle)), "ps_plan_create")
        self.handle = handle

    @property
    def n_instances(self):
        return len(self.sizes)

    def scratch_bytes(self):
        return lib().ps_plan_scratch_bytes(self.handle)

    def execute(self, log_unary, log_marginal, root_posterior=None, map_dev=None, stream=None, graph=False):
        """Launch the sweep asynchronously on `stream` (default: torch's current stream)."""
        fn = lib().ps_plan_execute_graph if graph else lib().ps_plan_execute
        check(fn(self.handle, _ptr(log_unary), _ptr(log_marginal), _ptr(root_posterior), _ptr(map_dev),
                 _stream_handle(stream)), "ps_plan_execute")

    KERNEL_CLASSES = ("reset", "leaf_max", "rowpass", "colpass", "root", "map_decode")

    METHODS = {"fp32": 0, "bf16": 1, "bf16x2": 2, "bf16x3": 3, "bf16mix": 4}

    def set_method(self, method):
        """Arithmetic of the spatial passes: 'fp32' (the drop-in), 'bf16',
        'bf16x2', 'bf16x3' or 'bf16mix' (banded-Toeplitz GEMMs on the matrix
        cores; 'bf16mix' = x pass 'bf16', y pass 'bf16x3').  Returns
        (levels, levels that use the method)."""
        check(lib().ps_plan_set_method(self.handle, self.METHODS[method]), "ps_plan_set_method")
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().ps_plan_method_levels(self.handle, ctypes.byref(a), ctypes.byref(b)), "ps_plan_method_levels")
        return a.value, b.value

    def set_profiling(self, enable=True):
        check(lib().ps_plan_set_profiling(self.handle, int(bool(enable))), "ps_plan_set_profiling")

    def read_profile(self):
        """{kernel class: (summed ms, launches)} of the profiled executions since the last read."""
        ms = (ctypes.c_double * 6)()
        n = (ctypes.c_int * 6)()
        check(lib().ps_plan_read_profile(self.handle, ms, n), "ps_plan_read_profile")
        return {k: (ms[i], n[i]) for i, k in enumerate(self.KERNEL_CLASSES)}

    def instance_view(self, buf, i, kind="unary"):
        """[P][S][R][H][W] (kind 'unary'/'marginal') or [S][H][W] ('root') view of instance i."""
        h, w = self.sizes[i]
        if kind == "root":
            o = self.root_offsets[i]
            return buf[o:o + self.S * h * w].view(self.S, h, w)
        o = self.unary_offsets[i]
        return buf[o:o + self.P * self.S * self.R * h * w].view(self.P, self.S, self.R, h, w)

    def close(self):
        if getattr(self, "handle", None):
            lib().ps_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MAP_DTYPE = np.dtype([("rotidx", np.int32), ("y", np.int32), ("x", np.int32), ("score", np.float32)])


def decode_maps(map_tensor, n_inst, S, P):
    """ps_map records (a torch int32 tensor of 4*n words, or bytes) -> structured numpy [n][S][P]."""
    raw = map_tensor.cpu().numpy() if hasattr(map_tensor, "cpu") else np.asarray(map_tensor)
    return np.frombuffer(raw.tobytes(), dtype=MAP_DTYPE).reshape(n_inst, S, P)


def compute_root_posterior_rot(model: PartModel, log_part_detections, joints: Optional[Sequence[Joint]] = None,
                               stream=None):
    """computeRootPosteriorRot for one instance.

    log_part_detections: torch CUDA float32 [P][S][R][H][W] (not modified; the
    reference masks it in place and its caller restores it, :847, :1001).
    Returns (root_part_posterior [S][H][W], log_part_posterior [P][S][R][H][W],
    best_part_hyp [P] (last scale), maps [S][P] structured array)."""
    import torch
    P, S, R, H, W = log_part_detections.shape
    plan = Plan(model, [(H, W)], [joints] if joints is not None else None)
    try:
        dev = log_part_detections.device
        marg = torch.empty_like(log_part_detections)
        root = torch.empty((S, H, W), dtype=torch.float32, device=dev)
        maps = torch.empty((S * P * 4,), dtype=torch.int32, device=dev)
        plan.execute(log_part_detections.contiguous(), marg, root, maps, stream=stream)
        torch.cuda.current_stream().synchronize() if stream is None else stream.synchronize()
        m = decode_maps(maps, 1, S, P)[0]
    finally:
        plan.close()
    ep = model.exp_param
    last = S - 1
    best = [PartHyp(last, int(m[last, p]["rotidx"]), int(m[last, p]["x"]), int(m[last, p]["y"]),
                    float(m[last, p]["score"]), ep.scale_from_index(last), ep.rot_from_index(int(m[last, p]["rotidx"])))
            for p in range(P)]
    return root, marg, best, m


def compute_rot_joint_marginal(model: PartModel, log_prob_child, offset_c, offset_p, C, rot_mean, rot_sigma,
                               scale, is_sparse=False, stream=None):
    """computeRotJointMarginal (objectdetect_findrot.cpp:292-456): child [R][H][W]
    (torch CUDA float32) -> parent message; is_sparse is bIsSparse."""
    import torch
    R, H, W = log_prob_child.shape
    child = log_prob_child.contiguous()
    parent = torch.empty_like(child)
    oc = (ctypes.c_double * 2)(*offset_c)
    op = (ctypes.c_double * 2)(*offset_p)
    cc = (ctypes.c_double * 4)(*C)
    check(lib().ps_message(model.handle, child.data_ptr(), parent.data_ptr(), H, W, oc, op, cc, float(rot_mean),
                           float(rot_sigma), float(scale), 1 if is_sparse else 0, _stream_handle(stream)),
          "ps_message")
    return parent


LOCAL_MAX_DTYPE = np.dtype([("d", np.int32), ("x", np.int32), ("y", np.int32), ("score", np.float32)])


def find_local_max(grids, max_n=1000, stream=None):
    """findLocalMax (objectdetect_aux.cpp:193-251) of device grids.

    grids: a torch CUDA float32 tensor [D][H][W] or a list of them (each
    contiguous).  Returns, per grid, a structured array of (d, x, y, score):
    all local maxima in scan order when there are at most max_n, else the
    max_n largest in decreasing order."""
    import torch
    single = hasattr(grids, "shape") and len(grids.shape) == 3
    gl = [grids] if single else list(grids)
    base = gl[0].data_ptr()
    dev = gl[0].device
    # one address space: offsets relative to the lowest-addressed grid
    lo = min(g.data_ptr() for g in gl)
    offs = (ctypes.c_longlong * len(gl))(*[(g.data_ptr() - lo) // 4 for g in gl])
    dims = (ctypes.c_int * (3 * len(gl)))(*[v for g in gl for v in g.shape])
    out = torch.empty((len(gl) * max_n * 4,), dtype=torch.int32, device=dev)
    cnt = torch.empty((len(gl),), dtype=torch.int32, device=dev)
    check(lib().ps_find_local_max(lo, offs, dims, len(gl), int(max_n), out.data_ptr(), cnt.data_ptr(),
                                  _stream_handle(stream)), "ps_find_local_max")
    del base
    rec = np.frombuffer(out.cpu().numpy().tobytes(), dtype=LOCAL_MAX_DTYPE).reshape(len(gl), max_n)
    n = cnt.cpu().numpy()
    res = [rec[i, :n[i]].copy() for i in range(len(gl))]
    return res[0] if single else res


def ingest_scores(score_grid, Tig, interpolate, height, width, stream=None):
    """PartApp::loadScoreGrid + clip_scores_fill + computeLogGrid
    (partapp.cpp:866-899, objectdetect_findrot.cpp:834-845) on the device.

    score_grid: torch CUDA float32 [n][gh][gw]; Tig: [n][3][3] (image = Tig *
    grid).  Returns the log unaries, torch CUDA float32 [n][height][width]."""
    import torch
    n, gh, gw = score_grid.shape
    T = np.ascontiguousarray(np.asarray(Tig, dtype=np.float64).reshape(n, 9))
    out = torch.empty((n, height, width), dtype=torch.float32, device=score_grid.device)
    src = score_grid.contiguous()
    check(lib().ps_ingest_scores(src.data_ptr(), n, gh, gw, T.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 1 if interpolate else 0, int(height), int(width), out.data_ptr(),
                                 _stream_handle(stream)), "ps_ingest_scores")
    return out


# ----------------------------------------------------------- synthetic data

PERSON_ROOT = 4


def synth_joints(seed, zero_rot_mean=False) -> List[Joint]:
    """Per-image spatial model of the 10-part person tree (ps_synth_joints)."""
    arr = (ps_joint * 9)()
    n = ctypes.c_int()
    check(lib().ps_synth_joints(int(seed), 0, int(bool(zero_rot_mean)), arr, ctypes.byref(n)), "ps_synth_joints")
    return [Joint.from_c(arr[i]) for i in range(n.value)]


def synth_unary(seed, P, S, R, H, W, threads=8):
    """Synthetic log unaries [P][S][R][H][W] float32 (ps_synth_unary)."""
    out = np.empty((P, S, R, H, W), dtype=np.float32)
    check(lib().ps_synth_unary(int(seed), P, S, R, H, W, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               int(threads)), "ps_synth_unary")
    return out


def person_model(exp_param: ExpParam, joints: Optional[Sequence[Joint]] = None, is_detect=None, is_upright=None,
                 is_sparse=True, seed=0) -> PartModel:
    """The 10-part person tree rooted at the torso (part 4)."""
    P = 10
    if joints is None:
        joints = synth_joints(seed)
    is_root = [1 if p == PERSON_ROOT else 0 for p in range(P)]
    return PartModel(is_detect if is_detect is not None else [1] * P,
                     is_upright if is_upright is not None else [0] * P, is_root, joints, exp_param, is_sparse)


def config_exp_param(cfg: int) -> ExpParam:
    """ExpParam of the benchmark configurations: 24 rotations over 360 degrees,
    one scale (cfg 0/1) or scales 0.8/1.0/1.2 (cfg 2+)."""
    if cfg in (0, 1):
        return ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    return ExpParam(-180.0, 180.0, 24, 0.7, 1.3, 3, 0.0)
