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

    def keep_layout(self):
        """(per-instance byte offsets, total bytes) of the keep flags of a
        pruned execution: one byte per 64 x 64 tile, [P][S][R][ty][tx] per
        instance."""
        offs = (ctypes.c_longlong * self.n_instances)()
        total = ctypes.c_longlong()
        check(lib().ps_plan_keep_layout(self.handle, offs, ctypes.byref(total)), "ps_plan_keep_layout")
        return list(offs), total.value

    def execute_pruned(self, log_unary, log_marginal, root_posterior, map_dev, keep, fallback=None, stream=None):
        """Pruned sweep: tiles whose keep byte is 0 are not computed (LOG_ZERO);
        instances with a nonzero fallback flag (uint32 per instance) run the
        whole sweep instead."""
        check(lib().ps_plan_execute_pruned(self.handle, _ptr(log_unary), _ptr(log_marginal), _ptr(root_posterior),
                                           _ptr(map_dev), _ptr(keep), _ptr(fallback), _stream_handle(stream)),
              "ps_plan_execute_pruned")

    KERNEL_CLASSES = ("reset", "leaf_max", "rowpass", "colpass", "root", "map_decode")

    METHODS = {"fp32": 0, "bf16": 1, "bf16x2": 2, "bf16x3": 3, "bf16mix": 4}

    def set_method(self, method):
        """Arithmetic of the spatial passes: 'fp32' (the drop-in), 'bf16',
        'bf16x2', 'bf16x3' or 'bf16mix' (banded-Toeplitz GEMMs on the matrix
        cores; 'bf16mix' runs the kernels of 'bf16').  Returns
        (levels, levels that use the method)."""
        check(lib().ps_plan_set_method(self.handle, self.METHODS[method]), "ps_plan_set_method")
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().ps_plan_method_levels(self.handle, ctypes.byref(a), ctypes.byref(b)), "ps_plan_method_levels")
        return a.value, b.value

    # certificate thresholds of the MAP guard per method: above twice the
    # largest spread of the log-marginal error (relative to the MAP state's)
    # measured against the FP64 oracle near each part's maximum
    MAP_GUARD_TAU = {"fp32": 0.0, "bf16": 4e-3, "bf16mix": 4e-3, "bf16x2": 4e-3, "bf16x3": 1e-4}

    def set_map_guard(self, tau):
        """MAP guard of the BF16 methods: instances whose best log marginal
        of some part beats the second best by at most tau are swept again in
        FP32 on the device (0 turns it off)."""
        check(lib().ps_plan_set_map_guard(self.handle, float(tau)), "ps_plan_set_map_guard")

    def read_map_guard(self, stream=None):
        """(instances re-evaluated by the last execution, per-instance flags)."""
        n = ctypes.c_int()
        flags = (ctypes.c_uint * self.n_instances)()
        check(lib().ps_plan_read_map_guard(self.handle, ctypes.byref(n), flags, _stream_handle(stream)),
              "ps_plan_read_map_guard")
        return n.value, [int(f) for f in flags]

    TUNE_MFMA_ROWS = 1
    TUNE_FP32_ROW_THREADS = 2
    TUNE_FP32_COL_ROWS = 3

    def tune(self, knob, value, cls=None):
        """Set a launch-shape knob for every launch class, or for one."""
        if cls is None:
            check(lib().ps_plan_tune(self.handle, knob, int(value)), "ps_plan_tune")
        else:
            check(lib().ps_plan_tune_class(self.handle, knob, int(cls), int(value)), "ps_plan_tune_class")

    def classes(self):
        """Row-pass launch classes in execution order: dicts of level, jobs,
        tallest and widest grid."""
        n = ctypes.c_int()
        check(lib().ps_plan_classes(self.handle, ctypes.byref(n), None, 0), "ps_plan_classes")
        info = (ctypes.c_int * (4 * max(1, n.value)))()
        check(lib().ps_plan_classes(self.handle, ctypes.byref(n), info, n.value), "ps_plan_classes")
        return [dict(level=info[4 * c], jobs=info[4 * c + 1], max_h=info[4 * c + 2], max_w=info[4 * c + 3])
                for c in range(n.value)]

    def _sweep_ms(self, unary, marg, root, maps, s, reps):
        import torch
        with torch.cuda.stream(s):
            self.execute(unary, marg, root, maps, stream=s)  # first use of this shape
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record(s)
            for _ in range(reps):
                self.execute(unary, marg, root, maps, stream=s)
            t1.record(s)
        t1.synchronize()
        return t0.elapsed_time(t1) / reps

    def autotune(self, unary, marg, root, maps, method_is_fp32=False, reps=2, stream=None, max_classes=8):
        """Choose the launch shapes of the current method by timing whole
        sweeps on the caller's buffers (HIP events on the execution stream):
        the BF16 methods' rows per x-pass block (candidates 2, 4, 8, 16); for
        the FP32 method the row pass's threads per block (one per column,
        128, 192, 256), then the column pass's output rows per lane (8, 16).
        Per knob: first one value for every launch class, then class by class
        (coordinate descent: each class keeps the value that gave the fastest
        sweep with the others fixed; skipped above max_classes classes, where
        it would take more sweeps than it can save).  Sets the choices and
        returns {"knobs": [{"knob": name, "all": {candidate: ms},
        "per_class": [(class, value, ms)]}], "ms": final ms per sweep}."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        if method_is_fp32:
            knobs = [(self.TUNE_FP32_ROW_THREADS, "fp32_row_threads", (0, 128, 192, 256)),
                     (self.TUNE_FP32_COL_ROWS, "fp32_col_rows_per_lane", (8, 16))]
        else:
            knobs = [(self.TUNE_MFMA_ROWS, "mfma_rows_per_block", (2, 4, 8, 16))]
        classes = self.classes()
        out = []
        cur = None
        for knob, name, cands in knobs:
            times = {}
            for c in cands:
                self.tune(knob, c)
                times[c] = self._sweep_ms(unary, marg, root, maps, s, reps)
            best = min(times, key=times.get)
            self.tune(knob, best)
            cur = times[best]
            per_class = []
            if 1 < len(classes) <= max_classes:
                for k in range(len(classes)):
                    pick = best
                    for c in cands:
                        if c == best:
                            continue
                        self.tune(knob, c, cls=k)
                        t = self._sweep_ms(unary, marg, root, maps, s, reps)
                        if t < cur:
                            cur, pick = t, c
                    self.tune(knob, pick, cls=k)
                    per_class.append((k, pick, round(cur, 3)))
            out.append({"knob": name, "all": {str(k): round(v, 3) for k, v in times.items()},
                        "per_class": per_class})
        return {"knobs": out, "ms": round(cur, 3)}

    def set_profiling(self, enable=True, classes=None):
        """Bracket the launches with HIP events: every kernel class, or only
        those named in `classes` (a subset of KERNEL_CLASSES)."""
        if classes is None:
            check(lib().ps_plan_set_profiling(self.handle, int(bool(enable))), "ps_plan_set_profiling")
            return
        mask = 0
        for c in classes:
            mask |= 1 << self.KERNEL_CLASSES.index(c)
        check(lib().ps_plan_set_profiling_classes(self.handle, ctypes.c_uint(mask if enable else 0)),
              "ps_plan_set_profiling_classes")

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
                               scale, is_sparse=False, stream=None, out=None, no_alloc=False):
    """computeRotJointMarginal (objectdetect_findrot.cpp:292-456): child [R][H][W]
    (torch CUDA float32) -> parent message; is_sparse is bIsSparse.  `out`:
    the parent tensor to write (default: a new one).  no_alloc: run through
    ps_message_noalloc (no device allocation inside the library)."""
    import torch
    R, H, W = log_prob_child.shape
    child = log_prob_child.contiguous()
    parent = torch.empty_like(child) if out is None else out
    assert parent.is_contiguous() and parent.shape == child.shape and parent.dtype == torch.float32
    oc = (ctypes.c_double * 2)(*offset_c)
    op = (ctypes.c_double * 2)(*offset_p)
    cc = (ctypes.c_double * 4)(*C)
    fn, name = (lib().ps_message_noalloc, "ps_message_noalloc") if no_alloc else (lib().ps_message, "ps_message")
    check(fn(model.handle, child.data_ptr(), parent.data_ptr(), H, W, oc, op, cc, float(rot_mean), float(rot_sigma),
             float(scale), 1 if is_sparse else 0, _stream_handle(stream)), name)
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


# ------------------------------------------------------------ extra unaries

DPM_RAW, DPM_LOG, DPM_LOG_OF_RAW = 0, 1, 2


def _grid_dims(t, what):
    assert t.is_cuda and t.dtype.is_floating_point and t.is_contiguous(), what + ": a contiguous CUDA float32 tensor"
    return t


def set_torso_pos_prior(root_unary, pos_prior_params, weight, stream=None):
    """setTorsoPosPrior (objectdetect_icps.cpp:137-191) in place on the root
    part's log unaries [S][R][H][W] (or [n][H][W]); pos_prior_params =
    (mu_x, mu_y, var_x, var_y), the `params` row of torso_pos_prior.mat."""
    _grid_dims(root_unary, "root_unary")
    H, W = root_unary.shape[-2:]
    n = root_unary.numel() // (H * W)
    pp = (ctypes.c_double * 4)(*[float(v) for v in np.asarray(pos_prior_params, dtype=np.float64).ravel()[:4]])
    check(lib().ps_add_torso_pos_prior(root_unary.data_ptr(), n, H, W, pp, float(weight), _stream_handle(stream)),
          "ps_add_torso_pos_prior")
    return root_unary


def add_dpm_score(part_unary, dpm, weight, mode=DPM_RAW, stream=None):
    """DPM scores into one part's log unaries [S][R][H][W] from dpm [n][H][W]
    (CUDA float32): mode DPM_RAW = addLoadDPMScore's accumulation
    (objectdetect_icps.cpp:469-484), DPM_LOG = addDPMScore (:490-524) on log
    grids, DPM_LOG_OF_RAW = computeLogGrid then addDPMScore (the DPM torso
    path, objectdetect_findrot.cpp:896-903)."""
    _grid_dims(part_unary, "part_unary")
    _grid_dims(dpm, "dpm")
    S, R, H, W = part_unary.shape
    if dpm.dim() == 2:
        dpm = dpm.unsqueeze(0)
    assert tuple(dpm.shape[1:]) == (H, W)
    check(lib().ps_add_dpm_score(part_unary.data_ptr(), S, R, H, W, dpm.data_ptr(), dpm.shape[0], float(weight),
                                 int(mode), _stream_handle(stream)), "ps_add_dpm_score")
    return part_unary


def add_extra_unary(log_unary, extra, weight, stream=None):
    """addExtraUnary (objectdetect_icps.cpp:526-548): log_unary += weight * extra."""
    _grid_dims(log_unary, "log_unary")
    _grid_dims(extra, "extra")
    assert extra.numel() == log_unary.numel()
    check(lib().ps_add_extra_unary(log_unary.data_ptr(), extra.data_ptr(), log_unary.numel(), float(weight),
                                   _stream_handle(stream)), "ps_add_extra_unary")
    return log_unary


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


def synth_unary_dev(seed, P, S, R, H, W, out, stream=None):
    """ps_synth_unary on the device into `out`, a torch CUDA float32 tensor (or
    a view) of P*S*R*H*W contiguous elements."""
    assert out.is_contiguous() and out.numel() == P * S * R * H * W
    check(lib().ps_synth_unary_dev(int(seed), P, S, R, H, W, out.data_ptr(), _stream_handle(stream)),
          "ps_synth_unary_dev")
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


def uneven_sizes(n, seed=20131201):
    """Sizes of the uneven batch: H and W uniform in {160, 168, ..., 480}."""
    rng = np.random.default_rng(seed)
    return [(8 * int(rng.integers(20, 61)), 8 * int(rng.integers(20, 61))) for _ in range(n)]


# ------------------------------------------------------------ MAT-file (level 5)

class MatFile:
    """A level-5 MAT file read into memory (ps_mat_open): variables by name,
    cell-array elements by their column-major cell index, numeric data as
    float64 numpy arrays in C order (matlab_io.hpp:154-230,
    matlab_cell_io.hpp:27-130)."""

    def __init__(self, path):
        h = ctypes.c_void_p()
        check(lib().ps_mat_open(str(path).encode(), ctypes.byref(h)), "ps_mat_open")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().ps_mat_close(self.handle)
            self.handle = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _info(self, name, cell_index):
        is_cell, nd = ctypes.c_int(), ctypes.c_int()
        dims = (ctypes.c_longlong * 8)()
        check(lib().ps_mat_var(self.handle, name.encode(), cell_index, ctypes.byref(is_cell), ctypes.byref(nd),
                               dims, None, 0), "ps_mat_var")
        return bool(is_cell.value), tuple(dims[i] for i in range(nd.value))

    def shape(self, name, cell_index=-1):
        return self._info(name, cell_index)[1]

    def is_cell(self, name):
        return self._info(name, -1)[0]

    def get(self, name, cell_index=-1):
        """float64 array in C order (a cell array needs a cell index)."""
        _, shape = self._info(name, cell_index)
        out = np.empty(shape, dtype=np.float64)
        nd = ctypes.c_int()
        dims = (ctypes.c_longlong * 8)()
        check(lib().ps_mat_var(self.handle, name.encode(), cell_index, None, ctypes.byref(nd), dims,
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size), "ps_mat_var")
        return out

    def cells(self, name):
        """A 2-D cell array as nested lists [i1][i2] (matlab_cell_io.hpp:54-64)."""
        n1, n2 = self.shape(name)
        return [[self.get(name, i1 + n1 * i2) for i2 in range(n2)] for i1 in range(n1)]


def mat_write(path, name, array, append=False, single=None, compress=True):
    """Write one real array (C order in memory; stored column-major) as class
    double, or single for float32 data (matlab_io.hpp:64-129)."""
    a = np.ascontiguousarray(np.asarray(array, dtype=np.float64))
    if a.ndim == 0:
        a = a.reshape(1, 1)
    if single is None:
        single = np.asarray(array).dtype == np.float32
    dims = (ctypes.c_longlong * a.ndim)(*a.shape)
    check(lib().ps_mat_write(str(path).encode(), int(bool(append)), name.encode(),
                             a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.ndim, dims, int(bool(single)),
                             int(bool(compress))), "ps_mat_write")


JOINT_POS_GAUSSIAN, JOINT_ROT_GAUSSIAN = 1, 2


def load_joint_mat(path, flip=False):
    """load_joint + loadJoints' flip and 0-based ids for one joint_<c>_<p>.mat
    (objectdetect_learnparam.cpp:126-170, objectdetect_aux.cpp:96-136).
    Returns (Joint, type)."""
    j = ps_joint()
    t = ctypes.c_int()
    check(lib().ps_load_joint_mat(str(path).encode(), int(bool(flip)), ctypes.byref(j), ctypes.byref(t)),
          "ps_load_joint_mat")
    return Joint.from_c(j), t.value


def save_joint_mat(path, child_id, parent_id, offset_c, offset_p, C, rot_mean=0.0, rot_sigma=0.0,
                   jtype=JOINT_ROT_GAUSSIAN):
    """save_joint (objectdetect_learnparam.cpp:61-89): part ids (1-based),
    scalars as 1x1, offsets as 2x1, C as 2x2, compressed."""
    mat_write(path, "type", float(jtype))
    mat_write(path, "child_idx", float(child_id), append=True)
    mat_write(path, "parent_idx", float(parent_id), append=True)
    mat_write(path, "offset_c", np.asarray(offset_c, np.float64).reshape(2, 1), append=True)
    mat_write(path, "offset_p", np.asarray(offset_p, np.float64).reshape(2, 1), append=True)
    mat_write(path, "C", np.asarray(C, np.float64).reshape(2, 2), append=True)
    mat_write(path, "rot_mean", float(rot_mean), append=True)
    mat_write(path, "rot_sigma", float(rot_sigma), append=True)


def load_joints(spatial_dir, joint_ids, flip=False):
    """loadJoints (objectdetect_aux.cpp:54-141) for the (child_id, parent_id)
    part-id pairs of a PartConfig, from <spatial_dir>/joint_<c>_<p>.mat."""
    import os
    return [load_joint_mat(os.path.join(spatial_dir, "joint_%d_%d.mat" % (c, p)), flip)[0] for c, p in joint_ids]


def save_part_marginal(path, log_prob_grid):
    """The log_part_posterior_final_*.mat dump (objectdetect_findrot.cpp:239-253):
    variable log_prob_grid, single [R][H][W]."""
    mat_write(path, "log_prob_grid", np.asarray(log_prob_grid, np.float32), single=True)


def load_part_marginal(path):
    """PartApp::loadPartMarginal (partapp.cpp:807-827): float32 [R][H][W]."""
    with MatFile(path) as m:
        return m.get("log_prob_grid").astype(np.float32)


def load_score_grid(path):
    """The score-grid file of loadScoreGrid (partapp.cpp:829-857): the
    cell_scoregrid cells [scale][rotation] as float32 grids and the
    transform_Ti2 / transform_T2g matrices [scale][rotation][3][3]."""
    with MatFile(path) as m:
        grids = [[g.astype(np.float32) for g in row] for row in m.cells("cell_scoregrid")]
        return grids, m.get("transform_Ti2"), m.get("transform_T2g")
