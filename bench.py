"""Benchmark of the pictorial-structure sweep on one MI355X.

Workload (the metric's configuration): 32 instances, each 300 x 300 positions
x 24 rotations x 1 scale, the 10-part person tree, synthetic log unaries and a
per-image spatial model (seeds 0..31).  One step = the whole sweep of the
batch (upward pass, root, downward pass, marginals of all parts, MAP, root
posterior) with the unaries already resident in HBM.

The metric compares the chosen method (the x, rotation and y filters as
banded-Toeplitz / circulant GEMMs on the matrix cores: the x filter bf16 x
bf16, the rotation filter with data and taps split into bf16 hi + lo, an fp16
intermediate, the y filter on fp16 data with split taps) with the plain FP32
sweep (the drop-in) on the same plan and inputs, in the same run; `value` is
the chosen method's images/s.  Its error against the FP64 oracle is measured
live on a sample of the instances (relative L1 of the normalised marginals
per part, MAP agreement).

Prints one JSON line.  `roofline` is measured live over the timed steps with
HIP events around each launch of the dominant kernel (on the stream the
kernels run on): its algorithmic HBM bytes per launch (the grids it must read
and write) over its average launch time, against the published HBM peak; a
separately profiled pass gives the per-kernel breakdown, and the issued MFMA
flops of the GEMMs over the BF16 dense peak are reported beside it.  `cpu_baseline` times the oracle (the float32 CPU
restatement of the reference) on a bounded sample of the same instances, one
instance per thread.
"""
import argparse
import concurrent.futures as cf
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "images/sec of 10-part marginals @300x300x24rot; chosen vs plain FP32 GPU at <=1e-3 L1"
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1, choices=[1, 2],
                    help="1: 32 x 300x300x24, 1 scale (the metric's configuration); 2: 256 instances of uneven "
                         "size (H, W in 160..480, multiples of 8) x 24 rotations x 3 scales")
    ap.add_argument("--instances", type=int, default=None, help="default 32 (config 1) / 256 (config 2)")
    ap.add_argument("--size", type=int, default=300, help="H = W of config 1")
    ap.add_argument("--rotations", type=int, default=24,
                    help="rotation steps over 360 degrees (24 in the metric; 48 is partapp's default)")
    ap.add_argument("--method", default="bf16mix", choices=["bf16mix", "bf16x3", "bf16x2", "bf16", "fp32"],
                    help="the chosen method (value); plain FP32 is timed beside it")
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 comparison")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--check", type=int, default=None,
                    help="instances checked against the oracles (default: all of config 1, 16 of config 2)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-guard", action="store_true", help="skip the MAP-guarded run of the chosen method")
    ap.add_argument("--no-prune", action="store_true", help="skip the learned-pruning run of the chosen method")
    ap.add_argument("--graph", action="store_true", help="time HIP-graph replays (no per-kernel roofline)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device copies of 256 MiB (the known-size kernel that scales the rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def message_plans(plan_sizes, joints, ep):
    """(H, W, rot taps, kx, ky) of every message of the sweep, both directions."""
    from partapp_amd import _lib
    L = _lib.lib()
    R = ep.num_rotation_steps
    out = []
    for (H, W), js in zip(plan_sizes, joints):
        for s in range(ep.num_scale_steps):
            scale = ep.scale_from_index(s)
            for j in js:
                for oc, op, rm in ((j.offset_c, j.offset_p, j.rot_mean), (j.offset_p, j.offset_c, -j.rot_mean)):
                    ints = [(ctypes.c_int * R)() for _ in range(4)]
                    d, rl, kx, ky = (ctypes.c_int() for _ in range(4))
                    st = L.ps_debug_message_plan(ctypes.byref(ep.to_c()), H, W, (ctypes.c_double * 2)(*oc),
                                                 (ctypes.c_double * 2)(*op), (ctypes.c_double * 4)(*j.C), rm,
                                                 j.rot_sigma, scale, ctypes.byref(d), ints[0], ints[1], ints[2],
                                                 ints[3], ctypes.byref(rl), None, ctypes.byref(kx),
                                                 ctypes.byref(ky))
                    assert st == 0
                    out.append((H, W, rl.value, kx.value, ky.value))
    return out


def workload_flops(msgs, R):
    """Algorithmic FP32 flops per kernel class, per sweep: 2 * states * taps.
    Row pass: rotation + x taps; column pass: y taps (one message each)."""
    row = sum(2.0 * R * H * W * (rl + 2 * kx + 1) for H, W, rl, kx, ky in msgs)
    col = sum(2.0 * R * H * W * (2 * ky + 1) for H, W, rl, kx, ky in msgs)
    return row, col


def mfma_flops(msgs, R, row_terms, col_products):
    """Issued MFMA flops of the Toeplitz GEMMs per sweep (32x32x16 tiles, K
    rounded up to steps of 16, R = 24 rotation rows padded to 32): the row
    pass's x filter with row_terms products per k-step, then the rotation
    filter as a 2-step GEMM per tile with data and taps split (3 products);
    the column pass with col_products per k-step (fp16 data times split fp16
    taps: 2; the bf16 three-term pass: 3)."""
    tile = 2.0 * 32 * 32 * 16
    fl = 0.0
    for H, W, rl, kx, ky in msgs:
        kkx, kky = (32 + 2 * kx + 15) // 16, (32 + 2 * ky + 15) // 16
        rot = 2 * 3
        fl += tile * (row_terms * kkx + rot) * ((W + 31) // 32) * H  # row pass: one 32-row tile per output row
        fl += tile * col_products * kky * 4 * R * ((W + 63) // 64) * ((H + 63) // 64)  # column pass: 4 waves / 64x64
    return fl


def grid_traffic(model, t_bytes=4):
    """Grids (R*H*W floats, 4 bytes per state) each kernel class must read or
    write per sweep and instance, given the row pass / column pass split: row
    pass = source read + row-filtered grid write; column pass = row-filtered
    grid read, the destination write(s) and the unary / stored belief its
    epilogue adds.  The row-filtered grid counts t_bytes / 4 of a grid (the
    BF16 methods with a 1- or 2-term row pass keep it in fp16)."""
    tg = t_bytes / 4.0
    P, root = model.num_parts, model.root
    parent = {j.child_idx: j.parent_idx for j in model.joints}
    children = {p: [c for c, q in parent.items() if q == p] for p in range(P)}
    row = col = 0
    for c, p in parent.items():
        leaf = not children[c]
        # upward message c -> p
        row += (0 if (leaf and not model.is_detect[c]) else 1) + tg
        col += tg + 1 + (1 if (p != root and model.is_detect[p]) else 0)
        # downward message p -> c
        row += 1 + tg
        # (grid read, marginal write, unary read; with a grandchild also the
        # stored belief read and the next source write)
        col += tg + 1 + 1 + (2 if children[c] else 0)
    k = len(children[root])
    root_grids = k + 1 + 1 + k  # messages + unary read, marginal + log_from_root writes
    leaf_grids = sum(1 for c in range(P) if c != root and not children[c] and model.is_detect[c])
    return {"rowpass": row, "colpass": col, "root": root_grids, "leaf_max": leaf_grids}


def device_copy_gbs(nbytes=1 << 30, reps=5):
    """Bandwidth of a plain device-to-device copy (read + write bytes / s),
    the practical HBM ceiling the sweep's kernels are compared with."""
    import torch
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    y.copy_(x)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for _ in range(reps):
        y.copy_(x)
    t1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (t0.elapsed_time(t1) / 1e3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    return round(gbs, 1)


def measured_traffic(method, kernel):
    """HBM bytes per launch of a kernel class measured by rocprofv3 (FETCH_SIZE
    and WRITE_SIZE, scaled by the same counters on a copy of known size), read
    from the summary committed under profiles/ for this workload, or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "traffic_%s_%s_32x300x300x24.txt" % (kernel, method))
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("bytes_per_launch"):
                    return int(float(line.split()[1])), os.path.relpath(path, os.path.dirname(path) + "/..")
    except OSError:
        pass
    return None, None


def host_threads():
    """(threads to use, host cores visible to this process): every core the
    process may run on (its affinity mask), capped by OMP_NUM_THREADS where
    the host sets it (the CPU share given to one GPU)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    n = min(visible, int(cap)) if cap.isdigit() and int(cap) > 0 else visible
    return max(1, n), visible


def cpu_baseline_timed(ep, sizes, joints, seeds, budget_s):
    """The float32 oracle on batches of the instances (cycling through them)
    until the time budget is spent, one instance per thread on the host cores;
    input generation is not timed."""
    import oracle
    from partapp_amd.api import synth_unary
    cores, _ = host_threads()
    oep = oracle.exp_param_from(ep)
    images = 0
    busy = 0.0
    i = 0
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        while busy < budget_s:
            batch = [(i + k) % len(sizes) for k in range(min(cores, len(sizes)))]
            us = [synth_unary(seeds[k], 10, ep.num_scale_steps, ep.num_rotation_steps, sizes[k][0], sizes[k][1],
                              threads=cores) for k in batch]
            tb = time.perf_counter()
            futs = [ex.submit(oracle.root_posterior_rot, oep, [1] * 10, [0] * 10, 4, joints[k], u, True, True)
                    for k, u in zip(batch, us)]
            for f in futs:
                f.result()
            busy += time.perf_counter() - tb
            images += len(batch)
            i = (i + len(batch)) % len(sizes)
    return images, busy, cores


def time_method(plan, method, unary, marg, root, maps, stream, args, steps=None, warmup=None, classes=(),
                guard=0.0):
    """Warmup, then EXACTLY `steps` (default args.steps) timed steps, the
    launches of the kernel classes in `classes` bracketed by HIP events on the
    stream they run on; `guard`: the MAP guard's threshold (0: off); returns
    (ms, profile or None)."""
    import torch
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    profile = bool(classes) and not args.graph
    plan.set_method(method)
    plan.set_map_guard(guard)
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            plan.execute(unary, marg, root, maps, stream=stream, graph=args.graph)
        stream.synchronize()
        if profile:
            plan.set_profiling(True, classes=classes)
            plan.read_profile()  # clear
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record(stream)
        for _ in range(steps):
            plan.execute(unary, marg, root, maps, stream=stream, graph=args.graph)
        t1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
    prof = None
    if profile:
        prof = plan.read_profile()
        plan.set_profiling(False)
    return t0.elapsed_time(t1), prof


def check_instances(plan, model, unary, legs, ep, joints, idx, pruned=None):
    """Error of each leg's last sweep on the instances `idx`, with the very
    unaries the GPU read.  legs: {name: (marginal tensor, map tensor)}; the
    leg named "fp32" is held to the FP32 bar against the float32 oracle (the
    CPU restatement of the reference: relative L1 of the normalised marginals
    per part <= 1e-5, identical MAP states), every leg is measured against
    the FP64 oracle (relative L1, MAP agreement, and for every MAP that
    differs the FP64 top-2 log gap, i.e. whether it is a near tie)."""
    import numpy as np

    import oracle
    from partapp_amd.api import decode_maps
    oep = oracle.exp_param_from(ep)
    maps = {k: decode_maps(m, plan.n_instances, plan.S, plan.P) for k, (_, m) in legs.items()}
    cores, _ = host_threads()
    lock = __import__("threading").Lock()

    def one(i):
        with lock:
            u = plan.instance_view(unary, i).cpu().numpy()
            mg = {k: plan.instance_view(g, i).cpu().numpy() for k, (g, _) in legs.items()}
        _, p64, b64, _ = oracle.root_posterior_rot(oep, model.is_detect, model.is_upright, model.root, joints[i],
                                                   u.astype(np.float64))
        ref32 = None
        if "fp32" in legs:
            _, p32, b32, _ = oracle.root_posterior_rot(oep, model.is_detect, model.is_upright, model.root,
                                                       joints[i], u)
        out = {k: {"l1_64": 0.0, "l1_32": 0.0, "agree64": 0, "agree32": 0, "ties": [], "far": [], "drop": 0.0}
               for k in legs}
        if pruned is not None:  # the FP64 mass of the pruned leg's dropped tiles (none for a fallback instance)
            name, keep_h, offs, fb_h = pruned
            H, W = plan.sizes[i]
            ty, tx = (H + 63) // 64, (W + 63) // 64
            kp = keep_h[offs[i]:offs[i] + plan.P * plan.S * plan.R * ty * tx].reshape(plan.P, plan.S, plan.R, ty, tx)
        for s in range(plan.S):
            for p in range(plan.P):
                n64 = oracle.normalized(p64[p, s])
                if pruned is not None and not fb_h[i]:
                    kf = np.repeat(np.repeat(kp[p, s], 64, axis=1), 64, axis=2)[:, :H, :W].astype(bool)
                    out[name]["drop"] += float(n64[~kf].sum())
                n32 = oracle.normalized(p32[p, s]) if "fp32" in legs else None
                for k in legs:
                    r = out[k]
                    nk = oracle.normalized(mg[k][p, s])
                    r["l1_64"] = max(r["l1_64"], oracle.rel_l1(nk, n64))
                    got = (int(maps[k][i, s, p]["rotidx"]), int(maps[k][i, s, p]["y"]), int(maps[k][i, s, p]["x"]))
                    if got == tuple(b64[s, p]):
                        r["agree64"] += 1
                    else:
                        gap = oracle.top2_gap(p64[p, s])
                        (r["ties"] if gap < 1e-3 else r["far"]).append((i, s, p, gap))
                    if k == "fp32":
                        r["l1_32"] = max(r["l1_32"], oracle.rel_l1(nk, n32))
                        r["agree32"] += int(got == tuple(b32[s, p]))
        return out

    with cf.ThreadPoolExecutor(max_workers=max(1, min(len(idx), cores))) as ex:
        per = list(ex.map(one, idx))
    total = len(idx) * plan.S * plan.P
    res = {}
    for k in legs:
        ties = [t for r in per for t in r[k]["ties"]]
        far = [t for r in per for t in r[k]["far"]]
        res[k] = {"max_rel_l1_vs_fp64": float("%.3g" % max(r[k]["l1_64"] for r in per)),
                  "map_agreement_vs_fp64": "%d/%d" % (sum(r[k]["agree64"] for r in per), total),
                  "map_mismatch_near_tie": len(ties), "map_mismatch_other": len(far),
                  "map_mismatches": [[int(a), int(b), int(c), float("%.3g" % g)] for a, b, c, g in (ties + far)[:8]],
                  "sample": "%d of %d instances (%d (instance, scale, part) MAPs)" % (len(idx), plan.n_instances,
                                                                                     total)}
        if pruned is not None and k == pruned[0]:
            drops = [r[k]["drop"] for r in per]
            res[k]["fp64_mass_of_dropped_tiles"] = {"max": float("%.3g" % max(drops)),
                                                    "median": float("%.3g" % float(np.median(drops)))}
            res[k]["bar"] = "rel L1 <= 1e-3 vs the FP64 oracle"
        if k == "fp32":
            res[k]["max_rel_l1_vs_f32_oracle"] = float("%.3g" % max(r[k]["l1_32"] for r in per))
            res[k]["map_identical_vs_f32_oracle"] = "%d/%d" % (sum(r[k]["agree32"] for r in per), total)
            res[k]["bar"] = "rel L1 <= 1e-5 vs the float32 oracle, MAP identical"
            res[k]["pass"] = bool(res[k]["max_rel_l1_vs_f32_oracle"] <= 1e-5 and
                                  sum(r[k]["agree32"] for r in per) == total)
        else:
            res[k]["bar"] = "rel L1 <= 1e-3 vs the FP64 oracle"
            res[k]["pass"] = bool(res[k]["max_rel_l1_vs_fp64"] <= 1e-3)
    return res


def calibrate_copy():
    import torch
    n = 64 * 1024 * 1024  # floats: 256 MiB
    x = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                      "launches": 3}), flush=True)


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: instances shard by image across GPUs with no exchange, "
              "and this benchmark measures one GPU's share (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from partapp_amd.api import (ExpParam, Plan, config_exp_param, person_model, synth_joints, synth_unary,
                                 uneven_sizes)

    from partapp_amd.api import synth_unary_dev

    dev = torch.device("cuda:0")
    ep = config_exp_param(args.config)
    if args.rotations != 24:
        ep = ExpParam(ep.min_part_rotation, ep.max_part_rotation, args.rotations, ep.min_object_scale,
                      ep.max_object_scale, ep.num_scale_steps, ep.strip_border_detections)
    R, S = ep.num_rotation_steps, ep.num_scale_steps
    if args.config == 1:
        n = args.instances or 32
        sizes = [(args.size, args.size)] * n
    else:
        n = args.instances or 256
        sizes = uneven_sizes(n)
        if args.check is None:
            args.check = 16
    H, W = sizes[0]
    seeds = list(range(n))
    joints = [synth_joints(s) for s in seeds]
    model = person_model(ep, joints[0])
    plan = Plan(model, sizes, joints)
    states = sum(S * R * h * w for h, w in sizes)
    print("plan: %d instances, %d scales, %.1f M states per part, scratch %.2f GB" % (
        n, S, states / 1e6, plan.scratch_bytes() / 1e9), file=sys.stderr, flush=True)
    unary = torch.empty((plan.unary_numel,), dtype=torch.float32, device=dev)
    threads, _ = host_threads()
    for i, s in enumerate(seeds):
        h, w = sizes[i]
        o = plan.unary_offsets[i]
        numel = 10 * S * R * h * w
        if args.config == 1:
            u = synth_unary(s, 10, S, R, h, w, threads=threads)
            unary[o:o + numel].copy_(torch.from_numpy(u.ravel()))
        else:
            synth_unary_dev(s, 10, S, R, h, w, unary[o:o + numel])
    marg = torch.empty_like(unary)
    root = torch.empty((plan.root_numel,), dtype=torch.float32, device=dev)
    maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    print("inputs resident: %.2f GB" % (unary.numel() * 4 / 1e9), file=sys.stderr, flush=True)
    copy_gbs = device_copy_gbs()
    print("device copy: %.1f GB/s" % copy_gbs, file=sys.stderr, flush=True)
    stream = torch.cuda.Stream() if args.graph else torch.cuda.current_stream()

    fp32 = None
    legs = {}
    tuned = {}
    if not args.no_fp32 and args.method != "fp32":
        # the FP32 leg writes its own outputs, so both legs can be checked (the
        # 256-instance batch has no room for a second 75 GB output: there the
        # FP32 leg is timed on the chosen leg's buffers and not checked)
        own = args.config == 1
        marg32 = torch.empty_like(unary) if own else marg
        maps32 = torch.zeros_like(maps) if own else maps
        if not args.no_autotune:
            plan.set_method("fp32")
            tuned["fp32"] = plan.autotune(unary, marg32, root, maps32, method_is_fp32=True, stream=stream)
            print("autotune fp32: %s" % tuned["fp32"], file=sys.stderr, flush=True)
        ms32, _ = time_method(plan, "fp32", unary, marg32, root, maps32, stream, args)
        if own:
            legs["fp32"] = (marg32, maps32)
        fp32 = {"value": round(n * args.steps / (ms32 / 1e3), 3), "ms_per_step": round(ms32 / args.steps, 4)}
        print("plain fp32: %.3f ms/step" % (ms32 / args.steps), file=sys.stderr, flush=True)

    if not args.no_autotune:
        plan.set_method(args.method)
        tuned[args.method] = plan.autotune(unary, marg, root, maps, method_is_fp32=args.method == "fp32",
                                           stream=stream)
        print("autotune %s: %s" % (args.method, tuned[args.method]), file=sys.stderr, flush=True)
    # per-kernel breakdown from a separate profiled pass (every class bracketed
    # by events), then the timed steps with events around the dominant kernel
    # only, so that the events barely perturb the measured throughput
    brk_steps = min(args.steps, 3)
    _, brk = time_method(plan, args.method, unary, marg, root, maps, stream, args, steps=brk_steps, warmup=1,
                         classes=Plan.KERNEL_CLASSES)
    dom = max(("rowpass", "colpass", "root"), key=lambda k: brk[k][0]) if brk else None
    ms, prof = time_method(plan, args.method, unary, marg, root, maps, stream, args, classes=(dom,) if dom else ())
    ms_per_step = ms / args.steps
    value = n * args.steps / (ms / 1e3)
    print("%s: timed %d steps, %.3f ms/step" % (args.method, args.steps, ms_per_step), file=sys.stderr, flush=True)

    # the same method with the MAP guard: instances whose MAP certificate
    # fails are swept again in FP32 within the step
    guarded = None
    early = {}  # checks made before a later leg reuses the chosen leg's buffers
    k_check = n if args.check is None else min(args.check, n)
    tau = Plan.MAP_GUARD_TAU.get(args.method, 0.0)
    if tau > 0 and not args.no_guard:
        if args.config == 1:
            margg = torch.empty_like(unary)
            mapsg = torch.zeros_like(maps)
        else:
            # no room for another 75 GB output: check the unguarded leg now,
            # then let the guarded leg overwrite its buffers
            if k_check > 0:
                early = check_instances(plan, model, unary, {args.method: (marg, maps)}, ep, joints,
                                        list(range(k_check)))
            margg, mapsg = marg, maps
        msg_, _ = time_method(plan, args.method, unary, margg, root, mapsg, stream, args, guard=tau)
        n_re, _ = plan.read_map_guard(stream)
        plan.set_map_guard(0.0)
        guarded = {"tau": tau, "value": round(n * args.steps / (msg_ / 1e3), 3),
                   "ms_per_step": round(msg_ / args.steps, 4), "reevaluated_in_fp32": "%d/%d" % (n_re, n)}
        legs[args.method + "+guard"] = (margg, mapsg)
        print("%s with the MAP guard (tau %g): %.3f ms/step, %d of %d instances re-evaluated" % (
            args.method, tau, msg_ / args.steps, n_re, n), file=sys.stderr, flush=True)

    # learned pruning: the net's keep masks (and fallback flags) recomputed
    # from the unaries inside every timed step, then the pruned sweep
    pruned, prune_info = None, None
    if not args.no_prune and args.method in ("bf16", "bf16x2", "bf16mix"):
        from partapp_amd import prune
        if os.path.exists(prune.WEIGHTS):
            pr = prune.Pruner(plan, dev)
            margp = torch.empty_like(unary)
            mapsp = torch.zeros_like(maps)
            plan.set_method(args.method)
            plan.set_map_guard(0.0)
            with torch.cuda.stream(stream):
                for _ in range(args.warmup):
                    pr.update(unary)
                    plan.execute_pruned(unary, margp, root, mapsp, pr.keep, pr.fallback, stream=stream)
                torch.cuda.synchronize()
                t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0.record(stream)
                for _ in range(args.steps):
                    pr.update(unary)
                    plan.execute_pruned(unary, margp, root, mapsp, pr.keep, pr.fallback, stream=stream)
                t1.record(stream)
                torch.cuda.synchronize()
            msp = t0.elapsed_time(t1)
            # the net and the keep rule alone (features, convolutions, flags)
            with torch.cuda.stream(stream):
                t2, t3 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t2.record(stream)
                for _ in range(args.steps):
                    pr.update(unary)
                t3.record(stream)
                torch.cuda.synchronize()
            ms_net = t2.elapsed_time(t3) / args.steps
            fb = pr.fallback.cpu().numpy()
            pruned = {"value": round(n * args.steps / (msp / 1e3), 3), "ms_per_step": round(msp / args.steps, 4),
                      "net_and_masks_ms_per_step": round(ms_net, 4),
                      "kept_tiles": round(pr.kept_fraction(), 4), "fallback_rate": "%d/%d" % (int(fb.sum()), n),
                      "predicted_dropped_mass_max": float("%.3g" % float(pr.predicted_dropped.max())),
                      "net": "3x3 convs 240->64->32->32, 1x1 -> 240, BF16 weights, on 4x max-pooled unaries; "
                             "per part, 64x64 tiles kept in decreasing predicted mass until at most 1e-4 is left "
                             "out; fall back to the exact sweep above 1e-3 predicted dropped mass per instance"}
            legs[args.method + "+prune"] = (margp, mapsp)
            prune_info = (args.method + "+prune", pr.keep.cpu().numpy(), pr.offsets, fb)
            print("%s with learned pruning: %.3f ms/step, kept tiles %.3f, fallback %d/%d" % (
                args.method, msp / args.steps, pr.kept_fraction(), int(fb.sum()), n), file=sys.stderr, flush=True)

    err = None
    k = k_check
    if k > 0:
        if args.method not in early:
            legs[args.method] = (marg, maps)
        tc = time.perf_counter()
        chk = check_instances(plan, model, unary, legs, ep, joints, list(range(k)), pruned=prune_info)
        chk.update(early)
        err = chk[args.method]
        if fp32 is not None and "fp32" in chk:
            fp32["parity"] = chk["fp32"]
        if guarded is not None:
            guarded["error_vs_fp64"] = chk[args.method + "+guard"]
        if pruned is not None:
            pruned["error_vs_fp64"] = chk[args.method + "+prune"]
        print("checked %d instances in %.1f s: %s" % (k, time.perf_counter() - tc, json.dumps(chk)), file=sys.stderr,
              flush=True)

    msgs = message_plans(sizes, joints, ep)
    roofline = None
    if prof is not None:
        print("breakdown over %d profiled sweeps:" % brk_steps, file=sys.stderr)
        for k, (t, c) in brk.items():
            print("  %-10s %9.3f ms total, %4d launches" % (k, t, c), file=sys.stderr)
        t_dom, n_dom = prof[dom]
        print("timed steps: %s %.3f ms total, %d launches" % (dom, t_dom, n_dom), file=sys.stderr)
        traffic_grids = grid_traffic(model, 2 if args.method in ("bf16", "bf16x2", "bf16mix") else 4)
        bytes_per_step = traffic_grids[dom] * 4.0 * states
        launches_per_step = n_dom / args.steps
        avg_ms = t_dom / n_dom
        achieved = bytes_per_step / launches_per_step / (avg_ms / 1e3) / 1e9
        traffic, tsrc = (measured_traffic(args.method, dom)
                         if args.config == 1 and n == 32 and args.size == 300 and R == 24 else (None, None))
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "algorithmic_bytes_per_launch": round(bytes_per_step / launches_per_step),
                    "avg_launch_ms": round(avg_ms, 4),
                    "kernel_ms_per_step": {k: round(v[0] / brk_steps, 4) for k, v in brk.items()},
                    "kernel_ms_source": "%d separately profiled sweeps (every launch bracketed by events); "
                                        "avg_launch_ms from the timed steps" % brk_steps}
        if tsrc:
            roofline["traffic_source"] = tsrc
        if copy_gbs:
            roofline["copy_gbs"] = copy_gbs
            roofline["frac_of_copy"] = round(achieved / copy_gbs, 4)
        # the whole sweep against the algorithmic minimum B_alg = 4 B * N_s * (2P + 4E)
        b_alg = 4.0 * states * (2 * 10 + 4 * 9)
        roofline["sweep_gbs"] = round(b_alg / (ms_per_step / 1e3) / 1e9, 1)
        if args.method == "fp32":
            row_f, col_f = workload_flops(msgs, R)
            f = {"rowpass": row_f, "colpass": col_f}.get(dom)
            if f:
                roofline["fp32_tflops"] = round(f / (t_dom / args.steps / 1e3) / 1e12, 3)
        else:
            terms = {"bf16x3": (3, 3), "bf16x2": (2, 2), "bf16": (1, 2), "bf16mix": (1, 2)}[args.method]
            t_gemm = (brk["rowpass"][0] + brk["colpass"][0]) / brk_steps / 1e3
            fl = mfma_flops(msgs, R, *terms)
            roofline["mfma"] = {"issued_tflops": round(fl / t_gemm / 1e12, 2), "peak": BF16_PEAK_TFLOPS,
                                "frac": round(fl / t_gemm / 1e12 / BF16_PEAK_TFLOPS, 4),
                                "kernels": "rowpass+colpass"}

    cpu = None
    if not args.no_cpu:
        k = min(n, 64)
        imgs, busy, cores = cpu_baseline_timed(ep, sizes[:k], joints[:k], seeds[:k], args.cpu_seconds)
        cap = os.environ.get("OMP_NUM_THREADS")
        cpu = {"value": round(imgs / busy, 4), "unit": "images/s", "cores": cores, "kind": "port",
               "per_core": round(imgs / busy / cores, 4), "host_cores_visible": host_threads()[1],
               "threads_rule": "every core in the process's affinity mask" + (
                   ", capped by OMP_NUM_THREADS=%s (this host's CPU share per GPU)" % cap if cap else ""),
               "sample": "%d sweeps over the first instances (float32 oracle, one instance per thread), %.1f s" % (
                   imgs, busy)}
        print("cpu baseline: %d images in %.2f s on %d threads" % (imgs, busy, cores), file=sys.stderr, flush=True)

    desc = {"bf16mix": "x, rotation and y filters as MFMA GEMMs (x bf16, rotation bf16 3-term split, fp16 "
                       "intermediate, y fp16 with split taps), FP32 accumulate",
            "bf16x3": "x, rotation and y filters as BF16 MFMA GEMMs, 3-term split, FP32 intermediate, FP32 accumulate",
            "bf16x2": "x, rotation and y filters as MFMA GEMMs (x data split hi+lo, fp16 intermediate), "
                      "FP32 accumulate",
            "bf16": "x, rotation and y filters as MFMA GEMMs (x bf16, fp16 intermediate), FP32 accumulate",
            "fp32": "plain FP32"}[args.method]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if args.method != "fp32" else "f32",
        "data": "synthetic",
        "config": {"workload": ("%s sweep, %d x (%dx%d positions x %d rotations x 1 scale), 10-part tree, "
                                "per-image spatial model" % (desc, n, H, W, R)) if args.config == 1 else
                               ("%s sweep, %d instances of uneven size (H, W in 160..480, multiples of 8) x %d "
                                "rotations x %d scales (0.8/1.0/1.2), 10-part tree, per-image spatial model"
                                % (desc, n, R, S)),
                   "instances": n, "height": H if args.config == 1 else "160..480",
                   "width": W if args.config == 1 else "160..480", "rotations": R, "scales": ep.num_scale_steps,
                   "states_per_part": states,
                   "parts": 10, "method": args.method, "execution": "hipgraph" if args.graph else "stream",
                   "autotune": tuned or None,
                   "launch_classes": plan.classes()},
        "plain_fp32": fp32,
        "vs_plain_fp32": round(value / fp32["value"], 3) if fp32 else None,
        "map_guard": guarded,
        "learned_pruning": pruned,
        "error_vs_fp64": err,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
