"""Benchmark of the pictorial-structure sweep on one MI355X.

Workload (the metric's configuration): 32 instances, each 300 x 300 positions
x 24 rotations x 1 scale, the 10-part person tree, synthetic log unaries and a
per-image spatial model (seeds 0..31).  One step = the whole sweep of the
batch (upward pass, root, downward pass, marginals of all parts, MAP, root
posterior) with the unaries already resident in HBM.

Prints one JSON line.  `roofline` is measured live over the timed steps with
HIP events around each launch (on the stream the kernels run on); the
dominant kernel's algorithmic FP32 flops per launch (2 x states x taps of the
filters it evaluates, per message) over its average launch time, against the
published FP32 peak.  `cpu_baseline` times the oracle (the float32 CPU
restatement of the reference) on a bounded sample of the same instances,
one instance per thread.
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
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--size", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", help="time HIP-graph replays (no per-kernel roofline)")
    return ap.parse_args()


def workload_flops(plan_sizes, joints, ep):
    """Algorithmic FP32 flops per kernel class, per sweep: 2 * states * taps.
    Row pass: rotation + x taps; column pass: y taps (one message each)."""
    from partapp_amd import _lib
    L = _lib.lib()
    R = ep.num_rotation_steps
    row = col = 0.0
    for (H, W), js in zip(plan_sizes, joints):
        n = R * H * W
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
                    row += 2.0 * n * (rl.value + 2 * kx.value + 1)
                    col += 2.0 * n * (2 * ky.value + 1)
    return row, col


def cpu_baseline_timed(ep, sizes, joints, seeds, budget_s):
    """The float32 oracle on as many instances as fit the time budget, one
    instance per thread on the host cores; input generation is not timed."""
    import oracle
    from partapp_amd.api import synth_unary
    cores = min(os.cpu_count() or 1, 16)
    oep = oracle.exp_param_from(ep)
    images = 0
    busy = 0.0
    i = 0
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        while i < len(sizes) and busy < budget_s:
            batch = list(range(i, min(i + cores, len(sizes))))
            us = [synth_unary(seeds[k], 10, ep.num_scale_steps, ep.num_rotation_steps, sizes[k][0], sizes[k][1],
                              threads=cores) for k in batch]
            tb = time.perf_counter()
            futs = [ex.submit(oracle.root_posterior_rot, oep, [1] * 10, [0] * 10, 4, joints[k], u, True, True)
                    for k, u in zip(batch, us)]
            for f in futs:
                f.result()
            busy += time.perf_counter() - tb
            images += len(batch)
            i += len(batch)
    return images, busy, cores


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: instances shard by image across GPUs with no exchange, "
              "and this build measures the single-GPU share (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from partapp_amd.api import Plan, config_exp_param, person_model, synth_joints, synth_unary

    dev = torch.device("cuda:0")
    n, H = args.instances, args.size
    W = H
    ep = config_exp_param(1)
    seeds = list(range(n))
    joints = [synth_joints(s) for s in seeds]
    sizes = [(H, W)] * n
    model = person_model(ep, joints[0])
    plan = Plan(model, sizes, joints)
    print("plan: %d instances %dx%dx%d, scratch %.2f GB" % (n, H, W, ep.num_rotation_steps,
                                                          plan.scratch_bytes() / 1e9), file=sys.stderr, flush=True)
    unary = torch.empty((plan.unary_numel,), dtype=torch.float32, device=dev)
    threads = min(os.cpu_count() or 1, 16)
    for i, s in enumerate(seeds):
        u = synth_unary(s, 10, 1, ep.num_rotation_steps, H, W, threads=threads)
        o = plan.unary_offsets[i]
        unary[o:o + u.size].copy_(torch.from_numpy(u.ravel()))
    marg = torch.empty_like(unary)
    root = torch.empty((plan.root_numel,), dtype=torch.float32, device=dev)
    maps = torch.zeros((plan.map_numel * 4,), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    print("inputs resident: %.2f GB" % (unary.numel() * 4 / 1e9), file=sys.stderr, flush=True)

    stream = torch.cuda.current_stream()
    if args.graph:
        stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            plan.execute(unary, marg, root, maps, stream=stream, graph=args.graph)
        stream.synchronize()
        print("warmup done", file=sys.stderr, flush=True)
        if not args.graph:
            plan.set_profiling(True)
            plan.read_profile()  # clear
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        t0.record(stream)
        for _ in range(args.steps):
            plan.execute(unary, marg, root, maps, stream=stream, graph=args.graph)
        t1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1)
    ms_per_step = ms / args.steps
    value = n * args.steps / (ms / 1e3)
    print("timed %d steps: %.3f ms/step (wall %.3f s)" % (args.steps, ms_per_step, wall), file=sys.stderr,
          flush=True)

    roofline = None
    if not args.graph:
        prof = plan.read_profile()
        plan.set_profiling(False)
        row_f, col_f = workload_flops(sizes, joints, ep)
        flops = {"rowpass": row_f, "colpass": col_f}
        for k, (t, c) in prof.items():
            print("  %-10s %9.3f ms total, %4d launches" % (k, t, c), file=sys.stderr)
        dom = max(("rowpass", "colpass"), key=lambda k: prof[k][0])
        t_dom, n_dom = prof[dom]
        launches_per_step = n_dom / args.steps
        flops_per_launch = flops[dom] / launches_per_step
        avg_s = t_dom / n_dom / 1e3
        achieved = flops_per_launch / avg_s / 1e12
        roofline = {"bound": "fp32", "kernel": dom, "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
                    "avg_launch_ms": round(t_dom / n_dom, 4),
                    "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in prof.items()}}

    cpu = None
    if not args.no_cpu:
        k = min(n, 64)
        imgs, busy, cores = cpu_baseline_timed(ep, sizes[:k], joints[:k], seeds[:k], args.cpu_seconds)
        cpu = {"value": round(imgs / busy, 4), "unit": "images/s", "cores": cores, "kind": "port",
               "sample": "%d of the %d instances (float32 oracle, one instance per thread), %.1f s" % (
                   imgs, n, busy)}
        print("cpu baseline: %d images in %.2f s on %d threads" % (imgs, busy, cores), file=sys.stderr, flush=True)

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
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain FP32 sweep, 32 x (300x300 positions x 24 rotations x 1 scale), "
                               "10-part tree, per-image spatial model",
                   "instances": n, "height": H, "width": W, "rotations": ep.num_rotation_steps,
                   "scales": ep.num_scale_steps, "parts": 10, "method": "plain-fp32",
                   "execution": "hipgraph" if args.graph else "stream"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
