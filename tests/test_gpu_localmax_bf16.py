"""findLocalMax (objectdetect_aux.cpp:193-251) on the root posterior of the
BF16 method bf16mix against the same search on the FP64 oracle's root
posterior.  The fp16 row-pass output flushes values more than ~e^-26 below a
message's bound, so the far tails of a posterior may differ; the local maxima
a caller takes as hypotheses, those within 20 nats of the best one, must be
the same points with the same scores (to the method's error), wherever a
point is a clear maximum (above its neighbours by more than the error) on
either side.
"""
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import ExpParam, find_local_max, person_model, synth_joints, synth_unary

from test_gpu_bf16 import run_method
from test_gpu_parity import oracle_run

pytestmark = pytest.mark.gpu

WINDOW = 20.0   # nats below the best local maximum
MARGIN = 1e-2   # a clear maximum: above every neighbour by more than this
SCORE_TOL = 5e-3
MAX_N = 4096    # ps_find_local_max's largest max_n: beyond it, the MAX_N best


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def margins(g):
    """value minus the largest neighbour of findLocalMax's test (3 x 3 in the
    slice, the same pixel in the neighbouring slices), per point."""
    D, H, W = g.shape
    pad = np.pad(g, ((1, 1), (1, 1), (1, 1)), constant_values=-np.inf)
    nb = np.full(g.shape, -np.inf)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy or dx:
                nb = np.maximum(nb, pad[1:-1, 1 + dy:H + 1 + dy, 1 + dx:W + 1 + dx])
    nb = np.maximum(nb, pad[:-2, 1:-1, 1:-1])
    nb = np.maximum(nb, pad[2:, 1:-1, 1:-1])
    return g - nb


def clear_maxima(points, g, cutoff):
    m = margins(g.astype(np.float64))
    return {(d, x, y): s for d, x, y, s in points if s >= cutoff and m[d, y, x] > MARGIN}


@pytest.mark.parametrize("scales", [1, 3])
def test_root_local_maxima_bf16mix_vs_f64(gpu, scales):
    import torch
    ep = ExpParam(-180.0, 180.0, 24, 0.8 if scales > 1 else 1.0, 1.2 if scales > 1 else 1.0, scales, 0.0)
    joints = synth_joints(340 + scales)
    model = person_model(ep, joints)
    u = synth_unary(341 + scales, 10, scales, 24, 104, 120)
    say("local maxima of the bf16mix root posterior, %d scale(s)" % scales)
    outs, _ = run_method(model, [u], [joints], "bf16mix")
    root = outs[0][1].astype(np.float32)
    r64 = oracle_run(model, u, joints, np.float64)[0]
    got = [(int(r["d"]), int(r["x"]), int(r["y"]), float(r["score"]))
           for r in find_local_max(torch.from_numpy(root).to(gpu), MAX_N)]
    ref = [(int(d), int(x), int(y), float(v)) for d, x, y, v in oracle.find_local_max(r64.astype(np.float32), MAX_N)]
    top = max(s for *_, s in ref)
    assert abs(max(s for *_, s in got) - top) <= SCORE_TOL
    # within the window, above the last kept score of a list cut at MAX_N
    cutoff = top - WINDOW
    for lst in (ref, got):
        if len(lst) == MAX_N:
            cutoff = max(cutoff, min(s for *_, s in lst) + 2 * SCORE_TOL)
    ref_clear, got_clear = clear_maxima(ref, r64, cutoff), clear_maxima(got, root, cutoff)
    got_all, ref_all = {p[:3]: p[3] for p in got}, {p[:3]: p[3] for p in ref}
    say("  %d / %d local maxima, %d / %d clear ones above %.2f (best %.2f)" % (
        len(got), len(ref), len(got_clear), len(ref_clear), cutoff, top))
    assert ref_clear
    for k, s in ref_clear.items():
        assert k in got_all, ("missing", k, s)
        assert abs(got_all[k] - s) <= SCORE_TOL, (k, got_all[k], s)
    for k, s in got_clear.items():
        assert k in ref_all, ("extra", k, s)
