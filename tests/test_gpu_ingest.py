"""Unary ingestion on the device (loadScoreGrid's grid-to-image mapping,
partapp.cpp:866-899, then clip_scores_fill + computeLogGrid,
objectdetect_findrot.cpp:834-845) against the oracle: the same LOG_ZERO
cells, the same values up to the float log's last bit."""
import math
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import ingest_scores

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def mats(kind, n):
    out = []
    for s in range(n):
        if kind == "stride4":
            T = [[4.0, 0, 2.0], [0, 4.0, 1.5 + 0.25 * s], [0, 0, 1]]
        elif kind == "rotated":
            a = 0.3 + 0.2 * s
            c, d = 2.5 * math.cos(a), 2.5 * math.sin(a)
            T = [[c, -d, 40.0], [d, c, -10.0 + s], [0, 0, 1]]
        else:  # "shrink": several grid cells land on one pixel, the last one wins
            T = [[0.5, 0.0, 0.2], [0.1 * s, 0.5, 0.3], [0, 0, 1]]
        out.append(T)
    return np.array(out, dtype=np.float64)


@pytest.mark.parametrize("interpolate", [False, True])
@pytest.mark.parametrize("kind", ["stride4", "rotated", "shrink"])
def test_ingest_matches_oracle(gpu, kind, interpolate):
    import torch
    n, gh, gw, H, W = 5, 37, 45, 150, 181
    rng = np.random.default_rng(sum(map(ord, kind)) + int(interpolate))
    sg = rng.normal(0.3, 1.0, (n, gh, gw)).astype(np.float32)
    sg[rng.random(sg.shape) < 0.1] = 0.0  # unevaluated cells
    T = mats(kind, n)
    say("ingest", kind, "bilinear" if interpolate else "direct")
    got = ingest_scores(torch.from_numpy(sg).to(gpu), T, interpolate, H, W).cpu().numpy()
    ref = oracle.ingest_scores(sg, T, interpolate, H, W)
    np.testing.assert_array_equal(got == -1e6, ref == -1e6)
    live = ref != -1e6
    assert live.mean() > 0.01
    np.testing.assert_allclose(got[live], ref[live], rtol=2e-7, atol=2e-7)
    say("  exact %.4f of %d live cells" % (np.mean(got[live] == ref[live]), live.sum()))


def test_ingest_rejects_singular_bilinear(gpu):
    import torch
    from partapp_amd._lib import PsError
    sg = torch.ones((1, 4, 4), device=gpu)
    with pytest.raises(PsError) as e:
        ingest_scores(sg, np.zeros((1, 3, 3)), True, 8, 8)
    assert e.value.status == -1
