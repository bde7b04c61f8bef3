"""findLocalMax on the GPU (ps_find_local_max) against the oracle's
restatement (objectdetect_aux.cpp:193-251), itself pinned on a hand-built grid
in test_oracle_pins.py: random grids with more and fewer maxima than max_n,
plateaus of equal values, a constant single-slice grid, several grids at once."""
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import find_local_max

pytestmark = pytest.mark.gpu


def as_tuples(a):
    return [(int(r["d"]), int(r["x"]), int(r["y"]), float(r["score"])) for r in a]


def ref_tuples(g, max_n):
    return [(int(d), int(x), int(y), float(np.float32(v))) for d, x, y, v in oracle.find_local_max(g, max_n)]


@pytest.mark.parametrize("shape,max_n,kind", [
    ((3, 40, 50), 1000, "random"),     # fewer maxima than max_n: scan order
    ((3, 40, 50), 50, "random"),       # more: the 50 largest
    ((24, 100, 120), 1000, "random"),  # ~10k candidates
    ((5, 30, 30), 200, "plateau"),     # many equal values
    ((1, 20, 30), 100, "constant"),    # every point is a local maximum
    ((8, 64, 64), 1000, "smooth"),     # few maxima
])
def test_local_max_matches_oracle(gpu, shape, max_n, kind):
    import torch
    rng = np.random.default_rng(hash((shape, max_n, kind)) % 2**32)
    if kind == "random":
        g = rng.normal(0, 1, shape).astype(np.float32)
    elif kind == "plateau":
        g = rng.integers(0, 4, shape).astype(np.float32)
    elif kind == "constant":
        g = np.full(shape, 2.5, dtype=np.float32)
    else:
        from scipy import ndimage
        g = ndimage.gaussian_filter(rng.normal(0, 1, shape), 3).astype(np.float32)
    print("local max", shape, max_n, kind, file=sys.stderr, flush=True)
    got = as_tuples(find_local_max(torch.from_numpy(g).to(gpu), max_n))
    assert got == ref_tuples(g, max_n)


def test_local_max_many_grids(gpu):
    import torch
    rng = np.random.default_rng(5)
    gs = [rng.normal(0, 1, s).astype(np.float32) for s in [(4, 33, 47), (2, 60, 20), (6, 25, 25)]]
    flat = torch.from_numpy(np.concatenate([g.ravel() for g in gs])).to(gpu)
    views, o = [], 0
    for g in gs:
        views.append(flat[o:o + g.size].view(*g.shape))
        o += g.size
    res = find_local_max(views, 40)
    for g, r in zip(gs, res):
        assert as_tuples(r) == ref_tuples(g, 40)
