"""ps_message_noalloc (the message entry point that allocates no device
memory) against ps_message, through the C ABI: bit-identical parent grids on
two calls in a row, the second on a smaller grid (the device array keeps the
first call's contents), for a diagonal and an off-diagonal covariance; a
message larger than the array is refused before any device work."""
import ctypes
import sys

import numpy as np
import pytest

from partapp_amd._lib import PsError, lib
from partapp_amd.api import ExpParam, compute_rot_joint_marginal, person_model

pytestmark = pytest.mark.gpu

CASES = [
    # (R, H, W), offset_c, offset_p, C, rot_mean, rot_sigma, scale
    ((24, 120, 136), (7.3, -11.2), (-3.9, 8.8), (30.0, 0.0, 0.0, 50.0), 0.3, 0.25, 1.0),
    ((24, 41, 57), (12.0, 4.0), (-6.5, 2.0), (40.0, 0.0, 0.0, 20.0), -0.2, 0.15, 1.2),
    ((24, 96, 104), (5.0, -3.0), (2.0, 7.0), (60.0, 25.0, 25.0, 30.0), 0.1, 0.2, 1.0),   # off-diagonal
    ((24, 48, 40), (-4.0, 9.0), (3.0, -2.0), (50.0, -15.0, -15.0, 20.0), 0.0, 0.3, 0.8),  # off-diagonal, smaller
]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("pair", [(0, 1), (2, 3)], ids=["diagonal", "off_diagonal"])
def test_noalloc_matches_message_twice(gpu, pair):
    import torch
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    model = person_model(ep)
    for k in pair:  # the second call on a smaller grid
        (R, H, W), oc, op, C, rm, rs, sc = CASES[k]
        say("no-alloc message case %d: %dx%dx%d" % (k, R, H, W))
        rng = np.random.default_rng(500 + k)
        child = torch.from_numpy(rng.normal(-3, 1.0, (R, H, W)).astype(np.float32)).to(gpu)
        ref = compute_rot_joint_marginal(model, child, oc, op, C, rm, rs, sc, is_sparse=(k == 2))
        got = torch.full_like(child, float("nan"))
        compute_rot_joint_marginal(model, child, oc, op, C, rm, rs, sc, is_sparse=(k == 2), out=got, no_alloc=True)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), "case %d: max |diff| %g" % (k, float((got - ref).abs().max()))
        assert bool((ref > -9e5).any())


def test_noalloc_refuses_messages_beyond_its_array(gpu):
    import torch
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    model = person_model(ep)
    # 24 x 600 x 600 floats of row-pass output alone is 34.6 MB > 32 MiB; the
    # size check runs on the host before anything touches the device
    H = W = 600
    child = torch.zeros((24, 4, 4), dtype=torch.float32, device=gpu)  # never read: the call is refused first
    oc = (ctypes.c_double * 2)(1.0, 2.0)
    op = (ctypes.c_double * 2)(0.0, 0.0)
    cc = (ctypes.c_double * 4)(16.0, 0.0, 0.0, 16.0)
    st = lib().ps_message_noalloc(model.handle, child.data_ptr(), child.data_ptr(), H, W, oc, op, cc, 0.0, 0.2, 1.0, 0,
                                  torch.cuda.current_stream().cuda_stream)
    assert st == -5  # PS_ERR_UNSUPPORTED
    assert b"ps_message_noalloc has" in lib().ps_last_error_string()
    with pytest.raises(PsError):
        compute_rot_joint_marginal(model, torch.zeros((24, 600, 600), dtype=torch.float32, device=gpu),
                                   (1.0, 2.0), (0.0, 0.0), (16.0, 0, 0, 16.0), 0.0, 0.2, 1.0, no_alloc=True)
