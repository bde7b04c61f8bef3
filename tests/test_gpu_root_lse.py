"""The root posterior: the log-sum-exp of the root belief over the valid root
rotations (objectdetect_findrot.cpp:713-726).

* The FP32 sweep sums expf without a shift, as the reference does: where the
  sum passes the float range (a root belief above ~88.7) it is +inf there
  and here, finite below.
* The BF16 methods accumulate it in FP64 with a running maximum shift: their
  root posterior stays finite on the same inputs and within the method's
  budget of the FP64 oracle's (relative L1 of the normalised posterior), as
  it is on unshifted inputs.
"""
import sys

import numpy as np
import pytest

import oracle
from partapp_amd.api import ExpParam, person_model, synth_joints, synth_unary

from test_gpu_bf16 import METHOD_L1, run_method
from test_gpu_parity import oracle_run

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("shift", [0.0, 15.0])
def test_root_posterior_vs_f64(gpu, shift):
    ep = ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0)
    joints = synth_joints(330)
    model = person_model(ep, joints)
    # every part's unary raised by `shift`: the root belief rises by about 10 x shift
    u = synth_unary(331, 10, 1, 24, 96, 112) + np.float32(shift)
    say("root posterior, unaries +%g" % shift)
    outs, _ = run_method(model, [u], [joints], "bf16mix")
    root = outs[0][1]
    r64 = oracle_run(model, u, joints, np.float64)[0]
    assert root.shape == r64.shape
    assert np.isfinite(root).all()
    d = oracle.rel_l1(oracle.normalized(root), oracle.normalized(r64))
    say("  bf16mix: rel L1 of the normalised root posterior vs f64 %.3g (f64 max %.1f)" % (d, r64.max()))
    assert d <= METHOD_L1["bf16mix"]

    outs32, _ = run_method(model, [u], [joints], "fp32")
    root32 = outs32[0][1]
    over, under = r64 > 89.5, r64 < 88.0
    assert np.isposinf(root32[over]).all()
    assert np.isfinite(root32[under]).all()
    if shift:
        assert over.any()
    else:
        assert not over.any()
        np.testing.assert_allclose(root32, oracle_run(model, u, joints)[0], rtol=2e-5, atol=2e-5)
