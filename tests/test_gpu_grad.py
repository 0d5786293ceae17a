"""Multi-step gradients through the adjoint entry points (small water-slab run).

The harness (tests/grad_harness.py) differentiates `steps` steps of the
water-slab scene with checkpointing and compares the adjoint gradient with:
  * central finite differences of the FP64 oracle replaying the run with the
    GPU run's discrete decisions frozen (bar: relative error <= 1e-4 per
    random direction);
  * the FP64 adjoint gradient, for the FP32 and mixed forms (relative L2,
    bounded by the forms' solver tolerance);
  * itself with a different checkpoint interval (same gradient up to the
    rounding of the FP64 atomics in the G2P adjoint).
"""
import numpy as np
import pytest

import grad_harness as gh
from helpers import say

pytestmark = pytest.mark.gpu

N, STEPS = 24, 4


@pytest.fixture(scope="module")
def setup(pkg, scenes):
    scene = scenes.make("diff_slab", n=N)
    u_init, s, target = gh.parameters(scene, 3)
    say("diff_slab n=%d particles=%d steps=%d" % (N, scene["n_particles"], STEPS))
    g64 = gh.adjoint_gradient(pkg, scene, pkg.FP64, STEPS, 2, u_init, s, target)
    say("  loss %.12e |g_u| %.6e |g_s| %.6e" % (g64["loss"], np.linalg.norm(g64["g_u"]), np.linalg.norm(g64["g_s"])))
    return scene, u_init, s, target, g64


def test_gradient_vs_fd(pkg, orc, setup):
    scene, u_init, s, target, g64 = setup
    errs = gh.fd_check(orc, pkg, scene, g64, u_init, s, target, ndirs=4, eps=1e-6, say=say)
    assert max(errs) <= 1e-4


def test_checkpoint_interval(pkg, setup):
    scene, u_init, s, target, g64 = setup
    g1 = gh.adjoint_gradient(pkg, scene, pkg.FP64, STEPS, 1, u_init, s, target)
    for key in ("g_u", "g_s"):
        e = np.linalg.norm(g1[key] - g64[key]) / np.linalg.norm(g64[key])
        say("  %s checkpoint 1 vs 2: rel L2 %.2e" % (key, e))
        assert e <= 1e-10


@pytest.mark.parametrize("form,tol,bar", [("FP32", 1e-7, 1e-3), ("MIXED", 1e-10, 1e-5)])
def test_reduced_precision_gradient(pkg, setup, form, tol, bar):
    scene, u_init, s, target, g64 = setup
    g = gh.adjoint_gradient(pkg, scene, getattr(pkg, form), STEPS, 2, u_init, s, target, tol=tol)
    for key in ("g_u", "g_s"):
        e = np.linalg.norm(g[key] - g64[key]) / np.linalg.norm(g64[key])
        say("  %s %s vs FP64 gradient: rel L2 %.2e" % (form, key, e))
        assert e <= bar
