"""The mixed form's single-kernel iteration with more slow rows (air cells
and the liquid cells beside them) than one launch of its slow-row blocks
covers at one row per thread (1024 blocks of 512 threads): the slow-row
kernels stride over their list, and the iteration sums as many slow-row
partials as blocks were launched.

Bubble column at 128^3 (14M particles) with the liquid between cell layers
k = 40 and 100 turned to air (about 950k air cells), one projection to 1e-8
in the mixed and in the FP64 form from the same inputs: both reach 1e-8 on
the FP64 true residual, and their pressures and liquid faces agree to 1e-6
(both solve the same system to 1e-8; about 1e-9 is typical).
"""
import numpy as np
import pytest

from helpers import load_scene, rel_l2, say

pytestmark = pytest.mark.gpu

FACE_W = ["WATER_UX", "WATER_UY", "WATER_UZ"]


def _prepared(pkg, s, precision):
    cfg = pkg.Config.from_buffer_copy(s["cfg"])
    cfg.precision = precision
    cfg.tol = 1e-8
    cfg.max_iters = 60000
    sim = pkg.Simulator(cfg)
    load_scene(sim, s)
    sim.hash_sort()
    sim.apply_forces()
    sim.levelset_update()
    sim.p2g()
    sim.extrapolate_bc(True)
    return sim


def test_mixed_projection_many_slow_rows(pkg, scenes):
    n = 128
    say("bubble column %d^3 with an air layer" % n)
    s = scenes.make("bubble_column", n=n)
    fl = s["particles"]["flag"]
    k = np.floor(s["particles"]["pos"][:, 2] / np.float32(s["cfg"].dx)).astype(np.int64)
    fl[(fl == pkg.TYPEFLUID) & (k >= 40) & (k < 100)] = pkg.TYPEAIR
    say("  %d particles, %d air" % (len(fl), int(np.count_nonzero(fl == pkg.TYPEAIR))))
    res = {}
    for name, prec in (("mixed", pkg.MIXED), ("fp64", pkg.FP64)):
        sim = _prepared(pkg, s, prec)
        nair = int(np.count_nonzero(sim.get_field("MARK") == pkg.TYPEAIR))
        info = sim.project()
        say("  %s: %d air cells, %s" % (name, nair, info))
        assert nair > 1024 * 512  # more slow rows than one row per thread of the slow-row launch
        assert info["converged"] == 1 and info["rel_residual"] <= 1e-8, name
        res[name] = {f: sim.get_field(f) for f in ["PRESSURE"] + FACE_W}
        sim.close()
    for f in ["PRESSURE"] + FACE_W:
        e = rel_l2(res["mixed"][f], res["fp64"][f])
        say("  %s rel L2 mixed vs FP64 form = %.2e" % (f, e))
        assert e <= 1e-6, f
