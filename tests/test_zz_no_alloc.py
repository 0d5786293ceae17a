"""mpf_hash_sort_noalloc against mpf_hash_sort, through the C ABI.

The no-allocation entry point takes the caller's device buffers and keeps
its scan scratch in a fixed per-GPU array; two calls in a row (the second on
a smaller grid with fewer particles, so stale scratch and cell lists from
the first would show) must each equal the context sort bit for bit.
"""
import importlib

import numpy as np
import pytest

from helpers import say

pytestmark = pytest.mark.gpu


def _scene(pkg, n, P, seed):
    rng = np.random.default_rng(seed)
    dx = np.float32(1.0 / 64)
    pos = (rng.random((P, 3), dtype=np.float32) * np.float32(n) * dx).astype(np.float32)
    # duplicates, face-exact and out-of-domain positions, shuffled
    pos[::11] = pos[::11].round(3)
    pos[5] = (n * dx, 0.5 * dx, 0.5 * dx)
    pos[6] = (-dx, 2 * dx, 3 * dx)
    pos = pos[rng.permutation(P)]
    cfg = pkg.default_config(n, n + 3, n - 2, particle_capacity=P + 16, mc_enabled=0, heat_enabled=0)
    # the particles fill the box scaled by n: a non-cubic grid leaves some outside on purpose
    return cfg, pos


def _reference(pkg, cfg, pos):
    sim = pkg.Simulator(cfg)
    sim.set_particles(len(pos), pos=pos)
    sim.hash_sort()
    out = {f: sim.get_field(f) for f in ("GRID_HASH", "GRID_INDEX", "CELL_START", "CELL_END")}
    sim.close()
    return out


def _noalloc(pkg, cfg, pos):
    """The entry point on device buffers from the library's own HIP runtime,
    pre-filled with garbage so that every output word must be written."""
    hm = importlib.import_module("2017multiphase_amd.hipmem")
    P = len(pos)
    G = cfg.nx * cfg.ny * cfg.nz
    dpos = hm.DeviceArray.from_host(pos.reshape(-1))
    out = {f: hm.DeviceArray(max(P, 1), np.uint32) for f in ("GRID_HASH", "GRID_INDEX")}
    out["CELL_START"] = hm.DeviceArray(G, np.uint32)
    out["CELL_END"] = hm.DeviceArray(G, np.uint32)
    for d in out.values():
        d.fill_bytes(0x5A)
    pkg.hash_sort_noalloc(cfg, P, dpos, out["GRID_HASH"], out["GRID_INDEX"], out["CELL_START"], out["CELL_END"])
    hm.synchronize()
    res = {f: d.download() for f, d in out.items()}
    res["GRID_HASH"] = res["GRID_HASH"][:P]
    res["GRID_INDEX"] = res["GRID_INDEX"][:P]
    return res


def test_noalloc_sort_matches_context_sort(pkg):
    for call, (n, P, seed) in enumerate(((40, 300000, 1), (12, 5000, 2))):
        cfg, pos = _scene(pkg, n, P, seed)
        say("no-alloc sort call %d: grid %dx%dx%d, %d particles" % (call, cfg.nx, cfg.ny, cfg.nz, P))
        ref = _reference(pkg, cfg, pos)
        got = _noalloc(pkg, cfg, pos)
        for f in ("GRID_HASH", "GRID_INDEX", "CELL_START"):
            assert np.array_equal(got[f], ref[f]), f
        ne = ref["CELL_START"] != pkg.CELL_UNDEF
        assert np.array_equal(got["CELL_END"][ne], ref["CELL_END"][ne])
        assert np.all(got["CELL_END"][~ne] == pkg.CELL_UNDEF)
        sizes = (ref["CELL_END"][ne] - ref["CELL_START"][ne]).astype(np.int64)
        assert sizes.sum() == P


def test_noalloc_sort_empty_and_limits(pkg):
    hm = importlib.import_module("2017multiphase_amd.hipmem")
    cfg = pkg.default_config(8, 8, 8, particle_capacity=16, mc_enabled=0)
    cs = hm.DeviceArray(512, np.uint32)
    ce = hm.DeviceArray(512, np.uint32)
    cs.fill_bytes(0)
    ce.fill_bytes(0)
    pkg.hash_sort_noalloc(cfg, 0, None, None, None, cs, ce)
    hm.synchronize()
    assert np.all(cs.download() == pkg.CELL_UNDEF)
    assert np.all(ce.download() == pkg.CELL_UNDEF)
    big = pkg.default_config(8, 8, 8)
    big.nx, big.ny, big.nz = 1024, 1024, 512  # beyond the fixed scratch: refused before any launch
    with pytest.raises(pkg.MpfError, match="no-allocation limit"):
        pkg.hash_sort_noalloc(big, 0, None, None, None, cs, ce)
