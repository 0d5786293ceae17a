"""C-ABI checks that need no GPU: the library loads, exports every function
include/mpflip.h declares, and the ctypes mirror of every struct has the C
layout (checked against offsetof from a C compiler)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mpflip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(pkg):
    L = ctypes.CDLL(pkg.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing
    assert sorted(pkg.ABI_FUNCTIONS) == declared_functions()


def test_abi_version(pkg):
    assert pkg.lib().mpf_abi_version() == 2


def test_config_default_without_gpu(pkg):
    cfg = pkg.default_config(64, 64, 64)
    assert cfg.dx == pytest.approx(1.0 / 64)
    assert cfg.samplespace == pytest.approx(0.5 / 64)
    assert cfg.m0 == pytest.approx(1000.0 * (0.5 / 64) ** 3, rel=1e-6)
    assert cfg.ref_max_iters == 800 and cfg.heat_max_iters == 3800
    with pytest.raises(ValueError):
        pkg.default_config(2, 64, 64)


def _c_offsets(struct_name, fields):
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    src.append('printf("%%zu\\n", sizeof(%s));' % struct_name)
    for f in fields:
        src.append('printf("%%zu\\n", offsetof(%s, %s));' % (struct_name, f))
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c99", "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return [int(x) for x in out]


@pytest.mark.parametrize("cname,pyname", [("mpf_config", "Config"), ("mpf_solve_info", "SolveInfo"),
                                          ("mpf_phase_info", "PhaseInfo"), ("mpf_mc_info", "McInfo"),
                                          ("mpf_step_info", "StepInfo")])
def test_struct_layout_matches_c(pkg, cname, pyname):
    S = getattr(pkg, pyname)
    names = [f for f, _ in S._fields_]
    got = _c_offsets(cname, names)
    assert got[0] == ctypes.sizeof(S)
    assert got[1:] == [getattr(S, f).offset for f in names]


def test_missing_library_fails_loudly(pkg, monkeypatch, tmp_path):
    monkeypatch.setattr(pkg, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(pkg, "_lib_handle", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        pkg.lib()
