"""The C-ABI library loads and exports every symbol include/cdft_hip.h declares.
No compute calls (CPU container)."""
import ctypes
import os
import re

import pytest

from cdft_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "cdft_hip.h")).read()
    return sorted(set(re.findall(r"\b(cdft_[a-z_0-9]+)\s*\(", txt)))


def test_library_exists_and_exports_header_symbols():
    assert os.path.exists(_lib.LIBPATH), "build with __graft_entry__.build()"
    lib = ctypes.CDLL(_lib.LIBPATH)
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(_lib.EXPORTS) == syms


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirrors agree with the C compiler's layout of the header structs."""
    src = tmp_path / "layout.c"
    fields = {"cdft_system_desc": _lib.SystemDesc, "cdft_solver_opts": _lib.SolverOpts, "cdft_result": _lib.Result,
              "cdft_tuning": _lib.Tuning}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cdft_hip.h"', "int main(void){"]
    for cname, py in fields.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, py in fields.items():
        assert int(out[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_no_silent_fallback_without_gpu():
    """Without a visible GPU the product path raises instead of computing on the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_lib.CDFTError):
        _lib.Context(0)


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; asking for more exits with a message, before any GPU use."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
