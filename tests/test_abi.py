"""The C ABI: struct layouts, exported symbols and defaults (CPU only; no
compute call needs a GPU here)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import cases

abi = cases.abi
ROOT = cases.ROOT
HEADER = os.path.join(ROOT, "include", "bubbles_rp.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(brp_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(brp):
    lib = brp.load_library()
    names = _declared_functions()
    assert len(names) >= 17 and "brp_adjoint" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _c_layout():
    structs = {"brp_config": abi.Config, "brp_step_stats": abi.StepStats, "brp_info": abi.Info,
               "brp_objective": abi.Objective, "brp_adjoint_stats": abi.AdjointStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return structs, out.split("\n")


def test_struct_layouts_match_header():
    structs, lines = _c_layout()
    for ln in lines:
        if not ln:
            continue
        cname, field, val = ln.split()
        py = structs[cname]
        if field == "size":
            assert C.sizeof(py) == int(val), cname
        else:
            assert getattr(py, field).offset == int(val), (cname, field)


def test_config_default_matches_python_twin(brp):
    lib = brp.load_library()
    c = abi.Config()
    lib.brp_config_default(C.byref(c))
    assert c.as_dict() == abi.default_config().as_dict()


def test_create_fails_loudly_or_runs(brp):
    lib = brp.load_library()
    cfg = cases.acoustic_case(16)
    h = C.c_void_p()
    rc = lib.brp_create(C.byref(cfg), C.byref(h))
    if rc == abi.ERR_NO_DEVICE:
        assert b"no HIP device" in lib.brp_last_error(None)
        assert not h
    else:
        assert rc == abi.OK
        lib.brp_destroy(h)


def test_create_rejects_bad_configs(brp):
    lib = brp.load_library()
    h = C.c_void_p()
    for bad in (dict(grid_x=2), dict(sim_order=4), dict(sim_cfl=0.0),
                dict(sim_smooth_delta_band=9), dict(precision=7)):
        cfg = cases.acoustic_case(16)
        cfg.update(**bad)
        assert lib.brp_create(C.byref(cfg), C.byref(h)) == abi.ERR_INVALID, bad
        assert lib.brp_last_error(None)


def test_missing_library_raises(brp, monkeypatch, tmp_path):
    monkeypatch.setattr(brp, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(brp, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        brp.load_library()
