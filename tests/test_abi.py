"""The C-ABI library loads without a GPU and exports every declared symbol;
the product never reaches the oracle."""
import ctypes
import glob
import os
import re

from maple3dmt_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(m3dh?_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(native.DEVICE_SYMBOLS + native.HOST_SYMBOLS) <= names


def test_declarations_cite_the_reference():
    text = open(os.path.join(ROOT, "include", "m3d_device.h")).read()
    for cite in ("forward_solver_fv.cpp:266-285", "forward_solver_fv.cpp:778-888",
                 "operators.cpp:556-837", "bicgstab.h:23-28"):
        assert cite.split(":")[0] in text


def test_product_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "maple3dmt_amd")
    for path in glob.glob(os.path.join(pkg, "**", "*"), recursive=True):
        if os.path.isdir(path) or path.endswith((".so", ".o", ".pyc")):
            continue
        text = open(path, errors="ignore").read()
        assert "oracle" not in text.lower(), path


def test_host_model_without_device():
    from maple3dmt_amd import configs
    m = configs.build_model(configs.halfspace())
    assert m.n == 5175 and m.nc == 1296
    try:
        m.device
    except native.NativeError:
        pass
    else:
        raise AssertionError("host-only model must not expose a device operator")


def test_error_reporting_is_a_string():
    lib = native.lib()
    assert isinstance(lib.m3d_last_error(), bytes)
    assert isinstance(lib.m3dh_last_error(), bytes)
    # Invalid arguments are rejected without touching a device.
    st = lib.m3d_operator_create(None, None, None)
    assert st != 0
    assert b"null" in lib.m3d_last_error()
    assert ctypes.sizeof(native.SolveResultC) > 0
