"""The C ABI: libhvc_engine.so loads and exports every entry point declared in
include/hvc_engine.h; the Python view declares the same set; argument errors
come back as status codes.  No device compute is needed here."""
import ctypes
import os
import re
import subprocess

import pytest

from oracle_lib import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("hvc_engine.h", "hvc_topology.h", "hvc_driver.h")]


def declared(headers=HEADERS):
    found = set()
    for h in headers:
        txt = open(h).read()
        found |= set(re.findall(r"^\s*(?:hvc_status|void|int|const char\*|hvc_ctx\*)\s+(hvc_\w+)\s*\(", txt, re.M))
    return sorted(found)


def test_header_and_python_view_agree():
    from unsupervised_growth_amd import _abi
    assert declared(HEADERS[:1]) == sorted(_abi.EXPORTS)
    assert declared(HEADERS[1:2]) == sorted(_abi.TOPOLOGY_EXPORTS)
    assert declared(HEADERS[2:]) == sorted(_abi.DRIVER_EXPORTS)
    assert set(os.listdir(os.path.join(ROOT, "include"))) == {os.path.basename(h) for h in HEADERS}


def test_library_exports_every_declared_symbol():
    from unsupervised_growth_amd import _abi
    assert os.path.exists(_abi.LIB_PATH), "build the engine first (python -m unsupervised_growth_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(hvc_\w+)$", out, re.M))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    lib = _abi.load()
    for name in declared():
        assert getattr(lib, name) is not None


def test_invalid_params_rejected_before_device_use():
    from unsupervised_growth_amd import _abi
    from unsupervised_growth_amd.engine import make_params
    lib = _abi.load()
    for bad in (dict(dt=0.0), dict(ra_model=3), dict(precision=7)):
        p = make_params(**bad)
        ctx = ctypes.c_void_p()
        st = lib.hvc_create(ctypes.byref(p), 0, ctypes.byref(ctx))
        assert st == _abi.HVC_E_ARG
        assert lib.hvc_last_error(ctx)
        lib.hvc_destroy(ctx)
    assert lib.hvc_create(None, 0, None) == _abi.HVC_E_ARG


def test_engine_reports_missing_device_or_builds_context():
    """Without a GPU hvc_create returns HVC_E_HIP (no CPU fallback exists)."""
    from unsupervised_growth_amd import _abi
    from unsupervised_growth_amd.engine import Engine, EngineError
    try:
        eng = Engine()
    except EngineError as e:
        assert e.status == _abi.HVC_E_HIP
    else:
        eng.close()
