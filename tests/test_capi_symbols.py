"""The C-ABI library loads and exports every entry point include/agal_capi.h declares; host-side
defaults and status strings match.  No device work (runs without a GPU)."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from agal_amd import capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "agal_capi.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(agal_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    return capi.load_library()


def test_every_declared_symbol_is_exported(lib):
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    assert sorted(capi.EXPORTED) == declared()


def test_params_default_matches_host_defaults(lib):
    a = capi.Params()
    lib.agal_params_default(C.byref(a))
    b = capi.default_params()
    for name, _ in capi.Params._fields_:
        assert getattr(a, name) == getattr(b, name), name


def test_status_strings(lib):
    for s in range(0, 9):
        assert lib.agal_status_string(s)
    assert lib.agal_status_string(capi.AGAL_ERR_POOL_EXHAUSTED).startswith(b"refinement cancelled")


def test_invalid_arguments_are_rejected_before_device_use(lib):
    fns = capi.bind(lib, "agal_", ["create"])
    h = C.c_void_p()
    bad = capi.default_params(velocity_set=15)
    assert fns["create"](C.byref(bad), 1000, 0, C.byref(h)) == capi.AGAL_ERR_ARG
    bad = capi.default_params(Nx=30)
    assert fns["create"](C.byref(bad), 1000, 0, C.byref(h)) == capi.AGAL_ERR_ARG
    assert fns["create"](C.byref(capi.default_params()), 0, 0, C.byref(h)) == capi.AGAL_ERR_ARG
    big = capi.default_params(Nx=64)
    assert fns["create"](C.byref(big), 100, 0, C.byref(h)) == capi.AGAL_ERR_CAPACITY
