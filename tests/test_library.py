"""The C-ABI library loads and exports every symbol include/steam_hip.h
declares; host-only entry points behave.  No GPU needed."""

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "steam_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(steam_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "steam_conv_plan_create" in names and "steam_flux_advance" in names
    assert len(names) >= 19


def test_library_exports_every_declared_symbol():
    from steam_amd import _lib
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    from steam_amd import _lib
    assert sorted(_lib.EXPORTED) == _declared()


def test_version_and_halfwidths_host_only():
    from steam_amd import _lib
    lib = _lib.load()
    assert lib.steam_version() >= 10000
    half = (ctypes.c_int64 * 3)()
    assert lib.steam_envelope_halfwidths(1.0, 0.125, 0.125, 0.125, 4.0, half) == 0
    assert list(half) == [32, 32, 32]
    assert lib.steam_envelope_halfwidths(1.0, 0.5, 0.5, 0.5, 3.0, half) == 0
    assert list(half) == [6, 6, 6]
    rc = lib.steam_envelope_halfwidths(1.0, 0.0, 0.5, 0.5, 3.0, half)
    assert rc == _lib.STEAM_EINVAL
    assert b"positive" in lib.steam_last_error()


def test_cpu_device_request_is_an_error():
    from steam_amd import _lib
    with pytest.raises(ValueError):
        _lib.require_gpu('cpu')
