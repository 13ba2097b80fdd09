"""The C-ABI libraries load and export every function include/*.h declares.
No compute call needs a GPU here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*\**\s*\**\s*(mrcpp_[a-z0-9_]+|mwg_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header", ["mwg.h", "mrcpp_amd.h"])
def test_every_declared_symbol_is_exported(mr, header):
    lib = ctypes.CDLL(mr.LIB_PATH)
    names = declared(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_host_abi(mr):
    names = declared("mrcpp_amd.h")
    assert set(names) <= set(mr.SIGNATURES), set(names) - set(mr.SIGNATURES)


def test_device_calls_fail_cleanly_without_init(mr):
    """Before mwg_init a device call returns an error code, never crashes."""
    lib = ctypes.CDLL(mr.LIB_PATH)
    lib.mwg_band_apply.restype = ctypes.c_int
    rc = lib.mwg_band_apply(None, None, ctypes.c_int(1), ctypes.c_int(0))
    assert rc != 0


def test_oracle_library_loads(orc):
    assert orc.Oracle.max_threads() >= 1


def test_band_launch_settings_are_checked(mr):
    """Launch-shape setters need no device: they check their arguments."""
    assert mr.band_launch() == (32, 0)
    for bad in ((4, 0), (2048, 0), (32, -1), (32, 65)):
        with pytest.raises(mr.MRCPPError):
            mr.set_band_launch(*bad)
    assert mr.band_launch() == (32, 0)
    with pytest.raises(ValueError):
        mr.set_band_layout("diagonal")
