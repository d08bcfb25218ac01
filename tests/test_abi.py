"""libjoker.so loads and exports every symbol include/joker.h declares (no
device needed; no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "joker.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(jk_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("jk_init", "jk_star_create", "jk_lnl_f64", "jk_lnl_mixed", "jk_reject",
              "jk_linear_posterior", "jk_status_string"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from thejoker_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(declared_symbols()) <= bound


def test_status_strings_and_version():
    from thejoker_amd import _lib
    lib = _lib.load()
    assert lib.jk_abi_version() == _lib.ABI_VERSION
    assert lib.jk_status_string(0) == b"ok"
    assert b"invalid" in lib.jk_status_string(1)
    assert lib.jk_status_string(99) == b"unknown status"


def test_compute_without_device_fails_loudly():
    import torch

    from thejoker_amd import _lib
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    lib = _lib.load()
    assert lib.jk_init(0) == 4  # JK_ENODEV
    with pytest.raises(_lib.JokerLibError):
        _lib.ensure_device(0)
