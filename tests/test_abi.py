"""CPU checks of the native boundary: the library loads and exports exactly
what include/vpm_hip.h declares, with the ctypes signatures the binding uses.
No kernel is launched.
"""
import ctypes
import re
from pathlib import Path

import pytest

from openonda_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "vpm_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(vpm_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_parses():
    names = declared_functions()
    assert "vpm_forward" in names and "vpm_get_selfinduced_velocity" in names
    assert len(names) >= 15


def test_every_declared_symbol_is_exported_and_bound():
    lib = ctypes.CDLL(str(L.lib_path()))
    names = declared_functions()
    for n in names:
        assert hasattr(lib, n), f"{n} declared in vpm_hip.h but not exported"
        assert n in L.SIGNATURES, f"{n} has no ctypes signature in openonda_amd/_lib.py"
    for n in L.SIGNATURES:
        assert n in names, f"{n} bound in _lib.py but not declared in vpm_hip.h"


def test_load_and_version():
    lib = L.load()
    assert lib.vpm_abi_version() == 3
    assert lib.vpm_device_count() >= 0
    assert lib.vpm_last_error() is not None


def test_host_side_queries_without_gpu():
    lib = L.load()
    assert lib.vpm_reduction_workspace_bytes(1000) >= 8
    # presorted: no radix-sort size query (that one asks the device)
    opts = L.ForwardOpts(L.OUT_VEL | L.OUT_STRETCH, L.PREC_F32C, L.SIGMA_UNIFORM, 1, -1)
    nb = lib.vpm_forward_workspace_bytes(1 << 20, ctypes.byref(opts))
    # sorted FP64 records (64 B) + keys/indices + scratch: a few hundred bytes per particle
    assert 64 * (1 << 20) < nb < 1024 * (1 << 20)
    assert lib.vpm_tile_config_count() >= 1


def test_argument_errors_return_status_without_launch():
    lib = L.load()
    # null particle set -> VPM_ERR_ARG, message set, nothing launched
    rc = lib.vpm_get_selfinduced_velocity(None, 0, None, 0, None)
    assert rc == L.VPM_ERR_ARG
    assert b"null" in lib.vpm_last_error()
    p = L.Particles(10, 5, None, None, None, None)  # ld < n
    rc = lib.vpm_get_strength_gradients(ctypes.byref(p), None, 10, None)
    assert rc == L.VPM_ERR_ARG
    assert b"ld" in lib.vpm_last_error()
    opts = L.ForwardOpts(0, L.PREC_F32C, L.SIGMA_UNIFORM, 0, -1)
    out = L.ForwardOut(0, None, None, None, None, None)
    p0 = L.Particles(0, 0, None, None, None, None)
    rc = lib.vpm_forward(ctypes.byref(p0), ctypes.byref(opts), ctypes.byref(out), None, 0, None)
    assert rc == L.VPM_ERR_ARG  # empty output mask
    with pytest.raises(L.VpmError):
        L.call("vpm_tile_config_info", 1000, None, None, None, None, None)


def test_tile_config_defaults():
    """tile_config = -1 picks one shape per output set: 512 lanes with
    128-source partials for velocity + stretching, 256 lanes with 256-source
    tiles, 128-source chunks and direct FP64 folds for velocity + stretching
    + PSE, 256 lanes for the other sets with PSE or the LES outputs; the
    batch uses the 64-source-chunk twin of the PSE shape with PSE or LES,
    512 lanes without."""
    lib = L.load()
    assert lib.vpm_tile_config_count() == 12
    assert lib.vpm_tile_config_default(L.OUT_VEL | L.OUT_STRETCH) == 5
    assert lib.vpm_tile_config_default(L.OUT_VEL | L.OUT_STRETCH | L.OUT_PSE) == 11
    assert lib.vpm_tile_config_default(31) == 1
    assert lib.vpm_tile_config_default(27) == 1
    assert lib.vpm_tile_config_default(L.OUT_VEL) == 0
    assert lib.vpm_batch_tile_config_default(31) == 4
    assert lib.vpm_batch_tile_config_default(27) == 4
    assert lib.vpm_batch_tile_config_default(L.OUT_VEL | L.OUT_STRETCH) == 0
    vals = [ctypes.c_int() for _ in range(5)]
    for idx, want_block in ((0, 512), (1, 256)):
        assert L.call("vpm_tile_config_info", idx, *[ctypes.byref(v) for v in vals]) == 0
        assert vals[0].value == want_block
