"""The C-ABI library: it loads, and it exports every entry point that
include/kennett_hip.h declares (symbol lookup only; no GPU work)."""

from __future__ import annotations

import ctypes
import os
import re

import pytest

from seismicinversion_amd import _abi, _build

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "kennett_hip.h")


def declared_entry_points():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kh_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_build.LIB_PATH):
        _build.build()
    return ctypes.CDLL(_build.LIB_PATH)


def test_header_declares_entry_points():
    names = declared_entry_points()
    assert "kh_forward" in names and "kh_last_error" in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_table_matches_header():
    assert sorted(_abi.EXPORTS) == declared_entry_points()


def test_version_and_error_strings(lib):
    lib.kh_version.restype = ctypes.c_char_p
    lib.kh_last_error.restype = ctypes.c_char_p
    assert b"gfx950" in lib.kh_version()
    assert lib.kh_last_error() == b""


def test_invalid_arguments_are_rejected_before_any_launch(lib):
    """Argument checks run on the host and return a status; no GPU needed."""
    lib.kh_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    rc = lib.kh_forward(vp(0), 1, vp(0), 4, vp(0), 4, 0, vp(0), vp(0))
    assert rc == 1 and b"at least 2 layers" in lib.kh_last_error()
    rc = lib.kh_forward(vp(0), 300, vp(0), 4, vp(0), 4, 0, vp(0), vp(0))
    assert rc == 4 and b"KH_MAX_LAYERS" in lib.kh_last_error()
    rc = lib.kh_forward_ex(vp(16), 3, vp(16), 4, vp(16), 4, 0, vp(16), 96, 1, vp(0))
    assert rc == 1 and b"block" in lib.kh_last_error()
    assert lib.kh_forward(vp(16), 3, vp(16), 0, vp(16), 4, 0, vp(16), vp(0)) == 0


def test_multi_tile_and_item_counts(lib):
    """Host-only helpers of the persistent launch: tiles per row (half-width
    tiles for short rows) and items per model (runs of tiles)."""
    from seismicinversion_amd import batch

    lib.kh_multi_items.restype = ctypes.c_longlong
    assert lib.kh_multi_tiles(1023, 64, 8, 0) == 2
    assert lib.kh_multi_tiles(256, 64, 8, 1) == 1 and lib.kh_multi_tiles(256, 64, 8, 0) == 1
    assert lib.kh_multi_tiles(700, 64, 8, 1) == 3
    assert lib.kh_multi_tiles(130, 64, 1, 1) == 3  # no half width at 1 per lane
    assert lib.kh_multi_tiles(0, 64, 8, 0) == 0
    assert lib.kh_multi_items(5, 8, 0) == 5 and lib.kh_multi_items(5, 8, 8) == 5
    assert lib.kh_multi_items(5, 8, 3) == 15 and lib.kh_multi_items(5, 8, 1) == 40
    assert lib.kh_multi_items(0, 8, 1) == 0
    for nw in (1, 100, 256, 257, 512, 513, 700, 1023, 1024, 3000, 4096):
        for block, wpt in ((64, 8), (128, 4), (256, 2), (64, 1)):
            f = batch.omega_per_lane(nw, block, wpt)
            assert f in (wpt, wpt // 2)
            assert -(-nw // (block * f)) == lib.kh_multi_tiles(nw, block, wpt, int(f < wpt))
    # kh_multi_workgroups checks its arguments before it asks the device
    lib.kh_last_error.restype = ctypes.c_char_p
    out = ctypes.c_int(0)
    assert lib.kh_multi_workgroups(96, 8, 20, 1, ctypes.byref(out)) == 1
    assert b"block" in lib.kh_last_error()
    assert lib.kh_multi_workgroups(64, 3, 20, 0, ctypes.byref(out)) == 1
    assert lib.kh_multi_workgroups(64, 8, 20, 1, None) == 1


def test_multi_table_workspace_is_checked_before_any_launch(lib):
    """With a table workspace, kh_forward_multi stages its per-(slowness,
    layer) tables there; one smaller than kh_multi_table_bytes(records) is
    refused with KH_ERR_OUT_OF_MEMORY, and one not aligned to 256 bytes
    with KH_ERR_INVALID_ARGUMENT, before anything is launched."""
    lib.kh_multi_table_bytes.restype = ctypes.c_size_t
    lib.kh_multi_table_bytes.argtypes = [ctypes.c_longlong]
    lib.kh_last_error.restype = ctypes.c_char_p
    assert lib.kh_multi_table_bytes(0) == 0
    need = lib.kh_multi_table_bytes(1000)
    assert need == 1000 * 256  # one 256-byte record per (model, slowness, layer)
    vp = ctypes.c_void_p
    f = lib.kh_forward_multi
    f.argtypes = [vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, vp, vp, vp, ctypes.c_int,
                  ctypes.c_int, ctypes.c_int, ctypes.c_longlong, vp, ctypes.c_size_t, vp, vp, vp]
    args = [vp(16), 2, 10, 20, vp(16), vp(16), vp(16), 64, 8, 0]
    rc = f(*args, 1000, vp(256), need - 1, vp(16), None, None)
    assert rc == 2 and b"table workspace" in lib.kh_last_error()
    rc = f(*args, 1000, vp(16), need, vp(16), None, None)
    assert rc == 1 and b"aligned" in lib.kh_last_error()
    rc = f(*args, 1000, vp(256), need, None, None, None)  # no counter
    assert rc == 1 and b"NULL" in lib.kh_last_error()
    rc = f(*args[:7], 96, 8, 0, 1000, vp(256), need, vp(16), None, None)
    assert rc == 1 and b"block" in lib.kh_last_error()
