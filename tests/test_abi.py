"""The C ABI library loads and exports every entry point include/chanles.h
declares; the Fortran host layer is built and linked against it.  No compute
call is made (these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import channel_rl_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "chanles.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(chan_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_declares_the_entry_points():
    names = declared()
    assert "chan_compute_time_step_RK3" in names and "chan_create" in names
    assert set(names) == set(channel_rl_amd.EXPORTS), set(names) ^ set(channel_rl_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(channel_rl_amd.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", channel_rl_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (chan_[A-Za-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_fortran_driver_links_the_library():
    exe = os.path.join(ROOT, "channel_rl_amd", "channel_driver")
    assert os.path.exists(exe), "channel_driver not built"
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libchanles.so" in out and "not found" not in out
    syms = subprocess.run(["nm", exe], capture_output=True, text=True).stdout
    for n in ("chan_create", "chan_compute_time_step_RK3", "chan_compute_dt"):
        assert n in syms


def test_python_front_is_loud_without_library(tmp_path):
    import pytest
    with pytest.raises(channel_rl_amd.ChanError):
        channel_rl_amd._LIB_saved = channel_rl_amd._LIB
        channel_rl_amd._LIB = None
        try:
            channel_rl_amd.load_library(str(tmp_path / "missing.so"))
        finally:
            channel_rl_amd._LIB = channel_rl_amd._LIB_saved


def test_create_rejects_invalid_grids():
    """chan_create's argument checks run before any device call, so they are
    exercised here: odd nx, too small a grid, periodic lengths that do not
    factor into 2, 3, 5 or exceed one block's line (2048)."""
    import numpy as np
    import pytest
    from channel_rl_amd import Channel, ChanError

    def grid(nx, ny, nz):
        return (np.linspace(0, 6.28, nx), 1.0 + np.tanh(1.5 * np.linspace(-1, 1, ny)) / np.tanh(1.5),
                np.linspace(0, 3.14, nz))

    for (nx, ny, nz), msg in [((33, 17, 34), "nx must be even"), ((4, 17, 34), "grid too small"),
                              ((16, 17, 36), "factor into 2, 3 and 5"), ((2162, 9, 34), "at most 2048"),
                              ((34, 9, 2162), "at most 2048")]:
        with pytest.raises(ChanError, match=msg):
            Channel(*grid(nx, ny, nz), nu=1e-3)
    with pytest.raises(ChanError, match="LES_model"):
        Channel(*grid(34, 17, 34), nu=1e-3, LES_model=1)
