"""Green's table cache files (initialize.f90:102-175), host side: the
reference's file names and the read entry point's checks of the file, which
run before any device work (CPU only)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cube2_amd import _lib
from cube2_amd.geometry import RATIO_SF
from cube2_amd.kick import green_file_name


def test_file_names_follow_the_reference():
    # output_dir()//'Gk3_<rsf>'//output_suffix(), output_suffix = '_<image>.bin'
    names = [green_file_name(f"Gk3_{int(RATIO_SF[L - 1])}") for L in range(2, 7)]
    assert names == ["Gk3_2_1.bin", "Gk3_4_1.bin", "Gk3_6_1.bin", "Gk3_8_1.bin", "Gk3_12_1.bin"]
    assert green_file_name("Gk2", image=3) == "Gk2_3.bin"


@pytest.fixture(scope="module")
def L():
    return _lib.lib()


def test_read_rejects_missing_and_short_files(L, tmp_path):
    n = 8
    dummy = C.c_void_p(16)          # never dereferenced: the file checks fail first
    assert L.cube2_green_file_read_dev(dummy, n, str(tmp_path / "none.bin").encode()) == _lib.ERR_ARG
    short = tmp_path / "short.bin"
    np.zeros(n * n * (n // 2 + 1) - 1, np.float32).tofile(short)
    assert L.cube2_green_file_read_dev(dummy, n, str(short).encode()) == _lib.ERR_ARG
    assert b"short.bin" in L.cube2_last_error()
    assert L.cube2_green_file_read_dev(dummy, 7, str(short).encode()) == _lib.ERR_ARG   # odd n
    assert L.cube2_green_file_write_dev(None, n, str(short).encode()) == _lib.ERR_ARG
