"""The C-ABI library loads and exports every declared symbol; host-side
argument validation runs without a GPU (CPU only)."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT


def declared_functions() -> list[str]:
    text = (ROOT / "include" / "cube2_short.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(c_[a-z0-9_]+|cube2_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from cube2_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert {"c_pp_force_kernel", "c_pm3_force_kernel", "c_pm3_fft_finalize",
            "c_gpu_velocity_finalize"} <= set(names)
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "Python bindings out of sync with the header"


def test_params_roundtrip_and_validation():
    from cube2_amd import _lib
    from cube2_amd.geometry import C1, REFERENCE_DEFAULT
    p = _lib.get_params()
    _lib.set_params(C1)
    q = _lib.get_params()
    assert (q.ng, q.nnt, q.nns, q.ngb) == (C1.ng, C1.nnt, C1.nns, C1.ngb)
    bad = _lib.Params(100, 3, 1, 4, 4, 16, 1, 0, 2.0, 3.0, 0.06)   # 25 coarse cells / 3 tiles
    assert _lib.lib().cube2_set_params(C.byref(bad)) == _lib.ERR_GEOMETRY
    _lib.set_params(REFERENCE_DEFAULT)
    del p


def test_pp_rejects_bad_metadata_before_device_work():
    from cube2_amd import _lib
    from cube2_amd.geometry import APM3, RATIO_SF, C0, ixyz3
    _lib.set_params(C0)
    ires = np.full(C0.ns3, 7, np.int32)          # outside 2..6
    ix = ixyz3(C0)
    f2 = np.zeros(1, np.float32)
    vm = np.zeros(3, np.float32)
    pt = np.zeros(C0.ns3)
    rc = _lib.lib().cube2_pp_force_dev(ires.ctypes.data, ix.ctypes.data, APM3.ctypes.data,
                                      RATIO_SF.ctypes.data, None, None, None, None, 0, 1.0, 0.02,
                                      1e-3, f2.ctypes.data, vm.ctypes.data, pt.ctypes.data, None,
                                      None)
    assert rc == _lib.ERR_ARG
    assert b"ires" in _lib.lib().cube2_last_error()
    ires[:] = 2
    ix2 = ix.copy()
    ix2[0, 3] = 5                                  # tile index beyond nnt
    rc = _lib.lib().cube2_pp_force_dev(ires.ctypes.data, ix2.ctypes.data, APM3.ctypes.data,
                                      RATIO_SF.ctypes.data, None, None, None, None, 0, 1.0, 0.02,
                                      1e-3, f2.ctypes.data, vm.ctypes.data, pt.ctypes.data, None,
                                      None)
    assert rc == _lib.ERR_ARG


def test_dropin_pp_exits_on_bad_metadata():
    # The reference terminates the process on invalid PP metadata
    # (pp_force_kernel.cu:979-987); the drop-in does the same.
    code = f"""
import sys, numpy as np
sys.path.insert(0, {str(ROOT)!r})
from cube2_amd import _lib
from cube2_amd.geometry import APM3, RATIO_SF, C0, ixyz3
_lib.set_params(C0)
ix = ixyz3(C0); ires = np.full(C0.ns3, 9, np.int32); isort = np.arange(1, C0.ns3+1, dtype=np.int32)
rhoc = np.zeros((1, C0.nte, C0.nte, C0.nte), np.int32); idx = np.zeros((1, C0.nte, C0.nte), np.int64)
xp = np.zeros((1, 3), np.int16); vp = np.zeros((1, 3), np.float32)
f2 = np.zeros(1, np.float32); vm = np.zeros(3, np.float32); pt = np.zeros(C0.ns3); tt = np.zeros(C0.ns3)
a = lambda x: x.ctypes.data
_lib.lib().c_pp_force_kernel(a(isort), a(ires), a(ix), a(APM3), a(RATIO_SF), a(rhoc), a(idx), a(xp), a(vp),
                             1.0, 0.02, 1e-3, a(f2), a(vm), a(pt), a(tt))
print("returned")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "returned" not in r.stdout
    assert "invalid PP subtile metadata" in r.stderr


def test_select_levels_matches_oracle(oracle_lib):
    from cube2_amd.geometry import REFERENCE_DEFAULT, ixyz3
    from cube2_amd.kick import select_levels
    g = REFERENCE_DEFAULT.with_(nnt=2, ng=128, nns=2)
    rng = np.random.default_rng(9)
    rhoc = rng.integers(0, 5, (g.ntile, g.nte, g.nte, g.nte)).astype(np.int32)
    hot = rng.integers(0, rhoc.size, 40)
    rhoc.reshape(-1)[hot] = rng.integers(1000, 100000, 40)
    ires, isort, napm = select_levels(g, rhoc)
    o_ires, o_isort, o_napm, _ = oracle_lib.select_levels(g, rhoc, ixyz3(g))
    assert np.array_equal(ires, o_ires)
    assert np.array_equal(isort, o_isort)
    assert np.array_equal(napm, o_napm)


def test_pp_accepts_reference_ratio_table_without_gpu():
    # ratio_sf carries 16 for the unused level 7 (parameters.f90:22); valid
    # metadata passes validation and only then needs the device.
    import torch
    from cube2_amd import _lib
    from cube2_amd.geometry import APM3, RATIO_SF, C0, ixyz3
    if torch.cuda.is_available():
        pytest.skip("device present: the call would run")
    _lib.set_params(C0)
    ires = np.full(C0.ns3, 2, np.int32)
    ix = ixyz3(C0)
    f2 = np.zeros(1, np.float32)
    vm = np.zeros(3, np.float32)
    pt = np.zeros(C0.ns3)
    rc = _lib.lib().cube2_pp_force_dev(ires.ctypes.data, ix.ctypes.data, APM3.ctypes.data,
                                      RATIO_SF.ctypes.data, None, None, None, None, 0, 1.0, 0.02,
                                      1e-3, f2.ctypes.data, vm.ctypes.data, pt.ctypes.data, None,
                                      None)
    assert rc == _lib.ERR_HIP, _lib.lib().cube2_last_error()


def test_kick_driver_validates_before_device_work():
    import ctypes as C
    from cube2_amd import _lib
    from cube2_amd.geometry import APM3, RATIO_SF, C0, ixyz3
    _lib.set_params(C0)
    ix = ixyz3(C0)
    ires = np.full(C0.ns3, 7, np.int32)
    isort = np.arange(1, C0.ns3 + 1, dtype=np.int32)
    out = _lib.KickResultC()
    a = lambda x: x.ctypes.data  # noqa: E731
    L = _lib.lib()
    args = (a(ix), a(APM3), a(RATIO_SF), None, None, None, None, 0, None, 1.0, 0.02, 1e-3,
            a(ires), a(isort), None, None, None, C.byref(out))
    assert L.cube2_kick_dev(_lib.KICK_PP | _lib.KICK_KEEP_LEVELS, *args) == _lib.ERR_ARG
    assert b"outside 2..6" in L.cube2_last_error()
    assert L.cube2_kick_dev(64, *args) == _lib.ERR_ARG          # unknown flag
    ires[:] = 2
    ix2 = ix.copy()
    ix2[0, 0] = 0
    args2 = (a(ix2),) + args[1:]
    assert L.cube2_kick_dev(_lib.KICK_PP | _lib.KICK_KEEP_LEVELS, *args2) == _lib.ERR_ARG
    assert b"ixyz3" in L.cube2_last_error()


def test_tuning_api_validates_and_round_trips():
    from cube2_amd import _lib
    L = _lib.lib()
    assert L.cube2_set_tuning(99, 1) == _lib.ERR_ARG
    assert L.cube2_set_tuning(_lib.TUNE_PP_DIRECT_LG, 9) == _lib.ERR_ARG
    assert L.cube2_set_tuning(_lib.TUNE_INTERP_SEG, 33) == _lib.ERR_ARG
    _lib.set_tuning(_lib.TUNE_INTERP_SEG, 16)
    assert _lib.get_tuning(_lib.TUNE_INTERP_SEG) == 16
    _lib.set_tuning(_lib.TUNE_INTERP_SEG, 0)
    assert _lib.get_tuning(_lib.TUNE_INTERP_SEG) == 4          # built-in choice
    assert _lib.get_tuning(_lib.TUNE_PP_DIRECT_LG) == 2
    assert _lib.get_tuning(_lib.TUNE_PP_DENSE_LG) == 4
    assert _lib.get_tuning(_lib.TUNE_PP_ORDER) == 1
    assert L.cube2_set_tuning(_lib.TUNE_PP_ORDER, 31) == _lib.ERR_ARG
    assert _lib.get_tuning(_lib.TUNE_PP_JBATCH) == 1
    assert L.cube2_set_tuning(_lib.TUNE_PP_JBATCH, 3) == _lib.ERR_ARG
    assert _lib.get_tuning(_lib.TUNE_PP_NEWTON) == 1
    assert L.cube2_set_tuning(_lib.TUNE_PP_NEWTON, 3) == _lib.ERR_ARG
    assert L.cube2_set_tuning(_lib.TUNE_PP_DENSE_LG, 2) == _lib.ERR_ARG
