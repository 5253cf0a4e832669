"""ctypes binding of libcube2_short.so (the HIP library).

The product path has no fallback: if the library is missing or fails to
load, every call raises.  Device buffers are passed as raw pointers
(e.g. ``torch.Tensor.data_ptr()`` of a ROCm tensor); no torch types cross
the C ABI.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libcube2_short.so"
_LIB = None

OK, ERR_ARG, ERR_GEOMETRY, ERR_CAPACITY, ERR_HIP, ERR_STATE = range(6)
PP_EXACT, PP_FAST, PP_REF, PP_PRECISE, PP_PLUMMER = 0, 1, 2, 3, 4
(TUNE_PP_BLOCKS_PER_CU, TUNE_PP_DIRECT_LG, TUNE_FFT_MIN_LGC, TUNE_INTERP_SEG, TUNE_PP_DENSE_LG,
 TUNE_KICK_OVERLAP, TUNE_PP_OVERLAP_BLOCKS, TUNE_PP_ORDER, TUNE_PP_JBATCH,
 TUNE_PP_NEWTON) = range(10)
TUNE_NAMES = {"pp_blocks_per_cu": 0, "pp_direct_lg": 1, "fft_min_lgc": 2, "interp_seg": 3,
              "pp_dense_lg": 4, "kick_overlap": 5, "pp_overlap_blocks": 6, "pp_order": 7,
              "pp_jbatch": 8, "pp_newton": 9}


class Cube2Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cube2 error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("ng", C.c_int), ("nnt", C.c_int), ("nns", C.c_int), ("ratio_cs", C.c_int),
                ("np_nc", C.c_int), ("ngb", C.c_int), ("ncore", C.c_int),
                ("body_centered_cubic", C.c_int), ("image_buffer", C.c_double),
                ("tile_buffer", C.c_double), ("app", C.c_double)]


class KickResultC(C.Structure):
    _fields_ = [("f2max_pm2", C.c_float), ("f2max_pm3", C.c_float), ("f2max_pp", C.c_float),
                ("vmax", C.c_float * 3),
                ("dt_pm2", C.c_double), ("dt_pm3", C.c_double), ("dt_pp", C.c_double),
                ("dt_vmax", C.c_double),
                ("pt_pm3", C.c_double), ("pt_pp", C.c_double), ("pair_counts", C.c_uint64 * 2),
                ("napm", C.c_int * 7)]


KICK_PM3, KICK_PP, KICK_KEEP_LEVELS, KICK_PM2, KICK_LEARNED = 1, 2, 4, 8, 16

vp_ = C.c_void_p
i32 = C.POINTER(C.c_int)

# name -> (restype, argtypes); the exported surface of include/cube2_short.h
SIGNATURES = {
    "c_pp_force_kernel": (None, [vp_] * 9 + [C.c_float] * 3 + [vp_] * 4),
    "c_pm3_force_kernel": (None, [vp_] * 13 + [C.c_float] * 3 + [vp_] * 3),
    "c_pm3_fft_finalize": (None, []),
    "c_gpu_velocity_finalize": (None, []),
    "cube2_set_params": (C.c_int, [C.POINTER(Params)]),
    "cube2_get_params": (C.c_int, [C.POINTER(Params)]),
    "cube2_set_pp_mode": (C.c_int, [C.c_int]),
    "cube2_set_tuning": (C.c_int, [C.c_int, C.c_int]),
    "cube2_get_tuning": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    "cube2_pm3_force_dev": (C.c_int, [vp_] * 7 + [C.c_int64, vp_] + [C.c_float] * 3 + [vp_, vp_]),
    "cube2_pp_force_dev": (C.c_int, [vp_] * 8 + [C.c_int64] + [C.c_float] * 3 + [vp_] * 5),
    "cube2_velocity_apply_dev": (C.c_int, [vp_, vp_, vp_, C.c_int64, vp_]),
    "cube2_pm2_force_dev": (C.c_int, [vp_] * 3 + [C.c_int64, vp_] + [C.c_float] * 3 + [vp_] * 2),
    "cube2_green_3d_dev": (C.c_int, [vp_, C.c_int, C.c_float, C.c_float, C.c_float]),
    "cube2_green_file_write_dev": (C.c_int, [vp_, C.c_int, C.c_char_p]),
    "cube2_green_file_read_dev": (C.c_int, [vp_, C.c_int, C.c_char_p]),
    "cube2_gemm_bf16_dev": (C.c_int, [vp_, C.c_int64, C.c_int64, vp_, C.c_int64, C.c_int64, vp_,
                                      C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cube2_gemm_bf16_noalloc_dev": (C.c_int, [vp_, C.c_int64, C.c_int64, vp_, C.c_int64, C.c_int64,
                                              vp_, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cube2_cnn_conv_pack_dev": (C.c_int, [vp_, vp_, C.c_int, C.c_int]),
    "cube2_cnn_conv3_dev": (C.c_int, [vp_, C.c_int, vp_, vp_, vp_, C.c_int] + [C.c_int] * 5),
    "cube2_learned_set": (C.c_int, [vp_, vp_, C.c_float, C.c_int]),
    "cube2_learned_clear": (C.c_int, []),
    "cube2_learned_inputs_dev": (C.c_int, [vp_, C.c_int, vp_, vp_]),
    "cube2_learned_samples_dev": (C.c_int, [C.c_int, C.c_int, vp_, vp_, vp_, vp_, vp_]),
    "cube2_learned_debug_last": (C.c_int, [vp_, vp_, C.c_int, vp_, vp_, vp_, vp_]),
    "cube2_learned_gather_dev": (C.c_int, [vp_, C.c_int, vp_, vp_, C.c_int, vp_]),
    "cube2_cnn_im2col_dev": (C.c_int, [vp_, vp_, C.c_int, C.c_int, C.c_int]),
    "cube2_cnn_col2im_dev": (C.c_int, [vp_, vp_, C.c_int, C.c_int, C.c_int]),
    "cube2_cnn_bias_act_dev": (C.c_int, [vp_, vp_, C.c_int64, C.c_int, C.c_int]),
    "cube2_cnn_act_bwd_dev": (C.c_int, [vp_, vp_, vp_, vp_, C.c_int64, C.c_int, C.c_int]),
    "cube2_to_bf16_dev": (C.c_int, [vp_, vp_, C.c_int64]),
    "cube2_cnn_adam_dev": (C.c_int, [vp_, vp_, vp_, vp_, vp_, C.c_int64, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_int]),
    "cube2_cnn_mse_dev": (C.c_int, [vp_, vp_, vp_, vp_, C.c_int64]),
    "cube2_cnn_cic_mse_dev": (C.c_int, [vp_, vp_, vp_, C.c_int, vp_, vp_, vp_]),
    "cube2_select_levels": (C.c_int, [vp_] * 5),
    "cube2_layout_build_dev": (C.c_int, [vp_, vp_, C.c_int64] + [vp_] * 4 + [C.c_int64, vp_]),
    "cube2_drift_dev": (C.c_int, [vp_] * 4 + [C.c_int64, C.c_float, C.c_float] + [vp_] * 4 +
                        [C.c_int64, vp_]),
    "cube2_slot_positions_dev": (C.c_int, [vp_] * 3 + [C.c_int64, vp_]),
    "cube2_halo_velocities_dev": (C.c_int, [vp_] * 3 + [C.c_int64]),
    "cube2_kick_dev": (C.c_int, [C.c_int] + [vp_] * 7 + [C.c_int64, vp_] + [C.c_float] * 3
                       + [vp_] * 5 + [C.POINTER(KickResultC)]),
    "cube2_kick_graph_dev": (C.c_int, [C.c_int] + [vp_] * 7 + [C.c_int64, vp_] + [C.c_float] * 3
                             + [vp_] * 5 + [C.POINTER(KickResultC), vp_]),
    "cube2_pp_debug_celllist": (C.c_int, [C.c_int] + [vp_] * 6),
    "cube2_pp_count_tiles": (C.c_int, [C.c_int]),
    "cube2_pp_debug_tile_pairs": (C.c_int, [C.c_int, vp_]),
    "cube2_pp_debug_pair_forces": (C.c_int, [C.c_int, vp_, C.c_int, C.c_float, vp_]),
    "cube2_pm3_debug_job": (C.c_int, [C.c_int] + [vp_] * 6 + [C.c_int64, vp_, C.c_float] + [vp_] * 4),
    "cube2_fft3d_dev": (C.c_int, [vp_, C.c_int, C.c_int, C.c_int]),
    "cube2_selftest_div": (C.c_int, [C.c_uint64, C.c_int, vp_]),
    "cube2_profile_enable": (C.c_int, [C.c_int]),
    "cube2_profile_reset": (C.c_int, []),
    "cube2_profile_get": (C.c_int, [C.c_char_p, vp_, vp_, vp_, vp_]),
    "cube2_profile_stages": (C.c_int, [C.c_char_p, C.c_int]),
    "cube2_synchronize": (C.c_int, []),
    "cube2_last_error": (C.c_char_p, []),
}


def _torch_runtime_first() -> None:
    """PyTorch-ROCm wheels carry their own HIP and HSA runtimes, so a Python
    process that uses torch holds two.  torch's runtime must open the device
    before the library's: with the library loaded first, its
    hipGetDeviceCount reports no device once torch has initialised.  Loading
    torch and querying its device first keeps the working order (the query
    returns False where there is no GPU)."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() or python cube2_amd/build.py")
        _torch_runtime_first()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != OK:
        raise Cube2Error(rc, lib().cube2_last_error().decode(errors="replace"))


def ptr(x) -> int | None:
    """Address of a numpy array or a torch tensor (host or device)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        assert x.flags["C_CONTIGUOUS"], "array must be C-contiguous"
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        assert x.is_contiguous(), "tensor must be contiguous"
        return x.data_ptr()
    if isinstance(x, int):
        return x
    raise TypeError(type(x))


def set_params(geom) -> None:
    p = Params(geom.ng, geom.nnt, geom.nns, geom.ratio_cs, geom.np_nc, geom.ngb, geom.ncore,
               int(geom.body_centered_cubic), geom.image_buffer, geom.tile_buffer, geom.app)
    check(lib().cube2_set_params(C.byref(p)))


def get_params() -> Params:
    p = Params()
    check(lib().cube2_get_params(C.byref(p)))
    return p


def set_pp_mode(mode: int) -> None:
    check(lib().cube2_set_pp_mode(int(mode)))


def set_tuning(key: int, value: int) -> None:
    check(lib().cube2_set_tuning(int(key), int(value)))


def get_tuning(key: int) -> int:
    v = C.c_int()
    check(lib().cube2_get_tuning(int(key), C.byref(v)))
    return v.value


def profile(on: bool = True) -> None:
    check(lib().cube2_profile_enable(int(on)))


def profile_reset() -> None:
    check(lib().cube2_profile_reset())


def profile_stats() -> dict:
    buf = C.create_string_buffer(4096)
    check(lib().cube2_profile_stages(buf, 4096))
    out = {}
    for name in filter(None, buf.value.decode().split(",")):
        ms, n, b, f = C.c_double(), C.c_int64(), C.c_double(), C.c_double()
        check(lib().cube2_profile_get(name.encode(), C.byref(ms), C.byref(n), C.byref(b), C.byref(f)))
        out[name] = dict(ms=ms.value, launches=n.value, bytes=b.value, flops=f.value)
    return out


def synchronize() -> None:
    check(lib().cube2_synchronize())


def pp_count_tiles(on: bool = True) -> None:
    check(lib().cube2_pp_count_tiles(int(on)))


def pp_tile_pairs(ntile: int) -> np.ndarray:
    """{P_cand, P_acc} per parent tile of the last PP call (counting on)."""
    out = np.zeros((ntile, 2), np.uint64)
    check(lib().cube2_pp_debug_tile_pairs(int(ntile), out.ctypes.data))
    return out


def pp_pair_forces(mode: int, d: np.ndarray, ranges: np.ndarray, app: float) -> np.ndarray:
    """The pair kernel's force for pairs d[n][3] with per-pair ranges."""
    n = d.shape[0]
    pairs = np.zeros((n, 4), np.float32)
    pairs[:, :3] = d
    pairs[:, 3] = ranges
    out = np.zeros((n, 3), np.float32)
    check(lib().cube2_pp_debug_pair_forces(int(mode), pairs.ctypes.data, n, float(app),
                                           out.ctypes.data))
    return out
