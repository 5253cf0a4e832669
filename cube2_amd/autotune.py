"""Autotuner for the launch-shape tunables of the short-range kick.

For one tunable at a time (in the order of ``CANDIDATES``) it times the
whole kick (``ShortRangeKick.kick``, eager dispatch) for every candidate
value, keeps the fastest and leaves it set in the library; later tunables
are tuned with the earlier choices in place.  A value of 0 is the library's
built-in choice.  The timings are returned so that a caller can report
them.
"""
from __future__ import annotations

import time

from . import _lib

CANDIDATES = {
    "pp_blocks_per_cu": [0, 24, 48, 128],
    "pp_overlap_blocks": [0, 32, 96],
    "pp_dense_lg": [0, 4, 6],
    "pp_direct_lg": [0, 3, 1],
    "fft_min_lgc": [0, 3, 4],
    "interp_seg": [0, 2, 8],
    "pp_order": [0, 20],
    "pp_jbatch": [0, 2],
    "pp_newton": [0, 2],
}


def _time_kick(kick, mass, a_mid, dt, reps):
    import torch
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kick.kick(mass, a_mid, dt)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def autotune(kick, mass: float, a_mid: float, dt: float, reps: int = 2, log=None) -> dict:
    """Tunes in place; returns {"chosen": {name: value}, "ms": {name: {value: ms}}}."""
    chosen, table = {}, {}
    for name, cands in CANDIDATES.items():
        key = _lib.TUNE_NAMES[name]
        table[name] = {}
        for v in cands:
            _lib.set_tuning(key, v)
            kick.kick(mass, a_mid, dt)  # settle buffers and plans for this shape
            table[name][v] = round(1e3 * _time_kick(kick, mass, a_mid, dt, reps), 3)
            if log:
                log(f"autotune {name}={v}: {table[name][v]} ms")
        best = min(table[name], key=table[name].get)
        _lib.set_tuning(key, best)
        chosen[name] = best
    return {"chosen": chosen, "ms": table}
