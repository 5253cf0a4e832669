"""Drift of one image on the device (drift.f90:1-158), through the library.

``drift_device`` moves every slot of the image (halo copies included) by
``dt_mid * vp`` and returns the new physical-only image: ``rhoc`` with
empty halo cells, ``idx_b_r`` (spine_image's running row totals) and the
compacted ``xp``/``vp`` in the reference's order.  ``next_layout`` turns
that image into the layout the next kick takes (halo copies rebuilt by the
device layout builder from the decoded positions), the device form of
``buffer_grid``/``buffer_x`` after ``drift``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .geometry import Geometry


def refresh_halo_velocities(rhoc, idx_b_r, vp, n_used: int, geom: Geometry) -> None:
    """buffer_v for one image (cube2_halo_velocities_dev): halo slots take
    their source particles' velocities, in place.  A kick updates physical
    slots only; the reference's step runs kick, then buffer_v, before the
    next drift (main.f90:26-31)."""
    _lib.set_params(geom)
    _lib.check(_lib.lib().cube2_halo_velocities_dev(rhoc.data_ptr(), idx_b_r.data_ptr(),
                                                     vp.data_ptr(), int(n_used)))


def drift_device(rhoc, idx_b_r, xp, vp, n_used: int, dt_mid: float, vz_max: float,
                 geom: Geometry):
    """Device tensors in, device tensors out: (rhoc, idx_b_r, xp, vp, n).
    vz_max must bound |vp_z| (sim%vz_max, kick.f90:239)."""
    import torch
    geom.validate()
    _lib.set_params(geom)
    dev = rhoc.device
    rhoc_out = torch.empty_like(rhoc)
    idx_out = torch.empty_like(idx_b_r)
    m = max(int(n_used), 1)
    xp_out = torch.empty((m, 3), dtype=torch.int16, device=dev)
    vp_out = torch.empty((m, 3), dtype=torch.float32, device=dev)
    n = C.c_int64(0)
    _lib.check(_lib.lib().cube2_drift_dev(
        rhoc.data_ptr(), idx_b_r.data_ptr(), xp.data_ptr(), vp.data_ptr(), int(n_used),
        float(dt_mid), float(vz_max), rhoc_out.data_ptr(), idx_out.data_ptr(), xp_out.data_ptr(),
        vp_out.data_ptr(), m, C.byref(n)))
    k = n.value
    return rhoc_out, idx_out, xp_out[:k], vp_out[:k], k


def slot_positions_device(rhoc, idx_b_r, xp, n_used: int, geom: Geometry):
    """FP64 positions (standard cells, in [0, ng)) of every slot."""
    import torch
    _lib.set_params(geom)
    pos = torch.empty((max(int(n_used), 1), 3), dtype=torch.float64, device=rhoc.device)
    _lib.check(_lib.lib().cube2_slot_positions_dev(rhoc.data_ptr(), idx_b_r.data_ptr(),
                                                    xp.data_ptr(), int(n_used), pos.data_ptr()))
    return pos[:int(n_used)]


def next_layout(rhoc, idx_b_r, xp, vp, n: int, geom: Geometry):
    """The kick's layout (halo copies included) of a physical-only image:
    its slots are the particles, in order, so their decoded positions and
    velocities rebuild it through the device layout builder."""
    from .layout import build_layout_device
    pos = slot_positions_device(rhoc, idx_b_r, xp, n, geom)
    return build_layout_device(pos, vp, geom)
