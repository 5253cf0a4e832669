"""Particle layout of one image: the inputs of the kick path.

Restates, for a single periodic image, what ``buffer_grid`` and
``spine_image`` produce (buffer_grid.f90:13-155, variables.f90:146-216):

* ``rhoc[tz][ty][tx][k][j][i]`` (int32): particles per coarse cell of each
  tile's extended region (``nte = nt + 2 ncb`` per dim, halo included);
* ``idx_b_r[tz][ty][tx][k][j]`` (int64): cumulative particle count at the end
  of each (j,k) row, tiles in (tz,ty,tx) order;
* ``xp`` (int16 [N_used][3]): position inside the coarse cell, decoded as
  ``((int16)(xp - 32768) + 32768.5)/65536`` (parameters.f90:87-88,118);
* ``vp`` (float32 [N_used][3]).

Halo cells hold copies of the periodic neighbours' particles, so a slot
index ``ip`` is unique per (tile, cell) and ``N_used = idx_b_r[-1,-1,-1,-1,-1]``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .geometry import Geometry


@dataclass
class Layout:
    geom: Geometry
    rhoc: np.ndarray      # int32 [ntile, nte, nte, nte]
    idx_b_r: np.ndarray   # int64 [ntile, nte, nte]
    xp: np.ndarray        # int16 [n_used, 3]
    vp: np.ndarray        # float32 [n_used, 3]
    src: np.ndarray       # int64 [n_used] source particle of each slot
    n_particles: int

    @property
    def n_used(self) -> int:
        return int(self.idx_b_r.reshape(-1)[-1]) if self.idx_b_r.size else 0

    def decoded_positions(self) -> np.ndarray:
        """Positions (standard cells, float64) of every slot, as the kernels decode them."""
        g = self.geom
        nte, ncb = g.nte, g.ncb
        counts = self.rhoc.reshape(-1).astype(np.int64)
        cell = np.repeat(np.arange(counts.size, dtype=np.int64), counts)
        i = cell % nte
        j = (cell // nte) % nte
        k = (cell // (nte * nte)) % nte
        t = cell // (nte ** 3)
        tx = t % g.nnt
        ty = (t // g.nnt) % g.nnt
        tz = t // (g.nnt * g.nnt)
        u = self.xp.view(np.uint16).astype(np.float64)
        frac = (u + 0.5) / 65536.0
        base = np.stack([tx * g.nt + i - ncb, ty * g.nt + j - ncb, tz * g.nt + k - ncb], axis=1)
        return (base + frac) * g.ratio_cs

    def physical_mask(self) -> np.ndarray:
        g = self.geom
        nte, ncb, nt = g.nte, g.ncb, g.nt
        idx = np.arange(nte)
        ph = (idx >= ncb) & (idx < ncb + nt)
        cellphys = (ph[None, :, None, None] & ph[None, None, :, None] & ph[None, None, None, :])
        cellphys = np.broadcast_to(cellphys, self.rhoc.shape).reshape(-1)
        return np.repeat(cellphys, self.rhoc.reshape(-1).astype(np.int64))


def build_layout_device(pos, vel, geom: Geometry):
    """The same layout built on the device by the library
    (cube2_layout_build_dev).  pos: float64 [n, 3] (numpy or a ROCm tensor),
    vel: float32 [n, 3] or None.  Returns device tensors
    (rhoc, idx_b_r, xp, vp) shaped as in ``Layout`` and n_used."""
    import ctypes as C

    import torch

    from . import _lib
    geom.validate()
    _lib.set_params(geom)
    dev = torch.device("cuda")
    d_pos = torch.as_tensor(np.ascontiguousarray(pos, np.float64) if isinstance(pos, np.ndarray)
                            else pos, dtype=torch.float64, device=dev).contiguous()
    n = d_pos.shape[0]
    d_vel = None
    if vel is not None:
        d_vel = torch.as_tensor(np.ascontiguousarray(vel, np.float32) if isinstance(vel, np.ndarray)
                                else vel, dtype=torch.float32, device=dev).contiguous()
    rhoc = torch.empty((geom.ntile, geom.nte, geom.nte, geom.nte), dtype=torch.int32, device=dev)
    idx = torch.empty((geom.ntile, geom.nte, geom.nte), dtype=torch.int64, device=dev)
    n_used = C.c_int64(0)
    L = _lib.lib()
    rc = L.cube2_layout_build_dev(d_pos.data_ptr(), None if d_vel is None else d_vel.data_ptr(), n,
                                  rhoc.data_ptr(), idx.data_ptr(), None, None, 0,
                                  C.byref(n_used))
    if rc not in (_lib.OK, _lib.ERR_CAPACITY):
        _lib.check(rc)
    m = max(n_used.value, 1)
    xp = torch.empty((m, 3), dtype=torch.int16, device=dev)
    vp = torch.empty((m, 3), dtype=torch.float32, device=dev)
    _lib.check(L.cube2_layout_build_dev(d_pos.data_ptr(), None if d_vel is None else d_vel.data_ptr(),
                                        n, rhoc.data_ptr(), idx.data_ptr(), xp.data_ptr(),
                                        vp.data_ptr(), m, C.byref(n_used)))
    k = n_used.value
    return rhoc, idx, xp[:k], vp[:k], k


def encode_positions(pos: np.ndarray, geom: Geometry):
    """Global coarse cell and int16 code of each position (standard cells)."""
    ng = geom.ng
    p = np.mod(np.asarray(pos, dtype=np.float64), ng)
    pc = p / geom.ratio_cs
    gc = np.floor(pc).astype(np.int64)
    gc = np.clip(gc, 0, geom.nc - 1)
    frac = pc - gc
    u = np.clip(np.floor(frac * 65536.0), 0, 65535).astype(np.uint16)
    return gc, u.view(np.int16)


def build_layout(pos: np.ndarray, vel: np.ndarray | None, geom: Geometry) -> Layout:
    """Lay out N particles (positions in standard cells, periodic in [0, ng))
    as one image with halos of ``ncb`` coarse cells around every tile."""
    geom.validate()
    n = pos.shape[0]
    nc, nt, ncb, nte, nnt = geom.nc, geom.nt, geom.ncb, geom.nte, geom.nnt
    gc, code = encode_positions(pos, geom)
    gid = (gc[:, 2] * nc + gc[:, 1]) * nc + gc[:, 0]
    order = np.argsort(gid, kind="stable")
    cnt = np.bincount(gid, minlength=nc ** 3).astype(np.int64)
    gstart = np.zeros(nc ** 3 + 1, dtype=np.int64)
    np.cumsum(cnt, out=gstart[1:])

    t = np.arange(nnt)
    e = np.arange(nte)
    # global coarse coordinate of each (tile, extended cell) along one axis
    ax = (t[:, None] * nt + e[None, :] - ncb) % nc            # [nnt, nte]
    gx = ax[None, None, :, None, None, :]
    gy = ax[None, :, None, None, :, None]
    gz = ax[:, None, None, :, None, None]
    g_of_cell = ((gz * nc + gy) * nc + gx).reshape(-1)          # [tz,ty,tx,k,j,i]
    rhoc = cnt[g_of_cell].astype(np.int32)
    counts = rhoc.astype(np.int64)
    slot_start = np.zeros(counts.size + 1, dtype=np.int64)
    np.cumsum(counts, out=slot_start[1:])
    n_used = int(slot_start[-1])
    rows_end = slot_start[1:].reshape(-1, nte)[:, -1]
    idx_b_r = rows_end.reshape(nnt, nnt, nnt, nte, nte).reshape(nnt ** 3, nte, nte)

    cell_of_slot = np.repeat(np.arange(counts.size, dtype=np.int64), counts)
    local = np.arange(n_used, dtype=np.int64) - slot_start[cell_of_slot]
    src = order[gstart[g_of_cell[cell_of_slot]] + local]
    del cell_of_slot, local
    xp = np.ascontiguousarray(code[src])
    if vel is None:
        vp = np.zeros((n_used, 3), dtype=np.float32)
    else:
        vp = np.ascontiguousarray(np.asarray(vel, dtype=np.float32)[src])
    return Layout(geom=geom, rhoc=rhoc.reshape(nnt ** 3, nte, nte, nte),
                  idx_b_r=np.ascontiguousarray(idx_b_r.astype(np.int64)),
                  xp=xp, vp=vp, src=src, n_particles=n)
