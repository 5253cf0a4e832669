"""The drift oracle (drift.f90:1-158 for one image) pinned by an identity,
an exact translation whose result the layout builder gives independently,
and per-particle conservation of a random drift (CPU only)."""
from __future__ import annotations

import numpy as np
import pytest

from cube2_amd.geometry import Geometry
from cube2_amd.layout import build_layout

G2 = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)   # ncb = 2, 8 tiles


def _clumpy(g, n, seed):
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0, g.ng, (n // 2, 3))]
    for c in rng.uniform(0, g.ng, (5, 3)):
        pos.append(c + rng.normal(0, 1.0, (n // 10, 3)))
    return np.concatenate(pos) % g.ng


def _row_totals(g, rhoc):
    return np.cumsum(rhoc.reshape(g.ntile * g.nte * g.nte, g.nte).sum(1)).reshape(
        g.ntile, g.nte, g.nte)


def _physical_rhoc(g, rhoc):
    r = np.zeros_like(rhoc).reshape(g.ntile, g.nte, g.nte, g.nte)
    s = slice(g.ncb, g.ncb + g.nt)
    r[:, s, s, s] = rhoc.reshape(r.shape)[:, s, s, s]
    return r.reshape(rhoc.shape)


def test_drift_with_zero_step_keeps_the_physical_image(oracle_lib):
    g = G2
    rng = np.random.default_rng(0)
    pos = _clumpy(g, 6000, 1)
    vel = rng.normal(0, 3, pos.shape).astype(np.float32)
    lay = build_layout(pos, vel, g)
    o = oracle_lib.drift(g, lay.rhoc, lay.idx_b_r, lay.xp, lay.vp, 0.0, 0.0)
    phys = lay.physical_mask()
    assert o["n"] == pos.shape[0]
    assert np.array_equal(o["xp"], lay.xp[phys])
    assert np.array_equal(o["vp"], lay.vp[phys])
    assert np.array_equal(o["rhoc"], _physical_rhoc(g, lay.rhoc))
    assert np.array_equal(o["idx_b_r"], _row_totals(g, o["rhoc"]))


@pytest.mark.parametrize("axis,sign", [(0, 1), (1, -1), (2, 1)])
def test_drift_by_one_coarse_cell_is_the_translated_image(oracle_lib, axis, sign):
    """dt_mid*vp/ratio_cs = +-1 exactly: every particle moves one coarse cell
    (codes wrap by 65536 to themselves), across tile faces through the halo
    copies.  The translated positions laid out from scratch give the same
    image, in the same order: a target cell's particles all come from one
    source cell, in their order there."""
    g = G2
    pos = _clumpy(g, 5000, 2)
    vel = np.zeros(pos.shape, np.float32)
    vel[:, axis] = 8.0 * sign                      # dt_mid 0.5: 0.5 * 8 / 4 = 1
    lay = build_layout(pos, vel, g)
    vz = 8.0 if axis == 2 else 0.0
    o = oracle_lib.drift(g, lay.rhoc, lay.idx_b_r, lay.xp, lay.vp, 0.5, vz)
    moved = pos.copy()
    moved[:, axis] = (moved[:, axis] + sign * g.ratio_cs) % g.ng
    want = build_layout(moved, vel, g)
    ph = want.physical_mask()
    assert o["n"] == pos.shape[0]
    assert np.array_equal(o["rhoc"], _physical_rhoc(g, want.rhoc))
    assert np.array_equal(o["xp"], want.xp[ph])
    assert np.array_equal(o["vp"], want.vp[ph])
    assert np.array_equal(o["idx_b_r"], _row_totals(g, o["rhoc"]))


def test_random_drift_moves_every_particle_once(oracle_lib):
    """Each particle ends in the cell ceiling(x + dt v / ratio_cs) (FP64 sum
    of the FP32 step) with code x + nint(dt v / (ratio_cs / 2^16)) mod 2^16,
    computed here per particle in global coordinates, without tiles, halos
    or spines; the drifted image holds exactly these (cell, code, velocity)
    triples."""
    g = G2
    rng = np.random.default_rng(3)
    pos = _clumpy(g, 6000, 4)
    vel = rng.uniform(-6, 6, pos.shape).astype(np.float32)
    lay = build_layout(pos, vel, g)
    dt = np.float32(0.7)
    vz_max = float(np.abs(vel[:, 2]).max())
    o = oracle_lib.drift(g, lay.rhoc, lay.idx_b_r, lay.xp, lay.vp, float(dt), vz_max)
    assert o["n"] == pos.shape[0]
    # expected per particle, from the physical slots of the input
    ph = lay.physical_mask()
    xp0 = lay.xp[ph].astype(np.int64)
    v0 = lay.vp[ph]
    cell0 = np.floor(lay.decoded_positions()[ph] / g.ratio_cs).astype(np.int64)
    u = (xp0 % 65536).astype(np.float64)                    # the code as uint16
    frac = (u + 0.5) / 65536.0
    step = (dt * v0) / np.float32(g.ratio_cs)               # FP32
    newcell = (np.ceil(cell0 + frac + step.astype(np.float64)).astype(np.int64) - 1) % g.nc
    sh = (dt * v0).astype(np.float64) / (g.ratio_cs / 65536.0)
    shift = np.floor(np.abs(sh) + 0.5) * np.sign(sh)         # nint: half away from zero
    code = ((xp0 + shift.astype(np.int64) + 32768) % 65536 - 32768).astype(np.int16)
    want = np.concatenate([newcell, code.astype(np.int64), v0.view(np.int32).astype(np.int64)], 1)
    # the oracle's image, decoded to (global cell, code, velocity bits)
    rho = o["rhoc"].reshape(g.ntile, g.nte, g.nte, g.nte)
    cells = []
    for t in range(g.ntile):
        tc = np.array([t % g.nnt, (t // g.nnt) % g.nnt, t // (g.nnt * g.nnt)])
        for k in range(g.nte):
            for j in range(g.nte):
                for i in range(g.nte):
                    c = rho[t, k, j, i]
                    if c:
                        gc = tc * g.nt + np.array([i, j, k]) - g.ncb
                        cells.append(np.repeat(gc[None], c, 0))
    cells = np.concatenate(cells) % g.nc
    got = np.concatenate([cells, o["xp"].astype(np.int64), o["vp"].view(np.int32).astype(np.int64)], 1)
    a = want[np.lexsort(want.T[::-1])]
    b = got[np.lexsort(got.T[::-1])]
    assert np.array_equal(a, b)
