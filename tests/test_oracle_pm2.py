"""The PM2 oracle (pm_thread with pm_layer 2, kick.f90:76-105) pinned by an
exact reduction to a PM3 job, the split force law it carries, translation
invariance across the tile's halo and momentum conservation (CPU only)."""
from __future__ import annotations

import numpy as np
import pytest

from cube2_amd.geometry import RATIO_SF, Geometry, ixyz3
from cube2_amd.layout import build_layout


def _fra(r, a):
    ep = 2 * r / a
    hi = (12 / ep ** 2 - 224 + 896 * ep - 840 * ep ** 2 + 224 * ep ** 3 + 70 * ep ** 4
          - 48 * ep ** 5 + 7 * ep ** 6) / 35 / a ** 2
    lo = (224 * ep - 224 * ep ** 3 + 70 * ep ** 4 + 48 * ep ** 5 - 21 * ep ** 6) / 35 / a ** 2
    return np.where(ep > 2, 1 / r ** 2, np.where(ep > 1, hi, lo))


def _green2(oracle_lib, g):
    # initialize.f90:135: Green_3D(ngt, apm2 = 3.5, apm1 = 3.5 ratio_cs, gridsize 1)
    return oracle_lib.green_3d(g.ngt, 3.5, 3.5 * g.ratio_cs, 1.0)


def _kick_to_force(dv, amid=0.02, dt=1e-3):
    return dv * 6 * np.pi / (amid * dt)


def test_pm2_equals_a_pm3_job_on_the_same_mesh(oracle_lib):
    """With ngb = ratio_cs (one halo cell) and one subtile per tile, a PM2
    job is a PM3 job at a level whose ratio_sf is 1: same mesh, same halo,
    same interpolation.  The two oracle paths must agree bit for bit."""
    g = Geometry(ng=64, nnt=2, nns=1, ratio_cs=4, ngb=4, app=0.05)
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, g.ng, (3000, 3))
    lay = build_layout(pos, None, g)
    G = _green2(oracle_lib, g)
    o2 = oracle_lib.pm2_force(g, lay.rhoc, lay.idx_b_r, lay.xp, G, 1.5, 0.02, 1e-3)
    rsf = RATIO_SF.copy()
    rsf[1] = 1                                   # level 2 on the standard grid
    assert g.ntt * g.ratio_cs * rsf[1] + 2 * g.ratio_cs * rsf[1] == g.ngt
    ires = np.full(g.ns3, 2, np.int32)
    isort = np.arange(1, g.ns3 + 1, dtype=np.int32)
    o3 = oracle_lib.pm3_force(g, isort, ires, ixyz3(g), rsf, lay.rhoc, lay.idx_b_r, lay.xp,
                              {2: G}, 1.5, 0.02, 1e-3)
    assert np.array_equal(o2["dv64"], o3["dv64"])
    assert o2["f2max"] == o3["f2max"]
    assert o2["ptotal"].sum() == o3["ptotal"].sum() == pos.shape[0]
    assert np.abs(o2["dv64"]).max() > 0


@pytest.mark.slow
def test_pm2_pair_force_follows_the_split_force_law(oracle_lib):
    """PM2 carries F_ra(r, 3.5) - F_ra(r, 3.5 ratio_cs) in standard cells:
    the coarse-grid softening of PM1 removed, the fine PM3 range cut off.
    Pairs at random orientations in the middle of one tile: the mean radial
    force matches the law to a few per cent, Newton's third law holds."""
    g = Geometry(ng=64, nnt=1, nns=1, ratio_cs=4, ngb=16, app=0.05)
    G = _green2(oracle_lib, g)
    rng = np.random.default_rng(1)
    for r in (3.0, 6.0, 10.0):
        ratios = []
        for q in range(6):
            c = np.full(3, 32.0) + rng.uniform(-2, 2, 3)
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            pos = np.array([c - 0.5 * r * u, c + 0.5 * r * u])
            lay = build_layout(pos, None, g)
            o = oracle_lib.pm2_force(g, lay.rhoc, lay.idx_b_r, lay.xp, G, 1.0, 0.02, 1e-3)
            phys = lay.physical_mask()
            p = lay.decoded_positions()
            i0 = np.flatnonzero(phys & (lay.src == 0))[0]
            i1 = np.flatnonzero(phys & (lay.src == 1))[0]
            d = p[i1] - p[i0]
            rr = np.linalg.norm(d)
            f0, f1 = _kick_to_force(o["dv64"][i0]), _kick_to_force(o["dv64"][i1])
            want = float(_fra(rr, 3.5) - _fra(rr, 3.5 * g.ratio_cs))
            ratios.append((f0 @ d / rr) / want)
            assert np.allclose(f1, -f0, atol=2e-3 * want)
        ratios = np.array(ratios)
        print(f"r={r}: force / law = {ratios.mean():.4f} +- {ratios.std():.4f}")
        assert abs(ratios.mean() - 1) < 0.03, (r, ratios)
        assert ratios.std() < 0.03, (r, ratios)


def test_pm2_source_in_the_halo_acts_like_one_inside(oracle_lib):
    """A source particle ncb - 1 coarse cells outside the tile (only in its
    halo copy) pulls a particle just inside the face exactly as the same pair
    moved 3 coarse cells into the tile: the deposit reaches all ncb halo
    cells (pm%nloop = ncb) and the mesh solve is translation invariant for
    whole-cell shifts.  One halo cell deep would see no source at all."""
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=16, app=0.05)
    assert g.ncb == 4
    G = _green2(oracle_lib, g)
    tgt = np.array([1.3, 15.2, 16.7])
    src = tgt + np.array([-8.0, 1.1, -0.6])      # x = -6.7: two coarse cells into the halo
    res = []
    for shift in (0.0, 12.0):                   # 12 = 3 coarse cells
        pos = np.array([tgt, src]) + np.array([shift, 0, 0])
        pos %= g.ng
        lay = build_layout(pos, None, g)
        o = oracle_lib.pm2_force(g, lay.rhoc, lay.idx_b_r, lay.xp, G, 1.0, 0.02, 1e-3)
        i = np.flatnonzero(lay.physical_mask() & (lay.src == 0))[0]
        res.append(o["dv64"][i])
    a, b = res
    assert np.linalg.norm(b) > 0
    assert np.allclose(a, b, rtol=0, atol=1e-9 * np.linalg.norm(b)), (a, b)
    assert a[0] < 0                             # attraction towards -x


def test_pm2_momentum_is_conserved_inside_a_tile(oracle_lib):
    """Particles deep inside one tile (no halo particles, force stencils
    inside the mesh): the same TSC kernel deposits and interpolates, the
    Green's function is symmetric and the difference operator odd, so the
    total Delta-v vanishes up to the FP32 rounding of the deposited mesh
    (~1e-7 of the summed magnitudes; a wrong sign or a shifted stencil
    leaves a total of the order of the magnitudes)."""
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=16, app=0.05)
    G = _green2(oracle_lib, g)
    rng = np.random.default_rng(7)
    pos = rng.uniform(4, 28, (400, 3))           # tile 0 spans [0, 32)
    lay = build_layout(pos, None, g)
    o = oracle_lib.pm2_force(g, lay.rhoc, lay.idx_b_r, lay.xp, G, 1.0, 0.02, 1e-3)
    phys = lay.physical_mask()
    dv = o["dv64"][phys]
    tot = dv.sum(axis=0)
    scale = np.abs(dv).sum(axis=0)
    assert np.all(np.abs(tot) < 2e-7 * scale), (tot, scale)
    assert o["ptotal"][0] == 400 and o["ptotal"][1:].sum() == 0
