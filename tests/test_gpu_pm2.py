"""PM2 (kick.f90:76-105) on the GPU through the C ABI against the oracle.
Tolerances (FP32 GPU path vs FP64 oracle): relative RMS <= 1e-5 for the
Delta-v, 3e-5 relative for f2max; particle counts exact."""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL_PM2 = 1e-5


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available()
    from cube2_amd import _lib
    _lib.lib()
    return _lib


def _clustered(g, n, seed):
    """Uniform background plus a few dense clumps, some across tile faces."""
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0, g.ng, (n // 2, 3))]
    for c in rng.uniform(0, g.ng, (6, 3)):
        pos.append(c + rng.normal(0, 1.5, (n // 12, 3)))
    return np.concatenate(pos) % g.ng


@pytest.mark.parametrize("ngb", [12, 16])
def test_pm2_matches_oracle(lib, oracle_lib, ngb):
    """ngb = 12: mesh 32 + 24 = 56 (radix-7 FFT); ngb = 16: mesh 64."""
    from cube2_amd.geometry import Geometry
    from cube2_amd.kick import KickResult, ShortRangeKick
    from cube2_amd.layout import build_layout
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=ngb, app=0.05)
    print(f"PM2 ng={g.ng} nnt={g.nnt} ngb={ngb} mesh {g.ngt}", flush=True)
    pos = _clustered(g, 12000, ngb)
    lay = build_layout(pos, None, g)
    k = ShortRangeKick(lay)
    mass, amid, dt = 1.25, 0.02, 1e-3
    r = k.kick(mass, amid, dt, pm3=False, pp=False, pm2=True)
    o = oracle_lib.pm2_force(g, lay.rhoc, lay.idx_b_r, lay.xp, k.green2().cpu().numpy(), mass,
                             amid, dt)
    phys = lay.physical_mask()
    dv = k.vp_host()[phys].astype(np.float64) - lay.vp[phys]
    err = rel_rms(dv, o["dv64"][phys])
    print(f"PM2 Delta-v relative RMS {err:.2e}; f2max {r.f2max_pm2:.6e} vs {o['f2max']:.6e}")
    assert err < RTOL_PM2
    assert abs(r.f2max_pm2 - o["f2max"]) < 3e-5 * o["f2max"]
    # the library's own entry point: per-tile counts of physical particles
    res = KickResult()
    k.pm2(mass, amid, dt, res)
    assert np.array_equal(res.ptotal_pm2, o["ptotal"])
    assert res.ptotal_pm2.sum() == pos.shape[0]
    k.apply_pending()                       # nothing left pending for the next test


def test_pm2_pm3_pp_driver_matches_separate_calls(lib):
    """The kick driver with PM2 | PM3 | PP equals pm2(), pm3(), pp() called
    one after the other (the mesh Delta-v adds up in FP64 and PP applies
    it), and its dt_pm2 is kick.f90:97's formula in FP64."""
    import torch
    from cube2_amd.geometry import Geometry
    from cube2_amd.kick import KickResult, ShortRangeKick
    from cube2_amd.layout import build_layout
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=12, app=0.05)
    lay = build_layout(_clustered(g, 8000, 3), None, g)
    mass, amid, dt = 1.0, 0.03, 2e-3
    k1 = ShortRangeKick(lay)
    r1 = k1.kick(mass, amid, dt, pm2=True)
    k2 = ShortRangeKick(lay, greens=k1.greens)
    r2 = KickResult()
    k2.pm2(mass, amid, dt, r2)
    k2.pm3(mass, amid, dt, r2)
    k2.pp(mass, amid, dt, r2)
    torch.cuda.synchronize()
    phys = lay.physical_mask()
    err = rel_rms(k1.vp_host()[phys] - lay.vp[phys], k2.vp_host()[phys] - lay.vp[phys])
    print(f"driver vs separate calls: Delta-v relative RMS {err:.2e}")
    assert err <= 1e-6
    assert r1.f2max_pm2 == pytest.approx(r2.f2max_pm2, rel=1e-5)
    a32 = float(np.float32(amid))
    assert r1.dt_native["dt_pm2"] == pytest.approx(
        0.5 * math.sqrt(1.0 / (math.sqrt(r1.f2max_pm2) * a32 / (6 * math.pi))), rel=1e-15)
    # without PM2 the velocities differ: PM2's share is not negligible
    k3 = ShortRangeKick(lay, greens=k1.greens)
    k3.kick(mass, amid, dt)
    assert rel_rms(k3.vp_host()[phys] - lay.vp[phys], k1.vp_host()[phys] - lay.vp[phys]) > 1e-3
    assert math.isinf(k3.kick(mass, amid, dt).dt_native["dt_pm2"])
