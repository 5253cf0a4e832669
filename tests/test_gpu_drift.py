"""Drift on the GPU (cube2_drift_dev) against the oracle: integer work, so
bit-exact (rhoc, idx_b_r, int16 codes, velocities, count and order)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available()
    from cube2_amd import _lib
    _lib.lib()
    return _lib


def _clumpy(g, n, seed):
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0, g.ng, (n // 2, 3))]
    for c in rng.uniform(0, g.ng, (5, 3)):
        pos.append(c + rng.normal(0, 1.0, (n // 10, 3)))
    return np.concatenate(pos) % g.ng


def _dev(lay):
    import torch
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t(lay.rhoc), t(lay.idx_b_r), t(lay.xp), t(lay.vp)


CASES = {
    "c0_like": dict(ng=64, nnt=1, nns=1, ratio_cs=4, ngb=16, n=20000, vmax=6.0, dt=0.7),
    "8_tiles": dict(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, n=30000, vmax=6.0, dt=0.7),
    "64_tiles": dict(ng=128, nnt=4, nns=1, ratio_cs=4, ngb=8, n=60000, vmax=9.0, dt=0.4),
    "zero_step": dict(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, n=8000, vmax=5.0, dt=0.0),
}


@pytest.mark.parametrize("case", list(CASES))
def test_drift_bit_exact_vs_oracle(lib, oracle_lib, case):
    from cube2_amd.drift import drift_device
    from cube2_amd.geometry import Geometry
    from cube2_amd.layout import build_layout
    c = CASES[case]
    g = Geometry(ng=c["ng"], nnt=c["nnt"], nns=c["nns"], ratio_cs=c["ratio_cs"], ngb=c["ngb"],
                 app=0.05)
    print(f"drift {case}: {g.ntile} tiles, {c['n']} particles", flush=True)
    rng = np.random.default_rng(7)
    pos = _clumpy(g, c["n"], 11)
    vel = rng.uniform(-c["vmax"], c["vmax"], pos.shape).astype(np.float32)
    lay = build_layout(pos, vel, g)
    dt = float(np.float32(c["dt"]))
    vz_max = float(np.abs(vel[:, 2]).max())
    o = oracle_lib.drift(g, lay.rhoc, lay.idx_b_r, lay.xp, lay.vp, dt, vz_max)
    r, i, x, v, n = drift_device(*_dev(lay), lay.n_used, dt, vz_max, g)
    assert n == o["n"] == pos.shape[0]
    assert np.array_equal(r.cpu().numpy(), o["rhoc"])
    assert np.array_equal(i.cpu().numpy(), o["idx_b_r"])
    assert np.array_equal(x.cpu().numpy(), o["xp"])
    assert np.array_equal(v.cpu().numpy().view(np.int32), o["vp"].view(np.int32))


def test_drift_then_layout_is_the_next_kick_image(lib, oracle_lib):
    """The drifted physical image plus rebuilt halo copies (decoded slot
    positions through the device layout builder) equals the host layout of
    the oracle's drifted particles, in order."""
    from cube2_amd.drift import drift_device, next_layout
    from cube2_amd.geometry import Geometry
    from cube2_amd.layout import build_layout
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)
    rng = np.random.default_rng(8)
    pos = _clumpy(g, 20000, 12)
    vel = rng.uniform(-6, 6, pos.shape).astype(np.float32)
    lay = build_layout(pos, vel, g)
    vz_max = float(np.abs(vel[:, 2]).max())
    o = oracle_lib.drift(g, lay.rhoc, lay.idx_b_r, lay.xp, lay.vp, 0.7, vz_max)
    img = drift_device(*_dev(lay), lay.n_used, 0.7, vz_max, g)
    rhoc, idx, xp, vp, n_used = next_layout(*img, g)
    # host: decode the oracle's image (all its slots are physical) and lay it out
    rho = o["rhoc"].reshape(g.ntile, g.nte, g.nte, g.nte)
    cells = []
    for t in range(g.ntile):
        tc = np.array([t % g.nnt, (t // g.nnt) % g.nnt, t // (g.nnt * g.nnt)])
        for k in range(g.nte):
            for j in range(g.nte):
                for i in range(g.nte):
                    if rho[t, k, j, i]:
                        gc = tc * g.nt + np.array([i, j, k]) - g.ncb
                        cells.append(np.repeat(gc[None], rho[t, k, j, i], 0))
    cells = np.concatenate(cells) % g.nc
    frac = ((o["xp"].astype(np.int64) % 65536) + 0.5) / 65536.0
    want = build_layout((cells + frac) * g.ratio_cs, o["vp"], g)
    assert n_used == want.n_used
    assert np.array_equal(rhoc.cpu().numpy(), want.rhoc)
    assert np.array_equal(idx.cpu().numpy(), want.idx_b_r)
    assert np.array_equal(xp.cpu().numpy(), want.xp)
    assert np.array_equal(vp.cpu().numpy(), want.vp)


def test_kick_halo_refresh_drift_keeps_every_particle(lib):
    """A kick updates the physical slots only; after buffer_v
    (cube2_halo_velocities_dev) every copy of a particle carries its
    velocity, and the drift keeps every particle exactly once."""
    import torch
    from cube2_amd.drift import drift_device, refresh_halo_velocities
    from cube2_amd.geometry import Geometry
    from cube2_amd.kick import ShortRangeKick
    from cube2_amd.layout import build_layout
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)
    rng = np.random.default_rng(10)
    pos = _clumpy(g, 20000, 13)
    vel = rng.uniform(-2, 2, pos.shape).astype(np.float32)
    lay = build_layout(pos, vel, g)
    k = ShortRangeKick(lay)
    k.kick(1.0, 0.02, 0.05)
    vp = k.vp_host()
    phys = lay.physical_mask()
    # physical velocity of each particle, by source index
    pv = np.zeros((pos.shape[0], 3), np.float32)
    pv[lay.src[phys]] = vp[phys]
    assert not np.array_equal(vp, pv[lay.src])          # halo copies are stale
    refresh_halo_velocities(k.d_rhoc, k.d_idx, k.d_vp, k.n_used, g)
    torch.cuda.synchronize()
    assert np.array_equal(k.vp_host(), pv[lay.src])
    vz = float(np.abs(pv[:, 2]).max())
    dt = float(np.float32(g.ratio_cs / np.abs(pv).max()))
    out = drift_device(k.d_rhoc, k.d_idx, k.d_xp, k.d_vp, k.n_used, dt, vz, g)
    assert out[4] == pos.shape[0]


def test_drift_rejects_a_step_beyond_vz_max(lib):
    from cube2_amd.drift import drift_device
    from cube2_amd.geometry import Geometry
    from cube2_amd.layout import build_layout
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)
    rng = np.random.default_rng(9)
    pos = rng.uniform(0, g.ng, (4000, 3))
    vel = rng.uniform(-20, 20, pos.shape).astype(np.float32)
    lay = build_layout(pos, vel, g)
    with pytest.raises(lib.Cube2Error) as e:
        drift_device(*_dev(lay), lay.n_used, 0.5, 0.1, g)   # |vz| up to 20, declared 0.1
    assert e.value.code == lib.ERR_ARG
