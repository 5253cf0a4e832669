"""Parity at the 256^3 headline configuration (config 1: 16,777,216
particles, levels 2..6 selected from the density) and at a scaled batching
workload, through the C ABI, against the oracle.

* The PM3 meshes of the two finest levels present (n = 320 at level 5 and
  n = 480 at level 6) through cube2_pm3_debug_job against the oracle's job
  (same Green's table): density, potential and force planes within 1e-5
  relative RMS (pm3_fft_cuda.cu:483-504, kick.f90:419-494).
* The Green's tables of those levels against an independent FP64
  evaluation of Green.f90:22-71 at sampled k.
* The default REF pair kernel on four parent tiles (a dense one among
  them) against the oracle restatement of pp_force_kernel.c:281-364: P_cand,
  P_acc and ptotal per tile exact, Delta-v within 1e-5 relative RMS (the
  pair forces are bit-identical; only the order of the per-particle sums
  differs).
* A 512-tile batching workload (per-tile counts lognormal, median 4000,
  sigma 1.5, capped at 200,000) through the recorded kick
  (cube2_kick_graph_dev): replays equal the eager kick bit for bit, and the
  integers equal the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import green_at

pytestmark = pytest.mark.gpu

TILES = [13, 21, 42, 63]
RTOL = 1e-5


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


def _green_dev(lib, n, rsf):
    import torch
    from cube2_amd.geometry import APM3F
    t = torch.empty((n, n, n // 2 + 1), dtype=torch.float32, device="cuda")
    lib.check(lib.lib().cube2_green_3d_dev(t.data_ptr(), n, float(APM3F), float(APM3F * rsf),
                                           float(np.float32(1) / np.float32(rsf))))
    return t


@pytest.mark.parametrize("level", [5, 6])
def test_c1_fine_level_job_matches_oracle(lib, oracle_lib, c1_case, level):
    import torch
    from cube2_amd.geometry import APM3F, RATIO_SF, ixyz3
    pos, lay, ires, _ = c1_case
    g = lay.geom
    lib.set_params(g)
    it = int(np.flatnonzero(ires == level)[0])
    n, nf, rsf = g.nft(level), g.nfp(level) + 2, int(RATIO_SF[level - 1])
    print(f"level {level}: subtile {it}, mesh {n}^3", flush=True)
    G = _green_dev(lib, n, rsf)
    Gh = G.cpu().numpy()
    rng = np.random.default_rng(level)
    idx = np.stack([rng.integers(0, n // 2 + 1, 300), rng.integers(0, n, 300),
                    rng.integers(0, n, 300)], 1)
    want_g = green_at(n, float(APM3F), float(APM3F * rsf), float(np.float32(1) / np.float32(rsf)), idx)
    got_g = Gh[idx[:, 2], idx[:, 1], idx[:, 0]].astype(np.float64)
    assert np.allclose(got_g, want_g, rtol=3e-7, atol=1e-7 * np.abs(want_g).max())
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in
         dict(rhoc=lay.rhoc, idx=lay.idx_b_r, xp=lay.xp).items()}
    rho = np.zeros((n, n, n), np.float32)
    phi = np.zeros((n, n, n), np.float32)
    force = np.zeros((3, nf, nf, nf), np.float32)
    f2 = np.zeros(1, np.float32)
    ix = ixyz3(g)  # held: the library reads it through its address
    lib.check(lib.lib().cube2_pm3_debug_job(it, ires.ctypes.data, ix.ctypes.data,
                                            RATIO_SF.ctypes.data, d["rhoc"].data_ptr(),
                                            d["idx"].data_ptr(), d["xp"].data_ptr(), lay.n_used,
                                            G.data_ptr(), 1.0, rho.ctypes.data, phi.ctypes.data,
                                            force.ctypes.data, f2.ctypes.data))
    del d, G
    o = oracle_lib.pm3_job(g, it, ires, ix, RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp, Gh, 1.0)
    e_rho, e_phi, e_f = rel_rms(rho, o["rho"]), rel_rms(phi, o["phi"]), rel_rms(force, o["force"])
    print(f"level {level}: rho {e_rho:.2e} phi {e_phi:.2e} force {e_f:.2e} "
          f"f2max {f2[0]:.6e} vs {o['f2max']:.6e}", flush=True)
    # The oracle sums each cell's FP32 TSC weights sequentially in FP32, as
    # the reference's real(4) density does (kick.f90:439-445); the device sums
    # them in 32.32 fixed point.  Cells in the clumps of the finest levels
    # take thousands of weights, so the oracle's own rounding reaches ~1e-6.
    assert e_rho < 5e-6
    assert e_phi < RTOL
    assert e_f < RTOL
    assert abs(f2[0] - o["f2max"]) < 3e-5 * o["f2max"]


def test_c1_ref_pair_kernel_matches_oracle_on_sampled_tiles(lib, oracle_lib, c1_case):
    from cube2_amd.geometry import APM3, RATIO_SF, ixyz3
    from cube2_amd.kick import ShortRangeKick
    pos, lay, ires, _ = c1_case
    g = lay.geom
    amid, dt = 0.02, 1e-3
    lib.set_pp_mode(lib.PP_REF)
    lib.pp_count_tiles(True)
    try:
        k = ShortRangeKick(lay)
        k.set_levels(ires)
        res = k.kick(1.0, amid, dt, pm3=False, pp=True)
        tp = lib.pp_tile_pairs(g.ntile)
    finally:
        lib.pp_count_tiles(False)
    assert int(tp[:, 0].sum()) == res.p_cand and int(tp[:, 1].sum()) == res.p_acc
    vp = k.vp_host()
    del k
    ix = ixyz3(g)
    ends = lay.idx_b_r.reshape(g.ntile, -1)[:, -1]
    phys = lay.physical_mask()
    tile_of_sub = ((ix[:, 5] - 1) * g.nnt + ix[:, 4] - 1) * g.nnt + ix[:, 3] - 1
    for t in TILES:
        o = oracle_lib.pp_force(g, ires, ix, APM3, RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp,
                                lay.vp, 1.0, amid, dt, tiles=(t, 1))
        sl = np.zeros(phys.size, bool)
        sl[(ends[t - 1] if t else 0):ends[t]] = True
        sl &= phys
        err = rel_rms(vp[sl], o["vp"][sl])
        print(f"tile {t}: {int(sl.sum())} particles, P_cand {int(tp[t, 0])} P_acc {int(tp[t, 1])}, "
              f"Delta-v relative RMS {err:.2e}", flush=True)
        assert int(tp[t, 0]) == o["p_cand"] and int(tp[t, 1]) == o["p_acc"]
        subs = tile_of_sub == t
        assert np.array_equal(res.ptotal_pp[subs], o["ptotal"][subs])
        assert err < RTOL


def test_c2_scaled_recorded_kick(lib, oracle_lib):
    """512 tiles of 32^3 fine cells with lognormal counts: the recorded kick
    equals the eager kick bit for bit on two replays (new dt on the second),
    and its pair counts and ptotal equal the oracle's."""
    import torch
    from cube2_amd.geometry import APM3, C2, RATIO_SF, ixyz3
    from cube2_amd.kick import ShortRangeKick
    from cube2_amd.layout import build_layout
    from cube2_amd.synth import config2
    g = C2.with_(ng=128, nnt=8)
    pos, vel, cnt = config2(seed=2, geom=g)
    lay = build_layout(pos, vel, g)
    print(f"C2 scaled: {g.ntile} tiles, {pos.shape[0]} particles, {lay.n_used} slots", flush=True)
    lib.set_pp_mode(lib.PP_REF)
    k = ShortRangeKick(lay)
    vp0 = k.d_vp.clone()
    runs = []
    for q, dt in enumerate((1e-3, 1e-3, 2e-3)):
        for mode in ("eager", "graph"):
            k.d_vp.copy_(vp0)
            torch.cuda.synchronize()
            r = k.kick(1.0, 0.02, dt) if mode == "eager" else k.kick_graph(1.0, 0.02, dt)
            runs.append((mode, dt, r, k.vp_host()))
    for (m1, dt1, r1, v1), (m2, dt2, r2, v2) in zip(runs[0::2], runs[1::2]):
        assert dt1 == dt2
        assert np.array_equal(v1.view(np.uint32), v2.view(np.uint32)), dt1
        assert (r1.p_cand, r1.p_acc) == (r2.p_cand, r2.p_acc)
        assert np.array_equal(r1.ptotal_pp, r2.ptotal_pp)
        assert np.array_equal(r1.ptotal_pm3, r2.ptotal_pm3)
        assert r1.f2max_pp == r2.f2max_pp and r1.f2max_pm3 == r2.f2max_pm3
    r = runs[1][2]
    o = oracle_lib.pp_force(g, k.ires, ixyz3(g), APM3, RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp,
                            lay.vp, 1.0, 0.02, 1e-3)
    print(f"C2 scaled: P_cand {r.p_cand} P_acc {r.p_acc} (oracle {o['p_cand']} {o['p_acc']})")
    assert (r.p_cand, r.p_acc) == (o["p_cand"], o["p_acc"])
    assert np.array_equal(r.ptotal_pp, o["ptotal"])
    assert np.array_equal(r.ptotal_pm3, o["ptotal"])
