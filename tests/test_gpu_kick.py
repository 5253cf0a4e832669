"""The library's kick driver (cube2_kick_dev, kick.f90:107-238) on the GPU:
device level selection equals the oracle's (integer, exact); one driver
call equals separate PM3 and PP calls; the dt limits equal the FP64
formulas of kick.f90:180,227,236."""
from __future__ import annotations

import ctypes as C
import math

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


@pytest.fixture(scope="module")
def mixed_layout():
    """64^3 box in 2^3 tiles of 2^3 subtiles, uniform background plus one
    clump; levels 2..4 drawn per subtile."""
    from cube2_amd.geometry import Geometry
    from cube2_amd.layout import build_layout
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)
    rng = np.random.default_rng(23)
    pos = np.concatenate([rng.uniform(0, 64, (20000, 3)),
                          np.array([31.5, 30.0, 33.0]) + rng.normal(0, 0.7, (3000, 3))]) % 64
    lay = build_layout(pos, None, g)
    ires = rng.integers(2, 5, g.ns3).astype(np.int32)
    return lay, ires


@pytest.mark.parametrize("seed", [9, 10])
def test_device_level_selection_matches_oracle(lib, oracle_lib, seed):
    import torch
    from cube2_amd.geometry import APM3, RATIO_SF, REFERENCE_DEFAULT, ixyz3
    g = REFERENCE_DEFAULT.with_(nnt=2, ng=128, nns=2)
    rng = np.random.default_rng(seed)
    rhoc = rng.integers(0, 5, (g.ntile, g.nte, g.nte, g.nte)).astype(np.int32)
    hot = rng.integers(0, rhoc.size, 60)            # halo and physical cells
    rhoc.reshape(-1)[hot] = rng.integers(1000, 100000, 60)
    rhoc.reshape(-1)[hot[:4]] = [1535, 1536, 7295, 7296]    # level thresholds
    lib.set_params(g)
    d_rhoc = torch.from_numpy(rhoc).cuda()
    ix = ixyz3(g)
    ires = np.zeros(g.ns3, np.int32)
    isort = np.zeros(g.ns3, np.int32)
    out = lib.KickResultC()
    lib.check(lib.lib().cube2_kick_dev(0, ix.ctypes.data, APM3.ctypes.data, RATIO_SF.ctypes.data,
                                       d_rhoc.data_ptr(), None, None, None, 0, None, 1.0, 1.0, 0.0,
                                       ires.ctypes.data, isort.ctypes.data, None, None, None,
                                       C.byref(out)))
    o_ires, o_isort, o_napm, _ = oracle_lib.select_levels(g, rhoc, ix)
    assert np.array_equal(ires, o_ires)
    assert np.array_equal(isort, o_isort)
    assert np.array_equal(np.array(out.napm[:]), o_napm)
    assert math.isinf(out.dt_pm3) and math.isinf(out.dt_pp) and math.isinf(out.dt_vmax)


def _rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def test_kick_driver_equals_separate_calls(lib, mixed_layout):
    """One driver call and the two device entry points: PP alone (exact
    mode) bit for bit; PM3 + PP within 1e-6, since the deposit's LDS float
    sums run in no fixed order (integers and pair counts exact)."""
    from cube2_amd.kick import KickResult, ShortRangeKick
    lay, ires = mixed_layout
    lib.set_pp_mode(lib.PP_EXACT)
    mass, amid, dt = 1.3, 0.05, 2e-3
    try:
        # PP only: deterministic, bit-identical
        k1 = ShortRangeKick(lay)
        k1.set_levels(ires)
        r1 = k1.kick(mass, amid, dt, pm3=False, pp=True)
        k2 = ShortRangeKick(lay)
        k2.set_levels(ires)
        r2 = KickResult()
        k2.pp(mass, amid, dt, r2)
        assert np.array_equal(k1.vp_host(), k2.vp_host())
        assert r1.f2max_pp == r2.f2max_pp and np.array_equal(r1.vmax, r2.vmax)
        assert np.array_equal(r1.ptotal_pp, r2.ptotal_pp)
        assert (r1.p_cand, r1.p_acc) == (r2.p_cand, r2.p_acc)
        assert math.isinf(r1.dt_native["dt_pm3"])
        # PM3 + PP
        k1 = ShortRangeKick(lay)
        k1.set_levels(ires)
        r1 = k1.kick(mass, amid, dt)
        k2 = ShortRangeKick(lay, greens=k1.greens)
        k2.set_levels(ires)
        r2 = KickResult()
        k2.pm3(mass, amid, dt, r2)
        k2.pp(mass, amid, dt, r2)
    finally:
        lib.set_pp_mode(lib.PP_REF)
    phys = lay.physical_mask()
    err = _rel_rms(k1.vp_host()[phys] - lay.vp[phys], k2.vp_host()[phys] - lay.vp[phys])
    print(f"driver vs separate calls: Delta-v relative RMS {err:.2e}")
    assert err <= 1e-6
    assert r1.f2max_pm3 == pytest.approx(r2.f2max_pm3, rel=1e-5)
    assert r1.f2max_pp == r2.f2max_pp
    assert np.array_equal(r1.ptotal_pm3, r2.ptotal_pm3)
    assert np.array_equal(r1.ptotal_pp, r2.ptotal_pp)
    assert (r1.p_cand, r1.p_acc) == (r2.p_cand, r2.p_acc)
    a32 = float(np.float32(amid))    # the library receives a_mid as a float
    want = r1.dt_limits(a32)
    for key in ("dt_pm3", "dt_pp", "dt_vmax"):
        assert r1.dt_native[key] == pytest.approx(want[key], rel=1e-15), key
    assert r1.dt_native["dt_pm3"] == pytest.approx(
        0.5 * math.sqrt(1.0 / (math.sqrt(r1.f2max_pm3) * a32 / (6 * math.pi))), rel=1e-15)
    assert r1.dt_native["dt_vmax"] == pytest.approx(0.9 * 14 / float(np.max(r1.vmax)), rel=1e-15)


def test_kick_driver_selects_levels_on_device(lib, mixed_layout):
    """select=True picks ires/isort from the device rhoc, as the host
    restatement does from the host rhoc."""
    from cube2_amd.kick import ShortRangeKick, select_levels
    lay, _ = mixed_layout
    k = ShortRangeKick(lay)
    h_ires, h_isort, h_napm = select_levels(lay.geom, lay.rhoc)
    res = k.kick(1.0, 0.02, 1e-3, select=True)
    assert np.array_equal(k.ires, h_ires)
    assert np.array_equal(k.isort, h_isort)
    assert np.array_equal(k.napm, h_napm)
    assert res.p_acc > 0 and res.f2max_pm3 > 0


def test_recorded_kick_replays_match_eager(lib, mixed_layout):
    """cube2_kick_graph_dev: the first call records, later calls replay.
    Every replay (with host work and synchronisation between replays, and
    with a changed dt) equals the eager kick: PP alone bit for bit, PM3 + PP
    within 1e-6; an eager library call in between forces a new recording."""
    import torch
    from cube2_amd.kick import ShortRangeKick
    lay, ires = mixed_layout
    phys = lay.physical_mask()
    mass, amid = 1.3, 0.05
    k = ShortRangeKick(lay)
    k.set_levels(ires)
    vp0 = k.d_vp.clone()

    def reset():
        k.d_vp.copy_(vp0)
        torch.cuda.synchronize()

    def eager(dt, pm3):
        reset()
        r = k.kick(mass, amid, dt, pm3=pm3, pp=True)
        return k.vp_host(), r

    for pm3 in (False, True):
        want, rw = eager(2e-3, pm3)
        want2, rw2 = eager(3e-3, pm3)
        seen = []
        for rep, dt in enumerate([2e-3, 2e-3, 3e-3, 2e-3]):
            reset()
            x = torch.randn(1 << 20, device="cuda")      # eager work between replays
            float((x * x).sum())
            r = k.kick_graph(mass, amid, dt, pm3=pm3)
            seen.append(r.recorded)
            got = k.vp_host()
            ref, rref = (want, rw) if dt == 2e-3 else (want2, rw2)
            if pm3:
                err = float(np.sqrt((((got - ref)[phys]) ** 2).sum() / ((ref - lay.vp)[phys] ** 2).sum()))
                print(f"pm3+pp replay {rep} (dt {dt}): Delta-v relative RMS vs eager {err:.2e}")
                assert err <= 1e-6
                assert r.f2max_pm3 == pytest.approx(rref.f2max_pm3, rel=1e-5)
                assert np.array_equal(r.ptotal_pm3, rref.ptotal_pm3)
            else:
                assert np.array_equal(got, ref), f"PP replay {rep} differs"
                assert r.f2max_pp == rref.f2max_pp and np.array_equal(r.vmax, rref.vmax)
            assert (r.p_cand, r.p_acc) == (rref.p_cand, rref.p_acc)
            assert np.array_equal(r.ptotal_pp, rref.ptotal_pp)
            assert r.dt_native["dt_pp"] == pytest.approx(rref.dt_native["dt_pp"], rel=1e-6)
        assert seen == [True, False, False, False]
        eager(2e-3, pm3)                                   # an eager library call in between
        reset()
        r = k.kick_graph(mass, amid, 2e-3, pm3=pm3)
        assert r.recorded
        if not pm3:
            assert np.array_equal(k.vp_host(), want)


def test_tunables_keep_results_and_rerecord(lib, mixed_layout):
    """Launch-shape tunables change no integer result and move Delta-v by at
    most the summation-order level (1e-6); a change makes the recorded kick
    record again."""
    import torch
    from cube2_amd.kick import ShortRangeKick
    lay, ires = mixed_layout
    phys = lay.physical_mask()
    k = ShortRangeKick(lay)
    k.set_levels(ires)
    vp0 = k.d_vp.clone()
    r0 = k.kick_graph(1.0, 0.02, 1e-3)
    base = k.vp_host()
    settings = [(lib.TUNE_PP_BLOCKS_PER_CU, 3), (lib.TUNE_PP_DIRECT_LG, 6),
                (lib.TUNE_FFT_MIN_LGC, 4), (lib.TUNE_INTERP_SEG, 2), (lib.TUNE_PP_DENSE_LG, 6),
                (lib.TUNE_PP_ORDER, 18), (lib.TUNE_PP_JBATCH, 2), (lib.TUNE_PP_NEWTON, 2)]
    try:
        for key, v in settings:
            lib.set_tuning(key, v)
            k.d_vp.copy_(vp0)
            torch.cuda.synchronize()
            r = k.kick_graph(1.0, 0.02, 1e-3)
            assert r.recorded
            got = k.vp_host()
            err = float(np.sqrt((((got - base)[phys]) ** 2).sum() / ((base - lay.vp)[phys] ** 2).sum()))
            assert err <= 1e-6, (key, v, err)
            assert (r.p_cand, r.p_acc) == (r0.p_cand, r0.p_acc)
            assert np.array_equal(r.ptotal_pm3, r0.ptotal_pm3)
    finally:
        for key, _ in settings:
            lib.set_tuning(key, 0)


def test_overlapped_kick_with_a_level_without_pp_range(lib, mixed_layout):
    """A level whose PP range is 0 leaves its particles out of PP (the unsplit
    pair kernel skips them, velocity update included); the overlapped kick's
    velocity update then runs in sorted order instead of over the physical
    rows and still equals the serial kick bit for bit."""
    import torch
    from cube2_amd.geometry import APM3
    from cube2_amd.kick import ShortRangeKick
    lay, ires = mixed_layout
    apm3 = np.array(APM3, np.float32).copy()
    apm3[3 - 1] = 0.0  # level 3
    assert (ires == 3).any()
    k = ShortRangeKick(lay)
    k.set_levels(ires)
    vp0 = k.d_vp.clone()
    runs = {}
    try:
        for ov in (2, 1):
            lib.set_tuning(lib.TUNE_KICK_OVERLAP, ov)
            k.d_vp.copy_(vp0)
            torch.cuda.synchronize()
            r = k.kick(1.3, 0.05, 2e-3, apm3=apm3)
            runs[ov] = (r, k.vp_host())
    finally:
        lib.set_tuning(lib.TUNE_KICK_OVERLAP, 0)
    (rs, vs), (ro, vo) = runs[2], runs[1]
    print(f"vp differs in {int((vo != vs).sum())} entries", flush=True)
    assert np.array_equal(vo.view(np.uint32), vs.view(np.uint32))
    assert rs.f2max_pp == ro.f2max_pp and np.array_equal(rs.vmax, ro.vmax)
    assert (rs.p_cand, rs.p_acc) == (ro.p_cand, ro.p_acc)
    # the level-3 particles kept their velocities through PP
    assert rs.p_acc > 0


@pytest.mark.parametrize("mode", ["ref", "fast"])
def test_chunk_order_changes_no_result(lib, mixed_layout, mode):
    """The pair kernel's chunks in cell order, all sorted heaviest first, or
    the heavy ones first: every particle is summed within its own chunk, so
    velocities, maxima and pair counts are the same bit for bit (a chunk lost
    or taken twice would change the velocities or the counts)."""
    import torch
    from cube2_amd.kick import ShortRangeKick
    lay, ires = mixed_layout
    lib.set_pp_mode({"ref": lib.PP_REF, "fast": lib.PP_FAST}[mode])
    k = ShortRangeKick(lay)
    k.set_levels(ires)
    vp0 = k.d_vp.clone()
    runs = {}
    try:
        for order in (1, 2, 9):
            lib.set_tuning(lib.TUNE_PP_ORDER, order)
            k.d_vp.copy_(vp0)
            torch.cuda.synchronize()
            r = k.kick(1.3, 0.05, 2e-3)
            runs[order] = (r, k.vp_host())
    finally:
        lib.set_tuning(lib.TUNE_PP_ORDER, 0)
        lib.set_pp_mode(lib.PP_REF)
    rc, vc = runs[1]
    for order in (2, 9):
        rs, vs = runs[order]
        print(f"{mode} order {order}: vp differs in {int((vs != vc).sum())} entries", flush=True)
        assert np.array_equal(vs.view(np.uint32), vc.view(np.uint32)), order
        assert rs.f2max_pp == rc.f2max_pp and np.array_equal(rs.vmax, rc.vmax)
        assert (rs.p_cand, rs.p_acc) == (rc.p_cand, rc.p_acc)
        assert np.array_equal(rs.ptotal_pp, rc.ptotal_pp)


@pytest.mark.parametrize("mode", ["ref", "fast", "precise"])
def test_overlapped_kick_equals_serial_kick(lib, mixed_layout, mode):
    """PP on the second stream beside PM3 (its pair sums apart from the
    velocity update, which follows both) gives the serial kick's results bit
    for bit: the update is vp = RN_f(vp + dv64 + PP term) in the same order,
    and neither pass reads what the other writes.  Eager and recorded."""
    import torch
    from cube2_amd.kick import ShortRangeKick
    lay, ires = mixed_layout
    lib.set_pp_mode({"ref": lib.PP_REF, "fast": lib.PP_FAST, "precise": lib.PP_PRECISE}[mode])
    k = ShortRangeKick(lay)
    k.set_levels(ires)
    vp0 = k.d_vp.clone()
    runs = {}
    try:
        for ov, graph in ((2, False), (1, False), (1, True)):
            lib.set_tuning(lib.TUNE_KICK_OVERLAP, ov)
            k.d_vp.copy_(vp0)
            torch.cuda.synchronize()
            r = k.kick_graph(1.3, 0.05, 2e-3) if graph else k.kick(1.3, 0.05, 2e-3)
            runs[(ov, graph)] = (r, k.vp_host())
    finally:
        lib.set_tuning(lib.TUNE_KICK_OVERLAP, 0)
        lib.set_pp_mode(lib.PP_REF)
    rs, vs = runs[(2, False)]
    for key in ((1, False), (1, True)):
        r, v = runs[key]
        print(f"{mode} overlap={key}: vp differs in {int((v != vs).sum())} entries", flush=True)
        assert np.array_equal(v.view(np.uint32), vs.view(np.uint32)), key
        assert r.f2max_pm3 == rs.f2max_pm3 and r.f2max_pp == rs.f2max_pp
        assert np.array_equal(r.vmax, rs.vmax)
        assert (r.p_cand, r.p_acc) == (rs.p_cand, rs.p_acc)
        assert np.array_equal(r.ptotal_pm3, rs.ptotal_pm3)
        assert np.array_equal(r.ptotal_pp, rs.ptotal_pp)
