"""The PP oracle pinned against the reference's golden vectors and against
closed forms (CPU only)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN, load_pp_case
from cube2_amd.geometry import APM3, RATIO_SF, REFERENCE_DEFAULT, C0, ixyz3
from cube2_amd.layout import build_layout


def test_fra_matches_reference_table(oracle_lib):
    z = np.load(GOLDEN / "fra_table.npz")
    got = oracle_lib.F_ra(z["r"], z["apm"])
    assert np.array_equal(got.view(np.uint32), z["f"].view(np.uint32))


@pytest.mark.parametrize("a", [0.05, 0.06, 0.4375, 1.75])
def test_fra_pieces_join_continuously(oracle_lib, a):
    # ep = 1 and ep = 2 are the joins of the three pieces; a wrong
    # coefficient shows up as a jump there.
    for ep in (1.0, 2.0):
        r0 = np.float32(ep * a / 2)
        lo = oracle_lib.F_ra([np.nextafter(r0, np.float32(0))], a)[0]
        hi = oracle_lib.F_ra([np.nextafter(r0, np.float32(9))], a)[0]
        assert abs(lo - hi) <= 2e-4 * abs(hi)
    r = np.float32(2 * a)
    assert oracle_lib.F_ra([r], a)[0] == np.float32(1) / (r * r)


def test_stable_sort_matches_reference_radix(oracle_lib):
    z = np.load(GOLDEN / "sort_cases.npz")
    for s in range(4):
        k, bound = z[f"keys{s}"], int(z[f"bound{s}"])
        ks, vs = oracle_lib.stable_sort(k, np.arange(k.size, dtype=np.uint32), bound)
        assert np.array_equal(ks, z[f"skeys{s}"])
        assert np.array_equal(vs, z[f"svals{s}"])
        assert np.array_equal(vs, np.argsort(k, kind="stable"))


@pytest.mark.parametrize("name", ["pp_case1", "pp_case2"])
def test_pp_oracle_bit_exact_vs_reference(oracle_lib, name):
    c = load_pp_case(name)
    g = REFERENCE_DEFAULT
    ires = np.full(g.ns3, 2, np.int32)
    o = oracle_lib.pp_force(g, ires, ixyz3(g), APM3, RATIO_SF, c["rhoc"], c["idx_b_r"], c["xp"],
                            c["vp_in"], c["mass"], c["a_mid"], c["dt"])
    assert np.array_equal(o["vp"].view(np.uint32), c["vp_out"].view(np.uint32))
    assert np.float32(o["f2max"]) == np.float32(c["f2max"])
    assert np.array_equal(o["vmax"], c["vmax"])
    # the reference C kernel counts per parent tile, numbered x-slowest
    per_tile = o["ptotal"].reshape(g.nnt, g.nnt, g.nnt, -1).sum(-1)      # [tz,ty,tx]
    assert np.array_equal(per_tile, c["ptotal_tile"].reshape(g.nnt, g.nnt, g.nnt).transpose(2, 1, 0))


def _two_body(oracle_lib, sep, mass=1.3):
    g = C0
    pos = np.array([[30.3, 31.1, 29.7], [30.3 + sep[0], 31.1 + sep[1], 29.7 + sep[2]]])
    lay = build_layout(pos, np.zeros((2, 3), np.float32), g)
    ires = np.full(g.ns3, 2, np.int32)
    o = oracle_lib.pp_force(g, ires, ixyz3(g), APM3, RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp,
                            lay.vp, mass, 0.02, 1e-3)
    return lay, o


def test_two_body_closed_form(oracle_lib):
    lay, o = _two_body(oracle_lib, (0.9, -0.4, 0.3))
    p = lay.decoded_positions()
    phys = lay.physical_mask()
    i0 = np.flatnonzero(phys & (lay.src == 0))[0]
    i1 = np.flatnonzero(phys & (lay.src == 1))[0]
    # FP64 evaluation of the same two-body kernel
    rng = float(APM3[1])
    app = C0.app
    acc = oracle_lib.direct_pp_f64(p[[i0, i1]], C0.ng, [0], [rng], app)[0]
    f = acc * 1.3 ** 2 * 0.02 * 1e-3 / (6 * np.pi)
    dv0 = o["vp"][i0].astype(np.float64)
    assert np.allclose(dv0, f, rtol=2e-5, atol=1e-12)
    assert np.allclose(o["vp"][i1], -o["vp"][i0], rtol=1e-6)
    assert o["p_acc"] == 2


def test_pair_sum_matches_fp64_direct_sum_and_conserves_momentum(oracle_lib):
    g = C0
    rng = np.random.default_rng(3)
    pos = np.concatenate([rng.uniform(0, 64, (20000, 3)), 32 + rng.normal(0, 1.0, (3000, 3))])
    lay = build_layout(pos, np.zeros((pos.shape[0], 3), np.float32), g)
    ires = np.full(g.ns3, 2, np.int32)
    mass, amid, dt = 1.0, 0.02, 1e-3
    o = oracle_lib.pp_force(g, ires, ixyz3(g), APM3, RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp,
                            lay.vp, mass, amid, dt)
    phys = np.flatnonzero(lay.physical_mask())
    dv = o["vp"][phys].astype(np.float64)
    # every pair is counted twice with opposite signs when all are active
    assert np.abs(dv.sum(0)).max() < 1e-6 * np.abs(dv).sum()
    # FP64 direct sum over the decoded positions of the physical slots
    p = lay.decoded_positions()[phys]
    sub = rng.choice(phys.size, 400, replace=False).astype(np.int32)
    acc = oracle_lib.direct_pp_f64(p, g.ng, sub, np.full(sub.size, float(APM3[1])), g.app)
    ref = acc * mass * mass * amid * dt / (6 * np.pi)
    err = np.sqrt(((dv[sub] - ref) ** 2).sum() / (ref ** 2).sum())
    assert err < 2e-6


def test_level_selection_thresholds_and_isort(oracle_lib):
    g = REFERENCE_DEFAULT.with_(nnt=1, ng=128, nns=4)      # 64 subtiles, ntt=8
    ix = ixyz3(g)
    rhoc = np.zeros((1, g.nte, g.nte, g.nte), np.int32)
    # subtile s gets max count m_s in one physical cell
    maxima = {0: 1535, 1: 1536, 2: 7295, 3: 7296, 4: 20479, 5: 20480, 6: 86399, 7: 86400, 8: 0}
    for s, m in maxima.items():
        sx, sy, sz = ix[s, :3] - 1
        rhoc[0, g.ncb + sz * g.ntt + 3, g.ncb + sy * g.ntt + 2, g.ncb + sx * g.ntt + 1] = m
    ires, isort, napm, npc = oracle_lib.select_levels(g, rhoc, ix)
    want = {0: 2, 1: 3, 2: 3, 3: 4, 4: 4, 5: 5, 6: 5, 7: 6, 8: 2}
    for s, L in want.items():
        assert ires[s] == L, (s, ires[s], L)
    assert napm.sum() == g.ns3
    # isort lists subtiles by non-increasing npc_max (heapsort of -npc_max)
    assert np.all(np.diff(npc[isort - 1]) <= 0)
    assert sorted(isort.tolist()) == list(range(1, g.ns3 + 1))


def test_indexed_sort_orders_and_permutes(oracle_lib):
    # Heapsort is not stable; ties only change the PM3 job order, never a
    # result, so the pin is: a permutation that orders the keys.
    rng = np.random.default_rng(4)
    arr = rng.integers(-6, 0, 301).astype(np.float32)
    idx = oracle_lib.indexed_sort(arr)
    assert np.all(np.diff(arr[idx - 1]) >= 0)
    assert sorted(idx.tolist()) == list(range(1, arr.size + 1))


def test_cell_binned_fp64_direct_sum_equals_all_pairs(oracle_lib):
    """The binned FP64 direct sum visits the same pairs as the all-pairs one
    (periodic minimum image, per-target ranges, a dense clump)."""
    rng = np.random.default_rng(0)
    pos = np.concatenate([rng.uniform(0, 32, (20000, 3)), 16 + rng.normal(0, 0.5, (3000, 3))]) % 32
    tg = rng.choice(pos.shape[0], 400, replace=False).astype(np.int32)
    rg = rng.choice([1.75, 0.875, 0.4375, 0.29], 400)
    a = oracle_lib.direct_pp_f64(pos, 32.0, tg, rg, 0.05)
    b = oracle_lib.direct_pp_f64(pos, 32.0, tg, rg, 0.05, cells=True)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_pair_forces_compose_the_reference_f_ra(oracle_lib):
    """or_pair_forces (the checker of the GPU pair law) against the pair
    expression of pp_force_kernel.c:308-318 evaluated in float32 numpy with
    F_ra taken from the compiled reference (oracle/_ref; where it is not
    built, from the oracle's F_ra, which the golden table pins)."""
    from conftest import pair_cases
    if oracle_lib.ref_available():
        R = oracle_lib.ref()
        fra = lambda r, a: np.float32(R.F_ra(float(r), float(a)))  # noqa: E731
    else:
        fra = lambda r, a: oracle_lib.F_ra([r], a)[0]  # noqa: E731
    rng = np.random.default_rng(31)
    for app in (0.05, 0.0):
        d, rg = pair_cases(rng, APM3[1:6:2], app)
        sel = rng.choice(d.shape[0], 3000, replace=False)
        d, rg = d[sel], rg[sel]
        got = oracle_lib.pair_forces(d, rg, app)
        f32 = np.float32
        for q in range(d.shape[0]):
            x, y, z = d[q]
            r2 = f32(f32(f32(x * x) + f32(y * y)) + f32(z * z))
            want = np.zeros(3, np.float32)
            if f32(0) < r2 < f32(rg[q] * rg[q]):
                rm = np.sqrt(r2)
                fpp = f32(fra(rm, app) - fra(rm, rg[q]))
                want = np.array([f32(fpp * c) / rm for c in (x, y, z)], np.float32)
            assert np.array_equal(got[q].view(np.uint32), want.view(np.uint32)), (q, d[q], rg[q])


def test_plummer_direct_sum_follows_its_potential(oracle_lib):
    """The oracle's Plummer law (law="plummer"): for one pair at distance r,
    the acceleration minus that of the unsoftened S2 law (app = 0) is
    F_pl(r) - 1/r^2 along the pair, with the attraction F_pl = dPhi/dr of
    Phi = -1/sqrt(r^2 + eps^2) taken here by central differences of the
    potential; at eps -> 0 it
    is the unsoftened law; the force peaks at r = eps/sqrt(2) with
    2/(3 sqrt(3) eps^2)."""
    eps, rg = 0.05, 1.75
    box = 32.0
    phi = lambda r: -1.0 / np.sqrt(r * r + eps * eps)  # noqa: E731
    rs = np.concatenate([np.linspace(0.004, 0.3, 60), np.linspace(0.3, 1.7, 30)])
    for r in rs:
        u = np.array([0.6, -0.48, 0.64])
        p = np.array([[10.0, 11.0, 12.0], [10.0, 11.0, 12.0] + r * u])
        a_pl = oracle_lib.direct_pp_f64(p, box, [0], [rg], eps, law="plummer")[0]
        a_0 = oracle_lib.direct_pp_f64(p, box, [0], [rg], 0.0)[0]
        h = 1e-5 * r
        f_fd = (phi(r + h) - phi(r - h)) / (2 * h)
        want = (f_fd - 1.0 / r ** 2) * u
        assert np.allclose(a_pl - a_0, want, rtol=1e-6, atol=1e-9 / r ** 2), r
        a_tiny = oracle_lib.direct_pp_f64(p, box, [0], [rg], 1e-9, law="plummer")[0]
        assert np.allclose(a_tiny, a_0, rtol=0, atol=1e-12 / r ** 2)
    # peak of the softened force (the S2 range term is smooth there)
    r = np.linspace(0.6, 0.8, 2001) * eps
    f = r / (r * r + eps * eps) ** 1.5
    assert abs(r[np.argmax(f)] - eps / np.sqrt(2)) < 2e-4 * eps
    assert np.isclose(f.max(), 2 / (3 * np.sqrt(3) * eps * eps), rtol=1e-7)
    p = np.array([[10.0, 11.0, 12.0], [10.0 + eps / np.sqrt(2), 11.0, 12.0]])
    a_pl = oracle_lib.direct_pp_f64(p, box, [0], [rg], eps, law="plummer")[0]
    a_0 = oracle_lib.direct_pp_f64(p, box, [0], [rg], 0.0)[0]
    assert np.isclose(a_pl[0] - a_0[0], 2 / (3 * np.sqrt(3) * eps * eps) - 2 / eps ** 2, rtol=1e-12)
