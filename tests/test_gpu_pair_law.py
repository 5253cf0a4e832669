"""The pair kernel's force law, pair by pair, through the C ABI
(cube2_pp_debug_pair_forces runs the kernel's own pair_force device
function and acceptance test).

REF must equal the reference's pair force bit for bit: the oracle's
(F_ra(r,app) - F_ra(r,range)) * d / r of pp_force_kernel.c:308-318, whose
F_ra is pinned to the compiled reference by tests/golden/fra_table.npz.
The cases straddle ep = 1 and ep = 2 for both the softening and the range,
and the r^2 gate above which the softening polynomial is skipped.  FAST and
PRECISE have their own rounding and are held to an FP64 evaluation of the
same law (bounds relative to 1/r^2, the scale of each term)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import pair_cases

pytestmark = pytest.mark.gpu

APPS = [0.05, 0.06, 0.01, 0.025, 0.0]


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available()
    from cube2_amd import _lib
    _lib.lib()
    return _lib


def _fra64(r, a):
    """F_ra of basic_functions.f08:95-109 in FP64 (the law, for FAST/PRECISE)."""
    r = np.asarray(r, np.float64)
    a = np.broadcast_to(np.asarray(a, np.float64), r.shape)
    soft = a > 0
    aa = np.where(soft, a, 1.0)
    ep = 2 * r / aa
    hi = (12 / ep ** 2 - 224 + 896 * ep - 840 * ep ** 2 + 224 * ep ** 3 + 70 * ep ** 4
          - 48 * ep ** 5 + 7 * ep ** 6) / (35 * aa * aa)
    lo = (224 * ep - 224 * ep ** 3 + 70 * ep ** 4 + 48 * ep ** 5 - 21 * ep ** 6) / (35 * aa * aa)
    return np.where(~soft | (ep > 2), 1.0 / r ** 2, np.where(ep > 1, hi, lo))


@pytest.mark.parametrize("app", APPS)
def test_ref_pair_forces_bit_identical(lib, oracle_lib, app):
    from cube2_amd.geometry import APM3
    rng = np.random.default_rng(int(app * 1000) + 5)
    d, rg = pair_cases(rng, APM3[1:6], app)
    got = lib.pp_pair_forces(lib.PP_REF, d, rg, app)
    want = oracle_lib.pair_forces(d, rg, app)
    nz = int((want != 0).any(1).sum())
    bad = np.flatnonzero((got.view(np.uint32) != want.view(np.uint32)).any(1))
    print(f"app {app}: {d.shape[0]} pairs, {nz} accepted, {bad.size} differ")
    assert nz > d.shape[0] // 2
    assert bad.size == 0, (d[bad[:5]], rg[bad[:5]], got[bad[:5]], want[bad[:5]])


def test_ref_pair_forces_on_golden_fra_points(lib, oracle_lib):
    """The (r, a) points of the reference's F_ra table as pairs along x: a
    as the range (with the C0 softening) and as the softening (range 1.75)."""
    from conftest import GOLDEN
    z = np.load(GOLDEN / "fra_table.npz")
    r, a = z["r"].astype(np.float32), z["apm"].astype(np.float32)
    d = np.zeros((r.size, 3), np.float32)
    d[:, 0] = r
    for app, rg in ((0.05, a), (0.0, a)):
        keep = rg > 0
        got = lib.pp_pair_forces(lib.PP_REF, d[keep], rg[keep], app)
        want = oracle_lib.pair_forces(d[keep], rg[keep], app)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    for ap in np.unique(a[a > 0])[:40]:
        rg = np.full(r.size, 1.75, np.float32)
        got = lib.pp_pair_forces(lib.PP_REF, d, rg, float(ap))
        want = oracle_lib.pair_forces(d, rg, float(ap))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), float(ap)


def _plummer64(r, eps):
    """r / (r^2 + eps^2)^(3/2) in FP64 (Plummer softening; eps = 0: 1/r^2)."""
    q = r * r + eps * eps
    return r / (q * np.sqrt(q))


@pytest.mark.parametrize("mode,bound", [("fast", 5e-5), ("precise", 1e-6), ("plummer", 1e-6)])
def test_other_arithmetics_against_fp64_law(lib, mode, bound):
    from cube2_amd.geometry import APM3
    m = {"fast": lib.PP_FAST, "precise": lib.PP_PRECISE, "plummer": lib.PP_PLUMMER}[mode]
    worst = 0.0
    for app in APPS:
        rng = np.random.default_rng(int(app * 1000) + 17)
        d, rg = pair_cases(rng, APM3[1:6], app)
        got = lib.pp_pair_forces(m, d, rg, app).astype(np.float64)
        d64 = d.astype(np.float64)
        r = np.sqrt((d64 ** 2).sum(1))
        r2f = (d * d).sum(1, dtype=np.float32)
        acc = (r2f > 0) & (r2f < rg * rg)
        soft = _plummer64(r, app) if mode == "plummer" else _fra64(r, app)
        f = soft - _fra64(r, rg)
        want = np.where(acc[:, None], (f / r)[:, None] * d64, 0.0)
        err = np.abs(got - want).max(1) * r ** 2
        worst = max(worst, float(err[acc].max()))
    print(f"{mode}: worst |error| * r^2 = {worst:.3e}")
    assert worst < bound
