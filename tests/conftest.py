"""Shared test setup: markers, repo on sys.path, golden-vector loaders."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def load_pp_case(name: str) -> dict:
    z = np.load(GOLDEN / f"{name}.npz")
    shape = tuple(int(s) for s in z["rhoc_shape"])
    rhoc = np.zeros(int(np.prod(shape)), np.int32)
    rhoc[z["rhoc_idx"]] = z["rhoc_val"]
    vp_out = z["vp_in"].copy()
    vp_out[z["vp_out_idx"]] = z["vp_out_val"]
    return dict(rhoc=rhoc.reshape(shape), idx_b_r=z["idx_b_r"], xp=z["xp"], vp_in=z["vp_in"],
                mass=float(z["mass"]), a_mid=float(z["a_mid"]), dt=float(z["dt"]),
                vp_out=vp_out, f2max=float(z["f2max"][0]), vmax=z["vmax"],
                ptotal_tile=z["ptotal_tile"])


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def c1_case():
    """The 256^3 particle set (seed 1) laid out for one image, with levels
    selected from its density: (positions, layout, ires, isort).  Built once
    per session on the host (about 15 s)."""
    from cube2_amd.geometry import C1
    from cube2_amd.kick import select_levels
    from cube2_amd.layout import build_layout
    from cube2_amd.synth import config1
    pos, _ = config1(seed=1)
    lay = build_layout(pos, None, C1)
    ires, isort, _ = select_levels(C1, lay.rhoc)
    return pos, lay, ires, isort


def pair_cases(rng, ranges, app):
    """Difference vectors d[n][3] (float32) and per-pair ranges that probe
    the pair law where its rounding is delicate: ep = 2r/a near 1 and 2 for
    both the softening app and the range, r^2 near the softening gate
    a^2 (1 + 2^-17), random pairs inside the range, and small r."""
    ds, rs = [], []

    def add(r, rg):
        r = np.asarray(r, np.float64)
        u = rng.standard_normal((r.size, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        axis = np.zeros((r.size, 3))
        axis[np.arange(r.size), rng.integers(0, 3, r.size)] = rng.choice([-1.0, 1.0], r.size)
        pick = rng.random(r.size) < 0.3
        u[pick] = axis[pick]
        ds.append((r[:, None] * u).astype(np.float32))
        rs.append(np.full(r.size, rg, np.float32))

    k = np.arange(-48, 49, dtype=np.float64) * 2.0 ** -23
    for rg in ranges:
        rg = float(rg)
        n = 1500
        add(rg * rng.random(n) ** (1 / 3), rg)
        add(np.exp(rng.uniform(np.log(2.0 ** -12), np.log(rg), 300)), rg)
        add(rg / 2 * (1 + k), rg)                     # range ep near 1
        add(rg * (1 + k[k < 0]), rg)                  # range ep near 2 (accepted side)
        if app > 0:
            add(app / 2 * (1 + k), rg)                # softening ep near 1
            add(app * (1 + k), rg)                    # softening ep near 2
            add(app * np.sqrt(1 + 2.0 ** -17) * (1 + k), rg)  # the softening gate
            add(app * rng.uniform(0.0, 1.2, 300), rg)
    return np.concatenate(ds), np.concatenate(rs)


def green_at(n, apm_soft, apm_range, gridsize, idx, n_int=2, sinc_pow=6):
    """Green_3D (Green/Green.f90:22-71, kshift 0) in FP64 numpy at the table
    indices idx[m] = (k1, k2, k3) of Gk(n/2+1, n, n), 0-based: the alias sum
    over n in [-n_int, n_int]^3 of U^2(k_n) R(k_n) projected on the
    4-point-gradient D(k), over |D|^2 (sum U^2)^2 gridsize^2; the real(4)
    constants (1./n, 2./15, alpha = 4./3, gridsize^2) rounded as the
    Fortran rounds them.  The 8 corners (::n/2)^3 are 0."""
    import itertools
    f32 = np.float32
    idx = np.asarray(idx, np.int64)
    step = 2 * np.pi * float(f32(1) / f32(n))
    kv = step * (((idx + n // 2) % n) - n // 2).astype(np.float64)      # (m, 3)
    c215 = float(f32(2) / f32(15))
    alpha = float(f32(4) / f32(3))
    oma = float(f32(1) - f32(4) / f32(3))
    s = np.sin(kv / 2)
    u2n = np.prod(1 - s ** 2 + c215 * s ** 4, axis=1)
    D = alpha * np.sin(kv) + oma * np.sin(2 * kv) / 2
    shifts = 2 * np.pi * np.array(list(itertools.product(range(-n_int, n_int + 1), repeat=3)),
                                  np.float64)
    kn = kv[:, None, :] + shifts[None, :, :]
    km = np.sqrt((kn ** 2).sum(2))
    zero = km == 0
    safe = np.where(zero, 1.0, km)
    with np.errstate(divide="ignore", invalid="ignore"):
        half = kn / 2
        sinc = np.where(kn == 0, 1.0, np.sin(half) / np.where(kn == 0, 1.0, half))
        U = np.prod(sinc, axis=2) ** sinc_pow

        def S(a):
            y = a * safe / 2
            return np.where(zero, 1.0, 12 * (2 - 2 * np.cos(y) - y * np.sin(y)) / y ** 4)
        ss = S(apm_soft)
        sr = np.zeros_like(km) if apm_range == 0 else S(apm_range)
        R = kn * np.where(zero, 0.0, (ss ** 2 - sr ** 2) / safe ** 2)[..., None]
        ur = (U[..., None] * R).sum(1)
        gs2 = float(f32(gridsize) * f32(gridsize))
        G = 4 * np.pi * (D * ur).sum(1) / (D ** 2).sum(1) / u2n ** 2 / gs2
    G[np.all(idx % (n // 2) == 0, axis=1)] = 0.0
    return G

