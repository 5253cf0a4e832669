"""Seeded synthetic particle sets for the measured configurations.

* Zel'dovich lattice: Lagrangian lattice at standard-cell centres (as
  utilities/ic.f90:547-551 places particles for np_nc = ratio_cs) displaced
  by the linear displacement of a Gaussian field with
  P(k) ~ k^n_s T_BBKS(k/Gamma)^2, normalised to sigma_8 with a top-hat and
  scaled by the linear growth factor to redshift z.
* Plummer spheres: centres uniform in the box, scale radii uniform in
  [a_min, a_max] fine cells, particle counts drawn from dN/dM ~ M^-1.9 and
  integerised to an exact total.

All positions are in standard cells in [0, ng).  No file or network access.
"""
from __future__ import annotations

import numpy as np

from .geometry import Geometry


def t_bbks(q: np.ndarray) -> np.ndarray:
    q = np.maximum(q, 1e-12)
    return (np.log1p(2.34 * q) / (2.34 * q)
            * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25)


def growth(z: float, om: float = 0.3, ol: float = 0.7) -> float:
    """Linear growth D(z)/D(0) (Carroll, Press & Turner 1992 fit)."""
    def g(a):
        omz = om / (om + ol * a ** 3)
        olz = 1 - omz
        return 2.5 * omz / (omz ** (4 / 7) - olz + (1 + omz / 2) * (1 + olz / 70)) * a
    return g(1 / (1 + z)) / g(1.0)


def _sigma8_norm(ns: float, gamma: float) -> float:
    k = np.logspace(-4, 2, 20000)
    x = 8.0 * k
    w = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
    p = k ** ns * t_bbks(k / gamma) ** 2
    return float(np.trapezoid(k ** 2 * p * w ** 2, k) / (2 * np.pi ** 2))


def zeldovich_lattice(npd: int, geom: Geometry, box_mpc: float, z: float, seed: int,
                      sigma8: float = 0.8, gamma: float = 0.21, ns: float = 0.96):
    """npd^3 particles on a lattice spanning the image, displaced to redshift z.
    Returns (pos [N,3] float64 standard cells, vel [N,3] float32)."""
    rng = np.random.default_rng(seed)
    ng = geom.ng
    spacing = ng / npd                       # standard cells per lattice site
    cell_mpc = box_mpc / npd
    white = rng.standard_normal((npd, npd, npd))
    kf = 2 * np.pi / box_mpc
    k1 = np.fft.fftfreq(npd, d=1.0 / npd) * kf
    k1r = np.fft.rfftfreq(npd, d=1.0 / npd) * kf
    kz, ky, kx = np.meshgrid(k1, k1, k1r, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    amp = sigma8 ** 2 / _sigma8_norm(ns, gamma) * growth(z) ** 2
    pk = np.zeros_like(kk)
    nz = kk > 0
    pk[nz] = amp * kk[nz] ** ns * t_bbks(kk[nz] / gamma) ** 2
    dk = np.fft.rfftn(white) * np.sqrt(pk / cell_mpc ** 3)
    k2 = np.where(nz, kk ** 2, 1.0)
    disp = []
    for kc in (kx, ky, kz):                 # psi_k = i k / k^2 delta_k (x, y, z order)
        psi = np.fft.irfftn(1j * kc / k2 * dk, s=(npd, npd, npd), axes=(0, 1, 2))
        disp.append(psi / (box_mpc / ng))    # Mpc/h -> standard cells
    disp = np.stack([d.reshape(-1) for d in disp], axis=1)
    iz, iy, ix = np.meshgrid(np.arange(npd), np.arange(npd), np.arange(npd), indexing="ij")
    q = (np.stack([ix, iy, iz], axis=-1).reshape(-1, 3) + 0.5) * spacing
    pos = np.mod(q + disp, ng)
    vel = disp.astype(np.float32)
    del cell_mpc
    return pos, vel


def plummer_spheres(n_total: int, n_spheres: int, geom: Geometry, seed: int,
                    a_min_fine: float = 0.5, a_max_fine: float = 3.0, slope: float = -1.9,
                    mass_range: float = 1000.0, trunc: float = 20.0, fine_per_std: int = 2):
    """n_total particles in n_spheres Plummer spheres (standard cells)."""
    rng = np.random.default_rng(seed)
    ng = geom.ng
    centres = rng.uniform(0, ng, size=(n_spheres, 3))
    a = rng.uniform(a_min_fine, a_max_fine, size=n_spheres) / fine_per_std
    # dN/dM ~ M^slope on [1, mass_range] by inverse-CDF sampling
    u = rng.uniform(size=n_spheres)
    e = slope + 1.0
    m = (1 + u * (mass_range ** e - 1)) ** (1 / e)
    share = m / m.sum() * n_total
    cnt = np.floor(share).astype(np.int64)
    rem = n_total - cnt.sum()
    if rem > 0:
        cnt[np.argsort(-(share - cnt), kind="stable")[:rem]] += 1
    sph = np.repeat(np.arange(n_spheres), cnt)
    umax = trunc ** 3 / (trunc ** 2 + 1) ** 1.5
    uu = rng.uniform(0, umax, size=n_total)
    r = a[sph] / np.sqrt(uu ** (-2.0 / 3.0) - 1.0)
    v = rng.standard_normal((n_total, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    pos = np.mod(centres[sph] + v * r[:, None], ng)
    vel = (rng.standard_normal((n_total, 3)) * 0.1).astype(np.float32)
    return pos, vel, cnt


def config0(seed: int = 0, geom: Geometry | None = None):
    """64^3 Zel'dovich lattice at z=49 in a 100 Mpc/h box (one tile)."""
    from .geometry import C0
    g = geom or C0
    return zeldovich_lattice(64, g, 100.0, 49.0, seed)


def config2(seed: int = 2, geom: Geometry | None = None, cap: int = 200_000):
    """Batching workload: per-tile particle counts lognormal (median 4000,
    sigma 1.5, capped at 200k), uniform within each tile of the C2 geometry.
    Returns positions, velocities and the counts (tile order z, y, x)."""
    from .geometry import C2
    g = geom or C2
    rng = np.random.default_rng(seed)
    cnt = np.minimum(np.rint(np.exp(rng.normal(np.log(4000.0), 1.5, g.ntile))), cap).astype(np.int64)
    edge = g.ng / g.nnt
    t = np.repeat(np.arange(g.ntile), cnt)
    corner = np.stack([t % g.nnt, (t // g.nnt) % g.nnt, t // (g.nnt * g.nnt)], 1) * edge
    pos = corner + rng.uniform(0.0, edge, (t.size, 3))
    vel = (rng.standard_normal((t.size, 3)) * 0.1).astype(np.float32)
    return pos, vel, cnt


def config1(seed: int = 1, geom: Geometry | None = None):
    """256^3 particles in 300 Mpc/h: 216^3 Zel'dovich lattice (60%) plus 2000
    Plummer spheres holding the remaining 40%."""
    from .geometry import C1
    g = geom or C1
    p1, v1 = zeldovich_lattice(216, g, 300.0, 49.0, seed)
    n_sph = 256 ** 3 - 216 ** 3
    p2, v2, _ = plummer_spheres(n_sph, 2000, g, seed + 1000)
    return np.concatenate([p1, p2]), np.concatenate([v1, v2])
